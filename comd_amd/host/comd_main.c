/*
 * comd_main.c -- command-line driver with the shape of src-flat/pmd.c:105-130:
 * initialise from the command line, run do_compute_work, destroy.
 */
#include <stdio.h>

#include "comd_host.h"

int main(int argc, char **argv)
{
    simflat_t *sim = initSimFromCmdLine(argc, argv);
    if (!sim) {
        fprintf(stderr, "initialisation failed (%d): %s\n", comd_host_status(), comd_host_error());
        return comd_host_status() ? comd_host_status() : 1;
    }
    const comd_run_options *o = comd_host_run_options();
    SimThreadData data = {0};
    data.sim = sim;
    data.niter = o->niter;
    data.nsteps = o->nsteps;
    data.eam_flag = sim->pot->force == eamForce;
    do_compute_work(&data);
    int rc = comd_host_status();
    if (rc) fprintf(stderr, "error (%d): %s\n", rc, comd_host_error());
    destroySimulation(&sim);
    return rc;
}
