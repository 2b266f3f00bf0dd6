/*
 * ref_harness.c -- drives the compiled CoMD src-flat reference (built from
 * its own sources under /root/reference by oracle/Makefile, never copied)
 * and dumps its state in a flat binary format, so that the oracle
 * restatement and the HIP path can be checked against the real thing.
 * TEST INFRASTRUCTURE ONLY.
 *
 * It calls the reference's own entry points: setEamPot (src-flat/eam.c:102),
 * getLJPot (src-flat/ljForce.c:75), create_fcc_lattice (src-flat/ic_fcc.c:63),
 * fromFileGzip (src-flat/read.c:155), setChebPot (src-flat/cheby.c:99),
 * reBoxAll (src-flat/domains.c:347), computeForce (src-flat/timestep.c:110)
 * and nTimeSteps (src-flat/timestep.c:114), in the order of
 * initSimFromCmdLine / do_compute_work (src-flat/docompute.c:68-218).
 *
 * usage:
 *   ref_harness (--eam DIR NAME | --lj SIGMA EPS CUTOFF MASS)
 *               (--fcc NX NY NZ LAT DEFGRAD | --input FILE)
 *               [--bf BF] [--noperiodic] [--steps K] [--dumps M]
 *               [--time] [--clsman FILE] [--gzwrite FILE] --out PREFIX
 * writes PREFIX.0.bin after the initial force and PREFIX.m.bin after each
 * further nTimeSteps(K) call; with --eam also PREFIX.pot.bin (tables and
 * Chebyshev coefficients).  --time prints the wall time of the step calls.
 * --clsman / --gzwrite write the final state with the reference's own
 * writeClsman / gzASCIIWrite (src-flat/write.c:112-220).
 *
 * Built with -DCOMD_DROPIN (oracle/_ref/ref_dropin, linked with
 * comd_amd/libcomd_dropin.so) it also takes
 *   --hip-force   store comd_hip_eam_force / comd_hip_lj_force in the
 *                 potential's force slot (the edit of eam.c:394 /
 *                 ljForce.c:79), so the reference's own computeForce and
 *                 nTimeSteps call the GPU force;
 *   --hip-steps   call comd_hip_nTimeSteps instead of nTimeSteps (the edit
 *                 of docompute.c:207): the whole step runs on the GPU.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "pmd.h"
#include "cheby.h"
#include "ic_fcc.h"
#include "read.h"
#include "eam.h"

#ifdef COMD_DROPIN
#include "comd_dropin.h"
#endif

extern int PERIODIC;
extern void gzASCIIWrite(simflat_t *s, char *fname); /* src-flat/write.c:175 (no header declares it) */

static double now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static void dump_state(simflat_t *s, const char *prefix, int idx)
{
    char path[2048];
    snprintf(path, sizeof path, "%s.%d.bin", prefix, idx);
    FILE *fp = fopen(path, "wb");
    if (!fp) { perror(path); exit(3); }
    fwrite("COMDREF1", 1, 8, fp);
    int hdr[5] = {s->nbx[0], s->nbx[1], s->nbx[2], s->nboxes, s->ntot};
    fwrite(hdr, sizeof(int), 5, fp);
    double dh[8] = {s->bounds[0], s->bounds[1], s->bounds[2], s->boxsize[0], s->boxsize[1], s->boxsize[2],
                    (double)s->e, (double)s->stress};
    fwrite(dh, sizeof(double), 8, fp);
    fwrite(s->natoms, sizeof(int), s->nboxes, fp);
    for (int c = 0; c < s->nboxes; c++)
        for (int i = 0; i < s->natoms[c]; i++) {
            int o = c * MAXATOMS + i;
            int ii[2] = {s->id[o], s->iType[o]};
            fwrite(ii, sizeof(int), 2, fp);
            double rec[12] = {s->r[o][0], s->r[o][1], s->r[o][2], s->p[o][0], s->p[o][1], s->p[o][2],
                              s->f[o][0], s->f[o][1], s->f[o][2], s->f[o][3], s->rho[o], s->fi[o]};
            fwrite(rec, sizeof(double), 12, fp);
        }
    fclose(fp);
}

static void dump_array(FILE *fp, potentialarray_t *a, int with_guards)
{
    int n = a->n;
    double h[3] = {a->x0, a->xn, with_guards ? a->invDx : 0.0};
    fwrite(&n, sizeof(int), 1, fp);
    fwrite(h, sizeof(double), 3, fp);
    if (with_guards) fwrite(a->values - 1, sizeof(double), n + 3, fp);
    else fwrite(a->values, sizeof(double), n, fp);
}

static void dump_pot(eampotential_t *pot, eam_cheby_t *ch, const char *prefix)
{
    char path[2048];
    snprintf(path, sizeof path, "%s.pot.bin", prefix);
    FILE *fp = fopen(path, "wb");
    if (!fp) { perror(path); exit(3); }
    double mh[3] = {pot->cutoff, pot->mass, pot->lat};
    fwrite(mh, sizeof(double), 3, fp);
    dump_array(fp, pot->phi, 1);
    dump_array(fp, pot->rho, 1);
    dump_array(fp, pot->f, 1);
    dump_array(fp, ch->phi, 0);
    dump_array(fp, ch->rho, 0);
    dump_array(fp, ch->f, 0);
    dump_array(fp, ch->dphi, 0);
    dump_array(fp, ch->drho, 0);
    dump_array(fp, ch->df, 0);
    fclose(fp);
}

int main(int argc, char **argv)
{
    command_t cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.bf = 1.0;
    cmd.defgrad = 1.0;
    int doeam = 0, dolj = 0, dofcc = 0, steps = 10, dumps = 0, dotime = 0, hipforce = 0, hipsteps = 0;
    double ljs = 0, lje = 0, ljc = 0, ljm = 0;
    const char *prefix = NULL;
    char *clsman = NULL, *gzwrite = NULL;
    char potdir[1024] = "", potname[1024] = "";
    PERIODIC = 1;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--eam") && a + 2 < argc) {
            doeam = 1;
            snprintf(potdir, sizeof potdir, "%s", argv[++a]);
            snprintf(potname, sizeof potname, "%s", argv[++a]);
        } else if (!strcmp(argv[a], "--lj") && a + 4 < argc) {
            dolj = 1;
            ljs = atof(argv[++a]); lje = atof(argv[++a]); ljc = atof(argv[++a]); ljm = atof(argv[++a]);
        } else if (!strcmp(argv[a], "--fcc") && a + 5 < argc) {
            dofcc = 1;
            cmd.nx = atoi(argv[++a]); cmd.ny = atoi(argv[++a]); cmd.nz = atoi(argv[++a]);
            cmd.lat = atof(argv[++a]); cmd.defgrad = atof(argv[++a]);
        } else if (!strcmp(argv[a], "--input") && a + 1 < argc) {
            snprintf(cmd.filename, sizeof cmd.filename, "%s", argv[++a]);
        } else if (!strcmp(argv[a], "--bf") && a + 1 < argc) {
            cmd.bf = atof(argv[++a]);
        } else if (!strcmp(argv[a], "--noperiodic")) {
            PERIODIC = 0;
        } else if (!strcmp(argv[a], "--steps") && a + 1 < argc) {
            steps = atoi(argv[++a]);
        } else if (!strcmp(argv[a], "--dumps") && a + 1 < argc) {
            dumps = atoi(argv[++a]);
        } else if (!strcmp(argv[a], "--time")) {
            dotime = 1;
        } else if (!strcmp(argv[a], "--out") && a + 1 < argc) {
            prefix = argv[++a];
        } else if (!strcmp(argv[a], "--clsman") && a + 1 < argc) {
            clsman = argv[++a];
        } else if (!strcmp(argv[a], "--gzwrite") && a + 1 < argc) {
            gzwrite = argv[++a];
#ifdef COMD_DROPIN
        } else if (!strcmp(argv[a], "--hip-force")) {
            hipforce = 1;
        } else if (!strcmp(argv[a], "--hip-steps")) {
            hipforce = hipsteps = 1;
#endif
        } else {
            fprintf(stderr, "bad argument %s\n", argv[a]);
            return 2;
        }
    }
    if ((!doeam && !dolj) || (!dofcc && !cmd.filename[0]) || (!prefix && !dotime)) {
        fprintf(stderr, "missing arguments\n");
        return 2;
    }
    cmd.periodic = PERIODIC;

    struct pmd_base_potential_t *pot;
    if (doeam) {
        pot = setEamPot(potdir, potname);
    } else {
        ljpotential_t *lj = getLJPot();
        lj->sigma = ljs;
        lj->epsilon = lje;
        lj->cutoff = ljc;
        lj->mass = ljm;
        pot = (struct pmd_base_potential_t *)lj;
    }
#ifdef COMD_DROPIN
    if (hipforce) pot->force = doeam ? comd_hip_eam_force : comd_hip_lj_force;
#endif
    simflat_t *sim = cmd.filename[0] ? fromFileGzip(cmd, pot) : create_fcc_lattice(cmd, pot);
    if (!sim) { fprintf(stderr, "no simulation\n"); return 3; }
    if (doeam) {
        sim->ch_pot = setChebPot((eampotential_t *)pot, 32);
        if (prefix) dump_pot((eampotential_t *)pot, sim->ch_pot, prefix);
    }
    reBoxAll(sim);
    double tf0 = now();
    computeForce(sim);
    double tf = now() - tf0;
    if (prefix) dump_state(sim, prefix, 0);
    double t0 = now();
    for (int m = 1; m <= dumps; m++) {
#ifdef COMD_DROPIN
        if (hipsteps) comd_hip_nTimeSteps(steps, sim, 1.0e-15);
        else
#endif
            nTimeSteps(steps, sim, 1.0e-15);
        if (prefix) dump_state(sim, prefix, m);
    }
#ifdef COMD_DROPIN
    if (comd_host_status()) {
        fprintf(stderr, "drop-in error %d: %s\n", comd_host_status(), comd_host_error());
        return 4;
    }
#endif
    double t1 = now();
    if (clsman) writeClsman(sim, clsman);
    if (gzwrite) gzASCIIWrite(sim, gzwrite);
    /* atom-steps/s excludes the extra force of each nTimeSteps call
     * (src-flat/timestep.c:136), estimated by the timed initial force */
    if (dotime)
        printf("REFTIME atoms=%d calls=%d steps=%d seconds=%.6f force_seconds=%.6f atom_steps_per_s=%.6e "
               "us_per_atom_step_ref_print=%.6f\n",
               sim->ntot, dumps, steps, t1 - t0, tf,
               (double)sim->ntot * steps * dumps / ((t1 - t0) - dumps * tf),
               1e6 * (t1 - t0) / ((double)dumps * (steps + 1) * sim->ntot));
    printf("REFDONE e=%.17g stress=%.17g ntot=%d\n", (double)sim->e, (double)sim->stress, sim->ntot);
    return 0;
}
