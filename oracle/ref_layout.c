/*
 * ref_layout.c -- prints the size and field offsets of the structs that
 * cross the drop-in boundary (simflat_t, pmd_base_potential_t,
 * potentialarray_t, eampotential_t, eam_cheby_t, ljpotential_t, command_t,
 * SimThreadData) as JSON.  TEST INFRASTRUCTURE ONLY.
 *
 * Compiled against the reference's headers (-DREF_HEADERS, by
 * oracle/Makefile target ref, into _ref/ref_layout) it records the
 * reference's layout (src-flat/include/pmdTypes.h:74-79,104-141,154-162,
 * eamTypes.h:68-112, ljTypes.h:72-81, mycommand.h:60-75) as a fixture;
 * compiled against include/comd_host.h by the tests it gives this library's,
 * and the two must be equal.
 */
#include <stddef.h>
#include <stdio.h>

#ifdef REF_HEADERS
#include "pmd.h"
#include "eam.h"
#include "ljTypes.h"
#include "mycommand.h"
#else
#include "comd_host.h"
#endif

static int first = 1;
#define SZ(T) (printf("%s\"%s\": {\"size\": %zu", first ? "" : ",\n", #T, sizeof(T)), first = 0)
#define OF(T, f) printf(", \"%s\": %zu", #f, offsetof(T, f))
#define END() printf("}")

int main(void)
{
    printf("{\n");
    SZ(simflat_t);
    OF(simflat_t, nbx); OF(simflat_t, nboxes); OF(simflat_t, ntot); OF(simflat_t, bounds);
    OF(simflat_t, boxsize); OF(simflat_t, dcenter); OF(simflat_t, natoms); OF(simflat_t, id);
    OF(simflat_t, iType); OF(simflat_t, mass); OF(simflat_t, r); OF(simflat_t, p); OF(simflat_t, f);
    OF(simflat_t, rho); OF(simflat_t, fi); OF(simflat_t, e); OF(simflat_t, nAtoms); OF(simflat_t, pot);
    OF(simflat_t, comment); OF(simflat_t, stress); OF(simflat_t, defgrad); OF(simflat_t, bf);
    OF(simflat_t, ch_pot);
    END();
    SZ(pmd_base_potential_t);
    OF(pmd_base_potential_t, cutoff); OF(pmd_base_potential_t, mass); OF(pmd_base_potential_t, force);
    OF(pmd_base_potential_t, destroy);
    END();
    SZ(potentialarray_t);
    OF(potentialarray_t, n); OF(potentialarray_t, x0); OF(potentialarray_t, xn); OF(potentialarray_t, invDx);
    OF(potentialarray_t, values);
    END();
    SZ(eampotential_t);
    OF(eampotential_t, cutoff); OF(eampotential_t, mass); OF(eampotential_t, force); OF(eampotential_t, destroy);
    OF(eampotential_t, phi); OF(eampotential_t, rho); OF(eampotential_t, f); OF(eampotential_t, lat);
    END();
    SZ(eam_cheby_t);
    OF(eam_cheby_t, cutoff); OF(eam_cheby_t, mass); OF(eam_cheby_t, phi); OF(eam_cheby_t, rho);
    OF(eam_cheby_t, f); OF(eam_cheby_t, dphi); OF(eam_cheby_t, drho); OF(eam_cheby_t, df); OF(eam_cheby_t, lat);
    END();
    SZ(ljpotential_t);
    OF(ljpotential_t, cutoff); OF(ljpotential_t, mass); OF(ljpotential_t, force); OF(ljpotential_t, destroy);
    OF(ljpotential_t, sigma); OF(ljpotential_t, epsilon);
    END();
    SZ(command_t);
    OF(command_t, filename); OF(command_t, potdir); OF(command_t, potname); OF(command_t, debug);
    OF(command_t, doeam); OF(command_t, periodic); OF(command_t, usegpu); OF(command_t, nx); OF(command_t, ny);
    OF(command_t, nz); OF(command_t, bf); OF(command_t, lat); OF(command_t, defgrad); OF(command_t, temp);
    END();
    SZ(SimThreadData);
    OF(SimThreadData, viz); OF(SimThreadData, sim); OF(SimThreadData, niter); OF(SimThreadData, nsteps);
    OF(SimThreadData, eam_flag); OF(SimThreadData, gpu_flag);
    END();
    printf("\n}\n");
    return 0;
}
