/*
 * comd_dropin.h -- the drop-in entry points of comd_amd/libcomd_dropin.so,
 * for a program built from CoMD src-flat's own sources.
 *
 * Only void pointers cross this header, so it can be included beside the
 * reference's headers (whose simflat_t it would otherwise clash with): the
 * pointer is the reference's simflat_t * (src-flat/include/pmdTypes.h:104-141)
 * of its double-precision build (real_t = double).  The library mirrors
 * the simulation's slot arrays to the GPU, runs there and writes the
 * results back into the same slots and scalars (f, rho, fi, e, stress; r, p,
 * natoms, id, iType after steps and rebinning), so the reference's callers
 * see what its own functions would have left.  Errors are recorded (see
 * comd_dropin_status) instead of exiting the process.
 */
#ifndef COMD_DROPIN_H
#define COMD_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

/* Force slots with the signature of pmd_base_potential_t::force
 * (src-flat/include/pmdTypes.h:77): replace eamForce (set at
 * src-flat/eam.c:394) and LJ (set at src-flat/ljForce.c:79).  Return 0, or
 * a COMD_ERR_* code of comd_hip.h (the reference ignores the value). */
int comd_hip_eam_force(void *simflat);
int comd_hip_lj_force(void *simflat);

/* nTimeSteps (src-flat/timestep.c:114-140), called from do_compute_work
 * (src-flat/docompute.c:207): n steps and the trailing force, all on the
 * GPU; returns the potential energy s->e. */
double comd_hip_nTimeSteps(int n, void *simflat, double dt);

/* reBoxAll (src-flat/domains.c:347-389) with the same cell assignment and
 * in-cell order as the reference's sequential scan. */
int comd_hip_reBoxAll(void *simflat);

/* Status of the last failed adapter call (0 if none) and its text. */
int comd_host_status(void);
const char *comd_host_error(void);

#ifdef __cplusplus
}
#endif
#endif
