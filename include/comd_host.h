/*
 * comd_host.h -- C host library (comd_amd/libcomd_host.so) that keeps the
 * simulation and potential entry points of CoMD src-flat and runs their hot
 * path on the GPU through the C-ABI of comd_hip.h.
 *
 * The types below are layout-compatible with the reference's double-precision
 * build (real_t = double, no in-situ visualisation):
 *   simflat_t             src-flat/include/pmdTypes.h:104-141
 *   pmd_base_potential_t  src-flat/include/pmdTypes.h:74-79
 *   potentialarray_t      src-flat/include/eamTypes.h:68-74
 *   eampotential_t        src-flat/include/eamTypes.h:90-99
 *   eam_cheby_t           src-flat/include/eamTypes.h:102-112
 *   ljpotential_t         src-flat/include/ljTypes.h:67-76
 *   command_t             src-flat/include/mycommand.h:52-67 (+ new fields at the end)
 * so that comd_hip_eam_force / comd_hip_lj_force can be stored in a reference
 * potential's force slot and comd_hip_nTimeSteps can replace nTimeSteps.
 *
 * No function here computes forces on the CPU.  Setup (table reading,
 * Chebyshev fit, lattice, input files) runs on the host as in the reference;
 * force, rebinning and integration run on the GPU.  Errors are returned as
 * COMD_ERR_* codes (comd_hip.h) instead of exiting the process.
 */
#ifndef COMD_HOST_H
#define COMD_HOST_H

#include "comd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef double real_t;
typedef real_t real3[3];
typedef real_t real4[4];

#define MAXATOMS 64

typedef struct pmd_base_potential_t {
    real_t cutoff;
    real_t mass;
    int (*force)(void *s);
    void (*destroy)(void *pot);
} pmd_base_potential_t;

typedef struct potentialarray_t {
    int n;
    real_t x0, xn, invDx;
    real_t *values; /* values[-1 .. n+1] valid */
} potentialarray_t;

typedef struct eampotential_t {
    real_t cutoff;
    real_t mass;
    int (*force)(void *s);
    void (*destroy)(void **pot);
    potentialarray_t *phi, *rho, *f;
    real_t lat;
} eampotential_t;

typedef struct eam_cheby_t {
    real_t cutoff;
    real_t mass;
    potentialarray_t *phi, *rho, *f, *dphi, *drho, *df;
    real_t lat;
} eam_cheby_t;

typedef struct ljpotential_t {
    real_t cutoff;
    real_t mass;
    int (*force)(void *s);
    void (*destroy)(void **pot);
    real_t sigma;
    real_t epsilon;
} ljpotential_t;

typedef struct simflat_t {
    int stateflag;
    int nbx[3];
    int nboxes;
    int ntot;
    real3 bounds;
    real3 boxsize;
    real3 *dcenter;
    int *natoms;
    int *id;
    int *iType;
    real_t *mass;
    real3 *r;
    real3 *p;
    real4 *f;
    real_t *rho;
    real_t *fi;
    real_t e;
    int nAtoms;
    pmd_base_potential_t *pot;
    char *comment;
    real_t stress;
    real_t defgrad;
    real_t bf;
    eam_cheby_t *ch_pot;
} simflat_t;

typedef struct command_t {
    char filename[1024];
    char potdir[1024];
    char potname[1024];
    int debug;
    int doeam;
    int periodic;
    int usegpu;
    int nx, ny, nz;
    double bf;
    double lat;
    double defgrad;
    double temp;
    /* additions of this library */
    int precision;         /* COMD_PREC_FP64 / COMD_PREC_MIXED */
    int cheby;             /* evaluate EAM through the Chebyshev fit */
    unsigned long long seed;
    double disp;           /* Gaussian displacement sigma, A */
    int niter, nsteps;     /* do_compute_work loop shape (20 x 10 in the reference) */
    double lj_sigma, lj_epsilon, lj_cutoff, lj_mass; /* 0: the reference's defaults */
} command_t;

typedef struct SimThreadData {
    void *viz;
    simflat_t *sim;
    int niter;
    int nsteps;
    int eam_flag;
    int gpu_flag;
} SimThreadData;

/* The reference's global periodic-boundary flag (src-flat/mycommand.c:67). */
extern int PERIODIC;

/* ---- potentials (src-flat/eam.c:102-398, src-flat/ljForce.c:75-86,
 *      src-flat/cheby.c:99-270) ---- */
pmd_base_potential_t *setEamPot(const char *dir, const char *name);
ljpotential_t *getLJPot(void);
eam_cheby_t *setChebPot(eampotential_t *pot, int n);
void destroyChebPot(eam_cheby_t **ch);

/* ---- link cells (src-flat/domains.c) ---- */
simflat_t *blankSimulation(pmd_base_potential_t *pot);
int allocDomains(simflat_t *s);
int putAtomInBox(simflat_t *s, int id, char flagMove, int iType, real_t mass, double x, double y, double z,
                 double px, double py, double pz, double fx, double fy, double fz);
void destroySimulation(simflat_t **ps);

/* ---- initial conditions (src-flat/ic_fcc.c:63-141, src-flat/read.c:155-241) ---- */
simflat_t *create_fcc_lattice(command_t cmd, pmd_base_potential_t *pot);
simflat_t *fromFileGzip(command_t cmd, pmd_base_potential_t *pot);

/* ---- hot path on the GPU ---- */
int eamForce(void *s);   /* force slot for EAM potentials, src-flat/eam.c:438 */
int LJ(void *s);         /* force slot for LJ, src-flat/ljForce.c:89 */
int computeForce(simflat_t *s);                      /* src-flat/timestep.c:110 */
double nTimeSteps(int n, simflat_t *s, real_t dt);   /* src-flat/timestep.c:114 */
void reBoxAll(simflat_t *s);                         /* src-flat/domains.c:347 */

/* Drop-in adapters for a reference build (same signatures as the force
 * slot and nTimeSteps); they mirror the simflat_t arrays to the device,
 * run there and write the results back into the same slots. */
int comd_hip_eam_force(void *simflat);
int comd_hip_lj_force(void *simflat);
double comd_hip_nTimeSteps(int n, void *simflat, double dt);
int comd_hip_reBoxAll(void *simflat);

/* ---- driver (src-flat/docompute.c:68-218, src-flat/mycommand.c:108-161) ---- */
int parseCommandLine(command_t *cmd, int argc, char **argv);
simflat_t *initSimFromCmdLine(int argc, char **argv);
void *do_compute_work(SimThreadData *data);

/* ---- additions ---- */
/* Status of the last failed call of this library (0 if none). */
int comd_host_status(void);
const char *comd_host_error(void);
/* Device context bound to a simulation (created on first use). */
comd_ctx *comd_host_ctx(simflat_t *s);
int comd_host_set_precision(simflat_t *s, int precision);
int comd_host_set_eval(simflat_t *s, int cheby);
/* Upload the host arrays to the device (done by the adapters on every call;
 * the driver does it once and then keeps the state device-resident). */
int comd_host_upload(simflat_t *s);
int comd_host_download(simflat_t *s);
/* Fill a comd_params from a simulation and its potential. */
int comd_host_params(const simflat_t *s, comd_params *p);

/* FCC positions in the order of create_fcc_lattice (src-flat/ic_fcc.c:115-132);
 * returns the atom count, writes at most max atoms (3 doubles each). */
int comd_fcc_positions(int nx, int ny, int nz, double lat, double defgrad, double *xyz, int max);
/* Seeded Gaussian displacements (sigma disp, A) and Maxwell-Boltzmann momenta
 * at temp kelvin for atoms of mass amu, with zero net momentum.  Streams are
 * SplitMix64 of (seed, atom index). */
void comd_seed_state(unsigned long long seed, int n, double mass_amu, double temp, double disp, double *xyz,
                     double *pxyz);
/* Write atoms in the reference's gz input format (src-flat/read.c:155-241)
 * at %.17g so that a reader recovers the doubles exactly. */
int comd_write_gz(const char *path, int n, const double bounds[3], const double *xyz, const double *pxyz);
/* Read the gz input format into arrays (n atoms, bounds); caller frees
 * *xyz and *pxyz with free(). */
int comd_read_gz(const char *path, int *n, double bounds[3], double **xyz, double **pxyz);

#ifdef __cplusplus
}
#endif
#endif
