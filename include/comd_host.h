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
 *   command_t             src-flat/include/mycommand.h:60-75
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
#include "comd_dropin.h"

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
} command_t;

/* Settings of this library's driver beyond the reference's command_t (kept
 * out of it, so that command_t, passed by value to create_fcc_lattice and
 * fromFileGzip, has the reference's layout).  No counterpart in the
 * reference.  parseCommandLine fills the process-wide copy from the options
 * --precision, --eval, --seed, --disp, --niter, --nsteps and --lj;
 * initSimFromCmdLine and the driver read it. */
typedef struct comd_run_options {
    int precision;         /* COMD_PREC_FP64 / COMD_PREC_MIXED */
    int cheby;             /* evaluate EAM through the Chebyshev fit */
    unsigned long long seed;
    double disp;           /* Gaussian displacement sigma, A */
    int niter, nsteps;     /* do_compute_work loop shape (20 x 10 in the reference) */
    double lj_sigma, lj_epsilon, lj_cutoff, lj_mass; /* 0: the reference's defaults */
} comd_run_options;

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
/* The reference's 0 K lattice: atoms at rest, type 1, ids in creation
 * order (cmd.temp is not used, as in ic_fcc.c). */
simflat_t *create_fcc_lattice(command_t cmd, pmd_base_potential_t *pot);
simflat_t *fromFileGzip(command_t cmd, pmd_base_potential_t *pot);

/* ---- output files (src-flat/write.c:112-220) ---- */
/* The reference's writers: clsman binary (Fortran records: natoms, an empty
 * comment, the bounds, then x y z type px py pz per atom in id order) and
 * the gz text read by fromFileGzip (%.8lf).  Positions are absolute (cell
 * corner + r), types are written as 1.  Ids must be 0 .. ntot-1; errors go
 * to comd_host_status() (the reference writes to stderr or exits). */
void writeClsman(simflat_t *s, char *fname);
void gzASCIIWrite(simflat_t *s, char *fname);

/* ---- hot path on the GPU ---- */
int eamForce(void *s);   /* force slot for EAM potentials, src-flat/eam.c:438 */
int LJ(void *s);         /* force slot for LJ, src-flat/ljForce.c:89 */
int computeForce(simflat_t *s);                      /* src-flat/timestep.c:110 */
double nTimeSteps(int n, simflat_t *s, real_t dt);   /* src-flat/timestep.c:114 */
void reBoxAll(simflat_t *s);                         /* src-flat/domains.c:347 */

/* The drop-in adapters comd_hip_eam_force, comd_hip_lj_force,
 * comd_hip_nTimeSteps and comd_hip_reBoxAll are declared in comd_dropin.h
 * (libcomd_dropin.so, a dependency of this library). */

/* ---- driver (src-flat/docompute.c:68-218, src-flat/mycommand.c:108-161) ---- */
/* As the reference's (mycommand.h:78, void); an unknown option or a bad
 * value is reported through comd_host_status() instead of exit(). */
void parseCommandLine(command_t *cmd, int argc, char **argv);
simflat_t *initSimFromCmdLine(int argc, char **argv);
void *do_compute_work(SimThreadData *data);

/* ---- additions ---- */
/* The driver settings parseCommandLine parsed last (defaults before any
 * call: FP64, tables, seed 1, 20 x 10 steps), and a setter for callers that
 * do not parse a command line. */
const comd_run_options *comd_host_run_options(void);
void comd_host_set_run_options(const comd_run_options *opts);
/* The FCC lattice of create_fcc_lattice with seeded Gaussian displacements
 * (sigma disp, A) and Maxwell-Boltzmann momenta at temp kelvin
 * (comd_seed_state), binned in id order as putAtomInBox would be called for
 * the displaced positions.  No counterpart in the reference (its lattice is
 * always at rest, and -t is parsed but unused). */
simflat_t *comd_host_seeded_lattice(command_t cmd, pmd_base_potential_t *pot, unsigned long long seed, double disp);
/* comd_host_status / comd_host_error (comd_dropin.h): status of the last
 * failed call of this library or of the adapters (0 if none). */
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
/* Read a clsman file (writeClsman's format) into arrays, as comd_read_gz.
 * No counterpart in the reference, which writes the format but has no
 * reader for it. */
int comd_read_clsman(const char *path, int *n, double bounds[3], double **xyz, double **pxyz);

#ifdef __cplusplus
}
#endif
#endif
