/*
 * oracle.h -- CPU oracle for the link-cell force and timestep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a fresh FP64 restatement of the
 * semantics of CoMD's src-flat force/timestep path, used as the checker
 * for the HIP implementation in comd_amd/.  Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it.  The product never
 * calls into it.
 *
 * Every function follows the reference file:line cited beside it, with
 * the same loop predicates and the same floating-point operation order,
 * compiled with -ffp-contract=off so that the results are bit-identical
 * with the reference built in double precision on x86-64.
 *
 * Pinning (what checks each part):
 *   table reader      decoded headers + guard slots vs the compiled reference
 *   interpolation     closed form at grid points, derivative-zero rule, clamps
 *   Chebyshev         node interpolation property, derivative vs finite
 *                     difference, Clenshaw vs direct sum, coefficient bits vs
 *                     the compiled reference
 *   EAM / LJ forces   perfect-FCC closed form, Newton-3, two-atom closed form,
 *                     per-atom bit equality with the compiled reference
 *   binning/rebox     brute force on tiny inputs, per-slot equality with the
 *                     compiled reference over trajectories
 *   integrator        free particle, time reversibility, reference trajectory
 */
#ifndef COMD_ORACLE_H
#define COMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXATOMS 64   /* slots per link cell, src-flat/include/pmdTypes.h:57 */

/* One tabulated function (src-flat/include/eamTypes.h:68-74).  values has
 * n+3 entries; vals points at entry 1 so that vals[-1], vals[n], vals[n+1]
 * are the guard slots of src-flat/eam.c:232-246. */
typedef struct orc_table {
    int n;
    double x0, xn, invDx;
    double *store;   /* n+3 doubles */
    double *vals;    /* store + 1 */
} orc_table;

/* Chebyshev series (src-flat/cheby.c:125-179): n coefficients on [x0,xn]. */
typedef struct orc_cheb {
    int n;
    double x0, xn;
    double c[64];
} orc_cheb;

enum { ORC_POT_EAM = 1, ORC_POT_LJ = 2 };
enum { ORC_EVAL_TABLE = 0, ORC_EVAL_CHEB = 1 };

typedef struct orc_pot {
    int kind;
    double cutoff, mass, lat;
    /* EAM tables (src-flat/eam.c:347-398) */
    orc_table phi, rho, f;
    /* Chebyshev fits and derivative fits (src-flat/cheby.c:99-121) */
    orc_cheb cphi, crho, cf, cdphi, cdrho, cdf;
    int eval;  /* ORC_EVAL_TABLE or ORC_EVAL_CHEB */
    /* LJ parameters (src-flat/ljForce.c:75-86) */
    double sigma, epsilon;
    /* sign of the stored LJ force: +1 keeps src-flat's -grad U
     * (ljForce.c:204-208); -1 stores +grad U, the EAM convention, under
     * which the integrator's subtraction (timestep.c:71) gives physical
     * motion.  A test-only variant, set by orc_pot_set_lj_sign. */
    double lj_sign;
} orc_pot;

/* The simulation in the reference's slot layout
 * (src-flat/include/pmdTypes.h:104-141): slot i of cell c is c*64+i. */
typedef struct orc_sim {
    int nbx[3];
    int nboxes;
    int ntot;
    double bounds[3];
    double boxsize[3];
    double defgrad, bf;
    int periodic;
    double *dcenter;   /* nboxes*3 */
    int *natoms;       /* nboxes */
    int *id;           /* slots */
    int *itype;        /* slots */
    double *mass;      /* slots */
    double *r;         /* slots*3 */
    double *p;         /* slots*3 */
    double *f;         /* slots*4 : fx fy fz e */
    double *rho;       /* slots */
    double *fi;        /* slots */
    double e;
    double stress;
    int err;           /* sticky error: 0 ok, else ORC_ERR_* */
    const orc_pot *pot;
} orc_sim;

enum {
    ORC_OK = 0,
    ORC_ERR_NOFILE = 1,
    ORC_ERR_TABLE = 2,
    ORC_ERR_GRID = 3,      /* nbx < 1 */
    ORC_ERR_OVERFLOW = 4,  /* cell reached 64 atoms in a move */
    ORC_ERR_ALLOC = 5,
    ORC_ERR_OUTSIDE = 6,   /* atom outside the cell grid after wrapping */
};

/* ---- potentials ---- */
int  orc_pot_read_eam(orc_pot *pot, const char *dir, const char *name);
void orc_pot_set_lj(orc_pot *pot, double sigma, double epsilon, double cutoff, double mass);
void orc_pot_set_lj_sign(orc_pot *pot, double sign);
void orc_pot_free(orc_pot *pot);
void orc_pot_set_eval(orc_pot *pot, int eval);
double orc_pot_get(const orc_pot *pot, int which); /* 0 cutoff, 1 mass, 2 lat */
void orc_pot_fit_cheb(orc_pot *pot, int n);
void orc_table_interp(const orc_table *a, double r, double *val, double *der);
double orc_cheb_eval(const orc_cheb *c, double x);
void orc_cheb_fit(const orc_table *t, int n, orc_cheb *out);
void orc_cheb_deriv(const orc_cheb *c, orc_cheb *out);

/* ---- domains ---- */
orc_sim *orc_sim_create(const orc_pot *pot, const double bounds[3], double defgrad, double bf, int periodic);
void orc_sim_free(orc_sim *s);
int  orc_put_atom(orc_sim *s, int id, int itype, double mass,
                  double x, double y, double z, double px, double py, double pz);
void orc_rebox_all(orc_sim *s);
int  orc_neighbor_boxes(const orc_sim *s, int ibox, int *nbrs /* 27 */);

/* ---- forces ---- */
int  orc_eam_force(orc_sim *s);
int  orc_lj_force(orc_sim *s);
void orc_virial_tensor(const orc_sim *s, double *out);
/* Centrosymmetry parameter per slot (src-viz/AtomVisualize.c:399-673). */
void orc_centrosymmetry(const orc_sim *s, double *out);
int  orc_compute_force(orc_sim *s);

/* ---- timestep ---- */
double orc_n_time_steps(orc_sim *s, int n, double dt_seconds);
double orc_kinetic_energy(const orc_sim *s);
void orc_advance_positions(orc_sim *s, double dt_internal);  /* src-flat/timestep.c:76-108 */
void orc_advance_velocity(orc_sim *s, double dt_internal);   /* src-flat/timestep.c:62-74 */
void orc_sim_set_state(orc_sim *s, const int *natoms, const int *id, const int *itype, const double *r, const double *p);

/* ---- inputs ---- */
orc_sim *orc_read_gz(const orc_pot *pot, const char *path, double bf, int periodic);
orc_sim *orc_create_fcc(const orc_pot *pot, int nx, int ny, int nz, double lat,
                        double defgrad, double bf, int periodic);
orc_sim *orc_from_atoms(const orc_pot *pot, const double bounds[3], double bf, int periodic,
                        int n, const double *xyz, const double *pxyz, int itype);

/* ---- accessors for ctypes ---- */
orc_pot *orc_pot_new(void);
void orc_pot_delete(orc_pot *p);
void orc_sim_info(const orc_sim *s, int *nbx3, int *nboxes, int *ntot,
                  double *bounds3, double *boxsize3, double *e, double *stress, int *err);
void orc_sim_get(const orc_sim *s, int *natoms, int *id, int *itype, double *r, double *p,
                 double *f, double *rho, double *fi);
void orc_sim_set_rp(orc_sim *s, const double *r, const double *p);
void orc_pot_get_cheb(const orc_pot *p, double *out /* 6*n */, int *n);
void orc_pot_get_table(const orc_pot *p, int which, int *n, double *x0xninv /* 3 */, double *vals /* n+3 */);
void orc_pot_eval(const orc_pot *p, int which, double x, double *val, double *der);
void orc_pot_set_cheb3(orc_pot *p, const double *coef /* 3*n */, int n);

#ifdef __cplusplus
}
#endif
#endif
