// comd_internal.h -- declarations shared by the translation units of the
// gfx950 library (not part of the C-ABI).
#pragma once

#include "comd_common.h"

// Potential description as the kernels see it.
struct PotDev {
    int kind;     // COMD_POT_EAM / COMD_POT_LJ
    int eval;     // COMD_EVAL_TABLE / COMD_EVAL_CHEB
    double rc2;   // cutoff squared
    // tables phi, rho, F: values[-1..n+1] at tab[k] (tab[k][0] is values[-1])
    const double *tab[3];
    double x0[3], xn[3], invdx[3];
    int n[3];
    // per-interval coefficients of phi and rho (value c0 + t(c1 + t c2),
    // derivative d0 + t d1): (d0 d1 c0 c1 c2 0) in double, the same
    // padded to 8 floats for the mixed path
    const double *cf64[2];
    const float *cf32[2];
    int same_grid;  // phi and rho tables share x0, xn, invDx and n
    // compact rows for the neighbour-list passes (only when same_grid and
    // cutoff <= xn), per interval and one more (the last interval re-centred
    // on its end, read at r = xn): density pass, FP64 (phi d0 d1 c0 c1 c2,
    // rho c0 c1 c2) and mixed (phi d0 d1 in double, then phi c0 c1 c2, rho
    // c0 c1 c2 in float: 40 bytes); embedding-force pass (rho d0 d1) in double
    const double *cd64, *cr64;
    const float *cdm;
    int balanced;   // drain an atom's queues round-robin over its lanes
    // Chebyshev: 6 series of cn coefficients (phi rho F dphi drho dF), each
    // at a stride of cstride (cn rounded up to a multiple of 8, zero-padded)
    const double *cheb;
    const float *cheb32;
    int cn, cstride;
    double ca[3], cb[3];   // domains of phi, rho, F
    // Lennard-Jones
    double lj_s6, lj_eps;
    // sign of the stored LJ force: +1 stores -grad U as src-flat/ljForce.c
    // does (its integrator then climbs the potential, timestep.c:71), -1
    // stores +grad U, the convention of the EAM force, for physical motion
    double lj_fsign;
    double mass_amu;
};

enum { CM_KIND_EAM1 = 0, CM_KIND_EAM3 = 1, CM_KIND_LJ = 2 };

// Address-space qualifiers: loads through CM_GLOBAL / CM_LDS pointers are
// global / LDS instructions rather than flat ones.
#define CM_GLOBAL __attribute__((address_space(1)))
#define CM_LDS __attribute__((address_space(3)))

// Read-only data read through the scalar unit when the address is the same
// in every lane (series coefficients).
#define CM_CONST __attribute__((address_space(4)))

// Clenshaw sum with the lower clamp only, src-flat/cheby.c:276-310, over a
// series zero-padded to m8 coefficients (a multiple of 8; leading zeros
// leave the sum exactly as it is).  The coefficients, in global memory at an
// address common to the wave, are read by scalar loads 8 at a time, each
// block's loads issued before the previous block's iterations so that their
// latency is hidden behind them.
__device__ __forceinline__ double cheb64(const double *cg, int m8, double a, double b, double x)
{
    const CM_CONST double *c = (const CM_CONST double *)cg;
    if (x < a) x = a;
    double d = 0.0, dd = 0.0;
    const double y = (2.0 * x - a - b) / (b - a);
    const double y2 = 2.0 * y;
    double cc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) cc[u] = c[m8 - 8 + u];
    for (int jb = m8 - 8; jb >= 0; jb -= 8) {
        double cn[8];
        const int jn = jb >= 8 ? jb - 8 : 0;  // the next block (block 0 again after the last)
#pragma unroll
        for (int u = 0; u < 8; u++) cn[u] = c[jn + u];
#pragma unroll
        for (int u = 7; u >= 0; u--) {
            if (jb + u > 0) {
                const double sv = d;
                d = y2 * d - dd + cc[u];
                dd = sv;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) cc[u] = cn[u];
    }
    return y * d - dd + 0.5 * cc[0];
}

__device__ __forceinline__ float cheb32(const float *cg, int m8, float a, float b, float x)
{
    const CM_CONST float *c = (const CM_CONST float *)cg;
    x = fmaxf(x, a);
    float d = 0.0f, dd = 0.0f;
    const float y = (2.0f * x - a - b) / (b - a);
    const float y2 = 2.0f * y;
    float cc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) cc[u] = c[m8 - 8 + u];
    for (int jb = m8 - 8; jb >= 0; jb -= 8) {
        float cn[8];
        const int jn = jb >= 8 ? jb - 8 : 0;
#pragma unroll
        for (int u = 0; u < 8; u++) cn[u] = c[jn + u];
#pragma unroll
        for (int u = 7; u >= 0; u--) {
            if (jb + u > 0) {
                const float sv = d;
                d = fmaf(y2, d, cc[u] - dd);
                dd = sv;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) cc[u] = cn[u];
    }
    return fmaf(y, d, 0.5f * cc[0] - dd);
}


// Launch shape of the force kernels (chosen by the tuner).
struct ForceCfg {
    int blocks;   // persistent grid size
    int waves;    // waves per block (cells processed concurrently per block)
    int tab_lds;  // stage the potential tables in LDS when they fit
    int max_lds;  // LDS bytes allowed per block
    int ncu;      // compute units of the device
};
int cm_force_lds_bytes(const PotDev &P, int kind, int mixed, int waves, int tab_lds);

void cm_launch_move(const DevState &d, hipStream_t st, double dtk, double dtd, int kick, int drift, int rebin);
void cm_launch_rebox(const DevState &d, hipStream_t st);
void cm_launch_bin(const DevState &d, const Grid &g, hipStream_t st, int n, const double *xyz, const double *pxyz,
                   int itype, int *cnt, int *tmp);
void cm_launch_fill_int(int *p, int v, long n, hipStream_t st);
void cm_launch_force(const DevState &d, const PotDev &pot, int kind, int mixed, const ForceCfg &cfg, hipStream_t st);
// All replicas of a batch in one launch (pass 1: EAM density or LJ; pass 3:
// EAM embedding force); pots: device array indexed by Grid::pot.
void cm_launch_force_batch(const DevState &d, const PotDev *pots, int pass, int mixed, const ForceCfg &cfg,
                           hipStream_t st);
// Energies of one replica's cells: per-block partial (PE, KE, virial).
void cm_launch_reduce(const DevState &d, const Grid &g, double *partials, int nparts, hipStream_t st);
void cm_launch_reduce_final(const double *partials, int nparts, double *out, hipStream_t st);
// Every replica's energies at once (grids: the device descriptors), the same
// sums as cm_launch_reduce + cm_launch_reduce_final per replica: partials
// (nrep * nparts * 3 doubles), then out + 3 r for replica r.
void cm_launch_reduce_batch(const DevState &d, const Grid *grids, int nrep, double *partials, int nparts, double *out,
                            hipStream_t st);
// Staged value and derivative series (FP64, then FP32; 72 m8 bytes per
// replica) copied to every replica's series buffers (pots: the device
// descriptors).
void cm_launch_scatter_series(const PotDev *pots, int nrep, const char *src, int m8, hipStream_t st);
// The adjoint's per-replica dL/dE and residual weights from the energies E
// (3 doubles per replica) on the device.
void cm_launch_adj_rep(const double *E, const double *w, const double *rw, int nrep, int emode, double w_e,
                       double *rep, hipStream_t st);

// Device buffers and scalars of the Chebyshev-coefficient adjoint
// (adjoint.hip).  out holds 5 x cm_adjoint_moments() doubles: the moments of
// stages 0-3 and, at out[4*64], the weighted sum of squared force residuals.
struct CmAdjoint {
    double *gx, *gy, *gz;  // per slot: dL/df
    double *lam, *mu;      // per slot
    const double *f2;      // F'' series
    const double *dlde;    // per replica: dL/dE of its energy term
    const double *rw;      // per replica: weight of its force residuals, 1 / (3 N)
    double *part;          // nblocks x 64 partial moments
    int nblocks;
    double *out;
};
int cm_adjoint_max_coefs();
int cm_adjoint_moments();
void cm_launch_adj_resid(const DevState &d, const CmAdjoint &J, const double *ft, hipStream_t st);
void cm_launch_adj_stage(const DevState &d, const PotDev &P, const CmAdjoint &J, int stage, int mixed,
                         hipStream_t st);
void cm_launch_adj_final(const CmAdjoint &J, int nblocks, int stage, hipStream_t st);

// Observables (observe.hip): per-block partial sums of CM_OBS_N doubles
// over one replica's cells (at most CM_OBS_MAXBLOCKS blocks, in part), then
// their sum in block order and the comd_observables formed from it, written
// to out (device memory).  pe_scale multiplies the per-atom energy sum (4
// eps sigma^6 for Lennard-Jones, 1 for EAM).
#define CM_OBS_N 16
#define CM_OBS_MAXBLOCKS 2048
#define CM_KB_HARTREE 3.166811563e-6  // Boltzmann constant, hartree per kelvin
int cm_observe_blocks(const Grid &g);
void cm_launch_observe(const DevState &d, const PotDev &P, const Grid &g, double pe_scale, double *part,
                       comd_observables *out, hipStream_t st);

// Centrosymmetry parameter (centro.hip) of one replica's atoms into out
// (device memory, one double per slot of the replica's cells).
void cm_launch_centro(const DevState &d, const Grid &g, double *out, hipStream_t st);

// Group gi of ng along a dimension of na auxiliary cells: its first cell x0
// and extent lx (extents differ by at most one; na / ng each when ng divides
// na).
__host__ __device__ inline void group_span(int gi, int na, int ng, int &x0, int &lx)
{
    x0 = gi * na / ng;
    lx = (gi + 1) * na / ng - x0;
}

// One replica's auxiliary grid (cells of edge >= cutoff + skin) and its
// groups of auxiliary cells, for the neighbour lists.
struct NlRep {
    double L0, L1, L2, iL0, iL1, iL2;        // bounds and their inverses
    double ab0, ab1, ab2, iab0, iab1, iab2;  // auxiliary cell edges and their inverses
    double qscale, qinv;                     // group-frame fixed point of the mixed passes: units per A, A per unit
    double rl2;                              // (cutoff + skin)^2
    double rc2;                              // cutoff^2
    int na0, na1, na2;                       // auxiliary cells per dimension
    int gd0, gd1, gd2;                       // cells per group (the largest extent)
    int ng0, ng1, ng2;                       // groups per dimension
    int aux_base, group_base;                // its first auxiliary cell and group in the global numbering
    int id_base, nid;                        // its atom ids i map to list ids id_base + i, 0 <= i < nid
};

// Verlet neighbour lists of periodic simulations (a single one, or every
// replica of a batch) (nlist.hip): per atom all atoms within cutoff + skin
// at the last build, rebuilt on the device whenever an atom has moved more
// than half the skin.  Atoms, auxiliary cells and groups are numbered over
// all replicas; a group's replica is group_rep[g].
struct CmList {
    int n;                   // list id capacity (ids 0 .. n-1, over all replicas)
    int kmax;                // list capacity per atom (a multiple of 16)
    const NlRep *rep;        // [nrep]
    int nrep;
    int naux;                // auxiliary cells, all replicas
    const int *group_rep;    // [ngroups]: replica of each group
    const int *aux_group;    // [naux]: group of each auxiliary cell
    int ngroups;             // groups, all replicas; the build order is group-major
    int hcap;                // staged halo capacity per group (atoms, <= 65535)
    int lds_avail;           // dynamic LDS a list kernel's block may use
    int rows_lds;            // table rows in LDS: 0 never, else whenever they fit
    int pass_blocks, pass_threads;  // one block per group (passes and build)
    int tailq;               // build: list entries per thread held back for the list's end (k_nl_build)
    int lanes;               // force passes: lanes per home atom (1, 2 or 4), each taking every
                             // lanes-th chunk of 8 entries of the atom's list
    const int *cell_order;   // [naux]: auxiliary cells in build order (group, then cell)
    const int *group_q0;     // [ngroups + 1]: a group's first position in cell_order
    CmPos *pos;              // per build position: absolute position (kept by the drift)
    double *fq;              // per position in build order: F' of the last density pass
    double *f1;              // per position in build order: the density pass's pair forces (x, y, z),
                             // completed by the embedding-force pass into the slots
    uint16_t *selfh;         // per t: the atom's own index in its group's staged halo
    int *slot;               // per id: current slot, -1 when absent
    int *order;              // atoms (ids) in build order
    uint16_t *nbl;           // list of the t-th atom in build order: indices into its group's
                             // staged halo, entry k at nl_entry(k, t, n) (groups of 8 entries
                             // per atom, consecutive over t: one 16-byte load per lane fetches
                             // 8 entries, and a wave's loads are contiguous)
    int *ncnt;               // per t
    int *aux_cnt;            // per auxiliary cell: atoms keyed (cleared after each build)
    int *aux_n;              // the same, capped at CM_NL_AUXCAP, kept until the next build
    int *aux_start;          // first atom of each auxiliary cell in build order
    int *aux_mem;            // [cell][CM_NL_AUXCAP] member ids
    int *gstart;             // [ngroups + 1]: first atom of each group in build order
    int *nonempty;           // per link cell: non-empty stencil cells (Lennard-Jones virial)
    int *count;              // [0]: atoms in build order, [1]: builds so far, [2]: build blocks done,
                             // [3]: group-sum blocks done
    unsigned *flags;         // CM_NL_REBUILD, CM_NL_INVALID, CM_NL_OVERFLOW (during a build)
};
#define CM_NL_AUXCAP 64
#define CM_NL_SLOTS 180                 // halo cells of a group, at most (4 + 2) x (4 + 2) x (3 + 2)
#define CM_NL_STATIC_LDS (6 << 10)      // static LDS of the list kernels, rounded up
__host__ __device__ inline size_t nl_entry(int k, int t, int n) { return ((size_t)(k >> 3) * n + t) * 8 + (k & 7); }
void cm_nl_init(const DevState &d, const CmList &l, hipStream_t st);
void cm_nl_build(const DevState &d, const CmList &l, hipStream_t st);
void cm_nl_sync(const DevState &d, const CmList &l, hipStream_t st);
size_t cm_nl_build_lds(const CmList &l);
// Entries per thread of the build's tail queue: the most (16, 12, 8 or 4)
// that fit without fewer resident build blocks per CU than with none.
int cm_nl_pick_tailq(CmList l);
size_t cm_nl_pass_lds(const PotDev &P, const CmList &l, int kind, int mixed, int *tables_in_lds);
size_t cm_nl_tab_bytes(const PotDev &P, int kind, int mixed);
// The list pass of one kind over groups [g0, g0 + ng) (a single
// simulation: all of them; a batch: the groups of its EAM or of its LJ
// replicas, which share one potential).
void cm_nl_force(const DevState &d, const PotDev &P, const CmList &l, int kind, int mixed, hipStream_t st, int g0,
                 int ng);
// Resident blocks per CU of a list pass (-1 if the runtime cannot tell) and
// its dynamic LDS bytes.
int cm_nl_occupancy(const PotDev &P, const CmList &l, int kind, int mixed, int *lds_bytes);

// Adjoint pair passes over the lists (nlist.hip): pass 1 (phi moments and
// the energy moments), 2 (rho' moments and lambda), 3 (rho moments), one
// block per group writing ANMOM partial moments to part[block]; gb holds
// dL/df in build order (cm_launch_adj_nl_gather fills it).
struct CmAdjNl {
    const double *gb;            // build order: dL/df, 3 per atom
    const double *gx, *gy, *gz;  // per slot: dL/df
    const double *mu;            // per slot: mu (pass 3)
    double *lam;                 // per slot: lambda (pass 2)
    const double *dlde;          // per replica: dL/dE
    double *part;                // per block: partial moments
};
size_t cm_adj_nl_lds(const CmList &l);
void cm_launch_adj_nl_gather(const CmList &l, const double *gx, const double *gy, const double *gz, double *gb,
                             hipStream_t st);
void cm_launch_adj_nl(const DevState &d, const PotDev &P, const CmList &l, const CmAdjNl &A, int pass, int mixed,
                      hipStream_t st);
