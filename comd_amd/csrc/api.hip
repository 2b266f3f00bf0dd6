// api.hip -- context management and the C-ABI entry points of
// include/comd_hip.h.
//
// A context holds one or more replicas (independent simulations).  Their
// link cells are numbered one after the other and share one set of slot
// arrays, so every kernel of a step covers all replicas in one launch; a
// single simulation is a context with one replica.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <string>
#include <vector>

#include "comd_internal.h"

// src-flat/include/constants.h:61-65
static const double kBohrPerAtuToAPerS = (double)0.529 / (double)2.418e-17;
static const double kAmuToMe = (double)1822.83;

static char g_create_err[512] = "";

// Scratch of comd_observe_noalloc: the block partials of the observables
// reduction (CM_OBS_MAXBLOCKS x CM_OBS_N doubles, 256 KiB), one copy per
// device, held until the process ends.  Its name is a host-side stand-in:
// host code uses only the device address that hipGetSymbolAddress gives
// for the current device (scratch_of).
static __device__ __align__(256) double scratch[CM_OBS_MAXBLOCKS * CM_OBS_N];

#define CM_MAX_DEVICES 64
static double *g_scratch_ptr[CM_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The address of scratch on device dev, looked up once per device with that
// device current (at context creation); null when the lookup failed.
static double *scratch_lookup(int dev)
{
    if (dev < 0 || dev >= CM_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}

static double *scratch_of(int dev)
{
    if (dev < 0 || dev >= CM_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[dev];
}

struct TimedEvent {
    int cls;
    hipEvent_t a, b;
};

// One replica: its parameters, grid, potential and the device memory its
// potential owns.
struct Replica {
    comd_params prm;  // table/series pointers cleared after creation
    Grid g;
    PotDev pot;
    int ntot;
    long nl_ids;                // id capacity of its atoms (max id + 1)
    uint64_t pot_hash;          // of the potential's parameters, tables and series
    std::vector<void *> owned;  // device buffers of the potential
    double *d_cheb;             // 6 series of cn at a stride of pot.cstride (FP64), or null
    float *d_cheb32;
};

struct comd_ctx {
    int nrep;
    std::vector<Replica> reps;
    int precision;         // COMD_PREC_*, shared by all replicas
    int ncells;            // cells of all replicas
    long nslots;
    int ntot;
    int have_atoms;
    int device;
    hipStream_t stream;
    // moving state (slot layout)
    int *natoms, *id, *itype;
    double *rx, *ry, *rz, *px, *py, *pz;
    // outputs
    double *fx, *fy, *fz, *fe, *rho, *fi, *vir;
    int *nnb;
    // rebin scratch
    unsigned long long *leave;
    int *mvidx, *meta, *counts, *touched, *tc, *arr_cnt, *arr;
    int *mv_slot, *mv_rank, *mv_id, *mv_itype;
    float4 *mv_nlr, *nlr;
    double *mv_rx, *mv_ry, *mv_rz, *mv_px, *mv_py, *mv_pz;
    uint8_t *perm;
    unsigned *flags;
    int *mh;  // per replica: multi-hop flag of the current rebin pass
    // pair lists of the EAM density pass (EAM contexts)
    uint16_t *plist;
    uint8_t *pcount;
    int *pvalid;
    int use_plist;
    double *partials;
    int nparts;
    // observables and the energy monitor (allocated on first use)
    double *obs_part, *mon;
    comd_observables *obs_out;
    long mon_cap;
    double *red_out;       // 3 doubles, device
    double *host_out;      // pinned 4 doubles
    unsigned *host_flags;  // pinned
    // replica descriptors on the device
    Grid *d_grids;
    int *d_cell_grid;
    PotDev *d_pots;
    // Chebyshev-coefficient adjoint (allocated on first use)
    double *adj_g[3], *adj_lam, *adj_mu, *adj_ft, *adj_f2, *adj_part, *adj_out;
    double *adj_rep;      // per replica: dL/dE, then the residual weight 1 / (3 N)
    int adj_nblocks;
    double *adj_gb, *adj_part_nl;  // list passes: dL/df in build order, partial moments
    long adj_gb_cap, adj_part_nl_cap;
    int adj_lists;        // COMD_TUNE_ADJOINT_LISTS: 0 link cells, 1 lists with a group per CU, 2 lists
    int adj_have_target;  // adj_ft holds a target (comd_eam_cheb_grad with f_target NULL reuses it)
    double *adj_E;        // per replica: the energy reduction's output (3 doubles each)
    double *adj_epart;    // per replica: its energy reduction's partials (nparts x 3 doubles)
    // staging for small per-call uploads (series, adjoint scalars): pinned
    // host memory and its device copy, stage_cap bytes each
    char *stage_h, *stage_d;
    size_t stage_cap;
    std::vector<double> cheb_host;  // current value series phi, rho, F (3 x cn) of replica 0
    // force-kernel launch shape (comd_tune / comd_autotune)
    int ncu, max_lds;
    int lds_block;         // LDS bytes a block may use on this device
    int force_waves, force_blocks_per_cu, tab_lds;
    // Verlet neighbour lists (single periodic simulations; comd_tune)
    CmList nl;
    int nl_skin_milli;  // skin in 1e-3 A, 0 = link-cell kernels only
    int nl_active;      // lists allocated and usable for the atoms held
    int nl_stale;       // list positions not updated by the last drift
    int nl_invalid_seen; // the lists were invalid when the host last looked
    int nl_group_opt;   // COMD_TUNE_LIST_GROUPS (-1: by precision)
    int nl_rows_lds;    // COMD_TUNE_LIST_ROWS_LDS (-1: whenever they fit)
    int nl_threads_div; // COMD_TUNE_LIST_THREADS_DIV: home atoms per thread of the list passes
    int nl_lanes;       // COMD_TUNE_LIST_LANES: lanes per home atom of the list passes (0: by size)
    long nl_alloc_n, nl_alloc_k, nl_alloc_aux, nl_alloc_groups, nl_alloc_rep;
    int nl_eam_groups, nl_lj_groups;  // the list groups of the EAM replicas (first), then of the LJ ones
    int nl_eam_rep, nl_lj_rep;        // a replica of each kind (their potentials stand for the kind's)
    // one DKD step captured as a HIP graph (comd_steps)
    int use_graph;
    hipGraphExec_t graph;
    double graph_dt;
    // timing
    int timing;
    std::vector<TimedEvent> events;
    std::vector<hipEvent_t> pool;  // reusable events
    size_t pool_used;
    double ms[COMD_K_NCLASS];
    int launches[COMD_K_NCLASS];
    std::string err;
};

static int set_err(comd_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else snprintf(g_create_err, sizeof g_create_err, "%s", buf);
    return code;
}

#define HIPCHK(ctx, call)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) return set_err((ctx), COMD_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static DevState dev_state(const comd_ctx *c)
{
    DevState d;
    d.ncells = c->ncells;
    d.ngrids = c->nrep;
    d.g0 = c->reps[0].g;
    d.grids = c->d_grids;
    d.cell_grid = c->d_cell_grid;
    d.natoms = c->natoms;
    d.id = c->id;
    d.itype = c->itype;
    d.rx = c->rx; d.ry = c->ry; d.rz = c->rz;
    d.px = c->px; d.py = c->py; d.pz = c->pz;
    d.fx = c->fx; d.fy = c->fy; d.fz = c->fz; d.fe = c->fe;
    d.rho = c->rho; d.fi = c->fi; d.vir = c->vir;
    d.nnb = c->nnb;
    d.leave = c->leave;
    d.mvidx = c->mvidx;
    d.counts = c->counts;
    d.touched = c->touched;
    d.tc = c->tc;
    d.arr_cnt = c->arr_cnt;
    d.arr = c->arr;
    d.mv_slot = c->mv_slot; d.mv_rank = c->mv_rank; d.mv_id = c->mv_id; d.mv_itype = c->mv_itype;
    d.mv_nlr = c->mv_nlr;
    d.nlr = c->nlr;
    d.mv_rx = c->mv_rx; d.mv_ry = c->mv_ry; d.mv_rz = c->mv_rz;
    d.mv_px = c->mv_px; d.mv_py = c->mv_py; d.mv_pz = c->mv_pz;
    d.perm = c->perm;
    d.meta = c->meta;
    d.flags = c->flags;
    d.mh = c->mh;
    const bool lists = c->use_plist && c->plist;
    d.plist = lists ? c->plist : nullptr;
    d.pcount = lists ? c->pcount : nullptr;
    d.pvalid = lists ? c->pvalid : nullptr;
    const bool nl = c->nl_active != 0;
    d.nl_pos = nl ? c->nl.pos : nullptr;
    d.nl_slot = nl ? c->nl.slot : nullptr;
    d.nl_flags = nl ? c->nl.flags : nullptr;
    // the displacement test reads the build's positions in FP32: half the
    // skin less their rounding (2^-23 of a coordinate, which stays within
    // about one box length; three coordinates)
    double lmax = 0.0;
    for (const Replica &r : c->reps) lmax = fmax(lmax, fmax(r.g.L0, fmax(r.g.L1, r.g.L2)));
    const double hs = 0.5e-3 * c->nl_skin_milli - 2.0 * 0x1p-23 * lmax;
    d.nl_half2 = hs > 0.0 ? hs * hs : 0.0;
    d.cell_gate = nl ? c->nl.flags : nullptr;
    return d;
}

// ---- timing helpers: HIP events recorded on the context's stream around
// each kernel group, taken from a reusable pool
static hipEvent_t pool_event(comd_ctx *c)
{
    if (c->pool_used == c->pool.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        c->pool.push_back(e);
    }
    return c->pool[c->pool_used++];
}
static void tbegin(comd_ctx *c, int cls, hipEvent_t *a)
{
    (void)cls;
    if (!c->timing) return;
    *a = pool_event(c);
    hipEventRecord(*a, c->stream);
}
static void tend(comd_ctx *c, int cls, hipEvent_t a)
{
    if (!c->timing) return;
    hipEvent_t b = pool_event(c);
    hipEventRecord(b, c->stream);
    c->events.push_back({cls, a, b});
}

// ---- library operations on the device (stream-ordered, no host sync)

static void invalidate_graph(comd_ctx *c)
{
    if (c->graph) hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
}

static void op_rebox(comd_ctx *c)
{
    hipEvent_t ev = nullptr;
    tbegin(c, COMD_K_REBIN, &ev);
    DevState d = dev_state(c);
    cm_launch_rebox(d, c->stream);
    tend(c, COMD_K_REBIN, ev);
}

// rebin: record movers for the op_rebox that follows (0 for a lone kick)
// nl_update = 0: the drift leaves the list positions stale (the step's
// second half drift: the next step's first one renews them, and a force
// call in between renews them first)
static void op_move(comd_ctx *c, double dtk, double dtd, int kick, int drift, int rebin = 1, int nl_update = 1)
{
    hipEvent_t ev = nullptr;
    tbegin(c, COMD_K_REBIN, &ev);
    DevState d = dev_state(c);
    if (!nl_update) d.nl_pos = nullptr;
    cm_launch_move(d, c->stream, dtk, dtd, kick, drift, rebin);
    tend(c, COMD_K_REBIN, ev);
    if (drift && c->nl_active) c->nl_stale = nl_update ? 0 : 1;
}

static ForceCfg force_cfg(const comd_ctx *c, int kind, int mixed)
{
    ForceCfg f;
    f.waves = c->nrep > 1 ? 1 : c->force_waves;
    f.tab_lds = c->tab_lds;
    f.max_lds = c->max_lds;
    f.ncu = c->ncu;
    // fewer waves per block when the per-wave LDS does not fit
    while (f.waves > 1 && cm_force_lds_bytes(c->reps[0].pot, kind, mixed, f.waves, 0) > f.max_lds) f.waves--;
    long b = (long)c->ncu * c->force_blocks_per_cu;
    const long need = ((long)c->ncells + f.waves - 1) / f.waves;
    f.blocks = (int)(b < need ? b : need);
    if (f.blocks < 1) f.blocks = 1;
    return f;
}

static bool any_eam(const comd_ctx *c)
{
    for (const Replica &r : c->reps)
        if (r.prm.potential == COMD_POT_EAM) return true;
    return false;
}

static void op_force(comd_ctx *c)
{
    DevState d = dev_state(c);
    const int mixed = c->precision == COMD_PREC_MIXED;
    hipEvent_t ev = nullptr;
    // with neighbour lists: the (usually skipped) list build, then each pass
    // over the lists; the link-cell kernels run only if the lists are invalid
    const bool nl = c->nl_active != 0;
    if (nl) {
        tbegin(c, COMD_K_LIST_BUILD, &ev);
        if (c->nl_stale) {
            cm_nl_sync(d, c->nl, c->stream);
            c->nl_stale = 0;
        }
        cm_nl_build(d, c->nl, c->stream);
        tend(c, COMD_K_LIST_BUILD, ev);
    }
    if (c->nrep > 1) {
        // lists: the EAM replicas' groups (density, then embedding force)
        // and the LJ replicas' groups, each kind with its shared potential;
        // then the link-cell passes over all replicas, which do the work
        // when there are no lists and return at once beside valid ones
        if (nl) {
            tbegin(c, COMD_K_DENSITY, &ev);
            if (c->nl_eam_groups > 0)
                cm_nl_force(d, c->reps[c->nl_eam_rep].pot, c->nl, CM_KIND_EAM1, mixed, c->stream, 0,
                            c->nl_eam_groups);
            if (c->nl_lj_groups > 0)
                cm_nl_force(d, c->reps[c->nl_lj_rep].pot, c->nl, CM_KIND_LJ, mixed, c->stream, c->nl_eam_groups,
                            c->nl_lj_groups);
            tend(c, COMD_K_DENSITY, ev);
            if (c->nl_eam_groups > 0) {
                tbegin(c, COMD_K_EMBED_FORCE, &ev);
                cm_nl_force(d, c->reps[c->nl_eam_rep].pot, c->nl, CM_KIND_EAM3, mixed, c->stream, 0,
                            c->nl_eam_groups);
                tend(c, COMD_K_EMBED_FORCE, ev);
            }
        }
        auto cells = [&](int pass, int kind, int cls) {
            tbegin(c, nl ? COMD_K_CELL_FALLBACK : cls, &ev);
            ForceCfg cfg = force_cfg(c, kind, mixed);
            if (nl && !c->nl_invalid_seen && cfg.blocks > c->ncu) cfg.blocks = c->ncu;
            cm_launch_force_batch(d, c->d_pots, pass, mixed, cfg, c->stream);
            tend(c, nl ? COMD_K_CELL_FALLBACK : cls, ev);
        };
        cells(1, CM_KIND_EAM1, COMD_K_DENSITY);
        if (any_eam(c)) cells(3, CM_KIND_EAM3, COMD_K_EMBED_FORCE);
        return;
    }
    const PotDev &P = c->reps[0].pot;
    // one pass: the list kernel, then the link-cell kernel (which does the
    // work when there are no lists, and returns at once while they are valid)
    auto pass = [&](int kind, int cls) {
        if (nl) {
            tbegin(c, cls, &ev);
            cm_nl_force(d, P, c->nl, kind, mixed, c->stream, 0, c->nl.ngroups);
            tend(c, cls, ev);
        }
        const int ccls = nl ? COMD_K_CELL_FALLBACK : cls;
        tbegin(c, ccls, &ev);
        ForceCfg cfg = force_cfg(c, kind, mixed);
        // beside valid lists the link-cell pass returns at once: a small grid
        // keeps that cheap (the full grid once the host has seen the lists
        // invalid)
        if (nl && !c->nl_invalid_seen && cfg.blocks > c->ncu) cfg.blocks = c->ncu;
        cm_launch_force(d, P, kind, mixed, cfg, c->stream);
        tend(c, ccls, ev);
    };
    if (P.kind == COMD_POT_LJ) {
        pass(CM_KIND_LJ, COMD_K_DENSITY);
    } else {
        pass(CM_KIND_EAM1, COMD_K_DENSITY);
        pass(CM_KIND_EAM3, COMD_K_EMBED_FORCE);
    }
}

// One velocity-Verlet step of src-flat/timestep.c:114-135 without the
// trailing force: half drift, rebin, force, kick + half drift, rebin.
static void op_step(comd_ctx *c, double dt)
{
    op_move(c, 0.0, dt / 2.0, 0, 1);
    op_rebox(c);
    op_force(c);
    op_move(c, dt, dt / 2.0, 1, 1, 1, 0);
    op_rebox(c);
}

// n steps: replays of the captured step graph when graphs are on and the
// kernels are not being timed, else launched one by one.
static int run_steps(comd_ctx *c, int n, double dt)
{
    if (!c->use_graph || c->timing || n < 2) {
        for (int i = 0; i < n; i++) op_step(c, dt);
        HIPCHK(c, hipGetLastError());
        return COMD_OK;
    }
    if (c->graph && c->graph_dt != dt) invalidate_graph(c);
    if (!c->graph) {
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        op_step(c, dt);
        HIPCHK(c, hipStreamEndCapture(c->stream, &g));
        const hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->graph = nullptr;
            return set_err(c, COMD_ERR_HIP, "graph instantiation: %s", hipGetErrorString(e));
        }
        c->graph_dt = dt;
    }
    for (int i = 0; i < n; i++) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
    // each step ends with a drift that leaves the list positions stale (the
    // capture recorded that on the host once; replays must too)
    if (c->nl_active) c->nl_stale = 1;
    return COMD_OK;
}

static int check_flags(comd_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->host_flags, c->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    if (c->nl_active)
        HIPCHK(c, hipMemcpyAsync(c->host_flags + 1, c->nl.flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const unsigned f = *c->host_flags;
    // lists found invalid: the link-cell passes beside the list passes get
    // their full grid until the lists are valid again
    if (c->nl_active) {
        const int inv = (c->host_flags[1] & CM_NL_INVALID) != 0;
        if (inv != c->nl_invalid_seen) invalidate_graph(c);
        c->nl_invalid_seen = inv;
    }
    if (f & CM_FLAG_OVERFLOW) return set_err(c, COMD_ERR_OVERFLOW, "a link cell reached %d atoms", CM_SLOTS);
    if (f & CM_FLAG_OUTSIDE) return set_err(c, COMD_ERR_OUTSIDE, "an atom fell outside the cell grid");
    return COMD_OK;
}

static int clear_flags(comd_ctx *c)
{
    HIPCHK(c, hipMemsetAsync(c->flags, 0, 4, c->stream));
    return COMD_OK;
}

// ---- creation

extern "C" {

int comd_grid_for(comd_params *p, const double bounds[3], double defgrad, double bf)
{
    if (!p || !bounds || !(p->cutoff > 0.0)) return COMD_ERR_ARG;
    double strain[3] = {defgrad, 1.0, 1.0};
    if (strain[0] == 0.0) strain[0] = 1.0;
    const double boxfactor = bf == 0.0 ? 1.0 : bf;
    for (int j = 0; j < 3; j++) {
        p->bounds[j] = bounds[j];
        p->nbx[j] = (int)floor(bounds[j] / (p->cutoff * boxfactor * strain[j]));
        if (p->nbx[j] < 1) return COMD_ERR_GRID;
        p->boxsize[j] = bounds[j] / (double)p->nbx[j];
    }
    return COMD_OK;
}

}

static void free_nl(comd_ctx *c)
{
    void *p[] = {c->nl.pos,      c->nl.selfh,           c->nl.slot,    c->nl.order,    c->nl.fq,    c->nl.f1,
                 c->nl.nbl,      c->nl.ncnt,     c->nl.aux_cnt,   c->nl.aux_n,   c->nl.aux_start,
                 c->nl.aux_mem,  c->nl.gstart,   c->nl.nonempty,  c->nl.count,   c->nl.flags,
                 (void *)c->nl.cell_order, (void *)c->nl.group_q0, (void *)c->nl.group_rep,
                 (void *)c->nl.aux_group, (void *)c->nl.rep};
    for (void *q : p)
        if (q) hipFree(q);
    memset(&c->nl, 0, sizeof c->nl);
    c->nl_alloc_n = c->nl_alloc_k = c->nl_alloc_aux = c->nl_alloc_groups = c->nl_alloc_rep = 0;
    c->nl_active = 0;
}

static void free_all(comd_ctx *c)
{
    if (c->graph) hipGraphExecDestroy(c->graph);
    free_nl(c);
    void *ptrs[] = {c->natoms, c->id, c->itype, c->rx, c->ry, c->rz, c->px, c->py, c->pz,
                    c->fx, c->fy, c->fz, c->fe, c->rho, c->fi, c->vir, c->nnb, c->leave, c->mvidx, c->meta,
                    c->counts, c->touched, c->tc, c->arr_cnt, c->arr, c->mv_slot, c->mv_rank, c->mv_id, c->mv_nlr, c->nlr,
                    c->mv_itype, c->mv_rx, c->mv_ry, c->mv_rz, c->mv_px, c->mv_py, c->mv_pz,
                    c->perm, c->flags, c->mh, c->plist, c->pcount, c->pvalid,
                    c->partials, c->red_out, c->d_grids, c->d_cell_grid, c->d_pots,
                    c->adj_g[0], c->adj_g[1], c->adj_g[2], c->adj_lam, c->adj_mu,
                    c->adj_ft, c->adj_f2, c->adj_part, c->adj_out, c->adj_rep, c->obs_part, c->obs_out, c->mon,
                    c->adj_gb, c->adj_part_nl, c->adj_E, c->adj_epart, c->stage_d};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (c->stage_h) hipHostFree(c->stage_h);
    for (Replica &r : c->reps)
        for (void *p : r.owned)
            if (p) hipFree(p);
    if (c->host_out) hipHostFree(c->host_out);
    if (c->host_flags) hipHostFree(c->host_flags);
    for (auto e : c->pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
}

// Per-interval coefficients of the table formula (src-flat/eam.c:418-433):
// on interval i the value is c0 + t(c1 + t c2) and the derivative d0 + t d1
// (0 for i <= 0), with c0 = v[i], c1 = (v[i+1]-v[i-1])/2,
// c2 = (v[i+1]+v[i-1]-2v[i])/2, d0 = c1 invDx,
// d1 = ((v[i+2]-v[i]) - (v[i+1]-v[i-1])) invDx / 2.  Layout per interval:
// (d0, d1, c0, c1, c2, 0) in double, (d0, d1, c0, c1, c2, 0, 0, 0) in float.
static void table_coefs(int n, double invdx, const double *vals /* values[-1] first */, std::vector<double> &c64,
                        std::vector<float> &c32)
{
    c64.assign((size_t)6 * n, 0.0);
    c32.assign((size_t)8 * n, 0.0f);
    const double *v = vals + 1;
    for (int i = 0; i < n; i++) {
        const double vm = v[i - 1], v0 = v[i], v1 = v[i + 1], v2 = v[i + 2];
        const double g0 = v1 - vm, g1 = v2 - v0;
        const double A = v1 + vm - 2.0 * v0;
        const double k[5] = {i <= 0 ? 0.0 : 0.5 * g0 * invdx, i <= 0 ? 0.0 : 0.5 * (g1 - g0) * invdx, v0, 0.5 * g0,
                             0.5 * A};
        for (int q = 0; q < 5; q++) {
            c64[6 * i + q] = k[q];
            c32[8 * i + q] = (float)k[q];
        }
    }
}

static int check_params(const comd_params *p)
{
    if (!p) return set_err(nullptr, COMD_ERR_ARG, "null parameters");
    if (p->potential != COMD_POT_EAM && p->potential != COMD_POT_LJ)
        return set_err(nullptr, COMD_ERR_ARG, "unknown potential %d", p->potential);
    if (p->precision != COMD_PREC_FP64 && p->precision != COMD_PREC_MIXED)
        return set_err(nullptr, COMD_ERR_ARG, "unknown precision %d", p->precision);
    for (int j = 0; j < 3; j++)
        if (p->nbx[j] < 1 || !(p->boxsize[j] > 0) || !(p->bounds[j] > 0))
            return set_err(nullptr, COMD_ERR_GRID, "bad grid in dimension %d", j);
    if (!(p->cutoff > 0) || !(p->mass > 0)) return set_err(nullptr, COMD_ERR_ARG, "bad cutoff or mass");
    if (p->potential == COMD_POT_EAM) {
        for (int k = 0; k < 3; k++)
            if (p->tab_n[k] < 2 || !p->tab_values[k]) return set_err(nullptr, COMD_ERR_ARG, "EAM table %d missing", k);
        if (p->eval == COMD_EVAL_CHEB && (p->cheb_n < 2 || p->cheb_n > 64 || !p->cheb_coef))
            return set_err(nullptr, COMD_ERR_ARG, "Chebyshev coefficients missing");
    }
    return COMD_OK;
}

// Device copy of one replica's potential (tables, per-interval coefficients,
// Chebyshev series) and its PotDev.
// FNV-1a over bytes, to tell replicas that share a potential.
static uint64_t fnv1a(uint64_t h, const void *data, size_t n)
{
    const unsigned char *b = (const unsigned char *)data;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

static uint64_t pot_hash(const comd_params *p)
{
    uint64_t h = 1469598103934665603ull;
    const double sc[5] = {p->cutoff, p->mass, p->lj_sigma, p->lj_epsilon, (double)p->eval};
    h = fnv1a(h, &p->potential, sizeof p->potential);
    h = fnv1a(h, sc, sizeof sc);
    if (p->potential == COMD_POT_EAM) {
        for (int k = 0; k < 3; k++) {
            const double hd[3] = {p->tab_x0[k], p->tab_xn[k], p->tab_invdx[k]};
            h = fnv1a(h, hd, sizeof hd);
            h = fnv1a(h, &p->tab_n[k], sizeof p->tab_n[k]);
            if (p->tab_values[k]) h = fnv1a(h, p->tab_values[k], (size_t)(p->tab_n[k] + 3) * 8);
        }
        if (p->cheb_n >= 2 && p->cheb_coef) {
            h = fnv1a(h, p->cheb_coef, (size_t)6 * p->cheb_n * 8);
            h = fnv1a(h, p->cheb_x0, sizeof p->cheb_x0);
            h = fnv1a(h, p->cheb_xn, sizeof p->cheb_xn);
        }
    }
    return h;
}

static bool setup_pot(const comd_params *p, Replica &R)
{
    R.pot_hash = pot_hash(p);
    PotDev &P = R.pot;
    memset(&P, 0, sizeof P);
    P.kind = p->potential;
    P.eval = p->eval;
    P.rc2 = p->cutoff * p->cutoff;
    P.mass_amu = p->mass;
    R.d_cheb = nullptr;
    R.d_cheb32 = nullptr;
    bool ok = true;
    auto alloc = [&](size_t bytes) -> void * {
        void *q = nullptr;
        if (!ok || hipMalloc(&q, bytes < 16 ? 16 : bytes) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        R.owned.push_back(q);
        return q;
    };
    if (p->potential == COMD_POT_LJ) {
        const double sg = p->lj_sigma;
        P.lj_s6 = sg * sg * sg * sg * sg * sg;  // src-flat/ljForce.c:143
        P.lj_eps = p->lj_epsilon;
        P.lj_fsign = 1.0;
        return true;
    }
    for (int k = 0; k < 3 && ok; k++) {
        const int n = p->tab_n[k];
        double *t = (double *)alloc((size_t)(n + 3) * 8);
        if (!ok) break;
        ok = hipMemcpy(t, p->tab_values[k], (size_t)(n + 3) * 8, hipMemcpyHostToDevice) == hipSuccess;
        P.tab[k] = t;
        P.x0[k] = p->tab_x0[k];
        P.xn[k] = p->tab_xn[k];
        P.invdx[k] = p->tab_invdx[k];
        P.n[k] = n;
    }
    for (int k = 0; ok && k < 2; k++) {
        std::vector<double> c64;
        std::vector<float> c32;
        table_coefs(p->tab_n[k], p->tab_invdx[k], p->tab_values[k], c64, c32);
        double *a = (double *)alloc(c64.size() * 8);
        float *b = (float *)alloc(c32.size() * 4);
        if (!ok) break;
        ok = hipMemcpy(a, c64.data(), c64.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(b, c32.data(), c32.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
        P.cf64[k] = a;
        P.cf32[k] = b;
    }
    P.same_grid = p->tab_n[0] == p->tab_n[1] && p->tab_x0[0] == p->tab_x0[1] && p->tab_xn[0] == p->tab_xn[1] &&
                  p->tab_invdx[0] == p->tab_invdx[1];
    // compact rows need every pair inside the cutoff on the tables
    // (r <= cutoff <= xn), so that the interval index stays <= n
    if (ok && P.same_grid && p->cutoff <= p->tab_xn[0]) {
        // the rows the list passes read, packed per interval, and one more:
        // the last interval's polynomials re-centred on its end (t -> 1 + t),
        // read at r = xn (interval n, t = 0: the reference's value there)
        std::vector<double> f64, g64, d64, r64;
        std::vector<float> f32, g32;
        table_coefs(p->tab_n[0], p->tab_invdx[0], p->tab_values[0], f64, f32);
        table_coefs(p->tab_n[1], p->tab_invdx[1], p->tab_values[1], g64, g32);
        const int n = p->tab_n[0];
        d64.resize((size_t)8 * (n + 1));
        r64.resize((size_t)2 * (n + 1));
        for (int i = 0; i < n; i++) {
            const double row[8] = {f64[6 * i + 0], f64[6 * i + 1], f64[6 * i + 2], f64[6 * i + 3],
                                   f64[6 * i + 4], g64[6 * i + 2], g64[6 * i + 3], g64[6 * i + 4]};
            for (int q = 0; q < 8; q++) d64[8 * i + q] = row[q];
            r64[2 * i + 0] = g64[6 * i + 0];
            r64[2 * i + 1] = g64[6 * i + 1];
        }
        {
            const double *a = &d64[8 * (n - 1)], *b = &r64[2 * (n - 1)];
            double *e = &d64[8 * n], *f = &r64[2 * n];
            e[0] = a[0] + a[1];  // phi': d0 + d1 t
            e[1] = a[1];
            e[2] = a[2] + a[3] + a[4];  // phi: c0 + t (c1 + t c2)
            e[3] = a[3] + 2.0 * a[4];
            e[4] = a[4];
            e[5] = a[5] + a[6] + a[7];  // rho
            e[6] = a[6] + 2.0 * a[7];
            e[7] = a[7];
            f[0] = b[0] + b[1];  // rho'
            f[1] = b[1];
        }
        // mixed density rows: phi d0 d1 (double), then 6 floats (40 bytes)
        std::vector<float> dm((size_t)10 * (n + 1));
        for (int i = 0; i <= n; i++) {
            memcpy(&dm[10 * i], &d64[8 * i], 2 * sizeof(double));
            for (int q = 0; q < 6; q++) dm[10 * i + 4 + q] = (float)d64[8 * i + 2 + q];
        }
        double *a = (double *)alloc(d64.size() * 8), *b = (double *)alloc(r64.size() * 8);
        float *e = (float *)alloc(dm.size() * 4);
        if (ok)
            ok = hipMemcpy(a, d64.data(), d64.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(b, r64.data(), r64.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(e, dm.data(), dm.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
        P.cd64 = a;
        P.cr64 = b;
        P.cdm = e;
    }
    if (ok && p->cheb_n >= 2 && p->cheb_coef) {
        const int m = p->cheb_n, m8 = (m + 7) / 8 * 8;
        R.d_cheb = (double *)alloc((size_t)6 * m8 * 8);
        R.d_cheb32 = (float *)alloc((size_t)6 * m8 * 4);
        if (ok) {
            // series at a stride of m8, zero-padded
            std::vector<double> c64((size_t)6 * m8, 0.0);
            std::vector<float> c32((size_t)6 * m8, 0.0f);
            for (int k = 0; k < 6; k++)
                for (int j = 0; j < m; j++) {
                    c64[(size_t)k * m8 + j] = p->cheb_coef[(size_t)k * m + j];
                    c32[(size_t)k * m8 + j] = (float)p->cheb_coef[(size_t)k * m + j];
                }
            ok = hipMemcpy(R.d_cheb, c64.data(), c64.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(R.d_cheb32, c32.data(), c32.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
            P.cheb = R.d_cheb;
            P.cheb32 = R.d_cheb32;
            P.cn = m;
            P.cstride = m8;
            for (int k = 0; k < 3; k++) {
                P.ca[k] = p->cheb_x0[k];
                P.cb[k] = p->cheb_xn[k];
            }
        }
    }
    return ok;
}

// Device copies of the replicas' grids and potentials.
static int upload_descriptors(comd_ctx *c)
{
    std::vector<Grid> gs(c->nrep);
    std::vector<PotDev> ps(c->nrep);
    for (int r = 0; r < c->nrep; r++) {
        gs[r] = c->reps[r].g;
        ps[r] = c->reps[r].pot;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_grids, gs.data(), gs.size() * sizeof(Grid), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_pots, ps.data(), ps.size() * sizeof(PotDev), hipMemcpyHostToDevice));
    invalidate_graph(c);
    return COMD_OK;
}

static int ctx_create(int nrep, const comd_params *prms, comd_ctx **out)
{
    if (!prms || !out || nrep < 1) return set_err(nullptr, COMD_ERR_ARG, "null argument");
    *out = nullptr;
    long ncl = 0;
    for (int r = 0; r < nrep; r++) {
        const int rc = check_params(&prms[r]);
        if (rc) return rc;
        if (prms[r].precision != prms[0].precision)
            return set_err(nullptr, COMD_ERR_ARG, "replicas of one context share their precision");
        ncl += (long)prms[r].nbx[0] * prms[r].nbx[1] * prms[r].nbx[2];
    }
    if (ncl > (1L << 26)) return set_err(nullptr, COMD_ERR_ARG, "too many cells");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(nullptr, COMD_ERR_NODEVICE, "no HIP device");

    comd_ctx *c = new comd_ctx();
    c->nrep = nrep;
    c->reps.resize(nrep);
    c->precision = prms[0].precision;
    c->ncells = (int)ncl;
    c->nslots = ncl * CM_SLOTS;
    hipGetDevice(&c->device);
    scratch_lookup(c->device);  // with the context's device current
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_err(nullptr, COMD_ERR_HIP, "stream creation failed");
    }
    bool ok = true;
    auto alloc = [&](void **ptr, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(ptr, bytes < 16 ? 16 : bytes) != hipSuccess) ok = false;
        else hipMemsetAsync(*ptr, 0, bytes < 16 ? 16 : bytes, c->stream);
    };
    const size_t S = (size_t)c->nslots, C = (size_t)c->ncells;
    alloc((void **)&c->natoms, C * 4);
    alloc((void **)&c->id, S * 4);
    alloc((void **)&c->itype, S * 4);
    alloc((void **)&c->rx, S * 8);
    alloc((void **)&c->ry, S * 8);
    alloc((void **)&c->rz, S * 8);
    alloc((void **)&c->px, S * 8);
    alloc((void **)&c->py, S * 8);
    alloc((void **)&c->pz, S * 8);
    alloc((void **)&c->fx, S * 8);
    alloc((void **)&c->fy, S * 8);
    alloc((void **)&c->fz, S * 8);
    alloc((void **)&c->fe, S * 8);
    alloc((void **)&c->rho, S * 8);
    alloc((void **)&c->fi, S * 8);
    alloc((void **)&c->vir, S * 8);
    alloc((void **)&c->nnb, S * 4);
    alloc((void **)&c->leave, C * 8);
    alloc((void **)&c->mvidx, S * 4);
    alloc((void **)&c->meta, C * 16);
    alloc((void **)&c->counts, 16);
    alloc((void **)&c->touched, C * 4);
    alloc((void **)&c->tc, C * 4);
    alloc((void **)&c->arr_cnt, C * 4);
    alloc((void **)&c->arr, S * 4);
    // one mover record per slot at most
    alloc((void **)&c->mv_slot, S * 4);
    alloc((void **)&c->mv_rank, S * 4);
    alloc((void **)&c->mv_id, S * 4);
    alloc((void **)&c->mv_nlr, S * 16);
    alloc((void **)&c->nlr, S * 16);
    alloc((void **)&c->mv_itype, S * 4);
    alloc((void **)&c->mv_rx, S * 8);
    alloc((void **)&c->mv_ry, S * 8);
    alloc((void **)&c->mv_rz, S * 8);
    alloc((void **)&c->mv_px, S * 8);
    alloc((void **)&c->mv_py, S * 8);
    alloc((void **)&c->mv_pz, S * 8);
    alloc((void **)&c->perm, S);
    alloc((void **)&c->flags, 16);
    alloc((void **)&c->mh, (size_t)nrep * 4);
    bool eam = false;
    for (int r = 0; r < nrep; r++) eam = eam || prms[r].potential == COMD_POT_EAM;
    if (eam) {
        alloc((void **)&c->plist, C * CM_PLIST * 64 * 2);
        alloc((void **)&c->pcount, S);
        alloc((void **)&c->pvalid, C * 4);
    }
    c->use_plist = 1;
    c->nparts = 1024;
    alloc((void **)&c->partials, (size_t)c->nparts * 3 * 8);
    alloc((void **)&c->red_out, 4 * 8);
    alloc((void **)&c->d_grids, (size_t)nrep * sizeof(Grid));
    alloc((void **)&c->d_pots, (size_t)nrep * sizeof(PotDev));
    if (nrep > 1) alloc((void **)&c->d_cell_grid, C * 4);
    if (ok && hipHostMalloc((void **)&c->host_out, 4 * 8, 0) != hipSuccess) ok = false;
    if (ok && hipHostMalloc((void **)&c->host_flags, 16, 0) != hipSuccess) ok = false;
    if (ok) ok = hipStreamSynchronize(c->stream) == hipSuccess;

    // replicas: grids (cells numbered replica after replica) and potentials
    std::vector<int> cell_grid;
    int base = 0;
    for (int r = 0; ok && r < nrep; r++) {
        Replica &R = c->reps[r];
        const comd_params *p = &prms[r];
        R.prm = *p;
        for (int k = 0; k < 3; k++) R.prm.tab_values[k] = nullptr;
        R.prm.cheb_coef = nullptr;
        R.ntot = 0;
        Grid &g = R.g;
        memset(&g, 0, sizeof g);
        g.nbx0 = p->nbx[0];
        g.nbx1 = p->nbx[1];
        g.nbx2 = p->nbx[2];
        g.ncells = p->nbx[0] * p->nbx[1] * p->nbx[2];
        g.cell_base = base;
        g.periodic = p->periodic;
        g.pot = r;
        g.bs0 = p->boxsize[0];
        g.bs1 = p->boxsize[1];
        g.bs2 = p->boxsize[2];
        g.L0 = p->bounds[0];
        g.L1 = p->bounds[1];
        g.L2 = p->bounds[2];
        g.iL0 = 1.0 / g.L0;
        g.iL1 = 1.0 / g.L1;
        g.iL2 = 1.0 / g.L2;
        g.rmass = kAmuToMe * p->mass;  // src-flat/timestep.c:90
        if (nrep > 1) cell_grid.insert(cell_grid.end(), (size_t)g.ncells, r);
        base += g.ncells;
        ok = setup_pot(p, R);
        if (ok && r == 0 && R.pot.cheb) c->cheb_host.assign(p->cheb_coef, p->cheb_coef + 3 * p->cheb_n);
    }
    if (ok && nrep > 1)
        ok = hipMemcpy(c->d_cell_grid, cell_grid.data(), cell_grid.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = upload_descriptors(c) == COMD_OK;
    if (!ok) {
        free_all(c);
        delete c;
        return set_err(nullptr, COMD_ERR_NOMEM, "device allocation failed");
    }
    hipDeviceProp_t prop;
    c->ncu = 256;
    c->max_lds = 65536;
    c->lds_block = 65536;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
        if (prop.multiProcessorCount > 0) c->ncu = prop.multiProcessorCount;
        if (prop.sharedMemPerBlock > 0 && prop.sharedMemPerBlock < (size_t)c->max_lds) c->max_lds = (int)prop.sharedMemPerBlock;
        if (prop.sharedMemPerBlock > 0) c->lds_block = (int)std::min<size_t>(prop.sharedMemPerBlock, 160 << 10);
    }
    // measured on MI355X (2.05M-atom EAM): single-wave blocks, 8 per CU
    c->force_waves = 1;
    c->force_blocks_per_cu = 8;
    c->tab_lds = 1;
    c->use_graph = nrep > 1;
    c->nl_skin_milli = 400;
    c->nl_group_opt = -1;
    c->nl_rows_lds = -1;
    c->nl_threads_div = 1;
    c->nl_lanes = 0;
    c->adj_lists = 1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        free_all(c);
        delete c;
        return set_err(nullptr, COMD_ERR_HIP, "initialisation failed");
    }
    *out = c;
    return COMD_OK;
}

static int check_rep(comd_ctx *c, int r)
{
    if (r < 0 || r >= c->nrep) return set_err(c, COMD_ERR_ARG, "replica %d out of range", r);
    return COMD_OK;
}

// Neighbour lists for the atoms just binned or uploaded (ids below nids):
// used when the context holds one periodic simulation whose box has at
// least three link cells and three (cutoff + skin) cells per dimension;
// otherwise the link-cell kernels do every pass.  Sets the id -> slot and
// id -> position maps and requests a build.
static int nl_setup(comd_ctx *c)
{
    c->nl_active = 0;
    invalidate_graph(c);
    const int nrep = c->nrep;
    if (c->nl_skin_milli <= 0) return COMD_OK;
    const double skin = 1e-3 * c->nl_skin_milli;
    // every replica periodic, with atoms and at least 3 auxiliary cells per
    // dimension; the replicas of a kind share one potential
    int eam_rep = -1, lj_rep = -1;
    std::vector<std::array<int, 3>> nmax(nrep);
    long kmax = 16;
    for (int r = 0; r < nrep; r++) {
        const Replica &R = c->reps[r];
        const comd_params &p = R.prm;
        if (R.ntot <= 0 || R.nl_ids <= 0) return COMD_OK;
        if (!p.periodic || p.nbx[0] < 3 || p.nbx[1] < 3 || p.nbx[2] < 3) return COMD_OK;
        int &kr = p.potential == COMD_POT_LJ ? lj_rep : eam_rep;
        if (kr < 0) kr = r;
        else if (c->reps[kr].pot_hash != R.pot_hash || c->reps[kr].pot.lj_fsign != R.pot.lj_fsign ||
                 c->reps[kr].pot.eval != R.pot.eval)
            return COMD_OK;
        const double rl = p.cutoff + skin;
        for (int j = 0; j < 3; j++) {
            nmax[r][j] = (int)floor(p.bounds[j] / rl);
            if (nmax[r][j] < 3) return COMD_OK;
        }
        const double vol = p.bounds[0] * p.bounds[1] * p.bounds[2];
        const double expect = 4.0 / 3.0 * 3.141592653589793 * rl * rl * rl * (double)R.ntot / vol;
        kmax = std::max<long>(kmax, (((long)(1.5 * expect) + 16 + 15) / 16) * 16);
    }
    // groups of auxiliary cells: of the options whose staged halo (1.2x the
    // mean occupancy, plus 64 atoms, in the densest replica) fits the LDS a
    // block may use in every pass (16 B per halo atom in mixed precision, up
    // to 32 B in FP64, and a byte of slot map) beside the kernels' static
    // LDS, the largest that still gives two groups (blocks) per CU, else the
    // one with the most groups; a tuned option is taken as the largest
    // allowed.  Each replica's auxiliary grid is rounded down to whole groups
    // per dimension (cells a little larger than cutoff + skin): a partial
    // group at an edge would take a whole block for a sliver of atoms.
    const int gopts[5][3] = {{4, 4, 3}, {4, 4, 2}, {4, 2, 2}, {2, 2, 2}, {2, 2, 1}};
    // the shapes of the per-replica search of batches (the halo of each,
    // (x + 2)(y + 2)(z + 2) slots, within CM_NL_SLOTS)
    const int bopts[10][3] = {{4, 4, 3}, {4, 3, 3}, {4, 4, 2}, {3, 3, 3}, {4, 3, 2},
                              {3, 3, 2}, {4, 2, 2}, {3, 2, 2}, {2, 2, 2}, {2, 2, 1}};
    const long lds_avail = (long)c->lds_block - CM_NL_STATIC_LDS;
    const long hb = c->precision == COMD_PREC_MIXED ? 16 : 32;
    std::vector<std::array<int, 3>> gd(nrep), na(nrep), ng(nrep);
    long hcap = 0, naux = 0, ngroups = 0, threads = 64;
    // lanes per home atom in the force passes: as tuned, else two when the
    // atoms would fill fewer than two waves per CU at one lane each (each
    // lane then walks half the list: 32 000 Lennard-Jones atoms, pass
    // 0.034 -> 0.026 ms; at 256 000 atoms and above two lanes measured no
    // better or slower)
    long natoms_all = 0;
    for (const Replica &R : c->reps) natoms_all += R.ntot;
    const int lanes = c->nl_lanes > 0 ? c->nl_lanes : (natoms_all < 128L * c->ncu ? 2 : 1);
    // o >= 0: that option for every replica; o < 0 (batches): per replica
    // the shape with the fewest groups (then the fewest staged halo atoms)
    // among those whose groups' atoms (10 % spare) fit 768 threads and whose
    // staged halo keeps two blocks per CU in the mixed passes, so that
    // replicas of different density and cell rounding get groups of similar
    // size (one block shape and one halo capacity serve them all).  Each
    // option is tried on the grid rounded down to whole groups and on the
    // grid of the smallest cells >= cutoff + skin split into groups whose
    // extents differ by at most one (rounding a small replica's grid down
    // can make its cells, and so its halos, much larger)
    long htarget = 1L << 30;
    if (c->precision == COMD_PREC_MIXED) {
        long tab = 0;
        for (const Replica &R : c->reps) tab = std::max<long>(tab, (long)cm_nl_tab_bytes(R.pot, CM_KIND_EAM1, 1));
        htarget = ((long)c->lds_block / 2 - CM_NL_STATIC_LDS - tab) / (hb + 1) / 64 * 64;
    }
    auto hcap_of = [](double occ, long slots) { return (((long)(1.2 * occ * (double)slots) + 64 + 63) / 64) * 64; };
    // replica r's grid and groups for option (opts, oo), rounded or balanced
    auto layout = [&](int r, const int (*opts)[3], int oo, int bal) {
        long nr = 1;
        for (int j = 0; j < 3; j++) {
            gd[r][j] = std::min(opts[oo][j], nmax[r][j]);
            if (bal) {
                na[r][j] = nmax[r][j];
                ng[r][j] = (na[r][j] + gd[r][j] - 1) / gd[r][j];
                gd[r][j] = (na[r][j] + ng[r][j] - 1) / ng[r][j];
            } else {
                na[r][j] = nmax[r][j] / gd[r][j] * gd[r][j];
                if (na[r][j] < 3) {  // keep >= 3 cells (one group across the dimension)
                    gd[r][j] = nmax[r][j];
                    na[r][j] = nmax[r][j];
                }
                ng[r][j] = na[r][j] / gd[r][j];
            }
            nr *= na[r][j];
        }
        return nr;
    };
    auto shape = [&](int o) {  // fills gd, na, ng, hcap, naux, ngroups, threads
        hcap = naux = ngroups = 0;
        threads = 64;
        for (int r = 0; r < nrep; r++) {
            long nr;
            if (o >= 0) {
                nr = layout(r, gopts, o, 0);
            } else {
                int bo = 9, bb = 1;  // none fits: the smallest shape
                long bg = -1;
                double bh = 0.0;
                for (int oo = 0; oo < 10; oo++)
                    for (int bal = 0; bal < 2; bal++) {
                        const long n = layout(r, bopts, oo, bal);
                        const double occ = (double)c->reps[r].ntot / (double)n;
                        const long slots = (long)(gd[r][0] + 2) * (gd[r][1] + 2) * (gd[r][2] + 2);
                        if (1.1 * occ * gd[r][0] * gd[r][1] * gd[r][2] > 768.0 || hcap_of(occ, slots) > htarget)
                            continue;
                        const long g = (long)ng[r][0] * ng[r][1] * ng[r][2];
                        double h = occ;  // staged halo atoms: prod_j (na_j + 2 ng_j) cells
                        for (int j = 0; j < 3; j++) h *= (double)(na[r][j] + 2 * ng[r][j]);
                        if (bg < 0 || g < bg || (g == bg && h < bh)) bo = oo, bb = bal, bg = g, bh = h;
                    }
                nr = layout(r, bopts, bo, bb);
            }
            const double occ = (double)c->reps[r].ntot / (double)nr;
            const long cells = (long)gd[r][0] * gd[r][1] * gd[r][2];
            const long slots = (long)(gd[r][0] + 2) * (gd[r][1] + 2) * (gd[r][2] + 2);
            hcap = std::max<long>(hcap, hcap_of(occ, slots));
            threads = std::max<long>(
                threads, ((long)(1.1 * occ * (double)cells * lanes / c->nl_threads_div) + 63) / 64 * 64);
            naux += nr;
            ngroups += (long)ng[r][0] * ng[r][1] * ng[r][2];
        }
    };
    int best = -1;
    long best_ng = -1;
    if (nrep > 1 && c->nl_group_opt < 0) {
        shape(-1);
        if (hcap * (hb + 1) <= lds_avail && hcap <= 65535) best = -1, best_ng = 0;
        else best_ng = -2;  // fall back to one option for all
    }
    if (best_ng != 0) {
        for (int o = c->nl_group_opt >= 0 ? c->nl_group_opt : 0; o < 5; o++) {
            shape(o);
            if (hcap * (hb + 1) > lds_avail || hcap > 65535) continue;
            if (c->nl_group_opt >= 0 || ngroups >= 2L * c->ncu) {
                best = o;
                break;
            }
            if (ngroups > best_ng) {
                best = o;
                best_ng = ngroups;
            }
        }
        if (best < 0) return COMD_OK;  // too dense for staged halos
        shape(best);
    }
    // replicas in group order: the EAM ones, then the LJ ones
    std::vector<int> gorder;
    for (int pass = 0; pass < 2; pass++)
        for (int r = 0; r < nrep; r++)
            if ((c->reps[r].prm.potential == COMD_POT_LJ) == (pass == 1)) gorder.push_back(r);
    long nids = 0;
    for (int r = 0; r < nrep; r++) nids += c->reps[r].nl_ids;
    if (nids > c->nl_alloc_n || kmax > c->nl_alloc_k || naux > c->nl_alloc_aux || ngroups > c->nl_alloc_groups ||
        nrep > c->nl_alloc_rep) {
        free_nl(c);
        bool ok = true;
        auto alloc = [&](void **ptr, size_t bytes) {
            if (ok && hipMalloc(ptr, bytes < 16 ? 16 : bytes) != hipSuccess) ok = false;
        };
        const size_t n = (size_t)nids;
        alloc((void **)&c->nl.pos, n * sizeof(CmPos));
        alloc((void **)&c->nl.selfh, n * 2);
        alloc((void **)&c->nl.slot, n * 4);
        alloc((void **)&c->nl.order, n * 4);
        alloc((void **)&c->nl.fq, n * 8);
        alloc((void **)&c->nl.f1, n * 24);
        alloc((void **)&c->nl.nbl, n * (size_t)kmax * 2);
        alloc((void **)&c->nl.ncnt, n * 4);
        alloc((void **)&c->nl.aux_cnt, (size_t)naux * 4);
        alloc((void **)&c->nl.aux_n, (size_t)naux * 4);
        alloc((void **)&c->nl.aux_start, (size_t)naux * 4);
        alloc((void **)&c->nl.aux_mem, (size_t)naux * CM_NL_AUXCAP * 4);
        alloc((void **)&c->nl.gstart, (size_t)(ngroups + 1) * 4);
        alloc((void **)&c->nl.cell_order, (size_t)naux * 4);
        alloc((void **)&c->nl.group_q0, (size_t)(ngroups + 1) * 4);
        alloc((void **)&c->nl.group_rep, (size_t)ngroups * 4);
        alloc((void **)&c->nl.aux_group, (size_t)naux * 4);
        alloc((void **)&c->nl.rep, (size_t)nrep * sizeof(NlRep));
        alloc((void **)&c->nl.nonempty, (size_t)c->ncells * 4);
        alloc((void **)&c->nl.count, 16);
        alloc((void **)&c->nl.flags, 16);
        if (!ok) {
            free_nl(c);
            return set_err(c, COMD_ERR_NOMEM, "neighbour-list buffers");
        }
        c->nl_alloc_n = nids;
        c->nl_alloc_k = kmax;
        c->nl_alloc_aux = naux;
        c->nl_alloc_groups = ngroups;
        c->nl_alloc_rep = nrep;
    }
    CmList &l = c->nl;
    l.n = (int)nids;
    l.kmax = (int)kmax;
    l.nrep = nrep;
    l.naux = (int)naux;
    l.ngroups = (int)ngroups;
    l.hcap = (int)hcap;
    l.lds_avail = (int)lds_avail;
    l.rows_lds = c->nl_rows_lds;
    l.pass_blocks = (int)ngroups;
    // threads per pass block: a group's home atoms with 10 % to spare
    l.pass_threads = (int)std::min<long>(768, threads);
    l.lanes = lanes;
    l.tailq = cm_nl_pick_tailq(l);
    std::vector<NlRep> reps(nrep);
    std::vector<int> corder((size_t)naux), gq0((size_t)ngroups + 1), grep((size_t)ngroups), agroup((size_t)naux);
    long q = 0, g = 0, abase = 0, ibase = 0;
    c->nl_eam_groups = c->nl_lj_groups = 0;
    for (int r : gorder) {
        Replica &Rp = c->reps[r];
        const comd_params &p = Rp.prm;
        NlRep &R = reps[r];
        R.L0 = p.bounds[0];
        R.L1 = p.bounds[1];
        R.L2 = p.bounds[2];
        R.iL0 = 1.0 / R.L0;
        R.iL1 = 1.0 / R.L1;
        R.iL2 = 1.0 / R.L2;
        R.na0 = na[r][0];
        R.na1 = na[r][1];
        R.na2 = na[r][2];
        R.ab0 = R.L0 / R.na0;
        R.ab1 = R.L1 / R.na1;
        R.ab2 = R.L2 / R.na2;
        R.iab0 = 1.0 / R.ab0;
        R.iab1 = 1.0 / R.ab1;
        R.iab2 = 1.0 / R.ab2;
        R.gd0 = gd[r][0];
        R.gd1 = gd[r][1];
        R.gd2 = gd[r][2];
        R.ng0 = ng[r][0];
        R.ng1 = ng[r][1];
        R.ng2 = ng[r][2];
        const double rl = p.cutoff + skin;
        R.rl2 = rl * rl;
        R.rc2 = p.cutoff * p.cutoff;
        {
            // 32-bit fixed point of the mixed passes' group frames: a halo
            // atom lies within (gd/2 + 1) cells of the group centre plus the
            // half skin it may have moved since the build; 5 % to spare.  A
            // power of two, so that scaling by it or its inverse is exact:
            // the passes then work in fixed-point units and fold the scale
            // into constants
            const double ab[3] = {R.ab0, R.ab1, R.ab2};
            double ext = 0.0;
            for (int j = 0; j < 3; j++) ext = std::max(ext, (0.5 * gd[r][j] + 1.0) * ab[j]);
            ext += skin;
            R.qscale = std::ldexp(1.0, (int)std::floor(std::log2(2147483648.0 / (1.05 * ext))));
            R.qinv = 1.0 / R.qscale;
        }
        R.aux_base = (int)abase;
        R.group_base = (int)g;
        R.id_base = (int)ibase;
        R.nid = (int)Rp.nl_ids;
        Rp.g.nl_id_base = R.id_base;
        // its auxiliary cells in build order: groups (x fastest), then the
        // group's cells (x fastest)
        for (int gz = 0; gz < R.ng2; gz++)
            for (int gy = 0; gy < R.ng1; gy++)
                for (int gx = 0; gx < R.ng0; gx++) {
                    gq0[g] = (int)q;
                    grep[g] = r;
                    int x0, lx, y0, ly, z0, lz;
                    group_span(gx, R.na0, R.ng0, x0, lx);
                    group_span(gy, R.na1, R.ng1, y0, ly);
                    group_span(gz, R.na2, R.ng2, z0, lz);
                    for (int z = z0; z < z0 + lz; z++)
                        for (int y = y0; y < y0 + ly; y++)
                            for (int x = x0; x < x0 + lx; x++) {
                                const long a = abase + x + R.na0 * (y + (long)R.na1 * z);
                                corder[q++] = (int)a;
                                agroup[a] = (int)g;
                            }
                    g++;
                }
        const int rg = R.ng0 * R.ng1 * R.ng2;
        if (p.potential == COMD_POT_LJ) c->nl_lj_groups += rg;
        else c->nl_eam_groups += rg;
        abase += (long)R.na0 * R.na1 * R.na2;
        ibase += Rp.nl_ids;
    }
    gq0[g] = (int)q;
    c->nl_eam_rep = eam_rep;
    c->nl_lj_rep = lj_rep;
    HIPCHK(c, hipMemcpyAsync((void *)l.cell_order, corder.data(), corder.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void *)l.group_q0, gq0.data(), gq0.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void *)l.group_rep, grep.data(), grep.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void *)l.aux_group, agroup.data(), agroup.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void *)l.rep, reps.data(), reps.size() * sizeof(NlRep), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc = upload_descriptors(c);  // the grids' list id bases
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(l.nbl, 0, (size_t)nids * (size_t)kmax * 2, c->stream));
    HIPCHK(c, hipMemsetAsync(l.flags, 0, 16, c->stream));
    HIPCHK(c, hipMemsetAsync(l.count, 0, 16, c->stream));
    HIPCHK(c, hipMemsetAsync(l.aux_cnt, 0, (size_t)naux * 4, c->stream));
    c->nl_active = 1;
    c->nl_stale = 0;
    c->nl_invalid_seen = 0;
    cm_nl_init(dev_state(c), l, c->stream);
    HIPCHK(c, hipGetLastError());
    return COMD_OK;
}

// ---- state transfer of one replica

static int upload_rep(comd_ctx *c, int r, const int *natoms, const int *id, const int *itype, const double *rr,
                      const double *p)
{
    if (!natoms || !id || !rr) return set_err(c, COMD_ERR_ARG, "null argument");
    const Grid &g = c->reps[r].g;
    const size_t C = (size_t)g.ncells, S = C * CM_SLOTS;
    const size_t c0 = (size_t)g.cell_base, s0 = c0 * CM_SLOTS;
    int tot = 0;
    long nids = 0;
    for (size_t b = 0; b < C; b++) {
        if (natoms[b] < 0 || natoms[b] > CM_SLOTS) return set_err(c, COMD_ERR_ARG, "cell %zu holds %d atoms", b, natoms[b]);
        tot += natoms[b];
        for (int i = 0; i < natoms[b]; i++) {
            const int a = id[b * CM_SLOTS + i];
            if (a < 0 || a >= (1 << 30)) return set_err(c, COMD_ERR_ARG, "atom id %d in cell %zu out of range", a, b);
            nids = std::max<long>(nids, (long)a + 1);
        }
    }
    // the neighbour lists and the id -> slot map index atoms by id: each id
    // must appear once
    {
        std::vector<uint8_t> seen((size_t)nids, 0);
        for (size_t b = 0; b < C; b++)
            for (int i = 0; i < natoms[b]; i++) {
                const int a = id[b * CM_SLOTS + i];
                if (seen[a]) return set_err(c, COMD_ERR_ARG, "atom id %d appears twice", a);
                seen[a] = 1;
            }
    }
    std::vector<double> buf(S);
    std::vector<int> it(S, 0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->natoms + c0, natoms, C * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->id + s0, id, S * 4, hipMemcpyHostToDevice));
    if (itype) memcpy(it.data(), itype, S * 4);
    HIPCHK(c, hipMemcpy(c->itype + s0, it.data(), S * 4, hipMemcpyHostToDevice));
    double *dst[6] = {c->rx, c->ry, c->rz, c->px, c->py, c->pz};
    for (int k = 0; k < 6; k++) {
        const double *src = k < 3 ? rr : p;
        if (!src) std::fill(buf.begin(), buf.end(), 0.0);
        else
            for (size_t s = 0; s < S; s++) buf[s] = src[3 * s + (k % 3)];
        HIPCHK(c, hipMemcpy(dst[k] + s0, buf.data(), S * 8, hipMemcpyHostToDevice));
    }
    c->ntot += tot - c->reps[r].ntot;
    c->reps[r].ntot = tot;
    c->have_atoms = 1;
    c->reps[r].nl_ids = nids;
    return nl_setup(c);
}

static int download_rep(comd_ctx *c, int r, int *natoms, int *id, int *itype, double *rr, double *p, double *f,
                        double *rho, double *fi)
{
    const Grid &g = c->reps[r].g;
    const size_t C = (size_t)g.ncells, S = C * CM_SLOTS;
    const size_t c0 = (size_t)g.cell_base, s0 = c0 * CM_SLOTS;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (natoms) HIPCHK(c, hipMemcpy(natoms, c->natoms + c0, C * 4, hipMemcpyDeviceToHost));
    if (id) HIPCHK(c, hipMemcpy(id, c->id + s0, S * 4, hipMemcpyDeviceToHost));
    if (itype) HIPCHK(c, hipMemcpy(itype, c->itype + s0, S * 4, hipMemcpyDeviceToHost));
    std::vector<double> buf(S);
    auto gather = [&](double *dst, int stride, int k, const double *src) -> int {
        if (hipMemcpy(buf.data(), src + s0, S * 8, hipMemcpyDeviceToHost) != hipSuccess) return COMD_ERR_HIP;
        for (size_t s = 0; s < S; s++) dst[stride * s + k] = buf[s];
        return COMD_OK;
    };
    int rc = COMD_OK;
    if (rr) {
        rc |= gather(rr, 3, 0, c->rx);
        rc |= gather(rr, 3, 1, c->ry);
        rc |= gather(rr, 3, 2, c->rz);
    }
    if (p) {
        rc |= gather(p, 3, 0, c->px);
        rc |= gather(p, 3, 1, c->py);
        rc |= gather(p, 3, 2, c->pz);
    }
    if (f) {
        rc |= gather(f, 4, 0, c->fx);
        rc |= gather(f, 4, 1, c->fy);
        rc |= gather(f, 4, 2, c->fz);
        rc |= gather(f, 4, 3, c->fe);
    }
    if (rho) HIPCHK(c, hipMemcpy(rho, c->rho + s0, S * 8, hipMemcpyDeviceToHost));
    if (fi) HIPCHK(c, hipMemcpy(fi, c->fi + s0, S * 8, hipMemcpyDeviceToHost));
    if (rc) return set_err(c, COMD_ERR_HIP, "download failed");
    return COMD_OK;
}

static int bin_rep_device(comd_ctx *c, int r, int n, const double *d_xyz, const double *d_pxyz, int itype)
{
    const Grid &g = c->reps[r].g;
    if (n < 0 || (n > 0 && !d_xyz)) return set_err(c, COMD_ERR_ARG, "bad atom list");
    if ((long)n > (long)g.ncells * CM_SLOTS) return set_err(c, COMD_ERR_OVERFLOW, "more atoms than slots");
    int rc = clear_flags(c);
    if (rc) return rc;
    DevState d = dev_state(c);
    cm_launch_bin(d, g, c->stream, n, d_xyz, d_pxyz, itype, c->meta, c->arr);
    rc = check_flags(c);
    if (rc) return rc;
    c->ntot += n - c->reps[r].ntot;
    c->reps[r].ntot = n;
    c->have_atoms = 1;
    c->reps[r].nl_ids = n;
    return nl_setup(c);
}

static int bin_rep(comd_ctx *c, int r, int n, const double *xyz, const double *pxyz, int itype)
{
    if (n < 0 || (n > 0 && !xyz)) return set_err(c, COMD_ERR_ARG, "bad atom list");
    double *dx = nullptr, *dp = nullptr;
    const size_t bytes = (size_t)(n > 0 ? n : 1) * 3 * 8;
    if (hipMalloc((void **)&dx, bytes) != hipSuccess) return set_err(c, COMD_ERR_NOMEM, "no memory for input");
    if (pxyz && hipMalloc((void **)&dp, bytes) != hipSuccess) {
        hipFree(dx);
        return set_err(c, COMD_ERR_NOMEM, "no memory for input");
    }
    int rc = COMD_OK;
    if (n > 0) {
        if (hipMemcpy(dx, xyz, (size_t)n * 24, hipMemcpyHostToDevice) != hipSuccess ||
            (dp && hipMemcpy(dp, pxyz, (size_t)n * 24, hipMemcpyHostToDevice) != hipSuccess))
            rc = set_err(c, COMD_ERR_HIP, "input copy failed");
    }
    if (!rc) rc = bin_rep_device(c, r, n, dx, dp, itype);
    hipStreamSynchronize(c->stream);
    hipFree(dx);
    if (dp) hipFree(dp);
    return rc;
}

// PE (the reference's s->e), KE and the virial "stress" of one replica.
static int energy_rep(comd_ctx *c, int r, double *pe, double *ke, double *stress)
{
    const Replica &R = c->reps[r];
    hipEvent_t ev = nullptr;
    tbegin(c, COMD_K_REDUCE, &ev);
    DevState d = dev_state(c);
    cm_launch_reduce(d, R.g, c->partials, c->nparts, c->stream);
    cm_launch_reduce_final(c->partials, c->nparts, c->red_out, c->stream);
    tend(c, COMD_K_REDUCE, ev);
    HIPCHK(c, hipMemcpyAsync(c->host_out, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double e = c->host_out[0];
    if (R.prm.potential == COMD_POT_LJ) e = e * 4.0 * R.prm.lj_epsilon * R.pot.lj_s6;  // ljForce.c:217
    if (pe) *pe = e;
    if (ke) *ke = c->host_out[1];
    if (stress) *stress = c->host_out[2] / (R.g.L0 * R.g.L1 * R.g.L2);
    return COMD_OK;
}

// Derivative series of a Chebyshev series on [a, b]: the recurrence of
// src-flat/cheby.c:219-270 (last coefficient 0, d[j-1] = d[j+1] + 2 j c[j],
// scaled by 2/(b-a)).
static void cheb_derivative(int n, double a, double b, const double *c, double *d)
{
    d[n - 1] = 0.0;
    d[n - 2] = 2.0 * (n - 1) * c[n - 1];
    for (int j = n - 2; j > 0; j--) d[j - 1] = d[j + 1] + 2 * (j)*c[j];
    const double con = 2.0 / (b - a);
    for (int j = 0; j < n; j++) d[j] = d[j] * con;
}

// ---- C-ABI

extern "C" {

int comd_create(const comd_params *p, comd_ctx **out) { return ctx_create(1, p, out); }

int comd_batch_create(int n, const comd_params *prms, comd_ctx **out) { return ctx_create(n, prms, out); }

int comd_destroy(comd_ctx *ctx)
{
    if (!ctx) return COMD_OK;
    hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return COMD_OK;
}

const char *comd_last_error(const comd_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_err;
}

int comd_sizes(const comd_ctx *ctx, int *ncells, int *natoms_total)
{
    if (!ctx) return COMD_ERR_ARG;
    if (ncells) *ncells = ctx->ncells;
    if (natoms_total) *natoms_total = ctx->ntot;
    return COMD_OK;
}

int comd_batch_info(comd_ctx *c, int r, int *cell_base, int *ncells, int *natoms)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (cell_base) *cell_base = c->reps[r].g.cell_base;
    if (ncells) *ncells = c->reps[r].g.ncells;
    if (natoms) *natoms = c->reps[r].ntot;
    return COMD_OK;
}

int comd_upload(comd_ctx *c, const int *natoms, const int *id, const int *itype, const double *r, const double *p)
{
    if (!c) return COMD_ERR_ARG;
    return upload_rep(c, 0, natoms, id, itype, r, p);
}

int comd_batch_upload(comd_ctx *c, int r, const int *natoms, const int *id, const int *itype, const double *rr,
                      const double *p)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    return rc ? rc : upload_rep(c, r, natoms, id, itype, rr, p);
}

int comd_download(comd_ctx *c, int *natoms, int *id, int *itype, double *r, double *p, double *f, double *rho, double *fi)
{
    if (!c) return COMD_ERR_ARG;
    return download_rep(c, 0, natoms, id, itype, r, p, f, rho, fi);
}

int comd_batch_download(comd_ctx *c, int r, int *natoms, int *id, int *itype, double *rr, double *p, double *f,
                        double *rho, double *fi)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    return rc ? rc : download_rep(c, r, natoms, id, itype, rr, p, f, rho, fi);
}

int comd_bin_atoms_device(comd_ctx *c, int n, const double *d_xyz, const double *d_pxyz, int itype)
{
    if (!c) return COMD_ERR_ARG;
    return bin_rep_device(c, 0, n, d_xyz, d_pxyz, itype);
}

int comd_bin_atoms(comd_ctx *c, int n, const double *xyz, const double *pxyz, int itype)
{
    if (!c) return COMD_ERR_ARG;
    return bin_rep(c, 0, n, xyz, pxyz, itype);
}

int comd_batch_bin_atoms(comd_ctx *c, int r, int n, const double *xyz, const double *pxyz, int itype)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    return rc ? rc : bin_rep(c, r, n, xyz, pxyz, itype);
}

int comd_rebox(comd_ctx *c)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    int rc = clear_flags(c);
    if (rc) return rc;
    op_move(c, 0.0, 0.0, 0, 0);
    op_rebox(c);
    return check_flags(c);
}

int comd_force(comd_ctx *c)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    op_force(c);
    HIPCHK(c, hipGetLastError());
    return COMD_OK;
}

int comd_energy(comd_ctx *c, double *pe, double *ke, double *stress)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    double a = 0, b = 0, s = 0;
    for (int r = 0; r < c->nrep; r++) {
        double x, y, z;
        int rc = energy_rep(c, r, &x, &y, &z);
        if (rc) return rc;
        a += x;
        b += y;
        s += z;
    }
    if (pe) *pe = a;
    if (ke) *ke = b;
    if (stress) *stress = s;
    return COMD_OK;
}

int comd_batch_energy(comd_ctx *c, int r, double *pe, double *ke, double *stress)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    return energy_rep(c, r, pe, ke, stress);
}

int comd_nsteps(comd_ctx *c, int n, double dt_seconds, double *pe_out)
{
    if (!c || n < 0) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    int rc = clear_flags(c);
    if (rc) return rc;
    const double dt = dt_seconds * kBohrPerAtuToAPerS;  // timestep.c:126
    rc = run_steps(c, n, dt);
    if (rc) return rc;
    op_force(c);  // timestep.c:136
    HIPCHK(c, hipGetLastError());
    rc = check_flags(c);
    if (rc) return rc;
    if (pe_out) return comd_energy(c, pe_out, nullptr, nullptr);
    return COMD_OK;
}

int comd_steps(comd_ctx *c, int n, double dt_seconds)
{
    if (!c || n < 0) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    int rc = clear_flags(c);
    if (rc) return rc;
    return run_steps(c, n, dt_seconds * kBohrPerAtuToAPerS);
}

int comd_batch_step(comd_ctx *c, int n, double dt_seconds) { return comd_steps(c, n, dt_seconds); }

int comd_advance_positions(comd_ctx *c, double dt_internal)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    int rc = clear_flags(c);
    if (rc) return rc;
    op_move(c, 0.0, dt_internal, 0, 1);
    op_rebox(c);
    return check_flags(c);
}

int comd_advance_velocity(comd_ctx *c, double dt_internal)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    // kick only: k_move with drift and rebin off leaves r and the rebin
    // state untouched
    op_move(c, dt_internal, 0.0, 1, 0, 0);
    HIPCHK(c, hipGetLastError());
    return COMD_OK;
}

int comd_set_precision(comd_ctx *c, int precision)
{
    if (!c || (precision != COMD_PREC_FP64 && precision != COMD_PREC_MIXED)) return COMD_ERR_ARG;
    if (precision != c->precision) invalidate_graph(c);
    const int old = c->precision;
    c->precision = precision;
    for (Replica &R : c->reps) R.prm.precision = precision;
    // the list groups depend on the precision (halo atom width): rebuild
    if (precision != old && c->nl_active) return nl_setup(c);
    return COMD_OK;
}

int comd_set_eval(comd_ctx *c, int eval)
{
    if (!c || (eval != COMD_EVAL_TABLE && eval != COMD_EVAL_CHEB)) return COMD_ERR_ARG;
    for (Replica &R : c->reps)
        if (R.prm.potential == COMD_POT_EAM && eval == COMD_EVAL_CHEB && !R.pot.cheb)
            return set_err(c, COMD_ERR_ARG, "no Chebyshev coefficients");
    for (Replica &R : c->reps) {
        R.prm.eval = eval;
        R.pot.eval = eval;
    }
    return upload_descriptors(c);
}

int comd_device_info(int *count, char *name, int name_len)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    if (name && name_len > 0) {
        name[0] = 0;
        int dev = 0;
        hipDeviceProp_t prop;
        if (n > 0 && hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            snprintf(name, (size_t)name_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    return n > 0 ? COMD_OK : COMD_ERR_NODEVICE;
}

int comd_tune(comd_ctx *c, int key, int value)
{
    if (!c) return COMD_ERR_ARG;
    switch (key) {
    case COMD_TUNE_WAVES_PER_BLOCK:
        if (value < 1 || value > 4) return set_err(c, COMD_ERR_ARG, "waves per block must be 1..4");
        c->force_waves = value;
        break;
    case COMD_TUNE_BLOCKS_PER_CU:
        if (value < 1 || value > 64) return set_err(c, COMD_ERR_ARG, "blocks per CU must be 1..64");
        c->force_blocks_per_cu = value;
        break;
    case COMD_TUNE_TABLES_IN_LDS:
        c->tab_lds = value ? 1 : 0;
        break;
    case COMD_TUNE_BALANCED_DRAIN:
        for (Replica &R : c->reps) R.pot.balanced = value ? 1 : 0;
        return upload_descriptors(c);
    case COMD_TUNE_GRAPH:
        c->use_graph = value ? 1 : 0;
        break;
    case COMD_TUNE_PAIR_LISTS:
        c->use_plist = value ? 1 : 0;
        break;
    case COMD_TUNE_LJ_PHYSICAL:
        for (Replica &R : c->reps)
            if (R.prm.potential == COMD_POT_LJ) R.pot.lj_fsign = value ? -1.0 : 1.0;
        invalidate_graph(c);
        return upload_descriptors(c);
    case COMD_TUNE_LIST_GROUPS:
        if (value < -1 || value > 4) return set_err(c, COMD_ERR_ARG, "list group option must be -1..4");
        c->nl_group_opt = value;
        if (c->have_atoms) return nl_setup(c);
        break;
    case COMD_TUNE_LIST_ROWS_LDS:
        if (value < -1 || value > 1) return set_err(c, COMD_ERR_ARG, "list table rows option must be -1..1");
        c->nl_rows_lds = value;
        c->nl.rows_lds = value;
        break;
    case COMD_TUNE_LIST_THREADS_DIV:
        if (value < 1 || value > 4) return set_err(c, COMD_ERR_ARG, "list threads divisor must be 1..4");
        c->nl_threads_div = value;
        if (c->have_atoms) return nl_setup(c);
        break;
    case COMD_TUNE_LIST_LANES:
        if (value != 0 && value != 1 && value != 2 && value != 4)
            return set_err(c, COMD_ERR_ARG, "list lanes must be 0 (by size), 1, 2 or 4");
        c->nl_lanes = value;
        if (c->have_atoms) return nl_setup(c);
        break;
    case COMD_TUNE_ADJOINT_LISTS:
        if (value < 0 || value > 2) return set_err(c, COMD_ERR_ARG, "adjoint lists option must be 0, 1 or 2");
        c->adj_lists = value;
        break;
    case COMD_TUNE_SKIN_MILLI:
        if (value < 0 || value > 3000) return set_err(c, COMD_ERR_ARG, "skin must be 0..3000 (1e-3 A)");
        c->nl_skin_milli = value;
        if (c->have_atoms) return nl_setup(c);
        break;
    default:
        return set_err(c, COMD_ERR_ARG, "unknown tuning key %d", key);
    }
    invalidate_graph(c);
    return COMD_OK;
}

// Time of the fastest of 3 force calls after one untimed call (the list
// build, when one is due, happens in the untimed call).
static int time_force(comd_ctx *c, hipEvent_t a, hipEvent_t b, float *ms)
{
    op_force(c);
    float t = 0, tmin = 1e30f;
    for (int k = 0; k < 3; k++) {
        hipEventRecord(a, c->stream);
        op_force(c);
        hipEventRecord(b, c->stream);
        if (hipEventSynchronize(b) != hipSuccess) return set_err(c, COMD_ERR_HIP, "autotune timing failed");
        hipEventElapsedTime(&t, a, b);
        tmin = std::min(tmin, t);
    }
    *ms = tmin;
    return COMD_OK;
}

// The list passes' launch shape by timing: the group of auxiliary cells
// (the per-replica search of batches, then the five fixed options), the
// table rows in LDS or read from global memory, and one or two home atoms
// per thread; each option rebuilds the lists (its groups, halo capacity
// and threads per block follow from the shape).  Options whose lists do
// not fit (no list passes, or a build that overflowed) are skipped.
static int autotune_lists(comd_ctx *c, int *group_opt, int *rows_lds, int *threads_div, double *best_ms)
{
    const int g0 = c->nl_group_opt, r0 = c->nl_rows_lds, d0 = c->nl_threads_div;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    bool rows_matter = false;
    for (const Replica &R : c->reps)
        rows_matter = rows_matter || cm_nl_tab_bytes(R.pot, CM_KIND_EAM1, c->precision == COMD_PREC_MIXED) > 0;
    double best = 1e300;
    int bg = g0, br = r0, bd = d0;
    int rc = COMD_OK;
    for (int g = -1; g <= 4 && !rc; g++)
        for (int r = rows_matter ? 0 : -1; r >= -1 && !rc; r--)
            for (int dv = 1; dv <= 2 && !rc; dv++) {
                c->nl_group_opt = g;
                c->nl_rows_lds = r;
                c->nl_threads_div = dv;
                rc = nl_setup(c);
                if (rc || !c->nl_active) continue;
                float ms = 0;
                rc = time_force(c, a, b, &ms);
                if (rc) break;
                unsigned f = 0;
                if (hipMemcpy(&f, c->nl.flags, 4, hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = set_err(c, COMD_ERR_HIP, "autotune flag read failed");
                    break;
                }
                if (f & CM_NL_INVALID) continue;
                if (ms < best) {
                    best = ms;
                    bg = g;
                    br = r;
                    bd = dv;
                }
            }
    hipEventDestroy(a);
    hipEventDestroy(b);
    c->nl_group_opt = bg;
    c->nl_rows_lds = br;
    c->nl_threads_div = bd;
    const int rc2 = nl_setup(c);
    if (rc) return rc;
    if (rc2) return rc2;
    op_force(c);  // the lists built again with the choice, and the forces of the state
    HIPCHK(c, hipGetLastError());
    if (group_opt) *group_opt = bg;
    if (rows_lds) *rows_lds = br;
    if (threads_div) *threads_div = bd;
    if (best_ms) *best_ms = best < 1e299 ? best : -1.0;
    return COMD_OK;
}

int comd_autotune_lists(comd_ctx *c, int *group_opt, int *rows_lds, int *threads_div, double *best_ms)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    if (!c->nl_active) return set_err(c, COMD_ERR_STATE, "no neighbour lists in this context");
    const int timing_saved = c->timing;
    c->timing = 0;
    const int rc = autotune_lists(c, group_opt, rows_lds, threads_div, best_ms);
    c->timing = timing_saved;
    return rc;
}

int comd_autotune(comd_ctx *c, int *waves, int *blocks_per_cu, int *balanced, double *best_ms)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    if (c->nl_active) {
        // the list passes are what runs: their shape is tuned
        const int rc = comd_autotune_lists(c, nullptr, nullptr, nullptr, best_ms);
        if (rc) return rc;
        if (waves) *waves = c->force_waves;
        if (blocks_per_cu) *blocks_per_cu = c->force_blocks_per_cu;
        if (balanced) *balanced = c->reps[0].pot.balanced;
        return COMD_OK;
    }
    const int timing_saved = c->timing;
    const int nl_saved = c->nl_active;
    c->timing = 0;
    c->nl_active = 0;  // the shapes tuned are those of the link-cell kernels
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    const int wave_opts[] = {1, 2, 4};
    const int bpc_opts[] = {4, 8, 12, 16};
    const int nw = c->nrep > 1 ? 1 : 3;
    double best = 1e300;
    int bw = c->force_waves, bb = c->force_blocks_per_cu, bd = c->reps[0].pot.balanced;
    int rc = COMD_OK;
    for (int dr = 0; dr < 2 && !rc; dr++) {
        for (Replica &R : c->reps) R.pot.balanced = dr;
        rc = upload_descriptors(c);
        for (int iw = 0; iw < nw && !rc; iw++)
            for (int ib = 0; ib < 4 && !rc; ib++) {
                c->force_waves = wave_opts[iw];
                c->force_blocks_per_cu = bpc_opts[ib];
                op_force(c);  // warm-up
                float t = 0, tmin = 1e30f;
                for (int k = 0; k < 3; k++) {
                    hipEventRecord(a, c->stream);
                    op_force(c);
                    hipEventRecord(b, c->stream);
                    if (hipEventSynchronize(b) != hipSuccess) {
                        rc = set_err(c, COMD_ERR_HIP, "autotune timing failed");
                        break;
                    }
                    hipEventElapsedTime(&t, a, b);
                    tmin = std::min(tmin, t);
                }
                if (tmin < best) {
                    best = tmin;
                    bw = c->force_waves;
                    bb = c->force_blocks_per_cu;
                    bd = dr;
                }
            }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    c->force_waves = bw;
    c->force_blocks_per_cu = bb;
    for (Replica &R : c->reps) R.pot.balanced = bd;
    if (!rc) rc = upload_descriptors(c);
    c->timing = timing_saved;
    c->nl_active = nl_saved;
    if (rc) return rc;
    if (waves) *waves = bw;
    if (blocks_per_cu) *blocks_per_cu = bb;
    if (balanced) *balanced = bd;
    if (best_ms) *best_ms = best;
    return COMD_OK;
}

int comd_list_info(comd_ctx *c, int *active, int *valid, int *builds, int *kmax)
{
    if (!c) return COMD_ERR_ARG;
    int v = 0, nb = 0;
    if (c->nl_active) {
        unsigned f = 0;
        int cnt[2] = {0, 0};
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(&f, c->nl.flags, 4, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(cnt, c->nl.count, 8, hipMemcpyDeviceToHost));
        v = (f & CM_NL_INVALID) ? 0 : 1;
        nb = cnt[1];
    }
    if (active) *active = c->nl_active;
    if (valid) *valid = v;
    if (builds) *builds = nb;
    if (kmax) *kmax = c->nl_active ? c->nl.kmax : 0;
    return COMD_OK;
}

int comd_list_occupancy(comd_ctx *c, int *blocks_per_cu, int *lds_bytes, int *threads)
{
    if (!c || !blocks_per_cu || !lds_bytes || !threads) return COMD_ERR_ARG;
    for (int k = 0; k < 2; k++) blocks_per_cu[k] = lds_bytes[k] = 0;
    *threads = 0;
    if (!c->nl_active) return COMD_OK;
    const PotDev &P = c->reps[0].pot;
    const int mixed = c->precision == COMD_PREC_MIXED;
    const int kinds[2] = {P.kind == COMD_POT_LJ ? CM_KIND_LJ : CM_KIND_EAM1, CM_KIND_EAM3};
    for (int k = 0; k < (P.kind == COMD_POT_LJ ? 1 : 2); k++)
        blocks_per_cu[k] = cm_nl_occupancy(P, c->nl, kinds[k], mixed, &lds_bytes[k]);
    *threads = c->nl.pass_threads;
    return COMD_OK;
}

int comd_list_lanes(comd_ctx *c, int *lanes)
{
    if (!c || !lanes) return COMD_ERR_ARG;
    *lanes = c->nl_active ? c->nl.lanes : 0;
    return COMD_OK;
}

int comd_download_aux(comd_ctx *c, int *nnb, double *vir)
{
    if (!c) return COMD_ERR_ARG;
    const size_t S = (size_t)c->nslots;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (nnb) HIPCHK(c, hipMemcpy(nnb, c->nnb, S * 4, hipMemcpyDeviceToHost));
    if (vir) HIPCHK(c, hipMemcpy(vir, c->vir, S * 8, hipMemcpyDeviceToHost));
    return COMD_OK;
}

int comd_pair_count(comd_ctx *c, long long *half_pairs)
{
    if (!c || !half_pairs) return COMD_ERR_ARG;
    const size_t S = (size_t)c->nslots, C = (size_t)c->ncells;
    std::vector<int> nn(S), na(C);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(nn.data(), c->nnb, S * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(na.data(), c->natoms, C * 4, hipMemcpyDeviceToHost));
    long long tot = 0;
    for (size_t b = 0; b < C; b++)
        for (int i = 0; i < na[b]; i++) tot += nn[b * CM_SLOTS + i];
    *half_pairs = tot / 2;
    return COMD_OK;
}

int comd_sync(comd_ctx *c)
{
    if (!c) return COMD_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_flags(c);
}

int comd_timing(comd_ctx *c, int on)
{
    if (!c) return COMD_ERR_ARG;
    hipStreamSynchronize(c->stream);
    c->events.clear();
    c->pool_used = 0;
    for (int k = 0; k < COMD_K_NCLASS; k++) {
        c->ms[k] = 0;
        c->launches[k] = 0;
    }
    c->timing = on ? 1 : 0;
    return COMD_OK;
}

int comd_timing_read(comd_ctx *c, double *ms, int *launches)
{
    if (!c) return COMD_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &e : c->events) {
        float t = 0;
        hipEventElapsedTime(&t, e.a, e.b);
        c->ms[e.cls] += t;
        c->launches[e.cls] += 1;
    }
    c->events.clear();
    c->pool_used = 0;
    for (int k = 0; k < COMD_K_NCLASS; k++) {
        if (ms) ms[k] = c->ms[k];
        if (launches) launches[k] = c->launches[k];
    }
    return COMD_OK;
}

// ---- snapshots

// Header of a snapshot file (little-endian ints after the 8-byte magic).
enum { SNAP_VERSION = 2, SNAP_NHDR = 16 };
static const char kSnapMagic[8] = {'C', 'O', 'M', 'D', 'S', 'N', 'A', 'P'};

// One device array to or from a file, through a host buffer.
static bool snap_io(FILE *fp, void *dev, size_t bytes, bool save)
{
    if (bytes == 0) return true;
    std::vector<char> h(bytes);
    if (save) {
        if (hipMemcpy(h.data(), dev, bytes, hipMemcpyDeviceToHost) != hipSuccess) return false;
        return fwrite(h.data(), 1, bytes, fp) == bytes;
    }
    if (fread(h.data(), 1, bytes, fp) != bytes) return false;
    return hipMemcpy(dev, h.data(), bytes, hipMemcpyHostToDevice) == hipSuccess;
}

// The list buffers a snapshot carries, with their byte counts.
static std::vector<std::pair<void *, size_t>> snap_list_buffers(comd_ctx *c)
{
    const CmList &l = c->nl;
    const size_t n = (size_t)l.n;
    return {{l.pos, n * sizeof(CmPos)}, {l.selfh, n * 2},        {l.slot, n * 4},
            {l.order, n * 4},           {l.fq, n * 8},           {l.nbl, n * (size_t)l.kmax * 2},
            {l.ncnt, n * 4},
            {l.aux_n, (size_t)l.naux * 4}, {l.aux_start, (size_t)l.naux * 4}, {l.gstart, (size_t)(l.ngroups + 1) * 4},
            {l.nonempty, (size_t)c->ncells * 4}, {l.count, 16}, {l.flags, 16}};
}

// ---- observables and the energy monitor

// Force at the current positions, then the observables of replica r into
// out_dev (device memory) with part as the block partials.
static int observe_rep(comd_ctx *c, int r, double *part, comd_observables *out_dev)
{
    const Replica &R = c->reps[r];
    if (R.g.nbx0 < 3 || R.g.nbx1 < 3 || R.g.nbx2 < 3)
        return set_err(c, COMD_ERR_GRID, "observables need at least 3 link cells per dimension");
    op_force(c);
    const double pe_scale = R.prm.potential == COMD_POT_LJ ? 4.0 * R.prm.lj_epsilon * R.pot.lj_s6 : 1.0;
    cm_launch_observe(dev_state(c), R.pot, R.g, pe_scale, part, out_dev, c->stream);
    HIPCHK(c, hipGetLastError());
    return COMD_OK;
}

static int observables_rep(comd_ctx *c, int r, comd_observables *out)
{
    if (!out) return set_err(c, COMD_ERR_ARG, "null argument");
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    if (!c->obs_part) {
        if (hipMalloc((void **)&c->obs_part, (size_t)CM_OBS_MAXBLOCKS * CM_OBS_N * 8) != hipSuccess ||
            hipMalloc((void **)&c->obs_out, sizeof(comd_observables)) != hipSuccess)
            return set_err(c, COMD_ERR_NOMEM, "observables buffers");
    }
    int rc = observe_rep(c, r, c->obs_part, c->obs_out);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->obs_out, sizeof *out, hipMemcpyDeviceToHost, c->stream));
    return check_flags(c);  // synchronises
}

extern "C" {

int comd_observe(comd_ctx *c, comd_observables *out)
{
    if (!c) return COMD_ERR_ARG;
    return observables_rep(c, 0, out);
}

int comd_batch_observe(comd_ctx *c, int r, comd_observables *out)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    return rc ? rc : observables_rep(c, r, out);
}

int comd_observe_noalloc(comd_ctx *c, comd_observables *d_out)
{
    if (!c || !d_out) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    double *ptr = scratch_of(c->device);
    if (!ptr) return set_err(c, COMD_ERR_HIP, "no scratch address for device %d", c->device);
    if (cm_observe_blocks(c->reps[0].g) > CM_OBS_MAXBLOCKS)
        return set_err(c, COMD_ERR_ARG, "more than %d reduction blocks", CM_OBS_MAXBLOCKS);
    const int timing = c->timing;  // timing events would be created on demand
    c->timing = 0;
    const int rc = observe_rep(c, 0, ptr, d_out);
    c->timing = timing;
    return rc;
}

static int centro_rep(comd_ctx *c, int r, double *out)
{
    if (!out) return set_err(c, COMD_ERR_ARG, "null argument");
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    const Grid &g = c->reps[r].g;
    const size_t n = (size_t)g.ncells * CM_SLOTS;
    double *dev = nullptr;
    if (hipMalloc((void **)&dev, n * 8) != hipSuccess) return set_err(c, COMD_ERR_NOMEM, "centrosymmetry buffer");
    int rc = COMD_OK;
    if (hipMemsetAsync(dev, 0, n * 8, c->stream) != hipSuccess) rc = set_err(c, COMD_ERR_HIP, "memset failed");
    if (!rc) {
        cm_launch_centro(dev_state(c), g, dev, c->stream);
        if (hipGetLastError() != hipSuccess) rc = set_err(c, COMD_ERR_HIP, "centrosymmetry launch failed");
    }
    if (!rc && (hipMemcpyAsync(out, dev, n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = set_err(c, COMD_ERR_HIP, "centrosymmetry copy failed");
    hipFree(dev);
    return rc;
}

int comd_centrosymmetry(comd_ctx *c, double *out)
{
    if (!c) return COMD_ERR_ARG;
    return centro_rep(c, 0, out);
}

int comd_batch_centrosymmetry(comd_ctx *c, int r, double *out)
{
    if (!c) return COMD_ERR_ARG;
    int rc = check_rep(c, r);
    return rc ? rc : centro_rep(c, r, out);
}

int comd_snapshot_save(comd_ctx *c, const char *path)
{
    if (!c || !path) return COMD_ERR_ARG;
    if (c->nrep != 1) return set_err(c, COMD_ERR_ARG, "snapshots hold single simulations");
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->nl_active && c->nl_stale) {  // list positions of the last drift
        cm_nl_sync(dev_state(c), c->nl, c->stream);
        c->nl_stale = 0;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_err(c, COMD_ERR_ARG, "cannot write %s", path);
    const Replica &R = c->reps[0];
    const CmList &l = c->nl;
    int hdr[SNAP_NHDR] = {SNAP_VERSION, c->ncells, c->precision, R.pot.eval, c->nl_active, c->nl_skin_milli,
                          l.n, l.kmax, 0, l.naux, l.ngroups, l.hcap, R.ntot, (int)R.nl_ids,
                          R.pot.lj_fsign < 0.0, c->nl_active ? l.lanes : 0};
    const size_t S = (size_t)c->nslots, C = (size_t)c->ncells;
    bool ok = fwrite(kSnapMagic, 1, 8, fp) == 8 && fwrite(hdr, sizeof(int), SNAP_NHDR, fp) == SNAP_NHDR;
    void *arr[] = {c->natoms, c->id, c->itype, c->rx, c->ry, c->rz, c->px, c->py, c->pz, c->fx, c->fy, c->fz,
                   c->fe, c->rho, c->fi, c->vir, c->nnb, c->nlr};
    const size_t sz[] = {C * 4, S * 4, S * 4, S * 8, S * 8, S * 8, S * 8, S * 8, S * 8, S * 8, S * 8, S * 8,
                         S * 8, S * 8, S * 8, S * 8, S * 4, S * 16};
    for (size_t k = 0; ok && k < sizeof sz / sizeof sz[0]; k++) ok = snap_io(fp, arr[k], sz[k], true);
    if (ok && c->nl_active)
        for (auto &b : snap_list_buffers(c)) ok = ok && snap_io(fp, b.first, b.second, true);
    ok = (fclose(fp) == 0) && ok;
    return ok ? COMD_OK : set_err(c, COMD_ERR_ARG, "write failed for %s", path);
}

int comd_snapshot_load(comd_ctx *c, const char *path)
{
    if (!c || !path) return COMD_ERR_ARG;
    if (c->nrep != 1) return set_err(c, COMD_ERR_ARG, "snapshots hold single simulations");
    FILE *fp = fopen(path, "rb");
    if (!fp) return set_err(c, COMD_ERR_ARG, "cannot open %s", path);
    char magic[8];
    int hdr[SNAP_NHDR];
    if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, kSnapMagic, 8) != 0 ||
        fread(hdr, sizeof(int), SNAP_NHDR, fp) != SNAP_NHDR || hdr[0] != SNAP_VERSION) {
        fclose(fp);
        return set_err(c, COMD_ERR_ARG, "%s is not a snapshot", path);
    }
    if (hdr[1] != c->ncells) {
        fclose(fp);
        return set_err(c, COMD_ERR_ARG, "snapshot of %d cells, context of %d", hdr[1], c->ncells);
    }
    // the settings that shape the state: precision, evaluation, skin, LJ
    // sign, and the list passes' lanes per atom (the order of their sums;
    // 0 in files that predate them: the context's own choice)
    int rc = comd_set_precision(c, hdr[2]);
    if (!rc) rc = comd_set_eval(c, hdr[3]);
    if (!rc && c->nl_skin_milli != hdr[5]) rc = comd_tune(c, COMD_TUNE_SKIN_MILLI, hdr[5]);
    if (!rc && (hdr[15] == 1 || hdr[15] == 2 || hdr[15] == 4)) c->nl_lanes = hdr[15];
    if (!rc && c->reps[0].prm.potential == COMD_POT_LJ) rc = comd_tune(c, COMD_TUNE_LJ_PHYSICAL, hdr[14]);
    if (rc) {
        fclose(fp);
        return rc;
    }
    const size_t S = (size_t)c->nslots, C = (size_t)c->ncells;
    // the slot state through the upload path (checks, counts, list setup) ...
    std::vector<int> natoms(C), id(S), itype(S);
    std::vector<double> r(3 * S), p(3 * S), col(S);
    bool ok = fread(natoms.data(), 4, C, fp) == C && fread(id.data(), 4, S, fp) == S && fread(itype.data(), 4, S, fp) == S;
    for (int k = 0; ok && k < 6; k++) {
        ok = fread(col.data(), 8, S, fp) == S;
        double *dst = k < 3 ? r.data() : p.data();
        for (size_t q = 0; ok && q < S; q++) dst[3 * q + (k % 3)] = col[q];
    }
    if (!ok) {
        fclose(fp);
        return set_err(c, COMD_ERR_ARG, "%s is truncated", path);
    }
    rc = upload_rep(c, 0, natoms.data(), id.data(), itype.data(), r.data(), p.data());
    if (rc) {
        fclose(fp);
        return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // ... then the forces, the list words and the lists as they were
    const CmList &l = c->nl;
    if (hdr[4] != c->nl_active ||
        (c->nl_active && (hdr[6] != l.n || hdr[7] != l.kmax || hdr[9] != l.naux ||
                          hdr[10] != l.ngroups || hdr[11] != l.hcap))) {
        fclose(fp);
        return set_err(c, COMD_ERR_STATE, "the snapshot's neighbour lists do not fit this context");
    }
    void *arr[] = {c->fx, c->fy, c->fz, c->fe, c->rho, c->fi, c->vir, c->nnb, c->nlr};
    const size_t sz[] = {S * 8, S * 8, S * 8, S * 8, S * 8, S * 8, S * 8, S * 4, S * 16};
    for (size_t k = 0; ok && k < sizeof sz / sizeof sz[0]; k++) ok = snap_io(fp, arr[k], sz[k], false);
    if (ok && c->nl_active)
        for (auto &b : snap_list_buffers(c)) ok = ok && snap_io(fp, b.first, b.second, false);
    fclose(fp);
    if (!ok) return set_err(c, COMD_ERR_ARG, "%s is truncated", path);
    c->nl_stale = 0;
    c->nl_invalid_seen = 0;
    invalidate_graph(c);
    return check_flags(c);
}

int comd_run_monitored(comd_ctx *c, int nblocks, int block_steps, double dt_seconds, double *pe_ke)
{
    if (!c || nblocks < 0 || block_steps < 0 || !pe_ke) return COMD_ERR_ARG;
    if (c->nrep != 1) return set_err(c, COMD_ERR_ARG, "the monitor runs single simulations");
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    const long need = 3L * (nblocks + 1);
    if (need > c->mon_cap) {
        if (c->mon) hipFree(c->mon);
        c->mon = nullptr;
        c->mon_cap = 0;
        if (hipMalloc((void **)&c->mon, (size_t)need * 8) != hipSuccess) return set_err(c, COMD_ERR_NOMEM, "monitor");
        c->mon_cap = need;
    }
    int rc = clear_flags(c);
    if (rc) return rc;
    const Replica &R = c->reps[0];
    const double dt = dt_seconds * kBohrPerAtuToAPerS;  // timestep.c:126
    for (int b = 0; b <= nblocks; b++) {
        if (b > 0) {
            rc = run_steps(c, block_steps, dt);
            if (rc) return rc;
        }
        op_force(c);  // the trailing force of nTimeSteps (timestep.c:136)
        cm_launch_reduce(dev_state(c), R.g, c->partials, c->nparts, c->stream);
        cm_launch_reduce_final(c->partials, c->nparts, c->mon + 3 * b, c->stream);
    }
    HIPCHK(c, hipGetLastError());
    std::vector<double> m((size_t)need);
    HIPCHK(c, hipMemcpyAsync(m.data(), c->mon, (size_t)need * 8, hipMemcpyDeviceToHost, c->stream));
    rc = check_flags(c);  // synchronises
    if (rc) return rc;
    const double scale = R.prm.potential == COMD_POT_LJ ? 4.0 * R.prm.lj_epsilon * R.pot.lj_s6 : 1.0;
    for (int b = 0; b <= nblocks; b++) {
        pe_ke[2 * b] = m[3 * b] * scale;
        pe_ke[2 * b + 1] = m[3 * b + 1];
    }
    return COMD_OK;
}

}  // extern "C"

// ---- Chebyshev coefficients and their adjoint (single-replica contexts)

// The same value series (3 x n coefficients) for every EAM replica, with the
// derivative series (src-flat/cheby.c:219-270), zero-padded to the stride.
// The staging buffers with room for bytes at least; the stream is idle on
// return (the host may write the pinned buffer).
static int stage_reserve(comd_ctx *c, size_t bytes)
{
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bytes <= c->stage_cap) return COMD_OK;
    if (c->stage_h) hipHostFree(c->stage_h);
    if (c->stage_d) hipFree(c->stage_d);
    c->stage_h = c->stage_d = nullptr;
    c->stage_cap = 0;
    bytes = (bytes + 4095) / 4096 * 4096;
    if (hipHostMalloc((void **)&c->stage_h, bytes, 0) != hipSuccess || hipMalloc((void **)&c->stage_d, bytes) != hipSuccess)
        return set_err(c, COMD_ERR_NOMEM, "staging buffers");
    c->stage_cap = bytes;
    return COMD_OK;
}

// New value series for every replica, with their derivative series and FP32
// copies: formed on the host (once per distinct domain), staged in pinned
// memory and copied to each replica's series by one upload and one kernel,
// without waiting for either.
static int set_cheb_all(comd_ctx *c, const double *coef, int n)
{
    for (Replica &R : c->reps) {
        if (R.prm.potential != COMD_POT_EAM || !R.pot.cheb)
            return set_err(c, COMD_ERR_ARG, "a replica has no Chebyshev series");
        if (n != R.pot.cn) return set_err(c, COMD_ERR_ARG, "expected %d coefficients per series, got %d", R.pot.cn, n);
    }
    const int m = n, m8 = c->reps[0].pot.cstride;
    const size_t per = (size_t)6 * m8 * 12;  // 6 m8 doubles, then 6 m8 floats
    int rc = stage_reserve(c, per * c->nrep);
    if (rc) return rc;
    for (int r = 0; r < c->nrep; r++) {
        const PotDev &P = c->reps[r].pot;
        double *all = (double *)(c->stage_h + per * r);
        float *all32 = (float *)(all + 6 * m8);
        if (r > 0) {
            const PotDev &Q = c->reps[r - 1].pot;
            bool same = true;
            for (int k = 0; k < 3; k++) same = same && P.ca[k] == Q.ca[k] && P.cb[k] == Q.cb[k];
            if (same) {
                memcpy(all, c->stage_h + per * (r - 1), per);
                continue;
            }
        }
        for (int i = 0; i < 6 * m8; i++) all[i] = 0.0;
        for (int k = 0; k < 3; k++) {
            for (int j = 0; j < m; j++) all[(size_t)k * m8 + j] = coef[(size_t)k * m + j];
            cheb_derivative(m, P.ca[k], P.cb[k], coef + (size_t)k * m, all + (size_t)(3 + k) * m8);
        }
        for (int i = 0; i < 6 * m8; i++) all32[i] = (float)all[i];
    }
    HIPCHK(c, hipMemcpyAsync(c->stage_d, c->stage_h, per * c->nrep, hipMemcpyHostToDevice, c->stream));
    cm_launch_scatter_series(c->d_pots, c->nrep, c->stage_d, m8, c->stream);
    HIPCHK(c, hipGetLastError());
    c->cheb_host.assign(coef, coef + 3 * n);
    return COMD_OK;
}

int comd_set_cheb(comd_ctx *c, const double *coef, int n)
{
    if (!c || !coef) return COMD_ERR_ARG;
    return set_cheb_all(c, coef, n);
}

int comd_get_cheb(comd_ctx *c, double *coef6)
{
    if (!c || !coef6) return COMD_ERR_ARG;
    const Replica &R = c->reps[0];
    if (!R.pot.cheb) return set_err(c, COMD_ERR_ARG, "the context has no Chebyshev series");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int m = R.pot.cn, m8 = R.pot.cstride;
    std::vector<double> all((size_t)6 * m8);
    HIPCHK(c, hipMemcpy(all.data(), R.d_cheb, all.size() * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; k++)
        for (int j = 0; j < m; j++) coef6[(size_t)k * m + j] = all[(size_t)k * m8 + j];
    return COMD_OK;
}

// Loss and gradient over every replica of the context (one replica for
// comd_eam_cheb_grad): L = sum_r [ sum_{i in r} |f_i - t_i|^2 / (3 N_r)
// + w_e (E_r - e_target[r])^2 ], the replicas' moments summed by the same
// kernels in one launch per stage.  Energy mode (e_weight non-null): L =
// sum_r e_weight[r] E_r, the same stages with no force term (dL/df = 0)
// and dL/dE_r = e_weight[r]; energies[r] receives E_r.
static int cheb_grad(comd_ctx *c, const double *coef, int n, const double *f_target, const double *e_target,
                     double w_e, double *loss, double *grad, const double *e_weight = nullptr,
                     double *energies = nullptr)
{
    const bool emode = e_weight != nullptr;
    if (!c->have_atoms || c->ntot < 1) return set_err(c, COMD_ERR_STATE, "no atoms");
    if (!emode && !f_target && !c->adj_have_target) return set_err(c, COMD_ERR_STATE, "no force target");
    if (n > cm_adjoint_max_coefs())
        return set_err(c, COMD_ERR_ARG, "the adjoint supports at most %d coefficients per series",
                       cm_adjoint_max_coefs());
    const Replica &R0 = c->reps[0];
    for (const Replica &R : c->reps)
        for (int k = 0; k < 3; k++)
            if (R.pot.ca[k] != R0.pot.ca[k] || R.pot.cb[k] != R0.pot.cb[k] || R.pot.rc2 != R0.pot.rc2)
                return set_err(c, COMD_ERR_ARG, "the replicas' series domains or cutoffs differ");
    int rc = COMD_OK;
    const int m = n;
    const int NM = cm_adjoint_moments();
    const size_t S = (size_t)c->nslots;
    if (!c->adj_out) {
        c->adj_nblocks = (int)std::min<long>((long)c->ncells, (long)c->ncu * 8);
        bool ok = true;
        auto alloc = [&](void **ptr, size_t bytes) {
            *ptr = nullptr;
            if (ok && hipMalloc(ptr, bytes) != hipSuccess) ok = false;
        };
        for (int k = 0; k < 3; k++) alloc((void **)&c->adj_g[k], S * 8);
        alloc((void **)&c->adj_lam, S * 8);
        alloc((void **)&c->adj_mu, S * 8);
        alloc((void **)&c->adj_ft, S * 3 * 8);
        alloc((void **)&c->adj_f2, 64 * 8);
        alloc((void **)&c->adj_part, (size_t)c->adj_nblocks * NM * 8);
        alloc((void **)&c->adj_out, (size_t)5 * NM * 8);
        alloc((void **)&c->adj_rep, (size_t)2 * c->nrep * 8);
        alloc((void **)&c->adj_E, (size_t)3 * c->nrep * 8);
        alloc((void **)&c->adj_epart, (size_t)3 * c->nrep * c->nparts * 8);
        if (!ok) return set_err(c, COMD_ERR_NOMEM, "adjoint buffers");
    }
    if (f_target && !emode) {
        HIPCHK(c, hipMemcpy(c->adj_ft, f_target, S * 3 * 8, hipMemcpyHostToDevice));
        c->adj_have_target = 1;
    }
    // staged after the series (set_cheb_all's upload reads the staging
    // buffer's head): F'' series (the derivative recurrence applied twice,
    // zero-padded), then per replica the energy target or weight and the
    // residual weight 1 / (3 N)
    const int m8 = R0.pot.cstride, nrep = c->nrep;
    const size_t off = (size_t)6 * m8 * 12 * nrep;
    rc = stage_reserve(c, off + ((size_t)m8 + 2 * nrep) * 8);
    if (rc) return rc;
    rc = set_cheb_all(c, coef, n);
    if (rc) return rc;
    {
        double *h = (double *)(c->stage_h + off);
        std::vector<double> d1(m);
        for (int j = 0; j < m8; j++) h[j] = 0.0;
        cheb_derivative(m, R0.pot.ca[2], R0.pot.cb[2], coef + 2 * m, d1.data());
        cheb_derivative(m, R0.pot.ca[2], R0.pot.cb[2], d1.data(), h);
        for (int r = 0; r < nrep; r++) {
            h[m8 + r] = emode ? e_weight[r] : e_target[r];
            h[m8 + nrep + r] = c->reps[r].ntot > 0 ? 1.0 / (3.0 * (double)c->reps[r].ntot) : 0.0;
        }
        HIPCHK(c, hipMemcpyAsync(c->stage_d + off, h, ((size_t)m8 + 2 * nrep) * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->adj_f2, c->stage_d + off, (size_t)m8 * 8, hipMemcpyDeviceToDevice, c->stream));
    }
    const double *d_w = (const double *)(c->stage_d + off) + m8, *d_rw = d_w + nrep;

    // forward: forces and energies through the series (the descriptors
    // switched to the series and back only for replicas that use tables)
    bool tables = false;
    for (const Replica &R : c->reps) tables = tables || R.pot.eval != COMD_EVAL_CHEB;
    std::vector<int> eval_saved(nrep);
    for (int r = 0; r < nrep; r++) {
        eval_saved[r] = c->reps[r].pot.eval;
        c->reps[r].pot.eval = COMD_EVAL_CHEB;
    }
    if (tables) {
        rc = upload_descriptors(c);
        if (rc) return rc;
    }
    op_force(c);
    for (int r = 0; r < nrep; r++) c->reps[r].pot.eval = eval_saved[r];
    if (tables) {
        rc = upload_descriptors(c);
        if (rc) return rc;
    }
    // each replica's energy into adj_E (the sums of comd_energy), then
    // dL/dE on the device
    cm_launch_reduce_batch(dev_state(c), c->d_grids, nrep, c->adj_epart, c->nparts, c->adj_E, c->stream);
    cm_launch_adj_rep(c->adj_E, d_w, d_rw, nrep, emode, w_e, c->adj_rep, c->stream);

    const int mixed = c->precision == COMD_PREC_MIXED;
    CmAdjoint J;
    J.gx = c->adj_g[0];
    J.gy = c->adj_g[1];
    J.gz = c->adj_g[2];
    J.lam = c->adj_lam;
    J.mu = c->adj_mu;
    J.f2 = c->adj_f2;
    J.dlde = c->adj_rep;
    J.rw = c->adj_rep + c->nrep;
    J.part = c->adj_part;
    J.nblocks = c->adj_nblocks;
    J.out = c->adj_out;
    PotDev P = R0.pot;  // series, domains and cutoff common to the replicas
    P.eval = COMD_EVAL_CHEB;
    DevState d = dev_state(c);
    if (emode) {  // no force term: dL/df = 0
        for (int k = 0; k < 3; k++) HIPCHK(c, hipMemsetAsync(c->adj_g[k], 0, S * 8, c->stream));
    } else {
        cm_launch_adj_resid(d, J, c->adj_ft, c->stream);
    }
    // the pair passes over the neighbour lists when the forward pass used
    // them (valid after it, every group an EAM one, the staged halo in LDS)
    // and, by default, there is a group per CU at least (a block per group:
    // fewer groups leave CUs idle, where the link-cell passes, a wave per
    // cell, measured faster); else over the link cells
    bool nl = false;
    if (c->adj_lists && c->nl_active && c->nl_lj_groups == 0 && (c->adj_lists == 2 || c->nl.pass_blocks >= c->ncu)) {
        unsigned f = 0;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(&f, c->nl.flags, 4, hipMemcpyDeviceToHost));
        nl = !(f & CM_NL_INVALID) && cm_adj_nl_lds(c->nl) + (12 << 10) <= (size_t)c->lds_block;
    }
    if (nl) {
        const CmList &l = c->nl;
        const long ng = l.pass_blocks;
        if ((long)l.n * 3 > c->adj_gb_cap) {
            if (c->adj_gb) hipFree(c->adj_gb);
            c->adj_gb = nullptr;
            c->adj_gb_cap = 0;
            if (hipMalloc((void **)&c->adj_gb, (size_t)l.n * 3 * 8) != hipSuccess)
                return set_err(c, COMD_ERR_NOMEM, "adjoint list buffers");
            c->adj_gb_cap = (long)l.n * 3;
        }
        if (2 * ng > c->adj_part_nl_cap) {
            if (c->adj_part_nl) hipFree(c->adj_part_nl);
            c->adj_part_nl = nullptr;
            c->adj_part_nl_cap = 0;
            if (hipMalloc((void **)&c->adj_part_nl, (size_t)2 * ng * NM * 8) != hipSuccess)
                return set_err(c, COMD_ERR_NOMEM, "adjoint list buffers");
            c->adj_part_nl_cap = 2 * ng;
        }
        CmAdjNl A;
        A.gb = c->adj_gb;
        A.gx = J.gx;
        A.gy = J.gy;
        A.gz = J.gz;
        A.mu = J.mu;
        A.lam = J.lam;
        A.dlde = J.dlde;
        A.part = c->adj_part_nl;
        CmAdjoint Jn = J;
        Jn.part = c->adj_part_nl;
        cm_launch_adj_nl_gather(l, J.gx, J.gy, J.gz, c->adj_gb, c->stream);
        cm_launch_adj_nl(d, P, l, A, 1, mixed, c->stream);  // the phi gradient's moments, one set
        cm_launch_adj_final(Jn, (int)ng, 0, c->stream);
        cm_launch_adj_nl(d, P, l, A, 2, mixed, c->stream);
        cm_launch_adj_final(Jn, (int)ng, 1, c->stream);
        cm_launch_adj_stage(d, P, J, 2, mixed, c->stream);  // the atom pass, per slot
        cm_launch_adj_nl(d, P, l, A, 3, mixed, c->stream);
        cm_launch_adj_final(Jn, (int)ng, 3, c->stream);
    } else {
        for (int stage = 0; stage < 4; stage++) cm_launch_adj_stage(d, P, J, stage, mixed, c->stream);
    }
    HIPCHK(c, hipGetLastError());
    std::vector<double> out((size_t)5 * NM), E3((size_t)3 * nrep);
    HIPCHK(c, hipMemcpyAsync(out.data(), c->adj_out, out.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(E3.data(), c->adj_E, E3.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double lossE = 0.0;
    for (int r = 0; r < nrep; r++) {
        const double E = E3[(size_t)3 * r];
        if (energies) energies[r] = E;
        lossE += emode ? e_weight[r] * E : w_e * (E - e_target[r]) * (E - e_target[r]);
    }

    if (loss) *loss = (emode ? 0.0 : out[(size_t)4 * NM]) + lossE;
    if (!grad) return COMD_OK;
    // B_0 = T_0 / 2 in the series convention of src-flat/cheby.c:309; the
    // energy moments (set 1 of stages 0 and 2) carry each replica's dL/dE
    auto mom = [&](int stage, int set, int k) { return out[(size_t)stage * NM + set * 32 + k] * (k == 0 ? 0.5 : 1.0); };
    // columns of the derivative operator: D e_j
    auto dcol = [&](int fn, int j, std::vector<double> &col) {
        std::vector<double> e(m, 0.0);
        e[j] = 1.0;
        col.assign(m, 0.0);
        cheb_derivative(m, R0.pot.ca[fn], R0.pot.cb[fn], e.data(), col.data());
    };
    std::vector<double> col;
    for (int j = 0; j < m; j++) {
        double gphi = mom(0, 1, j), grho = mom(3, 0, j), gF = mom(2, 1, j);
        if (nl) {
            // the list pass applied the derivative operator per pair and
            // added the energy moments: stage 0 holds dL/dc_phi itself
            gphi += mom(0, 0, j);
        } else {
            dcol(0, j, col);
            for (int k = 0; k < m; k++) gphi += col[k] * mom(0, 0, k);
        }
        dcol(1, j, col);
        for (int k = 0; k < m; k++) grho += col[k] * mom(1, 0, k);
        dcol(2, j, col);
        for (int k = 0; k < m; k++) gF += col[k] * mom(2, 0, k);
        grad[j] = gphi;
        grad[m + j] = grho;
        grad[2 * m + j] = gF;
    }
    return COMD_OK;
}

int comd_eam_cheb_grad(comd_ctx *c, const double *coef, int n, const double *f_target, double e_target, double w_e,
                       double *loss, double *grad)
{
    if (!c || !coef) return COMD_ERR_ARG;
    if (c->nrep != 1) return set_err(c, COMD_ERR_ARG, "a batch takes comd_batch_cheb_grad");
    return cheb_grad(c, coef, n, f_target, &e_target, w_e, loss, grad);
}

int comd_batch_cheb_grad(comd_ctx *c, const double *coef, int n, const double *f_target, const double *e_target,
                         double w_e, double *loss, double *grad)
{
    if (!c || !coef || !e_target) return COMD_ERR_ARG;
    return cheb_grad(c, coef, n, f_target, e_target, w_e, loss, grad);
}

int comd_eam_cheb_energy_grad(comd_ctx *c, const double *coef, int n, double *energy, double *grad)
{
    if (!c || !coef) return COMD_ERR_ARG;
    if (c->nrep != 1) return set_err(c, COMD_ERR_ARG, "a batch takes comd_batch_cheb_energy_grad");
    const double one = 1.0;
    return cheb_grad(c, coef, n, nullptr, nullptr, 0.0, nullptr, grad, &one, energy);
}

int comd_batch_cheb_energy_grad(comd_ctx *c, const double *coef, int n, const double *weights, double *energies,
                                double *grad)
{
    if (!c || !coef) return COMD_ERR_ARG;
    std::vector<double> w((size_t)c->nrep, 1.0);
    if (weights) w.assign(weights, weights + c->nrep);
    return cheb_grad(c, coef, n, nullptr, nullptr, 0.0, nullptr, grad, w.data(), energies);
}

}  // extern "C"
