// api.hip -- context management and the C-ABI entry points of
// include/comd_hip.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "comd_internal.h"

// src-flat/include/constants.h:61-65
static const double kBohrPerAtuToAPerS = (double)0.529 / (double)2.418e-17;
static const double kAmuToMe = (double)1822.83;

static char g_create_err[512] = "";

struct TimedEvent {
    int cls;
    hipEvent_t a, b;
};

struct comd_ctx {
    comd_params prm;
    int ncells;
    long nslots;
    int ntot;
    int cur;               // which of the two moving-state buffers is current
    int have_atoms;
    int device;
    hipStream_t stream;
    // moving state, two buffers
    int *natoms[2], *id[2], *itype[2];
    double *rx[2], *ry[2], *rz[2], *px[2], *py[2], *pz[2];
    // outputs
    double *fx, *fy, *fz, *fe, *rho, *fi, *vir;
    int *nnb;
    // scratch
    int *dest, *mvidx, *meta, *counts, *touched, *tc, *arr_cnt, *arr;
    int *mv_slot, *mv_rank, *mv_id, *mv_itype;
    double *mv_rx, *mv_ry, *mv_rz, *mv_px, *mv_py, *mv_pz;
    uint8_t *perm;
    unsigned *flags;
    double *partials;
    int nparts;
    double *red_out;       // 3 doubles, device
    double *host_out;      // pinned 4 doubles
    unsigned *host_flags;  // pinned
    // potential
    PotDev pot;
    double *d_tab[3];
    double *d_cf64[2];
    float *d_cf32[2];
    double *d_cheb;
    float *d_cheb32;
    // Chebyshev-coefficient adjoint (allocated on first use)
    double *adj_g[3], *adj_lam, *adj_mu, *adj_ft, *adj_f2, *adj_part, *adj_out;
    int adj_nblocks;
    std::vector<double> cheb_host;  // current value series phi, rho, F (3 x cn)
    // force-kernel launch shape (comd_tune)
    int ncu, max_lds;
    int force_waves, force_blocks_per_cu, tab_lds;
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
    const int a = c->cur;
    d.nbx0 = c->prm.nbx[0];
    d.nbx1 = c->prm.nbx[1];
    d.nbx2 = c->prm.nbx[2];
    d.ncells = c->ncells;
    d.periodic = c->prm.periodic;
    d.bs0 = c->prm.boxsize[0];
    d.bs1 = c->prm.boxsize[1];
    d.bs2 = c->prm.boxsize[2];
    d.L0 = c->prm.bounds[0];
    d.L1 = c->prm.bounds[1];
    d.L2 = c->prm.bounds[2];
    d.natoms = c->natoms[a];
    d.id = c->id[a];
    d.itype = c->itype[a];
    d.rx = c->rx[a]; d.ry = c->ry[a]; d.rz = c->rz[a];
    d.px = c->px[a]; d.py = c->py[a]; d.pz = c->pz[a];
    d.fx = c->fx; d.fy = c->fy; d.fz = c->fz; d.fe = c->fe;
    d.rho = c->rho; d.fi = c->fi; d.vir = c->vir;
    d.nnb = c->nnb;
    d.dest = c->dest;
    d.mvidx = c->mvidx;
    d.counts = c->counts;
    d.touched = c->touched;
    d.tc = c->tc;
    d.arr_cnt = c->arr_cnt;
    d.arr = c->arr;
    d.mv_slot = c->mv_slot; d.mv_rank = c->mv_rank; d.mv_id = c->mv_id; d.mv_itype = c->mv_itype;
    d.mv_rx = c->mv_rx; d.mv_ry = c->mv_ry; d.mv_rz = c->mv_rz;
    d.mv_px = c->mv_px; d.mv_py = c->mv_py; d.mv_pz = c->mv_pz;
    d.perm = c->perm;
    d.meta = c->meta;
    d.flags = c->flags;
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

static void op_rebox(comd_ctx *c)
{
    hipEvent_t ev = nullptr;
    tbegin(c, COMD_K_REBIN, &ev);
    DevState d = dev_state(c);
    cm_launch_rebox(d, c->stream);
    tend(c, COMD_K_REBIN, ev);
}

// rebin: record movers for the op_rebox that follows (0 for a lone kick)
static void op_move(comd_ctx *c, double dtk, double dtd, int kick, int drift, int rebin = 1)
{
    hipEvent_t ev = nullptr;
    tbegin(c, COMD_K_REBIN, &ev);
    DevState d = dev_state(c);
    cm_launch_move(d, c->stream, dtk, dtd, kAmuToMe * c->prm.mass, kick, drift, rebin);
    tend(c, COMD_K_REBIN, ev);
}

static ForceCfg force_cfg(const comd_ctx *c, int kind, int mixed)
{
    ForceCfg f;
    f.waves = c->force_waves;
    f.tab_lds = c->tab_lds;
    f.max_lds = c->max_lds;
    f.ncu = c->ncu;
    // fewer waves per block when the per-wave LDS does not fit
    while (f.waves > 1 && cm_force_lds_bytes(c->pot, kind, mixed, f.waves, 0) > f.max_lds) f.waves--;
    long b = (long)c->ncu * c->force_blocks_per_cu;
    const long need = ((long)c->ncells + f.waves - 1) / f.waves;
    f.blocks = (int)(b < need ? b : need);
    if (f.blocks < 1) f.blocks = 1;
    return f;
}

static void op_force(comd_ctx *c)
{
    DevState d = dev_state(c);
    const int mixed = c->prm.precision == COMD_PREC_MIXED;
    c->pot.eval = c->prm.eval;
    if (c->prm.potential == COMD_POT_LJ) {
        hipEvent_t ev = nullptr;
        tbegin(c, COMD_K_DENSITY, &ev);
        cm_launch_force(d, c->pot, CM_KIND_LJ, mixed, force_cfg(c, CM_KIND_LJ, mixed), c->stream);
        tend(c, COMD_K_DENSITY, ev);
    } else {
        hipEvent_t ev = nullptr;
        tbegin(c, COMD_K_DENSITY, &ev);
        cm_launch_force(d, c->pot, CM_KIND_EAM1, mixed, force_cfg(c, CM_KIND_EAM1, mixed), c->stream);
        tend(c, COMD_K_DENSITY, ev);
        tbegin(c, COMD_K_EMBED_FORCE, &ev);
        cm_launch_force(d, c->pot, CM_KIND_EAM3, mixed, force_cfg(c, CM_KIND_EAM3, mixed), c->stream);
        tend(c, COMD_K_EMBED_FORCE, ev);
    }
}

static int check_flags(comd_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->host_flags, c->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const unsigned f = *c->host_flags;
    if (f & CM_FLAG_OVERFLOW) return set_err(c, COMD_ERR_OVERFLOW, "a link cell reached %d atoms", CM_SLOTS);
    if (f & CM_FLAG_OUTSIDE) return set_err(c, COMD_ERR_OUTSIDE, "an atom fell outside the cell grid");
    return COMD_OK;
}

// ---- C-ABI

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

static void free_all(comd_ctx *c)
{
    void *ptrs[] = {c->fx, c->fy, c->fz, c->fe, c->rho, c->fi, c->vir, c->nnb, c->dest, c->mvidx, c->meta,
                    c->counts, c->touched, c->tc, c->arr_cnt, c->arr, c->mv_slot, c->mv_rank, c->mv_id,
                    c->mv_itype, c->mv_rx, c->mv_ry, c->mv_rz, c->mv_px, c->mv_py, c->mv_pz,
                    c->perm, c->flags, c->partials, c->red_out,
                    c->d_tab[0], c->d_tab[1], c->d_tab[2], c->d_cf64[0], c->d_cf64[1], c->d_cf32[0], c->d_cf32[1],
                    c->d_cheb, c->d_cheb32, c->adj_g[0], c->adj_g[1], c->adj_g[2], c->adj_lam, c->adj_mu,
                    c->adj_ft, c->adj_f2, c->adj_part, c->adj_out};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (int b = 0; b < 2; b++) {
        void *q[] = {c->natoms[b], c->id[b], c->itype[b], c->rx[b], c->ry[b], c->rz[b], c->px[b], c->py[b], c->pz[b]};
        for (void *p : q)
            if (p) hipFree(p);
    }
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

int comd_create(const comd_params *p, comd_ctx **out)
{
    if (!p || !out) return set_err(nullptr, COMD_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->potential != COMD_POT_EAM && p->potential != COMD_POT_LJ)
        return set_err(nullptr, COMD_ERR_ARG, "unknown potential %d", p->potential);
    for (int j = 0; j < 3; j++)
        if (p->nbx[j] < 1 || !(p->boxsize[j] > 0) || !(p->bounds[j] > 0))
            return set_err(nullptr, COMD_ERR_GRID, "bad grid in dimension %d", j);
    if (p->potential == COMD_POT_EAM) {
        for (int k = 0; k < 3; k++)
            if (p->tab_n[k] < 2 || !p->tab_values[k]) return set_err(nullptr, COMD_ERR_ARG, "EAM table %d missing", k);
        if (p->eval == COMD_EVAL_CHEB && (p->cheb_n < 2 || p->cheb_n > 64 || !p->cheb_coef))
            return set_err(nullptr, COMD_ERR_ARG, "Chebyshev coefficients missing");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(nullptr, COMD_ERR_NODEVICE, "no HIP device");

    comd_ctx *c = new comd_ctx();
    c->prm = *p;
    for (int k = 0; k < 3; k++) c->prm.tab_values[k] = nullptr;
    c->prm.cheb_coef = nullptr;
    const long ncl = (long)p->nbx[0] * p->nbx[1] * p->nbx[2];
    if (ncl > (1L << 26)) {
        delete c;
        return set_err(nullptr, COMD_ERR_ARG, "too many cells");
    }
    c->ncells = (int)ncl;
    c->nslots = ncl * CM_SLOTS;
    hipGetDevice(&c->device);
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
    for (int b = 0; b < 1; b++) {
        alloc((void **)&c->natoms[b], C * 4);
        alloc((void **)&c->id[b], S * 4);
        alloc((void **)&c->itype[b], S * 4);
        alloc((void **)&c->rx[b], S * 8);
        alloc((void **)&c->ry[b], S * 8);
        alloc((void **)&c->rz[b], S * 8);
        alloc((void **)&c->px[b], S * 8);
        alloc((void **)&c->py[b], S * 8);
        alloc((void **)&c->pz[b], S * 8);
    }
    alloc((void **)&c->fx, S * 8);
    alloc((void **)&c->fy, S * 8);
    alloc((void **)&c->fz, S * 8);
    alloc((void **)&c->fe, S * 8);
    alloc((void **)&c->rho, S * 8);
    alloc((void **)&c->fi, S * 8);
    alloc((void **)&c->vir, S * 8);
    alloc((void **)&c->nnb, S * 4);
    alloc((void **)&c->dest, S * 4);
    alloc((void **)&c->mvidx, S * 4);
    alloc((void **)&c->meta, C * 16);
    alloc((void **)&c->counts, 16);
    alloc((void **)&c->touched, C * 4);
    alloc((void **)&c->tc, C * 4);
    alloc((void **)&c->arr_cnt, C * 4);
    alloc((void **)&c->arr, S * 4);
    // one mover record per slot: a pass can never overflow the list
    alloc((void **)&c->mv_slot, S * 4);
    alloc((void **)&c->mv_rank, S * 4);
    alloc((void **)&c->mv_id, S * 4);
    alloc((void **)&c->mv_itype, S * 4);
    alloc((void **)&c->mv_rx, S * 8);
    alloc((void **)&c->mv_ry, S * 8);
    alloc((void **)&c->mv_rz, S * 8);
    alloc((void **)&c->mv_px, S * 8);
    alloc((void **)&c->mv_py, S * 8);
    alloc((void **)&c->mv_pz, S * 8);
    alloc((void **)&c->perm, S);
    alloc((void **)&c->flags, 16);
    c->nparts = 1024;
    alloc((void **)&c->partials, (size_t)c->nparts * 3 * 8);
    alloc((void **)&c->red_out, 4 * 8);
    if (ok && hipHostMalloc((void **)&c->host_out, 4 * 8, 0) != hipSuccess) ok = false;
    if (ok && hipHostMalloc((void **)&c->host_flags, 16, 0) != hipSuccess) ok = false;

    // potential
    PotDev &P = c->pot;
    memset(&P, 0, sizeof P);
    P.kind = p->potential;
    P.eval = p->eval;
    P.rc2 = p->cutoff * p->cutoff;
    P.mass_amu = p->mass;
    if (p->potential == COMD_POT_LJ) {
        const double sg = p->lj_sigma;
        P.lj_s6 = sg * sg * sg * sg * sg * sg;  // src-flat/ljForce.c:143
        P.lj_eps = p->lj_epsilon;
    } else if (ok) {
        for (int k = 0; k < 3; k++) {
            const int n = p->tab_n[k];
            alloc((void **)&c->d_tab[k], (size_t)(n + 3) * 8);
            if (!ok) break;
            hipMemcpyAsync(c->d_tab[k], p->tab_values[k], (size_t)(n + 3) * 8, hipMemcpyHostToDevice, c->stream);
            P.tab[k] = c->d_tab[k];
            P.x0[k] = p->tab_x0[k];
            P.xn[k] = p->tab_xn[k];
            P.invdx[k] = p->tab_invdx[k];
            P.n[k] = n;
        }
        for (int k = 0; ok && k < 2; k++) {
            std::vector<double> c64;
            std::vector<float> c32;
            table_coefs(p->tab_n[k], p->tab_invdx[k], p->tab_values[k], c64, c32);
            alloc((void **)&c->d_cf64[k], c64.size() * 8);
            alloc((void **)&c->d_cf32[k], c32.size() * 4);
            if (!ok) break;
            hipMemcpyAsync(c->d_cf64[k], c64.data(), c64.size() * 8, hipMemcpyHostToDevice, c->stream);
            hipMemcpyAsync(c->d_cf32[k], c32.data(), c32.size() * 4, hipMemcpyHostToDevice, c->stream);
            hipStreamSynchronize(c->stream);
            P.cf64[k] = c->d_cf64[k];
            P.cf32[k] = c->d_cf32[k];
        }
        P.same_grid = p->tab_n[0] == p->tab_n[1] && p->tab_x0[0] == p->tab_x0[1] && p->tab_xn[0] == p->tab_xn[1] &&
                      p->tab_invdx[0] == p->tab_invdx[1];
        if (ok && p->cheb_n >= 2 && p->cheb_coef) {
            const int m = p->cheb_n;
            alloc((void **)&c->d_cheb, (size_t)6 * m * 8);
            alloc((void **)&c->d_cheb32, (size_t)6 * m * 4);
            if (ok) {
                std::vector<float> c32(6 * m);
                for (int i = 0; i < 6 * m; i++) c32[i] = (float)p->cheb_coef[i];
                hipMemcpyAsync(c->d_cheb, p->cheb_coef, (size_t)6 * m * 8, hipMemcpyHostToDevice, c->stream);
                hipMemcpyAsync(c->d_cheb32, c32.data(), (size_t)6 * m * 4, hipMemcpyHostToDevice, c->stream);
                hipStreamSynchronize(c->stream);
                P.cheb = c->d_cheb;
                P.cheb32 = c->d_cheb32;
                P.cn = m;
                c->cheb_host.assign(p->cheb_coef, p->cheb_coef + 3 * m);
                for (int k = 0; k < 3; k++) {
                    P.ca[k] = p->cheb_x0[k];
                    P.cb[k] = p->cheb_xn[k];
                }
            }
        }
    }
    if (!ok) {
        free_all(c);
        delete c;
        return set_err(nullptr, COMD_ERR_NOMEM, "device allocation failed");
    }
    hipDeviceProp_t prop;
    c->ncu = 256;
    c->max_lds = 65536;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
        if (prop.multiProcessorCount > 0) c->ncu = prop.multiProcessorCount;
        if (prop.sharedMemPerBlock > 0 && prop.sharedMemPerBlock < (size_t)c->max_lds) c->max_lds = (int)prop.sharedMemPerBlock;
    }
    // measured on MI355X (2.05M-atom EAM): single-wave blocks, 8 per CU
    c->force_waves = 1;
    c->force_blocks_per_cu = 8;
    c->tab_lds = 1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        free_all(c);
        delete c;
        return set_err(nullptr, COMD_ERR_HIP, "initialisation failed");
    }
    *out = c;
    return COMD_OK;
}

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

int comd_upload(comd_ctx *c, const int *natoms, const int *id, const int *itype, const double *r, const double *p)
{
    if (!c || !natoms || !id || !r) return COMD_ERR_ARG;
    const size_t S = (size_t)c->nslots, C = (size_t)c->ncells;
    int tot = 0;
    for (size_t b = 0; b < C; b++) {
        if (natoms[b] < 0 || natoms[b] > CM_SLOTS) return set_err(c, COMD_ERR_ARG, "cell %zu holds %d atoms", b, natoms[b]);
        tot += natoms[b];
    }
    std::vector<double> buf(S);
    std::vector<int> it(S, 0);
    const int a = c->cur;
    HIPCHK(c, hipMemcpyAsync(c->natoms[a], natoms, C * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->id[a], id, S * 4, hipMemcpyHostToDevice, c->stream));
    if (itype) memcpy(it.data(), itype, S * 4);
    HIPCHK(c, hipMemcpyAsync(c->itype[a], it.data(), S * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double *dst[6] = {c->rx[a], c->ry[a], c->rz[a], c->px[a], c->py[a], c->pz[a]};
    for (int k = 0; k < 6; k++) {
        const double *src = k < 3 ? r : p;
        if (!src) {
            std::fill(buf.begin(), buf.end(), 0.0);
        } else {
            for (size_t s = 0; s < S; s++) buf[s] = src[3 * s + (k % 3)];
        }
        HIPCHK(c, hipMemcpyAsync(dst[k], buf.data(), S * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->ntot = tot;
    c->have_atoms = 1;
    return COMD_OK;
}

int comd_download(comd_ctx *c, int *natoms, int *id, int *itype, double *r, double *p, double *f, double *rho, double *fi)
{
    if (!c) return COMD_ERR_ARG;
    const size_t S = (size_t)c->nslots, C = (size_t)c->ncells;
    const int a = c->cur;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (natoms) HIPCHK(c, hipMemcpy(natoms, c->natoms[a], C * 4, hipMemcpyDeviceToHost));
    if (id) HIPCHK(c, hipMemcpy(id, c->id[a], S * 4, hipMemcpyDeviceToHost));
    if (itype) HIPCHK(c, hipMemcpy(itype, c->itype[a], S * 4, hipMemcpyDeviceToHost));
    std::vector<double> buf(S);
    auto gather = [&](double *dst, int stride, int k, const double *src) -> int {
        if (hipMemcpy(buf.data(), src, S * 8, hipMemcpyDeviceToHost) != hipSuccess) return COMD_ERR_HIP;
        for (size_t s = 0; s < S; s++) dst[stride * s + k] = buf[s];
        return COMD_OK;
    };
    int rc = COMD_OK;
    if (r) {
        rc |= gather(r, 3, 0, c->rx[a]);
        rc |= gather(r, 3, 1, c->ry[a]);
        rc |= gather(r, 3, 2, c->rz[a]);
    }
    if (p) {
        rc |= gather(p, 3, 0, c->px[a]);
        rc |= gather(p, 3, 1, c->py[a]);
        rc |= gather(p, 3, 2, c->pz[a]);
    }
    if (f) {
        rc |= gather(f, 4, 0, c->fx);
        rc |= gather(f, 4, 1, c->fy);
        rc |= gather(f, 4, 2, c->fz);
        rc |= gather(f, 4, 3, c->fe);
    }
    if (rho) HIPCHK(c, hipMemcpy(rho, c->rho, S * 8, hipMemcpyDeviceToHost));
    if (fi) HIPCHK(c, hipMemcpy(fi, c->fi, S * 8, hipMemcpyDeviceToHost));
    if (rc) return set_err(c, COMD_ERR_HIP, "download failed");
    return COMD_OK;
}

int comd_bin_atoms_device(comd_ctx *c, int n, const double *d_xyz, const double *d_pxyz, int itype)
{
    if (!c || n < 0 || (n > 0 && !d_xyz)) return COMD_ERR_ARG;
    if ((long)n > c->nslots) return set_err(c, COMD_ERR_OVERFLOW, "more atoms than slots");
    HIPCHK(c, hipMemsetAsync(c->flags, 0, 4, c->stream));
    DevState d = dev_state(c);
    cm_launch_bin(d, c->stream, n, d_xyz, d_pxyz, itype, c->meta, c->arr);
    int rc = check_flags(c);
    if (rc) return rc;
    c->ntot = n;
    c->have_atoms = 1;
    return COMD_OK;
}

int comd_bin_atoms(comd_ctx *c, int n, const double *xyz, const double *pxyz, int itype)
{
    if (!c || n < 0 || (n > 0 && !xyz)) return COMD_ERR_ARG;
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
    if (!rc) rc = comd_bin_atoms_device(c, n, dx, dp, itype);
    hipStreamSynchronize(c->stream);
    hipFree(dx);
    if (dp) hipFree(dp);
    return rc;
}

int comd_rebox(comd_ctx *c)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    HIPCHK(c, hipMemsetAsync(c->flags, 0, 4, c->stream));
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
    hipEvent_t ev = nullptr;
    tbegin(c, COMD_K_REDUCE, &ev);
    DevState d = dev_state(c);
    cm_launch_reduce(d, c->pot, c->partials, c->nparts, c->stream);
    cm_launch_reduce_final(c->partials, c->nparts, c->red_out, c->stream);
    tend(c, COMD_K_REDUCE, ev);
    HIPCHK(c, hipMemcpyAsync(c->host_out, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double e = c->host_out[0];
    if (c->prm.potential == COMD_POT_LJ) e = e * 4.0 * c->prm.lj_epsilon * c->pot.lj_s6;  // ljForce.c:217
    if (pe) *pe = e;
    if (ke) *ke = c->host_out[1];
    if (stress) *stress = c->host_out[2] / (c->prm.bounds[0] * c->prm.bounds[1] * c->prm.bounds[2]);
    return COMD_OK;
}

int comd_nsteps(comd_ctx *c, int n, double dt_seconds, double *pe_out)
{
    if (!c || n < 0) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    HIPCHK(c, hipMemsetAsync(c->flags, 0, 4, c->stream));
    const double dt = dt_seconds * kBohrPerAtuToAPerS;  // timestep.c:126
    for (int i = 0; i < n; i++) {
        op_move(c, 0.0, dt / 2.0, 0, 1);
        op_rebox(c);
        op_force(c);
        op_move(c, dt, dt / 2.0, 1, 1);
        op_rebox(c);
    }
    op_force(c);  // timestep.c:136
    HIPCHK(c, hipGetLastError());
    int rc = check_flags(c);
    if (rc) return rc;
    if (pe_out) return comd_energy(c, pe_out, nullptr, nullptr);
    return COMD_OK;
}

int comd_steps(comd_ctx *c, int n, double dt_seconds)
{
    if (!c || n < 0) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    HIPCHK(c, hipMemsetAsync(c->flags, 0, 4, c->stream));
    const double dt = dt_seconds * kBohrPerAtuToAPerS;
    for (int i = 0; i < n; i++) {
        op_move(c, 0.0, dt / 2.0, 0, 1);
        op_rebox(c);
        op_force(c);
        op_move(c, dt, dt / 2.0, 1, 1);
        op_rebox(c);
    }
    HIPCHK(c, hipGetLastError());
    return COMD_OK;
}

int comd_advance_positions(comd_ctx *c, double dt_internal)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    HIPCHK(c, hipMemsetAsync(c->flags, 0, 4, c->stream));
    op_move(c, 0.0, dt_internal, 0, 1);
    op_rebox(c);
    return check_flags(c);
}

int comd_advance_velocity(comd_ctx *c, double dt_internal)
{
    if (!c) return COMD_ERR_ARG;
    if (!c->have_atoms) return set_err(c, COMD_ERR_STATE, "no atoms");
    // kick only: k_move with drift off leaves r untouched; its destination
    // scratch is rewritten before the next rebin
    op_move(c, dt_internal, 0.0, 1, 0, 0);
    HIPCHK(c, hipGetLastError());
    return COMD_OK;
}

int comd_set_precision(comd_ctx *c, int precision)
{
    if (!c || (precision != COMD_PREC_FP64 && precision != COMD_PREC_MIXED)) return COMD_ERR_ARG;
    c->prm.precision = precision;
    return COMD_OK;
}

int comd_set_eval(comd_ctx *c, int eval)
{
    if (!c || (eval != COMD_EVAL_TABLE && eval != COMD_EVAL_CHEB)) return COMD_ERR_ARG;
    if (eval == COMD_EVAL_CHEB && !c->pot.cheb) return set_err(c, COMD_ERR_ARG, "no Chebyshev coefficients");
    c->prm.eval = eval;
    return COMD_OK;
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
        return COMD_OK;
    case COMD_TUNE_BLOCKS_PER_CU:
        if (value < 1 || value > 64) return set_err(c, COMD_ERR_ARG, "blocks per CU must be 1..64");
        c->force_blocks_per_cu = value;
        return COMD_OK;
    case COMD_TUNE_TABLES_IN_LDS:
        c->tab_lds = value ? 1 : 0;
        return COMD_OK;
    case COMD_TUNE_BALANCED_DRAIN:
        c->pot.balanced = value ? 1 : 0;
        return COMD_OK;
    default:
        return set_err(c, COMD_ERR_ARG, "unknown tuning key %d", key);
    }
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
    HIPCHK(c, hipMemcpy(na.data(), c->natoms[c->cur], C * 4, hipMemcpyDeviceToHost));
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


// ---- Chebyshev coefficients and their adjoint

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

int comd_set_cheb(comd_ctx *c, const double *coef, int n)
{
    if (!c || !coef) return COMD_ERR_ARG;
    if (c->prm.potential != COMD_POT_EAM || !c->pot.cheb)
        return set_err(c, COMD_ERR_ARG, "the context has no Chebyshev series");
    const int m = c->pot.cn;
    if (n != m) return set_err(c, COMD_ERR_ARG, "expected %d coefficients per series, got %d", m, n);
    std::vector<double> all((size_t)6 * m);
    std::vector<float> all32((size_t)6 * m);
    for (int k = 0; k < 3; k++) {
        for (int j = 0; j < m; j++) all[(size_t)k * m + j] = coef[(size_t)k * m + j];
        cheb_derivative(m, c->pot.ca[k], c->pot.cb[k], coef + (size_t)k * m, all.data() + (size_t)(3 + k) * m);
    }
    for (int i = 0; i < 6 * m; i++) all32[i] = (float)all[i];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_cheb, all.data(), all.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_cheb32, all32.data(), all32.size() * 4, hipMemcpyHostToDevice));
    c->cheb_host.assign(coef, coef + 3 * m);
    return COMD_OK;
}

int comd_get_cheb(comd_ctx *c, double *coef6)
{
    if (!c || !coef6) return COMD_ERR_ARG;
    if (!c->pot.cheb) return set_err(c, COMD_ERR_ARG, "the context has no Chebyshev series");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(coef6, c->d_cheb, (size_t)6 * c->pot.cn * 8, hipMemcpyDeviceToHost));
    return COMD_OK;
}

int comd_eam_cheb_grad(comd_ctx *c, const double *coef, int n, const double *f_target, double e_target, double w_e,
                       double *loss, double *grad)
{
    if (!c || !coef || !f_target) return COMD_ERR_ARG;
    if (!c->have_atoms || c->ntot < 1) return set_err(c, COMD_ERR_STATE, "no atoms");
    if (n > cm_adjoint_max_coefs())
        return set_err(c, COMD_ERR_ARG, "the adjoint supports at most %d coefficients per series",
                       cm_adjoint_max_coefs());
    int rc = comd_set_cheb(c, coef, n);
    if (rc) return rc;
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
        if (!ok) return set_err(c, COMD_ERR_NOMEM, "adjoint buffers");
    }
    // F'' series: the derivative recurrence applied twice
    std::vector<double> d1(m), d2(m);
    cheb_derivative(m, c->pot.ca[2], c->pot.cb[2], coef + 2 * m, d1.data());
    cheb_derivative(m, c->pot.ca[2], c->pot.cb[2], d1.data(), d2.data());
    HIPCHK(c, hipMemcpy(c->adj_f2, d2.data(), (size_t)m * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->adj_ft, f_target, S * 3 * 8, hipMemcpyHostToDevice));

    // forward: forces and energy through the series
    const int eval_saved = c->prm.eval;
    c->prm.eval = COMD_EVAL_CHEB;
    op_force(c);
    c->prm.eval = eval_saved;
    double E = 0.0;
    rc = comd_energy(c, &E, nullptr, nullptr);
    if (rc) return rc;

    const int mixed = c->prm.precision == COMD_PREC_MIXED;
    const double N = (double)c->ntot;
    CmAdjoint J;
    J.gx = c->adj_g[0];
    J.gy = c->adj_g[1];
    J.gz = c->adj_g[2];
    J.lam = c->adj_lam;
    J.mu = c->adj_mu;
    J.f2 = c->adj_f2;
    J.dlde = 2.0 * w_e * (E - e_target);
    J.part = c->adj_part;
    J.nblocks = c->adj_nblocks;
    J.out = c->adj_out;
    PotDev P = c->pot;
    P.eval = COMD_EVAL_CHEB;
    DevState d = dev_state(c);
    cm_launch_adj_resid(d, J, c->adj_ft, 2.0 / (3.0 * N), c->stream);
    for (int stage = 0; stage < 4; stage++) cm_launch_adj_stage(d, P, J, stage, mixed, c->stream);
    HIPCHK(c, hipGetLastError());
    std::vector<double> out((size_t)5 * NM);
    HIPCHK(c, hipMemcpyAsync(out.data(), c->adj_out, out.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));

    const double L = out[(size_t)4 * NM] / (3.0 * N) + w_e * (E - e_target) * (E - e_target);
    if (loss) *loss = L;
    if (!grad) return COMD_OK;
    // B_0 = T_0 / 2 in the series convention of src-flat/cheby.c:309
    auto mom = [&](int stage, int set, int k) { return out[(size_t)stage * NM + set * 32 + k] * (k == 0 ? 0.5 : 1.0); };
    // columns of the derivative operator: D e_j
    auto dcol = [&](int fn, int j, std::vector<double> &col) {
        std::vector<double> e(m, 0.0);
        e[j] = 1.0;
        col.assign(m, 0.0);
        cheb_derivative(m, c->pot.ca[fn], c->pot.cb[fn], e.data(), col.data());
    };
    std::vector<double> col;
    for (int j = 0; j < m; j++) {
        double gphi = J.dlde * mom(0, 1, j), grho = mom(3, 0, j), gF = J.dlde * mom(2, 1, j);
        dcol(0, j, col);
        for (int k = 0; k < m; k++) gphi += col[k] * mom(0, 0, k);
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

}  // extern "C"
