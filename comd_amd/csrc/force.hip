// force.hip -- link-cell force kernels for gfx950: EAM density + pair force
// + embedding (passes 1 and 2 of src-flat/eam.c:438-660, fused), EAM
// embedding force (pass 3), and Lennard-Jones (src-flat/ljForce.c:89-224).
//
// Compiled with floating-point contraction: the pair sums may differ from
// the reference in the last bits (the parity bar is 1e-10 relative), unlike
// the integrator and the binning, which round exactly as the reference does.
//
// Every kernel is full-shell: each atom sums over all of its neighbours, so
// no atom is written by two waves and nothing needs atomics; the reference's
// half-shell loops (jBox >= iBox, eam.c:494,609) count each pair once and add
// to both atoms, which gives the same sums.
//
// Work mapping.  Each wave owns one home cell at a time (persistent grid,
// cells of a z-slab per XCD so that neighbouring cells are processed under
// the same L2).  The atoms of the 27 stencil cells (the reference's order,
// src-flat/domains.c:419-426, with the duplicates it keeps when a dimension
// has fewer than 3 cells) are staged in the wave's LDS, shifted into the home
// cell's frame.  The lanes are split into S = 64/n_home groups; lane (a, s)
// tests home atom a against halo atoms s, s+S, ... and pushes the accepted
// ones into its own LDS queue.  The queues of one atom are then drained
// round-robin by that atom's S lanes, so the pair arithmetic runs on dense,
// balanced lanes rather than on the ~15% of candidates inside the cutoff.
// The S partial sums of each home atom are added in a fixed order
// (deterministic results).
#include "comd_internal.h"

#define HMAX 288    // halo atoms staged per batch
#define QCAP 24     // queue entries per lane
#define UNROLL 4    // candidates tested per lane between queue checks

// ---------------------------------------------------------------------------
// per-wave LDS: halo x,y,z (HMAX doubles each), halo F' (EAM pass 3), queue
// (QCAP x 64 uint16), per-lane queue counts, stencil cell ids / offsets /
// shifts.  The reduction scratch (7 x 64 values) reuses the halo area.
struct WaveLds {
    double *hx, *hy, *hz, *hf;
    uint16_t *q;
    int *qn, *cell, *off;
    double *drb;
};

__host__ __device__ constexpr int wave_lds_bytes(int kind)
{
    return HMAX * 8 * (kind == CM_KIND_EAM3 ? 4 : 3) + 27 * 3 * 8 + QCAP * 64 * 2 + 64 * 4 + 27 * 4 + 28 * 4;
}

__device__ __forceinline__ WaveLds carve(char *base, int kind)
{
    WaveLds w;
    w.hx = (double *)base;
    w.hy = w.hx + HMAX;
    w.hz = w.hy + HMAX;
    double *p = w.hz + HMAX;
    if (kind == CM_KIND_EAM3) {
        w.hf = p;
        p += HMAX;
    } else {
        w.hf = nullptr;
    }
    w.drb = p;
    w.q = (uint16_t *)(w.drb + 27 * 3);
    w.qn = (int *)(w.q + QCAP * 64);
    w.cell = w.qn + 64;
    w.off = w.cell + 27;
    return w;
}

// Completes this wave's LDS writes before its lanes read them (the waves of
// a block work independently, so no block barrier is used in the cell loop).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// potential evaluation

// 3-point quadratic value and interpolated central-difference derivative,
// the formula of src-flat/eam.c:401-434 in double precision (used once per
// atom for F).  v points at values[-1].
__device__ __forceinline__ void interp64(const double *v, double x0, double xn, double invdx, int n, double r,
                                         double &val, double &der)
{
    r = fmin(fmax(r, x0), xn);
    r = (r - x0) * invdx;
    const double fl = floor(r);
    int i = (int)fl;
    r = r - fl;
    // r <= xn keeps i <= n-1 for a consistent table header; the clamp only
    // keeps a rounding or a bad header inside values[-1..n+1]
    i = min(max(i, 0), n - 1);
    const double *w = v + 1 + i;
    const double vm = w[-1], v0 = w[0], v1 = w[1], v2 = w[2];
    const double g0 = v1 - vm;
    const double g1 = v2 - v0;
    val = v0 + 0.5 * r * (r * (v1 + vm - 2.0 * v0) + g0);
    der = (i <= 0) ? 0.0 : 0.5 * (g0 + r * (g1 - g0)) * invdx;
}

// Interval of the table formula that holds r (clamped to [x0, xn] as in
// src-flat/eam.c:413-414) and the fraction t inside it.
__device__ __forceinline__ int table_interval(double r, double x0, double xn, double invdx, int n, double &t)
{
    r = fmin(fmax(r, x0), xn);
    const double u = (r - x0) * invdx;
    const double fl = floor(u);
    t = u - fl;
    return min(max((int)fl, 0), n - 1);
}

// Table formula from per-interval coefficients (d0 d1 c0 c1 c2 0):
// value c0 + t(c1 + t c2), derivative d0 + t d1.
__device__ __forceinline__ double2 cf_der64(const double *cf, int i) { return ((const double2 *)(cf + 6 * i))[0]; }
__device__ __forceinline__ double cf_val64(const double *cf, int i, double t)
{
    const double2 a = ((const double2 *)(cf + 6 * i))[1];
    const double c2 = cf[6 * i + 4];
    return fma(t, fma(t, c2, a.y), a.x);
}
__device__ __forceinline__ float4 cf_row32(const float *cf, int i) { return ((const float4 *)(cf + 8 * i))[0]; }

// 1/sqrt(x) in double precision: hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// 1/sqrt(x): single-precision estimate and one Newton step in double.
__device__ __forceinline__ double rsqrt_mixed(double x)
{
    const double y = (double)__builtin_amdgcn_rsqf((float)x);
    const double e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// ---------------------------------------------------------------------------
// one accepted pair (home atom a, neighbour j): dx = r_a - r_j

struct Acc {
    double x, y, z, e, rho, vir;
};

template <int KIND, int EVAL, int MIXED>
__device__ __forceinline__ void pair_term(const PotDev &P, double dx, double dy, double dz, double r2, double fia,
                                          double fij, Acc &A)
{
    if (KIND == CM_KIND_LJ) {
        // src-flat/ljForce.c:196-210 (f holds -grad U, energy without 4 eps s6)
        const double s6 = P.lj_s6;
        double r2i;
        if (MIXED) {
            const double y = rsqrt_mixed(r2);
            r2i = y * y;
        } else {
            r2i = 1.0 / r2;
        }
        const double r6 = r2i * r2i * r2i;
        A.e += 0.5 * r6 * (s6 * r6 - 1.0);
        const double fr = 4.0 * P.lj_eps * s6 * r6 * r2i * (12.0 * s6 * r6 - 6.0);
        A.x += dx * fr;
        A.y += dy * fr;
        A.z += dz * fr;
        A.vir += fr * dx * dx;
        return;
    }
    const double y = MIXED ? rsqrt_mixed(r2) : rsqrt64(r2);  // 1/r
    const double r = r2 * y;
    if (KIND == CM_KIND_EAM1) {
        // src-flat/eam.c:521-545 (f holds +grad U)
        double phiv, dphi, rhov;
        if (EVAL == COMD_EVAL_CHEB) {
            const int m = P.cn;
            if (MIXED) {
                const float rf = (float)r;
                const float a0 = (float)P.ca[0], b0 = (float)P.cb[0], a1 = (float)P.ca[1], b1 = (float)P.cb[1];
                phiv = cheb32(P.cheb32 + 0 * m, m, a0, b0, rf);
                dphi = cheb32(P.cheb32 + 3 * m, m, a0, b0, rf);
                rhov = cheb32(P.cheb32 + 1 * m, m, a1, b1, rf);
            } else {
                phiv = cheb64(P.cheb + 0 * m, m, P.ca[0], P.cb[0], r);
                dphi = cheb64(P.cheb + 3 * m, m, P.ca[0], P.cb[0], r);
                rhov = cheb64(P.cheb + 1 * m, m, P.ca[1], P.cb[1], r);
            }
        } else {
            double t0, t1;
            const int i0 = table_interval(r, P.x0[0], P.xn[0], P.invdx[0], P.n[0], t0);
            int i1 = i0;
            t1 = t0;
            if (!P.same_grid) i1 = table_interval(r, P.x0[1], P.xn[1], P.invdx[1], P.n[1], t1);
            if (MIXED) {
                const float4 a = cf_row32(P.cf32[0], i0);
                const float a2 = P.cf32[0][8 * i0 + 4];
                const float4 b = cf_row32(P.cf32[1], i1);
                const float b2 = P.cf32[1][8 * i1 + 4];
                const float f0 = (float)t0, f1 = (float)t1;
                phiv = fmaf(f0, fmaf(f0, a2, a.w), a.z);
                dphi = fmaf(f0, a.y, a.x);
                rhov = fmaf(f1, fmaf(f1, b2, b.w), b.z);
            } else {
                const double2 dd = cf_der64(P.cf64[0], i0);
                phiv = cf_val64(P.cf64[0], i0, t0);
                dphi = fma(t0, dd.y, dd.x);
                rhov = cf_val64(P.cf64[1], i1, t1);
            }
        }
        const double s = dphi * y;
        A.x += s * dx;
        A.y += s * dy;
        A.z += s * dz;
        A.e += 0.5 * phiv;
        A.rho += rhov;
    } else {
        // src-flat/eam.c:633-649
        double drho;
        if (EVAL == COMD_EVAL_CHEB) {
            if (MIXED)
                drho = cheb32(P.cheb32 + 4 * P.cn, P.cn, (float)P.ca[1], (float)P.cb[1], (float)r);
            else
                drho = cheb64(P.cheb + 4 * P.cn, P.cn, P.ca[1], P.cb[1], r);
        } else {
            double t1;
            const int i1 = table_interval(r, P.x0[1], P.xn[1], P.invdx[1], P.n[1], t1);
            if (MIXED) {
                const float2 dd = ((const float2 *)(P.cf32[1] + 8 * i1))[0];
                drho = fmaf((float)t1, dd.y, dd.x);
            } else {
                const double2 dd = cf_der64(P.cf64[1], i1);
                drho = fma(t1, dd.y, dd.x);
            }
        }
        const double g = (fia + fij) * drho * y;
        A.x += g * dx;
        A.y += g * dy;
        A.z += g * dz;
        A.vir += g * dx * dx;
    }
}

// ---------------------------------------------------------------------------

// One home cell: stencil, halo staging, candidate tests, pair terms,
// reduction of the lanes' partial sums and the epilogue.  The 64 lanes of a
// wave call it together.
template <int KIND, int EVAL, int MIXED>
__device__ __forceinline__ void pair_cell(const DevState &d, const Grid &g, const PotDev &P, const WaveLds &w, int c,
                                          int L)
{
    const double rc2 = P.rc2;
    const int nh = d.natoms[c];
    if (nh == 0) return;

    // ---- stencil entries, counts and corner shifts
    const int nent = g.ncells == 1 ? 1 : 27;
    int ne = 0;
    if (L < nent) {
        const int b = stencil_cell(g, c, L);
        w.cell[L] = b;
        if (b >= 0) {
            ne = d.natoms[b];
            double sx, sy, sz;
            corner_shift(g, c, b, sx, sy, sz);
            w.drb[3 * L + 0] = sx;
            w.drb[3 * L + 1] = sy;
            w.drb[3 * L + 2] = sz;
        }
    }
    int incl = ne;
    for (int o = 1; o < 32; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (L >= o) incl += t;
    }
    const int H = __shfl(incl, 31, 64);
    const int nonempty = __popcll(__ballot(ne > 0));
    if (L < nent) w.off[L] = incl - ne;
    if (L == 0) w.off[nent] = H;

    // ---- home atoms: lane (a, s)
    const int S = 64 / nh;
    const bool act = L < S * nh;
    const int a = act ? L % nh : 0;
    const int sg = act ? L / nh : 0;
    const int hs = c * CM_SLOTS + a;
    const double xa = d.rx[hs], ya = d.ry[hs], za = d.rz[hs];
    const double fia = KIND == CM_KIND_EAM3 ? d.fi[hs] : 0.0;
    Acc A = {0, 0, 0, 0, 0, 0};
    int npair = 0;
    wave_sync();

    for (int h0 = 0; h0 < H; h0 += HMAX) {
        const int hb = H - h0 < HMAX ? H - h0 : HMAX;
        // stage the batch: lane L takes halo atoms L, L+64, ...; its
        // stencil entry only moves forward
        {
            int e = 0;
            for (int h = h0 + L; h < h0 + hb; h += 64) {
                while (w.off[e + 1] <= h) e++;
                const int sl = w.cell[e] * CM_SLOTS + (h - w.off[e]);
                const double x = d.rx[sl], y = d.ry[sl], z = d.rz[sl];
                w.hx[h - h0] = x - w.drb[3 * e + 0];
                w.hy[h - h0] = y - w.drb[3 * e + 1];
                w.hz[h - h0] = z - w.drb[3 * e + 2];
                if (KIND == CM_KIND_EAM3) w.hf[h - h0] = d.fi[sl];
            }
        }
        wave_sync();

        int cnt = 0;
        // drain all queues: atom a's entries (the concatenation of the
        // queues of lanes a, a+nh, ...) are taken round-robin by its lanes
        auto drain = [&]() {
            if (!P.balanced) {
                // each lane drains its own queue
                for (int k = 0; __ballot(k < cnt) != 0; k++) {
                    if (k < cnt) {
                        const int j = w.q[k * 64 + L];
                        const double dx = xa - w.hx[j], dy = ya - w.hy[j], dz = za - w.hz[j];
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        pair_term<KIND, EVAL, MIXED>(P, dx, dy, dz, r2, fia,
                                                     KIND == CM_KIND_EAM3 ? w.hf[j] : 0.0, A);
                        npair++;
                    }
                }
                cnt = 0;
                return;
            }
            w.qn[L] = act ? cnt : 0;
            wave_sync();
            int tot = 0;
            if (act)
                for (int s = 0; s < S; s++) tot += w.qn[a + s * nh];
            int v = sg, sp = 0, base = 0;
            while (__ballot(act && v < tot) != 0) {
                if (act && v < tot) {
                    int cs = w.qn[a + sp * nh];
                    while (v >= base + cs) {
                        base += cs;
                        sp++;
                        cs = w.qn[a + sp * nh];
                    }
                    const int j = w.q[(v - base) * 64 + a + sp * nh];
                    const double dx = xa - w.hx[j], dy = ya - w.hy[j], dz = za - w.hz[j];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    pair_term<KIND, EVAL, MIXED>(P, dx, dy, dz, r2, fia, KIND == CM_KIND_EAM3 ? w.hf[j] : 0.0, A);
                    npair++;
                    v += S;
                }
            }
            cnt = 0;
            wave_sync();
        };
        const int trips = (hb + S - 1) / S;
        for (int it0 = 0; it0 < trips; it0 += UNROLL) {
            bool acc[UNROLL];
            int jj[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int j = sg + (it0 + u) * S;
                const bool valid = act && j < hb;
                const int jc = valid ? j : 0;
                const double dx = xa - w.hx[jc], dy = ya - w.hy[jc], dz = za - w.hz[jc];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const bool in = KIND == CM_KIND_EAM3 ? (r2 < rc2) : (r2 <= rc2);
                acc[u] = valid && in && r2 > 0.0;
                jj[u] = j;
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++)
                if (acc[u]) {
                    w.q[cnt * 64 + L] = (uint16_t)jj[u];
                    cnt++;
                }
            if (__ballot(cnt > QCAP - UNROLL) != 0) drain();
        }
        drain();
    }

    // ---- combine the S partial sums of each home atom, in s order
    // (scratch in the halo area, free once the last batch is drained)
    double *red = w.hx;
    int *redn = (int *)(red + 6 * 64);
    red[0 * 64 + L] = act ? A.x : 0.0;
    red[1 * 64 + L] = act ? A.y : 0.0;
    red[2 * 64 + L] = act ? A.z : 0.0;
    red[3 * 64 + L] = act ? A.e : 0.0;
    red[4 * 64 + L] = act ? A.rho : 0.0;
    red[5 * 64 + L] = act ? A.vir : 0.0;
    redn[L] = act ? npair : 0;
    wave_sync();
    if (L < nh) {
        double sx = 0, sy = 0, sz = 0, se = 0, sr = 0, sv = 0;
        int sn = 0;
        for (int s = 0; s < S; s++) {
            const int l = L + s * nh;
            sx += red[0 * 64 + l];
            sy += red[1 * 64 + l];
            sz += red[2 * 64 + l];
            se += red[3 * 64 + l];
            sr += red[4 * 64 + l];
            sv += red[5 * 64 + l];
            sn += redn[l];
        }
        const int o = c * CM_SLOTS + L;
        if (KIND == CM_KIND_EAM1) {
            // pass 2, src-flat/eam.c:576-594
            double fv, fp;
            if (EVAL == COMD_EVAL_CHEB) {
                fv = cheb64(P.cheb + 2 * P.cn, P.cn, P.ca[2], P.cb[2], sr);
                fp = cheb64(P.cheb + 5 * P.cn, P.cn, P.ca[2], P.cb[2], sr);
            } else {
                interp64(P.tab[2], P.x0[2], P.xn[2], P.invdx[2], P.n[2], sr, fv, fp);
            }
            d.fx[o] = sx;
            d.fy[o] = sy;
            d.fz[o] = sz;
            d.fe[o] = se + fv;
            d.rho[o] = sr;
            d.fi[o] = fp;
            d.nnb[o] = sn;
        } else if (KIND == CM_KIND_EAM3) {
            d.fx[o] += sx;
            d.fy[o] += sy;
            d.fz[o] += sz;
            d.vir[o] = sv;
        } else {
            // kinetic term of src-flat/ljForce.c:175, once per non-empty
            // stencil cell, with the mass in amu
            const double px = d.px[o];
            d.fx[o] = sx;
            d.fy[o] = sy;
            d.fz[o] = sz;
            d.fe[o] = se;
            d.vir[o] = sv - (double)nonempty * (px * px / P.mass_amu);
            d.nnb[o] = sn;
        }
    }
    wave_sync();
}

template <int KIND, int EVAL, int MIXED>
__global__ __launch_bounds__(256) void k_pair(DevState d, PotDev P)
{
    extern __shared__ __align__(16) char lds[];
    const int W = blockDim.x >> 6;
    const int wv = threadIdx.x >> 6;
    const int L = threadIdx.x & 63;
    WaveLds w = carve(lds + wv * wave_lds_bytes(KIND), KIND);

    // ---- cells of this wave: a slab of consecutive cells per XCD (blocks are
    // dispatched round-robin over the 8 XCDs)
    const int nxcd = gridDim.x < 8 ? (int)gridDim.x : 8;
    const int xcd = blockIdx.x % nxcd;
    const int blocks_here = ((int)gridDim.x - xcd + nxcd - 1) / nxcd;
    const int workers = blocks_here * W;
    const int worker = (blockIdx.x / nxcd) * W + wv;
    const int c_begin = (int)((long)d.ncells * xcd / nxcd);
    const int c_end = (int)((long)d.ncells * (xcd + 1) / nxcd);
    for (int c = c_begin + worker; c < c_end; c += workers) pair_cell<KIND, EVAL, MIXED>(d, d.g0, P, w, c, L);
}

// Many replicas in one launch: the cells of all replicas are dealt to the
// waves with a global stride (replicas of different cost are spread evenly
// over the waves); each cell uses its replica's grid and potential.  PASS 1
// runs the EAM density pass or Lennard-Jones, PASS 3 the EAM embedding-force
// pass (Lennard-Jones cells have none).
template <int PASS, int MIXED>
__global__ __launch_bounds__(64) void k_pair_batch(DevState d, const PotDev *pots)
{
    extern __shared__ __align__(16) char lds[];
    const int L = threadIdx.x;
    WaveLds w = carve(lds, PASS == 1 ? CM_KIND_EAM1 : CM_KIND_EAM3);
    for (int c = blockIdx.x; c < d.ncells; c += gridDim.x) {
        const Grid g = d.grids[d.cell_grid[c]];
        const PotDev P = pots[g.pot];
        if (PASS == 1) {
            if (P.kind == COMD_POT_LJ) pair_cell<CM_KIND_LJ, COMD_EVAL_TABLE, MIXED>(d, g, P, w, c, L);
            else if (P.eval == COMD_EVAL_CHEB) pair_cell<CM_KIND_EAM1, COMD_EVAL_CHEB, MIXED>(d, g, P, w, c, L);
            else pair_cell<CM_KIND_EAM1, COMD_EVAL_TABLE, MIXED>(d, g, P, w, c, L);
        } else if (P.kind != COMD_POT_LJ) {
            if (P.eval == COMD_EVAL_CHEB) pair_cell<CM_KIND_EAM3, COMD_EVAL_CHEB, MIXED>(d, g, P, w, c, L);
            else pair_cell<CM_KIND_EAM3, COMD_EVAL_TABLE, MIXED>(d, g, P, w, c, L);
        }
    }
}

// Persistent grids never exceed the blocks that can be resident at once (or
// the last ones would start only when others finish).  The occupancy query
// is made once per kernel and block shape, so that launches made while a
// step is captured into a graph issue nothing else.
template <typename K>
static int resident_blocks(K kern, int threads, size_t lds, int ncu, int *cache)
{
    if (*cache == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess || per_cu < 1)
            per_cu = -1;
        *cache = per_cu;
    }
    return *cache > 0 && ncu > 0 ? *cache * ncu : 1 << 30;
}

template <int KIND, int EVAL, int MIXED>
static void launch3(const DevState &d, const PotDev &P, const ForceCfg &cfg, hipStream_t st)
{
    static int occ[5] = {0, 0, 0, 0, 0};
    const int W = cfg.waves;
    const size_t lds = (size_t)W * wave_lds_bytes(KIND);
    auto kern = k_pair<KIND, EVAL, MIXED>;
    const int cap = resident_blocks(kern, 64 * W, lds, cfg.ncu, &occ[W < 5 ? W : 4]);
    const int blocks = cfg.blocks < cap ? cfg.blocks : cap;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * W), lds, st, d, P);
}

template <int KIND, int EVAL>
static void launch2(const DevState &d, const PotDev &P, int mixed, const ForceCfg &cfg, hipStream_t st)
{
    if (mixed) launch3<KIND, EVAL, 1>(d, P, cfg, st);
    else launch3<KIND, EVAL, 0>(d, P, cfg, st);
}

int cm_force_lds_bytes(const PotDev &P, int kind, int mixed, int waves, int tab_lds)
{
    (void)P;
    (void)mixed;
    (void)tab_lds;
    return waves * wave_lds_bytes(kind);
}

template <int PASS, int MIXED>
static void launch_batch(const DevState &d, const PotDev *pots, const ForceCfg &cfg, hipStream_t st)
{
    static int occ = 0;
    const size_t lds = wave_lds_bytes(PASS == 1 ? CM_KIND_EAM1 : CM_KIND_EAM3);
    auto kern = k_pair_batch<PASS, MIXED>;
    const int cap = resident_blocks(kern, 64, lds, cfg.ncu, &occ);
    const int blocks = cfg.blocks < cap ? cfg.blocks : cap;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64), lds, st, d, pots);
}

void cm_launch_force_batch(const DevState &d, const PotDev *pots, int pass, int mixed, const ForceCfg &cfg,
                           hipStream_t st)
{
    if (pass == 1) {
        if (mixed) launch_batch<1, 1>(d, pots, cfg, st);
        else launch_batch<1, 0>(d, pots, cfg, st);
    } else {
        if (mixed) launch_batch<3, 1>(d, pots, cfg, st);
        else launch_batch<3, 0>(d, pots, cfg, st);
    }
}

void cm_launch_force(const DevState &d, const PotDev &P, int kind, int mixed, const ForceCfg &cfg, hipStream_t st)
{
    if (kind == CM_KIND_LJ) {
        launch2<CM_KIND_LJ, COMD_EVAL_TABLE>(d, P, mixed, cfg, st);
    } else if (P.eval == COMD_EVAL_CHEB) {
        if (kind == CM_KIND_EAM1) launch2<CM_KIND_EAM1, COMD_EVAL_CHEB>(d, P, mixed, cfg, st);
        else launch2<CM_KIND_EAM3, COMD_EVAL_CHEB>(d, P, mixed, cfg, st);
    } else {
        if (kind == CM_KIND_EAM1) launch2<CM_KIND_EAM1, COMD_EVAL_TABLE>(d, P, mixed, cfg, st);
        else launch2<CM_KIND_EAM3, COMD_EVAL_TABLE>(d, P, mixed, cfg, st);
    }
}
