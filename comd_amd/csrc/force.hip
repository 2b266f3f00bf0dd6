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
#include "pair_terms.h"

#define HMAX 288    // halo atoms staged per batch
#define QCAP CM_PLIST  // queue entries per lane (also the pair-list rows a cell keeps)
#define UNROLL 4    // candidates tested per lane between queue checks

// ---------------------------------------------------------------------------
// per-wave LDS: halo x,y,z (HMAX doubles each), halo F' (EAM pass 3), queue
// (QCAP x 64 uint16), per-lane queue counts, stencil cell ids / offsets /
// shifts.  The reduction scratch (7 x 64 values) reuses the halo area.
struct WaveLds {
    double *hx, *hy, *hz, *hf;
    uint16_t *q;
    uint16_t *lst;  // EAM density pass: the pair list being recorded (CM_PLIST x 64)
    int *qn, *cell, *off;
    double *drb;
};

__host__ __device__ constexpr int wave_lds_bytes(int kind)
{
    return HMAX * 8 * (kind == CM_KIND_EAM3 ? 4 : 3) + 27 * 3 * 8 + QCAP * 64 * 2 +
           (kind == CM_KIND_EAM1 ? CM_PLIST * 64 * 2 : 0) + 64 * 4 + 27 * 4 + 28 * 4;
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
    w.lst = w.q + QCAP * 64;
    w.qn = (int *)(w.lst + (kind == CM_KIND_EAM1 ? CM_PLIST * 64 : 0));
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

// One home cell: stencil, halo staging, candidate tests, pair terms,
// reduction of the lanes' partial sums and the epilogue.  The 64 lanes of a
// wave call it together.
template <int KIND, int EVAL, int MIXED>
__device__ __forceinline__ void pair_cell(const DevState &d, const Grid &g, const PotDev &P, const WaveLds &w, int c,
                                          int L)
{
    const double rc2 = P.rc2;
    const TabsGlobal T = global_tabs(P);
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

    // pair list of the density pass: the halo index of every pair this lane
    // evaluates, in order (entries past CM_PLIST are counted, not kept)
    const bool keep = KIND == CM_KIND_EAM1 && d.plist != nullptr;
    int nrec = 0;
    auto record = [&](int hj) {
        if (keep) {
            if (nrec < CM_PLIST) w.lst[nrec * 64 + L] = (uint16_t)hj;
            nrec++;
        }
    };
    // the embedding-force pass replays that list when every lane's fitted:
    // the same pairs in the same order as its own candidate search would give
    const bool replay = KIND == CM_KIND_EAM3 && d.plist != nullptr && d.pvalid[c] != 0;
    int ncnt = 0, kread = 0;
    if (replay) {
        ncnt = act ? d.pcount[c * CM_SLOTS + L] : 0;
        int kmax = ncnt;
        for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
        for (int k = 0; k < kmax; k++) w.q[k * 64 + L] = d.plist[((size_t)c * CM_PLIST + k) * 64 + L];
    }
    wave_sync();

    for (int h0 = 0; h0 < H; h0 += HMAX) {
        const int hb = H - h0 < HMAX ? H - h0 : HMAX;
        // stage the batch: lane L takes halo atoms L, L+64, ... (at most
        // SU); all of its global loads are issued before any LDS store so
        // that their latencies overlap; its stencil entry only moves forward
        {
            constexpr int SU = (HMAX + 63) / 64;
            int sl[SU], ee[SU];
            int e = 0;
#pragma unroll
            for (int u = 0; u < SU; u++) {
                const int h = h0 + L + 64 * u;
                sl[u] = -1;
                ee[u] = 0;
                if (h < h0 + hb) {
                    while (w.off[e + 1] <= h) e++;
                    sl[u] = w.cell[e] * CM_SLOTS + (h - w.off[e]);
                    ee[u] = e;
                }
            }
            double sx[SU], sy[SU], sz[SU], sf[SU];
#pragma unroll
            for (int u = 0; u < SU; u++) {
                const int q = sl[u] >= 0 ? sl[u] : 0;
                sx[u] = d.rx[q];
                sy[u] = d.ry[q];
                sz[u] = d.rz[q];
                sf[u] = KIND == CM_KIND_EAM3 ? d.fi[q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < SU; u++) {
                if (sl[u] >= 0) {
                    const int hh = L + 64 * u;
                    const int eu = ee[u];
                    w.hx[hh] = sx[u] - w.drb[3 * eu + 0];
                    w.hy[hh] = sy[u] - w.drb[3 * eu + 1];
                    w.hz[hh] = sz[u] - w.drb[3 * eu + 2];
                    if (KIND == CM_KIND_EAM3) w.hf[hh] = sf[u];
                }
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
                        pair_term<KIND, EVAL, MIXED>(P, T, dx, dy, dz, r2, fia,
                                                     KIND == CM_KIND_EAM3 ? w.hf[j] : 0.0, A);
                        npair++;
                        record(h0 + j);
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
                    pair_term<KIND, EVAL, MIXED>(P, T, dx, dy, dz, r2, fia, KIND == CM_KIND_EAM3 ? w.hf[j] : 0.0, A);
                    npair++;
                    record(h0 + j);
                    v += S;
                }
            }
            cnt = 0;
            wave_sync();
        };
        if (replay) {
            // this batch's entries of the lane's list (entries rise with
            // the batch), re-tested with the pass's own predicate
            while (true) {
                const int hj = kread < ncnt ? (int)w.q[kread * 64 + L] : 0;
                const bool more = kread < ncnt && hj < h0 + hb;
                if (__ballot(more) == 0) break;
                if (more) {
                    const int j = hj - h0;
                    const double dx = xa - w.hx[j], dy = ya - w.hy[j], dz = za - w.hz[j];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 < rc2) {
                        pair_term<KIND, EVAL, MIXED>(P, T, dx, dy, dz, r2, fia, KIND == CM_KIND_EAM3 ? w.hf[j] : 0.0,
                                                     A);
                        npair++;
                    }
                    kread++;
                }
            }
            wave_sync();
            continue;
        }
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

    if (keep) {
        // the list leaves LDS as whole 128-byte rows (entries past a lane's
        // count are never read)
        d.pcount[c * CM_SLOTS + L] = (uint8_t)(nrec < 255 ? nrec : 255);
        const bool over = __ballot(nrec > CM_PLIST) != 0;
        if (L == 0) d.pvalid[c] = over ? 0 : 1;
        int kmax = nrec < CM_PLIST ? nrec : CM_PLIST;
        for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
        for (int k = 0; k < kmax; k++) d.plist[((size_t)c * CM_PLIST + k) * 64 + L] = w.lst[k * 64 + L];
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
                fv = cheb64(P.cheb + 2 * P.cstride, P.cstride, P.ca[2], P.cb[2], sr);
                fp = cheb64(P.cheb + 5 * P.cstride, P.cstride, P.ca[2], P.cb[2], sr);
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
            d.fx[o] = P.lj_fsign * sx;
            d.fy[o] = P.lj_fsign * sy;
            d.fz[o] = P.lj_fsign * sz;
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
    // the neighbour-list kernels did this pass unless their lists are invalid
    if (d.cell_gate && !(*d.cell_gate & CM_NL_INVALID)) return;
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
    // beside valid neighbour lists there is nothing to do
    if (d.cell_gate && !(*d.cell_gate & CM_NL_INVALID)) return;
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
