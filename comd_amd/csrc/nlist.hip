// nlist.hip -- Verlet neighbour lists for the force passes of a single
// periodic simulation.
//
// The link-cell kernels test ~27 x 10.5 candidates per atom and pass in every
// force call.  Here each atom keeps the ids of all atoms that were within
// cutoff + skin when the lists were last built; the force passes test only
// those (about 55 per atom for Cu with a 0.4 A skin).  Every pair inside the
// cutoff is on the lists as long as no atom has moved more than half the skin
// since the build; the drift kernel checks that per atom and raises a device
// flag, and the build kernels, launched every step, return at once unless the
// flag is up.  Nothing waits on the host, so a step stays capturable.
//
// Lists are keyed by atom id, so the rebinning (which moves atoms between
// slots with the reference's swap-with-last order, src-flat/domains.c:
// 347-389) does not invalidate them: the rebin kernels keep an id -> slot map
// and the drift kernel an id -> absolute position map.  Outputs go to the
// slot arrays exactly as the link-cell kernels write them, with the same
// cutoff predicates (src-flat/eam.c:519,633, ljForce.c:188); only the order
// of the per-atom sums differs.  Positions are absolute (cell corner + r),
// pairs are imaged by the nearest periodic image, so the lists need at
// least three auxiliary cells of edge cutoff + skin per dimension; the
// host falls back to the link-cell kernels otherwise.
#include "comd_internal.h"
#include "pair_terms.h"

__device__ __forceinline__ bool nl_rebuild(const CmList &l) { return (l.flags[0] & CM_NL_REBUILD) != 0; }

// ---------------------------------------------------------------------------
// id -> slot and id -> position from the slot arrays (after binning or an
// upload), and a rebuild request

__global__ void k_nl_clear_ids(CmList l)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < l.n; i += gridDim.x * blockDim.x) l.slot[i] = -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(l.flags, (unsigned)CM_NL_REBUILD);
}

// check = 0: after binning or an upload (the ids are checked);
// check = 1: positions after steps that did not keep them, with the
// displacement test of the drift kernel
__global__ __launch_bounds__(64) void k_nl_from_slots(DevState d, CmList l, int check)
{
    const int c = blockIdx.x;
    const int t = threadIdx.x;
    bool far = false;
    if (t < d.natoms[c]) {
        const int s = c * CM_SLOTS + t;
        const int id = d.id[s];
        if (id < 0 || id >= l.n) {
            atomicOr(l.flags, (unsigned)CM_NL_INVALID);
        } else {
            double cx, cy, cz;
            cell_corner(d.g0, c, cx, cy, cz);
            const double x = cx + d.rx[s], y = cy + d.ry[s], z = cz + d.rz[s];
            l.slot[id] = s;
            l.pos[id] = make_double4(x, y, z, 0.0);
            l.pos32[id] = make_uint4(fixed32(x, l.qs0), fixed32(y, l.qs1), fixed32(z, l.qs2), 0u);
            if (check) {
                const double4 q = l.ref[id];
                double ex = x - q.x, ey = y - q.y, ez = z - q.z;
                ex -= l.L0 * rint(ex * l.iL0);
                ey -= l.L1 * rint(ey * l.iL1);
                ez -= l.L2 * rint(ez * l.iL2);
                far = ex * ex + ey * ey + ez * ez > d.nl_half2;
            }
        }
    }
    if (__ballot(far) != 0 && t == 0) atomicOr(l.flags, (unsigned)CM_NL_REBUILD);
}

// ---------------------------------------------------------------------------
// build (each kernel returns at once unless a rebuild was requested, so the
// launches that find no request cost only their dispatch: their grids are
// kept small, with strided loops)

// x wrapped into [0, L) (to rounding); iL = 1 / L
__device__ __forceinline__ double wrap_coord(double x, double L, double iL) { return x - L * floor(x * iL); }

__device__ __forceinline__ void nl_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Each atom into its auxiliary cell (counts start at zero: the previous
// build's last block cleared them).
__global__ __launch_bounds__(256) void k_nl_key(CmList l)
{
    if (!nl_rebuild(l)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < l.n; i += gridDim.x * blockDim.x) {
        if (l.slot[i] < 0) continue;
        const double4 p = l.pos[i];
        int ax = (int)(wrap_coord(p.x, l.L0, l.iL0) * l.iab0);
        int ay = (int)(wrap_coord(p.y, l.L1, l.iL1) * l.iab1);
        int az = (int)(wrap_coord(p.z, l.L2, l.iL2) * l.iab2);
        ax = min(max(ax, 0), l.na0 - 1);
        ay = min(max(ay, 0), l.na1 - 1);
        az = min(max(az, 0), l.na2 - 1);
        const int a = ax + l.na0 * (ay + l.na1 * az);
        const int k = atomicAdd(&l.aux_cnt[a], 1);
        if (k < CM_NL_AUXCAP) l.aux_mem[a * CM_NL_AUXCAP + k] = i;
        else atomicOr(l.flags, (unsigned)CM_NL_OVERFLOW);
    }
}

// Exclusive prefix sum of the auxiliary cells' counts, one block walking
// tiles of 8 x 1024 cells; count[0] = atoms listed.
__global__ __launch_bounds__(1024) void k_nl_scan(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ int wsum[16];
    __shared__ int carry_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int carry = 0;
    for (int base = 0; base < l.naux; base += 8 * 1024) {
        int v[8], s = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int a = base + 8 * t + k;
            v[k] = a < l.naux ? min(l.aux_cnt[a], CM_NL_AUXCAP) : 0;
            s += v[k];
        }
        // inclusive scan of the threads' sums: within the wave, then across
        int incl = s;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        if (t == 0) {
            int run = 0;
            for (int k = 0; k < 16; k++) {
                const int x = wsum[k];
                wsum[k] = run;
                run += x;
            }
            carry_s = run;
        }
        __syncthreads();
        int run = carry + wsum[w] + incl - s;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int a = base + 8 * t + k;
            if (a < l.naux) l.aux_start[a] = run;
            run += v[k];
        }
        carry += carry_s;
        __syncthreads();
    }
    if (t == 0) l.count[0] = carry;
}

// Members of each auxiliary cell in id order (deterministic lists), and the
// build order.  One wave per cell, four waves per block.
__global__ __launch_bounds__(256) void k_nl_sort(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ int mem_s[4][CM_NL_AUXCAP];
    const int L = threadIdx.x & 63, w = threadIdx.x >> 6;
    int *mem = mem_s[w];
    for (int a = blockIdx.x * 4 + w; a < l.naux; a += gridDim.x * 4) {
        const int n = min(l.aux_cnt[a], CM_NL_AUXCAP);
        if (n == 0) continue;
        const int me = L < n ? l.aux_mem[a * CM_NL_AUXCAP + L] : 0;
        if (L < n) mem[L] = me;
        nl_wave_sync();
        int rank = 0;
        for (int u = 0; u < n; u++) rank += mem[u] < me;
        nl_wave_sync();
        if (L < n) {
            l.aux_mem[a * CM_NL_AUXCAP + rank] = me;
            l.order[l.aux_start[a] + rank] = me;
        }
    }
}

// One wave per auxiliary cell: its 27 neighbours' atoms are staged in LDS at
// their nearest periodic image.  The nh home atoms are split over the lanes
// P = 64 / nh ways: lane (a, q) tests halo entries q, q + P, ...; a count
// pass and a prefix over the P lanes of an atom place each lane's entries,
// so an atom's list holds, per staged batch, partition 0's neighbours, then
// partition 1's, ... (a fixed order for a given build).  The last block to
// finish closes the build: overflow -> lists invalid, rebuild request
// cleared, counts cleared for the next build.  About 8 KB of LDS per wave,
// so that enough waves are resident to hide the gathers' latency.
#define NL_BATCH 256   // halo atoms staged per batch and wave
__global__ __launch_bounds__(256) void k_nl_build(CmList l)
{
    if (!nl_rebuild(l)) return;
    struct WaveBuf {
        double hx[NL_BATCH], hy[NL_BATCH], hz[NL_BATCH];
        int hid[NL_BATCH];
        double shift[81];
        int cell[27], off[28];
    };
    __shared__ WaveBuf buf[4];
    __shared__ int last;
    const int L = threadIdx.x & 63, w = threadIdx.x >> 6;
    WaveBuf &B = buf[w];
    // waves walk contiguous ranges of cells, and the blocks of one XCD
    // (blocks are dealt round-robin over the 8 XCDs) neighbouring ranges
    const int nb = gridDim.x, qb = nb / 8, rb = nb % 8;
    const int xcd = blockIdx.x % 8, bi = blockIdx.x / 8;
    const int blk = nb < 8 ? (int)blockIdx.x : xcd * qb + min(xcd, rb) + bi;
    const int nw = nb * 4, gw = blk * 4 + w;
    const int per = (l.naux + nw - 1) / nw;
    const int a0 = gw * per, a1 = min(a0 + per, l.naux);
    bool ovf = false;
    for (int a = a0; a < a1; a++) {
        const int nh = min(l.aux_cnt[a], CM_NL_AUXCAP);
        if (nh == 0) continue;
        const int ax = a % l.na0, ay = (a / l.na0) % l.na1, az = a / (l.na0 * l.na1);
        nl_wave_sync();
        int ne = 0;
        if (L < 27) {
            int jx = ax + L / 9 - 1, jy = ay + (L / 3) % 3 - 1, jz = az + L % 3 - 1;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            if (jx < 0) { jx += l.na0; sx = -l.L0; } else if (jx >= l.na0) { jx -= l.na0; sx = l.L0; }
            if (jy < 0) { jy += l.na1; sy = -l.L1; } else if (jy >= l.na1) { jy -= l.na1; sy = l.L1; }
            if (jz < 0) { jz += l.na2; sz = -l.L2; } else if (jz >= l.na2) { jz -= l.na2; sz = l.L2; }
            const int b = jx + l.na0 * (jy + l.na1 * jz);
            B.cell[L] = b;
            ne = min(l.aux_cnt[b], CM_NL_AUXCAP);
            B.shift[3 * L + 0] = sx;
            B.shift[3 * L + 1] = sy;
            B.shift[3 * L + 2] = sz;
        }
        int incl = ne;
        for (int o = 1; o < 32; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (L >= o) incl += v;
        }
        const int H = __shfl(incl, 31, 64);
        if (L < 27) B.off[L] = incl - ne;
        if (L == 0) B.off[27] = H;
        // lane (ha, q): home atom ha, partition q of P
        const int P = 64 / nh;
        const bool act = L < P * nh;
        const int ha = act ? L % nh : 0, q = act ? L / nh : 0;
        const int me = act ? l.aux_mem[a * CM_NL_AUXCAP + ha] : -1;
        double xi = 0, yi = 0, zi = 0;
        if (act) {
            const double4 p = l.pos[me];
            xi = wrap_coord(p.x, l.L0, l.iL0);
            yi = wrap_coord(p.y, l.L1, l.iL1);
            zi = wrap_coord(p.z, l.L2, l.iL2);
        }
        const int t = l.aux_start[a] + ha;
        int base = 0;  // entries of atom ha so far (the same in its P lanes)
        for (int h0 = 0; h0 < H; h0 += NL_BATCH) {
            const int hb = min(H - h0, NL_BATCH);
            nl_wave_sync();
            int e = 0;
            for (int h = h0 + L; h < h0 + hb; h += 64) {
                while (B.off[e + 1] <= h) e++;
                const int j = l.aux_mem[B.cell[e] * CM_NL_AUXCAP + (h - B.off[e])];
                const double4 p = l.pos[j];
                B.hx[h - h0] = wrap_coord(p.x, l.L0, l.iL0) + B.shift[3 * e + 0];
                B.hy[h - h0] = wrap_coord(p.y, l.L1, l.iL1) + B.shift[3 * e + 1];
                B.hz[h - h0] = wrap_coord(p.z, l.L2, l.iL2) + B.shift[3 * e + 2];
                B.hid[h - h0] = j;
            }
            nl_wave_sync();
            int c = 0;
            if (act) {
                for (int j = q; j < hb; j += P) {
                    const double dx = xi - B.hx[j], dy = yi - B.hy[j], dz = zi - B.hz[j];
                    c += (dx * dx + dy * dy + dz * dz <= l.rl2) && B.hid[j] != me;
                }
            }
            int pre = 0, tot = 0;
            for (int qq = 0; qq < P; qq++) {
                const int v = __shfl(c, qq * nh + ha, 64);
                pre += qq < q ? v : 0;
                tot += v;
            }
            if (act) {
                int k = base + pre;
                for (int j = q; j < hb; j += P) {
                    const double dx = xi - B.hx[j], dy = yi - B.hy[j], dz = zi - B.hz[j];
                    if ((dx * dx + dy * dy + dz * dz <= l.rl2) && B.hid[j] != me) {
                        if (k < l.kmax) l.nbl[(size_t)k * l.n + t] = B.hid[j];
                        k++;
                    }
                }
            }
            base += tot;
        }
        if (act && q == 0) {
            l.ncnt[t] = min(base, l.kmax);
            ovf |= base > l.kmax;
            l.ref[me] = l.pos[me];
        }
    }
    if (ovf) atomicOr(l.flags, (unsigned)CM_NL_OVERFLOW);
    // the last block to finish closes the build
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&l.count[2], 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int a = threadIdx.x; a < l.naux; a += blockDim.x) l.aux_cnt[a] = 0;
    if (threadIdx.x == 0) {
        l.count[2] = 0;
        l.count[1] += 1;  // builds so far
        const unsigned f = atomicOr(l.flags, 0u);
        if (f & CM_NL_OVERFLOW) atomicOr(l.flags, (unsigned)CM_NL_INVALID);
        else atomicAnd(l.flags, ~(unsigned)CM_NL_INVALID);
        atomicAnd(l.flags, ~(unsigned)(CM_NL_OVERFLOW | CM_NL_REBUILD));
    }
}

// ---------------------------------------------------------------------------
// force passes over the lists (return at once when the lists are invalid;
// the link-cell kernels then run instead)

// Non-empty stencil cells of each link cell: the reference adds the kinetic
// virial term once per non-empty neighbour cell (src-flat/ljForce.c:175).
__global__ void k_nl_nonempty(DevState d, CmList l)
{
    if (l.flags[0] & CM_NL_INVALID) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncells) return;
    int n = 0;
    for (int e = 0; e < 27; e++) {
        const int b = stencil_cell(d.g0, c, e);
        n += b >= 0 && d.natoms[b] > 0;
    }
    l.nonempty[c] = n;
}

// Rows of the phi and rho coefficient tables a pass reads (phi: density
// pass only), staged per block in LDS when they fit: every pair reads a
// different interval, and from LDS these reads do not load the vector L1.
__host__ __device__ inline int nl_rows(const PotDev &P, int kind) { return (kind == CM_KIND_EAM1 ? P.n[0] : 0) + P.n[1]; }

// One atom (thread t of the build order) over its list.  TB: the table
// pointers' type (global or LDS rows).
template <int KIND, int EVAL, int MIXED, class TB>
__device__ __forceinline__ void nl_atom(const DevState &d, const PotDev &P, const CmList &l, const TB &T, int t)
{
    const int i = l.order[t];
    const int s = l.slot[i];
    // clang vector types: loadable through address-space-qualified pointers
    typedef double v4d __attribute__((ext_vector_type(4)));
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const CM_GLOBAL v4d *pos = (const CM_GLOBAL v4d *)l.pos;
    const CM_GLOBAL v4u *pos32 = (const CM_GLOBAL v4u *)l.pos32;
    // positions: FP64 passes gather double4 (x, y, z, F'), nearest image by
    // rint; mixed passes gather 16-byte fixed32 (x, y, z, float F'), nearest
    // image by 32-bit wrap-around.  F' is there once the density pass ran.
    v4d pi = {0, 0, 0, 0};
    v4u qi = {0, 0, 0, 0};
    if (MIXED) qi = pos32[i];
    else pi = pos[i];
    const double fia = KIND != CM_KIND_EAM3 ? 0.0 : MIXED ? (double)__uint_as_float(qi.w) : pi.w;
    const double rc2 = P.rc2;
    const double L0 = l.L0, L1 = l.L1, L2 = l.L2;
    const double iL0 = l.iL0, iL1 = l.iL1, iL2 = l.iL2;
    const int cnt = l.ncnt[t];
    Acc A = {0, 0, 0, 0, 0, 0};
    int npair = 0;
    // NU neighbours per trip: their list entries, then their positions (and
    // F'), are loaded together so that the loads' latencies overlap
    constexpr int NU = 4;
    const CM_GLOBAL int *row = (const CM_GLOBAL int *)l.nbl + t;
    for (int k0 = 0; k0 < cnt; k0 += NU) {
        int jj[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) jj[u] = k0 + u < cnt ? row[(size_t)(k0 + u) * l.n] : i;
        v4d pj[NU];
        v4u qj[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
            if (MIXED) qj[u] = pos32[jj[u]];
            else pj[u] = pos[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
            double dx, dy, dz, fij;
            if (MIXED) {
                dx = (double)(int)(qi.x - qj[u].x) * l.qi0;
                dy = (double)(int)(qi.y - qj[u].y) * l.qi1;
                dz = (double)(int)(qi.z - qj[u].z) * l.qi2;
                fij = (double)__uint_as_float(qj[u].w);
            } else {
                dx = pi.x - pj[u].x;
                dy = pi.y - pj[u].y;
                dz = pi.z - pj[u].z;
                dx -= L0 * rint(dx * iL0);
                dy -= L1 * rint(dy * iL1);
                dz -= L2 * rint(dz * iL2);
                fij = pj[u].w;
            }
            const double r2 = dx * dx + dy * dy + dz * dz;
            // a padding entry (k >= cnt) is the atom itself: r2 = 0
            const bool in = r2 > 0.0 && (KIND == CM_KIND_EAM3 ? (r2 < rc2) : (r2 <= rc2));
            if (in) {
                pair_term<KIND, EVAL, MIXED>(P, T, dx, dy, dz, r2, fia, KIND == CM_KIND_EAM3 ? fij : 0.0, A);
                npair++;
            }
        }
    }
    if (KIND == CM_KIND_EAM1) {
        // pass 2, src-flat/eam.c:576-594
        double fv, fp;
        if (EVAL == COMD_EVAL_CHEB) {
            fv = cheb64(P.cheb + 2 * P.cn, P.cn, P.ca[2], P.cb[2], A.rho);
            fp = cheb64(P.cheb + 5 * P.cn, P.cn, P.ca[2], P.cb[2], A.rho);
        } else {
            interp64(P.tab[2], P.x0[2], P.xn[2], P.invdx[2], P.n[2], A.rho, fv, fp);
        }
        d.fx[s] = A.x;
        d.fy[s] = A.y;
        d.fz[s] = A.z;
        d.fe[s] = A.e + fv;
        d.rho[s] = A.rho;
        d.fi[s] = fp;
        d.nnb[s] = npair;
        if (MIXED) ((unsigned *)&l.pos32[i])[3] = __float_as_uint((float)fp);
        else ((double *)&l.pos[i])[3] = fp;
    } else if (KIND == CM_KIND_EAM3) {
        d.fx[s] += A.x;
        d.fy[s] += A.y;
        d.fz[s] += A.z;
        d.vir[s] = A.vir;
    } else {
        const double px = d.px[s];
        d.fx[s] = P.lj_fsign * A.x;
        d.fy[s] = P.lj_fsign * A.y;
        d.fz[s] = P.lj_fsign * A.z;
        d.fe[s] = A.e;
        d.vir[s] = A.vir - (double)l.nonempty[s / CM_SLOTS] * (px * px / P.mass_amu);
        d.nnb[s] = npair;
    }
}

// TLDS: the table rows are staged in LDS (1) or read from global memory (0);
// a compile-time choice, so the row reads are LDS or global loads.
template <int KIND, int EVAL, int MIXED, int TLDS>
__global__ __launch_bounds__(512) void k_nl_pair(DevState d, PotDev P, CmList l)
{
    extern __shared__ __align__(16) char nls[];
    if (l.flags[0] & CM_NL_INVALID) return;
    // blocks are dealt round-robin over the 8 XCDs; give each XCD a
    // contiguous range of the build order (a slab of space) so that the
    // neighbours it gathers stay in its own L2
    const int nb = gridDim.x, q = nb / 8, rem = nb % 8;
    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    const int blk = nb < 8 ? (int)blockIdx.x : xcd * q + min(xcd, rem) + idx;
    const int t = blk * blockDim.x + threadIdx.x;
    const bool live = t < l.count[0];
    if (TLDS) {
        const int n0 = KIND == CM_KIND_EAM1 ? P.n[0] : 0, n1 = P.n[1];
        TabsLds T;
        if (MIXED) {
            CM_LDS float *s0 = (CM_LDS float *)nls, *s1 = s0 + 8 * n0;
            for (int k = threadIdx.x; k < 8 * n0; k += blockDim.x) s0[k] = P.cf32[0][k];
            for (int k = threadIdx.x; k < 8 * n1; k += blockDim.x) s1[k] = P.cf32[1][k];
            T.c32[0] = s0;
            T.c32[1] = s1;
            T.c64[0] = T.c64[1] = nullptr;
        } else {
            CM_LDS double *s0 = (CM_LDS double *)nls, *s1 = s0 + 6 * n0;
            for (int k = threadIdx.x; k < 6 * n0; k += blockDim.x) s0[k] = P.cf64[0][k];
            for (int k = threadIdx.x; k < 6 * n1; k += blockDim.x) s1[k] = P.cf64[1][k];
            T.c64[0] = s0;
            T.c64[1] = s1;
            T.c32[0] = T.c32[1] = nullptr;
        }
        __syncthreads();
        if (live) nl_atom<KIND, EVAL, MIXED>(d, P, l, T, t);
    } else {
        if (live) nl_atom<KIND, EVAL, MIXED>(d, P, l, global_tabs(P), t);
    }
}

// ---------------------------------------------------------------------------
// host launchers

static unsigned grid_for(long n, int block)
{
    long b = (n + block - 1) / block;
    return (unsigned)(b < 1 ? 1 : b);
}

void cm_nl_init(const DevState &d, const CmList &l, hipStream_t st)
{
    hipLaunchKernelGGL(k_nl_clear_ids, dim3(grid_for(l.n, 256)), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_from_slots, dim3(d.ncells), dim3(64), 0, st, d, l, 0);
}

void cm_nl_sync(const DevState &d, const CmList &l, hipStream_t st)
{
    hipLaunchKernelGGL(k_nl_from_slots, dim3(d.ncells), dim3(64), 0, st, d, l, 1);
}

void cm_nl_build(const DevState &d, const CmList &l, hipStream_t st)
{
    (void)d;
    hipLaunchKernelGGL(k_nl_key, dim3(1024), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_scan, dim3(1), dim3(1024), 0, st, l);
    hipLaunchKernelGGL(k_nl_sort, dim3(1024), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_build, dim3(1024), dim3(256), 0, st, l);
}

template <int KIND, int EVAL, int MIXED>
static void nl_launch(const DevState &d, const PotDev &P, const CmList &l, hipStream_t st)
{
    size_t lds = 0;
    if (EVAL == COMD_EVAL_TABLE && KIND != CM_KIND_LJ) {
        lds = (size_t)nl_rows(P, KIND) * (MIXED ? 32 : 48);
        if (lds > 65536) lds = 0;  // tables too large: read them from global memory
    }
    const dim3 grid(grid_for(l.n, 512)), block(512);
    if (EVAL == COMD_EVAL_TABLE && KIND != CM_KIND_LJ && lds > 0)
        hipLaunchKernelGGL((k_nl_pair<KIND, EVAL, MIXED, 1>), grid, block, lds, st, d, P, l);
    else
        hipLaunchKernelGGL((k_nl_pair<KIND, EVAL, MIXED, 0>), grid, block, 0, st, d, P, l);
}

template <int KIND, int EVAL>
static void nl_launch2(const DevState &d, const PotDev &P, const CmList &l, int mixed, hipStream_t st)
{
    if (mixed) nl_launch<KIND, EVAL, 1>(d, P, l, st);
    else nl_launch<KIND, EVAL, 0>(d, P, l, st);
}

void cm_nl_force(const DevState &d, const PotDev &P, const CmList &l, int kind, int mixed, hipStream_t st)
{
    if (kind == CM_KIND_LJ) {
        hipLaunchKernelGGL(k_nl_nonempty, dim3(grid_for(d.ncells, 256)), dim3(256), 0, st, d, l);
        nl_launch2<CM_KIND_LJ, COMD_EVAL_TABLE>(d, P, l, mixed, st);
    } else if (P.eval == COMD_EVAL_CHEB) {
        if (kind == CM_KIND_EAM1) nl_launch2<CM_KIND_EAM1, COMD_EVAL_CHEB>(d, P, l, mixed, st);
        else nl_launch2<CM_KIND_EAM3, COMD_EVAL_CHEB>(d, P, l, mixed, st);
    } else {
        if (kind == CM_KIND_EAM1) nl_launch2<CM_KIND_EAM1, COMD_EVAL_TABLE>(d, P, l, mixed, st);
        else nl_launch2<CM_KIND_EAM3, COMD_EVAL_TABLE>(d, P, l, mixed, st);
    }
}
