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

#define NL_BATCH 448  // halo atoms staged per batch in the build

__device__ __forceinline__ bool nl_rebuild(const CmList &l) { return (l.flags[0] & CM_NL_REBUILD) != 0; }

// ---------------------------------------------------------------------------
// id -> slot and id -> position from the slot arrays (after binning or an
// upload), and a rebuild request

__global__ void k_nl_clear_ids(CmList l)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < l.n; i += gridDim.x * blockDim.x) l.slot[i] = -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(l.flags, (unsigned)CM_NL_REBUILD);
}

__global__ __launch_bounds__(64) void k_nl_from_slots(DevState d, CmList l)
{
    const int c = blockIdx.x;
    const int t = threadIdx.x;
    if (t >= d.natoms[c]) return;
    const int s = c * CM_SLOTS + t;
    const int id = d.id[s];
    if (id < 0 || id >= l.n) {
        atomicOr(l.flags, (unsigned)CM_NL_INVALID);
        return;
    }
    double cx, cy, cz;
    cell_corner(d.g0, c, cx, cy, cz);
    l.slot[id] = s;
    const double x = cx + d.rx[s], y = cy + d.ry[s], z = cz + d.rz[s];
    l.pos[id] = make_double4(x, y, z, 0.0);
    l.pos32[id] = make_uint4(fixed32(x, l.qs0), fixed32(y, l.qs1), fixed32(z, l.qs2), 0u);
}

// fixed32 positions from the double ones (before the first mixed-precision
// pass after FP64 steps, which keep only the double positions current)
__global__ void k_nl_refresh32(CmList l)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= l.n || l.slot[i] < 0) return;
    const double4 p = l.pos[i];
    l.pos32[i] = make_uint4(fixed32(p.x, l.qs0), fixed32(p.y, l.qs1), fixed32(p.z, l.qs2), 0u);
}

// ---------------------------------------------------------------------------
// build (each kernel returns at once unless a rebuild was requested)

__global__ void k_nl_zero(CmList l)
{
    if (!nl_rebuild(l)) return;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < l.naux; a += gridDim.x * blockDim.x) l.aux_cnt[a] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(l.flags, ~(unsigned)CM_NL_INVALID);
}

__device__ __forceinline__ double wrap_coord(double x, double L) { return x - L * floor(x / L); }

__global__ void k_nl_key(CmList l)
{
    if (!nl_rebuild(l)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= l.n || l.slot[i] < 0) return;
    const double4 p = l.pos[i];
    int ax = (int)(wrap_coord(p.x, l.L0) / l.ab0);
    int ay = (int)(wrap_coord(p.y, l.L1) / l.ab1);
    int az = (int)(wrap_coord(p.z, l.L2) / l.ab2);
    ax = min(max(ax, 0), l.na0 - 1);
    ay = min(max(ay, 0), l.na1 - 1);
    az = min(max(az, 0), l.na2 - 1);
    const int a = ax + l.na0 * (ay + l.na1 * az);
    const int k = atomicAdd(&l.aux_cnt[a], 1);
    if (k < CM_NL_AUXCAP) l.aux_mem[a * CM_NL_AUXCAP + k] = i;
    else atomicOr(l.flags, (unsigned)CM_NL_INVALID);
}

// Exclusive prefix sum of the auxiliary cells' counts, one block, in a
// fixed order; count[0] = atoms listed.
__global__ __launch_bounds__(1024) void k_nl_scan(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (l.naux + 1023) / 1024;
    const int a0 = t * per, a1 = min(a0 + per, l.naux);
    int s = 0;
    for (int a = a0; a < a1; a++) s += min(l.aux_cnt[a], CM_NL_AUXCAP);
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int a = a0; a < a1; a++) {
        l.aux_start[a] = run;
        run += min(l.aux_cnt[a], CM_NL_AUXCAP);
    }
    if (t == 1023) l.count[0] = part[1023];
}

// Members of each auxiliary cell in id order (deterministic lists), and the
// build order.
__global__ __launch_bounds__(64) void k_nl_sort(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ int mem[CM_NL_AUXCAP];
    const int t = threadIdx.x;
    for (int a = blockIdx.x; a < l.naux; a += gridDim.x) {
        const int n = min(l.aux_cnt[a], CM_NL_AUXCAP);
        const int me = t < n ? l.aux_mem[a * CM_NL_AUXCAP + t] : 0;
        if (t < n) mem[t] = me;
        __syncthreads();
        int rank = 0;
        for (int u = 0; u < n; u++) rank += mem[u] < me;
        __syncthreads();
        if (t < n) {
            l.aux_mem[a * CM_NL_AUXCAP + rank] = me;
            l.order[l.aux_start[a] + rank] = me;
        }
        __syncthreads();
    }
}

// One wave per auxiliary cell: its 27 neighbours' atoms are staged in LDS at
// their nearest periodic image; lane i (a home atom) keeps every atom within
// cutoff + skin.
__global__ __launch_bounds__(64) void k_nl_build(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ double hx[NL_BATCH], hy[NL_BATCH], hz[NL_BATCH];
    __shared__ int hid[NL_BATCH];
    __shared__ int cell[27], off[28];
    __shared__ double shift[81];
    const int L = threadIdx.x;
    for (int a = blockIdx.x; a < l.naux; a += gridDim.x) {
        const int nh = min(l.aux_cnt[a], CM_NL_AUXCAP);
        if (nh == 0) continue;
        const int ax = a % l.na0, ay = (a / l.na0) % l.na1, az = a / (l.na0 * l.na1);
        int ne = 0;
        if (L < 27) {
            int jx = ax + L / 9 - 1, jy = ay + (L / 3) % 3 - 1, jz = az + L % 3 - 1;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            if (jx < 0) { jx += l.na0; sx = -l.L0; } else if (jx >= l.na0) { jx -= l.na0; sx = l.L0; }
            if (jy < 0) { jy += l.na1; sy = -l.L1; } else if (jy >= l.na1) { jy -= l.na1; sy = l.L1; }
            if (jz < 0) { jz += l.na2; sz = -l.L2; } else if (jz >= l.na2) { jz -= l.na2; sz = l.L2; }
            const int b = jx + l.na0 * (jy + l.na1 * jz);
            cell[L] = b;
            ne = min(l.aux_cnt[b], CM_NL_AUXCAP);
            shift[3 * L + 0] = sx;
            shift[3 * L + 1] = sy;
            shift[3 * L + 2] = sz;
        }
        int incl = ne;
        for (int o = 1; o < 32; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (L >= o) incl += v;
        }
        const int H = __shfl(incl, 31, 64);
        if (L < 27) off[L] = incl - ne;
        if (L == 0) off[27] = H;
        // home atom of this lane, wrapped into the box like its cell
        const int me = L < nh ? l.aux_mem[a * CM_NL_AUXCAP + L] : -1;
        double xi = 0, yi = 0, zi = 0;
        if (me >= 0) {
            const double4 p = l.pos[me];
            xi = wrap_coord(p.x, l.L0);
            yi = wrap_coord(p.y, l.L1);
            zi = wrap_coord(p.z, l.L2);
        }
        const int t = me >= 0 ? l.aux_start[a] + L : 0;
        int k = 0;
        for (int h0 = 0; h0 < H; h0 += NL_BATCH) {
            const int hb = min(H - h0, NL_BATCH);
            __syncthreads();
            int e = 0;
            for (int h = h0 + L; h < h0 + hb; h += 64) {
                while (off[e + 1] <= h) e++;
                const int j = l.aux_mem[cell[e] * CM_NL_AUXCAP + (h - off[e])];
                const double4 p = l.pos[j];
                hx[h - h0] = wrap_coord(p.x, l.L0) + shift[3 * e + 0];
                hy[h - h0] = wrap_coord(p.y, l.L1) + shift[3 * e + 1];
                hz[h - h0] = wrap_coord(p.z, l.L2) + shift[3 * e + 2];
                hid[h - h0] = j;
            }
            __syncthreads();
            if (me >= 0) {
                for (int j = 0; j < hb; j++) {
                    const double dx = xi - hx[j], dy = yi - hy[j], dz = zi - hz[j];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 <= l.rl2 && hid[j] != me) {
                        if (k < l.kmax) l.nbl[(size_t)k * l.n + t] = hid[j];
                        k++;
                    }
                }
            }
        }
        if (me >= 0) {
            l.ncnt[t] = min(k, l.kmax);
            if (k > l.kmax) atomicOr(l.flags, (unsigned)CM_NL_INVALID);
        }
        __syncthreads();
    }
}

__global__ void k_nl_finish(CmList l)
{
    if (!nl_rebuild(l)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < l.n && l.slot[i] >= 0) l.ref[i] = l.pos[i];
}

__global__ void k_nl_done(CmList l)
{
    if (!nl_rebuild(l)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        l.count[1] += 1;  // builds so far
        atomicAnd(l.flags, ~(unsigned)CM_NL_REBUILD);
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
    const double iL0 = 1.0 / L0, iL1 = 1.0 / L1, iL2 = 1.0 / L2;
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
        d.fx[s] = A.x;
        d.fy[s] = A.y;
        d.fz[s] = A.z;
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
    hipLaunchKernelGGL(k_nl_from_slots, dim3(d.ncells), dim3(64), 0, st, d, l);
}

void cm_nl_refresh32(const CmList &l, hipStream_t st)
{
    hipLaunchKernelGGL(k_nl_refresh32, dim3(grid_for(l.n, 256)), dim3(256), 0, st, l);
}

void cm_nl_build(const DevState &d, const CmList &l, hipStream_t st)
{
    (void)d;
    const unsigned ga = grid_for(l.naux, 256), gi = grid_for(l.n, 256);
    const unsigned gc = (unsigned)(l.naux < 8192 ? l.naux : 8192);
    hipLaunchKernelGGL(k_nl_zero, dim3(ga), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_key, dim3(gi), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_scan, dim3(1), dim3(1024), 0, st, l);
    hipLaunchKernelGGL(k_nl_sort, dim3(gc), dim3(64), 0, st, l);
    hipLaunchKernelGGL(k_nl_build, dim3(gc), dim3(64), 0, st, l);
    hipLaunchKernelGGL(k_nl_finish, dim3(gi), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_done, dim3(1), dim3(64), 0, st, l);
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
