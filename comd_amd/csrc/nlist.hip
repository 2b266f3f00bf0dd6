// nlist.hip -- Verlet neighbour lists for the force passes of a single
// periodic simulation.
//
// The link-cell kernels test ~27 x 10.5 candidates per atom and pass in every
// force call.  Here each atom keeps all atoms that were within cutoff + skin
// when the lists were last built; the force passes test only those (about
// 55 per atom for Cu with a 0.4 A skin).  Every pair inside the cutoff is on
// the lists as long as no atom has moved more than half the skin since the
// build; the drift kernel checks that per atom and raises a device flag, and
// the build kernels, launched every step, return at once unless the flag is
// up.  Nothing waits on the host, so a step stays capturable.
//
// Atoms are binned into auxiliary cells of edge >= cutoff + skin, and the
// cells into groups of up to 4 x 4 x 2.  The build order (t) is group-major,
// then cell, then atom id.  A force pass stages a group's halo -- the atoms
// of its cells and of the cells around them -- in LDS once, and each home
// atom's list holds 16-bit indices into that halo: the neighbour positions
// are LDS reads, not gathers through the vector L1, and the list stream is
// half as wide as with atom ids.
//
// Lists refer to atoms through the build order and an id -> position map
// kept by the drift kernel, so the rebinning (which moves atoms between slots
// with the reference's swap-with-last order, src-flat/domains.c:347-389) does
// not invalidate them; the rebin kernels keep an id -> slot map.  Outputs go
// to the slot arrays exactly as the link-cell kernels write them, with the
// same cutoff predicates (src-flat/eam.c:519,633, ljForce.c:188); only the
// order of the per-atom sums differs.  Positions are absolute (cell corner +
// r) and pairs are imaged by the nearest periodic image, so the lists need at
// least three auxiliary cells per dimension; the host falls back to the
// link-cell kernels otherwise.
#include "comd_internal.h"
#include "adj_basis.h"
#include "pair_terms.h"

#include <type_traits>

__device__ __forceinline__ bool nl_rebuild(const CmList &l) { return (l.flags[0] & CM_NL_REBUILD) != 0; }

// x wrapped into [0, L) (to rounding); iL = 1 / L
__device__ __forceinline__ double wrap_coord(double x, double L, double iL) { return x - L * floor(x * iL); }

// List position q (build order), read through the global address space.
__device__ __forceinline__ CmPos load_pos(const CmPos *pos, int q)
{
    const CM_GLOBAL double *p = (const CM_GLOBAL double *)pos + 3 * q;
    return CmPos{p[0], p[1], p[2]};
}

// Absolute position of the atom in slot s of a cell of grid g: the cell's
// corner plus r, as the drift kernel forms it.
__device__ __forceinline__ CmPos slot_pos(const Grid &g, const DevState &d, int s)
{
    const int c = s / CM_SLOTS;
    double cx, cy, cz;
    cell_corner(g, c, cx, cy, cz);
    return CmPos{cx + d.rx[s], cy + d.ry[s], cz + d.rz[s]};
}

// Replica r's list geometry.
__device__ __forceinline__ NlRep nl_rep(const CmList &l, int r) { return l.rep[r]; }


// Auxiliary cell of a position within its replica (the same arithmetic
// wherever it is needed, so that the key and build kernels agree).
__device__ __forceinline__ void aux_coords(const NlRep &R, const CmPos &p, int &ax, int &ay, int &az)
{
    ax = min(max((int)(wrap_coord(p.x, R.L0, R.iL0) * R.iab0), 0), R.na0 - 1);
    ay = min(max((int)(wrap_coord(p.y, R.L1, R.iL1) * R.iab1), 0), R.na1 - 1);
    az = min(max((int)(wrap_coord(p.z, R.L2, R.iL2) * R.iab2), 0), R.na2 - 1);
}

__device__ __forceinline__ void nl_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// id -> slot and id -> position from the slot arrays (after binning or an
// upload), and a rebuild request

__global__ void k_nl_clear_ids(CmList l)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < l.n; i += gridDim.x * blockDim.x) l.slot[i] = -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(l.flags, (unsigned)CM_NL_REBUILD);
}

// check = 0: after binning or an upload (the ids are checked; a build
// follows, which takes the positions from the slots); check = 1: positions
// after steps that did not keep them, with the displacement test of the
// drift kernel.  Four consecutive cells per wave, 16 lanes each.
__global__ __launch_bounds__(64) void k_nl_from_slots(DevState d, CmList l, int check)
{
    const int c = blockIdx.x * 4 + (threadIdx.x >> 4);
    bool far = false;
    if (c < d.ncells) {
        const int n = d.natoms[c];
        const NlRep R = nl_rep(l, rep_of(d, c));
        double cx, cy, cz;
        cell_corner(grid_of(d, c), c, cx, cy, cz);
        for (int t = threadIdx.x & 15; t < n; t += 16) {
            const int s = c * CM_SLOTS + t;
            const int i = d.id[s];
            if (i < 0 || i >= R.nid) {
                atomicOr(l.flags, (unsigned)CM_NL_INVALID);
                continue;
            }
            const int id = R.id_base + i;
            l.slot[id] = s;
            // any build position in range until the first build
            if (!check) d.nlr[s] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(id));
            if (check) {
                const double x = cx + d.rx[s], y = cy + d.ry[s], z = cz + d.rz[s];
                const float4 q = d.nlr[s];
                l.pos[__float_as_int(q.w)] = CmPos{x, y, z};
                double ex = x - q.x, ey = y - q.y, ez = z - q.z;
                ex -= R.L0 * rint(ex * R.iL0);
                ey -= R.L1 * rint(ey * R.iL1);
                ez -= R.L2 * rint(ez * R.iL2);
                far = far || ex * ex + ey * ey + ez * ez > d.nl_half2;
            }
        }
    }
    if (__ballot(far) != 0 && threadIdx.x == 0) atomicOr(l.flags, (unsigned)CM_NL_REBUILD);
}

// ---------------------------------------------------------------------------
// groups of auxiliary cells and their halos

struct GroupGeo {
    int x0, y0, z0;  // first cell
    int lx, ly, lz;  // cells per dimension (fewer at the upper edges)
};

// Group gl of replica R (gl counted within the replica).
__device__ __forceinline__ GroupGeo group_geo(const NlRep &R, int gl)
{
    GroupGeo G;
    const int gx = gl % R.ng0, gy = (gl / R.ng0) % R.ng1, gz = gl / (R.ng0 * R.ng1);
    group_span(gx, R.na0, R.ng0, G.x0, G.lx);
    group_span(gy, R.na1, R.ng1, G.y0, G.ly);
    group_span(gz, R.na2, R.ng2, G.z0, G.lz);
    return G;
}

// Halo slot with coordinates (ux, uy, uz) in the group's (lx + 2) x (ly + 2)
// x (lz + 2) slots (x fastest): its cell (index, and coordinates jx, jy, jz)
// and the image shift that brings the cell next to the group.
__device__ __forceinline__ int halo_cell(const NlRep &R, const GroupGeo &G, int ux, int uy, int uz, double &sx,
                                         double &sy, double &sz, int &jx, int &jy, int &jz)
{
    jx = G.x0 - 1 + ux;
    jy = G.y0 - 1 + uy;
    jz = G.z0 - 1 + uz;
    sx = sy = sz = 0.0;
    if (jx < 0) { jx += R.na0; sx = -R.L0; } else if (jx >= R.na0) { jx -= R.na0; sx = R.L0; }
    if (jy < 0) { jy += R.na1; sy = -R.L1; } else if (jy >= R.na1) { jy -= R.na1; sy = R.L1; }
    if (jz < 0) { jz += R.na2; sz = -R.L2; } else if (jz >= R.na2) { jz -= R.na2; sz = R.L2; }
    return R.aux_base + jx + R.na0 * (jy + R.na1 * jz);  // global auxiliary cell
}

// A group's halo table in LDS: per slot its first atom in build order and
// its coordinates (ux | uy << 8 | uz << 16); off[] the slots' offsets in the
// staged halo (off[ns] = halo atoms).
struct GroupLds {
    int start[CM_NL_SLOTS], off[CM_NL_SLOTS + 1], uxyz[CM_NL_SLOTS];
    int ns, t0, t1;
};

// Fills S for group g (global index; G its geometry in replica R) (all
// threads of the block; starts and ends with a barrier), and, when slot_of
// is given, the slot of each staged halo entry (entries below hcap).
__device__ __forceinline__ void group_setup(const CmList &l, const NlRep &R, const GroupGeo &G, int g, GroupLds &S,
                                            CM_LDS uint8_t *slot_of = nullptr)
{
    __syncthreads();
    const int ns = (G.lx + 2) * (G.ly + 2) * (G.lz + 2);
    const int hx = G.lx + 2, hy = G.ly + 2;
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        const int ux = s % hx, uy = (s / hx) % hy, uz = s / (hx * hy);
        double sx, sy, sz;
        int jx, jy, jz;
        const int c = halo_cell(R, G, ux, uy, uz, sx, sy, sz, jx, jy, jz);
        S.start[s] = l.aux_start[c];
        S.uxyz[s] = ux | uy << 8 | uz << 16;
        S.off[s + 1] = l.aux_n[c];  // counts, scanned below
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int L = threadIdx.x;
        int carry = 0;
        for (int b = 0; b < ns; b += 64) {
            const int v = b + L < ns ? S.off[b + L + 1] : 0;
            int incl = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(incl, o, 64);
                if (L >= o) incl += u;
            }
            if (b + L < ns) S.off[b + L + 1] = carry + incl;
            carry += __shfl(incl, 63, 64);
        }
        if (L == 0) {
            S.off[0] = 0;
            S.ns = ns;
            S.t0 = l.gstart[g];
            S.t1 = l.gstart[g + 1];
        }
    }
    __syncthreads();
    if (slot_of) {
        for (int s = threadIdx.x; s < ns; s += blockDim.x) {
            const int h1 = min(S.off[s + 1], l.hcap);
            for (int h = S.off[s]; h < h1; h++) slot_of[h] = (uint8_t)s;
        }
        __syncthreads();
    }
}

typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v3u __attribute__((ext_vector_type(3)));

// Position of halo entry h (slot s) in its group's frame: the nearest image
// of the atom to its slot's cell, moved with the cell next to the group,
// relative to the frame origin o.  Atoms that crossed the periodic boundary
// since the build, and cells that appear twice in a small box's halo, are
// placed right.
__device__ __forceinline__ void frame_pos(const NlRep &R, const GroupGeo &G, int uxyz, const CmPos &p, double ox,
                                          double oy, double oz, double &x, double &y, double &z)
{
    double sx, sy, sz;
    int jx, jy, jz;
    halo_cell(R, G, uxyz & 0xff, (uxyz >> 8) & 0xff, uxyz >> 16, sx, sy, sz, jx, jy, jz);
    const double cx = (jx + 0.5) * R.ab0, cy = (jy + 0.5) * R.ab1, cz = (jz + 0.5) * R.ab2;
    double ex = p.x - cx, ey = p.y - cy, ez = p.z - cz;
    ex -= R.L0 * rint(ex * R.iL0);
    ey -= R.L1 * rint(ey * R.iL1);
    ez -= R.L2 * rint(ez * R.iL2);
    x = (cx + sx - ox) + ex;
    y = (cy + sy - oy) + ey;
    z = (cz + sz - oz) + ez;
}

// ---------------------------------------------------------------------------
// build (each kernel returns at once unless a rebuild was requested, so the
// launches that find no request cost only their dispatch: their grids are
// kept small, with strided loops)

// Each atom into its auxiliary cell (counts start at zero: the previous
// build's last block cleared them).  Atoms taken in slot order, 16 lanes per
// link cell (its atoms beyond 16 in further rounds), so that the positions
// and ids are read in whole lines; the order within an auxiliary cell is
// set by k_nl_sort.  Ids outside the replica's range are left out, as
// k_nl_from_slots leaves them out of the id -> slot map (and marks the
// lists invalid).
__global__ __launch_bounds__(256) void k_nl_key(DevState d, CmList l)
{
    if (!nl_rebuild(l)) return;
    const int sub = threadIdx.x & 15;
    const int c0 = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 4), dc = (int)((gridDim.x * blockDim.x) >> 4);
    for (int c = c0; c < d.ncells; c += dc) {
        const int n = d.natoms[c];
        if (sub >= n) continue;
        const NlRep R = nl_rep(l, rep_of(d, c));
        const Grid g = grid_of(d, c);
        for (int t = sub; t < n; t += 16) {
            const int s = c * CM_SLOTS + t;
            const int id = d.id[s];
            if (id < 0 || id >= R.nid) continue;
            int ax, ay, az;
            aux_coords(R, slot_pos(g, d, s), ax, ay, az);
            const int a = R.aux_base + ax + R.na0 * (ay + R.na1 * az);
            const int k = atomicAdd(&l.aux_cnt[a], 1);
            if (k < CM_NL_AUXCAP) l.aux_mem[a * CM_NL_AUXCAP + k] = R.id_base + id;
            else atomicOr(l.flags, (unsigned)CM_NL_OVERFLOW);
        }
    }
}

// Per group (one wave each; a group has at most 64 cells): its cells' atom
// counts (capped) and their exclusive offsets within the group in build
// order; the block that finishes last then turns the group totals into each
// group's first atom (an exclusive prefix sum, gstart) and count[0].
__global__ __launch_bounds__(256) void k_nl_gsum(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ int wsum[4];
    __shared__ int carry_s, last;
    const int t = threadIdx.x, L = t & 63, w = t >> 6;
    for (int g = blockIdx.x * 4 + w; g < l.ngroups; g += gridDim.x * 4) {
        const int q0 = l.group_q0[g], nq = l.group_q0[g + 1] - q0;
        int a = 0, v = 0;
        if (L < nq) {
            a = l.cell_order[q0 + L];
            v = min(l.aux_cnt[a], CM_NL_AUXCAP);
        }
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (L >= o) incl += u;
        }
        if (L < nq) {
            l.aux_start[a] = incl - v;
            l.aux_n[a] = v;
        }
        if (L == 63) l.gstart[g] = incl;  // the group's total, for now
    }
    __syncthreads();
    if (t == 0) {
        __threadfence();
        last = atomicAdd(&l.count[3], 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    int carry = 0;
    for (int base = 0; base < l.ngroups; base += 4 * 256) {
        int v[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = base + 4 * t + k;
            v[k] = q < l.ngroups ? l.gstart[q] : 0;
            s += v[k];
        }
        int incl = s;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (L >= o) incl += u;
        }
        if (L == 63) wsum[w] = incl;
        __syncthreads();
        if (t == 0) {
            int run = 0;
            for (int k = 0; k < 4; k++) {
                const int x = wsum[k];
                wsum[k] = run;
                run += x;
            }
            carry_s = run;
        }
        __syncthreads();
        int run = carry + wsum[w] + incl - s;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = base + 4 * t + k;
            if (q < l.ngroups) l.gstart[q] = run;
            run += v[k];
        }
        carry += carry_s;
        __syncthreads();
    }
    if (t == 0) {
        l.count[0] = carry;
        l.gstart[l.ngroups] = carry;
        l.count[3] = 0;
    }
}

// Members of each auxiliary cell in id order (deterministic lists), placed in
// the build order; each cell's first atom made global (its group's first
// atom added).  One wave per cell, four waves per block.
__global__ __launch_bounds__(256) void k_nl_sort(CmList l)
{
    if (!nl_rebuild(l)) return;
    __shared__ int mem_s[4][CM_NL_AUXCAP];
    const int L = threadIdx.x & 63, w = threadIdx.x >> 6;
    int *mem = mem_s[w];
    for (int a = blockIdx.x * 4 + w; a < l.naux; a += gridDim.x * 4) {
        const int g = l.aux_group[a];
        const int start = l.aux_start[a] + l.gstart[g];
        nl_wave_sync();
        if (L == 0) l.aux_start[a] = start;
        const int n = l.aux_n[a];
        if (n == 0) continue;
        const int me = L < n ? l.aux_mem[a * CM_NL_AUXCAP + L] : 0;
        if (L < n) mem[L] = me;
        nl_wave_sync();
        int rank = 0;
        for (int u = 0; u < n; u++) rank += mem[u] < me;
        nl_wave_sync();
        if (L < n) l.order[start + rank] = me;
    }
}

// Bytes of the staged halo entries' slot map (one byte per entry).
__host__ __device__ constexpr int nl_slotmap_bytes(int hcap) { return (hcap + 15) / 16 * 16; }

// One block per group (the passes' launch shape): the halo is staged in LDS
// in the group's frame, in single precision; each home atom (a thread) tests
// the atoms of its own 27-cell stencil, as 9 runs of 3 x-adjacent cells
// (contiguous in the halo), id order within a cell, and lists the halo
// indices within cutoff + skin.  Entries beyond the cutoff at the build go
// to the end of the list (up to tailq of them per thread are held in LDS
// until the stencil is done, as many as fit without fewer resident blocks
// per CU; any more stay in place): in a lattice those
// are the shell just beyond the cutoff, which the passes test and reject,
// and at the same list positions across a wave a rejection by every lane
// skips the pair term's evaluation.  The order of the entries only orders
// the passes' sums.  The single-precision test compares with
// (cutoff + skin)^2 x (1 + 4e-6), more than its rounding (frame coordinates
// within 20 A: at most 1.5e-6 of r^2), so no pair within cutoff + skin is
// left out; the few extra pairs are rejected by the passes' cutoff test.
// The last block to finish closes the build: overflow (a list past kmax, a
// halo past hcap) -> lists invalid, rebuild request cleared, counts cleared
// for the next build.
__global__ __launch_bounds__(1024) void k_nl_build(DevState d, CmList l)
{
    if (!nl_rebuild(l)) return;
    extern __shared__ float4 hl4[];  // [hcap] frame x, y, z; then the slot map
    CM_LDS uint8_t *slot_of = (CM_LDS uint8_t *)((CM_LDS char *)hl4 + (size_t)l.hcap * 16);
    // then the tail queue: entry j of thread x at [j * blockDim.x + x]
    CM_LDS uint16_t *tq = (CM_LDS uint16_t *)(slot_of + nl_slotmap_bytes(l.hcap)) + threadIdx.x;
    __shared__ GroupLds S;
    __shared__ int last;
    bool ovf = false;
    // one group per block, XCD-contiguous as in the passes
    const int nb = gridDim.x, qb = nb / 8, rem = nb % 8;
    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    const int g = nb < 8 ? (int)blockIdx.x : xcd * qb + min(xcd, rem) + idx;
    const NlRep R = nl_rep(l, l.group_rep[g]);
    const GroupGeo G = group_geo(R, g - R.group_base);
    const float rl2 = (float)(R.rl2 * (1.0 + 4e-6)), rc2 = (float)R.rc2;
    // the group's replica: the grid of every atom it stages
    const Grid gr = d.ngrids == 1 ? d.g0 : d.grids[l.group_rep[g]];
    group_setup(l, R, G, g, S, slot_of);
    const int H = S.off[S.ns];
    if (H > l.hcap) {  // the same in every thread of the block
        ovf = true;
    } else {
        const double ox = (G.x0 + 0.5 * G.lx) * R.ab0, oy = (G.y0 + 0.5 * G.ly) * R.ab1,
                     oz = (G.z0 + 0.5 * G.lz) * R.ab2;
        constexpr int SB = 2;  // staged entries per thread in flight (each: slot, r, its cell's corner)
        for (int h0 = threadIdx.x; h0 < H; h0 += SB * blockDim.x) {
            int sl[SB], q[SB], id[SB];
#pragma unroll
            for (int k = 0; k < SB; k++) {
                const int h = min(h0 + k * (int)blockDim.x, H - 1);
                sl[k] = slot_of[h];
                q[k] = S.start[sl[k]] + (h - S.off[sl[k]]);
            }
#pragma unroll
            for (int k = 0; k < SB; k++) id[k] = l.slot[l.order[q[k]]];
            CmPos p[SB];
#pragma unroll
            for (int k = 0; k < SB; k++) p[k] = slot_pos(gr, d, id[k]);
#pragma unroll
            for (int k = 0; k < SB; k++) {
                const int h = h0 + k * (int)blockDim.x;
                if (h >= H) break;
                double x, y, z;
                frame_pos(R, G, S.uxyz[sl[k]], p[k], ox, oy, oz, x, y, z);
                hl4[h] = make_float4((float)x, (float)y, (float)z, 0.0f);
            }
        }
        __syncthreads();
        const int shx = G.lx + 2, shy = G.ly + 2;
        for (int t = S.t0 + threadIdx.x; t < S.t1; t += blockDim.x) {
            // the positions from the slots (as k_nl_key took them), kept in
            // build order for the passes and the drift
            const int ts = l.slot[l.order[t]];
            const CmPos p = slot_pos(gr, d, ts);
            l.pos[t] = p;
            int ax, ay, az;
            aux_coords(R, p, ax, ay, az);
            // home slot: the atom's own cell, an interior slot of the halo
            const int cx = ax - G.x0 + 1, cy = ay - G.y0 + 1, cz = az - G.z0 + 1;
            const int hs = cx + shx * (cy + shy * cz);
            const int self = S.off[hs] + (t - S.start[hs]);
            l.selfh[t] = (uint16_t)self;
            const float4 pi = hl4[self];
            // the 27 stencil cells as 9 runs of three x-adjacent halo slots,
            // whose atoms are contiguous in the halo; entries gathered 8 at a
            // time and stored as one 16-byte group
            int k = 0, nt = 0;
            unsigned long long lo = 0, hi = 0;
            auto push = [&](int h) {
                const int c8 = k & 7;
                if (c8 < 4) lo |= (unsigned long long)h << (16 * c8);
                else hi |= (unsigned long long)h << (16 * (c8 - 4));
                if (c8 == 7 && k < l.kmax) {
                    ulonglong2 w;
                    w.x = lo;
                    w.y = hi;
                    *(ulonglong2 *)&l.nbl[nl_entry(k - 7, t, l.n)] = w;
                    lo = hi = 0;
                }
                k++;
            };
            for (int e = 0; e < 9; e++) {
                const int s0 = (cx - 1) + shx * ((cy + e / 3 - 1) + shy * (cz + e % 3 - 1));
                const int h1 = S.off[s0 + 3];
                for (int h = S.off[s0]; h < h1; h++) {
                    const float4 qh = hl4[h];
                    const float dx = pi.x - qh.x, dy = pi.y - qh.y, dz = pi.z - qh.z;
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    if (r2 <= rl2 && h != self) {
                        if (r2 > rc2 && nt < l.tailq) tq[(nt++) * blockDim.x] = (uint16_t)h;
                        else push(h);
                    }
                }
            }
            for (int j = 0; j < nt; j++) push(tq[j * blockDim.x]);
            if ((k & 7) != 0 && k <= l.kmax) {
                ulonglong2 w;
                w.x = lo;
                w.y = hi;
                *(ulonglong2 *)&l.nbl[nl_entry(k & ~7, t, l.n)] = w;
            }
            l.ncnt[t] = min(k, l.kmax);
            ovf |= k > l.kmax;
            d.nlr[ts] = make_float4((float)p.x, (float)p.y, (float)p.z, __int_as_float(t));
        }
    }
    // the last block to finish closes the build.  The ticket word carries
    // the overflow too (bit 30, set before the block's ticket is taken): two
    // atomics on one address by one thread stay in order, so the last
    // ticket sees every overflow without a fence per block (each fence
    // writes back the L2 on this part, and the other blocks' lists are read
    // only by later kernels)
    const int bovf = __syncthreads_or((int)ovf);
    if (threadIdx.x == 0) {
        if (bovf) atomicOr(&l.count[2], 1 << 30);
        const int old = atomicAdd(&l.count[2], 1);
        last = (old & 0xffffff) == (int)gridDim.x - 1 ? 1 + ((old >> 30) & 1) : 0;
    }
    __syncthreads();
    if (!last) return;
    for (int a = threadIdx.x; a < l.naux; a += blockDim.x) l.aux_cnt[a] = 0;
    if (threadIdx.x == 0) {
        l.count[2] = 0;
        l.count[1] += 1;  // builds so far
        const unsigned f = atomicOr(l.flags, 0u);  // and the earlier kernels' overflow
        if ((f & CM_NL_OVERFLOW) || last == 2) atomicOr(l.flags, (unsigned)CM_NL_INVALID);
        else atomicAnd(l.flags, ~(unsigned)CM_NL_INVALID);
        atomicAnd(l.flags, ~(unsigned)(CM_NL_OVERFLOW | CM_NL_REBUILD));
    }
}

// ---------------------------------------------------------------------------
// force passes over the lists (return at once when the lists are invalid;
// the link-cell kernels then run instead)

// Non-empty stencil cells of each link cell: the reference adds the kinetic
// virial term once per non-empty neighbour cell (src-flat/ljForce.c:175).
// Cells of the replicas whose groups start at g0 or later (the
// Lennard-Jones ones); the 27 entries of stencil_cell, duplicates included.
__global__ void k_nl_nonempty(DevState d, CmList l, int g0)
{
    if (l.flags[0] & CM_NL_INVALID) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncells) return;
    if (g0 > 0 && l.rep[rep_of(d, c)].group_base < g0) return;
    const Grid g = grid_of(d, c);
    int n = 0;
    if (g.ncells == 1) {
        n = d.natoms[c] > 0;
    } else {
        int ix, iy, iz;
        cell_xyz(g, c, ix, iy, iz);
        for (int ox = -1; ox <= 1; ox++)
            for (int oy = -1; oy <= 1; oy++)
                for (int oz = -1; oz <= 1; oz++) {
                    const int b = cell_wrap(g, ix + ox, iy + oy, iz + oz);
                    n += b >= 0 && d.natoms[b] > 0;
                }
    }
    l.nonempty[c] = n;
}

// Chebyshev passes: N pairs of a list chunk evaluated together, each series
// as N interleaved Clenshaw recurrences (src-flat/cheby.c:276-310) that
// share every coefficient load: evaluated pair by pair, each pair reloaded
// all coefficients and ran one serial recurrence at a time.  x[u] is
// clamped below at a as in cheb64; the series' argument is formed with
// 1 / (b - a) computed once rather than a division per pair.
template <int N>
__device__ __forceinline__ void cheb64_n(const double *cg, int m8, double a, double b, const double (&x)[N],
                                         double (&out)[N])
{
    const CM_CONST double *c = (const CM_CONST double *)cg;
    const double inv = 1.0 / (b - a);
    double y2[N], d[N], dd[N];
#pragma unroll
    for (int u = 0; u < N; u++) {
        const double xu = x[u] < a ? a : x[u];
        y2[u] = 2.0 * ((2.0 * xu - a - b) * inv);
        d[u] = dd[u] = 0.0;
    }
    double cc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cc[k] = c[m8 - 8 + k];
    for (int jb = m8 - 8; jb >= 0; jb -= 8) {
        double cn[8];
        const int jn = jb >= 8 ? jb - 8 : 0;  // the next block (block 0 again after the last)
#pragma unroll
        for (int k = 0; k < 8; k++) cn[k] = c[jn + k];
#pragma unroll
        for (int k = 7; k >= 0; k--) {
            if (jb + k > 0) {
#pragma unroll
                for (int u = 0; u < N; u++) {
                    const double sv = d[u];
                    d[u] = y2[u] * d[u] - dd[u] + cc[k];
                    dd[u] = sv;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) cc[k] = cn[k];
    }
#pragma unroll
    for (int u = 0; u < N; u++) out[u] = 0.5 * y2[u] * d[u] - dd[u] + 0.5 * cc[0];
}

// The same in single precision (cheb32's recurrence) for the values of the
// mixed density pass.
template <int N>
__device__ __forceinline__ void cheb32_n(const float *cg, int m8, float a, float b, const float (&x)[N],
                                         float (&out)[N])
{
    const CM_CONST float *c = (const CM_CONST float *)cg;
    const float inv = 1.0f / (b - a);
    float y2[N], d[N], dd[N];
#pragma unroll
    for (int u = 0; u < N; u++) {
        const float xu = fmaxf(x[u], a);
        y2[u] = 2.0f * ((2.0f * xu - a - b) * inv);
        d[u] = dd[u] = 0.0f;
    }
    float cc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cc[k] = c[m8 - 8 + k];
    for (int jb = m8 - 8; jb >= 0; jb -= 8) {
        float cn[8];
        const int jn = jb >= 8 ? jb - 8 : 0;
#pragma unroll
        for (int k = 0; k < 8; k++) cn[k] = c[jn + k];
#pragma unroll
        for (int k = 7; k >= 0; k--) {
            if (jb + k > 0) {
#pragma unroll
                for (int u = 0; u < N; u++) {
                    const float sv = d[u];
                    d[u] = fmaf(y2[u], d[u], cc[k] - dd[u]);
                    dd[u] = sv;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) cc[k] = cn[k];
    }
#pragma unroll
    for (int u = 0; u < N; u++) out[u] = fmaf(0.5f * y2[u], d[u], 0.5f * cc[0] - dd[u]);
}

// Compact table rows a pass reads (density: phi and rho, 8 values per
// interval, 64 bytes or 40 in mixed precision; embedding force: rho', 2
// doubles), staged per block in LDS: every
// pair reads a different interval, and from LDS these reads do not load the
// vector L1.
static inline size_t nl_tab_bytes(const PotDev &P, int kind, int mixed)
{
    const size_t b = (size_t)(P.n[0] + 1) * (kind == CM_KIND_EAM1 ? (mixed ? 40 : 64) : 16);
    return (b + 15) / 16 * 16;  // the halo after the rows stays 16-byte aligned
}
// Staged halo bytes per atom: group-frame fixed point x, y, z and float F'
// (mixed); group-frame double x, y, z and, for the embedding-force pass, F'
// (FP64).
__host__ __device__ constexpr int nl_halo_bytes(int kind, int mixed)
{
    return mixed ? (kind == CM_KIND_EAM3 ? 16 : 12) : (kind == CM_KIND_EAM3 ? 32 : 24);
}
// Mixed passes: 32-bit words per staged halo entry (x, y, z and, for the
// embedding-force pass, F').  The density and LJ passes' 12-byte entries put
// a wave's random neighbour reads on distinct banks more often than a
// 16-byte stride (3h mod 64 against 4h mod 64): 4 % off the density pass,
// though each read takes two LDS instructions.  The embedding-force pass
// keeps one 16-byte read (a 20-byte stride measured 2 % slower there).
template <int KIND> __host__ __device__ constexpr int nl_halo_words() { return KIND == CM_KIND_EAM3 ? 4 : 3; }

// What a pass reads of home atom t before its pair terms: its atom index,
// its own halo entry, its list length and the first 8-entry chunk of its
// list (lane sub's, with several lanes per atom).  Read by nl_head before the
// block stages its halo, so that these loads wait alongside the staging.
struct NlHead {
    int i, self, cnt;
    v4u e;
};

__device__ __forceinline__ NlHead nl_head(const CmList &l, int t, int sub)
{
    NlHead h;
    h.i = l.order[t];
    h.self = l.selfh[t];
    h.cnt = l.ncnt[t];
    h.e = ((const CM_GLOBAL v4u *)l.nbl)[(size_t)min(sub, l.kmax / 8 - 1) * l.n + t];
    return h;
}

// One home atom t over its list (hd: its head, nl_head(l, t, sub)).  TB: the
// table pointers (compact rows in LDS when COMPACT, else the full tables in
// global memory); hq (mixed) or hx, hy, hz, hf (FP64): the staged halo,
// which holds the atom itself too.
template <int KIND, int EVAL, int MIXED, int COMPACT, int LANES, class TB>
__device__ __forceinline__ void nl_atom(const DevState &d, const PotDev &P, const CmList &l, const TB &T, int t,
                                        int sub, const NlHead &hd, const CM_LDS v4u *hq, const CM_LDS double *hx,
                                        const CM_LDS double *hy, const CM_LDS double *hz, const CM_LDS double *hf,
                                        double qscale, double qinv)
{
    const int i = hd.i;
    const int s = l.slot[i];
    const int self = hd.self;
    v4u qi = {0, 0, 0, 0};
    double xi = 0.0, yi = 0.0, zi = 0.0, fia = 0.0;
    constexpr int HW = nl_halo_words<KIND>();
    const CM_LDS unsigned *hu = (const CM_LDS unsigned *)hq;
    if (MIXED) {
        if constexpr (HW == 4) {
            qi = hq[self];
        } else {
            qi = v4u{hu[HW * self], hu[HW * self + 1], hu[HW * self + 2], 0u};
        }
        fia = (double)__uint_as_float(qi.w);
    } else {
        xi = hx[self];
        yi = hy[self];
        zi = hz[self];
        if (KIND == CM_KIND_EAM3) fia = hf[self];
    }
    if (KIND != CM_KIND_EAM3) fia = 0.0;
    const double rc2 = P.rc2;
    // mixed passes over compact rows work in fixed-point units (qinv is a
    // power of two: the cutoff test and the force terms are the same as in
    // angstrom; the table coordinate takes the unit scale)
    constexpr bool UNITS = MIXED && COMPACT && KIND != CM_KIND_LJ;
    const double rc2u = rc2 * (qscale * qscale);
    const double iu = qinv * P.invdx[0], cu = -P.x0[0] * P.invdx[0];
    const int cnt = hd.cnt;
    Acc A = {0, 0, 0, 0, 0, 0};
    int npair = 0;
    // 8 entries per 16-byte load, then the 8 halo reads from LDS, all issued
    // before the first pair is evaluated.  Whole chunks of 8 first (the next
    // chunk's load issued before this one is processed), then the partial
    // last chunk, whose entries past cnt are not this build's: entry 0 of
    // the halo is read in their place and they count as out of range.
    const CM_GLOBAL v4u *row = (const CM_GLOBAL v4u *)l.nbl + t;
    auto chunk = [&](const v4u &e, int k0, auto tail) {
        constexpr bool TAIL = decltype(tail)::value;
        const unsigned ew[4] = {e.x, e.y, e.z, e.w};
        int hj[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            hj[u] = (int)((ew[u >> 1] >> (16 * (u & 1))) & 0xffffu);
            if (TAIL && k0 + u >= cnt) hj[u] = 0;
        }
        // separation of pair u (angstrom) from the staged halo
        auto sep = [&](int u, double &dx, double &dy, double &dz) {
            if constexpr (MIXED) {
                const CM_LDS unsigned *q = hu + HW * hj[u];
                dx = (double)(int)(qi.x - q[0]) * qinv;
                dy = (double)(int)(qi.y - q[1]) * qinv;
                dz = (double)(int)(qi.z - q[2]) * qinv;
            } else {
                dx = xi - hx[hj[u]];
                dy = yi - hy[hj[u]];
                dz = zi - hz[hj[u]];
            }
        };
        if constexpr (EVAL == COMD_EVAL_CHEB && KIND != CM_KIND_LJ) {
            // the chunk's pairs NB at a time (cheb64_n): r of each, the
            // series, then the sums in list order, the separations read
            // again from the halo.  Four in FP64, two in mixed precision,
            // where the registers of more recurrences cost more resident
            // waves than the wider interleave gains (measured, 256 000
            // atoms: FP64 passes 0.333 + 0.147 -> 0.240 + 0.118 ms, mixed
            // 0.245 + 0.113 -> 0.186 + 0.096 ms).
            constexpr int NB = MIXED ? 2 : 4;
#pragma unroll
            for (int h = 0; h < 8; h += NB) {
                double rr[NB];
                unsigned msk = 0;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const int u = h + j;
                    double r2;
                    bool in;
                    if constexpr (MIXED) {
                        const CM_LDS unsigned *q = hu + HW * hj[u];
                        const double dxu = (double)(int)(qi.x - q[0]), dyu = (double)(int)(qi.y - q[1]),
                                     dzu = (double)(int)(qi.z - q[2]);
                        const double r2u = dxu * dxu + dyu * dyu + dzu * dzu;
                        in = KIND == CM_KIND_EAM3 ? (r2u < rc2u) : (r2u <= rc2u);
                        r2 = r2u * (qinv * qinv);
                    } else {
                        double dx, dy, dz;
                        sep(u, dx, dy, dz);
                        r2 = dx * dx + dy * dy + dz * dz;
                        in = KIND == CM_KIND_EAM3 ? (r2 < rc2) : (r2 <= rc2);
                    }
                    in = in && (!TAIL || k0 + u < cnt);
                    rr[j] = in ? r2 * (MIXED ? rsqrt_mixed(r2) : rsqrt64(r2)) : 0.0;  // outside: the lower end
                    msk |= (unsigned)in << j;
                }
                if (msk == 0) continue;
                npair += __builtin_popcount(msk);
                const int m = P.cstride;
                double v[NB];
                if constexpr (KIND == CM_KIND_EAM1) {
                    if constexpr (MIXED) {
                        float rf[NB], vf[NB];
#pragma unroll
                        for (int j = 0; j < NB; j++) rf[j] = (float)rr[j];
                        cheb32_n<NB>(P.cheb32 + 0 * m, m, (float)P.ca[0], (float)P.cb[0], rf, vf);
#pragma unroll
                        for (int j = 0; j < NB; j++)
                            if ((msk >> j) & 1u) A.e += 0.5 * (double)vf[j];
                        cheb32_n<NB>(P.cheb32 + 1 * m, m, (float)P.ca[1], (float)P.cb[1], rf, vf);
#pragma unroll
                        for (int j = 0; j < NB; j++)
                            if ((msk >> j) & 1u) A.rho += (double)vf[j];
                    } else {
                        cheb64_n<NB>(P.cheb + 0 * m, m, P.ca[0], P.cb[0], rr, v);
#pragma unroll
                        for (int j = 0; j < NB; j++)
                            if ((msk >> j) & 1u) A.e += 0.5 * v[j];
                        cheb64_n<NB>(P.cheb + 1 * m, m, P.ca[1], P.cb[1], rr, v);
#pragma unroll
                        for (int j = 0; j < NB; j++)
                            if ((msk >> j) & 1u) A.rho += v[j];
                    }
                    cheb64_n<NB>(P.cheb + 3 * m, m, P.ca[0], P.cb[0], rr, v);  // phi', FP64 in both modes
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        if ((msk >> j) & 1u) {
                            double dx, dy, dz;
                            sep(h + j, dx, dy, dz);
                            const double sc = v[j] / rr[j];
                            A.x += sc * dx;
                            A.y += sc * dy;
                            A.z += sc * dz;
                        }
                    }
                } else {
                    cheb64_n<NB>(P.cheb + 4 * m, m, P.ca[1], P.cb[1], rr, v);  // rho', FP64 in both modes
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        if ((msk >> j) & 1u) {
                            double dx, dy, dz, fij;
                            sep(h + j, dx, dy, dz);
                            if constexpr (MIXED) fij = (double)__uint_as_float(hq[hj[h + j]].w);
                            else fij = hf[hj[h + j]];
                            const double g = (fia + fij) * v[j] / rr[j];
                            A.x += g * dx;
                            A.y += g * dy;
                            A.z += g * dz;
                            A.vir += g * dx * dx;
                        }
                    }
                }
            }
            return;
        }
        if constexpr (MIXED) {
            // x, y, z only unless F' is needed
            v4u qj[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                if constexpr (HW == 4) {
                    qj[u] = hq[hj[u]];
                } else {
                    const CM_LDS unsigned *e = hu + HW * hj[u];
                    qj[u] = v4u{e[0], e[1], e[2], 0u};
                }
            }
            float ef = 0.0f, rf = 0.0f;  // the chunk's FP32 value sums (density pass)
#pragma unroll
            for (int u = 0; u < 8; u++) {
                // fixed-point differences (modulo 2^32, exact for pairs on
                // the lists, which are well within range)
                const double dxu = (double)(int)(qi.x - qj[u].x);
                const double dyu = (double)(int)(qi.y - qj[u].y);
                const double dzu = (double)(int)(qi.z - qj[u].z);
                const double fij = (double)__uint_as_float(qj[u].w);
                const double r2u = dxu * dxu + dyu * dyu + dzu * dzu;
                const bool in = (!TAIL || k0 + u < cnt) && (KIND == CM_KIND_EAM3 ? (r2u < rc2u) : (r2u <= rc2u));
                if (in) {
                    if constexpr (UNITS) {
                        pair_term_units<KIND>(P, T, dxu, dyu, dzu, r2u, iu, cu, fia, fij, A, ef, rf);
                    } else {
                        const double dx = dxu * qinv, dy = dyu * qinv, dz = dzu * qinv;
                        const double r2 = r2u * (qinv * qinv);
                        if constexpr (COMPACT) pair_term_compact<KIND, MIXED>(P, T, dx, dy, dz, r2, fia, fij, A);
                        else pair_term<KIND, EVAL, MIXED>(P, T, dx, dy, dz, r2, fia, fij, A);
                    }
                    npair++;
                }
            }
            if (UNITS && KIND == CM_KIND_EAM1) {
                A.e += (double)ef;
                A.rho += (double)rf;
            }
        } else {
            double xj[8], yj[8], zj[8], fj[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                xj[u] = hx[hj[u]];
                yj[u] = hy[hj[u]];
                zj[u] = hz[hj[u]];
                fj[u] = KIND == CM_KIND_EAM3 ? hf[hj[u]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double dx = xi - xj[u], dy = yi - yj[u], dz = zi - zj[u];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const bool in = (!TAIL || k0 + u < cnt) && (KIND == CM_KIND_EAM3 ? (r2 < rc2) : (r2 <= rc2));
                if (in) {
                    if constexpr (COMPACT) pair_term_compact<KIND, MIXED>(P, T, dx, dy, dz, r2, fia, fj[u], A);
                    else pair_term<KIND, EVAL, MIXED>(P, T, dx, dy, dz, r2, fia, fj[u], A);
                    npair++;
                }
            }
        }
    };
    // LANES: lane sub of the atom's l.lanes takes chunks sub, sub + lanes,
    // ... (else one lane takes them all); the next chunk's index clamped to
    // the list capacity: always a valid load
    const int clast = l.kmax / 8 - 1, step = LANES ? l.lanes : 1;
    if (!LANES) sub = 0;
    v4u e = hd.e;
    int k0 = 8 * sub;
    for (; k0 + 8 <= cnt; k0 += 8 * step) {
        const v4u en = row[(size_t)min((k0 >> 3) + step, clast) * l.n];
        chunk(e, k0, std::false_type());
        e = en;
    }
    if (k0 < cnt) chunk(e, k0, std::true_type());
    if (UNITS) A.vir *= qinv;
    // the atom's lanes (adjacent, the same t) add their sums; the first
    // writes the results
    for (int o = 1; LANES && o < step; o <<= 1) {
        A.x += __shfl_xor(A.x, o);
        A.y += __shfl_xor(A.y, o);
        A.z += __shfl_xor(A.z, o);
        A.e += __shfl_xor(A.e, o);
        A.rho += __shfl_xor(A.rho, o);
        A.vir += __shfl_xor(A.vir, o);
        npair += __shfl_xor(npair, o);
    }
    if (LANES && sub != 0) return;
    if (KIND == CM_KIND_EAM1) {
        // pass 2, src-flat/eam.c:576-594
        double fv, fp;
        if (EVAL == COMD_EVAL_CHEB) {
            fv = cheb64(P.cheb + 2 * P.cstride, P.cstride, P.ca[2], P.cb[2], A.rho);
            fp = cheb64(P.cheb + 5 * P.cstride, P.cstride, P.ca[2], P.cb[2], A.rho);
        } else {
            interp64(P.tab[2], P.x0[2], P.xn[2], P.invdx[2], P.n[2], A.rho, fv, fp);
        }
        // the pair forces stay in build order (coalesced) for the
        // embedding-force pass, which writes the sum to the slots once
        l.f1[3L * t + 0] = A.x;
        l.f1[3L * t + 1] = A.y;
        l.f1[3L * t + 2] = A.z;
        d.fe[s] = A.e + fv;
        d.rho[s] = A.rho;
        d.fi[s] = fp;
        d.nnb[s] = npair;
        l.fq[t] = fp;  // staged with the position by the embedding-force pass (coalesced: t is the thread's)
    } else if (KIND == CM_KIND_EAM3) {
        // the three loads before any store: the slot arrays and the
        // build-order pair forces may alias as far as the compiler knows,
        // which would otherwise put a load's wait after each store
        const double px = l.f1[3L * t + 0], py = l.f1[3L * t + 1], pz = l.f1[3L * t + 2];
        d.fx[s] = px + A.x;
        d.fy[s] = py + A.y;
        d.fz[s] = pz + A.z;
        d.vir[s] = A.vir;
    } else {
        const double px = d.px[s];
        const int ne = l.nonempty[s / CM_SLOTS];  // read before the stores, as above
        d.fx[s] = P.lj_fsign * A.x;
        d.fy[s] = P.lj_fsign * A.y;
        d.fz[s] = P.lj_fsign * A.z;
        d.fe[s] = A.e;
        d.vir[s] = A.vir - (double)ne * (px * px / P.mass_amu);
        d.nnb[s] = npair;
    }
}

// A block per group: the halo is staged in LDS in the group's frame (origin
// at the centre of the group's cells), then each thread takes home atoms.
// TLDS: the compact table rows are staged in LDS (1), or the full tables are
// read from global memory (0); a compile-time choice, so the row reads are
// LDS or global loads.  Dynamic LDS: [rows (TLDS)] [halo, hcap atoms]
// [slot of each halo entry, hcap bytes].
// Blocks of at most 768 threads (a group's home atoms); the Chebyshev passes
// declare that bound, so that their four-pair recurrences (cheb64_n) get the
// registers three waves per SIMD allow.
template <int KIND, int EVAL, int MIXED, int TLDS, int LANES>
__global__ __attribute__((amdgpu_flat_work_group_size(1, EVAL == COMD_EVAL_CHEB ? 768 : 1024)))
__attribute__((amdgpu_waves_per_eu(MIXED && KIND != CM_KIND_LJ && TLDS ? 6 : 1)))
void k_nl_pair(DevState d, PotDev P, CmList l, int tab_bytes, int gbase)
{
    extern __shared__ __align__(16) char nls[];
    __shared__ GroupLds S;
    if (l.flags[0] & CM_NL_INVALID) return;
    CM_LDS char *halo = (CM_LDS char *)nls + tab_bytes;
    CM_LDS uint8_t *slot_of = (CM_LDS uint8_t *)(halo + (size_t)l.hcap * nl_halo_bytes(KIND, MIXED));
    CM_LDS v4u *hq = (CM_LDS v4u *)halo;
    CM_LDS double *hx = (CM_LDS double *)halo, *hy = hx + l.hcap, *hz = hy + l.hcap, *hf = hz + l.hcap;
    CompactLds TL;
    TL.d64 = TL.r64 = nullptr;
    TL.dm = nullptr;
    if (TLDS) {
        if (MIXED && KIND == CM_KIND_EAM1) {
            const int nv = (P.n[0] + 1) * 10;
            CM_LDS float *dst = (CM_LDS float *)nls;
            for (int k = threadIdx.x; k < nv; k += blockDim.x) dst[k] = P.cdm[k];
            TL.dm = dst;
            TL.d64 = (CM_LDS double *)dst;
        } else {
            const int nv = (P.n[0] + 1) * (KIND == CM_KIND_EAM1 ? 8 : 2);
            const double *src = KIND == CM_KIND_EAM1 ? P.cd64 : P.cr64;
            CM_LDS double *dst = (CM_LDS double *)nls;
            for (int k = threadIdx.x; k < nv; k += blockDim.x) dst[k] = src[k];
            TL.d64 = TL.r64 = dst;
        }
    }
    const TabsGlobal TG = global_tabs(P);
    // one group per block (the dispatcher balances the groups' uneven work
    // over the CUs); blocks are dealt round-robin over the 8 XCDs, and each
    // XCD takes a contiguous range of groups, so that neighbouring groups,
    // which share halo cells, stage them through the same L2
    {
        const int nb = gridDim.x, qb = nb / 8, rem = nb % 8;
        const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
        const int g = gbase + (nb < 8 ? (int)blockIdx.x : xcd * qb + min(xcd, rem) + idx);
        const NlRep R = nl_rep(l, l.group_rep[g]);
        const GroupGeo G = group_geo(R, g - R.group_base);
        // LANES: l.lanes adjacent lanes per home atom.  The head of the
        // thread's first atom is read now, its wait shared with the staging
        const int lanes = LANES ? l.lanes : 1;
        const int sub = (int)threadIdx.x % lanes;
        const int per = (int)blockDim.x / lanes;
        const int ta = l.gstart[g] + (int)threadIdx.x / lanes, tb = l.gstart[g + 1];
        NlHead hd = {0, 0, 0, {0u, 0u, 0u, 0u}};
        if (ta < tb) hd = nl_head(l, ta, sub);
        group_setup(l, R, G, g, S, slot_of);  // its first barrier also covers the table staging
        const double ox = (G.x0 + 0.5 * G.lx) * R.ab0, oy = (G.y0 + 0.5 * G.ly) * R.ab1,
                     oz = (G.z0 + 0.5 * G.lz) * R.ab2;
        const int H = min(S.off[S.ns], l.hcap);  // never above hcap while the lists are valid
        // up to SB entries per thread in flight: their id and position loads
        // are issued together (a thread stages about hcap / blockDim of
        // them, up to 3.8 with 4 x 4 x 3 groups; one at a time in the mixed
        // embedding-force pass and four in the other mixed table and
        // Lennard-Jones passes, which keep to 80 registers for six waves
        // per SIMD with the atom's head held meanwhile)
        constexpr int SB = MIXED ? (KIND == CM_KIND_EAM3 ? 1 : EVAL == COMD_EVAL_CHEB ? 6 : 4) : 6;
        for (int h0 = threadIdx.x; h0 < H; h0 += SB * blockDim.x) {
            // slots first, then the position loads (build order: a
            // halo slot's atoms are consecutive)
            int sl[SB], q[SB];
#pragma unroll
            for (int k = 0; k < SB; k++) {
                const int h = min(h0 + k * (int)blockDim.x, H - 1);
                sl[k] = slot_of[h];
                q[k] = S.start[sl[k]] + (h - S.off[sl[k]]);
            }
            CmPos p[SB];
            double fp[SB];
#pragma unroll
            for (int k = 0; k < SB; k++) {
                p[k] = load_pos(l.pos, q[k]);
                fp[k] = KIND == CM_KIND_EAM3 ? ((const CM_GLOBAL double *)l.fq)[q[k]] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < SB; k++) {
                const int h = h0 + k * (int)blockDim.x;
                if (h >= H) break;
                double x, y, z;
                frame_pos(R, G, S.uxyz[sl[k]], p[k], ox, oy, oz, x, y, z);
                if (MIXED) {
                    const float fj = (float)fp[k];
                    v4u q;
                    q.x = (unsigned)(int)rint(x * R.qscale);
                    q.y = (unsigned)(int)rint(y * R.qscale);
                    q.z = (unsigned)(int)rint(z * R.qscale);
                    q.w = __float_as_uint(fj);
                    if constexpr (nl_halo_words<KIND>() == 4) {
                        hq[h] = q;
                    } else {
                        CM_LDS unsigned *e = (CM_LDS unsigned *)hq + nl_halo_words<KIND>() * h;
                        e[0] = q.x;
                        e[1] = q.y;
                        e[2] = q.z;
                    }
                } else {
                    hx[h] = x;
                    hy[h] = y;
                    hz[h] = z;
                    if (KIND == CM_KIND_EAM3) hf[h] = fp[k];
                }
            }
        }
        __syncthreads();
        // the first atom with its head read above, then any further ones
        // (more home atoms than threads) with theirs read in turn
        auto atom = [&](int t, const NlHead &h) {
            if (TLDS)
                nl_atom<KIND, EVAL, MIXED, 1, LANES>(d, P, l, TL, t, sub, h, hq, hx, hy, hz, hf, R.qscale, R.qinv);
            else
                nl_atom<KIND, EVAL, MIXED, 0, LANES>(d, P, l, TG, t, sub, h, hq, hx, hy, hz, hf, R.qscale, R.qinv);
        };
        if (ta < tb) atom(ta, hd);
        for (int t = ta + per; t < tb; t += per) atom(t, nl_head(l, t, sub));
    }
}

// ---------------------------------------------------------------------------
// Adjoint pair passes over the lists (the sums are stated in
// csrc/adjoint.hip, whose link-cell kernels k_adj_pair compute the same
// moments): a block per group with the group's halo positions staged in LDS
// (FP64, group frame), each thread a home atom that evaluates its own
// accepted list entries -- about four in five entries are inside the
// cutoff, so the basis sums run on nearly full waves without the link-cell
// kernels' pair queue.  dL/df and F' of the neighbours come from
// build-order copies (gb, l.fq) through the L2; each thread keeps its
// moments in registers and the block sums them over its waves in a fixed
// order (block_moments), so the gradient stays deterministic.  Pass 1
// differs from the link-cell kernels': instead of the force moments
// (1/2 w B_k, contracted with the derivative operator on the host) and
// the energy moments (1/2 dL/dE B_k) as two sets, it sums
// 1/2 dL/dE T_k + w / (b - a) k U_{k-1} per pair, the energy term and the
// derivative operator's result in one set of 32 registers
// (add_basis_deriv): the derivative series of basis function k is
// d B_k(y(r)) / dr = (2 / (b - a)) k U_{k-1}(y).

// dL/df in build order (gb[3t + k]) from the slot arrays.
__global__ __launch_bounds__(256) void k_adj_nl_gather(CmList l, const double *gx, const double *gy,
                                                       const double *gz, double *gb)
{
    const int n = l.gstart[l.ngroups];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int s = l.slot[l.order[t]];
        gb[3L * t + 0] = gx[s];
        gb[3L * t + 1] = gy[s];
        gb[3L * t + 2] = gz[s];
    }
}

template <int PASS, int MIXED>
__global__ __launch_bounds__(768) void k_adj_nl(DevState d, PotDev P, CmList l, CmAdjNl A, int pbase)
{
    extern __shared__ __align__(16) char nls[];
    __shared__ GroupLds S;
    __shared__ double red[12][ANMOM];
    double *hx = (double *)nls, *hy = hx + l.hcap, *hz = hy + l.hcap;
    CM_LDS uint8_t *slot_of = (CM_LDS uint8_t *)(nls + (size_t)l.hcap * 24);
    const int m = P.cn;
    const int dom = PASS == 1 ? 0 : 1;  // phi's series domain, else rho's
    const double ca = P.ca[dom], cb = P.cb[dom];
    const double rc2 = P.rc2;
    const int m8 = P.cstride;
    const double *dser = P.cheb + 4 * m8;  // rho' series
    double m0[AMAX];
#pragma unroll
    for (int k = 0; k < AMAX; k++) m0[k] = 0.0;

    // the group of this block, XCD-contiguous as in the force passes
    const int nb = gridDim.x, qb = nb / 8, rem = nb % 8;
    const int xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    const int g = nb < 8 ? (int)blockIdx.x : xcd * qb + min(xcd, rem) + idx;
    const NlRep R = nl_rep(l, l.group_rep[g]);
    const GroupGeo G = group_geo(R, g - R.group_base);
    group_setup(l, R, G, g, S, slot_of);
    const double ox = (G.x0 + 0.5 * G.lx) * R.ab0, oy = (G.y0 + 0.5 * G.ly) * R.ab1,
                 oz = (G.z0 + 0.5 * G.lz) * R.ab2;
    const int H = min(S.off[S.ns], l.hcap);
    for (int h = threadIdx.x; h < H; h += blockDim.x) {
        const int sl = slot_of[h];
        const CmPos p = load_pos(l.pos, S.start[sl] + (h - S.off[sl]));
        frame_pos(R, G, S.uxyz[sl], p, ox, oy, oz, hx[h], hy[h], hz[h]);
    }
    __syncthreads();
    // pass 1: each pair's half dL/dE (its replica's) weights the energy
    // moments, and w / (b - a) the derivative of the basis (the force
    // moments through the derivative operator)
    const double de = PASS == 1 ? 0.5 * A.dlde[l.group_rep[g]] : 0.0;
    const double inv_ba = 1.0 / (cb - ca);
    for (int t = S.t0 + threadIdx.x; t < S.t1; t += blockDim.x) {
        const int s = l.slot[l.order[t]];
        const int self = l.selfh[t];
        const double xi = hx[self], yi = hy[self], zi = hz[self];
        double gix = 0.0, giy = 0.0, giz = 0.0, wi = 0.0;
        if (PASS != 3) {
            gix = A.gx[s];
            giy = A.gy[s];
            giz = A.gz[s];
        }
        if (PASS == 2) wi = l.fq[t];
        if (PASS == 3) wi = A.mu[s];
        double lam = 0.0;
        const int cnt = l.ncnt[t];
        const CM_GLOBAL v4u *row = (const CM_GLOBAL v4u *)l.nbl + t;
        for (int k0 = 0; k0 < cnt; k0 += 8) {
            const v4u e = row[(size_t)(k0 >> 3) * l.n];
            const unsigned ew[4] = {e.x, e.y, e.z, e.w};
            for (int u = 0; u < 8 && k0 + u < cnt; u++) {
                const int h = (int)((ew[u >> 1] >> (16 * (u & 1))) & 0xffffu);
                const double dx = xi - hx[h], dy = yi - hy[h], dz = zi - hz[h];
                const double r2 = dx * dx + dy * dy + dz * dz;
                // pair predicates of src-flat/eam.c:519 (phi, rho) and 633 (rho')
                const bool in = PASS == 2 ? (r2 < rc2) : (r2 <= rc2);
                if (!in || !(r2 > 0.0)) continue;
                const double r = sqrt(r2);
                const double y = cheb_arg(r, ca, cb);
                if (PASS >= 3) {
                    add_basis<MIXED, 0>(y, m, wi, 0.0, m0, m0);
                    continue;
                }
                const int sl = slot_of[h];
                const int q = S.start[sl] + (h - S.off[sl]);
                const double w = ((gix - A.gb[3L * q + 0]) * dx + (giy - A.gb[3L * q + 1]) * dy +
                                  (giz - A.gb[3L * q + 2]) * dz) / r;
                if (PASS == 1) {
                    add_basis_deriv<MIXED>(y, m, de, w * inv_ba, m0);
                } else {
                    // rho'(r) from the same basis values (rho's domain)
                    const double drho = add_basis_sum<MIXED>(y, m, 0.5 * (wi + l.fq[q]) * w, m0, dser);
                    lam += drho * w;
                }
            }
        }
        if (PASS == 2) A.lam[s] = lam;
    }
    block_moments<0>(m0, red, A.part + (size_t)(pbase + blockIdx.x) * ANMOM);
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
    hipLaunchKernelGGL(k_nl_from_slots, dim3(grid_for(d.ncells, 4)), dim3(64), 0, st, d, l, 0);
}

void cm_nl_sync(const DevState &d, const CmList &l, hipStream_t st)
{
    hipLaunchKernelGGL(k_nl_from_slots, dim3(grid_for(d.ncells, 4)), dim3(64), 0, st, d, l, 1);
}

size_t cm_nl_build_lds(const CmList &l)
{
    return (size_t)l.hcap * 16 + nl_slotmap_bytes(l.hcap) + (size_t)l.tailq * l.pass_threads * 2;
}

int cm_nl_pick_tailq(CmList l)
{
    l.tailq = 0;
    int nb0 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb0, k_nl_build, l.pass_threads, cm_nl_build_lds(l)) !=
            hipSuccess ||
        nb0 < 1)
        return 0;
    for (int q = 16; q >= 4; q -= 4) {
        l.tailq = q;
        const size_t lds = cm_nl_build_lds(l);
        int nb = 0;
        if ((long)lds <= (long)l.lds_avail &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_nl_build, l.pass_threads, lds) == hipSuccess &&
            nb >= nb0)
            return q;
    }
    return 0;
}

void cm_nl_build(const DevState &d, const CmList &l, hipStream_t st)
{
    hipLaunchKernelGGL(k_nl_key, dim3(1024), dim3(256), 0, st, d, l);
    hipLaunchKernelGGL(k_nl_gsum, dim3((unsigned)std::min(4096, (l.ngroups + 3) / 4)), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_sort, dim3(1024), dim3(256), 0, st, l);
    hipLaunchKernelGGL(k_nl_build, dim3(l.pass_blocks), dim3(l.pass_threads), cm_nl_build_lds(l), st, d, l);
}

size_t cm_nl_tab_bytes(const PotDev &P, int kind, int mixed)
{
    return P.eval == COMD_EVAL_TABLE && kind != CM_KIND_LJ && P.cd64 ? nl_tab_bytes(P, kind, mixed) : 0;
}

size_t cm_nl_pass_lds(const PotDev &P, const CmList &l, int kind, int mixed, int *tables_in_lds)
{
    const size_t halo = (size_t)l.hcap * nl_halo_bytes(kind, mixed) + nl_slotmap_bytes(l.hcap);
    size_t tab = 0;
    if (P.eval == COMD_EVAL_TABLE && kind != CM_KIND_LJ && P.cd64) tab = nl_tab_bytes(P, kind, mixed);
    // rows in LDS whenever they fit beside the halo (unless tuned off)
    const int tl = tab > 0 && l.rows_lds != 0 && halo + tab <= (size_t)l.lds_avail;
    if (tables_in_lds) *tables_in_lds = tl;
    return halo + (tl ? tab : 0);
}

template <int KIND, int EVAL, int MIXED>
static void nl_launch(const DevState &d, const PotDev &P, const CmList &l, hipStream_t st, int g0, int ng)
{
    int tl = 0;
    const size_t lds = cm_nl_pass_lds(P, l, KIND, MIXED, &tl);
    const int tab = tl ? (int)nl_tab_bytes(P, KIND, MIXED) : 0;
    const dim3 grid(ng), block(l.pass_threads);
    if constexpr (EVAL == COMD_EVAL_TABLE && KIND != CM_KIND_LJ) {
        if (tl) {
            if (l.lanes > 1) hipLaunchKernelGGL((k_nl_pair<KIND, EVAL, MIXED, 1, 1>), grid, block, lds, st, d, P, l, tab, g0);
            else hipLaunchKernelGGL((k_nl_pair<KIND, EVAL, MIXED, 1, 0>), grid, block, lds, st, d, P, l, tab, g0);
            return;
        }
    }
    if (l.lanes > 1) hipLaunchKernelGGL((k_nl_pair<KIND, EVAL, MIXED, 0, 1>), grid, block, lds, st, d, P, l, 0, g0);
    else hipLaunchKernelGGL((k_nl_pair<KIND, EVAL, MIXED, 0, 0>), grid, block, lds, st, d, P, l, 0, g0);
}

// Resident blocks per CU of the pass nl_launch would start, and its dynamic LDS.
template <int KIND, int EVAL, int MIXED>
static int nl_occupancy(const PotDev &P, const CmList &l, int *lds_bytes)
{
    int tl = 0;
    const size_t lds = cm_nl_pass_lds(P, l, KIND, MIXED, &tl);
    const bool ln = l.lanes > 1;
    const void *fn = ln ? (const void *)k_nl_pair<KIND, EVAL, MIXED, 0, 1> : (const void *)k_nl_pair<KIND, EVAL, MIXED, 0, 0>;
    if constexpr (EVAL == COMD_EVAL_TABLE && KIND != CM_KIND_LJ)
        if (tl)
            fn = ln ? (const void *)k_nl_pair<KIND, EVAL, MIXED, 1, 1> : (const void *)k_nl_pair<KIND, EVAL, MIXED, 1, 0>;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, l.pass_threads, lds) != hipSuccess) nb = -1;
    if (lds_bytes) *lds_bytes = (int)lds;
    return nb;
}

int cm_nl_occupancy(const PotDev &P, const CmList &l, int kind, int mixed, int *lds_bytes)
{
    if (kind == CM_KIND_LJ)
        return mixed ? nl_occupancy<CM_KIND_LJ, COMD_EVAL_TABLE, 1>(P, l, lds_bytes)
                     : nl_occupancy<CM_KIND_LJ, COMD_EVAL_TABLE, 0>(P, l, lds_bytes);
    if (P.eval == COMD_EVAL_CHEB) {
        if (kind == CM_KIND_EAM1)
            return mixed ? nl_occupancy<CM_KIND_EAM1, COMD_EVAL_CHEB, 1>(P, l, lds_bytes)
                         : nl_occupancy<CM_KIND_EAM1, COMD_EVAL_CHEB, 0>(P, l, lds_bytes);
        return mixed ? nl_occupancy<CM_KIND_EAM3, COMD_EVAL_CHEB, 1>(P, l, lds_bytes)
                     : nl_occupancy<CM_KIND_EAM3, COMD_EVAL_CHEB, 0>(P, l, lds_bytes);
    }
    if (kind == CM_KIND_EAM1)
        return mixed ? nl_occupancy<CM_KIND_EAM1, COMD_EVAL_TABLE, 1>(P, l, lds_bytes)
                     : nl_occupancy<CM_KIND_EAM1, COMD_EVAL_TABLE, 0>(P, l, lds_bytes);
    return mixed ? nl_occupancy<CM_KIND_EAM3, COMD_EVAL_TABLE, 1>(P, l, lds_bytes)
                 : nl_occupancy<CM_KIND_EAM3, COMD_EVAL_TABLE, 0>(P, l, lds_bytes);
}

template <int KIND, int EVAL>
static void nl_launch2(const DevState &d, const PotDev &P, const CmList &l, int mixed, hipStream_t st, int g0, int ng)
{
    if (mixed) nl_launch<KIND, EVAL, 1>(d, P, l, st, g0, ng);
    else nl_launch<KIND, EVAL, 0>(d, P, l, st, g0, ng);
}

void cm_nl_force(const DevState &d, const PotDev &P, const CmList &l, int kind, int mixed, hipStream_t st, int g0,
                 int ng)
{
    if (ng <= 0) return;
    if (kind == CM_KIND_LJ) {
        hipLaunchKernelGGL(k_nl_nonempty, dim3(grid_for(d.ncells, 256)), dim3(256), 0, st, d, l, g0);
        nl_launch2<CM_KIND_LJ, COMD_EVAL_TABLE>(d, P, l, mixed, st, g0, ng);
    } else if (P.eval == COMD_EVAL_CHEB) {
        if (kind == CM_KIND_EAM1) nl_launch2<CM_KIND_EAM1, COMD_EVAL_CHEB>(d, P, l, mixed, st, g0, ng);
        else nl_launch2<CM_KIND_EAM3, COMD_EVAL_CHEB>(d, P, l, mixed, st, g0, ng);
    } else {
        if (kind == CM_KIND_EAM1) nl_launch2<CM_KIND_EAM1, COMD_EVAL_TABLE>(d, P, l, mixed, st, g0, ng);
        else nl_launch2<CM_KIND_EAM3, COMD_EVAL_TABLE>(d, P, l, mixed, st, g0, ng);
    }
}

size_t cm_adj_nl_lds(const CmList &l) { return (size_t)l.hcap * 24 + nl_slotmap_bytes(l.hcap); }

template <int MIXED>
static void adj_nl_launch(const DevState &d, const PotDev &P, const CmList &l, const CmAdjNl &A, int pass,
                          hipStream_t st)
{
    const dim3 grid(l.pass_blocks), block(l.pass_threads);
    const size_t lds = cm_adj_nl_lds(l);
    switch (pass) {
    case 1: hipLaunchKernelGGL((k_adj_nl<1, MIXED>), grid, block, lds, st, d, P, l, A, 0); break;
    case 2: hipLaunchKernelGGL((k_adj_nl<2, MIXED>), grid, block, lds, st, d, P, l, A, 0); break;
    case 3: hipLaunchKernelGGL((k_adj_nl<3, MIXED>), grid, block, lds, st, d, P, l, A, 0); break;
    default: break;
    }
}

void cm_launch_adj_nl_gather(const CmList &l, const double *gx, const double *gy, const double *gz, double *gb,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_nl_gather, dim3(grid_for(l.n, 256)), dim3(256), 0, st, l, gx, gy, gz, gb);
}

void cm_launch_adj_nl(const DevState &d, const PotDev &P, const CmList &l, const CmAdjNl &A, int pass, int mixed,
                      hipStream_t st)
{
    if (mixed) adj_nl_launch<1>(d, P, l, A, pass, st);
    else adj_nl_launch<0>(d, P, l, A, pass, st);
}
