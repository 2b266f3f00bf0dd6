// bin_rebox.hip -- atom binning (GPU counting sort) and a parallel, in-place
// emulation of reBoxAll, with the kick/drift that precedes every rebin of the
// position-Verlet step fused in.
//
// reBoxAll (src-flat/domains.c:347-389) scans cells in ascending order; an
// atom that left its cell is appended to the destination (moveAtom,
// domains.c:224-236) and the source's last atom fills the hole, which is
// examined next.  Its result is reproduced without a serial scan:
//   * whether an atom leaves, and where to, depends only on its own r;
//   * the order of departures from cell c depends only on c's own list and
//     on the NUMBER n_low of atoms that arrived from lower-indexed cells
//     before c was scanned (they are appended and never leave again);
//   * arrivals into c come in (source cell, departure rank) order; those
//     from lower cells fill the n_low placeholder positions, those from
//     higher cells are appended after c's scan.
// Only cells that lose or gain atoms change, so the pass works on the list of
// movers and of touched cells (a few per thousand atoms per half step) and
// updates the slot arrays in place.  The argument holds while no arrival
// into a higher-indexed cell would move again in the same pass; k_move flags
// that case for the atom's replica, whose cells the parallel kernels then
// leave alone, and a serial kernel (one block per flagged replica) redoes
// that replica's pass exactly as the reference does.
#include "comd_internal.h"

#define CM_ACAP 64  // arrivals recorded per cell and pass

__device__ static inline void mark_touched(const DevState &d, int c)
{
    if (atomicExch(&d.touched[c], 1) == 0) {
        const int i = atomicAdd(&d.counts[1], 1);
        d.tc[i] = c;
    }
}

// ---------------------------------------------------------------------------
// kick (p -= dtk f, src-flat/timestep.c:62-74), drift (r += (dtd p)/rmass,
// timestep.c:76-104) and the per-atom destination of reBoxAll
// (domains.c:356-377) of the atom in slot s of cell c = (ix, iy, iz);
// departing atoms are recorded as movers.
__device__ __forceinline__ void move_atom(const DevState &d, const Grid &g, int c, int ix, int iy, int iz, int s,
                                          double dtk, double dtd, int kick, int drift, int rebin)
{
    const double rmass = g.rmass;
    // every load first (the arrays may alias as far as the compiler can
    // tell, so a load after a store would wait for it)
    double x = d.rx[s], y = d.ry[s], z = d.rz[s];
    double px = d.px[s], py = d.py[s], pz = d.pz[s];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (kick) {
        fx = d.fx[s];
        fy = d.fy[s];
        fz = d.fz[s];
    }
    const bool upkeep = drift && d.nl_pos;
    float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (upkeep) q = d.nlr[s];  // position at the last build, build position
    if (kick) {
        px = px - dtk * fx;
        py = py - dtk * fy;
        pz = pz - dtk * fz;
        d.px[s] = px;
        d.py[s] = py;
        d.pz[s] = pz;
    }
    if (drift) {
        x = x + dtd * px / rmass;
        y = y + dtd * py / rmass;
        z = z + dtd * pz / rmass;
        d.rx[s] = x;
        d.ry[s] = y;
        d.rz[s] = z;
        if (upkeep) {
            // absolute position for the neighbour lists, and the rebuild
            // request once the atom has moved half the skin (nearest image)
            double cx, cy, cz;
            corner_of(g, ix, iy, iz, cx, cy, cz);
            const double ax = cx + x, ay = cy + y, az = cz + z;
            d.nl_pos[__float_as_int(q.w)] = CmPos{ax, ay, az};
            double ex = ax - q.x, ey = ay - q.y, ez = az - q.z;
            ex -= g.L0 * rint(ex * g.iL0);
            ey -= g.L1 * rint(ey * g.iL1);
            ez -= g.L2 * rint(ez * g.iL2);
            // one request per wave is enough
            const bool far = ex * ex + ey * ey + ez * ez > d.nl_half2;
            const unsigned long long who = __ballot(far);
            if (who != 0 && (int)__ffsll((long long)who) - 1 == (int)(threadIdx.x & 63))
                atomicOr(d.nl_flags, (unsigned)CM_NL_REBUILD);
        }
    }
    if (!rebin) return;
    int nx = ix, ny = iy, nz = iz;
    double ax = x, ay = y, az = z;
    if (x < 0.0) { nx = ix - 1; ax = x + g.bs0; } else if (x >= g.bs0) { nx = ix + 1; ax = x - g.bs0; }
    if (y < 0.0) { ny = iy - 1; ay = y + g.bs1; } else if (y >= g.bs1) { ny = iy + 1; ay = y - g.bs1; }
    if (z < 0.0) { nz = iz - 1; az = z + g.bs2; } else if (z >= g.bs2) { nz = iz + 1; az = z - g.bs2; }
    const int j = cell_wrap(g, nx, ny, nz);
    if (j < 0 || j == c) return;
    const int dst = j;
    if (dst > c) {
        // the arrival is examined again when dst is scanned
        int jx, jy, jz;
        cell_xyz(g, dst, jx, jy, jz);
        int mx = jx, my = jy, mz = jz;
        if (ax < 0.0) mx = jx - 1; else if (ax >= g.bs0) mx = jx + 1;
        if (ay < 0.0) my = jy - 1; else if (ay >= g.bs1) my = jy + 1;
        if (az < 0.0) mz = jz - 1; else if (az >= g.bs2) mz = jz + 1;
        const int j2 = cell_wrap(g, mx, my, mz);
        if (j2 >= 0 && j2 != dst) atomicOr(&d.mh[g.pot], 1);
    }
    const int m = atomicAdd(&d.counts[0], 1);  // capacity: one record per slot
    d.mv_slot[m] = s;
    d.mv_id[m] = d.id[s];
    d.mv_nlr[m] = d.nlr[s];
    d.mv_itype[m] = d.itype[s];
    d.mv_rx[m] = ax;
    d.mv_ry[m] = ay;
    d.mv_rz[m] = az;
    d.mv_px[m] = px;
    d.mv_py[m] = py;
    d.mv_pz[m] = pz;
    d.mvidx[s] = m;
    atomicOr(&d.leave[c], 1ull << (s - c * CM_SLOTS));
    const int k = atomicAdd(&d.arr_cnt[dst], 1);
    if (k < CM_ACAP) d.arr[dst * CM_ACAP + k] = m;
    mark_touched(d, c);
    mark_touched(d, dst);
}

// Any number of replicas: four consecutive global cells per wave, 16 lanes
// each (a cell's atoms beyond 16 taken in further rounds).  Held to 96
// registers for five waves per SIMD instead of four (no spills; the
// 64-replica batch step 0.462 -> 0.459 ms).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
void k_move(DevState d, double dtk, double dtd, int kick, int drift, int rebin)
{
    const int c = blockIdx.x * 4 + (threadIdx.x >> 4);
    if (c >= d.ncells) return;
    const int n = d.natoms[c];
    if ((threadIdx.x & 15) >= n) return;
    const Grid g = grid_of(d, c);
    int ix, iy, iz;
    cell_xyz(g, c, ix, iy, iz);
    for (int t = threadIdx.x & 15; t < n; t += 16) move_atom(d, g, c, ix, iy, iz, c * CM_SLOTS + t, dtk, dtd, kick, drift, rebin);
}

// One replica: four cells per wave, 16 lanes each, cell coordinates from the
// 3-D grid (x: groups of four cells).
__global__ __launch_bounds__(64) void k_move4(DevState d, double dtk, double dtd, int kick, int drift, int rebin)
{
    const Grid &g = d.g0;
    const int ix = blockIdx.x * 4 + (threadIdx.x >> 4);
    const int iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= g.nbx0) return;
    const int c = ix + g.nbx0 * (iy + g.nbx1 * iz);
    const int n = d.natoms[c];
    for (int t = threadIdx.x & 15; t < n; t += 16)
        move_atom(d, g, c, ix, iy, iz, c * CM_SLOTS + t, dtk, dtd, kick, drift, rebin);
}

// Per touched cell (one wave each): count arrivals from lower/higher cells,
// then run the cell's own swap-with-last scan with n_low never-departing
// placeholders appended, recording each departure's rank and the resulting
// arrangement.  The lanes load the arrivals and the departing atoms' mover
// records together; lane 0 runs the scan in LDS.
__global__ __launch_bounds__(64) void k_rebox_scan(DevState d)
{
    __shared__ uint8_t perm[128];
    __shared__ int rank_of[CM_SLOTS];
    const int L = threadIdx.x;
    const int ntc = d.counts[1];
    for (int i = blockIdx.x; i < ntc; i += gridDim.x) {
        const int c = d.tc[i];
        if (d.mh[rep_of(d, c)]) continue;  // the same in every lane: replayed serially
        const int n0 = d.natoms[c];
        int na = d.arr_cnt[c];
        // which of the cell's own atoms leave (set by k_move per departure)
        const unsigned long long leaves = d.leave[c];
        unsigned fl = 0;
        if (na > CM_ACAP) {
            fl |= CM_FLAG_OVERFLOW;
            na = CM_ACAP;
        }
        bool low = false;
        if (L < na) low = (d.mv_slot[d.arr[c * CM_ACAP + L]] / CM_SLOTS) < c;
        const bool leaving = L < n0 && ((leaves >> L) & 1ull);
        const int mv = leaving ? d.mvidx[c * CM_SLOTS + L] : -1;
        const int nlow = __popcll(__ballot(low));
        const int nhigh = na - nlow;
        if (nlow > 0 && n0 + nlow >= CM_SLOTS) fl |= CM_FLAG_OVERFLOW;
        int len = n0 + nlow;
        if (len > 127) len = 127;
        for (int k = L; k < len; k += 64) perm[k] = (uint8_t)(k < n0 ? k : CM_SLOTS + (k - n0));
        __syncthreads();
        if (L == 0) {
            int p = 0, rk = 0;
            while (p < len) {
                const int e = perm[p];
                if (e < CM_SLOTS && ((leaves >> e) & 1ull)) {
                    rank_of[e] = rk++;
                    len--;
                    perm[p] = perm[len];
                } else {
                    p++;
                }
            }
        }
        len = __shfl(len, 0, 64);
        __syncthreads();
        if (leaving) d.mv_rank[mv] = rank_of[L];
        if (nhigh > 0 && len + nhigh >= CM_SLOTS) fl |= CM_FLAG_OVERFLOW;
        const int keep = len < CM_SLOTS ? len : CM_SLOTS;
        if (L < keep) d.perm[c * CM_SLOTS + L] = perm[L];
        if (L == 0) {
            d.meta[4 * c + 0] = len;
            d.meta[4 * c + 1] = nlow;
            d.meta[4 * c + 2] = nhigh;
            d.meta[4 * c + 3] = n0;
            if (fl) atomicOr(d.flags, fl);
        }
        __syncthreads();
    }
}

// Per touched cell (one wave each, lane = output slot): order the arrivals by
// (source cell, rank), then write survivors and low arrivals in the scanned
// arrangement followed by the high arrivals.  All lanes read before any lane
// writes, so the cell is updated in place.  Clears the cell's pass state.
__global__ __launch_bounds__(64) void k_rebox_gather(DevState d)
{
    __shared__ int sm[2][CM_SLOTS];  // movers by low / high arrival order
    const int L = threadIdx.x;
    const int ntc = d.counts[1];
    for (int i = blockIdx.x; i < ntc; i += gridDim.x) {
        const int c = d.tc[i];
        // cells of a replica whose pass is replayed serially are left as the
        // move kernel left them (the same in every lane)
        if (!d.mh[rep_of(d, c)]) {
            const int len = d.meta[4 * c + 0];
            const int nhigh = d.meta[4 * c + 2];
            int na = d.arr_cnt[c];
            na = na < CM_ACAP ? na : CM_ACAP;
            int m = -1, src = 0, rk = 0;
            if (L < na) {
                m = d.arr[c * CM_ACAP + L];
                src = d.mv_slot[m] / CM_SLOTS;
                rk = d.mv_rank[m];
            }
            // rank among arrivals of the same class, by (source cell, rank)
            const bool low = src < c;
            int pos = 0;
            for (int u = 0; u < na; u++) {
                const int su = __shfl(src, u, 64);
                const int ru = __shfl(rk, u, 64);
                const bool lu = su < c;
                pos += (lu == low) && (su < src || (su == src && ru < rk));
            }
            if (L < na && pos < CM_SLOTS) sm[low ? 0 : 1][pos] = m;
            __syncthreads();
            // gather this lane's output atom
            const int tot = len + nhigh;
            const int nout = tot < CM_SLOTS ? tot : CM_SLOTS;
            const int keep = len < CM_SLOTS ? len : CM_SLOTS;
            int gid = 0, git = 0;
            float4 gq = {0.0f, 0.0f, 0.0f, 0.0f};
            double gx = 0, gy = 0, gz = 0, gpx = 0, gpy = 0, gpz = 0;
            if (L < nout) {
                int mv = -1, own = -1;
                if (L < keep) {
                    const int e = d.perm[c * CM_SLOTS + L];
                    if (e < CM_SLOTS) own = c * CM_SLOTS + e;
                    else mv = sm[0][(e - CM_SLOTS) & (CM_SLOTS - 1)];
                } else {
                    mv = sm[1][(L - len) & (CM_SLOTS - 1)];
                }
                if (own >= 0) {
                    gid = d.id[own];
                    gq = d.nlr[own];
                    git = d.itype[own];
                    gx = d.rx[own];
                    gy = d.ry[own];
                    gz = d.rz[own];
                    gpx = d.px[own];
                    gpy = d.py[own];
                    gpz = d.pz[own];
                } else if (mv >= 0) {
                    gid = d.mv_id[mv];
                    gq = d.mv_nlr[mv];
                    git = d.mv_itype[mv];
                    gx = d.mv_rx[mv];
                    gy = d.mv_ry[mv];
                    gz = d.mv_rz[mv];
                    gpx = d.mv_px[mv];
                    gpy = d.mv_py[mv];
                    gpz = d.mv_pz[mv];
                }
            }
            __syncthreads();  // every lane has read its source before any slot is overwritten
            if (L < nout) {
                const int o = c * CM_SLOTS + L;
                if (d.nl_slot) d.nl_slot[grid_of(d, c).nl_id_base + gid] = o;
                d.id[o] = gid;
                d.nlr[o] = gq;
                d.itype[o] = git;
                d.rx[o] = gx;
                d.ry[o] = gy;
                d.rz[o] = gz;
                d.px[o] = gpx;
                d.py[o] = gpy;
                d.pz[o] = gpz;
            }
            if (L == 0) d.natoms[c] = nout;
        }
        if (L == 0) {
            d.arr_cnt[c] = 0;
            d.touched[c] = 0;
            d.leave[c] = 0;
        }
        __syncthreads();
    }
}

// Serial pass of the replicas in which k_move found an arrival that would
// move again (one block each, its first thread walking the replica's cells):
// reBoxAll exactly as the reference runs it, in place.  Block 0 also clears
// the pass counters.
__device__ static void serial_copy(const DevState &d, int from, int to)
{
    if (d.nl_slot) d.nl_slot[grid_of(d, to / CM_SLOTS).nl_id_base + d.id[from]] = to;
    d.id[to] = d.id[from];
    d.nlr[to] = d.nlr[from];
    d.itype[to] = d.itype[from];
    d.rx[to] = d.rx[from];
    d.ry[to] = d.ry[from];
    d.rz[to] = d.rz[from];
    d.px[to] = d.px[from];
    d.py[to] = d.py[from];
    d.pz[to] = d.pz[from];
}

__global__ void k_rebox_serial(DevState d)
{
    if (threadIdx.x != 0) return;
    const int r = blockIdx.x;
    if (r == 0) {
        d.counts[0] = 0;
        d.counts[1] = 0;
    }
    if (!d.mh[r]) return;
    const Grid g = d.ngrids == 1 ? d.g0 : d.grids[r];
    unsigned fl = 0;
    for (int c = g.cell_base; c < g.cell_base + g.ncells; c++) {
        int ix, iy, iz;
        cell_xyz(g, c, ix, iy, iz);
        int i = 0;
        while (i < d.natoms[c]) {
            const int o = c * CM_SLOTS + i;
            const double x = d.rx[o], y = d.ry[o], z = d.rz[o];
            int nx = ix, ny = iy, nz = iz;
            double ax = x, ay = y, az = z;
            if (x < 0.0) { nx = ix - 1; ax = x + g.bs0; } else if (x >= g.bs0) { nx = ix + 1; ax = x - g.bs0; }
            if (y < 0.0) { ny = iy - 1; ay = y + g.bs1; } else if (y >= g.bs1) { ny = iy + 1; ay = y - g.bs1; }
            if (z < 0.0) { nz = iz - 1; az = z + g.bs2; } else if (z >= g.bs2) { nz = iz + 1; az = z - g.bs2; }
            const int j = cell_wrap(g, nx, ny, nz);
            if (j < 0 || j == c) {
                i++;
                continue;
            }
            d.rx[o] = ax;
            d.ry[o] = ay;
            d.rz[o] = az;
            const int nj = d.natoms[j];
            if (nj >= CM_SLOTS) {  // no room: the reference aborts here
                fl |= CM_FLAG_OVERFLOW;
                i++;
                continue;
            }
            serial_copy(d, o, j * CM_SLOTS + nj);
            d.natoms[j] = nj + 1;
            if (nj + 1 >= CM_SLOTS) fl |= CM_FLAG_OVERFLOW;
            // the cell's last atom fills the hole; none to move when the
            // departed atom was the last (a self-copy would point its id ->
            // slot entry back at the slot it left)
            const int ni = --d.natoms[c];
            if (ni > i) serial_copy(d, c * CM_SLOTS + ni, o);
        }
    }
    d.mh[r] = 0;
    if (fl) atomicOr(d.flags, fl);
}

// ---------------------------------------------------------------------------
// Binning of a flat atom list as the sequence of putAtomInBox calls
// (src-flat/domains.c:237-291): wrap once into [0,L), cell =
// floor(x/boxsize), append.  Stable counting sort: atomic tickets per cell,
// then each cell's entries are ordered by input index.
__device__ static inline bool wrap_bin(const Grid &g, double &x, double &y, double &z, int &cell)
{
    if (x < 0.0) x += g.L0; else if (x >= g.L0) x -= g.L0;
    if (y < 0.0) y += g.L1; else if (y >= g.L1) y -= g.L1;
    if (z < 0.0) z += g.L2; else if (z >= g.L2) z -= g.L2;
    const int ix = (int)floor(x / g.bs0);
    const int iy = (int)floor(y / g.bs1);
    const int iz = (int)floor(z / g.bs2);
    if (ix < 0 || ix >= g.nbx0 || iy < 0 || iy >= g.nbx1 || iz < 0 || iz >= g.nbx2) return false;
    cell = g.cell_base + ix + g.nbx0 * iy + g.nbx0 * g.nbx1 * iz;
    return true;
}

__global__ void k_bin_key(DevState d, Grid g, int n, const double *xyz, int *cnt, int *tmp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    int cell;
    if (!wrap_bin(g, x, y, z, cell)) {
        atomicOr(d.flags, (unsigned)CM_FLAG_OUTSIDE);
        return;
    }
    const int t = atomicAdd(&cnt[cell], 1);
    if (t < CM_SLOTS) tmp[cell * CM_SLOTS + t] = i;
    else atomicOr(d.flags, (unsigned)CM_FLAG_OVERFLOW);
}

__global__ __launch_bounds__(64) void k_bin_place(DevState d, Grid g, const double *xyz, const double *pxyz,
                                                  int itype, const int *cnt, const int *tmp)
{
    __shared__ int sidx[CM_SLOTS];
    const int c = g.cell_base + blockIdx.x;
    const int t = threadIdx.x;
    int n = cnt[c];
    if (n > CM_SLOTS) n = CM_SLOTS;
    if (t < n) sidx[t] = tmp[c * CM_SLOTS + t];
    __syncthreads();
    if (t == 0) d.natoms[c] = n;
    if (t >= n) return;
    const int me = sidx[t];
    int rank = 0;
    for (int u = 0; u < n; u++) rank += (sidx[u] < me);
    double x = xyz[3 * me], y = xyz[3 * me + 1], z = xyz[3 * me + 2];
    int cell;
    wrap_bin(g, x, y, z, cell);
    double cx, cy, cz;
    cell_corner(g, c, cx, cy, cz);
    const int s = c * CM_SLOTS + rank;
    d.id[s] = me;
    d.itype[s] = itype;
    d.rx[s] = x - cx;
    d.ry[s] = y - cy;
    d.rz[s] = z - cz;
    d.px[s] = pxyz ? pxyz[3 * me] : 0.0;
    d.py[s] = pxyz ? pxyz[3 * me + 1] : 0.0;
    d.pz[s] = pxyz ? pxyz[3 * me + 2] : 0.0;
}

__global__ void k_fill_int(int *p, int v, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// host launchers

void cm_launch_move(const DevState &d, hipStream_t st, double dtk, double dtd, int kick, int drift, int rebin)
{
    if (d.ngrids == 1 && d.g0.cell_base == 0) {
        const dim3 grid((unsigned)((d.g0.nbx0 + 3) / 4), (unsigned)d.g0.nbx1, (unsigned)d.g0.nbx2);
        hipLaunchKernelGGL(k_move4, grid, dim3(64), 0, st, d, dtk, dtd, kick, drift, rebin);
    } else {
        hipLaunchKernelGGL(k_move, dim3((d.ncells + 3) / 4), dim3(64), 0, st, d, dtk, dtd, kick, drift, rebin);
    }
}

void cm_launch_rebox(const DevState &d, hipStream_t st)
{
    const int ns = d.ncells < 8192 ? d.ncells : 8192;
    hipLaunchKernelGGL(k_rebox_scan, dim3(ns), dim3(64), 0, st, d);
    const int ng = d.ncells < 4096 ? d.ncells : 4096;
    hipLaunchKernelGGL(k_rebox_gather, dim3(ng), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_rebox_serial, dim3((unsigned)d.ngrids), dim3(64), 0, st, d);
}

void cm_launch_fill_int(int *p, int v, long n, hipStream_t st)
{
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
}

void cm_launch_bin(const DevState &d, const Grid &g, hipStream_t st, int n, const double *xyz, const double *pxyz,
                   int itype, int *cnt, int *tmp)
{
    cm_launch_fill_int(cnt + g.cell_base, 0, g.ncells, st);
    if (n > 0) hipLaunchKernelGGL(k_bin_key, dim3((n + 255) / 256), dim3(256), 0, st, d, g, n, xyz, cnt, tmp);
    hipLaunchKernelGGL(k_bin_place, dim3(g.ncells), dim3(64), 0, st, d, g, xyz, pxyz, itype, cnt, tmp);
}
