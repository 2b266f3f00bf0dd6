// bin_rebox.hip -- atom binning (GPU counting sort) and the parallel
// emulation of reBoxAll, plus the fused kick/drift that precedes every
// rebin in the position-Verlet step.
//
// reBoxAll (src-flat/domains.c:347-389) scans cells in ascending order; an
// atom that left its cell is appended to the destination (moveAtom,
// domains.c:224-236) and the source's last atom fills the hole, which is
// examined next.  Its result is reproduced here without a serial scan:
//   * whether an atom leaves, and where to, depends only on its own r;
//   * the order of departures from cell c depends only on c's own list and
//     on the NUMBER n_low of atoms that arrived from lower-indexed cells
//     before c was scanned (they are appended and never leave again);
//   * arrivals into c come in (source cell, departure rank) order; those
//     from lower cells fill the n_low placeholder positions, those from
//     higher cells are appended after c's scan.
// That holds while no arrival into a higher-indexed cell would move again in
// the same pass (k_move flags it).  Then a serial kernel redoes the pass
// exactly as the reference does.
#include "comd_internal.h"

// ---------------------------------------------------------------------------
// kick (p -= dtk f, src-flat/timestep.c:62-74), drift (r += (dtd p)/rmass,
// timestep.c:76-104) and the per-atom destination of reBoxAll
// (domains.c:356-377).  One block of 64 threads per cell.
__global__ __launch_bounds__(64) void k_move(DevState d, double dtk, double dtd, double rmass, int kick, int drift)
{
    const int c = blockIdx.x;
    const int t = threadIdx.x;
    const int n = d.natoms[c];
    if (t >= n) return;
    const int s = c * CM_SLOTS + t;
    double x = d.rx[s], y = d.ry[s], z = d.rz[s];
    if (kick || drift) {
        double px = d.px[s], py = d.py[s], pz = d.pz[s];
        if (kick) {
            px = px - dtk * d.fx[s];
            py = py - dtk * d.fy[s];
            pz = pz - dtk * d.fz[s];
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
        }
    }
    int ix, iy, iz;
    cell_xyz(d, c, ix, iy, iz);
    int nx = ix, ny = iy, nz = iz;
    double ax = x, ay = y, az = z;
    if (x < 0.0) { nx = ix - 1; ax = x + d.bs0; } else if (x >= d.bs0) { nx = ix + 1; ax = x - d.bs0; }
    if (y < 0.0) { ny = iy - 1; ay = y + d.bs1; } else if (y >= d.bs1) { ny = iy + 1; ay = y - d.bs1; }
    if (z < 0.0) { nz = iz - 1; az = z + d.bs2; } else if (z >= d.bs2) { nz = iz + 1; az = z - d.bs2; }
    int j = cell_wrap(d, nx, ny, nz);
    int dst = (j < 0 || j == c) ? -1 : j;
    d.dest[s] = dst;
    if (dst >= 0) {
        d.rnx[s] = ax;
        d.rny[s] = ay;
        d.rnz[s] = az;
        if (dst > c) {
            // the arrival is examined again when dst is scanned
            int jx, jy, jz;
            cell_xyz(d, dst, jx, jy, jz);
            int mx = jx, my = jy, mz = jz;
            if (ax < 0.0) mx = jx - 1; else if (ax >= d.bs0) mx = jx + 1;
            if (ay < 0.0) my = jy - 1; else if (ay >= d.bs1) my = jy + 1;
            if (az < 0.0) mz = jz - 1; else if (az >= d.bs2) mz = jz + 1;
            int j2 = cell_wrap(d, mx, my, mz);
            if (j2 >= 0 && j2 != dst) atomicOr(d.flags, (unsigned)CM_FLAG_MULTIHOP);
        }
    }
}

// Distinct cells (other than c) that an atom can leave to reach c: the 26
// neighbours, wrapped and de-duplicated, ascending.
__device__ static int source_cells(const DevState &d, int c, int *src)
{
    int n = 0;
    if (d.ncells == 1) return 0;
    int ix, iy, iz;
    cell_xyz(d, c, ix, iy, iz);
    for (int e = 0; e < 27; e++) {
        if (e == 13) continue;
        int b = cell_wrap(d, ix + e / 9 - 1, iy + (e / 3) % 3 - 1, iz + e % 3 - 1);
        if (b < 0 || b == c) continue;
        int k = n;
        bool dup = false;
        for (int q = 0; q < n; q++) dup |= (src[q] == b);
        if (dup) continue;
        while (k > 0 && src[k - 1] > b) { src[k] = src[k - 1]; k--; }
        src[k] = b;
        n++;
    }
    return n;
}

// Per cell: count arrivals from lower/higher cells, then run the cell's own
// swap-with-last scan with n_low never-departing placeholders appended,
// recording each departure's rank and the surviving arrangement.
__global__ __launch_bounds__(64) void k_rebox_scan(DevState d)
{
    __shared__ uint8_t sperm[64][128];
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= d.ncells) return;
    uint8_t *perm = sperm[threadIdx.x];
    const int n0 = d.natoms[c];
    int src[26];
    const int nsrc = source_cells(d, c, src);
    int nlow = 0, nhigh = 0;
    for (int q = 0; q < nsrc; q++) {
        const int b = src[q];
        const int nb = d.natoms[b];
        int k = 0;
        for (int s = 0; s < nb; s++) k += (d.dest[b * CM_SLOTS + s] == c);
        if (b < c) nlow += k; else nhigh += k;
    }
    unsigned fl = 0;
    if (nlow > 0 && n0 + nlow >= CM_SLOTS) fl |= CM_FLAG_OVERFLOW;
    int len = n0 + nlow;
    if (len > 127) len = 127;
    for (int i = 0; i < len; i++) perm[i] = (uint8_t)(i < n0 ? i : CM_SLOTS + (i - n0));
    int i = 0, rk = 0;
    while (i < len) {
        const int e = perm[i];
        const bool dep = e < CM_SLOTS && d.dest[c * CM_SLOTS + e] >= 0;
        if (!dep) {
            i++;
        } else {
            d.rank[c * CM_SLOTS + e] = rk++;
            len--;
            perm[i] = perm[len];
        }
    }
    if (nhigh > 0 && len + nhigh >= CM_SLOTS) fl |= CM_FLAG_OVERFLOW;
    const int keep = len < CM_SLOTS ? len : CM_SLOTS;
    for (int k = 0; k < keep; k++) d.perm[c * CM_SLOTS + k] = perm[k];
    d.meta[4 * c + 0] = len;
    d.meta[4 * c + 1] = nlow;
    d.meta[4 * c + 2] = nhigh;
    d.meta[4 * c + 3] = n0;
    if (fl) atomicOr(d.flags, fl);
}

__device__ static inline void copy_slot_to_next(const DevState &d, int from, int to, bool moved)
{
    d.id_n[to] = d.id[from];
    d.itype_n[to] = d.itype[from];
    if (moved) {
        d.rx_n[to] = d.rnx[from];
        d.ry_n[to] = d.rny[from];
        d.rz_n[to] = d.rnz[from];
    } else {
        d.rx_n[to] = d.rx[from];
        d.ry_n[to] = d.ry[from];
        d.rz_n[to] = d.rz[from];
    }
    d.px_n[to] = d.px[from];
    d.py_n[to] = d.py[from];
    d.pz_n[to] = d.pz[from];
}

// Per cell: write the final list into the next buffers -- survivors and
// low arrivals in the scanned arrangement, then high arrivals in
// (source cell, rank) order.
__global__ __launch_bounds__(64) void k_rebox_gather(DevState d)
{
    __shared__ uint8_t sph[64][64];
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= d.ncells) return;
    uint8_t *phpos = sph[threadIdx.x];
    const int len = d.meta[4 * c + 0];
    const int nhigh = d.meta[4 * c + 2];
    const int keep = len < CM_SLOTS ? len : CM_SLOTS;
    for (int k = 0; k < keep; k++) {
        const int e = d.perm[c * CM_SLOTS + k];
        if (e < CM_SLOTS) copy_slot_to_next(d, c * CM_SLOTS + e, c * CM_SLOTS + k, false);
        else if (e - CM_SLOTS < CM_SLOTS) phpos[e - CM_SLOTS] = (uint8_t)k;
    }
    int src[26];
    const int nsrc = source_cells(d, c, src);
    int jl = 0, jh = 0;
    for (int q = 0; q < nsrc; q++) {
        const int b = src[q];
        const int nb = d.natoms[b];
        int last = -1;
        for (;;) {
            int best = -1, bestrank = 0x7fffffff;
            for (int s = 0; s < nb; s++) {
                const int o = b * CM_SLOTS + s;
                if (d.dest[o] == c) {
                    const int r = d.rank[o];
                    if (r > last && r < bestrank) { bestrank = r; best = o; }
                }
            }
            if (best < 0) break;
            last = bestrank;
            if (b < c) {
                if (jl < CM_SLOTS && jl < len) {
                    // placeholder jl sits at phpos[jl] unless it was cut off by overflow
                    const int k = phpos[jl];
                    if (k < keep) copy_slot_to_next(d, best, c * CM_SLOTS + k, true);
                }
                jl++;
            } else {
                const int k = len + jh;
                if (k < CM_SLOTS) copy_slot_to_next(d, best, c * CM_SLOTS + k, true);
                jh++;
            }
        }
    }
    int tot = len + nhigh;
    d.natoms_n[c] = tot < CM_SLOTS ? tot : CM_SLOTS;
}

// Fallback for a pass in which an arrival would move again: the whole pass
// is redone serially on the next buffers, exactly as reBoxAll does.
__global__ void k_rebox_serial_copy(DevState d)
{
    if (!(*d.flags & CM_FLAG_MULTIHOP)) return;
    const long ns = (long)d.ncells * CM_SLOTS;
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < ns; s += (long)gridDim.x * blockDim.x) {
        d.id_n[s] = d.id[s];
        d.itype_n[s] = d.itype[s];
        d.rx_n[s] = d.rx[s];
        d.ry_n[s] = d.ry[s];
        d.rz_n[s] = d.rz[s];
        d.px_n[s] = d.px[s];
        d.py_n[s] = d.py[s];
        d.pz_n[s] = d.pz[s];
        if ((s & (CM_SLOTS - 1)) == 0) d.natoms_n[s / CM_SLOTS] = d.natoms[s / CM_SLOTS];
    }
}

__device__ static void serial_copy(const DevState &d, int from, int to)
{
    d.id_n[to] = d.id_n[from];
    d.itype_n[to] = d.itype_n[from];
    d.rx_n[to] = d.rx_n[from];
    d.ry_n[to] = d.ry_n[from];
    d.rz_n[to] = d.rz_n[from];
    d.px_n[to] = d.px_n[from];
    d.py_n[to] = d.py_n[from];
    d.pz_n[to] = d.pz_n[from];
}

__global__ void k_rebox_serial(DevState d)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!(*d.flags & CM_FLAG_MULTIHOP)) return;
    unsigned fl = 0;
    for (int c = 0; c < d.ncells; c++) {
        int ix, iy, iz;
        cell_xyz(d, c, ix, iy, iz);
        int i = 0;
        while (i < d.natoms_n[c]) {
            const int o = c * CM_SLOTS + i;
            double x = d.rx_n[o], y = d.ry_n[o], z = d.rz_n[o];
            int nx = ix, ny = iy, nz = iz;
            double ax = x, ay = y, az = z;
            if (x < 0.0) { nx = ix - 1; ax = x + d.bs0; } else if (x >= d.bs0) { nx = ix + 1; ax = x - d.bs0; }
            if (y < 0.0) { ny = iy - 1; ay = y + d.bs1; } else if (y >= d.bs1) { ny = iy + 1; ay = y - d.bs1; }
            if (z < 0.0) { nz = iz - 1; az = z + d.bs2; } else if (z >= d.bs2) { nz = iz + 1; az = z - d.bs2; }
            const int j = cell_wrap(d, nx, ny, nz);
            if (j < 0 || j == c) {
                i++;
                continue;
            }
            d.rx_n[o] = ax;
            d.ry_n[o] = ay;
            d.rz_n[o] = az;
            const int nj = d.natoms_n[j];
            if (nj >= CM_SLOTS) {  // no room: the reference aborts here
                fl |= CM_FLAG_OVERFLOW;
                i++;
                continue;
            }
            serial_copy(d, o, j * CM_SLOTS + nj);
            d.natoms_n[j] = nj + 1;
            if (nj + 1 >= CM_SLOTS) fl |= CM_FLAG_OVERFLOW;
            const int ni = --d.natoms_n[c];
            if (ni) serial_copy(d, c * CM_SLOTS + ni, o);
        }
    }
    atomicAnd(d.flags, ~(unsigned)CM_FLAG_MULTIHOP);
    if (fl) atomicOr(d.flags, fl);
}

// ---------------------------------------------------------------------------
// Binning of a flat atom list as the sequence of putAtomInBox calls
// (src-flat/domains.c:237-291): wrap once into [0,L), cell =
// floor(x/boxsize), append.  Stable counting sort: atomic tickets per cell,
// then each cell's entries are ordered by input index.
__device__ static inline bool wrap_bin(const DevState &d, double &x, double &y, double &z, int &cell)
{
    if (x < 0.0) x += d.L0; else if (x >= d.L0) x -= d.L0;
    if (y < 0.0) y += d.L1; else if (y >= d.L1) y -= d.L1;
    if (z < 0.0) z += d.L2; else if (z >= d.L2) z -= d.L2;
    const int ix = (int)floor(x / d.bs0);
    const int iy = (int)floor(y / d.bs1);
    const int iz = (int)floor(z / d.bs2);
    if (ix < 0 || ix >= d.nbx0 || iy < 0 || iy >= d.nbx1 || iz < 0 || iz >= d.nbx2) return false;
    cell = ix + d.nbx0 * iy + d.nbx0 * d.nbx1 * iz;
    return true;
}

__global__ void k_bin_key(DevState d, int n, const double *xyz, int *cnt, int *tmp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    int cell;
    if (!wrap_bin(d, x, y, z, cell)) {
        atomicOr(d.flags, (unsigned)CM_FLAG_OUTSIDE);
        return;
    }
    const int t = atomicAdd(&cnt[cell], 1);
    if (t < CM_SLOTS) tmp[cell * CM_SLOTS + t] = i;
    else atomicOr(d.flags, (unsigned)CM_FLAG_OVERFLOW);
}

__global__ __launch_bounds__(64) void k_bin_place(DevState d, const double *xyz, const double *pxyz, int itype,
                                                  const int *cnt, const int *tmp)
{
    __shared__ int sidx[CM_SLOTS];
    const int c = blockIdx.x;
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
    wrap_bin(d, x, y, z, cell);
    double cx, cy, cz;
    cell_corner(d, c, cx, cy, cz);
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

void cm_launch_move(const DevState &d, hipStream_t st, double dtk, double dtd, double rmass, int kick, int drift)
{
    hipLaunchKernelGGL(k_move, dim3(d.ncells), dim3(64), 0, st, d, dtk, dtd, rmass, kick, drift);
}

void cm_launch_rebox(const DevState &d, hipStream_t st)
{
    const int nb = (d.ncells + 63) / 64;
    hipLaunchKernelGGL(k_rebox_scan, dim3(nb), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_rebox_gather, dim3(nb), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_rebox_serial_copy, dim3(1024), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_rebox_serial, dim3(1), dim3(64), 0, st, d);
}

void cm_launch_fill_int(int *p, int v, long n, hipStream_t st)
{
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
}

void cm_launch_bin(const DevState &d, hipStream_t st, int n, const double *xyz, const double *pxyz, int itype,
                   int *cnt, int *tmp)
{
    cm_launch_fill_int(cnt, 0, d.ncells, st);
    if (n > 0) hipLaunchKernelGGL(k_bin_key, dim3((n + 255) / 256), dim3(256), 0, st, d, n, xyz, cnt, tmp);
    hipLaunchKernelGGL(k_bin_place, dim3(d.ncells), dim3(64), 0, st, d, xyz, pxyz, itype, cnt, tmp);
}
