// centro.hip -- the centrosymmetry parameter of src-viz's in-situ viewer
// (computeCentrosymmetry, src-viz/AtomVisualize.c:399-673), the one other
// per-atom compute kernel of the reference, on the device state.
//
// Per atom: its 12 nearest neighbours, found as the reference finds them --
// the atoms of its own cell first (separations from r alone, 416-446), then
// those of the 26 surrounding cells, x outer, y, z inner, wrapped
// periodically with the box length rounded to float as the image offset
// (448-503); a candidate takes a free place while fewer than 12 are held,
// else replaces the farthest held one (the first of equal ones) unless it
// is farther.  Fewer than 12: the parameter is 1000 (508).  Otherwise the
// smallest, over the 10395 ways of pairing the 12 neighbours, of the sum of
// |r_j + r_k|^2 over the pairs, each pairing's sum accumulated in float
// with the pairs in increasing order of their first index (the exact branch,
// 540-651): enumerated depth first here, so that the float partial sums
// of pairings that share their first pairs are formed once.  Contraction
// is off in this file, so the sums round as the reference's do.
//
// out holds one value per slot of the replica's own cells (its first cell
// at 0).  Four cells per wave, 16 lanes each (a lane per atom, further rounds for
// cells above 16 atoms); each lane keeps its 12 neighbours and the 66 pair
// sums in LDS, lane-major.
#include "comd_internal.h"

#define CS_WAVES 1  // waves per block

struct CentroLds {
    double r[12][3][64];  // separations of the 12 neighbours, lane-major
    double rsq[12][64];
    double s[66][64];     // |r_j + r_k|^2, j < k, in row order
};

__device__ __forceinline__ int pair_index(int j, int k)  // j < k
{
    return j * 11 - (j * (j - 1)) / 2 + (k - j - 1);
}

__global__ __launch_bounds__(64) void k_centro(DevState d, Grid g, double *out)
{
    __shared__ CentroLds L;
    const int lane = threadIdx.x;
    const int c = g.cell_base + blockIdx.x * 4 + (lane >> 4);
    if (c >= g.cell_base + g.ncells) return;
    const int nbx = g.nbx0, nby = g.nbx1, nbz = g.nbx2;
    const int lc = c - g.cell_base;
    const int bz = lc / (nbx * nby), by = (lc - bz * nbx * nby) / nbx, bx = lc - bz * nbx * nby - by * nbx;
    const int na = d.natoms[c];
    double cix, ciy, ciz;
    corner_of(g, bx, by, bz, cix, ciy, ciz);
    // the box length rounded to float, the reference's image offsets
    const double Lx = (double)(float)g.L0, Ly = (double)(float)g.L1, Lz = (double)(float)g.L2;
    for (int t = lane & 15; t < na; t += 16) {
        const int io = c * CM_SLOTS + t;
        const double xi = d.rx[io], yi = d.ry[io], zi = d.rz[io];
        int n = 0, kmax = 0;
        double rmax = 0.0;
        auto offer = [&](double dx, double dy, double dz) {
            const double rsq = dx * dx + dy * dy + dz * dz;
            int i1;
            if (n < 12) {
                i1 = n++;
            } else {
                if (rsq > rmax) return;
                i1 = kmax;
            }
            L.r[i1][0][lane] = dx;
            L.r[i1][1][lane] = dy;
            L.r[i1][2][lane] = dz;
            L.rsq[i1][lane] = rsq;
            if (n == 12) {  // the farthest held one (the first of equal ones)
                kmax = 0;
                rmax = L.rsq[0][lane];
                for (int k = 1; k < 12; k++) {
                    const double v = L.rsq[k][lane];
                    if (v > rmax) {
                        kmax = k;
                        rmax = v;
                    }
                }
            }
        };
        for (int j = 0; j < na; j++) {
            if (j == t) continue;
            const int jo = c * CM_SLOTS + j;
            offer(d.rx[jo] - xi, d.ry[jo] - yi, d.rz[jo] - zi);
        }
        for (int cx = bx - 1; cx <= bx + 1; cx++)
            for (int cy = by - 1; cy <= by + 1; cy++)
                for (int cz = bz - 1; cz <= bz + 1; cz++) {
                    if (cx == bx && cy == by && cz == bz) continue;
                    int nx = cx, ny = cy, nz = cz;
                    double ox = 0.0, oy = 0.0, oz = 0.0;
                    if (nx < 0) { nx = nbx - 1; ox = -Lx; }
                    if (nx >= nbx) { nx = 0; ox = Lx; }
                    if (ny < 0) { ny = nby - 1; oy = -Ly; }
                    if (ny >= nby) { ny = 0; oy = Ly; }
                    if (nz < 0) { nz = nbz - 1; oz = -Lz; }
                    if (nz >= nbz) { nz = 0; oz = Lz; }
                    const int b2 = g.cell_base + nz * nbx * nby + ny * nbx + nx;
                    double cjx, cjy, cjz;
                    corner_of(g, nx, ny, nz, cjx, cjy, cjz);
                    const int nb = d.natoms[b2];
                    for (int j = 0; j < nb; j++) {
                        const int jo = b2 * CM_SLOTS + j;
                        offer((ox + cjx + d.rx[jo]) - (cix + xi), (oy + cjy + d.ry[jo]) - (ciy + yi),
                              (oz + cjz + d.rz[jo]) - (ciz + zi));
                    }
                }
        if (n < 12) {
            out[lc * CM_SLOTS + t] = 1000.0;
            continue;
        }
        for (int j = 0; j < 11; j++)
            for (int k = j + 1; k < 12; k++) {
                const double a = L.r[j][0][lane] + L.r[k][0][lane], b = L.r[j][1][lane] + L.r[k][1][lane],
                             e = L.r[j][2][lane] + L.r[k][2][lane];
                L.s[pair_index(j, k)][lane] = a * a + b * b + e * e;
            }
        // depth first over the pairings: at each level the lowest unpaired
        // neighbour pairs with each other unpaired one
        float best = 1000.0f;
        auto sp = [&](int j, int k) { return L.s[pair_index(j, k)][lane]; };
        auto addf = [](float acc, double v) { return __double2float_rn((double)acc + v); };
        const int m0 = 0xffe;  // neighbour 0 pairs first
        for (int q0 = m0; q0; q0 &= q0 - 1) {
            const int p0 = __builtin_ctz(q0);
            const float a1 = addf(0.0f, sp(0, p0));
            const int r1 = m0 & ~(1 << p0);
            const int n1 = __builtin_ctz(r1), m1 = r1 & ~(1 << n1);
            for (int q1 = m1; q1; q1 &= q1 - 1) {
                const int p1 = __builtin_ctz(q1);
                const float a2 = addf(a1, sp(n1, p1));
                const int r2 = m1 & ~(1 << p1);
                const int n2 = __builtin_ctz(r2), m2 = r2 & ~(1 << n2);
                for (int q2 = m2; q2; q2 &= q2 - 1) {
                    const int p2 = __builtin_ctz(q2);
                    const float a3 = addf(a2, sp(n2, p2));
                    const int r3 = m2 & ~(1 << p2);
                    const int n3 = __builtin_ctz(r3), m3 = r3 & ~(1 << n3);
                    for (int q3 = m3; q3; q3 &= q3 - 1) {
                        const int p3 = __builtin_ctz(q3);
                        const float a4 = addf(a3, sp(n3, p3));
                        const int r4 = m3 & ~(1 << p3);
                        const int n4 = __builtin_ctz(r4), m4 = r4 & ~(1 << n4);
                        for (int q4 = m4; q4; q4 &= q4 - 1) {
                            const int p4 = __builtin_ctz(q4);
                            const float a5 = addf(a4, sp(n4, p4));
                            const int r5 = m4 & ~(1 << p4);
                            const int n5 = __builtin_ctz(r5), p5 = __builtin_ctz(r5 & ~(1 << n5));
                            const float a6 = addf(a5, sp(n5, p5));
                            best = a6 < best ? a6 : best;
                        }
                    }
                }
            }
        }
        out[lc * CM_SLOTS + t] = (double)best;
    }
}

void cm_launch_centro(const DevState &d, const Grid &g, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_centro, dim3((unsigned)((g.ncells + 3) / 4)), dim3(64), 0, st, d, g, out);
}
