// observe.hip -- observables beyond the reference's: the full virial tensor
// (pair, embedding and kinetic terms), from which pressure and stress
// follow, with the potential and kinetic energy and the reference's scalar
// "stress" in one deterministic reduction.
//
// The reference computes only a partial virial: the xx component of the
// EAM embedding term (src-flat/eam.c:649, divided by the volume at 657) and,
// for Lennard-Jones, the pair xx term with a kinetic term counted once per
// non-empty neighbour cell (ljForce.c:175,210).  Here, with the physical
// force f_i = -grad_i U (whatever sign convention the force array uses),
//
//   W_ab = sum_i p_ia p_ib / m + sum_{i<j} r_ij,a f_ij,b ,   r_ij = r_i - r_j,
//
// where f_ij is the force on i from the pair term of i and j:
//   EAM  f_ij = -(phi'(r) + (F'_i + F'_j) rho'(r)) r_ij / r
//   LJ   f_ij = 4 eps s6 r^-8 (12 s6 r^-6 - 6) r_ij.
// The pair sum runs over the link-cell stencil of each atom with the cutoff
// predicates of the force passes (phi: r^2 <= rc^2, eam.c:519; rho':
// r^2 < rc^2, eam.c:633; LJ: r^2 <= rc^2, ljForce.c:188), half of each pair
// to each atom.  F' is that of the last force evaluation (the entry points
// run one first).  Not on the timestep path: a thread per atom, neighbours
// read through the caches, FP64 throughout.
#include <algorithm>

#include "comd_internal.h"
#include "pair_terms.h"

#define OBS_BLOCK 256

// One block's partial sums, in this order (CM_OBS_N doubles):
//   0-5   pair + embedding virial  xx yy zz xy xz yz
//   6-11  kinetic term             xx yy zz xy xz yz
//   12    potential energy (sum of per-atom energies, f[3])
//   13    kinetic energy
//   14    the reference's "stress" sum (before the volume)
//   15    atoms
template <int KIND, int EVAL>
__global__ __launch_bounds__(OBS_BLOCK) void k_observe(DevState d, PotDev P, Grid g, double *part)
{
    __shared__ double sh[CM_OBS_N][OBS_BLOCK];
    double a[CM_OBS_N];
#pragma unroll
    for (int k = 0; k < CM_OBS_N; k++) a[k] = 0.0;
    const long s0 = (long)g.cell_base * CM_SLOTS, ns = (long)g.ncells * CM_SLOTS;
    const double rc2 = P.rc2;
    for (long s = s0 + blockIdx.x * (long)OBS_BLOCK + threadIdx.x; s < s0 + ns; s += (long)gridDim.x * OBS_BLOCK) {
        const int c = (int)(s / CM_SLOTS), i = (int)(s % CM_SLOTS);
        if (i >= d.natoms[c]) continue;
        const double xi = d.rx[s], yi = d.ry[s], zi = d.rz[s];
        const double fpi = KIND == CM_KIND_LJ ? 0.0 : d.fi[s];
        double w[6] = {0, 0, 0, 0, 0, 0};
        for (int e = 0; e < 27; e++) {
            const int b = stencil_cell(g, c, e);
            if (b < 0) continue;
            double sx, sy, sz;
            corner_shift(g, c, b, sx, sy, sz);
            const int nb = d.natoms[b];
            for (int j = 0; j < nb; j++) {
                if (b == c && j == i) continue;
                const long t = (long)b * CM_SLOTS + j;
                const double dx = sx + xi - d.rx[t], dy = sy + yi - d.ry[t], dz = sz + zi - d.rz[t];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > rc2) continue;
                double k;  // r_ij (x) f_ij = k dr (x) dr
                if (KIND == CM_KIND_LJ) {
                    const double r2i = 1.0 / r2, r6 = r2i * r2i * r2i, s6 = P.lj_s6;
                    k = 4.0 * P.lj_eps * s6 * r6 * r2i * (12.0 * s6 * r6 - 6.0);
                } else {
                    const double r = sqrt(r2);
                    double v, dphi, drho = 0.0;
                    if (EVAL == COMD_EVAL_CHEB) {
                        dphi = cheb64(P.cheb + 3 * P.cstride, P.cstride, P.ca[0], P.cb[0], r);
                        if (r2 < rc2) drho = cheb64(P.cheb + 4 * P.cstride, P.cstride, P.ca[1], P.cb[1], r);
                    } else {
                        interp64(P.tab[0], P.x0[0], P.xn[0], P.invdx[0], P.n[0], r, v, dphi);
                        if (r2 < rc2) interp64(P.tab[1], P.x0[1], P.xn[1], P.invdx[1], P.n[1], r, v, drho);
                    }
                    k = -(dphi + (fpi + d.fi[t]) * drho) / r;
                }
                k *= 0.5;  // half of the pair to each atom
                w[0] += k * dx * dx;
                w[1] += k * dy * dy;
                w[2] += k * dz * dz;
                w[3] += k * dx * dy;
                w[4] += k * dx * dz;
                w[5] += k * dy * dz;
            }
        }
        const double px = d.px[s], py = d.py[s], pz = d.pz[s], im = 1.0 / g.rmass;
#pragma unroll
        for (int q = 0; q < 6; q++) a[q] += w[q];
        a[6] += px * px * im;
        a[7] += py * py * im;
        a[8] += pz * pz * im;
        a[9] += px * py * im;
        a[10] += px * pz * im;
        a[11] += py * pz * im;
        a[12] += d.fe[s];
        a[13] += (px * px + py * py + pz * pz) / (2.0 * g.rmass);
        a[14] += d.vir[s];
        a[15] += 1.0;
    }
#pragma unroll
    for (int k = 0; k < CM_OBS_N; k++) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = OBS_BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < CM_OBS_N; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < CM_OBS_N) part[(size_t)CM_OBS_N * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// The blocks' partial sums added in block order (one block of 256 threads:
// thread t sums blocks t, t + 256, ... then a tree over the threads), and
// the observables formed from the sums (pe_scale: 4 eps sigma^6 for LJ,
// whose per-atom energies lack it, src-flat/ljForce.c:200-201,217; 1 for
// EAM).
__global__ __launch_bounds__(OBS_BLOCK) void k_observe_finish(const double *part, int nparts, double volume,
                                                              double pe_scale, comd_observables *out)
{
    __shared__ double sh[CM_OBS_N][OBS_BLOCK];
    double a[CM_OBS_N];
    for (int k = 0; k < CM_OBS_N; k++) a[k] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += OBS_BLOCK)
        for (int k = 0; k < CM_OBS_N; k++) a[k] += part[(size_t)CM_OBS_N * b + k];
    for (int k = 0; k < CM_OBS_N; k++) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = OBS_BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < CM_OBS_N; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    auto at = [&](int k) { return sh[k][0]; };
    comd_observables r;
    r.pe = at(12) * pe_scale;
    r.ke = at(13);
    r.etot = r.pe + r.ke;
    r.natoms = (int)at(15);
    r.temperature = r.natoms > 0 ? 2.0 * r.ke / (3.0 * (double)r.natoms * CM_KB_HARTREE) : 0.0;
    // xx yy zz xy xz yz -> row-major
    const int map[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int k = 0; k < 9; k++) {
        r.virial_pair[k] = at(map[k]);
        r.virial[k] = at(map[k]) + at(6 + map[k]);
    }
    r.pressure = (r.virial[0] + r.virial[4] + r.virial[8]) / (3.0 * volume);
    r.stress_ref = at(14) / volume;
    r.volume = volume;
    *out = r;
}

int cm_observe_blocks(const Grid &g)
{
    const long b = ((long)g.ncells * CM_SLOTS + OBS_BLOCK - 1) / OBS_BLOCK;
    return (int)std::min<long>(std::max<long>(b, 1), CM_OBS_MAXBLOCKS);
}

void cm_launch_observe(const DevState &d, const PotDev &P, const Grid &g, double pe_scale, double *part,
                       comd_observables *out, hipStream_t st)
{
    const int nb = cm_observe_blocks(g);
    const dim3 grid(nb), block(OBS_BLOCK);
    if (P.kind == COMD_POT_LJ) hipLaunchKernelGGL((k_observe<CM_KIND_LJ, COMD_EVAL_TABLE>), grid, block, 0, st, d, P, g, part);
    else if (P.eval == COMD_EVAL_CHEB)
        hipLaunchKernelGGL((k_observe<CM_KIND_EAM1, COMD_EVAL_CHEB>), grid, block, 0, st, d, P, g, part);
    else hipLaunchKernelGGL((k_observe<CM_KIND_EAM1, COMD_EVAL_TABLE>), grid, block, 0, st, d, P, g, part);
    hipLaunchKernelGGL(k_observe_finish, dim3(1), block, 0, st, part, nb, g.L0 * g.L1 * g.L2, pe_scale, out);
}
