// reduce.hip -- deterministic two-stage reductions of per-atom potential
// energy, kinetic energy and virial: one partial triple per block, then one
// block sums the partials in a fixed order.
#include "comd_internal.h"

__global__ __launch_bounds__(256) void k_reduce(DevState d, double rmass, int cell0, int ncells, double *part)
{
    __shared__ double sh[3][256];
    double pe = 0, ke = 0, vi = 0;
    const long s0 = (long)cell0 * CM_SLOTS;
    const long ns = (long)ncells * CM_SLOTS;
    for (long s = s0 + blockIdx.x * 256L + threadIdx.x; s < s0 + ns; s += (long)gridDim.x * 256L) {
        const int c = (int)(s / CM_SLOTS);
        const int i = (int)(s % CM_SLOTS);
        if (i < d.natoms[c]) {
            pe += d.fe[s];
            vi += d.vir[s];
            const double px = d.px[s], py = d.py[s], pz = d.pz[s];
            ke += (px * px + py * py + pz * pz) / (2.0 * rmass);
        }
    }
    sh[0][threadIdx.x] = pe;
    sh[1][threadIdx.x] = ke;
    sh[2][threadIdx.x] = vi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[3 * blockIdx.x + k] = sh[k][0];
}

__global__ __launch_bounds__(256) void k_reduce_final(const double *part, int nparts, double *out)
{
    __shared__ double sh[3][256];
    double a[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += 256)
        for (int k = 0; k < 3; k++) a[k] += part[3 * b + k];
    for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[k] = sh[k][0];
}

void cm_launch_reduce(const DevState &d, const Grid &g, double *partials, int nparts, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce, dim3(nparts), dim3(256), 0, st, d, g.rmass, g.cell_base, g.ncells, partials);
}

void cm_launch_reduce_final(const double *partials, int nparts, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partials, nparts, out);
}
