// reduce.hip -- deterministic two-stage reductions of per-atom potential
// energy, kinetic energy and virial: one partial triple per block, then one
// block sums the partials in a fixed order.
#include "comd_internal.h"

// Block b of nb: the partial (PE, KE, virial) of the slots s0 + b * 256 +
// thread, stepping by nb * 256, written to part[0..2].
__device__ __forceinline__ void reduce_part(const DevState &d, double rmass, int cell0, int ncells, int b, int nb,
                                            double *part)
{
    __shared__ double sh[3][256];
    double pe = 0, ke = 0, vi = 0;
    const long s0 = (long)cell0 * CM_SLOTS;
    const long ns = (long)ncells * CM_SLOTS;
    for (long s = s0 + b * 256L + threadIdx.x; s < s0 + ns; s += (long)nb * 256L) {
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
        for (int k = 0; k < 3; k++) part[k] = sh[k][0];
}

__global__ __launch_bounds__(256) void k_reduce(DevState d, double rmass, int cell0, int ncells, double *part)
{
    reduce_part(d, rmass, cell0, ncells, blockIdx.x, gridDim.x, part + 3 * blockIdx.x);
}

// Every replica at once: block (b, r) is block b of replica r's k_reduce
// (the same partition, so the same sums), partials at part + 3 (r nb + b).
__global__ __launch_bounds__(256) void k_reduce_batch(DevState d, const Grid *grids, double *part)
{
    const Grid &g = grids[blockIdx.y];
    reduce_part(d, g.rmass, g.cell_base, g.ncells, blockIdx.x, gridDim.x,
                part + 3 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x));
}

__device__ __forceinline__ void reduce_final(const double *part, int nparts, double *out)
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

__global__ __launch_bounds__(256) void k_reduce_final(const double *part, int nparts, double *out)
{
    reduce_final(part, nparts, out);
}

// Block r: replica r's k_reduce_final, out at out + 3 r.
__global__ __launch_bounds__(256) void k_reduce_final_batch(const double *part, int nparts, double *out)
{
    reduce_final(part + (size_t)3 * nparts * blockIdx.x, nparts, out + 3 * blockIdx.x);
}

void cm_launch_reduce(const DevState &d, const Grid &g, double *partials, int nparts, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce, dim3(nparts), dim3(256), 0, st, d, g.rmass, g.cell_base, g.ncells, partials);
}

void cm_launch_reduce_final(const double *partials, int nparts, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partials, nparts, out);
}

void cm_launch_reduce_batch(const DevState &d, const Grid *grids, int nrep, double *partials, int nparts, double *out,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_batch, dim3((unsigned)nparts, (unsigned)nrep), dim3(256), 0, st, d, grids, partials);
    hipLaunchKernelGGL(k_reduce_final_batch, dim3((unsigned)nrep), dim3(256), 0, st, partials, nparts, out);
}

// Block r copies replica r's staged series (6 m8 doubles, then 6 m8 floats,
// at src + r * 72 m8 bytes) to its series buffers.
__global__ __launch_bounds__(256) void k_scatter_series(const PotDev *pots, const char *src, int m8)
{
    const PotDev &P = pots[blockIdx.x];
    const double *s64 = (const double *)(src + (size_t)blockIdx.x * 72 * m8);
    const float *s32 = (const float *)(s64 + 6 * m8);
    double *d64 = (double *)P.cheb;
    float *d32 = (float *)P.cheb32;
    for (int i = threadIdx.x; i < 6 * m8; i += 256) {
        d64[i] = s64[i];
        d32[i] = s32[i];
    }
}

void cm_launch_scatter_series(const PotDev *pots, int nrep, const char *src, int m8, hipStream_t st)
{
    hipLaunchKernelGGL(k_scatter_series, dim3((unsigned)nrep), dim3(256), 0, st, pots, src, m8);
}

// Per replica r of a gradient call: dL/dE (emode: its weight w[r]; else
// 2 w_e (E_r - E_target,r) with E_r = E[3 r], the potential energy of the
// energy reduction) and the force-residual weight (emode: 0; else rw[r]).
__global__ __launch_bounds__(256) void k_adj_rep(const double *E, const double *w, const double *rw, int nrep,
                                                 int emode, double w_e, double *rep)
{
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nrep; r += gridDim.x * 256) {
        rep[r] = emode ? w[r] : 2.0 * w_e * (E[3 * r] - w[r]);
        rep[nrep + r] = emode ? 0.0 : rw[r];
    }
}

void cm_launch_adj_rep(const double *E, const double *w, const double *rw, int nrep, int emode, double w_e,
                       double *rep, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_rep, dim3((unsigned)((nrep + 255) / 256)), dim3(256), 0, st, E, w, rw, nrep, emode,
                       w_e, rep);
}
