// adj_basis.h -- the Chebyshev basis sums shared by the adjoint kernels
// (csrc/adjoint.hip over link cells, csrc/nlist.hip over the neighbour
// lists).  Series in the convention of src-flat/cheby.c:290-310.
#pragma once

#include "comd_internal.h"

#define AMAX 32   // coefficients per series supported by the adjoint
#define ANMOM 64  // moments written per block (two sets of AMAX)

// T_0..T_{m-1} at y, each times wa added to a[] (and times wb to b[]).
template <int MIXED, int TWO>
__device__ __forceinline__ void add_basis(double y, int m, double wa, double wb, double *a, double *b)
{
    if (MIXED) {
        const float yf = (float)y, y2 = 2.0f * yf;
        float tm = 1.0f, t = yf;
        a[0] += wa;
        if (TWO) b[0] += wb;
#pragma unroll
        for (int k = 1; k < AMAX; k++) {
            if (k < m) {
                a[k] += wa * (double)t;
                if (TWO) b[k] += wb * (double)t;
            }
            const float tn = y2 * t - tm;
            tm = t;
            t = tn;
        }
    } else {
        const double y2 = 2.0 * y;
        double tm = 1.0, t = y;
        a[0] += wa;
        if (TWO) b[0] += wb;
#pragma unroll
        for (int k = 1; k < AMAX; k++) {
            if (k < m) {
                a[k] += wa * t;
                if (TWO) b[k] += wb * t;
            }
            const double tn = y2 * t - tm;
            tm = t;
            t = tn;
        }
    }
}

// add_basis<MIXED, 0> that also returns the series c at y from the same
// basis values: c_0 / 2 + sum_k c_k T_k(y) (the convention of
// src-flat/cheby.c:290-310; c zero-padded past m, at an address common to
// the wave), in place of a separate Clenshaw sum over the same argument.
template <int MIXED>
__device__ __forceinline__ double add_basis_sum(double y, int m, double wa, double *a, const double *cg)
{
    const CM_CONST double *c = (const CM_CONST double *)cg;
    double s = 0.5 * c[0];
    a[0] += wa;
    if (MIXED) {
        const float yf = (float)y, y2 = 2.0f * yf;
        float tm = 1.0f, t = yf;
#pragma unroll
        for (int k = 1; k < AMAX; k++) {
            if (k < m) {
                a[k] += wa * (double)t;
                s += c[k] * (double)t;
            }
            const float tn = y2 * t - tm;
            tm = t;
            t = tn;
        }
    } else {
        const double y2 = 2.0 * y;
        double tm = 1.0, t = y;
#pragma unroll
        for (int k = 1; k < AMAX; k++) {
            if (k < m) {
                a[k] += wa * t;
                s += c[k] * t;
            }
            const double tn = y2 * t - tm;
            tm = t;
            t = tn;
        }
    }
    return s;
}

// T_k(y) times wt plus the derivative of T_k with respect to y, k U_{k-1}(y)
// (U: the second-kind polynomials, U_0 = 1, U_1 = 2y), times wd, added to
// a[k], k < m.  With wd = w / (b - a) (the domain's dy/dr = 2 / (b - a)
// times w / 2) this is, for the series' basis B_k (B_0 = T_0 / 2 applied by
// the caller), the pair's term of the energy moments and of the derivative
// operator applied to the force moments in one set.
template <int MIXED>
__device__ __forceinline__ void add_basis_deriv(double y, int m, double wt, double wd, double *a)
{
    a[0] += wt;
    if (MIXED) {
        const float yf = (float)y, y2 = 2.0f * yf;
        float tm = 1.0f, t = yf, um = 1.0f, u = y2;  // T_{k-1}, T_k, U_{k-1}, U_k at k = 1
#pragma unroll
        for (int k = 1; k < AMAX; k++) {
            if (k < m) {
                a[k] += wt * (double)t;
                a[k] += wd * (double)k * (double)um;
            }
            const float tn = y2 * t - tm, un = y2 * u - um;
            tm = t;
            t = tn;
            um = u;
            u = un;
        }
    } else {
        const double y2 = 2.0 * y;
        double tm = 1.0, t = y, um = 1.0, u = y2;
#pragma unroll
        for (int k = 1; k < AMAX; k++) {
            if (k < m) {
                a[k] += wt * t;
                a[k] += wd * (double)k * um;
            }
            const double tn = y2 * t - tm, un = y2 * u - um;
            tm = t;
            t = tn;
            um = u;
            u = un;
        }
    }
}

__device__ __forceinline__ double cheb_arg(double x, double a, double b)
{
    if (x < a) x = a;  // lower clamp only (src-flat/cheby.c:290-293)
    return (2.0 * x - a - b) / (b - a);
}

// Sum of v over the wave's 64 lanes (result in every lane).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}


// The moments a[k] summed over the block's waves, in wave order through LDS
// (red: one row of ANMOM doubles per wave), and written by threads 0 .. 63
// to part[0 .. ANMOM): a block's partials in a fixed order.  SET 0: a[] is
// the first set of AMAX moments, the second is zero; SET 1: the reverse.
template <int SET>
__device__ __forceinline__ void block_moments(const double *a, double (*red)[ANMOM], double *part)
{
    const int L = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < AMAX; k++) {
        const double s = wave_sum(a[k]);
        if (L == 0) {
            red[w][SET * AMAX + k] = s;
            red[w][(1 - SET) * AMAX + k] = 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x < ANMOM) {
        double s = 0.0;
        for (int q = 0; q < nw; q++) s += red[q][threadIdx.x];
        part[threadIdx.x] = s;
    }
}
