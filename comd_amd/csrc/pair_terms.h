// pair_terms.h -- the per-pair arithmetic of the force kernels (table and
// Chebyshev evaluation, EAM density/pair and embedding-force terms,
// Lennard-Jones), shared by the link-cell kernels (force.hip) and the
// neighbour-list kernels (nlist.hip).  Internal to comd_amd/csrc.
#pragma once

#include "comd_internal.h"

// ---------------------------------------------------------------------------
// potential evaluation

// 3-point quadratic value and interpolated central-difference derivative,
// the formula of src-flat/eam.c:401-434 in double precision (used once per
// atom for F).  v points at values[-1].
__device__ __forceinline__ void interp64(const double *vg, double x0, double xn, double invdx, int n, double r,
                                         double &val, double &der)
{
    const CM_GLOBAL double *v = (const CM_GLOBAL double *)vg;
    r = fmin(fmax(r, x0), xn);
    r = (r - x0) * invdx;
    // the reference reads interval i = n at r = xn (t = 0); the same value
    // and derivative are the end (t = 1) of interval n - 1, whose stencil
    // stays inside values[-1..n+1]
    const int i = min(max((int)floor(r), 0), n - 1);
    r = r - (double)i;
    const CM_GLOBAL double *w = v + 1 + i;
    const double vm = w[-1], v0 = w[0], v1 = w[1], v2 = w[2];
    const double g0 = v1 - vm;
    const double g1 = v2 - v0;
    val = v0 + 0.5 * r * (r * (v1 + vm - 2.0 * v0) + g0);
    der = (i <= 0) ? 0.0 : 0.5 * (g0 + r * (g1 - g0)) * invdx;
}

// Interval of the table formula that holds r (clamped to [x0, xn] as in
// src-flat/eam.c:413-414) and the fraction t inside it.
__device__ __forceinline__ int table_interval(double r, double x0, double xn, double invdx, int n, double &t)
{
    r = fmin(fmax(r, x0), xn);
    const double u = (r - x0) * invdx;
    const int i = min(max((int)floor(u), 0), n - 1);  // r = xn: the end of interval n - 1
    t = u - (double)i;
    return i;
}

// Table formula from per-interval coefficients (d0 d1 c0 c1 c2 0):
// value c0 + t(c1 + t c2), derivative d0 + t d1.  PD / PF are pointers to
// double / float in any address space (generic, global or LDS), so that a
// caller that knows where the rows live gets global or LDS loads rather
// than flat ones (flat loads of LDS data still occupy the texture
// addresser).
template <class PD> __device__ __forceinline__ double2 cf_der64(PD cf, int i)
{
    return make_double2(cf[6 * i], cf[6 * i + 1]);
}
template <class PD> __device__ __forceinline__ double cf_val64(PD cf, int i, double t)
{
    return fma(t, fma(t, cf[6 * i + 4], cf[6 * i + 3]), cf[6 * i + 2]);
}
template <class PF> __device__ __forceinline__ float4 cf_row32(PF cf, int i)
{
    return make_float4(cf[8 * i], cf[8 * i + 1], cf[8 * i + 2], cf[8 * i + 3]);
}

// 1/sqrt(x) in double precision: hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// 1/sqrt(x): the hardware double estimate (about 5e-8 relative on gfx950,
// no better than the single-precision one but without the two conversions)
// and one Newton step in double (about 4e-15).
__device__ __forceinline__ double rsqrt_mixed(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// Compact rows of the neighbour-list passes (PotDev::cd64, cr64), behind
// pointers of a given address space.
template <class PD, class PF> struct CompactT {
    PD d64, r64;
    PF dm;  // mixed density rows (40 bytes: phi d0 d1 in double, 6 floats)
};
using CompactGlobal = CompactT<const CM_GLOBAL double *, const CM_GLOBAL float *>;
using CompactLds = CompactT<const CM_LDS double *, const CM_LDS float *>;

// ---------------------------------------------------------------------------
// one accepted pair (home atom a, neighbour j): dx = r_a - r_j

// Per-interval coefficient rows of the phi and rho tables, in global memory
// or staged in LDS by the caller, behind pointers of a given address space.
template <class PD, class PF> struct TabsT {
    PD c64[2];
    PF c32[2];
};
using Tabs = TabsT<const double *, const float *>;
using TabsGlobal = TabsT<const CM_GLOBAL double *, const CM_GLOBAL float *>;
using TabsLds = TabsT<const CM_LDS double *, const CM_LDS float *>;

__device__ __forceinline__ TabsGlobal global_tabs(const PotDev &P)
{
    TabsGlobal t;
    t.c64[0] = (const CM_GLOBAL double *)P.cf64[0];
    t.c64[1] = (const CM_GLOBAL double *)P.cf64[1];
    t.c32[0] = (const CM_GLOBAL float *)P.cf32[0];
    t.c32[1] = (const CM_GLOBAL float *)P.cf32[1];
    return t;
}

struct Acc {
    double x, y, z, e, rho, vir;
};

template <int KIND, int EVAL, int MIXED, class TB>
__device__ __forceinline__ void pair_term(const PotDev &P, const TB &T, double dx, double dy, double dz, double r2,
                                          double fia, double fij, Acc &A)
{
    if (KIND == CM_KIND_LJ) {
        // src-flat/ljForce.c:196-210 (f holds -grad U, energy without 4 eps s6)
        const double s6 = P.lj_s6;
        double r2i;
        if (MIXED) {
            const double y = rsqrt_mixed(r2);
            r2i = y * y;
        } else {
            r2i = 1.0 / r2;
        }
        const double r6 = r2i * r2i * r2i;
        A.e += 0.5 * r6 * (s6 * r6 - 1.0);
        const double fr = 4.0 * P.lj_eps * s6 * r6 * r2i * (12.0 * s6 * r6 - 6.0);
        A.x += dx * fr;
        A.y += dy * fr;
        A.z += dz * fr;
        A.vir += fr * dx * dx;
        return;
    }
    const double y = MIXED ? rsqrt_mixed(r2) : rsqrt64(r2);  // 1/r
    const double r = r2 * y;
    if (KIND == CM_KIND_EAM1) {
        // src-flat/eam.c:521-545 (f holds +grad U)
        double phiv, dphi, rhov;
        if (EVAL == COMD_EVAL_CHEB) {
            const int m = P.cstride;
            if (MIXED) {
                // values in FP32; the derivative, whose pair terms cancel in
                // the force sum, in FP64 (the same issue rate on this device)
                const float rf = (float)r;
                const float a0 = (float)P.ca[0], b0 = (float)P.cb[0], a1 = (float)P.ca[1], b1 = (float)P.cb[1];
                phiv = cheb32(P.cheb32 + 0 * m, m, a0, b0, rf);
                dphi = cheb64(P.cheb + 3 * m, m, P.ca[0], P.cb[0], r);
                rhov = cheb32(P.cheb32 + 1 * m, m, a1, b1, rf);
            } else {
                phiv = cheb64(P.cheb + 0 * m, m, P.ca[0], P.cb[0], r);
                dphi = cheb64(P.cheb + 3 * m, m, P.ca[0], P.cb[0], r);
                rhov = cheb64(P.cheb + 1 * m, m, P.ca[1], P.cb[1], r);
            }
        } else {
            double t0, t1;
            const int i0 = table_interval(r, P.x0[0], P.xn[0], P.invdx[0], P.n[0], t0);
            int i1 = i0;
            t1 = t0;
            if (!P.same_grid) i1 = table_interval(r, P.x0[1], P.xn[1], P.invdx[1], P.n[1], t1);
            if (MIXED) {
                // values in FP32, the derivative in FP64 (see above)
                const float4 a = cf_row32(T.c32[0], i0);
                const float a2 = T.c32[0][8 * i0 + 4];
                const float4 b = cf_row32(T.c32[1], i1);
                const float b2 = T.c32[1][8 * i1 + 4];
                const float f0 = (float)t0, f1 = (float)t1;
                const double2 dd = cf_der64(T.c64[0], i0);
                phiv = fmaf(f0, fmaf(f0, a2, a.w), a.z);
                dphi = fma(t0, dd.y, dd.x);
                rhov = fmaf(f1, fmaf(f1, b2, b.w), b.z);
            } else {
                const double2 dd = cf_der64(T.c64[0], i0);
                phiv = cf_val64(T.c64[0], i0, t0);
                dphi = fma(t0, dd.y, dd.x);
                rhov = cf_val64(T.c64[1], i1, t1);
            }
        }
        const double s = dphi * y;
        A.x += s * dx;
        A.y += s * dy;
        A.z += s * dz;
        A.e += 0.5 * phiv;
        A.rho += rhov;
    } else {
        // src-flat/eam.c:633-649
        double drho;
        if (EVAL == COMD_EVAL_CHEB) {
            drho = cheb64(P.cheb + 4 * P.cstride, P.cstride, P.ca[1], P.cb[1], r);  // FP64 in both modes
        } else {
            double t1;
            const int i1 = table_interval(r, P.x0[1], P.xn[1], P.invdx[1], P.n[1], t1);
            const double2 dd = cf_der64(T.c64[1], i1);  // FP64 in both modes
            drho = fma(t1, dd.y, dd.x);
        }
        const double g = (fia + fij) * drho * y;
        A.x += g * dx;
        A.y += g * dy;
        A.z += g * dz;
        A.vir += g * dx * dx;
    }
}

// ---------------------------------------------------------------------------

// The table branch of pair_term for EAM with phi and rho on one grid, from
// the compact FP64 rows (the formulas of src-flat/eam.c:401-434 and 521-545 /
// 633-649).  The interval and fraction: u = (r - x0) invDx as one fma,
// clamped below at 0; r <= rc <= xn holds for every accepted pair, so the
// interval is at most n (r = xn: row n, the last interval re-centred on its
// end, gives the reference's value at i = n, t = 0).  The derivatives, whose pair
// terms cancel in the force sums, are FP64 in both modes; MIXED evaluates the
// density pass's values (phi, rho) in FP32 from 40-byte rows, which halves
// the rows' LDS traffic.
__device__ __forceinline__ int compact_interval(double u, double &t)
{
    u = fmax(u, 0.0);
    t = __builtin_amdgcn_fract(u);  // u - floor(u), exact
    return (int)u;
}

template <int KIND, int MIXED, class CT>
__device__ __forceinline__ void pair_term_compact(const PotDev &P, const CT &T, double dx, double dy, double dz,
                                                  double r2, double fia, double fij, Acc &A)
{
    const double y = MIXED ? rsqrt_mixed(r2) : rsqrt64(r2);  // 1/r
    const double r = r2 * y;
    double t;
    const int i = compact_interval(fma(r, P.invdx[0], -P.x0[0] * P.invdx[0]), t);
    if (KIND == CM_KIND_EAM1) {
        double phiv, dphi, rhov;
        if (MIXED) {
            const double pd0 = T.d64[5 * i + 0], pd1 = T.d64[5 * i + 1];
            const float f = (float)t;
            const float pc0 = T.dm[10 * i + 4], pc1 = T.dm[10 * i + 5], pc2 = T.dm[10 * i + 6],
                        rc0 = T.dm[10 * i + 7], rc1 = T.dm[10 * i + 8], rc2 = T.dm[10 * i + 9];
            phiv = fmaf(f, fmaf(f, pc2, pc1), pc0);
            dphi = fma(t, pd1, pd0);
            rhov = fmaf(f, fmaf(f, rc2, rc1), rc0);
        } else {
            const double pd0 = T.d64[8 * i + 0], pd1 = T.d64[8 * i + 1], pc0 = T.d64[8 * i + 2],
                         pc1 = T.d64[8 * i + 3], pc2 = T.d64[8 * i + 4], rc0 = T.d64[8 * i + 5],
                         rc1 = T.d64[8 * i + 6], rc2 = T.d64[8 * i + 7];
            phiv = fma(t, fma(t, pc2, pc1), pc0);
            dphi = fma(t, pd1, pd0);
            rhov = fma(t, fma(t, rc2, rc1), rc0);
        }
        const double s = dphi * y;
        A.x += s * dx;
        A.y += s * dy;
        A.z += s * dz;
        A.e += 0.5 * phiv;
        A.rho += rhov;
    } else {
        const double drho = fma(t, T.r64[2 * i + 1], T.r64[2 * i]);
        const double g = (fia + fij) * drho * y;
        A.x += g * dx;
        A.y += g * dy;
        A.z += g * dz;
        A.vir += g * dx * dx;
    }
}

// The same for the mixed passes, with the pair's separation in fixed-point
// units (dx = dxu * qinv, r2 = r2u * qinv^2, qinv a power of two): 1/r in
// units yu = y * qinv, so the force terms dphi * y * dx equal dphi * yu * dxu
// exactly and the unit scale only enters the table coordinate, through
// iu = qinv * invDx.  The virial sum comes out in units of 1/qinv (the
// caller scales it).  The density pass's FP32 values (pair energy, density)
// go to the caller's FP32 sums ef, rf over one list chunk (8 pairs), which
// the caller adds to the FP64 sums.
template <int KIND, class CT>
__device__ __forceinline__ void pair_term_units(const PotDev &P, const CT &T, double dxu, double dyu, double dzu,
                                                double r2u, double iu, double cu, double fia, double fij, Acc &A,
                                                float &ef, float &rf)
{
    const double yu = rsqrt_mixed(r2u);
    double t;
    const int i = compact_interval(fma(r2u * yu, iu, cu), t);
    if (KIND == CM_KIND_EAM1) {
        // row offset in doubles by a 24-bit multiply (i <= n < 2^24)
        const int o = __mul24(i, 5);
        const double pd0 = T.d64[o + 0], pd1 = T.d64[o + 1];
        const float f = (float)t;
        const float pc0 = T.dm[2 * o + 4], pc1 = T.dm[2 * o + 5], pc2 = T.dm[2 * o + 6],
                    rc0 = T.dm[2 * o + 7], rc1 = T.dm[2 * o + 8], rc2 = T.dm[2 * o + 9];
        const double s = fma(t, pd1, pd0) * yu;
        A.x += s * dxu;
        A.y += s * dyu;
        A.z += s * dzu;
        ef = fmaf(0.5f, fmaf(f, fmaf(f, pc2, pc1), pc0), ef);
        rf += fmaf(f, fmaf(f, rc2, rc1), rc0);
    } else {
        const double g = (fia + fij) * fma(t, T.r64[2 * i + 1], T.r64[2 * i]) * yu;
        A.x += g * dxu;
        A.y += g * dyu;
        A.z += g * dzu;
        A.vir += g * dxu * dxu;
    }
}
