// force.hip -- link-cell force kernels for gfx950: EAM density + pair force
// + embedding (passes 1 and 2 of src-flat/eam.c:438-660, fused), EAM
// embedding force (pass 3), and Lennard-Jones (src-flat/ljForce.c:89-224).
//
// Every kernel is full-shell: each atom sums over all of its neighbours, so
// no atom is written by two waves and nothing needs atomics; the reference's
// half-shell sums give the same totals (its pair loop with jBox >= iBox,
// eam.c:494,609, counts each pair once and adds to both atoms).
//
// Work mapping.  One wave (one 64-thread block) owns one home cell at a
// time (persistent grid).  The atoms of the 27 stencil cells (the reference's
// order, src-flat/domains.c:419-426, duplicates kept as it keeps them when a
// dimension has fewer than 3 cells) are staged in LDS shifted into the home
// cell's frame.  The wave's lanes are split into S = 64/n_home groups; lane
// (a, s) tests home atom a against halo atoms s, s+S, ...  Accepted
// neighbours are pushed into the lane's own LDS queue, and the queues are
// drained together, so the expensive pair arithmetic runs on dense lanes
// rather than on the ~15% of candidates inside the cutoff.  The S partial
// sums of each home atom are then added in a fixed order.
#include "comd_internal.h"

#define HMAX 384  // halo atoms staged per batch
#define QCAP 32   // queue entries per lane

// ---------------------------------------------------------------------------
// potential evaluation

// 3-point quadratic value and interpolated central-difference derivative,
// the formula of src-flat/eam.c:401-434 in double precision.  v points at
// values[-1].
__device__ __forceinline__ void interp64(const double *v, double x0, double xn, double invdx, int n, double r,
                                         double &val, double &der)
{
    if (r < x0) r = x0;
    else if (r > xn) r = xn;
    r = (r - x0) * invdx;
    const double fl = floor(r);
    int i = (int)fl;
    r = r - fl;
    // r <= xn keeps i <= n-1 for a consistent table header; the clamp only
    // keeps a rounding or a bad header inside values[-1..n+1]
    i = i < n - 1 ? i : n - 1;
    i = i > 0 ? i : 0;
    const double *w = v + 1 + i;
    const double vm = w[-1], v0 = w[0], v1 = w[1], v2 = w[2];
    const double g0 = v1 - vm;
    const double g1 = v2 - v0;
    val = v0 + 0.5 * r * (r * (v1 + vm - 2.0 * v0) + g0);
    der = (i <= 0) ? 0.0 : 0.5 * (g0 + r * (g1 - g0)) * invdx;
}

// Same function from per-interval single-precision coefficients computed
// in double precision on the host (value c0 + t(c1 + t c2), derivative
// d0 + t d1), so no difference of nearby table values is taken in FP32.
__device__ __forceinline__ void interp32(const float *coef, float x0, float xn, float invdx, int n, float r,
                                         float &val, float &der)
{
    r = fminf(fmaxf(r, x0), xn);
    const float u = (r - x0) * invdx;
    const float fl = floorf(u);
    int i = (int)fl;
    i = i < n - 1 ? i : n - 1;
    i = i > 0 ? i : 0;
    const float t = u - (float)i;
    const float4 c = *reinterpret_cast<const float4 *>(coef + 8 * i);
    const float d1 = coef[8 * i + 4];
    val = fmaf(t, fmaf(t, c.z, c.y), c.x);
    der = fmaf(t, d1, c.w);
}

// Clenshaw sum with the lower clamp only, src-flat/cheby.c:276-310.
__device__ __forceinline__ double cheb64(const double *c, int m, double a, double b, double x)
{
    if (x < a) x = a;
    double d = 0.0, dd = 0.0;
    const double y = (2.0 * x - a - b) / (b - a);
    const double y2 = 2.0 * y;
    for (int j = m - 1; j > 0; j--) {
        const double sv = d;
        d = y2 * d - dd + c[j];
        dd = sv;
    }
    return y * d - dd + 0.5 * c[0];
}

__device__ __forceinline__ float cheb32(const float *c, int m, float a, float b, float x)
{
    x = fmaxf(x, a);
    float d = 0.0f, dd = 0.0f;
    const float y = (2.0f * x - a - b) / (b - a);
    const float y2 = 2.0f * y;
    for (int j = m - 1; j > 0; j--) {
        const float sv = d;
        d = fmaf(y2, d, c[j] - dd);
        dd = sv;
    }
    return fmaf(y, d, 0.5f * c[0] - dd);
}

// ---------------------------------------------------------------------------
// one accepted pair (home atom a, neighbour j): dx = r_a - r_j

struct Acc {
    double x, y, z, e, rho, vir;
};

template <int KIND, int EVAL, int MIXED>
__device__ __forceinline__ void pair_term(const PotDev &P, double dx, double dy, double dz, double r2, double fia,
                                          double fij, Acc &A)
{
    if (!MIXED) {
        if (KIND == CM_KIND_LJ) {
            // src-flat/ljForce.c:196-210 (f holds -grad U, energy without 4 eps s6)
            const double s6 = P.lj_s6;
            const double r2i = 1.0 / r2;
            const double r6 = r2i * r2i * r2i;
            A.e += 0.5 * r6 * (s6 * r6 - 1.0);
            const double fr = 4.0 * P.lj_eps * s6 * r6 * r2i * (12.0 * s6 * r6 - 6.0);
            A.x += dx * fr;
            A.y += dy * fr;
            A.z += dz * fr;
            A.vir += fr * dx * dx;
        } else {
            const double r = sqrt(r2);
            if (KIND == CM_KIND_EAM1) {
                // src-flat/eam.c:521-545 (f holds +grad U)
                double phiv, dphi, rhov, drho;
                if (EVAL == COMD_EVAL_CHEB) {
                    const int m = P.cn;
                    phiv = cheb64(P.cheb + 0 * m, m, P.ca[0], P.cb[0], r);
                    dphi = cheb64(P.cheb + 3 * m, m, P.ca[0], P.cb[0], r);
                    rhov = cheb64(P.cheb + 1 * m, m, P.ca[1], P.cb[1], r);
                } else {
                    interp64(P.tab[0], P.x0[0], P.xn[0], P.invdx[0], P.n[0], r, phiv, dphi);
                    interp64(P.tab[1], P.x0[1], P.xn[1], P.invdx[1], P.n[1], r, rhov, drho);
                }
                A.x += dphi * dx / r;
                A.y += dphi * dy / r;
                A.z += dphi * dz / r;
                A.e += phiv / 2.0;
                A.rho += rhov;
            } else {
                // src-flat/eam.c:633-649
                double rhov, drho;
                if (EVAL == COMD_EVAL_CHEB) drho = cheb64(P.cheb + 4 * P.cn, P.cn, P.ca[1], P.cb[1], r);
                else interp64(P.tab[1], P.x0[1], P.xn[1], P.invdx[1], P.n[1], r, rhov, drho);
                const double g = (fia + fij) * drho;
                A.x += g * dx / r;
                A.y += g * dy / r;
                A.z += g * dz / r;
                A.vir += g * dx / r * dx;
            }
        }
    } else {
        // mixed precision: pair arithmetic in FP32, sums in FP64
        const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
        const float r2f = (float)r2;
        if (KIND == CM_KIND_LJ) {
            const float s6 = (float)P.lj_s6;
            const float r2i = 1.0f / r2f;
            const float r6 = r2i * r2i * r2i;
            const float sr = s6 * r6;
            A.e += (double)(0.5f * r6 * (sr - 1.0f));
            const float fr = (float)(4.0 * P.lj_eps) * sr * r2i * (12.0f * sr - 6.0f);
            A.x += (double)(fx * fr);
            A.y += (double)(fy * fr);
            A.z += (double)(fz * fr);
            A.vir += (double)(fr * fx * fx);
        } else {
            const float rinv = __builtin_amdgcn_rsqf(r2f);
            const float r = r2f * rinv;
            if (KIND == CM_KIND_EAM1) {
                float phiv, dphi, rhov, drho;
                if (EVAL == COMD_EVAL_CHEB) {
                    const int m = P.cn;
                    const float a0 = (float)P.ca[0], b0 = (float)P.cb[0], a1 = (float)P.ca[1], b1 = (float)P.cb[1];
                    phiv = cheb32(P.cheb32 + 0 * m, m, a0, b0, r);
                    dphi = cheb32(P.cheb32 + 3 * m, m, a0, b0, r);
                    rhov = cheb32(P.cheb32 + 1 * m, m, a1, b1, r);
                } else {
                    interp32(P.coef32[0], P.x0f[0], P.xnf[0], P.invdxf[0], P.n[0], r, phiv, dphi);
                    interp32(P.coef32[1], P.x0f[1], P.xnf[1], P.invdxf[1], P.n[1], r, rhov, drho);
                }
                const float s = dphi * rinv;
                A.x += (double)(s * fx);
                A.y += (double)(s * fy);
                A.z += (double)(s * fz);
                A.e += (double)(0.5f * phiv);
                A.rho += (double)rhov;
            } else {
                float rhov, drho;
                if (EVAL == COMD_EVAL_CHEB)
                    drho = cheb32(P.cheb32 + 4 * P.cn, P.cn, (float)P.ca[1], (float)P.cb[1], r);
                else interp32(P.coef32[1], P.x0f[1], P.xnf[1], P.invdxf[1], P.n[1], r, rhov, drho);
                const float g = (float)(fia + fij) * drho * rinv;
                A.x += (double)(g * fx);
                A.y += (double)(g * fy);
                A.z += (double)(g * fz);
                A.vir += (double)(g * fx * fx);
            }
        }
    }
}

// ---------------------------------------------------------------------------

template <int KIND, int EVAL, int MIXED>
__global__ __launch_bounds__(64) void k_pair(DevState d, PotDev P)
{
    __shared__ double hx[HMAX], hy[HMAX], hz[HMAX];
    __shared__ double hf[KIND == CM_KIND_EAM3 ? HMAX : 1];
    __shared__ uint16_t q[QCAP][64];
    __shared__ int s_cell[27], s_off[28];
    __shared__ double s_drb[27][3];
    __shared__ double red[6][64];

    const int L = threadIdx.x;
    const double rc2 = P.rc2;
    for (int c = blockIdx.x; c < d.ncells; c += gridDim.x) {
        const int nh = d.natoms[c];
        if (nh == 0) continue;

        // ---- stencil entries, counts and corner shifts
        const int nent = d.ncells == 1 ? 1 : 27;
        int ne = 0;
        if (L < nent) {
            const int b = stencil_cell(d, c, L);
            s_cell[L] = b;
            if (b >= 0) {
                ne = d.natoms[b];
                double cix, ciy, ciz, cjx, cjy, cjz;
                cell_corner(d, c, cix, ciy, ciz);
                cell_corner(d, b, cjx, cjy, cjz);
                s_drb[L][0] = image1(cix - cjx, d.L0, d.periodic);
                s_drb[L][1] = image1(ciy - cjy, d.L1, d.periodic);
                s_drb[L][2] = image1(ciz - cjz, d.L2, d.periodic);
            }
        }
        int incl = ne;
        for (int o = 1; o < 32; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (L >= o) incl += t;
        }
        const int H = __shfl(incl, 31, 64);
        const int nonempty = __popcll(__ballot(ne > 0));
        if (L < nent) s_off[L] = incl - ne;
        if (L == 0) s_off[nent] = H;

        // ---- home atoms: lane (a, s)
        const int S = 64 / nh;
        const bool act = L < S * nh;
        const int a = act ? L % nh : 0;
        const int sg = act ? L / nh : 0;
        const int hs = c * CM_SLOTS + a;
        const double xa = d.rx[hs], ya = d.ry[hs], za = d.rz[hs];
        const double fia = KIND == CM_KIND_EAM3 ? d.fi[hs] : 0.0;
        Acc A = {0, 0, 0, 0, 0, 0};

        for (int h0 = 0; h0 < H; h0 += HMAX) {
            const int hb = H - h0 < HMAX ? H - h0 : HMAX;
            __syncthreads();
            for (int h = h0 + L; h < h0 + hb; h += 64) {
                int e = 0;
                while (s_off[e + 1] <= h) e++;
                const int sl = s_cell[e] * CM_SLOTS + (h - s_off[e]);
                hx[h - h0] = d.rx[sl] - s_drb[e][0];
                hy[h - h0] = d.ry[sl] - s_drb[e][1];
                hz[h - h0] = d.rz[sl] - s_drb[e][2];
                if (KIND == CM_KIND_EAM3) hf[h - h0] = d.fi[sl];
            }
            __syncthreads();

            int cnt = 0;
            auto drain = [&]() {
                for (int k = 0; __ballot(k < cnt) != 0; k++) {
                    if (k < cnt) {
                        const int j = q[k][L];
                        const double dx = xa - hx[j], dy = ya - hy[j], dz = za - hz[j];
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        pair_term<KIND, EVAL, MIXED>(P, dx, dy, dz, r2, fia, KIND == CM_KIND_EAM3 ? hf[j] : 0.0, A);
                    }
                }
                cnt = 0;
            };
            const int trips = (hb + S - 1) / S;
            for (int it = 0; it < trips; it++) {
                const int j = sg + it * S;
                if (act && j < hb) {
                    const double dx = xa - hx[j], dy = ya - hy[j], dz = za - hz[j];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const bool in = KIND == CM_KIND_EAM3 ? (r2 < rc2) : (r2 <= rc2);
                    if (in && r2 > 0.0) {
                        q[cnt][L] = (uint16_t)j;
                        cnt++;
                    }
                }
                if (__ballot(cnt == QCAP) != 0) drain();
            }
            drain();
        }

        // ---- combine the S partial sums of each home atom, in s order
        __syncthreads();
        red[0][L] = act ? A.x : 0.0;
        red[1][L] = act ? A.y : 0.0;
        red[2][L] = act ? A.z : 0.0;
        red[3][L] = act ? A.e : 0.0;
        red[4][L] = act ? A.rho : 0.0;
        red[5][L] = act ? A.vir : 0.0;
        __syncthreads();
        if (L < nh) {
            double sx = 0, sy = 0, sz = 0, se = 0, sr = 0, sv = 0;
            for (int s = 0; s < S; s++) {
                const int l = L + s * nh;
                sx += red[0][l];
                sy += red[1][l];
                sz += red[2][l];
                se += red[3][l];
                sr += red[4][l];
                sv += red[5][l];
            }
            const int o = c * CM_SLOTS + L;
            if (KIND == CM_KIND_EAM1) {
                // pass 2, src-flat/eam.c:576-594
                double fv, fp;
                if (EVAL == COMD_EVAL_CHEB) {
                    fv = cheb64(P.cheb + 2 * P.cn, P.cn, P.ca[2], P.cb[2], sr);
                    fp = cheb64(P.cheb + 5 * P.cn, P.cn, P.ca[2], P.cb[2], sr);
                } else {
                    interp64(P.tab[2], P.x0[2], P.xn[2], P.invdx[2], P.n[2], sr, fv, fp);
                }
                d.fx[o] = sx;
                d.fy[o] = sy;
                d.fz[o] = sz;
                d.fe[o] = se + fv;
                d.rho[o] = sr;
                d.fi[o] = fp;
            } else if (KIND == CM_KIND_EAM3) {
                d.fx[o] += sx;
                d.fy[o] += sy;
                d.fz[o] += sz;
                d.vir[o] = sv;
            } else {
                // kinetic term of src-flat/ljForce.c:175, once per non-empty
                // stencil cell, with the mass in amu
                const double px = d.px[o];
                d.fx[o] = sx;
                d.fy[o] = sy;
                d.fz[o] = sz;
                d.fe[o] = se;
                d.vir[o] = sv - (double)nonempty * (px * px / P.mass_amu);
            }
        }
    }
}

template <int KIND, int EVAL>
static void launch_kind(const DevState &d, const PotDev &P, int mixed, int nb, hipStream_t st)
{
    if (mixed) hipLaunchKernelGGL((k_pair<KIND, EVAL, 1>), dim3(nb), dim3(64), 0, st, d, P);
    else hipLaunchKernelGGL((k_pair<KIND, EVAL, 0>), dim3(nb), dim3(64), 0, st, d, P);
}

void cm_launch_force(const DevState &d, const PotDev &P, int kind, int mixed, int nblocks, hipStream_t st)
{
    int nb = nblocks < d.ncells ? nblocks : d.ncells;
    if (nb < 1) nb = 1;
    if (kind == CM_KIND_LJ) {
        launch_kind<CM_KIND_LJ, COMD_EVAL_TABLE>(d, P, mixed, nb, st);
    } else if (P.eval == COMD_EVAL_CHEB) {
        if (kind == CM_KIND_EAM1) launch_kind<CM_KIND_EAM1, COMD_EVAL_CHEB>(d, P, mixed, nb, st);
        else launch_kind<CM_KIND_EAM3, COMD_EVAL_CHEB>(d, P, mixed, nb, st);
    } else {
        if (kind == CM_KIND_EAM1) launch_kind<CM_KIND_EAM1, COMD_EVAL_TABLE>(d, P, mixed, nb, st);
        else launch_kind<CM_KIND_EAM3, COMD_EVAL_TABLE>(d, P, mixed, nb, st);
    }
}

// ---------------------------------------------------------------------------
// deterministic two-stage reductions of per-atom energy, kinetic energy and
// virial: one partial triple per block, then one block sums the partials.

__global__ __launch_bounds__(256) void k_reduce(DevState d, double rmass, double *part)
{
    __shared__ double sh[3][256];
    double pe = 0, ke = 0, vi = 0;
    const long ns = (long)d.ncells * CM_SLOTS;
    for (long s = blockIdx.x * 256L + threadIdx.x; s < ns; s += (long)gridDim.x * 256L) {
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

void cm_launch_reduce(const DevState &d, const PotDev &P, double *partials, int nparts, hipStream_t st)
{
    const double rmass = 1822.83 * P.mass_amu;
    hipLaunchKernelGGL(k_reduce, dim3(nparts), dim3(256), 0, st, d, rmass, partials);
}

void cm_launch_reduce_final(const double *partials, int nparts, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partials, nparts, out);
}
