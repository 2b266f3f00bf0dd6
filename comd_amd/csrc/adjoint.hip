// adjoint.hip -- gradient of a force-matching loss with respect to the
// Chebyshev coefficients of phi, rho and F (EAM evaluated through the
// series of src-flat/cheby.c:99-310, as src-flat/eam.c:523-527,582-584,
// 637-639 do under USE_CHEBY).
//
// Forward quantities (f holds +grad U, as src-flat/eam.c fills it):
//   rho_i = sum_j rho(r_ij)                          (pairs with r2 <= rc2)
//   E     = sum_i F(rho_i) + 1/2 sum_i sum_j phi(r_ij)
//   f_i   = sum_j phi'(r_ij) u_ij                    (r2 <= rc2)
//         + sum_j (F'(rho_i) + F'(rho_j)) rho'(r_ij) u_ij   (r2 < rc2)
//   u_ij  = (r_i - r_j) / r_ij
// with phi' = series(D c_phi), rho' = series(D c_rho), F' = series(D c_F),
// D the derivative recurrence of src-flat/cheby.c:219-270.
//
// Loss L = sum_i g0 |f_i - t_i|^2 + w_E (E - E_t)^2.  With g_i = dL/df_i,
// w_ij = (g_i - g_j) . u_ij and B_k(x) = T_k(y(x)) (B_0 = 1/2):
//   pair pass 1: M_phi'[k] = 1/2 sum_ij w_ij B_k(r_ij),  M_phiE[k] = 1/2 sum_ij B_k(r_ij)
//   pair pass 2: M_rho'[k] = 1/2 sum_ij (F'_i + F'_j) w_ij B_k(r_ij),
//                lambda_i = sum_j rho'(r_ij) w_ij            (dL/dF'_i through the forces)
//   atom pass:   mu_i = lambda_i F''(rho_i) + dL/dE F'(rho_i) (dL/drho_i)
//                M_F[k] = sum_i lambda_i B_k(rho_i), M_FE[k] = sum_i B_k(rho_i)
//   pair pass 3: M_rho[k] = sum_ij mu_i B_k(r_ij)
// (sums over ordered pairs, every atom seeing all its neighbours; rho_i is
// recomputed by the forward force pass, never stored across calls).  The host
// assembles  dL/dc_phi = D^T M_phi' + dL/dE M_phiE,
//            dL/dc_rho = D^T M_rho' + M_rho,
//            dL/dc_F   = D^T M_F + dL/dE M_FE.
//
// Every pass is full-shell over the 27-cell stencil with the halo staged in
// LDS, one wave per cell on a persistent grid; the pairs a wave finds are
// queued and evaluated 64 at a time; each lane keeps its moments in
// registers and the waves' partials are summed in a fixed order, so the
// gradient is deterministic.  MIXED evaluates the Chebyshev basis and rho'
// in single precision and accumulates in double.
#include "comd_internal.h"
#include "adj_basis.h"

#define AHMAX 320  // halo atoms staged per batch

struct AdjDev {
    const double *gx, *gy, *gz;  // per slot: dL/df
    double *lam;                 // per slot: lambda (pair pass 2)
    double *mu;                  // per slot: mu (atom pass)
    const double *f2;            // F'' series (cstride coefficients, zero-padded, FP64)
    const double *dlde;          // per replica: dL/dE of its energy term
    double *part;                // per block: ANMOM partial moments
};


template <int TWO>
__device__ __forceinline__ void write_moments(const double *a, const double *b, double *part)
{
    const int L = threadIdx.x;
#pragma unroll
    for (int k = 0; k < AMAX; k++) {
        const double s = wave_sum(a[k]);
        if (L == 0) part[k] = s;
        if (TWO) {
            const double t = wave_sum(b[k]);
            if (L == 0) part[AMAX + k] = t;
        } else if (L == 0) {
            part[AMAX + k] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// pair passes 1-3
//
// One wave per cell: the wave tests 64 candidates of one home atom at a time
// (home atoms in order) and queues the pairs inside the cutoff in LDS, in
// that order; every 64 queued pairs are evaluated one per lane.  The basis
// sums, the costly part, so run on full waves (tested lane by lane, about
// one candidate in eight is a pair, and a wave would run them for whichever
// few lanes found one).  Pass 2's lambda_i is summed per batch by a
// segmented scan over the lanes (the queue is in home-atom order, and is
// emptied after each staged halo batch so that it stays so), so every sum
// keeps a fixed order.

#define AQ 128  // queued pairs: a batch of 64 and up to 63 more

template <int PASS, int MIXED>
__global__ __launch_bounds__(64) void k_adj_pair(DevState d, PotDev P, AdjDev A)
{
    __shared__ double hx[AHMAX], hy[AHMAX], hz[AHMAX];
    __shared__ double hgx[PASS == 3 ? 1 : AHMAX], hgy[PASS == 3 ? 1 : AHMAX], hgz[PASS == 3 ? 1 : AHMAX];
    __shared__ double hf[PASS == 2 ? AHMAX : 1];
    __shared__ double drb[81];
    __shared__ int cell[27], off[28];
    // the cell's home atoms: position, dL/df (passes 1, 2), F' (pass 2) or
    // mu (pass 3), and lambda (pass 2)
    __shared__ double ox[CM_SLOTS], oy[CM_SLOTS], oz[CM_SLOTS];
    __shared__ double ogx[PASS == 3 ? 1 : CM_SLOTS], ogy[PASS == 3 ? 1 : CM_SLOTS], ogz[PASS == 3 ? 1 : CM_SLOTS];
    __shared__ double ow[PASS == 1 ? 1 : CM_SLOTS];
    __shared__ double lamacc[PASS == 2 ? CM_SLOTS : 1];
    // queued pairs: r^2; (g_i - g_j) . (r_i - r_j) (passes 1, 2) or mu_i
    // (pass 3); F'_i + F'_j and the home atom (pass 2)
    __shared__ double qr2[AQ], qv[AQ];
    __shared__ double qf[PASS == 2 ? AQ : 1];
    __shared__ int qh[PASS == 2 ? AQ : 1];

    const int L = threadIdx.x;
    const int m = P.cn;
    const int dom = PASS == 1 ? 0 : 1;
    const double ca = P.ca[dom], cb = P.cb[dom];
    const double rc2 = P.rc2;
    const int m8 = P.cstride;
    const double *dser = P.cheb + 4 * m8;      // rho' series

    double m0[AMAX], m1[AMAX];
#pragma unroll
    for (int k = 0; k < AMAX; k++) m0[k] = m1[k] = 0.0;

    // evaluates queue entries [0, n), n <= 64, one per lane
    auto flush = [&](int n, double de) {
        __syncthreads();
        double v = 0.0;
        int key = 1 << 30;
        if (L < n) {
            const double r = sqrt(qr2[L]);
            const double y = cheb_arg(r, ca, cb);
            if (PASS == 3) {
                add_basis<MIXED, 0>(y, m, qv[L], 0.0, m0, m1);
            } else {
                const double w = qv[L] / r;
                if (PASS == 1) {
                    add_basis<MIXED, 1>(y, m, 0.5 * w, 0.5 * de, m0, m1);
                } else {
                    // rho'(r) from the same basis values (rho's domain)
                    const double drho = add_basis_sum<MIXED>(y, m, 0.5 * qf[L] * w, m0, dser);
                    v = drho * w;
                    key = qh[L];
                }
            }
        }
        if (PASS == 2) {
            // lambda_a += the batch's rho'(r) w of home atom a
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double vu = __shfl_up(v, o, 64);
                const int ku = __shfl_up(key, o, 64);
                if (L >= o && ku == key) v += vu;
            }
            const int kn = __shfl_down(key, 1, 64);
            if (L < n && (L == 63 || kn != key)) lamacc[key] += v;
        }
        __syncthreads();
    };

    for (int c = blockIdx.x; c < d.ncells; c += gridDim.x) {
        const int nh = d.natoms[c];
        if (nh == 0) continue;
        const Grid g = grid_of(d, c);
        const double de = PASS == 1 ? A.dlde[rep_of(d, c)] : 0.0;  // weight of the energy moments
        const int nent = g.ncells == 1 ? 1 : 27;
        int ne = 0;
        if (L < nent) {
            const int b = stencil_cell(g, c, L);
            cell[L] = b;
            if (b >= 0) {
                ne = d.natoms[b];
                double sx, sy, sz;
                corner_shift(g, c, b, sx, sy, sz);
                drb[3 * L + 0] = sx;
                drb[3 * L + 1] = sy;
                drb[3 * L + 2] = sz;
            }
        }
        int incl = ne;
        for (int o = 1; o < 32; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (L >= o) incl += t;
        }
        const int H = __shfl(incl, 31, 64);
        if (L < nent) off[L] = incl - ne;
        if (L == 0) off[nent] = H;
        if (L < nh) {
            const int hs = c * CM_SLOTS + L;
            ox[L] = d.rx[hs];
            oy[L] = d.ry[hs];
            oz[L] = d.rz[hs];
            if (PASS != 3) {
                ogx[L] = A.gx[hs];
                ogy[L] = A.gy[hs];
                ogz[L] = A.gz[hs];
            }
            if (PASS == 2) {
                ow[L] = d.fi[hs];
                lamacc[L] = 0.0;
            }
            if (PASS == 3) ow[L] = A.mu[hs];
        }
        int qn = 0;

        for (int h0 = 0; h0 < H; h0 += AHMAX) {
            const int hb = H - h0 < AHMAX ? H - h0 : AHMAX;
            __syncthreads();
            {
                int e = 0;
                for (int h = h0 + L; h < h0 + hb; h += 64) {
                    while (off[e + 1] <= h) e++;
                    const int sl = cell[e] * CM_SLOTS + (h - off[e]);
                    hx[h - h0] = d.rx[sl] - drb[3 * e + 0];
                    hy[h - h0] = d.ry[sl] - drb[3 * e + 1];
                    hz[h - h0] = d.rz[sl] - drb[3 * e + 2];
                    if (PASS != 3) {
                        hgx[h - h0] = A.gx[sl];
                        hgy[h - h0] = A.gy[sl];
                        hgz[h - h0] = A.gz[sl];
                    }
                    if (PASS == 2) hf[h - h0] = d.fi[sl];
                }
            }
            __syncthreads();
            for (int a = 0; a < nh; a++) {
                const double xa = ox[a], ya = oy[a], za = oz[a];
                for (int j0 = 0; j0 < hb; j0 += 64) {
                    const int j = j0 + L;
                    double dx = 0.0, dy = 0.0, dz = 0.0, r2 = 0.0;
                    bool acc = false;
                    if (j < hb) {
                        dx = xa - hx[j];
                        dy = ya - hy[j];
                        dz = za - hz[j];
                        r2 = dx * dx + dy * dy + dz * dz;
                        // pair predicates of src-flat/eam.c:519 (phi, rho) and 633 (rho')
                        const bool in = PASS == 2 ? (r2 < rc2) : (r2 <= rc2);
                        acc = in && r2 > 0.0;
                    }
                    const unsigned long long mk = __ballot(acc);
                    if (acc) {
                        const int p = qn + __popcll(mk & ((1ull << L) - 1ull));
                        qr2[p] = r2;
                        if (PASS == 3) {
                            qv[p] = ow[a];
                        } else {
                            qv[p] = (ogx[a] - hgx[j]) * dx + (ogy[a] - hgy[j]) * dy + (ogz[a] - hgz[j]) * dz;
                            if (PASS == 2) {
                                qf[p] = ow[a] + hf[j];
                                qh[p] = a;
                            }
                        }
                    }
                    qn += __popcll(mk);
                    if (qn >= 64) {
                        flush(64, de);
                        // the rest (fewer than 64) to the front
                        if (L < qn - 64) {
                            qr2[L] = qr2[64 + L];
                            qv[L] = qv[64 + L];
                            if (PASS == 2) {
                                qf[L] = qf[64 + L];
                                qh[L] = qh[64 + L];
                            }
                        }
                        qn -= 64;
                    }
                }
            }
            // at the end of each staged halo batch: a batch holds each home
            // atom's pairs in one run (the scan of lambda relies on it)
            if (qn > 0) flush(qn, de);
            qn = 0;
        }
        __syncthreads();
        if (PASS == 2 && L < nh) A.lam[c * CM_SLOTS + L] = lamacc[L];
        __syncthreads();
    }
    write_moments<PASS == 1>(m0, m1, A.part + (size_t)blockIdx.x * ANMOM);
}

// ---------------------------------------------------------------------------
// atom pass: mu_i, M_F, M_FE

template <int MIXED>
__global__ __launch_bounds__(64) void k_adj_atom(DevState d, PotDev P, AdjDev A)
{
    const int L = threadIdx.x;
    const int m = P.cn;
    const double ca = P.ca[2], cb = P.cb[2];
    double m0[AMAX], m1[AMAX];
#pragma unroll
    for (int k = 0; k < AMAX; k++) m0[k] = m1[k] = 0.0;
    for (int c = blockIdx.x; c < d.ncells; c += gridDim.x) {
        if (L < d.natoms[c]) {
            const int s = c * CM_SLOTS + L;
            const double de = A.dlde[rep_of(d, c)];
            const double rho = d.rho[s];
            const double lam = A.lam[s];
            const double f2 = cheb64(A.f2, P.cstride, ca, cb, rho);  // zero-padded as the series
            A.mu[s] = lam * f2 + de * d.fi[s];
            add_basis<MIXED, 1>(cheb_arg(rho, ca, cb), m, lam, de, m0, m1);
        }
    }
    write_moments<1>(m0, m1, A.part + (size_t)blockIdx.x * ANMOM);
}

// ---------------------------------------------------------------------------
// residual: with w the atom's replica weight (1 / (3 N) of its replica),
// g_i = 2 w (f_i - t_i), and the partial sums of w |f_i - t_i|^2

__global__ __launch_bounds__(64) void k_adj_resid(DevState d, const double *ft, const double *rw, double *gx,
                                                  double *gy, double *gz, double *part)
{
    const int L = threadIdx.x;
    double ss = 0.0;
    for (int c = blockIdx.x; c < d.ncells; c += gridDim.x) {
        const int s = c * CM_SLOTS + L;
        if (L < d.natoms[c]) {
            const double w = rw[rep_of(d, c)];
            const double ex = d.fx[s] - ft[3L * s + 0];
            const double ey = d.fy[s] - ft[3L * s + 1];
            const double ez = d.fz[s] - ft[3L * s + 2];
            gx[s] = 2.0 * w * ex;
            gy[s] = 2.0 * w * ey;
            gz[s] = 2.0 * w * ez;
            ss += w * (ex * ex + ey * ey + ez * ez);
        } else {
            gx[s] = 0.0;
            gy[s] = 0.0;
            gz[s] = 0.0;
        }
    }
    const double t = wave_sum(ss);
    part[(size_t)blockIdx.x * ANMOM + L] = L == 0 ? t : 0.0;
}

// Sums the blocks' partial moments in a fixed order: out[k] = sum_b
// part[b][k].  Sixteen waves take interleaved runs of 8 blocks (their loads
// issued together), and their sums are added in wave order.
__global__ __launch_bounds__(1024) void k_adj_final(const double *part, int nblocks, double *out)
{
    __shared__ double acc[16][ANMOM];
    const int k = threadIdx.x & 63, j = threadIdx.x >> 6;
    double s = 0.0;
    for (int b0 = j * 8; b0 < nblocks; b0 += 16 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = b0 + u < nblocks ? part[(size_t)(b0 + u) * ANMOM + k] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    acc[j][k] = s;
    __syncthreads();
    if (j == 0) {
        double t = 0.0;
        for (int q = 0; q < 16; q++) t += acc[q][k];
        out[k] = t;
    }
}

// ---------------------------------------------------------------------------

int cm_adjoint_max_coefs() { return AMAX; }
int cm_adjoint_moments() { return ANMOM; }

template <int MIXED>
static void launch_all(const DevState &d, const PotDev &P, const CmAdjoint &J, int stage, hipStream_t st)
{
    AdjDev A;
    A.gx = J.gx;
    A.gy = J.gy;
    A.gz = J.gz;
    A.lam = J.lam;
    A.mu = J.mu;
    A.f2 = J.f2;
    A.dlde = J.dlde;
    A.part = J.part;
    const dim3 g(J.nblocks), b(64);
    double *o = J.out + (size_t)stage * ANMOM;
    switch (stage) {
    case 0: hipLaunchKernelGGL((k_adj_pair<1, MIXED>), g, b, 0, st, d, P, A); break;
    case 1: hipLaunchKernelGGL((k_adj_pair<2, MIXED>), g, b, 0, st, d, P, A); break;
    case 2: hipLaunchKernelGGL((k_adj_atom<MIXED>), g, b, 0, st, d, P, A); break;
    case 3: hipLaunchKernelGGL((k_adj_pair<3, MIXED>), g, b, 0, st, d, P, A); break;
    default: return;
    }
    hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(1024), 0, st, J.part, J.nblocks, o);
}

void cm_launch_adj_resid(const DevState &d, const CmAdjoint &J, const double *ft, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_resid, dim3(J.nblocks), dim3(64), 0, st, d, ft, J.rw, J.gx, J.gy, J.gz, J.part);
    hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(1024), 0, st, J.part, J.nblocks, J.out + (size_t)4 * ANMOM);
}

void cm_launch_adj_final(const CmAdjoint &J, int nblocks, int stage, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(1024), 0, st, J.part, nblocks, J.out + (size_t)stage * ANMOM);
}

void cm_launch_adj_stage(const DevState &d, const PotDev &P, const CmAdjoint &J, int stage, int mixed, hipStream_t st)
{
    if (mixed) launch_all<1>(d, P, J, stage, st);
    else launch_all<0>(d, P, J, stage, st);
}
