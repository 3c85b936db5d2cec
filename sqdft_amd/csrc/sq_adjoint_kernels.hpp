// Kernels of the Veff derivatives of the SQ density (no counterpart in the
// reference).  With the Lanczos bounds chi, zee frozen, rho_i depends on Veff
// through the Chebyshev moments of sq.cpp:568-655:
//   t_0 = e_c, t_1 = A t_0, t_{k+1} = 2 A t_k - t_{k-1},  A = (H_i - chi)/zee,
//   mu_k = t_k[c]/h^3 (k <= npl/2),
//   mu_2k = (2 t_k.t_k - 1)/h^3, mu_2k-1 = (2 t_k.t_{k-1} - t_1[c])/h^3 (k > npl/4),
// and rho_i = |sum_j 2 mu_j C_j| with the Fermi coefficients C_j of
// sq.cpp:1230-1259.  dA/dVeff(g) = diag(e_g)/zee on the node cube.
#pragma once
#include "sq_kernels.hpp"

namespace sqk {

// Per node at lambda_f: sign of rho_pc = sum_j 2 mu_j C_j and
// |d rho_pc / d lambda| = bet * sum_j 2 mu_j C'_j, C' the coefficients of
// f(1 - f) (sq.cpp:740-765, NeConstraint_derv sq.cpp:1407-1449).  Both sums
// use W[nd][k] = mu_0 + 2 sum_j mu_j cos(pi j (k + 1/2)/(npl + 1)).
__global__ void k_adj_node(const double *W, const double *chi, const double *zee, const double *ynode,
                           int nn, int npl, double lf, double bet0, double *sgn, double *dl)
{
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= nn) return;
    const int K = npl + 1;
    const double fac = 2.0 / (npl + 1);
    const double lfp = (lf - chi[nd]) / zee[nd];
    const double bet = bet0 * zee[nd];
    const double *w = W + (size_t)nd * K;
    double sr = 0.0, sd = 0.0;
    for (int k = 0; k < K; ++k) {
        const double f = fermi_dirac(ynode[k], lfp, bet);
        sr += f * w[k];
        sd += f * (1 - f) * w[k];
    }
    sgn[nd] = fac * sr < 0 ? -1.0 : 1.0;
    dl[nd] = bet0 * fac * sd;
}

// Reverse (adjoint) step k of the recurrence for B nodes, elementwise:
//   tb_k = [k < K2] (2 A tb_{k+1} - tb_{k+2}) + g_k,
//   g_k = b_k/h^3 e_c + [k > q4] (4 b_2k t_k + 2 b_2k-1 t_{k-1})/h^3
//         + [q4 < k+1 <= K2] 2 b_2k+1 t_{k+1}/h^3 - [k == 1] sum_odd/h^3 e_c,
//   G += [k < K2] 2 tb_{k+1} t_k,
// where b_j = 2 C_j wt_i is dL/dmu_j and sum_odd = sum_{k'>q4} b_{2k'-1}
// (each mu_2k-1 subtracts t_1[c]).  The z-march has already written
// 2 A tb_{k+1} - tb_{k+2} into tb when has_mv is set.
template <typename T>
struct AdjArgs {
    int M3, c, k, K2, q4, npl, has_mv;
    double inv_h3;
    const int *node_list;
    const double *Ci, *wt;        // [nn][npl+1], [nn]
    T *tb;                        // [B][M3]
    const T *tb1;                 // tb_{k+1}
    const T *tkm, *tk, *tkp;      // t_{k-1}, t_k, t_{k+1} (tkp null when k = K2)
    T *G;                         // [B][M3]
};

template <typename T>
__global__ void k_adj_update(AdjArgs<T> A)
{
    const int b = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.M3) return;
    const int nd = A.node_list[b];
    const double *ci = A.Ci + (size_t)nd * (A.npl + 1);
    const double wt = A.wt[nd];
    const int k = A.k;
    auto be = [&](int j) { return 2.0 * ci[j] * wt; };
    const size_t o = (size_t)b * A.M3 + p;
    double v = A.has_mv ? (double)A.tb[o] : 0.0;
    double g = 0.0;
    if (p == A.c) {
        g += be(k) * A.inv_h3;
        if (k == 1) {
            double so = 0.0;
            for (int kk = A.q4 + 1; kk <= A.K2; ++kk) so += be(2 * kk - 1);
            g -= so * A.inv_h3;
        }
    }
    const double tk = (double)A.tk[o];
    if (k > A.q4) g += (4.0 * be(2 * k) * tk + 2.0 * be(2 * k - 1) * (double)A.tkm[o]) * A.inv_h3;
    if (A.tkp && k + 1 > A.q4 && k + 1 <= A.K2) g += 2.0 * be(2 * k + 1) * (double)A.tkp[o] * A.inv_h3;
    A.tb[o] = (T)(v + g);
    if (k < A.K2) A.G[o] = (T)((double)A.G[o] + 2.0 * (double)A.tb1[o] * tk);
}

// t_1 = A t_0 with t_0 = e_c contributes tb_1[c] to G[c]
template <typename T>
__global__ void k_adj_center(int B, int M3, int c, const T *tb1, T *G)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const size_t o = (size_t)b * M3 + c;
    G[o] = (T)((double)G[o] + (double)tb1[o]);
}

// dL/dVeff(g) += sum over the batch's nodes i whose cube holds g of G_i(p)/zee_i.
// The batch is the nodes of z-planes [z0, z0 + nz) in natural order, so node
// (x, y, z) sits in slot ((z - z0) n + y) n + x.  Every grid point is owned by
// one thread, which visits the cube offsets in a fixed order: the sum is
// deterministic and needs no atomics.
template <typename T>
__global__ void k_adj_gather(int n, int nloc, int z0, int nz, const T *G, const double *zee, double *out)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = n * n * n;
    if (g >= nn) return;
    const int M = 2 * nloc + 1, M2 = M * M, M3 = M2 * M;
    const int gx = g % n, gy = (g / n) % n, gz = g / (n * n);
    double acc = 0.0;
    for (int dz = -nloc; dz <= nloc; ++dz) {
        const int bz = wrapn(gz - dz, n);
        const int zs = wrapn(bz - z0, n);
        if (zs >= nz) continue;
        for (int dy = -nloc; dy <= nloc; ++dy) {
            const int by = wrapn(gy - dy, n);
            const size_t rowslot = ((size_t)zs * n + by) * n;
            const int prow = ((dz + nloc) * M + (dy + nloc)) * M + nloc;
            for (int dx = -nloc; dx <= nloc; ++dx) {
                const int bx = wrapn(gx - dx, n);
                const size_t slot = rowslot + bx;
                const int node = (bz * n + by) * n + bx;
                acc += (double)G[slot * M3 + prow + dx] / zee[node];
            }
        }
    }
    out[g] += acc;
}

// ---------------------------------------------------------------- tangent
// Forward-mode derivative along dV: d_0 = 0, d_1 = dA t_0,
// d_{k+1} = 2 (dA t_k + A d_k) - d_{k-1},  dA = diag(dV on the cube)/zee.
// Per step the moments' derivatives need d_k[c], t_k.d_k and
// d_k.t_{k-1} + t_k.d_{k-1}.

// d_1 = dV(node)/zee e_c; rec[b][1] = (d_1[c], 0, 0)
template <typename T>
__global__ void k_tan_init(int B, int M3, int c, int K2, const int *node_list, const double *dV,
                           const double *zee, T *D1, double *rec)
{
    const int b = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M3) return;
    const int nd = node_list[b];
    const double v = dV[nd] / zee[nd];
    D1[(size_t)b * M3 + p] = p == c ? (T)v : (T)0;
    if (p == c) {
        double *r = rec + ((size_t)b * (K2 + 1) + 1) * 3;
        r[0] = (double)(T)v; r[1] = 0.0; r[2] = 0.0;
    }
}

// rec[b][k] = (d_k[c], t_k.d_k, d_k.t_{k-1} + t_k.d_{k-1}) from the tangent
// z-march step (MODE_TAN): the centre value and its two row halves' dots,
// added in half order.
__global__ void k_tan_rec(int B, int k, int K2, const double *center, const double *dots, double *rec)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double *d = dots + 4 * (size_t)b;
    double *r = rec + ((size_t)b * (K2 + 1) + k) * 3;
    r[0] = center[b];
    r[1] = d[0] + d[2];
    r[2] = d[1] + d[3];
}

// drho_pc = sum_j 2 C_j dmu_j with dmu_1 = d_1[c]/h^3, dmu_k = d_k[c]/h^3,
// dmu_2k = 4 t_k.d_k/h^3 and dmu_2k-1 = (2 (d_k.t_{k-1} + t_k.d_{k-1}) - d_1[c])/h^3
// for k > q4; drho = sgn * drho_pc
__global__ void k_tan_finish(int B, int K2, int q4, int npl, double inv_h3, const int *node_list,
                             const double *rec, const double *Ci, const double *sgn, double *drho)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int nd = node_list[b];
    const double *r = rec + (size_t)b * (K2 + 1) * 3;
    const double *ci = Ci + (size_t)nd * (npl + 1);
    const double dh = r[3 * 1];
    double s = 2.0 * ci[1] * (dh * inv_h3);
    for (int k = 2; k <= K2; ++k) {
        s += 2.0 * ci[k] * (r[3 * k] * inv_h3);
        if (k > q4) {
            s += 2.0 * ci[2 * k] * (4.0 * r[3 * k + 1] * inv_h3);
            s += 2.0 * ci[2 * k - 1] * ((2.0 * r[3 * k + 2] - dh) * inv_h3);
        }
    }
    drho[nd] = sgn[nd] * s;
}

// chi/zee of a batch gathered from the all-node arrays
__global__ void k_gather_bounds(int B, const int *node_list, const double *chi, const double *zee,
                                double *chiB, double *zeeB)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    chiB[b] = chi[node_list[b]];
    zeeB[b] = zee[node_list[b]];
}

}  // namespace sqk
