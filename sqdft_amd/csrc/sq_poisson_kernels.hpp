// sq_poisson_kernels.hpp -- FP64 device kernels of the SCF step around the SQ
// path: the AAJ Poisson iteration (PoissonSolver_AAJ, src/poisson.cpp:19-325,
// residual of PoissonResidual src/poisson.cpp:715-928) and the LDA
// exchange-correlation potential (src/scf.cpp:19-64).  All on the periodic
// n^3 grid, x fastest; every sum is reduced per block in a fixed tree and
// the block partials are added on the host in block order, so results do not
// depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

namespace sqk {

constexpr int PB = 256;   // threads per block of the grid kernels

// Sum over the block of NV values held per thread; thread 0 gets the totals.
template <int NV>
__device__ __forceinline__ void grid_block_sum(double (&v)[NV], double *red /* [PB/64][NV] */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v[i] += __shfl_xor(v[i], o, 64);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) red[wave * NV + i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = red[i];
            for (int w = 1; w < PB / 64; ++w) s += red[w * NV + i];
            v[i] = s;
        }
}

struct AajArgs {
    int n;
    double lap[FDN + 1];
    const double *phi, *rhs;
    double *phinew, *res;
    double *Xold, *Fold;          // previous iterate and residual
    double *DXk, *DFk;            // history column written this iteration (hist)
    int hist;
    double beta;
    double *part;                 // [gridDim.x] partial sums of res^2
};

// One Jacobi step: r = -(4 pi rhs + Lap_FD phi)/(3 lap0) with periodic
// neighbours, phi_new = phi + beta r, the history columns and ||r||^2.
__global__ void __launch_bounds__(PB) k_aaj_step(AajArgs A)
{
    __shared__ double red[PB / 64];
    const int n = A.n;
    const size_t N = (size_t)n * n * n;
    double acc[1] = {0.0};
    for (size_t c = blockIdx.x * (size_t)PB + threadIdx.x; c < N; c += (size_t)gridDim.x * PB) {
        const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((size_t)n * n));
        const double *P = A.phi;
        double L = P[c] * 3 * A.lap[0];
#pragma unroll
        for (int a = 1; a <= FDN; ++a) {
            const int im = i - a < 0 ? i - a + n : i - a, ip = i + a >= n ? i + a - n : i + a;
            const int jm = j - a < 0 ? j - a + n : j - a, jp = j + a >= n ? j + a - n : j + a;
            const int km = k - a < 0 ? k - a + n : k - a, kp = k + a >= n ? k + a - n : k + a;
            const size_t rowk = (size_t)k * n;
            L += (P[(rowk + j) * n + im] + P[(rowk + j) * n + ip] + P[(rowk + jm) * n + i] + P[(rowk + jp) * n + i] +
                  P[((size_t)km * n + j) * n + i] + P[((size_t)kp * n + j) * n + i]) *
                 A.lap[a];
        }
        const double r = -(((4 * M_PI) * (A.rhs[c]) + L) / (3 * A.lap[0]));
        A.res[c] = r;
        A.phinew[c] = P[c] + A.beta * r;
        if (A.hist) {
            A.DXk[c] = P[c] - A.Xold[c];
            A.DFk[c] = r - A.Fold[c];
        }
        A.Xold[c] = P[c];
        A.Fold[c] = r;
        acc[0] += r * r;
    }
    grid_block_sum<1>(acc, red);
    if (threadIdx.x == 0) A.part[blockIdx.x] = acc[0];
}

// Partial sums of DF_i . DF_j (i <= j < m) and DF_j . r for the Anderson step
// (AndersonExtrapolation, src/anderson.cpp:20-100).
template <int M>
__global__ void __launch_bounds__(PB) k_aaj_gram(const double *DF, const double *res, size_t N, double *part)
{
    constexpr int NV = M * (M + 1) / 2 + M;
    __shared__ double red[(PB / 64) * NV];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (size_t c = blockIdx.x * (size_t)PB + threadIdx.x; c < N; c += (size_t)gridDim.x * PB) {
        double f[M];
#pragma unroll
        for (int q = 0; q < M; ++q) f[q] = DF[q * N + c];
        const double r = res[c];
        int t = 0;
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
            for (int b2 = a; b2 < M; ++b2) acc[t++] += f[a] * f[b2];
#pragma unroll
        for (int a = 0; a < M; ++a) acc[t++] += f[a] * r;
    }
    grid_block_sum<NV>(acc, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) part[(size_t)blockIdx.x * NV + q] = acc[q];
}

struct AajUpd {
    double x[8];
    int m;
    double beta;
};

// phi_new -= sum_j (DX_j + beta DF_j) x_j (the Anderson update, anderson.cpp:88-97)
__global__ void __launch_bounds__(PB) k_aaj_update(const double *DX, const double *DF, size_t N, AajUpd U,
                                                  double *phinew)
{
    for (size_t c = blockIdx.x * (size_t)PB + threadIdx.x; c < N; c += (size_t)gridDim.x * PB) {
        double t = 0;
        for (int j = 0; j < U.m; ++j) t = t + (DX[j * N + c] + U.beta * DF[j * N + c]) * U.x[j];
        phinew[c] = phinew[c] - t;
    }
}

// LDA exchange-correlation potential: Perdew-Wang 92 correlation with the
// Slater exchange -(3/pi)^(1/3) rho^(1/3) (src/scf.cpp:19-64).
__global__ void __launch_bounds__(PB) k_vxc_lda(const double *rho, size_t N, double *vxc)
{
    const double p = 1.0, A = 0.031091, alpha1 = 0.21370, beta1 = 7.5957, beta2 = 3.5876, beta3 = 1.6382,
                 beta4 = 0.49294, C3 = 0.9847450218427;
    for (size_t c = blockIdx.x * (size_t)PB + threadIdx.x; c < N; c += (size_t)gridDim.x * PB) {
        const double rhoi = rho[c];
        double V = 0.0;
        if (rhoi != 0) {
            const double rs = pow((0.75 / (M_PI * rhoi)), (1.0 / 3.0));
            const double G = 2.0 * A * (beta1 * pow(rs, 0.5) + beta2 * rs + beta3 * pow(rs, 1.5) + beta4 * pow(rs, (p + 1.0)));
            const double dG = A * (beta1 * pow(rs, -0.5) + 2.0 * beta2 + 3.0 * beta3 * pow(rs, 0.5) +
                                   2.0 * (p + 1.0) * beta4 * pow(rs, p));
            V = (-2.0 * A * (1.0 + alpha1 * rs)) * log(1.0 + 1.0 / G) -
                (rs / 3.0) * (-2.0 * A * alpha1 * log(1.0 + 1.0 / G) - ((-2.0 * A * (1.0 + alpha1 * rs)) * dG) / (G * G + G));
        }
        vxc[c] = V - C3 * pow(rhoi, 1.0 / 3.0);
    }
}

// Grid terms of the total energy (EvaluateTotalEnergy, src/energy.cpp:19-87):
// per block the partial sums of 1/2 phi (b - rho) h^3, (eps_c rho + e_x) h^3
// and -Vxc rho h^3, with PW92 correlation and Slater exchange.
__global__ void __launch_bounds__(PB) k_energy_terms(const double *phi, const double *b, const double *rho,
                                                    const double *vxc, size_t N, double h, double *part)
{
    __shared__ double red[(PB / 64) * 3];
    const double p = 1.0, A = 0.031091, alpha1 = 0.21370, beta1 = 7.5957, beta2 = 3.5876, beta3 = 1.6382,
                 beta4 = 0.49294, C2 = 0.73855876638202;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t c = blockIdx.x * (size_t)PB + threadIdx.x; c < N; c += (size_t)gridDim.x * PB) {
        const double rhoi = rho[c];
        acc[0] += 0.5 * (phi[c] * (b[c] - rhoi)) * h * h * h;
        const double Ex = -C2 * pow(rhoi, 4.0 / 3.0);
        double Ec = 0.0;
        if (rhoi != 0) {
            const double rs = pow(0.75 / (M_PI * rhoi), (1.0 / 3.0));
            Ec = -2.0 * A * (1.0 + alpha1 * rs) *
                 log(1.0 + 1.0 / (2.0 * A * (beta1 * pow(rs, 0.5) + beta2 * rs + beta3 * pow(rs, 1.5) + beta4 * pow(rs, (p + 1.0)))));
        }
        acc[1] += (Ec * rhoi + Ex) * h * h * h;
        acc[2] += -1.0 * vxc[c] * rhoi * h * h * h;
    }
    grid_block_sum<3>(acc, red);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[(size_t)blockIdx.x * 3 + q] = acc[q];
}

}  // namespace sqk
