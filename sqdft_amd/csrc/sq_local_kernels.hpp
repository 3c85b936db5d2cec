// sq_local_kernels.hpp -- FP64 device kernels of the local (pseudocharge)
// forces, ForcesLocal (src/forces.cpp:20-315), on one rank.
//
// Per rb-replica J of an atom (Replica_atoms, src/initialize.cpp:21-200) the
// reference evaluates on the intersection of J's rb-box with the domain,
// widened by 2 FDn nodes: V_J (the spline of r V_loc over r) and the smooth
// reference potential V~_J; b_J = -(1/4pi) Lap_FD V_J and b~_J on the box
// widened by FDn; then at every intersection node the FD gradients of b_J,
// b~_J, V_J, V~_J and
//   f_loc  += grad(b_J) (phi - V_J) h^3
//   f_corr += 0.5 (grad b~_J (Vc - V~_J) + grad b_J (Vc + V_J)
//                  + (grad V~_J - grad V_J)(b~ + b) + b_J grad V_J - b~_J grad V~_J) h^3.
// Here the boxes of a batch of replicas live side by side in scratch memory
// (one box per blockIdx.y); each stage is one launch over all of them.  The
// force sums are reduced per block in a fixed tree and the block partials
// are added on the host in (replica, block) order, so the result does not
// depend on scheduling.  The arithmetic is kept unfused, in the reference's
// order, so that values agree with the FP64 CPU restatement to rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace sqk {

struct LImg {
    double x, y, z;     // replica position (bohr)
    int lo[3], hi[3];   // intersection of the rb-box with [0, n-1]^3
    int type, pad;
    long off;           // first point of this replica's box in the scratch arrays
};

struct LocArgs {
    int n, F, nimg;
    double h, h3;
    double lap[FDN + 1], grad[FDN + 1];
    const LImg *img;
    const double *R, *RV, *DV;   // radial grids, r V_loc, spline derivatives (all types)
    const int *toff, *tsize;     // per type: offset into R/RV/DV and length
    const double *Z;             // per type valence charge
    double *V, *Vt, *bJ, *bJt;   // scratch boxes (intersection + 2 FDn per side)
    const double *phi, *b, *bt, *vc;
    double *part;                // [nimg][gridDim.x][6] block partials
    int *err;                    // set to 1 when r leaves a radial grid
    // pseudocharge gather
    const double *RU, *DU;       // per type r-grid values of uu and their spline derivatives
    double *gb, *gbt, *gvc, *grho;  // n^3 accumulators (b, b~, Vc, rho_at)
};

// Interval of the reference's single-point search (src/spline.cpp:11-55):
// the last knot <= p, the last interval when p is the final knot.
__device__ __forceinline__ double spline_at(const double *X, const double *Y, const double *D, int n, double p,
                                            int *err)
{
#pragma clang fp contract(off)
    if (p < X[0] || p > X[n - 1]) { *err = 1; return 0.0; }
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (X[mid] <= p) lo = mid; else hi = mid;
    }
    const int j = lo > n - 2 ? n - 2 : lo;
    const double dx = 1.0 / (X[j + 1] - X[j]);
    const double dy = (Y[j + 1] - Y[j]) * dx;
    const double c0 = Y[j], c1 = D[j];
    const double c2 = dx * (3.0 * dy - 2.0 * D[j] - D[j + 1]);
    const double c3 = dx * dx * (-2.0 * dy + D[j] + D[j + 1]);
    const double t = p - X[j];
    return ((c3 * t + c2) * t + c1) * t + c0;
}

__device__ __forceinline__ void box_dims(const LImg &I, int w, int &nx, int &ny, int &nz)
{
    nx = I.hi[0] - I.lo[0] + 1 + 2 * w;
    ny = I.hi[1] - I.lo[1] + 1 + 2 * w;
    nz = I.hi[2] - I.lo[2] + 1 + 2 * w;
}

// V_J and V~_J (rc~ = 0.5) on the box widened by 2 FDn
// (src/forces.cpp:113-151; src/initialize.cpp:565-595).
__global__ void __launch_bounds__(PB) k_loc_potential(LocArgs A)
{
#pragma clang fp contract(off)
    const LImg I = A.img[blockIdx.y];
    const int F2 = 2 * A.F;
    int nx, ny, nz;
    box_dims(I, F2, nx, ny, nz);
    const long NB = (long)nx * ny * nz;
    const int t = I.type;
    const double *X = A.R + A.toff[t], *Y = A.RV + A.toff[t], *D = A.DV + A.toff[t];
    const int ns = A.tsize[t];
    const double Z = A.Z[t], rc = 0.5, h = A.h;
    const int i0 = I.lo[0] - F2, j0 = I.lo[1] - F2, k0 = I.lo[2] - F2;
    int bad = 0;
    for (long p = blockIdx.x * (long)PB + threadIdx.x; p < NB; p += (long)gridDim.x * PB) {
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long)nx * ny));
        const double dx = (i0 + i) * h - I.x, dy = (j0 + j) * h - I.y, dz = (k0 + k) * h - I.z;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double v = spline_at(X, Y, D, ns, r, &bad);
        v = fabs(r - 0.0) < 1e-12 ? Y[1] / X[1] : v / r;
        double vt;
        if (r < rc)
            vt = -Z * (9 * pow(r, 7) - 30 * rc * pow(r, 6) + 28 * rc * rc * pow(r, 5) - 14 * pow(rc, 5) * pow(r, 2) +
                       12 * pow(rc, 7)) / (5 * pow(rc, 8));
        else
            vt = -Z / r;
        A.V[I.off + p] = v;
        A.Vt[I.off + p] = vt;
    }
    if (bad) *A.err = 1;
}

// b_J, b~_J on the box widened by FDn (src/forces.cpp:156-177); stored at
// the same box index as V.
__global__ void __launch_bounds__(PB) k_loc_charge(LocArgs A)
{
#pragma clang fp contract(off)
    const LImg I = A.img[blockIdx.y];
    const int F = A.F;
    int nx, ny, nz;
    box_dims(I, 2 * F, nx, ny, nz);
    const int mx = nx - 2 * F, my = ny - 2 * F, mz = nz - 2 * F;
    const long NI = (long)mx * my * mz;
    const long sy = nx, sz = (long)nx * ny;
    const double q = -1 / (4 * M_PI);
    const double *V = A.V + I.off, *Vt = A.Vt + I.off;
    for (long p = blockIdx.x * (long)PB + threadIdx.x; p < NI; p += (long)gridDim.x * PB) {
        const int i = F + (int)(p % mx), j = F + (int)((p / mx) % my), k = F + (int)(p / ((long)mx * my));
        const long l = k * sz + j * sy + i;
        double s = V[l] * 3 * A.lap[0] * q;
        for (int a = 1; a <= F; ++a)
            s += (V[l - a] + V[l + a] + V[l - a * sy] + V[l + a * sy] + V[l - a * sz] + V[l + a * sz]) * A.lap[a] * q;
        double st = Vt[l] * 3 * A.lap[0] * q;
        for (int a = 1; a <= F; ++a)
            st += (Vt[l - a] + Vt[l + a] + Vt[l - a * sy] + Vt[l + a * sy] + Vt[l - a * sz] + Vt[l + a * sz]) *
                  A.lap[a] * q;
        A.bJ[I.off + l] = s;
        A.bJt[I.off + l] = st;
    }
}

// Gradients at the intersection nodes and the six force sums
// (src/forces.cpp:183-276); one partial per block.
__global__ void __launch_bounds__(PB) k_loc_force(LocArgs A)
{
#pragma clang fp contract(off)
    __shared__ double red[(PB / 64) * 6];
    const LImg I = A.img[blockIdx.y];
    const int F = A.F, n = A.n;
    int nx, ny, nz;
    box_dims(I, 2 * F, nx, ny, nz);
    const int mx = nx - 4 * F, my = ny - 4 * F, mz = nz - 4 * F;
    const long NI = (long)mx * my * mz;
    const long sy = nx, sz = (long)nx * ny;
    const long step[3] = {1, sy, sz};
    const double *V = A.V + I.off, *Vt = A.Vt + I.off, *bJ = A.bJ + I.off, *bJt = A.bJt + I.off;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long p = blockIdx.x * (long)PB + threadIdx.x; p < NI; p += (long)gridDim.x * PB) {
        const int ii = (int)(p % mx), jj = (int)((p / mx) % my), kk = (int)(p / ((long)mx * my));
        const long l = (kk + 2 * F) * sz + (jj + 2 * F) * sy + (ii + 2 * F);
        const long g = ((long)(I.lo[2] + kk) * n + (I.lo[1] + jj)) * n + (I.lo[0] + ii);
        const double ph = A.phi[g], vc = A.vc[g], bb = A.bt[g] + A.b[g];
        for (int d = 0; d < 3; ++d) {
            const long s = step[d];
            double Db = 0, Dbt = 0, DV = 0, DVt = 0;
            for (int a = 1; a <= F; ++a) Db += (-bJ[l - a * s] + bJ[l + a * s]) * A.grad[a];
            for (int a = 1; a <= F; ++a) Dbt += (-bJt[l - a * s] + bJt[l + a * s]) * A.grad[a];
            for (int a = 1; a <= F; ++a) DV += (-V[l - a * s] + V[l + a * s]) * A.grad[a];
            for (int a = 1; a <= F; ++a) DVt += (-Vt[l - a * s] + Vt[l + a * s]) * A.grad[a];
            acc[d] += Db * (ph - V[l]) * A.h3;
            acc[3 + d] += 0.5 * (Dbt * (vc - Vt[l]) + Db * (vc + V[l]) + (DVt - DV) * bb + bJ[l] * DV - bJt[l] * DVt) *
                          A.h3;
        }
    }
    grid_block_sum<6>(acc, red);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) A.part[((long)blockIdx.y * gridDim.x + blockIdx.x) * 6 + c] = acc[c];
}

// Pseudocharge gather (ChargeDensity, src/initialize.cpp:565-630): every grid
// node adds, replica by replica in list order, b_J, b~_J, V~_J - V_J and the
// atomic density uu(r) of the replicas whose intersection holds it -- the
// order the reference accumulates them in -- and the per-node terms of
// Eself, Ncharge, Eself~, Ncharge~ are reduced per block.
__global__ void __launch_bounds__(PB) k_pc_gather(LocArgs A, double *part4)
{
#pragma clang fp contract(off)
    __shared__ double red[(PB / 64) * 4];
    const int n = A.n, F = A.F;
    const long N = (long)n * n * n;
    const double h = A.h, h3 = h * h * h;
    double acc[4] = {0, 0, 0, 0};
    int bad = 0;
    for (long g = blockIdx.x * (long)PB + threadIdx.x; g < N; g += (long)gridDim.x * PB) {
        const int i = (int)(g % n), j = (int)((g / n) % n), k = (int)(g / ((long)n * n));
        double b = A.gb[g], bt = A.gbt[g], vc = A.gvc[g], ra = A.grho[g];
        for (int m = 0; m < A.nimg; ++m) {
            const LImg &I = A.img[m];
            if (i < I.lo[0] || i > I.hi[0] || j < I.lo[1] || j > I.hi[1] || k < I.lo[2] || k > I.hi[2]) continue;
            int nx, ny, nz;
            box_dims(I, 2 * F, nx, ny, nz);
            const long l = I.off + ((long)(k - I.lo[2] + 2 * F) * ny + (j - I.lo[1] + 2 * F)) * nx + (i - I.lo[0] + 2 * F);
            const double bj = A.bJ[l], btj = A.bJt[l], v = A.V[l], vt = A.Vt[l];
            b += bj;
            acc[0] += 0.5 * bj * v * h3;
            acc[1] += bj * h3;
            bt += btj;
            acc[2] += 0.5 * btj * vt * h3;
            acc[3] += btj * h3;
            vc += vt - v;
            const double dx = i * h - I.x, dy = j * h - I.y, dz = k * h - I.z;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const int t = I.type;
            ra += spline_at(A.R + A.toff[t], A.RU + A.toff[t], A.DU + A.toff[t], A.tsize[t], r, &bad);
        }
        A.gb[g] = b; A.gbt[g] = bt; A.gvc[g] = vc; A.grho[g] = ra;
    }
    if (bad) *A.err = 1;
    grid_block_sum<4>(acc, red);
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) part4[(long)blockIdx.x * 4 + c] = acc[c];
}

}  // namespace sqk
