// sq_kernels.hpp -- device kernels of the MI355X SQ hot path (gfx950).
//
// Data layout in HBM (per batch of B FD nodes, node-major):
//   vectors   X, Y      [B][M^3]  M = 2 nloc + 1, each node cube z-major (z, y, x)
//   projections s       [B][n_ip] FP64, one value per image-projector
//   image lists         [B][LCAP] image ids whose rc-cube meets the node cube
//   group image lists   [ceil(B/NG)][GLCAP] the same for NG consecutive slots
//   Veff                [n^3] periodic grid (x fastest), read through wrapped indices
//   projector ELL       over the extended grid (n + 2 nloc)^3: per point the
//                       (image-projector id, U value) pairs that cover it
// Two workgroups own one node cube, one per half of its rows: each marches
// through the cube along z, keeping a 13-plane register window per column
// (z-neighbours), staging the current plane in LDS (x/y-neighbours, with six
// halo rows beyond its half) and streaming each output plane to HBM as soon
// as it is complete.  The projections of 8 consecutive cubes (a 4x2 half
// brick) are formed by one workgroup from their union, slab by slab.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sqk {

constexpr int FDN = 6;          // half FD order (12th-order stencil)
constexpr int WIN = 2 * FDN + 1;
constexpr int LCAP = 512;       // max images meeting one node cube
constexpr int NG = 8;           // node cubes per workgroup of the grouped projector kernels
constexpr int SLW = 10;         // LDS row of k_nlapplyg: NG projections padded to 80 B (rows
                                // r and r + 4 would share banks at 64 B)
constexpr int GLCAP = 1024;     // max images meeting the cubes of one group
constexpr int MAXM = 43;        // largest cube side one workgroup marches ((M+1)/2*M <= 1024)
constexpr int MAXIT = 500;      // Lanczos max_iter (sq.cpp:967)

__device__ __forceinline__ int wrapn(int a, int n)
{
    int r = a % n;
    return r < 0 ? r + n : r;
}

template <typename T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

// Workgroup barrier that orders LDS traffic only: global loads issued
// before it stay in flight (a full __syncthreads() also waits for them).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums NV doubles over the workgroup; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *red /* [16][NV] */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) red[wave * NV + i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[w * NV + i];
            v[i] = s;
        }
    }
}

// ---------------------------------------------------------------- arguments
struct Geo {
    int n, nloc, M, M2, M3, E;    // E = n + 2 nloc (extended grid side)
    int pitch, plane;             // LDS plane pitch and size
    double lap[FDN + 1];          // coeff_lap
    const int *node_list;         // [B] node ids of this batch
    const int *done;              // [B] Lanczos freeze flags (nullable)
};

// Nonlocal apply data: for every point of the extended grid
// [-nloc, n - 1 + nloc]^3 the (image-projector id, U value) pairs of the
// projectors that cover it, in ELL form: slot j of point g at [j][g], with
// cnt[g] valid slots (images in list order, projectors in lm order).
template <typename T>
struct NL {
    const unsigned char *cnt;     // [E^3]
    const T *u;                   // [K][E^3]
    const int *ip;                // [K][E^3]
    const unsigned short *ip16;   // [K][E^3] the same ids in 16 bits (n_ip <= 65535), else null
    size_t E3;
    int k_alloc;                  // K rows allocated
    const double *s;              // [B][n_ip] projections of the input vector
    int n_ip;
};

// CHEBD: CHEB + D column; TAN: the forward-mode step d_k = 2 A d_{k-1} - d_{k-2} + 2 dA t_{k-1};
// ADJ: the reverse step tb_k = 2 A tb_{k+1} - tb_{k+2} + g_k with G += 2 tb_{k+1} t_k
enum { MODE_HSUB = 0, MODE_CHEB = 1, MODE_LANCZOS = 2, MODE_CHEBD = 3, MODE_TAN = 4, MODE_ADJ = 5 };

template <typename T>
struct ZArgs {
    Geo g;
    const T *nlv;                 // [B][M^3] nonlocal term of the input vector (k_nlapply)
    const T *veff;                // [n^3]
    T *X;                         // input (HSUB/CHEB: t_k; LANCZOS: r')
    T *Y;                         // HSUB: out; CHEB: t_{k-1} in / t_{k+1} out; LANCZOS: v
    T *O;                         // CHEB: t_{k+1} goes here instead of Y when set
    T *XO, *YO;                   // LANCZOS: r'_{k+1} and v_k out (the other pair of
                                  // buffers: the other half-workgroup reads X, Y as halo)
    // CHEB per node
    const double *chi, *zee;      // indexed by batch slot, or by node id if chi_by_node
    int chi_by_node;
    int first;                    // CHEB: t1 = (H t0 - chi t0)/zee, no t_{k-1}
    double *center;               // [B] t_{k+1}[c]
    double *dots;                 // [B][2 halves][2] (t_{k+1}.t_{k+1}, t_{k+1}.t_k)
    // LANCZOS per node
    const double *ca, *ib, *bprev;  // [B] alpha_{k-1}, 1/beta_k, beta_k
    double *lred;                 // [B][2 halves][4] v.w, r.r, r.v, v.v
    // CHEBD: CHEB with a density-matrix column (forces.cpp:410-460):
    // D += Ci[nd][k] * t_k / h^3 for the produced t_k
    T *D;                         // [B][M^3]
    const double *Ci;             // [nn][npl+1]
    int k, npl;
    double inv_h3;
    // TAN: dA = diag(dV)/zee on the cube; dots become (t_k.d_k, d_k.t_{k-1} + t_k.d_{k-1})
    const double *dV;             // [n^3] the direction
    const T *tk, *tkm;            // [B][M^3] t_k and t_{k-1} of the forward recurrence
    // ADJ (k < K2): the source g_k of dL/dmu (b_j = 2 C_j wt, Ci at k and the
    // doubled moments) and the gradient accumulator G
    const T *tkp;                 // [B][M^3] t_{k+1}
    T *G;                         // [B][M^3]
    const double *wt;             // [nn] dL/drho weights (after the Fermi shift)
    int K2, q4;
};

// ---------------------------------------------------------------- z-march
// Raw buffer loads with a per-workgroup resource: a lane offset at or past
// the range returns 0 without touching memory, which makes every load of
// the march branch-free (cube edges, idle lanes and empty ELL slots).
constexpr unsigned OOB = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *p, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ int bload_i32(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ int bload_u8(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
__device__ __forceinline__ int bload_u16(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}

// Nonlocal term of B node vectors, NL_b(p) = sum_{J,lm} U_J,lm(p) s_b[J,lm]
// with the projections s_b of the vector (VnlocsubTimesVec, nonlocal.cpp:173-191):
// one workgroup per node, the node's projections in LDS, one thread per
// point at a time through the ELL entries of the extended-grid point.
template <typename T>
struct NArgs {
    Geo g;
    NL<T> nl;
    T *nlv;                       // [B][M^3] out
};

template <typename T, int K>
__global__ void __launch_bounds__(512) k_nlapply(NArgs<T> A)
{
    constexpr int KS = K > 0 ? K : 1;
    constexpr unsigned TS = sizeof(T);
    const Geo &G = A.g;
    const int b = blockIdx.x;
    if (G.done && G.done[b]) return;
    const int node = G.node_list[b];
    const int n = G.n, M = G.M, M2 = G.M2, E = G.E;
    const int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    extern __shared__ unsigned char smem_raw[];
    double *sv = reinterpret_cast<double *>(smem_raw);
    {
        const double *__restrict__ sIn = A.nl.s + (size_t)b * A.nl.n_ip;
        for (int i = threadIdx.x; i < A.nl.n_ip; i += blockDim.x) sv[i] = sIn[i];
    }
    __syncthreads();
    const int E3 = (int)A.nl.E3;
    const int KA = A.nl.k_alloc;
    const __amdgpu_buffer_rsrc_t rEU = brsrc(A.nl.u, (size_t)KA * E3 * TS);
    const __amdgpu_buffer_rsrc_t rEI = brsrc(A.nl.ip, (size_t)KA * E3 * 4);
    const __amdgpu_buffer_rsrc_t rEC = brsrc(A.nl.cnt, (size_t)E3);
    T *out = A.nlv + (size_t)b * G.M3;
    const int g00 = (nk * E + nj) * E + ni;
#pragma unroll 2
    for (int p = threadIdx.x; p < G.M3; p += blockDim.x) {
        const int z = p / M2, r = p - z * M2, y = r / M, x = r - y * M;
        const int g = g00 + (z * E + y) * E + x;
        const int cn = bload_u8(rEC, (unsigned)g);
        double a = 0.0;
        if (K > 0) {
            T u[KS];
            int ip[KS];
#pragma unroll
            for (int j = 0; j < KS; ++j) {
                const bool ok = j < cn;
                u[j] = bload<T>(rEU, ok ? (unsigned)(j * E3 + g) * TS : OOB);
                ip[j] = bload_i32(rEI, ok ? (unsigned)(j * E3 + g) * 4u : OOB);
            }
#pragma unroll
            for (int j = 0; j < KS; ++j) a += sv[ip[j]] * (double)u[j];
        } else {
            for (int j = 0; j < cn; ++j) a += sv[A.nl.ip[j * E3 + g]] * (double)A.nl.u[j * E3 + g];
        }
        out[p] = (T)a;
    }
}

// Two workgroups per node cube, one per half of its rows: half h owns rows
// [y0, y1) (y0 = 0, y1 = ceil(M/2) for h = 0; the rest for h = 1) and thread
// (x, y) one column of them for every z.  The 13 z-neighbours of each column
// live in a register window that shifts by one plane per step; the current
// plane is staged in LDS for the x/y-neighbours, together with the six rows
// beyond the half (loaded as plain input values, no window: only the y-stencil
// reads them).  The pointwise inputs of plane z (Veff, the nonlocal term,
// t_{k-1} or v_{k-1}, D, the halo rows) are loaded during step z - 1 into one
// of two register sets that alternate between the steps of the loop (unrolled
// by two, so no copy waits on a load in flight); the window feed of plane
// z + 7 is issued at the start of step z and inserted at its end.  The
// per-plane barrier orders LDS traffic only.  Each half reduces its own dot
// products into its slot of lred / dots ([B][2][.]); the half that holds the
// cube centre writes the centre value.
constexpr int ZHALF = 2;
__host__ __device__ inline int zhalf_rows(int M, int h) { return h == 0 ? (M + 1) / 2 : M / 2; }

template <typename T, int MODE>
__device__ __forceinline__ void zmarch_body(const ZArgs<T> &A)
{
    constexpr bool TAN = MODE == MODE_TAN;
    constexpr bool ADJ = MODE == MODE_ADJ;
    constexpr bool CH = MODE == MODE_CHEB || MODE == MODE_CHEBD || TAN || ADJ;
    constexpr bool YQ = MODE == MODE_LANCZOS && sizeof(T) == 8;
    constexpr unsigned TS = sizeof(T);
    const Geo &G = A.g;
    const int b = blockIdx.x / ZHALF, h = blockIdx.x % ZHALF;
    if (G.done && G.done[b]) return;
    const int node = G.node_list[b];
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2, PP = G.pitch;
    const int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    const int y0 = h == 0 ? 0 : (M + 1) / 2, ny = zhalf_rows(M, h);
    const int tid = threadIdx.x;
    const int pyl = tid / M;
    const int x0 = tid - pyl * M;
    const int py = y0 + pyl;
    const bool act = pyl < ny;
    // halo point of this thread: row hy (6 rows beyond the half), column hx
    const int hr = tid / M, hx = tid - hr * M;
    const int hy = h == 0 ? y0 + ny + hr : y0 - FDN + hr;
    const bool hact = hr < FDN && hy >= 0 && hy < M;

    // LDS: two staged planes (rows y0 - 6 .. y0 + ny + 5) | reduction scratch.
    // The planes and the register window hold the stored type (FP32 in the
    // mixed mode: fewer registers, two workgroups per CU).
    typedef T W;
    extern __shared__ unsigned char smem_raw[];
    W *lds = reinterpret_cast<W *>(smem_raw);
    double *red = reinterpret_cast<double *>(smem_raw + 2 * (size_t)G.plane * sizeof(W));
    __shared__ double sh_center;
    for (int i = tid; i < 2 * G.plane; i += blockDim.x) lds[i] = (W)0;
    if (tid == 0) sh_center = 0.0;

    // buffer resources of this node's vectors and of Veff
    const size_t base = (size_t)b * G.M3;
    const size_t vb = (size_t)G.M3 * TS;
    T *Xw = A.X + base;
    T *Yb = A.Y + base;
    T *Yob = ((CH && A.O) ? A.O : A.Y) + base;
    T *Db = MODE == MODE_CHEBD ? A.D + base : nullptr;
    const __amdgpu_buffer_rsrc_t rX = brsrc(Xw, vb), rY = brsrc(Yb, vb);
    const __amdgpu_buffer_rsrc_t rN = brsrc(A.nlv + base, vb);
    const __amdgpu_buffer_rsrc_t rD = brsrc(MODE == MODE_CHEBD ? Db : Yb, vb);
    const __amdgpu_buffer_rsrc_t rV = brsrc(A.veff, (size_t)n * n * n * TS);
    const __amdgpu_buffer_rsrc_t rTk = brsrc((TAN || ADJ) ? A.tk + base : Yb, vb);
    const __amdgpu_buffer_rsrc_t rTm = brsrc((TAN || ADJ) ? A.tkm + base : Yb, vb);
    const __amdgpu_buffer_rsrc_t rTp = brsrc(ADJ ? A.tkp + base : Yb, vb), rG = brsrc(ADJ ? A.G + base : Yb, vb);
    const __amdgpu_buffer_rsrc_t rDV = brsrc(TAN ? (const void *)A.dV : (const void *)Yb,
                                             TAN ? (size_t)n * n * n * sizeof(double) : vb);

    const int col = py * M + x0;                 // offset of the column in a plane
    const int hcol = hy * M + hx;                // offset of the halo point in a plane
    const int gx = wrapn(ni - nloc + x0, n);
    const int gyo = wrapn(nj - nloc + py, n) * n;
    // Arithmetic is FP64 in both modes: the mixed mode stores vectors, Veff
    // and the nonlocal term in FP32 (half the HBM traffic) but forms H v, the
    // recurrence and the scalings in FP64, so that only the storage rounding of
    // each vector (not ~40 FP32 roundings of terms of size (chi + zee)|t| per
    // point) enters the moments.
    typedef double R;
    const R hl = -0.5;
    const R c0d = (R)(3 * G.lap[0]);
    R cl[FDN + 1];
#pragma unroll
    for (int a = 0; a <= FDN; ++a) cl[a] = (R)G.lap[a];

    // per-node scalars
    R chi = 0, izee = 0, two_izee = 0;
    T ca = 0, ib = 0;
    R bp = 0;
    R dk = 0, d0 = 0;   // DM-column weights of t_k (and t_0 on the first step)
    if (CH) {
        const int ci_idx = A.chi_by_node ? node : b;
        chi = A.chi[ci_idx];
        izee = 1.0 / A.zee[ci_idx];
        two_izee = 2.0 / A.zee[ci_idx];
        if (MODE == MODE_CHEBD) {
            const double *ci = A.Ci + (size_t)node * (A.npl + 1);
            dk = ci[A.k] * A.inv_h3;
            d0 = ci[0] * A.inv_h3;
        }
    }
    if (MODE == MODE_LANCZOS) {
        ca = (T)A.ca[b]; ib = (T)A.ib[b]; bp = A.bprev[b];
    }
    // ADJ: b_j = 2 C_j wt (dL/dmu_j) for the terms of g_k; the centre term
    // carries b_k (and, at k = 1, minus the b_{2k'-1} of every doubled odd
    // moment, each of which subtracts t_1[c])
    double a_gc = 0.0, a_b2k = 0.0, a_b2km = 0.0, a_b2kp = 0.0;
    bool adj_hi = false, adj_hp = false;
    if (ADJ) {
        const double *ci = A.Ci + (size_t)node * (A.npl + 1);
        const double wtn = A.wt[node];
        auto be = [&](int j) { return 2.0 * ci[j] * wtn; };
        const int k = A.k;
        a_gc = be(k) * A.inv_h3;
        if (k == 1) {
            double so = 0.0;
            for (int kk = A.q4 + 1; kk <= A.K2; ++kk) so += be(2 * kk - 1);
            a_gc -= so * A.inv_h3;
        }
        adj_hi = k > A.q4;
        adj_hp = k + 1 > A.q4 && k + 1 <= A.K2;
        if (adj_hi) { a_b2k = be(2 * k); a_b2km = be(2 * k - 1); }
        if (adj_hp) a_b2kp = be(2 * k + 1);
    }
    const bool cheb_prev = CH && !A.first;
    const bool dacc = MODE == MODE_CHEBD && !A.first;

    // byte offset of (plane q) in a node vector for this column / halo point, or OOB
    auto voff = [&](int q) -> unsigned {
        return (act && q >= 0 && q < M) ? (unsigned)(q * M2 + col) * TS : OOB;
    };
    auto hoff = [&](int q) -> unsigned {
        return (hact && q >= 0 && q < M) ? (unsigned)(q * M2 + hcol) * TS : OOB;
    };
    // window feed of plane q: the raw loads, formed into the input value
    // (LANCZOS: v = (r' - alpha v_{k-1}) / beta) only where the plane enters
    // the window, so that no wait for them comes earlier
    struct FD {
        T x, y;
    };
    auto load_feed = [&](int q) -> FD {
        const unsigned o = voff(q);
        FD f;
        f.x = bload<T>(rX, o);
        f.y = MODE == MODE_LANCZOS ? bload<T>(rY, o) : (T)0;
        return f;
    };
    auto formed = [&](const FD &f) -> T {
        if (MODE == MODE_LANCZOS) return (f.x - ca * f.y) * ib;
        return f.x;
    };
    struct PW {
        T ve, nl, pv, dv;   // Veff, nonlocal term, t_{k-1} / v_{k-1}, D
        T hx, hyv;          // halo point: the two stored vectors
        T tt, tm;           // TAN, ADJ: t_k, t_{k-1}
        double dvv;         // TAN: the direction dV at the point
        T tp, gv;           // ADJ: t_{k+1}, G
    };
    auto fill_pw = [&](int z, PW &P) {
        const unsigned o = voff(z);
        const int vo = (wrapn(nk - nloc + z, n) * n) * n + gyo + gx;
        P.ve = bload<T>(rV, (act && z < M) ? (unsigned)vo * TS : OOB);
        P.nl = bload<T>(rN, o);
        // FP64 LANCZOS takes v_{k-1} from the feed queue (yq) instead
        P.pv = (cheb_prev || (MODE == MODE_LANCZOS && !YQ)) ? bload<T>(rY, o) : (T)0;
        P.dv = dacc ? bload<T>(rD, o) : (T)0;
        const unsigned ho = hoff(z);
        P.hx = bload<T>(rX, ho);
        P.hyv = MODE == MODE_LANCZOS ? bload<T>(rY, ho) : (T)0;
        if (TAN) {
            P.tt = bload<T>(rTk, o);
            P.tm = bload<T>(rTm, o);
            P.dvv = bload<double>(rDV, (act && z < M) ? (unsigned)vo * 8u : OOB);
        } else {
            P.tt = ADJ ? bload<T>(rTk, o) : (T)0;
            P.tm = ADJ ? bload<T>(rTm, adj_hi ? o : OOB) : (T)0;
            P.dvv = 0.0;
        }
        P.tp = ADJ ? bload<T>(rTp, adj_hp ? o : OOB) : (T)0;
        P.gv = ADJ ? bload<T>(rG, o) : (T)0;
    };

    // window: w[s] holds plane z - 6 + s during step z; FP64 LANCZOS: yq[j]
    // holds the stored v_{k-1} of plane z + 1 + j, read by the feed six steps
    // before the plane's pointwise inputs are needed (one HBM read of that
    // vector; FP32 passes measured faster re-reading it from L2)
    W w[WIN];
    T yq[FDN];
    PW SA, SB;
#pragma unroll
    for (int s = 0; s < FDN; ++s) w[s] = 0;
    fill_pw(0, SA);
#pragma unroll
    for (int s = 0; s <= FDN; ++s) {   // planes 0..6
        const FD f = load_feed(s);
        w[FDN + s] = (W)formed(f);
        if (YQ) {
            if (s == 0) SA.pv = f.y;
            else yq[s - 1] = f.y;
        }
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    __syncthreads();

    const int lrow = (pyl + FDN) * PP + x0 + FDN;          // LDS slot of the column
    const int hrow = (hy - y0 + FDN) * PP + hx + FDN;      // LDS slot of the halo point
    // step z: S holds plane z's pointwise inputs; N receives plane z + 1's.
    auto step = [&](int z, PW &S, PW &N) {
        const FD qf = load_feed(z + FDN + 1);
        fill_pw(z + 1, N);
        if (YQ) N.pv = yq[0];
        // stage plane z
        W *pl = lds + (z & 1) * G.plane;
        if (act) pl[lrow] = w[FDN];
        if (hact) pl[hrow] = (W)(MODE == MODE_LANCZOS ? (T)((S.hx - ca * S.hyv) * ib) : S.hx);
        lds_barrier();
        if (act) {
            const W *row = pl + lrow;
            const R cv = (R)w[FDN];
            R t = cv * c0d;
#pragma unroll
            for (int a = 1; a <= FDN; ++a) {
                const R xl = (R)row[-a], xr = (R)row[a];
                const R ym = (R)row[-a * PP], yp = (R)row[a * PP];
                t += (xl + xr + ym + yp + (R)w[FDN - a] + (R)w[FDN + a]) * cl[a];
                // keep the LDS reads of one offset in flight at a time (register pressure)
                __builtin_amdgcn_sched_barrier(0);
            }
            const R hv = hl * t + (R)S.ve * cv + (R)S.nl;
            const int ozp = z * M2 + col;
            if (MODE == MODE_HSUB) {
                Yb[ozp] = (T)hv;
            } else if (CH) {
                R o;
                if (A.first) o = (hv - chi * cv) * izee;
                else o = (hv - chi * cv) * two_izee - (R)S.pv;
                const T os = (T)o;   // the stored value is the one reduced and projected
                if (ADJ) {
                    // tb_k = (2 A tb_{k+1} - tb_{k+2}) + g_k and G += 2 tb_{k+1} t_k
                    // (tb_{k+1} is this point's input), in the terms' order of
                    // the elementwise step k_adj_update
                    const bool ctr = z == nloc && py == nloc && x0 == nloc;
                    double g = ctr ? a_gc : 0.0;
                    const double tkv = (double)S.tt;
                    if (adj_hi) g += (4.0 * a_b2k * tkv + 2.0 * a_b2km * (double)S.tm) * A.inv_h3;
                    if (adj_hp) g += 2.0 * a_b2kp * (double)S.tp * A.inv_h3;
                    Yob[ozp] = (T)((double)os + g);
                    A.G[base + ozp] = (T)((double)S.gv + 2.0 * cv * tkv);
                } else if (TAN) {
                    // d_k = (2 A d_{k-1} - d_{k-2}) + 2 dA t_{k-1}; dots t_k.d_k and
                    // d_k.t_{k-1} + t_k.d_{k-1} (d_{k-1} is this point's input)
                    const T dn = (T)((double)os + 2.0 * (S.dvv * izee) * (double)S.tm);
                    Yob[ozp] = dn;
                    acc[0] += (double)S.tt * (double)dn;
                    acc[1] += (double)dn * (double)S.tm + (double)S.tt * cv;
                    if (z == nloc && py == nloc && x0 == nloc) sh_center = (double)dn;
                } else {
                    Yob[ozp] = os;
                    acc[0] += (double)os * (double)os;
                    acc[1] += (double)os * cv;
                    if (z == nloc && py == nloc && x0 == nloc) sh_center = (double)os;
                }
                if (MODE == MODE_CHEBD) {
                    if (A.first) {   // D = C0 t0/h^3 + C1 t1/h^3 with t0 = e_c
                        const bool ctr = z == nloc && py == nloc && x0 == nloc;
                        Db[ozp] = (T)((R)os * dk + (ctr ? d0 : 0.0));
                    } else {
                        Db[ozp] = (T)((R)S.dv + (R)os * dk);
                    }
                }
            } else {  // LANCZOS: r' = H v - beta_k v_{k-1}; store v and r'
                const T q = (T)(hv - bp * (R)S.pv);
                A.XO[base + ozp] = q;
                A.YO[base + ozp] = (T)cv;
                acc[0] += cv * hv;
                acc[1] += (double)q * (double)q;
                acc[2] += (double)q * cv;
                acc[3] += cv * cv;
            }
        }
        // shift the window and insert plane z + 7
#pragma unroll
        for (int s = 0; s < WIN - 1; ++s) w[s] = w[s + 1];
        w[WIN - 1] = (W)formed(qf);
        if (YQ) {
#pragma unroll
            for (int j = 0; j < FDN - 1; ++j) yq[j] = yq[j + 1];
            yq[FDN - 1] = qf.y;
        }
    };
    for (int z = 0; z < M; z += 2) {
        step(z, SA, SB);
        if (z + 1 < M) step(z + 1, SB, SA);
    }
    if (MODE == MODE_HSUB || ADJ) return;
    block_sum<4>(acc, red);
    if (tid == 0) {
        const bool has_ctr = nloc >= y0 && nloc < y0 + ny;
        if (CH) {
            if (has_ctr) A.center[b] = sh_center;
            A.dots[(2 * (size_t)b + h) * 2] = acc[0];
            A.dots[(2 * (size_t)b + h) * 2 + 1] = acc[1];
        } else {
            double *o = A.lred + (2 * (size_t)b + h) * 4;
            o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
        }
    }
}

// FP32, up to 640 threads (M <= 35): at most 80 VGPRs, six waves per SIMD,
// so that two 10-wave workgroups share a CU whatever SIMDs their waves land
// on (at 96 VGPRs, five per SIMD, the second one does not always fit).
// Otherwise up to 1024 threads and no register cap.
template <typename T, int MODE>
__global__ __attribute__((amdgpu_waves_per_eu(6, 8))) void __launch_bounds__(640) k_zmarch(ZArgs<T> A)
{
    zmarch_body<T, MODE>(A);
}
template <typename T, int MODE>
__global__ void __launch_bounds__(1024) k_zmarch_big(ZArgs<T> A)
{
    zmarch_body<T, MODE>(A);
}

// ---------------------------------------------------------------- projections
template <typename T>
struct PArgs {
    Geo g;
    const T *V;                   // [B][M^3] vector to project (LANCZOS: r')
    const T *V2;                  // LANCZOS: previous normalized vector
    double *s;                    // [B][n_ip] out
    int n_ip;
    const int *list, *list_cnt;   // [B][LCAP], [B]
    const int *img_type, *img_nn, *ip_base;
    const long *udv_off;
    const T *udv;
    const int *nlocc, *nproj;     // per type
    const double *inv_denom;      // [type][16]
    const double *ca, *ib;        // LANCZOS: projects v = (V - ca V2) * ib
    const int *glist, *glist_cnt; // group image lists (k_projg, k_projs)
    int B;                        // slots in the batch (k_projg, k_projs)
    int pu;                       // k_projs: largest union plane (points) staged at once
};

// s[b][ip] = <U_ip, v_b>/Denom over the intersection of the image's rc-cube
// and node b's cube (nonlocal.cpp:149-171).  One wave per image.  With
// FORMED, v is the Lanczos vector formed from the two stored vectors with the
// same rounding as the z-march forms it, so that the projections always
// belong to the vector being multiplied (a recursion on projections would
// amplify rounding by |alpha/beta| per pass).
template <typename T, int MAXP, int FORMED>
__global__ void __launch_bounds__(256) k_proj(PArgs<T> A)
{
    const Geo &G = A.g;
    const int b = blockIdx.x;
    if (G.done && G.done[b]) return;
    const int node = G.node_list[b];
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2;
    const int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
    const int cnt = A.list_cnt[b];
    const T *__restrict__ V = A.V + (size_t)b * G.M3;
    const T *__restrict__ V2 = FORMED ? A.V2 + (size_t)b * G.M3 : nullptr;
    double *sb = A.s + (size_t)b * A.n_ip;
    T cab = 0, ibb = 1;
    if (FORMED) { cab = (T)A.ca[b]; ibb = (T)A.ib[b]; }
    for (int li = wave; li < cnt; li += nwave) {
        const int J = A.list[(size_t)b * LCAP + li];
        const int t = A.img_type[J];
        const int nc = A.nlocc[t], np = A.nproj[t], w = 2 * nc + 1;
        const int cx = A.img_nn[3 * J], cy = A.img_nn[3 * J + 1], cz = A.img_nn[3 * J + 2];
        const int lox = max(cx - nc, ni - nloc), hix = min(cx + nc, ni + nloc);
        const int loy = max(cy - nc, nj - nloc), hiy = min(cy + nc, nj + nloc);
        const int loz = max(cz - nc, nk - nloc), hiz = min(cz + nc, nk + nloc);
        const int bx = hix - lox + 1, by = hiy - loy + 1;
        const int area = bx * by;
        const int w2 = w * w, w3 = w2 * w;
        const T *U = A.udv + A.udv_off[J];
        double acc[MAXP];
#pragma unroll
        for (int i = 0; i < MAXP; ++i) acc[i] = 0.0;
        for (int a = lane; a < area; a += 64) {
            const int iy = a / bx, ix = a - iy * bx;
            const int gx = lox + ix, gy = loy + iy;
            const int cidx = (gy - nj + nloc) * M + (gx - ni + nloc);
            const int ridx = (gy - cy + nc) * w + (gx - cx + nc);
            for (int gz = loz; gz <= hiz; ++gz) {
                const size_t o = (size_t)(gz - nk + nloc) * M2 + cidx;
                double v;
                if (FORMED) v = (double)((V[o] - cab * V2[o]) * ibb);
                else v = (double)V[o];
                const int q = (gz - cz + nc) * w2 + ridx;
#pragma unroll
                for (int i = 0; i < MAXP; ++i)
                    if (i < np) acc[i] += v * (double)U[(size_t)i * w3 + q];
            }
        }
#pragma unroll
        for (int i = 0; i < MAXP; ++i) {
            if (i < np) {
                const double r = wave_sum(acc[i]);
                if (lane == 0) sb[A.ip_base[J] + i] = r * A.inv_denom[16 * t + i];
            }
        }
    }
}

// The projections of k_proj for the NG consecutive slots of a group at once:
// one wave per image of the group list reads each U value once for the
// group (the cubes of neighbouring nodes overlap in all but a few planes)
// and accumulates it against every member's vector.  Lane l always takes
// the points of the image's rc-box whose xy index in the box is l mod 64,
// z ascending, and a point outside a member's cube adds an exact zero
// (buffer loads past the member's range read 0), so each member's sums do
// not depend on which slots share its group.  Members are handled GS at a
// time (GS accumulator sets of MAXP values per lane).
template <typename T, int MAXP, int FORMED>
__global__ void __launch_bounds__(256) k_projg(PArgs<T> A)
{
    constexpr int GS = MAXP <= 4 ? NG : MAXP <= 8 ? NG / 2 : NG / 4;
    constexpr unsigned TS = sizeof(T);
    const Geo &G = A.g;
    const int q = blockIdx.x;
    const int b0 = q * NG, gn = min(NG, A.B - b0);
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2, M3 = G.M3;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
    int ni[NG], nj[NG], nk[NG];
    bool on[NG];
    bool any = false;
#pragma unroll
    for (int m = 0; m < NG; ++m) {
        on[m] = m < gn && !(G.done && G.done[b0 + m]);
        const int node = G.node_list[b0 + min(m, gn - 1)];
        ni[m] = node % n; nj[m] = (node / n) % n; nk[m] = node / (n * n);
        any |= on[m];
    }
    if (!any) return;
    const size_t gb = (size_t)gn * M3 * TS;
    const __amdgpu_buffer_rsrc_t rX = brsrc(A.V + (size_t)b0 * M3, gb);
    const __amdgpu_buffer_rsrc_t rY = brsrc((FORMED ? A.V2 : A.V) + (size_t)b0 * M3, gb);
    const int cnt = A.glist_cnt[q];
    for (int li = wave; li < cnt; li += nwave) {
        const int J = A.glist[(size_t)q * GLCAP + li];
        const int t = A.img_type[J];
        const int nc = A.nlocc[t], np = A.nproj[t], w = 2 * nc + 1, w2 = w * w, w3 = w2 * w;
        const int ox = A.img_nn[3 * J] - nc, oy = A.img_nn[3 * J + 1] - nc, oz = A.img_nn[3 * J + 2] - nc;
        const __amdgpu_buffer_rsrc_t rU = brsrc(A.udv + A.udv_off[J], (size_t)np * w3 * TS);
        const double *idn = A.inv_denom + 16 * t;
        const int ipb = A.ip_base[J];
#pragma unroll 1
        for (int m0 = 0; m0 < NG; m0 += GS) {
            // member ranges in box coordinates, and their union
            int lx[GS], hx[GS], ly[GS], hy[GS], lz[GS], hz[GS];
            int ulx = w, uhx = -1, uly = w, uhy = -1, ulz = w, uhz = -1;
            bool ne[GS];
#pragma unroll
            for (int g = 0; g < GS; ++g) {
                const int m = m0 + g;
                lx[g] = max(0, ni[m] - nloc - ox); hx[g] = min(w - 1, ni[m] + nloc - ox);
                ly[g] = max(0, nj[m] - nloc - oy); hy[g] = min(w - 1, nj[m] + nloc - oy);
                lz[g] = max(0, nk[m] - nloc - oz); hz[g] = min(w - 1, nk[m] + nloc - oz);
                ne[g] = on[m] && lx[g] <= hx[g] && ly[g] <= hy[g] && lz[g] <= hz[g];
                if (!ne[g]) { lx[g] = w; hx[g] = -1; lz[g] = w; hz[g] = -1; }
                else {
                    ulx = min(ulx, lx[g]); uhx = max(uhx, hx[g]);
                    uly = min(uly, ly[g]); uhy = max(uhy, hy[g]);
                    ulz = min(ulz, lz[g]); uhz = max(uhz, hz[g]);
                }
            }
            if (uhx < 0) continue;
            T cab[GS], ibb[GS];
#pragma unroll
            for (int g = 0; g < GS; ++g) {
                cab[g] = 0; ibb[g] = 1;
                if (FORMED && ne[g]) { cab[g] = (T)A.ca[b0 + m0 + g]; ibb[g] = (T)A.ib[b0 + m0 + g]; }
            }
            double acc[GS][MAXP];
#pragma unroll
            for (int g = 0; g < GS; ++g)
#pragma unroll
                for (int i = 0; i < MAXP; ++i) acc[g][i] = 0.0;
            const int a_hi = (uhy + 1) * w;
            for (int a0 = (uly * w) & ~63; a0 < a_hi; a0 += 64) {
                const int a = a0 + lane;
                const int by = a / w, bx = a - by * w;
                const bool inu = a < w2 && bx >= ulx && bx <= uhx && by >= uly && by <= uhy;
                int base[GS];
                bool inm[GS];
#pragma unroll
                for (int g = 0; g < GS; ++g) {
                    const int m = m0 + g;
                    inm[g] = inu && bx >= lx[g] && bx <= hx[g] && by >= ly[g] && by <= hy[g];
                    base[g] = m * M3 + (oz - nk[m] + nloc) * M2 + (oy + by - nj[m] + nloc) * M +
                              (ox + bx - ni[m] + nloc);
                }
#pragma unroll 1
                for (int bz = ulz; bz <= uhz; ++bz) {
                    T u[MAXP];
#pragma unroll
                    for (int i = 0; i < MAXP; ++i)
                        u[i] = bload<T>(rU, (inu && i < np) ? (unsigned)(i * w3 + bz * w2 + a) * TS : OOB);
#pragma unroll
                    for (int g = 0; g < GS; ++g) {
                        const bool ok = inm[g] && bz >= lz[g] && bz <= hz[g];
                        const unsigned o = ok ? (unsigned)(base[g] + bz * M2) * TS : OOB;
                        double v;
                        if (FORMED) v = (double)((bload<T>(rX, o) - cab[g] * bload<T>(rY, o)) * ibb[g]);
                        else v = (double)bload<T>(rX, o);
#pragma unroll
                        for (int i = 0; i < MAXP; ++i) acc[g][i] += v * (double)u[i];
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < GS; ++g) {
                if (!ne[g]) continue;
                double *sb = A.s + (size_t)(b0 + m0 + g) * A.n_ip + ipb;
#pragma unroll
                for (int i = 0; i < MAXP; ++i) {
                    if (i < np) {
                        const double r = wave_sum(acc[g][i]);
                        if (lane == 0) sb[i] = r * idn[i];
                    }
                }
            }
        }
    }
}

// Wave sums of NV values at once (NV a power of two <= 32): the values are
// halved between lane pairs (xor 32, 16, ...) until each lane holds one, then
// summed over the remaining lane bits.  Every value goes through the same
// pairing tree as wave_sum (xor 32, 16, 8, 4, 2, 1), so each total is
// bitwise the one wave_sum would give.  Returns, in lane l, the total of
// value l >> (6 - log2 NV).
// One halving step between lanes l and l ^ 32 (O = 32) or l ^ 16 (O = 16)
// with the gfx950 half-wave swaps: permlane{32,16}_swap(p, q) leaves the
// lanes whose bit O is clear with (own p, partner's p) and the others with
// (partner's q, own q), so their sum is p or q summed over the lane pair.
template <int O>
__device__ __forceinline__ double swap_sum(double p, double q)
{
    const unsigned long long pb = __builtin_bit_cast(unsigned long long, p);
    const unsigned long long qb = __builtin_bit_cast(unsigned long long, q);
    unsigned lo0, lo1, hi0, hi1;
    if (O == 32) {
        const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)pb, (unsigned)qb, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(pb >> 32), (unsigned)(qb >> 32), false, false);
        lo0 = lo[0]; lo1 = lo[1]; hi0 = hi[0]; hi1 = hi[1];
    } else {
        const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)pb, (unsigned)qb, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(pb >> 32), (unsigned)(qb >> 32), false, false);
        lo0 = lo[0]; lo1 = lo[1]; hi0 = hi[0]; hi1 = hi[1];
    }
    const double a = __builtin_bit_cast(double, ((unsigned long long)hi0 << 32) | lo0);
    const double b = __builtin_bit_cast(double, ((unsigned long long)hi1 << 32) | lo1);
    return a + b;
}

template <int NV>
__device__ __forceinline__ double wave_sums(double (&v)[NV], int lane)
{
#pragma unroll
    for (int lv = 0; lv < 6; ++lv) {
        const int o = 32 >> lv, h = (NV >> lv) / 2;   // h values move to the partner
        if (h > 0 && o >= 16) {
#pragma unroll
            for (int j = 0; j < NV / 2; ++j)
                if (j < h) v[j] = o == 32 ? swap_sum<32>(v[j], v[j + h]) : swap_sum<16>(v[j], v[j + h]);
        } else if (h > 0) {
            const bool up = (lane & o) != 0;
#pragma unroll
            for (int j = 0; j < NV / 2; ++j) {
                if (j < h) {
                    // the empty asm keeps the selects on register values (a
                    // select between two array elements would otherwise become
                    // a load from a selected address and move v to scratch)
                    double lo = v[j], hi = v[j + h];
                    asm volatile("" : "+v"(lo), "+v"(hi));
                    const double send = up ? lo : hi;
                    const double keep = up ? hi : lo;
                    v[j] = keep + __shfl_xor(send, o, 64);
                }
            }
        } else {
            v[0] += __shfl_xor(v[0], o, 64);
        }
    }
    return v[0];
}

// The projections of k_projg with every vector value read from HBM once per
// group: the workgroup walks the union of the members' cubes plane by plane
// (global z ascending), stages each plane of the members' (formed) vectors
// in LDS as [point][member] with zeros outside a member's cube, and the
// images crossing the plane (one range of the z-sorted group list) are dealt
// to the waves: one wave sums an image's slice of the slab and adds it,
// divided by Denom, to the image's projections in memory with one hardware
// FP64 atomic per value (slabs are separated by a fence and a barrier, so
// each projection grows slab by slab in z order whichever wave did the sum).
// Slabs hold SZ planes whose first global z is a multiple of SZ.  Lane l takes the box points
// whose xy index in the image box is l mod 64; the plane's partial sums are
// combined by wave_sums and added plane after plane, so a member's totals
// depend only on its own cube.  A group whose union plane exceeds pu points
// is walked one member at a time, in the same order.

// Image data that k_projs reads per image and plane, cached in LDS per group
struct ImgMeta {
    int ox, oy, oz;               // rc-box origin (unwrapped grid coordinates)
    int w, np, t, ipb;            // box side, projectors, type, first image-projector id
    long uoff;                    // offset of the image's U block in udv
};

// First index of the z-sorted group list whose image centre z exceeds zc - 1
// (centre z = oz + (w - 1) / 2).
__device__ __forceinline__ int first_cz_at_least(const ImgMeta *im, int cnt, int zc)
{
    int lo = 0, hi = cnt;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (im[mid].oz + (im[mid].w - 1) / 2 < zc) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename T, int MAXP, int FORMED, int SZ, int NT, int NPU, int NM>
__global__ void __launch_bounds__(NT) k_projs(PArgs<T> A)
{
    // NM members per workgroup: all NG slots of the group, or a 2x2 block of
    // its 4x2 slots (two workgroups per group, half the LDS each)
    static_assert(NM == NG || NM == 4, "k_projs: NM is NG or 4");
    constexpr int SUBS = NG / NM;
    constexpr int NV = NM * MAXP;
    constexpr int LV = NV == 32 ? 5 : NV == 16 ? 4 : NV == 8 ? 3 : NV == 4 ? 2 : 0;
    static_assert(LV > 0, "k_projs: NM * MAXP must be 4, 8, 16 or 32");
    // projector values are loaded and summed for the first NPU of the MAXP
    // slots (3 for a table of p projectors); the other sums stay exact zeros
    static_assert(NPU >= 1 && NPU <= MAXP, "k_projs: NPU of MAXP slots");
    constexpr unsigned TS = sizeof(T);
    constexpr int RPI = SZ >= 4 ? 2 : 3;   // rounds of 64 box points per U load batch
    constexpr int PPT = SZ >= 4 ? 2 : 3;   // staged points per thread per load batch
    const Geo &G = A.g;
    const int q = blockIdx.x / SUBS, hsub = blockIdx.x % SUBS;
    const int b0 = q * NG, gn = min(NG, A.B - b0);
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2, M3 = G.M3;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwave = blockDim.x >> 6;
    int ni[NM], nj[NM], nk[NM], gs[NM];   // gs: the member's slot in the group
    bool on[NM];
    T cab[NM], ibb[NM];
    int nact = 0;
    unsigned onmask = 0;
    int x0 = 1 << 30, x1 = -(1 << 30), y0 = 1 << 30, y1 = -(1 << 30);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        gs[m] = NM == NG ? m : (m & 1) + 2 * hsub + 4 * (m >> 1);
        on[m] = gs[m] < gn && !(G.done && G.done[b0 + gs[m]]);
        const int node = G.node_list[b0 + min(gs[m], gn - 1)];
        ni[m] = node % n; nj[m] = (node / n) % n; nk[m] = node / (n * n);
        cab[m] = 0; ibb[m] = 1;
        if (on[m]) {
            ++nact;
            onmask |= 1u << m;
            if (FORMED) { cab[m] = (T)A.ca[b0 + gs[m]]; ibb[m] = (T)A.ib[b0 + gs[m]]; }
            x0 = min(x0, ni[m]); x1 = max(x1, ni[m]);
            y0 = min(y0, nj[m]); y1 = max(y1, nj[m]);
        }
    }
    const int cnt = A.glist_cnt[q];
    if (nact == 0 || cnt == 0) return;
    // LDS: image data [cnt] | staged slab [SZ][pu][NM]
    __shared__ int sh_ncmax, sh_lo, sh_hi;
    if (tid == 0) sh_ncmax = 0;
    __syncthreads();
    extern __shared__ unsigned char smem_raw[];
    ImgMeta *im = reinterpret_cast<ImgMeta *>(smem_raw);
    T *pl = reinterpret_cast<T *>(smem_raw + ((size_t)cnt * sizeof(ImgMeta) + 15) / 16 * 16);
    for (int li = tid; li < cnt; li += blockDim.x) {
        const int J = A.glist[(size_t)q * GLCAP + li];
        const int t = A.img_type[J], nc = A.nlocc[t];
        ImgMeta e;
        e.ox = A.img_nn[3 * J] - nc; e.oy = A.img_nn[3 * J + 1] - nc; e.oz = A.img_nn[3 * J + 2] - nc;
        e.w = 2 * nc + 1; e.np = A.nproj[t]; e.t = t; e.ipb = A.ip_base[J];
        e.uoff = A.udv_off[J];
        im[li] = e;
        atomicMax(&sh_ncmax, nc);
        // this group's projections start at zero; each slab adds its totals
#pragma unroll
        for (int m = 0; m < NM; ++m)
            if ((onmask >> m) & 1u)
                for (int i = 0; i < e.np; ++i) A.s[(size_t)(b0 + gs[m]) * A.n_ip + e.ipb + i] = 0.0;
    }
    const size_t gb = (size_t)gn * M3 * TS;
    const __amdgpu_buffer_rsrc_t rX = brsrc(A.V + (size_t)b0 * M3, gb);
    const __amdgpu_buffer_rsrc_t rY = brsrc((FORMED ? A.V2 : A.V) + (size_t)b0 * M3, gb);
    const bool together = (x1 - x0 + M) * (y1 - y0 + M) <= A.pu;
#pragma unroll 1
    for (int sub = 0; sub < (together ? 1 : NM); ++sub) {
        // members staged in this walk, and the union of their cubes
        bool in_sub[NM];
        int ux0 = 1 << 30, ux1 = -(1 << 30), uy0 = 1 << 30, uy1 = -(1 << 30), uz0 = 1 << 30, uz1 = -(1 << 30);
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            in_sub[m] = on[m] && (together || m == sub);
            if (in_sub[m]) {
                ux0 = min(ux0, ni[m] - nloc); ux1 = max(ux1, ni[m] + nloc);
                uy0 = min(uy0, nj[m] - nloc); uy1 = max(uy1, nj[m] + nloc);
                uz0 = min(uz0, nk[m] - nloc); uz1 = max(uz1, nk[m] + nloc);
            }
        }
        if (ux1 < ux0) continue;
        const int ux = ux1 - ux0 + 1, uy = uy1 - uy0 + 1, up = ux * uy, sp = SZ * up;
        // slabs of SZ planes whose first global z is a multiple of SZ (the same
        // slab boundaries for every grouping of the nodes)
        const int zs0 = (uz0 + 4096) / SZ * SZ - 4096;
#pragma unroll 1
        for (int gzs = zs0; gzs <= uz1; gzs += SZ) {
            // the previous slab is consumed, and this wave's atomics (and the
            // zeroing stores) have been performed in L2 (vmcnt 0) before any
            // wave adds the next slab's totals
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            // the images that can cross the slab form one range of the z-sorted list
            if (tid == 0) {
                sh_lo = first_cz_at_least(im, cnt, gzs - sh_ncmax);
                sh_hi = first_cz_at_least(im, cnt, gzs + SZ + sh_ncmax);
            }
            // stage the slab ([plane][point][member]): every load of the
            // thread's points issued before the first store
#pragma unroll 1
            for (int p0 = 0; p0 < sp; p0 += PPT * (int)blockDim.x) {
                T xv[PPT][NM], yv[PPT][NM];
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int p = p0 + r * (int)blockDim.x + tid;
                    const int sl = p / up, pq = p - sl * up;
                    const int yy = pq / ux, xx = pq - yy * ux;
                    const int gx = ux0 + xx, gy = uy0 + yy, gz = gzs + sl;
#pragma unroll
                    for (int m = 0; m < NM; ++m) {
                        const int lx = gx - ni[m] + nloc, ly = gy - nj[m] + nloc, lz = gz - nk[m] + nloc;
                        const bool ok = p < sp && in_sub[m] && (unsigned)lx < (unsigned)M &&
                                        (unsigned)ly < (unsigned)M && (unsigned)lz < (unsigned)M;
                        const unsigned o = ok ? (unsigned)(gs[m] * M3 + lz * M2 + ly * M + lx) * TS : OOB;
                        xv[r][m] = bload<T>(rX, o);
                        yv[r][m] = FORMED ? bload<T>(rY, o) : (T)0;
                    }
                }
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int p = p0 + r * (int)blockDim.x + tid;
                    if (p < sp) {
#pragma unroll
                        for (int m = 0; m < NM; ++m)
                            pl[(size_t)p * NM + m] = FORMED ? (xv[r][m] - cab[m] * yv[r][m]) * ibb[m] : xv[r][m];
                    }
                }
            }
            __syncthreads();
            const int li_lo = __builtin_amdgcn_readfirstlane(sh_lo), li_hi = __builtin_amdgcn_readfirstlane(sh_hi);
#pragma unroll 1
            for (int li = li_lo + wave; li < li_hi; li += nwave) {
                // the image's data is wave-uniform: keep it in scalar registers
                const ImgMeta &e = im[li];
                const int w = __builtin_amdgcn_readfirstlane(e.w), oz = __builtin_amdgcn_readfirstlane(e.oz);
                const int bz0 = gzs - oz;   // box z of the slab's first plane
                if (bz0 + SZ <= 0 || bz0 >= w) continue;
                const int ox = __builtin_amdgcn_readfirstlane(e.ox), oy = __builtin_amdgcn_readfirstlane(e.oy);
                const int rlo = max(0, uy0 - oy), rhi = min(w - 1, uy1 - oy);
                const int clo = max(0, ux0 - ox), chi = min(w - 1, ux1 - ox);
                if (rlo > rhi || clo > chi) continue;
                const int np = __builtin_amdgcn_readfirstlane(e.np), w2 = w * w, w3 = w2 * w;
                const long uoff = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(e.uoff >> 32)) << 32) |
                                         (unsigned)__builtin_amdgcn_readfirstlane((int)e.uoff));
                const __amdgpu_buffer_rsrc_t rU = brsrc(A.udv + uoff, (size_t)np * w3 * TS);
                // the value this lane adds at the end: member kw / MAXP, projector kw % MAXP
                const int kw = lane >> (6 - LV), mw = kw / MAXP, iw = kw - mw * MAXP;
                const bool writer = (lane & ((64 >> LV) - 1)) == 0 && ((onmask >> mw) & 1u) && iw < np;
                const int t = __builtin_amdgcn_readfirstlane(e.t), ipb = __builtin_amdgcn_readfirstlane(e.ipb);
                const double idn = writer ? A.inv_denom[16 * t + iw] : 0.0;
                double acc[NV];
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[k] = 0.0;
                const int a_hi = (rhi + 1) * w;
                // a / w through the float reciprocal: exact, since (a + 1/2)/w
                // stays at least 1/(2w) away from an integer (a < w^2, w < 1000)
                const float rw = 1.0f / (float)w;
                // RPI rounds of 64 box points at a time, their U loads issued together
#pragma unroll 1
                for (int c0 = (rlo * w) & ~63; c0 < a_hi; c0 += 64 * RPI) {
                    T u[RPI][SZ][NPU];
                    int pp[RPI];
                    bool in[RPI];
#pragma unroll
                    for (int r = 0; r < RPI; ++r) {
                        const int a = c0 + 64 * r + lane;
                        const int by = (int)(((float)a + 0.5f) * rw), bx = a - by * w;
                        in[r] = a < a_hi && by >= rlo && by <= rhi && bx >= clo && bx <= chi;
#pragma unroll
                        for (int z = 0; z < SZ; ++z) {
                            const bool okz = in[r] && (unsigned)(bz0 + z) < (unsigned)w;
#pragma unroll
                            for (int i = 0; i < NPU; ++i)
                                u[r][z][i] = bload<T>(rU, (okz && i < np) ? (unsigned)(i * w3 + (bz0 + z) * w2 + a) * TS : OOB);
                        }
                        pp[r] = in[r] ? (oy + by - uy0) * ux + (ox + bx - ux0) : 0;
                    }
#pragma unroll
                    for (int r = 0; r < RPI; ++r) {
#pragma unroll
                        for (int z = 0; z < SZ; ++z) {
                            const T *pv = pl + ((size_t)z * up + pp[r]) * NM;
                            T v[NM];
#pragma unroll
                            for (int m = 0; m < NM; ++m) v[m] = pv[m];
                            if (in[r]) {
#pragma unroll
                                for (int m = 0; m < NM; ++m)
#pragma unroll
                                    for (int i = 0; i < NPU; ++i)
                                        acc[m * MAXP + i] += (double)v[m] * (double)u[r][z][i];
                            }
                        }
                    }
                }
                const double tot = wave_sums<NV>(acc, lane);
                // one lane per value and slab; slabs are ordered by the fence
                // and barrier above, so the additions happen in slab order
                const int gw = NM == NG ? mw : (mw & 1) + 2 * hsub + 4 * (mw >> 1);   // gs[mw]
                if (writer) unsafeAtomicAdd(A.s + (size_t)(b0 + gw) * A.n_ip + ipb + iw, tot * idn);
            }
        }
    }
}

// The projections of k_projs with every image owned by one wave for its
// whole z-extent.  The workgroup walks the union of its members' cubes in
// slabs of SZ planes staged in LDS as in k_projs; image li of the z-sorted
// group list belongs to wave li mod nwave, and the k-th of a wave's images
// that meets the walk's union holds that wave's register slot k mod R from
// the slab where its box enters the union to the slab where it leaves.  Lane l keeps, for each slot, the partial sums of the
// box points whose xy index is l mod 64 over every plane, so the wave reduces
// (wave_sums) and stores each projection once, when its image leaves: no
// per-slab reduction, no atomics, no zeroing pass.  Per lane and member the
// products are added plane by plane, z ascending, in the same round order
// for every grouping of the nodes, so each member's totals depend only on its
// own cube.  An image whose box misses the union in z contributes zero.
// Requires one box size for every image of the list and R larger than the
// slot distance of any two images of a wave alive at once (k_group_images
// computes the R a batch needs; the host picks a compiled R at least that).
template <typename T, int FORMED, int SZ, int NT, int NM, int R, int RPI>
__global__ void __launch_bounds__(NT) k_projr(PArgs<T> A)
{
    static_assert(NM == NG || NM == 4, "k_projr: NM is NG or 4");
    constexpr int SUBS = NG / NM;
    constexpr int NV = NM;                    // one projector per image (s channel)
    constexpr int LV = NV == 8 ? 3 : 2;
    constexpr unsigned TS = sizeof(T);
    constexpr int PPT = 2;                    // staged points per thread per load batch
    const Geo &G = A.g;
    const int q = blockIdx.x / SUBS, hsub = blockIdx.x % SUBS;
    const int b0 = q * NG, gn = min(NG, A.B - b0);
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2, M3 = G.M3;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int NW = NT / 64;
    int ni[NM], nj[NM], nk[NM], gs[NM];
    bool on[NM];
    T cab[NM], ibb[NM];
    int nact = 0;
    unsigned onmask = 0;
    // extent of every member of the group (frozen ones included), so that the
    // choice between one walk and member-by-member walks is the same in every
    // pass (k_group_images checks the slot plan of that choice)
    int x0 = 1 << 30, x1 = -(1 << 30), y0 = 1 << 30, y1 = -(1 << 30);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        gs[m] = NM == NG ? m : (m & 1) + 2 * hsub + 4 * (m >> 1);
        on[m] = gs[m] < gn && !(G.done && G.done[b0 + gs[m]]);
        const int node = G.node_list[b0 + min(gs[m], gn - 1)];
        ni[m] = node % n; nj[m] = (node / n) % n; nk[m] = node / (n * n);
        cab[m] = 0; ibb[m] = 1;
        if (gs[m] < gn) {
            x0 = min(x0, ni[m]); x1 = max(x1, ni[m]);
            y0 = min(y0, nj[m]); y1 = max(y1, nj[m]);
        }
        if (on[m]) {
            ++nact;
            onmask |= 1u << m;
            if (FORMED) { cab[m] = (T)A.ca[b0 + gs[m]]; ibb[m] = (T)A.ib[b0 + gs[m]]; }
        }
    }
    const int cnt = A.glist_cnt[q];
    if (nact == 0 || cnt == 0) return;
    // LDS: image data [cnt] | staged slab [SZ][NM][padded plane]
    extern __shared__ unsigned char smem_raw[];
    ImgMeta *im = reinterpret_cast<ImgMeta *>(smem_raw);
    T *pl = reinterpret_cast<T *>(smem_raw + ((size_t)cnt * sizeof(ImgMeta) + 15) / 16 * 16);
    for (int li = tid; li < cnt; li += blockDim.x) {
        const int J = A.glist[(size_t)q * GLCAP + li];
        const int t = A.img_type[J], nc = A.nlocc[t];
        ImgMeta e;
        e.ox = A.img_nn[3 * J] - nc; e.oy = A.img_nn[3 * J + 1] - nc; e.oz = A.img_nn[3 * J + 2] - nc;
        e.w = 2 * nc + 1; e.np = A.nproj[t]; e.t = t; e.ipb = A.ip_base[J];
        e.uoff = A.udv_off[J];
        im[li] = e;
    }
    __syncthreads();
    const size_t gb = (size_t)gn * M3 * TS;
    const __amdgpu_buffer_rsrc_t rX = brsrc(A.V + (size_t)b0 * M3, gb);
    const __amdgpu_buffer_rsrc_t rY = brsrc((FORMED ? A.V2 : A.V) + (size_t)b0 * M3, gb);
    // LDS rows of the slab are padded to a pitch congruent to the box side
    // modulo 32 values, so that the box rows one wave reads fall on disjoint
    // banks; A.pu is the slab capacity per plane and member
    const int wbox = __builtin_amdgcn_readfirstlane(im[0].w);
    auto upitch_of = [&](int u) { return u + (((wbox - u) % 32) + 32) % 32; };
    const bool together = upitch_of(x1 - x0 + M) * (y1 - y0 + M) <= A.pu;
    // the value this lane stores for a leaving image: member kw
    const int kw = lane >> (6 - LV);
    const bool wlane = (lane & ((64 >> LV) - 1)) == 0;
    const int gw = NM == NG ? kw : (kw & 1) + 2 * hsub + 4 * (kw >> 1);   // gs[kw]
    double *sw = A.s + (size_t)(b0 + gw) * A.n_ip;
#pragma unroll 1
    for (int sub = 0; sub < (together ? 1 : NM); ++sub) {
        bool in_sub[NM];
        unsigned submask = 0;
        int ux0 = 1 << 30, ux1 = -(1 << 30), uy0 = 1 << 30, uy1 = -(1 << 30), uz0 = 1 << 30, uz1 = -(1 << 30);
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            in_sub[m] = on[m] && (together || m == sub);
            if (in_sub[m]) {
                submask |= 1u << m;
                ux0 = min(ux0, ni[m] - nloc); ux1 = max(ux1, ni[m] + nloc);
                uy0 = min(uy0, nj[m] - nloc); uy1 = max(uy1, nj[m] + nloc);
                uz0 = min(uz0, nk[m] - nloc); uz1 = max(uz1, nk[m] + nloc);
            }
        }
        if (ux1 < ux0) continue;
        const bool writer = wlane && ((submask >> kw) & 1u);
        const int ux = ux1 - ux0 + 1, uy = uy1 - uy0 + 1, upx = upitch_of(ux), up = upx * uy;
        const int np_ = ux * uy, sp = SZ * np_;   // staged points of a slab (unpadded)
        const int zs0 = (uz0 + 4096) / SZ * SZ - 4096;
        // per-wave slot state: images started (slotted or not), slotted
        // images, live-slot mask, per-slot image data
        int nstart = 0, nslot = 0;
        unsigned live = 0;
        int s_oz[R], s_ze[R], s_w[R], s_ipb[R], s_t[R];
        long s_uoff[R];
        int s_pp[R][RPI];                     // LDS point of the lane's box point, or -1
        double acc[R][NM];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            s_oz[r] = s_ze[r] = s_w[r] = s_ipb[r] = s_t[r] = 0;
            s_uoff[r] = 0;
#pragma unroll
            for (int k = 0; k < RPI; ++k) s_pp[r][k] = -1;
#pragma unroll
            for (int m = 0; m < NM; ++m) acc[r][m] = 0.0;
        }
#pragma unroll 1
        for (int gzs = zs0; gzs <= uz1; gzs += SZ) {
            // the previous slab is consumed before it is overwritten
            __syncthreads();
#pragma unroll 1
            for (int p0 = 0; p0 < sp; p0 += PPT * (int)blockDim.x) {
                T xv[PPT][NM], yv[PPT][NM];
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int p = p0 + r * (int)blockDim.x + tid;
                    const int sl = p / np_, pq = p - sl * np_;
                    const int yy = pq / ux, xx = pq - yy * ux;
                    const int gx = ux0 + xx, gy = uy0 + yy, gz = gzs + sl;
#pragma unroll
                    for (int m = 0; m < NM; ++m) {
                        const int lx = gx - ni[m] + nloc, ly = gy - nj[m] + nloc, lz = gz - nk[m] + nloc;
                        const bool ok = p < sp && in_sub[m] && (unsigned)lx < (unsigned)M &&
                                        (unsigned)ly < (unsigned)M && (unsigned)lz < (unsigned)M;
                        const unsigned o = ok ? (unsigned)(gs[m] * M3 + lz * M2 + ly * M + lx) * TS : OOB;
                        xv[r][m] = bload<T>(rX, o);
                        yv[r][m] = FORMED ? bload<T>(rY, o) : (T)0;
                    }
                }
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int p = p0 + r * (int)blockDim.x + tid;
                    if (p < sp) {
                        // member-major: [plane][member][row][padded column]
                        const int sl = p / np_, pq = p - sl * np_;
                        const int yy = pq / ux, xx = pq - yy * ux;
                        T *dst = pl + (size_t)sl * NM * up + yy * upx + xx;
#pragma unroll
                        for (int m = 0; m < NM; ++m)
                            dst[(size_t)m * up] = FORMED ? (xv[r][m] - cab[m] * yv[r][m]) * ibb[m] : xv[r][m];
                    }
                }
            }
            __syncthreads();
            const int zl = gzs + SZ - 1;   // last plane of the slab
            // images of this wave that enter the union in this slab take their slot
#pragma unroll 1
            for (;;) {
                const int li = wave + NW * nstart;
                if (li >= cnt) break;
                const int oz = __builtin_amdgcn_readfirstlane(im[li].oz);
                const int w = __builtin_amdgcn_readfirstlane(im[li].w);
                const int bx0 = __builtin_amdgcn_readfirstlane(im[li].ox), by0 = __builtin_amdgcn_readfirstlane(im[li].oy);
                const bool miss = oz + w - 1 < uz0 || oz > uz1 || bx0 + w - 1 < ux0 || bx0 > ux1 ||
                                  by0 + w - 1 < uy0 || by0 > uy1;
                const int zs = max(oz, uz0);
                if (!miss && zs > zl) break;
                if (miss) {   // the box misses the union: the projection is zero
                    const int ipb = __builtin_amdgcn_readfirstlane(im[li].ipb);
                    if (writer) sw[ipb] = 0.0;
                    ++nstart;
                    continue;
                }
                const int slot = nslot % R;
                ++nslot;
                const int ox = __builtin_amdgcn_readfirstlane(im[li].ox), oy = __builtin_amdgcn_readfirstlane(im[li].oy);
                const int rlo = max(0, uy0 - oy), rhi = min(w - 1, uy1 - oy);
                const int clo = max(0, ux0 - ox), chi = min(w - 1, ux1 - ox);
                const float rw = 1.0f / (float)w;
                int pp[RPI];
#pragma unroll
                for (int k = 0; k < RPI; ++k) {
                    const int a = 64 * k + lane;
                    const int by = (int)(((float)a + 0.5f) * rw), bx = a - by * w;
                    const bool in = a < w * w && by >= rlo && by <= rhi && bx >= clo && bx <= chi;
                    pp[k] = in ? (oy + by - uy0) * upx + (ox + bx - ux0) : -1;
                }
                const int ipb = __builtin_amdgcn_readfirstlane(im[li].ipb), t = __builtin_amdgcn_readfirstlane(im[li].t);
                const long uoff = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(im[li].uoff >> 32)) << 32) |
                                         (unsigned)__builtin_amdgcn_readfirstlane((int)im[li].uoff));
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r == slot) {
                        s_oz[r] = oz; s_ze[r] = min(oz + w - 1, uz1); s_w[r] = w; s_ipb[r] = ipb; s_t[r] = t;
                        s_uoff[r] = uoff;
#pragma unroll
                        for (int k = 0; k < RPI; ++k) s_pp[r][k] = pp[k];
#pragma unroll
                        for (int m = 0; m < NM; ++m) acc[r][m] = 0.0;
                    }
                }
                live |= 1u << slot;
                ++nstart;
            }
            // every live slot adds the slab's planes of its box
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!((live >> r) & 1u)) continue;
                const int w = s_w[r], w2 = w * w;
                const int za = max(gzs, max(s_oz[r], uz0)), zb = min(zl, s_ze[r]);
                const __amdgpu_buffer_rsrc_t rU = brsrc(A.udv + s_uoff[r], (size_t)w2 * w * TS);
                for (int z = za; z <= zb; ++z) {
                    const int bz = z - s_oz[r];
                    T u[RPI];
#pragma unroll
                    for (int k = 0; k < RPI; ++k)
                        u[k] = bload<T>(rU, s_pp[r][k] >= 0 ? (unsigned)(bz * w2 + 64 * k + lane) * TS : OOB);
#pragma unroll
                    for (int k = 0; k < RPI; ++k) {
                        const T *pv = pl + (size_t)(z - gzs) * NM * up + max(s_pp[r][k], 0);
                        T v[NM];
#pragma unroll
                        for (int m = 0; m < NM; ++m) v[m] = pv[(size_t)m * up];
                        if (s_pp[r][k] >= 0) {
#pragma unroll
                            for (int m = 0; m < NM; ++m) acc[r][m] += (double)v[m] * (double)u[k];
                        }
                    }
                }
                if (s_ze[r] <= zl) {   // the image leaves the union in this slab
                    const double tot = wave_sums<NV>(acc[r], lane);
                    if (writer) sw[s_ipb[r]] = tot * A.inv_denom[16 * s_t[r]];
                    live &= ~(1u << r);
                }
            }
        }
    }
}

// The projections of k_projr for tables of up to four projectors (the p
// channel), accumulated on the matrix cores: every live image of a wave owns
// one FP64 4x4x4 MFMA accumulator (four 4x4 blocks, one double per lane).
// Block b = q + 2 ph multiplies members 4q..4q+3 (rows) by the image's
// projectors (columns) over four box points (K) of phase ph: a box row of
// side w <= 16 is covered by two MFMAs, point x = 8 hh + 4 ph + k, each into
// its own accumulator (two independent chains).  The accumulators and the
// two phases are added when the image leaves the union (lane l + lane l ^ 8),
// so each total is one fixed sum of MFMA partial sums.  Operands: A = the
// member's staged value at the point (LDS, member-major, member stride
// congruent to 8 modulo 32 values so that the 8 members of one point fall on
// distinct banks), B = the projector value (zero past the box or the
// union).  A member's row of the product depends only on its own values, so
// each member's totals are independent of the grouping.  Slot plan, walks and
// staging as in k_projr.  Lane layout of v_mfma_f64_4x4x4f64 on gfx950:
// A[b][i][k] at lane 16k + 4b + i, B[b][k][j] at 16k + 4b + j, D[b][i][j] at
// 16i + 4b + j.
template <typename T, int FORMED, int SZ, int NT, int R>
__global__ void __launch_bounds__(NT) k_projm(PArgs<T> A)
{
    constexpr int NM = NG;
    constexpr unsigned TS = sizeof(T);
    constexpr int PPT = 2;
    const Geo &G = A.g;
    const int q = blockIdx.x;
    const int b0 = q * NG, gn = min(NG, A.B - b0);
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2, M3 = G.M3;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int NW = NT / 64;
    int ni[NM], nj[NM], nk[NM];
    bool on[NM];
    T cab[NM], ibb[NM];
    int nact = 0;
    unsigned onmask = 0;
    int x0 = 1 << 30, x1 = -(1 << 30), y0 = 1 << 30, y1 = -(1 << 30);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        on[m] = m < gn && !(G.done && G.done[b0 + m]);
        const int node = G.node_list[b0 + min(m, gn - 1)];
        ni[m] = node % n; nj[m] = (node / n) % n; nk[m] = node / (n * n);
        cab[m] = 0; ibb[m] = 1;
        if (m < gn) {
            x0 = min(x0, ni[m]); x1 = max(x1, ni[m]);
            y0 = min(y0, nj[m]); y1 = max(y1, nj[m]);
        }
        if (on[m]) {
            ++nact;
            onmask |= 1u << m;
            if (FORMED) { cab[m] = (T)A.ca[b0 + m]; ibb[m] = (T)A.ib[b0 + m]; }
        }
    }
    const int cnt = A.glist_cnt[q];
    if (nact == 0 || cnt == 0) return;
    // LDS: image data [cnt] | staged slab [SZ][NM][member stride]
    extern __shared__ unsigned char smem_raw[];
    ImgMeta *im = reinterpret_cast<ImgMeta *>(smem_raw);
    T *pl = reinterpret_cast<T *>(smem_raw + ((size_t)cnt * sizeof(ImgMeta) + 15) / 16 * 16);
    for (int li = tid; li < cnt; li += blockDim.x) {
        const int J = A.glist[(size_t)q * GLCAP + li];
        const int t = A.img_type[J], nc = A.nlocc[t];
        ImgMeta e;
        e.ox = A.img_nn[3 * J] - nc; e.oy = A.img_nn[3 * J + 1] - nc; e.oz = A.img_nn[3 * J + 2] - nc;
        e.w = 2 * nc + 1; e.np = A.nproj[t]; e.t = t; e.ipb = A.ip_base[J];
        e.uoff = A.udv_off[J];
        im[li] = e;
    }
    __syncthreads();
    const size_t gb = (size_t)gn * M3 * TS;
    const __amdgpu_buffer_rsrc_t rX = brsrc(A.V + (size_t)b0 * M3, gb);
    const __amdgpu_buffer_rsrc_t rY = brsrc((FORMED ? A.V2 : A.V) + (size_t)b0 * M3, gb);
    const int wbox = __builtin_amdgcn_readfirstlane(im[0].w);
    auto upitch_of = [&](int u) { return u + (((wbox - u) % 32) + 32) % 32; };
    auto mstride_of = [&](int u) { return u + (((8 - u) % 32) + 32) % 32; };
    const bool together = mstride_of(upitch_of(x1 - x0 + M) * (y1 - y0 + M)) <= A.pu;
    // lane roles in the MFMA: K index, block (member quad, point phase), row/column
    const int kk = lane >> 4, bb = (lane >> 2) & 3, ii = lane & 3;
    const int mA = 4 * (bb & 1) + ii;          // member of the lane's A value
    const int bxl0 = 4 * (bb >> 1) + kk;       // box x of the lane's point, first MFMA of a row
    // writers of the totals: lanes of phase 0; member 4q + i, projector j
    const int mw = 4 * ((lane >> 2) & 1) + (lane >> 4), jw = lane & 3;
    const bool wl = (lane & 8) == 0;
    double *sw = A.s + (size_t)(b0 + mw) * A.n_ip;
#pragma unroll 1
    for (int sub = 0; sub < (together ? 1 : NM); ++sub) {
        bool in_sub[NM];
        unsigned submask = 0;
        int ux0 = 1 << 30, ux1 = -(1 << 30), uy0 = 1 << 30, uy1 = -(1 << 30), uz0 = 1 << 30, uz1 = -(1 << 30);
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            in_sub[m] = on[m] && (together || m == sub);
            if (in_sub[m]) {
                submask |= 1u << m;
                ux0 = min(ux0, ni[m] - nloc); ux1 = max(ux1, ni[m] + nloc);
                uy0 = min(uy0, nj[m] - nloc); uy1 = max(uy1, nj[m] + nloc);
                uz0 = min(uz0, nk[m] - nloc); uz1 = max(uz1, nk[m] + nloc);
            }
        }
        if (ux1 < ux0) continue;
        const bool writer = wl && ((submask >> mw) & 1u);
        const int ux = ux1 - ux0 + 1, uy = uy1 - uy0 + 1, upx = upitch_of(ux), ms = mstride_of(upx * uy);
        const int np_ = ux * uy, sp = SZ * np_;
        const int zs0 = (uz0 + 4096) / SZ * SZ - 4096;
        int nstart = 0, nslot = 0;
        unsigned live = 0;
        int s_oz[R], s_ze[R], s_ipb[R], s_t[R], s_np[R], s_ox[R], s_oy[R];
        long s_uoff[R];
        double cacc[R][2];   // per slot: the first and second MFMA of each box row
#pragma unroll
        for (int r = 0; r < R; ++r) {
            s_oz[r] = s_ze[r] = s_ipb[r] = s_t[r] = s_np[r] = s_ox[r] = s_oy[r] = 0;
            s_uoff[r] = 0;
            cacc[r][0] = cacc[r][1] = 0.0;
        }
#pragma unroll 1
        for (int gzs = zs0; gzs <= uz1; gzs += SZ) {
            __syncthreads();
#pragma unroll 1
            for (int p0 = 0; p0 < sp; p0 += PPT * (int)blockDim.x) {
                T xv[PPT][NM], yv[PPT][NM];
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int p = p0 + r * (int)blockDim.x + tid;
                    const int sl = p / np_, pq = p - sl * np_;
                    const int yy = pq / ux, xx = pq - yy * ux;
                    const int gx = ux0 + xx, gy = uy0 + yy, gz = gzs + sl;
#pragma unroll
                    for (int m = 0; m < NM; ++m) {
                        const int lx = gx - ni[m] + nloc, ly = gy - nj[m] + nloc, lz = gz - nk[m] + nloc;
                        const bool ok = p < sp && in_sub[m] && (unsigned)lx < (unsigned)M &&
                                        (unsigned)ly < (unsigned)M && (unsigned)lz < (unsigned)M;
                        const unsigned o = ok ? (unsigned)(m * M3 + lz * M2 + ly * M + lx) * TS : OOB;
                        xv[r][m] = bload<T>(rX, o);
                        yv[r][m] = FORMED ? bload<T>(rY, o) : (T)0;
                    }
                }
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int p = p0 + r * (int)blockDim.x + tid;
                    if (p < sp) {
                        const int sl = p / np_, pq = p - sl * np_;
                        const int yy = pq / ux, xx = pq - yy * ux;
                        T *dst = pl + (size_t)sl * NM * ms + yy * upx + xx;
#pragma unroll
                        for (int m = 0; m < NM; ++m)
                            dst[(size_t)m * ms] = FORMED ? (xv[r][m] - cab[m] * yv[r][m]) * ibb[m] : xv[r][m];
                    }
                }
            }
            __syncthreads();
            const int zl = gzs + SZ - 1;
#pragma unroll 1
            for (;;) {
                const int li = wave + NW * nstart;
                if (li >= cnt) break;
                const int oz = __builtin_amdgcn_readfirstlane(im[li].oz);
                const int w = wbox;
                const int bx0 = __builtin_amdgcn_readfirstlane(im[li].ox), by0 = __builtin_amdgcn_readfirstlane(im[li].oy);
                const bool miss = oz + w - 1 < uz0 || oz > uz1 || bx0 + w - 1 < ux0 || bx0 > ux1 ||
                                  by0 + w - 1 < uy0 || by0 > uy1;
                const int zs = max(oz, uz0);
                if (!miss && zs > zl) break;
                const int ipb = __builtin_amdgcn_readfirstlane(im[li].ipb);
                const int np = __builtin_amdgcn_readfirstlane(im[li].np);
                if (miss) {   // the box misses the union: the projections are zero
                    if (writer && jw < np) sw[ipb + jw] = 0.0;
                    ++nstart;
                    continue;
                }
                const int slot = nslot % R;
                ++nslot;
                const int t = __builtin_amdgcn_readfirstlane(im[li].t);
                const long uoff = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(im[li].uoff >> 32)) << 32) |
                                         (unsigned)__builtin_amdgcn_readfirstlane((int)im[li].uoff));
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r == slot) {
                        s_oz[r] = oz; s_ze[r] = min(oz + w - 1, uz1); s_ipb[r] = ipb; s_t[r] = t; s_np[r] = np;
                        s_ox[r] = bx0; s_oy[r] = by0; s_uoff[r] = uoff;
                        cacc[r][0] = cacc[r][1] = 0.0;
                    }
                }
                live |= 1u << slot;
                ++nstart;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!((live >> r) & 1u)) continue;
                const int w = wbox, w2 = w * w, w3 = w2 * w;
                const int ox = s_ox[r], oy = s_oy[r];
                const int za = max(gzs, max(s_oz[r], uz0)), zb = min(zl, s_ze[r]);
                const int rlo = max(0, uy0 - oy), rhi = min(w - 1, uy1 - oy);
                const int clo = max(0, ux0 - ox), chi = min(w - 1, ux1 - ox);
                const __amdgpu_buffer_rsrc_t rU = brsrc(A.udv + s_uoff[r], (size_t)s_np[r] * w3 * TS);
                const bool colj = ii < s_np[r];
                bool okx[2];
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) okx[hh] = bxl0 + 8 * hh >= clo && bxl0 + 8 * hh <= chi;
                for (int z = za; z <= zb; ++z) {
                    const int bz = z - s_oz[r];
                    const T *pz = pl + (size_t)(z - gzs) * NM * ms + (size_t)mA * ms;
                    // rows in chunks of RC, the next chunk's operands loaded
                    // while the current chunk's MFMAs run
                    constexpr int RC = 4;
                    T an[RC][2], bn[RC][2];
                    auto load_rows = [&](int by0) {
#pragma unroll
                        for (int j = 0; j < RC; ++j) {
                            const int by = by0 + j;
                            const bool rowok = by <= rhi;
                            const int lrow = (oy + by - uy0) * upx + (ox - ux0);
                            const int urow = ii * w3 + bz * w2 + by * w;
#pragma unroll
                            for (int hh = 0; hh < 2; ++hh) {
                                const int bx = bxl0 + 8 * hh;
                                const bool ok = rowok && okx[hh];
                                an[j][hh] = ok ? pz[lrow + bx] : (T)0;
                                bn[j][hh] = bload<T>(rU, (ok && colj) ? (unsigned)(urow + bx) * TS : OOB);
                            }
                        }
                    };
                    load_rows(rlo);
                    for (int by0 = rlo; by0 <= rhi; by0 += RC) {
                        T ac[RC][2], bc[RC][2];
#pragma unroll
                        for (int j = 0; j < RC; ++j)
#pragma unroll
                            for (int hh = 0; hh < 2; ++hh) { ac[j][hh] = an[j][hh]; bc[j][hh] = bn[j][hh]; }
                        if (by0 + RC <= rhi) load_rows(by0 + RC);
#pragma unroll
                        for (int j = 0; j < RC; ++j) {
                            if (by0 + j <= rhi) {
#pragma unroll
                                for (int hh = 0; hh < 2; ++hh)
                                    cacc[r][hh] = __builtin_amdgcn_mfma_f64_4x4x4f64((double)ac[j][hh], (double)bc[j][hh],
                                                                                     cacc[r][hh], 0, 0, 0);
                            }
                        }
                    }
                }
                if (s_ze[r] <= zl) {   // the image leaves the union: add the partial sums and store
                    const double c = cacc[r][0] + cacc[r][1];
                    const double tot = c + __shfl_xor(c, 8, 64);
                    if (writer && jw < s_np[r]) sw[s_ipb[r] + jw] = tot * A.inv_denom[16 * s_t[r] + jw];
                    live &= ~(1u << r);
                }
            }
        }
    }
}

// The nonlocal term of k_nlapply for the NG consecutive slots of a group:
// each point of the union of the members' cubes reads its ELL entries once
// and applies them to every member's projections, which sit in LDS
// compacted to the images of the group list (loc maps an image-projector id
// to its row, sl holds NG values per row).  Each member's sum runs over the
// point's entries in slot order, as in k_nlapply.  When the members' cubes
// are far apart (a group that spans two bricks), each member's cube is
// walked on its own instead of the union box.
template <typename T>
struct NGArgs {
    Geo g;
    NL<T> nl;
    T *nlv;                       // [B][M^3] out
    const int *glist, *glist_cnt;
    const int *img_type, *nproj, *ip_base;
    int B;
};

template <typename T, bool IP16>
__global__ void __launch_bounds__(512, 4) k_nlapplyg(NGArgs<T> A)
{
    constexpr unsigned TS = sizeof(T);
    const Geo &G = A.g;
    const int q = blockIdx.x;
    const int b0 = q * NG, gn = min(NG, A.B - b0);
    const int n = G.n, M = G.M, M2 = G.M2, M3 = G.M3, E = G.E;
    int ni[NG], nj[NG], nk[NG];
    bool on[NG];
    int nact = 0;
    int ex0 = 1 << 30, ex1 = -1, ey0 = 1 << 30, ey1 = -1, ez0 = 1 << 30, ez1 = -1;
#pragma unroll
    for (int m = 0; m < NG; ++m) {
        on[m] = m < gn && !(G.done && G.done[b0 + m]);
        const int node = G.node_list[b0 + min(m, gn - 1)];
        ni[m] = node % n; nj[m] = (node / n) % n; nk[m] = node / (n * n);
        if (on[m]) {
            ++nact;
            ex0 = min(ex0, ni[m]); ex1 = max(ex1, ni[m] + M - 1);
            ey0 = min(ey0, nj[m]); ey1 = max(ey1, nj[m] + M - 1);
            ez0 = min(ez0, nk[m]); ez1 = max(ez1, nk[m] + M - 1);
        }
    }
    if (nact == 0) return;
    // LDS: loc [n_ip] | row offsets [cnt + 1] | sl [rows][SLW]
    extern __shared__ unsigned char smem_raw[];
    int *loc = reinterpret_cast<int *>(smem_raw);
    const int cnt = A.glist_cnt[q];
    int *roff = loc + A.nl.n_ip;
    double *sl = reinterpret_cast<double *>(smem_raw + (((size_t)A.nl.n_ip + cnt + 1) * 4 + 15) / 16 * 16);
    const int *gl = A.glist + (size_t)q * GLCAP;
    for (int i = threadIdx.x; i < A.nl.n_ip; i += blockDim.x) loc[i] = 0;
    if (threadIdx.x == 0) {
        int r = 0;
        for (int l = 0; l < cnt; ++l) { roff[l] = r; r += A.nproj[A.img_type[gl[l]]]; }
        roff[cnt] = r;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < cnt; l += blockDim.x) {
        const int J = gl[l], r0 = roff[l], np = roff[l + 1] - r0, ipb = A.ip_base[J];
        for (int i = 0; i < np; ++i) {
            loc[ipb + i] = r0 + i;
#pragma unroll
            for (int m = 0; m < NG; ++m)
                sl[(size_t)(r0 + i) * SLW + m] = on[m] ? A.nl.s[(size_t)(b0 + m) * A.nl.n_ip + ipb + i] : 0.0;
        }
    }
    if (cnt == 0 && threadIdx.x < SLW) sl[threadIdx.x] = 0.0;
    __syncthreads();
    const int E3 = (int)A.nl.E3;
    const int KA = A.nl.k_alloc;
    const __amdgpu_buffer_rsrc_t rEU = brsrc(A.nl.u, (size_t)KA * E3 * TS);
    constexpr unsigned IS = IP16 ? 2u : 4u;   // bytes per ELL id
    const __amdgpu_buffer_rsrc_t rEI =
        brsrc(IP16 ? (const void *)A.nl.ip16 : (const void *)A.nl.ip, (size_t)KA * E3 * IS);
    const __amdgpu_buffer_rsrc_t rEC = brsrc(A.nl.cnt, (size_t)E3);
    T *out = A.nlv + (size_t)b0 * M3;
    // one extended-grid point: sum its entries against every member's
    // projections, four slots at a time up to the wave's largest count (a
    // slot past the point's count loads u = 0 and adds exact zeros)
    constexpr int CH = 4;
    auto point = [&](int gx, int gy, int gz, double a[NG]) {
        const int g = (gz * E + gy) * E + gx;
        const int cn = bload_u8(rEC, (unsigned)g);
#pragma unroll
        for (int m = 0; m < NG; ++m) a[m] = 0.0;
#pragma unroll 1
        for (int j0 = 0; j0 < KA && __ballot(cn > j0) != 0; j0 += CH) {
            T u[CH];
            int ip[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = j0 + c;
                const bool ok = j < cn;
                u[c] = bload<T>(rEU, ok ? (unsigned)(j * E3 + g) * TS : OOB);
                ip[c] = IP16 ? bload_u16(rEI, ok ? (unsigned)(j * E3 + g) * IS : OOB)
                             : bload_i32(rEI, ok ? (unsigned)(j * E3 + g) * IS : OOB);
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const double *row = sl + (size_t)loc[ip[c]] * SLW;
#pragma unroll
                for (int m = 0; m < NG; ++m) a[m] += row[m] * (double)u[c];
            }
        }
    };
    const int ux = ex1 - ex0 + 1, uy = ey1 - ey0 + 1, uz = ez1 - ez0 + 1;
    if ((long)ux * uy * uz <= (long)nact * M3) {
        const int uxy = ux * uy, tot = uxy * uz;
#pragma unroll 1
        for (int p = threadIdx.x; p < tot; p += blockDim.x) {
            const int z = p / uxy, r = p - z * uxy, y = r / ux, x = r - y * ux;
            const int gx = ex0 + x, gy = ey0 + y, gz = ez0 + z;
            double a[NG];
            point(gx, gy, gz, a);
#pragma unroll
            for (int m = 0; m < NG; ++m) {
                const unsigned lx = gx - ni[m], ly = gy - nj[m], lz = gz - nk[m];
                if (on[m] && lx < (unsigned)M && ly < (unsigned)M && lz < (unsigned)M)
                    out[(size_t)m * M3 + (lz * M2 + ly * M + lx)] = (T)a[m];
            }
        }
    } else {
        for (int m = 0; m < NG; ++m) {
            if (!on[m]) continue;
#pragma unroll 1
            for (int p = threadIdx.x; p < M3; p += blockDim.x) {
                const int z = p / M2, r = p - z * M2, y = r / M, x = r - y * M;
                double a[NG];
                point(ni[m] + x, nj[m] + y, nk[m] + z, a);
                out[(size_t)m * M3 + p] = (T)a[m];
            }
        }
    }
}

// ---------------------------------------------------------------- nonlocal forces
constexpr int FSLOT = 32;   // images whose rc-box holds one node centre

template <typename T>
struct FArgs {
    Geo g;
    const T *D;                   // [B][M^3] density-matrix columns
    const int *list, *list_cnt;
    const int *img_type, *img_nn, *img_atom, *nlocc, *nproj;
    const long *udv_off;
    const T *udv;
    const double *inv_denom;      // [type][16]
    double grad[FDN + 1];         // coeff_grad
    double *contrib;              // [B][FSLOT][4]: atom, and sum_lm s_lm,a U_lm(centre) for a = x,y,z
    int *ncontrib, *overflow;
};

// For each image whose rc-box holds node b's centre:
//   s_lm,a = <U_lm, d_a D>/Denom over the rc-box cut to the node cube, with
//   the FD gradient of D taken with zero values outside the cube, and the
//   image's term of [V_nl,J d_a D](centre) = sum_lm s_lm,a U_lm(centre)
// (forces.cpp:463-529 with VnlocsubTimesVec_J, nonlocal.cpp:522-724).
// One wave per image; slots follow the order of the node's image list.
template <typename T, int MAXP>
__global__ void __launch_bounds__(256) k_force_contract(FArgs<T> A)
{
    const Geo &G = A.g;
    const int b = blockIdx.x;
    const int node = G.node_list[b];
    const int n = G.n, nloc = G.nloc, M = G.M, M2 = G.M2;
    const int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
    const int cnt = A.list_cnt[b];
    __shared__ int slot_of[LCAP];
    if (threadIdx.x == 0) {
        int r = 0;
        for (int li = 0; li < cnt; ++li) {
            const int J = A.list[(size_t)b * LCAP + li];
            const int nc = A.nlocc[A.img_type[J]];
            const bool in = abs(A.img_nn[3 * J] - ni) <= nc && abs(A.img_nn[3 * J + 1] - nj) <= nc &&
                            abs(A.img_nn[3 * J + 2] - nk) <= nc;
            slot_of[li] = (in && r < FSLOT) ? r : -1;
            if (in) ++r;
        }
        if (r > FSLOT) atomicOr(A.overflow, 1);
        A.ncontrib[b] = r < FSLOT ? r : FSLOT;
    }
    __syncthreads();
    const T *__restrict__ D = A.D + (size_t)b * G.M3;
    auto Dat = [&](int x, int y, int z) -> double {
        return (x < 0 || y < 0 || z < 0 || x >= M || y >= M || z >= M) ? 0.0
                                                                        : (double)D[(size_t)z * M2 + y * M + x];
    };
    for (int li = wave; li < cnt; li += nwave) {
        const int slot = slot_of[li];
        if (slot < 0) continue;
        const int J = A.list[(size_t)b * LCAP + li];
        const int t = A.img_type[J];
        const int nc = A.nlocc[t], np = A.nproj[t], w = 2 * nc + 1, w2 = w * w, w3 = w2 * w;
        const int cx = A.img_nn[3 * J], cy = A.img_nn[3 * J + 1], cz = A.img_nn[3 * J + 2];
        const int lox = max(cx - nc, ni - nloc), hix = min(cx + nc, ni + nloc);
        const int loy = max(cy - nc, nj - nloc), hiy = min(cy + nc, nj + nloc);
        const int loz = max(cz - nc, nk - nloc), hiz = min(cz + nc, nk + nloc);
        const int bx = hix - lox + 1, area = bx * (hiy - loy + 1);
        const T *U = A.udv + A.udv_off[J];
        double acc[MAXP][3];
#pragma unroll
        for (int i = 0; i < MAXP; ++i) acc[i][0] = acc[i][1] = acc[i][2] = 0.0;
        for (int a = lane; a < area; a += 64) {
            const int iy = a / bx, ix = a - iy * bx;
            const int gx = lox + ix, gy = loy + iy;
            const int x = gx - ni + nloc, y = gy - nj + nloc;
            const int ridx = (gy - cy + nc) * w + (gx - cx + nc);
            for (int gz = loz; gz <= hiz; ++gz) {
                const int z = gz - nk + nloc;
                double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
                for (int s = 1; s <= FDN; ++s) {
                    g0 += (-Dat(x - s, y, z) + Dat(x + s, y, z)) * A.grad[s];
                    g1 += (-Dat(x, y - s, z) + Dat(x, y + s, z)) * A.grad[s];
                    g2 += (-Dat(x, y, z - s) + Dat(x, y, z + s)) * A.grad[s];
                }
                const int q = (gz - cz + nc) * w2 + ridx;
#pragma unroll
                for (int i = 0; i < MAXP; ++i)
                    if (i < np) {
                        const double uq = (double)U[(size_t)i * w3 + q];
                        acc[i][0] += g0 * uq;
                        acc[i][1] += g1 * uq;
                        acc[i][2] += g2 * uq;
                    }
            }
        }
        double f[3] = {0.0, 0.0, 0.0};
        const int qc = (nk - cz + nc) * w2 + (nj - cy + nc) * w + (ni - cx + nc);
#pragma unroll
        for (int i = 0; i < MAXP; ++i)
            if (i < np) {
                const double uc = (double)U[(size_t)i * w3 + qc];
                const double id = A.inv_denom[16 * t + i];
#pragma unroll
                for (int d = 0; d < 3; ++d) f[d] += wave_sum(acc[i][d]) * id * uc;
            }
        if (lane == 0) {
            double *o = A.contrib + ((size_t)b * FSLOT + slot) * 4;
            o[0] = (double)A.img_atom[J];
            o[1] = f[0]; o[2] = f[1]; o[3] = f[2];
        }
    }
}

// ---------------------------------------------------------------- image lists
struct LArgs {
    Geo g;
    int n_img;
    const int *img_type, *img_nn, *nlocc, *nproj;
    int *list, *list_cnt, *overflow;
    int *glist, *glist_cnt;       // [ceil(B/NG)][GLCAP], [ceil(B/NG)]
    // register slots k_projr needs for any group of the batch (a maximum;
    // 99 when the list mixes box sizes) with pr_nw waves walking slabs of
    // pr_sz planes, pr_pu the slab capacity per plane and member
    int *pr_bad;
    int pr_nw, pr_sz, pr_pu;
    int pr_mfma;                  // k_projm's member stride (congruent to 8 modulo 32)
};

__global__ void k_node_images(LArgs A, int B)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int node = A.g.node_list[b];
    const int n = A.g.n, nloc = A.g.nloc;
    const int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    int cnt = 0;
    int *L = A.list + (size_t)b * LCAP;
    for (int J = 0; J < A.n_img; ++J) {
        const int t = A.img_type[J];
        if (A.nproj[t] == 0) continue;
        const int R = nloc + A.nlocc[t];
        const int dx = A.img_nn[3 * J] - ni, dy = A.img_nn[3 * J + 1] - nj, dz = A.img_nn[3 * J + 2] - nk;
        if (dx < -R || dx > R || dy < -R || dy > R || dz < -R || dz > R) continue;
        if (cnt < LCAP) L[cnt] = J;
        ++cnt;
    }
    if (cnt > LCAP) { atomicOr(A.overflow, 1); cnt = LCAP; }
    // stable insertion sort by the image's z (then list order): the waves of
    // the projections then sweep the cube in z and share its planes in cache
    for (int i = 1; i < cnt; ++i) {
        const int J = L[i], kz = A.img_nn[3 * J + 2];
        int j = i - 1;
        while (j >= 0 && A.img_nn[3 * L[j] + 2] > kz) { L[j + 1] = L[j]; --j; }
        L[j + 1] = J;
    }
    A.list_cnt[b] = cnt;
}

// Images meeting any of the cubes of the NG consecutive slots of a group
// (the union of their node lists), sorted by z like the node lists.
__global__ void k_group_images(LArgs A, int B)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (B + NG - 1) / NG) return;
    const int b0 = q * NG, gn = min(NG, B - b0);
    const int n = A.g.n, nloc = A.g.nloc;
    int ni[NG], nj[NG], nk[NG];
#pragma unroll
    for (int m = 0; m < NG; ++m) {
        const int node = A.g.node_list[b0 + min(m, gn - 1)];
        ni[m] = node % n; nj[m] = (node / n) % n; nk[m] = node / (n * n);
    }
    int cnt = 0;
    int *L = A.glist + (size_t)q * GLCAP;
    for (int J = 0; J < A.n_img; ++J) {
        const int t = A.img_type[J];
        if (A.nproj[t] == 0) continue;
        const int R = nloc + A.nlocc[t];
        const int cx = A.img_nn[3 * J], cy = A.img_nn[3 * J + 1], cz = A.img_nn[3 * J + 2];
        bool hit = false;
#pragma unroll
        for (int m = 0; m < NG; ++m)
            hit |= abs(cx - ni[m]) <= R && abs(cy - nj[m]) <= R && abs(cz - nk[m]) <= R;
        if (!hit) continue;
        if (cnt < GLCAP) L[cnt] = J;
        ++cnt;
    }
    if (cnt > GLCAP) { atomicOr(A.overflow, 1); cnt = GLCAP; }
    for (int i = 1; i < cnt; ++i) {
        const int J = L[i], kz = A.img_nn[3 * J + 2];
        int j = i - 1;
        while (j >= 0 && A.img_nn[3 * L[j] + 2] > kz) { L[j + 1] = L[j]; --j; }
        L[j + 1] = J;
    }
    A.glist_cnt[q] = cnt;
    if (A.pr_bad && cnt > 0) {
        // the register slots k_projr needs (written as a maximum into
        // pr_bad; 99 when it does not apply): the k-th image of a wave that
        // meets the walk's union takes slot k mod R, so R must exceed k - j
        // for every earlier one j still alive in the slab where k enters.  Over the union of the
        // group's cubes when k_projr walks them together, else over each
        // member's cube (a walk with fewer members clamps to a smaller
        // z-range and only shortens lives).
        int need = 1;
        const int nc0 = A.nlocc[A.img_type[L[0]]];
        for (int i = 1; i < cnt; ++i)
            if (A.nlocc[A.img_type[L[i]]] != nc0) need = 99;
        const int w = 2 * nc0 + 1, SZ = A.pr_sz;
        int x0 = 1 << 30, x1 = -(1 << 30), y0 = 1 << 30, y1 = -(1 << 30);
        for (int m = 0; m < gn; ++m) {
            x0 = min(x0, ni[m]); x1 = max(x1, ni[m]); y0 = min(y0, nj[m]); y1 = max(y1, nj[m]);
        }
        const int ux = x1 - x0 + 2 * nloc + 1;
        const int upx = ux + (((w - ux) % 32) + 32) % 32;
        int slab = upx * (y1 - y0 + 2 * nloc + 1);
        if (A.pr_mfma) slab += (((8 - slab) % 32) + 32) % 32;
        const bool together = slab <= A.pr_pu;
        for (int walk = 0; walk < (together ? 1 : gn) && need < 99; ++walk) {
            int uz0 = 1 << 30, uz1 = -(1 << 30), ux0 = 1 << 30, ux1 = -(1 << 30), uy0 = 1 << 30, uy1 = -(1 << 30);
            for (int m = 0; m < gn; ++m)
                if (together || m == walk) {
                    uz0 = min(uz0, nk[m] - nloc); uz1 = max(uz1, nk[m] + nloc);
                    ux0 = min(ux0, ni[m] - nloc); ux1 = max(ux1, ni[m] + nloc);
                    uy0 = min(uy0, nj[m] - nloc); uy1 = max(uy1, nj[m] + nloc);
                }
            const int zs0 = (uz0 + 4096) / SZ * SZ - 4096;
            for (int wv = 0; wv < A.pr_nw; ++wv) {
                constexpr int WN = 32;   // live images looked back on
                int kk[WN], ee[WN], nl = 0;
                int k = -1;
                for (int li = wv; li < cnt; li += A.pr_nw) {
                    const int oz = A.img_nn[3 * L[li] + 2] - nc0;
                    const int bx0 = A.img_nn[3 * L[li]] - nc0, by0 = A.img_nn[3 * L[li] + 1] - nc0;
                    if (oz + w - 1 < uz0 || oz > uz1 || bx0 + w - 1 < ux0 || bx0 > ux1 || by0 + w - 1 < uy0 ||
                        by0 > uy1)
                        continue;   // no slot
                    ++k;
                    const int S = (max(oz, uz0) - zs0) / SZ, E = (min(oz + w - 1, uz1) - zs0) / SZ;
                    // drop images that have left before this one enters
                    int j = 0;
                    for (int i = 0; i < nl; ++i)
                        if (ee[i] >= S) { kk[j] = kk[i]; ee[j] = ee[i]; ++j; }
                    nl = j;
                    if (nl > 0) need = max(need, k - kk[0] + 1);
                    if (nl == WN) { need = 99; break; }
                    kk[nl] = k; ee[nl] = E; ++nl;
                }
            }
        }
        atomicMax(A.pr_bad, need);
    }
}

// ---------------------------------------------------------------- fills
template <typename T>
__global__ void k_fill(T *p, size_t cnt, T v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
__global__ void k_fill_delta(T *p, int M3, int B, int c)
{
    const size_t tot = (size_t)M3 * B;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x)
        p[i] = ((int)(i % M3) == c) ? (T)1 : (T)0;
}

// ---------------------------------------------------------------- Lanczos scalars
// Per node after each Lanczos pass: alpha_p = v.w, beta_p = |r' - alpha_p v|,
// then (p >= 1) the QL extreme eigenvalues of the (p+1)-sized tridiagonal
// (sq.cpp:1025-1069, TridiagEigenSolve sq.cpp:1094-1180).
struct QArgs {
    int B, pass;
    double tol;
    const double *lred;       // [B][2 halves][4]
    double *aa, *bb;          // [MAXIT+2][B] (pass-major: the lanes of a wave read one row)
    double *d, *e;            // scratch [B][MAXIT+2]
    double *lmin, *lmax, *lmin_prev, *lmax_prev;
    double *lchg;             // [B][2] last change of the largest / smallest Ritz value
    int *count, *done;
    double *ca, *ib, *bprev;  // next pass scalars
    double *chi, *zee;        // outputs at convergence
    int *err;
};

__device__ int ql_extremes(double *d, double *e, int n, double *lmin, double *lmax)
{
    e[n - 1] = 0.0;
    int m, i;
    double s, r, p, g, f, dd, c, bq;
    for (int l = 0; l <= n - 1; l++) {
        int iter = 0;
        do {
            for (m = l; m <= n - 2; m++) {
                dd = fabs(d[m]) + fabs(d[m + 1]);
                if ((double)(fabs(e[m]) + dd) == dd) break;
            }
            if (m != l) {
                if (iter++ == 200) return -1;
                g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                r = sqrt(g * g + 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                s = c = 1.0;
                p = 0.0;
                for (i = m - 1; i >= l; i--) {
                    f = s * e[i];
                    bq = c * e[i];
                    e[i + 1] = (r = sqrt(g * g + f * f));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * bq;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - bq;
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    *lmin = d[0];
    *lmax = d[0];
    for (i = 1; i < n; i++) {
        if (d[i] > *lmax) *lmax = d[i];
        else if (d[i] < *lmin) *lmin = d[i];
    }
    return 0;
}

// Largest and smallest eigenvalue of the symmetric tridiagonal (d[0..n-1],
// e[0..n-2], strided by st elements) by Newton's method on det(xI - T).
// With q_i = p_i/p_{i-1} (Sturm ratios) and r_i = q_i'/q_i, p'/p = sum_i
// q_i'/q_i; past an extreme root the characteristic polynomial is monotone
// and convex (concave on the left), so iterates started outside the spectrum
// move monotonically onto the root, and each stops at the first step that no
// longer moves inward.  A start xh is outside (above) the spectrum iff every
// q_i(xh) > 0 (xh I - T positive definite), xl below iff every q_i(xl) < 0.
// Each side first tries the near start given (just past the previous pass's
// extreme, which the new one can only exceed), widening the gap 16-fold per
// failed sweep, and falls back to the far start (a Weyl bound, always
// outside).  Both sides share every sweep over the rows, which are read in
// chunks of CH with the next chunk's loads in flight.  Returns false when
// either needs more than 200 sweeps.
__device__ bool tridiag_extremes(const double *d, const double *e, size_t st, int n, double fh, double fl,
                                 double bh, double bl, double gh, double gl, double *hi, double *lo)
{
#pragma clang fp contract(off)
    constexpr int CH = 8;
    enum { PROBE = 0, NEWTON = 1, DONE = 2 };
    // each side starts just past the previous extreme (bh + gh, bl - gl) when
    // that lies inside the far start, else at the far start
    int mh = bh + gh < fh ? PROBE : NEWTON, ml = bl - gl > fl ? PROBE : NEWTON;
    double xh = mh == PROBE ? bh + gh : fh, xl = ml == PROBE ? bl - gl : fl;
    for (int it = 0; it < 200; ++it) {
        double qh = xh - d[0], ql = xl - d[0], dqh = 1.0, dql = 1.0;
        bool ph = qh > 0.0, pl = ql < 0.0;
        double iqh = 1.0 / qh, iql = 1.0 / ql;
        double sh = dqh * iqh, sl = dql * iql;
        double dn[CH], en[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int i = 1 + j;
            dn[j] = i < n ? d[(size_t)i * st] : 0.0;
            en[j] = i < n ? e[(size_t)(i - 1) * st] : 0.0;
        }
        for (int i0 = 1; i0 < n; i0 += CH) {
            double dc[CH], ec[CH];
#pragma unroll
            for (int j = 0; j < CH; ++j) { dc[j] = dn[j]; ec[j] = en[j]; }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int i = i0 + CH + j;
                dn[j] = i < n ? d[(size_t)i * st] : 0.0;
                en[j] = i < n ? e[(size_t)(i - 1) * st] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                if (i0 + j < n) {
                    const double em2 = ec[j] * ec[j];
                    const double rh = em2 * iqh, rl = em2 * iql;
                    dqh = 1.0 + rh * dqh * iqh;
                    dql = 1.0 + rl * dql * iql;
                    qh = (xh - dc[j]) - rh;
                    ql = (xl - dc[j]) - rl;
                    ph = ph && qh > 0.0;
                    pl = pl && ql < 0.0;
                    iqh = 1.0 / qh;
                    iql = 1.0 / ql;
                    sh += dqh * iqh;
                    sl += dql * iql;
                }
            }
        }
        // upper side
        bool step_h = mh == NEWTON;
        if (mh == PROBE) {
            if (ph) {
                mh = NEWTON;          // outside: Newton from this sweep on
                step_h = true;
            } else {
                gh *= 16.0;           // inside: widen, or take the far start
                xh = bh + gh;
                if (!(xh < fh)) { xh = fh; mh = NEWTON; }
            }
        }
        if (step_h) {
            const double xn = xh - 1.0 / sh;
            if (!(xn < xh)) { *hi = xh; mh = DONE; }
            else xh = xn;
        }
        // lower side
        bool step_l = ml == NEWTON;
        if (ml == PROBE) {
            if (pl) {
                ml = NEWTON;
                step_l = true;
            } else {
                gl *= 16.0;
                xl = bl - gl;
                if (!(xl > fl)) { xl = fl; ml = NEWTON; }
            }
        }
        if (step_l) {
            const double xn = xl - 1.0 / sl;
            if (!(xn > xl)) { *lo = xl; ml = DONE; }
            else xl = xn;
        }
        if (mh == DONE && ml == DONE) return true;
    }
    return false;
}

__global__ void k_lanczos_scalars(QArgs A)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.B || A.done[b]) return;
    const int S = MAXIT + 2;
    // the two row halves of the z-march, added in half order
    const double *r0 = A.lred + 8 * (size_t)b, *r1 = r0 + 4;
    const double r[4] = {r0[0] + r1[0], r0[1] + r1[1], r0[2] + r1[2], r0[3] + r1[3]};
    const int p = A.pass;
    const double al = r[0];
    double b2 = r[1] - 2.0 * al * r[2] + al * al * r[3];
    if (b2 < 0.0) b2 = 0.0;
    const double be = sqrt(b2);
    const size_t st = (size_t)A.B;
    A.aa[(size_t)p * st + b] = al;
    A.bb[(size_t)p * st + b] = be;
    A.ca[b] = al;
    A.ib[b] = 1.0 / be;
    A.bprev[b] = be;
    if (p == 0) return;
    const int nq = p + 1;
    // Extreme eigenvalues of the (p+1)-sized tridiagonal.  The reference runs
    // the full QL (TridiagEigenSolve, sq.cpp:1094-1180) every pass; only the
    // extremes are used, and adding row p moves them by at most |beta_{p-1}|
    // beyond max(lambda_max(T_p), alpha_p) (min for lambda_min), which gives
    // Newton a start outside the spectrum.  QL stays as the fallback.
    const double *ad = A.aa + b, *be2 = A.bb + b;   // row i at [i * st]
    double lo, hi, up, dn;
    if (p == 1) {   // Gershgorin
        const double r = fabs(be2[0]);
        up = fmax(ad[0], ad[st]) + r;
        dn = fmin(ad[0], ad[st]) - r;
    } else {
        const double r = fabs(be2[(size_t)(p - 1) * st]);
        up = fmax(A.lmax_prev[b], ad[(size_t)p * st]) + r;
        dn = fmin(A.lmin_prev[b], ad[(size_t)p * st]) - r;
    }
    up += 8.0 * 2.2e-16 * fabs(up) + 1e-300;
    dn -= 8.0 * 2.2e-16 * fabs(dn) + 1e-300;
    // near starts: past the previous extremes by twice their last change
    // (pass 1 has no previous extremes: the Gershgorin bounds)
    const double bh = p == 1 ? up : A.lmax_prev[b], bl = p == 1 ? dn : A.lmin_prev[b];
    const double gh = 2.0 * A.lchg[2 * (size_t)b] + 1e-13 * (fabs(bh) + 1.0);
    const double gl = 2.0 * A.lchg[2 * (size_t)b + 1] + 1e-13 * (fabs(bl) + 1.0);
    if (!tridiag_extremes(ad, be2, st, nq, up, dn, bh, bl, gh, gl, &hi, &lo)) {
        double *d = A.d + (size_t)b * S, *e = A.e + (size_t)b * S;
        for (int i = 0; i < nq; ++i) { d[i] = ad[(size_t)i * st]; e[i] = be2[(size_t)i * st]; }
        if (ql_extremes(d, e, nq, &lo, &hi)) { atomicOr(A.err, 1); A.done[b] = 1; return; }
    }
    const double dl = fabs(lo - A.lmin_prev[b]), dm = fabs(hi - A.lmax_prev[b]);
    A.lmin_prev[b] = lo;
    A.lmax_prev[b] = hi;
    A.lchg[2 * (size_t)b] = dm;
    A.lchg[2 * (size_t)b + 1] = dl;
    const int count = p;   // the reference's count after this iteration
    A.count[b] = count;
    if (!(count < MAXIT && (dl > A.tol || dm > A.tol))) {
        lo -= A.tol;
        hi += A.tol;
        A.lmin[b] = lo;
        A.lmax[b] = hi;
        A.chi[b] = (hi + lo) / 2;
        A.zee[b] = (hi - lo) / 2;
        A.done[b] = 1;
    }
}

// ---------------------------------------------------------------- moments
// mu[nd][k] from the k-th recurrence step (sq.cpp:604-653)
struct MArgs {
    int B, k, npl;
    double h3;
    const int *node_list;
    const double *center, *dots;
    double *htemp;            // [B] t1[c]
    double *mu;               // [nn][npl+1] (global node index)
};

__global__ void k_moments(MArgs A)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.B) return;
    const int nd = A.node_list[b];
    double *m = A.mu + (size_t)nd * (A.npl + 1);
    const int k = A.k;
    if (k == 1) {
        m[0] = (double)(1.0 / A.h3);
        A.htemp[b] = A.center[b];
    }
    m[k] = (double)(A.center[b] / A.h3);
    if (4 * k > A.npl) {
        const double *d0 = A.dots + 4 * (size_t)b, *d1 = d0 + 2;   // the two row halves
        m[2 * k] = (double)((2 * (d0[0] + d1[0]) - 1) / A.h3);
        m[2 * k - 1] = (double)((2 * (d0[1] + d1[1]) - A.htemp[b]) / A.h3);
    }
}

// ---------------------------------------------------------------- Fermi level
// W[nd][k] = mu_0 + 2 sum_{j>=1} mu_j cos(pi j (k+1/2)/(npl+1)), so that
// sum_j 2 mu_j C_j = (2/(npl+1)) sum_k f(y_k) W[nd][k] for the discrete
// Chebyshev coefficients of sq.cpp:1230-1259.
__global__ void k_cheb_W(const double *mu, const double *costab, int nn, int npl, double *W)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const int K = npl + 1;
    if (i >= (size_t)nn * K) return;
    const size_t nd = i / K;
    const int k = (int)(i - nd * K);
    const double *m = mu + nd * K;
    double s = m[0];
    for (int j = 1; j < K; ++j) s += 2.0 * m[j] * costab[(size_t)j * K + k];
    W[i] = s;
}

__device__ __forceinline__ double fermi_dirac(double t, double lf, double bet)
{
    return 1 / (1 + exp(bet * (t - lf)));
}

// per-node sum_j 2 mu_j C_j for FermiDirac (mode 0) or FermiDirac_derv (mode 1),
// reduced per block into part[blockIdx]
__global__ void k_fermi_sum(const double *W, const double *chi, const double *zee,
                            const double *ynode, int nn, int npl, double lf, double bet0,
                            int mode, double *part)
{
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (nd < nn) {
        const int K = npl + 1;
        const double lfp = (lf - chi[nd]) / zee[nd];
        const double bet = bet0 * zee[nd];
        const double *w = W + (size_t)nd * K;
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            double f = fermi_dirac(ynode[k], lfp, bet);
            if (mode == 1) f = f * (1 - f);
            s += f * w[k];
        }
        v[0] = (2.0 / (npl + 1)) * s;
    }
    __shared__ double red[16];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void k_sum_parts(const double *part, int np, double *out)
{
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < np; i += blockDim.x) v[0] += part[i];
    __shared__ double red[16];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) *out = v[0];
}

// rho, Ebs and Eent per node (sq.cpp:740-788); block partials of Ebs, Eent
__global__ void k_rho_energy(const double *W, const double *chi, const double *zee,
                             const double *ynode, int nn, int npl, double lf, double bet0,
                             double h3, double *rho, double *part, int *neg)
{
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (nd < nn) {
        const int K = npl + 1;
        const double fac = 2.0 / (npl + 1);
        const double lfp = (lf - chi[nd]) / zee[nd];
        const double bet = bet0 * zee[nd];
        const double *w = W + (size_t)nd * K;
        double sr = 0.0, se = 0.0, sn = 0.0;
        for (int k = 0; k < K; ++k) {
            const double y = ynode[k];
            const double f = fermi_dirac(y, lfp, bet);
            double en;
            if (fabs(f) < 0.01 * 1e-12 || fabs(f - 1.0) < 0.01 * 1e-12) en = 0.0;
            else en = (f * log(f) + (1 - f) * log(1 - f));
            sr += f * w[k];
            se += (y * f) * w[k];
            sn += en * w[k];
        }
        const double rho_pc = fac * sr;
        rho[nd] = fabs(rho_pc);
        if (rho_pc < 0 && fabs(rho_pc) > 1e-6) atomicOr(neg, 1);
        v[0] = (zee[nd] * (fac * se) + chi[nd] * rho_pc) * h3;
        v[1] = (1 / bet0) * (fac * sn) * h3;
    }
    __shared__ double red[32];
    block_sum<2>(v, red);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = v[0]; part[2 * blockIdx.x + 1] = v[1]; }
}

// explicit Fermi-Dirac coefficients Ci[nd][j] kept for the forces (sq.cpp:757)
__global__ void k_ci(const double *chi, const double *zee, const double *ynode,
                     const double *costab, int nn, int npl, double lf, double bet0, double *Ci)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const int K = npl + 1;
    if (i >= (size_t)nn * K) return;
    const size_t nd = i / K;
    const int j = (int)(i - nd * K);
    const double lfp = (lf - chi[nd]) / zee[nd];
    const double bet = bet0 * zee[nd];
    double s = 0.0;
    for (int k = 0; k < K; ++k) s = s + fermi_dirac(ynode[k], lfp, bet) * costab[(size_t)j * K + k];
    double c = (2.0 / (npl + 1)) * s;
    if (j == 0) c = c / 2.0;
    Ci[i] = c;
}

}  // namespace sqk
