// sq_gpu.hip -- context, orchestration and C-ABI of the MI355X SQ hot path.
//
// Per SQ density step (ClenshawCurtisSpectralQuadrature, sq.cpp:377-834):
//   for each batch of FD nodes (node-major cubes in HBM):
//     image lists  -> Lanczos passes (projections + nonlocal apply + z-march
//                     + FP64 extremes of the tridiagonal)
//                  -> Chebyshev steps (projections + nonlocal apply + fused
//                     z-march recurrence and moment dots)
//   Fermi level (Brent or Newton on the host, each g(lambda) one device sum)
//   rho, Ebs, Eent, Fermi coefficients.
#include "sqdft_gpu.h"
#include "sqdft_system.h"
#include "sq_kernels.hpp"
#include "sq_adjoint_kernels.hpp"
#include "sq_poisson_kernels.hpp"
#include "sq_local_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

using namespace sqk;

namespace {

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "sqdft_amd: %s failed: %s (%s:%d)\n", #x,         \
                         hipGetErrorString(e_), __FILE__, __LINE__);               \
            return SQG_E_HIP;                                                      \
        }                                                                          \
    } while (0)

// Owning device buffer.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t ensure(size_t b)
    {
        if (b <= bytes && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, b ? b : 8);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

struct EventPair { hipEvent_t a, b; double bytes; };

}  // namespace

struct sqg_ctx {
    sqg_params p{};
    int M = 0, M2 = 0, M3 = 0, E = 0, nn = 0, pitch = 0, plane = 0;
    hipStream_t st = nullptr;
    // geometry
    bool geom = false;
    int n_typ = 0, n_atm = 0, n_img = 0, n_ip = 0, max_nproj = 0;
    std::vector<int> h_nlocc, h_nproj, h_img_type, h_img_atom, h_img_nn, h_ip_base;
    std::vector<long> h_udv_off;
    std::vector<double> h_denom;
    DBuf img_type, img_nn, img_atom, ip_base, nlocc, nproj, udv_off, udv, inv_denom;
    DBuf ell_cnt, ell_u, ell_ip, ell_ip16;   // projector ELL over the extended grid
    DBuf udv64, ell_u64, veff64;             // FP64 copies of udv, ell_u, veff (SQG_MIXED_LANCZOS)
    int ell_k = 0;                 // largest entry count of any point
    DBuf Dm, contrib, ncontrib;   // nonlocal forces
    // batch buffers
    int Bcap = 0;
    DBuf X, Y, X2, Y2, NLv, s, node_list, list, list_cnt, flags;   // X2, Y2: Lanczos ping-pong
    DBuf glist, glist_cnt;         // group image lists (NG slots per group)
    int gl_max = 0;                // longest group list of the current batch
    int projr_R = 0;               // k_projr register slots for the batch (0: k_projs instead)
    DBuf lred, aa, bb, dq, eq, lmin, lmax, lmin_prev, lmax_prev, lchg, cnt, done, ca, ib, bprev;
    DBuf chiB, zeeB, center, dots, htemp;
    // all-node state of the last density step
    DBuf veff, chi_all, zee_all, lmin_all, lmax_all, iters_all, mu_all, W, Ci_all, rho_d, part;
    DBuf costab, ynode, scal;
    // Veff derivatives: vectors, per-node sign / d rho/d lambda / weight,
    // n^3 input and output, per-step tangent records
    DBuf adj, adj_sgn, adj_dl, adj_wt, adj_in, adj_out, adj_rec;
    // AAJ Poisson: iterates, residual, history, rhs, result, block partials
    DBuf pz_phi, pz_new, pz_res, pz_X, pz_F, pz_rhs, pz_out, pz_DX, pz_DF, pz_part;
    // local forces: replica boxes, radial tables, block partials, error flag
    DBuf lf_V, lf_Vt, lf_bJ, lf_bJt, lf_img, lf_R, lf_RV, lf_DV, lf_RU, lf_DU, lf_toff, lf_tsize, lf_Z, lf_part, lf_err;
    bool have_state = false;
    double lambda_f = 0.0;
    unsigned char *scratch_ptr = nullptr;   // this device's copy of `scratch` (sqg_hsub_apply_batch_noalloc)
    // device timing of the hot kernels over the last density call (which:
    // 0 Chebyshev z-march, 1 Lanczos z-march, 2 projections of the Lanczos
    // passes, 3 projections of the Chebyshev steps) and of the whole Lanczos
    // phase of each batch (EV_LANCZOS)
    static constexpr int NEV = 5, EV_LANCZOS = 4;
    std::vector<EventPair> ev[NEV];
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    long last_launches[NEV] = {};
    double last_ms[NEV] = {}, last_bytes[NEV] = {};

    size_t real_size() const { return p.precision == SQG_MIXED ? sizeof(float) : sizeof(double); }
    // SQG_MIXED_LANCZOS: FP32 Lanczos passes, everything else in FP64 (the
    // operator data are kept in both types)
    bool split() const { return p.precision == SQG_MIXED_LANCZOS; }
    ~sqg_ctx()
    {
        for (auto e : pool) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
    hipEvent_t ev_get()
    {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[pool_used++];
    }
};

namespace {

// Per-call buffers of sqg_hsub_apply_batch_noalloc: a file-scope device array
// (each GPU the library runs on holds its own copy until the process ends).
// Host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (sqg_ctx::scratch_ptr), never through its name.
__device__ __align__(256) unsigned char scratch[SQG_NOALLOC_SCRATCH_BYTES];

int grid1(size_t n, int bs) { return (int)std::min<size_t>((n + bs - 1) / bs, 1u << 20); }

Geo make_geo(const sqg_ctx *c, const int *dnodes, const int *ddone)
{
    Geo g{};
    g.n = c->p.n; g.nloc = c->p.nloc; g.M = c->M; g.M2 = c->M2; g.M3 = c->M3; g.E = c->E;
    g.pitch = c->pitch; g.plane = c->plane;
    for (int a = 0; a <= FDN; ++a) g.lap[a] = c->p.coeff_lap[a];
    g.node_list = dnodes;
    g.done = ddone;
    return g;
}

// Threads of one z-march workgroup: one column per thread over the larger
// half of the cube's rows (two workgroups per node cube).
int zmarch_threads(int M) { return (zhalf_rows(M, 0) * M + 63) / 64 * 64; }

// Visiting order of the FD nodes: 4x4x2 bricks, so that workgroups that run
// at the same time (consecutive batch slots) march overlapping cubes and
// share Veff, projector tables and ELL entries in the caches.
std::vector<int> node_order(int n)
{
    const int bx = 4, by = 4, bz = 2;
    std::vector<int> seq;
    seq.reserve((size_t)n * n * n);
    for (int z0 = 0; z0 < n; z0 += bz)
        for (int y0 = 0; y0 < n; y0 += by)
            for (int x0 = 0; x0 < n; x0 += bx)
                for (int z = z0; z < std::min(z0 + bz, n); ++z)
                    for (int y = y0; y < std::min(y0 + by, n); ++y)
                        for (int x = x0; x < std::min(x0 + bx, n); ++x) seq.push_back((z * n + y) * n + x);
    return seq;
}

// Allocates per-batch buffers for up to B nodes.
int ensure_batch(sqg_ctx *c, int B)
{
    if (B <= c->Bcap) return SQG_OK;
    const size_t rs = c->real_size();
    const size_t S = MAXIT + 2;
    HIPCHK(c->X.ensure((size_t)B * c->M3 * rs));
    HIPCHK(c->Y.ensure((size_t)B * c->M3 * rs));
    HIPCHK(c->X2.ensure((size_t)B * c->M3 * rs));
    HIPCHK(c->Y2.ensure((size_t)B * c->M3 * rs));
    HIPCHK(c->NLv.ensure((size_t)B * c->M3 * rs));
    HIPCHK(c->s.ensure((size_t)B * std::max(c->n_ip, 1) * sizeof(double)));
    HIPCHK(c->node_list.ensure((size_t)B * sizeof(int)));
    HIPCHK(c->list.ensure((size_t)B * LCAP * sizeof(int)));
    HIPCHK(c->list_cnt.ensure((size_t)B * sizeof(int)));
    HIPCHK(c->glist.ensure((size_t)((B + NG - 1) / NG) * GLCAP * sizeof(int)));
    HIPCHK(c->glist_cnt.ensure((size_t)((B + NG - 1) / NG) * sizeof(int)));
    HIPCHK(c->flags.ensure(16 * sizeof(int)));
    HIPCHK(c->lred.ensure((size_t)B * ZHALF * 4 * sizeof(double)));
    for (DBuf *d : {&c->aa, &c->bb, &c->dq, &c->eq}) HIPCHK(d->ensure((size_t)B * S * sizeof(double)));
    for (DBuf *d : {&c->lmin, &c->lmax, &c->lmin_prev, &c->lmax_prev, &c->ca, &c->ib, &c->bprev,
                    &c->chiB, &c->zeeB, &c->center, &c->htemp})
        HIPCHK(d->ensure((size_t)B * sizeof(double)));
    HIPCHK(c->dots.ensure((size_t)B * ZHALF * 2 * sizeof(double)));
    HIPCHK(c->lchg.ensure((size_t)B * 2 * sizeof(double)));
    HIPCHK(c->cnt.ensure((size_t)B * sizeof(int)));
    HIPCHK(c->done.ensure((size_t)B * sizeof(int)));
    c->Bcap = B;
    return SQG_OK;
}

int choose_batch(sqg_ctx *c, int want, int nvec = 2)
{
    int B = c->p.batch_nodes > 0 ? c->p.batch_nodes : c->nn;
    B = std::min(B, want);
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) == hipSuccess && freeb > 0) {
        const size_t per = (nvec + 3) * (size_t)c->M3 * c->real_size() + (size_t)std::max(c->n_ip, 1) * 8 +
                           LCAP * 4 + GLCAP * 4 / NG + 4 * (MAXIT + 2) * 8 + 256;
        const size_t cap = (size_t)(0.6 * (double)(freeb + (size_t)c->Bcap * per)) / per;
        if ((size_t)B > cap) B = (int)std::max<size_t>(cap, 1);
    }
    return std::max(B, 1);
}

// Nodes per batch of a derivative pass (adjoint, tangent) whose extra
// vectors, ext bytes per node, live in c->adj on top of the batch buffers the
// density step sized: at most that batch, and what free memory plus adj's
// current allocation (freed before it grows) holds.
int choose_batch_adj(sqg_ctx *c, int want, size_t ext)
{
    int B = choose_batch(c, want);
    if (c->Bcap > 0) B = std::min(B, c->Bcap);
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) == hipSuccess && freeb > 0) {
        const size_t cap = (size_t)(0.8 * (double)(freeb + c->adj.bytes)) / std::max<size_t>(ext, 1);
        if ((size_t)B > cap) B = (int)std::max<size_t>(cap, 1);
    }
    return std::max(B, 1);
}

// The operator data in the type a kernel of type T reads: the FP64 copies
// when an FP64 phase runs in a context that stores them in FP32.
template <typename T>
T *op_data(sqg_ctx *c, DBuf &d, DBuf &d64)
{
    return (sizeof(T) == 8 && c->split()) ? d64.as<T>() : d.as<T>();
}

template <typename T>
ZArgs<T> zargs(sqg_ctx *c, const int *dn, const int *dd)
{
    ZArgs<T> A{};
    A.g = make_geo(c, dn, dd);
    A.nlv = c->NLv.as<T>();
    A.veff = op_data<T>(c, c->veff, c->veff64);
    return A;
}

template <typename T>
PArgs<T> pargs(sqg_ctx *c, const int *dn, const int *dd, const T *V)
{
    PArgs<T> A{};
    A.g = make_geo(c, dn, dd);
    A.V = V;
    A.s = c->s.as<double>();
    A.n_ip = std::max(c->n_ip, 1);
    A.list = c->list.as<int>();
    A.list_cnt = c->list_cnt.as<int>();
    A.glist = c->glist.as<int>();
    A.glist_cnt = c->glist_cnt.as<int>();
    A.img_type = c->img_type.as<int>();
    A.img_nn = c->img_nn.as<int>();
    A.ip_base = c->ip_base.as<int>();
    A.udv_off = c->udv_off.as<long>();
    A.udv = op_data<T>(c, c->udv, c->udv64);
    A.nlocc = c->nlocc.as<int>();
    A.nproj = c->nproj.as<int>();
    A.inv_denom = c->inv_denom.as<double>();
    A.ca = c->ca.as<double>();
    A.ib = c->ib.as<double>();
    return A;
}

// Shape of k_projr (s-channel tables only): members per workgroup, register
// slots per wave and planes per slab; nm = 0 when it does not apply.
constexpr int PROJR_NT = 512, PROJR_RPI = 3;
constexpr int PROJR_SLOTS[3] = {5, 10, 16};   // compiled slot counts (five keep FP64 out of scratch)
constexpr int PROJM_SLOTS = 16;               // k_projm: one double per slot and lane
struct ProjrShape { int nm, R, sz, mfma; };
// ts: bytes per stored value of the phase (4 or 8).
ProjrShape projr_shape(const sqg_ctx *c, size_t ts)
{
    if (c->max_nproj < 1 || c->max_nproj > 4 || c->n_ip == 0) return {0, 0, 0, 0};
    // p channels in FP64: k_projm on the matrix cores.  FP32 tables (mixed
    // precision) stay with k_projs: its VALU sums of FP32 operands ran at
    // 1.18 ms a pass on the 216-atom LiH batches against k_projm's 1.67 ms
    // (FP64 MFMAs fed by conversions), while in FP64 the two tie (2.3 ms).
    if (c->max_nproj > 1 && ts == 4) return {0, 0, 0, 0};
    const int mfma = c->max_nproj > 1;
    int w = 0;
    for (int t = 0; t < c->n_typ; ++t)
        if (c->h_nproj[t] > 0) w = std::max(w, 2 * c->h_nlocc[t] + 1);
    if (mfma ? w > 16 : w * w > 64 * PROJR_RPI) return {0, 0, 0, 0};
    // eight members per workgroup; FP32 slabs of two planes fit the LDS
    return {8, 0, ts == 4 ? 2 : 1, mfma};
}

// k_projr's slab capacity per plane and member: the union plane of a 4x2
// (NM = 8) or 2x2 (NM = 4) block of cubes, or one cube, with rows padded to a
// pitch congruent to the box side modulo 32 (as k_projr pads them).
int projr_pu(const sqg_ctx *c, const ProjrShape &ps)
{
    int wbox = 0;
    for (int t = 0; t < c->n_typ; ++t)
        if (c->h_nproj[t] > 0) wbox = 2 * c->h_nlocc[t] + 1;
    auto upitch = [&](int u) { return u + (((wbox - u) % 32) + 32) % 32; };
    auto slab = [&](int u) { return ps.mfma ? u + (((8 - u) % 32) + 32) % 32 : u; };
    const int ux = ps.nm == 8 ? c->M + 3 : c->M + 1, uy = c->M + 1;
    return std::max(slab(upitch(ux) * uy), slab(upitch(c->M) * c->M));
}

// Image lists of the B nodes in node_list (per node and per group of NG
// slots); sets gl_max and projr_R.  Overflow goes to flags[0].
int node_images(sqg_ctx *c, int B)
{
    LArgs L{};
    L.g = make_geo(c, c->node_list.as<int>(), nullptr);
    L.n_img = c->n_img;
    L.img_type = c->img_type.as<int>();
    L.img_nn = c->img_nn.as<int>();
    L.nlocc = c->nlocc.as<int>();
    L.nproj = c->nproj.as<int>();
    L.list = c->list.as<int>();
    L.list_cnt = c->list_cnt.as<int>();
    L.overflow = c->flags.as<int>();
    L.glist = c->glist.as<int>();
    L.glist_cnt = c->glist_cnt.as<int>();
    // the projection shapes of the phases that will run on these lists (both
    // stored types in SQG_MIXED_LANCZOS): each plans its register slots
    const ProjrShape shapes[2] = {projr_shape(c, c->split() ? 4 : c->real_size()), projr_shape(c, 8)};
    const int nsh = c->split() ? 2 : 1;
    hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(64), 0, c->st, c->flags.as<int>() + 4, (size_t)1, 0);
    hipLaunchKernelGGL(k_node_images, dim3((B + 127) / 128), dim3(128), 0, c->st, L, B);
    const int ng = (B + NG - 1) / NG;
    bool any_nm = false, any_mfma = false;
    for (int i = 0; i < nsh; ++i) {
        const ProjrShape &ps = shapes[i];
        LArgs Li = L;
        if (ps.nm) {
            Li.pr_bad = c->flags.as<int>() + 4;
            Li.pr_nw = PROJR_NT / 64; Li.pr_sz = ps.sz; Li.pr_pu = projr_pu(c, ps); Li.pr_mfma = ps.mfma;
            any_nm = true;
            any_mfma |= ps.mfma != 0;
        }
        // a second launch rebuilds the same group lists and adds its slot plan
        if (i == 0 || ps.nm) hipLaunchKernelGGL(k_group_images, dim3((ng + 63) / 64), dim3(64), 0, c->st, Li, B);
    }
    std::vector<int> gc(ng + 1);
    HIPCHK(hipMemcpyAsync(gc.data(), c->glist_cnt.p, (size_t)ng * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(gc.data() + ng, c->flags.as<int>() + 4, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->gl_max = 0;
    for (int i = 0; i < ng; ++i) c->gl_max = std::max(c->gl_max, gc[i]);
    // the smallest slot count that holds every group of the batch: k_projr
    // gives the same sums with any slot count, so batches stay bitwise alike
    c->projr_R = 0;
    if (any_mfma) {
        if (gc[ng] <= PROJM_SLOTS) c->projr_R = PROJM_SLOTS;
    } else if (any_nm) {
        for (int r : PROJR_SLOTS)
            if (gc[ng] <= r) { c->projr_R = r; break; }
    }
    return SQG_OK;
}

// LDS bytes of one k_nlapplyg workgroup (loc table, row offsets, NG values per row)
size_t nlapplyg_lds(const sqg_ctx *c)
{
    const size_t head = ((size_t)(std::max(c->n_ip, 1) + c->gl_max + 1) * 4 + 15) / 16 * 16;
    return head + (size_t)std::max(c->gl_max * c->max_nproj, 1) * SLW * sizeof(double);
}

template <typename T, int K>
void nlapply_go(sqg_ctx *c, int B, const NArgs<T> &A)
{
    const size_t shg = nlapplyg_lds(c);
    if (shg <= 160 * 1024) {
        NGArgs<T> N{};
        N.g = A.g; N.nl = A.nl; N.nlv = A.nlv;
        N.glist = c->glist.as<int>(); N.glist_cnt = c->glist_cnt.as<int>();
        N.img_type = c->img_type.as<int>(); N.nproj = c->nproj.as<int>(); N.ip_base = c->ip_base.as<int>();
        N.B = B;
        auto fn = N.nl.ip16 ? k_nlapplyg<T, true> : k_nlapplyg<T, false>;
        if (shg > 64 * 1024) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shg);
        hipLaunchKernelGGL(fn, dim3((B + NG - 1) / NG), dim3(512), shg, c->st, N);
        return;
    }
    const size_t shm = (size_t)std::max(c->n_ip, 1) * sizeof(double);
    auto fn = k_nlapply<T, K>;
    if (shm > 64 * 1024) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL(fn, dim3(B), dim3(512), shm, c->st, A);
}

template <typename T, int MAXP, int FORMED, int SZ, int NT, int NPU, int NM>
void projs_go(sqg_ctx *c, dim3 grid, size_t shm, const PArgs<T> &P)
{
    auto fn = k_projs<T, MAXP, FORMED, SZ, NT, NPU, NM>;
    if (shm > 64 * 1024) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL(fn, dim3(grid.x * (NG / NM)), dim3(NT), shm, c->st, P);
}

template <typename T, int FORMED, int SZ, int NM>
void projs_np(sqg_ctx *c, dim3 grid, size_t shs, const PArgs<T> &P, int npu)
{
    constexpr int NT = 512;
    if (npu == 1) projs_go<T, 1, FORMED, SZ, NT, 1, NM>(c, grid, shs, P);
    else if (npu == 3) projs_go<T, 4, FORMED, SZ, NT, 3, NM>(c, grid, shs, P);
    else projs_go<T, 4, FORMED, SZ, NT, 4, NM>(c, grid, shs, P);
}

// Staged projections (k_projs, all NG members of a group per workgroup):
// planes per slab (sz) and the largest union plane (pu) staged at once.
// Returns false when not even a single cube plane fits in LDS.
template <typename T, int FORMED>
bool launch_projs(sqg_ctx *c, dim3 grid, PArgs<T> &P, int mp)
{
    constexpr int D = sizeof(T) == 4 ? 4 : 2;   // deep slab (planes)
    const size_t sa = ((size_t)std::max(c->gl_max, 1) * sizeof(ImgMeta) + 15) / 16 * 16;
    // projector values are loaded and summed for npu of the mp slots
    const int npu = mp == 1 ? 1 : c->max_nproj == 3 ? 3 : 4;
    for (int pu : {(c->M + 3) * (c->M + 3), c->M2}) {
        for (int sz : {D, D / 2}) {
            const size_t shs = sa + (size_t)pu * NG * sizeof(T) * sz;
            if (shs > 160 * 1024) continue;
            P.pu = pu;
            if (sz == 4) projs_np<T, FORMED, 4, NG>(c, grid, shs, P, npu);
            else if (sz == 2) projs_np<T, FORMED, 2, NG>(c, grid, shs, P, npu);
            else projs_np<T, FORMED, 1, NG>(c, grid, shs, P, npu);
            return true;
        }
    }
    return false;
}

template <typename T, int FORMED, int SZ, int NM, int R>
void projr_go(sqg_ctx *c, dim3 grid, size_t shm, const PArgs<T> &P)
{
    auto fn = k_projr<T, FORMED, SZ, PROJR_NT, NM, R, PROJR_RPI>;
    if (shm > 64 * 1024) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL(fn, dim3(grid.x * (NG / NM)), dim3(PROJR_NT), shm, c->st, P);
}

template <typename T, int FORMED>
bool launch_projr(sqg_ctx *c, dim3 grid, PArgs<T> &P)
{
    const ProjrShape ps = projr_shape(c, sizeof(T));
    const size_t sa = ((size_t)std::max(c->gl_max, 1) * sizeof(ImgMeta) + 15) / 16 * 16;
    const int pu = projr_pu(c, ps);
    const size_t shm = sa + (size_t)pu * ps.nm * sizeof(T) * ps.sz;
    if (shm > 160 * 1024) return false;
    P.pu = pu;
    constexpr int SZ = sizeof(T) == 8 ? 1 : 2;
    if (ps.sz != SZ) return false;
    if constexpr (sizeof(T) == 8) {
        if (ps.mfma) {
            auto fn = k_projm<T, FORMED, SZ, PROJR_NT, PROJM_SLOTS>;
            if (shm > 64 * 1024) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            hipLaunchKernelGGL(fn, grid, dim3(PROJR_NT), shm, c->st, P);
            return true;
        }
    }
    if (ps.mfma) return false;
    if (c->projr_R == 5) projr_go<T, FORMED, SZ, 8, 5>(c, grid, shm, P);
    else if (c->projr_R == 10) projr_go<T, FORMED, SZ, 8, 10>(c, grid, shm, P);
    else projr_go<T, FORMED, SZ, 8, 16>(c, grid, shm, P);
    return true;
}

// Projections s = <U, v>/Denom of the B vectors (k_proj), then their
// nonlocal term NL = sum U s into NLv (k_nlapply) for the next z-march.
template <typename T, int FORMED>
void launch_proj(sqg_ctx *c, int B, const PArgs<T> &A, int which = -1, double alg_bytes = 0.0)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (which >= 0 && c->n_ip > 0) {
        e0 = c->ev_get();
        e1 = c->ev_get();
        if (e0) (void)hipEventRecord(e0, c->st);
    }
    if (c->n_ip > 0) {
        PArgs<T> P = A;
        P.B = B;
        const dim3 grid((B + NG - 1) / NG);
        if (c->projr_R && launch_projr<T, FORMED>(c, grid, P)) goto apply;
        // staged projections (k_projs) with the deepest slab whose planes
        // (a union of a 4x4 footprint, else a single cube plane) fit in LDS
        const int mp = c->max_nproj <= 1 ? 1 : c->max_nproj <= 4 ? 4 : 0;
        if (mp && launch_projs<T, FORMED>(c, grid, P, mp)) goto apply;
        if (c->max_nproj <= 1)
            hipLaunchKernelGGL((k_projg<T, 1, FORMED>), grid, dim3(256), 0, c->st, P);
        else if (c->max_nproj <= 4)
            hipLaunchKernelGGL((k_projg<T, 4, FORMED>), grid, dim3(256), 0, c->st, P);
        else if (c->max_nproj <= 8)
            hipLaunchKernelGGL((k_projg<T, 8, FORMED>), grid, dim3(256), 0, c->st, P);
        else
            hipLaunchKernelGGL((k_projg<T, 16, FORMED>), grid, dim3(256), 0, c->st, P);
    }
apply:
    if (e0 && e1) {
        (void)hipEventRecord(e1, c->st);
        c->ev[which].push_back({e0, e1, alg_bytes});
    }
    NArgs<T> N{};
    N.g = A.g;
    N.nl.cnt = c->ell_cnt.as<unsigned char>();
    N.nl.u = op_data<T>(c, c->ell_u, c->ell_u64);
    N.nl.ip = c->ell_ip.as<int>();
    N.nl.ip16 = c->n_ip <= 65535 ? c->ell_ip16.as<unsigned short>() : nullptr;
    N.nl.E3 = (size_t)c->E * c->E * c->E;
    N.nl.k_alloc = std::max(c->ell_k, 1);
    N.nl.s = c->s.as<double>();
    N.nl.n_ip = std::max(c->n_ip, 1);
    N.nlv = c->NLv.as<T>();
    switch (c->ell_k <= 4 ? 4 : c->ell_k <= 8 ? 8 : c->ell_k <= 12 ? 12 : c->ell_k <= 16 ? 16 : 0) {
    case 4: nlapply_go<T, 4>(c, B, N); break;
    case 8: nlapply_go<T, 8>(c, B, N); break;
    case 12: nlapply_go<T, 12>(c, B, N); break;
    case 16: nlapply_go<T, 16>(c, B, N); break;
    default: nlapply_go<T, 0>(c, B, N); break;
    }
}

// LDS bytes of one z-march workgroup: two staged planes and the reduction scratch
size_t zmarch_lds(const sqg_ctx *c)
{
    return 2 * (size_t)c->plane * c->real_size() + 16 * 4 * sizeof(double);
}

template <typename T, int MODE, int LB>
void zmarch_k(sqg_ctx *c, int B, const ZArgs<T> &A, size_t shm)
{
    void (*fn)(ZArgs<T>);
    if constexpr (LB <= 640) fn = k_zmarch<T, MODE>;
    else fn = k_zmarch_big<T, MODE>;
    if (shm > 64 * 1024) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL(fn, dim3(B * ZHALF), dim3(zmarch_threads(c->M)), shm, c->st, A);
}

void finish_zmarch(sqg_ctx *c, int which, hipEvent_t e0, hipEvent_t e1, double alg_bytes)
{
    if (which >= 0 && e0 && e1) {
        (void)hipEventRecord(e1, c->st);
        c->ev[which].push_back({e0, e1, alg_bytes});
    }
}

template <typename T, int MODE>
void launch_zmarch(sqg_ctx *c, int B, const ZArgs<T> &A, int which, double alg_bytes)
{
    const size_t shm = zmarch_lds(c);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (which >= 0) {
        e0 = c->ev_get();
        e1 = c->ev_get();
        if (e0) (void)hipEventRecord(e0, c->st);
    }
    // FP32 workgroups of up to 640 threads (M <= 35) are held to 80 VGPRs so
    // that two share a CU; FP64 ones (and the FP32 force and derivative
    // passes, with their extra vectors) would spill there and run one per CU
    if constexpr (sizeof(T) == 4 && MODE != MODE_CHEBD && MODE != MODE_TAN && MODE != MODE_ADJ) {
        if (zmarch_threads(c->M) <= 640) {
            zmarch_k<T, MODE, 640>(c, B, A, shm);
            return finish_zmarch(c, which, e0, e1, alg_bytes);
        }
    }
    zmarch_k<T, MODE, 1024>(c, B, A, shm);
    finish_zmarch(c, which, e0, e1, alg_bytes);
}

// Lanczos bounds of the B nodes in node_list (image lists already built).
// Results land in chiB/zeeB/lmin/lmax/cnt.
template <typename T>
int lanczos_batch(sqg_ctx *c, int B, long *passes_out)
{
    const int *dn = c->node_list.as<int>();
    int *flags = c->flags.as<int>();
    const size_t vlen = (size_t)B * c->M3;
    T *X = c->X.as<T>(), *Y = c->Y.as<T>();
    const double rs = (double)sizeof(T);
    hipEvent_t ph0 = c->ev_get(), ph1 = c->ev_get();
    if (ph0) (void)hipEventRecord(ph0, c->st);
    // ---------------- Lanczos (sq.cpp:965-1092)
    hipLaunchKernelGGL((k_fill<T>), dim3(grid1(vlen, 256)), dim3(256), 0, c->st, X, vlen, (T)1);
    hipLaunchKernelGGL((k_fill<T>), dim3(grid1(vlen, 256)), dim3(256), 0, c->st, Y, vlen, (T)0);
    HIPCHK(hipMemsetAsync(c->s.p, 0, (size_t)B * std::max(c->n_ip, 1) * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(c->done.p, 0, (size_t)B * sizeof(int), c->st));
    HIPCHK(hipMemsetAsync(c->cnt.p, 0, (size_t)B * sizeof(int), c->st));
    HIPCHK(hipMemsetAsync(c->lmin_prev.p, 0, (size_t)B * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(c->lmax_prev.p, 0, (size_t)B * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(c->lchg.p, 0, (size_t)B * 2 * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(c->ca.p, 0, (size_t)B * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(c->bprev.p, 0, (size_t)B * sizeof(double), c->st));
    {
        const double ib0 = 1.0 / std::sqrt((double)c->M3);   // ones/|ones| (sq.cpp:988-998)
        hipLaunchKernelGGL((k_fill<double>), dim3(grid1(B, 256)), dim3(256), 0, c->st, c->ib.as<double>(),
                           (size_t)B, ib0);
    }
    int *ddone = c->done.as<int>();
    QArgs Q{};
    Q.B = B; Q.tol = c->p.lanczos_tol;
    Q.lred = c->lred.as<double>(); Q.aa = c->aa.as<double>(); Q.bb = c->bb.as<double>();
    Q.d = c->dq.as<double>(); Q.e = c->eq.as<double>();
    Q.lmin = c->lmin.as<double>(); Q.lmax = c->lmax.as<double>();
    Q.lmin_prev = c->lmin_prev.as<double>(); Q.lmax_prev = c->lmax_prev.as<double>();
    Q.lchg = c->lchg.as<double>();
    Q.count = c->cnt.as<int>(); Q.done = ddone;
    Q.ca = c->ca.as<double>(); Q.ib = c->ib.as<double>(); Q.bprev = c->bprev.as<double>();
    Q.chi = c->chiB.as<double>(); Q.zee = c->zeeB.as<double>();
    Q.err = flags + 1;
    std::vector<int> hdone(B);
    int active = B;
    long passes = 0;
    T *X2 = c->X2.as<T>(), *Y2 = c->Y2.as<T>();
    for (int pass = 0; pass <= MAXIT && active > 0; ++pass) {
        // (r', v_{k-1}) in one pair of buffers, (r'_{k+1}, v_k) out into the other
        T *Xi = (pass & 1) ? X2 : X, *Yi = (pass & 1) ? Y2 : Y;
        T *Xo = (pass & 1) ? X : X2, *Yo = (pass & 1) ? Y : Y2;
        PArgs<T> P = pargs<T>(c, dn, ddone, Xi);
        P.V2 = Yi;
        launch_proj<T, 1>(c, B, P, 2, 2.0 * active * c->M3 * rs);   // reads r' and v_{k-1}
        ZArgs<T> Z = zargs<T>(c, dn, ddone);
        Z.X = Xi; Z.Y = Yi; Z.XO = Xo; Z.YO = Yo;
        Z.ca = c->ca.as<double>(); Z.ib = c->ib.as<double>(); Z.bprev = c->bprev.as<double>();
        Z.lred = c->lred.as<double>();
        // reads r' and v_{k-1}, writes v_k and r'_{k+1}
        launch_zmarch<T, MODE_LANCZOS>(c, B, Z, 1, 4.0 * active * c->M3 * rs);
        passes += active;
        Q.pass = pass;
        hipLaunchKernelGGL(k_lanczos_scalars, dim3((B + 63) / 64), dim3(64), 0, c->st, Q);
        if (pass >= 1) {
            HIPCHK(hipMemcpyAsync(hdone.data(), ddone, (size_t)B * sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            active = 0;
            for (int i = 0; i < B; ++i) active += hdone[i] == 0;
        }
    }
    if (passes_out) *passes_out += passes;
    if (ph0 && ph1) {
        (void)hipEventRecord(ph1, c->st);
        c->ev[sqg_ctx::EV_LANCZOS].push_back({ph0, ph1, 0.0});
    }
    return SQG_OK;
}

// Lanczos + Chebyshev moments for B nodes already in node_list.
// Results land in chiB/zeeB/lmin/lmax/cnt (per batch) and mu_all (global).
// Without lanczos the bounds of the last density step are reused.
// TL: the stored type of the Lanczos passes, T: of the Chebyshev steps.
template <typename TL, typename T>
int run_batch(sqg_ctx *c, int B, long *passes_out, bool lanczos)
{
    const int *dn = c->node_list.as<int>();
    int *flags = c->flags.as<int>();
    HIPCHK(hipMemsetAsync(flags, 0, 16 * sizeof(int), c->st));
    // image lists
    {
        const int st = node_images(c, B);
        if (st) return st;
    }
    const size_t vlen = (size_t)B * c->M3;
    T *X = c->X.as<T>(), *Y = c->Y.as<T>();
    const double rs = (double)sizeof(T);
    if (lanczos) {
        const int st = lanczos_batch<TL>(c, B, passes_out);
        if (st) return st;
    } else {   // frozen bounds of the last density step
        hipLaunchKernelGGL(k_gather_bounds, dim3((B + 127) / 128), dim3(128), 0, c->st, B, dn,
                           c->chi_all.as<double>(), c->zee_all.as<double>(), c->chiB.as<double>(),
                           c->zeeB.as<double>());
    }
    // ---------------- Chebyshev moments (sq.cpp:564-655)
    const int c_idx = (c->p.nloc * c->M + c->p.nloc) * c->M + c->p.nloc;
    hipLaunchKernelGGL((k_fill_delta<T>), dim3(grid1(vlen, 256)), dim3(256), 0, c->st, X, c->M3, B, c_idx);
    // t_j lives in bufs[j % 3]: step k reads t_{k-1} and t_{k-2} and writes
    // t_k into the third buffer (no vector is overwritten while it is read)
    T *bufs[3] = {X, Y, c->X2.as<T>()};
    MArgs Mg{};
    Mg.B = B; Mg.npl = c->p.npl; Mg.h3 = std::pow(c->p.h, 3);
    Mg.node_list = dn; Mg.center = c->center.as<double>(); Mg.dots = c->dots.as<double>();
    Mg.htemp = c->htemp.as<double>(); Mg.mu = c->mu_all.as<double>();
    for (int k = 1; k <= c->p.npl / 2; ++k) {
        T *in = bufs[(k - 1) % 3], *prev = bufs[(k + 1) % 3], *out = bufs[k % 3];
        PArgs<T> P = pargs<T>(c, dn, nullptr, in);
        launch_proj<T, 0>(c, B, P, 3, (double)B * c->M3 * rs);   // reads t_k
        ZArgs<T> Z = zargs<T>(c, dn, nullptr);
        Z.X = in; Z.Y = prev; Z.O = out;
        Z.chi = c->chiB.as<double>(); Z.zee = c->zeeB.as<double>();
        Z.first = k == 1;
        Z.center = c->center.as<double>(); Z.dots = c->dots.as<double>();
        launch_zmarch<T, MODE_CHEB>(c, B, Z, 0, 3.0 * B * c->M3 * rs);
        Mg.k = k;
        hipLaunchKernelGGL(k_moments, dim3((B + 127) / 128), dim3(128), 0, c->st, Mg);
    }
    int hflags[4];
    HIPCHK(hipMemcpyAsync(hflags, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (hflags[0] || hflags[3]) return SQG_E_OVERFLOW;
    if (hflags[1]) return SQG_E_QL;
    return SQG_OK;
}

__global__ void k_scatter_batch(int B, const int *node_list, const double *chiB, const double *zeeB,
                                const double *lmn, const double *lmx, const int *cnt,
                                double *chi, double *zee, double *lmin, double *lmax, int *iters)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int nd = node_list[b];
    chi[nd] = chiB[b]; zee[nd] = zeeB[b]; lmin[nd] = lmn[b]; lmax[nd] = lmx[b]; iters[nd] = cnt[b];
}

// Per-node results of a node list, gathered from the all-node state
// (iteration counts as doubles, exact).
__global__ void k_gather_nodes(int nb, int K, const int *nodes, const double *chi, const double *zee,
                               const int *iters, const double *mu, double *ochi, double *ozee, double *oit,
                               double *omu)
{
    const size_t tot = (size_t)nb * K;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / K), k = (int)(i - (size_t)b * K);
        const int nd = nodes[b];
        omu[i] = mu[(size_t)nd * K + k];
        if (k == 0) {
            ochi[b] = chi[nd];
            ozee[b] = zee[nd];
            oit[b] = (double)iters[nd];
        }
    }
}

template <typename T>
__global__ void k_convert(const double *in, T *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (T)in[i];
}

template <typename T>
__global__ void k_to_double(const T *in, double *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

int upload_veff(sqg_ctx *c, const double *veff)
{
    const size_t N = (size_t)c->nn;
    DBuf tmp;
    if (c->p.precision == SQG_MIXED || c->split()) {
        HIPCHK(c->veff.ensure(N * sizeof(float)));
        DBuf *dst = &tmp;
        if (c->split()) {
            HIPCHK(c->veff64.ensure(N * sizeof(double)));
            dst = &c->veff64;
        } else {
            HIPCHK(tmp.ensure(N * sizeof(double)));
        }
        HIPCHK(hipMemcpyAsync(dst->p, veff, N * sizeof(double), hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL((k_convert<float>), dim3(grid1(N, 256)), dim3(256), 0, c->st, dst->as<double>(),
                           c->veff.as<float>(), N);
        HIPCHK(hipStreamSynchronize(c->st));
    } else {
        HIPCHK(c->veff.ensure(N * sizeof(double)));
        HIPCHK(hipMemcpyAsync(c->veff.p, veff, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    }
    return SQG_OK;
}

int ensure_node_state(sqg_ctx *c)
{
    const size_t N = (size_t)c->nn, K = (size_t)c->p.npl + 1;
    for (DBuf *d : {&c->chi_all, &c->zee_all, &c->lmin_all, &c->lmax_all, &c->rho_d})
        HIPCHK(d->ensure(N * sizeof(double)));
    HIPCHK(c->iters_all.ensure(N * sizeof(int)));
    HIPCHK(c->mu_all.ensure(N * K * sizeof(double)));
    return SQG_OK;
}

template <typename TL, typename T>
int moments_for_nodes(sqg_ctx *c, const int *nodes, int count, long *passes, int batch_cap, bool lanczos)
{
    int done_n = 0;
    while (done_n < count) {
        int B = std::min(batch_cap, count - done_n);
        int st = ensure_batch(c, B);
        if (st) return st;
        HIPCHK(hipMemcpyAsync(c->node_list.p, nodes + done_n, (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->st));
        st = run_batch<TL, T>(c, B, passes, lanczos);
        if (st) return st;
        if (lanczos)
            hipLaunchKernelGGL(k_scatter_batch, dim3((B + 127) / 128), dim3(128), 0, c->st, B,
                               c->node_list.as<int>(), c->chiB.as<double>(), c->zeeB.as<double>(),
                               c->lmin.as<double>(), c->lmax.as<double>(), c->cnt.as<int>(),
                               c->chi_all.as<double>(), c->zee_all.as<double>(), c->lmin_all.as<double>(),
                               c->lmax_all.as<double>(), c->iters_all.as<int>());
        done_n += B;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return SQG_OK;
}

int moments_dispatch(sqg_ctx *c, const int *nodes, int count, long *passes, int cap = 0, bool lanczos = true)
{
    if (cap <= 0) cap = choose_batch(c, count);
    if (c->p.precision == SQG_MIXED) return moments_for_nodes<float, float>(c, nodes, count, passes, cap, lanczos);
    if (c->split()) return moments_for_nodes<float, double>(c, nodes, count, passes, cap, lanczos);
    return moments_for_nodes<double, double>(c, nodes, count, passes, cap, lanczos);
}

// electron-count constraint g(lambda) (sq.cpp:1183-1228); mode 1 gives d/dlambda / bet
int fermi_eval(sqg_ctx *c, double lf, int mode, double *out)
{
    const int nn = c->nn, bs = 256, nb = (nn + bs - 1) / bs;
    HIPCHK(c->part.ensure((size_t)2 * nb * sizeof(double)));
    hipLaunchKernelGGL(k_fermi_sum, dim3(nb), dim3(bs), 0, c->st, c->W.as<double>(), c->chi_all.as<double>(),
                       c->zee_all.as<double>(), c->ynode.as<double>(), nn, c->p.npl, lf, c->p.bet, mode,
                       c->part.as<double>());
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->st, c->part.as<double>(), nb, c->scal.as<double>());
    HIPCHK(hipMemcpyAsync(out, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return SQG_OK;
}

int ne_constraint(sqg_ctx *c, double lf, double ncharge, double *g)
{
    double s;
    int st = fermi_eval(c, lf, 0, &s);
    if (st) return st;
    s = s * std::pow(c->p.h, 3);
    *g = s - (-ncharge);
    return SQG_OK;
}

double sign_of(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Brent's root finder with the reference's exact control flow (sq.cpp:1289-1404)
int brent(sqg_ctx *c, double x1, double x2, int scf_iter, double lf_prev, double ncharge,
          double *root, int *iters)
{
    const double EPS = 2.5e-16, tol = c->p.fermi_tol;
    const int ITMAX = 100;
    double a = x1, b = x2, cc = x2, d = 0.0, e = 0.0, fa, fb, fc, p, q, r, s, tol1, xm;
    if (scf_iter > 1) { a = lf_prev - 2; b = lf_prev + 2; }
    int st = ne_constraint(c, a, ncharge, &fa);
    if (!st) st = ne_constraint(c, b, ncharge, &fb);
    if (st) return st;
    if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0)) return SQG_E_BRACKET;
    fc = fb;
    for (int iter = 1; iter <= ITMAX; iter++) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { cc = a; fc = fa; e = d = b - a; }
        if (std::fabs(fc) < std::fabs(fb)) { a = b; b = cc; cc = a; fa = fb; fb = fc; fc = fa; }
        tol1 = 2.0 * EPS * std::fabs(b) + 0.5 * tol;
        xm = 0.5 * (cc - b);
        if (std::fabs(xm) <= tol1 || fb == 0.0) { *root = b; *iters = iter; return SQG_OK; }
        if (std::fabs(e) >= tol1 && std::fabs(fa) > std::fabs(fb)) {
            s = fb / fa;
            if (a == cc) { p = 2.0 * xm * s; q = 1.0 - s; }
            else {
                q = fa / fc; r = fb / fc;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = std::fabs(p);
            const double min1 = 3.0 * xm * q - std::fabs(tol1 * q);
            const double min2 = std::fabs(e * q);
            if (2.0 * p < (min1 < min2 ? min1 : min2)) { e = d; d = p / q; }
            else { d = xm; e = d; }
        } else { d = xm; e = d; }
        a = b; fa = fb;
        if (std::fabs(d) > tol1) b += d;
        else b += sign_of(tol1, xm);
        st = ne_constraint(c, b, ncharge, &fb);
        if (st) return st;
    }
    return SQG_E_BRENT_ITMAX;
}

// Newton-Raphson on the electron count (sq.cpp:1407-1490)
int newton(sqg_ctx *c, double lf0, double ncharge, double *root, int *iters, int *warn)
{
    const double tol = c->p.fermi_tol, h3 = std::pow(c->p.h, 3);
    double lnew = 0.0, lold = lf0, err = 1.0;
    int count = 0;
    const int ITMAX = 200;
    while (err > tol && count < ITMAX) {
        double g, dg;
        int st = fermi_eval(c, lold, 0, &g);
        if (!st) st = fermi_eval(c, lold, 1, &dg);
        if (st) return st;
        dg = dg * c->p.bet;
        dg = dg * h3;
        g = g * h3;
        const double fa = g - (-ncharge);
        lnew = lold - (fa / dg);
        err = std::fabs(lnew - lold);
        lold = lnew;
        count += 1;
    }
    *warn = count == ITMAX;
    *root = lnew;
    *iters = count;
    return SQG_OK;
}

// Density-matrix columns of B nodes (npl Chebyshev steps with the D
// accumulator, forces.cpp:376-460) and their contraction with the nonlocal
// projectors of the images that hold each node centre.  Appends
// (node, atom, fx, fy, fz) records in node order.
template <typename T>
int forces_batch(sqg_ctx *c, int B, const int *hnodes, std::vector<double> &rec)
{
    const int *dn = c->node_list.as<int>();
    int *flags = c->flags.as<int>();
    HIPCHK(hipMemsetAsync(flags, 0, 16 * sizeof(int), c->st));
    {
        const int st = node_images(c, B);
        if (st) return st;
    }
    const size_t vlen = (size_t)B * c->M3;
    HIPCHK(c->Dm.ensure(vlen * sizeof(T)));
    HIPCHK(c->contrib.ensure((size_t)B * FSLOT * 4 * sizeof(double)));
    HIPCHK(c->ncontrib.ensure((size_t)B * sizeof(int)));
    T *X = c->X.as<T>(), *Y = c->Y.as<T>(), *D = c->Dm.as<T>();
    const int c_idx = (c->p.nloc * c->M + c->p.nloc) * c->M + c->p.nloc;
    hipLaunchKernelGGL((k_fill_delta<T>), dim3(grid1(vlen, 256)), dim3(256), 0, c->st, X, c->M3, B, c_idx);
    // t_j lives in bufs[j % 3]: step k reads t_{k-1} and t_{k-2} and writes
    // t_k into the third buffer (no vector is overwritten while it is read)
    T *bufs[3] = {X, Y, c->X2.as<T>()};
    for (int k = 1; k <= c->p.npl; ++k) {
        T *in = bufs[(k - 1) % 3], *prev = bufs[(k + 1) % 3], *out = bufs[k % 3];
        PArgs<T> P = pargs<T>(c, dn, nullptr, in);
        launch_proj<T, 0>(c, B, P);
        ZArgs<T> Z = zargs<T>(c, dn, nullptr);
        Z.X = in; Z.Y = prev; Z.O = out;
        Z.chi = c->chi_all.as<double>(); Z.zee = c->zee_all.as<double>(); Z.chi_by_node = 1;
        Z.first = k == 1;
        Z.center = c->center.as<double>(); Z.dots = c->dots.as<double>();
        Z.D = D; Z.Ci = c->Ci_all.as<double>(); Z.k = k; Z.npl = c->p.npl;
        Z.inv_h3 = 1.0 / std::pow(c->p.h, 3);
        launch_zmarch<T, MODE_CHEBD>(c, B, Z, -1, 0.0);
    }
    FArgs<T> F{};
    F.g = make_geo(c, dn, nullptr);
    F.D = D;
    F.list = c->list.as<int>(); F.list_cnt = c->list_cnt.as<int>();
    F.img_type = c->img_type.as<int>(); F.img_nn = c->img_nn.as<int>(); F.img_atom = c->img_atom.as<int>();
    F.nlocc = c->nlocc.as<int>(); F.nproj = c->nproj.as<int>();
    F.udv_off = c->udv_off.as<long>(); F.udv = op_data<T>(c, c->udv, c->udv64); F.inv_denom = c->inv_denom.as<double>();
    for (int a = 0; a <= FDN; ++a) F.grad[a] = c->p.coeff_grad[a];
    F.contrib = c->contrib.as<double>(); F.ncontrib = c->ncontrib.as<int>(); F.overflow = flags + 2;
    if (c->max_nproj <= 1) hipLaunchKernelGGL((k_force_contract<T, 1>), dim3(B), dim3(256), 0, c->st, F);
    else if (c->max_nproj <= 4) hipLaunchKernelGGL((k_force_contract<T, 4>), dim3(B), dim3(256), 0, c->st, F);
    else hipLaunchKernelGGL((k_force_contract<T, 16>), dim3(B), dim3(256), 0, c->st, F);
    std::vector<double> hc((size_t)B * FSLOT * 4);
    std::vector<int> hn(B);
    int hf[4];
    HIPCHK(hipMemcpyAsync(hc.data(), c->contrib.p, hc.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(hn.data(), c->ncontrib.p, (size_t)B * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(hf, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (hf[0] || hf[2] || hf[3]) return SQG_E_OVERFLOW;
    for (int b = 0; b < B; ++b)
        for (int s = 0; s < hn[b]; ++s) {
            const double *o = &hc[((size_t)b * FSLOT + s) * 4];
            rec.insert(rec.end(), {(double)hnodes[b], o[0], o[1], o[2], o[3]});
        }
    return SQG_OK;
}

// rho, Ebs, Eent and the Fermi coefficients Ci at lambda_f from W, chi, zee
// (sq.cpp:740-797); the event, when given, is recorded before the final wait
int rho_assembly(sqg_ctx *c, double lf, double *rho, double *Ebs, double *Eent, int *neg, hipEvent_t tend)
{
    const size_t N = (size_t)c->nn, K = (size_t)c->p.npl + 1;
    const int bs = 256, nb = (int)((N + bs - 1) / bs);
    HIPCHK(c->part.ensure((size_t)2 * nb * sizeof(double)));
    HIPCHK(hipMemsetAsync(c->flags.p, 0, sizeof(int), c->st));
    hipLaunchKernelGGL(k_rho_energy, dim3(nb), dim3(bs), 0, c->st, c->W.as<double>(), c->chi_all.as<double>(),
                       c->zee_all.as<double>(), c->ynode.as<double>(), c->nn, c->p.npl, lf, c->p.bet,
                       std::pow(c->p.h, 3), c->rho_d.as<double>(), c->part.as<double>(), c->flags.as<int>());
    HIPCHK(c->Ci_all.ensure(N * K * sizeof(double)));
    hipLaunchKernelGGL(k_ci, dim3(grid1(N * K, 256)), dim3(256), 0, c->st, c->chi_all.as<double>(),
                       c->zee_all.as<double>(), c->ynode.as<double>(), c->costab.as<double>(), c->nn, c->p.npl, lf,
                       c->p.bet, c->Ci_all.as<double>());
    std::vector<double> part(2 * (size_t)nb);
    HIPCHK(hipMemcpyAsync(part.data(), c->part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(neg, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(rho, c->rho_d.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    if (tend) HIPCHK(hipEventRecord(tend, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double e1 = 0.0, e2 = 0.0;
    for (int i = 0; i < nb; ++i) { e1 += part[2 * i]; e2 += part[2 * i + 1]; }
    *Ebs = e1;
    *Eent = e2;
    return SQG_OK;
}

#include "sq_adjoint_host.hpp"
#include "sq_poisson_host.hpp"

// Event bookkeeping of one call: reset before its first launch, summed after
// its stream has synchronised.
void timing_reset(sqg_ctx *c)
{
    for (int w = 0; w < sqg_ctx::NEV; ++w) c->ev[w].clear();
    c->pool_used = 0;
}
double event_sum_ms(sqg_ctx *c, int which, double *bytes, long *launches);
void timing_sum(sqg_ctx *c)
{
    for (int w = 0; w < sqg_ctx::NEV; ++w) c->last_ms[w] = event_sum_ms(c, w, &c->last_bytes[w], &c->last_launches[w]);
}

double event_sum_ms(sqg_ctx *c, int which, double *bytes, long *launches)
{
    double ms = 0.0, by = 0.0;
    for (const EventPair &e : c->ev[which]) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess) ms += t;
        by += e.bytes;
    }
    *bytes = by;
    *launches = (long)c->ev[which].size();
    return ms;
}

}  // namespace

extern "C" {

int sqg_create(const sqg_params *p, sqg_ctx **out)
{
    if (!p || !out) return SQG_E_ARG;
    *out = nullptr;
    if (p->FDn != FDN) return SQG_E_FDN;
    if (p->npl <= 0 || p->npl % 4) return SQG_E_NPL;
    if (p->n < 1 || p->nloc < 0 || p->h <= 0) return SQG_E_ARG;
    if (2 * p->nloc + 1 > MAXM) return SQG_E_CUBE;
    if (p->precision != SQG_FP64 && p->precision != SQG_MIXED && p->precision != SQG_MIXED_LANCZOS)
        return SQG_E_ARG;
    // Veff is read through a buffer resource whose byte range is 32-bit
    if ((size_t)p->n * p->n * p->n * sizeof(double) >= 0x80000000u) return SQG_E_OVERFLOW;
    std::unique_ptr<sqg_ctx> c(new sqg_ctx());
    c->p = *p;
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamCreate(&c->st));
    // this device's address of the allocation-free entry point's scratch array
    {
        void *sp = nullptr;
        HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        c->scratch_ptr = static_cast<unsigned char *>(sp);
    }
    c->M = 2 * p->nloc + 1;
    c->M2 = c->M * c->M;
    c->M3 = c->M2 * c->M;
    c->E = p->n + 2 * p->nloc;
    c->nn = p->n * p->n * p->n;
    // LDS row pitch of the z-march plane: at least M + 12 (six zero columns
    // each side) and congruent to M modulo 32 doubles, so that the two row
    // pieces one wave covers fall on disjoint LDS banks
    c->pitch = c->M + 32;
    c->plane = c->pitch * (zhalf_rows(c->M, 0) + 2 * FDN);   // a half's rows and six beyond each side
    // Chebyshev-Gauss nodes and cosine table with the reference's expressions (sq.cpp:1241-1257)
    {
        const int K = p->npl + 1;
        std::vector<double> y(K), ct((size_t)K * K);
        for (int k = 0; k < K; ++k) y[k] = std::cos(M_PI * ((double)((k - 0.5 + 1) / (p->npl + 1))));
        for (int j = 0; j < K; ++j)
            for (int k = 0; k < K; ++k)
                ct[(size_t)j * K + k] = std::cos((M_PI * (j - 1 + 1)) * ((double)((k - 0.5 + 1) / (p->npl + 1))));
        HIPCHK(c->ynode.ensure(K * sizeof(double)));
        HIPCHK(c->costab.ensure((size_t)K * K * sizeof(double)));
        HIPCHK(hipMemcpy(c->ynode.p, y.data(), K * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->costab.p, ct.data(), (size_t)K * K * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c->scal.ensure(16 * sizeof(double)));
    }
    *out = c.release();
    return SQG_OK;
}

int sqg_destroy(sqg_ctx *ctx)
{
    delete ctx;
    return SQG_OK;
}

int sqg_set_geometry(sqg_ctx *c, int n_typ, const int *nlocc, const int *nproj, const double *denom,
                     int n_atm, int n_img, const int *img_type, const int *img_atom, const int *img_nn,
                     const double *udv)
{
    if (!c || n_typ < 1 || n_typ > 8 || n_img < 0 || !nlocc || !nproj || !denom) return SQG_E_ARG;
    if (n_img > 0 && (!img_type || !img_atom || !img_nn || !udv)) return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    c->geom = false;
    c->have_state = false;
    c->n_typ = n_typ; c->n_atm = n_atm; c->n_img = n_img;
    c->h_nlocc.assign(nlocc, nlocc + n_typ);
    c->h_nproj.assign(nproj, nproj + n_typ);
    c->h_denom.assign(denom, denom + 16 * n_typ);
    c->h_img_type.assign(img_type, img_type + n_img);
    c->h_img_atom.assign(img_atom, img_atom + n_img);
    c->h_img_nn.assign(img_nn, img_nn + 3 * n_img);
    std::vector<double> invd(16 * n_typ, 0.0);
    c->max_nproj = 0;
    for (int t = 0; t < n_typ; ++t) {
        if (nproj[t] < 0 || nproj[t] > 16 || nlocc[t] < 0) return SQG_E_ARG;
        c->max_nproj = std::max(c->max_nproj, nproj[t]);
        for (int i = 0; i < nproj[t]; ++i) {
            if (denom[16 * t + i] == 0) return SQG_E_DENOM;
            invd[16 * t + i] = 1.0 / denom[16 * t + i];
        }
    }
    c->h_udv_off.resize(n_img);
    c->h_ip_base.resize(n_img);
    long off = 0;
    int ip = 0;
    for (int J = 0; J < n_img; ++J) {
        const int t = img_type[J];
        if (t < 0 || t >= n_typ) return SQG_E_ARG;
        const long w = 2 * nlocc[t] + 1;
        c->h_udv_off[J] = off;
        c->h_ip_base[J] = ip;
        off += w * w * w * nproj[t];
        ip += nproj[t];
    }
    c->n_ip = ip;
    // projector ELL over the extended grid [-nloc, n-1+nloc]^3
    const int E = c->E, nloc = c->p.nloc;
    const size_t E3 = (size_t)E * E * E;
    auto for_each_point = [&](auto &&fn) {
        for (int J = 0; J < n_img; ++J) {
            const int t = img_type[J], nc = nlocc[t], np = nproj[t];
            if (np == 0) continue;
            const int w = 2 * nc + 1;
            int lo[3], hi[3];
            bool ok = true;
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::max(img_nn[3 * J + d] - nc, -nloc);
                hi[d] = std::min(img_nn[3 * J + d] + nc, c->p.n - 1 + nloc);
                if (lo[d] > hi[d]) ok = false;
            }
            if (!ok) continue;
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        const size_t g = ((size_t)(z + nloc) * E + (y + nloc)) * E + (x + nloc);
                        const long q = ((long)(z - img_nn[3 * J + 2] + nc) * w + (y - img_nn[3 * J + 1] + nc)) * w +
                                       (x - img_nn[3 * J] + nc);
                        fn(J, np, g, q, (long)w * w * w);
                    }
        }
    };
    std::vector<int> cnt(E3, 0);
    for_each_point([&](int, int np, size_t g, long, long) { cnt[g] += np; });
    int kmax = 0;
    for (size_t g = 0; g < E3; ++g) kmax = std::max(kmax, cnt[g]);
    if (kmax > 255) return SQG_E_OVERFLOW;
    // the kernels read the ELL through buffer resources with a 32-bit byte
    // range (out-of-range loads return zero), so each table stays below 2 GiB
    if ((size_t)std::max(kmax, 1) * E3 * sizeof(double) >= 0x80000000u) return SQG_E_OVERFLOW;
    const int KE = std::max(kmax, 1);
    std::vector<unsigned char> ecnt(E3, 0);
    std::vector<int> eip((size_t)KE * E3, 0);
    std::vector<double> eu((size_t)KE * E3, 0.0);
    for_each_point([&](int J, int np, size_t g, long q, long w3) {
        for (int i = 0; i < np; ++i) {
            const int k = ecnt[g]++;
            eip[(size_t)k * E3 + g] = c->h_ip_base[J] + i;
            eu[(size_t)k * E3 + g] = udv[c->h_udv_off[J] + i * w3 + q];
        }
    });
    c->ell_k = kmax;
    const bool mixed = c->p.precision == SQG_MIXED || c->split();
    auto up = [&](DBuf &d, const void *src, size_t bytes) -> int {
        HIPCHK(d.ensure(bytes));
        if (bytes) HIPCHK(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
        return SQG_OK;
    };
    int st = 0;
    st |= up(c->nlocc, nlocc, n_typ * sizeof(int));
    st |= up(c->nproj, nproj, n_typ * sizeof(int));
    st |= up(c->inv_denom, invd.data(), invd.size() * sizeof(double));
    st |= up(c->img_type, img_type, std::max(n_img, 1) * sizeof(int));
    st |= up(c->img_nn, img_nn, 3 * std::max(n_img, 1) * sizeof(int));
    st |= up(c->img_atom, img_atom, std::max(n_img, 1) * sizeof(int));
    st |= up(c->ip_base, c->h_ip_base.data(), std::max(n_img, 1) * sizeof(int));
    st |= up(c->udv_off, c->h_udv_off.data(), std::max(n_img, 1) * sizeof(long));
    st |= up(c->ell_cnt, ecnt.data(), ecnt.size());
    st |= up(c->ell_ip, eip.data(), eip.size() * sizeof(int));
    if (c->n_ip <= 65535) {   // the grouped apply reads the ids in 16 bits
        std::vector<unsigned short> e16(eip.begin(), eip.end());
        st |= up(c->ell_ip16, e16.data(), e16.size() * sizeof(unsigned short));
    }
    if (mixed) {
        std::vector<float> uf(std::max<long>(off, 1)), ef(eu.size());
        for (long i = 0; i < off; ++i) uf[i] = (float)udv[i];
        for (size_t i = 0; i < eu.size(); ++i) ef[i] = (float)eu[i];
        st |= up(c->udv, uf.data(), uf.size() * sizeof(float));
        st |= up(c->ell_u, ef.data(), ef.size() * sizeof(float));
    }
    if (!mixed || c->split()) {
        DBuf &du = c->split() ? c->udv64 : c->udv, &de = c->split() ? c->ell_u64 : c->ell_u;
        std::vector<double> ud(udv, udv + off);
        if (ud.empty()) ud.push_back(0.0);
        st |= up(du, ud.data(), ud.size() * sizeof(double));
        st |= up(de, eu.data(), eu.size() * sizeof(double));
    }
    if (st) return SQG_E_HIP;
    // the nonlocal apply keeps every projection of its node in LDS (160 KB per CU)
    if ((size_t)std::max(c->n_ip, 1) * sizeof(double) > 160 * 1024) return SQG_E_OVERFLOW;
    c->geom = true;
    return SQG_OK;
}

int sqg_node_moments(sqg_ctx *c, const double *veff, const int *nodes, int nb, double *chi, double *zee,
                     double *mu, int *iters)
{
    if (!c || !veff || !nodes || nb < 0) return SQG_E_ARG;
    if (!c->geom) return SQG_E_NOGEOM;
    for (int i = 0; i < nb; ++i)
        if (nodes[i] < 0 || nodes[i] >= c->nn) return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    int st = upload_veff(c, veff);
    if (!st) st = ensure_node_state(c);
    if (st) return st;
    c->have_state = false;
    timing_reset(c);
    long passes = 0;
    st = moments_dispatch(c, nodes, nb, &passes);
    if (st) return st;
    timing_sum(c);
    if (nb == 0) return SQG_OK;
    // gather the listed nodes' results on the device, then one copy each
    const size_t K = (size_t)c->p.npl + 1;
    DBuf dn, g;
    HIPCHK(dn.ensure((size_t)nb * sizeof(int)));
    HIPCHK(g.ensure((size_t)nb * (K + 3) * sizeof(double)));
    HIPCHK(hipMemcpyAsync(dn.p, nodes, (size_t)nb * sizeof(int), hipMemcpyHostToDevice, c->st));
    double *gd = g.as<double>();
    hipLaunchKernelGGL(k_gather_nodes, dim3(grid1((size_t)nb * K, 256)), dim3(256), 0, c->st, nb, (int)K,
                       dn.as<int>(), c->chi_all.as<double>(), c->zee_all.as<double>(), c->iters_all.as<int>(),
                       c->mu_all.as<double>(), gd, gd + nb, gd + 2 * (size_t)nb, gd + 3 * (size_t)nb);
    std::vector<double> h((size_t)nb * (K + 3));
    HIPCHK(hipMemcpyAsync(h.data(), gd, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int i = 0; i < nb; ++i) {
        if (chi) chi[i] = h[i];
        if (zee) zee[i] = h[nb + i];
        if (iters) iters[i] = (int)h[2 * (size_t)nb + i];
    }
    if (mu) std::memcpy(mu, h.data() + 3 * (size_t)nb, (size_t)nb * K * sizeof(double));
    return SQG_OK;
}

int sqg_density(sqg_ctx *c, const double *veff, int scf_iter, double ncharge, double *lambda_f, double *rho,
                double *ebs, double *eent, sqg_stats *stats)
{
    if (!c || !veff || !lambda_f || !rho) return SQG_E_ARG;
    if (!c->geom) return SQG_E_NOGEOM;
    HIPCHK(hipSetDevice(c->p.device));
    hipEvent_t t0, t1, t2, t3;
    HIPCHK(hipEventCreate(&t0)); HIPCHK(hipEventCreate(&t1)); HIPCHK(hipEventCreate(&t2)); HIPCHK(hipEventCreate(&t3));
    struct EvGuard { hipEvent_t *e; ~EvGuard() { for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]); } };
    hipEvent_t evs[4] = {t0, t1, t2, t3};
    EvGuard guard{evs};
    int st = upload_veff(c, veff);
    if (!st) st = ensure_node_state(c);
    if (st) return st;
    c->have_state = false;
    timing_reset(c);
    HIPCHK(hipEventRecord(t0, c->st));
    const int cap = choose_batch(c, c->nn);
    const std::vector<int> all = node_order(c->p.n);
    long passes = 0;
    st = moments_dispatch(c, all.data(), c->nn, &passes, cap);
    if (st) return st;
    HIPCHK(hipEventRecord(t1, c->st));
    // global spectrum bounds (sq.cpp:676-713)
    const size_t N = (size_t)c->nn, K = (size_t)c->p.npl + 1;
    std::vector<double> lmn(N), lmx(N);
    std::vector<int> its(N);
    HIPCHK(hipMemcpy(lmn.data(), c->lmin_all.p, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lmx.data(), c->lmax_all.p, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(its.data(), c->iters_all.p, N * sizeof(int), hipMemcpyDeviceToHost));
    double lmin = lmn[0], lmax = lmx[0];
    int itmin = its[0], itmax = its[0];
    for (size_t i = 1; i < N; ++i) {
        lmin = std::min(lmin, lmn[i]);
        lmax = std::max(lmax, lmx[i]);
        itmin = std::min(itmin, its[i]);
        itmax = std::max(itmax, its[i]);
    }
    // Fermi level
    HIPCHK(c->W.ensure(N * K * sizeof(double)));
    hipLaunchKernelGGL(k_cheb_W, dim3(grid1(N * K, 256)), dim3(256), 0, c->st, c->mu_all.as<double>(),
                       c->costab.as<double>(), c->nn, c->p.npl, c->W.as<double>());
    double lf = 0.0;
    int fit = 0, warn = 0;
    if (c->p.fermi_alg == 2) {
        const double lf0 = scf_iter <= 1 ? 0.2 : *lambda_f;
        st = newton(c, lf0, ncharge, &lf, &fit, &warn);
    } else {
        st = brent(c, lmin - 1, lmax + 1, scf_iter, *lambda_f, ncharge, &lf, &fit);
    }
    if (st) return st;
    HIPCHK(hipEventRecord(t2, c->st));
    double Ebs = 0.0, Eent = 0.0;
    int neg = 0;
    st = rho_assembly(c, lf, rho, &Ebs, &Eent, &neg, t3);
    if (st) return st;
    *lambda_f = lf;
    c->lambda_f = lf;
    if (ebs) *ebs = Ebs;
    if (eent) *eent = Eent;
    c->have_state = true;
    timing_sum(c);
    if (stats) {
        float a = 0, b = 0, d = 0;
        (void)hipEventElapsedTime(&a, t0, t1);
        (void)hipEventElapsedTime(&b, t1, t2);
        (void)hipEventElapsedTime(&d, t2, t3);
        std::memset(stats, 0, sizeof(*stats));
        stats->t_lanczos = c->last_ms[sqg_ctx::EV_LANCZOS] * 1e-3;
        stats->t_cheb = a * 1e-3 - stats->t_lanczos;
        stats->t_fermi = b * 1e-3;
        stats->t_rho = d * 1e-3;
        stats->t_total = (a + b + d) * 1e-3;
        stats->lanczos_iters_node0 = its[0];
        stats->lanczos_iters_min = itmin;
        stats->lanczos_iters_max = itmax;
        stats->lanczos_passes = passes;
        stats->fermi_iters = fit;
        stats->lambda_min = lmin;
        stats->lambda_max = lmax;
    }
    if (neg) return SQG_E_NEG_RHO;
    if (itmax >= MAXIT) return SQG_W_LANCZOS_MAXIT;
    if (warn) return SQG_W_NEWTON_MAXIT;
    return SQG_OK;
}

int sqg_get_node_data(sqg_ctx *c, double *chi, double *zee, double *mu, double *Ci, int *iters)
{
    if (!c) return SQG_E_ARG;
    if (!c->have_state) return SQG_E_NOSTATE;
    const size_t N = (size_t)c->nn, K = (size_t)c->p.npl + 1;
    if (chi) HIPCHK(hipMemcpy(chi, c->chi_all.p, N * sizeof(double), hipMemcpyDeviceToHost));
    if (zee) HIPCHK(hipMemcpy(zee, c->zee_all.p, N * sizeof(double), hipMemcpyDeviceToHost));
    if (mu) HIPCHK(hipMemcpy(mu, c->mu_all.p, N * K * sizeof(double), hipMemcpyDeviceToHost));
    if (Ci) HIPCHK(hipMemcpy(Ci, c->Ci_all.p, N * K * sizeof(double), hipMemcpyDeviceToHost));
    if (iters) HIPCHK(hipMemcpy(iters, c->iters_all.p, N * sizeof(int), hipMemcpyDeviceToHost));
    return SQG_OK;
}

int sqg_hsub_apply_batch(sqg_ctx *c, const double *veff, const int *nodes, int nb, const double *v, double *hv)
{
    if (!c || !veff || !nodes || nb < 0 || !v || !hv) return SQG_E_ARG;
    if (!c->geom) return SQG_E_NOGEOM;
    for (int i = 0; i < nb; ++i)
        if (nodes[i] < 0 || nodes[i] >= c->nn) return SQG_E_ARG;
    if (nb == 0) return SQG_OK;
    HIPCHK(hipSetDevice(c->p.device));
    int st = upload_veff(c, veff);
    if (!st) st = ensure_batch(c, nb);
    if (st) return st;
    const size_t len = (size_t)nb * c->M3;
    DBuf tmp;
    HIPCHK(tmp.ensure(len * sizeof(double)));
    HIPCHK(hipMemcpy(c->node_list.p, nodes, nb * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(tmp.p, v, len * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(c->flags.p, 0, 16 * sizeof(int), c->st));
    const int *dn = c->node_list.as<int>();
    st = node_images(c, nb);
    if (st) return st;
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        T *X = c->X.as<T>(), *Y = c->Y.as<T>();
        hipLaunchKernelGGL((k_convert<T>), dim3(grid1(len, 256)), dim3(256), 0, c->st, tmp.as<double>(), X, len);
        PArgs<T> P = pargs<T>(c, dn, nullptr, X);
        launch_proj<T, 0>(c, nb, P);
        ZArgs<T> Z = zargs<T>(c, dn, nullptr);
        Z.X = X; Z.Y = Y;
        launch_zmarch<T, MODE_HSUB>(c, nb, Z, -1, 0.0);
        hipLaunchKernelGGL((k_to_double<T>), dim3(grid1(len, 256)), dim3(256), 0, c->st, Y, tmp.as<double>(), len);
        return SQG_OK;
    };
    st = c->p.precision == SQG_MIXED ? run(float()) : run(double());
    if (st) return st;
    int ovf4[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(ovf4, c->flags.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    const int ovf = ovf4[0] | ovf4[2] | ovf4[3];
    HIPCHK(hipMemcpyAsync(hv, tmp.p, len * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return ovf ? SQG_E_OVERFLOW : SQG_OK;
}

// Points the context's per-batch buffers at regions of the scratch array for
// the duration of one call and restores them (the context keeps owning its
// own allocations; nothing here allocates or frees).
struct ScratchView {
    sqg_ctx *c;
    DBuf *bufs[11];
    void *saved_p[11];
    size_t saved_b[11];
    int n = 0;
    size_t off = 0;
    explicit ScratchView(sqg_ctx *c_) : c(c_) {}
    void put(DBuf &d, size_t bytes)
    {
        bufs[n] = &d; saved_p[n] = d.p; saved_b[n] = d.bytes; ++n;
        d.p = c->scratch_ptr + off;
        d.bytes = bytes;
        off += (bytes + 255) / 256 * 256;
    }
    ~ScratchView()
    {
        for (int i = 0; i < n; ++i) { bufs[i]->p = saved_p[i]; bufs[i]->bytes = saved_b[i]; }
    }
};

// Scratch bytes the allocation-free H v of nb nodes needs (per-call buffers
// rounded to 256 bytes each).
size_t noalloc_bytes(const sqg_ctx *c, int nb)
{
    auto r = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t rs = c->real_size(), len = (size_t)nb * c->M3, ng = (size_t)(nb + NG - 1) / NG;
    size_t t = r(nb * sizeof(int)) + r((size_t)nb * LCAP * sizeof(int)) + r(nb * sizeof(int)) +
               r(ng * GLCAP * sizeof(int)) + r(ng * sizeof(int)) + r(16 * sizeof(int)) +
               r((size_t)nb * std::max(c->n_ip, 1) * sizeof(double)) + r(len * rs);
    if (c->p.precision == SQG_MIXED) t += 2 * r(len * rs) + r((size_t)c->nn * rs);
    return t;
}


int sqg_hsub_apply_batch_noalloc(sqg_ctx *c, const double *d_veff, const int *nodes, int nb, const double *d_v,
                                 double *d_hv)
{
    if (!c || !d_veff || !nodes || nb < 0 || !d_v || !d_hv) return SQG_E_ARG;
    if (!c->geom) return SQG_E_NOGEOM;
    if (!c->scratch_ptr) return SQG_E_HIP;
    for (int i = 0; i < nb; ++i)
        if (nodes[i] < 0 || nodes[i] >= c->nn) return SQG_E_ARG;
    if (nb == 0) return SQG_OK;
    if (noalloc_bytes(c, nb) > (size_t)SQG_NOALLOC_SCRATCH_BYTES) return SQG_E_OVERFLOW;
    HIPCHK(hipSetDevice(c->p.device));
    const size_t len = (size_t)nb * c->M3, rs = c->real_size(), ng = (size_t)(nb + NG - 1) / NG;
    ScratchView V(c);
    V.put(c->node_list, nb * sizeof(int));
    V.put(c->list, (size_t)nb * LCAP * sizeof(int));
    V.put(c->list_cnt, nb * sizeof(int));
    V.put(c->glist, ng * GLCAP * sizeof(int));
    V.put(c->glist_cnt, ng * sizeof(int));
    V.put(c->flags, 16 * sizeof(int));
    V.put(c->s, (size_t)nb * std::max(c->n_ip, 1) * sizeof(double));
    V.put(c->NLv, len * rs);
    const bool mixed = c->p.precision == SQG_MIXED;
    if (mixed) {   // FP32 copies of the input, the output and Veff
        V.put(c->X, len * rs);
        V.put(c->Y, len * rs);
        V.put(c->veff, (size_t)c->nn * rs);
    }
    // flags start at zero (a fill kernel, not a memset node)
    hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(64), 0, c->st, c->flags.as<int>(), (size_t)16, 0);
    HIPCHK(hipMemcpyAsync(c->node_list.p, nodes, nb * sizeof(int), hipMemcpyHostToDevice, c->st));
    int st = node_images(c, nb);
    if (st) return st;
    const int *dn = c->node_list.as<int>();
    auto run = [&](auto tag) {
        using T = decltype(tag);
        const T *X;
        T *Y;
        if constexpr (sizeof(T) == 8) {
            X = d_v; Y = d_hv;
        } else {
            hipLaunchKernelGGL((k_convert<T>), dim3(grid1(len, 256)), dim3(256), 0, c->st, d_v, c->X.as<T>(), len);
            hipLaunchKernelGGL((k_convert<T>), dim3(grid1((size_t)c->nn, 256)), dim3(256), 0, c->st, d_veff,
                               c->veff.as<T>(), (size_t)c->nn);
            X = c->X.as<T>(); Y = c->Y.as<T>();
        }
        PArgs<T> P = pargs<T>(c, dn, nullptr, X);
        launch_proj<T, 0>(c, nb, P);
        ZArgs<T> Z = zargs<T>(c, dn, nullptr);
        if constexpr (sizeof(T) == 8) Z.veff = d_veff;
        Z.X = const_cast<T *>(X); Z.Y = Y;
        launch_zmarch<T, MODE_HSUB>(c, nb, Z, -1, 0.0);
        if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL((k_to_double<T>), dim3(grid1(len, 256)), dim3(256), 0, c->st, Y, d_hv, len);
    };
    if (mixed) run(float());
    else run(double());
    int ovf4[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(ovf4, c->flags.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return (ovf4[0] | ovf4[2] | ovf4[3]) ? SQG_E_OVERFLOW : SQG_OK;
}

int sqg_forces_nonlocal(sqg_ctx *c, double *f)
{
    if (!c || !f) return SQG_E_ARG;
    if (!c->have_state) return SQG_E_NOSTATE;
    HIPCHK(hipSetDevice(c->p.device));
    const int na = c->n_atm;
    for (int i = 0; i < 3 * na; ++i) f[i] = 0.0;
    if (c->n_ip == 0) return SQG_OK;
    std::vector<double> rec;   // node, atom, fx, fy, fz per contributing image
    const int cap = choose_batch(c, c->nn, 3);
    const std::vector<int> nodes = node_order(c->p.n);
    int done_n = 0;
    while (done_n < c->nn) {
        const int B = std::min(cap, c->nn - done_n);
        int st = ensure_batch(c, B);
        if (st) return st;
        HIPCHK(hipMemcpy(c->node_list.p, nodes.data() + done_n, (size_t)B * sizeof(int), hipMemcpyHostToDevice));
        st = c->p.precision == SQG_MIXED ? forces_batch<float>(c, B, nodes.data() + done_n, rec)
                                         : forces_batch<double>(c, B, nodes.data() + done_n, rec);
        if (st) return st;
        done_n += B;
    }
    // sum in node order; within a node, all images of one atom form one
    // [V_nl,J d D](centre) term as in forces.cpp:509-520
    {
        const size_t nr0 = rec.size() / 5;
        std::vector<size_t> idx(nr0);
        for (size_t q = 0; q < nr0; ++q) idx[q] = q;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b2) { return rec[5 * a] < rec[5 * b2]; });
        std::vector<double> sorted(rec.size());
        for (size_t q = 0; q < nr0; ++q)
            for (int d = 0; d < 5; ++d) sorted[5 * q + d] = rec[5 * idx[q] + d];
        rec.swap(sorted);
    }
    const double h3 = std::pow(c->p.h, 3);
    size_t i = 0;
    const size_t nr = rec.size() / 5;
    while (i < nr) {
        size_t j = i;
        const double nd = rec[5 * i];
        std::vector<std::pair<int, std::array<double, 3>>> per;
        while (j < nr && rec[5 * j] == nd) {
            const int a = (int)rec[5 * j + 1];
            bool found = false;
            for (auto &q : per)
                if (q.first == a) {
                    for (int d = 0; d < 3; ++d) q.second[d] += rec[5 * j + 2 + d];
                    found = true;
                }
            if (!found) per.push_back({a, {rec[5 * j + 2], rec[5 * j + 3], rec[5 * j + 4]}});
            ++j;
        }
        std::sort(per.begin(), per.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (auto &q : per)
            for (int d = 0; d < 3; ++d) f[d * na + q.first] += -4 * q.second[d] * h3;
        i = j;
    }
    return SQG_OK;
}

int sqg_last_kernel_timing(sqg_ctx *c, int which, long *launches, double *total_ms, double *alg_bytes)
{
    if (!c || which < 0 || which > 3) return SQG_E_ARG;
    if (launches) *launches = c->last_launches[which];
    if (total_ms) *total_ms = c->last_ms[which];
    if (alg_bytes) *alg_bytes = c->last_bytes[which];
    return SQG_OK;
}

int sqg_density_fixed(sqg_ctx *c, const double *veff, double ncharge, double lf_lo, double lf_hi,
                      double *lambda_f, double *rho)
{
    if (!c || !veff || !lambda_f || !rho) return SQG_E_ARG;
    if (!c->have_state) return SQG_E_NOSTATE;
    HIPCHK(hipSetDevice(c->p.device));
    int st = upload_veff(c, veff);
    if (st) return st;
    c->have_state = false;
    const int cap = choose_batch(c, c->nn);
    const std::vector<int> all = node_order(c->p.n);
    long passes = 0;
    timing_reset(c);
    st = moments_dispatch(c, all.data(), c->nn, &passes, cap, false);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(c->st));
    timing_sum(c);
    const size_t N = (size_t)c->nn, K = (size_t)c->p.npl + 1;
    HIPCHK(c->W.ensure(N * K * sizeof(double)));
    hipLaunchKernelGGL(k_cheb_W, dim3(grid1(N * K, 256)), dim3(256), 0, c->st, c->mu_all.as<double>(),
                       c->costab.as<double>(), c->nn, c->p.npl, c->W.as<double>());
    double lf = 0.0;
    int fit = 0;
    st = brent(c, lf_lo, lf_hi, 1, 0.0, ncharge, &lf, &fit);
    if (st) return st;
    double Ebs = 0.0, Eent = 0.0;
    int neg = 0;
    st = rho_assembly(c, lf, rho, &Ebs, &Eent, &neg, nullptr);
    if (st) return st;
    *lambda_f = lf;
    c->lambda_f = lf;
    c->have_state = true;
    return neg ? SQG_E_NEG_RHO : SQG_OK;
}

int sqg_adjoint_veff(sqg_ctx *c, const double *w, int fixed_n, double *dLdV)
{
    if (!c || !w || !dLdV) return SQG_E_ARG;
    if (!c->have_state) return SQG_E_NOSTATE;
    HIPCHK(hipSetDevice(c->p.device));
    const int n = c->p.n, nn = c->nn;
    std::vector<double> sgn, dl;
    int st = adj_node_terms(c, sgn, dl);
    if (st) return st;
    // fixed electron count: the Fermi level moves by dlambda = -sum dRho/sum d rho/d lambda,
    // which turns the weights into w - kappa
    double num = 0.0, den = 0.0;
    for (int i = 0; i < nn; ++i) {
        num += w[i] * sgn[i] * dl[i];
        den += sgn[i] * dl[i];
    }
    if (fixed_n && den == 0.0) return SQG_E_ARG;
    const double kappa = fixed_n ? num / den : 0.0;
    std::vector<double> wt(nn);
    for (int i = 0; i < nn; ++i) wt[i] = (w[i] - kappa) * sgn[i];
    HIPCHK(c->adj_wt.ensure((size_t)nn * sizeof(double)));
    HIPCHK(c->adj_out.ensure((size_t)nn * sizeof(double)));
    HIPCHK(hipMemcpyAsync(c->adj_wt.p, wt.data(), (size_t)nn * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemsetAsync(c->adj_out.p, 0, (size_t)nn * sizeof(double), c->st));
    const AdjPlan P = adj_plan(c);
    const int plane = n * n;
    const int cap = choose_batch_adj(c, nn, (size_t)P.nvec * c->M3 * c->real_size());
    int nz = std::max(1, std::min(n, cap / plane));
    nz = std::min(nz, std::max(1, 65535 / plane));
    for (int z0 = 0; z0 < n; z0 += nz) {
        const int nzb = std::min(nz, n - z0);
        st = c->p.precision == SQG_MIXED ? adjoint_batch<float>(c, z0, nzb, P) : adjoint_batch<double>(c, z0, nzb, P);
        if (st) return st;
    }
    HIPCHK(hipMemcpy(dLdV, c->adj_out.p, (size_t)nn * sizeof(double), hipMemcpyDeviceToHost));
    return SQG_OK;
}

int sqg_tangent_veff(sqg_ctx *c, const double *dV, int fixed_n, double *drho)
{
    if (!c || !dV || !drho) return SQG_E_ARG;
    if (!c->have_state) return SQG_E_NOSTATE;
    HIPCHK(hipSetDevice(c->p.device));
    const int nn = c->nn;
    std::vector<double> sgn, dl;
    int st = adj_node_terms(c, sgn, dl);
    if (st) return st;
    HIPCHK(c->adj_in.ensure((size_t)nn * sizeof(double)));
    HIPCHK(c->adj_out.ensure((size_t)nn * sizeof(double)));
    HIPCHK(hipMemcpyAsync(c->adj_in.p, dV, (size_t)nn * sizeof(double), hipMemcpyHostToDevice, c->st));
    const int cap = std::min(choose_batch_adj(c, nn, 6 * (size_t)c->M3 * c->real_size() +
                                                     (size_t)(c->p.npl / 2 + 1) * 3 * sizeof(double)),
                             65535);
    for (int s0 = 0; s0 < nn; s0 += cap) {
        const int B = std::min(cap, nn - s0);
        st = c->p.precision == SQG_MIXED ? tangent_batch<float>(c, s0, B) : tangent_batch<double>(c, s0, B);
        if (st) return st;
    }
    HIPCHK(hipMemcpy(drho, c->adj_out.p, (size_t)nn * sizeof(double), hipMemcpyDeviceToHost));
    if (fixed_n) {
        double num = 0.0, den = 0.0;
        for (int i = 0; i < nn; ++i) {
            num += drho[i];
            den += sgn[i] * dl[i];
        }
        if (den == 0.0) return SQG_E_ARG;
        const double dlam = -num / den;
        for (int i = 0; i < nn; ++i) drho[i] += sgn[i] * dl[i] * dlam;
    }
    return SQG_OK;
}

int sqg_poisson_aaj(sqg_ctx *c, const double *rhs, const double *phi_guess, double tol, int maxiter, double beta,
                    int m, int p, double *phi, int *iters, double *relres)
{
    // periodic stencil indices wrap once: the grid must span the stencil half-width
    if (!c || !rhs || !phi_guess || !phi || m < 1 || m > 8 || p < 1 || maxiter < 1 || !(tol > 0) || c->p.n < FDN)
        return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    return poisson_aaj(c, rhs, phi_guess, tol, maxiter, beta, m, p, phi, iters, relres);
}

int sqg_vxc_lda(sqg_ctx *c, const double *rho, double *vxc)
{
    if (!c || !rho || !vxc) return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    const size_t N = (size_t)c->p.n * c->p.n * c->p.n;
    HIPCHK(c->pz_rhs.ensure(N * sizeof(double)));
    HIPCHK(c->pz_out.ensure(N * sizeof(double)));
    HIPCHK(hipMemcpyAsync(c->pz_rhs.p, rho, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_vxc_lda, dim3((int)std::min<size_t>((N + PB - 1) / PB, 1024)), dim3(PB), 0, c->st,
                       c->pz_rhs.as<double>(), N, c->pz_out.as<double>());
    HIPCHK(hipMemcpyAsync(vxc, c->pz_out.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return SQG_OK;
}

int sqg_total_energy(sqg_ctx *c, const double *phi, const double *b, const double *rho, const double *vxc,
                     double ebs, double eent, double eself, double ecorr, int n_atm, double *etot)
{
    if (!c || !phi || !b || !rho || !vxc || !etot || n_atm < 1) return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    const size_t N = (size_t)c->p.n * c->p.n * c->p.n;
    const int nb = (int)std::min<size_t>((N + PB - 1) / PB, 1024);
    for (DBuf *d : {&c->pz_phi, &c->pz_new, &c->pz_res, &c->pz_X}) HIPCHK(d->ensure(N * sizeof(double)));
    HIPCHK(c->pz_part.ensure((size_t)nb * 3 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(c->pz_phi.p, phi, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_new.p, b, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_res.p, rho, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_X.p, vxc, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_energy_terms, dim3(nb), dim3(PB), 0, c->st, c->pz_phi.as<double>(), c->pz_new.as<double>(),
                       c->pz_res.as<double>(), c->pz_X.as<double>(), N, c->p.h, c->pz_part.as<double>());
    std::vector<double> part((size_t)nb * 3);
    HIPCHK(hipMemcpyAsync(part.data(), c->pz_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double t[3] = {0, 0, 0};
    for (int bb = 0; bb < nb; ++bb)
        for (int q = 0; q < 3; ++q) t[q] += part[(size_t)bb * 3 + q];
    double E = t[0] + t[1] + t[2];
    E = E - eself;
    E = E + ebs + eent + ecorr;
    *etot = E / n_atm;
    return SQG_OK;
}

}  // extern "C"

namespace {

// Replica boxes of a local-pseudopotential pass (ForcesLocal / ChargeDensity):
// validated spec, radial tables on the device, and batches of replicas whose
// boxes (intersection + 2 FDn per side) fit the scratch arrays.
struct LocPlan {
    std::vector<sqk::LImg> img;
    std::vector<long> npts;
    long cap = 0, bmax = 1;
    sqk::LocArgs A{};
};

int loc_prepare(sqg_ctx *c, const sqg_local_spec *L, bool with_uu, LocPlan &P)
{
    const int n = c->p.n, F = c->p.FDn;
    if (L->n_typ < 1 || L->n_typ > SQDFT_MAX_TYPES || L->n_img < 0 || L->n_atm < 1 || !L->size || !L->r ||
        !L->rvloc || !L->dvloc || !L->Z || (with_uu && (!L->uu || !L->duu)) ||
        (L->n_img > 0 && (!L->img_xyz || !L->img_type || !L->img_atom || !L->img_box)))
        return SQG_E_ARG;
    // radial tables of all types side by side
    std::vector<int> toff(L->n_typ), tsize(L->n_typ);
    std::vector<double> R, RV, DV, RU, DU;
    for (int t = 0; t < L->n_typ; ++t) {
        if (L->size[t] < 2 || !L->r[t] || !L->rvloc[t] || !L->dvloc[t]) return SQG_E_ARG;
        if (with_uu && (!L->uu[t] || !L->duu[t])) return SQG_E_ARG;
        toff[t] = (int)R.size();
        tsize[t] = L->size[t];
        R.insert(R.end(), L->r[t], L->r[t] + L->size[t]);
        RV.insert(RV.end(), L->rvloc[t], L->rvloc[t] + L->size[t]);
        DV.insert(DV.end(), L->dvloc[t], L->dvloc[t] + L->size[t]);
        if (with_uu) {
            RU.insert(RU.end(), L->uu[t], L->uu[t] + L->size[t]);
            DU.insert(DU.end(), L->duu[t], L->duu[t] + L->size[t]);
        }
    }
    // replicas: every box index must lie in [0, n-1] (the kernels address the
    // n^3 arrays there without wrapping)
    P.img.assign(L->n_img, sqk::LImg{});
    P.npts.assign(L->n_img, 0);
    for (int m = 0; m < L->n_img; ++m) {
        sqk::LImg &I = P.img[m];
        I.x = L->img_xyz[3 * m]; I.y = L->img_xyz[3 * m + 1]; I.z = L->img_xyz[3 * m + 2];
        I.type = L->img_type[m]; I.pad = 0;
        if (I.type < 0 || I.type >= L->n_typ || L->img_atom[m] < 0 || L->img_atom[m] >= L->n_atm) return SQG_E_ARG;
        long np = 1;
        for (int d = 0; d < 3; ++d) {
            I.lo[d] = L->img_box[6 * m + d];
            I.hi[d] = L->img_box[6 * m + 3 + d];
            if (I.lo[d] < 0 || I.hi[d] > n - 1 || I.lo[d] > I.hi[d]) return SQG_E_ARG;
            np *= I.hi[d] - I.lo[d] + 1 + 4 * F;
        }
        P.npts[m] = np;
        P.bmax = std::max(P.bmax, np);
    }
    P.cap = std::max(1L << 26, P.bmax);   // 64M points per scratch array
    for (DBuf *d : {&c->lf_V, &c->lf_Vt, &c->lf_bJ, &c->lf_bJt}) HIPCHK(d->ensure((size_t)P.cap * sizeof(double)));
    const size_t nr = R.size();
    for (DBuf *d : {&c->lf_R, &c->lf_RV, &c->lf_DV, &c->lf_RU, &c->lf_DU}) HIPCHK(d->ensure(nr * sizeof(double)));
    HIPCHK(c->lf_toff.ensure(L->n_typ * sizeof(int)));
    HIPCHK(c->lf_tsize.ensure(L->n_typ * sizeof(int)));
    HIPCHK(c->lf_Z.ensure(L->n_typ * sizeof(double)));
    HIPCHK(c->lf_img.ensure(std::max<size_t>(P.img.size(), 1) * sizeof(sqk::LImg)));
    HIPCHK(c->lf_err.ensure(sizeof(int)));
    HIPCHK(hipMemcpyAsync(c->lf_R.p, R.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->lf_RV.p, RV.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->lf_DV.p, DV.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (with_uu) {
        HIPCHK(hipMemcpyAsync(c->lf_RU.p, RU.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->lf_DU.p, DU.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->st));
    }
    HIPCHK(hipMemcpyAsync(c->lf_toff.p, toff.data(), L->n_typ * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->lf_tsize.p, tsize.data(), L->n_typ * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->lf_Z.p, L->Z, L->n_typ * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemsetAsync(c->lf_err.p, 0, sizeof(int), c->st));
    sqk::LocArgs &A = P.A;
    A.n = n; A.F = F; A.h = c->p.h; A.h3 = std::pow(c->p.h, 3);
    for (int a = 0; a <= F; ++a) { A.lap[a] = c->p.coeff_lap[a]; A.grad[a] = c->p.coeff_grad[a]; }
    A.R = c->lf_R.as<double>(); A.RV = c->lf_RV.as<double>(); A.DV = c->lf_DV.as<double>();
    A.RU = c->lf_RU.as<double>(); A.DU = c->lf_DU.as<double>();
    A.toff = c->lf_toff.as<int>(); A.tsize = c->lf_tsize.as<int>(); A.Z = c->lf_Z.as<double>();
    A.V = c->lf_V.as<double>(); A.Vt = c->lf_Vt.as<double>(); A.bJ = c->lf_bJ.as<double>();
    A.bJt = c->lf_bJt.as<double>();
    A.err = c->lf_err.as<int>();
    A.img = c->lf_img.as<sqk::LImg>();
    return SQG_OK;
}

// Runs the potential and charge kernels batch by batch, then body(m0, nb)
// with A.img / A.nimg set to the batch (replicas m0 .. m0 + nb - 1).
template <class Body>
int loc_batches(sqg_ctx *c, LocPlan &P, Body &&body)
{
    const int nimg = (int)P.img.size();
    for (int m0 = 0; m0 < nimg;) {
        int m1 = m0;
        long used = 0, bmax = 1;
        while (m1 < nimg && m1 - m0 < 65535 && used + P.npts[m1] <= P.cap) {
            P.img[m1].off = used;
            used += P.npts[m1];
            bmax = std::max(bmax, P.npts[m1]);
            ++m1;
        }
        const int nb = m1 - m0;
        HIPCHK(hipMemcpyAsync(c->lf_img.p, P.img.data() + m0, nb * sizeof(sqk::LImg), hipMemcpyHostToDevice, c->st));
        P.A.nimg = nb;
        const int gx = (int)std::min<long>((bmax + sqk::PB - 1) / sqk::PB, 64);
        hipLaunchKernelGGL(sqk::k_loc_potential, dim3(gx, nb), dim3(sqk::PB), 0, c->st, P.A);
        hipLaunchKernelGGL(sqk::k_loc_charge, dim3(gx, nb), dim3(sqk::PB), 0, c->st, P.A);
        HIPCHK(hipGetLastError());
        if (int st = body(m0, nb)) return st;
        m0 = m1;
    }
    return SQG_OK;
}

int loc_error(sqg_ctx *c)
{
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, c->lf_err.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return err ? SQG_E_SPLINE : SQG_OK;
}

}  // namespace

extern "C" {

int sqg_forces_local(sqg_ctx *c, const sqg_local_spec *L, const double *phi, const double *b, const double *bt,
                     const double *vc, double *floc, double *fcorr)
{
    if (!c || !L || !phi || !b || !bt || !vc || !floc || !fcorr) return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    LocPlan P;
    if (int st = loc_prepare(c, L, false, P)) return st;
    const size_t N = (size_t)c->p.n * c->p.n * c->p.n;
    for (DBuf *d : {&c->pz_phi, &c->pz_new, &c->pz_res, &c->pz_X}) HIPCHK(d->ensure(N * sizeof(double)));
    HIPCHK(hipMemcpyAsync(c->pz_phi.p, phi, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_new.p, b, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_res.p, bt, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_X.p, vc, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    P.A.phi = c->pz_phi.as<double>(); P.A.b = c->pz_new.as<double>(); P.A.bt = c->pz_res.as<double>();
    P.A.vc = c->pz_X.as<double>();
    for (int q = 0; q < 3 * L->n_atm; ++q) floc[q] = fcorr[q] = 0.0;
    const int GX = 64;
    std::vector<double> part;
    int st = loc_batches(c, P, [&](int m0, int nb) -> int {
        HIPCHK(c->lf_part.ensure((size_t)nb * GX * 6 * sizeof(double)));
        P.A.part = c->lf_part.as<double>();
        hipLaunchKernelGGL(sqk::k_loc_force, dim3(GX, nb), dim3(sqk::PB), 0, c->st, P.A);
        HIPCHK(hipGetLastError());
        part.resize((size_t)nb * GX * 6);
        HIPCHK(hipMemcpyAsync(part.data(), c->lf_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        // replica totals in block order, added to the atom in replica order
        for (int q = 0; q < nb; ++q) {
            double f[6] = {0, 0, 0, 0, 0, 0};
            for (int bx = 0; bx < GX; ++bx)
                for (int k = 0; k < 6; ++k) f[k] += part[((size_t)q * GX + bx) * 6 + k];
            const int a = L->img_atom[m0 + q];
            for (int d = 0; d < 3; ++d) {
                floc[d * L->n_atm + a] += f[d];
                fcorr[d * L->n_atm + a] += f[3 + d];
            }
        }
        return SQG_OK;
    });
    if (st) return st;
    return loc_error(c);
}

int sqg_pseudocharge(sqg_ctx *c, const sqg_local_spec *L, double *b_out, double *rho_out, double *ncharge,
                     double *eself, double *ecorr, double *bt_out, double *vc_out, double *rho_at_out)
{
    if (!c || !L) return SQG_E_ARG;
    HIPCHK(hipSetDevice(c->p.device));
    LocPlan P;
    if (int st = loc_prepare(c, L, true, P)) return st;
    const int n = c->p.n;
    const size_t N = (size_t)n * n * n;
    for (DBuf *d : {&c->pz_phi, &c->pz_new, &c->pz_res, &c->pz_X}) HIPCHK(d->ensure(N * sizeof(double)));
    // b, b~, Vc, rho_at accumulate over the batches in replica order
    for (DBuf *d : {&c->pz_phi, &c->pz_new, &c->pz_res, &c->pz_X})
        HIPCHK(hipMemsetAsync(d->p, 0, N * sizeof(double), c->st));
    P.A.gb = c->pz_phi.as<double>(); P.A.gbt = c->pz_new.as<double>(); P.A.gvc = c->pz_res.as<double>();
    P.A.grho = c->pz_X.as<double>();
    const int nbk = (int)std::min<size_t>((N + sqk::PB - 1) / sqk::PB, 1024);
    HIPCHK(c->lf_part.ensure((size_t)nbk * 4 * sizeof(double)));
    std::vector<double> part((size_t)nbk * 4);
    double es = 0, nc = 0, est = 0, nct = 0;
    int st = loc_batches(c, P, [&](int, int) -> int {
        hipLaunchKernelGGL(sqk::k_pc_gather, dim3(nbk), dim3(sqk::PB), 0, c->st, P.A, c->lf_part.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(part.data(), c->lf_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int q = 0; q < nbk; ++q) {
            es += part[4 * q]; nc += part[4 * q + 1]; est += part[4 * q + 2]; nct += part[4 * q + 3];
        }
        return SQG_OK;
    });
    if (st) return st;
    if ((st = loc_error(c))) return st;
    std::vector<double> b(N), bt(N), vc(N), ra(N);
    HIPCHK(hipMemcpyAsync(b.data(), c->pz_phi.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(bt.data(), c->pz_new.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(vc.data(), c->pz_res.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(ra.data(), c->pz_X.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    // scaling, Vc shift and Ecorr in grid order (src/initialize.cpp:699-779)
    const double h = c->p.h;
    double Nel = 0;
    for (size_t g = 0; g < N; ++g) Nel += ra[g] * h * h * h;
    if (rho_out)
        for (size_t g = 0; g < N; ++g) rho_out[g] = std::fabs(nc / Nel) * ra[g];
    for (size_t g = 0; g < N; ++g) bt[g] = std::fabs(nc / nct) * bt[g];
    // Vcon: V~ - V at node (0,0,0) summed over the replicas holding it, i.e. Vc(0,0,0)
    const double V000 = vc[0], Vcon = vc[0];
    double Ec = 0;
    for (size_t g = 0; g < N; ++g) {
        vc[g] += -V000 + Vcon;
        Ec += 0.5 * (bt[g] + b[g]) * vc[g] * std::pow(h, 3);
    }
    if (b_out) std::memcpy(b_out, b.data(), N * sizeof(double));
    if (bt_out) std::memcpy(bt_out, bt.data(), N * sizeof(double));
    if (vc_out) std::memcpy(vc_out, vc.data(), N * sizeof(double));
    if (rho_at_out) std::memcpy(rho_at_out, ra.data(), N * sizeof(double));
    if (ncharge) *ncharge = nc;
    if (eself) *eself = es;
    if (ecorr) *ecorr = Ec + es - est;
    return SQG_OK;
}

}  // extern "C"
