// Host drivers of the Veff derivatives (included by sq_gpu.hip inside its
// anonymous namespace; uses the context and launch helpers defined there).
// No counterpart in the reference: they differentiate the moment recurrence
// of sq.cpp:568-655 with chi, zee frozen.

// Chebyshev step out = 2 A x - y (first: out = A x) for the nodes of the
// current batch, with the projections of x computed first.
template <typename T>
void cheb_step(sqg_ctx *c, int B, T *x, T *y, T *out, bool first)
{
    const int *dn = c->node_list.as<int>();
    PArgs<T> P = pargs<T>(c, dn, nullptr, x);
    launch_proj<T, 0>(c, B, P);
    ZArgs<T> Z = zargs<T>(c, dn, nullptr);
    Z.X = x;
    Z.Y = y ? y : out;
    Z.O = out;
    Z.chi = c->chi_all.as<double>();
    Z.zee = c->zee_all.as<double>();
    Z.chi_by_node = 1;
    Z.first = first;
    Z.center = c->center.as<double>();
    Z.dots = c->dots.as<double>();
    launch_zmarch<T, MODE_CHEB>(c, B, Z, -1, 0.0);
}

int batch_images(sqg_ctx *c, int B)
{
    const int *dn = c->node_list.as<int>();
    HIPCHK(hipMemsetAsync(c->flags.p, 0, 16 * sizeof(int), c->st));
    (void)dn;
    return node_images(c, B);
}

// Sign of rho_pc and d rho_pc/d lambda per node at the last lambda_f
// (device arrays adj_sgn, adj_dl; copied to the host vectors).
int adj_node_terms(sqg_ctx *c, std::vector<double> &sgn, std::vector<double> &dl)
{
    const size_t N = (size_t)c->nn;
    HIPCHK(c->adj_sgn.ensure(N * sizeof(double)));
    HIPCHK(c->adj_dl.ensure(N * sizeof(double)));
    hipLaunchKernelGGL(k_adj_node, dim3(grid1(N, 256)), dim3(256), 0, c->st, c->W.as<double>(),
                       c->chi_all.as<double>(), c->zee_all.as<double>(), c->ynode.as<double>(), c->nn, c->p.npl,
                       c->lambda_f, c->p.bet, c->adj_sgn.as<double>(), c->adj_dl.as<double>());
    sgn.resize(N);
    dl.resize(N);
    HIPCHK(hipMemcpyAsync(sgn.data(), c->adj_sgn.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(dl.data(), c->adj_dl.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return SQG_OK;
}

// Checkpoint layout of the reverse sweep: t_k is kept for k = j m and j m + 1
// (j < J = ceil(K2/m)); segment j is recomputed from its pair into m + 2
// slots before its reverse steps k = j m + 1 .. min((j + 1) m, K2).
struct AdjPlan {
    int m, J, nvec;
};

AdjPlan adj_plan(const sqg_ctx *c)
{
    const int K2 = c->p.npl / 2;
    int m = c->p.ckpt_interval;
    if (m <= 0 || m > K2) m = K2;
    if (m < 2) m = 2;
    AdjPlan P;
    P.m = m;
    P.J = (K2 + m - 1) / m;
    P.nvec = 2 * P.J + (m + 2) + 3 + 1;
    return P;
}

// dL/dVeff contributions of the nodes of z-planes [z0, z0 + nz), added into
// adj_out (n^3 doubles).
template <typename T>
int adjoint_batch(sqg_ctx *c, int z0, int nz, const AdjPlan &P)
{
    const int n = c->p.n, B = nz * n * n, K2 = c->p.npl / 2, q4 = c->p.npl / 4, m = P.m, J = P.J;
    const size_t M3 = c->M3, vl = (size_t)B * M3;
    const int cidx = (c->p.nloc * c->M + c->p.nloc) * c->M + c->p.nloc;
    int st = ensure_batch(c, B);
    if (st) return st;
    std::vector<int> nodes(B);
    for (int i = 0; i < B; ++i) nodes[i] = z0 * n * n + i;
    HIPCHK(hipMemcpyAsync(c->node_list.p, nodes.data(), (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->st));
    batch_images(c, B);
    HIPCHK(c->adj.ensure((size_t)P.nvec * vl * sizeof(T)));
    T *base = c->adj.as<T>();
    auto pair = [&](int j, int w) { return base + (size_t)(2 * j + w) * vl; };
    T *seg0 = base + (size_t)(2 * J) * vl;
    auto seg = [&](int i) { return seg0 + (size_t)i * vl; };
    T *tbr = seg0 + (size_t)(m + 2) * vl;
    auto tb = [&](int k) { return tbr + (size_t)(k % 3) * vl; };
    T *G = tbr + 3 * vl;

    // forward sweep: only the checkpoint pairs are kept (the rest rotate
    // through three segment slots)
    auto fwd = [&](int k) -> T * {
        const int r = k % m;
        return r <= 1 ? pair(k / m, r) : seg(2 + k % 3);
    };
    hipLaunchKernelGGL((k_fill_delta<T>), dim3(grid1(vl, 256)), dim3(256), 0, c->st, pair(0, 0), (int)M3, B, cidx);
    cheb_step<T>(c, B, pair(0, 0), nullptr, pair(0, 1), true);
    const int kf = (J - 1) * m + 1;
    for (int k = 1; k < kf; ++k) cheb_step<T>(c, B, fwd(k), fwd(k - 1), fwd(k + 1), false);

    // reverse sweep
    hipLaunchKernelGGL((k_fill<T>), dim3(grid1(4 * vl, 256)), dim3(256), 0, c->st, tbr, 4 * vl, (T)0);
    const double inv_h3 = 1.0 / std::pow(c->p.h, 3);
    const dim3 eg((unsigned)((M3 + 255) / 256), (unsigned)B);
    for (int j = J - 1; j >= 0; --j) {
        const int k0 = j * m, khi = std::min((j + 1) * m, K2), top = std::min((j + 1) * m + 1, K2);
        auto sp = [&](int k) -> T * { return k - k0 <= 1 ? pair(j, k - k0) : seg(k - k0); };
        for (int kk = k0 + 1; kk < top; ++kk) cheb_step<T>(c, B, sp(kk), sp(kk - 1), sp(kk + 1), false);
        for (int k = khi; k >= k0 + 1; --k) {
            if (k < K2) {   // the step and its source in one z-march
                const int *dn = c->node_list.as<int>();
                PArgs<T> P = pargs<T>(c, dn, nullptr, tb(k + 1));
                launch_proj<T, 0>(c, B, P);
                ZArgs<T> Z = zargs<T>(c, dn, nullptr);
                Z.X = tb(k + 1);
                Z.Y = tb(k + 2);
                Z.O = tb(k);
                Z.chi = c->chi_all.as<double>();
                Z.zee = c->zee_all.as<double>();
                Z.chi_by_node = 1;
                Z.first = 0;
                Z.Ci = c->Ci_all.as<double>();
                Z.wt = c->adj_wt.as<double>();
                Z.k = k; Z.npl = c->p.npl; Z.K2 = K2; Z.q4 = q4;
                Z.inv_h3 = inv_h3;
                Z.tk = sp(k); Z.tkm = sp(k - 1); Z.tkp = sp(k + 1);
                Z.G = G;
                launch_zmarch<T, MODE_ADJ>(c, B, Z, -1, 0.0);
                continue;
            }
            AdjArgs<T> A{};
            A.M3 = (int)M3; A.c = cidx; A.k = k; A.K2 = K2; A.q4 = q4; A.npl = c->p.npl;
            A.has_mv = k < K2;
            A.inv_h3 = inv_h3;
            A.node_list = c->node_list.as<int>();
            A.Ci = c->Ci_all.as<double>();
            A.wt = c->adj_wt.as<double>();
            A.tb = tb(k); A.tb1 = tb(k + 1);
            A.tkm = sp(k - 1); A.tk = sp(k); A.tkp = k + 1 <= K2 ? sp(k + 1) : nullptr;
            A.G = G;
            hipLaunchKernelGGL((k_adj_update<T>), eg, dim3(256), 0, c->st, A);
        }
    }
    hipLaunchKernelGGL((k_adj_center<T>), dim3((B + 127) / 128), dim3(128), 0, c->st, B, (int)M3, cidx,
                       (const T *)tb(1), G);
    hipLaunchKernelGGL((k_adj_gather<T>), dim3(grid1(c->nn, 256)), dim3(256), 0, c->st, n, c->p.nloc, z0, nz,
                       (const T *)G, c->zee_all.as<double>(), c->adj_out.as<double>());
    int ovf4[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(ovf4, c->flags.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    const int ovf = ovf4[0] | ovf4[2] | ovf4[3];
    HIPCHK(hipStreamSynchronize(c->st));
    return ovf ? SQG_E_OVERFLOW : SQG_OK;
}

// d rho along dV (adj_in) for the B nodes starting at node s0, into adj_out.
template <typename T>
int tangent_batch(sqg_ctx *c, int s0, int B)
{
    const int K2 = c->p.npl / 2, q4 = c->p.npl / 4;
    const size_t M3 = c->M3, vl = (size_t)B * M3;
    const int cidx = (c->p.nloc * c->M + c->p.nloc) * c->M + c->p.nloc;
    int st = ensure_batch(c, B);
    if (st) return st;
    std::vector<int> nodes(B);
    for (int i = 0; i < B; ++i) nodes[i] = s0 + i;
    HIPCHK(hipMemcpyAsync(c->node_list.p, nodes.data(), (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->st));
    batch_images(c, B);
    HIPCHK(c->adj.ensure(6 * vl * sizeof(T)));
    HIPCHK(c->adj_rec.ensure((size_t)B * (K2 + 1) * 3 * sizeof(double)));
    T *tr[3], *dr[3];
    for (int i = 0; i < 3; ++i) {
        tr[i] = c->adj.as<T>() + (size_t)i * vl;
        dr[i] = c->adj.as<T>() + (size_t)(3 + i) * vl;
    }
    const int *dn = c->node_list.as<int>();
    double *rec = c->adj_rec.as<double>();
    const dim3 eg((unsigned)((M3 + 255) / 256), (unsigned)B);
    hipLaunchKernelGGL((k_fill_delta<T>), dim3(grid1(vl, 256)), dim3(256), 0, c->st, tr[0], (int)M3, B, cidx);
    hipLaunchKernelGGL((k_fill<T>), dim3(grid1(vl, 256)), dim3(256), 0, c->st, dr[0], vl, (T)0);
    cheb_step<T>(c, B, tr[0], nullptr, tr[1], true);
    hipLaunchKernelGGL((k_tan_init<T>), eg, dim3(256), 0, c->st, B, (int)M3, cidx, K2, dn,
                       c->adj_in.as<double>(), c->zee_all.as<double>(), dr[1], rec);
    for (int k = 2; k <= K2; ++k) {
        cheb_step<T>(c, B, tr[(k - 1) % 3], tr[(k - 2) % 3], tr[k % 3], false);
        // d_k = 2 A d_{k-1} - d_{k-2} + 2 dA t_{k-1} and its dots in one z-march
        PArgs<T> P = pargs<T>(c, dn, nullptr, dr[(k - 1) % 3]);
        launch_proj<T, 0>(c, B, P);
        ZArgs<T> Z = zargs<T>(c, dn, nullptr);
        Z.X = dr[(k - 1) % 3];
        Z.Y = dr[(k - 2) % 3];
        Z.O = dr[k % 3];
        Z.chi = c->chi_all.as<double>();
        Z.zee = c->zee_all.as<double>();
        Z.chi_by_node = 1;
        Z.first = 0;
        Z.center = c->center.as<double>();
        Z.dots = c->dots.as<double>();
        Z.dV = c->adj_in.as<double>();
        Z.tk = tr[k % 3];
        Z.tkm = tr[(k - 1) % 3];
        launch_zmarch<T, MODE_TAN>(c, B, Z, -1, 0.0);
        hipLaunchKernelGGL(k_tan_rec, dim3((B + 127) / 128), dim3(128), 0, c->st, B, k, K2,
                           c->center.as<double>(), c->dots.as<double>(), rec);
    }
    hipLaunchKernelGGL(k_tan_finish, dim3((B + 127) / 128), dim3(128), 0, c->st, B, K2, q4, c->p.npl,
                       1.0 / std::pow(c->p.h, 3), dn, (const double *)rec, c->Ci_all.as<double>(),
                       c->adj_sgn.as<double>(), c->adj_out.as<double>());
    int ovf4[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(ovf4, c->flags.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    const int ovf = ovf4[0] | ovf4[2] | ovf4[3];
    HIPCHK(hipStreamSynchronize(c->st));
    return ovf ? SQG_E_OVERFLOW : SQG_OK;
}
