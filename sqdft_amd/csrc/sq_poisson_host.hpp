// sq_poisson_host.hpp -- host side of the AAJ Poisson solver and the LDA
// potential (included by sq_gpu.hip inside its anonymous namespace).

// x = pinv(A) b for a symmetric m x m matrix (column major) through a
// one-sided Jacobi SVD.  Singular values below m * max(w) * 2.22044605e-16
// are set to zero and zero ones skipped, as SingularValueDecomp and
// PseudoInverseTimesVec do (src/anderson.cpp:392-397, :134-142).
void pinv_solve(const double *Ac, const double *b, double *x, int m)
{
    double W[64], V[64];
    for (int i = 0; i < m * m; ++i) W[i] = Ac[i];
    for (int i = 0; i < m * m; ++i) V[i] = 0.0;
    for (int i = 0; i < m; ++i) V[i * m + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double worst = 0.0;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                double app = 0, aqq = 0, apq = 0;
                for (int i = 0; i < m; ++i) {
                    app += W[p * m + i] * W[p * m + i];
                    aqq += W[q * m + i] * W[q * m + i];
                    apq += W[p * m + i] * W[q * m + i];
                }
                if (apq == 0.0) continue;
                worst = std::max(worst, std::fabs(apq) / std::sqrt(app * aqq));
                const double zeta = (aqq - app) / (2 * apq);
                const double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                for (int i = 0; i < m; ++i) {
                    const double wp = W[p * m + i], wq = W[q * m + i];
                    W[p * m + i] = cs * wp - sn * wq;
                    W[q * m + i] = sn * wp + cs * wq;
                    const double vp = V[p * m + i], vq = V[q * m + i];
                    V[p * m + i] = cs * vp - sn * vq;
                    V[q * m + i] = sn * vp + cs * vq;
                }
            }
        if (worst < 1e-15) break;
    }
    double coef[8], sv[8], wmax = 0.0;
    for (int j = 0; j < m; ++j) {
        double s2 = 0;
        for (int i = 0; i < m; ++i) s2 += W[j * m + i] * W[j * m + i];
        sv[j] = std::sqrt(s2);
        if (sv[j] > wmax) wmax = sv[j];
    }
    const double wmin = m * wmax * (2.22044605e-16);
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        if (sv[j] >= wmin && sv[j] != 0.0) {
            for (int i = 0; i < m; ++i) s += W[j * m + i] / sv[j] * b[i];
            s /= sv[j];
        }
        coef[j] = s;
    }
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += V[j * m + i] * coef[j];
        x[i] = s;
    }
}

template <int M>
void gram_go(sqg_ctx *c, int nb, size_t N)
{
    hipLaunchKernelGGL((k_aaj_gram<M>), dim3(nb), dim3(PB), 0, c->st, c->pz_DF.as<double>(), c->pz_res.as<double>(), N,
                       c->pz_part.as<double>());
}

// PoissonSolver_AAJ (src/poisson.cpp:19-325): Jacobi with relaxation beta and
// an Anderson step every p iterations over the last m (iterate, residual)
// differences; stops at ||r|| / (4 pi ||rhs|| / |3 lap0|) <= tol and
// returns the iterate whose residual met it, shifted so that phi(0,0,0) = 0.
int poisson_aaj(sqg_ctx *c, const double *rhs, const double *guess, double tol, int maxiter, double beta, int m,
                int p, double *phi, int *iters_out, double *relres_out)
{
    const int n = c->p.n;
    const size_t N = (size_t)n * n * n;
    const int nb = (int)std::min<size_t>((N + PB - 1) / PB, 1024);
    const int NV = m * (m + 1) / 2 + m;
    for (DBuf *d : {&c->pz_phi, &c->pz_new, &c->pz_res, &c->pz_X, &c->pz_F, &c->pz_rhs, &c->pz_out})
        HIPCHK(d->ensure(N * sizeof(double)));
    HIPCHK(c->pz_DX.ensure((size_t)m * N * sizeof(double)));
    HIPCHK(c->pz_DF.ensure((size_t)m * N * sizeof(double)));
    HIPCHK(c->pz_part.ensure((size_t)nb * std::max(NV, 1) * sizeof(double)));
    HIPCHK(hipMemcpyAsync(c->pz_rhs.p, rhs, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pz_phi.p, guess, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL((k_fill<double>), dim3(grid1(N, 256)), dim3(256), 0, c->st, c->pz_X.as<double>(), N, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3(grid1(N, 256)), dim3(256), 0, c->st, c->pz_F.as<double>(), N, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3(grid1(m * N, 256)), dim3(256), 0, c->st, c->pz_DX.as<double>(), m * N, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3(grid1(m * N, 256)), dim3(256), 0, c->st, c->pz_DF.as<double>(), m * N, 0.0);
    // relative-residual scale of poisson.cpp:125-140, summed in index order on the host
    double rn = 0;
    for (size_t i = 0; i < N; ++i) rn = rn + rhs[i] * rhs[i];
    rn = std::sqrt(rn);
    rn *= 4 * M_PI * std::fabs(-1 / (3 * c->p.coeff_lap[0]));
    std::vector<double> part((size_t)nb * std::max(NV, 1));
    double *old = c->pz_phi.as<double>(), *nw = c->pz_new.as<double>();
    bool have_out = false;
    int iter = 1;
    double r = tol + 1;
    while (r > tol && iter <= maxiter) {
        AajArgs A{};
        A.n = n;
        for (int a = 0; a <= FDN; ++a) A.lap[a] = c->p.coeff_lap[a];
        A.phi = old; A.rhs = c->pz_rhs.as<double>();
        A.phinew = nw; A.res = c->pz_res.as<double>();
        A.Xold = c->pz_X.as<double>(); A.Fold = c->pz_F.as<double>();
        A.hist = iter > 1;
        const int k = iter > 1 ? (iter - 2) % m : 0;
        A.DXk = c->pz_DX.as<double>() + (size_t)k * N;
        A.DFk = c->pz_DF.as<double>() + (size_t)k * N;
        A.beta = beta;
        A.part = c->pz_part.as<double>();
        hipLaunchKernelGGL(k_aaj_step, dim3(nb), dim3(PB), 0, c->st, A);
        HIPCHK(hipMemcpyAsync(part.data(), c->pz_part.p, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        double s = 0;
        for (int bb = 0; bb < nb; ++bb) s += part[bb];
        r = std::sqrt(s) / rn;
        if (iter % p == 0 && iter > 1) {
            switch (m) {
            case 1: gram_go<1>(c, nb, N); break;
            case 2: gram_go<2>(c, nb, N); break;
            case 3: gram_go<3>(c, nb, N); break;
            case 4: gram_go<4>(c, nb, N); break;
            case 5: gram_go<5>(c, nb, N); break;
            case 6: gram_go<6>(c, nb, N); break;
            case 7: gram_go<7>(c, nb, N); break;
            default: gram_go<8>(c, nb, N); break;
            }
            HIPCHK(hipMemcpyAsync(part.data(), c->pz_part.p, (size_t)nb * NV * sizeof(double), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            std::vector<double> tot(NV, 0.0);
            for (int bb = 0; bb < nb; ++bb)
                for (int q = 0; q < NV; ++q) tot[q] += part[(size_t)bb * NV + q];
            double FtF[64], Ftf[8];
            int t = 0;
            for (int a = 0; a < m; ++a)
                for (int b2 = a; b2 < m; ++b2, ++t) FtF[b2 * m + a] = FtF[a * m + b2] = tot[t];
            for (int a = 0; a < m; ++a) Ftf[a] = tot[t++];
            AajUpd U{};
            pinv_solve(FtF, Ftf, U.x, m);
            U.m = m;
            U.beta = beta;
            hipLaunchKernelGGL(k_aaj_update, dim3(nb), dim3(PB), 0, c->st, c->pz_DX.as<double>(), c->pz_DF.as<double>(),
                               N, U, nw);
        }
        if (r <= tol || iter == maxiter) {
            HIPCHK(hipMemcpyAsync(c->pz_out.p, old, N * sizeof(double), hipMemcpyDeviceToDevice, c->st));
            have_out = true;
        }
        std::swap(old, nw);
        iter = iter + 1;
    }
    if (!have_out) return SQG_E_ARG;   // maxiter < 1
    HIPCHK(hipMemcpyAsync(phi, c->pz_out.p, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const double p000 = phi[0];   // phi(0,0,0) = 0 (poisson.cpp:272-290)
    for (size_t i = 0; i < N; ++i) phi[i] = phi[i] + 0.0 - p000;
    if (iters_out) *iters_out = iter - 1;
    if (relres_out) *relres_out = r;
    return SQG_OK;
}
