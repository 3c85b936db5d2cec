/*
 * sqdft_gpu.h -- C-ABI of the MI355X spectral-quadrature (SQ) hot path.
 *
 * One context per GPU; not thread-safe; one HIP stream; every call is
 * synchronous at return.  Host buffers are owned by the caller and only
 * borrowed during a call; device memory is owned by the context.  Every
 * entry point returns an int status (SQG_OK or a negative SQG_E_* code)
 * instead of printing and calling exit() as the reference does.
 *
 * Reference interfaces replaced (phanish-suryanarayana/SQDFT):
 *   sqg_set_geometry      <- output of Replica_atoms_Rcut + NonlocalProjectors
 *                            (src/initialize.cpp:1003-1196, src/nonlocal.cpp:317-518;
 *                            protos inc/func_sq.h:54, :368), consumed by
 *                            VnlocsubTimesVec (src/nonlocal.cpp:20-215, inc/func_sq.h:341)
 *   sqg_density           <- ClenshawCurtisSpectralQuadrature(DS_SQ*, int scf_iter)
 *                            (src/sq.cpp:377-834, inc/func_sq.h:515), called by
 *                            SCF_iteration (src/scf.cpp:150)
 *   sqg_forces_nonlocal   <- ForcesNonLocal(DS_SQ*) (src/forces.cpp:318-578,
 *                            inc/func_sq.h:540) with VnlocsubTimesVec_J
 *                            (src/nonlocal.cpp:522-724, inc/func_sq.h:351)
 *   sqg_hsub_apply_batch  <- HsubTimesVec(DS_SQ*, double***, int,int,int, double***)
 *                            (src/sq.cpp:838-929, inc/func_sq.h:377)
 *   sqg_node_moments      <- the per-node body of the SQ loop: LanczosAlgorithm
 *                            (src/sq.cpp:965-1092, inc/func_sq.h:417) with
 *                            TridiagEigenSolve (src/sq.cpp:1094-1180, inc/func_sq.h:405)
 *                            and the Chebyshev moment recurrence (src/sq.cpp:564-655)
 *   sqg_get_node_data     <- reads DS_SQ::chi, zee, rho_pj, Ci (inc/ds_sq.h:195-204),
 *                            which forces.cpp:402-450 consumes after the SQ call
 */
#ifndef SQDFT_GPU_H
#define SQDFT_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SQG_OK 0
#define SQG_E_ARG           -1
#define SQG_E_HIP           -2    /* a HIP runtime call failed */
#define SQG_E_NOGEOM        -3    /* sqg_set_geometry not called */
#define SQG_E_NPL           -4    /* npl not a multiple of 4 (sq.cpp:523-524) */
#define SQG_E_NEG_RHO       -5    /* rho < -1e-6 at some node (sq.cpp:764-765) */
#define SQG_E_QL            -6    /* QL exceeded 200 sweeps (sq.cpp:1125) */
#define SQG_E_BRACKET       -7    /* Brent root not bracketed (sq.cpp:1307-1311) */
#define SQG_E_BRENT_ITMAX   -8    /* Brent exceeded 100 iterations (sq.cpp:1400-1401) */
#define SQG_E_DENOM         -9    /* a KB denominator is zero (nonlocal.cpp:163) */
#define SQG_E_CUBE          -10   /* 2*nloc+1 above what one workgroup can march (43) */
#define SQG_E_OVERFLOW      -11   /* a table above a build limit (image list, ELL
                                     slots, a 2 GiB buffer range) */
#define SQG_E_NOSTATE       -12   /* forces/getters called before sqg_density */
#define SQG_E_FDN           -13   /* FD order other than 12 */
#define SQG_E_SPLINE        -14   /* spline abscissa outside the radial grid (spline.cpp:17-22) */
#define SQG_W_LANCZOS_MAXIT  1    /* some node hit max_iter=500 (sq.cpp:1073-1074; warning) */
#define SQG_W_NEWTON_MAXIT   2    /* Newton hit ITMAX=200 (sq.cpp:1479-1482; warning) */

#define SQG_FP64  0   /* every vector, stencil and reduction in FP64 */
#define SQG_MIXED 1   /* FP32 vectors/stencil/Veff/projectors, FP64 reductions,
                         moments, coefficients, Lanczos tridiagonal and Fermi solve */
#define SQG_MIXED_LANCZOS 2   /* the Lanczos phase as SQG_MIXED, the Chebyshev
                         moments, forces and Veff derivatives as SQG_FP64 */

typedef struct sqg_params {
    int n;                    /* FD nodes per direction of the periodic cube */
    double h;                 /* mesh size (DS_SQ::delta) */
    int FDn;                  /* half FD order; must be 6 */
    double coeff_lap[16];     /* DS_SQ::coeff_lap[0..FDn] */
    double coeff_grad[16];    /* DS_SQ::coeff_grad[0..FDn] */
    int nloc;                 /* ceil(Rcut/h) */
    int npl;                  /* Chebyshev degree, multiple of 4 */
    double bet;               /* 1/(kB T) in 1/Ha */
    double lanczos_tol;
    double fermi_tol;
    int fermi_alg;            /* 1 Brent, 2 Newton-Raphson */
    int precision;            /* SQG_FP64, SQG_MIXED or SQG_MIXED_LANCZOS */
    int batch_nodes;          /* FD nodes per launch (0: all nodes, capped by memory) */
    int ckpt_interval;        /* adjoint checkpoint interval m: sqg_adjoint_veff keeps
                                 the recurrence pair every m steps and recomputes the
                                 vectors in between (<= 0 or > npl/2: npl/2) */
    int device;               /* HIP device ordinal */
} sqg_params;

typedef struct sqg_stats {
    /* seconds, GPU events: t_lanczos = every Lanczos pass (projections, nonlocal
       apply, z-march, eigenvalue bounds) summed over batches; t_cheb = the rest
       of the moments phase (image lists, Chebyshev recurrence, moments) */
    double t_lanczos, t_cheb, t_fermi, t_rho, t_total;
    int lanczos_iters_node0;  /* the count the reference prints for node 0 */
    int lanczos_iters_min, lanczos_iters_max;
    long lanczos_passes;      /* Hsub passes spent in Lanczos, summed over nodes */
    int fermi_iters;
    double lambda_min, lambda_max;   /* global padded spectrum bounds */
} sqg_stats;

typedef struct sqg_ctx sqg_ctx;

int sqg_create(const sqg_params *p, sqg_ctx **out);
int sqg_destroy(sqg_ctx *ctx);

/* Atom images and their projector tables (host FP64, copied to the device).
 *   per type t:  nlocc[t] = ceil(rc/h), nproj[t] = sum_{l<=lmax, l!=lloc}(2l+1),
 *                denom[16*t + i] = Denom_l of the i-th projector (l-major, m=-l..l)
 *   per image J: img_type[J], img_atom[J] (atom index in type order),
 *                img_nn[3J..3J+2] nearest node (unwrapped, may lie outside [0,n)),
 *                udv + sum of earlier images' sizes: nproj x (2nlocc+1)^3 values,
 *                projector-major then z, y, x. */
int sqg_set_geometry(sqg_ctx *ctx, int n_typ, const int *nlocc, const int *nproj,
                     const double *denom, int n_atm, int n_img, const int *img_type,
                     const int *img_atom, const int *img_nn, const double *udv);

/* One SQ density step on the periodic Veff grid (n^3, x fastest).
 * lambda_f: in = previous Fermi level (used when scf_iter > 1 or Newton),
 * out = new Fermi level.  rho: n^3 out.  The per-node chi, zee, moments and
 * Fermi coefficients stay on the device for sqg_forces_nonlocal and getters. */
int sqg_density(sqg_ctx *ctx, const double *veff, int scf_iter, double ncharge,
                double *lambda_f, double *rho, double *ebs, double *eent, sqg_stats *stats);

/* Nonlocal forces with the Veff/chi/zee/Ci of the last sqg_density call.
 * f: 3*n_atm out, laid out fx[n_atm], fy[n_atm], fz[n_atm]. */
int sqg_forces_nonlocal(sqg_ctx *ctx, double *f);

/* Copies the last step's per-node data (any pointer may be NULL):
 * chi[nn], zee[nn], mu[nn*(npl+1)], Ci[nn*(npl+1)], lanczos_iters[nn]. */
int sqg_get_node_data(sqg_ctx *ctx, double *chi, double *zee, double *mu, double *Ci,
                      int *lanczos_iters);

/* Lanczos bounds and Chebyshev moments for a list of nodes only
 * (no Fermi solve); outputs sized nb, nb, nb*(npl+1), nb. */
int sqg_node_moments(sqg_ctx *ctx, const double *veff, const int *nodes, int nb,
                     double *chi, double *zee, double *mu, int *lanczos_iters);

/* Hv = H_i v for nb nodes; v and hv hold nb vectors of (2 nloc+1)^3
 * values, each z-major ([kk][jj][ii]) like the reference's double***. */
int sqg_hsub_apply_batch(sqg_ctx *ctx, const double *veff, const int *nodes, int nb,
                         const double *v, double *hv);

/* The scratch array of sqg_hsub_apply_batch_noalloc: a file-scope device
 * array the library holds on every GPU it runs on until the process ends. */
#define SQG_NOALLOC_SCRATCH_BYTES (64 << 20)

/* sqg_hsub_apply_batch without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  veff (n^3), v and hv
 * (nb vectors of (2 nloc+1)^3 values, z-major) are FP64 device pointers;
 * nodes is a host array.  Every per-call buffer (image lists, projections,
 * the nonlocal term, the FP32 copies of the mixed mode) lives in the
 * library's SQG_NOALLOC_SCRATCH_BYTES scratch array: SQG_E_OVERFLOW when nb
 * nodes need more.  Runs on the context's stream and returns after it has
 * synchronised.  Calls that overlap on one GPU share that array and race. */
int sqg_hsub_apply_batch_noalloc(sqg_ctx *ctx, const double *veff, const int *nodes, int nb,
                                 const double *v, double *hv);

/* Device time (ms, HIP events) of the hot kernels over the last density
 * call: launches, summed duration and summed algorithmic bytes (N =
 * (2nloc+1)^3 values per node, s = sizeof(real)).  which: 0 Chebyshev
 * z-march (3 N s per node: t_k, t_{k-1} read, t_{k+1} written), 1 Lanczos
 * z-march (4 N s: r', v_{k-1} read, v_k, r'_{k+1} written), 2 projections of
 * the Lanczos passes (2 N s: the two stored vectors read), 3 projections of
 * the Chebyshev steps (N s). */
int sqg_last_kernel_timing(sqg_ctx *ctx, int which, long *launches, double *total_ms,
                           double *alg_bytes);

/* ---- Derivatives with respect to Veff (no counterpart in the reference) ----
 * All three use the state of the last density call (sqg_density or
 * sqg_density_fixed): Veff, chi, zee, moments, lambda_f and Ci.  chi and zee
 * stay frozen (the Lanczos bounds are not differentiated); the moments are
 * differentiated exactly as src/sq.cpp:568-655 computes them (direct centres
 * up to npl/2, the doubling dot products above npl/4). */

/* The SQ density step with the last step's chi, zee kept: moments, Brent
 * root in [lf_lo, lf_hi] (the scf_iter = 1 bracket of src/sq.cpp:1289-1404),
 * rho.  The finite-difference partner of the two calls below. */
int sqg_density_fixed(sqg_ctx *ctx, const double *veff, double ncharge, double lf_lo, double lf_hi,
                      double *lambda_f, double *rho);

/* dL/dVeff (n^3 out) for L = sum_i w[i] rho[i] (w: n^3).  Reverse mode
 * through the Chebyshev recurrence, recomputing vectors from checkpoints
 * every ckpt_interval steps.  fixed_n = 1 includes the Fermi-level shift that
 * keeps the electron count fixed. */
int sqg_adjoint_veff(sqg_ctx *ctx, const double *w, int fixed_n, double *dLdV);

/* drho (n^3 out) along dV (n^3): forward mode through the recurrence. */
int sqg_tangent_veff(sqg_ctx *ctx, const double *dV, int fixed_n, double *drho);

/* ---- The SCF step around the SQ call (FP64 on the device) ----
 * AAJ Poisson solve <- PoissonSolver_AAJ(DS_SQ*) (src/poisson.cpp:19-325,
 * inc/func_sq.h:303) with PoissonResidual (src/poisson.cpp:715-928) and
 * AndersonExtrapolation (src/anderson.cpp:20-163): phi of
 * -Lap_FD phi = 4 pi rhs on the periodic n^3 grid (rhs = rho + b, x fastest)
 * from phi_guess, Jacobi relaxation beta, an Anderson step every p
 * iterations over the last m (1..8) differences, stopped at relative
 * residual <= tol or maxiter; phi is shifted to phi(0,0,0) = 0.  iters and
 * relres are the counts the reference prints. */
int sqg_poisson_aaj(sqg_ctx *ctx, const double *rhs, const double *phi_guess, double tol, int maxiter,
                    double beta, int m, int p, double *phi, int *iters, double *relres);

/* LDA exchange-correlation potential <- ExchangeCorrelationPotential(DS_SQ*)
 * (src/scf.cpp:19-64, inc/func_sq.h:309): n^3 values of rho in, Vxc out. */
int sqg_vxc_lda(sqg_ctx *ctx, const double *rho, double *vxc);

/* Total energy per atom <- EvaluateTotalEnergy(DS_SQ*) (src/energy.cpp:19-87,
 * inc/func_sq.h:521): phi and vxc of the potential the SQ call used, rho its
 * output, ebs/eent from sqg_density, eself/ecorr from the pseudocharge. */
int sqg_total_energy(sqg_ctx *ctx, const double *phi, const double *b, const double *rho, const double *vxc,
                     double ebs, double eent, double eself, double ecorr, int n_atm, double *etot);

/* The pseudopotential data and rb-replicas ForcesLocal loops over. */
typedef struct sqg_local_spec {
    int n_typ;
    const int *size;              /* per type: radial grid length (>= 2) */
    const double *const *r;       /* per type: radial grid, increasing from 0 */
    const double *const *rvloc;   /* per type: r V_loc(r) */
    const double *const *dvloc;   /* per type: its spline derivative table (getYD_gen) */
    const double *Z;              /* per type: valence charge */
    int n_img;                    /* replicas (sqdft_system_rb_images) */
    const double *img_xyz;        /* 3 per replica (bohr) */
    const int *img_type, *img_atom;
    const int *img_box;           /* 6 per replica: intersection lo x,y,z, hi x,y,z */
    int n_atm;
    const double *const *uu;      /* per type: atomic valence density uu(r) (sqg_pseudocharge only) */
    const double *const *duu;     /* per type: its spline derivative table (sqg_pseudocharge only) */
} sqg_local_spec;

/* Local forces <- ForcesLocal(DS_SQ*) (src/forces.cpp:20-315, inc/func_sq.h:534)
 * on one rank: phi from the last Poisson solve, b, b_tilda, vc
 * from the pseudocharge (n^3 each, x fastest).  floc and fcorr get 3*n_atm
 * values laid out fx[n_atm], fy[n_atm], fz[n_atm]; ForcesTotal
 * (src/forces.cpp:613-619) adds floc + fnloc + fcorr.  SQG_E_ARG for an
 * inconsistent spec, SQG_E_SPLINE when a distance leaves a radial grid. */
int sqg_forces_local(sqg_ctx *ctx, const sqg_local_spec *spec, const double *phi, const double *b,
                     const double *b_tilda, const double *vc, double *floc, double *fcorr);

/* Pseudocharge and starting density <- ChargeDensity(DS_SQ*)
 * (src/initialize.cpp:471-784, first MD step, one rank; inc/func_sq.h:
 * ChargeDensity) over the replicas of spec (uu, duu required): b, rho =
 * rho_at scaled to -Ncharge electrons, b_tilda (scaled to Ncharge) and vc
 * (shifted) as n^3 arrays, and ncharge, eself, ecorr -- the outputs of
 * sqdft_system_pseudocharge, computed on the device, and rho_at, the unscaled
 * superposition of the atomic densities (what ChargeExtrapolation_MD
 * subtracts, src/md.cpp:185); any output may be NULL. */
int sqg_pseudocharge(sqg_ctx *ctx, const sqg_local_spec *spec, double *b, double *rho, double *ncharge,
                     double *eself, double *ecorr, double *b_tilda, double *vc, double *rho_at);

#ifdef __cplusplus
}
#endif
#endif
