/*
 * sq_oracle.h -- FP64 CPU oracle for the SQDFT spectral-quadrature hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (phanish-suryanarayana/SQDFT, src/sq.cpp, src/nonlocal.cpp,
 * src/spline.cpp, src/forces.cpp:318-578, src/readfiles.cpp:370-406 and
 * 605-1225, src/initialize.cpp:1003-1196).  It exists so that tests, the
 * smoke check and the CPU-baseline leg of the benchmark can check the HIP
 * product path.  Nothing in the product (sqdft_amd/) links or calls it.
 *
 * Every function cites the reference lines it follows.  Error paths that
 * call exit() in the reference return a negative status here instead.
 */
#ifndef SQ_ORACLE_H
#define SQ_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (mirror the reference's exit paths) */
#define SQO_OK 0
#define SQO_ERR_NPL_NOT_MULT4   -1   /* sq.cpp:523-524 */
#define SQO_ERR_NEG_RHO         -2   /* sq.cpp:764-765 */
#define SQO_ERR_QL_ITER         -3   /* sq.cpp:1125 */
#define SQO_ERR_BRENT_BRACKET   -4   /* sq.cpp:1307-1311 */
#define SQO_ERR_BRENT_ITMAX     -5   /* sq.cpp:1400-1401 */
#define SQO_ERR_DENOM_ZERO      -6   /* nonlocal.cpp:163 */
#define SQO_ERR_SPLINE_RANGE    -7   /* spline.cpp:17-22 */
#define SQO_ERR_IO              -8
#define SQO_ERR_ARG             -9
#define SQO_WARN_LANCZOS_MAXIT   1   /* sq.cpp:1073-1074 (warning only) */

/* ---------------- small math pieces ---------------- */
double sqo_fract(int n, int k);                                    /* readfiles.cpp:394-406 */
void   sqo_fd_coeffs(int FDn, double h, double *lap, double *grad); /* readfiles.cpp:370-389 */
void   sqo_spline_yd(const double *X, const double *Y, double *YD, int len); /* spline.cpp:57-126 */
int    sqo_spline_eval(const double *X, const double *Y, int len,
                       const double *x, double *y, int nx, const double *YD); /* spline.cpp:11-55 */
double sqo_sph_harm(double x, double y, double z, double r, int l, int m);   /* nonlocal.cpp:218-314 */
int    sqo_tridiag_eig(const double *diag, const double *sub, int n,
                       double *lmin, double *lmax);                           /* sq.cpp:1094-1180 */
/* func: 0 FermiDirac, 1 UbandFunc, 2 EentFunc, 3 FermiDirac_derv (sq.cpp:1261-1287,1451-1457) */
double sqo_func(int func, double t, double lf, double bet);
void   sqo_cheb_coeff(int npl, double lfp, double bet, double *C, int func); /* sq.cpp:1230-1259 */

/* ---------------- pseudopotential (readfiles.cpp:605-1225) ---------------- */
typedef struct {
    int size;           /* radial points incl. r=0 */
    double *r;          /* RadialGrid */
    double *rV[4];      /* rVs rVp rVd rVf (Ha) */
    double *rU[4];      /* rUs rUp rUd rUf */
    double *rVloc;
    double *uu;         /* valence charge / (4 pi r^2) */
    double *DV[4], *DU[4], *DVloc;  /* spline derivative tables */
    double rc_l[4];
    int lmax;
    double rc, rz;      /* rounded to the grid as in readfiles.cpp:1122-1124 */
    double Denom[4];
} sqo_psd;

int  sqo_read_psd(const char *path, const char *symb, int lloc, int Z, double h, sqo_psd *P);
void sqo_free_psd(sqo_psd *P);

/* ---------------- system ---------------- */
typedef struct {
    /* grid / SQ parameters */
    int n;              /* nodes per direction (periodic cube) */
    double L;           /* domain length */
    double h;           /* mesh size */
    int FDn;
    double coeff_lap[16], coeff_grad[16];
    int nloc;
    double Rcut;
    int npl;
    double bet;
    double lanczos_tol, fermi_tol;
    int fermi_alg;
    /* types */
    int n_typ;
    int Z[8], lloc[8], natm[8];
    sqo_psd psd[8];
    int nlocc[8], nproj[8];
    /* atoms: positions by type, concatenated in type order */
    int n_atm;
    double *ax, *ay, *az;
    /* periodic images whose rc-cube meets the domain+Rcut box (initialize.cpp:1003-1196) */
    int n_img;
    double *ix, *iy, *iz;
    int *ityp, *iatm, *inn;    /* inn: 3 per image (nearest node) */
    long *iudv;                /* offset of the image's projector table in udv */
    double *udv;               /* per image: nproj x (2 nlocc+1)^3 */
    long udv_len;
} sqo_sys;

/* Builds a system: reads pseudopotentials, images and projector cubes.
 * atom coordinates are given per type (concatenated in type order). */
/* h <= 0 -> L/n.  The reference computes h per unit cell (domain/n_int,
 * readfiles.cpp:294-303) before scaling domain and n_int by ncell. */
int  sqo_sys_build(sqo_sys *S, int n, double L, double h, int FDn, double Rcut, int npl, double T,
                   double lanczos_tol, double fermi_tol, int fermi_alg,
                   int n_typ, const char *const *psd_paths, const char *const *symbs,
                   const int *Z, const int *lloc, const int *natm,
                   const double *ax, const double *ay, const double *az);
void sqo_sys_free(sqo_sys *S);

/* nearest node with the left-tie rule (nonlocal.cpp:48-64) */
void sqo_nearest_node(double x, double y, double z, double h, int *ijk);

/* H_i v on node (i,j,k)'s cube; veff is the periodic n^3 grid, x fastest
 * (sq.cpp:838-929 + nonlocal.cpp:20-215).  v, hv have (2 nloc+1)^3 entries,
 * z-major ([kk][jj][ii]). */
int sqo_hsub(const sqo_sys *S, const double *veff, int node, const double *v, double *hv);
/* V_nl restricted to one atom (all its images), nonlocal.cpp:522-724 */
int sqo_vnl_atom(const sqo_sys *S, int node, int atom, const double *v, double *vv);

/* Lanczos bounds for one node (sq.cpp:965-1092); returns iteration count */
int sqo_lanczos(const sqo_sys *S, const double *veff, int node, double *lmin, double *lmax,
                int *iters);
/* Lanczos + Chebyshev moments for one node (sq.cpp:530-661) */
int sqo_node_moments(const sqo_sys *S, const double *veff, int node,
                     double *chi, double *zee, double *lmin, double *lmax, double *mu, int *iters);

/* The per-node loop body (Lanczos + moments) over a list of nodes, OpenMP
 * over nodes (sq.cpp:530-661).  Outputs are indexed by list position. */
int sqo_moments_nodes(const sqo_sys *S, const double *veff, const int *nodes, int count,
                      double *chi, double *zee, double *mu, int *iters, int nthreads);

/* Fermi constraint and root solvers (sq.cpp:1183-1228, 1289-1490) */
double sqo_ne_constraint(const sqo_sys *S, int nn, const double *chi, const double *zee,
                         const double *mu, double ncharge, double lf);
int sqo_brent(const sqo_sys *S, int nn, const double *chi, const double *zee, const double *mu,
              double ncharge, double x1, double x2, int scf_iter, double lf_prev,
              double *root, int *iters);
int sqo_newton(const sqo_sys *S, int nn, const double *chi, const double *zee, const double *mu,
               double ncharge, double lf0, double *root, int *iters);

/* The whole SQ density step for all n^3 nodes (sq.cpp:377-834).
 * Outputs may be NULL except rho/lambda_f.  nthreads <= 0 uses OpenMP default. */
int sqo_density(const sqo_sys *S, const double *veff, int scf_iter, double ncharge,
                double *lambda_f, double *rho, double *ebs, double *eent,
                double *chi, double *zee, double *mu, double *Ci,
                int *lanczos_iters, int *fermi_iters, int nthreads);

/* rho/Ebs/Eent assembly from given moments (sq.cpp:740-797) */
int sqo_rho_energy(const sqo_sys *S, int nn, const double *chi, const double *zee,
                   const double *mu, double lambda_f, double *rho, double *ebs, double *eent,
                   double *Ci);

/* ---- derivatives with respect to Veff (no counterpart in the reference) ----
 * chi/zee are frozen (the Lanczos bounds are not differentiated).  The
 * moments follow sq.cpp:568-655 exactly (centres up to npl/2, the doubling
 * dot products above npl/4), and these functions differentiate that
 * computation.  fixed_n = 1 adds the Fermi-level shift that keeps the
 * electron count fixed. */
/* Chebyshev moments of one node with given chi, zee (sq.cpp:568-655). */
int sqo_moments_fixed(const sqo_sys *S, const double *veff, int node, double chi, double zee,
                      double *mu);
/* density step with frozen chi/zee for all nodes: moments, Brent, rho */
int sqo_density_fixed(const sqo_sys *S, const double *veff, const double *chi, const double *zee,
                      double ncharge, double lf_lo, double lf_hi, double *lambda_f, double *rho,
                      double *mu, int nthreads);
/* dL/dVeff for L = sum_i w[i] rho[i] (reverse mode through the recurrence) */
int sqo_adjoint_veff(const sqo_sys *S, const double *veff, const double *chi, const double *zee,
                     const double *mu, double lambda_f, const double *w, int fixed_n,
                     double *dLdV, int nthreads);
/* drho along dV (forward mode through the recurrence) */
int sqo_tangent_veff(const sqo_sys *S, const double *veff, const double *chi, const double *zee,
                     const double *mu, double lambda_f, const double *dV, int fixed_n,
                     double *drho, int nthreads);

/* Nonlocal forces (forces.cpp:318-578), uses chi, zee, Ci of the last SQ call.
 * f has 3*n_atm entries laid out fx[n_atm], fy[n_atm], fz[n_atm].
 * nodes==NULL -> all n^3 nodes; otherwise only the listed nodes contribute. */
int sqo_forces_nonlocal(const sqo_sys *S, const double *veff, const double *chi,
                        const double *zee, const double *Ci, const int *nodes, int n_nodes,
                        double *f, int nthreads);


/* AAJ Poisson solver (poisson.cpp:19-325 with anderson.cpp:20-163): phi of
 * -Lap_FD phi = 4 pi rhs on the periodic n^3 grid from phi_guess, shifted so
 * that phi(0,0,0) = 0; iters and the final relative residual out. */
int sqo_poisson_aaj(int n, int FDn, const double *lap, const double *rhs, const double *phi_guess, double tol,
                    int maxiter, double beta, int m, int p, double *phi, int *iters, double *relres);
/* LDA exchange-correlation potential of scf.cpp:19-64 (N grid values). */
void sqo_vxc_lda(size_t N, const double *rho, double *vxc);
/* Total energy per atom of EvaluateTotalEnergy (energy.cpp:19-87). */
double sqo_total_energy(size_t N, double h, const double *phi, const double *b, const double *rho,
                        const double *vxc, double Ebs, double Eent, double Eself, double Ecorr, int n_atm);

/* Pseudocharge of ChargeDensity (initialize.cpp:471-784, first MD step) over
 * the rb-replicas of Replica_atoms (initialize.cpp:21-200), one rank: b, b~
 * (scaled to Ncharge), Vc (shifted), rho = rho_at scaled to Ncharge electrons;
 * scal = {Ncharge, Eself, Ncharge~, Eself~, Ecorr}.  rb per type (bohr). */
int sqo_charge_density(const sqo_sys *S, const double *rb, double *b, double *bt, double *vc,
                       double *rho, double *scal);
/* ForcesLocal (forces.cpp:20-315): floc = sum_J grad(b_J) (phi - V_J) h^3 and
 * the correction fcorr of the same loop; both laid out fx[n_atm], fy, fz. */
int sqo_forces_local(const sqo_sys *S, const double *rb, const double *phi, const double *b,
                     const double *bt, const double *vc, double *floc, double *fcorr);
#ifdef __cplusplus
}
#endif
#endif
