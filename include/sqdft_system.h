/*
 * sqdft_system.h -- host-side system setup for the SQ hot path (C-ABI).
 *
 * Replaces the reference's geometry/pseudopotential preparation that feeds
 * the spectral-quadrature kernels (phanish-suryanarayana/SQDFT):
 *   sqdft_system_create      <- Read_psd (src/readfiles.cpp:605-1225),
 *                               FD coefficients (src/readfiles.cpp:370-406),
 *                               Replica_atoms_Rcut (src/initialize.cpp:1003-1196,
 *                               proto inc/func_sq.h:54) and
 *                               NonlocalProjectors (src/nonlocal.cpp:317-518,
 *                               proto inc/func_sq.h:368)
 *   sqdft_system_from_files  <- Read_input + Read_atoms (src/readfiles.cpp:86-602)
 *
 * Conventions: plain pointers and sizes, caller-owned buffers, int status
 * (0 = OK, negative = error, see SQDFT_E_*), no exit().  All arithmetic is
 * FP64 and unfused so that the tables match the reference bit for bit.
 */
#ifndef SQDFT_SYSTEM_H
#define SQDFT_SYSTEM_H

#ifdef __cplusplus
extern "C" {
#endif

#define SQDFT_OK 0
#define SQDFT_E_ARG        -1
#define SQDFT_E_IO         -2
#define SQDFT_E_SPLINE     -3   /* spline.cpp:17-22 */
#define SQDFT_E_DENOM      -4   /* nonlocal.cpp:163 */
#define SQDFT_E_MESH       -5   /* readfiles.cpp:298-303 (anisotropic mesh) */

#define SQDFT_MAX_TYPES 8

/* One atom type: the ".atoms" "Atoms" block (readfiles.cpp:481-510). */
typedef struct sqdft_type_spec {
    const char *symbol;      /* must match the first token of the .pot file */
    int Z;                   /* valence charge */
    int lloc;                /* local channel */
    const char *psd_path;    /* Troullier-Martins table in the `atom` format */
    int natm;                /* atoms of this type */
    const double *xyz;       /* 3*natm Cartesian coordinates (bohr), x,y,z per atom */
    double rb;               /* pseudocharge radius (".atoms" rb); 0 -> no pseudocharge */
} sqdft_type_spec;

/* Grid and SQ parameters (".input" keywords, readfiles.cpp:86-316). */
typedef struct sqdft_spec {
    int n;                   /* FD intervals per direction of the whole periodic cube */
    double L;                /* cube edge (bohr) */
    double h;                /* mesh size; 0 -> L/n.  The reference computes it per unit
                                cell (domain/n_int) before multiplying by ncell
                                (readfiles.cpp:294-323), so pass that value for
                                bit-exact parity */
    int fd_order;            /* FD_order (12) */
    double Rcut;             /* truncation radius; nloc = ceil(Rcut/h) */
    int npl;                 /* Chebyshev degree (multiple of 4) */
    double T;                /* electronic temperature (K) */
    double lanczos_tol;
    double fermi_tol;
    int fermi_alg;           /* 1 Brent, 2 Newton */
    int n_typ;
    const sqdft_type_spec *types;
} sqdft_spec;

/* Derived scalars of a built system. */
typedef struct sqdft_info {
    int n, FDn, nloc, npl, fermi_alg, n_typ, n_atm, n_img;
    double L, h, Rcut, bet, lanczos_tol, fermi_tol;
    double coeff_lap[16], coeff_grad[16];
    long udv_len;            /* total doubles in the projector tables */
    int max_nproj;
} sqdft_info;

/* Per-type nonlocal data. */
typedef struct sqdft_type_info {
    int Z, lloc, lmax, natm, nlocc, nproj;
    double rc, rz;           /* grid-rounded radii (readfiles.cpp:1121-1124) */
    double denom[4];         /* Kleinman-Bylander denominators per l (carry 1/h^3) */
    int size;                /* radial grid length (incl. r = 0) */
    double rb;               /* pseudocharge radius as given (bohr); the replicas use
                                ceil(rb/h)*h (src/readfiles.cpp:512) */
} sqdft_type_info;

typedef struct sqdft_system sqdft_system;

int  sqdft_system_create(const sqdft_spec *spec, sqdft_system **out);
/* Reads "<dir>/<stem>.input" and the .atoms file it names, with the
 * reference's keyword semantics (domain and n_int are per unit cell and are
 * multiplied by ncell; perturb must be 0: displaced coordinates are written
 * explicitly instead of drawn from rand()). */
int  sqdft_system_from_files(const char *dir, const char *stem, sqdft_system **out);
void sqdft_system_destroy(sqdft_system *S);

int  sqdft_system_info(const sqdft_system *S, sqdft_info *info);
int  sqdft_system_type_info(const sqdft_system *S, int t, sqdft_type_info *ti);
/* Radial tables of type t: r, rV[l] (Ha), rU[l], rVloc, uu (each `size` long);
 * any pointer may be NULL. */
int  sqdft_system_radial(const sqdft_system *S, int t, double *r, double *rV, double *rU,
                         double *rVloc, double *uu);
/* Atom coordinates in type order: xyz[3*n_atm]; type[n_atm]. */
int  sqdft_system_atoms(const sqdft_system *S, double *xyz, int *type);
/* Periodic images whose rc-cube meets the (domain + Rcut) box, in the
 * reference's order (type, atom, z-shift, y-shift, x-shift):
 * xyz[3*n_img], type[n_img], atom[n_img] (global index in type order),
 * nn[3*n_img] nearest node (left-tie rule), udv_off[n_img]. */
int  sqdft_system_images(const sqdft_system *S, double *xyz, int *type, int *atom, int *nn,
                         long *udv_off);
/* Projector tables U_l Y_lm (V_l - V_loc) on each image's (2 nlocc+1)^3 cube,
 * projector-major then z,y,x (udv_len doubles). */
int  sqdft_system_udv(const sqdft_system *S, double *udv);

/* Pseudocharge and starting density of the SCF (ChargeDensity,
 * src/initialize.cpp:471-784, over the rb images of Replica_atoms,
 * src/initialize.cpp:21-150; first MD step, Correction 1).  Outputs, any of
 * which may be NULL: b[n^3] (x fastest), rho[n^3] = rho_at scaled to hold
 * -Ncharge electrons, ncharge = sum b h^3 (negative), eself, and ecorr,
 * the energy correction (the reference prints it divided by the atom count),
 * b_tilda[n^3] (the smooth reference charge scaled to Ncharge) and vc[n^3]
 * (V~ - V shifted as initialize.cpp:752-775), the inputs of ForcesLocal.
 * Returns SQDFT_E_ARG when a type has rb <= 0. */
int  sqdft_system_pseudocharge(const sqdft_system *S, double *b, double *rho, double *ncharge,
                               double *eself, double *ecorr, double *b_tilda, double *vc);

/* Periodic replicas whose rb-box (half-width round(rb'/h) nodes around the
 * nearest node, rb' = ceil(rb/h)*h as src/readfiles.cpp:512 rounds it)
 * meets the domain [0, n-1]^3, in Replica_atoms' order (type, atom,
 * z-shift, y-shift, x-shift; src/initialize.cpp:21-200).  Call with
 * NULL arrays to get the count in *n_img.  xyz[3*n_img], type[n_img],
 * atom[n_img] (global index in type order), box[6*n_img] = the intersection
 * lo x,y,z then hi x,y,z (inclusive node indices). */
int  sqdft_system_rb_images(const sqdft_system *S, int *n_img, double *xyz, int *type, int *atom,
                            int *box);
/* Moves the atoms (xyz[3*n_atm], type order, as sqdft_system_atoms) and
 * rebuilds the images and projector tables for the new positions (what
 * Replica_atoms_Rcut + NonlocalProjectors redo every MD step, src/scf.cpp:349-350). */
int  sqdft_system_set_positions(sqdft_system *S, const double *xyz);
/* Spline derivative table of type t's r V_loc(r) (getYD_gen, src/spline.cpp:57-126),
 * `size` doubles: with sqdft_system_radial's r and rVloc it defines V_loc. */
int  sqdft_system_vloc_deriv(const sqdft_system *S, int t, double *d);
/* The same for the atomic valence density uu(r) (ChargeDensity's Duu,
 * src/initialize.cpp:511-512). */
int  sqdft_system_uu_deriv(const sqdft_system *S, int t, double *d);

/* Poisson-solver and SCF parameters (".input" poisson_tol, poisson_maxiter,
 * beta_aaj, m_aaj, p_aaj, scf_tol, scf_miniter, scf_maxiter, beta_scf,
 * m_scf, p_scf; defaults 1e-6, 1000, 0.6, 7, 6, 1e-4, 3, 40, 0.1, 7, 2 for
 * systems built from a spec). */
typedef struct sqdft_scf_params {
    double tol;               /* Poisson (AAJ) */
    int maxiter;
    double beta;
    int m, p;
    double scf_tol;           /* SCF loop and Periodic Pulay mixing of Veff */
    int scf_miniter, scf_maxiter;
    double beta_scf;
    int m_scf, p_scf;
} sqdft_scf_params;
int  sqdft_system_scf_params(const sqdft_system *S, sqdft_scf_params *pp);
/* Cubic-spline evaluation with the reference's routine (tests). */
int  sqdft_spline_eval(const double *X, const double *Y, const double *YD, int len,
                       const double *x, double *y, int nx);
int  sqdft_spline_derivs(const double *X, const double *Y, double *YD, int len);

#ifdef __cplusplus
}
#endif
#endif
