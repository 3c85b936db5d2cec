/*
 * sq_oracle.c -- FP64 CPU oracle for the SQDFT spectral-quadrature hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see sq_oracle.h).  Plain-C restatement of
 * phanish-suryanarayana/SQDFT; each function cites the reference lines it
 * follows.  Build with -ffp-contract=off so that no multiply-add is fused:
 * the reference is built without FMA contraction (makefile:16, no -O, no
 * -march), so unfused arithmetic reproduces its rounding where the
 * evaluation order is the same.
 */
#include "sq_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TEMP_TOL 1e-12

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
static int wrapi(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }

/* ------------------------------------------------------------------ */
/* FD coefficients: readfiles.cpp:370-406                             */
/* ------------------------------------------------------------------ */
double sqo_fract(int n, int k)
{
    double Nr = 1, Dr = 1;
    for (int i = n - k + 1; i <= n; i++) Nr *= i;
    for (int i = n + 1; i <= n + k; i++) Dr *= i;
    return Nr / Dr;
}

void sqo_fd_coeffs(int FDn, double h, double *lap, double *grad)
{
    lap[0] = 0;
    for (int a = 1; a <= FDn; a++) lap[0] += -(2.0 / (a * a));
    lap[0] *= 1 / (h * h);
    for (int p = 1; p <= FDn; p++)
        lap[p] = (2 * pow(-1, p + 1) * sqo_fract(FDn, p) / (p * p * h * h));
    grad[0] = 0;
    for (int p = 1; p <= FDn; p++)
        grad[p] = (pow(-1, p + 1) * sqo_fract(FDn, p) / (p * h));
}

/* ------------------------------------------------------------------ */
/* Clamped cubic spline: spline.cpp:11-126                             */
/* ------------------------------------------------------------------ */
static void tridiag_solve(double *A, double *B, double *C, double *D, int len)
{   /* spline.cpp:99-126 */
    double *F = (double *)calloc((size_t)len, sizeof(double));
    double b = B[0];
    D[0] = D[0] / b;
    for (int i = 1; i < len; i++) {
        F[i] = C[i - 1] / b;
        b = B[i] - A[i] * F[i];
        D[i] = (D[i] - D[i - 1] * A[i]) / b;
    }
    for (int i = len - 2; i >= 0; i--) D[i] -= (D[i + 1] * F[i + 1]);
    free(F);
}

void sqo_spline_yd(const double *X, const double *Y, double *YD, int len)
{   /* spline.cpp:57-96 */
    double *A = (double *)calloc((size_t)len, sizeof(double));
    double *B = (double *)calloc((size_t)len, sizeof(double));
    double *C = (double *)calloc((size_t)len, sizeof(double));
    double h0 = X[1] - X[0], h1 = X[2] - X[1];
    double r0 = (Y[1] - Y[0]) / h0, r1 = (Y[2] - Y[1]) / h1;
    B[0] = h1 * (h0 + h1);
    C[0] = (h0 + h1) * (h0 + h1);
    YD[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0;
    int i;
    for (i = 1; i < len - 1; i++) {
        h0 = X[i] - X[i - 1]; h1 = X[i + 1] - X[i];
        r0 = (Y[i] - Y[i - 1]) / h0; r1 = (Y[i + 1] - Y[i]) / h1;
        A[i] = h1;
        B[i] = 2 * (h0 + h1);
        C[i] = h0;
        YD[i] = 3 * (r0 * h1 + r1 * h0);
    }
    A[i] = (h0 + h1) * (h0 + h1);
    B[i] = h0 * (h0 + h1);
    YD[i] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0);
    tridiag_solve(A, B, C, YD, len);
    free(A); free(B); free(C);
}

int sqo_spline_eval(const double *X1, const double *Y1, int len1,
                    const double *X2, double *Y2, int len2, const double *YD)
{   /* spline.cpp:11-55 */
    if (X2[0] < X1[0] || X2[len2 - 1] > X1[len1 - 1]) return SQO_ERR_SPLINE_RANGE;
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0, p1 = 0, p3 = X2[0] - 1;
    int j = 0;
    for (int i = 0; i < len2; i++) {
        double p2 = X2[i];
        if (p2 > p3) {
            for (; j < len1 && p2 > X1[j]; j++) ;
            if (p2 < X1[j]) j--;
            if (j > len1 - 2) j = len1 - 2; /* p2 on the last knot */
            p1 = X1[j];
            p3 = X1[j + 1];
            double dx = 1.0 / (X1[j + 1] - X1[j]);
            double dy = (Y1[j + 1] - Y1[j]) * dx;
            A0 = Y1[j];
            A1 = YD[j];
            A2 = dx * (3.0 * dy - 2.0 * YD[j] - YD[j + 1]);
            A3 = dx * dx * (-2.0 * dy + YD[j] + YD[j + 1]);
        }
        double x = p2 - p1;
        Y2[i] = ((A3 * x + A2) * x + A1) * x + A0;
    }
    return SQO_OK;
}

/* ------------------------------------------------------------------ */
/* Real spherical harmonics: nonlocal.cpp:218-314                      */
/* ------------------------------------------------------------------ */
double sqo_sph_harm(double x, double y, double z, double r, int l, int m)
{
    /* coefficients are the closed forms named in the reference comments,
       written out to the same 15 digits */
    const double C00 = 0.282094791773878;   /* 0.5 sqrt(1/pi)  */
    const double C1 = 0.488602511902920;    /* sqrt(3/(4 pi))  */
    const double C2a = 1.092548430592079;   /* 0.5 sqrt(15/pi) */
    const double C20 = 0.315391565252520;   /* 0.25 sqrt(5/pi) */
    const double C2p2 = 0.546274215296040;  /* 0.25 sqrt(15/pi) */
    const double C3m3 = 0.590043589926644;  /* 0.25 sqrt(35/(2pi)) */
    const double C3m2 = 2.890611442640554;  /* 0.5 sqrt(105/pi) */
    const double C3m1 = 0.457045799464466;  /* 0.25 sqrt(21/(2pi)) */
    const double C30 = 0.373176332590115;   /* 0.25 sqrt(7/pi) */
    const double C3p2 = 1.445305721320277;  /* 0.25 sqrt(105/pi) */
    double SH = 0.0;
    if (l == 0) SH = C00;
    if (fabs(r) > TEMP_TOL) {
        if (l == 0) SH = C00;
        else if (l == 1) {
            if (m == -1) SH = C1 * (y / r);
            else if (m == 0) SH = C1 * (z / r);
            else SH = C1 * (x / r);
        } else if (l == 2) {
            if (m == -2) SH = C2a * (x * y) / (r * r);
            else if (m == -1) SH = C2a * (y * z) / (r * r);
            else if (m == 0) SH = C20 * (-x * x - y * y + 2 * z * z) / (r * r);
            else if (m == 1) SH = C2a * (z * x) / (r * r);
            else SH = C2p2 * (x * x - y * y) / (r * r);
        } else if (l == 3) {
            if (m == -3) SH = C3m3 * (3 * x * x - y * y) * y / (r * r * r);
            else if (m == -2) SH = C3m2 * (x * y * z) / (r * r * r);
            else if (m == -1) SH = C3m1 * y * (4 * z * z - x * x - y * y) / (r * r * r);
            else if (m == 0) SH = C30 * z * (2 * z * z - 3 * x * x - 3 * y * y) / (r * r * r);
            else if (m == 1) SH = C3m1 * x * (4 * z * z - x * x - y * y) / (r * r * r);
            else if (m == 2) SH = C3p2 * z * (x * x - y * y) / (r * r * r);
            else SH = C3m3 * x * (x * x - 3 * y * y) / (r * r * r);
        }
    }
    return SH;
}

/* ------------------------------------------------------------------ */
/* Implicit-shift QL on a symmetric tridiagonal: sq.cpp:1094-1180      */
/* ------------------------------------------------------------------ */
int sqo_tridiag_eig(const double *diag, const double *sub, int n, double *lmin, double *lmax)
{
    double *d = (double *)malloc((size_t)n * sizeof(double));
    double *e = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) { d[i] = diag[i]; e[i] = sub[i]; }
    e[n - 1] = 0.0;
    int m, i;
    double s, r, p, g, f, dd, c, b;
    for (int l = 0; l <= n - 1; l++) {
        int iter = 0;
        do {
            for (m = l; m <= n - 2; m++) {
                dd = fabs(d[m]) + fabs(d[m + 1]);
                if ((double)(fabs(e[m]) + dd) == dd) break;
            }
            if (m != l) {
                if (iter++ == 200) { free(d); free(e); return SQO_ERR_QL_ITER; }
                g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                r = sqrt(g * g + 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                s = c = 1.0;
                p = 0.0;
                for (i = m - 1; i >= l; i--) {
                    f = s * e[i];
                    b = c * e[i];
                    e[i + 1] = (r = sqrt(g * g + f * f));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
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
    free(d); free(e);
    return SQO_OK;
}

/* ------------------------------------------------------------------ */
/* Functions interpolated by the Chebyshev coefficients                */
/* sq.cpp:1261-1287, 1451-1457                                         */
/* ------------------------------------------------------------------ */
double sqo_func(int func, double t, double lf, double bet)
{
    double v;
    switch (func) {
    case 0: return 1 / (1 + exp(bet * (t - lf)));
    case 1: return t * (1 / (1 + exp(bet * (t - lf))));
    case 2: {
        double focc = (1 / (1 + exp(bet * (t - lf))));
        if (fabs(focc) < 0.01 * TEMP_TOL || fabs(focc - 1.0) < 0.01 * TEMP_TOL) v = 0.0;
        else v = (focc * log(focc) + (1 - focc) * log(1 - focc));
        return v;
    }
    default:
        v = 1 / (1 + exp(bet * (t - lf)));
        return v * (1 - v);
    }
}

void sqo_cheb_coeff(int npl, double lfp, double bet, double *C, int func)
{   /* sq.cpp:1230-1259: interpolation at the npl+1 Chebyshev-Gauss nodes */
    double *d = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    for (int k = 0; k < npl + 1; k++) {
        double y = cos(M_PI * ((double)((k - 0.5 + 1) / (npl + 1))));
        d[k] = sqo_func(func, y, lfp, bet);
    }
    double fac = 2.0 / (npl + 1);
    for (int j = 0; j < npl + 1; j++) {
        double sum = 0.0;
        for (int k = 0; k < npl + 1; k++)
            sum = sum + d[k] * cos((M_PI * (j - 1 + 1)) * ((double)((k - 0.5 + 1) / (npl + 1))));
        C[j] = fac * sum;
    }
    C[0] = C[0] / 2.0;
    free(d);
}

/* ------------------------------------------------------------------ */
/* Pseudopotential reader: readfiles.cpp:605-1225                      */
/* ------------------------------------------------------------------ */
static int read_line(FILE *f, char *buf, int n)
{
    return fgets(buf, n, f) != NULL;
}

static int starts(const char *line, const char *key)
{
    return strncmp(line, key, strlen(key)) == 0;
}

/* reads nvals numbers with fscanf semantics into dst[1..nvals] */
static int read_vals(FILE *f, double *dst, int nvals, double scale)
{
    for (int i = 1; i <= nvals; i++) {
        double v;
        if (fscanf(f, "%lf", &v) != 1) return SQO_ERR_IO;
        dst[i] = scale == 1.0 ? v : scale * v;
    }
    return SQO_OK;
}

int sqo_read_psd(const char *path, const char *symb, int lloc, int Z, double h, sqo_psd *P)
{
    memset(P, 0, sizeof(*P));
    FILE *f = fopen(path, "r");
    if (!f) return SQO_ERR_IO;
    char tok[256], line[4096];
    if (fscanf(f, "%255s", tok) != 1 || strcmp(tok, symb) != 0) { fclose(f); return SQO_ERR_IO; }
    /* count grid values: tokens after the "Radial grid follows" line up to
       and including the word "Pseudopotential" (readfiles.cpp:634-644) */
    while (read_line(f, line, sizeof line))
        if (starts(line, " Radial grid follows")) break;
    int count = 0;
    do {
        if (fscanf(f, "%255s", tok) != 1) { fclose(f); return SQO_ERR_IO; }
        count++;
    } while (strcmp(tok, "Pseudopotential") != 0);
    P->size = count;
    int S = count;
    P->r = (double *)calloc((size_t)S, sizeof(double));
    P->rVloc = (double *)calloc((size_t)S, sizeof(double));
    P->uu = (double *)calloc((size_t)S, sizeof(double));
    P->DVloc = (double *)calloc((size_t)S, sizeof(double));
    for (int l = 0; l < 4; l++) {
        P->rV[l] = (double *)calloc((size_t)S, sizeof(double));
        P->rU[l] = (double *)calloc((size_t)S, sizeof(double));
        P->DV[l] = (double *)calloc((size_t)S, sizeof(double));
        P->DU[l] = (double *)calloc((size_t)S, sizeof(double));
    }
    rewind(f);
    while (read_line(f, line, sizeof line))
        if (starts(line, " Radial grid follows")) break;
    P->r[0] = 0;
    if (read_vals(f, P->r, S - 1, 1.0)) { fclose(f); return SQO_ERR_IO; }
    /* blocks are identified by their header lines */
    while (read_line(f, line, sizeof line)) {
        if (starts(line, " Pseudopotential follows")) {
            unsigned l;
            if (fscanf(f, "%u", &l) != 1 || l > 3) { fclose(f); return SQO_ERR_IO; }
            P->lmax = (int)l;
            if (read_vals(f, P->rV[l], S - 1, 0.5)) { fclose(f); return SQO_ERR_IO; } /* Ry -> Ha */
            P->rV[l][0] = P->rV[l][1];
            if ((int)l == lloc) memcpy(P->rVloc, P->rV[l], (size_t)S * sizeof(double));
        } else if (starts(line, " Valence charge follows")) {
            for (int ir = 1; ir < S; ir++) {
                double v;
                if (fscanf(f, "%lf", &v) != 1) { fclose(f); return SQO_ERR_IO; }
                P->uu[ir] = v / (4 * M_PI * P->r[ir] * P->r[ir]);
            }
            P->uu[0] = P->uu[1];
        } else if (starts(line, " Pseudo-wave-function follows")) {
            int l; double zel, rc;
            if (fscanf(f, "%d %lf %lf", &l, &zel, &rc) != 3 || l < 0 || l > 3) {
                fclose(f); return SQO_ERR_IO;
            }
            P->rc_l[l] = rc;
            if (read_vals(f, P->rU[l], S - 1, 1.0)) { fclose(f); return SQO_ERR_IO; }
            P->rU[l][0] = P->rU[l][1];
        }
    }
    fclose(f);
    /* rz: first grid radius where rVloc reaches -Z (readfiles.cpp:1113-1119) */
    int ir = 0;
    while (ir < S && fabs(P->rVloc[ir] + (double)Z) > TEMP_TOL) ir++;
    P->rz = ir > 0 ? P->r[ir - 1] : 0.0;
    double rc = fmax(fmax(P->rc_l[0], P->rc_l[1]), fmax(P->rc_l[2], P->rc_l[3]));
    P->rc = ceil(rc / h) * h;
    P->rz = (ceil(P->rz / h)) * h;
    /* spline derivative tables and KB denominators (readfiles.cpp:1126-1216) */
    sqo_spline_yd(P->r, P->rVloc, P->DVloc, S);
    double dr = 1e-4;
    int nr = 1 + (int)ceil((P->rc) / dr);
    for (int l = 0; l < 4; l++) {
        sqo_spline_yd(P->r, P->rU[l], P->DU[l], S);
        sqo_spline_yd(P->r, P->rV[l], P->DV[l], S);
        double den = 0.0;
        for (int k = 0; k < nr; k++) {
            double rr = k * dr, Urr, Vrr, Vlrr;
            sqo_spline_eval(P->r, P->rU[l], S, &rr, &Urr, 1, P->DU[l]);
            sqo_spline_eval(P->r, P->rV[l], S, &rr, &Vrr, 1, P->DV[l]);
            sqo_spline_eval(P->r, P->rVloc, S, &rr, &Vlrr, 1, P->DVloc);
            if (rr != 0) den += Urr * (Vrr - Vlrr) * Urr * dr / (rr * h * h * h);
        }
        P->Denom[l] = den;
    }
    return SQO_OK;
}

void sqo_free_psd(sqo_psd *P)
{
    free(P->r); free(P->rVloc); free(P->uu); free(P->DVloc);
    for (int l = 0; l < 4; l++) { free(P->rV[l]); free(P->rU[l]); free(P->DV[l]); free(P->DU[l]); }
    memset(P, 0, sizeof(*P));
}

/* ------------------------------------------------------------------ */
/* Nearest node with the half-spacing left-tie rule: nonlocal.cpp:48-64 */
/* ------------------------------------------------------------------ */
static int nn1(double x, double h)
{
    int i = (int)round(x / h);
    if (fabs(fabs(i * h - x) - h * 0.5) < TEMP_TOL) i = (int)floor((x / h) - 10 * TEMP_TOL);
    return i;
}

void sqo_nearest_node(double x, double y, double z, double h, int *ijk)
{
    ijk[0] = nn1(x, h); ijk[1] = nn1(y, h); ijk[2] = nn1(z, h);
}

/* ------------------------------------------------------------------ */
/* System: images (initialize.cpp:1003-1196), projectors               */
/* (nonlocal.cpp:317-518)                                               */
/* ------------------------------------------------------------------ */
static int type_nproj(int lmax, int lloc)
{
    int c = 0;
    for (int l = 0; l <= lmax; l++) if (l != lloc) c += 2 * l + 1;
    return c;
}

int sqo_sys_build(sqo_sys *S, int n, double L, double h, int FDn, double Rcut, int npl, double T,
                  double lanczos_tol, double fermi_tol, int fermi_alg,
                  int n_typ, const char *const *psd_paths, const char *const *symbs,
                  const int *Z, const int *lloc, const int *natm,
                  const double *ax, const double *ay, const double *az)
{
    memset(S, 0, sizeof(*S));
    if (n_typ < 1 || n_typ > 8 || FDn < 1 || FDn > 15) return SQO_ERR_ARG;
    S->n = n; S->L = L; S->h = h > 0 ? h : L / n; S->FDn = FDn;
    /* readfiles.cpp:294-316 */
    S->nloc = (int)ceil(Rcut / S->h);
    S->Rcut = S->nloc * S->h;
    S->npl = npl;
    double kB = 8.617343e-5 / 27.211384523;
    S->bet = (1 / (kB * T));
    S->lanczos_tol = lanczos_tol; S->fermi_tol = fermi_tol; S->fermi_alg = fermi_alg;
    sqo_fd_coeffs(FDn, S->h, S->coeff_lap, S->coeff_grad);
    S->n_typ = n_typ;
    int na = 0;
    for (int t = 0; t < n_typ; t++) {
        S->Z[t] = Z[t]; S->lloc[t] = lloc[t]; S->natm[t] = natm[t];
        int st = sqo_read_psd(psd_paths[t], symbs[t], lloc[t], Z[t], S->h, &S->psd[t]);
        if (st) return st;
        S->nlocc[t] = (int)ceil(S->psd[t].rc / S->h);
        S->nproj[t] = type_nproj(S->psd[t].lmax, lloc[t]);
        na += natm[t];
    }
    S->n_atm = na;
    S->ax = (double *)malloc((size_t)na * sizeof(double));
    S->ay = (double *)malloc((size_t)na * sizeof(double));
    S->az = (double *)malloc((size_t)na * sizeof(double));
    memcpy(S->ax, ax, (size_t)na * sizeof(double));
    memcpy(S->ay, ay, (size_t)na * sizeof(double));
    memcpy(S->az, az, (size_t)na * sizeof(double));

    /* images: two passes (count, then fill) exactly as the reference loops */
    for (int pass = 0; pass < 2; pass++) {
        int cnt = 0, a0 = 0;
        long off = 0;
        for (int t = 0; t < n_typ; t++) {
            double rc = S->psd[t].rc;
            int nc = (int)ceil(rc / S->h);
            for (int ja = 0; ja < natm[t]; ja++) {
                int a = a0 + ja;
                double x = ax[a], y = ay[a], z = az[a];
                int px = (int)floor(((L + S->Rcut + rc - x) / L) + TEMP_TOL);
                int py = (int)floor(((L + S->Rcut + rc - y) / L) + TEMP_TOL);
                int pz = (int)floor(((L + S->Rcut + rc - z) / L) + TEMP_TOL);
                int mx = (int)ceil(((-(S->Rcut + rc) - x) / L) - TEMP_TOL);
                int my = (int)ceil(((-(S->Rcut + rc) - y) / L) - TEMP_TOL);
                int mz = (int)ceil(((-(S->Rcut + rc) - z) / L) - TEMP_TOL);
                for (int rr = mz; rr <= pz; rr++)
                    for (int qq = my; qq <= py; qq++)
                        for (int pp = mx; pp <= px; pp++) {
                            double X = x + pp * L, Y = y + qq * L, Zc = z + rr * L;
                            int ijk[3];
                            sqo_nearest_node(X, Y, Zc, S->h, ijk);
                            int ok = 1;
                            for (int d = 0; d < 3; d++) {
                                int lo = ijk[d] - nc, hi = ijk[d] + nc;
                                if (imax(lo, -S->nloc) > imin(hi, n - 1 + S->nloc)) ok = 0;
                            }
                            if (!ok) continue;
                            if (pass == 1) {
                                S->ix[cnt] = X; S->iy[cnt] = Y; S->iz[cnt] = Zc;
                                S->ityp[cnt] = t; S->iatm[cnt] = a;
                                S->inn[3 * cnt] = ijk[0]; S->inn[3 * cnt + 1] = ijk[1];
                                S->inn[3 * cnt + 2] = ijk[2];
                                S->iudv[cnt] = off;
                            }
                            long nnd = (long)(2 * nc + 1) * (2 * nc + 1) * (2 * nc + 1);
                            off += nnd * S->nproj[t];
                            cnt++;
                        }
            }
            a0 += natm[t];
        }
        if (pass == 0) {
            S->n_img = cnt;
            S->ix = (double *)malloc((size_t)(cnt + 1) * sizeof(double));
            S->iy = (double *)malloc((size_t)(cnt + 1) * sizeof(double));
            S->iz = (double *)malloc((size_t)(cnt + 1) * sizeof(double));
            S->ityp = (int *)malloc((size_t)(cnt + 1) * sizeof(int));
            S->iatm = (int *)malloc((size_t)(cnt + 1) * sizeof(int));
            S->inn = (int *)malloc((size_t)(3 * cnt + 3) * sizeof(int));
            S->iudv = (long *)malloc((size_t)(cnt + 1) * sizeof(long));
            S->udv_len = off;
            S->udv = (double *)calloc((size_t)(off + 1), sizeof(double));
        }
    }

    /* projector cubes UdV = U_l(r) Y_lm (V_l(r) - V_loc(r)) (nonlocal.cpp:330-515) */
    for (int J = 0; J < S->n_img; J++) {
        int t = S->ityp[J];
        const sqo_psd *P = &S->psd[t];
        int nc = S->nlocc[t], w = 2 * nc + 1;
        long nnd = (long)w * w * w;
        double *U = S->udv + S->iudv[J];
        int cnt = 0;
        for (int l = 0; l <= P->lmax; l++) {
            if (l == S->lloc[t]) continue;
            for (int m = -l; m <= l; m++, cnt++) {
                long ctr = 0;
                for (int kk = -nc; kk <= nc; kk++)
                    for (int jj = -nc; jj <= nc; jj++)
                        for (int ii = -nc; ii <= nc; ii++, ctr++) {
                            double xrc = (S->inn[3 * J] + ii) * S->h;
                            double yrc = (S->inn[3 * J + 1] + jj) * S->h;
                            double zrc = (S->inn[3 * J + 2] + kk) * S->h;
                            double dx = xrc - S->ix[J], dy = yrc - S->iy[J], dz = zrc - S->iz[J];
                            double r = sqrt(dx * dx + dy * dy + dz * dz);
                            double Vloc, Vl, Ul;
                            int st = sqo_spline_eval(P->r, P->rVloc, P->size, &r, &Vloc, 1, P->DVloc);
                            st |= sqo_spline_eval(P->r, P->rV[l], P->size, &r, &Vl, 1, P->DV[l]);
                            st |= sqo_spline_eval(P->r, P->rU[l], P->size, &r, &Ul, 1, P->DU[l]);
                            if (st) return SQO_ERR_SPLINE_RANGE;
                            if (fabs(r - 0.0) < TEMP_TOL) {
                                Vloc = P->rVloc[1] / P->r[1];
                                Vl = P->rV[l][1] / P->r[1];
                                Ul = P->rU[l][1] / P->r[1];
                            } else {
                                Vloc = Vloc / r; Vl = Vl / r; Ul = Ul / r;
                            }
                            double SH = sqo_sph_harm(dx, dy, dz, r, l, m);
                            U[cnt * nnd + ctr] = Ul * SH * (Vl - Vloc);
                        }
            }
        }
    }
    return SQO_OK;
}

void sqo_sys_free(sqo_sys *S)
{
    for (int t = 0; t < S->n_typ; t++) sqo_free_psd(&S->psd[t]);
    free(S->ax); free(S->ay); free(S->az);
    free(S->ix); free(S->iy); free(S->iz); free(S->ityp); free(S->iatm); free(S->inn);
    free(S->iudv); free(S->udv);
    memset(S, 0, sizeof(*S));
}

/* ------------------------------------------------------------------ */
/* Nonlocal apply restricted to a node cube (nonlocal.cpp:20-215,      */
/* 522-724).  atom < 0: all atoms; otherwise only that atom's images.  */
/* Accumulates into vv.                                                 */
/* ------------------------------------------------------------------ */
static int vnl_accumulate(const sqo_sys *S, int node, int atom, const double *v, double *vv)
{
    int n = S->n, nloc = S->nloc, M = 2 * nloc + 1;
    int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    for (int J = 0; J < S->n_img; J++) {
        if (atom >= 0 && S->iatm[J] != atom) continue;
        int t = S->ityp[J];
        int nc = S->nlocc[t], w = 2 * nc + 1;
        long nnd = (long)w * w * w;
        const int *c = S->inn + 3 * J;
        int lo[3], hi[3], ok = 1;
        int nd3[3] = {ni, nj, nk};
        for (int d = 0; d < 3; d++) {
            lo[d] = imax(c[d] - nc, nd3[d] - nloc);
            hi[d] = imin(c[d] + nc, nd3[d] + nloc);
            if (lo[d] > hi[d]) ok = 0;
        }
        if (!ok) continue;
        const double *U = S->udv + S->iudv[J];
        int cnt = 0;
        for (int l = 0; l <= S->psd[t].lmax; l++) {
            if (l == S->lloc[t]) continue;
            double Denom = S->psd[t].Denom[l];
            if (Denom == 0) return SQO_ERR_DENOM_ZERO;
            for (int m = -l; m <= l; m++, cnt++) {
                const double *u = U + cnt * nnd;
                double ts = 0.0;
                for (int k = lo[2]; k <= hi[2]; k++)
                    for (int j = lo[1]; j <= hi[1]; j++)
                        for (int i = lo[0]; i <= hi[0]; i++) {
                            long p = ((long)(k - nk + nloc) * M + (j - nj + nloc)) * M + (i - ni + nloc);
                            long q = ((long)(k - c[2] + nc) * w + (j - c[1] + nc)) * w + (i - c[0] + nc);
                            ts = ts + v[p] * u[q] / Denom;
                        }
                for (int k = lo[2]; k <= hi[2]; k++)
                    for (int j = lo[1]; j <= hi[1]; j++)
                        for (int i = lo[0]; i <= hi[0]; i++) {
                            long p = ((long)(k - nk + nloc) * M + (j - nj + nloc)) * M + (i - ni + nloc);
                            long q = ((long)(k - c[2] + nc) * w + (j - c[1] + nc)) * w + (i - c[0] + nc);
                            vv[p] = vv[p] + ts * u[q];
                        }
            }
        }
    }
    return SQO_OK;
}

int sqo_vnl_atom(const sqo_sys *S, int node, int atom, const double *v, double *vv)
{
    long N = (long)(2 * S->nloc + 1) * (2 * S->nloc + 1) * (2 * S->nloc + 1);
    for (long p = 0; p < N; p++) vv[p] = 0.0;
    return vnl_accumulate(S, node, atom, v, vv);
}

/* H_i v (sq.cpp:838-929) */
int sqo_hsub(const sqo_sys *S, const double *veff, int node, const double *v, double *hv)
{
    int n = S->n, nloc = S->nloc, M = 2 * nloc + 1, F = S->FDn, MP = M + 2 * F;
    long N = (long)M * M * M;
    int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    double *pad = (double *)calloc((size_t)MP * MP * MP, sizeof(double));
    double *vv = (double *)calloc((size_t)N, sizeof(double));
    for (int kk = 0; kk < M; kk++)
        for (int jj = 0; jj < M; jj++)
            for (int ii = 0; ii < M; ii++)
                pad[((long)(kk + F) * MP + (jj + F)) * MP + ii + F] = v[((long)kk * M + jj) * M + ii];
    int st = vnl_accumulate(S, node, -1, v, vv);
    const double *c = S->coeff_lap;
    long sz = (long)MP * MP, sy = MP;
    for (int kk = 0; kk < M; kk++)
        for (int jj = 0; jj < M; jj++)
            for (int ii = 0; ii < M; ii++) {
                long q = ((long)(kk + F) * MP + (jj + F)) * MP + ii + F;
                double temp = pad[q] * 3 * c[0];
                for (int a = 1; a <= F; a++)
                    temp += (pad[q - a] + pad[q + a] + pad[q - a * sy] + pad[q + a * sy] +
                             pad[q - a * sz] + pad[q + a * sz]) * c[a];
                long p = ((long)kk * M + jj) * M + ii;
                int gx = wrapi(ni - nloc + ii, n), gy = wrapi(nj - nloc + jj, n), gz = wrapi(nk - nloc + kk, n);
                double ve = veff[((long)gz * n + gy) * n + gx];
                hv[p] = -0.5 * temp + ve * v[p] + vv[p];
            }
    free(pad); free(vv);
    return st;
}

/* ------------------------------------------------------------------ */
/* Lanczos (sq.cpp:931-1092)                                           */
/* ------------------------------------------------------------------ */
static double dot(const double *a, const double *b, long N)
{
    double s = 0.0;
    for (long i = 0; i < N; i++) s = s + a[i] * b[i];
    return s;
}

int sqo_lanczos(const sqo_sys *S, const double *veff, int node, double *lmin, double *lmax, int *iters)
{
    const int max_iter = 500;
    long N = (long)(2 * S->nloc + 1) * (2 * S->nloc + 1) * (2 * S->nloc + 1);
    double *vkm1 = (double *)malloc((size_t)N * sizeof(double));
    double *vk = (double *)malloc((size_t)N * sizeof(double));
    double *vkp1 = (double *)malloc((size_t)N * sizeof(double));
    double *aa = (double *)calloc((size_t)max_iter + 2, sizeof(double));
    double *bb = (double *)calloc((size_t)max_iter + 2, sizeof(double));
    for (long i = 0; i < N; i++) vkm1[i] = 1.0;
    double val = sqrt(dot(vkm1, vkm1, N));
    for (long i = 0; i < N; i++) vkm1[i] = vkm1[i] / val;
    int st = sqo_hsub(S, veff, node, vkm1, vk);
    val = dot(vkm1, vk, N);
    aa[0] = val;
    for (long i = 0; i < N; i++) vk[i] = vk[i] - aa[0] * vkm1[i];
    val = sqrt(dot(vk, vk, N));
    bb[0] = val;
    for (long i = 0; i < N; i++) vk[i] = vk[i] / bb[0];
    int count = 0;
    double lmin_prev = 0.0, lmax_prev = 0.0, dl = 1.0, dm = 1.0;
    while (count < max_iter && (dl > S->lanczos_tol || dm > S->lanczos_tol)) {
        st |= sqo_hsub(S, veff, node, vk, vkp1);
        val = dot(vk, vkp1, N);
        aa[count + 1] = val;
        for (long i = 0; i < N; i++)
            vkp1[i] = vkp1[i] - aa[count + 1] * vk[i] - bb[count] * vkm1[i];
        val = sqrt(dot(vkp1, vkp1, N));
        bb[count + 1] = val;
        for (long i = 0; i < N; i++) {
            vkm1[i] = vk[i];
            vk[i] = vkp1[i] / bb[count + 1];
        }
        int q = sqo_tridiag_eig(aa, bb, count + 2, lmin, lmax);
        if (q) { st = q; break; }
        dl = fabs((*lmin) - lmin_prev);
        dm = fabs((*lmax) - lmax_prev);
        lmin_prev = *lmin;
        lmax_prev = *lmax;
        count = count + 1;
    }
    *lmin -= S->lanczos_tol;
    *lmax += S->lanczos_tol;
    *iters = count;
    free(vkm1); free(vk); free(vkp1); free(aa); free(bb);
    if (st < 0) return st;
    return count == max_iter ? SQO_WARN_LANCZOS_MAXIT : SQO_OK;
}

/* ------------------------------------------------------------------ */
/* Per-node Lanczos + Chebyshev moments (sq.cpp:530-661)               */
/* ------------------------------------------------------------------ */
int sqo_node_moments(const sqo_sys *S, const double *veff, int node,
                     double *chi, double *zee, double *lmin_out, double *lmax_out,
                     double *mu, int *iters)
{
    int nloc = S->nloc, M = 2 * nloc + 1, npl = S->npl;
    if (npl % 4 != 0) return SQO_ERR_NPL_NOT_MULT4;
    long N = (long)M * M * M, c = ((long)nloc * M + nloc) * M + nloc;
    double lmin, lmax;
    int st = sqo_lanczos(S, veff, node, &lmin, &lmax, iters);
    if (st < 0) return st;
    if (lmin_out) *lmin_out = lmin;
    if (lmax_out) *lmax_out = lmax;
    double ch = (lmax + lmin) / 2, ze = (lmax - lmin) / 2;
    *chi = ch; *zee = ze;
    double *t0 = (double *)calloc((size_t)N, sizeof(double));
    double *t1 = (double *)malloc((size_t)N * sizeof(double));
    double *t2 = (double *)malloc((size_t)N * sizeof(double));
    double *Hv = (double *)malloc((size_t)N * sizeof(double));
    double h3 = pow(S->h, 3);
    t0[c] = 1.0;
    st = sqo_hsub(S, veff, node, t0, Hv);
    for (long i = 0; i < N; i++) t1[i] = (Hv[i] - ch * t0[i]) / ze;
    double htemp = t1[c];
    int q4 = (int)round(0.25 * npl), q2 = (int)round(0.5 * npl);
    for (int nq = 0; nq <= q4; nq++) {
        if (nq == 0) mu[nq] = (double)(t0[c] / h3);
        else if (nq == 1) mu[nq] = (double)(t1[c] / h3);
        else {
            st |= sqo_hsub(S, veff, node, t1, Hv);
            for (long i = 0; i < N; i++) {
                t2[i] = (2 * (Hv[i] - ch * t1[i]) / ze) - t0[i];
                t0[i] = t1[i];
                t1[i] = t2[i];
            }
            mu[nq] = (double)(t2[c] / h3);
        }
    }
    for (int nq = q4 + 1; nq <= q2; nq++) {
        double d1 = 0.0, d2 = 0.0;
        st |= sqo_hsub(S, veff, node, t1, Hv);
        for (long i = 0; i < N; i++) {
            t2[i] = (2 * (Hv[i] - ch * t1[i]) / ze) - t0[i];
            d1 += t2[i] * t2[i];
            d2 += t2[i] * t1[i];
            t0[i] = t1[i];
            t1[i] = t2[i];
        }
        mu[nq] = (double)(t2[c] / h3);
        mu[2 * nq] = (double)((2 * d1 - 1) / h3);
        mu[2 * nq - 1] = (double)((2 * d2 - htemp) / h3);
    }
    free(t0); free(t1); free(t2); free(Hv);
    return st;
}

/* ------------------------------------------------------------------ */
/* Derivatives with respect to Veff (chi, zee frozen)                  */
/* ------------------------------------------------------------------ */
int sqo_moments_fixed(const sqo_sys *S, const double *veff, int node, double ch, double ze, double *mu)
{   /* sq.cpp:568-655 with given chi, zee */
    int nloc = S->nloc, M = 2 * nloc + 1, npl = S->npl;
    long N = (long)M * M * M, c = ((long)nloc * M + nloc) * M + nloc;
    double *t0 = (double *)calloc((size_t)N, sizeof(double));
    double *t1 = (double *)malloc((size_t)N * sizeof(double));
    double *t2 = (double *)malloc((size_t)N * sizeof(double));
    double *Hv = (double *)malloc((size_t)N * sizeof(double));
    double h3 = pow(S->h, 3);
    t0[c] = 1.0;
    int st = sqo_hsub(S, veff, node, t0, Hv);
    for (long i = 0; i < N; i++) t1[i] = (Hv[i] - ch * t0[i]) / ze;
    double htemp = t1[c];
    int q4 = npl / 4, q2 = npl / 2;
    mu[0] = t0[c] / h3;
    mu[1] = t1[c] / h3;
    for (int nq = 2; nq <= q2; nq++) {
        double d1 = 0.0, d2 = 0.0;
        st |= sqo_hsub(S, veff, node, t1, Hv);
        for (long i = 0; i < N; i++) {
            t2[i] = (2 * (Hv[i] - ch * t1[i]) / ze) - t0[i];
            d1 += t2[i] * t2[i];
            d2 += t2[i] * t1[i];
            t0[i] = t1[i];
            t1[i] = t2[i];
        }
        mu[nq] = t2[c] / h3;
        if (nq > q4) {
            mu[2 * nq] = (2 * d1 - 1) / h3;
            mu[2 * nq - 1] = (2 * d2 - htemp) / h3;
        }
    }
    free(t0); free(t1); free(t2); free(Hv);
    return st;
}

int sqo_density_fixed(const sqo_sys *S, const double *veff, const double *chi, const double *zee,
                      double ncharge, double lf_lo, double lf_hi, double *lambda_f, double *rho,
                      double *mu, int nthreads)
{
    int n = S->n, npl = S->npl, nn = n * n * n, st = SQO_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int nd = 0; nd < nn; nd++) {
        int s = sqo_moments_fixed(S, veff, nd, chi[nd], zee[nd], mu + (long)nd * (npl + 1));
        if (s < 0) {
#pragma omp critical
            st = s;
        }
    }
    if (st < 0) return st;
    double lf;
    int it;
    st = sqo_brent(S, nn, chi, zee, mu, ncharge, lf_lo, lf_hi, 1, 0.0, &lf, &it);
    if (st < 0) return st;
    *lambda_f = lf;
    return sqo_rho_energy(S, nn, chi, zee, mu, lf, rho, NULL, NULL, NULL);
}

/* apply A = (H - chi)/zee */
static int apply_A(const sqo_sys *S, const double *veff, int node, double ch, double ze,
                   const double *v, double *out, double *tmp, long N)
{
    int st = sqo_hsub(S, veff, node, v, tmp);
    for (long i = 0; i < N; i++) out[i] = (tmp[i] - ch * v[i]) / ze;
    return st;
}

/* per-node rho_pc, its lambda-derivative and the Fermi coefficients */
static void node_fermi(const sqo_sys *S, const double *m, double ch, double ze, double lf,
                       double *rho_pc, double *drho_dl, double *C, double *Cd)
{
    int npl = S->npl;
    double lfp = (lf - ch) / ze, bet = S->bet * ze;
    sqo_cheb_coeff(npl, lfp, bet, C, 0);
    sqo_cheb_coeff(npl, lfp, bet, Cd, 3);
    double r = 0.0, d = 0.0;
    for (int j = 0; j <= npl; j++) { r += 2 * m[j] * C[j]; d += 2 * m[j] * Cd[j]; }
    *rho_pc = r;
    *drho_dl = d * S->bet;
}

static void map_cube_to_grid(const sqo_sys *S, int node, long *g)
{
    int n = S->n, nloc = S->nloc, M = 2 * nloc + 1;
    int ni = node % n, nj = (node / n) % n, nk = node / (n * n);
    long p = 0;
    for (int z = 0; z < M; z++)
        for (int y = 0; y < M; y++)
            for (int x = 0; x < M; x++, p++)
                g[p] = ((long)wrapi(nk - nloc + z, n) * n + wrapi(nj - nloc + y, n)) * n + wrapi(ni - nloc + x, n);
}

int sqo_adjoint_veff(const sqo_sys *S, const double *veff, const double *chi, const double *zee,
                     const double *mu, double lambda_f, const double *w, int fixed_n,
                     double *dLdV, int nthreads)
{
    int n = S->n, npl = S->npl, nn = n * n * n, nloc = S->nloc, M = 2 * nloc + 1;
    long N = (long)M * M * M, c = ((long)nloc * M + nloc) * M + nloc;
    int K2 = npl / 2, q4 = npl / 4;
    double h3 = pow(S->h, 3);
    double *sgn = (double *)malloc((size_t)nn * sizeof(double));
    double *dl = (double *)malloc((size_t)nn * sizeof(double));
    double *C = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    double *Cd = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    double num = 0.0, den = 0.0;
    for (int i = 0; i < nn; i++) {
        double r;
        node_fermi(S, mu + (long)i * (npl + 1), chi[i], zee[i], lambda_f, &r, &dl[i], C, Cd);
        sgn[i] = r < 0 ? -1.0 : 1.0;
        num += w[i] * sgn[i] * dl[i];
        den += sgn[i] * dl[i];
    }
    double kappa = fixed_n ? num / den : 0.0;
    free(C); free(Cd);
    int nth = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    nth = omp_get_max_threads();
#endif
    double *acc = (double *)calloc((size_t)nth * nn, sizeof(double));
    int st = SQO_OK;
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *T = (double *)malloc((size_t)(K2 + 1) * N * sizeof(double));
        double *tb0 = (double *)malloc((size_t)N * sizeof(double));   /* tbar_{k}   */
        double *tb1 = (double *)calloc((size_t)N, sizeof(double));    /* tbar_{k+1} */
        double *tb2 = (double *)calloc((size_t)N, sizeof(double));    /* tbar_{k+2} */
        double *G = (double *)malloc((size_t)N * sizeof(double));
        double *tmp = (double *)malloc((size_t)N * sizeof(double));
        double *tmp2 = (double *)malloc((size_t)N * sizeof(double));
        double *Cn = (double *)malloc((size_t)(npl + 1) * sizeof(double));
        double *Cdn = (double *)malloc((size_t)(npl + 1) * sizeof(double));
        double *be = (double *)malloc((size_t)(npl + 2) * sizeof(double));
        long *gidx = (long *)malloc((size_t)N * sizeof(long));
        double *my = acc + (size_t)tid * nn;
#pragma omp for schedule(static)
        for (int i = 0; i < nn; i++) {
            const double *m = mu + (long)i * (npl + 1);
            double r, d;
            node_fermi(S, m, chi[i], zee[i], lambda_f, &r, &d, Cn, Cdn);
            double wt = (w[i] - kappa) * sgn[i];
            for (int j = 0; j <= npl; j++) be[j] = 2 * Cn[j] * wt;
            /* forward vectors t_0..t_K2 */
            double *t = T;
            for (long p = 0; p < N; p++) t[p] = 0.0;
            t[c] = 1.0;
            int s = apply_A(S, veff, i, chi[i], zee[i], t, T + N, tmp, N);
            for (int k = 1; k < K2; k++) {
                s |= apply_A(S, veff, i, chi[i], zee[i], T + k * N, tmp2, tmp, N);
                for (long p = 0; p < N; p++) T[(k + 1) * N + p] = 2 * tmp2[p] - T[(k - 1) * N + p];
            }
            /* backward (Clenshaw-type) recurrence of the adjoint vectors */
            double sum_odd = 0.0;
            for (int k = q4 + 1; k <= K2; k++) sum_odd += be[2 * k - 1];
            for (long p = 0; p < N; p++) { tb1[p] = 0.0; tb2[p] = 0.0; G[p] = 0.0; }
            for (int k = K2; k >= 1; k--) {
                /* direct gradient g_k */
                const double *tk = T + (long)k * N;
                for (long p = 0; p < N; p++) tb0[p] = 0.0;
                tb0[c] += be[k] / h3;
                if (k > q4) {
                    const double *tkm = T + (long)(k - 1) * N;
                    for (long p = 0; p < N; p++) tb0[p] += (4 * be[2 * k] * tk[p] + 2 * be[2 * k - 1] * tkm[p]) / h3;
                }
                if (k + 1 > q4 && k + 1 <= K2) {
                    const double *tkp = T + (long)(k + 1) * N;
                    for (long p = 0; p < N; p++) tb0[p] += 2 * be[2 * k + 1] * tkp[p] / h3;
                }
                if (k == 1) tb0[c] -= sum_odd / h3;
                if (k < K2) {
                    s |= apply_A(S, veff, i, chi[i], zee[i], tb1, tmp2, tmp, N);
                    for (long p = 0; p < N; p++) tb0[p] += 2 * tmp2[p] - tb2[p];
                }
                /* G += 2 tbar_{k+1} t_k (k >= 1); the k = 0 term uses tbar_1 */
                if (k < K2)
                    for (long p = 0; p < N; p++) G[p] += 2 * tb1[p] * tk[p];
                double *r2 = tb2; tb2 = tb1; tb1 = tb0; tb0 = r2;
            }
            /* tb1 now holds tbar_1; t_1 = A t_0 contributes tbar_1 t_0^T */
            G[c] += tb1[c] * 1.0;
            map_cube_to_grid(S, i, gidx);
            for (long p = 0; p < N; p++) my[gidx[p]] += G[p] / zee[i];
            if (s < 0) {
#pragma omp critical
                st = s;
            }
        }
        free(T); free(tb0); free(tb1); free(tb2); free(G); free(tmp); free(tmp2);
        free(Cn); free(Cdn); free(be); free(gidx);
    }
    for (int i = 0; i < nn; i++) {
        double s = 0.0;
        for (int t = 0; t < nth; t++) s += acc[(size_t)t * nn + i];
        dLdV[i] = s;
    }
    free(acc); free(sgn); free(dl);
    return st;
}

int sqo_tangent_veff(const sqo_sys *S, const double *veff, const double *chi, const double *zee,
                     const double *mu, double lambda_f, const double *dV, int fixed_n,
                     double *drho, int nthreads)
{
    int n = S->n, npl = S->npl, nn = n * n * n, nloc = S->nloc, M = 2 * nloc + 1;
    long N = (long)M * M * M, c = ((long)nloc * M + nloc) * M + nloc;
    int K2 = npl / 2, q4 = npl / 4;
    double h3 = pow(S->h, 3);
    double *dl = (double *)malloc((size_t)nn * sizeof(double));
    double *sgn = (double *)malloc((size_t)nn * sizeof(double));
    int st = SQO_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *t0 = (double *)calloc((size_t)N, sizeof(double)), *t1 = (double *)malloc((size_t)N * 8);
        double *d0 = (double *)calloc((size_t)N, sizeof(double)), *d1 = (double *)malloc((size_t)N * 8);
        double *t2 = (double *)malloc((size_t)N * 8), *dd2 = (double *)malloc((size_t)N * 8);
        double *a = (double *)malloc((size_t)N * 8), *tmp = (double *)malloc((size_t)N * 8);
        double *dA = (double *)malloc((size_t)N * 8), *dmu = (double *)malloc((size_t)(npl + 1) * 8);
        double *Cn = (double *)malloc((size_t)(npl + 1) * 8), *Cdn = (double *)malloc((size_t)(npl + 1) * 8);
        long *gidx = (long *)malloc((size_t)N * sizeof(long));
#pragma omp for schedule(static)
        for (int i = 0; i < nn; i++) {
            map_cube_to_grid(S, i, gidx);
            for (long p = 0; p < N; p++) { dA[p] = dV[gidx[p]] / zee[i]; t0[p] = 0.0; d0[p] = 0.0; }
            t0[c] = 1.0;
            int s = apply_A(S, veff, i, chi[i], zee[i], t0, t1, tmp, N);
            for (long p = 0; p < N; p++) d1[p] = dA[p] * t0[p];
            double dh = d1[c];
            for (int j = 0; j <= npl; j++) dmu[j] = 0.0;
            dmu[1] = d1[c] / h3;
            for (int k = 2; k <= K2; k++) {
                s |= apply_A(S, veff, i, chi[i], zee[i], t1, a, tmp, N);
                for (long p = 0; p < N; p++) t2[p] = 2 * a[p] - t0[p];
                s |= apply_A(S, veff, i, chi[i], zee[i], d1, a, tmp, N);
                double s1 = 0.0, s2 = 0.0;
                for (long p = 0; p < N; p++) {
                    dd2[p] = 2 * (dA[p] * t1[p] + a[p]) - d0[p];
                    s1 += t2[p] * dd2[p];
                    s2 += dd2[p] * t1[p] + t2[p] * d1[p];
                }
                dmu[k] = dd2[c] / h3;
                if (k > q4) {
                    dmu[2 * k] = 4 * s1 / h3;
                    dmu[2 * k - 1] = (2 * s2 - dh) / h3;
                }
                double *r = t0; t0 = t1; t1 = t2; t2 = r;
                r = d0; d0 = d1; d1 = dd2; dd2 = r;
            }
            double rp, dlr;
            node_fermi(S, mu + (long)i * (npl + 1), chi[i], zee[i], lambda_f, &rp, &dlr, Cn, Cdn);
            double sg = rp < 0 ? -1.0 : 1.0, v = 0.0;
            for (int j = 0; j <= npl; j++) v += 2 * Cn[j] * dmu[j];
            drho[i] = sg * v;
            dl[i] = sg * dlr;
            sgn[i] = sg;
            if (s < 0) {
#pragma omp critical
                st = s;
            }
        }
        free(t0); free(t1); free(d0); free(d1); free(t2); free(dd2); free(a); free(tmp); free(dA);
        free(dmu); free(Cn); free(Cdn); free(gidx);
    }
    if (fixed_n) {
        double num = 0.0, den = 0.0;
        for (int i = 0; i < nn; i++) { num += drho[i]; den += dl[i]; }
        double dlam = -num / den;
        for (int i = 0; i < nn; i++) drho[i] += dl[i] * dlam;
    }
    free(dl); free(sgn);
    return st;
}

int sqo_moments_nodes(const sqo_sys *S, const double *veff, const int *nodes, int count,
                      double *chi, double *zee, double *mu, int *iters, int nthreads)
{
    int st = SQO_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < count; q++) {
        double lo, hi;
        int s = sqo_node_moments(S, veff, nodes[q], &chi[q], &zee[q], &lo, &hi,
                                 mu + (long)q * (S->npl + 1), &iters[q]);
        if (s < 0) {
#pragma omp critical
            st = s;
        }
    }
    return st;
}

/* ------------------------------------------------------------------ */
/* Fermi level (sq.cpp:1183-1228, 1289-1490)                           */
/* ------------------------------------------------------------------ */
double sqo_ne_constraint(const sqo_sys *S, int nn, const double *chi, const double *zee,
                         const double *mu, double ncharge, double lf)
{
    int npl = S->npl;
    double *C = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    double g = 0.0;
    for (int k = 0; k < nn; k++) {
        double lfp = (lf - chi[k]) / zee[k];
        double bet = S->bet * zee[k];
        sqo_cheb_coeff(npl, lfp, bet, C, 0);
        for (int j = 0; j < npl + 1; j++) g = g + 2 * mu[(long)k * (npl + 1) + j] * C[j];
    }
    g = g * pow(S->h, 3);
    free(C);
    return g - (-ncharge);
}

#define SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))

int sqo_brent(const sqo_sys *S, int nn, const double *chi, const double *zee, const double *mu,
              double ncharge, double x1, double x2, int scf_iter, double lf_prev,
              double *root, int *iters)
{
    double EPS = 2.5e-16, tol = S->fermi_tol;
    const int ITMAX = 100;
    double a = x1, b = x2, c = x2, d = 0.0, e = 0.0, min1, min2, tol1q, eq;
    if (scf_iter > 1) { a = lf_prev - 2; b = lf_prev + 2; }
    double fa = sqo_ne_constraint(S, nn, chi, zee, mu, ncharge, a);
    double fb = sqo_ne_constraint(S, nn, chi, zee, mu, ncharge, b);
    double fc, p, q, r, s, tol1, xm;
    if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0)) return SQO_ERR_BRENT_BRACKET;
    fc = fb;
    for (int iter = 1; iter <= ITMAX; iter++) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a; fc = fa; e = d = b - a;
        }
        if (fabs((double)fc) < fabs((double)fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        tol1 = 2.0 * EPS * fabs((double)b) + 0.5 * tol;
        xm = 0.5 * (c - b);
        if (fabs((double)xm) <= tol1 || fb == 0.0) {
            *root = b; *iters = iter;
            return SQO_OK;
        }
        if (fabs((double)e) >= tol1 && fabs((double)fa) > fabs((double)fb)) {
            s = fb / fa;
            if (a == c) {
                p = 2.0 * xm * s;
                q = 1.0 - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = fabs((double)p);
            tol1q = tol1 * q;
            min1 = 3.0 * xm * q - fabs((double)tol1q);
            eq = e * q;
            min2 = fabs((double)eq);
            if (2.0 * p < (min1 < min2 ? min1 : min2)) { e = d; d = p / q; }
            else { d = xm; e = d; }
        } else { d = xm; e = d; }
        a = b; fa = fb;
        if (fabs((double)d) > tol1) b += d;
        else b += SIGN(tol1, xm);
        fb = sqo_ne_constraint(S, nn, chi, zee, mu, ncharge, b);
    }
    return SQO_ERR_BRENT_ITMAX;
}

int sqo_newton(const sqo_sys *S, int nn, const double *chi, const double *zee, const double *mu,
               double ncharge, double lf0, double *root, int *iters)
{   /* sq.cpp:1407-1490 */
    int npl = S->npl;
    double *C = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    double tol = S->fermi_tol, lnew = 0.0, lold = lf0, err = 1.0;
    const int ITMAX = 200;
    int count = 0;
    double h3 = pow(S->h, 3);
    while (err > tol && count < ITMAX) {
        double g = 0.0, dg = 0.0;
        for (int k = 0; k < nn; k++) {
            double lfp = (lold - chi[k]) / zee[k];
            double bet = S->bet * zee[k];
            sqo_cheb_coeff(npl, lfp, bet, C, 0);
            for (int j = 0; j < npl + 1; j++) g = g + 2 * mu[(long)k * (npl + 1) + j] * C[j];
            sqo_cheb_coeff(npl, lfp, bet, C, 3);
            for (int j = 0; j < npl + 1; j++) dg = dg + 2 * mu[(long)k * (npl + 1) + j] * C[j];
        }
        dg = dg * S->bet;
        dg = dg * h3;
        g = g * h3;
        double fa = g - (-ncharge), dfa = dg;
        lnew = lold - (fa / dfa);
        err = fabs(lnew - lold);
        lold = lnew;
        count += 1;
    }
    free(C);
    *root = lnew; *iters = count;
    return SQO_OK;
}

/* rho / Ebs / Eent (sq.cpp:740-797) */
int sqo_rho_energy(const sqo_sys *S, int nn, const double *chi, const double *zee,
                   const double *mu, double lambda_f, double *rho, double *ebs_out,
                   double *eent_out, double *Ci_out)
{
    int npl = S->npl;
    double *Cf = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    double *C = (double *)malloc((size_t)(npl + 1) * sizeof(double));
    double bet0 = S->bet, h3 = pow(S->h, 3);
    double Ebs = 0.0, Eent = 0.0;
    int st = SQO_OK;
    for (int nd = 0; nd < nn; nd++) {
        const double *m = mu + (long)nd * (npl + 1);
        double bet = bet0 * zee[nd];
        double lfp = (lambda_f - chi[nd]) / zee[nd];
        sqo_cheb_coeff(npl, lfp, bet, Cf, 0);
        double rho_pc = 0.0;
        for (int j = 0; j < npl + 1; j++) rho_pc = rho_pc + 2 * m[j] * Cf[j];
        rho[nd] = fabs(rho_pc);
        if (rho_pc < 0 && fabs(rho_pc) > 1e-6) st = SQO_ERR_NEG_RHO;
        if (Ci_out) memcpy(Ci_out + (long)nd * (npl + 1), Cf, (size_t)(npl + 1) * sizeof(double));
        sqo_cheb_coeff(npl, lfp, bet, C, 1);
        double e = 0.0;
        for (int j = 0; j < npl + 1; j++) e = e + 2 * m[j] * C[j];
        Ebs = Ebs + (zee[nd] * e + chi[nd] * rho_pc) * h3;
        sqo_cheb_coeff(npl, lfp, bet, C, 2);
        double en = 0.0;
        for (int j = 0; j < npl + 1; j++) en = en + 2 * m[j] * C[j];
        Eent = Eent + (1 / bet0) * en * h3;
    }
    if (ebs_out) *ebs_out = Ebs;
    if (eent_out) *eent_out = Eent;
    free(Cf); free(C);
    return st;
}

/* ------------------------------------------------------------------ */
/* The SQ density step (sq.cpp:377-834)                                */
/* ------------------------------------------------------------------ */
int sqo_density(const sqo_sys *S, const double *veff, int scf_iter, double ncharge,
                double *lambda_f, double *rho, double *ebs, double *eent,
                double *chi_o, double *zee_o, double *mu_o, double *Ci_o,
                int *lanczos_iters, int *fermi_iters, int nthreads)
{
    int n = S->n, npl = S->npl, nn = n * n * n;
    if (npl % 4 != 0) return SQO_ERR_NPL_NOT_MULT4;
    double *chi = chi_o ? chi_o : (double *)malloc((size_t)nn * sizeof(double));
    double *zee = zee_o ? zee_o : (double *)malloc((size_t)nn * sizeof(double));
    double *mu = mu_o ? mu_o : (double *)malloc((size_t)nn * (npl + 1) * sizeof(double));
    double *lmn = (double *)malloc((size_t)nn * sizeof(double));
    double *lmx = (double *)malloc((size_t)nn * sizeof(double));
    int *its = (int *)malloc((size_t)nn * sizeof(int));
    int st = SQO_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int nd = 0; nd < nn; nd++) {
        int s = sqo_node_moments(S, veff, nd, &chi[nd], &zee[nd], &lmn[nd], &lmx[nd],
                                 mu + (long)nd * (npl + 1), &its[nd]);
        if (s < 0) {
#pragma omp critical
            st = s;
        }
    }
    if (lanczos_iters) memcpy(lanczos_iters, its, (size_t)nn * sizeof(int));
    if (st >= 0) {
        /* global spectrum bounds in node order (sq.cpp:657-713) */
        double lmin = lmn[0], lmax = lmx[0];
        for (int nd = 1; nd < nn; nd++) {
            if (lmn[nd] < lmin) lmin = lmn[nd];
            if (lmx[nd] > lmax) lmax = lmx[nd];
        }
        if (scf_iter > 1) {
            /* the reference's bracket ignores x1/x2 after the first SCF iteration */
        }
        double lf = 0.0;
        int fit = 0;
        if (S->fermi_alg == 2) {
            double lf0 = scf_iter <= 1 ? 0.2 : *lambda_f;
            st = sqo_newton(S, nn, chi, zee, mu, ncharge, lf0, &lf, &fit);
        } else {
            st = sqo_brent(S, nn, chi, zee, mu, ncharge, lmin - 1, lmax + 1, scf_iter, *lambda_f,
                           &lf, &fit);
        }
        if (fermi_iters) *fermi_iters = fit;
        if (st >= 0) {
            *lambda_f = lf;
            st = sqo_rho_energy(S, nn, chi, zee, mu, lf, rho, ebs, eent, Ci_o);
        }
    }
    if (!chi_o) free(chi);
    if (!zee_o) free(zee);
    if (!mu_o) free(mu);
    free(lmn); free(lmx); free(its);
    return st;
}

/* ------------------------------------------------------------------ */
/* Nonlocal forces (forces.cpp:318-578)                                */
/* ------------------------------------------------------------------ */
int sqo_forces_nonlocal(const sqo_sys *S, const double *veff, const double *chi,
                        const double *zee, const double *Ci, const int *nodes, int n_nodes,
                        double *f, int nthreads)
{
    int n = S->n, nloc = S->nloc, M = 2 * nloc + 1, F = S->FDn, MP = M + 2 * F, npl = S->npl;
    int na = S->n_atm;
    long N = (long)M * M * M;
    int nn = nodes ? n_nodes : n * n * n;
    double h3 = pow(S->h, 3);
    /* per-node contributions, summed in node order afterwards for a
       deterministic result independent of the thread count */
    double *part = (double *)calloc((size_t)nn * 3 * na, sizeof(double));
    int st = SQO_OK;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *t0 = (double *)malloc((size_t)N * sizeof(double));
        double *t1 = (double *)malloc((size_t)N * sizeof(double));
        double *t2 = (double *)malloc((size_t)N * sizeof(double));
        double *Hv = (double *)malloc((size_t)N * sizeof(double));
        double *DM = (double *)malloc((size_t)MP * MP * MP * sizeof(double));
        double *gr[3];
        for (int d = 0; d < 3; d++) gr[d] = (double *)malloc((size_t)N * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < nn; q++) {
            int nd = nodes ? nodes[q] : q;
            long c = ((long)nloc * M + nloc) * M + nloc;
            for (long i = 0; i < N; i++) t0[i] = 0.0;
            t0[c] = 1.0;
            int s = sqo_hsub(S, veff, nd, t0, Hv);
            for (long i = 0; i < N; i++) t1[i] = (Hv[i] - chi[nd] * t0[i]) / zee[nd];
            for (long i = 0; i < (long)MP * MP * MP; i++) DM[i] = 0.0;
            const double *C = Ci + (long)nd * (npl + 1);
            for (int nq = 0; nq <= npl; nq++) {
                if (nq >= 2) {
                    s |= sqo_hsub(S, veff, nd, t1, Hv);
                    for (long i = 0; i < N; i++) {
                        t2[i] = (2 * (Hv[i] - chi[nd] * t1[i]) / zee[nd]) - t0[i];
                    }
                }
                const double *tq = nq == 0 ? t0 : (nq == 1 ? t1 : t2);
                for (int kk = 0; kk < M; kk++)
                    for (int jj = 0; jj < M; jj++)
                        for (int ii = 0; ii < M; ii++) {
                            long p = ((long)kk * M + jj) * M + ii;
                            long pp = ((long)(kk + F) * MP + jj + F) * MP + ii + F;
                            DM[pp] += (double)(tq[p] * C[nq] / h3);
                        }
                if (nq >= 2)
                    for (long i = 0; i < N; i++) { t0[i] = t1[i]; t1[i] = t2[i]; }
            }
            long sz = (long)MP * MP, sy = MP;
            for (int kk = 0; kk < M; kk++)
                for (int jj = 0; jj < M; jj++)
                    for (int ii = 0; ii < M; ii++) {
                        long p = ((long)kk * M + jj) * M + ii;
                        long pp = ((long)(kk + F) * MP + jj + F) * MP + ii + F;
                        double gx = 0.0, gy = 0.0, gz = 0.0;
                        for (int a = 1; a <= F; a++)
                            gx += (-DM[pp - a] + DM[pp + a]) * S->coeff_grad[a];
                        for (int a = 1; a <= F; a++)
                            gy += (-DM[pp - a * sy] + DM[pp + a * sy]) * S->coeff_grad[a];
                        for (int a = 1; a <= F; a++)
                            gz += (-DM[pp - a * sz] + DM[pp + a * sz]) * S->coeff_grad[a];
                        gr[0][p] = gx; gr[1][p] = gy; gr[2][p] = gz;
                    }
            for (int a = 0; a < na; a++) {
                int has = 0;
                for (int J = 0; J < S->n_img; J++) if (S->iatm[J] == a) { has = 1; break; }
                if (!has) continue;
                for (int d = 0; d < 3; d++) {
                    s |= sqo_vnl_atom(S, nd, a, gr[d], Hv);
                    part[((long)q * 3 + d) * na + a] += -4 * Hv[c] * h3;
                }
            }
            if (s < 0) {
#pragma omp critical
                st = s;
            }
        }
        free(t0); free(t1); free(t2); free(Hv); free(DM);
        for (int d = 0; d < 3; d++) free(gr[d]);
    }
    for (int i = 0; i < 3 * na; i++) f[i] = 0.0;
    for (int q = 0; q < nn; q++)
        for (int d = 0; d < 3; d++)
            for (int a = 0; a < na; a++) f[d * na + a] += part[((long)q * 3 + d) * na + a];
    free(part);
    return st;
}

/* ------------------------------------------------------------------ */
/* Poisson: AAJ (Jacobi + periodic Anderson), poisson.cpp:19-325,     */
/* PoissonResidual poisson.cpp:715-928, AndersonExtrapolation and     */
/* PseudoInverseTimesVec anderson.cpp:20-163                          */
/* ------------------------------------------------------------------ */

/* x = pinv(A) b for the symmetric m x m matrix A (column major), by a
 * one-sided Jacobi SVD of A; singular values below m * max(w) *
 * 2.22044605e-16 are zeroed (anderson.cpp:392-397) and zero ones skipped
 * (anderson.cpp:134-142). */
static void pinv_times_vec(const double *Ac, const double *b, double *x, int m)
{
    double U[64], V[64], w[8];
    for (int i = 0; i < m * m; i++) U[i] = Ac[i];   /* columns of U start as columns of A */
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) V[j * m + i] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                double a = 0, bb = 0, c = 0;
                for (int i = 0; i < m; i++) {
                    a += U[p * m + i] * U[p * m + i];
                    bb += U[q * m + i] * U[q * m + i];
                    c += U[p * m + i] * U[q * m + i];
                }
                if (c == 0.0) continue;
                const double rel = fabs(c) / sqrt(a * bb);
                if (rel > off) off = rel;
                const double zeta = (bb - a) / (2 * c);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
                for (int i = 0; i < m; i++) {
                    const double up = U[p * m + i], uq = U[q * m + i];
                    U[p * m + i] = cs * up - sn * uq;
                    U[q * m + i] = sn * up + cs * uq;
                    const double vp = V[p * m + i], vq = V[q * m + i];
                    V[p * m + i] = cs * vp - sn * vq;
                    V[q * m + i] = sn * vp + cs * vq;
                }
            }
        if (off < 1e-15) break;
    }
    double wmax = 0.0;
    for (int j = 0; j < m; j++) {
        double s = 0;
        for (int i = 0; i < m; i++) s += U[j * m + i] * U[j * m + i];
        w[j] = sqrt(s);
        if (w[j] > wmax) wmax = w[j];
    }
    const double wmin = m * wmax * (2.22044605e-16);
    for (int j = 0; j < m; j++)
        if (w[j] < wmin) w[j] = 0.0;
    double tmp[8];
    for (int j = 0; j < m; j++) {   /* (U_j/w_j)' b / w_j */
        double s = 0.0;
        if (w[j] != 0.0) {
            for (int i = 0; i < m; i++) s += U[j * m + i] / w[j] * b[i];
            s /= w[j];
        }
        tmp[j] = s;
    }
    for (int i = 0; i < m; i++) {
        double s = 0.0;
        for (int j = 0; j < m; j++) s += V[j * m + i] * tmp[j];
        x[i] = s;
    }
}

/* Jacobi residual of -Lap phi = 4 pi rhs on the periodic n^3 grid:
 * r = -(4 pi rhs + Lap_FD phi)/(3 lap[0]) (poisson.cpp:779-786). */
static void aaj_residual(int n, int FDn, const double *lap, const double *phi, const double *rhs, double *r)
{
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t c = ((size_t)k * n + j) * n + i;
                double L = phi[c] * 3 * lap[0];
                for (int a = 1; a <= FDn; a++)
                    L += (phi[((size_t)k * n + j) * n + wrapi(i - a, n)] + phi[((size_t)k * n + j) * n + wrapi(i + a, n)] +
                          phi[((size_t)k * n + wrapi(j - a, n)) * n + i] + phi[((size_t)k * n + wrapi(j + a, n)) * n + i] +
                          phi[((size_t)wrapi(k - a, n) * n + j) * n + i] + phi[((size_t)wrapi(k + a, n) * n + j) * n + i]) *
                         lap[a];
                r[c] = -(((4 * M_PI) * (rhs[c]) + L) / (3 * lap[0]));
            }
}

int sqo_poisson_aaj(int n, int FDn, const double *lap, const double *rhs, const double *phi_guess, double tol,
                    int maxiter, double beta, int m, int p, double *phi, int *iters, double *relres)
{
    if (m < 1 || m > 8 || p < 1) return -1;
    const size_t N = (size_t)n * n * n;
    double *old = malloc(N * sizeof(double)), *nw = malloc(N * sizeof(double)), *res = malloc(N * sizeof(double));
    double *Xold = calloc(N, sizeof(double)), *Fold = calloc(N, sizeof(double)), *am = malloc(N * sizeof(double));
    double *DX = calloc(N * m, sizeof(double)), *DF = calloc(N * m, sizeof(double));
    if (!old || !nw || !res || !Xold || !Fold || !am || !DX || !DF) return -2;
    memcpy(old, phi_guess, N * sizeof(double));
    double rn = 0;
    for (size_t c = 0; c < N; c++) rn = rn + rhs[c] * rhs[c];
    rn = sqrt(rn);
    rn *= 4 * M_PI * fabs(-1 / (3 * lap[0]));
    int iter = 1;
    double r = tol + 1;
    while (r > tol && iter <= maxiter) {
        aaj_residual(n, FDn, lap, old, rhs, res);
        for (size_t c = 0; c < N; c++) nw[c] = old[c] + beta * res[c];
        double s = 0;
        for (size_t c = 0; c < N; c++) s = s + res[c] * res[c];
        r = sqrt(s) / rn;
        if (iter > 1) {
            const int k = ((iter - 2) % m);
            for (size_t c = 0; c < N; c++) {
                DX[k * N + c] = old[c] - Xold[c];
                DF[k * N + c] = res[c] - Fold[c];
            }
        }
        memcpy(Xold, old, N * sizeof(double));
        memcpy(Fold, res, N * sizeof(double));
        if (iter % p == 0 && iter > 1) {
            double FtF[64], Ftf[8], x[8];
            int ctr = 0;
            for (int jj = 0; jj < m; jj++)
                for (int ii = 0; ii < m; ii++) {
                    double t = 0;
                    for (size_t c = 0; c < N; c++) t = t + DF[ii * N + c] * DF[jj * N + c];
                    FtF[ctr++] = t;
                }
            for (int jj = 0; jj < m; jj++) {
                double t = 0;
                for (size_t c = 0; c < N; c++) t = t + DF[jj * N + c] * res[c];
                Ftf[jj] = t;
            }
            pinv_times_vec(FtF, Ftf, x, m);
            for (size_t c = 0; c < N; c++) {
                double t = 0;
                for (int jj = 0; jj < m; jj++) t = t + (DX[jj * N + c] + beta * DF[jj * N + c]) * x[jj];
                am[c] = t;
            }
            for (size_t c = 0; c < N; c++) nw[c] = nw[c] - am[c];
        }
        if (r <= tol || iter == maxiter) memcpy(phi, old, N * sizeof(double));
        memcpy(old, nw, N * sizeof(double));
        iter = iter + 1;
    }
    const double p000 = phi[0];   /* phi is defined up to a constant: phi(0,0,0) = 0 (poisson.cpp:272-290) */
    for (size_t c = 0; c < N; c++) phi[c] = phi[c] + 0.0 - p000;
    *iters = iter - 1;
    *relres = r;
    free(old); free(nw); free(res); free(Xold); free(Fold); free(am); free(DX); free(DF);
    return 0;
}

/* ------------------------------------------------------------------ */
/* LDA exchange-correlation potential (Perdew-Wang 92 correlation,     */
/* Slater exchange), scf.cpp:19-64                                     */
/* ------------------------------------------------------------------ */
void sqo_vxc_lda(size_t N, const double *rho, double *vxc)
{
    const double p = 1.0, A = 0.031091, alpha1 = 0.21370, beta1 = 7.5957, beta2 = 3.5876, beta3 = 1.6382,
                 beta4 = 0.49294, C3 = 0.9847450218427;
    for (size_t c = 0; c < N; c++) {
        const double rhoi = rho[c];
        double V;
        if (rhoi == 0) {
            V = 0.0;
        } else {
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

/* ------------------------------------------------------------------ */
/* Total energy per atom: EvaluateTotalEnergy, energy.cpp:19-87        */
/* (phi and vxc of the potential the SQ call used, rho its output)     */
/* ------------------------------------------------------------------ */
double sqo_total_energy(size_t N, double h, const double *phi, const double *b, const double *rho,
                        const double *vxc, double Ebs, double Eent, double Eself, double Ecorr, int n_atm)
{
    const double p = 1.0, A = 0.031091, alpha1 = 0.21370, beta1 = 7.5957, beta2 = 3.5876, beta3 = 1.6382,
                 beta4 = 0.49294, C2 = 0.73855876638202;
    double Eele = 0.0, Exc = 0.0, Exc_del = 0.0;
    for (size_t c = 0; c < N; c++) Eele = Eele + 0.5 * (phi[c] * (b[c] - rho[c])) * h * h * h;
    for (size_t c = 0; c < N; c++) {
        const double rhoi = rho[c];
        const double Ex = -C2 * pow(rhoi, 4.0 / 3.0);
        double Ec;
        if (rhoi == 0) {
            Ec = 0.0;
        } else {
            Ec = pow(0.75 / (M_PI * rhoi), (1.0 / 3.0));
            Ec = -2.0 * A * (1.0 + alpha1 * Ec) *
                 log(1.0 + 1.0 / (2.0 * A * (beta1 * pow(Ec, 0.5) + beta2 * Ec + beta3 * pow(Ec, 1.5) + beta4 * pow(Ec, (p + 1.0)))));
        }
        Exc = Exc + (Ec * rhoi + Ex) * h * h * h;
        Exc_del = Exc_del + -1.0 * vxc[c] * rhoi * h * h * h;
    }
    double Etot = Eele + Exc + Exc_del;
    Etot = Etot - Eself;
    Etot = Etot + Ebs + Eent + Ecorr;
    return Etot / n_atm;
}

/* ------------------------------------------------------------------ */
/* Pseudocharge and local forces: Replica_atoms (initialize.cpp:21-200),*/
/* ChargeDensity (initialize.cpp:471-784), ForcesLocal (forces.cpp:20-315)*/
/* on one rank: the processor domain is [0, n-1]^3.                      */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; int lo[3], hi[3]; int type, atom; } rb_img;

/* periodic replicas whose rb-box around the nearest node meets the domain,
   in the reference's (type, atom, z, y, x) order (initialize.cpp:34-122) */
static int rb_images(const sqo_sys *S, const double *rb, rb_img **out)
{
    int cap = 64, cnt = 0, a = 0;
    rb_img *v = (rb_img *)malloc(cap * sizeof(rb_img));
    const double L = S->L, h = S->h;
    for (int t = 0; t < S->n_typ; t++) {
        const double rbt = ceil(rb[t] / h) * h;   /* rb rounded up to the mesh when read (readfiles.cpp:512) */
        int Rb = (int)round(rbt / h);
        for (int q = 0; q < S->natm[t]; q++, a++) {
            double c[3] = {S->ax[a], S->ay[a], S->az[a]};
            int pm[3], mm[3];
            for (int d = 0; d < 3; d++) {
                pm[d] = (int)floor(((L + rbt - c[d]) / L) + TEMP_TOL);
                mm[d] = (int)ceil(((-rbt - c[d]) / L) - TEMP_TOL);
            }
            for (int rr = mm[2]; rr <= pm[2]; rr++)
                for (int qq = mm[1]; qq <= pm[1]; qq++)
                    for (int pp = mm[0]; pp <= pm[0]; pp++) {
                        rb_img I;
                        I.x = c[0] + pp * L; I.y = c[1] + qq * L; I.z = c[2] + rr * L;
                        double p3[3] = {I.x, I.y, I.z};
                        int ok = 1;
                        for (int d = 0; d < 3; d++) {
                            int nn = nn1(p3[d], h);
                            I.lo[d] = imax(nn - Rb, 0);
                            I.hi[d] = imin(nn + Rb, S->n - 1);
                            if (I.lo[d] > I.hi[d]) ok = 0;
                        }
                        if (!ok) continue;
                        I.type = t; I.atom = a;
                        if (cnt == cap) { cap *= 2; v = (rb_img *)realloc(v, cap * sizeof(rb_img)); }
                        v[cnt++] = I;
                    }
        }
    }
    *out = v;
    return cnt;
}

/* V_J and the smooth reference potential V~_J (rc~ = 0.5) on the image's
   intersection box widened by mg nodes (initialize.cpp:565-595) */
static int vloc_box(const sqo_sys *S, const rb_img *I, int mg, const double *DVloc, double *V, double *Vt)
{
    const sqo_psd *P = &S->psd[I->type];
    const double h = S->h, rc = 0.5, Z = S->Z[I->type];
    int nx = I->hi[0] - I->lo[0] + 1 + 2 * mg, ny = I->hi[1] - I->lo[1] + 1 + 2 * mg;
    int nz = I->hi[2] - I->lo[2] + 1 + 2 * mg;
    int i0 = I->lo[0] - mg, j0 = I->lo[1] - mg, k0 = I->lo[2] - mg;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double dx = (i0 + i) * h - I->x, dy = (j0 + j) * h - I->y, dz = (k0 + k) * h - I->z;
                double r = sqrt(dx * dx + dy * dy + dz * dz), v;
                if (sqo_spline_eval(P->r, P->rVloc, P->size, &r, &v, 1, DVloc)) return SQO_ERR_SPLINE_RANGE;
                long g = ((long)k * ny + j) * nx + i;
                V[g] = fabs(r - 0.0) < TEMP_TOL ? P->rVloc[1] / P->r[1] : v / r;
                Vt[g] = r < rc ? -Z * (9 * pow(r, 7) - 30 * rc * pow(r, 6) + 28 * rc * rc * pow(r, 5) -
                                       14 * pow(rc, 5) * pow(r, 2) + 12 * pow(rc, 7)) / (5 * pow(rc, 8))
                               : -Z / r;
            }
    return SQO_OK;
}

/* -(1/4pi) Lap_FD of a box array at interior point (i,j,k) */
static double neg_lap_4pi(const sqo_sys *S, const double *V, int nx, int ny, int i, int j, int k)
{
    const double q = -1 / (4 * M_PI);
#define AT(kk, jj, ii) V[((long)(kk) * ny + (jj)) * nx + (ii)]
    double s = AT(k, j, i) * 3 * S->coeff_lap[0] * q;
    for (int a = 1; a <= S->FDn; a++)
        s += (AT(k, j, i - a) + AT(k, j, i + a) + AT(k, j - a, i) + AT(k, j + a, i) + AT(k - a, j, i) +
              AT(k + a, j, i)) * S->coeff_lap[a] * q;
#undef AT
    return s;
}

int sqo_charge_density(const sqo_sys *S, const double *rb, double *b, double *bt, double *vc,
                       double *rho, double *scal)
{
    const int n = S->n, F = S->FDn;
    const double h = S->h, h3 = h * h * h;
    const long N3 = (long)n * n * n;
    double *rho_at = (double *)calloc(N3, sizeof(double));
    for (long g = 0; g < N3; g++) b[g] = bt[g] = vc[g] = 0.0;
    double Eself = 0, Ncharge = 0, Eself_t = 0, Ncharge_t = 0, Vcon = 0;
    rb_img *im;
    int nim = rb_images(S, rb, &im), st = SQO_OK;
    double *DV = NULL, *Duu = NULL;
    int cur_t = -1;
    for (int m = 0; m < nim && st == SQO_OK; m++) {
        const rb_img *I = &im[m];
        const sqo_psd *P = &S->psd[I->type];
        if (I->type != cur_t) {   /* initialize.cpp:508-512 */
            free(DV); free(Duu);
            DV = (double *)malloc(P->size * sizeof(double));
            Duu = (double *)malloc(P->size * sizeof(double));
            sqo_spline_yd(P->r, P->rVloc, DV, P->size);
            sqo_spline_yd(P->r, P->uu, Duu, P->size);
            cur_t = I->type;
        }
        int nx = I->hi[0] - I->lo[0] + 1 + 2 * F, ny = I->hi[1] - I->lo[1] + 1 + 2 * F;
        int nz = I->hi[2] - I->lo[2] + 1 + 2 * F;
        int i0 = I->lo[0] - F, j0 = I->lo[1] - F, k0 = I->lo[2] - F;
        double *V = (double *)malloc((size_t)nx * ny * nz * sizeof(double));
        double *Vt = (double *)malloc((size_t)nx * ny * nz * sizeof(double));
        st = vloc_box(S, I, F, DV, V, Vt);
        /* atomic valence density on the intersection (initialize.cpp:582-587) */
        for (int k = F; k < nz - F && st == SQO_OK; k++)
            for (int j = F; j < ny - F; j++)
                for (int i = F; i < nx - F; i++) {
                    double dx = (i0 + i) * h - I->x, dy = (j0 + j) * h - I->y, dz = (k0 + k) * h - I->z;
                    double r = sqrt(dx * dx + dy * dy + dz * dz), u;
                    if (sqo_spline_eval(P->r, P->uu, P->size, &r, &u, 1, Duu)) { st = SQO_ERR_SPLINE_RANGE; break; }
                    rho_at[((long)(k + k0) * n + (j + j0)) * n + (i + i0)] += u;
                }
        /* b_J, b~_J and their sums (initialize.cpp:599-630) */
        for (int k = I->lo[2]; k <= I->hi[2] && st == SQO_OK; k++)
            for (int j = I->lo[1]; j <= I->hi[1]; j++)
                for (int i = I->lo[0]; i <= I->hi[0]; i++) {
                    int li = i - i0, lj = j - j0, lk = k - k0;
                    long l = ((long)lk * ny + lj) * nx + li, g = ((long)k * n + j) * n + i;
                    double bj = neg_lap_4pi(S, V, nx, ny, li, lj, lk);
                    double btj = neg_lap_4pi(S, Vt, nx, ny, li, lj, lk);
                    b[g] += bj;
                    Eself += 0.5 * bj * V[l] * h3;
                    Ncharge += bj * h3;
                    bt[g] += btj;
                    Eself_t += 0.5 * btj * Vt[l] * h3;
                    Ncharge_t += btj * h3;
                    vc[g] += Vt[l] - V[l];
                }
        if (I->lo[0] <= 0 && I->lo[1] <= 0 && I->lo[2] <= 0)   /* node (0,0,0), initialize.cpp:633-636 */
            Vcon += Vt[((long)(0 - k0) * ny + (0 - j0)) * nx + (0 - i0)] - V[((long)(0 - k0) * ny + (0 - j0)) * nx + (0 - i0)];
        free(V); free(Vt);
    }
    free(DV); free(Duu); free(im);
    if (st != SQO_OK) { free(rho_at); return st; }
    /* rho = rho_at scaled to Ncharge electrons, b~ scaled to Ncharge, Vc
       shifted, Ecorr (initialize.cpp:661-779 at the first MD step) */
    double Nel = 0;
    for (long g = 0; g < N3; g++) Nel += rho_at[g] * h3;
    for (long g = 0; g < N3; g++) rho[g] = fabs(Ncharge / Nel) * rho_at[g];
    for (long g = 0; g < N3; g++) bt[g] = fabs(Ncharge / Ncharge_t) * bt[g];
    double V000 = vc[0], Ec = 0;
    for (long g = 0; g < N3; g++) {
        vc[g] += -V000 + Vcon;
        Ec += 0.5 * (bt[g] + b[g]) * vc[g] * pow(h, 3);
    }
    free(rho_at);
    scal[0] = Ncharge; scal[1] = Eself; scal[2] = Ncharge_t; scal[3] = Eself_t;
    scal[4] = Ec + Eself - Eself_t;
    return SQO_OK;
}

int sqo_forces_local(const sqo_sys *S, const double *rb, const double *phi, const double *b,
                     const double *bt, const double *vc, double *floc, double *fcorr)
{
    const int n = S->n, F = S->FDn, na = S->n_atm;
    const double h3 = pow(S->h, 3);
    for (int a = 0; a < 3 * na; a++) floc[a] = fcorr[a] = 0.0;
    rb_img *im;
    int nim = rb_images(S, rb, &im), st = SQO_OK;
    double *DV = NULL;
    int cur_t = -1;
    for (int m = 0; m < nim && st == SQO_OK; m++) {
        const rb_img *I = &im[m];
        const sqo_psd *P = &S->psd[I->type];
        if (I->type != cur_t) {   /* forces.cpp:51-52 */
            free(DV);
            DV = (double *)malloc(P->size * sizeof(double));
            sqo_spline_yd(P->r, P->rVloc, DV, P->size);
            cur_t = I->type;
        }
        /* V on the intersection widened by 2F, b_J on it widened by F */
        int nx = I->hi[0] - I->lo[0] + 1 + 4 * F, ny = I->hi[1] - I->lo[1] + 1 + 4 * F;
        int nz = I->hi[2] - I->lo[2] + 1 + 4 * F;
        int i0 = I->lo[0] - 2 * F, j0 = I->lo[1] - 2 * F, k0 = I->lo[2] - 2 * F;
        size_t NB = (size_t)nx * ny * nz;
        double *V = (double *)malloc(NB * sizeof(double)), *Vt = (double *)malloc(NB * sizeof(double));
        double *bJ = (double *)calloc(NB, sizeof(double)), *bJt = (double *)calloc(NB, sizeof(double));
        st = vloc_box(S, I, 2 * F, DV, V, Vt);
        for (int k = F; k < nz - F && st == SQO_OK; k++)   /* forces.cpp:158-177 */
            for (int j = F; j < ny - F; j++)
                for (int i = F; i < nx - F; i++) {
                    long l = ((long)k * ny + j) * nx + i;
                    bJ[l] = neg_lap_4pi(S, V, nx, ny, i, j, k);
                    bJt[l] = neg_lap_4pi(S, Vt, nx, ny, i, j, k);
                }
        const long sx = 1, sy = nx, sz = (long)nx * ny;
        const long step[3] = {sx, sy, sz};
        double *fl[3] = {floc, floc + na, floc + 2 * na}, *fc[3] = {fcorr, fcorr + na, fcorr + 2 * na};
        for (int k = 2 * F; k < nz - 2 * F && st == SQO_OK; k++)   /* forces.cpp:183-276 */
            for (int j = 2 * F; j < ny - 2 * F; j++)
                for (int i = 2 * F; i < nx - 2 * F; i++) {
                    long l = ((long)k * ny + j) * nx + i;
                    long g = ((long)(k + k0) * n + (j + j0)) * n + (i + i0);
                    for (int d = 0; d < 3; d++) {
                        long s = step[d];
                        double Db = 0, Dbt = 0, DVv = 0, DVt = 0;
                        for (int a = 1; a <= F; a++) Db += (-bJ[l - a * s] + bJ[l + a * s]) * S->coeff_grad[a];
                        fl[d][I->atom] += Db * (phi[g] - V[l]) * h3;
                        for (int a = 1; a <= F; a++) Dbt += (-bJt[l - a * s] + bJt[l + a * s]) * S->coeff_grad[a];
                        for (int a = 1; a <= F; a++) DVv += (-V[l - a * s] + V[l + a * s]) * S->coeff_grad[a];
                        for (int a = 1; a <= F; a++) DVt += (-Vt[l - a * s] + Vt[l + a * s]) * S->coeff_grad[a];
                        fc[d][I->atom] += 0.5 * (Dbt * (vc[g] - Vt[l]) + Db * (vc[g] + V[l]) +
                                                 (DVt - DVv) * (bt[g] + b[g]) + bJ[l] * DVv - bJt[l] * DVt) * h3;
                    }
                }
        free(V); free(Vt); free(bJ); free(bJt);
    }
    free(DV); free(im);
    return st;
}
