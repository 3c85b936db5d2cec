// system.cpp -- host-side preparation of the SQ hot path's inputs.
//
// Builds everything the device kernels consume from a geometry and the
// Troullier-Martins tables: FD weights, KB denominators, the periodic images
// that can touch any node cube, their nearest nodes and their projector
// tables.  Restates (phanish-suryanarayana/SQDFT):
//   FD weights          src/readfiles.cpp:370-406
//   pseudopotential     src/readfiles.cpp:605-1225 (units, r=0 rows, rc/rz rounding, Denom)
//   cubic spline        src/spline.cpp:11-126
//   images              src/initialize.cpp:1003-1196 (Replica_atoms_Rcut)
//   projector tables    src/nonlocal.cpp:218-518 (SphericalHarmonics, NonlocalProjectors)
//   .input / .atoms     src/readfiles.cpp:86-602
// Compiled with -ffp-contract=off: every table must equal the reference's
// (unfused FP64) values bit for bit.
#include "sqdft_system.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

constexpr double kTinyR = 1e-12;  // TEMP_TOL of the reference

// ---------------------------------------------------------------- spline
// Clamped cubic spline on a non-uniform grid; derivative table from the
// tridiagonal system of spline.cpp:57-126, evaluation as spline.cpp:11-55.
void spline_derivs(const double *X, const double *Y, double *YD, int len)
{
    std::vector<double> lo(len, 0.0), di(len, 0.0), up(len, 0.0);
    double h0 = X[1] - X[0], h1 = X[2] - X[1];
    double r0 = (Y[1] - Y[0]) / h0, r1 = (Y[2] - Y[1]) / h1;
    di[0] = h1 * (h0 + h1);
    up[0] = (h0 + h1) * (h0 + h1);
    YD[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0;
    int i = 1;
    for (; i < len - 1; ++i) {
        h0 = X[i] - X[i - 1];
        h1 = X[i + 1] - X[i];
        r0 = (Y[i] - Y[i - 1]) / h0;
        r1 = (Y[i + 1] - Y[i]) / h1;
        lo[i] = h1;
        di[i] = 2 * (h0 + h1);
        up[i] = h0;
        YD[i] = 3 * (r0 * h1 + r1 * h0);
    }
    lo[i] = (h0 + h1) * (h0 + h1);
    di[i] = h0 * (h0 + h1);
    YD[i] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0);
    // Thomas elimination
    std::vector<double> fw(len, 0.0);
    double piv = di[0];
    YD[0] = YD[0] / piv;
    for (int k = 1; k < len; ++k) {
        fw[k] = up[k - 1] / piv;
        piv = di[k] - lo[k] * fw[k];
        YD[k] = (YD[k] - YD[k - 1] * lo[k]) / piv;
    }
    for (int k = len - 2; k >= 0; --k) YD[k] -= (YD[k + 1] * fw[k + 1]);
}

struct Spline {
    const double *X = nullptr, *Y = nullptr;
    std::vector<double> D;
    int n = 0;
    void build(const double *x, const double *y, int len)
    {
        X = x; Y = y; n = len;
        D.assign(len, 0.0);
        spline_derivs(x, y, D.data(), len);
    }
    // single-point evaluation; the reference's interval search for one
    // abscissa starts from the first knot
    bool eval(double p, double *out) const
    {
        if (p < X[0] || p > X[n - 1]) return false;
        int j = 0;
        while (j < n && p > X[j]) ++j;
        if (p < X[j]) --j;
        if (j > n - 2) j = n - 2;  // p == last knot: use the last interval
        const double left = X[j];
        const double dx = 1.0 / (X[j + 1] - X[j]);
        const double dy = (Y[j + 1] - Y[j]) * dx;
        const double c0 = Y[j], c1 = D[j];
        const double c2 = dx * (3.0 * dy - 2.0 * D[j] - D[j + 1]);
        const double c3 = dx * dx * (-2.0 * dy + D[j] + D[j + 1]);
        const double t = p - left;
        *out = ((c3 * t + c2) * t + c1) * t + c0;
        return true;
    }
};

// ------------------------------------------------------ real harmonics
double real_ylm(double x, double y, double z, double r, int l, int m)
{
    // normalisation constants of nonlocal.cpp:221-240 (15 significant digits)
    static const double k00 = 0.282094791773878;
    static const double k1 = 0.488602511902920;
    static const double k2 = 1.092548430592079, k20 = 0.315391565252520, k22 = 0.546274215296040;
    static const double k33 = 0.590043589926644, k32m = 2.890611442640554;
    static const double k31 = 0.457045799464466, k30 = 0.373176332590115;
    static const double k32p = 1.445305721320277;
    if (l == 0) return k00;
    if (!(std::fabs(r) > kTinyR)) return 0.0;
    const double r2 = r * r, r3 = r * r * r;
    switch (l * 10 + (m + 3)) {
    case 12: return k1 * (y / r);
    case 13: return k1 * (z / r);
    case 14: return k1 * (x / r);
    case 21: return k2 * (x * y) / r2;
    case 22: return k2 * (y * z) / r2;
    case 23: return k20 * (-x * x - y * y + 2 * z * z) / r2;
    case 24: return k2 * (z * x) / r2;
    case 25: return k22 * (x * x - y * y) / r2;
    case 30: return k33 * (3 * x * x - y * y) * y / r3;
    case 31: return k32m * (x * y * z) / r3;
    case 32: return k31 * y * (4 * z * z - x * x - y * y) / r3;
    case 33: return k30 * z * (2 * z * z - 3 * x * x - 3 * y * y) / r3;
    case 34: return k31 * x * (4 * z * z - x * x - y * y) / r3;
    case 35: return k32p * z * (x * x - y * y) / r3;
    case 36: return k33 * x * (x * x - 3 * y * y) / r3;
    default: return 0.0;
    }
}

// ------------------------------------------------- pseudopotential table
struct PsdTable {
    int size = 0, lmax = 0, Z = 0, lloc = 0;
    std::vector<double> r, rVloc, uu;
    std::vector<double> rV[4], rU[4];
    double rc_l[4] = {0, 0, 0, 0};
    double rc = 0, rz = 0;
    double denom[4] = {0, 0, 0, 0};
    Spline sVloc, sV[4], sU[4];
};

bool has_prefix(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }

int read_numbers(std::istream &in, std::vector<double> &dst, int count, double scale)
{
    std::string tok;
    for (int i = 1; i <= count; ++i) {
        if (!(in >> tok)) return SQDFT_E_IO;
        const double v = std::strtod(tok.c_str(), nullptr);
        dst[i] = (scale == 1.0) ? v : scale * v;
    }
    return SQDFT_OK;
}

int load_psd(const std::string &path, const std::string &symbol, int Z, int lloc, double h,
             PsdTable &P)
{
    std::ifstream f(path);
    if (!f) return SQDFT_E_IO;
    std::string line, tok;
    if (!(f >> tok) || tok != symbol) return SQDFT_E_IO;
    std::getline(f, line);
    while (std::getline(f, line) && !has_prefix(line, " Radial grid follows")) {}
    if (!f) return SQDFT_E_IO;
    // the grid runs up to the first "Pseudopotential" token; the table keeps
    // an extra leading row for r = 0 (readfiles.cpp:634-700)
    std::vector<double> grid;
    while (f >> tok && tok != "Pseudopotential") grid.push_back(std::strtod(tok.c_str(), nullptr));
    if (!f) return SQDFT_E_IO;
    const int S = (int)grid.size() + 1;
    P.size = S; P.Z = Z; P.lloc = lloc;
    P.r.assign(S, 0.0);
    for (int i = 1; i < S; ++i) P.r[i] = grid[i - 1];
    P.rVloc.assign(S, 0.0);
    P.uu.assign(S, 0.0);
    for (int l = 0; l < 4; ++l) { P.rV[l].assign(S, 0.0); P.rU[l].assign(S, 0.0); }
    // second pass from the top: every block is identified by its header line
    f.clear();
    f.seekg(0);
    while (std::getline(f, line) && !has_prefix(line, " Radial grid follows")) {}
    for (;;) {
        if (!std::getline(f, line)) break;
        if (has_prefix(line, " Pseudopotential follows")) {
            int l;
            if (!(f >> l) || l < 0 || l > 3) return SQDFT_E_IO;
            P.lmax = l;
            if (read_numbers(f, P.rV[l], S - 1, 0.5)) return SQDFT_E_IO;  // Ry -> Ha
            P.rV[l][0] = P.rV[l][1];
            if (l == lloc) P.rVloc = P.rV[l];
        } else if (has_prefix(line, " Valence charge follows")) {
            std::vector<double> raw(S, 0.0);
            if (read_numbers(f, raw, S - 1, 1.0)) return SQDFT_E_IO;
            for (int i = 1; i < S; ++i) P.uu[i] = raw[i] / (4 * M_PI * P.r[i] * P.r[i]);
            P.uu[0] = P.uu[1];
        } else if (has_prefix(line, " Pseudo-wave-function follows")) {
            int l;
            double zel, rc;
            if (!(f >> l >> zel >> rc) || l < 0 || l > 3) return SQDFT_E_IO;
            P.rc_l[l] = rc;
            if (read_numbers(f, P.rU[l], S - 1, 1.0)) return SQDFT_E_IO;
            P.rU[l][0] = P.rU[l][1];
        }
    }
    // radius where rVloc reaches -Z, and the grid-rounded radii (readfiles.cpp:1110-1124)
    int ir = 0;
    while (ir < S && std::fabs(P.rVloc[ir] + (double)Z) > kTinyR) ++ir;
    P.rz = ir > 0 ? P.r[ir - 1] : 0.0;
    const double rcmax = std::max(std::max(P.rc_l[0], P.rc_l[1]), std::max(P.rc_l[2], P.rc_l[3]));
    P.rc = std::ceil(rcmax / h) * h;
    P.rz = (std::ceil(P.rz / h)) * h;
    P.sVloc.build(P.r.data(), P.rVloc.data(), S);
    for (int l = 0; l < 4; ++l) {
        P.sV[l].build(P.r.data(), P.rV[l].data(), S);
        P.sU[l].build(P.r.data(), P.rU[l].data(), S);
    }
    // KB denominators: rectangle rule with dr = 1e-4 up to rc (readfiles.cpp:1126-1216)
    const double dr = 1e-4;
    const int nr = 1 + (int)std::ceil(P.rc / dr);
    for (int l = 0; l < 4; ++l) {
        double acc = 0.0;
        for (int k = 0; k < nr; ++k) {
            const double rr = k * dr;
            double U, V, Vl;
            if (!P.sU[l].eval(rr, &U) || !P.sV[l].eval(rr, &V) || !P.sVloc.eval(rr, &Vl))
                return SQDFT_E_SPLINE;
            if (rr != 0) acc += U * (V - Vl) * U * dr / (rr * h * h * h);
        }
        P.denom[l] = acc;
    }
    return SQDFT_OK;
}

int nearest_node(double x, double h)
{
    int i = (int)std::round(x / h);
    if (std::fabs(std::fabs(i * h - x) - h * 0.5) < kTinyR) i = (int)std::floor((x / h) - 10 * kTinyR);
    return i;
}

}  // namespace

struct sqdft_system {
    sqdft_info info{};
    std::vector<PsdTable> psd;
    std::vector<sqdft_type_info> tinfo;
    std::vector<double> atom_xyz;  // 3 per atom, type order
    std::vector<int> atom_type;
    std::vector<double> img_xyz;
    std::vector<int> img_type, img_atom, img_nn;
    std::vector<long> img_off;
    std::vector<double> udv;
    std::vector<double> rb;        // per type
    sqdft_scf_params pois{1e-6, 1000, 0.6, 7, 6, 1e-4, 3, 40, 0.1, 7, 2};
};

namespace {

int build_geometry(sqdft_system *S);

int build(const sqdft_spec *spec, sqdft_system *S)
{
    if (!spec || spec->n < 1 || spec->L <= 0 || spec->n_typ < 1 || spec->n_typ > SQDFT_MAX_TYPES ||
        spec->fd_order < 2 || spec->fd_order > 30 || spec->fd_order % 2)
        return SQDFT_E_ARG;
    sqdft_info &I = S->info;
    I.n = spec->n;
    I.L = spec->L;
    I.h = spec->h > 0 ? spec->h : spec->L / spec->n;
    I.FDn = spec->fd_order / 2;
    I.nloc = (int)std::ceil(spec->Rcut / I.h);
    I.Rcut = I.nloc * I.h;
    I.npl = spec->npl;
    const double kB = 8.617343e-5 / 27.211384523;  // Ha/K (readfiles.cpp:309-311)
    I.bet = (1 / (kB * spec->T));
    I.lanczos_tol = spec->lanczos_tol;
    I.fermi_tol = spec->fermi_tol;
    I.fermi_alg = spec->fermi_alg;
    I.n_typ = spec->n_typ;
    // FD weights of the (2 FDn)-th order central Laplacian and gradient
    {
        const int F = I.FDn;
        auto ratio = [](int nn, int k) {
            double num = 1, den = 1;
            for (int i = nn - k + 1; i <= nn; ++i) num *= i;
            for (int i = nn + 1; i <= nn + k; ++i) den *= i;
            return num / den;
        };
        double c0 = 0;
        for (int a = 1; a <= F; ++a) c0 += -(2.0 / (a * a));
        c0 *= 1 / (I.h * I.h);
        I.coeff_lap[0] = c0;
        I.coeff_grad[0] = 0;
        for (int p = 1; p <= F; ++p) {
            const double sgn = std::pow(-1, p + 1);
            I.coeff_lap[p] = (2 * sgn * ratio(F, p) / (p * p * I.h * I.h));
            I.coeff_grad[p] = (sgn * ratio(F, p) / (p * I.h));
        }
    }
    S->psd.resize(spec->n_typ);
    S->tinfo.resize(spec->n_typ);
    I.max_nproj = 0;
    int na = 0;
    for (int t = 0; t < spec->n_typ; ++t) {
        const sqdft_type_spec &T = spec->types[t];
        if (!T.symbol || !T.psd_path || T.natm < 0 || (T.natm > 0 && !T.xyz)) return SQDFT_E_ARG;
        int st = load_psd(T.psd_path, T.symbol, T.Z, T.lloc, I.h, S->psd[t]);
        if (st) return st;
        const PsdTable &P = S->psd[t];
        sqdft_type_info &ti = S->tinfo[t];
        ti.Z = T.Z; ti.lloc = T.lloc; ti.lmax = P.lmax; ti.natm = T.natm;
        ti.rc = P.rc; ti.rz = P.rz; ti.size = P.size;
        ti.nlocc = (int)std::ceil(P.rc / I.h);
        ti.nproj = 0;
        for (int l = 0; l <= P.lmax; ++l)
            if (l != T.lloc) {
                ti.nproj += 2 * l + 1;
                if (P.denom[l] == 0) return SQDFT_E_DENOM;
            }
        for (int l = 0; l < 4; ++l) ti.denom[l] = P.denom[l];
        if (ti.nproj > I.max_nproj) I.max_nproj = ti.nproj;
        // rb is rounded up to a whole number of mesh intervals when the atoms
        // are read (src/readfiles.cpp:512); the replica boxes then use
        // round(rb/h) nodes (src/initialize.cpp:68-70)
        ti.rb = T.rb;   // as given; the replica search uses the rounded value
        S->rb.push_back(T.rb > 0 ? std::ceil(T.rb / I.h) * I.h : T.rb);
        for (int a = 0; a < T.natm; ++a) {
            S->atom_xyz.push_back(T.xyz[3 * a]);
            S->atom_xyz.push_back(T.xyz[3 * a + 1]);
            S->atom_xyz.push_back(T.xyz[3 * a + 2]);
            S->atom_type.push_back(t);
        }
        na += T.natm;
    }
    I.n_atm = na;

    return build_geometry(S);
}

// Images whose rc-cube meets [-nloc, n-1+nloc]^3 and their projector tables,
// from the current atom positions (Replica_atoms_Rcut + NonlocalProjectors).
int build_geometry(sqdft_system *S)
{
    sqdft_info &I = S->info;
    const int na = I.n_atm;
    S->img_xyz.clear(); S->img_nn.clear(); S->img_type.clear(); S->img_atom.clear(); S->img_off.clear();
    // periodic images (initialize.cpp:1003-1196)
    const double L = I.L, h = I.h;
    long off = 0;
    for (int a = 0; a < na; ++a) {
        const int t = S->atom_type[a];
        const double rc = S->psd[t].rc;
        const int nc = S->tinfo[t].nlocc;
        const double *p = &S->atom_xyz[3 * a];
        int lo_shift[3], hi_shift[3];
        for (int d = 0; d < 3; ++d) {
            hi_shift[d] = (int)std::floor(((L + I.Rcut + rc - p[d]) / L) + kTinyR);
            lo_shift[d] = (int)std::ceil(((-(I.Rcut + rc) - p[d]) / L) - kTinyR);
        }
        const long cube = (long)(2 * nc + 1) * (2 * nc + 1) * (2 * nc + 1);
        for (int sz = lo_shift[2]; sz <= hi_shift[2]; ++sz)
            for (int sy = lo_shift[1]; sy <= hi_shift[1]; ++sy)
                for (int sx = lo_shift[0]; sx <= hi_shift[0]; ++sx) {
                    const double q[3] = {p[0] + sx * L, p[1] + sy * L, p[2] + sz * L};
                    int nn[3];
                    bool meets = true;
                    for (int d = 0; d < 3; ++d) {
                        nn[d] = nearest_node(q[d], h);
                        const int lo = std::max(nn[d] - nc, -I.nloc);
                        const int hi = std::min(nn[d] + nc, I.n - 1 + I.nloc);
                        if (lo > hi) meets = false;
                    }
                    if (!meets) continue;
                    for (int d = 0; d < 3; ++d) {
                        S->img_xyz.push_back(q[d]);
                        S->img_nn.push_back(nn[d]);
                    }
                    S->img_type.push_back(t);
                    S->img_atom.push_back(a);
                    S->img_off.push_back(off);
                    off += cube * S->tinfo[t].nproj;
                }
    }
    I.n_img = (int)S->img_type.size();
    I.udv_len = off;

    // projector tables U_l(r) Y_lm (V_l(r) - V_loc(r)) (nonlocal.cpp:330-515)
    S->udv.assign((size_t)off, 0.0);
    for (int J = 0; J < I.n_img; ++J) {
        const int t = S->img_type[J];
        const PsdTable &P = S->psd[t];
        const int nc = S->tinfo[t].nlocc, w = 2 * nc + 1;
        const long cube = (long)w * w * w;
        const double *c = &S->img_xyz[3 * J];
        const int *nn = &S->img_nn[3 * J];
        // radial quantities are shared by all (l,m) of one grid point
        std::vector<double> dx(cube), dy(cube), dz(cube), rad(cube);
        long k = 0;
        for (int kz = -nc; kz <= nc; ++kz)
            for (int ky = -nc; ky <= nc; ++ky)
                for (int kx = -nc; kx <= nc; ++kx, ++k) {
                    dx[k] = (nn[0] + kx) * h - c[0];
                    dy[k] = (nn[1] + ky) * h - c[1];
                    dz[k] = (nn[2] + kz) * h - c[2];
                    rad[k] = std::sqrt(dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k]);
                }
        double *out = S->udv.data() + S->img_off[J];
        int proj = 0;
        for (int l = 0; l <= P.lmax; ++l) {
            if (l == P.lloc) continue;
            for (int m = -l; m <= l; ++m, ++proj) {
                for (long q = 0; q < cube; ++q) {
                    const double r = rad[q];
                    double vloc, vl, ul;
                    if (!P.sVloc.eval(r, &vloc) || !P.sV[l].eval(r, &vl) || !P.sU[l].eval(r, &ul))
                        return SQDFT_E_SPLINE;
                    if (std::fabs(r - 0.0) < kTinyR) {
                        vloc = P.rVloc[1] / P.r[1];
                        vl = P.rV[l][1] / P.r[1];
                        ul = P.rU[l][1] / P.r[1];
                    } else {
                        vloc = vloc / r;
                        vl = vl / r;
                        ul = ul / r;
                    }
                    out[proj * cube + q] = ul * real_ylm(dx[q], dy[q], dz[q], r, l, m) * (vl - vloc);
                }
            }
        }
    }
    return SQDFT_OK;
}

// ------------------------------------------------------------ .input
struct InputKeys {
    double domain[3] = {0, 0, 0};
    int n_int[3] = {0, 0, 0};
    int ncell = 1, FD_order = 12, npl = 0, fermi_alg = 1, rand_seed = 0;
    double T = 0, Rcut = 0, lanczos_tol = 0, fermi_tol = 0, perturb = 0;
    double poisson_tol = 1e-6, beta_aaj = 0.6, scf_tol = 1e-4, beta_scf = 0.1;
    int poisson_maxiter = 1000, m_aaj = 7, p_aaj = 6, scf_miniter = 3, scf_maxiter = 40, m_scf = 7, p_scf = 2;
    std::string atoms_file;
};

std::string join_path(const std::string &dir, const std::string &f)
{
    if (f.empty() || f[0] == '/' || dir.empty()) return f;
    return dir + "/" + f;
}

// Whitespace token cursor with fscanf-like numeric reads: a keyword that is
// followed by a non-number (e.g. a keyword word inside a comment) leaves the
// target unchanged and consumes nothing, as fscanf does in readfiles.cpp.
struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    explicit Tokens(std::istream &in) { std::string w; while (in >> w) t.push_back(w); }
    bool next(std::string &w) { if (i >= t.size()) return false; w = t[i++]; return true; }
    template <typename V> bool num(V &v)
    {
        if (i >= t.size()) return false;
        const char *s = t[i].c_str();
        char *e = nullptr;
        const double d = std::strtod(s, &e);
        if (e == s) return false;
        v = (V)d;
        ++i;
        return true;
    }
};

// Periodic replicas whose rb-box meets the domain [0, n-1]^3 (Replica_atoms,
// src/initialize.cpp:21-200, one rank), in its (type, atom, z, y, x) order.
struct RbImage { double x, y, z; int lo[3], hi[3]; int type, atom; };

int rb_image_list(const sqdft_system *S, std::vector<RbImage> &out)
{
    const sqdft_info &I = S->info;
    const int n = I.n;
    const double h = I.h, L = I.L;
    out.clear();
    size_t ia = 0;
    for (int t = 0; t < I.n_typ; ++t) {
        const double rb = S->rb[t];
        if (!(rb > 0)) return SQDFT_E_ARG;
        const int Rb = (int)std::round(rb / h);
        for (int a = 0; a < S->tinfo[t].natm; ++a, ++ia) {
            const double c[3] = {S->atom_xyz[3 * ia], S->atom_xyz[3 * ia + 1], S->atom_xyz[3 * ia + 2]};
            int pm[3], mm[3];
            for (int d = 0; d < 3; ++d) {
                pm[d] = (int)std::floor(((L + rb - c[d]) / L) + kTinyR);
                mm[d] = (int)std::ceil(((-rb - c[d]) / L) - kTinyR);
            }
            for (int rr = mm[2]; rr <= pm[2]; ++rr)
                for (int qq = mm[1]; qq <= pm[1]; ++qq)
                    for (int pp = mm[0]; pp <= pm[0]; ++pp) {
                        RbImage R;
                        R.x = c[0] + pp * L; R.y = c[1] + qq * L; R.z = c[2] + rr * L;
                        const double A[3] = {R.x, R.y, R.z};
                        bool meet = true;
                        for (int d = 0; d < 3; ++d) {
                            const int nn = nearest_node(A[d], h);
                            R.lo[d] = std::max(nn - Rb, 0);
                            R.hi[d] = std::min(nn + Rb, n - 1);
                            meet = meet && R.lo[d] <= R.hi[d];
                        }
                        if (!meet) continue;
                        R.type = t; R.atom = (int)ia;
                        out.push_back(R);
                    }
        }
    }
    return SQDFT_OK;
}

}  // namespace

extern "C" {

int sqdft_system_create(const sqdft_spec *spec, sqdft_system **out)
{
    if (!out) return SQDFT_E_ARG;
    *out = nullptr;
    auto *S = new sqdft_system();
    const int st = build(spec, S);
    if (st) { delete S; return st; }
    *out = S;
    return SQDFT_OK;
}

int sqdft_system_from_files(const char *dir_c, const char *stem, sqdft_system **out)
{
    if (!dir_c || !stem || !out) return SQDFT_E_ARG;
    *out = nullptr;
    const std::string dir(dir_c);
    std::ifstream in(join_path(dir, std::string(stem) + ".input"));
    if (!in) return SQDFT_E_IO;
    InputKeys K;
    std::string w;
    Tokens tk(in);
    while (tk.next(w)) {
        if (w == "FD_order") tk.num(K.FD_order);
        else if (w == "domain") { tk.num(K.domain[0]) && tk.num(K.domain[1]) && tk.num(K.domain[2]); }
        else if (w == "n_int") { tk.num(K.n_int[0]) && tk.num(K.n_int[1]) && tk.num(K.n_int[2]); }
        else if (w == "atoms_file") tk.next(K.atoms_file);
        else if (w == "ncell") tk.num(K.ncell);
        else if (w == "perturb") tk.num(K.perturb);
        else if (w == "poisson_tol") tk.num(K.poisson_tol);
        else if (w == "poisson_maxiter") tk.num(K.poisson_maxiter);
        else if (w == "beta_aaj") tk.num(K.beta_aaj);
        else if (w == "m_aaj") tk.num(K.m_aaj);
        else if (w == "p_aaj") tk.num(K.p_aaj);
        else if (w == "scf_tol") tk.num(K.scf_tol);
        else if (w == "scf_miniter") tk.num(K.scf_miniter);
        else if (w == "scf_maxiter") tk.num(K.scf_maxiter);
        else if (w == "beta_scf") tk.num(K.beta_scf);
        else if (w == "m_scf") tk.num(K.m_scf);
        else if (w == "p_scf") tk.num(K.p_scf);
        else if (w == "rand_seed") tk.num(K.rand_seed);
        else if (w == "T") tk.num(K.T);
        else if (w == "npl") tk.num(K.npl);
        else if (w == "Rcut") tk.num(K.Rcut);
        else if (w == "lanczos_tol") tk.num(K.lanczos_tol);
        else if (w == "fermi_tol") tk.num(K.fermi_tol);
        else if (w == "fermi_alg") tk.num(K.fermi_alg);
    }
    if (K.n_int[0] < 1 || K.n_int[1] < 1 || K.n_int[2] < 1 || K.atoms_file.empty()) return SQDFT_E_IO;
    // isotropic mesh, per unit cell (readfiles.cpp:294-304)
    const double hx = K.domain[0] / K.n_int[0], hy = K.domain[1] / K.n_int[1],
                 hz = K.domain[2] / K.n_int[2];
    if (std::fabs(hx - hy) >= kTinyR || std::fabs(hx - hz) >= kTinyR || std::fabs(hy - hz) >= kTinyR)
        return SQDFT_E_MESH;
    const double latconst = K.domain[0];
    // .atoms (readfiles.cpp:409-602)
    std::ifstream at(join_path(dir, K.atoms_file + ".atoms"));
    if (!at) return SQDFT_E_IO;
    int n_atm_cell = 0, n_typ = 0, frac = 0;
    Tokens ta(at);
    ta.num(n_atm_cell);
    struct TypeRec { std::string sym, path; int Z = 0, natm = 0, lloc = 0; double mass = 0, rb = 0; std::vector<double> pos; };
    std::vector<TypeRec> types;
    while (ta.next(w)) {
        if (w == "n_typ") ta.num(n_typ);
        else if (w == "frac_coord") ta.num(frac);
        else if (w == "Atoms") {
            for (int t = 0; t < n_typ; ++t) {
                TypeRec R;
                bool ok = ta.next(R.sym) && ta.num(R.Z) && ta.num(R.mass) && ta.num(R.natm) &&
                          ta.num(R.lloc) && ta.num(R.rb) && ta.next(R.path);
                if (!ok || R.natm < 0) return SQDFT_E_IO;
                R.pos.resize(3 * R.natm);
                for (int a = 0; a < 3 * R.natm; ++a)
                    if (!ta.num(R.pos[a])) return SQDFT_E_IO;
                types.push_back(R);
            }
            break;
        }
    }
    if ((int)types.size() != n_typ || n_typ < 1) return SQDFT_E_IO;
    const double scale = frac == 1 ? latconst : 1.0;
    const int nc = K.ncell;
    std::vector<std::vector<double>> xyz(n_typ);
    for (int t = 0; t < n_typ; ++t) {
        const TypeRec &R = types[t];
        std::srand(t + K.rand_seed);  // per-type seed of readfiles.cpp:526
        for (int pp = 0; pp < nc; ++pp)
            for (int qq = 0; qq < nc; ++qq)
                for (int rr = 0; rr < nc; ++rr)
                    for (int a = 0; a < R.natm; ++a) {
                        const double rx = (2 * ((double)(std::rand()) / (double)(RAND_MAX)) - 1);
                        const double ry = (2 * ((double)(std::rand()) / (double)(RAND_MAX)) - 1);
                        const double rz = (2 * ((double)(std::rand()) / (double)(RAND_MAX)) - 1);
                        xyz[t].push_back((double)(pp)*latconst + scale * R.pos[3 * a] + rx * std::fabs(K.perturb));
                        xyz[t].push_back((double)(qq)*latconst + scale * R.pos[3 * a + 1] + ry * std::fabs(K.perturb));
                        xyz[t].push_back((double)(rr)*latconst + scale * R.pos[3 * a + 2] + rz * std::fabs(K.perturb));
                    }
    }
    std::vector<std::string> paths(n_typ);
    std::vector<sqdft_type_spec> ts(n_typ);
    for (int t = 0; t < n_typ; ++t) {
        paths[t] = join_path(dir, types[t].path);
        ts[t].symbol = types[t].sym.c_str();
        ts[t].Z = types[t].Z;
        ts[t].lloc = types[t].lloc;
        ts[t].psd_path = paths[t].c_str();
        ts[t].natm = types[t].natm * nc * nc * nc;
        ts[t].xyz = xyz[t].data();
        ts[t].rb = types[t].rb;
    }
    sqdft_spec sp{};
    sp.n = K.n_int[0] * nc;
    sp.L = K.domain[0] * nc;
    sp.h = hx;
    sp.fd_order = K.FD_order;
    sp.Rcut = K.Rcut;
    sp.npl = K.npl;
    sp.T = K.T;
    sp.lanczos_tol = K.lanczos_tol;
    sp.fermi_tol = K.fermi_tol;
    sp.fermi_alg = K.fermi_alg;
    sp.n_typ = n_typ;
    sp.types = ts.data();
    const int st = sqdft_system_create(&sp, out);
    if (st == SQDFT_OK)
        (*out)->pois = {K.poisson_tol, K.poisson_maxiter, K.beta_aaj, K.m_aaj,    K.p_aaj,
                        K.scf_tol,     K.scf_miniter,     K.scf_maxiter, K.beta_scf, K.m_scf, K.p_scf};
    return st;
}

void sqdft_system_destroy(sqdft_system *S) { delete S; }

int sqdft_system_info(const sqdft_system *S, sqdft_info *info)
{
    if (!S || !info) return SQDFT_E_ARG;
    *info = S->info;
    return SQDFT_OK;
}

int sqdft_system_type_info(const sqdft_system *S, int t, sqdft_type_info *ti)
{
    if (!S || !ti || t < 0 || t >= S->info.n_typ) return SQDFT_E_ARG;
    *ti = S->tinfo[t];
    return SQDFT_OK;
}

int sqdft_system_radial(const sqdft_system *S, int t, double *r, double *rV, double *rU,
                        double *rVloc, double *uu)
{
    if (!S || t < 0 || t >= S->info.n_typ) return SQDFT_E_ARG;
    const PsdTable &P = S->psd[t];
    const size_t n = (size_t)P.size;
    if (r) std::memcpy(r, P.r.data(), n * sizeof(double));
    for (int l = 0; l < 4; ++l) {
        if (rV) std::memcpy(rV + l * n, P.rV[l].data(), n * sizeof(double));
        if (rU) std::memcpy(rU + l * n, P.rU[l].data(), n * sizeof(double));
    }
    if (rVloc) std::memcpy(rVloc, P.rVloc.data(), n * sizeof(double));
    if (uu) std::memcpy(uu, P.uu.data(), n * sizeof(double));
    return SQDFT_OK;
}

int sqdft_system_atoms(const sqdft_system *S, double *xyz, int *type)
{
    if (!S) return SQDFT_E_ARG;
    if (xyz) std::memcpy(xyz, S->atom_xyz.data(), S->atom_xyz.size() * sizeof(double));
    if (type) std::memcpy(type, S->atom_type.data(), S->atom_type.size() * sizeof(int));
    return SQDFT_OK;
}

int sqdft_system_images(const sqdft_system *S, double *xyz, int *type, int *atom, int *nn,
                        long *udv_off)
{
    if (!S) return SQDFT_E_ARG;
    const size_t n = (size_t)S->info.n_img;
    if (xyz) std::memcpy(xyz, S->img_xyz.data(), 3 * n * sizeof(double));
    if (type) std::memcpy(type, S->img_type.data(), n * sizeof(int));
    if (atom) std::memcpy(atom, S->img_atom.data(), n * sizeof(int));
    if (nn) std::memcpy(nn, S->img_nn.data(), 3 * n * sizeof(int));
    if (udv_off) std::memcpy(udv_off, S->img_off.data(), n * sizeof(long));
    return SQDFT_OK;
}

int sqdft_system_udv(const sqdft_system *S, double *udv)
{
    if (!S || !udv) return SQDFT_E_ARG;
    std::memcpy(udv, S->udv.data(), S->udv.size() * sizeof(double));
    return SQDFT_OK;
}

int sqdft_spline_derivs(const double *X, const double *Y, double *YD, int len)
{
    if (!X || !Y || !YD || len < 3) return SQDFT_E_ARG;
    spline_derivs(X, Y, YD, len);
    return SQDFT_OK;
}

int sqdft_spline_eval(const double *X, const double *Y, const double *YD, int len,
                      const double *x, double *y, int nx)
{
    if (!X || !Y || !YD || !x || !y || len < 2) return SQDFT_E_ARG;
    Spline s;
    s.X = X; s.Y = Y; s.n = len;
    s.D.assign(YD, YD + len);
    for (int i = 0; i < nx; ++i)
        if (!s.eval(x[i], &y[i])) return SQDFT_E_SPLINE;
    return SQDFT_OK;
}

int sqdft_system_scf_params(const sqdft_system *S, sqdft_scf_params *pp)
{
    if (!S || !pp) return SQDFT_E_ARG;
    *pp = S->pois;
    return SQDFT_OK;
}

// ChargeDensity (src/initialize.cpp:471-784) over the rb images of
// Replica_atoms (src/initialize.cpp:21-150), one domain covering the whole
// periodic cell.  Per image whose rb-box meets [0, n-1]^3: Vloc_J and the
// smooth reference potential Vtilde_J on the intersection widened by FDn,
// b_J = -(1/4 pi) Lap_FD Vloc_J on the intersection (b~ likewise), Vc +=
// Vtilde_J - Vloc_J, Eself += 1/2 b_J Vloc_J h^3, and rho_at from the
// atomic density table on the intersection.  The products and sums follow
// the reference's order term by term so that the printed energy correction
// is reproduced to its last digit.
int sqdft_system_pseudocharge(const sqdft_system *S, double *b_out, double *rho_out, double *ncharge,
                              double *eself, double *ecorr, double *bt_out, double *vc_out)
{
    if (!S) return SQDFT_E_ARG;
    const sqdft_info &I = S->info;
    const int n = I.n, F = I.FDn;
    const double h = I.h, L = I.L;
    for (int t = 0; t < I.n_typ; ++t)
        if (!(S->rb[t] > 0)) return SQDFT_E_ARG;
    const size_t N3 = (size_t)n * n * n;
    std::vector<double> b(N3, 0.0), bt(N3, 0.0), Vc(N3, 0.0), rho_at(N3, 0.0);
    const double rc_t = 0.5;   // rc_tilda
    const double k4pi = -1 / (4 * M_PI);
    double Eself = 0, Ncharge = 0, Nel = 0, Eself_t = 0, Ncharge_t = 0, Vcon = 0;
    std::vector<RbImage> images;
    if (int st = rb_image_list(S, images)) return st;
    std::vector<Spline> sUU(I.n_typ);
    for (int t = 0; t < I.n_typ; ++t) sUU[t].build(S->psd[t].r.data(), S->psd[t].uu.data(), S->psd[t].size);
    for (const RbImage &R : images) {
        const PsdTable &P = S->psd[R.type];
        const double Z = (double)P.Z;
        const double Ax = R.x, Ay = R.y, Az = R.z;
        const int xs = R.lo[0], xe = R.hi[0], ys = R.lo[1], ye = R.hi[1], zs = R.lo[2], ze = R.hi[2];
        const int nx = xe - xs + 1 + 2 * F, ny = ye - ys + 1 + 2 * F, nz = ze - zs + 1 + 2 * F;
        const int i0 = xs - F, j0 = ys - F, k0 = zs - F;
        std::vector<double> V((size_t)nx * ny * nz), Vt((size_t)nx * ny * nz);
        auto at = [&](int k, int j, int i) { return ((size_t)k * ny + j) * nx + i; };
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double dx = (i0 + i) * h - Ax, dy = (j0 + j) * h - Ay, dz = (k0 + k) * h - Az;
                    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                    double v = 0;
                    if (!P.sVloc.eval(r, &v)) return SQDFT_E_SPLINE;
                    if (std::fabs(r - 0.0) < kTinyR) v = P.rVloc[1] / P.r[1];
                    else v = v / r;
                    V[at(k, j, i)] = v;
                    if (i >= F && i < nx - F && j >= F && j < ny - F && k >= F && k < nz - F) {
                        double u = 0;
                        if (!sUU[R.type].eval(r, &u)) return SQDFT_E_SPLINE;
                        rho_at[((size_t)(k + k0) * n + (j + j0)) * n + (i + i0)] += u;
                        Nel += u * h * h * h;
                    }
                    if (r < rc_t)
                        Vt[at(k, j, i)] = -Z * (9 * std::pow(r, 7) - 30 * rc_t * std::pow(r, 6) +
                                                28 * rc_t * rc_t * std::pow(r, 5) -
                                                14 * std::pow(rc_t, 5) * std::pow(r, 2) + 12 * std::pow(rc_t, 7)) /
                                          (5 * std::pow(rc_t, 8));
                    else
                        Vt[at(k, j, i)] = -Z / r;
                }
        for (int k = zs; k <= ze; ++k)
            for (int j = ys; j <= ye; ++j)
                for (int i = xs; i <= xe; ++i) {
                    const int lk = k - k0, lj = j - j0, li = i - i0;
                    double bv = V[at(lk, lj, li)] * 3 * I.coeff_lap[0] * k4pi;
                    double btv = Vt[at(lk, lj, li)] * 3 * I.coeff_lap[0] * k4pi;
                    for (int s = 1; s <= F; ++s) {
                        bv += (V[at(lk, lj, li - s)] + V[at(lk, lj, li + s)] + V[at(lk, lj - s, li)] +
                               V[at(lk, lj + s, li)] + V[at(lk - s, lj, li)] + V[at(lk + s, lj, li)]) *
                              I.coeff_lap[s] * k4pi;
                        btv += (Vt[at(lk, lj, li - s)] + Vt[at(lk, lj, li + s)] + Vt[at(lk, lj - s, li)] +
                                Vt[at(lk, lj + s, li)] + Vt[at(lk - s, lj, li)] + Vt[at(lk + s, lj, li)]) *
                               I.coeff_lap[s] * k4pi;
                    }
                    const size_t g = ((size_t)k * n + j) * n + i;
                    b[g] += bv;
                    Eself += 0.5 * bv * V[at(lk, lj, li)] * h * h * h;
                    Ncharge += bv * h * h * h;
                    bt[g] += btv;
                    Eself_t += 0.5 * btv * Vt[at(lk, lj, li)] * h * h * h;
                    Ncharge_t += btv * h * h * h;
                    Vc[g] += Vt[at(lk, lj, li)] - V[at(lk, lj, li)];
                }
        if (xs <= 0 && xe >= 0 && ys <= 0 && ye >= 0 && zs <= 0 && ze >= 0)
            Vcon += Vt[at(0 - k0, 0 - j0, 0 - i0)] - V[at(0 - k0, 0 - j0, 0 - i0)];
    }
    // rho_guess = rho_at scaled to -Ncharge electrons; b~ scaled to Ncharge
    double Nel2 = 0;
    for (size_t g = 0; g < N3; ++g) Nel2 += rho_at[g] * h * h * h;
    (void)Nel;
    std::vector<double> rho(N3);
    for (size_t g = 0; g < N3; ++g) rho[g] = std::fabs(Ncharge / Nel2) * rho_at[g];
    for (size_t g = 0; g < N3; ++g) bt[g] = std::fabs(Ncharge / Ncharge_t) * bt[g];
    const double V000 = Vc[0];
    double Ec = 0;
    for (size_t g = 0; g < N3; ++g) {
        Vc[g] += -V000 + Vcon;
        Ec += 0.5 * (bt[g] + b[g]) * Vc[g] * std::pow(h, 3);
    }
    if (b_out) std::memcpy(b_out, b.data(), N3 * sizeof(double));
    if (rho_out) std::memcpy(rho_out, rho.data(), N3 * sizeof(double));
    if (ncharge) *ncharge = Ncharge;
    if (eself) *eself = Eself;
    if (ecorr) *ecorr = Ec + Eself - Eself_t;
    if (bt_out) std::memcpy(bt_out, bt.data(), N3 * sizeof(double));
    if (vc_out) std::memcpy(vc_out, Vc.data(), N3 * sizeof(double));
    return SQDFT_OK;
}

int sqdft_system_rb_images(const sqdft_system *S, int *n_img, double *xyz, int *type, int *atom, int *box)
{
    if (!S || !n_img) return SQDFT_E_ARG;
    std::vector<RbImage> im;
    const int st = rb_image_list(S, im);
    if (st) return st;
    for (size_t m = 0; m < im.size(); ++m) {
        const RbImage &R = im[m];
        if (xyz) { xyz[3 * m] = R.x; xyz[3 * m + 1] = R.y; xyz[3 * m + 2] = R.z; }
        if (type) type[m] = R.type;
        if (atom) atom[m] = R.atom;
        if (box) for (int d = 0; d < 3; ++d) { box[6 * m + d] = R.lo[d]; box[6 * m + 3 + d] = R.hi[d]; }
    }
    *n_img = (int)im.size();
    return SQDFT_OK;
}

int sqdft_system_set_positions(sqdft_system *S, const double *xyz)
{
    if (!S || !xyz) return SQDFT_E_ARG;
    std::memcpy(S->atom_xyz.data(), xyz, S->atom_xyz.size() * sizeof(double));
    return build_geometry(S);
}

int sqdft_system_vloc_deriv(const sqdft_system *S, int t, double *d)
{
    if (!S || !d || t < 0 || t >= S->info.n_typ) return SQDFT_E_ARG;
    const PsdTable &P = S->psd[t];
    spline_derivs(P.r.data(), P.rVloc.data(), d, P.size);
    return SQDFT_OK;
}

int sqdft_system_uu_deriv(const sqdft_system *S, int t, double *d)
{
    if (!S || !d || t < 0 || t >= S->info.n_typ) return SQDFT_E_ARG;
    const PsdTable &P = S->psd[t];
    spline_derivs(P.r.data(), P.uu.data(), d, P.size);
    return SQDFT_OK;
}

}  // extern "C"
