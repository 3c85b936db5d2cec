"""ctypes view of the FP64 CPU oracle (``oracle/libsq_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke`` and
the CPU-baseline leg of bench.py, as the checker.  The product path never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsq_oracle.so")

c_int_p = C.POINTER(C.c_int)
c_dbl_p = C.POINTER(C.c_double)
c_long_p = C.POINTER(C.c_long)


class Psd(C.Structure):
    _fields_ = [("size", C.c_int), ("r", c_dbl_p), ("rV", c_dbl_p * 4), ("rU", c_dbl_p * 4),
                ("rVloc", c_dbl_p), ("uu", c_dbl_p), ("DV", c_dbl_p * 4), ("DU", c_dbl_p * 4),
                ("DVloc", c_dbl_p), ("rc_l", C.c_double * 4), ("lmax", C.c_int),
                ("rc", C.c_double), ("rz", C.c_double), ("Denom", C.c_double * 4)]


class Sys(C.Structure):
    _fields_ = [("n", C.c_int), ("L", C.c_double), ("h", C.c_double), ("FDn", C.c_int),
                ("coeff_lap", C.c_double * 16), ("coeff_grad", C.c_double * 16),
                ("nloc", C.c_int), ("Rcut", C.c_double), ("npl", C.c_int), ("bet", C.c_double),
                ("lanczos_tol", C.c_double), ("fermi_tol", C.c_double), ("fermi_alg", C.c_int),
                ("n_typ", C.c_int), ("Z", C.c_int * 8), ("lloc", C.c_int * 8), ("natm", C.c_int * 8),
                ("psd", Psd * 8), ("nlocc", C.c_int * 8), ("nproj", C.c_int * 8),
                ("n_atm", C.c_int), ("ax", c_dbl_p), ("ay", c_dbl_p), ("az", c_dbl_p),
                ("n_img", C.c_int), ("ix", c_dbl_p), ("iy", c_dbl_p), ("iz", c_dbl_p),
                ("ityp", c_int_p), ("iatm", c_int_p), ("inn", c_int_p), ("iudv", c_long_p),
                ("udv", c_dbl_p), ("udv_len", C.c_long)]


_lib = None


def _cpu_flags() -> set:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def use_timing_build() -> str:
    """Switches this process to an -O3 timing build of the same source (for
    the CPU baseline only; the checker stays the -ffp-contract=off build).
    Must be called before the first load(); returns the library path."""
    global LIB
    if _lib is not None:
        raise RuntimeError("oracle already loaded; call use_timing_build() first")
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= _cpu_flags()
    LIB = os.path.join(HERE, "libsq_oracle_v4.so" if v4 else "libsq_oracle_v3.so")
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(LIB)
    P = C.POINTER(Sys)
    sig = {
        "sqo_fract": (C.c_double, [C.c_int, C.c_int]),
        "sqo_fd_coeffs": (None, [C.c_int, C.c_double, c_dbl_p, c_dbl_p]),
        "sqo_spline_yd": (None, [c_dbl_p, c_dbl_p, c_dbl_p, C.c_int]),
        "sqo_spline_eval": (C.c_int, [c_dbl_p, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p, C.c_int, c_dbl_p]),
        "sqo_sph_harm": (C.c_double, [C.c_double] * 4 + [C.c_int, C.c_int]),
        "sqo_tridiag_eig": (C.c_int, [c_dbl_p, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p]),
        "sqo_func": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double]),
        "sqo_cheb_coeff": (None, [C.c_int, C.c_double, C.c_double, c_dbl_p, C.c_int]),
        "sqo_sys_build": (C.c_int, [P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), c_int_p, c_int_p,
                                    c_int_p, c_dbl_p, c_dbl_p, c_dbl_p]),
        "sqo_sys_free": (None, [P]),
        "sqo_hsub": (C.c_int, [P, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p]),
        "sqo_vnl_atom": (C.c_int, [P, C.c_int, C.c_int, c_dbl_p, c_dbl_p]),
        "sqo_lanczos": (C.c_int, [P, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p, c_int_p]),
        "sqo_node_moments": (C.c_int, [P, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                       c_int_p]),
        "sqo_moments_nodes": (C.c_int, [P, c_dbl_p, c_int_p, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p,
                                        C.c_int]),
        "sqo_moments_fixed": (C.c_int, [P, c_dbl_p, C.c_int, C.c_double, C.c_double, c_dbl_p]),
        "sqo_density_fixed": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, C.c_double, C.c_double,
                                        c_dbl_p, c_dbl_p, c_dbl_p, C.c_int]),
        "sqo_adjoint_veff": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, c_dbl_p, C.c_int,
                                       c_dbl_p, C.c_int]),
        "sqo_tangent_veff": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, c_dbl_p, C.c_int,
                                       c_dbl_p, C.c_int]),
        "sqo_ne_constraint": (C.c_double, [P, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, C.c_double]),
        "sqo_brent": (C.c_int, [P, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, C.c_double, C.c_double,
                                C.c_int, C.c_double, c_dbl_p, c_int_p]),
        "sqo_newton": (C.c_int, [P, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, C.c_double, c_dbl_p,
                                 c_int_p]),
        "sqo_density": (C.c_int, [P, c_dbl_p, C.c_int, C.c_double, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                  c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p, c_int_p, C.c_int]),
        "sqo_rho_energy": (C.c_int, [P, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, c_dbl_p, c_dbl_p,
                                     c_dbl_p, c_dbl_p]),
        "sqo_forces_nonlocal": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p, C.c_int, c_dbl_p,
                                          C.c_int]),
        "sqo_poisson_aaj": (C.c_int, [C.c_int, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, C.c_int,
                                      C.c_double, C.c_int, C.c_int, c_dbl_p, c_int_p, c_dbl_p]),
        "sqo_vxc_lda": (None, [C.c_size_t, c_dbl_p, c_dbl_p]),
        "sqo_charge_density": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
        "sqo_forces_local": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
        "sqo_total_energy": (C.c_double, [C.c_size_t, C.c_double, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_int]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(c_dbl_p)


def _i(a):
    return a.ctypes.data_as(c_int_p)


class OracleSystem:
    """FP64 oracle system built from the same inputs as ``sqdft_amd.System``."""

    def __init__(self, n, L_box, types, Rcut, npl, T, h=0.0, fd_order=12, lanczos_tol=1e-6,
                 fermi_tol=1e-6, fermi_alg=1):
        lib = load()
        self._lib = lib
        self.S = Sys()
        nt = len(types)
        paths = (C.c_char_p * nt)(*[str(t.psd_path).encode() for t in types])
        syms = (C.c_char_p * nt)(*[t.symbol.encode() for t in types])
        Z = np.array([t.Z for t in types], dtype=np.int32)
        lloc = np.array([t.lloc for t in types], dtype=np.int32)
        xyz = [np.asarray(t.xyz, dtype=np.float64).reshape(-1, 3) for t in types]
        natm = np.array([x.shape[0] for x in xyz], dtype=np.int32)
        allx = np.concatenate(xyz, axis=0)
        ax, ay, az = (np.ascontiguousarray(allx[:, k]) for k in range(3))
        st = lib.sqo_sys_build(C.byref(self.S), n, L_box, h, fd_order // 2, Rcut, npl, T, lanczos_tol,
                               fermi_tol, fermi_alg, nt, paths, syms, _i(Z), _i(lloc), _i(natm),
                               _d(ax), _d(ay), _d(az))
        if st != 0:
            raise RuntimeError(f"sqo_sys_build failed: {st}")
        self.n = n
        self.nloc = self.S.nloc
        self.npl = npl
        self.M = 2 * self.nloc + 1

    @classmethod
    def like(cls, system, types, Rcut, npl, T, h):
        info = system.info
        return cls(info["n"], info["L"], types, Rcut, npl, T, h=h, lanczos_tol=info["lanczos_tol"],
                   fermi_tol=info["fermi_tol"], fermi_alg=info["fermi_alg"])

    def __del__(self):
        try:
            self._lib.sqo_sys_free(C.byref(self.S))
        except Exception:
            pass

    # --------------------------------------------------------------- views
    def images(self) -> dict:
        m = self.S.n_img
        g = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy()  # noqa: E731
        return {"xyz": np.stack([g(self.S.ix, m), g(self.S.iy, m), g(self.S.iz, m)], axis=1),
                "type": g(self.S.ityp, m), "atom": g(self.S.iatm, m),
                "nn": g(self.S.inn, 3 * m).reshape(-1, 3), "udv_off": g(self.S.iudv, m)}

    def udv(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.S.udv, shape=(self.S.udv_len,)).copy()

    def denom(self, t: int) -> np.ndarray:
        return np.array(self.S.psd[t].Denom[:])

    # --------------------------------------------------------------- path
    def hsub(self, veff, node, v) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        hv = np.zeros_like(v)
        st = self._lib.sqo_hsub(C.byref(self.S), _d(np.ascontiguousarray(veff, dtype=np.float64)), int(node),
                                _d(v), _d(hv))
        if st < 0:
            raise RuntimeError(f"sqo_hsub: {st}")
        return hv

    def vnl_atom(self, node, atom, v) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        out = np.zeros_like(v)
        self._lib.sqo_vnl_atom(C.byref(self.S), int(node), int(atom), _d(v), _d(out))
        return out

    def lanczos(self, veff, node):
        lo, hi, it = C.c_double(), C.c_double(), C.c_int()
        st = self._lib.sqo_lanczos(C.byref(self.S), _d(np.ascontiguousarray(veff)), int(node), C.byref(lo),
                                   C.byref(hi), C.byref(it))
        if st < 0:
            raise RuntimeError(f"sqo_lanczos: {st}")
        return lo.value, hi.value, it.value

    def node_moments(self, veff, node) -> dict:
        chi, zee, lo, hi = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        mu = np.zeros(self.npl + 1)
        it = C.c_int()
        st = self._lib.sqo_node_moments(C.byref(self.S), _d(np.ascontiguousarray(veff)), int(node),
                                        C.byref(chi), C.byref(zee), C.byref(lo), C.byref(hi), _d(mu),
                                        C.byref(it))
        if st < 0:
            raise RuntimeError(f"sqo_node_moments: {st}")
        return {"chi": chi.value, "zee": zee.value, "mu": mu, "iters": it.value, "lmin": lo.value,
                "lmax": hi.value}

    def moments_nodes(self, veff, nodes, nthreads=0) -> dict:
        nd = np.ascontiguousarray(nodes, dtype=np.int32)
        K = self.npl + 1
        chi = np.zeros(nd.size); zee = np.zeros(nd.size); mu = np.zeros(nd.size * K)
        it = np.zeros(nd.size, dtype=np.int32)
        st = self._lib.sqo_moments_nodes(C.byref(self.S), _d(np.ascontiguousarray(veff)), _i(nd), nd.size,
                                         _d(chi), _d(zee), _d(mu), _i(it), int(nthreads))
        if st < 0:
            raise RuntimeError(f"sqo_moments_nodes: {st}")
        return {"chi": chi, "zee": zee, "mu": mu.reshape(-1, K), "iters": it}

    def density(self, veff, scf_iter=1, ncharge=None, lambda_f=0.0, nthreads=0) -> dict:
        nn, K = self.n ** 3, self.npl + 1
        if ncharge is None:
            ncharge = -float(sum(self.S.Z[t] * self.S.natm[t] for t in range(self.S.n_typ)))
        lf = C.c_double(lambda_f)
        rho = np.zeros(nn); chi = np.zeros(nn); zee = np.zeros(nn); mu = np.zeros(nn * K)
        ci = np.zeros(nn * K); its = np.zeros(nn, dtype=np.int32)
        ebs, eent, fit = C.c_double(), C.c_double(), C.c_int()
        st = self._lib.sqo_density(C.byref(self.S), _d(np.ascontiguousarray(veff)), int(scf_iter),
                                   float(ncharge), C.byref(lf), _d(rho), C.byref(ebs), C.byref(eent),
                                   _d(chi), _d(zee), _d(mu), _d(ci), _i(its), C.byref(fit), int(nthreads))
        return {"status": st, "rho": rho, "lambda_f": lf.value, "ebs": ebs.value, "eent": eent.value,
                "chi": chi, "zee": zee, "mu": mu.reshape(nn, K), "Ci": ci.reshape(nn, K), "iters": its,
                "fermi_iters": fit.value}

    def density_fixed(self, veff, chi, zee, lf_lo, lf_hi, ncharge=None, nthreads=0) -> dict:
        nn, K = self.n ** 3, self.npl + 1
        if ncharge is None:
            ncharge = -float(sum(self.S.Z[t] * self.S.natm[t] for t in range(self.S.n_typ)))
        lf = C.c_double()
        rho = np.zeros(nn); mu = np.zeros(nn * K)
        st = self._lib.sqo_density_fixed(C.byref(self.S), _d(np.ascontiguousarray(veff)),
                                         _d(np.ascontiguousarray(chi)), _d(np.ascontiguousarray(zee)),
                                         float(ncharge), lf_lo, lf_hi, C.byref(lf), _d(rho), _d(mu), int(nthreads))
        if st < 0:
            raise RuntimeError(f"sqo_density_fixed: {st}")
        return {"rho": rho, "lambda_f": lf.value, "mu": mu.reshape(nn, K)}

    def adjoint_veff(self, veff, chi, zee, mu, lambda_f, w, fixed_n=1, nthreads=0) -> np.ndarray:
        out = np.zeros(self.n ** 3)
        st = self._lib.sqo_adjoint_veff(C.byref(self.S), _d(np.ascontiguousarray(veff)),
                                        _d(np.ascontiguousarray(chi)), _d(np.ascontiguousarray(zee)),
                                        _d(np.ascontiguousarray(mu)), lambda_f, _d(np.ascontiguousarray(w)),
                                        int(fixed_n), _d(out), int(nthreads))
        if st < 0:
            raise RuntimeError(f"sqo_adjoint_veff: {st}")
        return out

    def tangent_veff(self, veff, chi, zee, mu, lambda_f, dv, fixed_n=1, nthreads=0) -> np.ndarray:
        out = np.zeros(self.n ** 3)
        st = self._lib.sqo_tangent_veff(C.byref(self.S), _d(np.ascontiguousarray(veff)),
                                        _d(np.ascontiguousarray(chi)), _d(np.ascontiguousarray(zee)),
                                        _d(np.ascontiguousarray(mu)), lambda_f, _d(np.ascontiguousarray(dv)),
                                        int(fixed_n), _d(out), int(nthreads))
        if st < 0:
            raise RuntimeError(f"sqo_tangent_veff: {st}")
        return out

    def forces_nonlocal(self, veff, chi, zee, Ci, nodes=None, nthreads=0) -> np.ndarray:
        na = self.S.n_atm
        f = np.zeros(3 * na)
        if nodes is None:
            np_nodes, cnt = None, 0
            ptr = C.cast(None, c_int_p)
        else:
            np_nodes = np.ascontiguousarray(nodes, dtype=np.int32)
            cnt = np_nodes.size
            ptr = _i(np_nodes)
        st = self._lib.sqo_forces_nonlocal(C.byref(self.S), _d(np.ascontiguousarray(veff)),
                                           _d(np.ascontiguousarray(chi)), _d(np.ascontiguousarray(zee)),
                                           _d(np.ascontiguousarray(Ci)), ptr, cnt, _d(f), int(nthreads))
        if st < 0:
            raise RuntimeError(f"sqo_forces_nonlocal: {st}")
        return f.reshape(3, na).T.copy()

    def charge_density(self, rb) -> dict:
        """ChargeDensity at the first MD step (initialize.cpp:471-784)."""
        N = self.n ** 3
        rb = np.ascontiguousarray(rb, dtype=np.float64)
        b, bt, vc, rho, sc = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(5)
        st = self._lib.sqo_charge_density(C.byref(self.S), _d(rb), _d(b), _d(bt), _d(vc), _d(rho), _d(sc))
        if st < 0:
            raise RuntimeError(f"sqo_charge_density: {st}")
        return {"b": b, "b_tilda": bt, "vc": vc, "rho": rho, "ncharge": sc[0], "eself": sc[1],
                "ncharge_tilda": sc[2], "eself_tilda": sc[3], "ecorr": sc[4]}

    def forces_local(self, rb, phi, b, bt, vc):
        """ForcesLocal (forces.cpp:20-315): (floc, fcorr), each (n_atm, 3)."""
        na = self.S.n_atm
        fl, fc = np.zeros(3 * na), np.zeros(3 * na)
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (rb, phi, b, bt, vc)]
        st = self._lib.sqo_forces_local(C.byref(self.S), *[_d(a) for a in arr], _d(fl), _d(fc))
        if st < 0:
            raise RuntimeError(f"sqo_forces_local: {st}")
        return fl.reshape(3, na).T.copy(), fc.reshape(3, na).T.copy()


def cheb_coeff(npl, lfp, bet, func=0) -> np.ndarray:
    c = np.zeros(npl + 1)
    load().sqo_cheb_coeff(npl, lfp, bet, _d(c), func)
    return c


def tridiag_eig(diag, sub):
    lo, hi = C.c_double(), C.c_double()
    d = np.ascontiguousarray(diag, dtype=np.float64)
    s = np.ascontiguousarray(sub, dtype=np.float64)
    st = load().sqo_tridiag_eig(_d(d), _d(s), d.size, C.byref(lo), C.byref(hi))
    return st, lo.value, hi.value


def fd_coeffs(FDn, h):
    lap, grad = np.zeros(16), np.zeros(16)
    load().sqo_fd_coeffs(FDn, h, _d(lap), _d(grad))
    return lap[: FDn + 1], grad[: FDn + 1]


def poisson_aaj(n, FDn, lap, rhs, phi_guess, tol=1e-6, maxiter=1000, beta=0.6, m=7, p=6):
    """AAJ solve of -Lap_FD phi = 4 pi rhs (poisson.cpp:19-325): (phi, iters, relres)."""
    lp = np.ascontiguousarray(np.asarray(lap, dtype=np.float64))
    r = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64).reshape(-1))
    g = np.ascontiguousarray(np.asarray(phi_guess, dtype=np.float64).reshape(-1))
    phi = np.zeros(n ** 3)
    it, rr = C.c_int(), C.c_double()
    st = load().sqo_poisson_aaj(n, FDn, _d(lp), _d(r), _d(g), tol, maxiter, beta, m, p, _d(phi), C.byref(it),
                                C.byref(rr))
    if st < 0:
        raise RuntimeError(f"sqo_poisson_aaj: {st}")
    return phi, it.value, rr.value


def vxc_lda(rho):
    """LDA exchange-correlation potential of scf.cpp:19-64."""
    r = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1))
    v = np.zeros_like(r)
    load().sqo_vxc_lda(r.size, _d(r), _d(v))
    return v


def total_energy(h, phi, b, rho, vxc, ebs, eent, eself, ecorr, n_atm):
    """Total energy per atom of EvaluateTotalEnergy (energy.cpp:19-87)."""
    a = [np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1)) for x in (phi, b, rho, vxc)]
    return load().sqo_total_energy(a[0].size, h, *[_d(x) for x in a], ebs, eent, eself, ecorr, n_atm)


def periodic_pulay(DX, DF, res, beta):
    """The Anderson update of AndersonExtrapolation (anderson.cpp:20-100) in
    the reference's loop order: FtF_ij = sum_k DF_i[k] DF_j[k], Ftf_j =
    sum_k DF_j[k] res[k], x = pinv(FtF) Ftf with the singular values below
    m * max(w) * 2.22044605e-16 zeroed (anderson.cpp:392-397) and zero ones
    dropped (anderson.cpp:134-142), am[k] = sum_j (DX_j[k] + beta DF_j[k]) x_j."""
    m = DF.shape[0]
    FtF = np.array([[float(np.dot(DF[i], DF[j])) for j in range(m)] for i in range(m)])
    Ftf = np.array([float(np.dot(DF[j], res)) for j in range(m)])
    U, s, Vt = np.linalg.svd(FtF)
    s = np.where(s < m * s.max() * 2.22044605e-16, 0.0, s)
    c = np.array([(U[:, j] @ Ftf) / s[j] if s[j] != 0 else 0.0 for j in range(m)])
    x = Vt.T @ c
    am = np.zeros_like(res)
    for j in range(m):
        am += (DX[j] + beta * DF[j]) * x[j]
    return am
