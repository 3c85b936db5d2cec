"""Spectral-quadrature hot path on the GPU (Python view of include/sqdft_gpu.h).

``SpectralQuadrature`` owns one ``sqg_ctx``.  Its methods keep the
reference's call points (phanish-suryanarayana/SQDFT):

====================================  =======================================
method                                reference
====================================  =======================================
``ClenshawCurtisSpectralQuadrature``  sq.cpp:377-834, called at scf.cpp:150
``ForcesNonLocal``                    forces.cpp:318-578, called at forces.cpp:605
``HsubTimesVec``                      sq.cpp:838-929
``node_moments``                      the per-node body sq.cpp:530-661
====================================  =======================================

Errors that the reference reports with ``exit()`` raise ``SQError`` with the
C-ABI status code.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .system import System

PRECISIONS = {"fp64": 0, "mixed": 1, "mixed_lanczos": 2}


def node_order(n: int) -> np.ndarray:
    """The order in which sqg_density visits the n^3 FD nodes and cuts them
    into batches: 4x4x2 bricks of nodes (x fastest inside a brick, bricks x,
    then y, then z), so that consecutive batch slots hold overlapping cubes."""
    out = []
    for z0 in range(0, n, 2):
        for y0 in range(0, n, 4):
            for x0 in range(0, n, 4):
                for z in range(z0, min(z0 + 2, n)):
                    for y in range(y0, min(y0 + 4, n)):
                        for x in range(x0, min(x0 + 4, n)):
                            out.append((z * n + y) * n + x)
    return np.asarray(out, dtype=np.int32)


class SpectralQuadrature:
    def __init__(self, system: System, precision: str = "fp64", batch_nodes: int = 0,
                 device: int = 0, ckpt_interval: int = 8, fermi_tol: float | None = None,
                 fermi_alg: int | None = None, lanczos_tol: float | None = None):
        self._lib = L.load()
        self.system = system
        info = system.info
        self.n = info["n"]
        self.nloc = info["nloc"]
        self.npl = info["npl"]
        self.M = 2 * self.nloc + 1
        self.precision = precision
        p = L.Params()
        p.n = info["n"]
        p.h = info["h"]
        p.FDn = info["FDn"]
        for i, c in enumerate(info["coeff_lap"]):
            p.coeff_lap[i] = c
        for i, c in enumerate(info["coeff_grad"]):
            p.coeff_grad[i] = c
        p.nloc = info["nloc"]
        p.npl = info["npl"]
        p.bet = info["bet"]
        p.lanczos_tol = info["lanczos_tol"] if lanczos_tol is None else lanczos_tol
        p.fermi_tol = info["fermi_tol"] if fermi_tol is None else fermi_tol
        p.fermi_alg = info["fermi_alg"] if fermi_alg is None else int(fermi_alg)
        p.precision = PRECISIONS[precision]
        p.batch_nodes = batch_nodes
        p.ckpt_interval = ckpt_interval
        p.device = device
        self._ctx = C.c_void_p()
        L.check("sqg_create", self._lib.sqg_create(C.byref(p), C.byref(self._ctx)))
        self.set_geometry()
        self.lambda_f = 0.0
        self.last_status = 0

    def set_geometry(self) -> None:
        """Uploads the system's current images and projector tables (after
        System.set_positions); the next density step starts afresh."""
        g = self.system.geometry_arrays()
        L.check("sqg_set_geometry", self._lib.sqg_set_geometry(
            self._ctx, len(g["nlocc"]), L.iptr(g["nlocc"]), L.iptr(g["nproj"]), L.dptr(g["denom"]),
            int(g["n_atm"]), len(g["img_type"]), L.iptr(g["img_type"]), L.iptr(g["img_atom"]),
            L.iptr(g["img_nn"]), L.dptr(g["udv"])))

    def close(self) -> None:
        if self._ctx:
            self._lib.sqg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def _veff(self, veff) -> np.ndarray:
        v = np.ascontiguousarray(np.asarray(veff, dtype=np.float64).reshape(-1))
        if v.size != self.n ** 3:
            raise ValueError(f"veff must have n^3 = {self.n ** 3} entries")
        return v

    def ClenshawCurtisSpectralQuadrature(self, veff, scf_iter: int = 1, ncharge: float | None = None,
                                         lambda_f: float | None = None) -> dict:
        """One SQ density step; returns rho (n^3), lambda_f, Ebs, Eent and stats."""
        v = self._veff(veff)
        if ncharge is None:
            ncharge = -float(self.system.total_valence())
        lf = C.c_double(self.lambda_f if lambda_f is None else lambda_f)
        rho = np.zeros(self.n ** 3)
        ebs = C.c_double()
        eent = C.c_double()
        st = L.Stats()
        code = self._lib.sqg_density(self._ctx, L.dptr(v), int(scf_iter), float(ncharge), C.byref(lf),
                                     L.dptr(rho), C.byref(ebs), C.byref(eent), C.byref(st))
        self.last_status = L.check("sqg_density", code)
        self.lambda_f = lf.value
        return {"rho": rho, "lambda_f": lf.value, "ebs": ebs.value, "eent": eent.value,
                "stats": st.as_dict(), "status": code}

    density = ClenshawCurtisSpectralQuadrature

    def node_data(self) -> dict:
        nn, K = self.n ** 3, self.npl + 1
        chi = np.zeros(nn); zee = np.zeros(nn); mu = np.zeros(nn * K); ci = np.zeros(nn * K)
        it = np.zeros(nn, dtype=np.int32)
        L.check("sqg_get_node_data", self._lib.sqg_get_node_data(
            self._ctx, L.dptr(chi), L.dptr(zee), L.dptr(mu), L.dptr(ci), L.iptr(it)))
        return {"chi": chi, "zee": zee, "mu": mu.reshape(nn, K), "Ci": ci.reshape(nn, K), "iters": it}

    def node_moments(self, veff, nodes) -> dict:
        v = self._veff(veff)
        nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32).reshape(-1))
        nb, K = nd.size, self.npl + 1
        chi = np.zeros(nb); zee = np.zeros(nb); mu = np.zeros(nb * K)
        it = np.zeros(nb, dtype=np.int32)
        L.check("sqg_node_moments", self._lib.sqg_node_moments(
            self._ctx, L.dptr(v), L.iptr(nd), nb, L.dptr(chi), L.dptr(zee), L.dptr(mu), L.iptr(it)))
        return {"chi": chi, "zee": zee, "mu": mu.reshape(nb, K), "iters": it}

    def HsubTimesVec(self, veff, nodes, vecs) -> np.ndarray:
        """H_i v for each node i in ``nodes``; ``vecs`` is (nb, M, M, M) z-major."""
        v = self._veff(veff)
        nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32).reshape(-1))
        x = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64).reshape(nd.size, -1))
        if x.shape[1] != self.M ** 3:
            raise ValueError("each vector must have (2 nloc + 1)^3 entries")
        hv = np.zeros_like(x)
        L.check("sqg_hsub_apply_batch", self._lib.sqg_hsub_apply_batch(
            self._ctx, L.dptr(v), L.iptr(nd), nd.size, L.dptr(x), L.dptr(hv)))
        return hv.reshape(nd.size, self.M, self.M, self.M)

    def ForcesNonLocal(self) -> np.ndarray:
        """Nonlocal forces (n_atm, 3) from the state of the last density step."""
        na = self.system.n_atm
        f = np.zeros(3 * na)
        L.check("sqg_forces_nonlocal", self._lib.sqg_forces_nonlocal(self._ctx, L.dptr(f)))
        return f.reshape(3, na).T.copy()

    # ------------------------------------------------------------------
    # Veff derivatives (chi, zee of the last density step frozen)
    def density_fixed(self, veff, lf_lo: float, lf_hi: float, ncharge: float | None = None) -> dict:
        """Density step with the last step's Lanczos bounds kept; Brent in [lf_lo, lf_hi]."""
        v = self._veff(veff)
        if ncharge is None:
            ncharge = -float(self.system.total_valence())
        lf = C.c_double()
        rho = np.zeros(self.n ** 3)
        L.check("sqg_density_fixed", self._lib.sqg_density_fixed(
            self._ctx, L.dptr(v), float(ncharge), float(lf_lo), float(lf_hi), C.byref(lf), L.dptr(rho)))
        self.lambda_f = lf.value
        return {"rho": rho, "lambda_f": lf.value}

    def adjoint_veff(self, w, fixed_n: bool = True) -> np.ndarray:
        """dL/dVeff for L = sum_i w_i rho_i at the state of the last density step."""
        wv = np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(-1))
        if wv.size != self.n ** 3:
            raise ValueError(f"w must have n^3 = {self.n ** 3} entries")
        out = np.zeros(self.n ** 3)
        L.check("sqg_adjoint_veff", self._lib.sqg_adjoint_veff(self._ctx, L.dptr(wv), int(bool(fixed_n)),
                                                               L.dptr(out)))
        return out

    def tangent_veff(self, dv, fixed_n: bool = True) -> np.ndarray:
        """d rho along dv at the state of the last density step (forward mode)."""
        d = np.ascontiguousarray(np.asarray(dv, dtype=np.float64).reshape(-1))
        if d.size != self.n ** 3:
            raise ValueError(f"dv must have n^3 = {self.n ** 3} entries")
        out = np.zeros(self.n ** 3)
        L.check("sqg_tangent_veff", self._lib.sqg_tangent_veff(self._ctx, L.dptr(d), int(bool(fixed_n)),
                                                               L.dptr(out)))
        return out

    def kernel_timing(self, which: int = 0) -> dict:
        """Device time of the last density call's hot kernels.  which: 0 Chebyshev
        z-march, 1 Lanczos z-march, 2 projections of the Lanczos passes, 3
        projections of the Chebyshev steps."""
        n = C.c_long(); ms = C.c_double(); by = C.c_double()
        L.check("sqg_last_kernel_timing", self._lib.sqg_last_kernel_timing(
            self._ctx, which, C.byref(n), C.byref(ms), C.byref(by)))
        return {"launches": n.value, "total_ms": ms.value, "alg_bytes": by.value}
