"""The SCF step around the SQ call, on the device in FP64.

Mirrors the reference's SQDFT_forces / SCF_iteration sequence
(src/scf.cpp:337-375, :87-200): the pseudocharge b and the starting
density (ChargeDensity, src/initialize.cpp:471-784, host), the AAJ Poisson
solve for phi (PoissonSolver_AAJ, src/poisson.cpp:19-325), the LDA
potential (ExchangeCorrelationPotential, src/scf.cpp:19-64), Veff = phi + Vxc
(EvaluateEffectivePotential, src/scf.cpp:66-82), then per SCF iteration the
SQ density, the total energy (EvaluateTotalEnergy, src/energy.cpp:19-87),
the next Poisson solve from the previous phi and the Veff residual the
reference prints as SCF_err.  ``iteration`` returns the unmixed Veff_new;
``run`` adds the reference's mixing (simple mixing with a Periodic Pulay /
Anderson step every p_scf iterations, src/scf.cpp:200-292) on the host.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L
from .sq import SpectralQuadrature
from .system import System


def gmres(matvec, rhs: np.ndarray, k: int) -> tuple[np.ndarray, float]:
    """k steps of GMRES from x0 = 0 (modified Gram-Schmidt Arnoldi, small
    least-squares solve); returns x and the relative residual."""
    beta0 = float(np.linalg.norm(rhs))
    if beta0 == 0.0:
        return np.zeros_like(rhs), 0.0
    Q = [rhs / beta0]
    Hm = np.zeros((k + 1, k))
    kk = k
    for j in range(k):
        w = matvec(Q[j])
        for i in range(j + 1):
            Hm[i, j] = Q[i] @ w
            w = w - Hm[i, j] * Q[i]
        Hm[j + 1, j] = np.linalg.norm(w)
        if Hm[j + 1, j] <= 1e-14 * beta0:      # invariant subspace: exact solve
            kk = j + 1
            break
        Q.append(w / Hm[j + 1, j])
    e1 = np.zeros(kk + 1)
    e1[0] = beta0
    y = np.linalg.lstsq(Hm[: kk + 1, :kk], e1, rcond=None)[0]
    res = float(np.linalg.norm(Hm[: kk + 1, :kk] @ y - e1)) / beta0
    return sum(y[i] * Q[i] for i in range(kk)), res


def pinv_times_vec(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x = pinv(A) b for a small symmetric matrix as PseudoInverseTimesVec
    (src/anderson.cpp:100-163) computes it: singular values below
    m * max(w) * 2.22044605e-16 are set to zero (SingularValueDecomp,
    src/anderson.cpp:392-397) and zero ones dropped."""
    U, s, Vt = np.linalg.svd(A)
    s = np.where(s < len(s) * s.max() * 2.22044605e-16, 0.0, s)
    c = U.T @ b
    c = np.where(s != 0, c / np.where(s != 0, s, 1.0), 0.0)
    return Vt.T @ c


def anderson_extrapolation(DX: np.ndarray, DF: np.ndarray, res: np.ndarray, beta: float) -> np.ndarray:
    """Anderson update of AndersonExtrapolation (src/anderson.cpp:20-100):
    x = pinv(DF DF') DF res, update = sum_j (DX_j + beta DF_j) x_j."""
    x = pinv_times_vec(DF @ DF.T, DF @ res)
    return (DX + beta * DF).T @ x


class SCF:
    """Owns a SpectralQuadrature context and the SCF state of one geometry."""

    def __init__(self, system: System, precision: str = "fp64", pseudocharge: str = "device", **sq_kw):
        """pseudocharge: "device" (sqg_pseudocharge) or "host"
        (sqdft_system_pseudocharge); the two agree to rounding."""
        self.S = system
        self.sq = SpectralQuadrature(system, precision=precision, **sq_kw)
        self._lib = self.sq._lib
        self._spec = None
        if pseudocharge not in ("device", "host"):
            raise ValueError(f"unknown pseudocharge {pseudocharge!r}")
        pc = self.Pseudocharge() if pseudocharge == "device" else system.pseudocharge()
        self.b, self.rho = pc["b"], pc["rho"]
        self.rho_at = pc.get("rho_at")
        self.ncharge, self.eself, self.ecorr = pc["ncharge"], pc["eself"], pc["ecorr"]
        self.b_tilda, self.vc = pc["b_tilda"], pc["vc"]
        self.pp = system.poisson_params()
        self.n3 = system.n ** 3
        self.phi = None
        self.vxc = None
        self.veff = None
        self.scf_iter = 0
        self.poisson_log = []   # (iterations, relative residual) of every solve
        self.prnt_scf = None    # path: the density of every SCF iteration goes there (.restart layout)

    def close(self):
        self.sq.close()

    # ------------------------------------------------------------ kernels
    def PoissonSolver_AAJ(self, rhs, phi_guess, tol: float | None = None, log: bool = True) -> np.ndarray:
        r = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        g = np.ascontiguousarray(phi_guess, dtype=np.float64).reshape(-1)
        phi = np.zeros(self.n3)
        it, rr = C.c_int(), C.c_double()
        pp = self.pp
        L.check("sqg_poisson_aaj", self._lib.sqg_poisson_aaj(
            self.sq._ctx, L.dptr(r), L.dptr(g), pp["tol"] if tol is None else tol, pp["maxiter"], pp["beta"],
            pp["m"], pp["p"], L.dptr(phi), C.byref(it), C.byref(rr)))
        if log:
            self.poisson_log.append((it.value, rr.value))
        return phi

    def ExchangeCorrelationPotential(self, rho) -> np.ndarray:
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        v = np.zeros(self.n3)
        L.check("sqg_vxc_lda", self._lib.sqg_vxc_lda(self.sq._ctx, L.dptr(r), L.dptr(v)))
        return v

    def EvaluateTotalEnergy(self, rho_new, ebs, eent) -> float:
        """Per atom, with the phi and Vxc of the Veff the SQ call used."""
        e = C.c_double()
        arrs = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (self.phi, self.b, rho_new, self.vxc)]
        L.check("sqg_total_energy", self._lib.sqg_total_energy(
            self.sq._ctx, *[L.dptr(a) for a in arrs], ebs, eent, self.eself, self.ecorr, self.S.n_atm, C.byref(e)))
        return e.value

    def _local_spec(self):
        """The replicas and radial tables of the local pseudopotential passes
        (built once per geometry; the arrays stay referenced by the spec)."""
        if self._spec is None:
            S, nt = self.S, self.S.info["n_typ"]
            tabs = [S.vloc_spline(t) for t in range(nt)]
            keys = ("r", "rVloc", "d", "uu", "duu")
            keep = [np.ascontiguousarray(tb[k], dtype=np.float64) for tb in tabs for k in keys]
            ptrs = lambda k: (L.c_dbl_p * nt)(*[L.dptr(keep[len(keys) * t + k]) for t in range(nt)])  # noqa: E731
            size = np.array([tb["r"].size for tb in tabs], dtype=np.int32)
            Z = np.array([float(S.type_info(t)["Z"]) for t in range(nt)])
            im = S.rb_images()
            xyz = np.ascontiguousarray(im["xyz"], dtype=np.float64).reshape(-1)
            typ, atm = np.ascontiguousarray(im["type"]), np.ascontiguousarray(im["atom"])
            box = np.ascontiguousarray(im["box"], dtype=np.int32).reshape(-1)
            arrs = (ptrs(0), ptrs(1), ptrs(2), ptrs(3), ptrs(4))
            spec = L.LocalSpec(n_typ=nt, size=L.iptr(size), r=arrs[0], rvloc=arrs[1], dvloc=arrs[2], Z=L.dptr(Z),
                               n_img=typ.size, img_xyz=L.dptr(xyz), img_type=L.iptr(typ), img_atom=L.iptr(atm),
                               img_box=L.iptr(box), n_atm=S.n_atm, uu=arrs[3], duu=arrs[4])
            self._spec = (spec, keep, arrs, size, Z, xyz, typ, atm, box)
        return self._spec[0]

    def Pseudocharge(self) -> dict:
        """ChargeDensity on the device (sqg_pseudocharge): b, rho (rho_at
        scaled), ncharge, eself, ecorr, b_tilda, vc."""
        n3 = self.S.n ** 3
        b, rho, bt, vc, ra = np.zeros(n3), np.zeros(n3), np.zeros(n3), np.zeros(n3), np.zeros(n3)
        nc, es, ec = C.c_double(), C.c_double(), C.c_double()
        L.check("sqg_pseudocharge", self._lib.sqg_pseudocharge(
            self.sq._ctx, C.byref(self._local_spec()), L.dptr(b), L.dptr(rho), C.byref(nc), C.byref(es),
            C.byref(ec), L.dptr(bt), L.dptr(vc), L.dptr(ra)))
        return {"b": b, "rho": rho, "ncharge": nc.value, "eself": es.value, "ecorr": ec.value,
                "b_tilda": bt, "vc": vc, "rho_at": ra}

    def ForcesLocal(self) -> tuple[np.ndarray, np.ndarray]:
        """Local forces and their correction term, each (n_atm, 3), from the
        current phi and the pseudocharge (ForcesLocal, src/forces.cpp:20-315)."""
        na = self.S.n_atm
        arrs = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (self.phi, self.b, self.b_tilda, self.vc)]
        fl, fc = np.zeros(3 * na), np.zeros(3 * na)
        L.check("sqg_forces_local", self._lib.sqg_forces_local(self.sq._ctx, C.byref(self._local_spec()),
                                                               *[L.dptr(a) for a in arrs], L.dptr(fl), L.dptr(fc)))
        return fl.reshape(3, na).T.copy(), fc.reshape(3, na).T.copy()

    def ForcesTotal(self) -> np.ndarray:
        """Atomic forces (n_atm, 3) after the SCF: local + nonlocal + the
        correction term, before the net-force removal (src/forces.cpp:580-647)."""
        fl, fc = self.ForcesLocal()
        return fl + self.sq.ForcesNonLocal() + fc

    # ------------------------------------------------------------ response
    def vxc_derivative(self, rho, eta: float = 1e-4) -> np.ndarray:
        """dVxc/drho per grid point, by a central difference in log rho on
        the device LDA kernel (relative step eta, error O(eta^2))."""
        r = np.abs(np.asarray(rho, dtype=np.float64).reshape(-1))
        return (self.ExchangeCorrelationPotential(r * (1 + eta)) -
                self.ExchangeCorrelationPotential(r * (1 - eta))) / (2 * eta * r)

    def response_step(self, res, rho, krylov: int = 4, products: str = "adjoint") -> tuple[np.ndarray, float]:
        """Newton step x of the SCF fixed point Veff = G(rho(Veff)), G = the
        Poisson + LDA map: solves (I - G' chi) x = res with `krylov` GMRES
        steps.  chi v is the fixed-N density response at the last density
        step, from the hand-written adjoint (chi^T v, sqg_adjoint_veff; chi of
        the truncated cubes is symmetric up to the truncation) or the
        tangent (chi v, sqg_tangent_veff); G' d = Poisson(d) + Vxc'(rho) d."""
        D = self.vxc_derivative(rho)
        zero = np.zeros(self.n3)
        resp = self.sq.adjoint_veff if products == "adjoint" else self.sq.tangent_veff

        def mv(v):
            drho = resp(v, fixed_n=True)
            dphi = self.PoissonSolver_AAJ(drho, zero, tol=1e-8, log=False)
            return v - (dphi + D * drho)

        return gmres(mv, np.asarray(res, dtype=np.float64).reshape(-1), krylov)

    # ------------------------------------------------------------ driver
    def set_positions(self, xyz) -> None:
        """New atom positions ((n_atm, 3), type order): images and projector
        tables on the host and the device, and the pseudocharge (what
        SQDFT_forces redoes per MD step, src/scf.cpp:346-350).  The starting
        density becomes rho_at scaled to Ncharge electrons until
        set_density_guess says otherwise."""
        self.S.set_positions(xyz)
        self.sq.set_geometry()
        self._spec = None
        pc = self.Pseudocharge()
        self.b, self.b_tilda, self.vc, self.rho_at = pc["b"], pc["b_tilda"], pc["vc"], pc["rho_at"]
        self.ncharge, self.eself, self.ecorr = pc["ncharge"], pc["eself"], pc["ecorr"]
        self.rho = pc["rho"]
        self.veff = None
        self.scf_iter = 0

    def restart_density(self, path: str, nproc: int | None = None) -> None:
        """Starting density from a .restart file (RestartSCF,
        src/initialize.cpp:675-678), scaled like any guess."""
        from . import files
        self.set_density_guess(files.read_restart(path, self.S.n, nproc), keep_phi=False)

    def set_density_guess(self, rho, keep_phi: bool = True) -> None:
        """Starting density: rho scaled to hold -Ncharge electrons
        (src/initialize.cpp:699-735); keep_phi: the next Poisson solve starts
        from the last phi instead of phi = 1 (src/initialize.cpp:764-765)."""
        r = np.asarray(rho, dtype=np.float64).reshape(-1)
        nel = float(np.sum(r)) * self.S.info["h"] ** 3
        self.rho = abs(self.ncharge / nel) * r
        if not keep_phi:
            self.phi = None

    def initial_potential(self) -> np.ndarray:
        """phi from rho + b (phi_guess = 1, or the last phi after
        set_density_guess(keep_phi=True)), Vxc(rho), Veff (src/scf.cpp:353-372)."""
        guess = np.ones(self.n3) if self.phi is None else self.phi
        self.phi = self.PoissonSolver_AAJ(self.b + self.rho, guess)
        self.vxc = self.ExchangeCorrelationPotential(self.rho)
        self.veff = self.phi + self.vxc
        return self.veff

    def iteration(self) -> dict:
        """One SCF iteration on the current Veff (src/scf.cpp:145-200, without mixing)."""
        if self.veff is None:
            self.initial_potential()
        self.scf_iter += 1
        t0 = time.perf_counter()
        out = self.sq.density(self.veff, scf_iter=self.scf_iter, ncharge=self.ncharge)
        t1 = time.perf_counter()
        etot = self.EvaluateTotalEnergy(out["rho"], out["ebs"], out["eent"])
        if self.prnt_scf:      # PrintSCF after the energy (src/scf.cpp:162-163)
            from . import files
            files.write_restart(self.prnt_scf, out["rho"], self.S.n)
        phi = self.PoissonSolver_AAJ(self.b + out["rho"], self.phi)
        vxc = self.ExchangeCorrelationPotential(out["rho"])
        veff_new = phi + vxc
        err = float(np.sqrt(np.sum((veff_new - self.veff) ** 2)) / np.sqrt(np.sum(veff_new ** 2)))
        t2 = time.perf_counter()
        self.phi, self.vxc, self.rho = phi, vxc, out["rho"]
        return {"lambda_f": out["lambda_f"], "etot": etot, "scf_err": err, "rho": out["rho"], "veff_new": veff_new,
                "ebs": out["ebs"], "eent": out["eent"], "stats": out["stats"],
                "poisson": self.poisson_log[-1], "t_sq": t1 - t0, "t_rest": t2 - t1}

    def run(self, mixing: str = "pulay", scf_tol: float | None = None, krylov: int = 4,
            products: str = "adjoint", progress=None) -> dict:
        """SCF loop to scf_tol (SCF_iteration, src/scf.cpp:84-330): per
        iteration the SQ density, energy, Poisson and Vxc, then the Veff
        update.  mixing="pulay": the reference's Veff_old + beta_scf res with
        an Anderson step every p_scf iterations over the last m_scf
        differences.  mixing="response": the Newton step of response_step
        (no counterpart in the reference).  progress(iteration, scf_err), when
        given, is called after every iteration."""
        if mixing not in ("pulay", "response"):
            raise ValueError(f"unknown mixing {mixing!r}")
        pp = dict(self.pp)
        if scf_tol is not None:
            pp["scf_tol"] = scf_tol
        m, p, beta = pp["m_scf"], pp["p_scf"], pp["beta_scf"]
        if self.veff is None:
            self.initial_potential()
        N = self.n3
        DX, DF = np.zeros((m, N)), np.zeros((m, N))
        Xold, Fold = np.zeros(N), np.zeros(N)
        veff_old = self.veff.copy()
        history = []
        err, count, last = pp["scf_tol"] + 1.0, 1, None
        n_prod = 0
        while err > pp["scf_tol"] and count <= pp["scf_maxiter"]:
            last = self.iteration()
            err = last["scf_err"]
            res = last["veff_new"] - veff_old
            gres = None
            if mixing == "response":
                if err > pp["scf_tol"]:
                    x, gres = self.response_step(res, last["rho"], krylov, products)
                    n_prod += krylov
                    vnew = veff_old + x
                else:
                    vnew = veff_old
            else:
                vnew = veff_old + beta * res
                if count > 1:
                    k = (count - 2) % m
                    DX[k] = veff_old - Xold
                    DF[k] = res - Fold
                Xold, Fold = veff_old.copy(), res.copy()
                if count % p == 0 and count > 1:
                    vnew = vnew - anderson_extrapolation(DX, DF, res, beta)
            veff_old = vnew
            if err > pp["scf_tol"] and count < pp["scf_maxiter"]:
                self.veff = vnew
            history.append({"lambda_f": last["lambda_f"], "etot": last["etot"], "scf_err": err,
                            "poisson": last["poisson"], "fermi_iters": last["stats"]["fermi_iters"],
                            "lanczos0": last["stats"]["lanczos_iters_node0"], "brent": last["stats"]["fermi_iters"],
                            "t_sq": last["t_sq"], "t_rest": last["t_rest"], "gmres_res": gres})
            if progress is not None:
                progress(count, err)
            count += 1
        return {"iterations": count - 1, "converged": err <= pp["scf_tol"], "free_energy": last["etot"],
                "lambda_f": last["lambda_f"], "history": history, "response_products": n_prod}
