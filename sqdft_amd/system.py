"""Host-side system description: geometry, pseudopotentials, images, projectors.

Python view of ``include/sqdft_system.h``.  ``System.build`` replaces the
reference's input chain that feeds the SQ kernels (Read_psd,
Replica_atoms_Rcut and NonlocalProjectors: src/readfiles.cpp:605-1225,
src/initialize.cpp:1003-1196, src/nonlocal.cpp:317-518); ``System.from_files``
reads a ``.input``/``.atoms`` pair with the keyword semantics of
src/readfiles.cpp:86-602.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib as L


@dataclass
class AtomType:
    symbol: str
    Z: int
    lloc: int
    psd_path: str
    xyz: np.ndarray          # (natm, 3) Cartesian, bohr
    rb: float = 0.0          # pseudocharge radius (bohr); 0: no pseudocharge


class System:
    """Owns a ``sqdft_system`` handle."""

    def __init__(self, handle: C.c_void_p):
        self._lib = L.load()
        self._h = handle
        info = L.Info()
        L.check("sqdft_system_info", self._lib.sqdft_system_info(self._h, C.byref(info)))
        self._info = info

    # ------------------------------------------------------------ builders
    @classmethod
    def build(cls, n: int, L_box: float, types: Sequence[AtomType], Rcut: float, npl: int, T: float,
              fd_order: int = 12, lanczos_tol: float = 1e-6, fermi_tol: float = 1e-6,
              fermi_alg: int = 1, h: float = 0.0) -> "System":
        lib = L.load()
        keep = []
        ts = (L.TypeSpec * len(types))()
        for i, t in enumerate(types):
            xyz = np.ascontiguousarray(np.asarray(t.xyz, dtype=np.float64).reshape(-1, 3))
            keep.append(xyz)
            ts[i].symbol = t.symbol.encode()
            ts[i].Z = int(t.Z)
            ts[i].lloc = int(t.lloc)
            ts[i].psd_path = str(t.psd_path).encode()
            ts[i].natm = xyz.shape[0]
            ts[i].xyz = L.dptr(xyz)
            ts[i].rb = float(t.rb)
        spec = L.Spec(n=n, L=L_box, h=h, fd_order=fd_order, Rcut=Rcut, npl=npl, T=T,
                      lanczos_tol=lanczos_tol, fermi_tol=fermi_tol, fermi_alg=fermi_alg,
                      n_typ=len(types), types=ts)
        hnd = C.c_void_p()
        L.check("sqdft_system_create", lib.sqdft_system_create(C.byref(spec), C.byref(hnd)))
        return cls(hnd)

    @classmethod
    def from_files(cls, directory: str, stem: str) -> "System":
        lib = L.load()
        hnd = C.c_void_p()
        L.check("sqdft_system_from_files",
                lib.sqdft_system_from_files(directory.encode(), stem.encode(), C.byref(hnd)))
        return cls(hnd)

    def close(self) -> None:
        if self._h:
            self._lib.sqdft_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ accessors
    @property
    def info(self) -> dict:
        i = self._info
        d = {k: getattr(i, k) for k, _ in L.Info._fields_ if k not in ("coeff_lap", "coeff_grad")}
        d["coeff_lap"] = np.array(i.coeff_lap[: i.FDn + 1])
        d["coeff_grad"] = np.array(i.coeff_grad[: i.FDn + 1])
        return d

    @property
    def n(self) -> int:
        return self._info.n

    @property
    def nn(self) -> int:
        return self._info.n ** 3

    @property
    def n_atm(self) -> int:
        return self._info.n_atm

    def type_info(self, t: int) -> dict:
        ti = L.TypeInfo()
        L.check("sqdft_system_type_info", self._lib.sqdft_system_type_info(self._h, t, C.byref(ti)))
        d = {k: getattr(ti, k) for k, _ in L.TypeInfo._fields_ if k != "denom"}
        d["denom"] = np.array(ti.denom[:])
        return d

    def radial(self, t: int) -> dict:
        size = self.type_info(t)["size"]
        r = np.zeros(size); rv = np.zeros(4 * size); ru = np.zeros(4 * size)
        vl = np.zeros(size); uu = np.zeros(size)
        L.check("sqdft_system_radial", self._lib.sqdft_system_radial(
            self._h, t, L.dptr(r), L.dptr(rv), L.dptr(ru), L.dptr(vl), L.dptr(uu)))
        return {"r": r, "rV": rv.reshape(4, size), "rU": ru.reshape(4, size), "rVloc": vl, "uu": uu}

    def atoms(self) -> tuple[np.ndarray, np.ndarray]:
        xyz = np.zeros(3 * self.n_atm); typ = np.zeros(self.n_atm, dtype=np.int32)
        L.check("sqdft_system_atoms", self._lib.sqdft_system_atoms(self._h, L.dptr(xyz), L.iptr(typ)))
        return xyz.reshape(-1, 3), typ

    def images(self) -> dict:
        m = self._info.n_img
        xyz = np.zeros(3 * max(m, 1)); typ = np.zeros(max(m, 1), dtype=np.int32)
        atm = np.zeros(max(m, 1), dtype=np.int32); nn = np.zeros(3 * max(m, 1), dtype=np.int32)
        off = np.zeros(max(m, 1), dtype=np.int64)
        L.check("sqdft_system_images", self._lib.sqdft_system_images(
            self._h, L.dptr(xyz), L.iptr(typ), L.iptr(atm), L.iptr(nn), L.lptr(off)))
        return {"xyz": xyz[: 3 * m].reshape(-1, 3), "type": typ[:m], "atom": atm[:m],
                "nn": nn[: 3 * m].reshape(-1, 3), "udv_off": off[:m]}

    def udv(self) -> np.ndarray:
        u = np.zeros(max(self._info.udv_len, 1))
        L.check("sqdft_system_udv", self._lib.sqdft_system_udv(self._h, L.dptr(u)))
        return u[: self._info.udv_len]

    def pseudocharge(self) -> dict:
        """Pseudocharge b, starting density rho (rho_at scaled to -Ncharge
        electrons), Ncharge, Eself, the energy correction Ecorr, and b_tilda,
        vc for the local forces (ChargeDensity, src/initialize.cpp:471-784);
        arrays are n^3, x fastest."""
        n3 = self.n ** 3
        b, rho, bt, vc = np.zeros(n3), np.zeros(n3), np.zeros(n3), np.zeros(n3)
        nc = C.c_double(); es = C.c_double(); ec = C.c_double()
        L.check("sqdft_system_pseudocharge", self._lib.sqdft_system_pseudocharge(
            self._h, L.dptr(b), L.dptr(rho), C.byref(nc), C.byref(es), C.byref(ec), L.dptr(bt), L.dptr(vc)))
        return {"b": b, "rho": rho, "ncharge": nc.value, "eself": es.value, "ecorr": ec.value,
                "b_tilda": bt, "vc": vc}

    def set_positions(self, xyz) -> None:
        """Moves the atoms ((n_atm, 3), type order) and rebuilds the images and
        projector tables."""
        x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1))
        if x.size != 3 * self.n_atm:
            raise ValueError(f"xyz must hold 3 * n_atm = {3 * self.n_atm} values")
        L.check("sqdft_system_set_positions", self._lib.sqdft_system_set_positions(self._h, L.dptr(x)))
        info = L.Info()   # image count and table length change with the positions
        L.check("sqdft_system_info", self._lib.sqdft_system_info(self._h, C.byref(info)))
        self._info = info

    def rb_images(self) -> dict:
        """Replicas whose rb-box meets the domain (Replica_atoms,
        src/initialize.cpp:21-200): xyz, type, atom, box (lo xyz, hi xyz)."""
        cnt = C.c_int()
        null = C.POINTER(C.c_double)(), C.POINTER(C.c_int)()
        L.check("sqdft_system_rb_images", self._lib.sqdft_system_rb_images(
            self._h, C.byref(cnt), null[0], null[1], null[1], null[1]))
        m = cnt.value
        xyz = np.zeros(3 * m); typ = np.zeros(m, np.int32); atm = np.zeros(m, np.int32)
        box = np.zeros(6 * m, np.int32)
        L.check("sqdft_system_rb_images", self._lib.sqdft_system_rb_images(
            self._h, C.byref(cnt), L.dptr(xyz), L.iptr(typ), L.iptr(atm), L.iptr(box)))
        return {"xyz": xyz.reshape(m, 3), "type": typ, "atom": atm, "box": box.reshape(m, 6)}

    def vloc_spline(self, t: int) -> dict:
        """r, r V_loc(r), uu(r) and their spline derivative tables for type t."""
        rad = self.radial(t)
        d, du = np.zeros(rad["r"].size), np.zeros(rad["r"].size)
        L.check("sqdft_system_vloc_deriv", self._lib.sqdft_system_vloc_deriv(self._h, t, L.dptr(d)))
        L.check("sqdft_system_uu_deriv", self._lib.sqdft_system_uu_deriv(self._h, t, L.dptr(du)))
        return {"r": rad["r"], "rVloc": rad["rVloc"], "d": d, "uu": rad["uu"], "duu": du}

    def poisson_params(self) -> dict:
        """AAJ and SCF parameters of the input: tol, maxiter, beta, m, p of the
        Poisson solver; scf_tol, scf_miniter, scf_maxiter, beta_scf, m_scf, p_scf."""
        pp = L.ScfParams()
        L.check("sqdft_system_scf_params", self._lib.sqdft_system_scf_params(self._h, C.byref(pp)))
        return {k: getattr(pp, k) for k, _ in pp._fields_}

    def total_valence(self) -> int:
        return int(sum(self.type_info(t)["Z"] * self.type_info(t)["natm"] for t in range(self._info.n_typ)))

    def geometry_arrays(self) -> dict:
        """Flat arrays in the layout ``sqg_set_geometry`` expects."""
        nt = self._info.n_typ
        nlocc = np.zeros(nt, dtype=np.int32)
        nproj = np.zeros(nt, dtype=np.int32)
        denom = np.zeros(16 * nt)
        for t in range(nt):
            ti = self.type_info(t)
            nlocc[t] = ti["nlocc"]
            nproj[t] = ti["nproj"]
            i = 0
            for l in range(ti["lmax"] + 1):
                if l == ti["lloc"]:
                    continue
                for _ in range(2 * l + 1):
                    denom[16 * t + i] = ti["denom"][l]
                    i += 1
        im = self.images()
        return {"nlocc": nlocc, "nproj": nproj, "denom": denom, "n_atm": self.n_atm,
                "img_type": np.ascontiguousarray(im["type"], dtype=np.int32),
                "img_atom": np.ascontiguousarray(im["atom"], dtype=np.int32),
                "img_nn": np.ascontiguousarray(im["nn"].reshape(-1), dtype=np.int32),
                "udv": np.ascontiguousarray(self.udv())}
