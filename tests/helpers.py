"""Builds matched (product System, oracle system) pairs for the tests."""
from __future__ import annotations

import os

import numpy as np

from sqdft_amd import configs
from sqdft_amd.system import AtomType, System

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PSD = os.path.join(GOLDEN, "pseudopotentials")


def types_of(system: System):
    """AtomType list reconstructed from a built system (same coordinates)."""
    xyz, typ = system.atoms()
    out = []
    syms = {3: ("Al", "psd_Al.pot")}
    for t in range(system.info["n_typ"]):
        ti = system.type_info(t)
        sel = xyz[typ == t]
        if ti["Z"] == 3:
            sym, f = "Al", "psd_Al.pot"
        elif ti["lmax"] >= 1:
            sym, f = "Li", "psd_Li.pot"
        else:
            sym, f = "H", "psd_H.pot"
        out.append(AtomType(sym, ti["Z"], ti["lloc"], os.path.join(PSD, f), sel, rb=ti["rb"]))
    del syms
    return out


def oracle_for(system: System, Rcut: float, T: float):
    from oracle import OracleSystem

    info = system.info
    return OracleSystem(info["n"], info["L"], types_of(system), Rcut, info["npl"], T, h=info["h"],
                        lanczos_tol=info["lanczos_tol"], fermi_tol=info["fermi_tol"],
                        fermi_alg=info["fermi_alg"])


def test_al():
    """The reference's tests/Al input (4 atoms, n=10, nloc=8, npl=28)."""
    S = System.from_files(os.path.join(GOLDEN, "Al"), "test_Al")
    return S, 6.0, 46418.02


def test_lih():
    """The reference's tests/LiH input (8 atoms, n=15, nloc=13, npl=40)."""
    S = System.from_files(os.path.join(GOLDEN, "LiH"), "test_LiH")
    return S, 6.0, 46418.02


def small_al(**kw):
    Rcut = kw.pop("Rcut", 2.0)
    T = kw.pop("T", 58000.0)
    return configs.small_al(Rcut=Rcut, T=T, **kw), Rcut, T


def small_lih(**kw):
    Rcut = kw.pop("Rcut", 2.0)
    T = kw.pop("T", 40000.0)
    return configs.small_lih(Rcut=Rcut, T=T, **kw), Rcut, T


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
