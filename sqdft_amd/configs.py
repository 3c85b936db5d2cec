"""Synthetic atomic configurations of the SQ workloads.

Coordinates are generated here from a seed (numpy PCG64) and handed to the
system builder explicitly, so no run depends on the C library's rand().

=========  ============================================================  =====
name       system                                                        nodes
=========  ============================================================  =====
al32       FCC Al 2x2x2 cells (32 atoms), a=7.78, n=32, Rcut=8, npl=48   32^3
lih216     rocksalt LiH 3x3x3 (216 atoms), a=7.6, n=48, Rcut=7, npl=64   48^3
al256      FCC Al 4x4x4 (256 atoms), n=64, Rcut=8, npl=48                64^3
al32_adj   al32 at npl=64                                                32^3
al864      FCC Al 6x6x6 (864 atoms), n=96, Rcut=8, npl=48                96^3
=========  ============================================================  =====

Electronic temperatures: 58,000 K (Al), 40,000 K (LiH); FD order 12;
lanczos_tol = fermi_tol = 1e-6; Brent root finder.  Displacements are
uniform in (-d, d) per coordinate with d = 0.1 bohr (Al) and 0.05 bohr (LiH).
"""
from __future__ import annotations

import os

import numpy as np

from .system import AtomType, System

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PSD_DIR = os.path.join(ROOT, "tests", "golden", "pseudopotentials")

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])


def _tile(basis: np.ndarray, ncell: int, a: float) -> np.ndarray:
    out = []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                out.append((basis + np.array([i, j, k])) * a)
    return np.concatenate(out, axis=0)


def fcc_al(ncell: int, a: float = 7.78, disp: float = 0.1, seed: int = 0) -> np.ndarray:
    pos = _tile(FCC_BASIS, ncell, a)
    rng = np.random.Generator(np.random.PCG64(seed))
    return pos + rng.uniform(-disp, disp, size=pos.shape)


def rocksalt_lih(ncell: int, a: float = 7.6, disp: float = 0.05, seed: int = 0):
    li = _tile(FCC_BASIS, ncell, a)
    h = _tile(FCC_BASIS + np.array([0.5, 0.0, 0.0]), ncell, a)
    rng = np.random.Generator(np.random.PCG64(seed))
    li = li + rng.uniform(-disp, disp, size=li.shape)
    h = h + rng.uniform(-disp, disp, size=h.shape)
    return li, h


CONFIGS = {
    "al32": dict(kind="al", ncell=2, n_cell=16, npl=48, Rcut=8.0, T=58000.0),
    "lih216": dict(kind="lih", ncell=3, n_cell=16, npl=64, Rcut=7.0, T=40000.0),
    "al256": dict(kind="al", ncell=4, n_cell=16, npl=48, Rcut=8.0, T=58000.0),
    "al32_adj": dict(kind="al", ncell=2, n_cell=16, npl=64, Rcut=8.0, T=58000.0),
    "al864": dict(kind="al", ncell=6, n_cell=16, npl=48, Rcut=8.0, T=58000.0),
}


def build(name: str, seed: int = 0, psd_dir: str = PSD_DIR, **over) -> System:
    """Builds one of the named workloads (keyword overrides: npl, Rcut, T, n_cell, ncell)."""
    c = dict(CONFIGS[name])
    c.update(over)
    ncell, n_cell = c["ncell"], c["n_cell"]
    if c["kind"] == "al":
        a = 7.78
        types = [AtomType("Al", 3, 1, os.path.join(psd_dir, "psd_Al.pot"), fcc_al(ncell, a, 0.1, seed), rb=6.0)]
    else:
        a = 7.6
        li, h = rocksalt_lih(ncell, a, 0.05, seed)
        types = [AtomType("Li", 1, 0, os.path.join(psd_dir, "psd_Li.pot"), li, rb=8.0),
                 AtomType("H", 1, 0, os.path.join(psd_dir, "psd_H.pot"), h, rb=8.0)]
    # the mesh is computed per unit cell, as the reference does (readfiles.cpp:294-303)
    return System.build(n=n_cell * ncell, L_box=a * ncell, types=types, Rcut=c["Rcut"], npl=c["npl"],
                        T=c["T"], h=a / n_cell)


def small_al(ncell: int = 1, n_cell: int = 8, Rcut: float = 2.0, npl: int = 16, T: float = 58000.0,
             seed: int = 0, psd_dir: str = PSD_DIR, disp: float = 0.1) -> System:
    """A reduced Al cell for fast parity tests (small node cubes)."""
    a = 7.78
    types = [AtomType("Al", 3, 1, os.path.join(psd_dir, "psd_Al.pot"), fcc_al(ncell, a, disp, seed))]
    return System.build(n=n_cell * ncell, L_box=a * ncell, types=types, Rcut=Rcut, npl=npl, T=T,
                        h=a / n_cell)


def small_lih(ncell: int = 1, n_cell: int = 8, Rcut: float = 2.0, npl: int = 16, T: float = 40000.0,
              seed: int = 0, psd_dir: str = PSD_DIR) -> System:
    """A reduced LiH cell (3 p-projectors per Li image, none for H)."""
    a = 7.6
    li, h = rocksalt_lih(ncell, a, 0.05, seed)
    types = [AtomType("Li", 1, 0, os.path.join(psd_dir, "psd_Li.pot"), li),
             AtomType("H", 1, 0, os.path.join(psd_dir, "psd_H.pot"), h)]
    return System.build(n=n_cell * ncell, L_box=a * ncell, types=types, Rcut=Rcut, npl=npl, T=T,
                        h=a / n_cell)


def smooth_veff(system: System, depth: float = 0.9, width: float = 1.1, shift: float = 0.15,
                seed: int = 1, ripple: float = 0.02) -> np.ndarray:
    """A smooth periodic effective potential (n^3, x fastest): attractive
    Gaussian wells on every atom plus a weak seeded plane-wave ripple.  Used
    as the input of SQ steps where the SCF potential itself is not the
    object under test."""
    info = system.info
    n, L_box = info["n"], info["L"]
    h = info["h"]
    xyz, _ = system.atoms()
    g = np.arange(n) * h
    v = np.full((n, n, n), shift)  # [z][y][x]
    for p in xyz:
        d = [(g - p[k] + 0.5 * L_box) % L_box - 0.5 * L_box for k in range(3)]
        ex = np.exp(-d[0] ** 2 / (2 * width ** 2))
        ey = np.exp(-d[1] ** 2 / (2 * width ** 2))
        ez = np.exp(-d[2] ** 2 / (2 * width ** 2))
        v -= depth * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(3):
        kv = rng.integers(-2, 3, size=3)
        ph = rng.uniform(0, 2 * np.pi)
        zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
        v += ripple * np.cos(2 * np.pi * (kv[0] * xx + kv[1] * yy + kv[2] * zz) / L_box + ph)
    return np.ascontiguousarray(v.reshape(-1))
