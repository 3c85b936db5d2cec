"""The reference's output and restart files, read and written on the host.

* ``.aout`` — atoms and forces after the SCF (PrintAtoms,
  src/readfiles.cpp:1230-1287);
* ``.restart`` — the electron density of the latest SCF iteration, in the
  block order of the reference's MPI ranks (PrintSCF / RestartSCF,
  src/readfiles.cpp:1324-1440);
* ``.restartMD`` — MD step count, velocities, forces, positions and the
  running energy sums (PrintMD / RestartMD, src/readfiles.cpp:1443-1646);
* ``.vout`` — atom speeds per MD step (PrintVels, src/readfiles.cpp:1290-1322);
* the trajectory file (datacollect, src/md.cpp:22-50).

Vectors of atoms are (n_atm, 3) arrays in type order; the density is the
n^3 grid with x fastest.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class TypeMeta:
    """One atom type as the .atoms file gives it (src/readfiles.cpp:388-520)."""
    symbol: str
    Z: int
    mass: float            # amu
    natm: int
    lloc: int
    rb: float              # bohr
    psd_path: str
    xyz: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))


def _tokens(path: str) -> list[str]:
    """Whitespace tokens of a file with // comments removed, one list per line."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.split("//", 1)[0].strip()
            if line:
                out.append(line)
    return out


def read_atoms_file(path: str) -> tuple[list[TypeMeta], int]:
    """Types and Cartesian positions of a .atoms file; returns (types,
    frac_coord).  Fractional coordinates are returned as written (the caller
    multiplies by the cell side)."""
    lines = _tokens(path)
    n_atm = int(lines[0].split()[0])
    n_typ, frac = None, 0
    i = 1
    while i < len(lines) and lines[i].split()[0] != "Atoms":
        t = lines[i].split()
        if t[0] == "n_typ":
            n_typ = int(t[1])
        elif t[0] == "frac_coord":
            frac = int(t[1])
        i += 1
    if n_typ is None or i == len(lines):
        raise ValueError(f"{path}: no n_typ or Atoms section")
    i += 1
    types = []
    for _ in range(n_typ):
        t = lines[i].split()
        meta = TypeMeta(t[0], int(t[1]), float(t[2]), int(t[3]), int(t[4]), float(t[5]), lines[i + 1].split()[0])
        xyz = np.array([[float(v) for v in lines[i + 2 + a].split()[:3]] for a in range(meta.natm)])
        meta.xyz = xyz.reshape(meta.natm, 3)
        types.append(meta)
        i += 2 + meta.natm
    if sum(m.natm for m in types) != n_atm:
        raise ValueError(f"{path}: {n_atm} atoms declared, {sum(m.natm for m in types)} listed")
    return types, frac


# ------------------------------------------------------------------ .aout

def write_aout(path: str, types: list[TypeMeta], xyz, forces, frac_coord: int = 0, latconst: float = 1.0) -> None:
    """Atoms and forces in PrintAtoms' layout: positions (bohr, or fractions
    of latconst when frac_coord is 1) per type, then the forces per type."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    forces = np.asarray(forces, dtype=np.float64).reshape(-1, 3)
    n_atm = sum(t.natm for t in types)
    if xyz.shape[0] != n_atm or forces.shape[0] != n_atm:
        raise ValueError("positions / forces do not match the atom count")
    lc = latconst if frac_coord == 1 else 1.0
    with open(path, "w") as f:
        f.write(f"{n_atm}\n")
        f.write(f"n_typ {len(types)}\n")
        f.write(f"frac_coord {frac_coord}\n")
        if frac_coord == 1:
            f.write("Note: Atom positions below are in fractional coordinates. \n")
        f.write("\nAtoms\n")
        a = 0
        for t in types:
            f.write(f"{t.symbol} {t.Z} {t.mass:f} {t.natm} {t.lloc} {t.rb:f}\n")
            f.write(f"{t.psd_path}\n")
            for _ in range(t.natm):
                x, y, z = xyz[a] / lc
                f.write(f"{x:.16f} {y:.16f} {z:.16f}\n")
                a += 1
        f.write(" \n")
        f.write("Atomic Forces (Ha/Bohr): \n")
        f.write(" \n")
        a = 0
        for t in types:
            f.write(f"{t.symbol}\n")
            for _ in range(t.natm):
                fx, fy, fz = forces[a]
                f.write(f"{fx:.14f} {fy:.14f} {fz:.14f}\n")
                a += 1


def read_aout(path: str) -> dict:
    """Types, positions (as written) and forces of a .aout file."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    n_atm = int(lines[0].split()[0])
    n_typ = int(lines[1].split()[1])
    frac = int(lines[2].split()[1])
    i = lines.index("Atoms") + 1
    types, xyz = [], []
    for _ in range(n_typ):
        t = lines[i].split()
        meta = TypeMeta(t[0], int(t[1]), float(t[2]), int(t[3]), int(t[4]), float(t[5]), lines[i + 1].strip())
        rows = [[float(v) for v in lines[i + 2 + a].split()] for a in range(meta.natm)]
        meta.xyz = np.array(rows).reshape(meta.natm, 3)
        xyz.extend(rows)
        types.append(meta)
        i += 2 + meta.natm
    j = next(k for k in range(i, len(lines)) if lines[k].startswith("Atomic Forces"))
    j += 2
    forces = []
    for t in types:
        if lines[j].strip() != t.symbol:
            raise ValueError(f"{path}: forces of {t.symbol} expected at line {j + 1}")
        forces.extend([[float(v) for v in lines[j + 1 + a].split()] for a in range(t.natm)])
        j += 1 + t.natm
    if len(xyz) != n_atm:
        raise ValueError(f"{path}: {n_atm} atoms declared, {len(xyz)} listed")
    return {"types": types, "frac_coord": frac, "xyz": np.array(xyz), "forces": np.array(forces)}


# ------------------------------------------------------------------ .restart

def _rank_grid(n: int, nproc: int) -> int:
    """Ranks per direction of the reference's cubic decomposition
    (src/readfiles.cpp:45-63): nproc = p^3 with p dividing n."""
    p = int(round(nproc ** (1.0 / 3.0)))
    if p < 1 or p ** 3 != nproc:
        raise ValueError(f"{nproc} ranks are not a cubic decomposition")
    if n % p:
        raise ValueError(f"{n} nodes per direction do not divide among {p} ranks")
    return p


def density_to_rank_order(rho, n: int, nproc: int = 1) -> np.ndarray:
    """The grid (x fastest) in the order PrintSCF writes it: rank by rank
    (ranks counted x-block outermost, z-block innermost, as
    Processor_domain, src/initialize.cpp:218-274), each rank's block with x
    fastest."""
    p = _rank_grid(n, nproc)
    b = n // p
    g = np.asarray(rho, dtype=np.float64).reshape(p, b, p, b, p, b)   # [bz][z][by][y][bx][x]
    return g.transpose(4, 2, 0, 1, 3, 5).reshape(-1)                   # [bx][by][bz][z][y][x]


def density_from_rank_order(vals, n: int, nproc: int = 1) -> np.ndarray:
    """Inverse of density_to_rank_order."""
    p = _rank_grid(n, nproc)
    b = n // p
    g = np.asarray(vals, dtype=np.float64).reshape(p, p, p, b, b, b)  # [bx][by][bz][z][y][x]
    return g.transpose(2, 3, 1, 4, 0, 5).reshape(-1)


def write_restart(path: str, rho, n: int, nproc: int = 1) -> None:
    """The density in PrintSCF's layout: the rank count, then one value per
    line (%.14f) in rank-block order."""
    vals = density_to_rank_order(rho, n, nproc)
    with open(path, "w") as f:
        f.write(f"{nproc}\n")
        f.write("".join(f"{v:.14f}\n" for v in vals))


def read_restart(path: str, n: int, nproc: int | None = None) -> np.ndarray:
    """The density of a .restart file on the n^3 grid (x fastest).  The
    reference refuses a file written on another rank count
    (src/readfiles.cpp:1407); here any cubic decomposition that divides n is
    reordered, and nproc, when given, must match the file."""
    with open(path) as f:
        np_file = int(f.readline().split()[0])
        vals = np.array(f.read().split(), dtype=np.float64)
    if nproc is not None and nproc != np_file:
        raise ValueError(f"{path} was written by {np_file} ranks, not {nproc}")
    if vals.size != n ** 3:
        raise ValueError(f"{path} holds {vals.size} values, not {n}^3")
    return density_from_rank_order(vals, n, np_file)


# ------------------------------------------------------------------ MD files

@dataclass
class MDState:
    """What .restartMD carries: the MD steps done, velocities, forces,
    positions (bohr) and the running sums of TE, PE, KE, T and their squares
    (the reference's mean_* / std_* fields hold sums, not means)."""
    step_count: int
    vels: np.ndarray
    forces: np.ndarray
    xyz: np.ndarray
    sums: np.ndarray       # [TE, TE^2, PE, PE^2, KE, KE^2, T, T^2]


def write_restart_md(path: str, st: MDState) -> None:
    """PrintMD's layout (%.15f throughout)."""
    def rows(a):
        return "".join(f"{x:.15f} {y:.15f} {z:.15f}\n" for x, y, z in np.asarray(a).reshape(-1, 3))
    s = np.asarray(st.sums, dtype=np.float64)
    with open(path, "w") as f:
        f.write("MDstepCount\n")
        f.write(f"{st.step_count}\n")
        f.write("MD velocities\n")
        f.write(rows(st.vels))
        f.write("MD forces\n")
        f.write(rows(st.forces))
        f.write("Atom positions\n")
        f.write(rows(st.xyz))
        f.write("MD stats\n")
        for i in range(4):
            f.write(f"{s[2 * i]:.15f} {s[2 * i + 1]:.15f}\n")


def read_restart_md(path: str, n_atm: int) -> MDState:
    """RestartMD: each section found by its header line."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]

    def after(h):
        try:
            return lines.index(h) + 1
        except ValueError:
            raise ValueError(f"{path}: no '{h}' section") from None

    def block(h):
        i = after(h)
        return np.array([[float(v) for v in lines[i + a].split()[:3]] for a in range(n_atm)])

    count = int(lines[after("MDstepCount")].split()[0])
    i = after("MD stats")
    sums = np.array([float(v) for k in range(4) for v in lines[i + k].split()[:2]])
    return MDState(count, block("MD velocities"), block("MD forces"), block("Atom positions"), sums)


def write_vout(path: str, step: int, vels) -> None:
    """PrintVels: the step number, then each atom's speed (%.6f); the file is
    started at step 1 and appended to after."""
    v = np.asarray(vels, dtype=np.float64).reshape(-1, 3)
    with open(path, "w" if step == 1 else "a") as f:
        f.write(f"{step}\n")
        f.write("".join(f"{s:.6f}\n" for s in np.sqrt(np.sum(v * v, axis=1))))


def _ostream(x: float) -> str:
    """A double as std::ostream prints it by default (%g, 6 significant digits)."""
    return f"{x:g}"


def append_trajectory(path: str, step: int, max_steps: int, domain, xyz) -> None:
    """datacollect: a header at step 1, then the positions of every step and
    a blank line."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    with open(path, "a") as f:
        if step == 1:
            f.write("Trajectory Data\n")
            f.write("---------------.\n\n")
            f.write(f"{max_steps}\n")
            f.write(" ".join(_ostream(d) for d in domain) + "\n")
            f.write(f"{xyz.shape[0]}\n\n")
        for x, y, z in xyz:
            f.write(f" {_ostream(x)}  {_ostream(y)}  {_ostream(z)}\n")
        f.write("\n")


def exists(path: str | None) -> bool:
    return bool(path) and os.path.exists(path)
