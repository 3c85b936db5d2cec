"""NVE quantum molecular dynamics around the SCF (NVE_MD, src/md.cpp:53-82).

Each step follows the reference: leapfrog part 1 (v += dt/2 a; r += dt v,
src/md.cpp:83-126), wrap-around into the periodic cell (WrapAround_PBC,
:498-540), charge extrapolation for the starting density
(ChargeExtrapolation_MD, :171-317), the SCF and forces at the new positions
(SQDFT_forces, src/scf.cpp:337-395, through ``SCF``), leapfrog part 2
(a = F/m; v += dt/2 a, :128-169) and the energies (MD_energies, :319-364).
Velocities start as in Initialize_MD (:366-496): Maxwell-Boltzmann
(Box-Muller on the C library's rand(), reseeded per atom type with
type + rand_seed) or uniformly random directions, net momentum removed and
rescaled to the requested temperature.  The reference takes that temperature
from the electronic one (src/md.cpp:385-386); ``T_ion`` overrides it.

Units: bohr, fs, Ha; masses in amu are converted with 1 amu =
1.066572305 Ha fs^2 / bohr^2 (src/md.cpp:389).

Files (``sqdft_amd.files``): ``restart_md`` continues from a .restartMD
file (RestartMD, src/main.cpp:91-100: step count, velocities, forces ->
accelerations, positions, energy sums); ``prnt_md`` writes one after every
step (PrintMD, src/main.cpp:139-140); ``trajectory`` appends the positions
of every step (datacollect, src/md.cpp:76-77); ``vout`` the atom speeds
(PrintVels).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import files
from .scf import SCF, pinv_times_vec

AMU = 1.066572305                     # Ha fs^2 / bohr^2 per amu
KB = 8.617343e-5 / 27.211384523       # Ha / K (src/readfiles.cpp:309-311)
RAND_MAX = 2147483647                 # the C library's rand() range

_libc = None


def _crand():
    global _libc
    if _libc is None:
        _libc = C.CDLL(None)
        _libc.rand.restype = C.c_int
        _libc.srand.argtypes = [C.c_uint]
    return _libc


def initial_velocities(masses: np.ndarray, types: np.ndarray, T: float, rand_seed: int, vds: int = 1) -> np.ndarray:
    """Initialize_MD's velocities (src/md.cpp:400-496): (n_atm, 3) in bohr/fs
    for atoms in type order; masses in Ha fs^2/bohr^2 per atom."""
    na = masses.size
    dof = 3 * (na - 1)
    v = np.zeros((na, 3))
    lib = _crand()
    U = lambda: lib.rand() / RAND_MAX  # noqa: E731
    vmag = np.sqrt(dof * T * KB / np.sum(masses))
    seeded = -1
    for k in range(na):
        if types[k] != seeded:            # srand(type + rand_seed) at each type's first atom
            seeded = int(types[k])
            lib.srand(seeded + rand_seed)
        if vds == 2:
            s = 2.0
            while s > 1.0:
                x = 2.0 * U() - 1
                y = 2.0 * U() - 1
                s = x * x + y * y
            v[k, 2] = vmag * (1.0 - 2.0 * s)
            s = 2.0 * np.sqrt(1.0 - s)
            v[k, 0] = s * x * vmag
            v[k, 1] = s * y * vmag
        else:
            sd = np.sqrt(KB * T / masses[k])
            for d in range(3):
                c = sd * np.cos(2 * np.pi * U())
                v[k, d] = c * np.sqrt(-2.0 * np.log(U()))
    v -= v.sum(axis=0) / na               # no net momentum (equal-weight mean, as the reference)
    ke = 0.5 * np.sum(masses[:, None] * v * v)
    return v * np.sqrt(dof * KB * T / (2.0 * ke))


def wrap_around(xyz: np.ndarray, L: float) -> np.ndarray:
    """WrapAround_PBC (src/md.cpp:498-540): one shift by L per coordinate."""
    x = xyz.copy()
    x[x >= L] -= L
    x[x < 0] += L
    return x


def extrapolation_coefficients(r0: np.ndarray, r1: np.ndarray, r2: np.ndarray, rnew: np.ndarray) -> tuple:
    """alpha, beta of ChargeExtrapolation_MD (src/md.cpp:186-276): least
    squares of rnew - r0 on (r0 - r1, r1 - r2) over all coordinates."""
    d01, d12, dn = (r0 - r1).ravel(), (r1 - r2).ravel(), (rnew - r0).ravel()
    FtF = np.array([[d01 @ d01, d01 @ d12], [d01 @ d12, d12 @ d12]])
    Ftf = np.array([d01 @ dn, d12 @ dn])
    a, b = pinv_times_vec(FtF, Ftf)
    return float(a), float(b)


class NVE:
    """NVE MD of the system an ``SCF`` object holds (positions are moved in
    place).  masses: amu per atom type."""

    def __init__(self, scf: SCF, masses, time_step: float = 1.0, T_ion: float | None = None,
                 rand_seed: int = 2, vds: int = 1, chg_extrap: bool = True, restart_md: str | None = None,
                 prnt_md: str | None = None, trajectory: str | None = None, vout: str | None = None,
                 max_steps: int = 0, **scf_kw):
        self.scf = scf
        S = scf.S
        xyz, typ = S.atoms()
        self.types = typ
        self.na = S.n_atm
        self.L = S.info["L"]
        self.dt = float(time_step)
        self.m = np.array([masses[t] for t in typ], dtype=np.float64) * AMU
        self.dof = 3 * (self.na - 1)
        T_el = 1.0 / (KB * S.info["bet"])
        self.T0 = T_el if T_ion is None else float(T_ion)
        self.chg_extrap = chg_extrap
        self.scf_kw = scf_kw
        self.x = xyz.copy()                       # mapped into the cell
        self.x_nm = xyz.copy()                    # not mapped (for the extrapolation)
        self.v = initial_velocities(self.m, typ, self.T0, rand_seed, vds)
        self.a = np.zeros_like(self.x)            # the reference's forces are zero here
        self.step_count = 0
        n3 = S.n ** 3
        self.rho_at_sum = np.zeros(n3)            # rho_at of every geometry so far: the reference never clears it
        self.drho_dt, self.drho_2dt, self.drho_new = np.zeros(n3), np.zeros(n3), np.zeros(n3)
        self.r0, self.r1, self.r2 = np.zeros_like(self.x), np.zeros_like(self.x), np.zeros_like(self.x)
        self.history = []
        self.prnt_md, self.trajectory, self.vout, self.max_steps = prnt_md, trajectory, vout, max_steps
        self.restart_count = 0
        self.sums = np.zeros(8)                   # TE, TE^2, PE, PE^2, KE, KE^2, T, T^2 over all steps
        if restart_md:
            st = files.read_restart_md(restart_md, self.na)
            if max_steps and st.step_count >= max_steps:
                raise ValueError(f"{restart_md}: {st.step_count} of {max_steps} MD steps already done")
            self.restart_count = st.step_count
            self.v = st.vels.copy()
            self.a = st.forces / self.m[:, None]
            self.x = st.xyz.copy()
            self.x_nm = st.xyz.copy()
            self.sums = st.sums.copy()

    def step(self) -> dict:
        t0 = time.perf_counter()
        self.step_count += 1
        k = self.step_count
        dt = self.dt
        # leapfrog part 1
        self.v = self.v + dt * 0.5 * self.a
        self.x = self.x + dt * self.v
        self.x_nm = self.x_nm + dt * self.v
        self.x = wrap_around(self.x, self.L)
        # charge extrapolation, with the density and rho_at of the last SCF
        scf = self.scf
        if k > 1:
            drho = scf.rho - self.rho_at_sum
            if k > 3:
                al, be = extrapolation_coefficients(self.r0, self.r1, self.r2, self.x_nm)
                self.drho_new = (1 + al) * drho + (be - al) * self.drho_dt - be * self.drho_2dt
            self.drho_2dt, self.drho_dt = self.drho_dt, drho
        self.r2, self.r1, self.r0 = self.r1, self.r0, self.x_nm.copy()
        # SCF and forces at the new positions
        last_rho = scf.rho
        scf.set_positions(self.x)
        self.rho_at_sum = self.rho_at_sum + scf.rho_at
        if k == 1:
            guess = self.rho_at_sum
        elif k > 3 and self.chg_extrap:
            guess = self.rho_at_sum + self.drho_new
            guess = np.where(guess < 0.0, 1e-6, guess)
        else:
            guess = last_rho
        scf.set_density_guess(guess, keep_phi=k > 1)
        scf.initial_potential()
        out = scf.run(**self.scf_kw)
        f = scf.ForcesTotal()
        f = f - f.mean(axis=0)                    # net force removed (src/forces.cpp:713-741)
        # leapfrog part 2
        self.a = f / self.m[:, None]
        self.v = self.v + dt * 0.5 * self.a
        # energies (MD_energies)
        pe = out["free_energy"]
        ke = 0.5 * np.sum(self.m[:, None] * self.v * self.v) / self.na
        rec = {"step": k, "TE": pe + ke, "PE": pe, "KE": ke, "T": 2 * ke * self.na / (KB * self.dof),
               "scf_iterations": out["iterations"], "scf_converged": out["converged"],
               "force_sum": f.sum(axis=0), "velocity_sum": self.v.sum(axis=0), "forces": f,
               "positions": self.x.copy(),
               "seconds": time.perf_counter() - t0}
        self.history.append(rec)
        for i, q in enumerate((rec["TE"], rec["PE"], rec["KE"], rec["T"])):
            self.sums[2 * i] += q
            self.sums[2 * i + 1] += q * q
        total = k + self.restart_count
        rec["md_step"] = total
        if self.trajectory:
            files.append_trajectory(self.trajectory, k, self.max_steps, [self.L] * 3, self.x)
        if self.vout:
            files.write_vout(self.vout, k, self.v)
        if self.prnt_md:
            files.write_restart_md(self.prnt_md, files.MDState(total, self.v, f, self.x, self.sums))
        return rec

    def means(self) -> dict:
        """Mean and standard deviation of TE, PE, KE and T over every step so
        far, restarted runs included (src/main.cpp:144-148)."""
        n = max(self.step_count + self.restart_count, 1)
        out = {}
        for i, name in enumerate(("TE", "PE", "KE", "T")):
            m = self.sums[2 * i] / n
            out[name] = (m, float(np.sqrt(abs(self.sums[2 * i + 1] / n - m * m))))
        return out

    def run(self, steps: int) -> list:
        return [self.step() for _ in range(steps)]
