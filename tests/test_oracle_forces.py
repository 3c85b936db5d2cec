"""Pins the oracle's nonlocal-force formula (forces.cpp:318-578) physically:
at fixed Veff only the nonlocal projectors move with an atom, so the force
must equal minus the derivative of the Mermin band free energy
F = Ebs + Eent at fixed electron count.  Central differences, delta = 1e-3
bohr, on the reference's Al test cell (Rcut = 6.2 bohr); agreement is limited
by the spectral-quadrature truncation, about 1e-4 Ha/bohr.  CPU only."""
import numpy as np
import pytest

import helpers as H
from oracle import OracleSystem
from sqdft_amd.configs import smooth_veff
from sqdft_amd.system import AtomType


def test_nonlocal_force_is_minus_free_energy_gradient():
    S, Rcut, T = H.test_al()
    info = S.info
    veff = smooth_veff(S)
    types = H.types_of(S)

    def free_energy(ts):
        o = OracleSystem(info["n"], info["L"], ts, Rcut, info["npl"], T, h=info["h"])
        d = o.density(veff, 1, nthreads=0)
        assert d["status"] == 0
        return d["ebs"] + d["eent"], o, d

    _, o, d = free_energy(types)
    f = o.forces_nonlocal(veff, d["chi"], d["zee"], d["Ci"], nthreads=0)
    atom, delta = 0, 1e-3
    fd = np.zeros(3)
    for ax in range(3):
        vals = []
        for sgn in (1, -1):
            ts = [AtomType(t.symbol, t.Z, t.lloc, t.psd_path, np.array(t.xyz, copy=True)) for t in types]
            ts[0].xyz[atom, ax] += sgn * delta
            vals.append(free_energy(ts)[0])
        fd[ax] = -(vals[0] - vals[1]) / (2 * delta)
    assert np.max(np.abs(fd - f[atom])) < 4e-4
    assert np.linalg.norm(fd - f[atom]) < 0.1 * np.linalg.norm(fd)
