"""NVE molecular dynamics on the GPU path (NVE_MD, src/md.cpp:53-82) on the
reference's Al test cell with its MD defaults (1 fs, Maxwell-Boltzmann
velocities at the electronic temperature, charge extrapolation)."""
import sys

import numpy as np
import pytest

import helpers as H
from sqdft_amd.md import NVE
from sqdft_amd.scf import SCF

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_nve_al_cell():
    S = H.test_al()[0]
    say("md: 4 NVE steps on the Al test cell")
    scf = SCF(S)
    dyn = NVE(scf, masses=[26.9815385], time_step=1.0)
    recs = dyn.run(4)
    for r in recs:
        say("step", r["step"], "TE", r["TE"], "PE", r["PE"], "KE", r["KE"], "T", r["T"],
            "scf", r["scf_iterations"], f"{r['seconds']:.2f}s")
        assert r["scf_converged"]
        # net force removed and equal masses: the momentum stays zero
        assert np.max(np.abs(r["velocity_sum"])) < 1e-12
    # On this 10^3 grid the reference's total forces differ from -dF/dR of its
    # own free energy by ~5 % (central differences of the converged SCF), so
    # TE drifts by that fraction of the potential-energy change it tracks
    te = np.array([r["TE"] for r in recs])
    dpe = np.sum(np.abs(np.diff([r["PE"] for r in recs])))
    say("TE drift", float(np.max(np.abs(te - te[0]))), "sum |dPE|", float(dpe))
    assert np.max(np.abs(te - te[0])) < 0.15 * dpe
    # leapfrog: positions advance by dt times the half-step velocity
    L = dyn.L
    for a, b in zip(recs[:-1], recs[1:]):
        dx = b["positions"] - a["positions"]
        dx -= np.round(dx / L) * L
        assert np.all(np.abs(dx) < 1.0)
    # step 1 is a plain SCF + forces at the moved positions
    S2 = H.test_al()[0]
    S2.set_positions(recs[0]["positions"])
    fresh = SCF(S2)
    out = fresh.run()
    f = fresh.ForcesTotal()
    f = f - f.mean(axis=0)
    say("step-1 forces vs a fresh SCF", float(np.max(np.abs(f - recs[0]["forces"]))))
    assert abs(out["free_energy"] - recs[0]["PE"]) < 1e-10
    assert np.max(np.abs(f - recs[0]["forces"])) < 1e-9
    fresh.close()
    scf.close()
