"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle on the
same inputs.  FP64 path: 1e-10 relative or better; mixed path: stated
tolerances below."""
import os
import sys

import numpy as np
import pytest

import helpers as H
from sqdft_amd import configs
from sqdft_amd.configs import smooth_veff
from sqdft_amd.sq import SpectralQuadrature

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SYSTEMS = {
    "small_al": H.small_al,
    "small_lih": H.small_lih,
    "test_al": H.test_al,
    "test_lih": H.test_lih,
}


@pytest.fixture(scope="module")
def cases():
    out = {}
    for k, f in SYSTEMS.items():
        S, Rcut, T = f()
        out[k] = (S, H.oracle_for(S, Rcut, T), smooth_veff(S))
    return out


@pytest.mark.parametrize("name", ["small_al", "small_lih", "test_al", "test_lih"])
@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed_lanczos"])
def test_hsub_matches_oracle(cases, name, prec):
    S, osys, veff = cases[name]
    say("hsub", name, prec)
    sq = SpectralQuadrature(S, precision=prec)
    rng = np.random.Generator(np.random.PCG64(11))
    nodes = np.array([0, 1, S.n + 3, S.nn // 2, S.nn - 1], dtype=np.int32)
    v = rng.normal(size=(nodes.size, sq.M ** 3))
    hv = sq.HsubTimesVec(veff, nodes, v).reshape(nodes.size, -1)
    for i, nd in enumerate(nodes):
        ref = osys.hsub(veff, int(nd), v[i])
        # mixed_lanczos applies H in FP64 (its FP64 copies of the operator data)
        assert H.rel(hv[i], ref) < (2e-6 if prec == "mixed" else 1e-13), (name, nd)


@pytest.mark.parametrize("name", ["small_al", "small_lih", "test_al", "test_lih"])
def test_node_moments_match_oracle(cases, name):
    S, osys, veff = cases[name]
    say("moments", name)
    sq = SpectralQuadrature(S, precision="fp64")
    nodes = np.array([0, 5, S.nn // 3, S.nn - 2], dtype=np.int32)
    g = sq.node_moments(veff, nodes)
    for i, nd in enumerate(nodes):
        r = osys.node_moments(veff, int(nd))
        assert g["iters"][i] == r["iters"], (name, nd)
        assert g["chi"][i] == pytest.approx(r["chi"], rel=FP64_TOL)
        assert g["zee"][i] == pytest.approx(r["zee"], rel=FP64_TOL)
        assert H.rel(g["mu"][i], r["mu"]) < FP64_TOL


@pytest.mark.parametrize("name", ["small_al", "small_lih", "test_al"])
def test_density_matches_oracle(cases, name):
    S, osys, veff = cases[name]
    say("density", name)
    ref = osys.density(veff, scf_iter=1, nthreads=0)
    assert ref["status"] == 0
    sq = SpectralQuadrature(S, precision="fp64")
    out = sq.density(veff, scf_iter=1)
    assert out["lambda_f"] == pytest.approx(ref["lambda_f"], abs=FP64_TOL)
    assert H.rel(out["rho"], ref["rho"]) < FP64_TOL
    assert out["ebs"] == pytest.approx(ref["ebs"], rel=FP64_TOL)
    assert out["eent"] == pytest.approx(ref["eent"], rel=FP64_TOL)
    nd = sq.node_data()
    assert np.array_equal(nd["iters"], ref["iters"])
    assert H.rel(nd["chi"], ref["chi"]) < FP64_TOL
    assert H.rel(nd["mu"], ref["mu"]) < FP64_TOL
    assert H.rel(nd["Ci"], ref["Ci"]) < FP64_TOL
    assert out["stats"]["fermi_iters"] == ref["fermi_iters"]
    # second SCF iteration: bracket lambda_f +- 2 (sq.cpp:1301-1304)
    ref2 = osys.density(veff + 0.01, scf_iter=2, lambda_f=ref["lambda_f"], nthreads=0)
    out2 = sq.density(veff + 0.01, scf_iter=2, lambda_f=out["lambda_f"])
    assert out2["lambda_f"] == pytest.approx(ref2["lambda_f"], abs=FP64_TOL)
    assert H.rel(out2["rho"], ref2["rho"]) < FP64_TOL


@pytest.mark.parametrize("name", ["small_al", "test_lih"])
def test_newton_fermi_matches_oracle(cases, name):
    """fermi_alg = 2: NewtonRaphson on NeConstraint_derv (src/sq.cpp:1407-1490),
    started at 0.2 in the first SCF iteration and at the previous Fermi level
    afterwards (src/sq.cpp:722-725)."""
    from oracle import OracleSystem
    S, osys1, veff = cases[name]
    say("newton", name)
    Rcut, T = SYSTEMS[name]()[1:]
    info = S.info
    osys = OracleSystem(info["n"], info["L"], H.types_of(S), Rcut, info["npl"], T, h=info["h"],
                        lanczos_tol=info["lanczos_tol"], fermi_tol=info["fermi_tol"], fermi_alg=2)
    ref = osys.density(veff, scf_iter=1, nthreads=0)
    sq = SpectralQuadrature(S, precision="fp64", fermi_alg=2)
    out = sq.density(veff, scf_iter=1)
    say("newton", name, "its", out["stats"]["fermi_iters"], ref["fermi_iters"], "lambda_f", out["lambda_f"],
        ref["lambda_f"])
    assert out["stats"]["fermi_iters"] == ref["fermi_iters"]
    assert out["lambda_f"] == pytest.approx(ref["lambda_f"], abs=FP64_TOL)
    assert H.rel(out["rho"], ref["rho"]) < FP64_TOL
    # Newton and Brent find the same root to the Fermi tolerance
    b = osys1.density(veff, scf_iter=1, nthreads=0)
    assert abs(out["lambda_f"] - b["lambda_f"]) < 10 * info["fermi_tol"]
    ref2 = osys.density(veff + 0.01, scf_iter=2, lambda_f=ref["lambda_f"], nthreads=0)
    out2 = sq.density(veff + 0.01, scf_iter=2, lambda_f=out["lambda_f"])
    assert out2["stats"]["fermi_iters"] == ref2["fermi_iters"]
    assert out2["lambda_f"] == pytest.approx(ref2["lambda_f"], abs=FP64_TOL)
    assert H.rel(out2["rho"], ref2["rho"]) < FP64_TOL
    sq.close()


def test_exit_paths_return_status(cases):
    """The reference's exit paths as status codes, each forced from the host
    with valid memory and the same condition the oracle reports in
    test_status_codes.py: Brent bracket (sq.cpp:1307-1311), Brent ITMAX
    (:1400-1401), negative density (:764-765) and the Lanczos max_iter
    warning (:1073-1074)."""
    from sqdft_amd._lib import SQError
    from test_status_codes import neg_rho_case
    S, _, veff = cases["small_al"]
    say("status: bracket")
    sq = SpectralQuadrature(S, precision="fp64")
    good = sq.density(veff, scf_iter=1)
    with pytest.raises(SQError) as e:
        sq.density(veff, scf_iter=2, lambda_f=good["lambda_f"] + 10.0)
    assert e.value.code == -7
    # the context stays usable after an error return
    again = sq.density(veff, scf_iter=2, lambda_f=good["lambda_f"])
    assert again["lambda_f"] == pytest.approx(good["lambda_f"], abs=1e-6)
    sq.close()
    say("status: brent itmax")
    sq = SpectralQuadrature(S, precision="fp64", fermi_tol=-1.0)
    with pytest.raises(SQError) as e:
        sq.density(veff, scf_iter=1)
    assert e.value.code == -8
    sq.close()
    say("status: negative density")
    Sn, _, _, vn = neg_rho_case()
    sq = SpectralQuadrature(Sn, precision="fp64")
    with pytest.raises(SQError) as e:
        sq.density(vn, scf_iter=1)
    assert e.value.code == -5
    sq.close()
    say("status: lanczos max_iter")
    Sl, _, _ = H.small_al(npl=8)
    vl = smooth_veff(Sl)
    sq = SpectralQuadrature(Sl, precision="fp64", lanczos_tol=-1.0)
    out = sq.density(vl, scf_iter=1)
    assert out["status"] == 1
    assert np.all(sq.node_data()["iters"] == 500)
    sq.close()


@pytest.mark.parametrize("name", ["small_al", "small_lih", "test_al", "test_lih"])
def test_nonlocal_forces_match_oracle(cases, name):
    S, osys, veff = cases[name]
    say("forces", name)
    ref = osys.density(veff, scf_iter=1, nthreads=0)
    fref = osys.forces_nonlocal(veff, ref["chi"], ref["zee"], ref["Ci"], nthreads=0)
    sq = SpectralQuadrature(S, precision="fp64")
    sq.density(veff, scf_iter=1)
    f = sq.ForcesNonLocal()
    assert np.max(np.abs(f - fref)) <= 1e-10 * np.max(np.abs(fref)) + 1e-14
    # mixed path: FP32 recurrence and accumulator
    sqm = SpectralQuadrature(S, precision="mixed")
    sqm.density(veff, scf_iter=1)
    fm = sqm.ForcesNonLocal()
    say("mixed force err", np.max(np.abs(fm - fref)))
    assert np.max(np.abs(fm - fref)) < 1e-4
    # FP32 Lanczos only: the force pass is FP64, with the FP32 passes' bounds
    sql = SpectralQuadrature(S, precision="mixed_lanczos")
    sql.density(veff, scf_iter=1)
    fl = sql.ForcesNonLocal()
    say("mixed_lanczos force err", np.max(np.abs(fl - fref)))
    assert np.max(np.abs(fl - fref)) < 1e-6


def test_density_batching_is_invisible(cases):
    S, osys, veff = cases["small_al"]
    say("batching")
    for prec in ("fp64", "mixed_lanczos"):
        a = SpectralQuadrature(S, precision=prec).density(veff)
        b = SpectralQuadrature(S, precision=prec, batch_nodes=37).density(veff)
        assert np.array_equal(a["rho"], b["rho"]), prec
        assert a["lambda_f"] == b["lambda_f"], prec


def test_mixed_precision_density_error(cases):
    # FP32 vectors/stencil, FP64 moments and Fermi solve.  Both sides solve the
    # Fermi level to 1e-12: at the reference's fermi_tol of 1e-6 two runs may
    # sit ~1e-6 apart in lambda_f, which alone moves rho by ~3e-6 here
    S, osys, veff = cases["test_al"]
    say("mixed density")
    tol0 = osys.S.fermi_tol
    osys.S.fermi_tol = 1e-12
    try:
        ref = osys.density(veff, scf_iter=1, nthreads=0)
    finally:
        osys.S.fermi_tol = tol0
    out = SpectralQuadrature(S, precision="mixed", fermi_tol=1e-12).density(veff)
    err = H.rel(out["rho"], ref["rho"])
    say("mixed rho rel err", err, "dlf", out["lambda_f"] - ref["lambda_f"])
    assert err < 2e-6
    assert out["lambda_f"] == pytest.approx(ref["lambda_f"], abs=1e-6)
    # FP32 Lanczos passes, FP64 Chebyshev moments: the FP32 storage of the
    # Chebyshev vectors is what costs the mixed mode its 1e-6
    lo = SpectralQuadrature(S, precision="mixed_lanczos", fermi_tol=1e-12).density(veff)
    errl = H.rel(lo["rho"], ref["rho"])
    say("mixed_lanczos rho rel err", errl, "dlf", lo["lambda_f"] - ref["lambda_f"])
    assert errl < 1e-6
    assert lo["lambda_f"] == pytest.approx(ref["lambda_f"], abs=1e-8)


def test_electron_count_at_al32_geometry():
    # size-independent property on the 32-atom Al cell (n=32, nloc=17):
    # sum rho h^3 equals the electron count to the Fermi tolerance
    S = configs.build("al32", seed=0)
    say("al32 moments subset")
    veff = smooth_veff(S)
    sq = SpectralQuadrature(S, precision="fp64")
    osys = H.oracle_for(S, 8.0, 58000.0)
    nodes = np.array([0, 4321, 20000], dtype=np.int32)
    g = sq.node_moments(veff, nodes)
    for i, nd in enumerate(nodes):
        r = osys.node_moments(veff, int(nd))
        assert g["iters"][i] == r["iters"]
        assert g["chi"][i] == pytest.approx(r["chi"], rel=FP64_TOL)
        assert H.rel(g["mu"][i], r["mu"]) < FP64_TOL


@pytest.mark.parametrize("name", ["al32", "lih216", "al256"])
def test_full_size_density_matches_oracle_fixture(name):
    """One SQ density step over every node of the 32-atom Al cell (32^3
    nodes, 35^3 cubes, npl 48) and the 216-atom LiH cell (48^3 nodes, 31^3
    cubes, npl 64) against the FP64 oracle's run on the same inputs
    (tests/golden/sq_<name>.npz, written by make_sq_fixtures.py): per-node
    rho, Lanczos counts, Fermi level, Ebs, Eent; mixed precision within its
    stated 5e-6, FP32 Lanczos with FP64 moments within 1e-6 (Fermi
    tolerance 1e-6 on both sides)."""
    path = os.path.join(H.GOLDEN, f"sq_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated (tests/golden/make_sq_fixtures.py {name})")
    fx = np.load(path)
    k = int(fx["stride"]) if "stride" in fx else 1   # every k-th node stored
    S = configs.build(name, seed=0)
    veff = smooth_veff(S)
    say(f"{name}: full density step vs oracle fixture")
    sq = SpectralQuadrature(S, precision="fp64")
    out = sq.density(veff, scf_iter=1)
    it = sq.node_data()["iters"]
    sq.close()
    err = H.rel(out["rho"][::k], fx["rho"])
    say(name, "fp64 rho rel err", err, "dlf", out["lambda_f"] - float(fx["lambda_f"]))
    assert err < FP64_TOL
    assert np.array_equal(it[::k], fx["iters"].astype(it.dtype))
    assert out["stats"]["fermi_iters"] == int(fx["fermi_iters"])
    assert out["lambda_f"] == pytest.approx(float(fx["lambda_f"]), abs=1e-10)
    assert out["ebs"] == pytest.approx(float(fx["ebs"]), rel=FP64_TOL)
    assert out["eent"] == pytest.approx(float(fx["eent"]), rel=FP64_TOL)
    sqm = SpectralQuadrature(S, precision="mixed")
    mo = sqm.density(veff, scf_iter=1)
    sqm.close()
    errm = H.rel(mo["rho"][::k], fx["rho"])
    say(name, "mixed rho rel err", errm)
    assert errm < 5e-6
    assert mo["lambda_f"] == pytest.approx(float(fx["lambda_f"]), abs=2e-6)
    sql = SpectralQuadrature(S, precision="mixed_lanczos")
    lo = sql.density(veff, scf_iter=1)
    sql.close()
    errl = H.rel(lo["rho"][::k], fx["rho"])
    say(name, "mixed_lanczos rho rel err", errl, "dlf", lo["lambda_f"] - float(fx["lambda_f"]))
    assert errl < 1e-6
    assert lo["lambda_f"] == pytest.approx(float(fx["lambda_f"]), abs=1e-6)
