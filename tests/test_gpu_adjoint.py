"""GPU Veff derivatives (sqg_adjoint_veff, sqg_tangent_veff, sqg_density_fixed)
against the FP64 oracle's reverse and forward modes at the same chi, zee,
moments and Fermi level, against each other (<g, dv> = <w, drho>), and
against central differences of the GPU's own frozen-bounds density step.
FP64 path: 1e-10 relative; mixed path: stated tolerance."""
import sys

import numpy as np
import pytest

import helpers as H
from sqdft_amd.configs import smooth_veff
from sqdft_amd.sq import SpectralQuadrature

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SYSTEMS = {"small_al": H.small_al, "small_lih": H.small_lih, "test_al": H.test_al}


@pytest.fixture(scope="module")
def cases():
    out = {}
    for k, f in SYSTEMS.items():
        S, Rcut, T = f()
        out[k] = (S, H.oracle_for(S, Rcut, T), smooth_veff(S))
    return out


def gpu_state(S, veff, prec="fp64", ckpt=8):
    sq = SpectralQuadrature(S, precision=prec, ckpt_interval=ckpt)
    d = sq.density(veff, scf_iter=1)
    nd = sq.node_data()
    return sq, d, nd


@pytest.mark.parametrize("name", ["small_al", "small_lih", "test_al"])
@pytest.mark.parametrize("fixed_n", [0, 1])
def test_adjoint_matches_oracle(cases, name, fixed_n):
    S, osys, veff = cases[name]
    say("adjoint", name, fixed_n)
    sq, d, nd = gpu_state(S, veff)
    w = np.random.Generator(np.random.PCG64(21)).normal(size=S.nn)
    g = sq.adjoint_veff(w, fixed_n=fixed_n)
    ref = osys.adjoint_veff(veff, nd["chi"], nd["zee"], nd["mu"].reshape(-1), d["lambda_f"], w,
                            fixed_n=fixed_n)
    err = H.rel(g, ref)
    say("  rel err", err)
    assert err < 1e-10


@pytest.mark.parametrize("name", ["small_al", "small_lih", "test_al"])
def test_tangent_matches_oracle(cases, name):
    S, osys, veff = cases[name]
    say("tangent", name)
    sq, d, nd = gpu_state(S, veff)
    dv = np.random.Generator(np.random.PCG64(22)).normal(size=S.nn)
    dr = sq.tangent_veff(dv, fixed_n=True)
    ref = osys.tangent_veff(veff, nd["chi"], nd["zee"], nd["mu"].reshape(-1), d["lambda_f"], dv, fixed_n=1)
    err = H.rel(dr, ref)
    say("  rel err", err)
    assert err < 1e-10
    assert abs(dr.sum()) < 1e-10 * np.abs(dr).sum()


def test_checkpointing_is_invisible(cases):
    # ckpt 3 (recomputed segments) against every vector kept: same kernels on
    # the same inputs, so the result is bitwise the same
    S, osys, veff = cases["small_lih"]
    say("checkpoint")
    w = np.random.Generator(np.random.PCG64(23)).normal(size=S.nn)
    a = gpu_state(S, veff, ckpt=3)[0].adjoint_veff(w)
    b = gpu_state(S, veff, ckpt=0)[0].adjoint_veff(w)
    assert np.array_equal(a, b)


def test_adjoint_tangent_duality_and_differences(cases):
    S, osys, veff = cases["small_al"]
    say("duality + central differences")
    sq, d, nd = gpu_state(S, veff)
    rng = np.random.Generator(np.random.PCG64(24))
    w = rng.normal(size=S.nn)
    g = sq.adjoint_veff(w)
    for _ in range(3):
        dv = rng.normal(size=S.nn)
        assert g @ dv == pytest.approx(w @ sq.tangent_veff(dv), rel=1e-10, abs=1e-12 * np.abs(g).sum())
    lf = d["lambda_f"]
    eps = 1e-5
    # Brent's tolerance has to sit far below eps for a 2e-5 difference
    tight = SpectralQuadrature(S, precision="fp64", fermi_tol=1e-14)
    tight.density(veff, scf_iter=1)
    for node in rng.choice(S.nn, size=4, replace=False):
        dv = np.zeros(S.nn)
        dv[node] = 1.0
        rp = tight.density_fixed(veff + eps * dv, lf - 1, lf + 1)["rho"]
        rm = tight.density_fixed(veff - eps * dv, lf - 1, lf + 1)["rho"]
        fd = w @ (rp - rm) / (2 * eps)
        assert g[node] == pytest.approx(fd, rel=2e-5, abs=2e-6 * np.abs(g).max())


@pytest.mark.parametrize("name", ["small_al", "small_lih"])
def test_mixed_precision_adjoint(cases, name):
    # FP32 vectors and recurrence, FP64 weights and accumulation
    S, osys, veff = cases[name]
    say("mixed adjoint", name)
    w = np.random.Generator(np.random.PCG64(25)).normal(size=S.nn)
    ref = gpu_state(S, veff)[0].adjoint_veff(w)
    g = gpu_state(S, veff, prec="mixed")[0].adjoint_veff(w)
    err = H.rel(g, ref)
    say("  mixed rel err", err)
    assert err < 1e-4


def test_gradcheck_autograd_function(cases):
    # torch.autograd.gradcheck of the wrapped density (fast mode: one random
    # direction, central differences against the adjoint's vector-Jacobian product)
    import torch

    from sqdft_amd.autograd import SQDensity
    S, osys, veff = cases["small_lih"]
    say("gradcheck")
    sq = SpectralQuadrature(S, precision="fp64", fermi_tol=1e-14)
    lf = sq.density(veff, scf_iter=1)["lambda_f"]
    v = torch.tensor(veff, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: SQDensity.apply(x, sq, lf - 1, lf + 1), (v,), eps=1e-5,
                                    atol=1e-6, rtol=1e-4, fast_mode=True)
