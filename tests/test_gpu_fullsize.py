"""The adjoint workload (32-atom FCC Al cell, 32^3 nodes, npl 64, checkpoint
interval 8) and the large single-GPU workload (864-atom FCC Al cell, 96^3 =
884,736 nodes, batches of 8,192 nodes) at their full sizes, through the C-ABI.

Adjoint: dL/dVeff for L = sum_i w_i rho_i (w ~ N(0,1), fixed electron count,
so the Fermi-level shift is included) in FP64 and in mixed precision, checked
against central differences of the GPU's own frozen-bounds density step
(delta = 1e-5 Ha) on 64 seeded nodes and against the FP64 forward mode along 8
seeded directions (<dL/dV, dv> = <w, drho(dv)>).  The oracle's reverse mode
covers the same arithmetic on small cells (test_gpu_adjoint.py); at this size
it would take hours on the host.

Large workload: the FP64 per-node results of a seeded node subset against the
oracle (1e-10), the same nodes bit for bit whatever the batch size, and one
full density step in 8,192-node batches: the electron count it conserves and
its per-node data equal, bit for bit, to the subset computed on its own."""
import sys
import time

import numpy as np
import pytest

import helpers as H
from sqdft_amd import configs
from sqdft_amd.configs import smooth_veff
from sqdft_amd.sq import SpectralQuadrature

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def adj_case():
    S = configs.build("al32_adj", seed=0)
    assert S.info["npl"] == 64 and S.nn == 32 ** 3
    veff = smooth_veff(S)
    rng = np.random.Generator(np.random.PCG64(3))
    w = rng.normal(size=S.nn)
    say("al32 adjoint: density step (fp64, fermi_tol 1e-14)")
    sq = SpectralQuadrature(S, precision="fp64", ckpt_interval=8, fermi_tol=1e-14)
    d = sq.density(veff, scf_iter=1)
    t0 = time.perf_counter()
    g64 = sq.adjoint_veff(w, fixed_n=True)
    say(f"  fp64 adjoint {time.perf_counter() - t0:.2f} s")
    yield S, veff, w, sq, d, g64, rng
    sq.close()


def test_al32_adjoint_vs_forward_mode(adj_case):
    S, veff, w, sq, d, g64, rng = adj_case
    say("al32 adjoint: 8 forward-mode directions")
    rng = np.random.Generator(np.random.PCG64(31))
    for i in range(8):
        dv = rng.normal(size=S.nn)
        t0 = time.perf_counter()
        lhs, rhs = g64 @ dv, w @ sq.tangent_veff(dv, fixed_n=True)
        say(f"  direction {i}: <g,dv> {lhs:.15e} <w,drho> {rhs:.15e} ({time.perf_counter() - t0:.2f} s)")
        assert lhs == pytest.approx(rhs, rel=1e-10, abs=1e-12 * np.abs(g64).sum() * np.abs(dv).max())


def test_al32_adjoint_vs_central_differences(adj_case):
    S, veff, w, sq, d, g64, rng = adj_case
    lf = d["lambda_f"]
    eps = 1e-5
    nodes = np.random.Generator(np.random.PCG64(32)).choice(S.nn, size=64, replace=False)
    say("al32 adjoint: central differences on 64 nodes")
    worst = 0.0
    gmax = np.abs(g64).max()
    for node in nodes:
        dv = np.zeros(S.nn)
        dv[node] = eps
        rp = sq.density_fixed(veff + dv, lf - 1, lf + 1)["rho"]
        rm = sq.density_fixed(veff - dv, lf - 1, lf + 1)["rho"]
        fd = w @ (rp - rm) / (2 * eps)
        worst = max(worst, abs(g64[node] - fd) / max(abs(fd), 1e-3 * gmax))
        assert g64[node] == pytest.approx(fd, rel=2e-5, abs=2e-6 * gmax), node
    say(f"  worst relative difference {worst:.2e}")


def test_al32_mixed_adjoint(adj_case):
    # FP32 vectors, FP64 arithmetic and accumulation; the bounds chi, zee of
    # the mixed step's own Lanczos phase
    S, veff, w, sq, d, g64, rng = adj_case
    say("al32 adjoint: mixed adjoint")
    sqm = SpectralQuadrature(S, precision="mixed", ckpt_interval=8, fermi_tol=1e-14)
    sqm.density(veff, scf_iter=1)
    t0 = time.perf_counter()
    g32 = sqm.adjoint_veff(w, fixed_n=True)
    say(f"  mixed adjoint {time.perf_counter() - t0:.2f} s")
    err = H.rel(g32, g64)
    say(f"  rel err vs fp64 {err:.3e}")
    assert err < 1e-4
    # and against the FP64 forward mode along the same directions
    rng = np.random.Generator(np.random.PCG64(31))
    for _ in range(2):
        dv = rng.normal(size=S.nn)
        assert g32 @ dv == pytest.approx(w @ sq.tangent_veff(dv, fixed_n=True), rel=1e-4)
    sqm.close()


# ---------------------------------------------------------------- Al864

@pytest.fixture(scope="module")
def al864():
    t0 = time.perf_counter()
    S = configs.build("al864", seed=0)
    assert S.nn == 96 ** 3
    veff = smooth_veff(S)
    nodes = np.sort(np.random.Generator(np.random.PCG64(41)).choice(S.nn, size=24, replace=False)).astype(np.int32)
    say(f"al864: built in {time.perf_counter() - t0:.1f} s")
    return S, veff, nodes


def test_al864_node_subset_vs_oracle(al864):
    S, veff, nodes = al864
    say("al864: 24 seeded nodes on the GPU (batch 8192) and in the oracle")
    sq = SpectralQuadrature(S, precision="fp64", batch_nodes=8192)
    g = sq.node_moments(veff, nodes)
    sq.close()
    osys = H.oracle_for(S, 8.0, 58000.0)
    t0 = time.perf_counter()
    ref = osys.moments_nodes(veff, nodes)
    say(f"  oracle {time.perf_counter() - t0:.1f} s; Lanczos counts {g['iters'].min()}..{g['iters'].max()}")
    assert np.array_equal(g["iters"], ref["iters"])
    assert H.rel(g["chi"], ref["chi"]) < 1e-10
    assert H.rel(g["zee"], ref["zee"]) < 1e-10
    for i in range(nodes.size):
        assert H.rel(g["mu"][i], ref["mu"][i]) < 1e-10, int(nodes[i])


def test_al864_batch_size_invariance(al864):
    S, veff, nodes = al864
    say("al864: the subset in batches of 8192 and of 5")
    a = SpectralQuadrature(S, precision="fp64", batch_nodes=8192)
    ga = a.node_moments(veff, nodes)
    a.close()
    b = SpectralQuadrature(S, precision="fp64", batch_nodes=5)
    gb = b.node_moments(veff, nodes)
    b.close()
    for k in ("chi", "zee", "mu", "iters"):
        assert np.array_equal(ga[k], gb[k]), k


def test_al864_full_density_step(al864):
    S, veff, nodes = al864
    say("al864: full FP64 density step in 8192-node batches")
    sq = SpectralQuadrature(S, precision="fp64", batch_nodes=8192)
    t0 = time.perf_counter()
    d = sq.density(veff, scf_iter=1)
    dt = time.perf_counter() - t0
    st = d["stats"]
    say(f"  {dt:.1f} s ({S.nn / dt:.0f} nodes/s), lambda_f {d['lambda_f']:.12f}, "
        f"Lanczos {st['lanczos_iters_min']}..{st['lanczos_iters_max']}, Brent {st['fermi_iters']}")
    ne = d["rho"].sum() * S.info["h"] ** 3
    say(f"  electrons {ne:.10f} of {S.total_valence()}")
    assert ne == pytest.approx(S.total_valence(), rel=1e-5)
    nd = sq.node_data()
    sub = SpectralQuadrature(S, precision="fp64", batch_nodes=8192)
    g = sub.node_moments(veff, nodes)
    sub.close()
    assert np.array_equal(nd["iters"][nodes], g["iters"])
    assert np.array_equal(nd["chi"][nodes], g["chi"])
    assert np.array_equal(nd["zee"][nodes], g["zee"])
    assert np.array_equal(nd["mu"][nodes], g["mu"])
    sq.close()
