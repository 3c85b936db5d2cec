"""Pins the oracle's sub-Hamiltonian, Lanczos bounds and Chebyshev moments
against a dense H_i assembled from its definition on tiny node cubes
(brute force with numpy), plus invariants.  CPU only."""
import numpy as np
import pytest

import helpers as H


def dense_hsub(osys, veff, node):
    """H_i = -1/2 Lap_FD + diag(Veff) + sum_J sum_lm u u^T / Denom on the
    (2 nloc+1)^3 cube with zero Dirichlet values outside (sq.cpp:838-929)."""
    S = osys.S
    n, nloc, M = S.n, S.nloc, 2 * S.nloc + 1
    F = S.FDn
    c = np.array(S.coeff_lap[: F + 1])
    D1 = np.zeros((M, M))
    for i in range(M):
        D1[i, i] = c[0]
        for a in range(1, F + 1):
            if i - a >= 0:
                D1[i, i - a] = c[a]
            if i + a < M:
                D1[i, i + a] = c[a]
    I = np.eye(M)
    lap = np.kron(np.kron(D1, I), I) + np.kron(np.kron(I, D1), I) + np.kron(np.kron(I, I), D1)
    ni, nj, nk = node % n, (node // n) % n, node // (n * n)
    z, y, x = np.meshgrid(np.arange(M), np.arange(M), np.arange(M), indexing="ij")
    gi = ((nk - nloc + z) % n) * n * n + ((nj - nloc + y) % n) * n + ((ni - nloc + x) % n)
    Hd = -0.5 * lap + np.diag(veff[gi.reshape(-1)])
    im = osys.images()
    udv = osys.udv()
    for J in range(len(im["type"])):
        t = im["type"][J]
        nc = S.nlocc[t]
        w = 2 * nc + 1
        cn = im["nn"][J]
        gx = ni - nloc + x
        gy = nj - nloc + y
        gz = nk - nloc + z
        inside = (np.abs(gx - cn[0]) <= nc) & (np.abs(gy - cn[1]) <= nc) & (np.abs(gz - cn[2]) <= nc)
        if not inside.any():
            continue
        q = ((gz - cn[2] + nc) * w + (gy - cn[1] + nc)) * w + (gx - cn[0] + nc)
        pr = 0
        psd = S.psd[t]
        for l in range(psd.lmax + 1):
            if l == S.lloc[t]:
                continue
            for _ in range(2 * l + 1):
                u = np.zeros(M ** 3)
                tab = udv[im["udv_off"][J] + pr * w ** 3: im["udv_off"][J] + (pr + 1) * w ** 3]
                u[inside.reshape(-1)] = tab[q.reshape(-1)[inside.reshape(-1)]]
                Hd += np.outer(u, u) / psd.Denom[l]
                pr += 1
    return Hd


SYSTEMS = {"al": lambda: H.small_al(), "lih": lambda: H.small_lih()}


@pytest.fixture(scope="module", params=list(SYSTEMS))
def sysdata(request):
    S, Rcut, T = SYSTEMS[request.param]()
    from sqdft_amd.configs import smooth_veff
    osys = H.oracle_for(S, Rcut, T)
    return S, osys, smooth_veff(S)


def test_hsub_matches_dense(sysdata):
    S, osys, veff = sysdata
    rng = np.random.Generator(np.random.PCG64(3))
    for node in (0, 7, 100, S.nn - 1):
        Hd = dense_hsub(osys, veff, node)
        v = rng.normal(size=osys.M ** 3)
        assert H.rel(osys.hsub(veff, node, v), Hd @ v) < 1e-13
        assert np.max(np.abs(Hd - Hd.T)) < 1e-12 * np.max(np.abs(Hd))


def test_hsub_shift_and_symmetry(sysdata):
    S, osys, veff = sysdata
    rng = np.random.Generator(np.random.PCG64(4))
    u, v = rng.normal(size=(2, osys.M ** 3))
    node = 33
    hu, hv = osys.hsub(veff, node, u), osys.hsub(veff, node, v)
    assert abs(u @ hv - hu @ v) < 1e-12 * abs(u @ hv)
    hv2 = osys.hsub(veff + 0.37, node, v)
    assert H.rel(hv2 - hv, 0.37 * v) < 1e-12


def test_lanczos_bounds_bracket_dense_spectrum(sysdata):
    S, osys, veff = sysdata
    tol = S.info["lanczos_tol"]
    for node in (0, 51):
        ev = np.linalg.eigvalsh(dense_hsub(osys, veff, node))
        lo, hi, it = osys.lanczos(veff, node)
        # padded by +-tol (sq.cpp:1067-1068).  The stopping rule bounds the
        # change of the estimates, not their error, so a bound may sit
        # slightly inside the spectrum; it may not sit far from it.
        assert lo - ev[0] < 1e-5 and ev[-1] - hi < 1e-5
        assert ev[0] - lo < 1e-3 and hi - ev[-1] < 1e-3
        assert it >= 1


def test_moments_match_dense_chebyshev(sysdata):
    S, osys, veff = sysdata
    h3 = S.info["h"] ** 3
    npl = osys.npl
    for node in (0, 64):
        r = osys.node_moments(veff, node)
        Hd = dense_hsub(osys, veff, node)
        A = (Hd - r["chi"] * np.eye(Hd.shape[0])) / r["zee"]
        lam, V = np.linalg.eigh(A)
        c = (osys.M ** 3) // 2
        wts = V[c] ** 2
        # T_j on the eigenvalues by the scalar three-term recurrence
        T0, T1 = np.ones_like(lam), lam.copy()
        mu = [wts @ T0, wts @ T1]
        for _ in range(2, npl + 1):
            T0, T1 = T1, 2 * lam * T1 - T0
            mu.append(wts @ T1)
        mu = np.array(mu) / h3
        assert np.max(np.abs(r["mu"] - mu)) < 1e-9 * np.max(np.abs(mu))
        assert r["mu"][0] == pytest.approx(1 / h3, rel=1e-15)
