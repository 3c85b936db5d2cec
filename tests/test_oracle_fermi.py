"""Fermi level, density assembly and their invariants on the oracle (CPU)."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import oracle as O
from sqdft_amd.configs import smooth_veff


@pytest.fixture(scope="module")
def al_density():
    S, Rcut, T = H.small_al()
    osys = H.oracle_for(S, Rcut, T)
    veff = smooth_veff(S)
    d = osys.density(veff, scf_iter=1, nthreads=4)
    assert d["status"] == 0
    return S, osys, veff, d


def test_electron_count_conserved(al_density):
    S, osys, veff, d = al_density
    h3 = S.info["h"] ** 3
    # Brent stops within fermi_tol of the root of the (signed) constraint
    g = osys._lib.sqo_ne_constraint(C.byref(osys.S), S.nn, O._d(d["chi"]), O._d(d["zee"]),
                                    O._d(np.ascontiguousarray(d["mu"])), -12.0, d["lambda_f"])
    assert abs(g) < 1e-4
    assert d["rho"].sum() * h3 == pytest.approx(12.0, abs=1e-4)


def test_brent_root_matches_bisection(al_density):
    S, osys, veff, d = al_density
    f = lambda lf: osys._lib.sqo_ne_constraint(  # noqa: E731
        C.byref(osys.S), S.nn, O._d(d["chi"]), O._d(d["zee"]), O._d(np.ascontiguousarray(d["mu"])),
        -12.0, lf)
    a, b = d["lambda_f"] - 1.0, d["lambda_f"] + 1.0
    fa = f(a)
    for _ in range(80):
        m = 0.5 * (a + b)
        fm = f(m)
        if (fm > 0) == (fa > 0):
            a, fa = m, fm
        else:
            b = m
    assert d["lambda_f"] == pytest.approx(0.5 * (a + b), abs=S.info["fermi_tol"])


def test_newton_agrees_with_brent(al_density):
    S, osys, veff, d = al_density
    root, it = C.c_double(), C.c_int()
    st = osys._lib.sqo_newton(C.byref(osys.S), S.nn, O._d(d["chi"]), O._d(d["zee"]),
                              O._d(np.ascontiguousarray(d["mu"])), -12.0, 0.2, C.byref(root), C.byref(it))
    assert st == 0
    assert root.value == pytest.approx(d["lambda_f"], abs=2 * S.info["fermi_tol"])


def test_local_only_closed_form():
    # coeff_lap = 0 and no projectors: H_i = diag(Veff on the cube), so
    # rho_i = 2 p(Veff_i)/h^3 with p the Chebyshev interpolant of the
    # Fermi-Dirac function on [lmin, lmax] (sq.cpp:740-765)
    S, Rcut, T = H.small_al()
    osys = H.oracle_for(S, Rcut, T)
    for a in range(16):
        osys.S.coeff_lap[a] = 0.0
    osys.S.n_img = 0
    veff = smooth_veff(S)
    d = osys.density(veff, scf_iter=1, nthreads=4)
    assert d["status"] == 0
    h3 = S.info["h"] ** 3
    bet = S.info["bet"]
    lf = d["lambda_f"]
    fd = 1 / (1 + np.exp(bet * (veff - lf)))
    # the Chebyshev interpolant of degree npl approximates f to ~1e-3 here
    assert np.max(np.abs(d["rho"] * h3 / 2 - fd)) < 5e-3
    # exact statement: the same interpolant evaluated at the node's own Veff
    nd = 17
    chi, zee = d["chi"][nd], d["zee"][nd]
    Cc = O.cheb_coeff(osys.npl, (lf - chi) / zee, bet * zee, 0)
    x = (veff[nd] - chi) / zee
    T0, T1, p = 1.0, x, Cc[0] + Cc[1] * x
    for j in range(2, osys.npl + 1):
        T0, T1 = T1, 2 * x * T1 - T0
        p += Cc[j] * T1
    assert d["rho"][nd] == pytest.approx(2 * p / h3, rel=1e-10)
