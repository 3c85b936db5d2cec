"""Pins the oracle's small math pieces against closed forms and independent
library routines (CPU only)."""
from fractions import Fraction

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.linalg import eigvalsh_tridiagonal

import oracle as O

# 12th-order central-difference weights (textbook values, Fornberg 1988)
LAP12 = [Fraction(-5369, 1800), Fraction(12, 7), Fraction(-15, 56), Fraction(10, 189),
         Fraction(-1, 112), Fraction(2, 1925), Fraction(-1, 16632)]
GRAD12 = [Fraction(0), Fraction(6, 7), Fraction(-15, 56), Fraction(5, 63), Fraction(-1, 56),
          Fraction(1, 385), Fraction(-1, 5544)]


def test_fd_coefficients_match_textbook_rationals():
    # readfiles.cpp:370-389: coeff_lap[0] is the per-axis diagonal weight
    h = 0.48625
    lap, grad = O.fd_coeffs(6, h)
    for p in range(7):
        assert lap[p] * h * h == pytest.approx(float(LAP12[p]), rel=1e-14, abs=1e-15)
        assert grad[p] * h == pytest.approx(float(GRAD12[p]), rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("k", range(0, 14))
def test_fd_stencils_exact_on_polynomials(k):
    h = 0.3
    lap, grad = O.fd_coeffs(6, h)
    x0 = 0.7
    f = lambda x: x ** k  # noqa: E731
    d2 = lap[0] * f(x0) + sum(lap[a] * (f(x0 + a * h) + f(x0 - a * h)) for a in range(1, 7))
    d1 = sum(grad[a] * (f(x0 + a * h) - f(x0 - a * h)) for a in range(1, 7))
    ex2 = k * (k - 1) * x0 ** (k - 2) if k >= 2 else 0.0
    ex1 = k * x0 ** (k - 1) if k >= 1 else 0.0
    # the symmetric second-derivative stencil is exact through degree 13,
    # the antisymmetric first-derivative stencil through degree 12
    assert d2 == pytest.approx(ex2, rel=1e-8, abs=1e-8)
    if k <= 12:
        assert d1 == pytest.approx(ex1, rel=1e-8, abs=1e-9)


def test_tridiag_ql_matches_lapack():
    rng = np.random.Generator(np.random.PCG64(7))
    for n in (2, 3, 5, 17, 40, 63):
        d = rng.normal(20, 10, n)
        e = np.abs(rng.normal(10, 3, n))  # Lanczos betas are positive
        st, lo, hi = O.tridiag_eig(d, e)
        assert st == 0
        ref = eigvalsh_tridiagonal(d, e[: n - 1])
        assert lo == pytest.approx(ref[0], rel=1e-13, abs=1e-12)
        assert hi == pytest.approx(ref[-1], rel=1e-13, abs=1e-12)


@pytest.mark.parametrize("func", [0, 1, 2, 3])
@pytest.mark.parametrize("npl", [8, 28, 48, 64])
def test_cheb_coeff_interpolates_at_gauss_nodes(func, npl):
    # discrete Chebyshev interpolation: sum_j C_j T_j(y_k) = f(y_k) at the
    # npl+1 Chebyshev-Gauss nodes (the defining property of sq.cpp:1230-1259)
    lfp, bet = 0.13, 7.5
    C = O.cheb_coeff(npl, lfp, bet, func)
    k = np.arange(npl + 1)
    y = np.cos(np.pi * (k + 0.5) / (npl + 1))
    T = np.cos(np.outer(np.arccos(y), np.arange(npl + 1)))
    p = T @ C
    f = np.array([O.load().sqo_func(func, float(v), lfp, bet) for v in y])
    assert np.max(np.abs(p - f)) < 1e-13


def test_cheb_coeff_of_constant_function():
    # bet -> 0: FermiDirac == 1/2 everywhere -> C_0 = 1/2, C_j = 0
    C = O.cheb_coeff(32, 0.0, 0.0, 0)
    assert C[0] == pytest.approx(0.5, abs=1e-15)
    assert np.max(np.abs(C[1:])) < 1e-14


def test_entropy_function_clamp():
    f = O.load().sqo_func
    # focc within 1e-14 of 0 or 1 gives exactly 0 (sq.cpp:1276-1279)
    assert f(2, 10.0, 0.0, 10.0) == 0.0
    assert f(2, -10.0, 0.0, 10.0) == 0.0
    v = f(2, 0.1, 0.0, 3.0)
    fo = 1 / (1 + np.exp(0.3))
    assert v == pytest.approx(fo * np.log(fo) + (1 - fo) * np.log(1 - fo), rel=1e-14)


def _radial(psd_dir):
    import os
    from sqdft_amd.configs import small_al
    S = small_al(psd_dir=psd_dir)
    R = S.radial(0)
    return S, R


def test_spline_exact_on_cubics(psd_dir):
    _, R = _radial(psd_dir)
    r = R["r"]
    y = 0.3 - 1.1 * r + 0.25 * r ** 2 - 0.02 * r ** 3
    lib = O.load()
    yd = np.zeros_like(r)
    lib.sqo_spline_yd(O._d(r), O._d(y), O._d(yd), r.size)
    xs = np.linspace(0.0, 9.0, 57)
    out = np.zeros_like(xs)
    for i, x in enumerate(xs):
        xa = np.array([x]); ya = np.zeros(1)
        assert lib.sqo_spline_eval(O._d(r), O._d(y), r.size, O._d(xa), O._d(ya), 1, O._d(yd)) == 0
        out[i] = ya[0]
    ex = 0.3 - 1.1 * xs + 0.25 * xs ** 2 - 0.02 * xs ** 3
    assert np.max(np.abs(out - ex)) < 1e-9


def test_spline_matches_scipy_not_a_knot(psd_dir):
    # the reference's end rows are the not-a-knot conditions (spline.cpp:65-94)
    _, R = _radial(psd_dir)
    r, y = R["r"][1:], R["rU"][0][1:]
    lib = O.load()
    yd = np.zeros_like(r)
    lib.sqo_spline_yd(O._d(r), O._d(y), O._d(yd), r.size)
    cs = CubicSpline(r, y, bc_type="not-a-knot")
    assert np.max(np.abs(yd - cs(r, 1))) < 1e-7 * np.max(np.abs(cs(r, 1)))


def test_real_spherical_harmonics_orthonormal():
    f = O.load().sqo_sph_harm
    xg, wg = np.polynomial.legendre.leggauss(24)
    phi = np.linspace(0, 2 * np.pi, 48, endpoint=False)
    lm = [(l, m) for l in range(4) for m in range(-l, l + 1)]
    G = np.zeros((len(lm), len(lm)))
    for ct, w in zip(xg, wg):
        st = np.sqrt(1 - ct * ct)
        for p in phi:
            x, y, z = st * np.cos(p), st * np.sin(p), ct
            v = np.array([f(x, y, z, 1.0, l, m) for l, m in lm])
            G += w * (2 * np.pi / phi.size) * np.outer(v, v)
    assert np.max(np.abs(G - np.eye(len(lm)))) < 1e-12


def test_kb_denominator_against_scipy_spline(psd_dir):
    # Denom_l = sum_r (rU)^2 (rV_l - rV_loc)/r dr / h^3 with dr = 1e-4 up to rc
    S, R = _radial(psd_dir)
    ti = S.type_info(0)
    h = S.info["h"]
    r = R["r"]
    l = 0
    su = CubicSpline(r, R["rU"][l], bc_type="not-a-knot")
    sv = CubicSpline(r, R["rV"][l], bc_type="not-a-knot")
    sl = CubicSpline(r, R["rVloc"], bc_type="not-a-knot")
    dr = 1e-4
    rr = np.arange(1, int(np.ceil(ti["rc"] / dr)) + 1) * dr
    den = np.sum(su(rr) ** 2 * (sv(rr) - sl(rr)) / rr) * dr / h ** 3
    assert ti["denom"][l] == pytest.approx(den, rel=1e-8)


def _lih():
    import helpers as H
    S, _, _ = H.test_lih()
    return S


def test_kb_denominator_p_channel_li():
    # the same rectangle rule for Li's l = 1 channel (lloc = 0): the
    # projector channel of the two-type cell
    S = _lih()
    ti, R, h = S.type_info(0), S.radial(0), S.info["h"]
    assert (ti["lloc"], ti["lmax"], ti["nproj"]) == (0, 1, 3)
    r = R["r"]
    su = CubicSpline(r, R["rU"][1], bc_type="not-a-knot")
    sv = CubicSpline(r, R["rV"][1], bc_type="not-a-knot")
    sl = CubicSpline(r, R["rVloc"], bc_type="not-a-knot")
    dr = 1e-4
    rr = np.arange(1, int(np.ceil(ti["rc"] / dr)) + 1) * dr
    den = np.sum(su(rr) ** 2 * (sv(rr) - sl(rr)) / rr) * dr / h ** 3
    assert ti["denom"][1] == pytest.approx(den, rel=1e-8)
    assert ti["denom"][0] == 0.0   # the local channel has no projector


def test_p_projector_tables_against_analytic_form():
    """UdV_{J,1m}(p) = U_1(r) Y_1m(p - R_J) (V_1(r) - V_loc(r)) at points of
    Li image tables (src/nonlocal.cpp:317-518), with an independent
    not-a-knot spline of rU, rV and the real l = 1 harmonics written from
    sqrt(3/(4 pi)) (y, z, x)/r for m = -1, 0, 1 (nonlocal.cpp:255-262)."""
    S = _lih()
    ti, R, h = S.type_info(0), S.radial(0), S.info["h"]
    nc = ti["nlocc"]
    w = 2 * nc + 1
    r = R["r"]
    su = CubicSpline(r, R["rU"][1], bc_type="not-a-knot")
    sv = CubicSpline(r, R["rV"][1], bc_type="not-a-knot")
    sl = CubicSpline(r, R["rVloc"], bc_type="not-a-knot")
    im, udv = S.images(), S.udv()
    li = np.flatnonzero(im["type"] == 0)
    rng = np.random.Generator(np.random.PCG64(3))
    c1 = np.sqrt(3.0 / (4.0 * np.pi))
    checked = 0
    for J in rng.choice(li, size=6, replace=False):
        tab = udv[im["udv_off"][J]: im["udv_off"][J] + 3 * w ** 3].reshape(3, w, w, w)
        scale = np.max(np.abs(tab))
        assert scale > 0
        for _ in range(40):
            k, j, i = rng.integers(0, w, size=3)
            p = (im["nn"][J] + np.array([i, j, k]) - nc) * h
            d = p - im["xyz"][J]
            rad = np.sqrt(d @ d)
            if rad < 1e-6:
                continue
            radial = su(rad) / rad * (sv(rad) - sl(rad)) / rad
            want = c1 * radial * np.array([d[1], d[2], d[0]]) / rad
            got = tab[:, k, j, i]
            assert np.max(np.abs(got - want)) < 1e-6 * scale, (J, i, j, k, got, want)
            checked += 1
    assert checked > 200
