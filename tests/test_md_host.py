"""Host pieces of the NVE driver (src/md.cpp): velocity initialisation,
wrap-around and the charge-extrapolation coefficients, checked against
properties the mathematics fixes (no GPU)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from sqdft_amd import md  # noqa: E402


def test_rand_is_the_c_library_sequence():
    # glibc's rand() after srand(1) (the generator Initialize_MD draws from)
    lib = md._crand()
    lib.srand(1)
    assert [lib.rand() for _ in range(3)] == [1804289383, 846930886, 1681692777]


def test_initial_velocities_momentum_and_temperature():
    for vds in (1, 2):
        types = np.array([0] * 5 + [1] * 3)
        m = np.where(types == 0, 26.98, 6.94) * md.AMU
        v = md.initial_velocities(m, types, 5000.0, 2, vds)
        assert np.max(np.abs(v.sum(axis=0))) < 1e-12
        ke = 0.5 * np.sum(m[:, None] * v * v)
        T = 2 * ke / (md.KB * 3 * (types.size - 1))
        assert abs(T / 5000.0 - 1) < 1e-12
        v2 = md.initial_velocities(m, types, 5000.0, 2, vds)
        assert np.array_equal(v, v2)                      # reseeded per type: reproducible
        assert not np.array_equal(v, md.initial_velocities(m, types, 5000.0, 3, vds))


def test_maxwell_boltzmann_components_are_gaussian():
    n = 20000
    types = np.zeros(n, dtype=int)
    m = np.full(n, 26.98 * md.AMU)
    v = md.initial_velocities(m, types, 1000.0, 7, 1).ravel()
    var = md.KB * 1000.0 / m[0]
    assert abs(np.var(v) / var - 1) < 0.02
    kurt = np.mean(v ** 4) / np.mean(v ** 2) ** 2
    assert abs(kurt - 3) < 0.1


def test_wrap_around():
    L = 7.78
    x = np.array([[-0.1, 7.78, 3.0], [7.9, -7.0, 0.0]])
    w = md.wrap_around(x, L)
    assert np.all((w >= 0) & (w < L))
    assert np.allclose(w, [[L - 0.1, 0.0, 3.0], [7.9 - L, L - 7.0, 0.0]], atol=0, rtol=0)


def test_extrapolation_coefficients():
    rng = np.random.default_rng(0)
    a, c = rng.normal(size=(4, 3)), rng.normal(size=(4, 3))
    r = lambda t: a * t * t + c * t  # noqa: E731
    # quadratic motion: r(3) = r(2) + 2 (r(2) - r(1)) - (r(1) - r(0)) exactly
    al, be = md.extrapolation_coefficients(r(2), r(1), r(0), r(3))
    assert abs(al - 2) < 1e-10 and abs(be + 1) < 1e-10
    # linear motion: the two columns coincide; the minimum-norm solution has
    # alpha + beta = 1, which extrapolates a linear drho exactly
    al, be = md.extrapolation_coefficients(c * 2, c * 1, c * 0, c * 3)
    assert abs(al + be - 1) < 1e-12
    D0, s = rng.normal(size=50), rng.normal(size=50)
    dn = (1 + al) * D0 + (be - al) * (D0 - s) - be * (D0 - 2 * s)
    assert np.allclose(dn, D0 + s, rtol=0, atol=1e-12)
