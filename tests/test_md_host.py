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


def test_extrapolation_near_straight_line():
    """Free flight with positions accumulated step by step as NVE.step does:
    the two history columns agree up to rounding, the Gram matrix is singular
    to working precision and its small singular value is zeroed
    (src/anderson.cpp:392-397), so the coefficients are the minimum-norm
    alpha = beta = 1/2 of the exact straight line."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        x = rng.uniform(0, 20, size=(8, 3))
        v = rng.normal(scale=1e-3, size=(8, 3))
        hist = [x.copy()]
        for _k in range(3):
            x = x + v * 41.341
            hist.append(x.copy())
        al, be = md.extrapolation_coefficients(hist[2], hist[1], hist[0], hist[3])
        assert abs(al - 0.5) < 1e-6 and abs(be - 0.5) < 1e-6, (al, be)


def test_pinv_zeroes_small_singular_values():
    """A Pulay history with two columns equal up to rounding gives the same
    update as the exactly rank-deficient history (the reference's threshold
    m * max(w) * eps), not one blown up by 1/w_min."""
    from sqdft_amd.scf import anderson_extrapolation, pinv_times_vec
    rng = np.random.default_rng(6)
    N = 400
    f, g, res = rng.normal(size=N), rng.normal(size=N), rng.normal(size=N)
    DF_exact = np.stack([f, g, f])
    DF_near = np.stack([f, g, f * (1 + 1e-16 * rng.choice([-1.0, 1.0], size=N))])
    DX = rng.normal(size=(3, N))
    a = anderson_extrapolation(DX, DF_exact, res, 0.1)
    b = anderson_extrapolation(DX, DF_near, res, 0.1)
    assert np.max(np.abs(a - b)) < 1e-8 * np.max(np.abs(a))
    # a full-rank system is solved exactly
    A = rng.normal(size=(4, 4))
    A = A @ A.T + np.eye(4)
    y = rng.normal(size=4)
    assert np.allclose(A @ pinv_times_vec(A, y), y, rtol=0, atol=1e-12)


class _System:
    """Stand-in for sqdft_amd.System: 2 atoms of one type in a 20-bohr cell."""
    n, n_atm = 4, 2

    def __init__(self, x0):
        self._x = np.array(x0, dtype=float)
        self.info = {"L": 20.0, "bet": 1.0 / (md.KB * 3000.0)}

    def atoms(self):
        return self._x.copy(), np.zeros(2, dtype=np.int32)


class _SCF:
    """Stand-in for SCF with a harmonic potential energy and a density whose
    difference from the accumulated atomic densities grows linearly with the
    step (so a linear extrapolation of it is exact)."""
    K = 0.05

    def __init__(self, x0):
        self.S = _System(x0)
        self.x0 = np.array(x0, dtype=float)
        self.rho = np.zeros(64)
        self.step = 0
        self.acc = np.zeros(64)
        self.guesses = []

    def set_positions(self, x):
        self.x = np.array(x)
        self.step += 1
        self.rho_at = np.full(64, 0.5) + 0.01 * self.step
        self.acc = self.acc + self.rho_at

    def set_density_guess(self, g, keep_phi=True):
        self.guesses.append((self.step, np.array(g), keep_phi))

    def initial_potential(self):
        pass

    def run(self, **kw):
        # converged density: accumulated rho_at + D0 + s * step
        self.rho = self.acc + 1.0 + 0.1 * self.step
        d = self.x - self.x0
        return {"free_energy": 0.5 * self.K * np.sum(d * d) / 2, "iterations": 1, "converged": True}

    def ForcesTotal(self):
        return -self.K * (self.x - self.x0)


def test_nve_density_guesses():
    # free flight (no force): the atoms move on straight lines, so the fitted
    # alpha + beta is 1 and a linearly growing density difference is
    # extrapolated exactly
    x0 = [[5.0, 5.0, 5.0], [9.0, 9.0, 9.0]]
    scf = _SCF(x0)
    scf.K = 0.0
    dyn = md.NVE(scf, masses=[1.0], time_step=0.5, T_ion=3000.0)
    dyn.run(8)
    # rho_at at step 1, the last density at steps 2-3, then the extrapolation
    acc = np.zeros(64)
    for k, g, keep in scf.guesses:
        acc = acc + (0.5 + 0.01 * k)
        if k == 1:
            assert np.allclose(g, acc) and not keep
        elif k <= 3:
            assert np.allclose(g, acc - (0.5 + 0.01 * k) + 1.0 + 0.1 * (k - 1)) and keep
        else:
            assert np.allclose(g, acc + 1.0 + 0.1 * k, rtol=0, atol=1e-6), k


def test_nve_leapfrog_conserves_energy():
    # leapfrog on a harmonic potential: the energy oscillates at O(dt^2)
    # without drift, and the momentum stays zero (net force removed)
    scf = _SCF([[5.0, 5.0, 5.0], [9.0, 9.0, 9.0]])
    dyn = md.NVE(scf, masses=[1.0], time_step=0.5, T_ion=3000.0)
    recs = dyn.run(40)
    te = np.array([r["TE"] for r in recs])
    wdt = np.sqrt(_SCF.K / md.AMU) * 0.5          # omega dt of the oscillator
    assert np.max(np.abs(te - te[0])) < 0.5 * wdt ** 2 * np.max([r["KE"] for r in recs])
    assert max(np.max(np.abs(r["velocity_sum"])) for r in recs) < 1e-12
