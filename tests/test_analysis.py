"""tau_int and C_v estimators against closed forms."""
import numpy as np

from ising_amd.analysis import consistent, independent_samples, mean_error, specific_heat, tau_int


def ar1(phi, n, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(n)
    x[0] = rng.normal()
    eps = rng.normal(size=n) * np.sqrt(1 - phi * phi)
    for k in range(1, n):
        x[k] = phi * x[k - 1] + eps[k]
    return x


def test_tau_int_ar1():
    for phi in (0.0, 0.5, 0.9):
        exact = 0.5 * (1 + phi) / (1 - phi)
        t = tau_int(ar1(phi, 200000, 1))
        assert abs(t - exact) < 0.06 * exact + 0.02, (phi, t, exact)


def test_mean_error_covers_truth():
    hits = 0
    for seed in range(40):
        m, err, _ = mean_error(ar1(0.8, 20000, seed) + 3.0)
        hits += abs(m - 3.0) < 2 * err
    assert hits >= 32


def test_independent_samples():
    x = np.stack([ar1(0.0, 10000, 2), ar1(0.9, 10000, 3)], axis=1)
    n, taus = independent_samples(x)
    assert abs(taus[0] - 0.5) < 0.05 and abs(taus[1] - 9.5) < 1.5
    assert abs(n - (10000 + 10000 / 19)) < 0.1 * n


def test_specific_heat_gaussian():
    rng = np.random.default_rng(0)
    e = rng.normal(scale=3.0, size=100000)
    assert abs(specific_heat(e, 2.0, 10) - 4 * 9 / 10) < 0.05
    assert consistent(1.0, 0.1, 1.25, 0.05) and not consistent(1.0, 0.01, 1.25, 0.01)
