"""tau_int and C_v estimators against closed forms."""
import numpy as np

from ising_amd.analysis import (consistent, independent_samples, mean_error, ratio_bootstrap, specific_heat, tau_int,
                                tau_int_checked)


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
    n, taus, conv = independent_samples(x)
    assert abs(taus[0] - 0.5) < 0.05 and abs(taus[1] - 9.5) < 1.5
    assert conv.all()
    assert abs(n - (10000 + 10000 / 19)) < 0.1 * n


def test_short_series_is_flagged_unconverged():
    """tau far above n / c: the window still 'closes' at the tail (the
    mean-subtracted autocorrelations sum to -1/2), but the estimate is a
    gross underestimate and must be flagged; the same process on a long
    series converges to the closed form."""
    for phi, n in ((0.99, 100), (0.996, 500), (0.998, 2000)):
        exact = 0.5 * (1 + phi) / (1 - phi)
        assert exact > n / 5
        flags = [tau_int_checked(ar1(phi, n, seed))[1] for seed in range(20)]
        assert not any(flags), (phi, n)
    t, ok = tau_int_checked(ar1(0.99, 400000, 5))
    assert ok and abs(t - 99.5) < 0.15 * 99.5, t


def test_unconverged_columns_leave_the_sample_count():
    x = np.stack([ar1(0.0, 2000, 2), ar1(0.995, 2000, 3)], axis=1)
    n, taus, conv = independent_samples(x)
    assert conv.tolist() == [True, False]
    assert abs(n - 2000 / (2 * taus[0])) < 1e-9


def test_ratio_bootstrap_band():
    rng = np.random.default_rng(1)
    num = rng.uniform(1.0, 3.0, size=64)
    den = rng.uniform(1.0, 2.0, size=64)
    r, lo, hi = ratio_bootstrap(num, den)
    assert abs(r - num.sum() / den.sum()) < 1e-12
    assert lo < r < hi and hi - lo < 0.2 * r
    assert ratio_bootstrap(2 * den, den)[0] == 2.0


def test_specific_heat_gaussian():
    rng = np.random.default_rng(0)
    e = rng.normal(scale=3.0, size=100000)
    assert abs(specific_heat(e, 2.0, 10) - 4 * 9 / 10) < 0.05
    assert consistent(1.0, 0.1, 1.25, 0.05) and not consistent(1.0, 0.01, 1.25, 0.01)
