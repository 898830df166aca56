"""Host-side statistics of Monte Carlo time series.

* integrated autocorrelation time with Sokal's automatic window
  (A. D. Sokal, "Monte Carlo methods in statistical mechanics", 1996:
  tau_int(W) = 1/2 + sum_{t=1}^{W} rho(t), smallest W with W >= c*tau_int(W));
* specific heat C_v = beta^2 (<E^2> - <E>^2) / N;
* error bars of a mean corrected by 2 tau_int;
* tau_int-corrected independent samples, the throughput metric.

The reference has no time series (statistics are taken once from the final
configuration, src/PT.f90:312-364), so these routines have no reference
counterpart; they are pinned by the AR(1) closed form in the tests.
"""
from __future__ import annotations

import numpy as np


def autocorrelation(x: np.ndarray) -> np.ndarray:
    """Normalised autocorrelation rho(t), t = 0..n-1, via FFT (biased estimator)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 2:
        return np.ones(max(n, 1))
    y = x - x.mean()
    nfft = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(y, nfft)
    acf = np.fft.irfft(f * np.conj(f), nfft)[:n]
    if acf[0] <= 0:
        return np.concatenate([[1.0], np.zeros(n - 1)])
    return acf / acf[0]


# A window is trusted only when the series spans at least this many
# autocorrelation times (Sokal's estimator is biased low on shorter ones).
MIN_LENGTH_OVER_TAU = 50.0


def tau_int_checked(x: np.ndarray, c: float = 5.0, min_ratio: float = MIN_LENGTH_OVER_TAU) -> tuple[float, bool]:
    """(tau_int, converged): Sokal's windowed estimate in units of the
    sampling interval, and whether it can be trusted.

    On a mean-subtracted series the biased autocorrelations sum to exactly
    -1/2 over all lags t >= 1, so the partial sums fall back to 0 at the tail
    and some window W >= c tau(W) is always found, even when tau is far above
    the series length. The estimate is marked converged only when the window
    closes while the series is at least min_ratio times longer than the
    estimate (n >= min_ratio tau); otherwise tau is a lower bound."""
    rho = autocorrelation(x)
    n = len(rho)
    if n < 2:
        return 0.5, False
    taus = 0.5 + np.cumsum(rho[1:])
    w = np.arange(1, n)
    ok = w >= c * taus
    if not ok.any():
        return float(max(taus[-1], 0.5)), False
    t = float(max(taus[np.argmax(ok)], 0.5))
    return t, bool(n >= min_ratio * t)


def tau_int(x: np.ndarray, c: float = 5.0) -> float:
    """Integrated autocorrelation time in units of the sampling interval
    (Sokal's window; see tau_int_checked for whether it converged)."""
    return tau_int_checked(x, c)[0]


def mean_error(x: np.ndarray, c: float = 5.0) -> tuple[float, float, float]:
    """(mean, standard error corrected for autocorrelation, tau_int)."""
    x = np.asarray(x, dtype=np.float64)
    t = tau_int(x, c)
    n = len(x)
    var = x.var(ddof=1) if n > 1 else 0.0
    return float(x.mean()), float(np.sqrt(var * 2.0 * t / max(n, 1))), t


def specific_heat(e: np.ndarray, beta: float, n_sites: int) -> float:
    """C_v per site from a total-energy series at inverse temperature beta."""
    e = np.asarray(e, dtype=np.float64)
    return float(beta * beta * e.var() / n_sites)


def specific_heat_error(e: np.ndarray, beta: float, n_sites: int, n_blocks: int = 20) -> float:
    """Blocking (jackknife over contiguous blocks) error of specific_heat."""
    e = np.asarray(e, dtype=np.float64)
    nb = min(n_blocks, len(e))
    if nb < 2:
        return float("nan")
    blocks = np.array_split(e, nb)
    vals = []
    for k in range(nb):
        rest = np.concatenate([b for j, b in enumerate(blocks) if j != k])
        vals.append(specific_heat(rest, beta, n_sites))
    vals = np.array(vals)
    return float(np.sqrt((nb - 1) / nb * np.sum((vals - vals.mean()) ** 2)))


def independent_samples(series: np.ndarray, c: float = 5.0) -> tuple[float, np.ndarray, np.ndarray]:
    """Per column (replica) n / (2 tau_int); returns (sum over the columns
    whose tau converged, tau per column, converged flag per column)."""
    s = np.asarray(series, dtype=np.float64)
    if s.ndim == 1:
        s = s[:, None]
    res = [tau_int_checked(s[:, k], c) for k in range(s.shape[1])]
    taus = np.array([r[0] for r in res])
    conv = np.array([r[1] for r in res], dtype=bool)
    return float(np.sum(s.shape[0] / (2.0 * taus[conv]))), taus, conv


def ratio_bootstrap(num: np.ndarray, den: np.ndarray, n_boot: int = 2000, seed: int = 0) -> tuple[float, float, float]:
    """sum(num) / sum(den) over columns, with the 16th and 84th percentiles
    of the same ratio over columns resampled with replacement (a 1-sigma
    band over replicas)."""
    num = np.asarray(num, dtype=np.float64)
    den = np.asarray(den, dtype=np.float64)
    r = float(num.sum() / den.sum())
    if len(num) < 2:
        return r, float("nan"), float("nan")
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(num), size=(n_boot, len(num)))
    b = num[idx].sum(axis=1) / den[idx].sum(axis=1)
    return r, float(np.percentile(b, 16)), float(np.percentile(b, 84))


def consistent(a: float, ea: float, b: float, eb: float, nsigma: float = 3.0) -> bool:
    """|a - b| <= nsigma * sqrt(ea^2 + eb^2)."""
    return abs(a - b) <= nsigma * np.sqrt(ea * ea + eb * eb)
