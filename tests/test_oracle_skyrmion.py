"""The skyrmion-count oracle (oracle/skyrmion_oracle.c) against what the
mathematics fixes: band-limited interpolation by an independent numpy FFT
route and exactly at the samples; the fine density's total is the integer
charge of smooth textures; the filter against scipy.ndimage; the count of
textures with a known number of separated skyrmions, its sign under
S -> -S and its translation invariance; the search of
compute_skyrmion_distribution (src/reciprocalSpaceProcesses.f90:930-983)
on hand-built count tables."""
import numpy as np
import pytest
from scipy import ndimage

from textures import skyrmions


def fft_upsample(x: np.ndarray, s: int, axis: int) -> np.ndarray:
    """Zero padding of the spectrum along one axis, the Nyquist term of an
    even length split between +-N/2 (independent of the oracle's sums)."""
    x = np.moveaxis(x, axis, -1)
    n = x.shape[-1]
    X = np.fft.fft(x, axis=-1)
    Y = np.zeros(x.shape[:-1] + (s * n,), dtype=complex)
    h = n // 2
    if n % 2 == 0:
        Y[..., :h] = X[..., :h]
        Y[..., s * n - h + 1:] = X[..., h + 1:]
        Y[..., h] = X[..., h] / 2
        Y[..., s * n - h] = X[..., h] / 2
    else:
        Y[..., :h + 1] = X[..., :h + 1]
        Y[..., s * n - h:] = X[..., h + 1:]
    y = np.fft.ifft(Y, axis=-1).real * s
    return np.moveaxis(y, -1, axis)


@pytest.mark.parametrize("nx,ny,sx,sy", [(12, 10, 3, 4), (9, 8, 5, 2), (7, 11, 2, 3)])
def test_interpolation_is_band_limited_zero_padding(orc, nx, ny, sx, sy):
    rng = np.random.default_rng(nx * ny)
    S = rng.normal(size=(3, ny, nx))
    F = orc.sk_interpolate(S, sx, sy)
    ref = fft_upsample(fft_upsample(S, sx, axis=2), sy, axis=1)
    assert np.max(np.abs(F - ref)) < 1e-12
    assert np.max(np.abs(F[:, ::sy, ::sx] - S)) < 1e-13


def test_coarse_grid_is_the_corner_spins(orc):
    lat = orc.lattice(8, 6, 4)
    s = skyrmions(8, 6, 4, [(3, 2)], 2.5, noise=0.1, seed=1)
    S = orc.sk_coarse(lat, s, 2)
    site = lambda i, j, z: (i * 6 + j) * 4 + z
    for i in range(8):
        for j in range(6):
            assert np.array_equal(S[:, j, i], s[:, site(i, j, 2)].astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("centres", [[(16, 16)], [(8, 8), (24, 22)]])
def test_fine_density_sums_to_the_charge(orc, centres):
    n = 32
    lat = orc.lattice(n, n, 1, (True, True, False))
    s = skyrmions(n, n, 1, centres, 6.0)
    q_coarse = orc.winding_layer(lat, s, 0)[1]
    d = orc.sk_density(orc.sk_interpolate(orc.sk_coarse(lat, s, 0), 4, 4))
    assert abs(abs(q_coarse) - len(centres)) < 1e-6
    assert abs(d.sum() - round(q_coarse)) < 1e-6
    fm = skyrmions(n, n, 1, [], 6.0, noise=0.02, seed=3)
    d0 = orc.sk_density(orc.sk_interpolate(orc.sk_coarse(lat, fm, 0), 4, 4))
    assert abs(d0.sum()) < 1e-9


def test_filter_matches_scipy_wrap(orc):
    rng = np.random.default_rng(2)
    d = rng.normal(size=(30, 22))
    for sigma in (0.95, 0.4, 1e-5):
        G = np.zeros(25)
        orc.lib().orc_sk_gauss_weights(sigma, orc._dp(G))
        ref = ndimage.convolve(d, G.reshape(5, 5).T, mode="wrap")
        assert np.max(np.abs(orc.sk_filter(d, sigma) - ref)) < 1e-13
        assert abs(orc.sk_filter(d, sigma).sum() - d.sum()) < 1e-10
    assert np.max(np.abs(orc.sk_filter(d, 1e-5) - d)) < 1e-15


TEXTURES = {
    "one": (32, [(16, 16)], 1),
    "two": (40, [(10, 10), (30, 28)], 2),
    "three": (48, [(10, 10), (34, 14), (22, 36)], 3),
    "ferromagnet": (32, [], 0),
}


@pytest.mark.parametrize("name", sorted(TEXTURES))
def test_count_of_known_textures(orc, name):
    n, centres, k = TEXTURES[name]
    lat = orc.lattice(n, n, 2)
    s = skyrmions(n, n, 2, centres, 5.5, noise=0.001, seed=7)
    counts, table, _ = orc.sk_distribution(lat, s, 1, 5, 5, n_part=2)
    q = orc.winding_layer(lat, s, 1)[1]
    sign = int(np.sign(round(q))) if k else 1
    assert counts.tolist() == [sign * k, 0], (counts, q)
    flipped, _, _ = orc.sk_distribution(lat, np.ascontiguousarray(-s), 1, 5, 5, n_part=2)
    assert flipped.tolist() == [-sign * k, 0]
    if k:
        assert np.any(table[:, :, 0] == sign * k)


def test_count_is_translation_invariant(orc):
    n = 32
    lat = orc.lattice(n, n, 1, (True, True, False))
    a = skyrmions(n, n, 1, [(9, 12), (25, 27)], 4.5, noise=0.001, seed=9)
    b = skyrmions(n, n, 1, [(14, 5), (30, 20)], 4.5, noise=0.001, seed=9)
    ca, ta, _ = orc.sk_distribution(lat, a, 0, 4, 4)
    cb, tb, _ = orc.sk_distribution(lat, b, 0, 4, 4)
    assert ca.tolist() == cb.tolist() and abs(ca[0]) == 2


def test_selection_follows_the_reference_search(orc):
    T, S = 4, 3
    t = np.zeros((T, S, 2), dtype=np.int32)
    # the first table entry within 0.1 of Q is returned, in threshold-major order
    t[1, 2] = [2, 0]; t[2, 0] = [0, 1]; t[3, 1] = [2, 0]
    assert orc.sk_select(t, 2.02).tolist() == [2, 0]
    t2 = t.copy(); t2[1, 2] = [1, 0]
    assert orc.sk_select(t2, 2.02).tolist() == [0, 1]       # 1*0 + 2*1 = 2
    # none within 0.1: the first entry whose nint came closest, strictly
    t3 = np.zeros((T, S, 1), dtype=np.int32)
    t3[0, 1] = [1]; t3[2, 2] = [3]; t3[3, 0] = [1]
    assert orc.sk_select(t3, 2.6).tolist() == [3]
    assert orc.sk_select(t3, 1.7).tolist() == [1]
    assert orc.sk_select(np.zeros((T, S, 1), dtype=np.int32), 0.3).tolist() == [0]
