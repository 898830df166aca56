"""Topological charge: exact invariants and analytic textures, for the
reference restatement (calculate_winding_number2) and the build form."""
import numpy as np
import pytest


def skyrmion(lx, ly, lz, radius=3.5, sign=1, rotate=None, cx=None, cy=None):
    s = np.zeros((3, lx * ly * lz))
    cx = lx / 2 - 0.3 if cx is None else cx
    cy = ly / 2 - 0.2 if cy is None else cy
    for ix in range(lx):
        for iy in range(ly):
            dx, dy = ix - cx, iy - cy
            th = np.pi * np.exp(-(np.hypot(dx, dy) / radius) ** 2)
            ph = sign * np.arctan2(dy, dx)
            v = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
            if rotate is not None:
                v = rotate @ v
            for iz in range(lz):
                s[:, (ix * ly + iy) * lz + iz] = v
    return s


def rand_rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_single_triangle_octant(orc):
    assert orc.ref_arc_winding([1, 0, 0], [0, 1, 0], [0, 0, 1]) == pytest.approx(np.pi / 2, abs=1e-7)


def test_ferromagnet_zero(orc):
    L = orc.lattice(16, 16, 2)
    s = np.zeros((3, 512)); s[2] = 1
    assert orc.winding_layer(L, s, 0) == (0, 0.0)
    assert orc.ref_winding_number(L, s, 1) == 0.0


@pytest.mark.parametrize("sign", [1, -1])
def test_skyrmion_charge(orc, sign):
    L = orc.lattice(24, 24, 2)
    s = skyrmion(24, 24, 2, sign=sign)
    q, raw = orc.winding_layer(L, s, 1)
    assert abs(q) == 1 and abs(raw - q) < 1e-9
    qs = [orc.winding_layer(L, skyrmion(24, 24, 2, sign=-sign), 1)[0]]
    assert qs[0] == -q
    assert abs(orc.ref_winding_number(L, s, 2) - q) < 1e-5


def test_two_skyrmions_add(orc):
    L = orc.lattice(32, 16, 1, (True, True, False))
    a = skyrmion(32, 16, 1, radius=2.5, cx=7.7, cy=7.6)
    b = skyrmion(32, 16, 1, radius=2.5, cx=23.6, cy=8.3)
    s = np.where(np.arange(32 * 16)[None, :] < 16 * 16, a, b)
    s /= np.linalg.norm(s, axis=0)
    q, raw = orc.winding_layer(L, s, 0)
    assert abs(q) == 2


def test_invariants_random(orc):
    rng = np.random.default_rng(3)
    L = orc.lattice(12, 10, 3)
    for _ in range(5):
        v = rng.normal(size=(3, 360)); v /= np.linalg.norm(v, axis=0)
        v = v.astype(np.float32).astype(np.float64)
        for z in range(3):
            q, raw = orc.winding_layer(L, v, z)
            assert abs(raw - q) < 1e-9                     # integer on a closed surface
            qm, rawm = orc.winding_layer(L, -v, z)
            assert qm == -q and rawm == -raw               # Q(-S) = -Q(S) bit for bit
            R = rand_rot(rng)
            qr, rawr = orc.winding_layer(L, R @ v, z)
            assert qr == q                                  # global rotation
            ref = orc.ref_winding_number(L, v, z + 1)
            assert abs(ref - raw) < 1e-5 and int(np.rint(ref)) == q


def test_mirror_flips_sign(orc):
    """Mirror x -> -x of the lattice combined with the axial-vector transform
    of the spins (S -> (Sx, -Sy, -Sz)) reverses Q. A smooth texture is used:
    the mirror swaps the plaquette diagonal, which only smooth fields ignore."""
    rng = np.random.default_rng(8)
    lx, ly, lz = 24, 24, 1
    L = orc.lattice(lx, ly, lz)
    v = skyrmion(lx, ly, lz, radius=4.0, rotate=rand_rot(rng))
    m = np.zeros_like(v)
    for ix in range(lx):
        for iy in range(ly):
            src = (ix * ly + iy)
            dst = (((lx - ix) % lx) * ly + iy)
            m[:, dst] = [v[0, src], -v[1, src], -v[2, src]]
    q, _ = orc.winding_layer(L, v, 0)
    qm, _ = orc.winding_layer(L, m, 0)
    assert abs(q) == 1 and qm == -q


def test_density_ferromagnet_and_sums(orc):
    """Charge density per plaquette (calculate_winding_number_density): zero
    for a ferromagnet; per layer it sums to the layer's charge, for both the
    reference restatement (single-precision pi) and the build form."""
    L = orc.lattice(16, 16, 2)
    s = np.zeros((3, 512)); s[2] = 1
    assert not np.any(orc.winding_density(L, s, 0))
    assert not np.any(orc.ref_winding_density(L, s, 1))
    rng = np.random.default_rng(12)
    L = orc.lattice(12, 10, 3)
    v = rng.normal(size=(3, 360)); v /= np.linalg.norm(v, axis=0)
    v = v.astype(np.float32).astype(np.float64)
    for z in range(3):
        d = orc.winding_density(L, v, z)
        q, raw = orc.winding_layer(L, v, z)
        assert d.shape == (10, 12) and abs(d.sum() - raw) < 1e-12
        dr = orc.ref_winding_density(L, v, z + 1)
        assert abs(dr.sum() - orc.ref_winding_number(L, v, z + 1)) < 1e-12
        assert np.max(np.abs(dr - d)) < 1e-6                     # float pi, float Im in the reference
        assert np.array_equal(orc.winding_density(L, -v, z), -d)  # odd under S -> -S, bit for bit


def test_density_locates_skyrmion(orc):
    """A skyrmion centred at (cx, cy): the density integrates to -1/+1, is
    concentrated within its radius, and its centroid sits at the centre."""
    L = orc.lattice(24, 24, 1)
    cx, cy = 10.7, 13.4
    s = skyrmion(24, 24, 1, radius=3.0, cx=cx, cy=cy)
    d = orc.winding_density(L, s, 0)
    q, _ = orc.winding_layer(L, s, 0)
    assert abs(q) == 1 and abs(d.sum() - q) < 1e-9
    iy, ix = np.mgrid[0:24, 0:24]
    w = d / d.sum()
    # plaquette (i, j) spans corners i..i+1, j..j+1: its centre is (i + 1/2, j + 1/2)
    assert abs((w * (ix + 0.5)).sum() - cx) < 0.3 and abs((w * (iy + 0.5)).sum() - cy) < 0.3
    near = np.hypot(ix + 0.5 - cx, iy + 0.5 - cy) < 6
    assert abs(w[near].sum() - 1) < 0.02
