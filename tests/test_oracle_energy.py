"""Oracle energies pinned by closed forms and by single-site / total
consistency; the reference-mode restatement of calculateHeisenbergEnergy and
TotalHeisenbergEnergy (src/energyMin.f90:86-229, 499-528) pinned the same way
and against the per-axis DMI form used on the GPU."""
import numpy as np
import pytest


def helix(dims, axis, n, plane_axes):
    """Spins rotating right-handedly about `axis` with wave number q = 2 pi n / L."""
    lx, ly, lz = dims
    q = 2 * np.pi * n / dims[axis]
    s = np.zeros((3, lx * ly * lz))
    a1, a2 = plane_axes
    for ix in range(lx):
        for iy in range(ly):
            for iz in range(lz):
                c = (ix, iy, iz)
                i = (ix * ly + iy) * lz + iz
                s[a1, i] = np.cos(q * c[axis])
                s[a2, i] = np.sin(q * c[axis])
    return s, q


def test_ferromagnet_periodic(orc):
    L = orc.lattice(6, 4, 8)
    N = 6 * 4 * 8
    for J, D, B, K in [(1.0, 0.5, 0.2, 0.05), (-0.7, 0.3, -0.4, 0.0), (2.0, 0.0, 0.0, -0.3)]:
        s = np.zeros((3, N)); s[2] = 1.0
        assert orc.total_energy(L, orc.ham(J, D, "bulk", B, K), s) == pytest.approx(-N * (3 * J + B + K), abs=1e-10)


def test_ferromagnet_open_bond_count(orc):
    dims = (6, 5, 3)
    L = orc.lattice(*dims, (False, True, False))
    lx, ly, lz = dims
    bonds = (lx - 1) * ly * lz + lx * ly * lz + lx * ly * (lz - 1)
    s = np.zeros((3, lx * ly * lz)); s[0] = 1.0
    assert orc.total_energy(L, orc.ham(1.3, 0.4, "bulk"), s) == pytest.approx(-1.3 * bonds, abs=1e-10)


@pytest.mark.parametrize("axis,plane", [(0, (1, 2)), (1, (2, 0)), (2, (0, 1))])
@pytest.mark.parametrize("n", [1, 3])
def test_bulk_helix(orc, axis, plane, n):
    dims = (8, 8, 8)
    L = orc.lattice(*dims)
    J, D, B, K = 1.0, 0.5, 0.2, 0.05
    s, q = helix(dims, axis, n, plane)
    N = s.shape[1]
    e = orc.total_energy(L, orc.ham(J, D, "bulk", B, K), s)
    expect = N * (-J * (2 + np.cos(q)) + D * np.sin(q)) - B * s[2].sum() - K * (s[2] ** 2).sum()
    assert e == pytest.approx(expect, abs=1e-9)


def test_reference_neel_spiral(orc):
    """Reference units: E_ref/N = J_r (2 + cos q) - Dz sin q for S = (sin qx, 0, cos qx)."""
    dims = (8, 6, 4)
    L = orc.lattice(*dims)
    J_r, Dz, B = -1.0, 0.6, 0.3
    s, q = helix(dims, 0, 1, (2, 0))          # (sin qx, 0, cos qx)
    N = s.shape[1]
    e = orc.ref_total_energy(L, J_r, Dz, B, s)
    assert e == pytest.approx(N * (J_r * (2 + np.cos(q)) - Dz * np.sin(q)), abs=1e-9)
    # the per-axis form with the reference mapping gives the same number
    e2 = orc.total_energy(L, orc.ham(-J_r, Dz, "interfacial_ref", B, 0.0, 2.0), s)
    assert e2 == pytest.approx(e, abs=1e-9)


def random_spins(n, rng):
    v = rng.normal(size=(3, n))
    return v / np.linalg.norm(v, axis=0)


@pytest.mark.parametrize("dmi", ["bulk", "interfacial_ref", "none"])
@pytest.mark.parametrize("per", [(1, 1, 1), (1, 1, 0), (0, 0, 0)])
def test_delta_energy_is_total_energy_difference(orc, dmi, per):
    dims = (6, 4, 4)
    L = orc.lattice(*dims, per)
    h = orc.ham(0.9, 0.45, dmi, 0.25, 0.07)
    rng = np.random.default_rng(1)
    s = random_spins(96, rng)
    for i in rng.integers(0, 96, 40):
        sn = random_spins(1, rng)[:, 0]
        e0 = orc.total_energy(L, h, s)
        t = s.copy(); t[:, i] = sn
        e1 = orc.total_energy(L, h, t)
        assert orc.delta_energy(L, h, s, int(i), sn) == pytest.approx(e1 - e0, abs=1e-11)


def test_local_field_is_gradient(orc):
    """dE is linear in S' with slope -h apart from the anisotropy term."""
    L = orc.lattice(4, 4, 4)
    h = orc.ham(1.0, 0.7, "bulk", 0.3, 0.0)
    rng = np.random.default_rng(2)
    s = random_spins(64, rng)
    for i in range(0, 64, 7):
        hf = orc.local_field(L, h, s, i)
        for _ in range(3):
            sn = random_spins(1, rng)[:, 0]
            assert orc.delta_energy(L, h, s, i, sn) == pytest.approx(-(sn - s[:, i]) @ hf, abs=1e-12)


def test_reference_local_energy_matches_per_axis_form(orc):
    """calculateHeisenbergEnergy's geometry (min-image r, D = Dz c x r) against
    d_x = -D y, d_y = D x in the oracle's per-axis form (Zeeman coefficient 1)."""
    dims = (6, 6, 4)
    L = orc.lattice(*dims)
    J_r, Dz, B = 1.2, -0.8, 0.5
    h = orc.ham(-J_r, Dz, "interfacial_ref", B, 0.0)
    rng = np.random.default_rng(4)
    s = random_spins(144, rng)
    for i in range(0, 144, 5):
        sn = random_spins(1, rng)[:, 0]
        old, new = orc.ref_local_energy(L, J_r, Dz, B, s, i, sn)
        assert new - old == pytest.approx(orc.delta_energy(L, h, s, i, sn), abs=1e-12)


def test_reference_total_energy_matches_per_axis_form(orc):
    L = orc.lattice(6, 4, 4)
    rng = np.random.default_rng(6)
    s = random_spins(96, rng)
    J_r, Dz, B = -1.0, 0.4, 0.35
    a = orc.ref_total_energy(L, J_r, Dz, B, s)
    b = orc.total_energy(L, orc.ham(-J_r, Dz, "interfacial_ref", B, 0.0, 2.0), s)
    assert a == pytest.approx(b, abs=1e-11)


def test_magnetisation_closed_forms(orc):
    L = orc.lattice(8, 8, 8)
    s = np.zeros((3, 512)); s[2] = 1.0
    assert np.array_equal(orc.magnetisation(L, s), [0.0, 0.0, 1.0])
    h, _ = helix((8, 8, 8), 0, 1, (1, 2))            # full periods: M = 0
    assert np.allclose(orc.magnetisation(L, h), 0.0, atol=1e-15)


def test_initial_states(orc):
    L = orc.lattice(32, 32, 16)
    s = orc.init_random(L, 3, 1)
    assert np.allclose(np.linalg.norm(s, axis=0), 1.0, atol=1e-15)
    # uniform on the sphere: <z> = 0, <z^2> = 1/3, azimuth uniform
    assert abs(s[2].mean()) < 0.01 and abs((s[2] ** 2).mean() - 1 / 3) < 0.01
    assert abs(np.mean(np.arctan2(s[1], s[0]) > 0) - 0.5) < 0.01
    f = orc.init_ferro(L, 3, 1, 0.1)
    assert set(np.unique(f[2])) <= {-1.0, 1.0} and np.all(f[:2] == 0)
    assert abs(np.mean(f[2] < 0) - 0.1) < 0.005
