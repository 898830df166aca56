"""The skyrmion count of a layer on the device against the oracle
(oracle/skyrmion_oracle.c): the interpolated charge density
(ising_winding_density_interpolated, replacing
calculate_winding_density_interpolated, src/reciprocalSpaceProcesses.f90:
365-431) to 1e-12, every (threshold, width) entry of the count table and the
selected counts of compute_skyrmion_distribution (:930-983) exactly, on
textures with a known number of skyrmions (the count must also equal that
number, with the sign of the charge), on simple-cubic and multi-atom cells;
the thin-film shape of the skyrmion-lattice configuration (128 x 128 x 4,
16 replicas) at the reference's 10 x 10 upsampling; refusals."""
import numpy as np
import pytest

from gpu_util import engine, olat
from ising_amd import Hamiltonian, IsingError, Lattice
from oracle import pyoracle as orc
from textures import skyrmions

pytestmark = pytest.mark.gpu

HAM = Hamiltonian(J=1.0, D=0.6, dmi="bulk", Bz=0.3)


def to_cell_sites(lat: Lattice, s_cells: np.ndarray) -> np.ndarray:
    """A texture given per cell on every atom of the cell."""
    a = lat.sites // (lat.lx * lat.ly * lat.lz)
    return np.ascontiguousarray(np.repeat(s_cells, a, axis=1))


FILM = (True, True, False)
CASES = {
    "sc_one": (Lattice(32, 32, 2, FILM), [(16, 16)], 5.5, 5),
    "sc_two": (Lattice(40, 40, 2, FILM), [(10, 10), (30, 28)], 5.5, 5),
    "sc_three_rect": (Lattice(48, 40, 3, FILM), [(10, 10), (34, 14), (22, 30)], 5.5, 4),
    "sc_ferro": (Lattice(24, 20, 4), [], 5.0, 4),
    "bcc_two": (Lattice(36, 36, 2, cell="bcc"), [(9, 9), (27, 26)], 5.5, 4),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_counts_match_oracle_and_truth(name):
    lat, centres, radius, sc = CASES[name]
    z = lat.lz - 1
    s = to_cell_sites(lat, skyrmions(lat.lx, lat.ly, lat.lz, centres, radius, noise=0.001, seed=3))
    with engine(lat, HAM, [1.0, 1.0]) as e:
        e.set_spins(0, s)
        e.set_spins(1, np.ascontiguousarray(-s))
        dens = e.winding_density_interpolated(0, z, sc, sc)
        got = [e.skyrmion_distribution(r, z, n_particles=2, scale=(sc, sc)) for r in range(2)]
        spins = [e.get_spins(r) for r in range(2)]
        q = e.winding_all()[1][:, z]
    lo = olat(lat)
    for r in range(2):
        counts, table, d = orc.sk_distribution(lo, spins[r], z, sc, sc, n_part=2)
        if r == 0:
            assert dens.shape == d.shape
            assert np.max(np.abs(dens - d)) < 1e-12
            assert abs(dens.sum() - round(q[0])) < 1e-6
        assert np.array_equal(got[r][1], table), (r, np.argwhere(got[r][1] != table)[:5])
        assert np.array_equal(got[r][0], counts)
        sign = (1 if r == 0 else -1) * (int(np.sign(round(q[0]))) if centres else 1)
        assert got[r][0].tolist() == [sign * len(centres), 0], (r, got[r][0], q)


def test_thin_film_skyrmion_lattice_at_reference_upsampling():
    """128 x 128 x 4 (periodic in-plane, open z), 16 replicas: a 4 x 4
    lattice of skyrmions in the middle layer of replica 5, counted at 10 x 10
    upsampling; every replica's count equals the texture's."""
    lat = Lattice(128, 128, 4, (True, True, False))
    centres = [(16 + 32 * i, 16 + 32 * j) for i in range(4) for j in range(4)]
    s = skyrmions(128, 128, 4, centres, 5.5, noise=1e-4, seed=11)
    with engine(lat, HAM, [1.0] * 16) as e:
        e.init_ferro(-1, 0.0)
        e.set_spins(5, s)
        counts, table = e.skyrmion_distribution(5, 2)
        ferro, _ = e.skyrmion_distribution(0, 2)
        q = e.winding_all()[1][5, 2]
    assert abs(abs(q) - 16) < 1e-6
    assert counts.tolist() == [int(round(q))]
    assert ferro.tolist() == [0]
    assert np.any(table[:, :, 0] == int(round(q)))


def test_monte_carlo_film_against_oracle():
    """A thin film relaxed by the sampler from the skyrmion texture (thermal
    noise in every spin): the device table against the oracle's."""
    lat = Lattice(48, 48, 2, (True, True, False))
    s = skyrmions(48, 48, 2, [(12, 12), (36, 30)], 6.0)
    with engine(lat, HAM, [8.0], cone=0.2) as e:
        e.set_spins(0, s)
        e.sweep(5)
        counts, table = e.skyrmion_distribution(0, 1, scale=(4, 4))
        spins = e.get_spins(0)
    oc, ot, _ = orc.sk_distribution(olat(lat), spins, 1, 4, 4)
    print(f"MC film: device {counts.tolist()}, oracle {oc.tolist()}, "
          f"{int(np.sum(table != ot))} of {table.size} table entries differ", flush=True)
    assert np.array_equal(counts, oc)
    assert np.sum(table != ot) <= 0.01 * table.size


def test_refusals():
    with engine(Lattice(16, 16, 4, (True, False, True)), HAM, [1.0]) as e:
        e.init_random()
        with pytest.raises(IsingError) as ex:
            e.skyrmion_distribution(0, 1)
        assert "UNSUPPORTED" in str(ex.value)
    with engine(Lattice(16, 16, 4), HAM, [1.0]) as e:
        e.init_random()
        for kw in ({"num_thresholds": 0}, {"n_particles": 0}, {"min_threshold": 0.5, "max_threshold": 0.2},
                   {"scale": (0, 10)}):
            with pytest.raises(IsingError):
                e.skyrmion_distribution(0, 1, **kw)
        with pytest.raises(IsingError):
            e.winding_density_interpolated(0, 4)
