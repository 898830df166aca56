"""GPU parity on body-centred cubic cells (the cell of the reference's PT
driver, src/PT.f90:145; src/unit_cells.f90:7-30): the two atoms of a cell
are the two colours of the checkerboard sweep, each site has eight
nearest neighbours, and the DMI vector of a bond (1/sqrt 3)(+-1, +-1, +-1)
is d_b = sum_a sgn_a d_a / sqrt 3.

Tolerances as in test_gpu_parity.py: single-site dE within 1e-5 of its
floating-point scale; FP64 totals of FP32 spins to 1e-12 relative; Q exact
as an integer; decision replay with <= 2 sites differing where
|beta dE - ln u| is within FP32 error; statistics within 4 sigma."""
import numpy as np
import pytest

from gpu_util import de_tolerance, engine, olat, random_unit
from ising_amd import Engine, Hamiltonian, IsingError, Ladder, Lattice, geometric_temperatures
from ising_amd._capi import tune_fused
from ising_amd.analysis import mean_error
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

HAMS = {
    "bulk_K": Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05),
    "interfacial_ref": Hamiltonian(J=-1.0, D=0.7, dmi="interfacial_ref", Bz=0.3),
}
LATS = {
    "periodic_6x5x4": Lattice(6, 5, 4, cell="bcc"),
    "open_z_5x4x3": Lattice(5, 4, 3, (True, True, False), cell="bcc"),
    "open_all_3x4x2": Lattice(3, 4, 2, (False, False, False), cell="bcc"),
}


def test_bcc_init_and_layout():
    lat = Lattice(5, 6, 3, cell="bcc")
    assert lat.sites == 2 * 5 * 6 * 3
    with engine(lat, HAMS["bulk_K"], [1.0, 2.0]) as e:
        e.init_random()
        for r in range(2):
            s = e.get_spins(r)
            assert s.shape == (3, lat.sites)
            assert np.max(np.abs(s - orc.init_random(olat(lat), e.seed, r))) < 2e-6
        rng = np.random.default_rng(1)
        t = random_unit(rng, lat.sites)
        e.set_spins(1, t)
        assert np.array_equal(e.get_spins(1), t.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("hname", sorted(HAMS))
@pytest.mark.parametrize("lname", sorted(LATS))
def test_bcc_delta_energy_parity(hname, lname):
    lat, ham = LATS[lname], HAMS[hname]
    rng = np.random.default_rng(7)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        sites = rng.integers(0, lat.sites, size=300).astype(np.int32)
        props = []
        for k, i in enumerate(sites):
            v = s[:, i] + [1.0, 0.1, 1e-3][k % 3] * rng.normal(size=3)
            props.append(v / np.linalg.norm(v))
        props = np.array(props).astype(np.float32).astype(np.float64)
        de = e.delta_energy(0, sites, props)
    lo, ho = olat(lat), orc.ham_from(ham)
    ref = np.array([orc.delta_energy(lo, ho, s, int(i), p) for i, p in zip(sites, props)])
    tol = de_tolerance(lo, ham, s, sites, props)
    assert np.all(np.abs(de - ref) <= tol + 1e-12), np.max(np.abs(de - ref) - tol)


def test_bcc_reference_mode_delta_energy():
    """calculateHeisenbergEnergy restated on bcc positions (min-image bonds,
    D = Dz c x r) against the GPU's signed per-axis form."""
    lat = Lattice(4, 4, 4, cell="bcc")
    J_r, Dz, B = -1.0, 0.8, 0.4
    ham = Hamiltonian.reference(J_r, Dz, B, total_energy_convention=False)
    rng = np.random.default_rng(3)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        sites = rng.integers(0, lat.sites, size=200).astype(np.int32)
        props = random_unit(rng, len(sites)).T.astype(np.float32).astype(np.float64)
        de = e.delta_energy(0, sites, props)
    lo = olat(lat)
    ref = np.array([np.subtract(*orc.ref_local_energy(lo, J_r, Dz, B, s, int(i), p)[::-1])
                    for i, p in zip(sites, props)])
    assert np.all(np.abs(de - ref) <= de_tolerance(lo, ham, s, sites, props) + 1e-12)


@pytest.mark.parametrize("hname", sorted(HAMS))
@pytest.mark.parametrize("lname", sorted(LATS))
def test_bcc_energy_and_magnetisation(hname, lname):
    lat, ham = LATS[lname], HAMS[hname]
    rng = np.random.default_rng(11)
    with engine(lat, ham, [1.0, 0.5]) as e:
        states = []
        for r in range(2):
            e.set_spins(r, random_unit(rng, lat.sites))
            states.append(e.get_spins(r))
        en, m = e.energy(), e.magnetisation()
    lo, ho = olat(lat), orc.ham_from(ham)
    for r in range(2):
        assert abs(en[r] - orc.total_energy(lo, ho, states[r])) <= 1e-11 * lat.sites
        assert np.allclose(m[r], orc.magnetisation(lo, states[r]), rtol=0, atol=1e-14)


def test_bcc_sweep_decision_replay():
    lat = Lattice(8, 8, 6, cell="bcc")
    ham = HAMS["bulk_K"]
    betas = [2.0, 0.7]
    with engine(lat, ham, betas, cone=0.5) as e:
        e.init_random()
        s0 = [e.get_spins(r) for r in range(2)]
        e.sweep(3)
        s1 = [e.get_spins(r) for r in range(2)]
        acc = e.accept_counts()
    lo, ho = olat(lat), orc.ham_from(ham)
    for r in range(2):
        s = np.ascontiguousarray(s0[r])
        n_acc = orc.sweeps(lo, ho, s, betas[r], 0.5, e.seed, r, 0, 3)
        diff = np.max(np.abs(s - s1[r]), axis=0)
        mism = np.sum(diff > 1e-3)
        assert mism <= 2, f"replica {r}: {mism} sites disagree"
        assert np.median(diff) < 1e-6
        assert abs(int(acc[r]) - n_acc) <= 2 * (mism + 1)


def test_bcc_winding_exact():
    """Q per layer from the square lattices of body-centre atoms (each cell's
    firstAtomInMeshCell): bit-exact integer, FP64 value and density against
    the oracle for a skyrmion on atom 1."""
    lx = ly = 16
    lat = Lattice(lx, ly, 2, (True, True, False), cell="bcc")
    lo = olat(lat)
    rng = np.random.default_rng(5)
    st = random_unit(rng, lat.sites)
    for ix in range(lx):
        for iy in range(ly):
            dx, dy = ix - 7.6, iy - 8.3
            th = np.pi * np.exp(-(np.hypot(dx, dy) / 2.8) ** 2)
            ph = np.arctan2(dy, dx)
            for iz in range(2):
                st[:, orc.bcc_site(lo, ix, iy, iz, 1)] = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]
    with engine(lat, HAMS["bulk_K"], [1.0]) as e:
        e.set_spins(0, st)
        s = e.get_spins(0)
        q, qr = e.winding_all()
        dens = [e.winding_density(0, z) for z in range(2)]
    for z in range(2):
        qo, rawo = orc.winding_layer(lo, s, z)
        assert q[0, z] == qo and abs(qo) == 1
        assert abs(qr[0, z] - rawo) < 1e-9
        assert np.max(np.abs(dens[z] - orc.winding_density(lo, s, z))) <= 1e-15


def test_bcc_paramagnet_langevin():
    lat = Lattice(16, 16, 16, cell="bcc")
    B = 1.0
    betas = [0.5, 1.5, 4.0]
    with engine(lat, Hamiltonian(J=0.0, D=0.0, dmi="none", Bz=B), betas, cone=0.7) as e:
        e.init_random()
        e.sweep(40)
        o = e.pt_run(n_sweeps=200, swap_every=0, measure_every=1)
    mz = o["magnetisation"][:, :, 2]
    for k, b in enumerate(betas):
        exact = 1.0 / np.tanh(b * B) - 1.0 / (b * B)
        m, err, _ = mean_error(mz[:, k])
        assert abs(m - exact) < max(4 * err, 1.5e-3), (b, m, exact, err)


def test_bcc_and_sc_ladders_share_launches():
    """An SC ladder and a bcc ladder in one context: every replica as in its
    own context (Philox counters depend only on seed, slot, site, sweep), and
    the PT schedule with graphs runs over both."""
    lads = [Ladder(Lattice(8, 8, 8), HAMS["bulk_K"], [1.0, 0.5], 0.5),
            Ladder(Lattice(6, 6, 4, cell="bcc"), HAMS["bulk_K"], [2.0, 0.8, 0.4], 0.4)]
    with Engine(lads, seed=99) as e:
        e.init_random()
        o = e.pt_run(n_sweeps=30, swap_every=5, measure_every=5, winding_every=10)
        both = [e.get_spins(r) for r in range(5)]
        en = e.energy()
        assert all((f >> 8) == 0 for f in e.tuning_candidates())      # no fused sweep with bcc replicas
        with pytest.raises(IsingError):
            e.set_tuning(tune_fused(4))
        with pytest.raises(IsingError) as ei:
            e.mlp_enable(True, ladder=1)
        assert "UNSUPPORTED" in str(ei.value)
    assert np.all(np.isfinite(o["energy"]))
    lo = [olat(l.lattice) for l in lads]
    ho = orc.ham_from(HAMS["bulk_K"])
    for r in range(5):
        li = 0 if r < 2 else 1
        assert abs(en[r] - orc.total_energy(lo[li], ho, both[r])) < 1e-9 * lads[li].lattice.sites
        assert np.allclose(np.linalg.norm(both[r], axis=0), 1.0, atol=1e-5)


def test_bcc_learned_matches_fixed_observables_not_applicable():
    """The learned proposal reads six SC neighbours: training records from a
    bcc ladder are refused."""
    with engine(Lattice(4, 4, 4, cell="bcc"), HAMS["bulk_K"], [1.0]) as e:
        with pytest.raises(IsingError):
            e.mlp_collect(8)
