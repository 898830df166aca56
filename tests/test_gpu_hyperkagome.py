"""GPU parity on hyper-kagome cells (src/unit_cells.f90:32-77, the lattice of
the reference's ISING driver, src/ISING.f90:105): twelve atoms per cubic
cell, four nearest neighbours each (the reference's first atom shell,
src/chainMesh.f90:495-591), swept in three colours of four atom types; the
DMI vector of a bond with unit vector r is d_b = sum_a r_a d_a.

Tolerances as in test_gpu_parity.py: single-site dE within 1e-5 of its
floating-point scale; FP64 totals of FP32 spins to 1e-12 relative (1e-11
per site absolute); Q exact as an integer; decision replay with <= 2 sites
differing where |beta dE - ln u| is within FP32 error; the demagnetising
field to 1e-12 of its scale."""
import numpy as np
import pytest

from gpu_util import de_tolerance, engine, olat, random_unit
from ising_amd import Engine, Hamiltonian, IsingError, Ladder, Lattice, geometric_temperatures
from ising_amd._capi import tune_fused
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

HAMS = {
    "bulk_K": Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05),
    "interfacial_ref": Hamiltonian(J=-1.0, D=0.7, dmi="interfacial_ref", Bz=0.3),
}
LATS = {
    "periodic_3x4x2": Lattice(3, 4, 2, cell="hyperkagome"),
    "open_z_4x3x3": Lattice(4, 3, 3, (True, True, False), cell="hyperkagome"),
    "open_all_2x3x2": Lattice(2, 3, 2, (False, False, False), cell="hyperkagome"),
}


def test_init_and_layout():
    lat = Lattice(3, 5, 2, cell="hyperkagome")
    assert lat.sites == 12 * 3 * 5 * 2
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
def test_delta_energy_parity(hname, lname):
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


def test_reference_mode_delta_energy():
    """calculateHeisenbergEnergy restated on hyper-kagome positions
    (minimum-image bonds, D = Dz c x r_{j->i}) against the GPU's table form."""
    lat = Lattice(3, 3, 2, cell="hyperkagome")
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
def test_energy_and_magnetisation(hname, lname):
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


def test_sweep_decision_replay():
    lat = Lattice(6, 5, 4, cell="hyperkagome")
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


def test_winding_exact():
    """Q per layer from the square lattices of atom 12 (each cell's
    firstAtomInMeshCell): bit-exact integer, FP64 value and density against
    the oracle for a skyrmion on that atom."""
    lx = ly = 12
    lat = Lattice(lx, ly, 2, (True, True, False), cell="hyperkagome")
    lo = olat(lat)
    rng = np.random.default_rng(5)
    st = random_unit(rng, lat.sites)
    for ix in range(lx):
        for iy in range(ly):
            dx, dy = ix - 5.6, iy - 6.3
            th = np.pi * np.exp(-(np.hypot(dx, dy) / 2.5) ** 2)
            ph = np.arctan2(dy, dx)
            for iz in range(2):
                st[:, orc.cell_site(lo, ix, iy, iz, 11)] = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph),
                                                            np.cos(th)]
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


def test_pt_run_records_and_swaps():
    """Three-colour sweeps in the PT schedule (graph replay): E and M
    records come from the measure kernel, the swap exponent takes its FP64
    energies; with a zero cone the spins stay put and the swap rounds replay
    on the oracle exactly."""
    lat = Lattice(3, 3, 2, cell="hyperkagome")
    ham = HAMS["bulk_K"]
    betas = 1.0 / geometric_temperatures(0.5, 3.0, 5)
    rng = np.random.default_rng(21)
    with engine(lat, ham, betas, cone=0.0) as e:
        for r in range(5):
            e.set_spins(r, random_unit(rng, lat.sites))
        en = e.energy()
        o = e.pt_run(n_sweeps=30, swap_every=1, measure_every=3, winding_every=0)
        sot = e.temperature_map()
        att, acc = e.swap_stats()
    ref_sot = np.arange(5, dtype=np.int32)
    ra, rc = np.zeros(5, np.int64), np.zeros(5, np.int64)
    for rnd in range(30):
        orc.pt_swap_round(betas, en, ref_sot, e.seed, 0, rnd, ra, rc)
        if (rnd + 1) % 3 == 0:      # the record after round rnd: energies at each temperature
            k = (rnd + 1) // 3 - 1
            assert np.allclose(o["energy"][k], en[ref_sot], rtol=1e-5, atol=0)
    assert np.array_equal(sot, ref_sot)
    assert np.array_equal(att, ra) and np.array_equal(acc, rc)


def test_graph_replay_equals_eager():
    lat = Lattice(4, 4, 3, cell="hyperkagome")
    betas = 1.0 / geometric_temperatures(0.3, 2.0, 4)
    outs = []
    for g in (False, True):
        with engine(lat, HAMS["interfacial_ref"], betas) as e:
            e.init_random()
            o = e.pt_run(n_sweeps=40, swap_every=5, measure_every=2, winding_every=10, use_graph=g)
            o["map"] = e.temperature_map()
            o["spins"] = e.get_spins(1)
            outs.append(o)
    for k in ("energy", "magnetisation", "winding", "winding_raw", "map", "spins"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_mixed_with_sc_and_bcc_ladders():
    """SC, bcc and hyper-kagome ladders in one context: every sweep is three
    colour launches (the third leaves SC and bcc replicas alone); each replica
    follows its own oracle chain; no fused sweep, no learned proposal on the
    hyper-kagome ladder."""
    ham = HAMS["bulk_K"]
    lads = [Ladder(Lattice(8, 8, 4), ham, [1.0, 0.5], 0.5),
            Ladder(Lattice(4, 4, 4, cell="bcc"), ham, [1.5], 0.4),
            Ladder(Lattice(3, 4, 3, cell="hyperkagome"), ham, [2.0, 0.6], 0.5)]
    with Engine(lads, seed=77) as e:
        e.init_random()
        s0 = [e.get_spins(r) for r in range(5)]
        e.sweep(3)
        s1 = [e.get_spins(r) for r in range(5)]
        en = e.energy()
        assert all((f >> 8) == 0 for f in e.tuning_candidates())
        with pytest.raises(IsingError):
            e.set_tuning(tune_fused(4))
        with pytest.raises(IsingError):
            e.mlp_enable(True, ladder=2)
        # the learned proposal on the SC ladder: the hyper-kagome replicas
        # keep their fixed cone in the same (three-colour) launches
        e.mlp_enable(True, ladder=0)
        e.sweep(2)
        s2 = [e.get_spins(r) for r in range(5)]
        e.mlp_enable(False)
        o = e.pt_run(n_sweeps=20, swap_every=5, measure_every=5, winding_every=10)
        assert np.all(np.isfinite(o["energy"]))
    ho = orc.ham_from(ham)
    for r, beta in ((3, 2.0), (4, 0.6)):
        lo = olat(lads[2].lattice)
        s = s1[r].copy()          # the oracle sweeps in place
        orc.sweeps(lo, ho, s, beta, 0.5, 77, r, 3, 2)
        assert np.sum(np.max(np.abs(s - s2[r]), axis=0) > 1e-3) <= 2, r
    assert np.allclose(np.linalg.norm(s2[0], axis=0), 1.0, atol=1e-5)
    slots = [(0, 0, 1.0, 0.5), (1, 0, 0.5, 0.5), (2, 1, 1.5, 0.4), (3, 2, 2.0, 0.5), (4, 2, 0.6, 0.5)]
    for r, li, beta, cone in slots:
        lo = olat(lads[li].lattice)
        s = np.ascontiguousarray(s0[r])
        orc.sweeps(lo, ho, s, beta, cone, 77, r, 0, 3)
        assert np.sum(np.max(np.abs(s - s1[r]), axis=0) > 1e-3) <= 2, r
        assert abs(en[r] - orc.total_energy(lo, ho, s1[r])) < 1e-11 * lads[li].lattice.sites


def test_demagnetisation_field_and_energy():
    """calculate_demagnetisation_field on hyper-kagome atoms (fractions down
    to -3/4: the deposit keeps the reference's 27-cell stencil) against the
    oracle; the total with the term."""
    lat = Lattice(3, 4, 2, cell="hyperkagome")
    rng = np.random.default_rng(4)
    ham = Hamiltonian(J=1.0, D=0.3, Bz=0.2, demag=True)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        H = e.demag_field(0)
        en = e.energy()[0]
    lo = olat(lat)
    want = orc.demag_field(lo, s)
    assert np.max(np.abs(H - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    assert np.max(np.abs(want)) > 1e-3
    e_ref = orc.total_energy(lo, orc.ham(1.0, 0.3, "bulk", Bz=0.2), s) + orc.demag_energy(lo, s, want, orc.MU0_REF)
    assert abs(en - e_ref) < 1e-12 * abs(e_ref)


@pytest.mark.parametrize("interval", [1, 2])
def test_demagnetisation_sweeps_replay(interval):
    """Sweeps with the term (three colour launches adding mu0 H of each site)
    against the oracle's Metropolis_mcs restatement with the same Philox
    counters and field schedule."""
    lat = Lattice(3, 2, 2, cell="hyperkagome")
    ham = Hamiltonian(J=1.0, D=0.5, Bz=0.2, demag=True)
    n = 4
    with engine(lat, ham, [1.3], seed=31) as e:
        e.set_demag_interval(interval)
        e.init_random()
        s = e.get_spins(0)
        e.sweep(n)
        got = e.get_spins(0)
        stored = e.stored_demag_field(0)
    lo = olat(lat)
    ref = s.copy()
    _, H = orc.sweeps_demag(lo, orc.ham(1.0, 0.5, "bulk", Bz=0.2), ref, 1.3, 0.5, 31, 0, 0, n, orc.MU0_REF, interval)
    bad = int(np.sum(np.max(np.abs(ref - got), axis=0) > 1e-3))
    assert bad <= 2, bad
    assert np.max(np.abs(stored - orc.demag_field(lo, got))) < 1e-12     # n is a multiple of the interval


def test_lattice_checks():
    with pytest.raises(IsingError):
        engine(Lattice(1, 3, 3, cell="hyperkagome"), HAMS["bulk_K"], [1.0])
    # an open axis may have one cell
    with engine(Lattice(1, 3, 2, (False, True, True), cell="hyperkagome"), HAMS["bulk_K"], [1.0]) as e:
        e.init_random()
        e.sweep(1)
        s = e.get_spins(0)
        assert np.allclose(np.linalg.norm(s, axis=0), 1.0, atol=1e-5)
