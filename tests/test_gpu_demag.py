"""GPU parity of the demagnetisation term (ising_hamiltonian.demag):
calculate_demagnetisation_field (src/reciprocalSpaceProcesses.f90:202-281)
computed on the device against the oracle's FP64 restatement
(oracle/demag_oracle.c), and its use in single-site energies
(src/energyMin.f90:113-116,163-166), totals (:223-226) and the field
schedule of Metropolis_mcs (:258-259,362-369).

Tolerances: the field is FP64 from FP32 spins, so it matches the oracle on
the same (FP32-rounded) spins to 1e-12 of its scale; single-site dE within
1e-5 of its floating-point scale (the term mu0 H is added to the FP32 local
field); totals to 1e-12 relative, and the demagnetisation part of a total
(E with the term minus E without, ~1e-6) to 1e-6 of itself."""
import numpy as np
import pytest

from gpu_util import de_tolerance, engine, olat, random_unit
from ising_amd import Engine, Hamiltonian, IsingError, Ladder, Lattice
from ising_amd._capi import MU0_REF, tune_fused
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

MU0 = orc.MU0_REF
LATS = {
    "sc_8x6x4": Lattice(8, 6, 4),
    "sc_4x4x6": Lattice(4, 4, 6),
    "bcc_4x3x3": Lattice(4, 3, 3, cell="bcc"),
    "bcc_5x2x3": Lattice(5, 2, 3, cell="bcc"),     # an axis of two cells: repeated cells in the deposit
    "hk_3x2x2": Lattice(3, 2, 2, cell="hyperkagome"),
    # line transforms L = P Q: P = 3, 3, 4 and several workgroups per axis
    "sc_12x18x16": Lattice(12, 18, 16),
    "bcc_7x13x5": Lattice(7, 13, 5, cell="bcc"),    # prime lengths: P = 1
    "sc_64x16x8": Lattice(64, 16, 8),                # 8 x 8, 4 x 4, 2 x 4: register FFTs
    # one line per workgroup; 1024 takes 80 KB of LDS
    "bcc_600x2x2": Lattice(600, 2, 2, cell="bcc"),
    "bcc_2x2x1024": Lattice(2, 2, 1024, cell="bcc"),
    "bcc_1024x2x2": Lattice(1024, 2, 2, cell="bcc"),
}


def test_mu0_constant():
    assert MU0_REF == MU0


@pytest.mark.parametrize("lname", sorted(LATS))
def test_field_matches_oracle(lname):
    lat = LATS[lname]
    rng = np.random.default_rng(3)
    ham = Hamiltonian(J=1.0, D=0.3)
    with engine(lat, ham, [1.0, 0.5]) as e:
        for r in range(2):
            e.set_spins(r, random_unit(rng, lat.sites))
        for r in range(2):
            s = e.get_spins(r)
            H = e.demag_field(r)
            want = orc.demag_field(olat(lat), s)
            assert np.max(np.abs(H - want)) < 1e-12 * max(1.0, np.max(np.abs(want))), (lname, r)
            assert np.max(np.abs(want)) > 1e-3


def test_field_closed_form_conical_spiral():
    """S = (sin t cos qx, sin t sin qx, cos t) with q along x: the x part is a
    longitudinal mode, the y part a transverse one, z uniform, so
    H = -x sin t P^2 cos(qx), P = cos(q/2) on simple-cubic cells."""
    lat = Lattice(8, 4, 4)
    t, q = 0.9, 2 * np.pi / 8
    ix = np.repeat(np.arange(8), 16)
    s = np.stack([np.sin(t) * np.cos(q * ix), np.sin(t) * np.sin(q * ix), np.full(128, np.cos(t))])
    with engine(lat, Hamiltonian(), [1.0]) as e:
        e.set_spins(0, s)
        H = e.demag_field(0)
    want = np.zeros_like(s)
    want[0] = -np.sin(t) * np.cos(q / 2) ** 2 * np.cos(q * ix)
    assert np.max(np.abs(H - want)) < 1e-6


def test_uneven_ladders_and_selection():
    """Ladders of different cells and sizes in one context: each replica's
    field is its own lattice's; ladders without the term keep no field in
    the sweeps, and ising_demag_field computes any periodic replica's."""
    lads = [Ladder(Lattice(6, 4, 4), Hamiltonian(J=1.0, demag=True), [1.0, 2.0]),
            Ladder(Lattice(3, 3, 2, cell="bcc"), Hamiltonian(J=1.0), [1.0]),
            Ladder(Lattice(4, 4, 4, (True, True, False)), Hamiltonian(J=1.0), [1.0])]
    rng = np.random.default_rng(5)
    with Engine(lads, seed=11) as e:
        for r in range(e.n_replicas):
            e.set_spins(r, random_unit(rng, e.sites(r)))
        e.sweep(1)
        for r in (0, 1, 2):
            lat = e.replica_lattice[r]
            s = e.get_spins(r)
            want = orc.demag_field(olat(lat), s)
            assert np.max(np.abs(e.demag_field(r) - want)) < 1e-12
        with pytest.raises(IsingError):
            e.demag_field(3)          # open z


@pytest.mark.parametrize("lname", ["sc_8x6x4", "bcc_4x3x3", "hk_3x2x2"])
def test_delta_energy_with_field(lname):
    lat = LATS[lname]
    ham = Hamiltonian(J=1.0, D=0.4, Bz=0.1, Kz=0.05, demag=True)
    rng = np.random.default_rng(9)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        # a field large enough for mu0 H to be seen at FP32 resolution
        H = rng.normal(size=(3, lat.sites)) * 3e5
        e.set_demag_field(0, H)
        assert np.array_equal(e.stored_demag_field(0), H)
        sites = rng.integers(0, lat.sites, size=200).astype(np.int32)
        props = random_unit(rng, len(sites)).T
        got = e.delta_energy(0, sites, props)
        oh = orc.ham(1.0, 0.4, "bulk", Bz=0.1, Kz=0.05)
        lo = olat(lat)
        want = np.array([orc.delta_energy_field(lo, oh, s, int(i), p, H, MU0) for i, p in zip(sites, props)])
        base = np.array([orc.delta_energy(lo, oh, s, int(i), p) for i, p in zip(sites, props)])
        tol = de_tolerance(lo, oh, s, sites, props) + 1e-5 * MU0 * np.abs(H[:, sites]).sum(axis=0) * 2
        assert np.all(np.abs(got - want) <= tol)
        assert np.max(np.abs(want - base)) > 0.1


@pytest.mark.parametrize("lname", ["sc_8x6x4", "bcc_5x2x3"])
def test_total_energy_with_field(lname):
    lat = LATS[lname]
    ham = Hamiltonian(J=1.0, D=0.3, Bz=0.2, demag=True)
    off = Hamiltonian(J=1.0, D=0.3, Bz=0.2)
    rng = np.random.default_rng(13)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        lo = olat(lat)
        H = orc.demag_field(lo, s)
        e_ref = orc.total_energy(lo, orc.ham(1.0, 0.3, "bulk", Bz=0.2), s) + orc.demag_energy(lo, s, H, MU0)
        got = e.energy()[0]
        assert abs(got - e_ref) < 1e-12 * abs(e_ref)
        part = e.energy_with(0, ham) - e.energy_with(0, off)
        want = orc.demag_energy(lo, s, H, MU0)
        # a difference of two FP64 totals of size ~N: FP64 rounding ~1e-13
        assert abs(part - want) < 1e-6 * abs(want) + 1e-12
        # ising_energy recomputed the field of the current spins
        assert np.max(np.abs(e.stored_demag_field(0) - H)) < 1e-12


@pytest.mark.parametrize("interval", [1, 3, 0])
def test_sweeps_replay_field_schedule(interval):
    """Decisions of ising_sweep with the term match the oracle's
    Metropolis_mcs restatement (same Philox counters); the stored field is
    the last one the schedule computed."""
    lat = Lattice(6, 4, 4)
    ham = Hamiltonian(J=1.0, D=0.5, Bz=0.2, demag=True)
    n = 6
    with engine(lat, ham, [1.5], seed=21) as e:
        e.set_demag_interval(interval)
        e.init_random()
        s = e.get_spins(0)
        e.sweep(n)
        got = e.get_spins(0)
        stored = e.stored_demag_field(0)
    lo = olat(lat)
    ref = s.copy()              # the oracle sweeps in place
    _, H = orc.sweeps_demag(lo, orc.ham(1.0, 0.5, "bulk", Bz=0.2), ref, 1.5, 0.5, 21, 0, 0, n, MU0, interval)
    bad = int(np.sum(np.max(np.abs(ref - got), axis=0) > 1e-3))
    assert bad <= 2, bad
    if bad == 0:      # same chain: the oracle's last field is the stored one
        assert np.max(np.abs(stored - H)) < 1e-5
    if interval == 0:
        assert np.max(np.abs(stored - orc.demag_field(lo, s))) < 1e-12
    elif n % interval == 0:
        assert np.max(np.abs(stored - orc.demag_field(lo, got))) < 1e-12


def test_fused_flag_then_demag_sweeps_in_place():
    """A context tuned to fused sweeps (which swap the spin buffers) that
    then turns the term on runs in-place half sweeps: the buffer parity must
    not advance, so the chain continues from the right spins."""
    lat = Lattice(6, 4, 4)
    plain = Hamiltonian(J=1.0, D=0.5, Bz=0.2)
    with engine(lat, plain, [1.2], seed=8) as e:
        e.init_random()
        e.set_tuning(tune_fused(2))
        e.sweep(1)
        s = e.get_spins(0)
        e.set_hamiltonian(0, Hamiltonian(J=1.0, D=0.5, Bz=0.2, demag=True))
        e.sweep(3)
        got = e.get_spins(0)
        e.set_hamiltonian(0, plain)
        e.sweep(2)                    # fused again
        got2 = e.get_spins(0)
    lo, oh = olat(lat), orc.ham(1.0, 0.5, "bulk", Bz=0.2)
    ref = s.copy()
    orc.sweeps_demag(lo, oh, ref, 1.2, 0.5, 8, 0, 1, 3, MU0, 1)
    assert int(np.sum(np.max(np.abs(ref - got), axis=0) > 1e-3)) <= 2
    ref2 = got.copy()
    orc.sweeps(lo, oh, ref2, 1.2, 0.5, 8, 0, 4, 2)
    assert int(np.sum(np.max(np.abs(ref2 - got2), axis=0) > 1e-3)) <= 2


def test_time_demag():
    """ising_time_demag recomputes the field of every replica with the term
    from its current spins and leaves it stored; refused without the term."""
    lat = Lattice(8, 6, 4)
    lads = [Ladder(lat, Hamiltonian(J=1.0, D=0.3, demag=True), [1.0, 0.5]), Ladder(lat, Hamiltonian(J=1.0), [0.7])]
    with Engine(lads, seed=5) as e:
        e.init_random()
        assert e.time_demag(3) > 0
        for r in range(2):
            want = orc.demag_field(olat(lat), e.get_spins(r))
            assert np.max(np.abs(e.stored_demag_field(r) - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
        assert np.max(np.abs(e.stored_demag_field(2))) == 0.0
    with engine(lat, Hamiltonian(J=1.0), [1.0]) as e:
        with pytest.raises(IsingError):
            e.time_demag(1)


def test_refusals():
    with pytest.raises(IsingError) as ex:
        engine(Lattice(4, 4, 4, (True, True, False)), Hamiltonian(demag=True), [1.0])
    assert ex.value.code == -6
    with pytest.raises(IsingError) as ex:     # at most 1024 cells per axis
        engine(Lattice(2, 1026, 2, cell="bcc"), Hamiltonian(demag=True), [1.0])
    assert ex.value.code == -6
    with engine(Lattice(4, 4, 4), Hamiltonian(J=1.0, demag=True), [1.0, 2.0]) as e:
        with pytest.raises(IsingError):
            e.pt_run(10, swap_every=5, measure_every=5)
        with pytest.raises(IsingError):
            e.set_tuning(tune_fused(4))
        with pytest.raises(IsingError):
            e.mlp_enable(True)
        with pytest.raises(IsingError):
            e.set_demag_interval(-1)
        assert all((f >> 8) == 0 for f in e.tuning_candidates())
    with engine(Lattice(4, 4, 4, (True, True, False)), Hamiltonian(J=1.0), [1.0]) as e:
        with pytest.raises(IsingError):
            e.set_hamiltonian(0, Hamiltonian(J=1.0, demag=True))
    # the learned sweep kernel carries no demagnetising field: no mix in one context
    lads = [Ladder(Lattice(4, 4, 4), Hamiltonian(J=1.0), [1.0]), Ladder(Lattice(4, 4, 4), Hamiltonian(J=1.0), [2.0])]
    with Engine(lads, seed=3) as e:
        e.mlp_enable(True, ladder=0)
        with pytest.raises(IsingError):
            e.set_hamiltonian(1, Hamiltonian(J=1.0, demag=True))
        e.mlp_enable(False)
        e.set_hamiltonian(1, Hamiltonian(J=1.0, demag=True))
        with pytest.raises(IsingError):
            e.mlp_enable(True, ladder=0)
