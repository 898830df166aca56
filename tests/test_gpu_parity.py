"""GPU parity: the HIP kernels against the FP64 oracle on identical inputs.

Tolerances (stated per test):
  * single-site dE (FP32 on the GPU): |dE_gpu - dE_64| <= 1e-5 x the
    floating-point scale of the term, |S'-S| (sum_j (|J|+2|D|)|S_j| + |Bz|)
    + |Kz||dSz||S'z+Sz|; and plain relative error <= 1e-5 where |dE| > 1e-2;
  * total energy / magnetisation (FP64 reductions of FP32 spins): relative
    1e-12 (only the summation order differs);
  * topological charge: nint(Q) bit-exact as an integer; FP64 Q within 1e-9.
"""
import numpy as np
import pytest

from gpu_util import de_tolerance, engine, olat, random_unit
from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
from ising_amd._capi import tune_fused
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

HAMS = {
    "bulk_K": Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05),
    "interfacial_ref": Hamiltonian(J=-1.0, D=0.7, dmi="interfacial_ref", Bz=0.3),
    "heisenberg": Hamiltonian(J=1.0, D=0.0, dmi="none", Bz=0.0),
}
LATS = {
    "periodic_8x6x4": Lattice(8, 6, 4),
    "open_z_8x8x5": Lattice(8, 8, 5, (True, True, False)),
    "open_all_6x5x3": Lattice(6, 5, 3, (False, False, False)),
}


def test_init_random_matches_oracle():
    lat = Lattice(16, 16, 16)
    with engine(lat, HAMS["bulk_K"], [1.0, 2.0]) as e:
        e.init_random()
        for r in range(2):
            s = e.get_spins(r)
            ref = orc.init_random(olat(lat), e.seed, r)
            assert np.max(np.abs(s - ref)) < 2e-6
            assert np.max(np.abs(np.linalg.norm(s, axis=0) - 1.0)) < 1e-6


def test_init_ferro_exact():
    lat = Lattice(32, 16, 4, (True, True, False))
    with engine(lat, HAMS["bulk_K"], [1.0]) as e:
        e.init_ferro(-1, 0.1)
        s = e.get_spins(0)
        ref = orc.init_ferro(olat(lat), e.seed, 0, 0.1)
        assert np.array_equal(s, ref)
        assert 0.05 < np.mean(s[2] < 0) < 0.15


def test_set_get_roundtrip_is_fp32():
    lat = Lattice(8, 6, 4)
    rng = np.random.default_rng(1)
    s = random_unit(rng, lat.sites)
    with engine(lat, HAMS["bulk_K"], [1.0]) as e:
        e.set_spins(0, s)
        back = e.get_spins(0)
    assert np.array_equal(back, s.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("hname", sorted(HAMS))
@pytest.mark.parametrize("lname", sorted(LATS))
def test_delta_energy_parity(hname, lname):
    lat, ham = LATS[lname], HAMS[hname]
    rng = np.random.default_rng(7)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        sites = rng.integers(0, lat.sites, size=400).astype(np.int32)
        # Mix of large and small moves; proposals rounded to FP32 as the kernel sees them.
        props = []
        for k, i in enumerate(sites):
            step = [1.0, 0.1, 1e-3][k % 3]
            v = s[:, i] + step * rng.normal(size=3)
            props.append(v / np.linalg.norm(v))
        props = np.array(props).astype(np.float32).astype(np.float64)
        de = e.delta_energy(0, sites, props)
    lo, ho = olat(lat), orc.ham_from(ham)
    ref = np.array([orc.delta_energy(lo, ho, s, int(i), p) for i, p in zip(sites, props)])
    tol = de_tolerance(lo, ham, s, sites, props)
    assert np.all(np.abs(de - ref) <= tol + 1e-12), np.max(np.abs(de - ref) - tol)
    big = np.abs(ref) > 1e-2
    assert np.all(np.abs(de[big] - ref[big]) <= 1e-5 * np.abs(ref[big]))


def test_reference_mode_delta_energy_matches_reference_restatement():
    """calculateHeisenbergEnergy (src/energyMin.f90:86-171) restated with its own
    geometry (min-image bonds, D = Dz c x r) against the GPU's per-axis form."""
    lat = Lattice(8, 8, 6)
    J_r, Dz, B = -1.0, 0.8, 0.4
    ham = Hamiltonian.reference(J_r, Dz, B, total_energy_convention=False)
    rng = np.random.default_rng(3)
    with engine(lat, ham, [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        sites = rng.integers(0, lat.sites, size=300).astype(np.int32)
        props = random_unit(rng, len(sites)).T.astype(np.float32).astype(np.float64)
        de = e.delta_energy(0, sites, props)
    lo = olat(lat)
    ref = np.array([np.subtract(*orc.ref_local_energy(lo, J_r, Dz, B, s, int(i), p)[::-1])
                    for i, p in zip(sites, props)])
    tol = de_tolerance(lo, ham, s, sites, props)
    assert np.all(np.abs(de - ref) <= tol + 1e-12)


@pytest.mark.parametrize("hname", sorted(HAMS))
@pytest.mark.parametrize("lname", sorted(LATS))
def test_energy_and_magnetisation_parity(hname, lname):
    lat, ham = LATS[lname], HAMS[hname]
    rng = np.random.default_rng(11)
    with engine(lat, ham, [1.0, 0.5, 0.25]) as e:
        states = []
        for r in range(3):
            e.set_spins(r, random_unit(rng, lat.sites))
            states.append(e.get_spins(r))
        en = e.energy()
        m = e.magnetisation()
    lo, ho = olat(lat), orc.ham_from(ham)
    for r in range(3):
        ref = orc.total_energy(lo, ho, states[r])
        assert abs(en[r] - ref) <= 1e-12 * max(1.0, np.sum(np.abs(states[r]))) * 10
        assert np.allclose(m[r], orc.magnetisation(lo, states[r]), rtol=0, atol=1e-14)


def test_reference_total_energy_matches_reference_restatement():
    """TotalHeisenbergEnergy (src/energyMin.f90:499-528), Zeeman coefficient 2."""
    lat = Lattice(8, 6, 6)
    J_r, Dz, B = -1.0, 0.6, 0.3
    rng = np.random.default_rng(5)
    with engine(lat, HAMS["bulk_K"], [1.0]) as e:
        e.set_spins(0, random_unit(rng, lat.sites))
        s = e.get_spins(0)
        got = e.energy_with(0, Hamiltonian.reference(J_r, Dz, B))
    ref = orc.ref_total_energy(olat(lat), J_r, Dz, B, s)
    assert abs(got - ref) <= 1e-11 * lat.sites


def skyrmion_texture(lx, ly, lz, radius=4.0, charge_sign=1):
    """Neel skyrmion centred in each layer: core down, far field up."""
    s = np.zeros((3, lx * ly * lz))
    cx, cy = lx / 2 - 0.3, ly / 2 - 0.2
    for ix in range(lx):
        for iy in range(ly):
            dx, dy = ix - cx, iy - cy
            rr = np.hypot(dx, dy)
            theta = np.pi * np.exp(-(rr / radius) ** 2)
            phi = np.arctan2(dy, dx) * charge_sign
            v = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
            for iz in range(lz):
                s[:, (ix * ly + iy) * lz + iz] = v
    return s


@pytest.mark.parametrize("kind", ["random", "skyrmion+", "skyrmion-", "ferro"])
def test_winding_exact(kind):
    lat = Lattice(32, 32, 4, (True, True, False))
    rng = np.random.default_rng(13)
    if kind == "random":
        st = random_unit(rng, lat.sites)
    elif kind == "ferro":
        st = np.zeros((3, lat.sites)); st[2] = 1.0
    else:
        st = skyrmion_texture(32, 32, 4, charge_sign=1 if kind == "skyrmion+" else -1)
    with engine(lat, HAMS["bulk_K"], [1.0]) as e:
        e.set_spins(0, st)
        s = e.get_spins(0)
        q, qr = e.winding_all()
        q1, qr1 = e.winding(0, 2)
    lo = olat(lat)
    for z in range(lat.lz):
        qo, rawo = orc.winding_layer(lo, s, z)
        assert q[0, z] == qo
        assert abs(qr[0, z] - rawo) < 1e-9
        ref = orc.ref_winding_number(lo, s, z + 1)
        assert abs(ref - rawo) < 1e-5
        assert int(np.rint(ref)) == qo
    assert (q1, qr1) == (q[0, 2], qr[0, 2])
    if kind.startswith("skyrmion"):
        assert abs(q[0, 0]) == 1


@pytest.mark.parametrize("lname", ["periodic_8x6x4", "open_z_8x8x5", "open_all_6x5x3"])
def test_winding_density_exact(lname):
    """ising_winding_density against the oracle's per-plaquette density (same
    FP64 operation order: equal to 1e-15 absolute), its layer sums against
    the layer charge (1e-12), and the reference restatement (1e-6)."""
    lat = LATS[lname]
    rng = np.random.default_rng(21)
    st = random_unit(rng, lat.sites)
    with engine(lat, HAMS["bulk_K"], [1.0, 0.5]) as e:
        e.set_spins(1, st)
        e.set_tuning(tune_fused(2))
        e.sweep(1)                      # a fused sweep: the state moves to the other spin buffer
        s = e.get_spins(1)
        dens = [e.winding_density(1, z) for z in range(lat.lz)]
        _, qr = e.winding_all()
    lo = olat(lat)
    for z in range(lat.lz):
        d = dens[z]
        assert d.shape == (lat.ly, lat.lx)
        assert np.max(np.abs(d - orc.winding_density(lo, s, z))) <= 1e-15
        assert abs(d.sum() - qr[1, z]) < 1e-12
        if all(lat.periodic):
            assert np.max(np.abs(d - orc.ref_winding_density(lo, s, z + 1))) < 1e-6


def test_sweep_decision_replay():
    """The same Philox counters on GPU and oracle: GPU FP32 sweeps against the
    FP64 oracle started from the GPU's own FP32 state. Decisions may only
    differ where |beta dE - ln u| is within FP32 error, which over a few
    thousand proposals essentially never happens."""
    lat = Lattice(16, 16, 16)
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


def test_pt_swap_replay():
    """With a zero cone radius the spins do not move, so the energies stay fixed
    and the device swap rounds can be replayed on the oracle exactly."""
    lat = Lattice(4, 4, 4)          # small: energy gaps give intermediate swap odds
    ham = HAMS["bulk_K"]
    T = geometric_temperatures(0.5, 3.0, 6)
    betas = 1.0 / T
    rng = np.random.default_rng(17)
    with engine(lat, ham, betas, cone=0.0) as e:
        for r in range(6):
            e.set_spins(r, random_unit(rng, lat.sites))
        en = e.energy()
        e.pt_run(n_sweeps=40, swap_every=1, measure_every=0, winding_every=0, use_graph=False)
        sot = e.temperature_map()
        att, acc = e.swap_stats()
        en_after = e.energy()
        gb = e.betas()
    assert np.allclose(en_after, en, rtol=1e-5)
    ref_sot = np.arange(6, dtype=np.int32)
    ra, rc = np.zeros(6, np.int64), np.zeros(6, np.int64)
    for rnd in range(40):
        orc.pt_swap_round(betas, en, ref_sot, e.seed, 0, rnd, ra, rc)
    assert np.array_equal(sot, ref_sot)
    assert np.array_equal(att, ra) and np.array_equal(acc, rc)
    assert np.allclose(gb[sot], betas.astype(np.float32))


def test_graph_replay_equals_eager():
    lat = Lattice(16, 16, 8)
    ham = HAMS["bulk_K"]
    betas = 1.0 / geometric_temperatures(0.3, 2.0, 4)
    outs = []
    for g in (False, True):
        with engine(lat, ham, betas) as e:
            e.init_random()
            o = e.pt_run(n_sweeps=60, swap_every=5, measure_every=2, winding_every=10, use_graph=g)
            o["map"] = e.temperature_map()
            o["spins"] = e.get_spins(1)
            outs.append(o)
    for k in ("energy", "magnetisation", "winding", "winding_raw", "map", "spins"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_uneven_ladders_one_launch():
    """Ladders of different lattice sizes in one context match separate contexts."""
    ham = HAMS["bulk_K"]
    lads = [Ladder(Lattice(8, 8, 8), ham, [1.0, 0.5], 0.5), Ladder(Lattice(12, 12, 12), ham, [2.0, 0.8, 0.4], 0.5)]
    with Engine(lads, seed=99) as e:
        e.init_random()
        e.sweep(4)
        both = [e.get_spins(r) for r in range(5)]
        en = e.energy()
    lo = [olat(l.lattice) for l in lads]
    ho = orc.ham_from(ham)
    for r in range(5):
        li = 0 if r < 2 else 1
        assert abs(en[r] - orc.total_energy(lo[li], ho, both[r])) < 1e-9 * lads[li].lattice.sites
    # replay replica 3 (ladder 1, beta 0.8) on the oracle from the init state
    s = orc.init_random(lo[1], 99, 3)
    with Engine(lads, seed=99) as e2:
        e2.init_random()
        s = np.ascontiguousarray(e2.get_spins(3))
    orc.sweeps(lo[1], ho, s, 0.8, 0.5, 99, 3, 0, 4)
    assert np.sum(np.max(np.abs(s - both[3]), axis=0) > 1e-3) <= 2


def test_paramagnet_langevin():
    """J = D = K = 0: independent spins, <S_z> = coth(beta B) - 1/(beta B)."""
    lat = Lattice(32, 32, 32)
    B = 1.0
    betas = [0.5, 1.0, 3.0]
    ham = Hamiltonian(J=0.0, D=0.0, dmi="none", Bz=B)
    with engine(lat, ham, betas, cone=1.5) as e:
        e.init_random()
        e.sweep(50)
        o = e.pt_run(n_sweeps=200, swap_every=0, measure_every=1)
    mz = o["magnetisation"][:, :, 2]
    for k, b in enumerate(betas):
        exact = 1.0 / np.tanh(b * B) - 1.0 / (b * B)
        # each sample averages 32768 spins; sigma of one sample ~ 0.0055/sqrt(2)
        assert abs(mz[:, k].mean() - exact) < 1.5e-3, (b, mz[:, k].mean(), exact)


def test_errors_are_reported():
    with pytest.raises(Exception):
        Engine([Ladder(Lattice(7, 8, 8), HAMS["bulk_K"], [1.0])])
    lat = Lattice(8, 8, 8)
    with engine(lat, HAMS["bulk_K"], [1.0]) as e:
        bad = np.zeros((3, lat.sites)); bad[2] = 1.0; bad[0, 5] = np.nan
        with pytest.raises(Exception) as ei:
            e.set_spins(0, bad)
        assert "BAD_SPIN" in str(ei.value)
        with pytest.raises(Exception):
            e.winding(0, 99)


@pytest.mark.parametrize("lname", sorted(LATS))
@pytest.mark.parametrize("hname", ["bulk_K", "interfacial_ref"])
def test_fused_measurement_matches_exact(lname, hname):
    """E and M recorded by the colour-1 launch (per-site FP32, FP64 sums)
    against the FP64 measure kernel on the same configuration:
    relative 1e-6 for E (FP32 per-site terms), 1e-6 absolute for M."""
    lat, ham = LATS[lname], HAMS[hname]
    with engine(lat, ham, [1.5, 0.6, 0.3]) as e:
        e.init_random()
        for _ in range(3):
            o = e.pt_run(n_sweeps=2, swap_every=0, measure_every=2, winding_every=0)
            sot = e.temperature_map()
            exact = e.energy()[sot]
            mag = e.magnetisation()[sot]
            scale = np.maximum(np.abs(exact), lat.sites)
            assert np.all(np.abs(o["energy"][-1] - exact) <= 1e-6 * scale)
            assert np.allclose(o["magnetisation"][-1], mag, atol=1e-6)


SWAP_LATS = dict(LATS, bcc_periodic_4x3x3=Lattice(4, 3, 3, cell="bcc"),
                 bcc_open_z_3x4x3=Lattice(3, 4, 3, (True, True, False), cell="bcc"))


@pytest.mark.parametrize("hname", sorted(HAMS))
@pytest.mark.parametrize("lname", sorted(SWAP_LATS))
def test_swap_sweep_energies_are_exact(lname, hname):
    """The sweep before a swap records FP64 per-site energies from its
    colour-1 launch (the swap exponent's energies): they equal the FP64
    measure kernel's on the same configuration to 1e-12 relative, and the
    magnetisation to 1e-12."""
    lat, ham = SWAP_LATS[lname], HAMS[hname]
    with engine(lat, ham, [1.5, 0.6, 0.3]) as e:
        e.init_random()
        for _ in range(3):
            o = e.pt_run(n_sweeps=2, swap_every=2, measure_every=2, winding_every=0, use_graph=False)
            sot = e.temperature_map()
            exact = e.energy()[sot]
            mag = e.magnetisation()[sot]
            assert np.all(np.abs(o["energy"][-1] - exact) <= 1e-12 * np.maximum(np.abs(exact), lat.sites))
            assert np.allclose(o["magnetisation"][-1], mag, rtol=0, atol=1e-12)


@pytest.mark.parametrize("tuning", [0, 1, tune_fused(4)])
def test_combined_finalize_records(tuning):
    """Sweeps that record E/M on their launches and the charge finish both
    sums in one finalize launch: the records equal separate measurements of
    the final state (E and M to the FP32-per-site bar, Q exactly), with
    graph replay and eager launches giving identical series."""
    lat = Lattice(16, 16, 8)
    betas = 1.0 / geometric_temperatures(0.5, 2.0, 3)
    outs = []
    for g in (False, True):
        with engine(lat, HAMS["bulk_K"], betas, seed=13) as e:
            e.set_tuning(tuning)
            e.init_random()
            o = e.pt_run(n_sweeps=12, swap_every=0, measure_every=2, winding_every=2, use_graph=g)
            en, m = e.energy(), e.magnetisation()
            q, qr = e.winding_all()
            sot = e.temperature_map()
        assert np.allclose(o["energy"][-1], en[sot], rtol=1e-6, atol=0)
        assert np.allclose(o["magnetisation"][-1], m[sot], rtol=0, atol=1e-6)
        assert np.array_equal(o["winding"][-1], q[sot])
        assert np.allclose(o["winding_raw"][-1], qr[sot], rtol=0, atol=1e-12)
        outs.append(o)
    for k in ("energy", "magnetisation", "winding", "winding_raw"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_time_pt_sweeps_advances_like_pt_run():
    """The in-schedule kernel timing runs the same chain as ising_pt_run:
    from one checkpoint both end in the same spins and temperature map."""
    lat = Lattice(12, 12, 8)
    betas = 1.0 / geometric_temperatures(0.4, 2.0, 4)
    with engine(lat, HAMS["bulk_K"], betas) as e:
        e.init_random()
        e.pt_run(n_sweeps=20, swap_every=5, measure_every=0)
        e.save_state()
        e.pt_run(n_sweeps=40, swap_every=10, measure_every=2, winding_every=4, winding_layer=3)
        a = [e.get_spins(r) for r in range(4)] + [e.temperature_map()]
        e.restore_state()
        ms = e.time_pt_sweeps(40, 10, 2, 4, winding_layer=3)
        b = [e.get_spins(r) for r in range(4)] + [e.temperature_map()]
    assert ms > 0
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_single_layer_series_are_that_layer_of_the_full_series():
    """A schedule that names one layer records that layer's charges alone
    ([n_q, R]), equal to its column of the every-layer series from the same
    checkpoint, through the allocating and the allocation-free entry points."""
    lat = Lattice(12, 12, 8)
    betas = 1.0 / geometric_temperatures(0.4, 2.0, 4)
    with engine(lat, HAMS["bulk_K"], betas) as e:
        e.init_random()
        e.pt_run(n_sweeps=20, swap_every=5, measure_every=0)
        e.save_state()
        full = e.pt_run(n_sweeps=40, swap_every=10, measure_every=2, winding_every=4)
        e.restore_state()
        one = e.pt_run(n_sweeps=40, swap_every=10, measure_every=2, winding_every=4, winding_layer=3)
        e.restore_state()
        st = e.pt_run(n_sweeps=40, swap_every=10, measure_every=2, winding_every=4, winding_layer=3, static=True)
    assert full["winding"].shape == (10, 4, 8) and one["winding"].shape == (10, 4)
    assert np.array_equal(one["winding"], full["winding"][:, :, 3])
    assert np.array_equal(one["winding_raw"], full["winding_raw"][:, :, 3])
    for k in ("winding", "winding_raw", "energy", "magnetisation"):
        assert np.array_equal(st[k], one[k]), k
    assert np.array_equal(one["energy"], full["energy"])


def test_checkpoint_restore_replays_exactly():
    """ising_save_state / ising_restore_state: a restored run repeats the
    same random numbers, swaps and records bit for bit."""
    lat = Lattice(12, 12, 8)
    betas = 1.0 / geometric_temperatures(0.4, 2.0, 4)
    with engine(lat, HAMS["bulk_K"], betas) as e:
        e.init_random()
        e.pt_run(n_sweeps=30, swap_every=5, measure_every=0)
        e.save_state()
        a = e.pt_run(n_sweeps=40, swap_every=5, measure_every=2, winding_every=10)
        sa, ma = e.get_spins(2), e.temperature_map()
        e.restore_state()
        b = e.pt_run(n_sweeps=40, swap_every=5, measure_every=2, winding_every=10)
        sb, mb = e.get_spins(2), e.temperature_map()
    assert np.array_equal(sa, sb) and np.array_equal(ma, mb)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
