"""The workloads of the other configurations at their real shapes, checked
against the oracle:

* thin film 128x128x4 (periodic in-plane, open z), 16 replicas, ferromagnet
  with 10% reversed spins, per-layer Q every 50 sweeps: Q exact as integers;
* 32^3 PT with anisotropy, 32 replicas: energies exact, <E> at hot
  temperatures within 3 sigma of FP64 oracle chains, swap bookkeeping;
* 256 replicas of mixed 24^3 / 32^3 / 48^3 ladders in one launch: energies
  of every lattice size exact, launch tuning keeps results bit-identical.
"""
import numpy as np
import pytest

from gpu_util import olat
from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
from ising_amd.analysis import mean_error
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def test_thin_film_winding_exact():
    lat = Lattice(128, 128, 4, (True, True, False))
    ham = Hamiltonian(J=1.0, D=0.6, dmi="bulk", Bz=0.3)
    T = geometric_temperatures(0.1, 1.0, 16)
    with Engine([Ladder(lat, ham, list(1.0 / T), 0.5)], seed=404) as e:
        e.init_ferro(-1, 0.1)
        o = e.pt_run(n_sweeps=200, swap_every=10, measure_every=0, winding_every=50)
        sot = e.temperature_map()
        q_end, qr_end = e.winding_all()
        spins = [e.get_spins(r) for r in range(16)]
    assert o["winding"].shape == (4, 16, 4)
    lo = olat(lat)
    for t in range(16):
        slot = sot[t]
        assert np.array_equal(o["winding"][-1, t], q_end[slot])
        for z in range(4):
            qo, raw = orc.winding_layer(lo, spins[slot], z)
            assert q_end[slot, z] == qo
            assert abs(qr_end[slot, z] - raw) < 1e-9


def test_pt_32_cubed_with_anisotropy():
    lat = Lattice(32, 32, 32)
    ham = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05)
    T = geometric_temperatures(0.2, 2.0, 32)
    betas = 1.0 / T
    with Engine([Ladder(lat, ham, list(betas), 0.5)], seed=202) as e:
        e.init_random()
        e.pt_run(n_sweeps=600, swap_every=10, measure_every=0)
        o = e.pt_run(n_sweeps=3000, swap_every=10, measure_every=2)
        att, acc = e.swap_stats()
        sot = e.temperature_map()
        en = e.energy()
        spins = {s: e.get_spins(s) for s in (0, 17, 31)}
    lo, ho = olat(lat), orc.ham_from(ham)
    for s, st in spins.items():
        assert abs(en[s] - orc.total_energy(lo, ho, st)) < 1e-9 * lat.sites
    # 360 rounds of alternating pairs: even pairs in even rounds, odd in odd
    assert att[:31].sum() == 180 * 16 + 180 * 15 and att[31] == 0
    assert np.all(acc <= att) and sorted(sot.tolist()) == list(range(32))
    # <E> at the two hottest temperatures against FP64 oracle chains
    orc.set_threads(8)
    for t in (31, 27):
        s = orc.init_random(lo, 7, t)
        orc.sweeps(lo, ho, s, betas[t], 0.5, 7, t, 0, 300)
        es = []
        for k in range(1200):
            orc.sweeps(lo, ho, s, betas[t], 0.5, 7, t, 300 + k, 1)
            if k % 2 == 0:
                es.append(orc.total_energy(lo, ho, s))
        a = mean_error(np.array(es))
        b = mean_error(o["energy"][:, t])
        assert abs(a[0] - b[0]) < 3 * np.hypot(a[1], b[1]), (t, a, b)


def test_batched_uneven_256_replicas():
    sizes = [24, 32, 48]
    T = geometric_temperatures(0.2, 2.0, 32)
    ladders = []
    for b in range(8):
        L = sizes[b % 3]
        ladders.append(Ladder(Lattice(L, L, L), Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.05 * b), list(1.0 / T), 0.5))
    runs = []
    for tuned in (False, True):
        with Engine(ladders, seed=303) as e:
            assert e.n_replicas == 256
            e.init_random()
            n_cand = len(e.tuning_candidates())
            assert any(f >> 8 for f in e.tuning_candidates())      # 48^3 planes fit: fused candidates
            if tuned:
                flags, ms = e.autotune(2)     # 2 timing sweeps per candidate, then 2 more
                assert len(ms) == n_cand
                e.sweep(2)
            else:
                e.sweep(2 * n_cand + 2)
            e.pt_run(n_sweeps=40, swap_every=10, measure_every=10)
            en = e.energy()
            picks = {0: e.get_spins(0), 32 + 5: e.get_spins(37), 64 + 31: e.get_spins(95)}
            runs.append((en, picks))
    (en0, p0), (en1, p1) = runs
    assert np.array_equal(en0, en1)                  # tuning flags change nothing but speed
    for slot, st in p0.items():
        b = slot // 32
        lo = olat(ladders[b].lattice)
        ho = orc.ham_from(ladders[b].ham)
        assert np.array_equal(st, p1[slot])
        assert abs(en0[slot] - orc.total_energy(lo, ho, st)) < 1e-9 * st.shape[1]


def test_schedule_autotune_rule():
    """The sweeps-per-swap tuner returns the most frequent exchange within the
    overhead budget of the fastest candidate; with no budget the fastest
    candidate, with an unbounded one the most frequent."""
    T = geometric_temperatures(0.3, 2.0, 8)
    ladders = [Ladder(Lattice(L, L, L), Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.1), list(1.0 / T), 0.5)
               for L in (16, 24)]
    cand = (1, 2, 5, 10)
    with Engine(ladders, seed=5) as e:
        e.init_random()
        k, ms = e.autotune_schedule(cand, n_sweeps=20, max_overhead=0.05, measure_every=2)
        assert set(ms) == set(cand) and all(t > 0 for t in ms.values())
        fastest = min(ms.values())
        assert k == min(c for c in cand if ms[c] <= 1.05 * fastest)
        k1, _ = e.autotune_schedule(cand, n_sweeps=10, max_overhead=1e9)
        assert k1 == 1
        k0, ms0 = e.autotune_schedule(cand, n_sweeps=10, max_overhead=0.0)
        assert ms0[k0] == min(ms0.values())
        # the tuning runs real sweeps: the chains stay valid
        en = e.energy()
        s = e.get_spins(9)
        lo = olat(ladders[1].lattice)
        assert abs(en[9] - orc.total_energy(lo, orc.ham_from(ladders[1].ham), s)) < 1e-9 * s.shape[1]

