"""The two configurations no other GPU test runs, at their exact shapes.

C5, the headline: 64^3 periodic simple cubic, 64 replicas on a geometric
ladder T = 0.2..2.0, J = 1, bulk D = 0.5, Bz = 0.2, cone r = 0.5; a fixed-
radius segment and a learned-radius segment (a random network), each with
swaps every 10 sweeps, E/M every 2 and the middle layer's Q every 2, as the
benchmark runs them, with half-sweep launches and with the fused sweep the
benchmark uses at this shape. Checked against the FP64 oracle: the last records' E
and M of every temperature against orc.total_energy / orc.magnetisation of
the final spins (E to 1e-9 N, M to 1e-12), the middle layer's Q as an
integer against orc.winding_layer, one sweep of four replicas replayed on
the oracle with the same Philox counters (orc.sweeps, and
orc.sweeps_learned for the learned radius), and swap rounds replayed on
orc.pt_swap_round from exact energies.

C1: 16^3, J = 1, bulk D = 0.3, B = K = 0, T = 0.5, r = 0.5, 1000 sweeps:
the first three sweeps replayed decision by decision, and <E> and <|M|>
over eight device chains within 3 sigma of four oracle chains.

Reference routines these configurations exercise: Metropolis_mcs
(src/energyMin.f90:234-374), TotalHeisenbergEnergy (:499-528),
calculate_winding_number2 (src/reciprocalSpaceProcesses.f90:320-364) and the
PT loop (src/PT.f90:219-287).
"""
import numpy as np
import pytest

from gpu_util import olat
from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
from ising_amd._capi import TUNE_XCD_REMAP, tune_fused
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

L5 = 64
R5 = 64
MID = L5 // 2
HAM5 = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2)
SEED5 = 20261203
CHECK = (0, 21, 42, 63)          # temperature positions whose spins are checked


def random_network(seed=77, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=scale, size=1793).astype(np.float32)
    p[1792] = np.log(0.5)
    return p


def check_segment(e: Engine, o: dict, lo, ho) -> None:
    """Last E/M/Q records of a segment against the oracle on the final spins."""
    sot = e.temperature_map()
    n = lo.lx * lo.ly * lo.lz
    for t in CHECK:
        slot = int(sot[t])
        s = e.get_spins(slot)
        e_ref = orc.total_energy(lo, ho, s)
        assert abs(o["energy"][-1, t] - e_ref) < 1e-9 * n, (t, o["energy"][-1, t], e_ref)
        m_ref = orc.magnetisation(lo, s)
        assert np.max(np.abs(o["magnetisation"][-1, t] - m_ref)) < 1e-12, (t, o["magnetisation"][-1, t], m_ref)
        q_ref, raw = orc.winding_layer(lo, s, MID)
        # the schedule names the middle layer: its charges alone, [n_q, R]
        assert o["winding"][-1, t] == q_ref, (t, o["winding"][-1, t], q_ref)
        assert abs(o["winding_raw"][-1, t] - raw) < 1e-9


def replay_one_sweep(e: Engine, lo, ho, sweep_index: int, params=None) -> None:
    """One sweep of every replica on device; four of them replayed on the
    oracle from the device's own state with the same counters. Decisions may
    differ only where |beta dE - ln u| is within FP32 error (and, for the
    learned radius, where a bfloat16 rounding of the network lands on the
    other side of a tie), each such flip also changing its neighbours' fields
    in the second colour."""
    slots = [int(x) for x in e.temperature_map()[list(CHECK)]]
    before = {s: np.ascontiguousarray(e.get_spins(s)) for s in slots}
    beta = e.betas()                      # the device's FP32 betas per slot
    e.sweep(1)
    n = lo.lx * lo.ly * lo.lz
    for s in slots:
        ref = before[s].copy()
        if params is None:
            orc.sweeps(lo, ho, ref, float(beta[s]), 0.5, SEED5, s, sweep_index, 1)
            bar = 2e-5
        else:
            orc.sweeps_learned(lo, ho, ref, float(beta[s]), params, SEED5, s, sweep_index, 1)
            bar = 1e-3
        diff = np.max(np.abs(ref - e.get_spins(s)), axis=0)
        bad = int(np.sum(diff > 1e-3))
        print(f"replay slot {s} beta {beta[s]:.4f}: {bad} of {n} sites differ, median {np.median(diff):.2e}",
              flush=True)
        assert bad <= bar * n, (s, bad)
        assert np.median(diff) < (1e-6 if params is None else 1e-4)


def replay_swap_rounds(e: Engine, rounds_done: int, n_rounds: int = 6) -> None:
    """With a zero radius the spins stay put (up to the FP32 renormalisation
    of an unmoved spin), so the exact energies of the current state fix every
    swap decision; the device rounds replayed on the oracle must move the
    same labels and count the same attempts."""
    betas_t = 1.0 / geometric_temperatures(0.2, 2.0, R5)
    e_slot = e.energy()
    sot = e.temperature_map().astype(np.int32)
    att0, acc0 = e.swap_stats()
    e.set_cone(0, 0.0)
    e.pt_run(n_sweeps=n_rounds, swap_every=1, measure_every=0, winding_every=0, use_graph=False)
    e.set_cone(0, 0.5)
    att1, acc1 = e.swap_stats()
    assert np.allclose(e.energy(), e_slot, rtol=1e-6, atol=0)
    ra, rc = np.zeros(R5, np.int64), np.zeros(R5, np.int64)
    ref = sot.copy()
    for k in range(n_rounds):
        orc.pt_swap_round(betas_t, e_slot, ref, SEED5, 0, rounds_done + k, ra, rc)
    assert np.array_equal(e.temperature_map(), ref)
    assert np.array_equal(att1 - att0, ra) and np.array_equal(acc1 - acc0, rc)


@pytest.fixture(scope="module", params=["half_sweeps", "fused16"])
def c5(request):
    """The C5 context and the device counters the replays need: sweeps and
    swap rounds run so far; once with two half-sweep launches per sweep and
    once with the fused sweep the benchmark's autotuner keeps at this shape
    (16 z planes per workgroup, die-contiguous blocks)."""
    T = geometric_temperatures(0.2, 2.0, R5)
    with Engine([Ladder(Lattice(L5, L5, L5), HAM5, list(1.0 / T), 0.5)], seed=SEED5) as e:
        if request.param == "fused16":
            e.set_tuning(tune_fused(16) | TUNE_XCD_REMAP)
        e.init_random()
        yield e, {"sweeps": 0, "rounds": 0}


def run_segment(e: Engine, done: dict, lo, ho, params=None) -> None:
    o = e.pt_run(n_sweeps=40, swap_every=10, measure_every=2, winding_every=2, winding_layer=MID)
    done["sweeps"] += 40
    done["rounds"] += 4
    assert o["energy"].shape == (20, R5) and o["winding"].shape == (20, R5)
    check_segment(e, o, lo, ho)
    replay_one_sweep(e, lo, ho, sweep_index=done["sweeps"], params=params)
    done["sweeps"] += 1
    e.mlp_enable(False)          # the zero radius of the swap replay must hold at every site
    replay_swap_rounds(e, rounds_done=done["rounds"], n_rounds=6)
    done["sweeps"] += 6
    done["rounds"] += 6


def test_c5_fixed_segment_against_oracle(c5):
    e, done = c5
    orc.set_threads(16)
    run_segment(e, done, olat(Lattice(L5, L5, L5)), orc.ham_from(HAM5))


def test_c5_learned_segment_against_oracle(c5):
    e, done = c5
    orc.set_threads(16)
    p = random_network()
    e.mlp_set_params(p)
    e.mlp_enable(True)
    run_segment(e, done, olat(Lattice(L5, L5, L5)), orc.ham_from(HAM5), params=p.astype(np.float64))


# ---- C1 -----------------------------------------------------------------

HAM1 = Hamiltonian(J=1.0, D=0.3, dmi="bulk", Bz=0.0, Kz=0.0)
BETA1 = 2.0
N1_SWEEPS = 1000
BURN1 = 200


def test_c1_first_sweeps_replayed():
    lat = Lattice(16, 16, 16)
    with Engine([Ladder(lat, HAM1, [BETA1], 0.5)], seed=1) as e:
        e.init_random()
        s0 = np.ascontiguousarray(e.get_spins(0))
        e.sweep(3)
        s1 = e.get_spins(0)
    lo, ho = olat(lat), orc.ham_from(HAM1)
    assert np.max(np.abs(s0 - orc.init_random(lo, 1, 0))) < 2e-6
    orc.sweeps(lo, ho, s0, BETA1, 0.5, 1, 0, 0, 3)
    diff = np.max(np.abs(s0 - s1), axis=0)
    assert np.sum(diff > 1e-3) <= 2 and np.median(diff) < 1e-6


def chain_means(e_series: np.ndarray, m_abs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    return e_series[BURN1:].mean(axis=0), m_abs[BURN1:].mean(axis=0)


def test_c1_observables_match_oracle_chains():
    """Eight device chains (slots of one context at T = 0.5, independent
    Philox keys) and four FP64 oracle chains of a different seed, 1000 sweeps
    each from random starts; per chain the mean over sweeps 200..1000; the
    two ensembles of chain means agree within 3 sigma (chain-to-chain spread,
    which also covers chains that settle into different helix domains)."""
    lat = Lattice(16, 16, 16)
    n_dev, n_orc = 8, 4
    with Engine([Ladder(lat, HAM1, [BETA1] * n_dev, 0.5)], seed=1) as e:
        e.init_random()
        o = e.pt_run(n_sweeps=N1_SWEEPS, swap_every=0, measure_every=1, winding_every=0)
    n = lat.sites
    ge, gm = chain_means(o["energy"] / n, np.linalg.norm(o["magnetisation"], axis=2))
    lo, ho = olat(lat), orc.ham_from(HAM1)
    orc.set_threads(8)
    oe = np.zeros((N1_SWEEPS, n_orc))
    om = np.zeros((N1_SWEEPS, n_orc))
    for c in range(n_orc):
        s = orc.init_random(lo, 555, c)
        for k in range(N1_SWEEPS):
            orc.sweeps(lo, ho, s, BETA1, 0.5, 555, c, k, 1)
            oe[k, c] = orc.total_energy(lo, ho, s) / n
            om[k, c] = np.linalg.norm(orc.magnetisation(lo, s))
    re, rm = chain_means(oe, om)
    for name, a, b in (("E", ge, re), ("|M|", gm, rm)):
        sa = a.std(ddof=1) / np.sqrt(len(a))
        sb = b.std(ddof=1) / np.sqrt(len(b))
        print(f"C1 <{name}>: device {a.mean():.6f} +- {sa:.2e}, oracle {b.mean():.6f} +- {sb:.2e}", flush=True)
        assert abs(a.mean() - b.mean()) <= 3 * np.hypot(sa, sb) + 1e-12, (name, a, b)
