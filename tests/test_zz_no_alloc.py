"""ising_pt_run_static (series staged in a static device array, nothing
allocated) against ising_pt_run on the same chains: two calls in a row, the
second with a smaller schedule, each pair started from one device
checkpoint, so records, final spins, temperature maps and swap counters must
agree bit for bit, with half sweeps and with the fused sweep. Also the size
limit of the static staging."""
import numpy as np
import pytest

from ising_amd import Engine, Hamiltonian, IsingError, Ladder, Lattice, geometric_temperatures
from ising_amd import _capi

pytestmark = pytest.mark.gpu


def _state(e: Engine, swaps0) -> dict:
    """State after a run; swap counters as increments over the run (they are
    cumulative and not part of a checkpoint)."""
    att, acc = e.swap_stats()
    return {"spins": [e.get_spins(r) for r in range(e.n_replicas)], "map": e.temperature_map(),
            "swap_attempts": att - swaps0[0], "swap_accepts": acc - swaps0[1], "energy": e.energy()}


def _same(a: dict, b: dict) -> None:
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("tuning", [0, _capi.tune_fused(4)], ids=["half_sweeps", "fused4"])
def test_static_matches_allocating_entry_point(tuning):
    lat = Lattice(16, 16, 8)
    ham = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2)
    T = geometric_temperatures(0.5, 2.0, 6)
    runs = [dict(n_sweeps=60, swap_every=5, measure_every=2, winding_every=3),
            dict(n_sweeps=24, swap_every=4, measure_every=3, winding_every=4, winding_layer=3)]
    with Engine([Ladder(lat, ham, list(1.0 / T), 0.5)], seed=31) as e:
        e.set_tuning(tuning)
        e.init_random()
        e.pt_run(20, 5, 0, 0)
        for k, kw in enumerate(runs):
            e.save_state()
            sw = e.swap_stats()
            ref = e.pt_run(use_graph=True, **kw)
            st_ref = _state(e, sw)
            e.restore_state()
            sw = e.swap_stats()
            got = e.pt_run(static=True, **kw)
            st_got = _state(e, sw)
            print(f"call {k}: {kw}", flush=True)
            _same(ref, got)
            _same(st_ref, st_got)
            assert ref["energy"].shape[0] == kw["n_sweeps"] // kw["measure_every"]
            assert np.all(np.isfinite(ref["energy"]))


def test_static_rejects_schedules_beyond_its_staging():
    lat = Lattice(8, 8, 8)
    with Engine([Ladder(lat, Hamiltonian(J=1.0), [1.0] * 4, 0.5)], seed=5) as e:
        e.init_random()
        # one E/M record takes 4 doubles per replica: more records than fit
        n = _capi.STATIC_SCRATCH_BYTES // (4 * 8 * 4) + 8
        with pytest.raises(IsingError, match="static scratch"):
            e.pt_run(n, 0, 1, 0, static=True)
        out = e.pt_run(16, 4, 2, 0, static=True)     # and the context still works
        assert out["energy"].shape == (8, 4)
