"""GPU tests of the fused sweep: one launch per sweep, z planes of each colour
staged in LDS (whole planes, or y slabs that recompute two halo rows on
either side), z-chunked workgroups that recompute their halo planes, results
written to the other spin buffer.

The fused launch must give the same chains bit for bit as the two half-sweep
launches (same Philox counters, same per-site arithmetic), for every chunk
size, boundary condition and lattice mix, plain and learned, eager and in
replayed graphs with odd periods; its E/M record (per-site FP32, FP64 sums)
must match the FP64 measure kernel to 1e-6 relative, like the half-sweep
record; acceptance counters must survive layout changes.
"""
import numpy as np
import pytest

from gpu_util import engine
from ising_amd import Engine, Hamiltonian, IsingError, Ladder, Lattice, geometric_temperatures
from ising_amd._capi import TUNE_XCD_REMAP, tune_fused, tune_slab
from ising_amd.learned import initial_params

pytestmark = pytest.mark.gpu

HAM = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05)
HAM_IF = Hamiltonian(J=-1.0, D=0.7, dmi="interfacial_ref", Bz=0.3)
LATS = {
    "cube_16": Lattice(16, 16, 16),
    "open_z_12x10x6": Lattice(12, 10, 6, (True, True, False)),
    "open_z_odd_8x6x5": Lattice(8, 6, 5, (True, True, False)),
    "open_xy_10x7x8": Lattice(10, 7, 8, (False, False, True)),
    "film_32x32x4": Lattice(32, 32, 4, (True, True, False)),
    "monolayer_16x12x1": Lattice(16, 12, 1, (True, True, False)),
    "bilayer_10x8x2": Lattice(10, 8, 2, (True, True, False)),
}
BETAS = [2.0, 0.9, 0.4]


def run_sweeps(lat, ham, flags, n=5, seed=77, learned=None):
    with engine(lat, ham, BETAS, seed=seed) as e:
        e.set_tuning(flags)
        if learned is not None:
            e.mlp_set_params(learned)
            e.mlp_enable(True)
        e.init_random()
        e.sweep(n)
        spins = [e.get_spins(r) for r in range(len(BETAS))]
        return spins, e.accept_counts(), e.energy()


@pytest.mark.parametrize("lname", sorted(LATS))
@pytest.mark.parametrize("lc", [1, 3, 4, 255])
def test_fused_equals_half_sweeps(lname, lc):
    lat = LATS[lname]
    ham = HAM_IF if "open_xy" in lname else HAM
    ref = run_sweeps(lat, ham, 0)
    got = run_sweeps(lat, ham, tune_fused(lc) | TUNE_XCD_REMAP)
    for r in range(len(BETAS)):
        assert np.array_equal(ref[0][r], got[0][r]), r
    assert np.array_equal(ref[1], got[1])
    assert np.array_equal(ref[2], got[2])


def test_fused_uneven_ladders():
    lads = [Ladder(Lattice(8, 8, 8), HAM, [1.0, 0.5], 0.5),
            Ladder(Lattice(12, 12, 12), HAM, [2.0, 0.8, 0.4], 0.3),
            Ladder(Lattice(16, 16, 4, (True, True, False)), HAM_IF, [1.5], 0.5)]
    out = []
    for flags in (0, tune_fused(5)):
        with Engine(lads, seed=99) as e:
            e.set_tuning(flags)
            e.init_random()
            e.sweep(3)
            out.append(([e.get_spins(r) for r in range(e.n_replicas)], e.accept_counts()))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])


def test_fused_learned_equals_half_sweeps():
    """A fused tuning with the learned proposal on: the learned sampler keeps
    the half sweeps, so the chains equal the untuned ones."""
    rng = np.random.default_rng(4)
    p = initial_params(scale=0.3, seed=4)
    p[1760:1792] = rng.normal(scale=0.1, size=32)
    lat = Lattice(16, 16, 8)
    ref = run_sweeps(lat, HAM, 0, n=3, learned=p)
    got = run_sweeps(lat, HAM, tune_fused(3), n=3, learned=p)
    for r in range(len(BETAS)):
        assert np.array_equal(ref[0][r], got[0][r]), r
    assert np.array_equal(ref[1], got[1])


def test_switch_between_fused_and_learned_sweeps():
    """An odd count of fused sweeps leaves the state in the second spin
    buffer; learned sweeps (half sweeps) continue from it, eager and in
    replayed PT periods, and fused sweeps take over again afterwards: the
    chains equal those of half sweeps throughout."""
    p = initial_params(scale=0.3, seed=5)
    lat = Lattice(16, 16, 8)
    outs = []
    for flags in (0, tune_fused(3)):
        with engine(lat, HAM, BETAS, seed=31) as e:
            e.set_tuning(flags)
            e.mlp_set_params(p)
            e.init_random()
            e.sweep(3)
            e.mlp_enable(True)
            e.sweep(2)
            e.pt_run(n_sweeps=6, swap_every=3, measure_every=1, winding_every=0)
            e.mlp_enable(False)
            o = e.pt_run(n_sweeps=9, swap_every=3, measure_every=1, winding_every=0)
            outs.append(([e.get_spins(r) for r in range(len(BETAS))], e.accept_counts(), e.temperature_map(),
                         o["energy"]))
    for a, b in zip(outs[0][0], outs[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    assert np.allclose(outs[0][3], outs[1][3], rtol=1e-12, atol=1e-9)


def test_fused_graph_odd_period():
    """Replayed graphs of a 3-sweep period (the spin-buffer parity alternates
    between replays) against eager half sweeps: spins, swaps, Q series bit
    for bit; E/M records to FP64 summation order."""
    lat = Lattice(16, 16, 8)
    betas = 1.0 / geometric_temperatures(0.3, 2.0, 4)
    outs = []
    for flags, graph in ((0, False), (tune_fused(4) | TUNE_XCD_REMAP, True)):
        with engine(lat, HAM, betas, seed=5) as e:
            e.set_tuning(flags)
            e.init_random()
            o = e.pt_run(n_sweeps=33, swap_every=3, measure_every=1, winding_every=3, use_graph=graph)
            o["map"] = e.temperature_map()
            o["spins"] = e.get_spins(2)
            outs.append(o)
    for k in ("winding", "winding_raw", "map", "spins"):
        assert np.array_equal(outs[0][k], outs[1][k]), k
    for k in ("energy", "magnetisation"):
        assert np.allclose(outs[0][k], outs[1][k], rtol=1e-12, atol=1e-9), k


@pytest.mark.parametrize("lname", ["cube_16", "open_z_odd_8x6x5", "open_xy_10x7x8"])
def test_fused_measurement_matches_exact(lname):
    """E and M recorded by the fused launch (per-site FP32, FP64 sums) against
    the FP64 measure kernel: relative 1e-6 for E, 1e-6 absolute for M."""
    lat = LATS[lname]
    with engine(lat, HAM, [1.5, 0.6, 0.3]) as e:
        e.set_tuning(tune_fused(2))
        e.init_random()
        for _ in range(3):
            o = e.pt_run(n_sweeps=3, swap_every=0, measure_every=1, winding_every=0)
            sot = e.temperature_map()
            exact = e.energy()[sot]
            mag = e.magnetisation()[sot]
            scale = np.maximum(np.abs(exact), lat.sites)
            assert np.all(np.abs(o["energy"][-1] - exact) <= 1e-6 * scale)
            assert np.allclose(o["magnetisation"][-1], mag, atol=1e-6)


def test_accept_counts_survive_layout_changes():
    lat = Lattice(16, 16, 16)
    ref = run_sweeps(lat, HAM, 0, n=8)
    with engine(lat, HAM, BETAS, seed=77) as e:
        e.init_random()
        for flags in (tune_fused(4), tune_fused(8) | TUNE_XCD_REMAP, 0, tune_fused(16), tune_fused(3) | tune_slab(5),
                      tune_fused(3), tune_fused(2) | tune_slab(2), 0):
            e.set_tuning(flags)
            e.sweep(1)
        assert np.array_equal(e.accept_counts(), ref[1])
        assert np.array_equal(e.get_spins(1), ref[0][1])


# y slabs: (lattice, Hamiltonian, owned rows per slab (0: the tallest window
# that fits), z planes per workgroup). Windows cross periodic images and open
# faces; the last slab of a lattice owns fewer rows.
SLABS = {
    "cube_16_slab2": (Lattice(16, 16, 16), HAM, 2, 4),
    "cube_16_slab5": (Lattice(16, 16, 16), HAM, 5, 255),
    "open_y_12x21x6_slab4": (Lattice(12, 21, 6, (True, False, False)), HAM, 4, 3),
    "open_xy_10x15x8_slab3": (Lattice(10, 15, 8, (False, False, True)), HAM_IF, 3, 5),
    "film_128x128x4_auto": (Lattice(128, 128, 4, (True, True, False)), HAM, 0, 4),
    "open_y_128x37x5_auto": (Lattice(128, 37, 5, (True, False, False)), HAM, 0, 2),
    "wide_600x10x4_auto": (Lattice(600, 10, 4), HAM, 0, 4),
}


@pytest.mark.parametrize("name", sorted(SLABS))
def test_fused_slabs_equal_half_sweeps(name):
    lat, ham, hy, lc = SLABS[name]
    ref = run_sweeps(lat, ham, 0, n=4)
    got = run_sweeps(lat, ham, tune_fused(lc) | tune_slab(hy) | TUNE_XCD_REMAP, n=4)
    for r in range(len(BETAS)):
        assert np.array_equal(ref[0][r], got[0][r]), r
    assert np.array_equal(ref[1], got[1])
    assert np.array_equal(ref[2], got[2])


def test_fused_slabs_beside_whole_planes():
    """One context holding replicas whose planes fit a workgroup and a film
    whose planes are cut into slabs: one launch, one workgroup shape."""
    lads = [Ladder(Lattice(16, 16, 16), HAM, [1.0, 0.5], 0.5),
            Ladder(Lattice(128, 128, 4, (True, True, False)), HAM, [2.0, 0.7], 0.4),
            Ladder(Lattice(12, 21, 6, (True, False, False)), HAM_IF, [1.5], 0.5)]
    out = []
    for flags in (0, tune_fused(4)):
        with Engine(lads, seed=12) as e:
            e.set_tuning(flags)
            e.init_random()
            e.sweep(3)
            out.append(([e.get_spins(r) for r in range(e.n_replicas)], e.accept_counts(), e.energy()))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_fused_slabs_graph_and_records():
    """Replayed graphs of a 3-sweep period on y slabs against eager half
    sweeps (spins, swaps and Q series bit for bit), and the slab launch's E/M
    records against the FP64 measure kernel."""
    lat = Lattice(16, 20, 8, (True, False, True))
    betas = 1.0 / geometric_temperatures(0.3, 2.0, 4)
    outs = []
    for flags, graph in ((0, False), (tune_fused(4) | tune_slab(3), True)):
        with engine(lat, HAM, betas, seed=6) as e:
            e.set_tuning(flags)
            e.init_random()
            o = e.pt_run(n_sweeps=33, swap_every=3, measure_every=1, winding_every=3, use_graph=graph)
            o["map"] = e.temperature_map()
            o["spins"] = e.get_spins(1)
            sot = e.temperature_map()
            o["exact"] = e.energy()[sot]
            o["mag"] = e.magnetisation()[sot]
            outs.append(o)
    for k in ("winding", "winding_raw", "map", "spins"):
        assert np.array_equal(outs[0][k], outs[1][k]), k
    for k in ("energy", "magnetisation"):
        assert np.allclose(outs[0][k], outs[1][k], rtol=1e-12, atol=1e-9), k
    got = outs[1]
    scale = np.maximum(np.abs(got["exact"]), lat.sites)
    assert np.all(np.abs(got["energy"][-1] - got["exact"]) <= 1e-6 * scale)
    assert np.allclose(got["magnetisation"][-1], got["mag"], atol=1e-6)


def test_autotune_ranks_slab_shapes():
    """A context whose planes all exceed one workgroup: the candidates carry
    cost-ranked slab heights for each chunk size (and 24-plane chunks), every
    candidate is timed, and the chain after autotuning equals the half
    sweeps' chain over the same sweeps (tuning changes no result)."""
    lat = Lattice(80, 80, 16)                    # 40 x 80 = 3200 sites of one colour per plane
    cand = None
    with engine(lat, HAM, BETAS, seed=21) as e:
        cand = e.tuning_candidates()
        fused = [f for f in cand if (f >> 8) & 0xFF]
        assert fused and all((f >> 16) & 0xFF >= 2 for f in fused)
        assert any(((f >> 8) & 0xFF) == 16 for f in fused)
        e.init_random()
        best, ms = e.autotune(2)
        assert set(ms) == set(cand) and best in cand
        e.sweep(3)
        got = [e.get_spins(r) for r in range(len(BETAS))], e.accept_counts()
    with engine(lat, HAM, BETAS, seed=21) as e:
        e.init_random()
        e.sweep(2 * len(cand) + 3)
        ref = [e.get_spins(r) for r in range(len(BETAS))], e.accept_counts()
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    assert np.array_equal(ref[1], got[1])


def test_tuning_flags_are_checked():
    # planes beyond one workgroup run as y slabs
    with engine(Lattice(128, 128, 4, (True, True, False)), HAM, [1.0]) as e:
        e.set_tuning(tune_fused(2))
        assert any((f >> 8) & 0xFF for f in e.tuning_candidates())
        with pytest.raises(IsingError):
            e.set_tuning(tune_fused(2) | tune_slab(1))      # a slab owns at least two rows
        with pytest.raises(IsingError):
            e.set_tuning(tune_slab(4))                      # slabs only with a fused chunk
    # the limit: a window of six rows of one colour (341 column pairs a row)
    with engine(Lattice(684, 8, 4), HAM, [1.0]) as e:
        with pytest.raises(IsingError) as ei:
            e.set_tuning(tune_fused(2))
        assert "UNSUPPORTED" in str(ei.value)
        assert all((f >> 8) == 0 for f in e.tuning_candidates())
    with engine(Lattice(682, 8, 4), HAM, [1.0]) as e:
        e.set_tuning(tune_fused(2))
    with engine(Lattice(64, 64, 3, (True, True, False)), HAM, [1.0]) as e:
        e.set_tuning(tune_fused(2))
        e.set_tuning(tune_fused(2) | tune_slab(62))
    with engine(Lattice(8, 8, 8), HAM, [1.0]) as e:
        with pytest.raises(IsingError):
            e.set_tuning(1 << 24)
        cand = e.tuning_candidates()
        assert tune_fused(8) in cand and 0 in cand
