"""Oracle RNGs: Philox4x32-10 against the Random123 known answers; the
reference generator (src/rand.f90) against draws of the reference's own
rand.f90 compiled into oracle/_ref (committed as tests/golden/ref_leaf.json)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).parent / "golden"


def test_philox_known_answers(orc):
    kat = json.loads((GOLD / "philox_kat.json").read_text())
    for v in kat["vectors"]:
        got = orc.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert [int(x) for x in got] == [int(x, 16) for x in v["out"]]


def test_metropolis_counter_layout(orc):
    seed, slot, site, sweep = 0x1234567890ABCDEF, 5, 77, 9
    a = orc.metropolis_random(seed, slot, site, sweep)
    b = orc.philox([site, sweep, 0, seed >> 32], [seed & 0xFFFFFFFF, slot])
    assert np.array_equal(a, b)


def test_reference_generator_matches_reference_build(orc):
    gold = json.loads((GOLD / "ref_leaf.json").read_text())
    for seed, draws in gold["uniform_streams"].items():
        r = orc.ref_rng(int(seed))
        got = [orc.ref_uniform(r) for _ in range(len(draws))]
        assert got == draws, seed


def test_reference_generator_range(orc):
    r = orc.ref_rng(3)
    u = np.array([orc.ref_uniform(r) for _ in range(20000)])
    assert u.min() > 0 and u.max() <= 1
    assert abs(u.mean() - 0.5) < 0.01


@pytest.mark.skipif(not (Path(__file__).parents[1] / "oracle" / "_ref" / "libref_leaf.so").exists(),
                    reason="oracle/_ref is built only where the reference sources are mounted")
def test_reference_build_live(orc):
    L = C.CDLL(str(Path(__file__).parents[1] / "oracle" / "_ref" / "libref_leaf.so"))
    buf = (C.c_double * 100)()
    L.ref_uniform_stream(C.c_int32(99), 100, buf)
    r = orc.ref_rng(99)
    assert [orc.ref_uniform(r) for _ in range(100)] == list(buf)
