"""Facts about the reference's arithmetic that the oracle's restatements rely
on, pinned by the reference's own leaf modules (tests/golden/ref_leaf.json,
written by tests/golden/make_ref_leaf_fixtures.py from oracle/_ref)."""
import json
from pathlib import Path

import numpy as np

GOLD = json.loads((Path(__file__).parent / "golden" / "ref_leaf.json").read_text())


def test_reference_pi_is_single_precision():
    # src/constants.f90:16 has no kind suffix on the literal.
    assert GOLD["pi"] == float(np.float32(np.pi))


def test_cmplx_without_kind_rounds_to_float(orc):
    for x, y in GOLD["cmplx_imag"]:
        assert y == float(np.float32(x))


def test_cross_and_dot_orientation(orc):
    for e in GOLD["cross_dot"]:
        a, b = np.array(e["a"]), np.array(e["b"])
        assert np.allclose(np.cross(a, b), e["cross"], rtol=0, atol=1e-15)
        assert abs(np.dot(a, b) - e["dot"]) < 1e-15


def test_ref_arc_winding_uses_float_imaginary_part(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.normal(size=(3, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        s1, s2, s3 = v
        re = 1 + s1 @ s2 + s2 @ s3 + s3 @ s1
        im = float(np.float32(s1 @ np.cross(s2, s3)))
        norm = np.sqrt(2 * (1 + s1 @ s2) * (1 + s2 @ s3) * (1 + s3 @ s1))
        assert abs(orc.ref_arc_winding(s1, s2, s3) - 2 * np.arctan2(im / norm, re / norm)) < 1e-14
