"""Writes tests/golden/ref_leaf.json from the reference's own leaf modules.

Run in a container where /root/reference is mounted, after
`make -C oracle/ref` has compiled constants.f90, rand.f90, vecNd.f90,
atom.f90 and unit_cells.f90 from /root/reference/src into
oracle/_ref/libref_leaf.so. The fixture holds only numbers (draws of
algor_uniform_random, constants::pi, vecNd cross and dot products, the
imaginary part of a kind-less cmplx(), constants::mu_0, the edge lengths,
angles and atom fractions of bcc_unit_cell and hyper_kagome_unit_cell).
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
LIB = ROOT / "oracle" / "_ref" / "libref_leaf.so"


def main():
    L = C.CDLL(str(LIB))
    L.ref_pi.restype = C.c_double
    L.ref_mu0.restype = C.c_double
    L.ref_cmplx_default_kind_imag.restype = C.c_double
    L.ref_cmplx_default_kind_imag.argtypes = [C.c_double]
    out = {"source": "reference src/constants.f90, src/rand.f90, src/vecNd.f90, src/atom.f90, src/unit_cells.f90 built by oracle/ref/Makefile"}
    streams = {}
    for seed in (0, 1, 42, 123456, -7, 2147483647):
        buf = (C.c_double * 64)()
        L.ref_uniform_stream(C.c_int32(seed), 64, buf)
        streams[str(seed)] = [float(v) for v in buf]
    out["uniform_streams"] = streams
    out["pi"] = L.ref_pi()
    out["mu0"] = L.ref_mu0()
    rng = np.random.default_rng(5)
    cd = []
    for _ in range(16):
        a, b = rng.normal(size=3), rng.normal(size=3)
        c = (C.c_double * 3)()
        d = C.c_double()
        L.ref_cross_dot(np.ascontiguousarray(a).ctypes.data_as(C.POINTER(C.c_double)),
                        np.ascontiguousarray(b).ctypes.data_as(C.POINTER(C.c_double)), c, C.byref(d))
        cd.append({"a": a.tolist(), "b": b.tolist(), "cross": list(c), "dot": d.value})
    out["cross_dot"] = cd
    out["cmplx_imag"] = [[x, L.ref_cmplx_default_kind_imag(x)] for x in (0.1, 1.0 / 3.0, -2.7182818284590452, 1e-3)]
    cells = {}
    for kind, name in ((0, "bcc"), (1, "hyperkagome")):
        abc, ang, xyz = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 48)()
        n = C.c_int()
        L.ref_unit_cell(kind, abc, ang, C.byref(n), xyz)
        cells[name] = {"abc": list(abc), "angles_deg": list(ang),
                       "frac": [list(xyz[3 * i:3 * i + 3]) for i in range(n.value)]}
    out["unit_cells"] = cells
    (Path(__file__).parent / "ref_leaf.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
