"""The C-ABI library without a GPU: it loads, exports every function the
header declares, and reports errors (bad arguments, bad dims, no device)
without touching a device."""
import ctypes as C
import re
from pathlib import Path

import pytest

from ising_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    txt = (ROOT / "include" / "ising_gpu.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ising_[a-z_0-9]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = _capi.load()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_capi.SIGNATURES), set(names) ^ set(_capi.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _capi.load(tmp_path / "nope.so")


def _desc(lx=8, ly=8, lz=8, per=(1, 1, 1), nrep=1):
    betas = (C.c_double * nrep)(*([1.0] * nrep))
    d = _capi.ising_ladder_desc()
    d.lattice = _capi.ising_lattice(lx, ly, lz)
    for a in range(3):
        d.lattice.periodic[a] = per[a]
    d.ham = _capi.ising_hamiltonian(1.0, 0.5, 1, 0, 0.2, 0.0, 1.0)
    d.n_replicas = nrep
    d.cone_radius = 0.5
    d.betas = C.cast(betas, C.POINTER(C.c_double))
    return d, betas


@pytest.mark.parametrize("dims,per", [((7, 8, 8), (1, 1, 1)), ((8, 2, 8), (1, 1, 1)), ((8, 8, 5), (1, 1, 1)),
                                      ((5, 8, 4), (0, 0, 0)), ((8, 0, 4), (0, 0, 0))])
def test_bad_dims_rejected_before_device(dims, per):
    lib = _capi.load()
    d, keep = _desc(*dims, per=per)
    h = C.c_void_p()
    rc = lib.ising_create(C.byref(h), C.byref(d), 1, 1, 0)
    assert rc == -2 and not h.value
    assert b"" != lib.ising_last_error(None)


def test_no_device_or_ok():
    lib = _capi.load()
    n = C.c_int(-1)
    assert lib.ising_device_count(C.byref(n)) == 0
    d, keep = _desc()
    h = C.c_void_p()
    rc = lib.ising_create(C.byref(h), C.byref(d), 1, 1, 0)
    if n.value == 0:
        assert rc == -4 and not h.value
        assert b"device" in lib.ising_last_error(None)
    else:
        assert rc == 0
        lib.ising_destroy(h)


def test_null_arguments():
    lib = _capi.load()
    assert lib.ising_create(None, None, 0, 0, 0) == -1
    lib.ising_destroy(None)
    assert lib.ising_sweep(None, 1) == -1
    assert lib.ising_energy(None, None) == -1
    assert lib.ising_pt_run(None, None, None) == -1
    assert lib.ising_version().startswith(b"ising_hip")
