"""ctypes declarations of the C ABI in include/ising_gpu.h.

The structures below mirror the header field for field; the library is
loaded from ``ising_amd/lib/libising_hip.so`` (built in-tree by
``ising_amd.build``). Loading fails loudly when the library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libising_hip.so"

ISING_OK = 0
ERRORS = {
    -1: "ISING_ERR_INVALID_ARG",
    -2: "ISING_ERR_BAD_DIMS",
    -3: "ISING_ERR_BAD_SPIN",
    -4: "ISING_ERR_NO_DEVICE",
    -5: "ISING_ERR_HIP",
    -6: "ISING_ERR_UNSUPPORTED",
    -7: "ISING_ERR_OUT_OF_MEMORY",
    -8: "ISING_ERR_STATE",
}
DMI_NONE, DMI_BULK, DMI_INTERFACIAL_REF = 0, 1, 2
CELL_SC, CELL_BCC, CELL_HYPERKAGOME = 0, 1, 2
MLP_NPARAM = 1793
MLP_RECORD = 32
TUNE_XCD_REMAP = 1
TUNE_NT_OWN = 2
TUNE_FUSED_SHIFT = 8
TUNE_SLAB_SHIFT = 16
TUNE_MAX_CANDIDATES = 16
MU0_REF = 1.256637006008532e-06
STATIC_SCRATCH_BYTES = 16 << 20


def tune_fused(planes: int) -> int:
    """Tuning flag: one fused launch per sweep, `planes` z planes per workgroup."""
    return int(planes) << TUNE_FUSED_SHIFT


def tune_slab(rows: int) -> int:
    """Tuning flag (with tune_fused): y slabs of `rows` owned rows per fused workgroup."""
    return int(rows) << TUNE_SLAB_SHIFT



class ising_lattice(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32), ("lz", C.c_int32), ("periodic", C.c_int32 * 3),
                ("cell", C.c_int32)]


class ising_hamiltonian(C.Structure):
    _fields_ = [("J", C.c_double), ("D", C.c_double), ("dmi_type", C.c_int32), ("demag", C.c_int32),
                ("Bz", C.c_double), ("Kz", C.c_double), ("zeeman_total_scale", C.c_double)]


class ising_ladder_desc(C.Structure):
    _fields_ = [("lattice", ising_lattice), ("ham", ising_hamiltonian), ("n_replicas", C.c_int32),
                ("reserved", C.c_int32), ("cone_radius", C.c_double), ("betas", C.POINTER(C.c_double))]


class ising_schedule(C.Structure):
    _fields_ = [("n_sweeps", C.c_int32), ("swap_every", C.c_int32), ("measure_every", C.c_int32),
                ("winding_every", C.c_int32), ("use_graph", C.c_int32), ("winding_layer", C.c_int32)]


class ising_series(C.Structure):
    _fields_ = [("energy", C.POINTER(C.c_double)), ("magnetisation", C.POINTER(C.c_double)),
                ("winding", C.POINTER(C.c_int64)), ("winding_raw", C.POINTER(C.c_double))]


class ising_skyrmion_params(C.Structure):
    _fields_ = [("scale_x", C.c_int32), ("scale_y", C.c_int32), ("n_particles", C.c_int32),
                ("num_thresholds", C.c_int32), ("num_sigmas", C.c_int32), ("reserved", C.c_int32),
                ("min_threshold", C.c_double), ("max_threshold", C.c_double)]


_P = C.c_void_p
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)

# name -> (restype, argtypes); exactly the functions the header declares.
SIGNATURES = {
    "ising_version": (C.c_char_p, []),
    "ising_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ising_create": (C.c_int, [C.POINTER(_P), C.POINTER(ising_ladder_desc), C.c_int, C.c_uint64, C.c_int]),
    "ising_destroy": (None, [_P]),
    "ising_last_error": (C.c_char_p, [_P]),
    "ising_num_replicas": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "ising_replica_sites": (C.c_int, [_P, C.c_int, _i64p]),
    "ising_set_hamiltonian": (C.c_int, [_P, C.c_int, C.POINTER(ising_hamiltonian)]),
    "ising_set_cone": (C.c_int, [_P, C.c_int, C.c_double]),
    "ising_set_beta": (C.c_int, [_P, C.c_int, C.c_double]),
    "ising_get_betas": (C.c_int, [_P, _dp]),
    "ising_set_spins": (C.c_int, [_P, C.c_int, _dp, C.c_int64]),
    "ising_get_spins": (C.c_int, [_P, C.c_int, _dp, C.c_int64]),
    "ising_init_random": (C.c_int, [_P, C.c_int]),
    "ising_init_ferro": (C.c_int, [_P, C.c_int, C.c_double]),
    "ising_sweep": (C.c_int, [_P, C.c_int]),
    "ising_energy": (C.c_int, [_P, _dp]),
    "ising_energy_with": (C.c_int, [_P, C.c_int, C.POINTER(ising_hamiltonian), _dp]),
    "ising_magnetisation": (C.c_int, [_P, _dp]),
    "ising_winding": (C.c_int, [_P, C.c_int, C.c_int, _i64p, _dp]),
    "ising_winding_all": (C.c_int, [_P, _i64p, _dp]),
    "ising_winding_density": (C.c_int, [_P, C.c_int, C.c_int, _dp]),
    "ising_delta_energy": (C.c_int, [_P, C.c_int, _i32p, _dp, C.c_int, _dp]),
    "ising_winding_density_interpolated": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "ising_skyrmion_distribution": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(ising_skyrmion_params), _i32p, _i32p]),
    "ising_demag_field": (C.c_int, [_P, C.c_int, _dp, C.c_int64]),
    "ising_get_demag_field": (C.c_int, [_P, C.c_int, _dp, C.c_int64]),
    "ising_set_demag_field": (C.c_int, [_P, C.c_int, _dp, C.c_int64]),
    "ising_set_demag_interval": (C.c_int, [_P, C.c_int]),
    "ising_pt_run": (C.c_int, [_P, C.POINTER(ising_schedule), C.POINTER(ising_series)]),
    "ising_pt_run_static": (C.c_int, [_P, C.POINTER(ising_schedule), C.POINTER(ising_series)]),
    "ising_temperature_map": (C.c_int, [_P, _i32p]),
    "ising_save_state": (C.c_int, [_P]),
    "ising_restore_state": (C.c_int, [_P]),
    "ising_swap_stats": (C.c_int, [_P, _i64p, _i64p]),
    "ising_accept_counts": (C.c_int, [_P, _i64p]),
    "ising_time_sweeps": (C.c_int, [_P, C.c_int, _dp]),
    "ising_time_pt_sweeps": (C.c_int, [_P, C.POINTER(ising_schedule), C.c_int, _dp]),
    "ising_time_demag": (C.c_int, [_P, C.c_int, _dp]),
    "ising_synchronize": (C.c_int, [_P]),
    "ising_bandwidth_probe": (C.c_int, [_P, C.c_int64, _dp]),
    "ising_set_tuning": (C.c_int, [_P, C.c_uint32]),
    "ising_get_tuning": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "ising_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint32), _dp]),
    "ising_tuning_candidates": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
    "ising_autotune_schedule": (C.c_int, [_P, C.POINTER(ising_schedule), _i32p, C.c_int, C.c_int, C.c_double, _dp]),
    "ising_mlp_set_params": (C.c_int, [_P, _fp, C.c_int]),
    "ising_mlp_get_params": (C.c_int, [_P, _fp, C.c_int]),
    "ising_mlp_enable": (C.c_int, [_P, C.c_int, C.c_int]),
    "ising_mlp_collect": (C.c_int, [_P, C.c_int]),
    "ising_mlp_num_records": (C.c_int, [_P, _i64p]),
    "ising_mlp_clear_records": (C.c_int, [_P]),
    "ising_mlp_get_records": (C.c_int, [_P, _fp, C.c_int64]),
    "ising_mlp_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, _dp]),
    "ising_mlp_forward": (C.c_int, [_P, _fp, C.c_int, _fp]),
    "ising_mlp_grad": (C.c_int, [_P, _fp, C.c_int, C.c_uint32, _dp, _fp]),
    "ising_mfma_probe": (C.c_int, [_P, _fp, _fp, _fp]),
}

_lib = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load libising_hip.so and attach the signatures (cached)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m ising_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
