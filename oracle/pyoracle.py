"""ctypes wrapper of the FP64 CPU oracle (oracle/libising_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result and by bench.py's cpu_baseline leg. The product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libising_oracle.so"

DMI = {"none": 0, "bulk": 1, "interfacial_ref": 2}


class orc_lattice(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32), ("lz", C.c_int32), ("periodic", C.c_int32 * 3),
                ("cell", C.c_int32)]


CELL_SC, CELL_BCC, CELL_HYPERKAGOME = 0, 1, 2
CELLS = {"sc": CELL_SC, "bcc": CELL_BCC, "hyperkagome": CELL_HYPERKAGOME}


class orc_ham(C.Structure):
    _fields_ = [("J", C.c_double), ("D", C.c_double), ("dmi_type", C.c_int32), ("reserved", C.c_int32),
                ("Bz", C.c_double), ("Kz", C.c_double), ("zeeman_total_scale", C.c_double)]


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        u32p = C.POINTER(C.c_uint32)
        dp = C.POINTER(C.c_double)
        lp, hp = C.POINTER(orc_lattice), C.POINTER(orc_ham)
        sig = {
            "orc_philox4x32_10": (None, [u32p, u32p, u32p]),
            "orc_metropolis_random": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, u32p]),
            "orc_site": (C.c_int64, [lp, C.c_int, C.c_int, C.c_int]),
            "orc_neighbour": (C.c_int64, [lp, C.c_int64, C.c_int, C.c_int]),
            "orc_local_field": (None, [lp, hp, dp, C.c_int64, dp]),
            "orc_delta_energy": (C.c_double, [lp, hp, dp, C.c_int64, dp]),
            "orc_total_energy": (C.c_double, [lp, hp, dp]),
            "orc_magnetisation": (None, [lp, dp, dp]),
            "orc_ref_local_energy": (None, [lp, C.c_double, C.c_double, C.c_double, dp, C.c_int64, dp, dp, dp]),
            "orc_ref_total_energy": (C.c_double, [lp, C.c_double, C.c_double, C.c_double, dp]),
            "orc_ref_arc_winding": (C.c_double, [dp, dp, dp]),
            "orc_ref_winding_number": (C.c_double, [lp, dp, C.c_int]),
            "orc_winding_layer": (C.c_double, [lp, dp, C.c_int, C.POINTER(C.c_int64)]),
            "orc_winding_density": (None, [lp, dp, C.c_int, dp]),
            "orc_bcc_bonds": (C.c_int, [lp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
            "orc_atoms_per_cell": (C.c_int, [lp]),
            "orc_atom_frac": (None, [lp, C.c_int, dp]),
            "orc_cell_site": (C.c_int64, [lp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_hk_bonds": (C.c_int, [lp, C.c_int64, C.POINTER(C.c_int64), dp]),
            "orc_num_colours": (C.c_int, [lp]),
            "orc_site_colour": (C.c_int, [lp, C.c_int64]),
            "orc_plaquette": (C.c_int, [lp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
            "orc_ref_winding_density": (None, [lp, dp, C.c_int, dp]),
            "orc_propose": (C.c_int, [dp, C.c_double, u32p, dp, dp]),
            "orc_half_sweep": (C.c_int64, [lp, hp, dp, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                                           C.c_uint32, C.c_int]),
            "orc_sweeps": (C.c_int64, [lp, hp, dp, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                                       C.c_uint32, C.c_int]),
            "orc_init_random": (None, [lp, dp, C.c_uint64, C.c_uint32]),
            "orc_init_ferro": (None, [lp, dp, C.c_uint64, C.c_uint32, C.c_double]),
            "orc_swap_uniform": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_pt_swap_round": (C.c_int, [C.c_int, dp, dp, C.POINTER(C.c_int32), C.c_uint64, C.c_uint32,
                                            C.c_uint32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
            "orc_ref_demag_field": (C.c_int, [lp, dp, dp]),
            "orc_demag_deposit": (None, [lp, dp, dp]),
            "orc_demag_mesh_field": (None, [lp, dp, dp]),
            "orc_demag_gather": (None, [lp, dp, dp]),
            "orc_demag_energy": (C.c_double, [lp, dp, dp, C.c_double]),
            "orc_delta_energy_field": (C.c_double, [lp, hp, dp, C.c_int64, dp, dp, C.c_double]),
            "orc_sweeps_demag": (C.c_int64, [lp, hp, dp, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                                             C.c_uint32, C.c_int, C.c_double, C.c_int, dp]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
            "orc_mlp_forward": (C.c_double, [dp, dp, C.c_int]),
            "orc_mlp_loss_grad": (C.c_double, [dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_double, dp]),
            "orc_mlp_unit_ball": (None, [C.c_uint64, C.c_uint32, C.c_uint32, dp]),
            "orc_mlp_nparam": (C.c_int, []),
            "orc_sweeps_learned": (C.c_int64, [lp, hp, dp, C.c_double, dp, C.c_int, C.c_double, C.c_double,
                                               C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]),
            "orc_sk_kernel_table": (None, [C.c_int, C.c_int, dp]),
            "orc_sk_coarse": (None, [lp, dp, C.c_int, dp]),
            "orc_sk_interpolate": (None, [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]),
            "orc_sk_density": (None, [C.c_int, C.c_int, dp, dp]),
            "orc_sk_gauss_weights": (None, [C.c_double, dp]),
            "orc_sk_filter": (None, [C.c_int, C.c_int, dp, C.c_double, dp]),
            "orc_sk_count_table": (None, [C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                          C.POINTER(C.c_int32)]),
            "orc_sk_select": (None, [C.POINTER(C.c_int32), C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.POINTER(C.c_int32)]),
            "orc_sk_distribution": (C.c_int, [lp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp]),
        }

        class ref_rng(C.Structure):
            _fields_ = [("seed", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32)]
        L.ref_rng = ref_rng
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.orc_ref_make_random.restype = ref_rng
        L.orc_ref_make_random.argtypes = [C.c_int32]
        L.orc_ref_uniform.restype = C.c_double
        L.orc_ref_uniform.argtypes = [C.POINTER(ref_rng)]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lattice(lx, ly, lz, periodic=(True, True, True), cell: str = "sc") -> orc_lattice:
    L = orc_lattice(lx, ly, lz)
    for a in range(3):
        L.periodic[a] = 1 if periodic[a] else 0
    L.cell = CELLS[cell]
    return L


def atoms_per_cell(L) -> int:
    return lib().orc_atoms_per_cell(C.byref(L))


def nsites(L) -> int:
    return L.lx * L.ly * L.lz * atoms_per_cell(L)


def cell_site(L, ix, iy, iz, atom) -> int:
    return ((ix * L.ly + iy) * L.lz + iz) * atoms_per_cell(L) + atom


def atom_frac(L, atom) -> np.ndarray:
    f = np.zeros(3)
    lib().orc_atom_frac(C.byref(L), atom, _dp(f))
    return f


def hk_bonds(L, site) -> tuple[np.ndarray, np.ndarray]:
    """Nearest neighbours of a hyper-kagome site and the unit bond vectors."""
    nb = (C.c_int64 * 12)()
    r = np.zeros((12, 3))
    n = lib().orc_hk_bonds(C.byref(L), site, nb, _dp(r))
    return np.array(nb[:n]), r[:n].copy()


def num_colours(L) -> int:
    return lib().orc_num_colours(C.byref(L))


def site_colour(L, site) -> int:
    return lib().orc_site_colour(C.byref(L), site)


def plaquette(L, i, j, z):
    """Sites (s1, s2 = +x, s3 = +y, s4 = +x+y) of plaquette (i, j) of layer z
    (0-based), or None across an open face."""
    ids = (C.c_int64 * 4)()
    return list(ids) if lib().orc_plaquette(C.byref(L), i, j, z, ids) else None


def bcc_site(L, ix, iy, iz, atom) -> int:
    return ((ix * L.ly + iy) * L.lz + iz) * 2 + atom


def bcc_bonds(L, site) -> tuple[np.ndarray, np.ndarray]:
    """Nearest neighbours of a bcc site and the signs of the bond vectors."""
    nb = (C.c_int64 * 8)()
    sg = (C.c_int * 24)()
    n = lib().orc_bcc_bonds(C.byref(L), site, nb, sg)
    return np.array(nb[:n]), np.array(sg[:3 * n]).reshape(n, 3)


def ham(J=1.0, D=0.0, dmi="bulk", Bz=0.0, Kz=0.0, zeeman_total_scale=1.0) -> orc_ham:
    return orc_ham(J, D, DMI[dmi], 0, Bz, Kz, zeeman_total_scale)


def ham_from(h) -> orc_ham:
    """Oracle Hamiltonian from an object with the ising_amd.Hamiltonian fields."""
    return ham(h.J, h.D, h.dmi, h.Bz, h.Kz, h.zeeman_total_scale)


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def metropolis_random(seed, slot, site, sweep) -> np.ndarray:
    o = (C.c_uint32 * 4)()
    lib().orc_metropolis_random(seed, slot, site, sweep, o)
    return np.array(list(o), dtype=np.uint32)


def site(L, ix, iy, iz) -> int:
    return lib().orc_site(C.byref(L), ix, iy, iz)


def neighbour(L, s, axis, d) -> int:
    return lib().orc_neighbour(C.byref(L), s, axis, d)


def _spins(s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    assert s.ndim == 2 and s.shape[0] == 3
    return s


def local_field(L, h, s, i) -> np.ndarray:
    s = _spins(s)
    o = np.zeros(3)
    lib().orc_local_field(C.byref(L), C.byref(h), _dp(s), i, _dp(o))
    return o


def delta_energy(L, h, s, i, sn) -> float:
    s = _spins(s)
    sn = np.ascontiguousarray(sn, dtype=np.float64)
    return lib().orc_delta_energy(C.byref(L), C.byref(h), _dp(s), i, _dp(sn))


def total_energy(L, h, s) -> float:
    s = _spins(s)
    return lib().orc_total_energy(C.byref(L), C.byref(h), _dp(s))


def magnetisation(L, s) -> np.ndarray:
    s = _spins(s)
    m = np.zeros(3)
    lib().orc_magnetisation(C.byref(L), _dp(s), _dp(m))
    return m


def ref_local_energy(L, J_r, Dz, B, s, i, sp) -> tuple[float, float]:
    s = _spins(s)
    sp = np.ascontiguousarray(sp, dtype=np.float64)
    o, n = C.c_double(), C.c_double()
    lib().orc_ref_local_energy(C.byref(L), J_r, Dz, B, _dp(s), i, _dp(sp), C.byref(o), C.byref(n))
    return o.value, n.value


def ref_total_energy(L, J_r, Dz, B, s) -> float:
    s = _spins(s)
    return lib().orc_ref_total_energy(C.byref(L), J_r, Dz, B, _dp(s))


def ref_arc_winding(a, b, c) -> float:
    a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c))
    return lib().orc_ref_arc_winding(_dp(a), _dp(b), _dp(c))


def ref_winding_number(L, s, z1) -> float:
    s = _spins(s)
    return lib().orc_ref_winding_number(C.byref(L), _dp(s), z1)


def winding_density(L, s, z) -> np.ndarray:
    """Per-plaquette charge density of layer z (0-based), [ly, lx] (x fastest)."""
    s = _spins(s)
    out = np.zeros(L.lx * L.ly)
    lib().orc_winding_density(C.byref(L), _dp(s), z, _dp(out))
    return out.reshape(L.ly, L.lx)


def ref_winding_density(L, s, z1) -> np.ndarray:
    """calculate_winding_number_density restated (z1 1-based), [ly, lx]."""
    s = _spins(s)
    out = np.zeros(L.lx * L.ly)
    lib().orc_ref_winding_density(C.byref(L), _dp(s), z1, _dp(out))
    return out.reshape(L.ly, L.lx)


def winding_layer(L, s, z) -> tuple[int, float]:
    s = _spins(s)
    q = C.c_int64()
    r = lib().orc_winding_layer(C.byref(L), _dp(s), z, C.byref(q))
    return q.value, r


def propose(sold, radius, rnd) -> tuple[int, np.ndarray, float]:
    so = np.ascontiguousarray(sold, dtype=np.float64)
    sn = np.zeros(3)
    u = C.c_double()
    r = (C.c_uint32 * 4)(*[int(v) for v in rnd])
    ok = lib().orc_propose(_dp(so), radius, r, _dp(sn), C.byref(u))
    return ok, sn, u.value


def half_sweep(L, h, s, beta, radius, seed, slot, sweep, colour) -> int:
    """In place on s (shape (3,N) float64, C-contiguous)."""
    assert s.flags.c_contiguous and s.dtype == np.float64
    return lib().orc_half_sweep(C.byref(L), C.byref(h), _dp(s), beta, radius, seed, slot, sweep, colour)


def sweeps(L, h, s, beta, radius, seed, slot, sweep0, n) -> int:
    assert s.flags.c_contiguous and s.dtype == np.float64
    return lib().orc_sweeps(C.byref(L), C.byref(h), _dp(s), beta, radius, seed, slot, sweep0, n)


def init_random(L, seed, slot) -> np.ndarray:
    s = np.zeros((3, nsites(L)))
    lib().orc_init_random(C.byref(L), _dp(s), seed, slot)
    return s


def init_ferro(L, seed, slot, frac) -> np.ndarray:
    s = np.zeros((3, nsites(L)))
    lib().orc_init_ferro(C.byref(L), _dp(s), seed, slot, frac)
    return s


def swap_uniform(seed, ladder, pair, rnd) -> float:
    return lib().orc_swap_uniform(seed, ladder, pair, rnd)


def pt_swap_round(beta_of_temp, e_of_slot, slot_of_temp, seed, ladder, rnd, attempts=None, accepts=None) -> int:
    b = np.ascontiguousarray(beta_of_temp, dtype=np.float64)
    e = np.ascontiguousarray(e_of_slot, dtype=np.float64)
    assert slot_of_temp.dtype == np.int32 and slot_of_temp.flags.c_contiguous
    ap = attempts.ctypes.data_as(C.POINTER(C.c_int64)) if attempts is not None else None
    cp = accepts.ctypes.data_as(C.POINTER(C.c_int64)) if accepts is not None else None
    return lib().orc_pt_swap_round(len(b), _dp(b), _dp(e), slot_of_temp.ctypes.data_as(C.POINTER(C.c_int32)),
                                   seed, ladder, rnd, ap, cp)


def ref_rng(seed):
    return lib().orc_ref_make_random(seed)


def ref_uniform(rng) -> float:
    return lib().orc_ref_uniform(C.byref(rng))


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


def num_threads() -> int:
    return lib().orc_num_threads()


MLP_NPARAM = 1793
MLP_LOG_RMIN, MLP_LOG_RMAX = float(np.log(0.01)), float(np.log(8.0))


def mlp_forward(params, x, quantize=True) -> float:
    p = np.ascontiguousarray(params, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_mlp_forward(_dp(p), _dp(xx), 1 if quantize else 0)


def sweeps_learned(L, h, s, beta, params, seed, slot, sweep0, n, quantize=True, lo=None, hi=None) -> int:
    """n sweeps with the network's radius at every site (orc_sweeps_learned)."""
    assert s.flags.c_contiguous and s.dtype == np.float64
    p = np.ascontiguousarray(params, dtype=np.float64)
    lo = MLP_LOG_RMIN if lo is None else lo
    hi = MLP_LOG_RMAX if hi is None else hi
    r = lib().orc_sweeps_learned(C.byref(L), C.byref(h), _dp(s), beta, _dp(p), 1 if quantize else 0, lo, hi,
                                 seed, slot, sweep0, n)
    if r < 0:
        raise ValueError("orc_sweeps_learned: simple-cubic lattices only")
    return r


def mlp_unit_ball(seed, rec, step) -> np.ndarray:
    u = np.zeros(3)
    lib().orc_mlp_unit_ball(seed, rec, step, _dp(u))
    return u


def mlp_loss_grad(params, recs, uvecs, quantize=True, lo=MLP_LOG_RMIN, hi=MLP_LOG_RMAX):
    p = np.ascontiguousarray(params, dtype=np.float64)
    r = np.ascontiguousarray(recs, dtype=np.float64)
    u = np.ascontiguousarray(uvecs, dtype=np.float64)
    assert r.shape[1] == 32 and u.shape == (r.shape[0], 3)
    g = np.zeros(MLP_NPARAM)
    loss = lib().orc_mlp_loss_grad(_dp(p), _dp(r), _dp(u), r.shape[0], 1 if quantize else 0, lo, hi, _dp(g))
    return loss, g


# ---- demagnetisation ------------------------------------------------------
MU0_REF = 1.256637006008532e-06   # the reference's constants::mu_0 (single-precision literal)


def ncells(L) -> int:
    return L.lx * L.ly * L.lz


def demag_field(L, s) -> np.ndarray:
    """calculate_demagnetisation_field restated: H, shape (3, N)."""
    s = _spins(s)
    H = np.zeros_like(s)
    rc = lib().orc_ref_demag_field(C.byref(L), _dp(s), _dp(H))
    if rc:
        raise ValueError("demagnetisation needs a fully periodic lattice")
    return H


def demag_deposit(L, s) -> np.ndarray:
    """Cloud-in-cell mesh, shape (3, lx, ly, lz)."""
    s = _spins(s)
    m = np.zeros(3 * ncells(L))
    lib().orc_demag_deposit(C.byref(L), _dp(s), _dp(m))
    return m.reshape(3, L.lx, L.ly, L.lz)


def demag_mesh_field(L, mesh) -> np.ndarray:
    mesh = np.ascontiguousarray(mesh, dtype=np.float64)
    h = np.zeros(3 * ncells(L))
    lib().orc_demag_mesh_field(C.byref(L), _dp(mesh), _dp(h))
    return h.reshape(3, L.lx, L.ly, L.lz)


def demag_gather(L, hmesh) -> np.ndarray:
    hmesh = np.ascontiguousarray(hmesh, dtype=np.float64)
    H = np.zeros((3, nsites(L)))
    lib().orc_demag_gather(C.byref(L), _dp(hmesh), _dp(H))
    return H


def demag_energy(L, s, H, mu0=MU0_REF) -> float:
    s, H = _spins(s), _spins(H)
    return lib().orc_demag_energy(C.byref(L), _dp(s), _dp(H), mu0)


def delta_energy_field(L, h, s, i, sn, H, mu0=MU0_REF) -> float:
    s, H = _spins(s), _spins(H)
    sn = np.ascontiguousarray(sn, dtype=np.float64)
    return lib().orc_delta_energy_field(C.byref(L), C.byref(h), _dp(s), i, _dp(sn), _dp(H), mu0)


def sweeps_demag(L, h, s, beta, radius, seed, slot, sweep0, n, mu0=MU0_REF, interval=1) -> tuple[int, np.ndarray]:
    """Metropolis_mcs with demagnetisation, in place on s; returns (accepts, last field)."""
    assert s.flags.c_contiguous and s.dtype == np.float64
    H = np.zeros_like(s)
    acc = lib().orc_sweeps_demag(C.byref(L), C.byref(h), _dp(s), beta, radius, seed, slot, sweep0, n, mu0,
                                 interval, _dp(H))
    if acc < 0:
        raise ValueError("demagnetisation needs a fully periodic lattice")
    return acc, H


# ---- skyrmion count of a layer (skyrmion_oracle.c) ----------------------

SK_THRESHOLDS, SK_SIGMAS = 60, 20
SK_MIN_T, SK_MAX_T = 1e-5, 0.95


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def sk_coarse(L, s, z) -> np.ndarray:
    """Corner spins of layer z, [3, ly, lx]."""
    out = np.zeros((3, L.ly, L.lx))
    lib().orc_sk_coarse(C.byref(L), _dp(np.ascontiguousarray(s, dtype=np.float64)), z, _dp(out))
    return out


def sk_interpolate(S, sx, sy) -> np.ndarray:
    """Band-limited interpolation of a coarse grid [3, ny, nx] to [3, sy ny, sx nx]."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    _, ny, nx = S.shape
    F = np.zeros((3, sy * ny, sx * nx))
    lib().orc_sk_interpolate(nx, ny, sx, sy, _dp(S), _dp(F))
    return F


def sk_density(F) -> np.ndarray:
    F = np.ascontiguousarray(F, dtype=np.float64)
    d = np.zeros(F.shape[1:])
    lib().orc_sk_density(F.shape[2], F.shape[1], _dp(F), _dp(d))
    return d


def sk_filter(d, sigma) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.float64)
    f = np.zeros_like(d)
    lib().orc_sk_filter(d.shape[1], d.shape[0], _dp(d), sigma, _dp(f))
    return f


def sk_count_table(d, n_part=1, T=SK_THRESHOLDS, S=SK_SIGMAS, min_t=SK_MIN_T, max_t=SK_MAX_T) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.zeros((T, S, n_part), dtype=np.int32)
    lib().orc_sk_count_table(d.shape[1], d.shape[0], _dp(d), n_part, T, S, min_t, max_t, _ip(out))
    return out


def sk_select(table, q) -> np.ndarray:
    t = np.ascontiguousarray(table, dtype=np.int32)
    out = np.zeros(t.shape[2], dtype=np.int32)
    lib().orc_sk_select(_ip(t), t.shape[2], t.shape[0], t.shape[1], q, _ip(out))
    return out


def sk_distribution(L, s, z, sx=10, sy=10, n_part=1, T=SK_THRESHOLDS, S=SK_SIGMAS, min_t=SK_MIN_T, max_t=SK_MAX_T):
    """(counts[n_part], table[T, S, n_part], fine density [sy ly, sx lx])."""
    out = np.zeros(n_part, dtype=np.int32)
    tab = np.zeros((T, S, n_part), dtype=np.int32)
    d = np.zeros((sy * L.ly, sx * L.lx))
    r = lib().orc_sk_distribution(C.byref(L), _dp(np.ascontiguousarray(s, dtype=np.float64)), z, sx, sy, n_part,
                                  min_t, max_t, T, S, _ip(out), _ip(tab), _dp(d))
    if r < 0:
        raise ValueError("skyrmion count: x and y must be periodic")
    return out, tab, d
