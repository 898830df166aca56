"""Python host mirror of the C ABI (include/ising_gpu.h).

The classes keep the reference's vocabulary (ladders of temperatures,
replicas, sweeps, swaps) and map one to one onto the C entry points:

    Engine.sweep            -> ising_sweep           (Metropolis_mcs, src/energyMin.f90:234-374)
    Engine.energy           -> ising_energy          (TotalHeisenbergEnergy, src/energyMin.f90:499-528)
    Engine.delta_energy     -> ising_delta_energy    (calculateHeisenbergEnergy, src/energyMin.f90:86-171)
    Engine.winding          -> ising_winding         (calculate_winding_number2, src/reciprocalSpaceProcesses.f90:320-364)
    Engine.magnetisation    -> ising_magnetisation   (src/PT-Utils.f90:59-65)
    Engine.pt_run           -> ising_pt_run          (PT loop, src/PT.f90:219-273)
    Engine.demag_field      -> ising_demag_field     (calculate_demagnetisation_field,
                                                      src/reciprocalSpaceProcesses.f90:202-281)

Spin arrays use the reference's atomSpins(numAtoms,3) column-major layout,
i.e. numpy arrays of shape (3, N), sites in the order (ix*Ly + iy)*Lz + iz.
Every error from the library raises IsingError; nothing falls back to CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _capi
from ._capi import DMI_BULK, DMI_INTERFACIAL_REF, DMI_NONE, TUNE_MAX_CANDIDATES

_DMI = {"none": DMI_NONE, "bulk": DMI_BULK, "interfacial_ref": DMI_INTERFACIAL_REF}


class IsingError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_capi.ERRORS.get(code, code)}: {msg}")
        self.code = code


@dataclass
class Lattice:
    lx: int
    ly: int
    lz: int
    periodic: tuple = (True, True, True)
    cell: str = "sc"          # "sc", "bcc" (two atoms per cell) or "hyperkagome" (twelve)

    @property
    def atoms_per_cell(self) -> int:
        return _ATOMS[self.cell]

    @property
    def sites(self) -> int:
        return self.lx * self.ly * self.lz * self.atoms_per_cell

    def _c(self) -> _capi.ising_lattice:
        if self.cell not in _ATOMS:
            raise ValueError("cell must be 'sc', 'bcc' or 'hyperkagome'")
        lat = _capi.ising_lattice(self.lx, self.ly, self.lz)
        for a in range(3):
            lat.periodic[a] = 1 if self.periodic[a] else 0
        lat.cell = _CELLS[self.cell]
        return lat


_ATOMS = {"sc": 1, "bcc": 2, "hyperkagome": 12}
_CELLS = {"sc": _capi.CELL_SC, "bcc": _capi.CELL_BCC, "hyperkagome": _capi.CELL_HYPERKAGOME}


@dataclass
class Hamiltonian:
    """E = -J sum S_i.S_j + sum d_a.(S_i x S_{i+a}) - s*Bz sum S_z - Kz sum S_z^2."""
    J: float = 1.0
    D: float = 0.0
    dmi: str = "bulk"
    Bz: float = 0.0
    Kz: float = 0.0
    zeeman_total_scale: float = 1.0
    demag: bool = False       # the reference's demagnetisation term (mu0 = MU0_REF)

    @classmethod
    def reference(cls, J: float, Dz: float, B: float, total_energy_convention: bool = True,
                  demag: bool = False) -> "Hamiltonian":
        """The reference's parameters: exchange +J S.S (src/energyMin.f90:139),
        interfacial DMI Dz (c x r) (:114,137), Zeeman -B S_z in single-site
        changes (:161) and -2 B S_z in total energies (:222), optional
        demagnetisation term -mu0 S.H (:113-116,163-166)."""
        return cls(J=-J, D=Dz, dmi="interfacial_ref", Bz=B, Kz=0.0,
                   zeeman_total_scale=2.0 if total_energy_convention else 1.0, demag=demag)

    def _c(self) -> _capi.ising_hamiltonian:
        if self.dmi not in _DMI:
            raise ValueError(f"dmi must be one of {sorted(_DMI)}")
        return _capi.ising_hamiltonian(J=self.J, D=self.D, dmi_type=_DMI[self.dmi], demag=1 if self.demag else 0,
                                       Bz=self.Bz, Kz=self.Kz, zeeman_total_scale=self.zeeman_total_scale)


@dataclass
class Ladder:
    lattice: Lattice
    ham: Hamiltonian
    betas: Sequence[float]
    cone_radius: float = 0.5


def geometric_temperatures(t_min: float, t_max: float, n: int) -> np.ndarray:
    """T_j = Tmin (Tmax/Tmin)^(j/(n-1)), the ladder of src/PT.f90:202."""
    if n == 1:
        return np.array([t_min], dtype=np.float64)
    j = np.arange(n, dtype=np.float64)
    return t_min * (t_max / t_min) ** (j / (n - 1))


class Engine:
    """A context of one or more PT ladders resident on one GPU."""

    def __init__(self, ladders: Sequence[Ladder] | Ladder, seed: int = 20261203, device: int = 0,
                 lib_path=None):
        if isinstance(ladders, Ladder):
            ladders = [ladders]
        self._lib = _capi.load(lib_path)
        self.ladders = list(ladders)
        descs = (_capi.ising_ladder_desc * len(self.ladders))()
        self._beta_bufs = []
        for k, L in enumerate(self.ladders):
            b = np.ascontiguousarray(L.betas, dtype=np.float64)
            self._beta_bufs.append(b)
            descs[k].lattice = L.lattice._c()
            descs[k].ham = L.ham._c()
            descs[k].n_replicas = len(b)
            descs[k].cone_radius = float(L.cone_radius)
            descs[k].betas = b.ctypes.data_as(C.POINTER(C.c_double))
        h = C.c_void_p()
        rc = self._lib.ising_create(C.byref(h), descs, len(self.ladders), C.c_uint64(seed), device)
        if rc != 0:
            raise IsingError(rc, self._lib.ising_last_error(None).decode())
        self._h = h
        self.seed = seed
        n = C.c_int()
        self._check(self._lib.ising_num_replicas(h, C.byref(n)))
        self.n_replicas = n.value
        self.replica_lattice = []
        self.replica_ladder = []
        for li, L in enumerate(self.ladders):
            for _ in range(len(L.betas)):
                self.replica_lattice.append(L.lattice)
                self.replica_ladder.append(li)
        self.zmax = max(L.lattice.lz for L in self.ladders)

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != 0:
            raise IsingError(rc, self._lib.ising_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ising_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters ---------------------------------------------------------
    def set_hamiltonian(self, ladder: int, ham: Hamiltonian) -> None:
        hc = ham._c()
        self._check(self._lib.ising_set_hamiltonian(self._h, ladder, C.byref(hc)))
        self.ladders[ladder].ham = ham

    def set_cone(self, ladder: int, radius: float) -> None:
        self._check(self._lib.ising_set_cone(self._h, ladder, float(radius)))
        self.ladders[ladder].cone_radius = radius

    def set_beta(self, replica: int, beta: float) -> None:
        self._check(self._lib.ising_set_beta(self._h, replica, float(beta)))

    def betas(self) -> np.ndarray:
        out = np.zeros(self.n_replicas)
        self._check(self._lib.ising_get_betas(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- spins --------------------------------------------------------------
    def sites(self, replica: int) -> int:
        return self.replica_lattice[replica].sites

    def set_spins(self, replica: int, spins: np.ndarray) -> None:
        s = np.ascontiguousarray(spins, dtype=np.float64)
        n = self.sites(replica)
        if s.shape != (3, n):
            raise ValueError(f"spins must have shape (3, {n})")
        self._check(self._lib.ising_set_spins(self._h, replica, s.ctypes.data_as(C.POINTER(C.c_double)), n))

    def get_spins(self, replica: int) -> np.ndarray:
        n = self.sites(replica)
        s = np.empty((3, n), dtype=np.float64)
        self._check(self._lib.ising_get_spins(self._h, replica, s.ctypes.data_as(C.POINTER(C.c_double)), n))
        return s

    def init_random(self, replica: int = -1) -> None:
        self._check(self._lib.ising_init_random(self._h, replica))

    def init_ferro(self, replica: int = -1, flip_fraction: float = 0.0) -> None:
        self._check(self._lib.ising_init_ferro(self._h, replica, float(flip_fraction)))

    # -- hot path -----------------------------------------------------------
    def sweep(self, n_sweeps: int = 1) -> None:
        self._check(self._lib.ising_sweep(self._h, int(n_sweeps)))

    def energy(self) -> np.ndarray:
        e = np.empty(self.n_replicas)
        self._check(self._lib.ising_energy(self._h, e.ctypes.data_as(C.POINTER(C.c_double))))
        return e

    def energy_with(self, replica: int, ham: Hamiltonian) -> float:
        e = C.c_double()
        hc = ham._c()
        self._check(self._lib.ising_energy_with(self._h, replica, C.byref(hc), C.byref(e)))
        return e.value

    def magnetisation(self) -> np.ndarray:
        m = np.empty((self.n_replicas, 3))
        self._check(self._lib.ising_magnetisation(self._h, m.ctypes.data_as(C.POINTER(C.c_double))))
        return m

    def winding(self, replica: int, z: int) -> tuple[int, float]:
        q = C.c_int64()
        qr = C.c_double()
        self._check(self._lib.ising_winding(self._h, replica, z, C.byref(q), C.byref(qr)))
        return q.value, qr.value

    def winding_all(self) -> tuple[np.ndarray, np.ndarray]:
        q = np.empty((self.n_replicas, self.zmax), dtype=np.int64)
        qr = np.empty((self.n_replicas, self.zmax))
        self._check(self._lib.ising_winding_all(self._h, q.ctypes.data_as(C.POINTER(C.c_int64)),
                                                qr.ctypes.data_as(C.POINTER(C.c_double))))
        return q, qr

    def winding_density(self, replica: int, z: int) -> np.ndarray:
        """Per-plaquette charge density of layer z, [ly, lx] (x fastest)."""
        lat = self.replica_lattice[replica]
        d = np.empty(lat.lx * lat.ly)
        self._check(self._lib.ising_winding_density(self._h, replica, z, d.ctypes.data_as(C.POINTER(C.c_double))))
        return d.reshape(lat.ly, lat.lx)

    def winding_density_interpolated(self, replica: int, z: int, scale_x: int = 10, scale_y: int = 10) -> np.ndarray:
        """Charge density of layer z on the interpolated grid, [sy*ly, sx*lx]."""
        lat = self.replica_lattice[replica]
        d = np.empty(lat.lx * scale_x * lat.ly * scale_y)
        self._check(self._lib.ising_winding_density_interpolated(self._h, replica, z, scale_x, scale_y,
                                                                 d.ctypes.data_as(C.POINTER(C.c_double))))
        return d.reshape(lat.ly * scale_y, lat.lx * scale_x)

    def skyrmion_distribution(self, replica: int, z: int, n_particles: int = 1, scale: tuple[int, int] = (10, 10),
                              num_thresholds: int = 60, num_sigmas: int = 20, min_threshold: float = 1e-5,
                              max_threshold: float = 0.95) -> tuple[np.ndarray, np.ndarray]:
        """(counts[n_particles], table[num_thresholds, num_sigmas, n_particles]):
        the skyrmion count of layer z (compute_skyrmion_distribution) and every
        (threshold, width) entry it searched."""
        prm = _capi.ising_skyrmion_params(scale[0], scale[1], n_particles, num_thresholds, num_sigmas, 0,
                                          min_threshold, max_threshold)
        counts = np.zeros(n_particles, dtype=np.int32)
        table = np.zeros((num_thresholds, num_sigmas, n_particles), dtype=np.int32)
        self._check(self._lib.ising_skyrmion_distribution(self._h, replica, z, C.byref(prm),
                                                          counts.ctypes.data_as(C.POINTER(C.c_int32)),
                                                          table.ctypes.data_as(C.POINTER(C.c_int32))))
        return counts, table

    def delta_energy(self, replica: int, sites: np.ndarray, props: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(sites, dtype=np.int32)
        p = np.ascontiguousarray(props, dtype=np.float64).reshape(-1, 3)
        if len(s) != len(p):
            raise ValueError("sites and proposals differ in length")
        out = np.empty(len(s))
        self._check(self._lib.ising_delta_energy(self._h, replica, s.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 p.ctypes.data_as(C.POINTER(C.c_double)), len(s),
                                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- demagnetisation -----------------------------------------------------
    def demag_field(self, replica: int) -> np.ndarray:
        """Field H of the current spins, shape (3, N) (the reference's
        demagnetisation_array, without mu0)."""
        n = self.sites(replica)
        h = np.empty((3, n), dtype=np.float64)
        self._check(self._lib.ising_demag_field(self._h, replica, h.ctypes.data_as(C.POINTER(C.c_double)), n))
        return h

    def stored_demag_field(self, replica: int) -> np.ndarray:
        """The field the replica's sweeps use now (last computed or set)."""
        n = self.sites(replica)
        h = np.empty((3, n), dtype=np.float64)
        self._check(self._lib.ising_get_demag_field(self._h, replica, h.ctypes.data_as(C.POINTER(C.c_double)), n))
        return h

    def set_demag_field(self, replica: int, h: np.ndarray) -> None:
        n = self.sites(replica)
        a = np.ascontiguousarray(h, dtype=np.float64)
        if a.shape != (3, n):
            raise ValueError(f"field must have shape (3, {n})")
        self._check(self._lib.ising_set_demag_field(self._h, replica, a.ctypes.data_as(C.POINTER(C.c_double)), n))

    def set_demag_interval(self, sweeps: int) -> None:
        self._check(self._lib.ising_set_demag_interval(self._h, int(sweeps)))

    # -- parallel tempering -------------------------------------------------
    def pt_run(self, n_sweeps: int, swap_every: int = 10, measure_every: int = 1, winding_every: int = 0,
               use_graph: bool = True, want: Sequence[str] = ("energy", "magnetisation", "winding"),
               winding_layer: int | None = None, static: bool = False) -> dict:
        """Run the PT schedule on device; returns series indexed by temperature
        position: energy [n_meas, R], magnetisation [n_meas, R, 3],
        winding / winding_raw [n_q, R, Zmax], or [n_q, R] with winding_layer = z
        (only that layer is computed). static=True calls ising_pt_run_static
        (no device allocation, eager periods)."""
        R = self.n_replicas
        Z = (self.zmax,) if winding_layer is None else ()
        n_meas = n_sweeps // measure_every if measure_every else 0
        n_q = n_sweeps // winding_every if winding_every else 0
        out = {}
        ser = _capi.ising_series()
        if n_meas and "energy" in want:
            out["energy"] = np.zeros((n_meas, R))
            ser.energy = out["energy"].ctypes.data_as(C.POINTER(C.c_double))
        if n_meas and "magnetisation" in want:
            out["magnetisation"] = np.zeros((n_meas, R, 3))
            ser.magnetisation = out["magnetisation"].ctypes.data_as(C.POINTER(C.c_double))
        if n_q and "winding" in want:
            out["winding"] = np.zeros((n_q, R) + Z, dtype=np.int64)
            out["winding_raw"] = np.zeros((n_q, R) + Z)
            ser.winding = out["winding"].ctypes.data_as(C.POINTER(C.c_int64))
            ser.winding_raw = out["winding_raw"].ctypes.data_as(C.POINTER(C.c_double))
        sch = _capi.ising_schedule(n_sweeps, swap_every, measure_every, winding_every, 1 if use_graph else 0,
                                   0 if winding_layer is None else int(winding_layer) + 1)
        fn = self._lib.ising_pt_run_static if static else self._lib.ising_pt_run
        self._check(fn(self._h, C.byref(sch), C.byref(ser)))
        return out

    def save_state(self) -> None:
        """Device-resident checkpoint: spins, temperature map, betas, counters."""
        self._check(self._lib.ising_save_state(self._h))

    def restore_state(self) -> None:
        self._check(self._lib.ising_restore_state(self._h))

    def temperature_map(self) -> np.ndarray:
        m = np.empty(self.n_replicas, dtype=np.int32)
        self._check(self._lib.ising_temperature_map(self._h, m.ctypes.data_as(C.POINTER(C.c_int32))))
        return m

    def swap_stats(self) -> tuple[np.ndarray, np.ndarray]:
        a = np.zeros(self.n_replicas, dtype=np.int64)
        b = np.zeros(self.n_replicas, dtype=np.int64)
        self._check(self._lib.ising_swap_stats(self._h, a.ctypes.data_as(C.POINTER(C.c_int64)),
                                               b.ctypes.data_as(C.POINTER(C.c_int64))))
        return a, b

    def time_sweeps(self, n_sweeps: int) -> float:
        """Average device time (ms) of one half-sweep launch over n_sweeps sweeps."""
        ms = C.c_double()
        self._check(self._lib.ising_time_sweeps(self._h, int(n_sweeps), C.byref(ms)))
        return ms.value

    def time_pt_sweeps(self, n_sweeps: int, swap_every: int = 10, measure_every: int = 2, winding_every: int = 0,
                       winding_layer: int | None = None) -> float:
        """Mean device time (ms) per half sweep of the plain sweep launches of
        n_sweeps sweeps of a PT schedule, timed inside that schedule."""
        ms = C.c_double()
        sch = _capi.ising_schedule(int(n_sweeps), swap_every, measure_every, winding_every, 0,
                                   0 if winding_layer is None else int(winding_layer) + 1)
        self._check(self._lib.ising_time_pt_sweeps(self._h, C.byref(sch), int(n_sweeps), C.byref(ms)))
        return ms.value

    def time_demag(self, n_fields: int) -> float:
        """Mean device time (ms) of one recomputation of the demagnetising
        field of every replica whose ladder has the term."""
        ms = C.c_double()
        self._check(self._lib.ising_time_demag(self._h, int(n_fields), C.byref(ms)))
        return ms.value

    def bandwidth_probe(self, nbytes: int = 256 << 20) -> float:
        """Device time (ms) of one plain copy of nbytes (read nbytes + write nbytes)."""
        ms = C.c_double()
        self._check(self._lib.ising_bandwidth_probe(self._h, int(nbytes), C.byref(ms)))
        return ms.value

    def set_tuning(self, flags: int) -> None:
        self._check(self._lib.ising_set_tuning(self._h, int(flags)))

    def tuning(self) -> int:
        f = C.c_uint32()
        self._check(self._lib.ising_get_tuning(self._h, C.byref(f)))
        return f.value

    def tuning_candidates(self) -> list[int]:
        """Flag sets that autotune() times, in order."""
        buf = (C.c_uint32 * TUNE_MAX_CANDIDATES)()
        n = C.c_int()
        self._check(self._lib.ising_tuning_candidates(self._h, buf, C.byref(n)))
        return [int(buf[k]) for k in range(n.value)]

    def autotune(self, n_sweeps: int = 20) -> tuple[int, dict[int, float]]:
        """Times each sweep-tuning flag set over n_sweeps real sweeps and keeps
        the fastest; returns (flags, {flags: ms per half sweep})."""
        cand = self.tuning_candidates()
        best = C.c_uint32()
        ms = np.zeros(TUNE_MAX_CANDIDATES)
        self._check(self._lib.ising_autotune(self._h, int(n_sweeps), C.byref(best),
                                             ms.ctypes.data_as(C.POINTER(C.c_double))))
        return best.value, {f: float(t) for f, t in zip(cand, ms)}

    def autotune_schedule(self, candidates: Sequence[int] = (1, 2, 5, 10, 20, 50), n_sweeps: int = 20,
                          max_overhead: float = 0.05, measure_every: int = 0, winding_every: int = 0,
                          winding_layer: int | None = None) -> tuple[int, dict[int, float]]:
        """Sweeps-per-swap interval: the most frequent exchange whose time per
        sweep is within (1 + max_overhead) of the fastest candidate's;
        returns (interval, {candidate: ms per sweep})."""
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(cand))
        sch = _capi.ising_schedule(0, 0, measure_every, winding_every, 0,
                                   0 if winding_layer is None else int(winding_layer) + 1)
        self._check(self._lib.ising_autotune_schedule(self._h, C.byref(sch), cand.ctypes.data_as(C.POINTER(C.c_int32)),
                                                      len(cand), int(n_sweeps), float(max_overhead),
                                                      ms.ctypes.data_as(C.POINTER(C.c_double))))
        return sch.swap_every, {int(c): float(t) for c, t in zip(cand, ms)}

    # -- learned proposal ---------------------------------------------------
    @staticmethod
    def _f32(a) -> np.ndarray:
        return np.ascontiguousarray(a, dtype=np.float32)

    def mlp_set_params(self, params) -> None:
        p = self._f32(params)
        self._check(self._lib.ising_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))

    def mlp_params(self) -> np.ndarray:
        p = np.zeros(_capi.MLP_NPARAM, dtype=np.float32)
        self._check(self._lib.ising_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))
        return p

    def mlp_enable(self, on: bool = True, ladder: int = -1) -> None:
        self._check(self._lib.ising_mlp_enable(self._h, int(ladder), 1 if on else 0))

    def mlp_collect(self, per_replica: int) -> None:
        self._check(self._lib.ising_mlp_collect(self._h, int(per_replica)))

    def mlp_num_records(self) -> int:
        n = C.c_int64()
        self._check(self._lib.ising_mlp_num_records(self._h, C.byref(n)))
        return n.value

    def mlp_clear_records(self) -> None:
        self._check(self._lib.ising_mlp_clear_records(self._h))

    def mlp_records(self, n: int | None = None) -> np.ndarray:
        n = self.mlp_num_records() if n is None else n
        out = np.zeros((n, _capi.MLP_RECORD), dtype=np.float32)
        self._check(self._lib.ising_mlp_get_records(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), n))
        return out

    def mlp_train(self, steps: int, batch: int = 16384, lr: float = 3e-3) -> np.ndarray:
        loss = np.zeros(steps)
        self._check(self._lib.ising_mlp_train(self._h, int(steps), int(batch), float(lr),
                                              loss.ctypes.data_as(C.POINTER(C.c_double))))
        return loss

    def mlp_forward(self, x) -> np.ndarray:
        xx = self._f32(x).reshape(-1, 21)
        f = np.zeros(len(xx), dtype=np.float32)
        self._check(self._lib.ising_mlp_forward(self._h, xx.ctypes.data_as(C.POINTER(C.c_float)), len(xx),
                                                f.ctypes.data_as(C.POINTER(C.c_float))))
        return f

    def mlp_grad(self, records, step: int = 0) -> tuple[float, np.ndarray]:
        r = self._f32(records).reshape(-1, _capi.MLP_RECORD)
        g = np.zeros(_capi.MLP_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._check(self._lib.ising_mlp_grad(self._h, r.ctypes.data_as(C.POINTER(C.c_float)), len(r), int(step),
                                             C.byref(loss), g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def mfma_probe(self, A, B) -> np.ndarray:
        a, b = self._f32(A).reshape(32, 16), self._f32(B).reshape(16, 32)
        d = np.zeros((32, 32), dtype=np.float32)
        self._check(self._lib.ising_mfma_probe(self._h, a.ctypes.data_as(C.POINTER(C.c_float)),
                                               b.ctypes.data_as(C.POINTER(C.c_float)),
                                               d.ctypes.data_as(C.POINTER(C.c_float))))
        return d

    def synchronize(self) -> None:
        self._check(self._lib.ising_synchronize(self._h))

    def accept_counts(self) -> np.ndarray:
        a = np.zeros(self.n_replicas, dtype=np.int64)
        self._check(self._lib.ising_accept_counts(self._h, a.ctypes.data_as(C.POINTER(C.c_int64))))
        return a
