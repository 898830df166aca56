"""Shared helpers of the GPU parity tests (imported by tests only)."""
from __future__ import annotations

import numpy as np

from ising_amd import Engine, Hamiltonian, Ladder, Lattice
from oracle import pyoracle as orc


def engine(lat: Lattice, ham: Hamiltonian, betas, cone=0.5, seed=20261203) -> Engine:
    return Engine([Ladder(lat, ham, list(betas), cone)], seed=seed)


def olat(lat: Lattice):
    return orc.lattice(lat.lx, lat.ly, lat.lz, lat.periodic, cell=lat.cell)


def random_unit(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.normal(size=(3, n))
    return v / np.linalg.norm(v, axis=0)


def de_tolerance(lat_o, ham, s, sites, props) -> np.ndarray:
    """1e-5 times the floating-point scale of each single-site dE:
    |S'-S| * (sum_j (|J| + 2|D|)|S_j| + |Bz|) + |Kz| |dSz| |S'z + Sz|."""
    n = s.shape[1]
    tol = np.empty(len(sites))
    for k, i in enumerate(sites):
        acc = abs(ham.Bz)
        if lat_o.cell == orc.CELL_BCC:
            nbrs = orc.bcc_bonds(lat_o, int(i))[0].tolist()
        elif lat_o.cell == orc.CELL_HYPERKAGOME:
            nbrs = orc.hk_bonds(lat_o, int(i))[0].tolist()
        else:
            nbrs = [orc.neighbour(lat_o, int(i), a, d) for a in range(3) for d in (-1, 1)]
        for j in nbrs:
            if j >= 0:
                acc += (abs(ham.J) + 2 * abs(ham.D)) * np.linalg.norm(s[:, j])
        dl = props[k] - s[:, i]
        tol[k] = 1e-5 * (np.linalg.norm(dl) * acc + abs(ham.Kz) * abs(dl[2]) * abs(props[k][2] + s[2, i]))
    return tol
