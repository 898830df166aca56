"""Spin textures with known topology for the skyrmion-count tests (imported
by tests only): compact skyrmions with a smooth radial profile on a periodic
simple-cubic lattice, in the reference's site order (ix*ly + iy)*lz + iz."""
from __future__ import annotations

import numpy as np


def skyrmions(lx: int, ly: int, lz: int, centres, radius: float, helicity: float = 0.0,
              noise: float = 0.0, seed: int = 0) -> np.ndarray:
    """Spins [3, N]: +z background with one skyrmion of the given radius (in
    cells) at each centre (x, y); polar angle pi (1 - r/R) inside, so each
    skyrmion carries charge of magnitude 1, its charge density peaks at the
    centre (as a Belavin-Polyakov skyrmion's does), and the supports of
    centres more than 2R apart do not touch. Optional Gaussian noise on every
    component before renormalisation."""
    ix, iy, iz = np.meshgrid(np.arange(lx), np.arange(ly), np.arange(lz), indexing="ij")
    theta = np.zeros(ix.shape)
    phi = np.zeros(ix.shape)
    for cx, cy in centres:
        dx = (ix - cx + lx / 2) % lx - lx / 2
        dy = (iy - cy + ly / 2) % ly - ly / 2
        r = np.hypot(dx, dy)
        t = np.clip(r / radius, 0.0, 1.0)
        inside = r < radius
        theta = np.where(inside, np.pi * (1.0 - t), theta)
        phi = np.where(inside, np.arctan2(dy, dx) + helicity, phi)
    s = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    if noise > 0:
        s = s + np.random.default_rng(seed).normal(scale=noise, size=s.shape)
        s /= np.linalg.norm(s, axis=0)
    return np.ascontiguousarray(s.reshape(3, -1))
