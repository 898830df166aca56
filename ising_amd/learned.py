"""Learned proposal: data generation, training and evaluation protocol.

The network (include/ising_gpu.h, ising_amd/csrc/mlp.hpp) maps a site's six
neighbour spins and beta*h to the log of its proposal radius through two
hidden layers of 32 units with SiLU in base 2, z / (1 + 2^-z). Protocol:

1. run the plain fixed-radius sampler and collect records (neighbours,
   beta*h, S, h, beta, K) of random sites every `collect_every` sweeps;
2. train with Adam on the negative expected squared jump distance
   -E[|S'-S|^2 min(1, exp(-beta dE))] of fresh proposals on those records;
3. enable the network in the sweep kernel; the fixed-radius sampler stays
   available as the fallback and as the yardstick (observables and tau_int).

The radius never depends on the site's own spin, so the proposal density is
symmetric and Metropolis remains exact for any parameters.
"""
from __future__ import annotations

import numpy as np

from .engine import Engine

NPARAM = 1793


def initial_params(log_r0: float = float(np.log(0.5)), seed: int = 1, scale: float = 0.05) -> np.ndarray:
    """Small random weights and output bias log r0: starts near the fixed radius r0."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    p[0:672] = rng.normal(scale=scale, size=672)                  # W1
    p[704:1728] = rng.normal(scale=scale, size=1024)             # W2
    p[1760:1792] = rng.normal(scale=scale, size=32)              # w3
    p[1792] = log_r0
    return p


def generate_data(engine: Engine, sweeps: int = 2000, collect_every: int = 20, per_replica: int = 4096,
                  swap_every: int = 10, use_graph: bool = True) -> int:
    """Plain PT sweeps with record collection; returns the number of records."""
    engine.mlp_enable(False)
    engine.mlp_clear_records()
    done = 0
    while done < sweeps:
        n = min(collect_every, sweeps - done)
        engine.pt_run(n, swap_every=swap_every if n >= swap_every else 0, measure_every=0, winding_every=0,
                      use_graph=use_graph)
        done += n
        engine.mlp_collect(per_replica)
    return engine.mlp_num_records()


def train(engine: Engine, steps: int = 1500, batch: int = 16384, lr: float = 3e-3,
          params: np.ndarray | None = None) -> np.ndarray:
    """Adam on the collected records; returns the per-step mean loss."""
    engine.mlp_set_params(initial_params() if params is None else params)
    return engine.mlp_train(steps, batch, lr)
