"""Replica exchange of the oracle: detailed balance of the swap step."""
import itertools

import numpy as np


def test_two_replica_acceptance(orc):
    beta = np.array([1.0, 0.6])
    for e_a, e_b in [(-10.0, -9.0), (-9.0, -10.0), (-5.0, -3.5)]:
        acc = 0
        n = 20000
        for k in range(n):
            sot = np.array([0, 1], dtype=np.int32)
            acc += orc.pt_swap_round(beta, [e_a, e_b], sot, 3, 0, 2 * k)
        p = min(1.0, np.exp((beta[0] - beta[1]) * (e_a - e_b)))
        assert abs(acc / n - p) < 4 * np.sqrt(p * (1 - p) / n) + 1e-9


def test_permutation_chain_is_boltzmann(orc):
    """With fixed energies of 3 replicas, the swap chain over temperature
    assignments is stationary at weight exp(-sum_t beta_t E_slot(t))."""
    beta = np.array([1.2, 0.8, 0.5])
    e = np.array([-3.0, -1.5, -2.2])
    sot = np.array([0, 1, 2], dtype=np.int32)
    counts = {}
    n = 60000
    for k in range(n):
        orc.pt_swap_round(beta, e, sot, 7, 1, k)
        key = tuple(sot)
        counts[key] = counts.get(key, 0) + 1
    perms = list(itertools.permutations(range(3)))
    w = np.array([np.exp(-np.sum(beta * e[list(p)])) for p in perms])
    w /= w.sum()
    freq = np.array([counts.get(p, 0) / n for p in perms])
    assert np.all(np.abs(freq - w) < 0.02), (freq, w)


def test_swap_uniform_range(orc):
    u = np.array([orc.swap_uniform(1, 2, k, 5) for k in range(20000)])
    assert u.min() > 0 and u.max() <= 1 and abs(u.mean() - 0.5) < 0.01
