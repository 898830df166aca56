"""Neighbour tables of the oracle against brute-force distance enumeration
on tiny lattices (periodic and open axes)."""
import itertools

import numpy as np
import pytest


@pytest.mark.parametrize("dims,per", [((4, 4, 4), (1, 1, 1)), ((6, 4, 5), (1, 1, 0)), ((4, 3, 2), (0, 0, 0)),
                                      ((8, 6, 4), (1, 0, 1))])
def test_neighbours_brute_force(orc, dims, per):
    L = orc.lattice(*dims, per)
    n = np.array(dims)
    coords = list(itertools.product(*[range(k) for k in dims]))
    for c in coords:
        i = orc.site(L, *c)
        assert i == (c[0] * dims[1] + c[1]) * dims[2] + c[2]
        brute = set()
        for c2 in coords:
            d = np.array(c2) - np.array(c)
            for a in range(3):
                if per[a]:
                    d[a] -= n[a] * np.round(d[a] / n[a])
            if np.sum(np.abs(d)) == 1:
                brute.add(orc.site(L, *c2))
        got = {orc.neighbour(L, i, a, s) for a in range(3) for s in (-1, 1)} - {-1}
        assert got == brute, c
