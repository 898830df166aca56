"""The oracle on the hyper-kagome cell (src/unit_cells.f90:32-77, the
lattice of the reference's ISING driver, src/ISING.f90:105), and the
plaquette atom of every multi-atom cell.

* atom fractions and cell edges against the reference's own
  hyper_kagome_unit_cell / bcc_unit_cell, compiled from its sources
  (tests/golden/ref_leaf.json);
* nearest neighbours against a brute-force minimum-image search over every
  atom of the lattice with initAtomShells' first-shell rule
  (src/chainMesh.f90:495-591, threshold 0.1 a);
* the sweep's colouring: no bond joins two sites of one colour, and the bond
  graph has triangles, so three colours are the fewest;
* energies: ferromagnet; a helix along z (bulk DMI) and a cycloid along x
  (the reference's interfacial DMI), whose energies follow from the bond
  geometry alone (every bond has two components of a/4 and one of zero);
  dE = E(after) - E(before);
* the Metropolis sweep over three colours: single spin in a field follows
  the Langevin law;
* plaquettes: the reference's cell lists (addAtomToChainCell puts each new
  atom at the head, src/chainMesh.f90:663-671) and computeAdjacentAtoms
  (:956-1031) restated on atom positions pick the corners the oracle uses;
  a skyrmion on those atoms has |Q| = 1.
"""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).parent / "golden" / "ref_leaf.json"


def positions(orc, L):
    """Cell-unit positions of every site in oracle order."""
    n = orc.nsites(L)
    na = orc.atoms_per_cell(L)
    fr = [orc.atom_frac(L, a) for a in range(na)]
    p = np.zeros((n, 3))
    for ix, iy, iz, a in itertools.product(range(L.lx), range(L.ly), range(L.lz), range(na)):
        p[orc.cell_site(L, ix, iy, iz, a)] = np.array([ix, iy, iz]) + fr[a]
    return p


def nint(x):
    """Fortran nint: half away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def min_image(d, L):
    d = d.copy()
    for ax, n in enumerate((L.lx, L.ly, L.lz)):
        if L.periodic[ax]:
            d[..., ax] -= n * nint(d[..., ax] / n)
    return d


@pytest.mark.parametrize("cell,kind", [("bcc", "bcc"), ("hyperkagome", "hyperkagome")])
def test_cell_matches_reference(orc, cell, kind):
    ref = json.loads(GOLDEN.read_text())["unit_cells"][kind]
    L = orc.lattice(2, 2, 2, cell=cell)
    assert orc.atoms_per_cell(L) == len(ref["frac"])
    for a, f in enumerate(ref["frac"]):
        assert np.array_equal(orc.atom_frac(L, a), np.array(f))
    assert ref["abc"][0] == ref["abc"][1] == ref["abc"][2] and ref["angles_deg"] == [90.0, 90.0, 90.0]


@pytest.mark.parametrize("dims,per", [((3, 3, 3), (True, True, True)), ((2, 2, 2), (True, True, True)),
                                      ((3, 2, 4), (True, False, True)), ((2, 2, 2), (False, False, False))])
def test_bonds_brute_force(orc, dims, per):
    L = orc.lattice(*dims, per, cell="hyperkagome")
    p = positions(orc, L)
    thr = 0.1            # 0.1 a in cell units (src/chainMesh.f90:783)
    for i in range(len(p)):
        d = min_image(p - p[i], L)
        dist = np.linalg.norm(d, axis=1)
        dist[i] = np.inf
        order = np.argsort(dist, kind="stable")
        shell = [order[0]]
        for k in range(1, len(order)):
            if dist[order[k]] - dist[order[k - 1]] > thr or not np.isfinite(dist[order[k]]):
                break
            shell.append(order[k])
        nb, r = orc.hk_bonds(L, i)
        assert set(nb.tolist()) == set(int(j) for j in shell) and len(nb) == len(shell)
        assert len(nb) <= 4
        if all(per) and min(dims) >= 2:
            assert len(nb) == 4 and np.allclose(dist[nb], np.sqrt(2) / 4)
        for j, rv in zip(nb, r):
            assert np.allclose(rv, d[j] / np.linalg.norm(d[j]), atol=1e-15)


def test_colouring(orc):
    L = orc.lattice(3, 3, 3, cell="hyperkagome")
    n = orc.nsites(L)
    assert orc.num_colours(L) == 3
    col = np.array([orc.site_colour(L, i) for i in range(n)])
    adj = {i: set(orc.hk_bonds(L, i)[0].tolist()) for i in range(n)}
    for i in range(n):
        for j in adj[i]:
            assert col[i] != col[j]
            assert i in adj[j]                   # every bond seen from both ends
    # a triangle: two colours cannot do
    assert any(adj[i] & adj[j] for i in range(n) for j in adj[i])
    # colours are unions of atom types with four types each
    for c in range(3):
        assert len({i % 12 for i in range(n) if col[i] == c}) == 4


def test_closed_forms(orc):
    lx, ly, lz = 3, 4, 5
    L = orc.lattice(lx, ly, lz, cell="hyperkagome")
    n = orc.nsites(L)
    cells = lx * ly * lz
    p = positions(orc, L)
    J, Bz, Kz, D = 0.8, 0.3, 0.2, 0.45
    # ferromagnet along n: 2N bonds
    u = np.array([0.3, -0.4, 0.866])
    u /= np.linalg.norm(u)
    s = np.repeat(u[:, None], n, axis=1)
    e = orc.total_energy(L, orc.ham(J, D, "bulk", Bz=Bz, Kz=Kz, zeeman_total_scale=2.0), s)
    assert abs(e - (-J * 2 * n - 2.0 * Bz * u[2] * n - Kz * u[2] ** 2 * n)) < 1e-10
    # helix along z: S = (cos qz, sin qz, 0); 24 bonds per cell, 16 with
    # |dz| = 1/4 (r_z = +-1/sqrt 2, same sign as dz), 8 with dz = 0
    for m in (1, 2):
        q = 2 * np.pi * m / lz
        s = np.stack([np.cos(q * p[:, 2]), np.sin(q * p[:, 2]), np.zeros(n)])
        e = orc.total_energy(L, orc.ham(J, D, "bulk"), s)
        want = -J * cells * (16 * np.cos(q / 4) + 8) + D * cells * 16 * np.sin(q / 4) / np.sqrt(2)
        assert abs(e - want) < 1e-10, (m, e, want)
    # cycloid along x in the x-z plane, interfacial d_b = D (r_y, -r_x, 0)
    for m in (1, 2):
        q = 2 * np.pi * m / lx
        s = np.stack([np.sin(q * p[:, 0]), np.zeros(n), np.cos(q * p[:, 0])])
        e = orc.total_energy(L, orc.ham(J, D, "interfacial_ref"), s)
        want = -J * cells * (16 * np.cos(q / 4) + 8) - D * cells * 16 * np.sin(q / 4) / np.sqrt(2)
        assert abs(e - want) < 1e-10, (m, e, want)


@pytest.mark.parametrize("dmi", ["bulk", "interfacial_ref"])
@pytest.mark.parametrize("per", [(True, True, True), (True, False, False)])
def test_delta_energy_is_energy_difference(orc, dmi, per):
    L = orc.lattice(3, 2, 3, per, cell="hyperkagome")
    h = orc.ham(1.1, 0.6, dmi, Bz=0.3, Kz=0.15)
    rng = np.random.default_rng(17)
    n = orc.nsites(L)
    s = rng.normal(size=(3, n))
    s /= np.linalg.norm(s, axis=0)
    e0 = orc.total_energy(L, h, s)
    for i in rng.integers(0, n, size=40):
        sn = rng.normal(size=3)
        sn /= np.linalg.norm(sn)
        t = s.copy()
        t[:, i] = sn
        assert abs(orc.delta_energy(L, h, s, int(i), sn) - (orc.total_energy(L, h, t) - e0)) < 1e-11


def test_reference_restatement(orc):
    """calculateHeisenbergEnergy / TotalHeisenbergEnergy restated on
    hyper-kagome positions (minimum-image bond vectors, D = Dz c x r_{j->i})
    against the build form in reference mode (J = -J_ref, interfacial DMI,
    Zeeman coefficient 1 per site change, 2 in totals)."""
    L = orc.lattice(3, 2, 3, cell="hyperkagome")
    Jr, Dz, B = 0.8, 0.45, 0.3
    rng = np.random.default_rng(9)
    n = orc.nsites(L)
    s = rng.normal(size=(3, n))
    s /= np.linalg.norm(s, axis=0)
    h1 = orc.ham(-Jr, Dz, "interfacial_ref", B, 0.0, zeeman_total_scale=1.0)
    for i in rng.choice(n, 20, replace=False):
        sp = rng.normal(size=3)
        sp /= np.linalg.norm(sp)
        oe, ne = orc.ref_local_energy(L, Jr, Dz, B, s, int(i), sp)
        assert ne - oe == pytest.approx(orc.delta_energy(L, h1, s, int(i), sp), abs=1e-12)
    h2 = orc.ham(-Jr, Dz, "interfacial_ref", B, 0.0, zeeman_total_scale=2.0)
    assert orc.ref_total_energy(L, Jr, Dz, B, s) == pytest.approx(orc.total_energy(L, h2, s), abs=1e-10)


def test_langevin_every_colour(orc):
    """J = D = 0: each spin in Bz follows <S_z> = coth(beta B) - 1/(beta B),
    on each of the three colours separately."""
    L = orc.lattice(3, 3, 2, cell="hyperkagome")
    n = orc.nsites(L)
    B, beta = 1.0, 1.5
    h = orc.ham(0.0, 0.0, "none", Bz=B)
    s = orc.init_random(L, 4, 0)
    col = np.array([orc.site_colour(L, i) for i in range(n)])
    orc.sweeps(L, h, s, beta, 1.0, 4, 0, 0, 200)
    acc = np.zeros(3)
    m = 3000
    for k in range(m):
        orc.sweeps(L, h, s, beta, 1.0, 4, 0, 200 + k, 1)
        acc += [s[2, col == c].mean() for c in range(3)]
    want = 1 / np.tanh(beta * B) - 1 / (beta * B)
    # 72 spins per colour, 3000 sweeps, tau ~ 2: error of the mean ~ 0.003
    assert np.all(np.abs(acc / m - want) < 0.015), (acc / m, want)


def reference_plaquettes(orc, L, z):
    """calculate_winding_number2's corners (src/reciprocalSpaceProcesses.f90:341-345)
    from the reference's data structures restated on positions: the head of
    each cell's atom list (atoms prepended in cell order) and
    computeAdjacentAtoms' +x / +y neighbour (another cell's atom whose other
    two Bravais coefficients vanish, the nearest on the positive side)."""
    na = orc.atoms_per_cell(L)
    p = positions(orc, L)
    dims = (L.lx, L.ly, L.lz)

    def cell_of(site):
        c = site // na
        return c // (L.ly * L.lz), (c // L.lz) % L.ly, c % L.lz

    def head(ix, iy, iz):
        lst = []
        for a in range(na):
            lst.insert(0, orc.cell_site(L, ix, iy, iz, a))   # addAtomToChainCell
        return lst[0]

    def higher(site, d):
        c = cell_of(site)
        best, bd = None, np.inf
        for off in itertools.product((-1, 0, 1), repeat=3):
            nc = tuple((c[k] + off[k]) % dims[k] for k in range(3))
            if nc == c:
                continue
            for a in range(na):
                t = orc.cell_site(L, *nc, a)
                v = min_image(p[t] - p[site], L)
                if any(abs(v[k]) > 1e-4 for k in range(3) if k != d):
                    continue
                if v[d] > 0 and v[d] <= bd:
                    best, bd = t, v[d]
        return best

    out = {}
    for i, j in itertools.product(range(L.lx), range(L.ly)):
        a1 = head(i, j, z)
        a2 = higher(a1, 0)
        a3 = higher(a1, 1)
        a4 = higher(a2, 1)
        out[(i, j)] = [a1, a2, a3, a4]
    return out


@pytest.mark.parametrize("cell,dims", [("bcc", (4, 3, 3)), ("hyperkagome", (3, 4, 3)), ("sc", (4, 4, 3))])
def test_plaquette_atoms_follow_reference_lists(orc, cell, dims):
    L = orc.lattice(*dims, cell=cell)
    for z in range(dims[2]):
        ref = reference_plaquettes(orc, L, z)
        for (i, j), ids in ref.items():
            assert orc.plaquette(L, i, j, z) == ids, (cell, i, j, z)


@pytest.mark.parametrize("cell", ["bcc", "hyperkagome"])
def test_skyrmion_on_plaquette_atoms(orc, cell):
    lx = ly = 16
    L = orc.lattice(lx, ly, 2, (True, True, False), cell=cell)
    na = orc.atoms_per_cell(L)
    rng = np.random.default_rng(8)
    s = rng.normal(size=(3, orc.nsites(L)))
    s /= np.linalg.norm(s, axis=0)
    for ix, iy, iz in itertools.product(range(lx), range(ly), range(2)):
        dx, dy = ix - 7.7, iy - 8.2
        th = np.pi * np.exp(-(np.hypot(dx, dy) / 2.8) ** 2)
        ph = np.arctan2(dy, dx)
        s[:, orc.cell_site(L, ix, iy, iz, na - 1)] = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]
    for z in range(2):
        q, raw = orc.winding_layer(L, s, z)
        assert abs(q) == 1 and abs(raw - q) < 1e-9
        assert abs(orc.winding_density(L, s, z).sum() - raw) < 1e-12
        assert abs(orc.ref_winding_number(L, s, z + 1) - raw) < 1e-5
