"""Pins of the oracle's demagnetising field (oracle/demag_oracle.c), the
restatement of calculate_demagnetisation_field
(src/reciprocalSpaceProcesses.f90:202-281), and of the demagnetisation terms
of the energies (src/energyMin.f90:113-116,163-166,223-226).

* closed forms of one magnetisation mode m0 cos(q.r + phi): the cloud-in-cell
  deposit to cell centres averages a mode by prod_a cos(q_a/2), the
  projection keeps its part along q, the gather averages again, so on
  simple-cubic cells H(r) = -q^(q^.m0) P^2 cos(q.r + phi) with
  P = prod_a cos(q_a/2); on bcc cells (atom 1 at the cell centre, weight 1)
  H = -q^(q^.m0)(1 + P) w cos(q.r + phi), w = 1 on atom 1 and P on atom 0;
* transform and projection against numpy.fft (rfft/fft/ifft/irfft), an
  independent implementation, on random meshes with Nyquist planes;
* the deposit on axes of one and two cells, where the reference's 27-cell
  list repeats cells, against a direct enumeration, also on hyper-kagome
  cells, whose atoms at -3/4 of an edge keep only the weight of the 27 cells
  around their own cell; the gather is the transpose of the deposit;
* mu_0 from the reference's own constants.f90 (tests/golden/ref_leaf.json).
"""
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).parent / "golden" / "ref_leaf.json"


def positions(orc, L):
    na = 2 if L.cell == orc.CELL_BCC else 1
    idx = np.arange(orc.nsites(L))
    cell, atom = idx // na, idx % na
    k = cell % L.lz
    j = (cell // L.lz) % L.ly
    i = cell // (L.lz * L.ly)
    return np.stack([i, j, k]).astype(float) + 0.5 * atom, atom


def mode(orc, L, n, m0, phi):
    q = 2 * np.pi * np.asarray(n, float) / np.array([L.lx, L.ly, L.lz], float)
    r, atom = positions(orc, L)
    s = np.asarray(m0, float)[:, None] * np.cos(q @ r + phi)[None, :]
    return np.ascontiguousarray(s), q, r, atom


def expected_mode_field(orc, L, q, r, atom, m0, phi):
    qh = q / np.linalg.norm(q)
    P = np.prod(np.cos(q / 2))
    if L.cell == orc.CELL_BCC:
        amp = (1 + P) * np.where(atom == 1, 1.0, P)
    else:
        amp = P * P
    return -qh[:, None] * (qh @ np.asarray(m0, float)) * (amp * np.cos(q @ r + phi))[None, :]


def test_mu0_is_reference_constant(orc):
    ref = json.loads(GOLDEN.read_text())
    assert orc.MU0_REF == ref["mu0"]
    assert orc.MU0_REF == float(np.float32(1.25663706127e-6))


CASES = [
    ("sc", (8, 6, 5), (1, 2, -1)),
    ("sc", (6, 6, 4), (-2, 1, 1)),
    ("sc", (5, 7, 3), (2, -3, 1)),
    ("bcc", (6, 4, 5), (2, -1, 1)),
    ("bcc", (4, 5, 3), (1, 2, -1)),
]


@pytest.mark.parametrize("cell,dims,n", CASES)
def test_single_mode_closed_form(orc, cell, dims, n):
    L = orc.lattice(*dims, cell=cell)
    qn = np.asarray(n, float) / np.asarray(dims, float)
    q = 2 * np.pi * qn
    qh = q / np.linalg.norm(q)
    perp = np.cross(qh, [0.3, -0.8, 0.5])
    perp /= np.linalg.norm(perp)
    for m0, nonzero in ((qh, True), (perp, False), (np.array([0.4, -0.7, 0.2]), True)):
        s, q, r, atom = mode(orc, L, n, m0, 0.7)
        H = orc.demag_field(L, s)
        want = expected_mode_field(orc, L, q, r, atom, m0, 0.7)
        assert np.max(np.abs(H - want)) < 1e-12, (cell, dims, n, m0)
        assert (np.max(np.abs(want)) > 1e-6) == nonzero


def test_uniform_magnetisation_has_no_field(orc):
    for cell, dims in (("sc", (6, 4, 5)), ("bcc", (4, 4, 3))):
        L = orc.lattice(*dims, cell=cell)
        s = np.ascontiguousarray(np.tile(np.array([[0.2], [-0.5], [0.84]]), (1, orc.nsites(L))))
        assert np.max(np.abs(orc.demag_field(L, s))) < 1e-13


def test_phase_sign_convention(orc):
    """A forward and inverse transform with the same sign would reflect the
    mode (r -> -r); with phi != 0 that changes the field."""
    L = orc.lattice(6, 5, 4)
    m0 = np.array([1.0, 0.5, -0.2])
    s, q, r, atom = mode(orc, L, (1, 1, 1), m0, 1.1)
    H = orc.demag_field(L, s)
    reflected = expected_mode_field(orc, L, -q, r, atom, m0, 1.1)
    assert np.max(np.abs(H - reflected)) > 1e-2
    assert np.max(np.abs(H - expected_mode_field(orc, L, q, r, atom, m0, 1.1))) < 1e-12


def numpy_mesh_field(mesh):
    """Steps 2-4 of calculate_demagnetisation_field with numpy.fft: forward
    r2c (x halved), projection with the reference's folding (n/2 stays
    positive), complex-to-real inverse scaled by 1/(lx ly lz)."""
    _, lx, ly, lz = mesh.shape
    F = np.fft.fft(np.fft.fft(np.fft.rfft(mesh, axis=1), axis=2), axis=3)

    def fold(n):
        w = np.arange(n)
        return np.where(w > n // 2, w - n, w)

    kx = 2 * np.pi * np.arange(lx // 2 + 1) / lx
    ky = 2 * np.pi * fold(ly) / ly
    kz = 2 * np.pi * fold(lz) / lz
    K = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    k2[0, 0, 0] = 1.0
    kdm = K[0] * F[0] + K[1] * F[1] + K[2] * F[2]
    G = np.stack([-(kdm / k2) * K[c] for c in range(3)])
    G[:, 0, 0, 0] = 0.0
    return np.fft.irfft(np.fft.ifft(np.fft.ifft(G, axis=3), axis=2), n=lx, axis=1)


@pytest.mark.parametrize("dims", [(6, 4, 5), (5, 6, 4), (8, 8, 2), (4, 3, 6), (7, 5, 3)])
def test_transform_and_projection_match_numpy(orc, dims):
    L = orc.lattice(*dims)
    rng = np.random.default_rng(sum(dims))
    mesh = rng.normal(size=(3,) + dims)
    got = orc.demag_mesh_field(L, mesh)
    want = numpy_mesh_field(mesh)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))


def wrap(i, n):
    return 0 if i == n else (n - 1 if i == -1 else i)


def brute_deposit(orc, L, s):
    """interpolate_to_fft_array as written: 27 listed cells (repeats kept);
    atom positions are the cell corner plus the atom's fraction (pinned
    against the reference's unit cells in test_oracle_hyperkagome.py)."""
    na = orc.atoms_per_cell(L)
    fr = [orc.atom_frac(L, a) for a in range(na)]
    dims = (L.lx, L.ly, L.lz)
    mesh = np.zeros((3,) + dims)
    for i in range(L.lx):
        for j in range(L.ly):
            for k in range(L.lz):
                for di in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for dk in (-1, 0, 1):
                            nc = (wrap(i + di, L.lx), wrap(j + dj, L.ly), wrap(k + dk, L.lz))
                            for a in range(na):
                                w = 1.0
                                for ax in range(3):
                                    d = (nc[ax] + fr[a][ax]) - ((i, j, k)[ax] + 0.5)
                                    n = dims[ax]
                                    q = d / n
                                    d -= n * (np.floor(q + 0.5) if q >= 0 else -np.floor(-q + 0.5))
                                    w *= max(1.0 - abs(d), 0.0)
                                site = ((nc[0] * L.ly + nc[1]) * L.lz + nc[2]) * na + a
                                mesh[:, i, j, k] += w * s[:, site]
    return mesh


@pytest.mark.parametrize("cell,dims", [("sc", (2, 3, 1)), ("sc", (1, 2, 4)), ("bcc", (2, 2, 3)),
                                       ("bcc", (1, 3, 2)), ("sc", (4, 3, 5)), ("hyperkagome", (3, 2, 4)),
                                       ("hyperkagome", (1, 3, 2))])
def test_deposit_small_axes_brute_force(orc, cell, dims):
    L = orc.lattice(*dims, cell=cell)
    rng = np.random.default_rng(3)
    s = rng.normal(size=(3, orc.nsites(L)))
    got = orc.demag_deposit(L, s)
    want = brute_deposit(orc, L, s)
    assert np.max(np.abs(got - want)) < 1e-13


@pytest.mark.parametrize("cell,dims", [("sc", (2, 3, 1)), ("bcc", (2, 2, 3)), ("sc", (5, 4, 3)), ("bcc", (3, 4, 4)),
                                       ("hyperkagome", (3, 4, 2))])
def test_gather_is_transpose_of_deposit(orc, cell, dims):
    L = orc.lattice(*dims, cell=cell)
    rng = np.random.default_rng(11)
    s = rng.normal(size=(3, orc.nsites(L)))
    m = rng.normal(size=(3,) + dims)
    lhs = np.sum(orc.demag_deposit(L, s) * m)
    rhs = np.sum(s * orc.demag_gather(L, m))
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_weights_sum_to_one_per_atom(orc):
    for cell, dims in (("sc", (4, 5, 3)), ("bcc", (3, 4, 5))):
        L = orc.lattice(*dims, cell=cell)
        ones = np.ones((3, orc.nsites(L)))
        assert np.allclose(orc.demag_deposit(L, ones).sum(axis=(1, 2, 3)), orc.nsites(L), rtol=0, atol=1e-10)


def test_field_operator_symmetric_on_odd_lattices(orc):
    """With no Nyquist planes the field is H = A s with A symmetric, so the
    demagnetisation energy -mu0/2 s.A s is a quadratic form."""
    L = orc.lattice(5, 3, 7, cell="bcc")
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=(2, 3, orc.nsites(L)))
    assert abs(np.sum(a * orc.demag_field(L, b)) - np.sum(b * orc.demag_field(L, a))) < 1e-12


def test_open_axes_rejected(orc):
    L = orc.lattice(4, 4, 4, periodic=(True, True, False))
    with pytest.raises(ValueError):
        orc.demag_field(L, np.zeros((3, 64)))


def test_energy_and_delta_energy_terms(orc):
    L = orc.lattice(4, 4, 3)
    rng = np.random.default_rng(4)
    s = rng.normal(size=(3, 48))
    s /= np.linalg.norm(s, axis=0)
    H = orc.demag_field(L, s)
    mu0 = 0.37
    assert abs(orc.demag_energy(L, s, H, mu0) - (-0.5 * mu0 * np.sum(s * H))) < 1e-12
    h = orc.ham(1.0, 0.4, "bulk", Bz=0.2)
    sn = np.array([0.6, 0.0, 0.8])
    for i in (0, 17, 47):
        base = orc.delta_energy(L, h, s, i, sn)
        got = orc.delta_energy_field(L, h, s, i, sn, H, mu0)
        assert abs(got - (base - mu0 * np.dot(sn - s[:, i], H[:, i]))) < 1e-13


def test_sweeps_demag_schedule(orc):
    L = orc.lattice(4, 4, 4)
    h = orc.ham(1.0, 0.3, "bulk", Bz=0.1)
    s0 = orc.init_random(L, 5, 0)
    a = s0.copy()
    orc.sweeps(L, h, a, 1.5, 0.5, 5, 0, 0, 4)
    b = s0.copy()
    orc.sweeps_demag(L, h, b, 1.5, 0.5, 5, 0, 0, 4, mu0=0.0, interval=2)
    assert np.array_equal(a, b)
    c = s0.copy()
    _, H = orc.sweeps_demag(L, h, c, 1.5, 0.5, 5, 0, 0, 6, mu0=2.0, interval=3)
    assert np.max(np.abs(H - orc.demag_field(L, c))) < 1e-15
    assert not np.array_equal(c[:, :], a)  # a strong field changes decisions
