"""GPU tests of the learned cone-width network.

* the MFMA operand layout the kernels assume, against numpy;
* the device forward pass against the FP64 oracle (same bfloat16 roundings):
  median |df| < 1e-4, max < 1e-2 (a bfloat16 tie may round differently);
* the device gradient of the training objective against the oracle gradient:
  relative L2 error < 2e-2 per parameter group, loss relative error < 1e-3;
* records collected by the device hold the site's neighbours and field;
* exactness: with an arbitrary network the paramagnet still obeys the
  Langevin law, and an interacting lattice matches the fixed sampler's <E>;
* training lowers the objective, and the trained sampler decorrelates E
  faster than the fixed one at high temperature.
"""
import numpy as np
import pytest

from gpu_util import engine, olat, random_unit
from ising_amd import Engine, Hamiltonian, Ladder, Lattice, geometric_temperatures
from ising_amd.analysis import mean_error, specific_heat, specific_heat_error, tau_int
from ising_amd.learned import initial_params
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

GROUPS = {"W1": slice(0, 672), "b1": slice(672, 704), "W2": slice(704, 1728), "b2": slice(1728, 1760),
          "w3": slice(1760, 1792), "b3": slice(1792, 1793)}
HAM = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05)


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def random_params(rng, scale=0.3):
    p = rng.normal(scale=scale, size=1793).astype(np.float32)
    p[1792] = np.log(0.5)
    return p


def test_mfma_operand_layout():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(32, 16)).astype(np.float32)
    B = rng.normal(size=(16, 32)).astype(np.float32)
    with engine(Lattice(4, 4, 4), HAM, [1.0]) as e:
        D = e.mfma_probe(A, B)
    ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
    assert np.allclose(D, ref, rtol=1e-5, atol=1e-5), np.argwhere(~np.isclose(D, ref, rtol=1e-5, atol=1e-5))[:8]


def test_forward_matches_oracle():
    rng = np.random.default_rng(1)
    p = random_params(rng)
    x = rng.normal(size=(1000, 21)).astype(np.float32)
    with engine(Lattice(4, 4, 4), HAM, [1.0]) as e:
        e.mlp_set_params(p)
        f = e.mlp_forward(x)
        back = e.mlp_params()
    assert np.array_equal(back, p)
    ref = np.array([orc.mlp_forward(p.astype(np.float64), xi.astype(np.float64), quantize=True) for xi in x])
    err = np.abs(f - ref)
    assert np.median(err) < 1e-4 and err.max() < 1e-2, (np.median(err), err.max())


def random_records(rng, n):
    rec = np.zeros((n, 32), dtype=np.float32)
    nb = rng.normal(size=(n, 6, 3)); nb /= np.linalg.norm(nb, axis=2)[:, :, None]
    S = rng.normal(size=(n, 3)); S /= np.linalg.norm(S, axis=1)[:, None]
    beta = rng.uniform(0.3, 3.0, size=n)
    h = nb.sum(axis=1) + np.array([0, 0, 0.2])
    rec[:, :18] = nb.reshape(n, 18)
    rec[:, 18:21] = beta[:, None] * h
    rec[:, 21:24] = S
    rec[:, 24:27] = h
    rec[:, 27] = beta
    rec[:, 28] = 0.05
    return rec


@pytest.mark.parametrize("n", [64, 1000])
def test_gradient_matches_oracle(n):
    rng = np.random.default_rng(2)
    p = random_params(rng, 0.25)
    rec = random_records(rng, n)
    step = 7
    with engine(Lattice(4, 4, 4), HAM, [1.0], seed=4242) as e:
        e.mlp_set_params(p)
        loss, g = e.mlp_grad(rec, step)
        seed = e.seed
    u = np.array([orc.mlp_unit_ball(seed, k, step) for k in range(n)])
    lo, go = orc.mlp_loss_grad(p.astype(np.float64), rec.astype(np.float64), u, quantize=True)
    assert abs(loss - lo) < 1e-3 * abs(lo)
    for name, sl in GROUPS.items():
        rel = np.linalg.norm(g[sl] - go[sl]) / max(np.linalg.norm(go[sl]), 1e-12)
        assert rel < 2e-2, (name, rel)


def test_collected_records_describe_the_site():
    lat = Lattice(8, 8, 6)
    betas = [0.5, 1.5]
    with engine(lat, HAM, betas) as e:
        e.init_random()
        e.sweep(3)
        spins = [e.get_spins(r) for r in range(2)]
        e.mlp_collect(50)
        recs = e.mlp_records()
    assert recs.shape == (100, 32)
    lo, ho = olat(lat), orc.ham_from(HAM)
    for k, rc in enumerate(recs):
        r = k // 50
        s = spins[r]
        site = int(np.argmin(np.sum((s.T - rc[21:24]) ** 2, axis=1)))
        assert np.array_equal(s[:, site].astype(np.float32), rc[21:24])
        nbrs = [orc.neighbour(lo, site, a, d) for a in range(3) for d in (1, -1)]
        for j, nb in enumerate(nbrs):
            assert np.array_equal(s[:, nb].astype(np.float32), rc[3 * j:3 * j + 3])
        hf = orc.local_field(lo, ho, s, site)
        assert np.allclose(rc[24:27], hf, atol=2e-5)
        assert np.allclose(rc[18:21], betas[r] * hf, atol=2e-5 * max(1, betas[r]))
        assert rc[27] == np.float32(betas[r]) and rc[28] == np.float32(HAM.Kz)


def test_learned_sampler_exact_on_paramagnet():
    """Any width network leaves the Langevin law of independent spins intact."""
    lat = Lattice(32, 32, 32)
    B = 1.0
    betas = [0.5, 1.5, 4.0]
    ham = Hamiltonian(J=0.0, D=0.0, dmi="none", Bz=B)
    rng = np.random.default_rng(5)
    with engine(lat, ham, betas, cone=0.5) as e:
        e.mlp_set_params(random_params(rng, 0.6))
        e.mlp_enable(True)
        e.init_random()
        e.sweep(60)
        o = e.pt_run(n_sweeps=200, swap_every=0, measure_every=1)
    mz = o["magnetisation"][:, :, 2]
    for k, b in enumerate(betas):
        exact = 1.0 / np.tanh(b * B) - 1.0 / (b * B)
        m, err, _ = mean_error(mz[:, k])
        assert abs(m - exact) < max(4 * err, 1.5e-3), (b, m, exact, err)


def test_learned_matches_fixed_observables():
    """Interacting 12^3 lattice: <E>, <|M|>, C_v and the middle layer's <Q> of
    the learned sampler (a random network whose radius varies site by site
    around 0.4) within 3 sigma of the fixed sampler's; both start from the
    same equilibrated configuration."""
    lat = Lattice(12, 12, 12)
    betas = [2.0, 0.9]
    rng = np.random.default_rng(6)
    p = random_params(rng, 0.4)
    p[1760:1792] = rng.normal(scale=0.08, size=32)
    p[1792] = np.log(0.4)
    res = {}
    with engine(lat, HAM, betas, cone=0.5, seed=11) as e:
        e.init_random()
        e.sweep(600)
        state = [e.get_spins(r) for r in range(2)]
        e.mlp_set_params(p)
        for learned in (False, True):
            for r in range(2):
                e.set_spins(r, state[r])
            e.mlp_enable(learned)
            e.sweep(100)
            res[learned] = e.pt_run(n_sweeps=4000, swap_every=0, measure_every=2, winding_every=2,
                                    winding_layer=6)
    for k in range(2):
        a = mean_error(res[False]["energy"][:, k])
        b = mean_error(res[True]["energy"][:, k])
        assert abs(a[0] - b[0]) < 3 * np.hypot(a[1], b[1]) + 1e-9, (k, a, b)
        ma = mean_error(np.linalg.norm(res[False]["magnetisation"][:, k], axis=1))
        mb = mean_error(np.linalg.norm(res[True]["magnetisation"][:, k], axis=1))
        assert abs(ma[0] - mb[0]) < 3 * np.hypot(ma[1], mb[1]) + 1e-9, (k, ma, mb)
        qa = mean_error(res[False]["winding"][:, k].astype(float))
        qb = mean_error(res[True]["winding"][:, k].astype(float))
        assert abs(qa[0] - qb[0]) < 3 * np.hypot(qa[1], qb[1]) + 1e-9, (k, qa, qb)
        ea, eb = res[False]["energy"][:, k], res[True]["energy"][:, k]
        ca, cea = specific_heat(ea, betas[k], lat.sites), specific_heat_error(ea, betas[k], lat.sites)
        cb, ceb = specific_heat(eb, betas[k], lat.sites), specific_heat_error(eb, betas[k], lat.sites)
        assert abs(ca - cb) < 3 * np.hypot(cea, ceb), (k, ca, cea, cb, ceb)


def test_training_improves_and_decorrelates_faster():
    """Train on plain-sampler data of a small ladder; the objective falls and
    tau_int(E) at the hottest temperature drops below the fixed sampler's."""
    from ising_amd.learned import generate_data, train
    lat = Lattice(16, 16, 16)
    T = geometric_temperatures(0.3, 2.0, 4)
    betas = 1.0 / T
    with engine(lat, HAM, betas, cone=0.5, seed=21) as e:
        e.init_random()
        n = generate_data(e, sweeps=300, collect_every=30, per_replica=2048)
        assert n == 10 * 4 * 2048
        loss = train(e, steps=300, batch=8192, lr=3e-3)
        assert loss[-20:].mean() < loss[:5].mean() - 0.02, (loss[:5].mean(), loss[-20:].mean())
        state = [e.get_spins(r) for r in range(4)]
        series = {}
        for learned in (False, True):
            for r in range(4):
                e.set_spins(r, state[r])
            e.mlp_enable(learned)
            e.sweep(100)
            o = e.pt_run(n_sweeps=2000, swap_every=0, measure_every=1)
            series[learned] = o["energy"]
    hot = 3
    t_fixed = tau_int(series[False][:, hot])
    t_learned = tau_int(series[True][:, hot])
    assert t_learned < t_fixed, (t_learned, t_fixed)
