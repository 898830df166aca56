"""The learned-width network and its objective in the FP64 oracle: gradients
against central finite differences, forward against closed-form special
cases, objective against its limits."""
import numpy as np
import pytest

NP = 1793
SL = {"W1": slice(0, 672), "b1": slice(672, 704), "W2": slice(704, 1728), "b2": slice(1728, 1760),
      "w3": slice(1760, 1792), "b3": slice(1792, 1793)}


def silu(z):
    """SiLU in base 2, the network's activation: z / (1 + 2^-z)."""
    return z / (1 + np.exp2(-z))


def random_params(rng, scale=0.4):
    p = rng.normal(scale=scale, size=NP)
    p[SL["b3"]] = np.log(0.5)
    return p


def random_records(rng, n):
    rec = np.zeros((n, 32))
    nb = rng.normal(size=(n, 6, 3))
    nb /= np.linalg.norm(nb, axis=2)[:, :, None]
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


def test_zero_weights_give_bias(orc):
    p = np.zeros(NP); p[SL["b3"]] = -0.7
    assert orc.mlp_forward(p, np.ones(21), quantize=False) == -0.7


def test_base2_silu_is_rescaled_silu():
    z = np.linspace(-6, 6, 101)
    assert np.allclose(silu(z), np.log2(np.e) * (z * np.log(2)) / (1 + np.exp(-z * np.log(2))))


def test_single_path_closed_form(orc):
    rng = np.random.default_rng(0)
    p = np.zeros(NP)
    W1 = np.zeros((32, 21)); W1[3, 5] = 0.8; W1[3, 19] = -0.3
    b1 = np.zeros(32); b1[3] = 0.1
    W2 = np.zeros((32, 32)); W2[7, 3] = 1.3
    b2 = np.zeros(32); b2[7] = -0.2
    w3 = np.zeros(32); w3[7] = 0.9
    p[SL["W1"]] = W1.ravel(); p[SL["b1"]] = b1; p[SL["W2"]] = W2.ravel(); p[SL["b2"]] = b2
    p[SL["w3"]] = w3; p[SL["b3"]] = 0.25
    x = rng.normal(size=21)
    # silu(0) = 0 for the inactive units, so only one path contributes.
    z1 = 0.8 * x[5] - 0.3 * x[19] + 0.1
    z2 = 1.3 * silu(z1) - 0.2
    expect = 0.9 * silu(z2) + 0.25 + np.sum(w3 * silu(b2)) - 0.9 * silu(-0.2)
    assert orc.mlp_forward(p, x, quantize=False) == pytest.approx(expect, abs=1e-14)


def test_objective_limits(orc):
    rng = np.random.default_rng(1)
    rec = random_records(rng, 50)
    u = np.array([orc.mlp_unit_ball(3, k, 0) for k in range(50)])
    p = np.zeros(NP)
    p[SL["b3"]] = np.log(0.01)              # smallest radius: tiny moves
    l_small, _ = orc.mlp_loss_grad(p, rec, u, quantize=False)
    assert -1e-3 < l_small <= 0
    rec0 = rec.copy(); rec0[:, 27] = 0.0     # beta = 0: every move accepted
    p[SL["b3"]] = np.log(0.7)
    l0, _ = orc.mlp_loss_grad(p, rec0, u, quantize=False)
    S = rec[:, 21:24]
    t = S + 0.7 * u
    sn = t / np.linalg.norm(t, axis=1)[:, None]
    assert l0 == pytest.approx(-np.mean(np.sum((sn - S) ** 2, axis=1)), rel=1e-12)


def test_unit_ball_is_uniform(orc):
    u = np.array([orc.mlp_unit_ball(9, k, 2) for k in range(20000)])
    r = np.linalg.norm(u, axis=1)
    assert r.max() <= 1.0
    # P(|u| < 0.5) = 1/8, mean direction ~ 0
    assert abs(np.mean(r < 0.5) - 0.125) < 0.01
    assert np.all(np.abs(u.mean(axis=0)) < 0.02)


@pytest.mark.parametrize("group", list(SL))
def test_gradient_finite_differences(orc, group):
    rng = np.random.default_rng(2)
    p = random_params(rng)
    rec = random_records(rng, 64)
    u = np.array([orc.mlp_unit_ball(5, k, 1) for k in range(64)])
    _, g = orc.mlp_loss_grad(p, rec, u, quantize=False)
    idx = np.arange(NP)[SL[group]]
    idx = rng.choice(idx, size=min(8, len(idx)), replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps; pm[i] -= eps
        lp, _ = orc.mlp_loss_grad(pp, rec, u, quantize=False)
        lm, _ = orc.mlp_loss_grad(pm, rec, u, quantize=False)
        fd = (lp - lm) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=1e-8), (group, i)


def test_quantized_close_to_exact(orc):
    rng = np.random.default_rng(4)
    p = random_params(rng)
    rec = random_records(rng, 128)
    u = np.array([orc.mlp_unit_ball(5, k, 1) for k in range(128)])
    l0, g0 = orc.mlp_loss_grad(p, rec, u, quantize=False)
    l1, g1 = orc.mlp_loss_grad(p, rec, u, quantize=True)
    assert abs(l0 - l1) < 0.02 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


def test_learned_sweep_with_constant_network_is_the_fixed_sweep(orc):
    """Zero weights and b3 = ln r: every site gets radius exp(b3), so the
    learned sweep repeats orc_sweeps with that radius bit for bit."""
    lat = orc.lattice(6, 6, 4, (True, True, False))
    h = orc.ham(1.0, 0.5, "bulk", 0.2, 0.05)
    p = np.zeros(NP); p[SL["b3"]] = np.log(0.37)
    a = orc.init_random(lat, 9, 2)
    b = a.copy()
    na = orc.sweeps_learned(lat, h, a, 1.3, p, 9, 2, 5, 3)
    nb = orc.sweeps(lat, h, b, 1.3, float(np.exp(p[SL["b3"]][0])), 9, 2, 5, 3)
    assert na == nb and np.array_equal(a, b)


def test_learned_sweep_keeps_the_langevin_law(orc):
    """A paramagnet in a field under an arbitrary network: every single-spin
    stationary law is exact for a symmetric proposal, <Sz> = coth(bB) - 1/(bB).
    The network's inputs are the neighbour spins, which keep changing, so the
    radius varies from step to step; a radius that read the site's own spin
    would make the proposal asymmetric and bias <Sz>."""
    rng = np.random.default_rng(12)
    p = random_params(rng, 0.8)
    p[SL["w3"]] *= 3.0
    lat = orc.lattice(8, 8, 8)
    B, beta = 1.0, 1.5
    h = orc.ham(0.0, 0.0, "none", B, 0.0)
    s = orc.init_random(lat, 3, 0)
    orc.sweeps_learned(lat, h, s, beta, p, 3, 0, 0, 50)
    mz = []
    for k in range(1500):
        orc.sweeps_learned(lat, h, s, beta, p, 3, 0, 50 + k, 1)
        mz.append(s[2].mean())
    mz = np.array(mz)
    exact = 1.0 / np.tanh(beta * B) - 1.0 / (beta * B)
    from ising_amd.analysis import mean_error
    m, err, _ = mean_error(mz)
    assert abs(m - exact) < 4 * err + 1e-3, (m, exact, err)


def test_learned_sweep_refuses_other_cells(orc):
    lat = orc.lattice(4, 4, 4, cell="bcc")
    s = orc.init_random(lat, 1, 0)
    with pytest.raises(ValueError):
        orc.sweeps_learned(lat, orc.ham(1.0), s, 1.0, np.zeros(NP), 1, 0, 0, 1)
