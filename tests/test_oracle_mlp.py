"""The float64 MLP oracle against what the mathematics fixes: central
finite differences, closed forms of a zero-weight network, Adam's first
step.  CPU only."""
import numpy as np
import pytest

from oracle import mlp


def _case(seed, n=64, censored=True):
    r = np.random.default_rng(seed)
    params = r.normal(0, 0.3, mlp.NPARAM)
    xn = r.normal(0, 1, (n, mlp.NF))
    y = r.normal(0, 1, (n, mlp.NO))
    mask = np.ones((n, mlp.NO))
    if censored:
        mask[r.random((n, mlp.NO)) < 0.3] = -1.0
        mask[r.random((n, mlp.NO)) < 0.1] = 0.0
    return params, xn, y, mask


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gradient_matches_finite_differences(seed):
    params, xn, y, mask = _case(seed)
    L, g = mlp.grad(params, xn, y, mask)
    assert L == pytest.approx(mlp.loss(params, xn, y, mask), rel=1e-12)
    r = np.random.default_rng(seed + 100)
    idx = np.concatenate([r.choice(mlp.NPARAM, 200, replace=False), np.arange(mlp.NPARAM - 3, mlp.NPARAM)])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (mlp.loss(p1, xn, y, mask) - mlp.loss(p2, xn, y, mask)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_zero_network_closed_form():
    # all weights zero: scores are b3, hidden layers are relu(b) with zero
    # weight gradients through the dead ReLUs
    P = {k: np.zeros(s) for k, s in mlp.SHAPES}
    P["b3"] = np.array([1.0, -2.0, 0.5])
    params = mlp.pack(P)
    r = np.random.default_rng(5)
    xn = r.normal(0, 1, (10, mlp.NF))
    y = r.normal(0, 1, (10, mlp.NO))
    mask = np.ones((10, mlp.NO))
    out, a1, a2 = mlp.forward(params, xn)
    assert np.all(out == P["b3"]) and not a1.any() and not a2.any()
    L, g = mlp.grad(params, xn, y, mask)
    e = P["b3"] - y
    assert L == pytest.approx((e * e).sum() / 30)
    G = mlp.unpack(g)
    np.testing.assert_allclose(G["b3"], 2 * e.sum(0) / 30, rtol=1e-12)
    assert not G["W1"].any() and not G["W2"].any() and not G["W3"].any()


def test_lower_bound_labels():
    # a censored label is met when the score is at or above it: no loss,
    # no gradient; below it, it counts like an exact label
    P = {k: np.zeros(s) for k, s in mlp.SHAPES}
    P["b3"] = np.array([3.0, 3.0, 3.0])
    params = mlp.pack(P)
    xn = np.zeros((1, mlp.NF))
    mask = np.array([[-1.0, -1.0, 1.0]])
    L, g = mlp.grad(params, xn, np.array([[2.0, 5.0, 3.0]]), mask)
    # plan 0: bound 2 met; plan 1: bound 5 missed by 2; plan 2: exact hit
    assert L == pytest.approx(4.0 / 3)
    np.testing.assert_allclose(mlp.unpack(g)["b3"], [0.0, 2 * -2.0 / 3, 0.0])


def test_adam_first_step_is_sign_like():
    r = np.random.default_rng(9)
    p = r.normal(0, 1, 50)
    g = r.normal(0, 1, 50)
    p1, m, v = mlp.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    np.testing.assert_allclose(p - p1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_planted_mistake_is_caught():
    # a gradient with the ReLU mask of layer 2 dropped must fail the
    # finite-difference check
    params, xn, y, mask = _case(4, censored=False)
    P = mlp.unpack(params)
    out, a1, a2 = mlp.forward(params, xn)
    e, w = mlp.residual(out, y, mask)
    d3 = 2 * w * e / w.sum()
    d2 = d3 @ P["W3"]                      # mask (a2 > 0) dropped
    gW2 = d2.T @ a1
    h = 1e-6
    bad = 0
    for i in range(0, mlp.H2 * mlp.H1, 97):
        j = mlp.H1 * mlp.NF + mlp.H1 + i
        p1, p2 = params.copy(), params.copy()
        p1[j] += h
        p2[j] -= h
        fd = (mlp.loss(p1, xn, y, mask) - mlp.loss(p2, xn, y, mask)) / (2 * h)
        bad += abs(fd - gW2.ravel()[i]) > 1e-6 + 1e-5 * abs(fd)
    assert bad > 0
