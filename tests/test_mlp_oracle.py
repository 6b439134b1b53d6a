"""The CPU definition of the surrogate network (oracle/mlp.py) pinned by
central finite differences: the analytic float64 gradient of the loss
against the numerical one, for parameters drawn from every layer."""

import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = mlp.WIDTHS
    params = mlp.init_params(rng) * 0.5
    params += rng.normal(0, 0.01, params.size)  # non-zero biases
    x = rng.uniform(-1, 1, (16, 5))
    t = rng.normal(0, 1, (16, 3))
    w = np.array([1.0, 2.0, 0.5])
    loss, g = mlp.loss_and_grad(params, x, t, w)
    sizes = [int(np.prod(s)) for s in mlp.param_shapes(widths)]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    h = 1e-6
    for blk in range(len(sizes)):
        for idx in rng.choice(sizes[blk], size=min(6, sizes[blk]), replace=False):
            k = starts[blk] + idx
            p1, p2 = params.copy(), params.copy()
            p1[k] += h
            p2[k] -= h
            num = (mlp.loss_and_grad(p1, x, t, w)[0] - mlp.loss_and_grad(p2, x, t, w)[0]) / (2 * h)
            assert abs(num - g[k]) <= 1e-6 * max(1.0, abs(num)), (blk, idx, num, g[k])


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -3.0, 0.0])
    got = mlp.bf16_round(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0);
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to 1 + 2^-6
    assert got.tolist() == [1.0, 1.0, 1.0 + 2 ** -6, -3.0, 0.0]


def test_parameter_count():
    assert sum(int(np.prod(s)) for s in mlp.param_shapes()) == 199683


def test_polynomial_monomials_and_codes():
    """The 56 monomials of total degree <= 3 in 5 inputs: distinct, packed
    codes that decode to the exponents, and a fit on them that recovers a
    cubic exactly (normal equations, as the device accumulates them)."""
    from optika_amd import surrogate

    ex = surrogate.EXPONENTS
    assert len(ex) == 56 == len(set(ex)) and all(sum(e) <= 3 for e in ex)
    for e, c in zip(ex, surrogate._EXP_CODES):
        assert tuple((int(c) >> (2 * k)) & 3 for k in range(5)) == tuple(e)
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, (4000, 5))
    F = mlp.poly_features(u, ex)
    # independent statement of one cubic: 0.3 + u0 - 2 u1 u4 + 0.5 u2^3 - u0 u3^2
    y = 0.3 + u[:, 0] - 2 * u[:, 1] * u[:, 4] + 0.5 * u[:, 2] ** 3 - u[:, 0] * u[:, 3] ** 2
    coef = np.linalg.solve(F.T @ F, F.T @ y)
    want = np.zeros(56)
    for e, c in [((0, 0, 0, 0, 0), 0.3), ((1, 0, 0, 0, 0), 1.0), ((0, 1, 0, 0, 1), -2.0),
                 ((0, 0, 3, 0, 0), 0.5), ((1, 0, 0, 2, 0), -1.0)]:
        want[ex.index(e)] = c
    np.testing.assert_allclose(coef, want, atol=1e-10)
