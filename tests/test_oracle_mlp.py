"""Pins the numpy restatement of the warm-start network (oracle/mlp.py): its analytic
gradient equals central finite differences of the loss in float64 without BF16
rounding, and BF16 rounding is round-to-nearest-even."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 0.0], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie -> even (1.0)
    assert r[2] == 1.0 + 2.0 ** -7  # above the tie -> up
    assert r[3] == -2.5 and r[4] == 0.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init_params(1).astype(np.float64)
    p[mlp.OB1:mlp.OW2] = rng.normal(0, 0.1, mlp.H)
    p[mlp.OB2:mlp.OW3] = rng.normal(0, 0.1, mlp.H)
    x = rng.normal(size=(40, mlp.IN))
    t = rng.normal(size=(40, mlp.OUT))
    loss, g = mlp.loss_and_grad(p, x, t, rounding=False)
    h = 1e-3
    idx = [mlp.OW1 + 3 * mlp.H + 5, mlp.OB1 + 7, mlp.OW2 + 100, mlp.OB2 + 3, mlp.OW3 + 10 * mlp.OUTP + 1,
           mlp.OB3 + 2]
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        lp = mlp.loss_and_grad(pp.astype(np.float32), x, t, rounding=False)[0]
        lm = mlp.loss_and_grad(pm.astype(np.float32), x, t, rounding=False)[0]
        assert abs(g[i] - (lp - lm) / (2 * h)) <= 2e-3 * max(abs(g[i]), 1e-3), i
    # padding parameters have zero gradient
    assert np.abs(g[mlp.OW1 + mlp.IN * mlp.H:mlp.OB1]).max() == 0
