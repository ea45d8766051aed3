"""The FP64 network oracle against central finite differences and closed forms."""
import numpy as np

from oracle import bh_oracle as O


def _params(rng):
    n = O.mlp_num_params()
    assert n == 64 * 12 + 64 + 64 * 64 + 64 + 3 * 64 + 3
    return (rng.normal(size=n) * 0.3).astype(np.float32)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.normal(size=(7, 12)).astype(np.float32)
    t = rng.normal(size=(7, 3)).astype(np.float32)
    loss, g = O.mlp_loss_grad(p, x, t)
    # loss is the mean squared error of the forward pass
    y = O.mlp_forward(p, x)
    np.testing.assert_allclose(loss, ((y - t) ** 2).sum(1).mean(), rtol=1e-12)
    # central differences on a float64 copy of the loss (parameters are float32,
    # so perturb by steps that float32 represents exactly)
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        h = np.float32(2.0 ** -10)
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        lp, _ = O.mlp_loss_grad(pp, x, t)
        lm, _ = O.mlp_loss_grad(pm, x, t)
        fd = (lp - lm) / (2 * float(h))
        assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_zero_network_outputs_bias():
    n = O.mlp_num_params()
    p = np.zeros(n, np.float32)
    p[-3:] = [1.5, -2.0, 0.25]
    x = np.random.default_rng(1).normal(size=(4, 12)).astype(np.float32)
    np.testing.assert_allclose(O.mlp_forward(p, x), np.tile([1.5, -2.0, 0.25], (4, 1)))
