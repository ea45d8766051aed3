"""The FP64 network oracle against central finite differences and closed forms.

The network's 6 outputs are a quadrupole tensor o; a training row's target is
(u, y, w) and the loss w |B(u) o - y|^2, with B(u) o the softened quadrupole's
shape (pinned physically in test_oracle_quadrupole.py)."""
import numpy as np

from oracle import bh_oracle as O


def _params(rng):
    n = O.mlp_num_params()
    assert n == 64 * 12 + 64 + 64 * 64 + 64 + 6 * 64 + 6
    return (rng.normal(size=n) * 0.3).astype(np.float32)


def _targets(rng, rows):
    u = rng.normal(size=(rows, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u *= rng.uniform(0.5, 1.0, (rows, 1))            # |u| = |r| / s <= 1
    y = rng.normal(size=(rows, 3)) * 0.3
    w = rng.uniform(0.2, 3.0, (rows, 1))
    return np.concatenate([u, y, w], axis=1).astype(np.float32)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.normal(size=(7, 12)).astype(np.float32)
    t = _targets(rng, 7)
    loss, g = O.mlp_loss_grad(p, x, t)
    # the loss is the mean squared error of B(u) o against y
    o = O.mlp_forward(p, x)
    want = np.mean([t[b, 6] * ((O.mlp_quad_basis(t[b, :3], o[b]) - t[b, 3:6]) ** 2).sum() for b in range(7)])
    np.testing.assert_allclose(loss, want, rtol=1e-12)
    # central differences on a float64 copy of the loss (parameters are float32,
    # so perturb by steps that float32 represents exactly)
    idx = np.concatenate([rng.choice(p.size, 60, replace=False), np.arange(p.size - 6, p.size)])
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
    p[-6:] = [1.5, -2.0, 0.25, 0.5, 0.0, -1.0]
    x = np.random.default_rng(1).normal(size=(4, 12)).astype(np.float32)
    np.testing.assert_allclose(O.mlp_forward(p, x), np.tile([1.5, -2.0, 0.25, 0.5, 0.0, -1.0], (4, 1)))


def test_basis_closed_forms():
    """B(u) o for simple tensors: an isotropic tensor (o = I) gives
    (-3 - 4.5 + 7.5 |u|^2) u; a single xx component seen along x gives 3."""
    u = np.array([0.6, 0.0, 0.8])
    np.testing.assert_allclose(O.mlp_quad_basis(u, [1, 0, 0, 1, 0, 1]), (-7.5 + 7.5) * u, atol=1e-15)
    u = np.array([0.3, 0.4, 0.0])
    np.testing.assert_allclose(O.mlp_quad_basis(u, [1, 0, 0, 1, 0, 1]), (-7.5 + 7.5 * 0.25) * u, atol=1e-15)
    np.testing.assert_allclose(O.mlp_quad_basis([1, 0, 0], [1, 0, 0, 0, 0, 0]), [3, 0, 0], atol=1e-15)
    # linear in o
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=6), rng.normal(size=6)
    u = rng.normal(size=3) * 0.5
    np.testing.assert_allclose(O.mlp_quad_basis(u, a + 2 * b), O.mlp_quad_basis(u, a) + 2 * O.mlp_quad_basis(u, b),
                               rtol=1e-12, atol=1e-14)
