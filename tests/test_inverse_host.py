"""Inverse-loop pieces that need no GPU: the losses and the activation maps.

PointToPointLoss and UniformActivationLoss follow the explicit objective of
exp/2025/08/27/inverse/src/30-inverse.py:100-110 (1/2 sum |u - target|^2 over
the selected points, sum |a - mean a|^2); the activation maps follow
exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:106-147.
Checked against closed forms written out here.
"""

import numpy as np
import pytest
import torch

from apple_amd import inverse


def test_point_to_point_value_and_grad():
    rng = np.random.default_rng(0)
    u = torch.as_tensor(rng.standard_normal((10, 3)))
    idx = torch.tensor([1, 4, 7])
    tgt = torch.as_tensor(rng.standard_normal((3, 3)))
    loss = inverse.PointToPointLoss(idx, tgt, weight=2.0)
    L, dLdu, _ = loss.value_and_grad(u, {})
    d = (u[idx] - tgt).numpy()
    assert float(L) == pytest.approx(0.5 * 2.0 * (d * d).sum(), rel=1e-14)
    g = np.zeros((10, 3))
    g[[1, 4, 7]] = 2.0 * d
    np.testing.assert_allclose(dLdu.numpy(), g, rtol=1e-14)


def test_uniform_activation_value_and_grad():
    rng = np.random.default_rng(1)
    a = torch.as_tensor(rng.standard_normal((8, 6)))
    sel = torch.tensor([0, 2, 3, 5])
    loss = inverse.UniformActivationLoss(sel, weight=0.5)
    L, _, dq = loss.value_and_grad(torch.zeros(4, 3, dtype=torch.float64), {"activation": a})
    an = a.numpy()[[0, 2, 3, 5]]
    d = an - an.mean(axis=0)
    assert float(L) == pytest.approx(0.5 * (d * d).sum(), rel=1e-14)
    # d/da of sum |a - mean|^2 is 2 (a - mean): the mean's own derivative sums to zero
    g = np.zeros((8, 6))
    g[[0, 2, 3, 5]] = 2 * 0.5 * d
    np.testing.assert_allclose(dq["activation"].numpy(), g, rtol=1e-12, atol=1e-14)


def test_activation_maps():
    x = torch.tensor([1.2, 0.9], dtype=torch.float64)
    a = inverse.directional_activation(x, 2 * x, 3 * x)
    np.testing.assert_allclose(a.numpy(), [[0.2, 1.4, 2.6, 0, 0, 0], [-0.1, 0.8, 1.7, 0, 0, 0]], rtol=1e-14)
    lr = torch.tensor([0.0, np.log(1.5), 5.0], dtype=torch.float64)
    m = inverse.muscle_x_activation(lr, prestrain=0.8, log_min=-2.0, log_max=2.0).numpy()
    rel = np.exp([0.0, np.log(1.5), 2.0])  # clipped at log_max
    np.testing.assert_allclose(m[:, 0], rel * 0.8 - 1.0, rtol=1e-14)
    np.testing.assert_allclose(m[:, 1], 1.0 / 0.64 - 1.0, rtol=1e-14)
    np.testing.assert_allclose(m[:, 2], 0.8 - 1.0, rtol=1e-14)
    np.testing.assert_allclose(m[:, 3:], 0.0)
