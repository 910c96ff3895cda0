"""CPU checks of the autograd front end's pose algebra (atac.autograd): the torch
se3_exp against the oracle's (Utils.h:126-159 restated in FP64), and the matrix
gradient G = E^-T M(g) pairing with every tangent direction E xi^ as g . xi."""
import numpy as np
import torch

from atac.autograd import se3_exp, tangent_pairing_matrix
from oracle import oracle as O


def _hat(xi):
    w, v = xi[:3], xi[3:]
    m = np.zeros((4, 4))
    m[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    m[:3, 3] = v
    return m


def test_se3_exp_matches_oracle():
    rng = np.random.default_rng(3)
    xi = rng.normal(0, 0.7, (64, 6))
    xi[:4, :3] = 0.0                      # pure translations
    got = se3_exp(torch.tensor(xi)).numpy()
    for k in range(len(xi)):
        assert np.allclose(got[k], O.se3_exp(xi[k]), rtol=0, atol=1e-13)


def test_se3_exp_small_angle_series():
    """|w| ~ 1e-10: exp is I + xi^ + xi^2 / 2 to within O(|xi|^3) (the closed form
    the oracle restates loses ~1e-11 to cancellation there)."""
    rng = np.random.default_rng(5)
    xi = rng.normal(0, 1.0, (16, 6))
    xi[:, :3] *= 1e-10
    got = se3_exp(torch.tensor(xi)).numpy()
    for k in range(len(xi)):
        h = _hat(xi[k])
        # rotation part to second order; translation V v with V = I + W / 2 + W^2 / 6
        want = np.eye(4) + h + h @ h / 2.0
        W = h[:3, :3]
        want[:3, 3] = (np.eye(3) + W / 2.0 + W @ W / 6.0) @ xi[k, 3:]
        assert np.allclose(got[k], want, rtol=0, atol=1e-15)


def test_tangent_pairing():
    rng = np.random.default_rng(4)
    n = 32
    E = se3_exp(torch.tensor(rng.normal(0, 1.0, (n, 6)))).numpy()
    g = rng.normal(size=(n, 6))
    G = tangent_pairing_matrix(torch.tensor(E), torch.tensor(g)).numpy()
    for k in range(n):
        for _ in range(4):
            xi = rng.normal(size=6)
            dE = E[k] @ _hat(xi)
            assert abs(np.sum(G[k] * dE) - g[k] @ xi) < 1e-10 * (1 + abs(g[k] @ xi))
