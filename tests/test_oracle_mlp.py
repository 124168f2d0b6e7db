"""Pins of the CPU reference of the learned model (oracle/mlp.py).

The network has no counterpart in the reference crate; its reference is the
mathematics: bf16 round-to-nearest-even (checked against torch's bf16
conversion), the gradient of the float64 network (checked against finite
differences and torch autograd), and the occupancy features (checked against
a direct count of points per grid cell).
"""
import numpy as np
import pytest
import torch

from oracle import mlp
from oracle import oracle as orc


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, np.inf, -np.inf, 3.4e38]])
    x = x.astype(np.float32)
    ours = mlp.bf16_bits(x)
    theirs = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, theirs)
    # ties go to even: 1 + 2^-8 -> 1.0, 1 + 3*2^-8 -> 1 + 2^-6
    assert mlp.bf16_value(mlp.bf16_bits(np.float32([1.0 + 2.0 ** -8])))[0] == 1.0
    assert mlp.bf16_value(mlp.bf16_bits(np.float32([1.0 + 3 * 2.0 ** -8])))[0] == 1.0 + 2.0 ** -6
    assert np.isnan(mlp.bf16_value(mlp.bf16_bits(np.float32([np.nan])))[0])


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, mlp.IN), np.float32)
    x[:, :mlp.FEATURES] = np.log1p(rng.poisson(5.0, (n, mlp.FEATURES)))
    t = rng.normal(0, 1, (n, 2))
    return mlp.bf16_bits(x), t


def test_gradient_matches_finite_differences():
    p = mlp.init_params(1).astype(np.float64)
    x, t = _batch(64, 2)
    loss, g = mlp.gradient(p, x, t, rounded=False)
    assert loss == pytest.approx(mlp.loss_only(p, x, t), rel=1e-14)
    rng = np.random.default_rng(3)
    for _ in range(5):  # directional derivatives
        v = rng.normal(0, 1, p.size)
        h = 1e-6
        fd = (mlp.loss_only(p + h * v, x, t) - mlp.loss_only(p - h * v, x, t)) / (2 * h)
        assert fd == pytest.approx(g @ v, rel=1e-6, abs=1e-10)
    # single coordinates in every block (weights and biases of all layers)
    offs = [0, mlp.N_W1, mlp.N_W1 + mlp.N_W2, mlp.WEIGHTS, mlp.WEIGHTS + mlp.H, mlp.WEIGHTS + 2 * mlp.H]
    for o in offs:
        for k in range(o, min(o + 40, p.size), 7):
            e = np.zeros(p.size)
            e[k] = 1.0
            h = 1e-6
            fd = (mlp.loss_only(p + h * e, x, t) - mlp.loss_only(p - h * e, x, t)) / (2 * h)
            assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9)


def test_gradient_matches_torch_autograd():
    p = mlp.init_params(4).astype(np.float64)
    x, t = _batch(96, 5)
    loss, g = mlp.gradient(p, x, t, rounded=False)
    W1, W2, W3, b1, b2, b3 = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in mlp.unpack(p))
    X = torch.tensor(mlp.bf16_value(x))
    y = torch.relu(torch.relu(X @ W1 + b1) @ W2 + b2) @ W3 + b3
    L = 0.5 * ((y[:, :2] - torch.tensor(t)) ** 2).sum() / len(t)
    L.backward()
    gt = np.concatenate([a.grad.numpy().ravel() for a in (W1, W2, W3, b1, b2, b3)])
    assert float(L.detach()) == pytest.approx(loss, rel=1e-13)
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-14)


def test_rounded_network_stays_close_to_float():
    p = mlp.init_params(6)
    x, t = _batch(256, 7)
    y0 = mlp.forward(p, x, rounded=False)
    y1 = mlp.forward(p, x, rounded=True)
    assert np.max(np.abs(y1 - y0)) < 0.05 * (1 + np.max(np.abs(y0)))
    _, g0 = mlp.gradient(p, x, t, rounded=False)
    _, g1 = mlp.gradient(p, x, t, rounded=True)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.99


def test_features_count_points_per_grid_cell():
    rng = np.random.default_rng(8)
    xyz = rng.random((3000, 3))
    g, bins, srt, delim = orc.grid(xyz, density=2.0)
    cpd = int(g.cpd)
    counts = np.bincount(bins, minlength=cpd ** 3)
    f = mlp.features(bins, delim, cpd)
    vals = mlp.bf16_value(f)
    for i in range(0, 3000, 97):
        b = int(bins[i])
        ix, iy, iz = b // (cpd * cpd), (b // cpd) % cpd, b % cpd
        j = 0
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    x, y, z = ix + dx, iy + dy, iz + dz
                    inside = 0 <= x < cpd and 0 <= y < cpd and 0 <= z < cpd
                    c = counts[(x * cpd + y) * cpd + z] if inside else 0
                    assert vals[i, j] == mlp.bf16(np.float32(np.log1p(c)))
                    j += 1
        assert np.all(f[i, mlp.FEATURES:] == 0)
    # the own cell's count appears in the centre feature
    assert np.all(vals[:, 13] == mlp.bf16(np.log1p(counts[bins]).astype(np.float32)))
