"""The indicator oracle (oracle/mlp.py) pinned against independent routines:
torch autograd and torch.optim.Adam in float64, torch's bfloat16 rounding, and
the closed-form behaviour of the WENO5 weights on linear data and on jumps."""
import numpy as np
import pytest
import torch

from oracle import mlp as M
from oracle import oracle as O


def _sample(n, seed):
    rng = np.random.default_rng(seed)
    feat = np.zeros((n, 8), np.float32)
    feat[:, :6] = rng.uniform(-1, 1, (n, 6))
    feat[:, 6] = 1.0
    label = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    return feat, label


def _torch_net(params):
    W1, W2, b2, w3, b3 = [torch.tensor(np.asarray(a, np.float64), requires_grad=True)
                          for a in M.unpack(params)]
    return W1, W2, b2, w3, b3


def test_bf16_matches_torch_rounding():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 3, 4000), rng.normal(0, 1e-3, 1000),
                        [1.00390625, 1.01171875, -2.0078125, 0.0, -0.0]]).astype(np.float32)
    ours = M.bf16(x)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)


def test_gradient_matches_autograd():
    params = M.init_params(3)
    params[M.P_B2:M.P_W3] = np.random.default_rng(4).normal(0, 0.1, M.NH)
    params[M.P_W1 + 6 * M.NH:M.P_W1 + 7 * M.NH] = np.random.default_rng(5).normal(0, 0.1, M.NH)
    feat, label = _sample(300, 1)
    idx = np.arange(0, 300, 2)
    pw = 3.0
    loss, g = M.loss_grad(params, feat, label, idx, pw, quantize=False)
    W1, W2, b2, w3, b3 = _torch_net(params)
    X = torch.zeros(len(idx), M.NF, dtype=torch.float64)
    X[:, :8] = torch.tensor(feat[idx], dtype=torch.float64)
    z = torch.relu(torch.relu(X @ W1) @ W2 + b2) @ w3 + b3
    y = torch.tensor(label[idx], dtype=torch.float64)
    tl = torch.nn.functional.binary_cross_entropy_with_logits(
        z, y, pos_weight=torch.tensor(pw, dtype=torch.float64))
    tl.backward()
    assert abs(loss - tl.item()) < 1e-12
    tg = np.concatenate([W1.grad.numpy().ravel(), W2.grad.numpy().ravel(), b2.grad.numpy(),
                         w3.grad.numpy(), b3.grad.numpy().ravel()])
    assert np.allclose(g, tg, rtol=1e-10, atol=1e-14)


def test_adam_matches_torch():
    params = M.init_params(1).astype(np.float64)
    rng = np.random.default_rng(2)
    t = torch.tensor(params.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(params)
    v = np.zeros_like(params)
    for step in range(1, 6):
        g = rng.normal(0, 1, M.P_N)
        g[M.P_W1 + 7 * M.NH:M.P_W2] = 0.0
        M.adam(params, g, m, v, step, 1e-2)
        t.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(params, t.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_weno_weights_linear_on_linear_data():
    rng = np.random.default_rng(0)
    v = rng.uniform(-1, 1, (200, 1)) * np.arange(6)[None, :] + rng.uniform(-1, 1, (200, 1))
    assert M.departure(v).max() < 1e-12


def test_weno_weights_depart_at_a_jump():
    # a jump inside the stencil zeroes the weights of the sub-stencils that cross
    # it: at the interface itself (k = 3) only one sub-stencil per edge is
    # smooth; at the stencil's ends (k = 1, 5) one edge loses its far sub-stencil
    # (linear weight 0.1)
    expect = {1: 0.1, 2: 0.7, 3: 0.9, 4: 0.7, 5: 0.1}
    for k, d in expect.items():
        v = np.zeros(6)
        v[k:] = 1.0
        assert abs(M.departure(v[None])[0] - d) < 1e-9, k


def test_smooth_profile_departs_little():
    x = np.linspace(0, 1, 2049)
    u = np.sin(2 * np.pi * x)
    v = np.stack([u[i:i + 6] for i in range(0, 2040, 7)])
    assert M.departure(v).max() < 1e-2


def test_features_and_labels_of_a_step():
    mx, my = 12, 5
    q = np.ones((1, my, mx))
    q[0, :, 6:] = 2.0
    feat, label = M.features(q, [O.EXTRAP] * 2, [O.EXTRAP] * 2, 1e-8, 0.05)
    nxint = (mx + 1) * my
    assert feat.shape == ((mx + 1) * my + mx * (my + 1), 8)
    assert np.all(np.abs(feat[:, :6]) <= 1.0) and np.all(feat[:, 6] == 1.0)
    lx = label[:nxint].reshape(my, mx + 1)
    # x-interfaces whose stencil (cells c-3..c+2) straddles the jump at cells 6|7
    for c in range(1, mx + 2):
        straddles = c - 3 <= 6 and c + 2 >= 7
        assert np.all(lx[:, c - 1] == (1 if straddles else 0)), c
    # y-interfaces see constant data: flat stencils, zero features, label 0
    assert np.all(label[nxint:] == 0) and np.all(feat[nxint:, :6] == 0)
