"""Pins of the CNN oracle (oracle/cnn_oracle.py), which has no counterpart
in the reference: a brute-force convolution, torch's CPU conv3d as a second
implementation, central finite differences for the backward pass,
torch.optim.Adam for the optimiser and torch's bfloat16 cast for rounding."""
import numpy as np
import pytest

import cnn_oracle as co


def rand_params(rng, scale=0.3):
    return rng.standard_normal(co.N_PARAM) * scale


def test_parameter_count_and_flops():
    assert co.N_PARAM == 15601
    # multiply-adds per voxel (27 taps): 2 * 27 * (3*16 + 16*16 + 16*16 + 16*1) = 31104 flops
    assert 2 * 27 * sum(a * b for a, b in zip(co.CIN, co.COUT)) == 31104


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 4, 5, 3))
    W = rng.standard_normal((27, 3, 4))
    b = rng.standard_normal(4)
    y = co.conv_same(x, W, b)
    nb, vz, vy, vx, _ = x.shape
    ref = np.zeros((nb, vz, vy, vx, 4))
    for q in range(nb):
        for z in range(vz):
            for yy in range(vy):
                for xx in range(vx):
                    acc = b.copy()
                    for kz in range(3):
                        for ky in range(3):
                            for kx in range(3):
                                zs, ys, xs = z + kz - 1, yy + ky - 1, xx + kx - 1
                                if 0 <= zs < vz and 0 <= ys < vy and 0 <= xs < vx:
                                    acc += x[q, zs, ys, xs] @ W[kz * 9 + ky * 3 + kx]
                    ref[q, z, yy, xx] = acc
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_forward_matches_torch_conv3d():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    p = rand_params(rng)
    x = rng.standard_normal((2, 5, 6, 7, 3))
    y, _ = co.forward(p, x)
    a = torch.tensor(np.moveaxis(x, -1, 1), dtype=torch.float64)
    for layer in range(4):
        W, b = co.weights(p, layer)
        # W[t, ci, co] with t = kz*9+ky*3+kx  ->  torch weight [co, ci, kz, ky, kx]
        wt = torch.tensor(W.reshape(3, 3, 3, co.CIN[layer], co.COUT[layer]).transpose(4, 3, 0, 1, 2).copy())
        a = torch.nn.functional.conv3d(a, wt, torch.tensor(b), padding=1)
        if layer < 3:
            a = torch.relu(a)
    np.testing.assert_allclose(y, a[:, 0].numpy(), rtol=1e-10, atol=1e-10)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rand_params(rng)
    # biases > 0 keep most ReLUs away from their kink
    for layer in range(3):
        p[co.SLICES[layer][1]] = 0.5
    x = rng.standard_normal((1, 4, 5, 6, 3))
    dy = rng.standard_normal((1, 4, 5, 6))
    g = co.gradient(p, x, dy)
    f = lambda q: float(np.sum(dy * co.forward(q, x)[0]))
    h = 1e-6
    for layer in range(4):
        ws, bs = co.SLICES[layer]
        picks = list(rng.choice(np.arange(ws.start, ws.stop), 6, replace=False)) + [bs.start, bs.stop - 1]
        for k in picks:
            e = np.zeros_like(p)
            e[k] = h
            fd = (f(p + e) - f(p - e)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (layer, k, fd, g[k])


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(5)]
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(grads, start=1):
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = co.adam_step(p, g, m, v, t)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)
    # closed form of the first step: m_hat = g, v_hat = g^2, so every
    # parameter moves by lr * g / (|g| + eps)
    g0 = grads[0]
    p1, _, _ = co.adam_step(p0, g0, np.zeros(50), np.zeros(50), 1)
    np.testing.assert_allclose(p0 - p1, 1e-3 * g0 / (np.abs(g0) + 1e-8), rtol=1e-9)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-20, 20, 10000),
                        [0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9]]).astype(np.float32)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(co.bf16_round(a), ref)
