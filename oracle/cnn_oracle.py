"""CPU oracle of the learned-initial-guess network (test infrastructure only).

A plain numpy restatement, in FP64, of what the HIP library's CNN computes.
The reference has no learned part, so this is pinned by mathematics rather
than by reference outputs: tests/test_cnn_oracle.py checks the forward pass
against a brute-force loop and against torch's CPU conv3d, the backward pass
against central finite differences of the forward pass, and Adam against
torch.optim.Adam.

Network: 3 -> 16 -> 16 -> 16 -> 1 channels, 3x3x3 kernels, zero "same"
padding (every layer keeps the volume size), ReLU after the first three
layers.  Tensors are channels-last: x[b, z, y, x, c].  Parameters are one
flat vector; per layer W[tap][cin][cout] with tap = kz*9 + ky*3 + kx
(offsets kx-1, ky-1, kz-1), then bias[cout]:
    y[v, co] = b[co] + sum_t sum_ci W[t, ci, co] * x[v + off(t), ci]

`bf16=True` rounds what the GPU stores in BF16 (inputs, weights, hidden
activations, output gradients) with round-to-nearest-even, so the GPU path
can be compared at accumulation-order tolerance.
"""
import numpy as np

CIN = (3, 16, 16, 16)
COUT = (16, 16, 16, 1)
N_PARAM = 15601


def layer_slices():
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        w = slice(off, off + 27 * ci * co)
        off += 27 * ci * co
        b = slice(off, off + co)
        off += co
        out.append((w, b))
    assert off == N_PARAM
    return out


SLICES = layer_slices()


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def weights(params, layer, bf16=False):
    ws, bs = SLICES[layer]
    W = np.asarray(params, dtype=np.float64)[ws].reshape(27, CIN[layer], COUT[layer])
    b = np.asarray(params, dtype=np.float64)[bs]
    if bf16:
        W = bf16_round(W)
        b = np.asarray(params, dtype=np.float32)[bs].astype(np.float64)
    return W, b


def offsets():
    return [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]


def conv_same(x, W, b=None):
    """y[v, co] = b[co] + sum_t x[v + off(t)] @ W[t]   (zero padding)."""
    nb, vz, vy, vx, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((nb, vz, vy, vx, W.shape[2]))
    for t, (dx, dy, dz) in enumerate(offsets()):
        src = xp[:, 1 + dz:1 + dz + vz, 1 + dy:1 + dy + vy, 1 + dx:1 + dx + vx, :]
        y += src @ W[t]
    if b is not None:
        y += b
    return y


def conv_transpose_same(dz, W):
    """Adjoint of conv_same in x: dx[u] = sum_t dz[u - off(t)] @ W[t]^T."""
    nb, vz, vy, vx, _ = dz.shape
    dp = np.pad(dz, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nb, vz, vy, vx, W.shape[1]))
    for t, (dx, dy, dzz) in enumerate(offsets()):
        src = dp[:, 1 - dzz:1 - dzz + vz, 1 - dy:1 - dy + vy, 1 - dx:1 - dx + vx, :]
        out += src @ W[t].T
    return out


def conv_weight_grad(x, dz):
    """dW[t, ci, co] = sum_v x[v + off(t), ci] * dz[v, co]."""
    nb, vz, vy, vx, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    g = np.zeros((27, ci, dz.shape[-1]))
    dzf = dz.reshape(-1, dz.shape[-1])
    for t, (dx, dy, dzz) in enumerate(offsets()):
        src = xp[:, 1 + dzz:1 + dzz + vz, 1 + dy:1 + dy + vy, 1 + dx:1 + dx + vx, :].reshape(-1, ci)
        g[t] = src.T @ dzf
    return g


def forward(params, x, bf16=False):
    """Returns (output [nb,vz,vy,vx], activations [x, a1, a2, a3])."""
    a = bf16_round(x) if bf16 else np.asarray(x, dtype=np.float64)
    acts = [a]
    for layer in range(4):
        W, b = weights(params, layer, bf16)
        z = conv_same(a, W, b)
        if layer < 3:
            a = np.maximum(z, 0.0)
            if bf16:
                a = bf16_round(a)
            acts.append(a)
        else:
            return z[..., 0], acts


def gradient(params, x, dy, bf16=False):
    """Gradient of sum(dy * forward(x)) with respect to the parameters."""
    _, acts = forward(params, x, bf16)
    g = np.zeros(N_PARAM)
    dz = np.asarray(dy, dtype=np.float64)[..., None]
    if bf16:
        dz = bf16_round(dz)
    for layer in range(3, -1, -1):
        ws, bs = SLICES[layer]
        g[ws] = conv_weight_grad(acts[layer], dz).reshape(-1)
        g[bs] = dz.reshape(-1, dz.shape[-1]).sum(axis=0)
        if layer > 0:
            W, _ = weights(params, layer, bf16)
            dx = conv_transpose_same(dz, W)
            dz = dx * (acts[layer] > 0)
            if bf16:
                dz = bf16_round(dz)
    return g


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, Algorithm 1); t is the 1-based step."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
