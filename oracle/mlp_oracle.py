"""TEST INFRASTRUCTURE ONLY — FP64 numpy oracle of the radiance-cache MLP (the checker of the
library's BF16-MFMA kernels).  Imported only by tests/ and bench.py's checks.

Network: 96(84 used)-64-64-64-64-64-16(3 used), ReLU, parameters W_l [out][in] then b_l.
The MFMA kernels feed bf16 operands (weights, inputs, every layer's ReLU output and every
back-propagated dZ) and accumulate in FP32; the oracle rounds the same tensors to bf16
(round-to-nearest-even) and accumulates in FP64.
"""
import numpy as np

IN, HID, OUT, LAYERS, FREQ = 96, 64, 16, 6, 12
DIMS = [(IN, HID)] + [(HID, HID)] * 4 + [(HID, OUT)]


def layer_offsets():
    offs, o = [], 0
    for i, j in DIMS:
        offs.append((o, o + i * j))
        o += i * j + j
    return offs, o


OFFS, NPARAM = layer_offsets()


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64).reshape(np.shape(x))


def unpack(params):
    p = np.asarray(params, np.float64)
    out = []
    for (i, j), (wo, bo) in zip(DIMS, OFFS):
        out.append((p[wo:wo + i * j].reshape(j, i), p[bo:bo + j]))
    return out


def forward(params, feats):
    """feats [n][96] -> outputs [n][3]; also returns the per-layer bf16 inputs."""
    layers = unpack(params)
    x = bf16(feats).T  # [96][n]
    acts = [x]
    z = None
    for l, (w, b) in enumerate(layers):
        z = bf16(w) @ x + b[:, None]
        if l < LAYERS - 1:
            x = bf16(np.maximum(z, 0.0))
            acts.append(x)
    return z[:3].T, acts


def loss_and_grad(params, feats, target):
    """Mean over the batch of |y - t|^2 and its gradient (bf16 dZ as on the MFMA path)."""
    layers = unpack(params)
    y, acts = forward(params, feats)
    B = feats.shape[0]
    diff = y - target
    loss = float((diff ** 2).sum() / B)
    dz = np.zeros((OUT, B))
    dz[:3] = bf16(2.0 * diff.T / B)
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        w, b = layers[l]
        grads[l] = (dz @ acts[l].T, dz.sum(1))
        if l == 0:
            break
        dx = bf16(w).T @ dz
        dz = bf16(dx * (acts[l] > 0))
    g = np.zeros(NPARAM)
    for (wo, bo), (gw, gb), (i, j) in zip(OFFS, grads, DIMS):
        g[wo:wo + i * j] = gw.ravel()
        g[bo:bo + j] = gb
    return loss, g


def encode(p, n, a, v):
    """84 features (+12 zero pads) of vertices: p, n, a, v are [m][3] arrays."""
    p = np.asarray(p, np.float64)
    m = p.shape[0]
    f = np.zeros((m, IN))
    f[:, 0:3] = np.clip(p, -2, 2)
    for q in range(6 * FREQ):
        c, k, cosine = q % 3, (q // 3) >> 1, (q // 3) & 1
        t = np.ldexp(p[:, c].astype(np.float32), k - 1).astype(np.float64)
        fr = t - np.floor(t)
        f[:, 3 + q] = np.cos(2 * np.pi * fr) if cosine else np.sin(2 * np.pi * fr)
    f[:, 75:78] = n
    f[:, 78:81] = a
    f[:, 81:84] = v
    return f


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
