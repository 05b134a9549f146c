"""TEST INFRASTRUCTURE ONLY — FP64 numpy oracle of the radiance-cache MLP (the checker of the
library's BF16-MFMA kernels).  Imported only by tests/ and bench.py's checks.

Network: 96(84 used, column layout in column_layout())-64-64-64-64-64-16(3 used), ReLU,
parameters W_l [out][in] then b_l.
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


def column_layout():
    """What each of the 96 input columns (the columns of W_0) holds: ("sin"|"cos", k, c) for
    sin/cos(2^k pi p_c), ("p"|"n"|"a"|"v", c) for the raw features, ("pad",) for zeros.
    Column P belongs to MFMA lane group g = (P >> 2) & 3 as its slot
    u = 8 (P // 32) + (4 if P % 32 >= 16 else 0) + P % 4; slots u < 18 hold octave
    k = 3 g + u // 6 of coordinate c = (u % 6) // 2 (sin for even u, cos for odd u); slots
    18..23 hold p, n (g = 0), albedo, view (g = 1) and padding (g = 2, 3)."""
    cols = []
    for P in range(IN):
        g = (P >> 2) & 3
        u = 8 * (P // 32) + (4 if P % 32 >= 16 else 0) + P % 4
        if u < 18:
            cols.append(("cos" if u & 1 else "sin", 3 * g + u // 6, (u % 6) // 2))
        elif g < 2:
            r = u - 18
            cols.append(((("p", "n"), ("a", "v"))[g][r // 3], r % 3))
        else:
            cols.append(("pad",))
    return cols


COLUMNS = column_layout()
PAD = np.array([c[0] == "pad" for c in COLUMNS])


def encode(p, n, a, v):
    """96 input columns (84 features + 12 zero pads) of vertices: p, n, a, v are [m][3]."""
    p = np.asarray(p, np.float64)
    raw = {"p": np.clip(p, -2, 2), "n": np.asarray(n, np.float64), "a": np.asarray(a, np.float64),
           "v": np.asarray(v, np.float64)}
    f = np.zeros((p.shape[0], IN))
    for P, col in enumerate(COLUMNS):
        if col[0] in ("sin", "cos"):
            _, k, c = col
            # 2^k pi p through the exactly reduced fraction of 2^(k-1) p (p is float32 data)
            t = np.ldexp(p[:, c].astype(np.float32), k - 1).astype(np.float64)
            fr = t - np.floor(t)
            f[:, P] = np.cos(2 * np.pi * fr) if col[0] == "cos" else np.sin(2 * np.pi * fr)
        elif col[0] != "pad":
            f[:, P] = raw[col[0]][:, col[1]]
    return f


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
