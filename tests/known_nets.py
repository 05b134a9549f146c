"""Hand-built radiance-cache networks whose outputs are known in closed form, shared by the
oracle's and the GPU's forward tests.

abs_net(col) reads input column `col` as x and returns (|x|, x, 2 relu(x) + 0.5):
  layer 0   unit 0 = relu(x), unit 1 = relu(-x), unit 63 = relu(-1) = 0
  layer l   (1..4) moves the two carriers to new units (0,1) -> (5,9) -> (17,30) -> (41,2)
            -> (11,60), never along the diagonal, so a transposed weight loses them; every
            carrier also adds the previous layer's unit 63 and every layer's unit 63 has
            bias -1, so a layer without its ReLU shifts every output by -1 or more
  layer 5   out0 = a + b, out1 = a - b, out2 = 2 a + 0.5, no ReLU (out1 < 0 for x < 0)
Every weight, bias and activation is exactly representable in bf16 for x in X_VALUES, so the
bf16 roundings of the MFMA path change nothing."""
import numpy as np

import mlp_oracle as M

CARRIERS = [(0, 1), (5, 9), (17, 30), (41, 2), (11, 60)]
X_VALUES = np.array([-1.5, -0.25, 0.0, 0.75, 2.0, -3.0, 0.5, 1.0])


def abs_net(col):
    p = np.zeros(M.NPARAM, np.float32)
    layers = []
    for (i, j), (wo, bo) in zip(M.DIMS, M.OFFS):
        layers.append((p[wo:wo + i * j].reshape(j, i), p[bo:bo + j]))
    w0, b0 = layers[0]
    w0[0, col], w0[1, col] = 1.0, -1.0
    b0[63] = -1.0
    for l in range(1, 5):
        w, b = layers[l]
        (pa, pb), (qa, qb) = CARRIERS[l - 1], CARRIERS[l]
        w[qa, pa] = w[qb, pb] = 1.0
        w[qa, 63] = w[qb, 63] = 1.0
        b[63] = -1.0
    w5, b5 = layers[5]
    a, b = CARRIERS[4]
    w5[0, a], w5[0, b] = 1.0, 1.0
    w5[1, a], w5[1, b] = 1.0, -1.0
    w5[2, a], b5[2] = 2.0, 0.5
    w5[0, 63] = w5[1, 63] = w5[2, 63] = 1.0
    return p


def abs_net_expected(x):
    x = np.asarray(x, np.float64)
    return np.stack([np.abs(x), x, 2 * np.maximum(x, 0) + 0.5], axis=1)
