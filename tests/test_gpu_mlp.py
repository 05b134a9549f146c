"""GPU parity of the BF16-MFMA radiance-cache MLP (through the C ABI) against the FP64 numpy
oracle: forward, gradient, one Adam step, the record encoder and the FP64 reference kernel;
plus a training run that must reduce the loss."""
import numpy as np
import pytest

import mlp_oracle as M
import nlpf as N
from gpu_util import gen_screen, host_records_fp32

pytestmark = pytest.mark.gpu


def random_params(rng):
    p = np.zeros(M.NPARAM, np.float32)
    for l, ((i, j), (wo, bo)) in enumerate(zip(M.DIMS, M.OFFS)):
        w = rng.normal(0, np.sqrt(2.0 / i), (j, i))
        if l == 0:
            w[:, M.PAD] = 0
        if l == M.LAYERS - 1:
            w[3:] = 0
        p[wo:wo + i * j] = w.ravel()
        p[bo:bo + j] = rng.normal(0, 0.05, j) * (np.arange(j) < (3 if l == M.LAYERS - 1 else j))
    return p


def random_feats(rng, n):
    f = rng.uniform(-1, 1, (n, M.IN)).astype(np.float32)
    f[:, M.PAD] = 0
    return f


@pytest.fixture(scope="module")
def h():
    with N.Handle({"type": "pf", "mode": "screen"}) as hh:
        hh.mlp_init(1)
        yield hh


def test_forward_matches_oracle(h):
    rng = np.random.default_rng(0)
    params = random_params(rng)
    h.mlp_set_params(params)
    assert np.array_equal(h.mlp_get_params(), params)
    n = 1000  # not a multiple of the 64-sample wave tile
    feats = random_feats(rng, n)
    df = h.to_device(feats)
    out = h.alloc(n * 12)
    h.mlp_infer(None, None, None, 0, None, df, n, out)
    y = out.download(np.float32, 3 * n).reshape(n, 3).astype(np.float64)
    ref, _ = M.forward(params, feats)
    err = np.abs(y - ref) / (np.abs(ref).max() + 1e-12)
    # bf16 rounding of a hidden activation can land on the other side of a tie when the FP32
    # and FP64 sums differ in the last bits: rare and small
    assert np.median(err) < 1e-6 and (err < 1e-4).mean() > 0.999 and err.max() < 1e-2, (np.median(err), err.max())


def test_forward_closed_form_network(h):
    """The hand-built |x| network (every value exact in bf16): the MFMA forward must give
    (|x|, x, 2 relu(x) + 0.5) bit for bit, reading x from any non-padding input column."""
    from known_nets import X_VALUES, abs_net, abs_net_expected
    rng = np.random.default_rng(5)
    n = 256
    x = np.resize(X_VALUES, n)
    for col in (int(np.flatnonzero(~M.PAD)[0]), int(np.flatnonzero(~M.PAD)[-1])):
        h.mlp_set_params(abs_net(col))
        feats = random_feats(rng, n)
        feats[:, col] = x
        out = h.alloc(n * 12)
        h.mlp_infer(None, None, None, 0, None, h.to_device(feats), n, out)
        y = out.download(np.float32, 3 * n).reshape(n, 3).astype(np.float64)
        assert np.array_equal(y, abs_net_expected(x)), (col, np.abs(y - abs_net_expected(x)).max())


def test_gradient_and_adam_match_oracle(h):
    rng = np.random.default_rng(1)
    params = random_params(rng)
    h.mlp_set_params(params)
    B = 1024
    feats = random_feats(rng, B)
    target = rng.normal(0, 1, (B, 3)).astype(np.float32)
    df, dt = h.to_device(feats), h.to_device(target)
    grad = h.alloc(M.NPARAM * 4)
    loss = h.mlp_train_step(None, None, None, 0, None, df, dt, B, 1e-2, apply=False, grad=grad)
    g = grad.download(np.float32, M.NPARAM).astype(np.float64)
    rl, rg = M.loss_and_grad(params, feats, target)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    assert np.linalg.norm(g - rg) <= 2e-3 * np.linalg.norm(rg), np.linalg.norm(g - rg) / np.linalg.norm(rg)
    # per layer as well (a transposed operand would break one layer's block only)
    for (i, j), (wo, bo) in zip(M.DIMS, M.OFFS):
        blk = slice(wo, bo + j)
        assert np.linalg.norm(g[blk] - rg[blk]) <= 5e-3 * np.linalg.norm(rg[blk]) + 1e-12
    # one Adam step (lr 1e-2) on the same batch
    h.mlp_train_step(None, None, None, 0, None, df, dt, B, 1e-2, apply=True)
    p1 = h.mlp_get_params().astype(np.float64)
    e1, _, _ = M.adam(params.astype(np.float64), g, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, 1e-2)
    pad = np.zeros(M.NPARAM, bool)
    wo, bo = M.OFFS[0]
    pad[wo:bo] = M.PAD[np.arange(96 * 64) % 96]
    wo, bo = M.OFFS[-1]
    pad[wo + 3 * 64:bo] = True
    pad[bo + 3:bo + 16] = True
    assert np.allclose(p1[~pad], e1[~pad], atol=1e-6)
    assert np.all(p1[pad] == 0)


def test_encoder_and_fp64_reference_on_records(h):
    rng = np.random.default_rng(2)
    W, H, S, L, d = 64, 40, 2, 4, 2
    desc, rec, _ = gen_screen(h, W, H, S, L, N.FP32, seed=4)
    host = host_records_fp32(rec, W, H, S, L)
    n = 700
    idx = rng.choice(S * W * H, n, replace=False).astype(np.uint32)
    didx = h.to_device(idx)
    feats = h.alloc(n * M.IN * 4)
    h.mlp_encode(desc, rec, None, d, didx, n, feats)
    gf = feats.download(np.float32, n * M.IN).reshape(n, M.IN)
    s, px = idx // (W * H), idx % (W * H)
    y, x = px // W, px % W
    p = host[s, d, 0:3, y, x]
    nn = host[s, d, 3:6, y, x]
    a = host[s, d, 6:9, y, x]
    par = host[s, d - 1, 0:3, y, x].astype(np.float64)
    v = par - p
    ln = np.linalg.norm(v, axis=1, keepdims=True)
    v = np.where(ln > 0, v / np.where(ln > 0, ln, 1), 0)  # escaped paths repeat the sky point
    ref = M.encode(p, nn, a, v)
    assert np.abs(gf - ref).max() < 2e-5
    params = random_params(rng)
    h.mlp_set_params(params)
    out64 = h.alloc(n * 24)
    h.mlp_reference(desc, rec, None, d, didx, n, out64)
    r64 = out64.download(np.float64, 3 * n).reshape(n, 3)
    o, _ = M.forward(params, gf.astype(np.float64))
    assert np.allclose(r64, o, rtol=1e-9, atol=1e-9)
    out = h.alloc(n * 12)
    h.mlp_infer(desc, rec, None, d, didx, None, n, out)
    y32 = out.download(np.float32, 3 * n).reshape(n, 3)
    err = np.abs(y32 - r64) / (np.abs(r64).max() + 1e-12)
    # as in the forward test: rare bf16 tie flips of a hidden activation (3 of 2100 seen)
    assert np.median(err) < 1e-6 and (err < 1e-4).mean() > 0.995 and err.max() < 1e-2, (np.median(err), err.max())


def test_training_reduces_loss(h):
    rng = np.random.default_rng(3)
    h.mlp_init(7)
    B = 4096
    feats = random_feats(rng, 4 * B)
    x = feats[:, :3].astype(np.float64)
    target = np.stack([np.sin(3 * x[:, 0]) + x[:, 1] ** 2, np.cos(2 * x[:, 1]) * x[:, 2], 0.5 + 0.3 * x[:, 0]], 1)
    df, dt = h.to_device(feats), h.to_device(target.astype(np.float32))
    losses = []
    for step in range(300):
        k = step % 4
        losses.append(h.mlp_train_step(None, None, None, 0, None, df.ptr + k * B * M.IN * 4,
                                       dt.ptr + k * B * 12, B, 1e-2))
    assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:5]), (losses[:5], losses[-5:])


def test_ema_inference_weights(h):
    """With a moving average on, inference uses the bias-corrected EMA of the parameter
    trajectory (checked against the FP64 oracle's forward on the numpy EMA of the downloaded
    parameters), the FP64 reference kernel uses the same weights, and turning it off returns
    to the parameters themselves."""
    rng = np.random.default_rng(5)
    params = random_params(rng)
    h.mlp_set_params(params)
    decay = 0.9
    h.mlp_set_ema(decay)
    B = 1024
    feats = random_feats(rng, B)
    target = rng.normal(0, 1, (B, 3)).astype(np.float32)
    df, dt = h.to_device(feats), h.to_device(target)
    ema = np.zeros(M.NPARAM)
    for t in range(1, 7):
        h.mlp_train_step(None, None, None, 0, None, df, dt, B, 1e-2)
        ema = decay * ema + (1 - decay) * h.mlp_get_params().astype(np.float64)
    ema_hat = ema / (1 - decay ** 6)
    raw = h.mlp_get_params()
    out = h.alloc(B * 12)
    h.mlp_infer(None, None, None, 0, None, df, B, out)
    y = out.download(np.float32, 3 * B).reshape(B, 3).astype(np.float64)
    ref_ema, _ = M.forward(ema_hat, feats)
    ref_raw, _ = M.forward(raw, feats)
    scale = np.abs(ref_ema).max()
    e_ema = np.median(np.abs(y - ref_ema)) / scale
    e_raw = np.median(np.abs(y - ref_raw)) / scale
    assert e_ema < 1e-3 and e_raw > 10 * e_ema, (e_ema, e_raw)
    h.mlp_set_ema(0.0)
    h.mlp_infer(None, None, None, 0, None, df, B, out)
    y0 = out.download(np.float32, 3 * B).reshape(B, 3).astype(np.float64)
    assert np.median(np.abs(y0 - ref_raw)) / scale < 1e-5
    with pytest.raises(N.NlpfError):
        h.mlp_set_ema(1.0)
