"""Pins the numpy MLP oracle: its analytic gradient against central finite differences of
its own loss (with the bf16 roundings switched to identity so the loss is smooth), the bf16
rounding helper against exact bit patterns, and the encoder against textbook values."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding_bits():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40])
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 2 ** -7  # above the tie -> up
    assert r[3] == -2.5
    assert np.isfinite(r[4])


def test_gradient_matches_finite_differences(monkeypatch):
    monkeypatch.setattr(M, "bf16", lambda x: np.asarray(x, np.float64))
    rng = np.random.default_rng(0)
    params = np.zeros(M.NPARAM)
    for (i, j), (wo, bo) in zip(M.DIMS, M.OFFS):
        params[wo:wo + i * j] = rng.normal(0, 1 / np.sqrt(i), i * j)
        params[bo:bo + j] = rng.normal(0, 0.1, j)
    feats = rng.normal(size=(8, M.IN))
    target = rng.normal(size=(8, 3))
    loss, g = M.loss_and_grad(params, feats, target)
    # output rows 3..15 carry no loss: their gradients vanish
    wo, bo = M.OFFS[-1]
    assert np.all(g[wo + 3 * 64:wo + 16 * 64] == 0) and np.all(g[bo + 3:bo + 16] == 0)
    idx = rng.choice(M.NPARAM, 40, replace=False)
    for k in idx:
        h = 1e-6
        p1, p2 = params.copy(), params.copy()
        p1[k] += h
        p2[k] -= h
        fd = (M.loss_and_grad(p1, feats, target)[0] - M.loss_and_grad(p2, feats, target)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_encoder_known_values():
    p = np.array([[0.25, -0.5, 0.125]])
    n, a, v = np.array([[0, 0, 1.0]]), np.array([[0.2, 0.3, 0.4]]), np.array([[0, 1.0, 0]])
    f = M.encode(p, n, a, v)[0]
    raw = {"p": p[0], "n": n[0], "a": a[0], "v": v[0]}
    seen = set()
    for P, col in enumerate(M.COLUMNS):
        seen.add(col if col[0] != "pad" else ("pad", P))
        if col[0] == "sin":
            assert np.isclose(f[P], np.sin(2 ** col[1] * np.pi * p[0, col[2]]), atol=1e-12)
        elif col[0] == "cos":
            assert np.isclose(f[P], np.cos(2 ** col[1] * np.pi * p[0, col[2]]), atol=1e-12)
        elif col[0] == "pad":
            assert f[P] == 0
        else:
            assert f[P] == raw[col[0]][col[1]]
    # the layout is a permutation of the 84 features plus 12 pads
    want = {(t, k, c) for t in ("sin", "cos") for k in range(M.FREQ) for c in range(3)}
    want |= {(t, c) for t in "pnav" for c in range(3)}
    assert {c for c in seen if c[0] != "pad"} == want and M.PAD.sum() == 12


def test_encoder_lane_groups_share_code():
    """Column P is computed by MFMA lane group (P >> 2) & 3; every group holds 18 sin/cos
    columns of 3 consecutive octaves, so the GPU encoder runs one code path per lane."""
    for g in range(4):
        cols = [M.COLUMNS[P] for P in range(M.IN) if (P >> 2) & 3 == g]
        trig = [c for c in cols if c[0] in ("sin", "cos")]
        assert len(cols) == 24 and len(trig) == 18
        assert {c[1] for c in trig} == {3 * g, 3 * g + 1, 3 * g + 2}
