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


def test_forward_closed_form_network(monkeypatch):
    """The hand-built |x| network: the oracle's forward must give (|x|, x, 2 relu(x) + 0.5)
    exactly, whatever the other input columns hold."""
    from known_nets import X_VALUES, abs_net, abs_net_expected
    rng = np.random.default_rng(3)
    for col in (0, int(np.flatnonzero(~M.PAD)[-1])):
        feats = rng.uniform(-1, 1, (len(X_VALUES), M.IN))
        feats[:, col] = X_VALUES
        y, _ = M.forward(abs_net(col), feats)
        assert np.array_equal(y, abs_net_expected(X_VALUES)), (col, y)


def test_forward_closed_form_catches_planted_mistakes(monkeypatch):
    """Mistakes planted in the oracle (a transposed hidden weight, a hidden layer without its
    ReLU) each change the closed-form network's output."""
    from known_nets import X_VALUES, abs_net, abs_net_expected
    feats = np.zeros((len(X_VALUES), M.IN))
    feats[:, 0] = X_VALUES
    want = abs_net_expected(X_VALUES)
    p = abs_net(0)
    for l in range(1, 5):
        q = p.copy()
        wo, _ = M.OFFS[l]
        q[wo:wo + 64 * 64] = q[wo:wo + 64 * 64].reshape(64, 64).T.ravel()
        assert not np.array_equal(M.forward(q, feats)[0], want), l
    monkeypatch.setattr(M, "LAYERS", M.LAYERS - 1)  # layer 4 without its ReLU
    assert not np.array_equal(M.forward(p, feats)[0], want)


def test_adam_first_steps_closed_form():
    """Step 1 of Adam moves every parameter by -lr g / (|g| + eps) (the bias-corrected moments
    are g and g^2); step 2 from g1 = 2, g2 = -1 at lr 0.01: m2 = 0.9*0.2 - 0.1 = 0.08,
    v2 = 0.999*0.004 + 0.001 = 0.004996, m2/(1 - 0.81) = 0.42105263157894735,
    v2/(1 - 0.998001) = 2.4992496248124062, update -0.01*0.42105263157894735/1.5809015227...
    = -0.0026633."""
    g = np.array([2.0, -0.5, 1e-3, 0.0, -7.0])
    p0 = np.array([0.1, 0.2, -0.3, 0.4, 0.0])
    lr = 0.01
    p1, m, v = M.adam(p0, g, np.zeros(5), np.zeros(5), 1, lr)
    np.testing.assert_allclose(p1 - p0, -lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18)
    assert np.allclose(np.abs(p1 - p0)[g != 0], lr, rtol=1e-4)  # the sign step of Adam's first update
    p2, _, _ = M.adam(p1[:1], np.array([-1.0]), m[:1], v[:1], 2, lr)
    assert abs((p2[0] - p1[0]) - (-0.01 * 0.42105263157894735 / (np.sqrt(2.4992496248124062) + 1e-8))) < 1e-15
    assert abs((p2[0] - p1[0]) - (-0.0026633)) < 1e-7
