"""Pins the FP64 reference gradient of the screen-mode NL-PF against central finite
differences of the FP64 oracle's own forward pass (no GPU).  The ramp + hue-shift chain is
piecewise linear with kinks at ramp stops and HSV sector edges, so the finite-difference
step is small (1e-8 relative for sigmas, 1e-5 for stop values) and the tiny image keeps kink crossings unlikely."""
import numpy as np
import pytest

import nlpf as N
from test_oracle_screen import rand_records


def setup(seed):
    rng = np.random.default_rng(seed)
    W, H, S, L, R = 10, 8, 2, 3, 2
    rec = rand_records(rng, W, H, S, L)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    table = N.style_table([N.ramp(pos, val), N.hue_shift(0.1, inner=0)])
    target = rng.random((H, W, 3)) * 0.5
    return rec, W, H, S, L, R, table, target


def loss(oracle, rec, W, H, S, L, R, sig, table, target):
    film = oracle.screen_run(rec, W, H, S, L, R, sig, nodes=table, root=1)
    return float(((film - target) ** 2).sum()), film


@pytest.mark.parametrize("seed", [1, 2])
def test_reference_gradient_matches_finite_differences(oracle, seed):
    rec, W, H, S, L, R, table, target = setup(seed)
    sig = np.array([0.05, 0.9, 0.5])
    l0, film = loss(oracle, rec, W, H, S, L, R, sig, table, target)
    g = oracle.screen_grad(rec, W, H, S, L, R, sig, table, 1, 2 * (film - target))
    assert g.shape == (19,)
    for k in range(3):
        h = 1e-8 * sig[k]
        sp, sm = sig.copy(), sig.copy()
        sp[k] += h
        sm[k] -= h
        lp = loss(oracle, rec, W, H, S, L, R, sp, table, target)[0]
        lm = loss(oracle, rec, W, H, S, L, R, sm, table, target)[0]
        # a kink of the chain inside [sigma-h, sigma+h] makes the central difference average
        # two slopes; the derivative at sigma is then one of the one-sided differences
        cands = [(lp - lm) / (2 * h), (lp - l0) / h, (l0 - lm) / h]
        err = min(abs(c - g[k]) for c in cands)
        assert err <= 1e-5 * max(1.0, abs(cands[0])), (k, cands, g[k])
    for m in range(16):
        tp, tm = table.copy(), table.copy()
        h = 1e-5  # the table is FP32: the quotient uses the exactly representable step
        tp[0]["p"][16 + m] += h
        tm[0]["p"][16 + m] -= h
        v0 = float(table[0]["p"][16 + m])
        up, um = float(tp[0]["p"][16 + m]) - v0, v0 - float(tm[0]["p"][16 + m])
        lp = loss(oracle, rec, W, H, S, L, R, sig, tp, target)[0]
        lm = loss(oracle, rec, W, H, S, L, R, sig, tm, target)[0]
        cands = [(lp - lm) / (up + um), (lp - l0) / up, (l0 - lm) / um]
        err = min(abs(c - g[3 + m]) for c in cands)
        assert err <= 5e-5 * max(1.0, abs(cands[0])), (m, cands, g[3 + m])
