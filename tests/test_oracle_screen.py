"""Pins the screen-mode oracle (FP64 and FP32 instances) against a hand-computed two-pixel
case, an independent numpy brute force on tiny images, the sigma -> infinity box reduction,
the constant-input invariant and the depth recursion (no GPU)."""
import numpy as np
import pytest

import nlpf as N

W709 = np.array([0.2126, 0.7152, 0.0722])


def rand_records(rng, W, H, S, L, pos_scale=0.05):
    rec = np.empty((S, L, 15, H, W), np.float32)
    yy, xx = np.mgrid[0:H, 0:W]
    for s in range(S):
        for d in range(L):
            # a tilted plane with noise so that position weights span (0, 1)
            rec[s, d, 0] = xx * pos_scale + rng.normal(0, 0.01, (H, W))
            rec[s, d, 1] = yy * pos_scale + rng.normal(0, 0.01, (H, W))
            rec[s, d, 2] = 0.3 * xx * pos_scale + rng.normal(0, 0.01, (H, W))
            n = rng.normal(0, 0.4, (3, H, W)) + np.array([0, 0, 1.0])[:, None, None]
            rec[s, d, 3:6] = n / np.linalg.norm(n, axis=0)
            rec[s, d, 6:9] = rng.uniform(0.1, 0.9, (3, H, W))
            rec[s, d, 9:12] = np.exp(-1 + 1.5 * rng.normal(size=(3, H, W)))
            rec[s, d, 12:15] = rng.uniform(0, 1, (3, H, W))
    return rec


def brute(rec, W, H, S, L, radius, sig, style=None, vstyle=None):
    """Independent numpy restatement: explicit per-pixel window loops."""
    sp, sn, sl = sig
    rec = rec.astype(np.float64)
    child = None
    levels = np.zeros((S, L, 3, H, W))
    for d in range(L - 1, -1, -1):
        for s in range(S):
            r = rec[s, d]
            inp = r[9:12].copy()
            if d + 1 < L:
                inp += r[12:15] * levels[s, d + 1]
            lum = np.tensordot(W709, r[6:9], axes=1)
            for y in range(H):
                for x in range(W):
                    ys = slice(max(0, y - radius), min(H, y + radius + 1))
                    xs = slice(max(0, x - radius), min(W, x + radius + 1))
                    dp = r[0:3, ys, xs] - r[0:3, y, x][:, None, None]
                    nn = (r[3:6, ys, xs] * r[3:6, y, x][:, None, None]).sum(0)
                    dl = lum[ys, xs] - lum[y, x]
                    w = np.exp(-(dp ** 2).sum(0) / (2 * sp * sp) - (1 - nn) ** 2 / (2 * sn * sn)
                               - dl ** 2 / (2 * sl * sl))
                    f = (w * inp[:, ys, xs]).sum((1, 2)) / w.sum()
                    if vstyle:
                        f = vstyle(f, s, d, y, x)
                    elif style:
                        f = style(f, d)
                    levels[s, d, :, y, x] = f
    film = levels[:, 0].mean(0).transpose(1, 2, 0)
    return film, levels


def test_two_pixel_closed_form(oracle):
    # W=2, H=1: out_0 = (in_0 + w in_1) / (1 + w) with w from the three guide distances
    rec = np.zeros((1, 1, 15, 1, 2), np.float32)
    rec[0, 0, 0:3, 0, 0] = [0.0, 0.0, 0.0]
    rec[0, 0, 0:3, 0, 1] = [0.03, 0.04, 0.0]        # |dp| = 0.05
    rec[0, 0, 3:6, 0, 0] = [0.0, 0.0, 1.0]
    rec[0, 0, 3:6, 0, 1] = [0.6, 0.0, 0.8]          # n.n = 0.8
    rec[0, 0, 6:9, 0, 0] = [0.5, 0.5, 0.5]
    rec[0, 0, 6:9, 0, 1] = [0.25, 0.25, 0.25]       # dl = 0.25
    rec[0, 0, 9:12, 0, 0] = [1.0, 2.0, 3.0]
    rec[0, 0, 9:12, 0, 1] = [5.0, 7.0, 11.0]
    sp, sn, sl = 0.05, 0.9, 0.5
    # the guides are float32 records, so the distances are those of the rounded values
    dx, dy, nz = (float(np.float32(v)) for v in (0.03, 0.04, 0.8))
    w = np.exp(-(dx * dx + dy * dy) / (2 * sp * sp) - (1 - nz) ** 2 / (2 * sn * sn) - 0.0625 / (2 * sl * sl))
    film = oracle.screen_run(rec, 2, 1, 1, 1, 1, (sp, sn, sl))
    e0 = (np.array([1.0, 2.0, 3.0]) + w * np.array([5.0, 7.0, 11.0])) / (1 + w)
    e1 = (np.array([5.0, 7.0, 11.0]) + w * np.array([1.0, 2.0, 3.0])) / (1 + w)
    assert np.allclose(film[0, 0], e0, rtol=1e-12)
    assert np.allclose(film[0, 1], e1, rtol=1e-12)


@pytest.mark.parametrize("radius", [1, 2, 3])
def test_matches_numpy_brute_force(oracle, radius):
    rng = np.random.default_rng(10 + radius)
    W, H, S, L = 9, 7, 2, 3
    rec = rand_records(rng, W, H, S, L)
    sig = (0.05, 0.9, 0.5)
    film, lv = oracle.screen_run(rec, W, H, S, L, radius, sig, want_levels=True)
    bf, blv = brute(rec, W, H, S, L, radius, sig)
    assert np.allclose(lv, blv, rtol=1e-11, atol=1e-13)
    assert np.allclose(film, bf, rtol=1e-11, atol=1e-13)
    # the FP32 instance (CPU baseline) stays within FP32 accumulation error
    f32 = oracle.screen_run(rec, W, H, S, L, radius, sig, precision="f32")
    assert np.allclose(f32, film, rtol=2e-5, atol=1e-6)


def test_style_epilogue_and_recursion(oracle):
    rng = np.random.default_rng(20)
    W, H, S, L = 8, 6, 1, 3
    rec = rand_records(rng, W, H, S, L)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    table = N.style_table([N.ramp(pos, val), N.hue_shift(0.1, inner=0)])
    sig = (0.05, 0.9, 0.5)
    film, lv = oracle.screen_run(rec, W, H, S, L, 2, sig, nodes=table, root=1, want_levels=True)

    def style(f, d):
        return oracle.stylize(table, [1], f[None], depth=d)[0]

    bf, blv = brute(rec, W, H, S, L, 2, sig, style=style)
    assert np.allclose(lv, blv, rtol=1e-11, atol=1e-13)
    assert np.allclose(film, bf, rtol=1e-11, atol=1e-13)


def test_per_vertex_materials_and_flags(oracle):
    """Material and flags planes: each vertex is stylized by its own material's tree with its own
    flags, at its own slice and depth (materials past the table and root -1 give the identity).
    The brute force applies the oracle's stylize per vertex; all-zero planes reproduce the
    one-tree run bit for bit."""
    rng = np.random.default_rng(11)
    W, H, S, L = 9, 7, 2, 3
    sig = (0.08, 0.4, 0.25)
    rec = rand_records(rng, W, H, S, L)
    table = N.style_table([N.rex(0.6), N.enable_if(2, inner=0), N.saturation(0.5),
                           N.enable_if(0, inner=2, else_=1), N.hue_shift(0.25), N.set_idx(2, inner=4),
                           N.toon(0.3)])
    roots = [3, 5, -1, 6]
    material = rng.integers(0, 5, (S, L, H, W)).astype(np.uint16)  # 4 is past the table
    flags = rng.choice([0, N.FLAG_STYLE2, N.FLAG_STYLE3, N.FLAG_STYLE4, N.FLAG_DISABLED],
                       (S, L, H, W)).astype(np.uint8)
    film, lv = oracle.screen_run(rec, W, H, S, L, 2, sig, nodes=table, want_levels=True, material=material,
                                 flags=flags, roots=roots)

    def vstyle(f, s, d, y, x):
        return oracle.stylize(table, roots, f[None], depth=d, flags=[flags[s, d, y, x]],
                              material=[material[s, d, y, x]])[0]

    bf, blv = brute(rec, W, H, S, L, 2, sig, vstyle=vstyle)
    assert np.abs(lv - blv).max() <= 1e-12 * max(1.0, np.abs(blv).max())
    assert np.abs(film - bf).max() <= 1e-12 * max(1.0, np.abs(bf).max())
    # the planes matter: the same run with every vertex on material 0 and flags 0 differs
    zero_m = np.zeros_like(material)
    f0, lv0 = oracle.screen_run(rec, W, H, S, L, 2, sig, nodes=table, want_levels=True, material=zero_m,
                                flags=np.zeros_like(flags), roots=roots)
    assert np.abs(f0 - film).max() > 1e-3
    one, lv1 = oracle.screen_run(rec, W, H, S, L, 2, sig, nodes=table, root=roots[0], want_levels=True)
    assert np.array_equal(f0, one) and np.array_equal(lv0, lv1)
    for prec in ("f64", "f32"):
        a = oracle.screen_run(rec, W, H, S, L, 2, sig, nodes=table, root=roots[0], precision=prec)
        b = oracle.screen_run(rec, W, H, S, L, 2, sig, nodes=table, precision=prec, material=zero_m,
                              flags=None, roots=roots)
        assert np.array_equal(a, b)


def test_box_limit_and_constant_invariant(oracle):
    rng = np.random.default_rng(30)
    W, H, S, L = 10, 8, 1, 1
    rec = rand_records(rng, W, H, S, L)
    # sigma -> infinity: every weight is 1, the filter is a clipped box mean
    film = oracle.screen_run(rec, W, H, S, L, 2, (1e8, 1e8, 1e8))
    inp = rec[0, 0, 9:12].astype(np.float64)
    box = np.zeros_like(inp)
    for y in range(H):
        for x in range(W):
            box[:, y, x] = inp[:, max(0, y - 2):y + 3, max(0, x - 2):x + 3].mean((1, 2))
    assert np.allclose(film.transpose(2, 0, 1), box, rtol=1e-12)
    # a constant input image is a fixed point whatever the guides
    rec[0, 0, 9:12] = np.array([0.3, 0.6, 0.9], np.float32)[:, None, None]
    film = oracle.screen_run(rec, W, H, S, L, 3, (0.05, 0.9, 0.5))
    assert np.allclose(film, np.array([0.3, 0.6, 0.9], np.float32).astype(np.float64), rtol=1e-12)


def test_zero_throughput_decouples_levels(oracle):
    rng = np.random.default_rng(40)
    W, H, S, L = 7, 5, 2, 3
    rec = rand_records(rng, W, H, S, L)
    rec[:, :, 12:15] = 0
    film, lv = oracle.screen_run(rec, W, H, S, L, 1, (0.05, 0.9, 0.5), want_levels=True)
    single = oracle.screen_run(np.ascontiguousarray(rec[:, :1]), W, H, S, 1, 1, (0.05, 0.9, 0.5))
    assert np.allclose(film, single, rtol=1e-13)
    # film is the slice mean of level 0
    assert np.allclose(film, lv[:, 0].mean(0).transpose(1, 2, 0), rtol=1e-13)
