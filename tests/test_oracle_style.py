"""Pins the CPU oracle's stylization arithmetic against closed forms, textbook values and
brute force (no GPU).  Each check is chosen so that a plausible slip in the oracle (a wrong
luminance weight, a swapped HSV sector, an sRGB constant, an inner/outer order) fails it."""
import colorsys

import numpy as np
import pytest

import nlpf as N


def sty(oracle, nodes, rgb, depth=0, flags=None, root=0, pos=None):
    table = N.style_table(nodes)
    return oracle.stylize(table, [root], np.asarray(rgb, np.float64).reshape(-1, 3), depth=depth,
                          flags=flags, pos=pos)


def srgb_encode(x):  # IEC 61966-2-1
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(np.maximum(x, 0), 1 / 2.4) - 0.055)


def srgb_decode(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.04045, x / 12.92, np.power((x + 0.055) / 1.055, 2.4))


def test_hsv_matches_textbook(oracle):
    rng = np.random.default_rng(0)
    for c in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0.2, 0.5, 0.9), (0.3, 0.3, 0.3)] + \
            [tuple(v) for v in rng.random((50, 3))]:
        hsv = oracle.rgb_to_hsv(np.array(c, float))
        ref = colorsys.rgb_to_hsv(*c)
        assert np.allclose(hsv, ref, atol=1e-12), (c, hsv, ref)
        back = oracle.hsv_to_rgb(hsv)
        assert np.allclose(back, c, atol=1e-12)


def test_toon_closed_form(oracle):
    w = 0.25
    nodes = [N.toon(w)]
    # black stays black; negative input clamps to black
    assert np.allclose(sty(oracle, nodes, [0, 0, 0]), 0)
    assert np.allclose(sty(oracle, nodes, [-1, -2, -3]), 0)
    # lum > 1 branch: luminance snaps down to a multiple of w, hue ratio preserved
    c = np.array([2.0, 1.5, 0.7])
    lum = 0.2126 * 2.0 + 0.7152 * 1.5 + 0.0722 * 0.7
    out = sty(oracle, nodes, c)[0]
    assert np.allclose(out, c / lum * (np.floor(lum / w) * w), rtol=1e-14)
    # lum <= 1: the band edge lives in sRGB space
    for lum_target in [0.01, 0.05, 0.2, 0.5, 0.9]:
        grey = np.full(3, lum_target)  # lum(grey) = grey (weights sum to 1)
        expect = srgb_decode(np.floor(srgb_encode(lum_target) / w) * w)
        assert np.allclose(sty(oracle, nodes, grey)[0], expect, rtol=1e-12)
    # the Rec.709 weights: each primary alone has lum = its weight
    for i, wt in enumerate([0.2126, 0.7152, 0.0722]):
        c = np.zeros(3)
        c[i] = 3.0 / wt  # lum 3 -> band 3.0 exactly, so output = input
        assert np.allclose(sty(oracle, nodes, c)[0], c, rtol=1e-12)
    # outputs sit on band values: lum/w (lum > 1) or srgb(lum)/w (lum <= 1) is an integer,
    # never above the input's band
    rng = np.random.default_rng(1)
    x = rng.random((200, 3)) * 3
    y = sty(oracle, nodes, x)
    wts = np.array([0.2126, 0.7152, 0.0722])
    lin, lout = x @ wts, y @ wts
    q = np.where(lin > 1, lout / w, srgb_encode(lout) / w)
    assert np.allclose(q, np.round(q), atol=1e-9)
    assert np.all(lout <= lin + 1e-12)


def test_ramp_and_hue_shift(oracle):
    pos = np.sort(np.random.default_rng(2).random(16)).astype(np.float32)
    val = np.sort(np.random.default_rng(3).random(16)).astype(np.float32)
    nodes = [N.ramp(pos, val)]
    # knots of grey inputs reproduce the stop values (grey lum = value)
    for k in range(16):
        g = np.full(3, float(pos[k]))
        assert np.allclose(sty(oracle, nodes, g)[0], np.full(3, float(val[k])), rtol=1e-6, atol=1e-7)
    # midpoint of a segment is the mean of its stops (linear interpolation)
    mid = 0.5 * (float(pos[4]) + float(pos[5]))
    assert np.allclose(sty(oracle, nodes, np.full(3, mid))[0], 0.5 * (float(val[4]) + float(val[5])), rtol=1e-6)
    # clamped outside [pos0, pos15], hue preserved
    assert np.allclose(sty(oracle, nodes, np.full(3, 50.0))[0], float(val[15]))
    c = np.array([0.3, 0.1, 0.05])
    out = sty(oracle, nodes, c)[0]
    assert np.allclose(out / out.sum(), c / c.sum(), rtol=1e-12)
    # hue shift: 1.0 is the identity; 1/3 maps red -> green -> blue (textbook HSV)
    assert np.allclose(sty(oracle, [N.hue_shift(1.0)], [0.2, 0.5, 0.9])[0], [0.2, 0.5, 0.9], atol=1e-6)
    assert np.allclose(sty(oracle, [N.hue_shift(1.0 / 3.0)], [1, 0, 0])[0], [0, 1, 0], atol=1e-6)
    assert np.allclose(sty(oracle, [N.hue_shift(2.0 / 3.0)], [1, 0, 0])[0], [0, 0, 1], atol=1e-6)
    # against colorsys on random colours, HDR scale equivariance (V carries the scale)
    rng = np.random.default_rng(4)
    for c in rng.random((40, 3)) * 5:
        h, s, v = colorsys.rgb_to_hsv(*c)
        ref = np.array(colorsys.hsv_to_rgb((h + np.float32(0.1)) % 1.0, s, v))
        assert np.allclose(sty(oracle, [N.hue_shift(0.1)], c)[0], ref, rtol=1e-6, atol=1e-9)


def test_saturation_rex_palette_eightbit(oracle):
    rng = np.random.default_rng(5)
    x = rng.random((100, 3))
    # saturation power 1 = identity (up to the sRGB round trip)
    assert np.allclose(sty(oracle, [N.saturation(1.0)], x), x, atol=1e-12)
    # very large power drives saturation to 0 -> grey at the sRGB-encoded max
    y = sty(oracle, [N.saturation(1e6)], x)
    assert np.allclose(y[:, 0], y[:, 1], atol=1e-9) and np.allclose(y[:, 1], y[:, 2], atol=1e-9)
    assert np.allclose(y[:, 0], x.max(axis=1), rtol=1e-9)
    # depth > 0 divides by indirect_scale then rescales: a scale-2 input at depth 1 equals 2x depth 0
    y0 = sty(oracle, [N.saturation(0.5, indirect_scale=2.0)], x, depth=0)
    y1 = sty(oracle, [N.saturation(0.5, indirect_scale=2.0)], 2 * x, depth=1)
    assert np.allclose(y1, 2 * y0, rtol=1e-12)
    # rex: blend 1 -> grey at the mean; blend 0 -> normalised I_max scaled to the mean
    r1 = sty(oracle, [N.rex(1.0)], x)
    assert np.allclose(r1, np.repeat(x.mean(axis=1, keepdims=True), 3, axis=1))
    r0 = sty(oracle, [N.rex(0.0)], x)
    imax = (x - x.min(axis=1, keepdims=True)) / x.max(axis=1, keepdims=True)
    exp0 = imax * x.mean(axis=1, keepdims=True) / imax.mean(axis=1, keepdims=True)
    assert np.allclose(r0, exp0)
    # palette: teaser palette hex strings (scenes/veach-ajar-sty-v2/scene-sty.json:2686-2700)
    hexes = ["0x1a1c2c", "0xef7d57", "0xffcd75", "0xa7f070", "0x38b764", "0x257179", "0x29366f",
             "0x3b5dc9", "0x41a6f6", "0x73eff7", "0xf4f4f4", "0x94b0c2", "0x566c86", "0x333c57"]
    pal = N.palette_hex(hexes)
    assert np.allclose(pal[0], np.array([0x1A, 0x1C, 0x2C]) / 255.0)
    nodes = [N.palette(pal, gamma=1.0)]
    assert np.allclose(sty(oracle, nodes, pal.astype(np.float64)), pal, atol=1e-7)
    # brute-force nearest entry at gamma 0.5, depth 0
    nodes = [N.palette(pal, gamma=0.5)]
    out = sty(oracle, nodes, x)
    g = np.power(np.clip(x, 0, 1), 0.5)
    pal64 = pal.astype(np.float64)
    best = np.argmin(((pal64[None] - g[:, None]) ** 2).sum(-1), axis=1)
    assert np.allclose(out, np.power(pal64[best], 2.0), rtol=1e-12)
    # eight-bit: sRGB grid points (r,g on 1/8, b on 1/4) are fixed points at depth 0
    grid = srgb_decode(np.array([[3 / 8, 5 / 8, 1 / 4], [1 / 8, 7 / 8, 3 / 4], [1.0, 0.5, 0.5]]))
    assert np.allclose(sty(oracle, [N.eight_bit()], grid), grid, rtol=1e-12)


def test_enable_if_set_idx_and_disabled_flag(oracle):
    # enable_if(idx 0, toon 0.25, else enable_if(idx 2, rex(1))) like the teaser chains
    nodes = [N.toon(0.25), N.rex(1.0), N.enable_if(2, inner=1), N.enable_if(0, inner=0, else_=2)]
    c = np.array([[0.6, 0.2, 0.1]])
    t = sty(oracle, nodes, c, root=3, flags=[0])
    assert np.allclose(t, sty(oracle, [N.toon(0.25)], c))
    r = sty(oracle, nodes, c, root=3, flags=[N.FLAG_STYLE3])
    assert np.allclose(r, np.full((1, 3), c.mean()))
    # style index 1 matches neither branch -> null surface -> identity
    assert np.allclose(sty(oracle, nodes, c, root=3, flags=[N.FLAG_STYLE2]), c)
    # stylization_disabled bypasses everything (SurfaceClosure, svm/surface/mod.rs:1833)
    assert np.allclose(sty(oracle, nodes, c, root=3, flags=[N.FLAG_DISABLED]), c)
    # the wrapper applies inner first: toon(rex(x)) != rex(toon(x))
    a = sty(oracle, [N.rex(0.3), N.toon(0.25, inner=0)], c, root=1)
    b = sty(oracle, [N.toon(0.25), N.rex(0.3, inner=0)], c, root=1)
    ta = sty(oracle, [N.toon(0.25)], sty(oracle, [N.rex(0.3)], c))
    assert np.allclose(a, ta) and not np.allclose(a, b)
    # only a SetStyleIdx root rewrites flags; index 0 clears the one-hot bits, disabled kept
    tab = N.style_table([N.toon(0.2), N.set_idx(2, inner=0), N.toon(0.2, inner=1)])
    assert oracle.update_flags(tab, 1, 0) == N.FLAG_STYLE3
    assert oracle.update_flags(tab, 1, N.FLAG_DISABLED | N.FLAG_STYLE2) == N.FLAG_DISABLED | N.FLAG_STYLE3
    assert oracle.update_flags(tab, 2, N.FLAG_STYLE2) == N.FLAG_STYLE2
    tab0 = N.style_table([N.set_idx(0)])
    assert oracle.update_flags(tab0, 0, N.FLAG_STYLE4) == 0


def test_depth_mask_gates_own_transform(oracle):
    c = np.array([[0.7, 0.4, 0.2]])
    nodes = [N.toon(0.25, depth_mask=N.depth_mask(1))]
    assert np.allclose(sty(oracle, nodes, c, depth=0), c)
    assert np.allclose(sty(oracle, nodes, c, depth=1), sty(oracle, [N.toon(0.25)], c))
    assert np.allclose(sty(oracle, nodes, c, depth=2), c)


def test_halftone_nearest_grid_point(oracle):
    # identity affine, grid distance 1: the nearest lattice point of p=(0.1,0.1,0.1) is the
    # origin at distance sqrt(0.03); dark tone applies when that is below r
    aff = np.eye(3, 4)
    c = np.array([[0.2, 0.2, 0.2]])
    p = np.array([[0.1, 0.1, 0.1]])
    node = N.halftone(1.0, 1.0, 1.0, 1.0, 1.0, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], aff, max_depth=4)
    # r = (1 - lum) * d * 0.5 * radius = 0.8 * 0.5 = 0.4 > 0.1732 -> dark (black)
    assert np.allclose(sty(oracle, [node], c, pos=p), 0)
    p2 = np.array([[0.5, 0.5, 0.5]])  # distance sqrt(0.75) = 0.866 > 0.4 -> light (unchanged)
    assert np.allclose(sty(oracle, [node], c, pos=p2), c)
    # beyond max_depth the node is inactive
    assert np.allclose(sty(oracle, [node], c, pos=p, depth=5), c)
