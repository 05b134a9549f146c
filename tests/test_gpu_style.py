"""GPU parity of every stylization node (nlpf_stylize through the C ABI) against the FP64
oracle; tolerance 1e-5 relative per pixel, with inputs kept away from the quantisers'
decision boundaries (toon bands, 8-bit rounding, palette ties) where one FP32 ulp flips the
result."""
import numpy as np
import pytest

import nlpf as N
import oracle as O
from gpu_util import pixel_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    with N.Handle({}) as hh:
        yield hh


def srgb_enc(x):
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(np.maximum(x, 0), 1 / 2.4) - 0.055)


def run_both(h, nodes, rgb, depth=0, flags=None, root=None, pos=None, material=None, roots=None):
    table = N.style_table(nodes)
    if roots is None:
        roots = [len(nodes) - 1 if root is None else root]
    h.set_style(table, roots)
    g = h.stylize(rgb, depth=depth, flags=flags, material=material, pos=pos)
    o = O.stylize(table, roots, rgb.astype(np.float32).astype(np.float64), depth=depth, flags=flags,
                  material=material, pos=None if pos is None else pos.astype(np.float32).astype(np.float64))
    return g, o


def test_toon(h):
    rng = np.random.default_rng(0)
    x = (rng.random((20000, 3)) * rng.choice([0.3, 1.0, 4.0], (20000, 1))).astype(np.float32)
    w = 0.25
    lum = x.astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    q = np.where(lum > 1, lum / w, srgb_enc(lum) / w)
    keep = np.abs(q - np.round(q)) > 1e-4
    g, o = run_both(h, [N.toon(w)], x)
    assert pixel_rel_err(g[keep], o[keep]).max() <= 1e-5


@pytest.mark.parametrize("depth", [0, 2])
def test_ramp_hue_saturation_rex(h, depth):
    rng = np.random.default_rng(1 + depth)
    x = (rng.random((20000, 3)) * 3).astype(np.float32)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    for nodes in ([N.ramp(pos, val)], [N.hue_shift(0.1)], [N.hue_shift(0.77)],
                  [N.saturation(0.6, indirect_scale=2.0)], [N.rex(0.35)],
                  [N.ramp(pos, val), N.hue_shift(0.1, inner=0)],
                  [N.rex(0.5), N.saturation(1.7, inner=0), N.hue_shift(0.3, inner=1)]):
        g, o = run_both(h, nodes, x, depth=depth)
        assert pixel_rel_err(g, o).max() <= 1e-5, nodes[-1]["type"]


def test_palette_and_eight_bit(h):
    rng = np.random.default_rng(3)
    x = (rng.random((20000, 3)) * 1.5).astype(np.float32)
    pal = N.palette_hex(["0x1a1c2c", "0xef7d57", "0xffcd75", "0xa7f070", "0x38b764", "0x257179",
                         "0x29366f", "0x3b5dc9", "0x41a6f6", "0x73eff7", "0xf4f4f4", "0x94b0c2"])
    for depth in (0, 1):
        g, o = run_both(h, [N.palette(pal, gamma=0.5, indirect_scale=2.0)], x, depth=depth)
        # drop near-ties between the two closest palette entries
        c = np.power(np.clip(x.astype(np.float64) / (2.0 if depth else 1.0), 0, 1), 0.5)
        d = np.sort(((pal.astype(np.float64)[None] - c[:, None]) ** 2).sum(-1), axis=1)
        keep = (d[:, 1] - d[:, 0]) > 1e-5
        assert pixel_rel_err(g[keep], o[keep]).max() <= 1e-5
        g, o = run_both(h, [N.eight_bit(indirect_scale=2.0)], x, depth=depth)
        s = srgb_enc(np.clip(x.astype(np.float64) / (2.0 if depth else np.maximum(1, x.max(1, keepdims=True))), 0, None))
        qs = s * np.array([8, 8, 4])
        keep = (np.abs(qs - np.floor(qs) - 0.5) > 1e-4).all(1)
        assert pixel_rel_err(g[keep], o[keep]).max() <= 1e-5


def test_halftone(h):
    rng = np.random.default_rng(4)
    x = (rng.random((5000, 3)) * 0.8).astype(np.float32)
    p = (rng.random((5000, 3)) * 2 - 1).astype(np.float32)
    aff = np.array([[1, 0.2, 0, 0.1], [0, 1, 0.3, 0], [0.1, 0, 1, -0.2]], np.float32)
    node = N.halftone(1.0, 0.8, 1.2, 0.15, 0.7, [0.1, 0.1, 0.2], [1.0, 0.95, 0.9], aff, max_depth=3)
    g, o = run_both(h, [node], x, depth=0, pos=p)
    err = pixel_rel_err(g, o)
    assert (err <= 1e-5).mean() > 0.999  # a dot-radius tie flips a tone; allow a handful
    g, o = run_both(h, [node], x, depth=4, pos=p)  # beyond max_depth: identity
    assert np.array_equal(g, x)


def test_flags_materials_and_branches(h):
    rng = np.random.default_rng(5)
    n = 4000
    x = (rng.random((n, 3)) * 2).astype(np.float32)
    nodes = [N.rex(0.6), N.enable_if(2, inner=0), N.saturation(0.5), N.enable_if(0, inner=2, else_=1),
             N.hue_shift(0.25), N.set_idx(2, inner=4)]
    flags = rng.choice([0, N.FLAG_STYLE2, N.FLAG_STYLE3, N.FLAG_STYLE4, N.FLAG_DISABLED], n).astype(np.uint8)
    material = rng.integers(0, 3, n).astype(np.uint16)
    depth = rng.integers(0, 4, n).astype(np.uint32)
    g, o = run_both(h, nodes, x, depth=depth, flags=flags, material=material, roots=[3, 5, -1])
    assert pixel_rel_err(g, o).max() <= 1e-5
    dis = flags == N.FLAG_DISABLED
    assert np.array_equal(g[dis], x[dis])
    assert np.array_equal(g[material == 2], x[material == 2])


def test_style_table_validation(h):
    with pytest.raises(N.NlpfError):
        h.set_style(N.style_table([N.toon(0.2, inner=5)]), [0])
    with pytest.raises(N.NlpfError):
        h.set_style(N.style_table([N.toon(0.2, inner=0)]), [0])  # cycle
    with pytest.raises(N.NlpfError):
        h.set_style(N.style_table([N.ramp([0.5, 0.2], [0.1, 0.2])]), [0])  # descending stops
    with pytest.raises(N.NlpfError):
        h.set_style(N.style_table([N.toon(0.0)]), [0])


def test_reference_scene_style_table(h):
    """The style table imported from the reference's stylized scene (veach-ajar-sty-v2), over all
    33 materials, random style flags and depths. Outputs within 1e-5 of the oracle, except where a
    quantiser (toon band, palette entry) flips under a 4-ulp change of the input, which is then
    checked against the oracle at that perturbed input."""
    import os
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "veach_ajar_sty_materials.json")
    nodes, roots, mats = N.style_from_scene(fixture)
    rng = np.random.default_rng(7)
    n = 20000
    x = (rng.random((n, 3)) * rng.choice([0.2, 1.0, 3.0], (n, 1))).astype(np.float32)
    flags = rng.choice([0, N.FLAG_STYLE2, N.FLAG_STYLE3, N.FLAG_STYLE4, N.FLAG_STYLE2 | N.FLAG_STYLE4],
                       n).astype(np.uint8)
    material = rng.integers(0, len(mats), n).astype(np.uint16)
    depth = rng.integers(0, 4, n).astype(np.uint32)
    h.set_style(nodes, roots)
    g = h.stylize(x, depth=depth, flags=flags, material=material)
    xo = x.astype(np.float64)
    o = O.stylize(nodes, roots, xo, depth=depth, flags=flags, material=material)
    bad = pixel_rel_err(g, o) > 1e-5
    assert bad.mean() < 0.01
    idx = np.flatnonzero(bad)
    explained = np.zeros(len(idx), bool)
    for e in (4.0, -4.0):
        xp = (x[idx] * np.float32(1 + e * 2.0 ** -24)).astype(np.float64)
        op = O.stylize(nodes, roots, xp, depth=depth[idx], flags=flags[idx], material=material[idx])
        explained |= pixel_rel_err(g[idx], op) <= 1e-5
    assert explained.all(), idx[~explained][:5]
    # materials without stylization pass through
    plain = np.isin(material, [i for i, r in enumerate(roots) if r < 0])
    assert plain.any() and np.array_equal(g[plain], x[plain])
