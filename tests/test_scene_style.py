"""Style tables imported from reference scene files (nlpf_style_from_scene, host only).

The fixture tests/golden/veach_ajar_sty_materials.json holds the materials of the reference's
stylized scene scenes/veach-ajar-sty-v2/scene-sty.json, unchanged. Every material's imported
tree is compared with an independent Python lowering written here from shader.rs:195-263 and
svm/compiler.rs:337-477. Known values are taken by reading the fixture. Synthetic graphs cover
the node kinds and the error paths that the scene does not use."""
import json
import math
import os

import numpy as np
import pytest

import nlpf as N

f32 = np.float32
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "veach_ajar_sty_materials.json")
NAMES = {v: k for k, v in N.STY.items()}


def lower_py(nodes, nid):
    """Nested-tuple form of the stylization chain below node nid (None = identity)."""
    n = nodes[nid]
    t = n["type"]
    ref = lambda f: n[f]["id"]  # noqa: E731
    val = lambda f: float(f32(nodes[ref(f)]["value"]))  # noqa: E731
    if t == "output":
        return lower_py(nodes, ref("node"))
    if not t.startswith("sty_"):
        return None
    inner = lower_py(nodes, ref("surface"))
    if t == "sty_toon":
        return ("TOON", val("color_band_width"), inner)
    if t == "sty_color_palette":
        cols = []
        for e in n["palette"]:
            x = int(e[2:] if e.startswith("0x") else e, 16)
            cols.append(tuple(float(f32((x >> s) & 0xFF) / f32(255)) for s in (16, 8, 0)))
        return ("PALETTE", val("indirect_scale"), val("gamma"), tuple(cols), inner)
    if t == "sty_saturation_adjust_rex":
        return ("REX", val("blend"), inner)
    if t == "sty_saturation_adjust":
        return ("SATURATION", val("power"), val("indirect_scale"), n["max_depth"], inner)
    if t == "sty_8bit":
        return ("EIGHT_BIT", val("indirect_scale"), inner)
    if t == "sty_hue_shift":
        return ("HUE_SHIFT", val("shift"), val("indirect_scale"), inner)
    if t == "sty_enable_if":
        e = n.get("else")
        return ("ENABLE_IF", n["idx"], inner, lower_py(nodes, e["id"]) if e else None)
    if t == "sty_set_idx":
        return ("SET_IDX", n["idx"], inner)
    raise ValueError(t)


def tree(table, i):
    """Nested-tuple form of the library's table below index i."""
    if i < 0:
        return None
    nd = table[i]
    t = NAMES[int(nd["type"])]
    p = [float(v) for v in nd["p"]]
    inner = tree(table, int(nd["inner"]))
    if t == "TOON":
        return (t, p[0], inner)
    if t == "PALETTE":
        k = int(nd["count"])
        return (t, p[0], p[1], tuple(tuple(p[2 + 3 * j:5 + 3 * j]) for j in range(k)), inner)
    if t == "REX":
        return (t, p[0], inner)
    if t == "SATURATION":
        return (t, p[0], p[1], int(nd["max_depth"]), inner)
    if t == "EIGHT_BIT":
        return (t, p[0], inner)
    if t == "HUE_SHIFT":
        return (t, p[0], p[1], inner)
    if t == "ENABLE_IF":
        return (t, int(nd["idx"]), inner, tree(table, int(nd["else_"])))
    if t == "SET_IDX":
        return (t, int(nd["idx"]), inner)
    if t == "HALFTONE":
        return (t, tuple(p[:23]), int(nd["max_depth"]), inner)
    raise ValueError(t)


@pytest.fixture(scope="module")
def scene():
    with open(FIXTURE) as f:
        return json.load(f)


def test_every_material_matches_independent_lowering(scene):
    nodes, roots, mats = N.style_from_scene(FIXTURE)
    assert mats == sorted(scene["materials"], key=str.encode) and len(mats) == 33
    for m, r in zip(mats, roots):
        g = scene["materials"][m]["shader"]
        assert tree(nodes, int(r)) == lower_py(g["nodes"], g["output"]["id"]), m
    # shared subtrees: the 31 stylized materials need 11 nodes
    assert len(nodes) == 11
    assert all(nd["depth_mask"] == N.ALL_DEPTHS for nd in nodes)


def test_known_materials(scene):
    nodes, roots, mats = N.style_from_scene(FIXTURE)
    r = dict(zip(mats, (int(x) for x in roots)))
    # Dragon: rex 0.5 for style index 0, else toon 0.2 (over the null surface) for index 2
    assert tree(nodes, r["Dragon"]) == ("ENABLE_IF", 0, ("REX", 0.5, None),
                                        ("ENABLE_IF", 2, ("TOON", float(f32(0.2)), None), None))
    # the lights carry no stylization; the mirrors only set the style index of their subtree
    assert r["MiniLightLeft"] == r["MiniLightRight"] == -1
    assert tree(nodes, r["Mirror"]) == ("SET_IDX", 1, None)
    assert tree(nodes, r["Mirror2"]) == ("SET_IDX", 2, None)
    # Floor: a 14-entry palette (gamma 0.5, indirect scale 1) for index 0, toon otherwise
    fl = tree(nodes, r["Floor"])
    assert fl[0] == "ENABLE_IF" and fl[2][0] == "ENABLE_IF" and fl[2][2][0] == "PALETTE"
    pal = fl[2][2]
    assert (pal[1], pal[2], len(pal[3])) == (1.0, 0.5, 14)
    assert pal[3][0] == tuple(float(f32(v) / f32(255)) for v in (0x1a, 0x1c, 0x2c))
    # Texture_base: a 17-entry palette for style index 1
    tb = tree(nodes, r["Texture_base"])
    assert tb[1] == 1 and len(tb[2][3]) == 17
    # a subset in the caller's order
    n2, r2, m2 = N.style_from_scene(FIXTURE, ["Mirror2", "Dragon"])
    assert tree(n2, int(r2[0])) == ("SET_IDX", 2, None) and tree(n2, int(r2[1])) == tree(nodes, r["Dragon"])


def graph(*entries, out="o"):
    nodes = {"o": {"type": "output", "node": {"id": out if out != "o" else entries[-1][0]}}}
    for nid, n in entries:
        nodes[nid] = n
    return {"materials": {"m": {"shader": {"nodes": nodes, "output": {"id": "o"}, "kind": "surface"}}}}


def ref(i):
    return {"id": i}


def flt(v):
    return {"type": "float", "value": v}


def test_all_node_kinds_match_builders():
    g = graph(("base", {"type": "diffuse", "color": ref("c")}), ("c", {"type": "rgb", "value": [1, 1, 1]}),
              ("a", flt(0.7)), ("b", flt(2.0)), ("w", flt(0.3)),
              ("s1", {"type": "sty_saturation_adjust", "surface": ref("base"), "power": ref("a"),
                      "indirect_scale": ref("b"), "max_depth": 3}),
              ("s2", {"type": "sty_8bit", "surface": ref("s1"), "indirect_scale": ref("b")}),
              ("s3", {"type": "sty_hue_shift", "surface": ref("s2"), "shift": ref("w"), "indirect_scale": ref("b")}),
              ("s4", {"type": "sty_toon", "surface": ref("s3"), "color_band_width": ref("w")}))
    nodes, roots, _ = N.style_from_scene(json.dumps(g))
    want = N.style_table([N.saturation(0.7, indirect_scale=2.0, max_depth=3), N.eight_bit(2.0, inner=0),
                          N.hue_shift(0.3, inner=1, indirect_scale=2.0), N.toon(0.3, inner=2)])
    assert roots.tolist() == [3]
    assert nodes.tobytes() == want.tobytes()


def rot(axis, a):
    c, s = math.cos(a), math.sin(a)
    m = np.eye(4)
    i, j = (axis + 1) % 3, (axis + 2) % 3
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


@pytest.mark.parametrize("system", ["Akari", "Blender"])
def test_half_tone_trs(system):
    t, r, s = [0.5, -1.0, 2.0], [0.3, -0.7, 1.1], [2.0, 0.5, 1.5]
    g = graph(("k", flt(1.0)), ("rs", flt(0.8)), ("rad", flt(1.2)), ("gd", flt(0.15)), ("bl", flt(0.7)),
              ("dk", {"type": "spectral_uplift", "rgb": ref("dkc")}), ("dkc", {"type": "rgb", "value": [0.1, 0.1, 0.2]}),
              ("lt", {"type": "float3", "value": [1.0, 0.95, 0.9]}),
              ("h", {"type": "sty_half_tone", "surface": ref("k"), "trs": {"translation": t, "rotation": r, "scale": s,
                     "coordinate_system": system}, "radius_scale": ref("rs"), "radius": ref("rad"),
                     "indirect_scale": ref("k"), "grid_distance": ref("gd"), "max_depth": 2, "blend": ref("bl"),
                     "light_tone": ref("lt"), "dark_tone": ref("dk")}))
    nodes, roots, _ = N.style_from_scene(json.dumps(g))
    # TRS::to_matrix(false), akari_scenegraph/src/scene.rs:31-61, written as matrix products
    S = np.diag([*s, 1.0])
    if system == "Akari":
        T = np.eye(4)
        T[:3, 3] = t
        M = T @ rot(1, r[1]) @ rot(0, r[0]) @ rot(2, r[2]) @ S
    else:
        T = np.eye(4)
        T[:3, 3] = [t[0], t[2], -t[1]]
        M = T @ rot(1, r[2]) @ rot(2, -r[1]) @ rot(0, r[0]) @ S
    want = N.halftone(1.0, 0.8, 1.2, 0.15, 0.7, [0.1, 0.1, 0.2], [1.0, 0.95, 0.9], M[:3], max_depth=2)
    got = nodes[int(roots[0])]
    assert got["type"] == N.STY["HALFTONE"] and got["max_depth"] == 2
    assert np.allclose(got["p"][:23], want["p"][:23], rtol=0, atol=2e-7)
    # a point check that pins the rotation order: the x axis goes where the matrix sends it
    assert not np.allclose(M, T @ rot(0, r[0]) @ rot(1, r[1]) @ rot(2, r[2]) @ S)


def test_errors():
    base = ("base", {"type": "diffuse", "color": ref("c")})
    col = ("c", {"type": "rgb", "value": [1, 1, 1]})
    cases = [
        (graph(base, col, ("tex", {"type": "image", "image": {}}),
               ("t", {"type": "sty_toon", "surface": ref("base"), "color_band_width": ref("tex")})),
         "constant float"),
        (graph(base, col, ("w", flt(0.2)), ("t", {"type": "sty_toon", "surface": ref("base"), "color_band_width": ref("w")}),
               ("mx", {"type": "mix", "first": ref("t"), "second": ref("base"), "factor": ref("w")})),
         "below it"),
        (graph(("e", {"type": "sty_enable_if", "surface": ref("e2"), "idx": 0}),
               ("e2", {"type": "sty_enable_if", "surface": ref("e"), "idx": 0}), out="e"), "deeper than 16"),
        (graph(base, col, ("e", {"type": "sty_set_idx", "surface": ref("base"), "idx": 4})), "0..3"),
        (graph(base, col, ("w", flt(0.0)), ("t", {"type": "sty_toon", "surface": ref("base"), "color_band_width": ref("w")})),
         "> 0"),
        (graph(base, col, ("z", {"type": "sty_bogus", "surface": ref("base")})), "unknown stylization"),
    ]
    for g, msg in cases:
        with pytest.raises(N.NlpfError) as e:
            N.style_from_scene(json.dumps(g))
        assert msg in str(e.value), (msg, str(e.value))
    with pytest.raises(N.NlpfError, match="no shader graph"):
        N.style_from_scene(FIXTURE, ["NoSuchMaterial"])
    with pytest.raises(N.NlpfError, match="capacity"):
        N.style_from_scene(FIXTURE, capacity=3)
    with pytest.raises(N.NlpfError, match="scene JSON"):
        N.style_from_scene("{", ["m"])
    # a hostile nesting depth is an error, not a stack overflow
    with pytest.raises(N.NlpfError, match="nested deeper"):
        N.style_from_scene('{"materials":' + "[" * 100000 + "]" * 100000 + "}", ["m"])
