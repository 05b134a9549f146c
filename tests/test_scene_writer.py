"""The procedural scene written as a reference scene file (nlpf_write_world_scene, host only).
The JSON and its buffer are read back here the way the reference's loader reads them:
vertices [f32; 3], indices [u32; 3], normals [f32; 3], per-triangle u32 material slots
(akari_render/src/load.rs:478-505), buffer paths relative to the JSON (scene.rs:641-690).
The geometry is checked against the analytic scene, the style table against a round trip
through nlpf_style_from_scene, and the camera's field of view against its closed form."""
import json
import math
import os

import numpy as np
import pytest

import nlpf as N
from test_scene_style import tree

f32 = np.float32
CENTERS = {"sphere0": ((-0.22, -0.3, -0.18), 0.2), "sphere1": ((0.17, -0.2, 0.08), 0.3)}


def load_scene(path):
    with open(path) as f:
        sc = json.load(f)
    base = os.path.dirname(path)
    blobs = {}
    for bid, b in sc["buffers"].items():
        assert b["type"] == "path"
        data = open(os.path.join(base, b["path"]), "rb").read()
        assert len(data) == b["length"]
        blobs[bid] = data

    def view(vid, dtype, k):
        v = sc["buffer_views"][vid]
        data = blobs[v["buffer"]["id"]]
        assert v["offset"] % 4 == 0 and v["offset"] + v["length"] <= len(data)
        a = np.frombuffer(data, dtype, v["length"] // np.dtype(dtype).itemsize, v["offset"])
        return a.reshape(-1, k) if k > 1 else a

    meshes = {}
    for gid, g in sc["geometries"].items():
        assert g["type"] == "mesh" and g["uvs"] is None and g["tangents"] is None
        meshes[gid] = dict(v=view(g["vertices"]["id"], np.float32, 3), n=view(g["normals"]["id"], np.float32, 3),
                           i=view(g["indices"]["id"], np.uint32, 3), m=view(g["materials"]["id"], np.uint32, 1))
    return sc, meshes


def style_nodes():
    aff = np.zeros((3, 4), np.float32)
    c, s = math.cos(0.4), math.sin(0.4)
    aff[:3, :3] = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]) @ np.diag([2.0, 1.0, 0.5])
    aff[:, 3] = [0.1, -0.2, 0.3]
    pal = N.palette_hex(["0x1a1c2c", "0xef7d57", "0xffcd75", "0xa7f070"])
    return N.style_table([
        N.toon(0.2),                                              # 0
        N.palette(pal, gamma=0.5, indirect_scale=2.0),            # 1
        N.enable_if(0, inner=1, else_=0),                         # 2
        N.rex(0.4, inner=2),                                      # 3
        N.saturation(0.7, indirect_scale=1.5, max_depth=2),       # 4
        N.eight_bit(2.0, inner=4),                                # 5
        N.halftone(1.0, 0.8, 1.2, 0.15, 0.7, [0.1, 0.1, 0.2], [1.0, 0.95, 0.9], aff, max_depth=3),  # 6
        N.set_idx(2, inner=6),                                    # 7
    ]), [3, 3, 5, 3, -1, 7, 2, 0]


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    d = tmp_path_factory.mktemp("scene")
    nodes, roots = style_nodes()
    path = str(d / "sub" / "box.json")
    N.write_world_scene(path, 64, 48, seed=5, nodes=nodes, roots=roots, sphere_segments=64)
    return path, nodes, roots


def test_structure_and_walls(written):
    path, _, _ = written
    sc, meshes = load_scene(path)
    assert os.path.exists(os.path.join(os.path.dirname(path), "box.bin"))
    assert sorted(sc["materials"]) == [f"mat{m}_{n}" for m, n in enumerate(N.SCENE_MATERIALS)]
    assert sc["lights"] == {}
    for iid, inst in sc["instances"].items():
        g = meshes[inst["geometry"]["id"]]
        assert len(inst["materials"]) == 1 and (g["m"] == 0).all() and len(g["m"]) == len(g["i"])
        assert g["i"].max() < len(g["v"])
    walls = {"floor": (1, -0.5, (0, 1, 0)), "ceiling": (1, 0.5, (0, -1, 0)), "back": (2, -0.5, (0, 0, 1)),
             "left": (0, -0.5, (1, 0, 0)), "right": (0, 0.5, (-1, 0, 0)), "light": (1, 0.499, (0, -1, 0))}
    for name, (axis, c, nrm) in walls.items():
        m = N.SCENE_MATERIALS.index(name)
        g = meshes[f"geom{m}_{name}"]
        assert len(g["v"]) == 4 and len(g["i"]) == 2
        assert np.all(g["v"][:, axis] == f32(c))
        ext = 0.12 if name == "light" else 0.5
        others = [a for a in range(3) if a != axis]
        assert np.array_equal(np.sort(np.abs(g["v"][:, others]).ravel()), np.full(8, f32(ext)))
        assert np.all(g["n"] == np.array(nrm, f32))
        v = g["v"].astype(np.float64)
        for t in g["i"]:
            cr = np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])
            assert np.dot(cr, nrm) > 0  # winding gives the facing (the light emits downwards)


def test_spheres(written):
    path, _, _ = written
    _, meshes = load_scene(path)
    for name, (c, r) in CENTERS.items():
        m = N.SCENE_MATERIALS.index(name)
        g = meshes[f"geom{m}_{name}"]
        v = g["v"].astype(np.float64) - c
        d = np.linalg.norm(v, axis=1)
        assert np.abs(d - r).max() < 1e-6
        assert np.abs(g["n"] - v / d[:, None]).max() < 1e-6
        assert len(g["v"]) == 2 + 31 * 64 and len(g["i"]) == 2 * 64 * 31
        tri = g["v"][g["i"]].astype(np.float64) - c
        cr = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        assert (np.einsum("ij,ij->i", cr, tri.mean(1)) > 0).all()  # outward winding
        # tessellation error of the chords: r (1 - cos(pi / 64)) at most
        assert (np.linalg.norm(tri.mean(1), axis=1) > r * math.cos(math.pi / 32)).all()


def test_style_round_trip(written):
    path, nodes, roots = written
    back, broots, mats = N.style_from_scene(path)
    assert mats == sorted(mats)
    for m in range(len(N.SCENE_MATERIALS)):
        want = tree(nodes, roots[m]) if m < len(roots) else None
        got = tree(back, int(broots[m]))
        if want is not None and want[0] == "HALFTONE":
            assert got[0] == "HALFTONE" and got[2:] == want[2:]
            assert np.allclose(got[1], want[1], atol=2e-6)
        elif m == 5:  # set_idx over the half tone
            assert got[0] == "SET_IDX" and got[2][0] == "HALFTONE"
            assert np.allclose(got[2][1], want[2][1], atol=2e-6) and got[1] == want[1]
        else:
            assert got == want, N.SCENE_MATERIALS[m]


def test_materials_and_camera(written):
    path, _, _ = written
    sc, _ = load_scene(path)
    light = sc["materials"]["mat7_light"]["shader"]["nodes"]
    assert light["bsdf"]["type"] == "emission" and light["estr"]["value"] == 15.0
    floor = sc["materials"]["mat0_floor"]["shader"]["nodes"]
    alb = np.array(floor["albedo"]["value"])
    assert floor["bsdf"]["type"] == "diffuse" and ((alb >= 0.2) & (alb < 0.8)).all()
    cam = sc["camera"]["data"]
    assert cam["sensor_width"] == 64 and cam["sensor_height"] == 48 and not cam["depth_of_field"]
    assert np.array_equal(np.array(cam["transform"]["data"]["translation"], f32), np.array([0, 0, 1.6], f32))
    # the reference scales the longer side by tan(fov / 2); the generator's vertical half
    # extent is tan(22.5 deg), so the horizontal one is that times 64 / 48
    assert math.isclose(math.tan(math.radians(cam["fov"]) / 2), math.tan(math.pi / 8) * 64 / 48, rel_tol=1e-6)
    N.write_world_scene(os.path.join(os.path.dirname(path), "tall.json"), 32, 64)
    with open(os.path.join(os.path.dirname(path), "tall.json")) as f:
        assert math.isclose(json.load(f)["camera"]["data"]["fov"], 45.0, rel_tol=1e-6)


def test_unsupported_tables(tmp_path):
    p = str(tmp_path / "x.json")
    for nodes, msg in ((N.style_table([N.ramp([0.0, 1.0], [0.0, 1.0])]), "no compiled counterpart"),
                       (N.style_table([N.hue_shift(0.1)]), "no compiled counterpart"),
                       (N.style_table([N.toon(0.2, depth_mask=N.depth_mask(1))]), "depth gating"),
                       (N.style_table([N.palette([[0.3, 0.3, 0.3]])]), "1/255")):
        with pytest.raises(N.NlpfError) as e:
            N.write_world_scene(p, 8, 8, nodes=nodes, roots=[0])
        assert e.value.code == 5 and msg in str(e.value)
    shear = np.array([[1, 0.3, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
    with pytest.raises(N.NlpfError, match="not translation"):
        N.write_world_scene(p, 8, 8, nodes=N.style_table([N.halftone(1, 1, 1, 0.1, 0.5, [0] * 3, [1] * 3, shear)]),
                            roots=[0])
    with pytest.raises(N.NlpfError, match="sphere_segments"):
        N.write_world_scene(p, 8, 8, sphere_segments=7)


def test_bsdf_table_round_trip(tmp_path):
    """principled and emission-only entries written as the reference's nodes read back as the
    same table (diffuse where the table has no entry); the principled node carries every input
    of the reference's node"""
    table = N.bsdf_table([
        {"kind": N.BSDF_PRINCIPLED, "metallic": 0.7, "roughness": 0.35},
        {},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.3, "specular_tint": (1.0, 0.8, 0.6),
         "specular_ior_level": 0.5, "transmission": 1.0},
        {"kind": N.BSDF_EMISSION},
    ])
    path = str(tmp_path / "box.json")
    N.write_world_scene(path, 16, 12, bsdfs=table, sphere_segments=8)
    names = [f"mat{m}_{n}" for m, n in enumerate(N.SCENE_MATERIALS)]
    back = N.bsdfs_from_scene(path, names)
    assert back[0] == table[0] and back[2] == table[2]
    assert list(back["kind"]) == [N.BSDF_PRINCIPLED, N.BSDF_DIFFUSE, N.BSDF_PRINCIPLED, N.BSDF_EMISSION] + \
        [N.BSDF_DIFFUSE] * 3 + [N.BSDF_EMISSION]
    with open(path) as f:
        sc = json.load(f)
    node = next(v for v in sc["materials"][names[0]]["shader"]["nodes"].values() if v["type"] == "principled")
    want = {"alpha", "anisotropic", "anisotropic_rotation", "base_color", "coat_ior", "coat_normal", "coat_roughness",
            "coat_tint", "coat_weight", "emission_color", "emission_strength", "ior", "metallic", "normal",
            "preference", "roughness", "sheen_tint", "sheen_weight", "specular_ior_level", "specular_tint",
            "subsurface_anisotropy", "subsurface_radius", "subsurface_scale", "subsurface_weight", "tangent",
            "transmission_weight", "type"}
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "veach_ajar_sty_materials.json")))
    gnode = next(v for v in golden["materials"]["Gold"]["shader"]["nodes"].values() if v["type"] == "principled")
    assert set(node) == set(gnode) == want
    with pytest.raises(N.NlpfError):
        N.write_world_scene(path, 8, 8, bsdfs=N.bsdf_table([{}] * 7 + [{"kind": N.BSDF_PRINCIPLED}]))
