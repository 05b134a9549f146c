"""The written scene file (nlpf_write_world_scene) describes the scene the GPU generator traces.
Every first hit must lie in its pixel's footprint under the reference's camera model
(PerspectiveCameraData::new, camera/mod.rs:139-152, restated here) applied to the written
camera. Every hit must lie on the written geometry of its material. Every vertex's reflectance
must be the written diffuse colour of its material."""
import json
import math

import numpy as np
import pytest

import nlpf as N

pytestmark = pytest.mark.gpu
CENTERS = {5: ((-0.22, -0.3, -0.18), 0.2), 6: ((0.17, -0.2, 0.08), 0.3)}
PLANES = {0: (1, -0.5), 1: (1, 0.5), 2: (2, -0.5), 3: (0, -0.5), 4: (0, 0.5), 7: (1, 0.499)}


@pytest.mark.parametrize("W,H", [(64, 48), (40, 72)])
def test_generator_traces_the_written_scene(tmp_path, W, H):
    branch, L, seed = 2, 3, 5
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch}
    path = str(tmp_path / "box.json")
    N.write_world_scene(path, W, H, seed=seed)
    with open(path) as f:
        sc = json.load(f)
    with N.Handle(cfg) as h:
        desc = h.world_desc(W, H, branch, L)
        bufs = h.alloc_world(desc)
        h.generate_world_records(desc, bufs, seed=seed)
        host = h.download_world(desc, bufs)
    n = W * H * branch
    rec = host["records"].reshape(L, 28, n)
    mat = host["material"].reshape(L, n)
    pd = host["path_depth"]
    hit = pd >= 0
    assert hit.mean() > 0.5

    # camera: raster position of each first hit under the reference's model
    cam = sc["camera"]["data"]
    o = np.array(cam["transform"]["data"]["translation"], np.float64)
    assert cam["transform"]["data"]["rotation"] == [0, 0, 0]
    s = math.tan(math.radians(cam["fov"]) / 2)
    sx, sy = (s, s * H / W) if W > H else (s * W / H, s)
    d = rec[0, 0:3].T.astype(np.float64) - o
    xc, yc = d[:, 0] / -d[:, 2], d[:, 1] / -d[:, 2]  # camera looks down -z (r2c puts z = -1)
    rx = (xc / sx + 1) * W / 2
    ry = (1 - yc / sy) * H / 2
    pix = np.arange(n) % (W * H)
    px, py = pix % W, pix // W
    eps = 1e-3
    assert (rx[hit] >= px[hit] - eps).all() and (rx[hit] <= px[hit] + 1 + eps).all()
    assert (ry[hit] >= py[hit] - eps).all() and (ry[hit] <= py[hit] + 1 + eps).all()

    names = sorted(sc["materials"])
    for dd in range(L):
        on = pd >= dd
        p = rec[dd, 0:3].T.astype(np.float64)
        for m in range(8):
            sel = on & (mat[dd] == m)
            if not sel.any():
                continue
            if m in PLANES:
                axis, c = PLANES[m]
                assert np.abs(p[sel, axis] - c).max() < 1e-5, m
            else:
                c, r = CENTERS[m]
                assert np.abs(np.linalg.norm(p[sel] - c, axis=1) - r).max() < 1e-5, m
            nodes = sc["materials"][names[m]]["shader"]["nodes"]
            alb = rec[dd, 12:15].T[sel]
            if m == 7:
                assert nodes["bsdf"]["type"] == "emission" and (alb == 0).all()
                assert nodes["estr"]["value"] == 15.0
            else:
                want = np.array(nodes["albedo"]["value"], np.float32)
                assert np.abs(alb - want).max() <= 2 * np.spacing(np.float32(1.0)), m
