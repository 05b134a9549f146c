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


def _first_sphere_hit(o, d, tmax):
    """True where the segment o + t d, 0 < t < tmax, meets one of the two spheres."""
    hit = np.zeros(len(o), bool)
    for c, r in CENTERS.values():
        oc = o - np.array(c)
        b = (oc * d).sum(1)
        cc = (oc * oc).sum(1) - r * r
        disc = b * b - cc
        sq = np.sqrt(np.maximum(disc, 0))
        t0, t1 = -b - sq, -b + sq
        hit |= (disc > 0) & (((t0 > 1e-4) & (t0 < tmax)) | ((t1 > 1e-4) & (t1 < tmax)))
    return hit


def test_generator_light_transport_closed_forms():
    """The tracer's light transport, from the stored geometry and the reference's formulas
    (lm_path_filtering.rs:1062-1218 kernel_mk_trace, :1260-1316 sample_light, :388-426
    handle_surface_light, light/area.rs:50-131 AreaLight):
      direct.wi      unit vector towards a point of the light quad, not blocked by a sphere;
      direct.pdf     dist^2 / (light area x |cos| at the light);
      direct.irr.    the light's emission (15 on its lower face, 0 behind it) stylized by the
                     light material's tree at depth d+1 with the vertex's new flags;
      direct.mis     1 when a sample is stored (all direct planes 0 otherwise);
      le, mis        15 on light hits from below, 0 elsewhere; the balance heuristic of the
                     previous bsdf pdf against the light pdf for light hits at depth > 0, else 1;
      wo             towards the previous vertex; bsdf_pdf |cos(ns, wi_next)| / pi of the
                     direction to the next vertex."""
    W, H, branch, L, seed = 64, 48, 2, 4, 11
    nodes = N.style_table([
        N.toon(0.4, depth_mask=N.depth_mask(2)),  # 0: light (15 -> 14.8), stylized at depth 2 only
        N.saturation(0.5, inner=-1),       # 1
        N.enable_if(2, inner=0, else_=-1), # 2: light root: toon only for style index 2
        N.set_idx(2, inner=1),             # 3: sphere0 root sets style index 2
    ])
    roots = [1, 1, 1, 1, 1, 3, 1, 2]
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch}
    with N.Handle(cfg) as h:
        h.set_style(nodes, roots)
        desc = h.world_desc(W, H, branch, L)
        bufs = h.alloc_world(desc)
        h.generate_world_records(desc, bufs, seed=seed)
        host = h.download_world(desc, bufs)
    check_light_transport(host, nodes, roots, W, H, branch, L)


def _ray_origin(rec, d, sel, target):
    """The tracer's ray origin off vertex d towards target: 1e-4 along the geometric normal
    turned to the target's side (offset_ray_origin / face_forward in the reference)."""
    p, ng = rec[d, 0:3].T[sel], rec[d, 6:9].T[sel]
    s = np.where((ng * (target - p)).sum(1) >= 0, 1.0, -1.0)
    return p + ng * s[:, None] * 1e-4


def check_light_transport(host, nodes, roots, W, H, branch, L, bsdfs=None):
    """bsdfs: the table the records were traced with; its principled materials' continuation
    pdf is the mixture pdf of the oracle (oracle_bsdf.c) at the traced direction."""
    import oracle as O
    n = W * H * branch
    rec = host["records"].reshape(L, 28, n).astype(np.float64)
    mat = host["material"].reshape(L, n)
    nfl = host["new_flags"].reshape(L, n)
    pd = host["path_depth"]
    LY, LH, AREA = 0.499, 0.12, (2 * 0.12) ** 2
    n_direct = n_occl = 0
    for d in range(L):
        on = pd >= d
        p = rec[d, 0:3].T
        dirp = rec[d, 19:28].T
        # ---- next-event estimation records ----
        cand = on & (d + 1 < L) & (mat[d] != 7)
        stored = cand & (dirp[:, 7] == 1)
        empty = cand & (dirp[:, 7] == 0)
        assert (dirp[empty] == 0)[:, :8].all()
        assert (dirp[on & ~cand][:, :8] == 0).all()
        assert set(np.unique(dirp[cand, 7])) <= {0.0, 1.0}
        n_direct += int(stored.sum())
        n_occl += int(empty.sum())
        wi = dirp[stored, 0:3]
        ps = p[stored]
        assert (np.abs(np.linalg.norm(wi, axis=1) - 1) < 1e-5).all()
        t = (LY - ps[:, 1]) / wi[:, 1]
        q = ps + t[:, None] * wi
        assert (t > 0).all()
        assert (np.abs(q[:, 0]) <= LH + 1e-4).all() and (np.abs(q[:, 2]) <= LH + 1e-4).all()
        pdf = t * t / (AREA * np.abs(wi[:, 1]))
        steep = np.abs(wi[:, 1]) > 0.05  # the plane intersection amplifies FP32 rounding at grazing angles
        assert np.allclose(dirp[stored, 6][steep], pdf[steep], rtol=1e-4), np.abs(dirp[stored, 6] / pdf - 1).max()
        # the shadow ray leaves from the tracer's offset origin (1e-4 along the normal turned to wi)
        assert not _first_sphere_hit(_ray_origin(rec, d, stored, q), wi, 0.99 * t).any()
        e = np.where(wi[:, 1] > 0, 15.0, 0.0)
        want = O.stylize(nodes, roots, np.stack([e, e, e], 1), depth=d + 1, flags=nfl[d][stored],
                         material=np.full(int(stored.sum()), 7), pos=q)
        got = dirp[stored, 3:6]
        assert (np.abs(got - want) <= 1e-5 * max(1.0, np.abs(want).max(initial=0.0))).all()
        # ---- emission and its MIS weight ----
        wo = rec[d, 9:12].T
        le, mis = rec[d, 15:18].T, rec[d, 18]
        light = on & (mat[d] == 7)
        assert (le[on & ~light] == 0).all() and (mis[on & ~light] == 1).all()
        front = light & (wo[:, 1] < 0)
        assert (le[front] == 15).all() and (le[light & ~front] == 0).all()
        if d == 0:
            assert (mis[light] == 1).all()
        elif light.any():
            # the light pdf from the previous ray's origin
            o = _ray_origin(rec, d - 1, light, p[light])
            dist2 = ((p[light] - o) ** 2).sum(1)
            lpdf = dist2 / (AREA * np.maximum(np.abs(wo[light, 1]), 1e-6))
            bp = rec[d - 1, 27][light]
            assert np.allclose(mis[light], bp / (bp + lpdf), rtol=1e-3, atol=1e-6)
        # ---- wo and the bsdf pdf of the continuation ----
        nxt = on & (pd >= d + 1)
        if nxt.any():
            tgt = rec[d + 1, 0:3].T[nxt]
            fwd = tgt - _ray_origin(rec, d, nxt, tgt)
            fwd /= np.linalg.norm(fwd, axis=1)[:, None]
            cosi = np.abs((fwd * rec[d, 3:6].T[nxt]).sum(1))
            glossy = np.zeros(int(nxt.sum()), bool)
            if bsdfs is not None:
                mn = mat[d][nxt]
                glossy = (mn < len(bsdfs)) & (bsdfs["kind"][np.minimum(mn, len(bsdfs) - 1)] == O.BSDF_PRINCIPLED)
                idx = np.flatnonzero(nxt)
                for k in np.flatnonzero(glossy):
                    t = idx[k]
                    want, _ = O.bsdf_eval(bsdfs[mn[k]], rec[d, 12:15, t], rec[d, 3:6, t], rec[d, 6:9, t],
                                          rec[d, 9:12, t], -rec[d + 1, 9:12, t])
                    assert abs(rec[d, 27, t] - want) <= 1e-5 * want, (d, t, rec[d, 27, t], want)
            ok = (cosi > 0.02) & ~glossy
            assert np.allclose(rec[d, 27][nxt][ok], cosi[ok] / np.pi, rtol=1e-3)
            assert np.abs(rec[d + 1, 9:12].T[nxt] + fwd).max() < 1e-3  # wo at d+1
    # both kinds occur, and the style table acts on the light (toon under style index 2)
    assert n_direct > 1000 and n_occl > 100


def _principled_scene_table():
    # floor: diffuse / metal mixture; ceiling: metal; back: tinted metal; the rest diffuse
    return N.bsdf_table([
        {"kind": N.BSDF_PRINCIPLED, "metallic": 0.7, "roughness": 0.35},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.5},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.3, "specular_tint": (1.0, 0.8, 0.6)},
    ])


def test_generator_light_transport_with_principled_materials():
    """the same closed forms with principled materials: every continuation pdf is the mixture
    pdf of the sampled direction (MicrofacetReflection / DiffuseBsdf through BsdfMixture)"""
    W, H, branch, L, seed = 64, 48, 2, 4, 13
    nodes = N.style_table([N.saturation(0.5, inner=-1)])
    roots = [0] * 8
    bs = _principled_scene_table()
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch}
    with N.Handle(cfg) as h:
        h.set_style(nodes, roots)
        h.set_bsdfs(bs)
        desc = h.world_desc(W, H, branch, L)
        bufs = h.alloc_world(desc)
        h.generate_world_records(desc, bufs, seed=seed)
        host = h.download_world(desc, bufs)
    check_light_transport(host, nodes, roots, W, H, branch, L, bsdfs=bs)


def _cap_inside(r, beta, u, c):
    """the cap {w : angle(w, r) < beta} lies inside {w : u . w < c}"""
    ang = np.arccos(np.clip(r @ u, -1, 1))
    return np.where(ang > beta, np.cos(ang - beta), 1.0) < c


def test_generator_samples_the_principled_lobes():
    """Directions drawn from a density p satisfy E[1{wi in A} / p(wi)] = |A| for any region A
    (fixed per vertex) inside p's support. A is chosen so that no continuation in it can leave
    the box through its open front (z = 0.5), so every sample in A is recorded at depth 1:
      floor (metallic 0.7): the directions above 0.1 of the normal heading to -z, |A| = 0.9 pi;
      floor and ceiling (metal): the cone of 10 degrees around wo's mirror direction (about the
      lobes' width), kept for the vertices whose cone lies above the surface and heads to -z,
      |A| = 2 pi (1 - cos 10)."""
    W, H, branch, L, seed = 320, 240, 4, 2, 21
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch}
    with N.Handle(cfg) as h:
        h.set_bsdfs(_principled_scene_table())
        desc = h.world_desc(W, H, branch, L)
        bufs = h.alloc_world(desc)
        h.generate_world_records(desc, bufs, seed=seed)
        host = h.download_world(desc, bufs)
    n = W * H * branch
    rec = host["records"].reshape(L, 28, n).astype(np.float64)
    mat = host["material"].reshape(L, n)
    pd = host["path_depth"]
    nxt = pd >= 1
    wi = np.where(nxt[:, None], -rec[1, 9:12].T, 0.0)
    pdf = rec[0, 27]

    def expect(sel, inA, area, what):
        x = np.where(inA[sel], 1.0 / np.where(pdf[sel] > 0, pdf[sel], 1.0), 0.0)
        m, se = x.mean(), x.std() / np.sqrt(len(x))
        print(f"{what}: {len(x)} vertices, E[1_A / p] = {m:.4f} +- {se:.4f}, |A| = {area:.4f}")
        assert se < 0.03 * area and abs(m - area) < 5 * se, (what, m, se, area)

    floor = (pd >= 0) & (mat[0] == 0)
    inA = nxt & (wi[:, 1] > 0.1) & (wi[:, 2] < 0)
    expect(floor, inA, 0.9 * np.pi, "floor, upper back quarter")
    beta = np.pi / 18
    cap = 2 * np.pi * (1 - np.cos(beta))
    for m_id, up, what in ((0, 1.0, "floor"), (1, -1.0, "ceiling")):
        on = (pd >= 0) & (mat[0] == m_id)
        wo = rec[0, 9:12].T
        nrm = np.array([0.0, up, 0.0])
        r = 2 * (wo @ nrm)[:, None] * nrm - wo
        keep = on & _cap_inside(r, beta, -nrm, 0.0) & _cap_inside(r, beta, np.array([0.0, 0.0, 1.0]), 0.0)
        inC = nxt & ((wi * r).sum(1) > np.cos(beta))
        assert keep.sum() > 2000, (what, int(keep.sum()))
        expect(keep, inC, cap, f"{what}, 10 degree cone around the mirror direction")
