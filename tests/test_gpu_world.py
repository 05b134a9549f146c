"""GPU parity of world mode (the reference's hash-grid NL-PF) against the FP64 oracle on the
same GPU-traced records of the procedural box scene.  Neighbour counts (integers) must match
exactly.  Tolerance 1e-5 relative per vertex and per film pixel, except where the vertex is
ill-conditioned: grazing neighbour directions make a vertex's FP64 result move by up to ~1e-5
when its float inputs move by one FP32 ulp, so no FP32 evaluation order can promise 1e-5
there.  Such a vertex is held to 16x its measured sensitivity (the largest change over three
oracle runs with every non-position input perturbed by a random relative 2^-24), and at
most 1% of the vertices may exceed 1e-5."""
import numpy as np
import pytest

import nlpf as N
import oracle as O
from gpu_util import pixel_rel_err

pytestmark = pytest.mark.gpu
AMIN = [-0.5, -0.5, -0.5]
AMAX = [0.5, 0.5, 0.5]


def run_world(cfg, W, H, branch, L, nodes=None, roots=(), seed=3, bsdfs=None):
    with N.Handle(cfg) as h:
        if nodes is not None:
            h.set_style(nodes, roots)
        if bsdfs is not None:
            h.set_bsdfs(bsdfs)
        desc = h.world_desc(W, H, branch, L, AMIN, AMAX)
        recs = h.alloc_world(desc)
        h.generate_world_records(desc, recs, seed=seed, albedo_lo=0.2, albedo_hi=0.8)
        film, rad = h.world_run(desc, recs, want_radiance=True)
        host = h.download_world(desc, recs)
        stats = h.stats()
    return film, rad, host, stats


def oracle_world(host, W, H, branch, L, cfg, nodes=None, roots=(), bsdfs=None):
    return O.world_run(host["records"], host["flags"], host["new_flags"], host["material"], host["path_depth"],
                       W, H, branch, L, AMIN, AMAX, cfg, nodes=nodes, roots=roots, want_radiance=True, bsdfs=bsdfs)


def styles():
    # walls/spheres: saturation + rex, light: hue shift; sphere0 sets style index 2 for its subtree
    nodes = N.style_table([
        N.saturation(0.7, indirect_scale=2.0),          # 0
        N.rex(0.3, inner=0),                            # 1
        N.hue_shift(0.15),                              # 2 (light)
        N.enable_if(2, inner=2, else_=1),               # 3
        N.set_idx(2, inner=1),                          # 4 sphere0 root
    ])
    roots = [3, 3, 3, 3, 3, 4, 3, 2]
    return nodes, roots


def sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad, trials=3, seed=0, bsdfs=None):
    """Per-vertex and per-pixel relative change of the FP64 oracle under 1-ulp input noise
    (planes 3..27: normals, directions, albedo, emission, direct light, pdfs, MIS weights)."""
    rng = np.random.default_rng(seed)
    n = W * H * branch
    rec0 = host["records"].reshape(L, 28, n)
    s_rad = np.zeros((L, n))
    s_film = np.zeros(W * H)
    for _ in range(trials):
        r = rec0.astype(np.float64)
        r[:, 3:] *= 1 + rng.uniform(-1, 1, r[:, 3:].shape) * 2.0 ** -24
        h2 = dict(host, records=r.astype(np.float32).ravel())
        f, rad, _ = oracle_world(h2, W, H, branch, L, cfg, nodes, roots, bsdfs=bsdfs)
        for d in range(L):
            s_rad[d] = np.maximum(s_rad[d], np.nan_to_num(pixel_rel_err(rad[d], orad[d])))
        s_film = np.maximum(s_film, pixel_rel_err(f.reshape(-1, 3), ofilm.reshape(-1, 3)))
    return s_rad, s_film


def check(film, rad, ofilm, orad, host, L, tol=1e-5, sens=None):
    pd = host["path_depth"]
    over = total = 0
    for d in range(L):
        m = pd >= d
        err = pixel_rel_err(rad[d][m], orad[d][m])
        bound = tol if sens is None else np.maximum(tol, 16 * sens[0][d][m])
        assert (err <= bound).all(), (d, err.max(), int(np.flatnonzero(m)[np.argmax(err / bound)]))
        over += int((err > tol).sum())
        total += int(m.sum())
    assert over <= 0.01 * total, (over, total)
    ef = pixel_rel_err(film.reshape(-1, 3), ofilm.reshape(-1, 3))
    bound = tol if sens is None else np.maximum(tol, 16 * sens[1])
    assert (ef <= bound).all(), (ef.max(), int(np.argmax(ef / bound)))


@pytest.mark.parametrize("kernel", ["thread", "wave"])
@pytest.mark.parametrize("smis", [True, False])
def test_world_parity_config1_like(smis, kernel):
    W, H, branch, L = 48, 40, 4, 4
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "smis": smis, "cluster": False,
           "radius_scale": [2.0, 3.0], "reuse": [1 << 16], "world_kernel": kernel}
    nodes, roots = styles()
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots)
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    assert [st["points"][d] for d in range(L)] == [int((host["path_depth"] >= d).sum()) for d in range(L)]
    sens = sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad)
    check(film, rad, ofilm, orad, host, L, sens=sens)
    assert st["neighbors"][0] > st["points"][0]  # reuse actually happens


@pytest.mark.parametrize("kernel", ["thread", "wave"])
def test_world_reservoir_and_batches_match_oracle(kernel):
    """reuse cap below the neighbour count (reservoir path) and a vram budget forcing several
    batches; the reservoir RNG streams are part of the deterministic contract."""
    W, H, branch, L = 32, 32, 2, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "smis": True, "cluster": False,
           "radius_scale": [3.0], "reuse": [6, 4], "vram": 1, "no_first_hit_stylization": True,
           "world_kernel": kernel}
    film, rad, host, st = run_world(cfg, W, H, branch, L)
    ofilm, orad, _ = oracle_world(host, W, H, branch, L, cfg)
    check(film, rad, ofilm, orad, host, L)


def test_world_bucket_collisions_and_flag_gate():
    W, H, branch, L = 24, 24, 2, 3
    nodes, roots = styles()
    for extra in ({"hash_table_size": 13}, {"disable_flag_check": True}):
        cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": False, "radius_scale": [2.5],
               "reuse": [1 << 16]}
        cfg.update(extra)
        film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, seed=9)
        ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
        check(film, rad, ofilm, orad, host, L)
        assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]


@pytest.mark.parametrize("extra", [
    {"cluster": False, "radius_scale": [2.0, 3.0]},
    {"cluster": True, "radius_scale": [2.5], "reuse": [24], "hash_table_size": 13},
    {"cluster": False, "radius_scale": [2.5], "hash_table_size": 1},
])
def test_world_grid_builds_agree(extra):
    """The counting grid build and the radix-sort build ("world_grid": "sort") give bit-identical
    films, radiance and neighbour counts: they differ only in where whole buckets sit in the entry
    arrays.  With one bucket for the whole grid (1536 entries at depth 0, above the counting
    build's 1024) the default build sorts as well, checked against the oracle."""
    W, H, branch, L = 32, 24, 2, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "reuse": [1 << 16]}
    cfg.update(extra)
    nodes, roots = styles()
    film0, rad0, host, st0 = run_world(dict(cfg, world_grid="auto"), W, H, branch, L, nodes, roots, seed=9)
    film1, rad1, _, st1 = run_world(dict(cfg, world_grid="sort"), W, H, branch, L, nodes, roots, seed=9)
    assert np.array_equal(film0, film1) and np.array_equal(rad0, rad1)
    assert st0["neighbors"][:L] == st1["neighbors"][:L] and sum(st0["neighbors"][:L]) > 0
    if extra.get("hash_table_size") == 1:
        assert st0["points"][0] > 1024
        ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
        check(film0, rad0, ofilm, orad, host, L)
        assert [st0["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]


def test_world_toon_depth1_mostly_exact():
    """Config 1: 4-level toon (w = 0.25) on every diffuse material, active at depth 1 only.
    Band-edge flips (1 FP32 ulp) may change a handful of vertices; everything else is exact."""
    W, H, branch, L = 64, 64, 4, 4
    nodes = N.style_table([N.toon(0.25, depth_mask=N.depth_mask(1))])
    roots = [0, 0, 0, 0, 0, 0, 0, -1]
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "smis": True, "cluster": False,
           "radius_scale": [1.0], "reuse": [1 << 16]}
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots)
    ofilm, orad, _ = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
    pd = host["path_depth"]
    for d in (3, 2):
        m = pd >= d
        assert pixel_rel_err(rad[d][m], orad[d][m]).max() <= 1e-5
    e = pixel_rel_err(film.reshape(-1, 3), ofilm.reshape(-1, 3))
    assert (e <= 1e-5).mean() >= 0.99, (e <= 1e-5).mean()


@pytest.mark.parametrize("kernel,styled", [("thread", False), ("wave", False), ("wave", True)])
def test_world_cluster_mode_matches_oracle(kernel, styled):
    """cluster = true: deterministic clustering (bucket-minimum seeds, ordered joining up to K),
    the gather around the seed restricted to the vertex's cluster, the reservoir over the K+1
    members; bitwise reproducible, and equal to the oracle's restatement. The wave kernel
    shares the SMIS denominators of uncapped lists within a cluster; with the style table,
    members of one cluster carry different flags and so form different lists."""
    W, H, branch, L = 32, 32, 2, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": True, "radius_scale": [3.0],
           "reuse": [16], "world_kernel": kernel}
    nodes, roots = styles() if styled else (None, ())
    f1, r1, host, st = run_world(cfg, W, H, branch, L, nodes, roots)
    f2, r2, _, _ = run_world(cfg, W, H, branch, L, nodes, roots)
    assert np.array_equal(f1, f2) and np.array_equal(r1, r2)
    assert np.isfinite(f1).all()
    # clusters cap the neighbour sets at reuse + 1 members (seed + K joiners), each visited at
    # most once per walked cell that hashes to its bucket
    for d in range(L):
        if st["points"][d]:
            assert st["neighbors"][d] <= st["points"][d] * 17 * 27
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    sens = sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad)
    check(f1, r1, ofilm, orad, host, L, sens=sens)


@pytest.mark.parametrize("kernel", ["thread", "wave"])
def test_world_clusters_with_bucket_collisions(kernel):
    """cluster = true on a 13-bucket table: nearly every walked cell collides with another, so
    a bucket (and the cluster members in it) is visited several times per walk and candidate
    counts exceed the cluster size. The reuse cap sits above the largest cluster but below
    27 times it, so only the exact candidate count may decide whether a depth can run without
    reservoir draws."""
    W, H, branch, L = 24, 24, 2, 3
    K = 24
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": True, "radius_scale": [2.5],
           "reuse": [K], "hash_table_size": 13, "world_kernel": kernel}
    nodes, roots = styles()
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, seed=9)
    ofilm, orad, onbr, cid = O.world_run(host["records"], host["flags"], host["new_flags"], host["material"],
                                         host["path_depth"], W, H, branch, L, AMIN, AMAX, cfg, nodes=nodes,
                                         roots=roots, want_radiance=True, want_clusters=True)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    for d in range(L):
        m = host["path_depth"] >= d
        if m.any():
            sizes = np.bincount(cid[d][m]).astype(np.int64)
            assert sizes.max() <= K < 27 * sizes.max()
            if d == 0:
                # without duplicate visits a vertex meets at most its own cluster's members
                assert st["neighbors"][0] > int(sizes @ sizes)
    sens = sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad)
    check(film, rad, ofilm, orad, host, L, sens=sens)


def test_world_large_clusters_share_smis_sums():
    """Large clusters under a reuse cap they never reach, so every list is uncapped: the SMIS
    denominators are formed once per list (by its first entry) and looked up by every member.
    A list of n members then costs n^2 pdf evaluations once instead of n^2 per member, so the
    evaluations equal the neighbour visits. (Unstylized: with the style table this density
    leaves 0.1 % of the depth-0 vertices up to 1.8e-5 from the FP64 oracle, the same with
    and without the shared sums, as FP32 error of depth 1 is carried up.)"""
    W, H, branch, L = 48, 40, 4, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": True, "radius_scale": [8.0],
           "reuse": [1024], "world_kernel": "wave"}
    film, rad, host, st = run_world(cfg, W, H, branch, L)
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    assert st["neighbors"][0] > 16 * st["points"][0]  # large lists
    assert [st["smis_pairs"][d] for d in range(L)] == [st["neighbors"][d] for d in range(L)]
    sens = sensitivity(host, W, H, branch, L, cfg, None, (), ofilm, orad)
    check(film, rad, ofilm, orad, host, L, sens=sens)


def test_world_clusters_over_batches():
    """Clustering with the wavefront filter over several reservoir batches (vram 1 MB at K = 4
    is 65536 points per batch): full clusters (K + 1 members) draw and sum their own SMIS
    denominators, uncapped lists share theirs, and a list's first entry can sit in another
    batch than its members."""
    W, H, branch, L = 160, 128, 4, 2
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": True, "radius_scale": [3.0],
           "reuse": [4], "vram": 1, "world_kernel": "wave"}
    film, rad, host, st = run_world(cfg, W, H, branch, L)
    assert st["points"][0] > 65536  # more than one batch
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    assert st["smis_pairs"][0] < 16 * st["points"][0]  # some lists shared
    sens = sensitivity(host, W, H, branch, L, cfg, None, (), ofilm, orad)
    check(film, rad, ofilm, orad, host, L, sens=sens)


def test_world_single_level_matches_oracle():
    """max_depth 0: one level, no propagation, only the first hits are filtered."""
    W, H, branch, L = 40, 30, 3, 1
    cfg = {"type": "pf", "max_depth": 0, "branch": branch, "cluster": False, "radius_scale": [2.0],
           "reuse": [1 << 16]}
    nodes, roots = styles()
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, seed=12)
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
    assert st["neighbors"][0] == int(onbr[0]) > 0
    check(film, rad, ofilm, orad, host, L)


def test_world_pass_without_hits():
    """Every camera ray missed: no points at any depth, an all-zero film and no NaN."""
    W, H, branch, L = 16, 8, 2, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": True}
    n = W * H * branch
    host = {"records": np.zeros(L * 28 * n, np.float32), "flags": np.zeros(L * n, np.uint8),
            "new_flags": np.zeros(L * n, np.uint8), "material": np.zeros(L * n, np.uint16),
            "path_depth": np.full(n, -1, np.int32)}
    with N.Handle(cfg) as h:
        desc = h.world_desc(W, H, branch, L, AMIN, AMAX)
        bufs = h.upload_world(desc, host)
        film, rad = h.world_run(desc, bufs, want_radiance=True)
        st = h.stats()
    assert (film == 0).all() and (rad == 0).all()
    assert st["vertices"] == 0 and all(v == 0 for v in st["neighbors"][:L])


@pytest.mark.parametrize("rs", [None, 2.0])
def test_world_config1_full_frame(rs):
    """Config 1 at its full size: 256x256 pixels x 4 paths, max_depth 3 (4 levels), the 4-level
    toon at depth 1, SMIS, no clustering and a reuse cap above every neighbour count, at the
    reference's default radius_scale [0.1, 0.05] and at 2.0; every vertex against the FP64
    oracle.  Every vertex and film pixel is held to 1e-5 relative (measured: below 4e-7), except
    toon band flips: at depth 1 only where the quantiser input lies within 1e-5 of a band edge,
    at depth 0 only on the flipped paths and their neighbours (within the grid radius)."""
    from scipy.spatial import cKDTree
    W, H, branch, L = 256, 256, 4, 4
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "smis": True, "cluster": False,
           "reuse": [1 << 16]}
    if rs is not None:
        cfg["radius_scale"] = [rs]
    nodes = N.style_table([N.toon(0.25, depth_mask=N.depth_mask(1))])
    off = N.style_table([N.toon(0.25, depth_mask=0)])
    roots = [0, 0, 0, 0, 0, 0, 0, -1]
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, seed=1)
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    _, prad, _ = oracle_world(host, W, H, branch, L, cfg, off, roots)
    pd = host["path_depth"]
    n = W * H * branch
    rec = host["records"].reshape(L, 28, n)
    mat = host["material"].reshape(L, n)
    report = []
    flipped = np.zeros(n, bool)
    for d in (3, 2, 1, 0):
        m = pd >= d
        err = pixel_rel_err(rad[d], orad[d])
        bad = m & ~(err <= 1e-5)
        if d == 1:
            near = toon_edge_world(prad[1]) & (mat[1] != 7)
            assert not (bad & ~near).any(), (d, err[bad & ~near].max())
            flipped = bad
        elif d == 0:
            allowed = flipped.copy()
            if flipped.any():
                radius = float(np.mean(np.array(AMAX) - np.array(AMIN))) * 0.01 * (0.1 if rs is None else rs)
                tree = cKDTree(rec[0, 0:3].T[m])
                idx = np.flatnonzero(m)
                for j in tree.query_ball_point(rec[0, 0:3].T[flipped & m], 1.0001 * radius):
                    allowed[idx[j]] = True
            assert not (bad & ~allowed).any(), (d, err[bad & ~allowed].max())
        else:
            assert not bad.any(), (d, err[bad].max())
        report.append((d, float(err[m].max()), float((err[m] > 1e-5).mean()), int(bad.sum())))
    ef = pixel_rel_err(film.reshape(-1, 3), ofilm.reshape(-1, 3))
    fb = ~(ef <= 1e-5)
    pix_allowed = np.zeros(W * H, bool)
    pix_allowed[np.flatnonzero(allowed) % (W * H)] = True
    assert not (fb & ~pix_allowed).any(), ef[fb & ~pix_allowed].max()
    print(f"config 1 world frame radius_scale {rs or 'default'}: per depth (max err, fraction > 1e-5, "
          f"outside bound) {report}; film max {ef.max():.2e}, fraction > 1e-5 {(ef > 1e-5).mean():.2e}, "
          f"toon flips {int(flipped.sum())}")
    assert flipped.mean() < 1e-3


def toon_edge_world(pre, w=0.25, tol=1e-5):
    lum = np.clip(pre, 0, 1e5) @ np.array([0.2126, 0.7152, 0.0722])
    enc = np.where(lum > 1, lum / w, np.where(lum <= 0.0031308, lum * 12.92,
                                               1.055 * np.power(np.maximum(lum, 0), 1 / 2.4) - 0.055) / w)
    return np.abs(enc - np.round(enc)) < tol * np.maximum(1, np.abs(enc))


def glossy_table():
    """BSDFs of the box scene's materials (floor, ceiling, back, left, right, sphere0, sphere1,
    light): the principled subset the filter evaluates exactly, from a near mirror (roughness 0,
    alpha 1e-4, as the reference scene's Gold) to diffuse / metal mixtures"""
    return N.bsdf_table([
        {"kind": N.BSDF_PRINCIPLED, "metallic": 0.5, "roughness": 0.4},
        {},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.3, "specular_tint": (1.0, 0.8, 0.6)},
        {},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.0},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.2, "specular_tint": (0.9, 0.9, 1.0)},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 0.3, "roughness": 0.6},
        {"kind": N.BSDF_EMISSION},
    ])


@pytest.mark.parametrize("kernel", ["thread", "wave"])
@pytest.mark.parametrize("smis", [True, False])
def test_world_glossy_parity(smis, kernel):
    """principled (metal and diffuse / metal) materials at the centre and, for SMIS, at every
    reused neighbour, against the oracle's restatement of the reference's BSDFs"""
    W, H, branch, L = 48, 40, 4, 4
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "smis": smis, "cluster": False,
           "radius_scale": [2.0, 3.0], "reuse": [1 << 16], "world_kernel": kernel}
    nodes, roots = styles()
    bs = glossy_table()
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, bsdfs=bs)
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots, bsdfs=bs)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    sens = sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad, bsdfs=bs)
    check(film, rad, ofilm, orad, host, L, sens=sens)
    # the table changes the result materially (so the test exercises the glossy lobes)
    dfilm = oracle_world(host, W, H, branch, L, cfg, nodes, roots)[0]
    assert np.abs(dfilm - ofilm).max() > 1e-2 * np.abs(ofilm).max()
    mat = host["material"].reshape(L, -1)
    assert (mat[host["path_depth"][None, :] >= np.arange(L)[:, None]] == 4).any()  # the roughness-0 wall


@pytest.mark.parametrize("kernel", ["thread", "wave"])
def test_world_glossy_clusters(kernel):
    """the reference's default settings (clustering, reuse [128, 64], SMIS) with glossy
    materials: on the wave kernel the SMIS denominators are shared within each cluster's list"""
    W, H, branch, L = 40, 32, 4, 4
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "radius_scale": [3.0, 4.0], "world_kernel": kernel}
    nodes, roots = styles()
    bs = glossy_table()
    film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, bsdfs=bs)
    ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots, bsdfs=bs)
    assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]]
    sens = sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad, bsdfs=bs)
    check(film, rad, ofilm, orad, host, L, sens=sens)


def test_world_diffuse_tables_are_the_default():
    """a table of diffuse entries runs the plain kernels (bit for bit the run without a table);
    principled with metallic 0 and no specular layer is that same diffuse BSDF"""
    W, H, branch, L = 32, 24, 2, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": False, "radius_scale": [2.5],
           "reuse": [1 << 16]}
    film0, rad0, host, _ = run_world(cfg, W, H, branch, L)
    film1, rad1, _, _ = run_world(cfg, W, H, branch, L, bsdfs=N.bsdf_table([{}] * 8))
    assert np.array_equal(film0, film1) and np.array_equal(rad0, rad1)
    film2, rad2, _, _ = run_world(cfg, W, H, branch, L,
                                  bsdfs=N.bsdf_table([{"kind": N.BSDF_PRINCIPLED, "metallic": 0.0}] * 7))
    for d in range(L):
        m = host["path_depth"] >= d
        assert pixel_rel_err(rad2[d][m], rad0[d][m]).max() < 1e-6
    assert pixel_rel_err(film2.reshape(-1, 3), film0.reshape(-1, 3)).max() < 1e-6


def test_set_bsdfs_rejects_what_needs_the_albedo_table():
    bad = [
        ({"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "coat_weight": 0.2}, 5, "coat"),
        ({"kind": N.BSDF_PRINCIPLED, "metallic": 0.5, "specular_ior_level": 0.5}, 5, "specular_ior_level"),
        ({"kind": N.BSDF_PRINCIPLED, "metallic": 0.0, "transmission": 1.0}, 5, "transmission"),
        ({"kind": N.BSDF_PRINCIPLED, "metallic": 1.5}, 2, "metallic"),
        ({"kind": N.BSDF_PRINCIPLED, "roughness": float("nan")}, 2, "non-finite"),
        ({"kind": 3}, 2, "kind"),
        ({"reserved": 1}, 2, "reserved"),
    ]
    with N.Handle({"type": "pf"}) as h:
        for entry, code, msg in bad:
            with pytest.raises(N.NlpfError) as e:
                h.set_bsdfs(N.bsdf_table([{}, entry]))
            assert e.value.code == code and msg in str(e.value) and "material 1" in str(e.value), entry
        # metallic 1 ignores the specular layer and the transmission side
        h.set_bsdfs(N.bsdf_table([{"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "specular_ior_level": 0.5,
                                   "transmission": 1.0}]))
        h.set_bsdfs(N.bsdf_table([]))


def test_world_random_configs():
    """Eight random world-mode cases against the oracle: frame shape (width 3..36, height
    2..30, 1..3 paths per pixel), 1..4 levels, radius_scale 1.5..3, clustering and SMIS on or
    off, a colliding, a sparse or the default bucket table, either filter kernel, reuse caps
    from 6 (reservoir draws over batches of 1 MB) to 2^16.  Neighbour counts exactly, values by
    the world-mode rule."""
    rng = np.random.default_rng(2027)
    nodes, roots = styles()
    for case in range(8):
        W, H = int(rng.integers(3, 37)), int(rng.integers(2, 31))
        branch, L = int(rng.integers(1, 4)), int(rng.integers(1, 5))
        cfg = {"type": "pf", "max_depth": L - 1, "branch": branch,
               "radius_scale": [float(rng.uniform(1.5, 3.0))],
               "cluster": bool(rng.integers(0, 2)), "smis": bool(rng.integers(0, 2)),
               "world_kernel": ["auto", "thread", "wave"][int(rng.integers(0, 3))],
               "reuse": [[1 << 16], [64], [6]][int(rng.integers(0, 3))]}
        table = [13, 1000003, None][int(rng.integers(0, 3))]
        if table is not None:
            cfg["hash_table_size"] = table
        if cfg["reuse"] == [6]:
            cfg["vram"] = 1
        seed = int(rng.integers(1, 1000))
        print("case", case, W, H, branch, L, cfg, "seed", seed, flush=True)
        film, rad, host, st = run_world(cfg, W, H, branch, L, nodes, roots, seed=seed)
        ofilm, orad, onbr = oracle_world(host, W, H, branch, L, cfg, nodes, roots)
        assert [st["neighbors"][d] for d in range(L)] == [int(v) for v in onbr[:L]], case
        sens = sensitivity(host, W, H, branch, L, cfg, nodes, roots, ofilm, orad, trials=2)
        check(film, rad, ofilm, orad, host, L, sens=sens)
