"""Pins the world-mode oracle (the reference's hash-grid NL-PF) against an independent numpy
restatement on tiny inputs, the brute-force radius search (with and without bucket
collisions), and closed forms of the Lambertian reuse estimator (no GPU)."""
import numpy as np
import pytest

import nlpf as N

NP = 28
AMIN = np.array([-0.5, -0.5, -0.5], np.float32)
AMAX = np.array([0.5, 0.5, 0.5], np.float32)


def teschner(c, table):
    c = np.asarray(c, np.uint64)
    return ((c[..., 0] * np.uint64(73856093)) ^ (c[..., 1] * np.uint64(19349663)) ^ (c[..., 2] * np.uint64(83492791))) \
        % np.uint64(table)


def cells(pos, res):
    t = (pos - AMIN) / (AMAX - AMIN)
    t = np.clip(t, 0, 1).astype(np.float32) * res.astype(np.float32)
    return np.minimum(t.astype(np.uint32), res - 1)


@pytest.mark.parametrize("table", [64 * 1024 * 1024, 7, 1])
def test_grid_neighbour_multiset(oracle, table):
    rng = np.random.default_rng(table)
    pos = (rng.random((600, 3)) - 0.5).astype(np.float32)
    radius = np.float32(0.12)
    got = oracle.world_neighbor_counts(pos, AMIN, AMAX, radius, table)
    res = np.floor((AMAX - AMIN) / radius).astype(np.uint32)
    c = cells(pos, res).astype(np.int64)
    b = teschner(c, table)
    dist = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1, dtype=np.float32))
    within = dist <= radius
    expect = np.zeros(len(pos), np.int64)
    for t in range(len(pos)):
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nc = c[t] + np.array([dx, dy, dz])
                    if (nc < 0).any() or (nc >= res).any():
                        continue
                    bb = teschner(nc, table)
                    expect[t] += int((within[t] & (b == bb)).sum())
    assert np.array_equal(got, expect)
    if table == 64 * 1024 * 1024:  # no collisions: exactly the brute-force radius search
        assert np.array_equal(got, within.sum(1))
    else:  # collisions only add duplicates
        assert (got >= within.sum(1)).all() and (got > within.sum(1)).any()


def make_records(rng, W, H, branch, L, same_normal=False, tight=False):
    Np = W * H * branch
    rec = np.zeros((L, NP, Np), np.float32)
    for d in range(L):
        if tight:
            rec[d, 0:3] = (rng.random((3, Np)) * 0.02 - 0.01).astype(np.float32)
        else:
            rec[d, 0:3] = (rng.random((3, Np)) - 0.5).astype(np.float32) * 0.3
        if same_normal:
            n = np.tile(np.array([[0.0], [0.0], [1.0]]), (1, Np))
        else:
            n = rng.normal(0, 0.3, (3, Np)) + np.array([[0.0], [0.0], [1.0]])
            n /= np.linalg.norm(n, axis=0)
        rec[d, 3:6] = n
        if d == 0:
            wo = rng.normal(0, 0.3, (3, Np)) + np.array([[0.0], [0.0], [1.0]])
            rec[d, 9:12] = wo / np.linalg.norm(wo, axis=0)
        else:
            # the child is reached along the parent's sampled direction: wo_child = -wi_parent,
            # with wi_parent on the parent's side of its normal (as BSDF sampling produces)
            pn = rec[d - 1, 3:6].astype(np.float64)
            wi = rng.normal(0, 0.4, (3, Np)) + pn
            wi /= np.linalg.norm(wi, axis=0)
            c = (wi * pn).sum(0)
            wi = np.where(c < 0, wi - 2 * c * pn, wi)  # reflect into the parent's hemisphere
            rec[d, 9:12] = -wi
        wo_d = rec[d, 9:12].astype(np.float64)
        flip = (wo_d * rec[d, 3:6]).sum(0) < 0
        rec[d, 3:6, flip] *= -1  # keep wo strictly on the normal's side
        rec[d, 6:9] = rec[d, 3:6]
        rec[d, 12:15] = rng.uniform(0.2, 0.8, (3, Np))
        rec[d, 15:18] = rng.uniform(0, 0.5, (3, Np)) * (rng.random(Np) < 0.2)
        rec[d, 18] = rng.uniform(0.5, 1.0, Np)
        wi = rng.normal(0, 0.5, (3, Np)) + np.array([[0.0], [0.0], [1.0]])
        rec[d, 19:22] = wi / np.linalg.norm(wi, axis=0)
        rec[d, 22:25] = rng.uniform(0, 3, (3, Np))
        rec[d, 25] = rng.uniform(0.1, 2, Np) * (rng.random(Np) < 0.8)
        rec[d, 26] = 1.0
        rec[d, 27] = rng.uniform(0.1, 0.3, Np)
    depth = rng.integers(-1, L, Np).astype(np.int32)
    return rec, depth


def numpy_world(rec, depth, W, H, branch, L, radius_scale, smis, nodes=None, roots=(), flags=None, new_flags=None,
                material=None, no_first_hit_stylization=True, clusters=None, bsdfs=None):
    """Independent restatement for cnt <= K (no reservoir) and no bucket collisions: the
    neighbour set of a vertex is every depth-d vertex within the radius with equal flags.
    With bsdfs, each vertex's BSDF is its material's entry, evaluated by the oracle's
    or_bsdf_eval (pinned on its own in test_oracle_bsdf.py)."""
    Np = W * H * branch
    ext = (AMAX - AMIN).astype(np.float32)
    avg = np.float32((ext[0] + ext[1] + ext[2]) / np.float32(3))
    ind = np.zeros((8, Np))
    rad = np.zeros((L, Np, 3))
    nf = np.zeros((L, Np), np.uint8) if new_flags is None else new_flags
    mat = np.zeros((L, Np), np.uint16) if material is None else material
    for d in range(L - 1, -1, -1):
        r = rec[d].astype(np.float64)
        pts = np.nonzero(depth >= d)[0]
        radius = np.float32(avg * np.float32(0.01) * np.float32(radius_scale))
        P = rec[d, 0:3, pts].astype(np.float32)  # (n, 3)
        dist = np.sqrt(((P[:, None, :] - P[None, :, :]) ** 2).sum(-1, dtype=np.float32))
        for a, c in enumerate(pts):
            if clusters is None:
                nb = pts[(dist[a] <= radius) & (nf[d, pts] == nf[d, c])]
            else:  # the gather runs around the seed and keeps the vertex's own cluster
                seed = clusters[d, c]
                sa = int(np.nonzero(pts == seed)[0][0])
                nb = pts[(dist[sa] <= radius) & (clusters[d, pts] == seed) & (nf[d, pts] == nf[d, c])]
            ns, ng, wo, R = r[3:6, c], r[6:9, c], r[9:12, c], r[12:15, c]

            def ev(ns_, ng_, wo_, wi_, R_, m_):
                """(pdf, f |cos_i| per channel)"""
                if bsdfs is not None:
                    import oracle as O
                    return O.bsdf_eval(bsdfs[m_] if m_ < len(bsdfs) else None, R_, ns_, ng_, wo_, wi_)
                sg = lambda x: 1.0 if x > 0 else -1.0
                fl = sg(ng_ @ ns_)
                ok = sg(fl * (wo_ @ ns_)) * sg(wo_ @ ng_) > 0 and sg(fl * (wi_ @ ns_)) * sg(wi_ @ ng_) > 0
                if not ok or not ((wo_ @ ns_) * (wi_ @ ns_) >= 0):
                    return 0.0, np.zeros(3)
                p = abs(wi_ @ ns_) / np.pi
                return p, R_ * p

            mc = mat[d, c]
            dsum = np.zeros(3)
            isum = np.zeros(3)
            for j in nb:
                pd = r[25, j]
                if pd > 0:
                    b, fc = ev(ns, ng, wo, r[19:22, j], R, mc)
                    dsum += fc * r[22:25, j] * r[26, j] * (pd / (pd + b)) / pd / len(nb)
                wi = ind[0:3, j]
                if np.linalg.norm(wi) == 0:
                    continue
                b, fc = ev(ns, ng, wo, wi, R, mc)
                reused = fc * ind[3:6, j] * ind[6, j]
                if smis:
                    s = sum(ev(r[3:6, k], r[6:9, k], r[9:12, k], wi, r[12:15, k], mat[d, k])[0] for k in nb)
                    isum += reused / s
                elif b > 0:
                    isum += reused / b / len(nb)
            Lr = r[15:18, c] + dsum + isum
            if nodes is not None and roots[mat[d, c]] >= 0 and (d > 0 or not no_first_hit_stylization):
                import oracle as O
                Lr = O.stylize(nodes, roots, Lr[None], depth=d, flags=[nf[d, c]], material=[mat[d, c]],
                               pos=r[0:3, c][None])[0]
            rad[d, c] = Lr
        ind[:] = 0
        if d > 0:
            for c in pts:
                ind[0:3, c] = -r[9:12, c]
                ind[3:6, c] = rad[d, c]
                ind[6, c] = r[18, c]
                ind[7, c] = rec[d - 1, 27, c]
    return rad


@pytest.mark.parametrize("smis", [True, False])
def test_glossy_materials_match_numpy_restatement(oracle, smis):
    """each vertex's own material BSDF at the centre, each reused neighbour's own BSDF in the SMIS
    denominators, RGB f for metals (conductor Fresnel per channel)"""
    rng = np.random.default_rng(17 + smis)
    W, H, branch, L = 6, 5, 2, 3
    rec, depth = make_records(rng, W, H, branch, L)
    Np = W * H * branch
    z8 = np.zeros((L, Np), np.uint8)
    mat = rng.integers(0, 4, (L, Np)).astype(np.uint16)
    table = np.concatenate([oracle.bsdf(kind=oracle.BSDF_DIFFUSE),
                            oracle.bsdf(metallic=1.0, roughness=0.3, specular_tint=(1.0, 0.8, 0.5)),
                            oracle.bsdf(metallic=0.4, roughness=0.6)])  # material 3: no entry (diffuse)
    cfg = {"cluster": False, "radius_scale": [8.0], "reuse": [1 << 20], "smis": smis, "no_first_hit_stylization": True}
    film, rad, nbr = oracle.world_run(rec, z8, z8, mat, depth, W, H, branch, L, AMIN, AMAX, cfg,
                                      want_radiance=True, bsdfs=table)
    ref = numpy_world(rec, depth, W, H, branch, L, 8.0, smis, material=mat, bsdfs=table)
    for d in range(L):
        m = depth >= d
        np.testing.assert_allclose(rad[d][m], ref[d][m], rtol=1e-12, atol=1e-14)
    plain = oracle.world_run(rec, z8, z8, mat, depth, W, H, branch, L, AMIN, AMAX, cfg, want_radiance=True)[1]
    assert np.abs(plain[0][depth >= 0] - rad[0][depth >= 0]).max() > 1e-3  # the table matters


@pytest.mark.parametrize("smis", [True, False])
def test_matches_numpy_restatement(oracle, smis):
    rng = np.random.default_rng(7 + smis)
    W, H, branch, L = 6, 5, 2, 3
    rec, depth = make_records(rng, W, H, branch, L)
    cfg = {"cluster": False, "radius_scale": [8.0], "reuse": [1 << 20], "smis": smis, "no_first_hit_stylization": True}
    Np = W * H * branch
    z8 = np.zeros((L, Np), np.uint8)
    z16 = np.zeros((L, Np), np.uint16)
    film, rad, nbr = oracle.world_run(rec, z8, z8, z16, depth, W, H, branch, L, AMIN, AMAX, cfg,
                                      want_radiance=True)
    ref = numpy_world(rec, depth, W, H, branch, L, 8.0, smis)
    for d in range(L):
        m = depth >= d
        assert np.allclose(rad[d][m], ref[d][m], rtol=1e-12, atol=1e-14), d
    # film: mean of valid branches (no first-hit stylization)
    rr = rad[0].reshape(branch, H * W, 3)
    valid = (depth >= 0).reshape(branch, H * W)
    expect = np.where(valid.any(0)[:, None], (rr * valid[..., None]).sum(0) / branch, 0)
    assert np.allclose(film.reshape(-1, 3), expect, rtol=1e-12)


def test_flags_gate_and_stylization(oracle):
    rng = np.random.default_rng(11)
    W, H, branch, L = 5, 4, 2, 2
    rec, depth = make_records(rng, W, H, branch, L)
    Np = W * H * branch
    nf = rng.choice([0, N.FLAG_STYLE2, N.FLAG_STYLE3], (L, Np)).astype(np.uint8)
    mat = rng.integers(0, 2, (L, Np)).astype(np.uint16)
    nodes = N.style_table([N.rex(0.4), N.hue_shift(0.2), N.enable_if(1, inner=1, else_=0)])
    roots = [2, 0]
    for nfhs in (True, False):
        cfg = {"cluster": False, "radius_scale": [8.0], "reuse": [1 << 20], "smis": True, "no_first_hit_stylization": nfhs}
        film, rad, _ = oracle.world_run(rec, nf, nf, mat, depth, W, H, branch, L, AMIN, AMAX, cfg, nodes=nodes,
                                        roots=roots, want_radiance=True)
        ref = numpy_world(rec, depth, W, H, branch, L, 8.0, True, nodes=nodes, roots=roots, new_flags=nf,
                          material=mat, no_first_hit_stylization=nfhs)
        for d in range(L):
            m = depth >= d
            assert np.allclose(rad[d][m], ref[d][m], rtol=1e-12, atol=1e-14)


def test_lambertian_closed_forms(oracle):
    """Same normals, wo and wi in the upper hemisphere: SMIS and 1/(pdf|N|) both reduce the
    indirect term to R * mean_j(L_j mis_j); the direct term is the balance-heuristic sum."""
    rng = np.random.default_rng(3)
    W, H, branch, L = 4, 4, 1, 2
    rec, _ = make_records(rng, W, H, branch, L, same_normal=True, tight=True)
    depth = np.full(W * H, 1, np.int32)
    rec[1, 25] = 0.0  # no direct light at the top level: its output is le
    cfg = {"cluster": False, "radius_scale": [8.0], "reuse": [1 << 20], "no_first_hit_stylization": True}
    Np = W * H
    z8 = np.zeros((L, Np), np.uint8)
    z16 = np.zeros((L, Np), np.uint16)
    for smis in (True, False):
        cfg["smis"] = smis
        _, rad, _ = oracle.world_run(rec, z8, z8, z16, depth, W, H, branch, L, AMIN, AMAX, cfg,
                                     want_radiance=True)
        r = rec.astype(np.float64)
        assert np.allclose(rad[1], r[1, 15:18].T, rtol=1e-12)
        ind = (r[1, 15:18] * r[1, 18]).mean(1)  # mean_j L_j mis_j over the whole (tight) cluster
        cosd = r[0, 21]  # wi.z with n = +z
        pd = r[0, 25]
        dterm = ((cosd / np.pi) * r[0, 22:25] * r[0, 26] / (pd + cosd / np.pi) * (pd > 0)).mean(1)
        expect = r[0, 15:18].T + r[0, 12:15].T * (ind + dterm)[None, :]
        assert np.allclose(rad[0], expect, rtol=1e-10)


def test_reservoir_cap_and_film_restylization(oracle):
    rng = np.random.default_rng(5)
    W, H, branch, L = 3, 3, 3, 1
    rec, _ = make_records(rng, W, H, branch, L, same_normal=True, tight=True)
    depth = np.zeros(W * H * branch, np.int32)
    Np = W * H * branch
    z8 = np.zeros((L, Np), np.uint8)
    z16 = np.zeros((L, Np), np.uint16)
    # K = 1: every vertex reuses exactly one of the 27 candidates
    cfg = {"cluster": False, "radius_scale": [8.0], "reuse": [1], "smis": False, "no_first_hit_stylization": True}
    _, rad, nbr = oracle.world_run(rec, z8, z8, z16, depth, W, H, branch, L, AMIN, AMAX, cfg, want_radiance=True)
    assert nbr[0] == Np * Np  # every vertex sees all of them
    r = rec[0].astype(np.float64)
    single = r[15:18].T[:, None, :] + r[12:15].T[:, None, :] * (
        (r[21] / np.pi) * r[22:25] * r[26] / (r[25] + r[21] / np.pi) * (r[25] > 0))[None].transpose(0, 2, 1)
    for c in range(Np):
        assert np.isclose(single[c], rad[0, c], rtol=1e-10).all(1).any()
    # film re-applies depth-0 stylization to the branch mean (kernel_update_film :451-470)
    nodes = N.style_table([N.rex(0.5)])
    cfg = {"cluster": False, "radius_scale": [8.0], "reuse": [1 << 20], "smis": True, "no_first_hit_stylization": False}
    film, rad, _ = oracle.world_run(rec, z8, z8, z16, depth, W, H, branch, L, AMIN, AMAX, cfg, nodes=nodes,
                                    roots=[0], want_radiance=True)
    mean = rad[0].reshape(branch, W * H, 3).mean(0)
    again = oracle.stylize(nodes, [0], mean, depth=0)
    assert np.allclose(film.reshape(-1, 3), again, rtol=1e-12)


def numpy_clusters(rec, depth, L, radius_scale, K, table=64 * 1024 * 1024):
    """Independent restatement of the deterministic clustering: bucket-minimum seeds, then
    rounds of proposals to the first seed with room (27-cell order, bucket entries by path id),
    each seed accepting its proposals in ascending vertex order up to K joiners."""
    UNSET = 0xFFFFFFFF
    ext = (AMAX - AMIN).astype(np.float32)
    avg = np.float32((ext[0] + ext[1] + ext[2]) / np.float32(3))
    out = np.full((L, rec.shape[2]), UNSET, np.uint32)
    for d in range(L):
        pts = np.nonzero(depth >= d)[0]
        radius = np.float32(avg * np.float32(0.01) * np.float32(radius_scale))
        res = np.floor(ext / radius).astype(np.uint32)
        P = rec[d, 0:3].T.astype(np.float32)
        cell = {int(i): cells(P[i], res).astype(np.int64) for i in pts}
        bucket = {i: int(teschner(cell[i], table)) for i in cell}
        members = {}
        for i in sorted(cell):
            members.setdefault(bucket[i], []).append(i)
        cid = {i: UNSET for i in cell}
        size = {i: 0 for i in cell}
        while True:
            seeds = [min([j for j in m if cid[j] == UNSET], default=None) for m in members.values()]
            for sd in seeds:
                if sd is not None:
                    cid[sd] = sd
            while True:
                props, n_uncl = [], 0
                for i in sorted(cell):
                    if cid[i] != UNSET:
                        continue
                    n_uncl += 1
                    found = None
                    for dz in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                if found is not None:
                                    continue
                                nc = cell[i] + np.array([dx, dy, dz])
                                if (nc < 0).any() or (nc >= res).any():
                                    continue
                                for j in members.get(int(teschner(nc, table)), []):
                                    dd = np.sqrt(np.float32(((P[j] - P[i]) ** 2).sum(dtype=np.float32)))
                                    if dd <= radius and cid[j] == j and size[j] < K:
                                        found = j
                                        break
                    if found is not None:
                        props.append((found, i))
                if not props:
                    break
                props.sort()
                for sd in sorted(set(p[0] for p in props)):
                    mine = [i for (t, i) in props if t == sd]
                    room = max(K - size[sd], 0)
                    for i in mine[:room]:
                        cid[i] = sd
                    size[sd] += min(len(mine), room)
            if n_uncl == 0:
                break
        for i, c in cid.items():
            out[d, i] = c
    return out


@pytest.mark.parametrize("smis,K", [(True, 1), (False, 1), (True, 64), (False, 64)])
def test_clustering_matches_restatement_and_invariants(oracle, smis, K):
    """cluster = true: the oracle's cluster ids equal the independent restatement; every vertex
    sits in exactly one cluster whose seed is within the radius and has at most K joiners.  With
    a cap the clusters never fill (K = 64, no reservoir draws), the filtered radiance equals the
    numpy restatement gathering each vertex's own cluster."""
    rng = np.random.default_rng(17)
    W, H, branch, L, rs = 8, 8, 2, 2, 8.0
    rec, depth = make_records(rng, W, H, branch, L)
    Np = W * H * branch
    z8 = np.zeros((L, Np), np.uint8)
    z16 = np.zeros((L, Np), np.uint16)
    cfg = {"cluster": True, "radius_scale": [rs], "reuse": [K], "smis": smis, "no_first_hit_stylization": True}
    film, rad, nbr, cl = oracle.world_run(rec, z8, z8, z16, depth, W, H, branch, L, AMIN, AMAX, cfg,
                                          want_radiance=True, want_clusters=True)
    expect = numpy_clusters(rec, depth, L, rs, K)
    assert np.array_equal(cl, expect)
    ext = (AMAX - AMIN).astype(np.float32)
    radius = np.float32(np.float32((ext[0] + ext[1] + ext[2]) / np.float32(3)) * np.float32(0.01) * np.float32(rs))
    for d in range(L):
        pts = np.nonzero(depth >= d)[0]
        c = cl[d, pts]
        assert (c != 0xFFFFFFFF).all() and (cl[d, c] == c).all()  # seeds are their own cluster
        P = rec[d, 0:3].T
        dd = np.sqrt(((P[pts] - P[c]) ** 2).sum(-1, dtype=np.float32))
        assert (dd <= radius).all()
        seeds, counts = np.unique(c, return_counts=True)
        assert (counts <= K + 1).all() and len(seeds) < len(pts)  # capped, and some clusters merge
        if K == 1:
            assert (counts == K + 1).any()  # the cap binds
    if K == 1:
        return
    assert max(np.unique(cl[d][depth >= d], return_counts=True)[1].max() for d in range(L)) <= K
    ref = numpy_world(rec, depth, W, H, branch, L, rs, smis, clusters=cl)
    for d in range(L):
        m = depth >= d
        assert np.allclose(rad[d][m], ref[d][m], rtol=1e-12, atol=1e-14), d


def test_result_independent_of_thread_count(oracle, tmp_path):
    """The oracle filters reservoir slots in parallel and each slot's points in order. With
    more points than one batch (vram 1 MB at K = 4 gives 65536 slots) and neighbour counts
    above K, its output must equal a one-thread run bit for bit."""
    import os
    import subprocess
    import sys
    rng = np.random.default_rng(11)
    W, H, branch, L = 96, 96, 8, 1
    rec, _ = make_records(rng, W, H, branch, L)
    Np = W * H * branch
    depth = np.zeros(Np, np.int32)
    z8 = np.zeros((L, Np), np.uint8)
    z16 = np.zeros((L, Np), np.uint16)
    cfg = {"cluster": False, "radius_scale": [1.2], "reuse": [4], "vram": 1, "smis": True,
           "no_first_hit_stylization": True}
    film, rad, nbr = oracle.world_run(rec, z8, z8, z16, depth, W, H, branch, L, AMIN, AMAX, cfg, want_radiance=True)
    assert nbr[0] > 4 * Np  # the reservoir draws
    np.save(tmp_path / "rec.npy", rec)
    script = (
        "import sys, numpy as np, oracle as O\n"
        f"rec = np.load(sys.argv[1]); Np = {Np}; L = {L}\n"
        "z8 = np.zeros((L, Np), np.uint8); z16 = np.zeros((L, Np), np.uint16)\n"
        f"f, r, n = O.world_run(rec, z8, z8, z16, np.zeros(Np, np.int32), {W}, {H}, {branch}, L,\n"
        "    np.full(3, -0.5, np.float32), np.full(3, 0.5, np.float32), " + repr(cfg) + ", want_radiance=True)\n"
        "np.save(sys.argv[2], r)\n")
    env = dict(os.environ, OMP_NUM_THREADS="1",
               PYTHONPATH=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    subprocess.run([sys.executable, "-c", script, str(tmp_path / "rec.npy"), str(tmp_path / "rad1.npy")],
                   check=True, env=env)
    assert np.array_equal(np.load(tmp_path / "rad1.npy"), rad)


# ---- random streams: published known answers ----
# PCG32 (O'Neill's pcg32-demo: pcg32_srandom_r(42, 54), the seeding of new_seq_offset(54, 42))
PCG_DEMO = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
CHACHA_CONST = [0x61707865, 0x3320646e, 0x79622d32, 0x6b206574]
# RFC 7539 2.3.2: ChaCha20 block of key 00..1f, counter 1, nonce 00000009 0000004a 00000000
RFC7539_232 = [0xe4e7f110, 0x15593bd1, 0x1fdd0f50, 0xc47120a3, 0xc7f4d1c7, 0x0368c033, 0x9aaa2204, 0x4e6cd4c3,
               0x466482d2, 0x09aa9f07, 0x05d7c214, 0xa2028bd9, 0xd19c12b5, 0xb94e16de, 0xe883d0cb, 0x4e3c50a2]
# RFC 7539 A.1 test vector 1: all-zero key, nonce and counter
RFC7539_A1 = bytes.fromhex("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                           "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")


def test_pcg32_known_answer(oracle):
    assert [int(v) for v in oracle.pcg32(54, 42, 6)] == PCG_DEMO


def test_chacha_block_known_answers(oracle):
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    got = oracle.chacha_block(CHACHA_CONST + key + [1, 0x09000000, 0x4a000000, 0], 20)
    assert [int(v) for v in got] == RFC7539_232
    assert oracle.chacha_block(CHACHA_CONST + [0] * 12, 20).astype("<u4").tobytes() == RFC7539_A1


def test_std_rng_known_answers(oracle):
    """rand 0.8.5 StdRng (ChaCha12) value-stability test (rngs/std.rs test_stdrng_construction):
    from_seed([1,0,0,0, 23,0,0,0, 200,1,0,0, 210,30,0,0, 0...]) gives next_u64
    10719222850664546238; from_rng of that generator (its next 32 bytes as the key) gives
    14064965282130556830.  rand_core 0.6.4 seed_from_u64(0) into an 8-byte seed reads
    5029875928683246316 (its value-breakage test)."""
    x = oracle.chacha12_u64([1, 23, 456, 7890, 0, 0, 0, 0], 0, 5)
    assert int(x[0]) == 10719222850664546238
    key = [w for v in x[1:5] for w in (int(v) & 0xFFFFFFFF, int(v) >> 32)]
    assert int(oracle.chacha12_u64(key, 0, 1)[0]) == 14064965282130556830
    k = oracle.std_seed_from_u64(0)
    assert int(k[0]) | (int(k[1]) << 32) == 5029875928683246316
    # block boundaries: the t-th u64 does not depend on where the stream is entered
    full = oracle.chacha12_u64(k, 0, 40)
    assert np.array_equal(oracle.chacha12_u64(k, 13, 27), full[13:])


def test_reservoir_streams_follow_reference_seeding(oracle):
    """stream t = Pcg32 new_seq_offset(t, t-th u64 of StdRng::seed_from_u64(seed))"""
    seeds = oracle.chacha12_u64(oracle.std_seed_from_u64(0), 0, 20)
    got = oracle.reservoir_streams(0, 3, 17, 4)
    for t in range(3, 20):
        assert np.array_equal(got[t - 3], oracle.pcg32(t, int(seeds[t]), 4))
