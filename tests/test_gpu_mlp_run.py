"""GPU tests of the screen-mode NL-PF with the learned cache (nlpf_screen_run_mlp).

What the pass must do, checked piece by piece:
  * every tile falling back (threshold < 0) gives the plain NL-PF result;
  * with no fallback, the audit slice is the plain result, the other slices hold the stylized
    network estimate at the cache depth, and the levels above it are the plain NL-PF applied to
    that mixed depth-k output (FP64 oracle on records whose depth k-1 radiance carries the
    GPU's depth-k values: in_{k-1} = radiance + throughput * out_k);
  * with a threshold between tile errors, exactly the tiles whose relative L2 error on the
    audit slice exceeds it are filtered exactly (halo included), the rest keep the estimate;
  * training on the audit slice makes the estimate better than truncating the paths at depth k;
  * the pass is deterministic.
FP32 records, tolerance 1e-5 relative per pixel against the FP64 oracle."""
import numpy as np
import pytest

import nlpf as N
import oracle as O
from gpu_util import gen_screen, host_records_fp32, pixel_rel_err, ramp_stops, run_screen

pytestmark = pytest.mark.gpu
SIG = (0.05, 0.9, 0.5)
W, H, S, L, K, R = 150, 70, 3, 4, 2, 7
PLANE = W * H


@pytest.fixture(scope="module")
def h():
    with N.Handle({"type": "pf", "mode": "screen"}) as hh:
        hh.set_sigmas(*SIG)
        hh.mlp_init(5)
        yield hh


@pytest.fixture(scope="module")
def recs(h):
    desc, rec, _ = gen_screen(h, W, H, S, L, N.FP32, seed=17)
    yield desc, rec, host_records_fp32(rec, W, H, S, L)
    rec.free()


def style(h, on):
    if on:
        pos, val = ramp_stops(np.random.default_rng(8))
        table = N.style_table([N.ramp(pos, val), N.hue_shift(0.1, inner=0)])
        h.set_style(table, [1])
        return table, 1
    h.set_style(N.style_table([N.toon(0.25)]), [-1])
    return None, -1


def run_mlp(h, desc, rec, **kw):
    film = h.alloc(PLANE * 3 * 4)
    lv = h.alloc(S * L * 3 * PLANE * 4)
    st = h.screen_run_mlp(desc, rec, None, film, lv, depth=K, **kw)
    f = film.download(np.float32, PLANE * 3).reshape(H, W, 3)
    levels = lv.download(np.float32, S * L * 3 * PLANE).reshape(S, L, 3, H, W)
    film.free()
    lv.free()
    return f, levels, st


def predict(h, desc, rec, slices):
    idx = np.concatenate([s * PLANE + np.arange(PLANE) for s in slices]).astype(np.uint32)
    didx = h.to_device(idx)
    out = h.alloc(len(idx) * 12)
    h.mlp_infer(desc, rec, None, K, didx, None, len(idx), out)
    y = out.download(np.float32, 3 * len(idx)).reshape(len(slices), H, W, 3)
    didx.free()
    out.free()
    return np.moveaxis(y, -1, 1)  # (slices, 3, H, W)


def check_upper_levels(host, lv, film, table, root, tol=1e-5, radius=R):
    """Levels k-1..0 and the film from the FP64 oracle run on the GPU's depth-k output."""
    mod = host[:, :K].copy()
    thr = mod[:, K - 1, 12:15].astype(np.float64)
    mod[:, K - 1, 9:12] = (mod[:, K - 1, 9:12].astype(np.float64) + thr * lv[:, K]).astype(np.float32)
    of, olv = O.screen_run(mod, W, H, S, K, radius, SIG, nodes=table, root=root, want_levels=True)
    e = pixel_rel_err(np.moveaxis(lv[:, :K], 2, -1), np.moveaxis(olv, 2, -1))
    assert e.max() <= tol, e.max()
    assert pixel_rel_err(film, of).max() <= tol


def tile_rel_l2(pred, ref):
    """Per 64x8 tile relative L2 error (pred, ref: (3, H, W)), as the library defines it."""
    tx, ty = (W + 63) // 64, (H + 7) // 8
    out = np.zeros((ty, tx))
    for j in range(ty):
        for i in range(tx):
            p = pred[:, 8 * j:8 * j + 8, 64 * i:64 * i + 64].astype(np.float64)
            r = ref[:, 8 * j:8 * j + 8, 64 * i:64 * i + 64].astype(np.float64)
            num, den = ((p - r) ** 2).sum(), (r ** 2).sum()
            # all-sky tiles filter to exactly 0 on both sides: error 0
            out[j, i] = np.sqrt(num / den) if den > 0 else (np.inf if num > 0 else 0.0)
    return out


def test_all_fallback_equals_plain(h, recs):
    desc, rec, _ = recs
    style(h, True)
    film0, lv0 = run_screen(h, desc, rec)
    film, lv, st = run_mlp(h, desc, rec, audit_slice=1, train_records=8192, batch=1024, tile_threshold=-1.0)
    assert st["fallback_tiles"] == st["tiles"] == ((W + 63) // 64) * ((H + 7) // 8)
    assert pixel_rel_err(np.moveaxis(lv[:, :K + 1], 2, -1), np.moveaxis(lv0[:, :K + 1], 2, -1)).max() <= 1e-6
    assert pixel_rel_err(film, film0).max() <= 1e-6


def test_no_fallback_uses_the_cache(h, recs):
    desc, rec, host = recs
    table, root = style(h, True)
    h.mlp_init(5)
    film0, lv0 = run_screen(h, desc, rec)
    sj = 2
    film, lv, st = run_mlp(h, desc, rec, audit_slice=sj, train_records=8192, batch=1024, tile_threshold=1e30)
    assert st["fallback_tiles"] == 0 and np.isfinite(st["train_loss"]) and np.isfinite(st["audit_rel_l2"])
    # the audit slice is filtered exactly at every level
    assert pixel_rel_err(np.moveaxis(lv[sj], 1, -1), np.moveaxis(lv0[sj], 1, -1)).max() <= 1e-6
    # the other slices hold the stylized estimate of the trained network
    others = [s for s in range(S) if s != sj]
    pred = predict(h, desc, rec, others)
    sty = O.stylize(table, [root], np.moveaxis(pred, 1, -1).reshape(-1, 3), depth=K)
    sty = sty.reshape(len(others), H, W, 3)
    assert pixel_rel_err(np.moveaxis(lv[others, K], 1, -1), sty).max() <= 1e-5
    check_upper_levels(host, lv, film, table, root)


@pytest.mark.parametrize("estimator", ["zero", "mean"])
def test_control_estimators(h, recs, estimator):
    """The controls the cache is measured against, run through the same pass: "zero" cuts the
    paths at depth k (out_k = 0, not stylized), "mean" puts the audit slice's mean exact
    unstylized depth-k value at every vertex, stylized at depth k.  The audit slice stays exact
    and the levels above k are the plain NL-PF of that depth-k output."""
    desc, rec, host = recs
    table, root = style(h, True)
    sj = 1
    film0, lv0 = run_screen(h, desc, rec)
    film, lv, st = run_mlp(h, desc, rec, audit_slice=sj, train_records=8192, batch=1024, tile_threshold=1e30,
                           estimator=estimator)
    assert st["fallback_tiles"] == 0
    assert pixel_rel_err(np.moveaxis(lv[sj], 1, -1), np.moveaxis(lv0[sj], 1, -1)).max() <= 1e-6
    others = [s for s in range(S) if s != sj]
    if estimator == "zero":
        assert (lv[others, K] == 0).all()
    else:
        # exact unstylized depth-k values of the audit slice: the same tree with depth k unstylized
        off = table.copy()
        off[0]["depth_mask"] = off[1]["depth_mask"] = N.ALL_DEPTHS & ~(1 << K)
        _, ulv = O.screen_run(host[sj:sj + 1], W, H, 1, L, R, SIG, nodes=off, root=root, want_levels=True)
        c = ulv[0, K].reshape(3, -1).mean(1)
        want = O.stylize(table, [root], c[None, :], depth=K)[0]
        got = np.moveaxis(lv[others, K], 1, -1).reshape(-1, 3)
        assert pixel_rel_err(got, np.broadcast_to(want, got.shape)).max() <= 1e-5
        assert np.ptp(got, axis=0).max() == 0  # one value everywhere
    # levels k-1..0: the plain NL-PF of the depth-k output (zero: the paths cut at depth k)
    check_upper_levels(host, lv, film, table, root)


def dilate(mask, times):
    from scipy.ndimage import binary_dilation
    return binary_dilation(mask, structure=np.ones((3, 3), bool), iterations=times) if times else mask


@pytest.mark.parametrize("radius", [R, 3])
def test_partial_fallback_selects_tiles(h, recs, radius):
    """At radius 3 the plain levels run 64x12 tiles while the masked fallback launches run the
    64x8 tiles the mask describes."""
    _, rec, host = recs
    desc = h.screen_desc(W, H, S, L, N.FP32, radius)
    table, root = style(h, False)
    h.mlp_init(11)
    sj = 0
    film0, lv0 = run_screen(h, desc, rec)
    # lr 0 keeps the weights, so the audit's image from the estimate is known before the call:
    # levels k-1..0 of the audit slice on records carrying the estimate at depth k
    pred = predict(h, desc, rec, list(range(S)))
    mod = host[sj:sj + 1, :K].copy()
    thr_k = mod[0, K - 1, 12:15].astype(np.float64)
    mod[0, K - 1, 9:12] = (mod[0, K - 1, 9:12].astype(np.float64) + thr_k * pred[sj]).astype(np.float32)
    _, plv = O.screen_run(mod, W, H, 1, K, radius, SIG, want_levels=True)
    err = tile_rel_l2(plv[0, 0], lv0[sj, 0])
    e = np.sort(err.ravel())
    gaps = np.diff(e)
    mid = len(e) // 2
    # a threshold in the middle (fallback mask covers the image once dilated), and one just
    # below the worst tile (a single level-0 tile: its dilated depth-k mask leaves tiles out)
    for lo, hi in ((mid - 5, mid + 5), (len(gaps) - 1, len(gaps))):
        i = lo + int(np.argmax(gaps[lo:hi]))
        thr = 0.5 * (e[i] + e[i + 1])
        assert gaps[i] > 1e-3 * thr, (e[lo:hi + 1], thr)
        mask0 = err > thr
        film, lv, st = run_mlp(h, desc, rec, audit_slice=sj, train_records=4096, batch=1024, lr=0.0,
                               tile_threshold=float(thr))
        assert st["fallback_tiles"] == mask0.sum() and 0 < mask0.sum() < mask0.size
        jrel = np.linalg.norm(plv[0, 0] - lv0[sj, 0]) / np.linalg.norm(lv0[sj, 0])
        assert abs(st["audit_rel_l2"] - jrel) < 1e-4 * jrel
        mk = dilate(mask0, K)  # depth-k tiles a level-0 fallback tile depends on
        pm = np.repeat(np.repeat(mk, 8, 0), 64, 1)[:H, :W]
        for s in range(S):
            if s == sj:
                continue
            e_fb = pixel_rel_err(np.moveaxis(lv[s, K], 0, -1), np.moveaxis(lv0[s, K], 0, -1))
            e_ml = pixel_rel_err(np.moveaxis(lv[s, K], 0, -1), np.moveaxis(pred[s], 0, -1))
            assert e_fb[pm].max() <= 1e-6  # fallback tiles: exact (their halo was refiltered)
            if (~pm).any():
                assert e_ml[~pm].max() <= 1e-6  # the rest: the network's estimate
        check_upper_levels(host, lv, film, table, root, radius=radius)
        # the film inside the level-0 fallback tiles is the plain NL-PF's
        p0 = np.repeat(np.repeat(mask0, 8, 0), 64, 1)[:H, :W]
        assert pixel_rel_err(film[p0], film0[p0]).max() <= 1e-5
    assert mask0.sum() == 1 and not mk.all()


@pytest.fixture(scope="module")
def smooth_recs(h):
    """Records without fireflies and with milder noise: the expectation the cache learns is not
    buried under per-vertex outliers no estimator can predict."""
    desc = h.screen_desc(W, H, S, L, N.FP32, R)
    rec = h.alloc(S * L * 15 * PLANE * 4)
    h.generate_screen_records(desc, rec, seed=23, sigma=0.5, firefly_prob=0.0)
    yield desc, rec, host_records_fp32(rec, W, H, S, L)
    rec.free()


def test_training_beats_truncation(h, smooth_recs):
    """Film relative L2 against the exact NL-PF: the trained cache at depth k must beat cutting
    the paths at depth k (what the pass costs without a network) and its own untrained start."""
    desc, rec, host = smooth_recs
    style(h, False)
    film0, _ = run_screen(h, desc, rec)
    h.mlp_init(3)
    film_un, _, _ = run_mlp(h, desc, rec, audit_slice=0, train_records=8192, batch=1024, lr=0.0,
                            tile_threshold=1e30)
    st = None
    for frame in range(4):  # training continues across frames
        film_ml, _, st = run_mlp(h, desc, rec, audit_slice=0, train_records=8192, batch=1024, tile_threshold=1e30,
                                 seed=frame)
    cut = host[:, :K].copy()
    cut[:, K - 1, 12:15] = 0
    film_tr = O.screen_run(cut, W, H, S, K, R, SIG)
    film_mean, _, _ = run_mlp(h, desc, rec, audit_slice=0, train_records=8192, batch=1024, tile_threshold=1e30,
                              estimator="mean")
    exact = film0.astype(np.float64)

    def rel(x):
        return np.linalg.norm(x - exact) / np.linalg.norm(exact)

    e_ml, e_tr, e_un, e_mean = rel(film_ml), rel(film_tr), rel(film_un), rel(film_mean)
    print(f"film rel L2: cache {e_ml:.4f}, truncation {e_tr:.4f}, untrained {e_un:.4f}, audit-slice mean "
          f"{e_mean:.4f}; cache fit {st['cache_rel_l2']:.4f}, audit {st['audit_rel_l2']:.4f}")
    assert e_ml < e_tr and e_ml < e_un


def test_deterministic(h, recs):
    desc, rec, _ = recs
    style(h, True)
    outs = []
    for _ in range(2):
        h.mlp_init(9)
        f, lv, st = run_mlp(h, desc, rec, audit_slice=1, train_records=8192, batch=1024, tile_threshold=0.5)
        outs.append((f, lv[:, :K + 1], st["fallback_tiles"], st["train_loss"]))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2:] == outs[1][2:]


def test_invalid_arguments(h, recs):
    desc, rec, _ = recs
    film = h.alloc(PLANE * 12)
    for kw in ({"depth": 0}, {"depth": L}, {"audit_slice": S}, {"batch": 100}, {"train_records": 3000, "batch": 1024},
               {"train_records": 2 * PLANE, "batch": 1024}):
        args = dict(depth=K, audit_slice=0, train_records=4096, batch=1024)
        args.update(kw)
        with pytest.raises(N.NlpfError):
            h.screen_run_mlp(desc, rec, None, film, None, **args)
    film.free()
