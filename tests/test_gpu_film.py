"""GPU parity of the film (crates/akari_render/src/film.rs) against the numpy film oracle.
Sample and splat accumulation must match bit for bit: the library adds each pixel's samples in
input order, as the oracle does, so sums are the sequential FP32 sums. The HDR resolve and the
world-mode film accumulation across passes must also match bit for bit. The sRGB resolve is
allowed 2 ulp of its power term, because the GPU's powf and the host's may differ there. The
image is written from device memory and read back."""
import numpy as np
import pytest
from PIL import Image

import film_oracle as F
import nlpf as N
from test_oracle_film import samples

pytestmark = pytest.mark.gpu
f32 = np.float32


def same(a, b):
    """Bitwise equality, with any NaN equal to any NaN."""
    a, b = np.asarray(a, f32).ravel(), np.asarray(b, f32).ravel()
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32)))


@pytest.mark.parametrize("weighted", [False, True])
def test_samples_and_splats_match_oracle(weighted):
    rng = np.random.default_rng(11 + weighted)
    W, H = 13, 7
    ref = F.film_new(W, H)
    with N.Handle() as h:
        film = h.film_alloc(W, H)
        for k, splat in enumerate((False, True, False, True)):
            xy, rgb, w = samples(rng, 4000 + 1000 * k, W, H)
            w = w if weighted else None
            if splat:
                got = h.film_add_splats(film, W, H, xy, rgb, w)
                want = F.add_splats(ref, W, H, xy, rgb, w)
            else:
                got = h.film_add_samples(film, W, H, xy, rgb, w)
                want = F.add_samples(ref, W, H, xy, rgb, w)
            assert got == want > 0
        out = film.download(np.float32, N.film_floats(W, H))
    assert same(out, ref)


def test_large_batch_and_one_long_run():
    """2^21 samples over a 1024x512 film plus 100000 samples on one pixel (one thread adds a
    long run in order)."""
    rng = np.random.default_rng(5)
    W, H, n = 1024, 512, 1 << 21
    xy = np.stack([rng.uniform(-2, W + 2, n), rng.uniform(-2, H + 2, n)], 1).astype(f32)
    rgb = rng.random((n, 3)).astype(f32)
    hot = np.tile(np.array([[3.5, 2.25]], f32), (100000, 1))
    hot_rgb = rng.random((100000, 3)).astype(f32)
    ref = F.film_new(W, H)
    d0 = F.add_samples(ref, W, H, xy, rgb)
    F.add_samples(ref, W, H, hot, hot_rgb)
    with N.Handle() as h:
        film = h.film_alloc(W, H)
        assert h.film_add_samples(film, W, H, xy, rgb) == d0 > 0
        assert h.film_add_samples(film, W, H, hot, hot_rgb) == 0
        out = film.download(np.float32, N.film_floats(W, H))
    assert same(out, ref)
    _, _, wt = F.planes(out, W, H)
    assert wt.sum(dtype=np.float64) == n - d0 + 100000


def test_resolve_hdr_exact_and_srgb():
    rng = np.random.default_rng(2)
    W, H = 97, 33
    film = rng.uniform(-0.5, 3.0, N.film_floats(W, H)).astype(f32)
    _, sp, wt = F.planes(film, W, H)
    wt[::7] = 0
    wt[1::11] = rng.uniform(1, 64, len(wt[1::11])).astype(f32).round()
    sp[5::13] = np.nan
    with N.Handle() as h:
        d = h.to_device(film)
        for scale in (1.0, 0.25):
            assert same(h.film_resolve(d, W, H, scale, hdr=True), F.resolve(film, W, H, scale, hdr=True))
            g = h.film_resolve(d, W, H, scale, hdr=False)
            o = F.resolve(film, W, H, scale, hdr=False)
            assert np.array_equal(np.isnan(g), np.isnan(o))
            m = ~np.isnan(o)
            g, o = g[m], o[m]
            # on the power branch the result is 1.055 p - 0.055 with p = v^(1/2.4); the GPU's and
            # the host's powf may differ by 2 ulp of p, which the subtraction does not shrink
            p = (np.abs(o) + f32(0.055)) / f32(1.055)
            tol = 2 * f32(1.055) * np.spacing(p) + np.spacing(np.abs(o))
            assert (np.abs(g - o) <= tol).all(), float(np.max(np.abs(g - o) / np.spacing(p)))


def world_pass(h, desc, seed, film):
    recs = h.alloc_world(desc)
    h.generate_world_records(desc, recs, seed=seed)
    _, rad = h.world_run(desc, recs, want_radiance=True, film=film)
    pd = recs["path_depth"].download(np.int32)
    for b in recs.values():
        b.free()
    return rad[0], pd


def test_world_film_accumulates_passes():
    """Three world-mode passes (different records) accumulate into one film exactly as the
    reference's repeats loop adds samples (empty stylization table: the sample of a path is the
    branch mean of its pixel's first-hit radiance)."""
    W, H, branch, L = 40, 24, 2, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": False, "radius_scale": [2.0],
           "reuse": [1 << 16], "no_first_hit_stylization": True}
    ref = F.film_new(W, H)
    with N.Handle(cfg) as h:
        desc = h.world_desc(W, H, branch, L, film_accumulate=1)
        film = h.film_alloc(W, H)
        for seed in (1, 2, 3):
            rad0, pd = world_pass(h, desc, seed, film)
            F.world_film_accumulate(ref, rad0, pd, W, H, branch)
        out = film.download(np.float32, N.film_floats(W, H))
    assert same(out, ref)
    _, _, wt = F.planes(out, W, H)
    assert wt.max() == 3 * branch and wt.min() < wt.max()  # some camera rays miss


def test_world_single_pass_resolve_equals_pass_film():
    """With a stylization table: one accumulated pass, resolved, is the pass film bit for bit."""
    W, H, branch, L = 32, 20, 3, 3
    cfg = {"type": "pf", "max_depth": L - 1, "branch": branch, "cluster": False, "radius_scale": [2.0],
           "reuse": [1 << 16]}
    nodes = N.style_table([N.saturation(0.6), N.toon(0.25, inner=0)])
    roots = [1, 1, 1, 1, 1, 0, 0, -1]
    with N.Handle(cfg) as h:
        h.set_style(nodes, roots)
        recs = h.alloc_world(h.world_desc(W, H, branch, L))
        h.generate_world_records(h.world_desc(W, H, branch, L), recs, seed=4)
        pass_film, _ = h.world_run(h.world_desc(W, H, branch, L), recs)
        film = h.film_alloc(W, H)
        h.world_run(h.world_desc(W, H, branch, L, film_accumulate=1), recs, film=film)
        rgba = h.film_resolve(film, W, H, 1.0, hdr=True)
    assert same(rgba[..., :3], pass_film)
    assert (rgba[..., 3] == 1).all()


def test_write_image_from_device(tmp_path):
    rng = np.random.default_rng(4)
    W, H = 45, 17
    film = rng.uniform(0, 2, N.film_floats(W, H)).astype(f32)
    with N.Handle() as h:
        d = h.to_device(film)
        rgba = h.alloc(16 * W * H)
        h.film_resolve(d, W, H, 0.5, hdr=True, rgba=rgba)
        host = rgba.download(np.float32).reshape(H, W, 4)
        h.write_image(rgba, W, H, str(tmp_path / "a" / "img.png"))
        h.write_image(rgba, W, H, str(tmp_path / "img.exr"))
    with Image.open(tmp_path / "a" / "img.png") as im:
        assert np.array_equal(np.asarray(im), F.quantize_ldr(host))
    assert same(host, F.resolve(film, W, H, 0.5, hdr=True))
    from test_image_io import read_exr
    names, ch = read_exr(tmp_path / "img.exr")
    assert names == ["B", "G", "R"]
    for c, name in enumerate("RGB"):
        assert same(ch[name], host[..., c])


def test_film_argument_errors():
    with N.Handle() as h:
        f = h.film_alloc(4, 4)
        lib = h.lib
        assert lib.nlpf_film_clear(h.h, None, 4, 4) == 2
        assert lib.nlpf_film_clear(h.h, N._ptr(f), 0, 4) == 2
        assert lib.nlpf_film_add_samples(h.h, N._ptr(f), 4, 4, None, None, None, 3, None) == 2
        assert lib.nlpf_film_add_samples(h.h, N._ptr(f), 4, 4, None, None, None, 0, None) == 0
        assert lib.nlpf_film_resolve(h.h, N._ptr(f), 4, 4, 1.0, 1, None) == 2
        assert lib.nlpf_film_resolve(h.h, N._ptr(f), 1 << 16, 1 << 15, 1.0, 1, N._ptr(f)) == 2
        assert "width * height" in lib.nlpf_last_error(h.h).decode()
        assert h.film_add_samples(f, 4, 4, np.zeros((0, 2), f32), np.zeros((0, 3), f32)) == 0
        with pytest.raises(ValueError):
            h.world_run(h.world_desc(4, 4, 1, 8, film_accumulate=1), {})
