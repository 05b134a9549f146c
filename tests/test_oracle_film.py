"""Pins the film oracle (oracle/film_oracle.py) without the reference: a per-sample Python loop
written from film.rs's definitions, sRGB closed forms, the quantiser's known values, and the
world-mode film against a loop restatement of kernel_update_film."""
import math

import numpy as np
import pytest

import film_oracle as F

f32 = np.float32


def loop_add(W, H, xy, rgb, weight, splat):
    """film.rs:162-222 one sample at a time (floor, NaN handling, sequential FP32 adds)."""
    film = [f32(0)] * (7 * W * H)
    P = W * H
    dropped = 0
    for i in range(len(xy)):
        x, y = float(xy[i][0]), float(xy[i][1])
        if math.isnan(x) or math.isnan(y):
            dropped += 1
            continue
        ix, iy = math.floor(x), math.floor(y)
        if ix < 0 or iy < 0 or ix >= W or iy >= H:
            dropped += 1
            continue
        p = ix + iy * W
        c = [f32(v) for v in rgb[i]]
        if any(math.isnan(v) for v in c):
            c = [f32(0)] * 3
        w = f32(1) if weight is None else f32(weight[i])
        with np.errstate(invalid="ignore", over="ignore"):
            c = [f32(v * w) for v in c]
        if splat:
            c = [f32(0) if math.isnan(v) else v for v in c]
            for ch in range(3):
                film[3 * P + 3 * p + ch] = f32(film[3 * P + 3 * p + ch] + c[ch])
        else:
            for ch in range(3):
                film[3 * p + ch] = f32(film[3 * p + ch] + c[ch])
            film[6 * P + p] = f32(film[6 * P + p] + w)
    return np.array(film, f32), dropped


def samples(rng, n, W, H):
    xy = np.stack([rng.uniform(-1.5, W + 1.5, n), rng.uniform(-1.5, H + 1.5, n)], 1).astype(f32)
    xy[::17] = np.floor(xy[::17])           # exact pixel corners
    xy[5::29, 0] = -0.25                     # floor -> -1: off the film
    xy[7::31, 1] = np.nan
    rgb = rng.standard_normal((n, 3)).astype(f32) * f32(3)
    rgb[3::13, 1] = np.nan
    w = rng.uniform(0.1, 2.0, n).astype(f32)
    w[11::37] = np.nan
    w[13::41] = 0
    return xy, rgb, w


@pytest.mark.parametrize("splat", [False, True])
@pytest.mark.parametrize("weighted", [False, True])
def test_add_matches_sequential_loop(splat, weighted):
    rng = np.random.default_rng(3 + 2 * splat + weighted)
    W, H, n = 5, 4, 600
    xy, rgb, w = samples(rng, n, W, H)
    w = w if weighted else None
    want, dropped = loop_add(W, H, xy, rgb, w, splat)
    film = F.film_new(W, H)
    got = (F.add_splats if splat else F.add_samples)(film, W, H, xy, rgb, w)
    assert got == dropped > 0
    assert np.array_equal(film.view(np.uint32), want.view(np.uint32))


def test_add_accumulates_onto_existing_film():
    rng = np.random.default_rng(5)
    W, H = 3, 3
    xy, rgb, w = samples(rng, 200, W, H)
    film = F.film_new(W, H)
    F.add_samples(film, W, H, xy[:100], rgb[:100], w[:100])
    F.add_samples(film, W, H, xy[100:], rgb[100:], w[100:])
    want, _ = loop_add(W, H, xy, rgb, w, False)
    assert np.array_equal(film.view(np.uint32), want.view(np.uint32))


def test_resolve_closed_forms():
    W, H = 2, 2
    film = F.film_new(W, H)
    px, sp, wt = F.planes(film, W, H)
    px[:] = [2, 4, 6, 1, 1, 1, 3, 3, 3, 0.5, 0.25, 0.125]
    wt[:] = [2, 0, 4, 1]
    sp[:] = [1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 8]
    out = F.resolve(film, W, H, splat_scale=0.5, hdr=True)
    assert np.array_equal(out.reshape(4, 4), np.array([[1.5, 2, 3, 1], [1, 1, 1, 1], [0.75, 0.75, 0.75, 1],
                                                        [0.5, 0.25, 4.125, 1]], f32))


def test_srgb_closed_forms():
    v = np.array([0.0, 1e-4, 0.0031308, 0.0031309, 0.18, 0.5, 1.0, 4.0], f32)
    s = F.linear_to_srgb(v)
    exact = np.where(v.astype(np.float64) <= 0.0031308, v * 12.92,
                     1.055 * np.power(v.astype(np.float64), 1 / 2.4) - 0.055)
    assert np.allclose(s, exact, rtol=1e-6, atol=1e-7)
    assert s[0] == 0 and s[1] == f32(1e-4) * f32(12.92)
    # the two branches meet at the threshold (IEC 61966-2-1): one step of 1e-7 past it moves the
    # value by about the slope there (12.92) times the step
    assert abs(float(s[3]) - float(s[2]) - 12.92e-7) < 2e-7
    # FP32 arithmetic of the reference: 1.055f - 0.055f rounds below one
    assert s[6] == f32(0.99999994)


def test_quantizer_known_values():
    # (linear, 8-bit code) of write_image_ldr; linear 1.0 maps to 254 because 1.055f * 1 - 0.055f
    # is 0.99999994 in FP32 and the cast truncates (rounding, or FP64, would give 255)
    cases = [(0.0, 0), (-1.0, 0), (np.nan, 0), (np.inf, 255), (1.0, 254), (1.01, 255), (0.5, 187),
             (0.2, 123), (0.0031308, 10)]
    rgba = np.array([[[v, v, v, 1.0] for v, _ in cases]], f32)
    q = F.quantize_ldr(rgba)[0, :, 0]
    assert q.tolist() == [c for _, c in cases]


def test_world_film_matches_loop():
    rng = np.random.default_rng(8)
    W, H, B = 3, 2, 3
    P = W * H
    rad0 = rng.random((B * P, 3)).astype(f32)
    rad0[4] = np.nan
    depth = rng.integers(-1, 3, B * P)
    film = F.film_new(W, H)
    F.world_film_accumulate(film, rad0, depth, W, H, B)
    F.world_film_accumulate(film, rad0[::-1].copy(), depth[::-1].copy(), W, H, B)
    want = np.zeros(7 * P, f32)
    for r, dp in ((rad0, depth), (rad0[::-1], depth[::-1])):
        for p in range(P):
            s = [f32(0)] * 3
            for b in range(B):
                if dp[b * P + p] >= 0:
                    s = [f32(s[c] + r[b * P + p][c]) for c in range(3)]
            L = [f32(v * (f32(1) / f32(B))) for v in s]
            if any(math.isnan(v) for v in L):
                L = [f32(0)] * 3
            for b in range(B):
                if dp[b * P + p] >= 0:
                    for c in range(3):
                        want[3 * p + c] = f32(want[3 * p + c] + L[c])
                    want[6 * P + p] = f32(want[6 * P + p] + 1)
    assert np.array_equal(film.view(np.uint32), want.view(np.uint32))
