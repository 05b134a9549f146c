"""Helpers shared by the GPU parity tests: generate records on the GPU, bring them to the
host, decode FP16 records exactly as the kernel does, compare with the oracle."""
import numpy as np

import nlpf as N


def ramp_stops(rng, n=16, min_gap=1.0 / 64):
    """Sorted U[0,1] stop positions pushed apart to at least min_gap: a near-vertical ramp
    segment amplifies one FP32 ulp of luminance beyond any fixed per-pixel tolerance."""
    pos = np.sort(rng.random(n))
    pos = pos * (1 - min_gap * (n - 1)) + min_gap * np.arange(n)
    val = np.sort(rng.random(n))
    return pos.astype(np.float32), val.astype(np.float32)


def gen_screen(h, W, H, S, L, precision=N.FP32, seed=1, radius=7):
    desc = h.screen_desc(W, H, S, L, precision, radius)
    plane = W * H
    esz = 2 if precision == N.FP16 else 4
    rec = h.alloc(S * L * 15 * plane * esz)
    anc = None
    if precision == N.FP16:
        anc = h.alloc(S * L * ((W + 7) // 8) * ((H + 7) // 8) * 3 * 4)
    h.generate_screen_records(desc, rec, anc, seed=seed)
    return desc, rec, anc


def host_records_fp32(rec, W, H, S, L):
    return rec.download(np.float32, S * L * 15 * W * H).reshape(S, L, 15, H, W)


def decode_fp16(rec, anc, W, H, S, L):
    """FP16 planes + anchors -> the FP32 values the kernel computes with."""
    half = rec.download(np.float16, S * L * 15 * W * H).reshape(S, L, 15, H, W)
    abx, aby = (W + 7) // 8, (H + 7) // 8
    a = anc.download(np.float32, S * L * abx * aby * 3).reshape(S, L, aby, abx, 3)
    out = half.astype(np.float32)
    ay = np.arange(H) // 8
    ax = np.arange(W) // 8
    for c in range(3):
        out[:, :, c] = out[:, :, c] + a[:, :, ay][:, :, :, ax][..., c]  # float32 add, as on the GPU
    out[:, :, 9:12] *= np.float32(N.FP16_RADIANCE_SCALE)
    return out


def run_screen(h, desc, rec, anc=None, want_levels=True):
    W, H, S, L = desc.width, desc.height, desc.spp, desc.levels
    dt = np.float64 if desc.precision == N.FP64 else np.float32
    film = h.alloc(H * W * 3 * dt().itemsize)
    lv = h.alloc(S * L * 3 * H * W * dt().itemsize) if want_levels else None
    h.screen_run(desc, rec, anc, film, lv)
    f = film.download(dt, H * W * 3).reshape(H, W, 3)
    levels = lv.download(dt, S * L * 3 * H * W).reshape(S, L, 3, H, W) if want_levels else None
    film.free()
    if lv is not None:
        lv.free()
    return f, levels


def pixel_rel_err(a, b, axis=-1, floor=1e-8):
    """Per-pixel relative error: max over channels of |a-b| divided by the pixel's largest
    |b| channel (channels of one pixel share one scale; HSV remaps can drive single channels
    to ~0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max(axis=axis) / np.maximum(np.abs(b).max(axis=axis), floor)
