"""TEST INFRASTRUCTURE ONLY — numpy restatement of the reference film
(crates/akari_render/src/film.rs) and of util::write_image's pixel mapping
(crates/akari_render/src/util/mod.rs:57-121), the checker of libnlpf's film kernels and image
writers.  Imported only by tests/.

All arithmetic is FP32, one rounding per reference operation.  The reference adds samples with
float atomics, whose order depends on scheduling; this oracle (like the library) adds every
pixel's samples in input order: np.add.at applies its updates one index at a time, in order.
"""
import numpy as np

f32 = np.float32


def film_new(W, H):
    """Film::data after Film::clear (film.rs:108-110, 224-227): | pixels | splat | weights |."""
    return np.zeros(7 * W * H, f32)


def planes(film, W, H):
    P = W * H
    return film[:3 * P], film[3 * P:6 * P], film[6 * P:]


def linear_index(xy, W, H):
    """Film::linear_index (film.rs:150-160): floor of the position; off-film (or NaN) -> -1."""
    xy = np.asarray(xy, f32).reshape(-1, 2)
    x, y = np.floor(xy[:, 0]), np.floor(xy[:, 1])
    with np.errstate(invalid="ignore"):
        ok = (x >= 0) & (y >= 0) & (x < W) & (y < H)
    idx = np.full(len(xy), -1, np.int64)
    idx[ok] = x[ok].astype(np.int64) + y[ok].astype(np.int64) * W
    return idx


def remove_nan(rgb):
    """Color::remove_nan (color.rs:319-327): a colour with any NaN channel becomes zero."""
    rgb = np.array(rgb, f32).reshape(-1, 3)
    rgb[np.isnan(rgb).any(axis=1)] = 0
    return rgb


def _weighted(rgb, weight, n):
    w = np.ones(n, f32) if weight is None else np.asarray(weight, f32).reshape(-1)
    with np.errstate(invalid="ignore", over="ignore"):
        c = remove_nan(rgb) * w[:, None]
    return c, w


def add_samples(film, W, H, xy, rgb, weight=None):
    """Film::add_sample (film.rs:187-222): c = remove_nan(rgb) * w; pixels += c, weights += w.
    Updates film in place; returns the number of off-film samples."""
    idx = linear_index(xy, W, H)
    c, w = _weighted(rgb, weight, len(idx))
    px, _, wt = planes(film, W, H)
    ok = idx >= 0
    with np.errstate(invalid="ignore", over="ignore"):
        for ch in range(3):
            np.add.at(px, 3 * idx[ok] + ch, c[ok, ch])
        np.add.at(wt, idx[ok], w[ok])
    return int((~ok).sum())


def add_splats(film, W, H, xy, rgb, weight=None):
    """Film::add_splat (film.rs:162-186): c = remove_nan(rgb) * w, NaN channels of c -> 0;
    splat += c."""
    idx = linear_index(xy, W, H)
    c, _ = _weighted(rgb, weight, len(idx))
    c[np.isnan(c)] = 0
    _, sp, _ = planes(film, W, H)
    ok = idx >= 0
    with np.errstate(invalid="ignore", over="ignore"):
        for ch in range(3):
            np.add.at(sp, 3 * idx[ok] + ch, c[ok, ch])
    return int((~ok).sum())


def linear_to_srgb(v):
    """linear_to_srgb (color.rs:536-539, 590-596) in FP32."""
    v = np.asarray(v, f32)
    with np.errstate(invalid="ignore"):
        hi = np.power(v, f32(1.0) / f32(2.4)) * f32(1.055) - f32(0.055)
    return np.where(v <= f32(0.0031308), v * f32(12.92), hi).astype(f32)


def resolve(film, W, H, splat_scale=1.0, hdr=True):
    """The copy_to_rgba_image kernel (film.rs:120-147): [H, W, 4]."""
    px, sp, wt = planes(film, W, H)
    d = np.where(wt == 0, f32(1), wt).astype(f32)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        rgb = px.reshape(-1, 3) / d[:, None] + sp.reshape(-1, 3) * f32(splat_scale)
    if not hdr:
        rgb = linear_to_srgb(rgb)
    out = np.ones((W * H, 4), f32)
    out[:, :3] = rgb
    return out.reshape(H, W, 4)


def quantize_ldr(rgba):
    """write_image_ldr's mapping (util/mod.rs:85-93): (linear_to_srgb(c) * 255).clamp(0, 255)
    as u8, i.e. truncation, NaN -> 0.  Returns uint8 [H, W, 3]."""
    s = linear_to_srgb(np.asarray(rgba, f32)[..., :3]) * f32(255)
    with np.errstate(invalid="ignore"):
        s = np.where(np.isnan(s), f32(0), np.clip(s, f32(0), f32(255)))
    return np.trunc(s).astype(np.uint8)


def world_film_accumulate(film, rad0, path_depth, W, H, branch):
    """kernel_splat + kernel_update_film + Film::add_sample (lm_path_filtering.rs:428-506) for an
    empty stylization table: every path that hit adds the branch mean of its pixel's first-hit
    radiance (sum over the hitting branches in branch order, times 1/branch) with weight 1.
    rad0: first-vertex radiance [branch * W * H, 3] (path t = b * W * H + pixel)."""
    P = W * H
    rad0 = np.asarray(rad0, f32).reshape(branch, P, 3)
    hit = np.asarray(path_depth).reshape(branch, P) >= 0
    s = np.zeros((P, 3), f32)
    for b in range(branch):
        s = np.where(hit[b][:, None], s + rad0[b], s)
    L = s * (f32(1) / f32(branch))
    L[np.isnan(L).any(axis=1)] = 0
    px, _, wt = planes(film, W, H)
    px = px.reshape(P, 3)
    for b in range(branch):
        px[:] = np.where(hit[b][:, None], px + L, px)
        wt[:] = np.where(hit[b], wt + f32(1), wt)
