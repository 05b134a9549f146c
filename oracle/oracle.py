"""TEST INFRASTRUCTURE ONLY — ctypes loader of the CPU oracle (oracle/liboracle.so).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker.  Built by oracle/Makefile (gcc).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NLPF_ORACLE_LIB selects another build of the same sources (the host-sanitizer build)
LIB_PATH = os.environ.get("NLPF_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    lib.or_apply_style.argtypes = [vp, i32, u32, u32, vp, vp, vp]
    lib.or_update_flags.argtypes = [vp, i32, u32]
    lib.or_update_flags.restype = u32
    lib.or_rgb_to_hsv.argtypes = [vp, vp]
    lib.or_hsv_to_rgb.argtypes = [vp, vp]
    lib.or_stylize_batch.argtypes = [vp, vp, u32, vp, vp, u64, vp, u32, vp, vp, vp]
    for name in ("or_screen_run_f64", "or_screen_run_f32"):
        fn = getattr(lib, name)
        fn.argtypes = [u32, u32, u32, u32, i32, dbl, dbl, dbl, vp, vp, i32, vp, vp, i32]
        fn.restype = i32
    for name in ("or_screen_run_shaded_f64", "or_screen_run_shaded_f32"):
        fn = getattr(lib, name)
        fn.argtypes = [u32, u32, u32, u32, i32, dbl, dbl, dbl, vp, vp, i32, vp, u32, vp, vp, vp, vp, i32]
        fn.restype = i32
    lib.or_world_neighbor_counts.argtypes = [vp, u32, vp, vp, C.c_float, u64, vp]
    lib.or_world_neighbor_counts.restype = i32
    lib.or_world_run.argtypes = [u32, u32, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp, u32, vp, u32, u64, i32, i32,
                                 i32, i32, u64, u64, vp, vp, u32, vp, vp, vp, vp, vp, u32]
    lib.or_ggx_d.argtypes = [dbl, dbl]
    lib.or_ggx_d.restype = dbl
    lib.or_ggx_lambda.argtypes = [dbl, dbl]
    lib.or_ggx_lambda.restype = dbl
    lib.or_fr_complex.argtypes = [dbl, dbl, dbl]
    lib.or_fr_complex.restype = dbl
    lib.or_conductor_nk.argtypes = [dbl, dbl, vp, vp]
    lib.or_bsdf_eval.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.or_bsdf_eval.restype = dbl
    lib.or_world_run.restype = i32
    lib.or_screen_grad_f64.argtypes = [u32, u32, u32, u32, i32, vp, vp, vp, u32, i32, vp, vp, vp]
    lib.or_screen_grad_f64.restype = i32
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# world-mode BSDF table entry (the library's nlpf_bsdf layout); kinds
BSDF_DTYPE = np.dtype([("kind", "<u4"), ("metallic", "<f4"), ("roughness", "<f4"), ("specular_tint", "<f4", 3),
                       ("specular_ior_level", "<f4"), ("transmission", "<f4"), ("coat_weight", "<f4"),
                       ("reserved", "<u4")])
BSDF_DIFFUSE, BSDF_PRINCIPLED, BSDF_EMISSION = 0, 1, 2


def bsdf(kind=BSDF_PRINCIPLED, metallic=1.0, roughness=0.5, specular_tint=(1.0, 1.0, 1.0), specular_ior_level=0.0,
         transmission=0.0, coat_weight=0.0):
    e = np.zeros(1, BSDF_DTYPE)
    e["kind"], e["metallic"], e["roughness"] = kind, metallic, roughness
    e["specular_tint"] = specular_tint
    e["specular_ior_level"], e["transmission"], e["coat_weight"] = specular_ior_level, transmission, coat_weight
    return e


def ggx_d(cos_h, alpha):
    return load().or_ggx_d(float(cos_h), float(alpha))


def ggx_lambda(cos_w, alpha):
    return load().or_ggx_lambda(float(cos_w), float(alpha))


def fr_complex(cos_i, n, k):
    return load().or_fr_complex(float(cos_i), float(n), float(k))


def conductor_nk(r, g):
    n, k = np.zeros(1), np.zeros(1)
    load().or_conductor_nk(float(r), float(g), _p(n), _p(k))
    return float(n[0]), float(k[0])


def bsdf_eval(entry, R, ns, ng, wo, wi):
    """(pdf, f |cos_i| per channel) of one table entry (None: diffuse) for world-space vectors."""
    fc = np.zeros(3)
    e = None if entry is None else np.ascontiguousarray(entry, BSDF_DTYPE)
    v = [np.ascontiguousarray(x, np.float64) for x in (R, ns, ng, wo, wi)]
    pdf = load().or_bsdf_eval(_p(e), *[_p(x) for x in v], _p(fc))
    return pdf, fc


def stylize(nodes, roots, rgb, depth=0, flags=None, material=None, pos=None):
    lib = load()
    rgb = np.ascontiguousarray(rgb, np.float64).reshape(-1, 3)
    n = len(rgb)
    out = np.empty_like(rgb)
    dd, d0 = (None, int(depth)) if np.ndim(depth) == 0 else (np.ascontiguousarray(depth, np.uint32), 0)
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    mt = None if material is None else np.ascontiguousarray(material, np.uint16)
    ps = None if pos is None else np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    nodes = np.ascontiguousarray(nodes)
    roots = np.ascontiguousarray(roots, np.int32)
    lib.or_stylize_batch(_p(nodes), _p(roots), len(roots), _p(rgb), _p(out), n, _p(dd), d0, _p(fl),
                         _p(mt), _p(ps))
    return out


def update_flags(nodes, root, flags):
    return int(load().or_update_flags(_p(np.ascontiguousarray(nodes)), int(root), int(flags)))


def rgb_to_hsv(c):
    c = np.ascontiguousarray(c, np.float64)
    o = np.empty(3)
    load().or_rgb_to_hsv(_p(c), _p(o))
    return o


def hsv_to_rgb(c):
    c = np.ascontiguousarray(c, np.float64)
    o = np.empty(3)
    load().or_hsv_to_rgb(_p(c), _p(o))
    return o


def screen_run(records, W, H, S, L, radius, sigmas, nodes=None, root=-1, precision="f64",
               want_levels=False, threads=0, material=None, flags=None, roots=None):
    """records: float32 planes [S][L][15][H][W] with absolute positions.  With material
    (uint16 [S][L][H][W]) each vertex takes roots[material]'s tree with flags (uint8, or 0)."""
    lib = load()
    records = np.ascontiguousarray(records, np.float32)
    assert records.size == S * L * 15 * H * W
    dt = np.float64 if precision == "f64" else np.float32
    film = np.empty((H, W, 3), dt)
    lv = np.empty((S, L, 3, H, W), dt) if want_levels else None
    if nodes is None or len(nodes) == 0:
        nodes_arr, root = np.zeros(1, dtype=[("b", "V256")]), -1
    else:
        nodes_arr = np.ascontiguousarray(nodes)
    if material is None:
        fn = lib.or_screen_run_f64 if precision == "f64" else lib.or_screen_run_f32
        rc = fn(W, H, S, L, radius, sigmas[0], sigmas[1], sigmas[2], _p(records), _p(nodes_arr), int(root),
                _p(film), _p(lv), int(threads))
    else:
        material = np.ascontiguousarray(material, np.uint16)
        assert material.size == S * L * H * W
        if flags is not None:
            flags = np.ascontiguousarray(flags, np.uint8)
            assert flags.size == S * L * H * W
        roots = np.ascontiguousarray(roots if roots is not None else [], np.int32)
        if nodes is None or len(nodes) == 0:
            roots = np.full(len(roots), -1, np.int32)
        fn = lib.or_screen_run_shaded_f64 if precision == "f64" else lib.or_screen_run_shaded_f32
        rc = fn(W, H, S, L, radius, sigmas[0], sigmas[1], sigmas[2], _p(records), _p(nodes_arr), -1,
                _p(roots), len(roots), _p(material), _p(flags), _p(film), _p(lv), int(threads))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return (film, lv) if want_levels else film


def pcg32(seq, offset, n):
    """n draws of Pcg32 new_seq_offset(seq, offset)."""
    out = np.empty(n, np.uint32)
    load().or_pcg32_fill(C.c_uint64(seq), C.c_uint64(offset), C.c_uint32(n), _p(out))
    return out


def chacha_block(words, rounds):
    inp = np.ascontiguousarray(words, np.uint32)
    out = np.empty(16, np.uint32)
    load().or_chacha_block(_p(inp), C.c_int(rounds), _p(out))
    return out


def std_seed_from_u64(state):
    key = np.empty(8, np.uint32)
    load().or_std_seed_from_u64(C.c_uint64(state), _p(key))
    return key


def chacha12_u64(key, first, n):
    key = np.ascontiguousarray(key, np.uint32)
    out = np.empty(n, np.uint64)
    load().or_chacha12_u64(_p(key), C.c_uint64(first), C.c_uint64(n), _p(out))
    return out


def reservoir_streams(seed, first, count, draws):
    out = np.empty((count, draws), np.uint32)
    rc = load().or_reservoir_streams(C.c_uint64(seed), C.c_uint64(first), C.c_uint64(count), C.c_uint32(draws),
                                     _p(out))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out


def world_neighbor_counts(pos, pmin, pmax, radius, table):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    out = np.empty(len(pos), np.uint32)
    pmin = np.ascontiguousarray(pmin, np.float32)
    pmax = np.ascontiguousarray(pmax, np.float32)
    rc = load().or_world_neighbor_counts(_p(pos), len(pos), _p(pmin), _p(pmax), float(radius), int(table), _p(out))
    assert rc == 0
    return out


def world_run(rec, flags, new_flags, material, path_depth, W, H, branch, L, aabb_min, aabb_max, cfg,
              nodes=None, roots=(), want_radiance=False, want_clusters=False, bsdfs=None):
    """cfg: dict with the "pf" fields (radius_scale, reuse, vram, smis, no_first_hit_stylization,
    disable_flag_check, cluster, seed, hash_table_size).  want_clusters adds the cluster ids
    [L][N] (0xFFFFFFFF off-level) to the returned tuple.  bsdfs: BSDF_DTYPE table by material
    (materials without an entry are diffuse)."""
    bs = None if bsdfs is None or len(bsdfs) == 0 else np.ascontiguousarray(bsdfs, BSDF_DTYPE)
    lib = load()
    N = W * H * branch
    rec = np.ascontiguousarray(rec, np.float32)
    assert rec.size == L * 28 * N
    film = np.empty((H, W, 3), np.float64)
    rad = np.empty((L, N, 3), np.float64) if want_radiance else None
    nbr = np.zeros(32, np.uint64)
    cl = np.empty((L, N), np.uint32) if want_clusters else None
    rs = np.ascontiguousarray(cfg.get("radius_scale", [0.1, 0.05]), np.float32)
    ru = np.ascontiguousarray(cfg.get("reuse", [128, 64]), np.uint32)
    if nodes is None or len(nodes) == 0:
        nodes_arr = np.zeros(1, dtype=[("b", "V256")])
        roots = np.full(1, -1, np.int32)
    else:
        nodes_arr = np.ascontiguousarray(nodes)
        roots = np.ascontiguousarray(roots, np.int32)
    rc = lib.or_world_run(W, H, branch, L, _p(np.ascontiguousarray(aabb_min, np.float32)),
                          _p(np.ascontiguousarray(aabb_max, np.float32)), _p(rec),
                          _p(np.ascontiguousarray(flags, np.uint8)), _p(np.ascontiguousarray(new_flags, np.uint8)),
                          _p(np.ascontiguousarray(material, np.uint16)),
                          _p(np.ascontiguousarray(path_depth, np.int32)), _p(rs), len(rs), _p(ru), len(ru),
                          int(cfg.get("vram", 2048)), int(cfg.get("smis", True)),
                          int(cfg.get("no_first_hit_stylization", False)), int(cfg.get("disable_flag_check", False)),
                          int(cfg.get("cluster", True)), int(cfg.get("seed", 0)),
                          int(cfg.get("hash_table_size", 64 * 1024 * 1024)),
                          _p(nodes_arr), _p(roots), len(roots), _p(film), _p(rad), _p(nbr), _p(cl),
                          _p(bs), 0 if bs is None else len(bs))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    out = (film, rad, nbr) if want_radiance else (film,)
    if want_clusters:
        out = out + (cl,)
    return out if len(out) > 1 else out[0]


def screen_grad(records, W, H, S, L, radius, sigmas, nodes, root, dfilm):
    """FP64 reference gradient: [dL/dsigma_pos, dL/dsigma_normal, dL/dsigma_lum, ramp stops...]."""
    lib = load()
    records = np.ascontiguousarray(records, np.float32)
    nodes = np.ascontiguousarray(nodes)
    nr = int((nodes["type"] == 8).sum()) if len(nodes) else 0
    grad = np.zeros(3 + 16 * nr)
    sig = np.ascontiguousarray(sigmas, np.float64)
    dfilm = np.ascontiguousarray(dfilm, np.float64)
    n = lib.or_screen_grad_f64(W, H, S, L, radius, _p(sig), _p(records), _p(nodes), len(nodes), int(root),
                               _p(dfilm), _p(grad), None)
    if n < 0:
        raise MemoryError("oracle allocation failed")
    return grad
