"""Python host mirror of libnlpf's C ABI (ctypes).

This is the binding a Python caller (tests, bench.py) uses; a Rust caller binds the same
symbols through an `extern "C"` block over include/nlpf.h.  It keeps the reference
integrator's configuration surface: `Config` carries the fields and defaults of
`lm_path_filtering::Config` (crates/akari_integrator/src/lm_path_filtering.rs:264-298).

The library must be built (practical-stylized_amd/build.py); loading fails loudly otherwise —
there is no CPU fallback.
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NLPF_LIB selects another build of the same sources (the host-sanitizer build)
LIB_PATH = os.environ.get("NLPF_LIB") or os.path.join(HERE, "libnlpf.so")

NLPF_OK = 0
FP32, FP16, FP64 = 0, 1, 2
PRECISIONS = {"fp32": FP32, "fp16": FP16, "fp64": FP64}
SCREEN_PLANES = 15
FP16_RADIANCE_SCALE = 64.0
NOALLOC_SCRATCH_BYTES = 256 << 20  # NLPF_NOALLOC_SCRATCH_BYTES
CACHE_ESTIMATORS = {"network": 0, "zero": 1, "mean": 2}  # NLPF_CACHE_*

STY = dict(BASE=0, TOON=1, PALETTE=2, SATURATION=3, REX=4, EIGHT_BIT=5, HALFTONE=6,
           HUE_SHIFT=7, RAMP=8, ENABLE_IF=9, SET_IDX=10)
FLAG_DISABLED, FLAG_STYLE2, FLAG_STYLE3, FLAG_STYLE4 = 1, 2, 4, 8
ALL_DEPTHS = 0xFFFFFFFF

# nlpf_style_node, 256 bytes (include/nlpf.h)
STYLE_NODE_DTYPE = np.dtype([
    ("type", "<i4"), ("inner", "<i4"), ("else_", "<i4"), ("idx", "<u4"),
    ("depth_mask", "<u4"), ("max_depth", "<u4"), ("count", "<u4"), ("reserved", "<u4"),
    ("p", "<f4", (56,)),
])
assert STYLE_NODE_DTYPE.itemsize == 256

# nlpf_bsdf, 40 bytes (include/nlpf.h): a material's BSDF in world mode
BSDF_DTYPE = np.dtype([
    ("kind", "<u4"), ("metallic", "<f4"), ("roughness", "<f4"), ("specular_tint", "<f4", (3,)),
    ("specular_ior_level", "<f4"), ("transmission", "<f4"), ("coat_weight", "<f4"), ("reserved", "<u4"),
])
assert BSDF_DTYPE.itemsize == 40
BSDF_DIFFUSE, BSDF_PRINCIPLED, BSDF_EMISSION = 0, 1, 2


def bsdf_table(entries):
    """BSDF table from dicts ({"kind": ..., "metallic": ..., ...}; missing fields 0, tint 1)."""
    t = np.zeros(len(entries), BSDF_DTYPE)
    t["specular_tint"] = 1.0
    for i, e in enumerate(entries):
        for k, v in e.items():
            t[i][k] = v
    return t

# defaults of lm_path_filtering::Config (lm_path_filtering.rs:280-298) + screen-mode fields
DEFAULT_CONFIG = {
    "max_depth": 7, "repeats": 1, "branch": 1, "radius_scale": [0.1, 0.05],
    "reuse": [128, 64], "vram": 2048, "smis": True, "no_first_hit_stylization": False,
    "disable_flag_check": False, "cluster": True, "scene_bounds": None,
    "mode": "world", "pixel_radius": 7, "sigma_pos": 0.05, "sigma_normal": 0.9,
    "sigma_lum": 0.5, "precision": "fp32", "seed": 0, "hash_table_size": 64 * 1024 * 1024,
    "world_kernel": "auto", "world_grid": "auto", "screen_kernel": "auto",
}


class NlpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nlpf error {code}: {msg}")
        self.code = code


class _ScreenDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("levels", C.c_uint32), ("precision", C.c_int32), ("radius", C.c_int32)]


class _GenDesc(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("mu", C.c_float), ("sigma", C.c_float),
                ("firefly_prob", C.c_float), ("firefly_scale", C.c_float),
                ("albedo_lo", C.c_float), ("albedo_hi", C.c_float)]


class _Stats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("level_ms", C.c_double * 32), ("film_ms", C.c_double),
                ("vertices", C.c_uint64), ("bytes_algorithmic", C.c_uint64),
                ("flops_algorithmic", C.c_double), ("points", C.c_uint64 * 32),
                ("neighbors", C.c_uint64 * 32), ("smis_pairs", C.c_uint64 * 32)]


class _WorldDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("branch", C.c_uint32),
                ("levels", C.c_uint32), ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3),
                ("film_accumulate", C.c_uint32)]


class _MlpRunDesc(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("audit_slice", C.c_uint32), ("train_records", C.c_uint32),
                ("batch", C.c_uint32), ("lr", C.c_float), ("tile_threshold", C.c_float), ("seed", C.c_uint64),
                ("estimator", C.c_uint32), ("reserved", C.c_uint32)]


class _MlpRunStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("audit_ms", C.c_double), ("train_ms", C.c_double),
                ("infer_ms", C.c_double), ("fallback_ms", C.c_double), ("filter_ms", C.c_double),
                ("tiles", C.c_uint64), ("fallback_tiles", C.c_uint64), ("fallback_tiles_filtered", C.c_uint64),
                ("train_loss", C.c_float), ("cache_rel_l2", C.c_float), ("audit_rel_l2", C.c_float)]


WORLD_PLANES = 28
NLPF_E_IO = 7


def film_floats(width, height):
    """Floats of a film buffer: | pixels rgb | splat rgb | weights | (film.rs:66-67, 108-110)."""
    return 7 * int(width) * int(height)


_lib = None

# (name, restype, argtypes)
_SIGS = [
    ("nlpf_create", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("nlpf_destroy", C.c_int, [C.c_void_p]),
    ("nlpf_last_error", C.c_char_p, [C.c_void_p]),
    ("nlpf_get_config_json", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("nlpf_set_sigmas", C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    ("nlpf_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("nlpf_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("nlpf_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("nlpf_download", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("nlpf_memset", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]),
    ("nlpf_copy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("nlpf_sync", C.c_int, [C.c_void_p]),
    ("nlpf_set_style", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    ("nlpf_stylize", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                               C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nlpf_screen_run", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    ("nlpf_screen_run_materials", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nlpf_screen_run_chunk", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("nlpf_screen_run_noalloc", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    ("nlpf_get_stats", C.c_int, [C.c_void_p, C.POINTER(_Stats)]),
    ("nlpf_generate_screen_records", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc),
                                               C.POINTER(_GenDesc), C.c_void_p, C.c_void_p]),
    ("nlpf_convert_records_fp16", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p,
                                            C.c_void_p, C.c_void_p]),
    ("nlpf_screen_backward", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    ("nlpf_mlp_param_count", C.c_uint64, []),
    ("nlpf_mlp_init", C.c_int, [C.c_void_p, C.c_uint64]),
    ("nlpf_mlp_get_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nlpf_mlp_set_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nlpf_mlp_encode", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p, C.c_uint32,
                                  C.c_void_p, C.c_uint64, C.c_void_p]),
    ("nlpf_mlp_infer", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p, C.c_uint32,
                                 C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("nlpf_mlp_train_step", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_int,
                                      C.POINTER(C.c_float), C.c_void_p]),
    ("nlpf_mlp_set_ema", C.c_int, [C.c_void_p, C.c_float]),
    ("nlpf_mlp_reference", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.c_void_p, C.c_uint64, C.c_void_p]),
    ("nlpf_screen_run_mlp", C.c_int, [C.c_void_p, C.POINTER(_ScreenDesc), C.c_void_p, C.c_void_p,
                                      C.POINTER(_MlpRunDesc), C.c_void_p, C.c_void_p, C.POINTER(_MlpRunStats)]),
    ("nlpf_world_run", C.c_int, [C.c_void_p, C.POINTER(_WorldDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nlpf_generate_world_records", C.c_int, [C.c_void_p, C.POINTER(_WorldDesc), C.POINTER(_GenDesc),
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nlpf_film_clear", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("nlpf_film_add_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("nlpf_film_add_splats", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("nlpf_film_resolve", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_int,
                                    C.c_void_p]),
    ("nlpf_write_image", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_char_p]),
    ("nlpf_write_image_host", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_char_p]),
    ("nlpf_style_from_scene", C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_void_p,
                                        C.POINTER(C.c_uint32), C.c_void_p]),
    ("nlpf_write_world_scene", C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(_GenDesc), C.c_void_p,
                                         C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]),
    ("nlpf_pcg32_fill", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p]),
    ("nlpf_set_bsdfs", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("nlpf_bsdfs_from_scene", C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_void_p]),
    ("nlpf_reservoir_streams", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p]),
]


def exported_symbols():
    return [s[0] for s in _SIGS]


def load(path=LIB_PATH):
    """Load libnlpf.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libnlpf.so not built at {path}: run practical-stylized_amd/build.py")
    lib = C.CDLL(path)
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def write_image_host(rgba, path):
    """util::write_image (crates/akari_render/src/util/mod.rs:57-121) of a host image
    rgba[H, W, 4]: ".exr" -> float RGB OpenEXR, otherwise 8-bit sRGB PNG. Needs no GPU."""
    lib = load()
    rgba = np.ascontiguousarray(rgba, np.float32)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError("rgba must be [H, W, 4]")
    H, W = rgba.shape[:2]
    rc = lib.nlpf_write_image_host(rgba.ctypes.data_as(C.c_void_p), W, H, os.fsencode(path))
    if rc != NLPF_OK:
        raise NlpfError(rc, lib.nlpf_last_error(None).decode())


def pcg32(seq, offset, n):
    """n draws of Pcg32 new_seq_offset(seq, offset) (crates/akari_render/src/sampler/mod.rs:91-117),
    computed on the host by the library's own generator code. Needs no GPU."""
    out = np.empty(n, np.uint32)
    rc = load().nlpf_pcg32_fill(seq, offset, n, out.ctypes.data_as(C.c_void_p))
    if rc != NLPF_OK:
        raise NlpfError(rc, "nlpf_pcg32_fill")
    return out


def reservoir_streams(seed, first, count, draws):
    """[count][draws]: the first draws of world-mode reservoir streams first..first+count-1 for
    config seed `seed`, seeded as init_pcg32_buffer_with_seed (sampler/mod.rs:157-174) seeds
    the reference's. Host side; needs no GPU."""
    out = np.empty((count, draws), np.uint32)
    rc = load().nlpf_reservoir_streams(seed, first, count, draws, out.ctypes.data_as(C.c_void_p))
    if rc != NLPF_OK:
        raise NlpfError(rc, "nlpf_reservoir_streams")
    return out


def style_from_scene(scene, materials=None, capacity=1024):
    """Style table of a reference scene file (akari_scenegraph JSON): returns (nodes, roots,
    materials), roots[m] being the root of the material named materials[m]. `scene` is a path
    or the JSON text; `materials` defaults to every material of the scene in ascending name
    order (the order of the reference's Collection, a BTreeMap). Needs no GPU."""
    lib = load()
    if isinstance(scene, (bytes, str)) and not str(scene).lstrip().startswith("{"):
        with open(scene, "rb") as f:
            text = f.read()
    else:
        text = scene.encode() if isinstance(scene, str) else bytes(scene)
    if materials is None:
        materials = sorted(json.loads(text)["materials"].keys(), key=lambda k: k.encode())
    names = (C.c_char_p * max(1, len(materials)))(*[m.encode() for m in materials])
    nodes = np.zeros(capacity, STYLE_NODE_DTYPE)
    roots = np.zeros(max(1, len(materials)), np.int32)
    n = C.c_uint32(capacity)
    rc = lib.nlpf_style_from_scene(text, names, len(materials), nodes.ctypes.data_as(C.c_void_p), C.byref(n),
                                   roots.ctypes.data_as(C.c_void_p))
    if rc != NLPF_OK:
        raise NlpfError(rc, lib.nlpf_last_error(None).decode())
    return nodes[:n.value].copy(), roots[:len(materials)].copy(), list(materials)


def bsdfs_from_scene(scene, materials):
    """BSDF table (BSDF_DTYPE) of the named materials of a reference scene file (path or JSON
    text); raises NlpfError naming the first material outside the supported kinds. Needs no GPU."""
    lib = load()
    if isinstance(scene, (bytes, str)) and not str(scene).lstrip().startswith("{"):
        with open(scene, "rb") as f:
            text = f.read()
    else:
        text = scene.encode() if isinstance(scene, str) else bytes(scene)
    names = (C.c_char_p * max(1, len(materials)))(*[m.encode() for m in materials])
    out = np.zeros(max(1, len(materials)), BSDF_DTYPE)
    rc = lib.nlpf_bsdfs_from_scene(text, names, len(materials), out.ctypes.data_as(C.c_void_p))
    if rc != NLPF_OK:
        raise NlpfError(rc, lib.nlpf_last_error(None).decode())
    return out[:len(materials)].copy()


SCENE_MATERIALS = ["floor", "ceiling", "back", "left", "right", "sphere0", "sphere1", "light"]


def write_world_scene(path, W, H, seed=1, albedo_lo=0.2, albedo_hi=0.8, nodes=None, roots=(),
                      sphere_segments=128, bsdfs=None):
    """The procedural scene of generate_world_records as a reference scene file (JSON + .bin),
    with the BSDF table's principled / emission materials where it has them. Needs no GPU."""
    lib = load()
    g = _GenDesc(seed, 0.0, 0.0, 0.0, 0.0, albedo_lo, albedo_hi)
    nodes = np.zeros(0, STYLE_NODE_DTYPE) if nodes is None else np.ascontiguousarray(nodes, STYLE_NODE_DTYPE)
    roots = np.ascontiguousarray(roots, np.int32)
    bs = np.zeros(0, BSDF_DTYPE) if bsdfs is None else np.ascontiguousarray(bsdfs, BSDF_DTYPE)
    rc = lib.nlpf_write_world_scene(os.fsencode(path), W, H, C.byref(g),
                                    nodes.ctypes.data_as(C.c_void_p) if len(nodes) else None, len(nodes),
                                    roots.ctypes.data_as(C.c_void_p) if len(roots) else None, len(roots),
                                    sphere_segments,
                                    bs.ctypes.data_as(C.c_void_p) if len(bs) else None, len(bs))
    if rc != NLPF_OK:
        raise NlpfError(rc, lib.nlpf_last_error(None).decode())


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return C.c_void_p(x.ptr)
    if isinstance(x, int):
        return C.c_void_p(x)
    raise TypeError(type(x))


class DeviceBuffer:
    """A hipMalloc'd block owned by a Handle."""

    def __init__(self, handle, nbytes):
        self.handle = handle
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        handle._check(handle.lib.nlpf_alloc(handle.h, C.c_size_t(max(1, self.nbytes)), C.byref(p)))
        self.ptr = p.value

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes, (offset, arr.nbytes, self.nbytes)
        self.handle._check(self.handle.lib.nlpf_upload(
            self.handle.h, C.c_void_p(self.ptr + offset), arr.ctypes.data_as(C.c_void_p), arr.nbytes))
        return self

    def download(self, dtype, count=None, offset=0):
        dtype = np.dtype(dtype)
        count = (self.nbytes - offset) // dtype.itemsize if count is None else int(count)
        assert offset + count * dtype.itemsize <= self.nbytes
        out = np.empty(count, dtype)
        self.handle._check(self.handle.lib.nlpf_download(
            self.handle.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.handle.lib.nlpf_free(self.handle.h, C.c_void_p(self.ptr))
            self.ptr = 0


class Handle:
    """One NL-PF context (one HIP stream), configured by the reference's "pf" JSON."""

    def __init__(self, config=None):
        self.lib = load()
        if config is None:
            config = {}
        js = config if isinstance(config, str) else json.dumps(config)
        h = C.c_void_p()
        rc = self.lib.nlpf_create(js.encode(), C.byref(h))
        if rc != NLPF_OK:
            raise NlpfError(rc, self.lib.nlpf_last_error(None).decode())
        self.h = h
        self._bufs = []

    def _check(self, rc):
        if rc != NLPF_OK:
            raise NlpfError(rc, self.lib.nlpf_last_error(self.h).decode())

    def close(self):
        if self.h:
            for b in self._bufs:
                b.free()
            self.lib.nlpf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def config(self):
        buf = C.create_string_buffer(4096)
        self._check(self.lib.nlpf_get_config_json(self.h, buf, 4096))
        return json.loads(buf.value.decode())

    def set_sigmas(self, sp, sn, sl):
        self._check(self.lib.nlpf_set_sigmas(self.h, sp, sn, sl))

    def alloc(self, nbytes):
        b = DeviceBuffer(self, nbytes)
        self._bufs.append(b)
        return b

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        return self.alloc(arr.nbytes).upload(arr)

    def copy(self, dst, src, nbytes):
        self._check(self.lib.nlpf_copy(self.h, _ptr(dst), _ptr(src), nbytes))

    def memset(self, dst, value, nbytes):
        self._check(self.lib.nlpf_memset(self.h, _ptr(dst), value, nbytes))

    def sync(self):
        self._check(self.lib.nlpf_sync(self.h))

    def set_style(self, nodes, roots):
        nodes = np.ascontiguousarray(nodes, dtype=STYLE_NODE_DTYPE)
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        self._check(self.lib.nlpf_set_style(self.h, nodes.ctypes.data_as(C.c_void_p), len(nodes),
                                            roots.ctypes.data_as(C.c_void_p), len(roots)))

    def set_bsdfs(self, table):
        """World-mode BSDF of each material (BSDF_DTYPE entries; materials beyond are diffuse)."""
        table = np.ascontiguousarray(table, dtype=BSDF_DTYPE)
        self._check(self.lib.nlpf_set_bsdfs(self.h, table.ctypes.data_as(C.c_void_p) if len(table) else None,
                                            len(table)))

    def stylize(self, rgb, depth=0, flags=None, material=None, pos=None):
        """Stylize host rgb[n,3] on the GPU; depth is an int or an array of n depths."""
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        n = len(rgb)
        din = self.to_device(rgb)
        dout = self.alloc(rgb.nbytes)
        dd = None
        d0 = 0
        if np.ndim(depth) == 0:
            d0 = int(depth)
        else:
            dd = self.to_device(np.asarray(depth, np.uint32))
        df = None if flags is None else self.to_device(np.asarray(flags, np.uint8))
        dm = None if material is None else self.to_device(np.asarray(material, np.uint16))
        dp = None if pos is None else self.to_device(np.asarray(pos, np.float32).reshape(-1, 3))
        self._check(self.lib.nlpf_stylize(self.h, _ptr(din), _ptr(dout), n, _ptr(dd), d0, _ptr(df),
                                          _ptr(dm), _ptr(dp)))
        out = dout.download(np.float32, 3 * n).reshape(n, 3)
        for b in (din, dout, dd, df, dm, dp):
            if b is not None:
                b.free()
        return out

    # ---- screen mode ----
    @staticmethod
    def screen_desc(W, H, S, L, precision=FP32, radius=7):
        return _ScreenDesc(W, H, S, L, precision, radius)

    def screen_run(self, desc, records, anchors=None, film=None, level_out=None):
        self._check(self.lib.nlpf_screen_run(self.h, C.byref(desc), _ptr(records), _ptr(anchors),
                                             _ptr(film), _ptr(level_out)))

    def screen_run_materials(self, desc, records, anchors, material, flags, film, level_out=None):
        """screen_run with per-vertex material (uint16) and flags (uint8 or None) planes laid out
        [spp][levels][H][W]; vertex i takes material[i]'s stylization tree with flags[i]."""
        self._check(self.lib.nlpf_screen_run_materials(self.h, C.byref(desc), _ptr(records), _ptr(anchors),
                                                       _ptr(material), _ptr(flags), _ptr(film), _ptr(level_out)))

    def screen_run_chunk(self, desc, records, anchors, film, slice0, total_spp, level_out=None):
        """Slices [slice0, slice0 + desc.spp) of a total_spp-slice frame; film carries the running
        sums between chunks and holds the frame's film after the last one."""
        self._check(self.lib.nlpf_screen_run_chunk(self.h, C.byref(desc), _ptr(records), _ptr(anchors),
                                                   _ptr(film), _ptr(level_out), slice0, total_spp))

    def screen_run_noalloc(self, desc, records, anchors, film):
        """nlpf_screen_run without level outputs and without any device allocation; asynchronous
        (call sync() before reading film)."""
        self._check(self.lib.nlpf_screen_run_noalloc(self.h, C.byref(desc), _ptr(records), _ptr(anchors),
                                                     _ptr(film)))

    def screen_backward(self, desc, records, level_out, dfilm, grad):
        """Adjoint pass; grad receives [dsigma_pos, dsigma_normal, dsigma_lum, ramp stops...]."""
        self._check(self.lib.nlpf_screen_backward(self.h, C.byref(desc), _ptr(records), _ptr(level_out),
                                                  _ptr(dfilm), _ptr(grad)))

    def generate_screen_records(self, desc, records, anchors=None, seed=1, mu=-1.0, sigma=1.5,
                                firefly_prob=1e-3, firefly_scale=1e3, albedo_lo=0.1, albedo_hi=0.9):
        g = _GenDesc(seed, mu, sigma, firefly_prob, firefly_scale, albedo_lo, albedo_hi)
        self._check(self.lib.nlpf_generate_screen_records(self.h, C.byref(desc), C.byref(g),
                                                          _ptr(records), _ptr(anchors)))

    def convert_records_fp16(self, desc, records32, records16, anchors):
        self._check(self.lib.nlpf_convert_records_fp16(self.h, C.byref(desc), _ptr(records32),
                                                       _ptr(records16), _ptr(anchors)))

    # ---- learned cache ----
    def mlp_param_count(self):
        return int(self.lib.nlpf_mlp_param_count())

    def mlp_init(self, seed=0):
        self._check(self.lib.nlpf_mlp_init(self.h, seed))

    def mlp_get_params(self):
        out = np.empty(self.mlp_param_count(), np.float32)
        self._check(self.lib.nlpf_mlp_get_params(self.h, out.ctypes.data_as(C.c_void_p), len(out)))
        return out

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.nlpf_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p), len(p)))

    def mlp_set_ema(self, decay):
        self._check(self.lib.nlpf_mlp_set_ema(self.h, decay))

    def mlp_encode(self, desc, records, anchors, depth, idx, n, feats):
        self._check(self.lib.nlpf_mlp_encode(self.h, C.byref(desc), _ptr(records), _ptr(anchors), depth,
                                             _ptr(idx), n, _ptr(feats)))

    def mlp_infer(self, desc, records, anchors, depth, idx, feats, n, out):
        self._check(self.lib.nlpf_mlp_infer(self.h, None if desc is None else C.byref(desc), _ptr(records),
                                            _ptr(anchors), depth, _ptr(idx), _ptr(feats), n, _ptr(out)))

    def mlp_train_step(self, desc, records, anchors, depth, idx, feats, target, batch, lr, apply=True, grad=None):
        loss = C.c_float()
        self._check(self.lib.nlpf_mlp_train_step(self.h, None if desc is None else C.byref(desc), _ptr(records),
                                                 _ptr(anchors), depth, _ptr(idx), _ptr(feats), _ptr(target), batch,
                                                 lr, int(apply), C.byref(loss), _ptr(grad)))
        return loss.value

    def mlp_reference(self, desc, records, anchors, depth, idx, n, out):
        self._check(self.lib.nlpf_mlp_reference(self.h, C.byref(desc), _ptr(records), _ptr(anchors), depth,
                                                _ptr(idx), n, _ptr(out)))

    def screen_run_mlp(self, desc, records, anchors, film, level_out=None, depth=2, audit_slice=0,
                       train_records=1 << 16, batch=1 << 14, lr=1e-2, tile_threshold=0.05, seed=0,
                       estimator="network"):
        """Screen NL-PF with the learned cache at `depth` (one frame; training continues across
        calls).  estimator: "network", or the controls "zero" (paths cut at `depth`) and "mean"
        (the audit slice's mean depth-`depth` value).  Returns the run statistics as a dict."""
        r = _MlpRunDesc(depth, audit_slice, train_records, batch, lr, tile_threshold, seed,
                        CACHE_ESTIMATORS[estimator], 0)
        st = _MlpRunStats()
        self._check(self.lib.nlpf_screen_run_mlp(self.h, C.byref(desc), _ptr(records), _ptr(anchors), C.byref(r),
                                                 _ptr(film), _ptr(level_out), C.byref(st)))
        return {k: getattr(st, k) for k, _ in _MlpRunStats._fields_}

    def stats(self):
        s = _Stats()
        self._check(self.lib.nlpf_get_stats(self.h, C.byref(s)))
        return {"total_ms": s.total_ms, "level_ms": list(s.level_ms), "film_ms": s.film_ms,
                "vertices": s.vertices, "bytes_algorithmic": s.bytes_algorithmic,
                "flops_algorithmic": s.flops_algorithmic, "points": list(s.points),
                "neighbors": list(s.neighbors), "smis_pairs": list(s.smis_pairs)}

    # ---- world mode ----
    @staticmethod
    def world_desc(W, H, branch, L, aabb_min=(-0.5, -0.5, -0.5), aabb_max=(0.5, 0.5, 0.5), film_accumulate=0):
        return _WorldDesc(W, H, branch, L, (C.c_float * 3)(*aabb_min), (C.c_float * 3)(*aabb_max),
                          int(film_accumulate))

    def alloc_world(self, desc):
        n = desc.width * desc.height * desc.branch
        L = desc.levels
        return {"records": self.alloc(L * WORLD_PLANES * n * 4), "flags": self.alloc(L * n),
                "new_flags": self.alloc(L * n), "material": self.alloc(L * n * 2),
                "path_depth": self.alloc(n * 4)}

    def generate_world_records(self, desc, bufs, seed=1, albedo_lo=0.2, albedo_hi=0.8):
        g = _GenDesc(seed, 0.0, 0.0, 0.0, 0.0, albedo_lo, albedo_hi)
        self._check(self.lib.nlpf_generate_world_records(
            self.h, C.byref(desc), C.byref(g), _ptr(bufs["records"]), _ptr(bufs["flags"]),
            _ptr(bufs["new_flags"]), _ptr(bufs["material"]), _ptr(bufs["path_depth"])))

    def upload_world(self, desc, host):
        bufs = self.alloc_world(desc)
        for k in bufs:
            bufs[k].upload(np.ascontiguousarray(host[k]))
        return bufs

    def download_world(self, desc, bufs):
        n = desc.width * desc.height * desc.branch
        L = desc.levels
        return {"records": bufs["records"].download(np.float32, L * WORLD_PLANES * n),
                "flags": bufs["flags"].download(np.uint8, L * n),
                "new_flags": bufs["new_flags"].download(np.uint8, L * n),
                "material": bufs["material"].download(np.uint16, L * n),
                "path_depth": bufs["path_depth"].download(np.int32, n)}

    def world_run(self, desc, bufs, want_radiance=False, film=None):
        """One world-mode pass. With desc.film_accumulate == 0 returns (film [H, W, 3], radiance);
        otherwise `film` must be a film buffer (film_alloc) that the pass adds its samples to,
        and the first element returned is None."""
        n = desc.width * desc.height * desc.branch
        L = desc.levels
        own = film is None
        if desc.film_accumulate and own:
            raise ValueError("film_accumulate needs a film buffer")
        if own:
            film = self.alloc(desc.width * desc.height * 12)
        rad = self.alloc(L * n * 12) if want_radiance else None
        self._check(self.lib.nlpf_world_run(
            self.h, C.byref(desc), _ptr(bufs["records"]), _ptr(bufs["flags"]), _ptr(bufs["new_flags"]),
            _ptr(bufs["material"]), _ptr(bufs["path_depth"]), _ptr(film), _ptr(rad)))
        f = None
        if own:
            f = film.download(np.float32).reshape(desc.height, desc.width, 3)
            film.free()
        r = rad.download(np.float32).reshape(L, n, 3) if want_radiance else None
        if rad is not None:
            rad.free()
        return f, r

    # ---- film (crates/akari_render/src/film.rs) ----
    def film_alloc(self, W, H, clear=True):
        f = self.alloc(4 * film_floats(W, H))
        if clear:
            self.film_clear(f, W, H)
        return f

    def film_clear(self, film, W, H):
        self._check(self.lib.nlpf_film_clear(self.h, _ptr(film), W, H))

    def _film_add(self, fn, film, W, H, xy, rgb, weight):
        n = int(xy.nbytes // 8) if isinstance(xy, np.ndarray) else None
        bufs = []

        def dev(a, dt, k):
            if a is None or isinstance(a, (DeviceBuffer, int)):
                return a
            b = self.to_device(np.ascontiguousarray(a, dt).reshape(-1, k))
            bufs.append(b)
            return b
        if n is None:
            raise TypeError("xy must be a host array [n, 2]")
        d = C.c_uint64(0)
        rc = fn(self.h, _ptr(film), W, H, _ptr(dev(xy, np.float32, 2)), _ptr(dev(rgb, np.float32, 3)),
                _ptr(dev(weight, np.float32, 1)), n, C.byref(d))
        for b in bufs:
            b.free()
        self._check(rc)
        return d.value

    def film_add_samples(self, film, W, H, xy, rgb, weight=None):
        """Film::add_sample for host samples xy[n,2], rgb[n,3], weight[n] (None = 1); returns the
        number of samples that fell outside the film."""
        return self._film_add(self.lib.nlpf_film_add_samples, film, W, H, xy, rgb, weight)

    def film_add_splats(self, film, W, H, xy, rgb, weight=None):
        return self._film_add(self.lib.nlpf_film_add_splats, film, W, H, xy, rgb, weight)

    def film_resolve(self, film, W, H, splat_scale=1.0, hdr=True, rgba=None):
        """Film::copy_to_rgba_image; returns the host image [H, W, 4] unless a device rgba is given."""
        own = rgba is None
        if own:
            rgba = self.alloc(16 * W * H)
        self._check(self.lib.nlpf_film_resolve(self.h, _ptr(film), W, H, float(splat_scale), int(bool(hdr)),
                                               _ptr(rgba)))
        if not own:
            return None
        out = rgba.download(np.float32).reshape(H, W, 4)
        rgba.free()
        return out

    def film_download(self, film, W, H):
        """(pixels [H,W,3], splat [H,W,3], weights [H,W]) of a film buffer."""
        a = film.download(np.float32, film_floats(W, H))
        P = W * H
        return a[:3 * P].reshape(H, W, 3), a[3 * P:6 * P].reshape(H, W, 3), a[6 * P:].reshape(H, W)

    def write_image(self, rgba, W, H, path):
        """util::write_image of a device image rgba (e.g. film_resolve's output buffer)."""
        self._check(self.lib.nlpf_write_image(self.h, _ptr(rgba), W, H, os.fsencode(path)))


# ---- stylization node builders ----
def node(kind, inner=-1, else_=-1, idx=0, depth_mask=ALL_DEPTHS, max_depth=0, count=0, params=()):
    n = np.zeros((), STYLE_NODE_DTYPE)
    n["type"] = STY[kind]
    n["inner"] = inner
    n["else_"] = else_
    n["idx"] = idx
    n["depth_mask"] = depth_mask
    n["max_depth"] = max_depth
    n["count"] = count
    p = np.zeros(56, np.float32)
    p[: len(params)] = params
    n["p"] = p
    return n


def toon(width, inner=-1, depth_mask=ALL_DEPTHS):
    return node("TOON", inner=inner, depth_mask=depth_mask, params=[width])


def ramp(positions, values, inner=-1, depth_mask=ALL_DEPTHS):
    positions = np.asarray(positions, np.float32)
    values = np.asarray(values, np.float32)
    assert len(positions) == len(values) <= 16
    p = np.zeros(32, np.float32)
    p[: len(positions)] = positions
    p[16: 16 + len(values)] = values
    return node("RAMP", inner=inner, depth_mask=depth_mask, count=len(positions), params=p)


def hue_shift(shift, inner=-1, depth_mask=ALL_DEPTHS, indirect_scale=1.0):
    return node("HUE_SHIFT", inner=inner, depth_mask=depth_mask, params=[shift, indirect_scale])


def saturation(power, indirect_scale=1.0, max_depth=0, inner=-1, depth_mask=ALL_DEPTHS):
    return node("SATURATION", inner=inner, depth_mask=depth_mask, max_depth=max_depth,
                params=[power, indirect_scale])


def rex(blend, inner=-1, depth_mask=ALL_DEPTHS):
    return node("REX", inner=inner, depth_mask=depth_mask, params=[blend])


def eight_bit(indirect_scale=1.0, inner=-1, depth_mask=ALL_DEPTHS):
    return node("EIGHT_BIT", inner=inner, depth_mask=depth_mask, params=[indirect_scale])


def palette_hex(entries):
    """Palette hex strings -> rgb/255, as the SVM compiler does (svm/compiler.rs:420-453)."""
    out = []
    for e in entries:
        x = int(e.replace("0x", ""), 16)
        out.append([((x >> 16) & 0xFF) / 255.0, ((x >> 8) & 0xFF) / 255.0, (x & 0xFF) / 255.0])
    return np.asarray(out, np.float32)


def palette(colors, gamma=1.0, indirect_scale=1.0, inner=-1, depth_mask=ALL_DEPTHS):
    colors = np.asarray(colors, np.float32).reshape(-1, 3)
    assert len(colors) <= 18
    return node("PALETTE", inner=inner, depth_mask=depth_mask, count=len(colors),
                params=np.concatenate([[indirect_scale, gamma], colors.ravel()]))


def halftone(indirect_scale, radius_scale, radius, grid_distance, blend, dark, light, affine,
             max_depth=0, inner=-1, depth_mask=ALL_DEPTHS):
    affine = np.asarray(affine, np.float32).reshape(3, 4)
    params = np.concatenate([[indirect_scale, radius_scale, radius, grid_distance, blend],
                             np.asarray(dark, np.float32), np.asarray(light, np.float32),
                             affine.ravel()])
    return node("HALFTONE", inner=inner, depth_mask=depth_mask, max_depth=max_depth, params=params)


def enable_if(idx, inner=-1, else_=-1):
    return node("ENABLE_IF", inner=inner, else_=else_, idx=idx)


def set_idx(idx, inner=-1):
    return node("SET_IDX", inner=inner, idx=idx)


def depth_mask(*depths):
    m = 0
    for d in depths:
        m |= 1 << min(d, 31)
    return m


def style_table(nodes):
    return np.array(list(nodes), dtype=STYLE_NODE_DTYPE)
