"""The C-ABI library loads and exports every entry point include/nlpf.h declares; struct
layouts of the Python mirror match the header; config errors are reported before any GPU
work (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import nlpf as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "nlpf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nlpf_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = N.load()
    names = header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(N.exported_symbols()) == set(names)


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "nlpf.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){'
                   'printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(nlpf_style_node), sizeof(nlpf_screen_desc),'
                   ' sizeof(nlpf_gen_desc), sizeof(nlpf_stats), offsetof(nlpf_style_node, p),'
                   ' sizeof(nlpf_world_desc), offsetof(nlpf_world_desc, film_accumulate));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out]
    assert sizes[0] == N.STYLE_NODE_DTYPE.itemsize == 256
    assert sizes[1] == C.sizeof(N._ScreenDesc)
    assert sizes[2] == C.sizeof(N._GenDesc)
    assert sizes[3] == C.sizeof(N._Stats)
    assert sizes[4] == N.STYLE_NODE_DTYPE.fields["p"][1]
    assert sizes[5] == C.sizeof(N._WorldDesc)
    assert sizes[6] == N._WorldDesc.film_accumulate.offset


@pytest.mark.parametrize("cfg,msg", [
    ("{", "config JSON"),
    ('{"type":"pt"}', "type"),
    ('{"max_depth":-1}', "max_depth"),
    ('{"reuse":[]}', "reuse"),
    ('{"radius_scale":[0]}', "radius_scale"),
    ('{"branch":0}', "branch"),
    ('{"smis":1}', "smis"),
    ('{"scene_bounds":[[0,0,0],[1,1]]}', "scene_bounds"),
    ('{"mode":"tiles"}', "mode"),
    ('{"pixel_radius":9}', "pixel_radius"),
    ('{"sigma_pos":0}', "sigma_pos"),
    ('{"precision":"bf16"}', "precision"),
    ('{"world_kernel":"warp"}', "world_kernel"),
    ('{"world_grid":"hash"}', "world_grid"),
    ('{"screen_kernel":"rows"}', "screen_kernel"),
    ('{"hash_table_size":0}', "hash_table_size"),
])
def test_invalid_config_rejected_before_gpu(cfg, msg):
    lib = N.load()
    h = C.c_void_p()
    rc = lib.nlpf_create(cfg.encode(), C.byref(h))
    assert rc == 1  # NLPF_E_INVALID_CONFIG
    assert msg in lib.nlpf_last_error(None).decode()
    assert not h.value


def test_null_arguments_rejected():
    lib = N.load()
    assert lib.nlpf_create(b"{}", None) == 2
    assert lib.nlpf_destroy(None) == 2
    assert lib.nlpf_sync(None) == 2
    assert lib.nlpf_screen_run_materials(None, None, None, None, None, None, None, None) == 2


def rust_block():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```rust\n(.*?)```", txt, flags=re.S)
    return next(b for b in blocks if 'extern "C"' in b)


def rust_struct_size(body):
    """Size of a #[repr(C)] struct from its field list (C layout: each field at its alignment,
    the total padded to the largest alignment)."""
    prim = {"f32": 4, "i32": 4, "u32": 4, "f64": 8, "u64": 8, "usize": 8}
    off = align_max = 0
    for ty in re.findall(r"pub\s+\w+\s*:\s*(\[[^\]]+\]|\w+)", body):
        m = re.match(r"\[\s*(\w+)\s*;\s*(\d+)\s*\]", ty)
        size, align = (prim[m.group(1)] * int(m.group(2)), prim[m.group(1)]) if m else (prim[ty], prim[ty])
        off = (off + align - 1) // align * align + size
        align_max = max(align_max, align)
    return (off + align_max - 1) // align_max * align_max


def test_documented_rust_binding_matches_header(tmp_path):
    """The Rust -sys declarations a caller would copy declare every entry point of the header,
    and each documented #[repr(C)] struct has the header struct's size."""
    block = rust_block()
    declared = set(re.findall(r"pub fn (nlpf_[a-z0-9_]+)\s*\(", block))
    assert declared == set(header_functions())
    structs = dict(re.findall(r"pub struct (nlpf_\w+)\s*\{(.*?)\}", block, flags=re.S))
    structs.pop("nlpf_handle", None)
    assert len(structs) >= 7
    names = sorted(structs)
    src = tmp_path / "rs.c"
    src.write_text('#include "nlpf.h"\n#include <stdio.h>\nint main(void){' +
                   "".join(f'printf("%zu\\n", sizeof({n}));' for n in names) + "return 0;}\n")
    exe = tmp_path / "rs"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    for n, c_size in zip(names, sizes):
        assert rust_struct_size(structs[n]) == c_size, n


def test_python_mirror_structs_match_header(tmp_path):
    """Every ctypes struct of the Python mirror has its header struct's size and field offsets."""
    mirror = {"nlpf_screen_desc": N._ScreenDesc, "nlpf_gen_desc": N._GenDesc, "nlpf_stats": N._Stats,
              "nlpf_world_desc": N._WorldDesc, "nlpf_mlp_run_desc": N._MlpRunDesc,
              "nlpf_mlp_run_stats": N._MlpRunStats}
    lines = []
    for cname, cls in mirror.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "py.c"
    src.write_text('#include "nlpf.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){' + "".join(lines) +
                   "return 0;}\n")
    exe = tmp_path / "py"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = iter(int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
    for cname, cls in mirror.items():
        assert next(vals) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert next(vals) == getattr(cls, fname).offset, (cname, fname)


def test_bsdf_table_layouts_match_header(tmp_path):
    """The Python mirror's BSDF_DTYPE and the oracle's or_bsdf have nlpf_bsdf's size and offsets."""
    import oracle as O
    fields = list(N.BSDF_DTYPE.names)
    src = tmp_path / "bsdf.c"
    src.write_text('#include "nlpf.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){' +
                   'printf("%zu\\n", sizeof(nlpf_bsdf));' +
                   "".join(f'printf("%zu\\n", offsetof(nlpf_bsdf, {f}));' for f in fields) + "return 0;}\n")
    exe = tmp_path / "bsdf"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == N.BSDF_DTYPE.itemsize == O.BSDF_DTYPE.itemsize
    for f, off in zip(fields, vals[1:]):
        assert N.BSDF_DTYPE.fields[f][1] == off == O.BSDF_DTYPE.fields[f][1], f


def header_prototypes():
    """{name: (return type, [parameter types])} from include/nlpf.h (comments and names
    stripped; pointer parameters reduced to "ptr")."""
    txt = open(os.path.join(ROOT, "include", "nlpf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    protos = {}
    for ret, name, params in re.findall(r"([A-Za-z_][\w \*]*?)\s*\b(nlpf_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", txt):
        kinds = []
        for p in [q.strip() for q in params.split(",") if q.strip() and q.strip() != "void"]:
            if "*" in p:
                kinds.append("ptr")
            else:
                kinds.append(re.sub(r"\s+\w+$", "", p).replace("const ", "").strip())
        protos[name] = ("ptr" if "*" in ret else ret.strip(), kinds)
    return protos


def test_python_mirror_signatures_match_header():
    """Each ctypes signature of the mirror has the header's parameter count, and integer and
    float parameters of the same width."""
    width = {C.c_int: ("int", 4), C.c_uint32: ("uint32_t", 4), C.c_int32: ("int32_t", 4),
             C.c_uint64: ("uint64_t", 8), C.c_size_t: ("size_t", 8), C.c_float: ("float", 4),
             C.c_double: ("double", 8), C.c_uint16: ("uint16_t", 2), C.c_uint8: ("uint8_t", 1)}
    c_width = {"int": 4, "uint32_t": 4, "int32_t": 4, "uint64_t": 8, "size_t": 8, "float": 4, "double": 8,
               "uint16_t": 2, "uint8_t": 1}
    protos = header_prototypes()
    for name, res, args in N._SIGS:
        ret, params = protos[name]
        assert len(args) == len(params), name
        for i, (a, p) in enumerate(zip(args, params)):
            if p == "ptr":
                assert a in (C.c_void_p, C.c_char_p) or hasattr(a, "_type_") and a not in width, (name, i, a)
            else:
                assert a in width and width[a][1] == c_width[p], (name, i, a, p)


def test_library_random_streams_known_answers(oracle):
    """The library's own generator code (host copy of the kernels' functions): the pcg32-demo
    outputs, and reservoir streams equal to the oracle's (which the published ChaCha12 /
    StdRng vectors pin)."""
    from test_oracle_world import PCG_DEMO
    assert [int(v) for v in N.pcg32(54, 42, 6)] == PCG_DEMO
    for seed in (0, 7):
        assert np.array_equal(N.reservoir_streams(seed, 0, 70, 3), oracle.reservoir_streams(seed, 0, 70, 3))
        assert np.array_equal(N.reservoir_streams(seed, 1000, 9, 2), oracle.reservoir_streams(seed, 1000, 9, 2))


def test_mirror_constants_match_header():
    txt = open(os.path.join(ROOT, "include", "nlpf.h")).read()
    m = re.search(r"#define NLPF_NOALLOC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", txt)
    assert m and int(m.group(1)) << int(m.group(2)) == N.NOALLOC_SCRATCH_BYTES
    assert re.search(r"#define NLPF_SCREEN_PLANES (\d+)", txt).group(1) == str(N.SCREEN_PLANES)


def test_documented_rust_config_covers_every_field():
    """The Rust Config facade of the integration notes names every field the library parses, with
    the Python mirror's defaults."""
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = re.search(r"pub struct Config \{(.*?)\n\}", txt, flags=re.S).group(1)
    fields = set(re.findall(r"pub (\w+)\s*:", body))
    assert fields == set(N.DEFAULT_CONFIG), fields ^ set(N.DEFAULT_CONFIG)
    dflt = re.search(r"impl Default for Config \{.*?Self \{(.*?)\}\s*\}\s*\}", txt, flags=re.S).group(1)
    for k, v in N.DEFAULT_CONFIG.items():
        m = re.search(rf"\b{k}: (vec!\[[^\]]*\]|\"[^\"]*\"\.into\(\)|[^,}}]+)", dflt)
        assert m, k
        got = m.group(1).strip()
        if isinstance(v, bool):
            assert got == str(v).lower(), (k, got)
        elif isinstance(v, str):
            assert got == f'"{v}".into()', (k, got)
        elif v is None:
            assert got == "None", (k, got)
        elif isinstance(v, list):
            assert [float(x) for x in re.findall(r"[\d.]+", got)] == [float(x) for x in v], (k, got)
        elif k == "hash_table_size":
            assert eval(got) == v, (k, got)
        else:
            assert float(got) == float(v), (k, got)
