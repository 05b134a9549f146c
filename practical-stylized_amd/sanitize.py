"""Host-side AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU test suite.

Builds libnlpf with the sanitizers on its host code only (device code is compiled as usual;
GPU sanitizers are not used) and the CPU oracle with the same clang runtime, then runs
`pytest -m "not gpu"` against both builds with the ASan runtime preloaded into Python.  The
host code exercised is the JSON config and scene parsers, the style-table lowering, the scene
writer, the image writers and the oracle.  Any report fails the run
(-fno-sanitize-recover; leak checks are off, Python itself is not instrumented).

    python practical-stylized_amd/sanitize.py
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import build as B  # noqa: E402

OUT = os.path.join(HERE, "_san")
CLANG = "/opt/rocm/llvm/bin/clang"
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
            "-Xarch_host", "-g"]


def asan_runtime():
    found = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not found:
        raise RuntimeError("clang ASan runtime not found under /opt/rocm/llvm")
    return found[0]


def build_library():
    objdir = os.path.join(OUT, "obj")
    os.makedirs(objdir, exist_ok=True)

    def one(src):
        obj = os.path.join(objdir, src + ".o")
        cmd = [B.HIPCC] + B.COMMON + HOST_SAN + ["-c", os.path.join(B.CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f"{src}:\n{r.stderr[-4000:]}")
        return obj

    with ThreadPoolExecutor(8) as ex:
        objs = list(ex.map(one, B.sources()))
    lib = os.path.join(OUT, "libnlpf.so")
    subprocess.run([B.HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-shared-libasan",
                    "-fsanitize=address,undefined", "-o", lib] + objs, check=True)
    return lib


def build_oracle():
    src = [os.path.join(ROOT, "oracle", f) for f in ("oracle_style.c", "oracle_screen.c", "oracle_world.c",
                                                      "oracle_grad.c", "oracle_bsdf.c")]
    lib = os.path.join(OUT, "liboracle.so")
    subprocess.run([CLANG, "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-shared-libasan", "-fopenmp", "-fPIC", "-std=c11",
                    "-shared", "-o", lib] + src + ["-lm", "-L/opt/rocm/llvm/lib", "-Wl,-rpath,/opt/rocm/llvm/lib"],
                   check=True)
    return lib


def main():
    os.makedirs(OUT, exist_ok=True)
    nlpf_lib, oracle_lib = build_library(), build_oracle()
    env = dict(os.environ, LD_PRELOAD=asan_runtime(), NLPF_LIB=nlpf_lib, NLPF_ORACLE_LIB=oracle_lib,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider"], env=env, cwd=ROOT)
    sys.exit(r.returncode)


if __name__ == "__main__":
    main()
