"""Instruction mix of the screen-mode level kernel's inner loop, from this repository's own
gfx950 assembly (hipcc -S of csrc/screen_fp32.hip). Prints, for every basic block of the chosen
kernel instantiation that holds at least 8 v_exp_f32, the VALU count per v_exp_f32 (one
v_exp_f32 per filtered pair) and the issue slots per pair with v_exp_f32 counted as 2.

    python practical-stylized_amd/probes/isa_mix.py [RAD NW R]   (default 7 4 2)
"""
import collections
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "csrc", "screen_fp32.hip")


def main(rad=7, nw=4, r=2):
    with tempfile.TemporaryDirectory() as d:
        asm = os.path.join(d, "k.s")
        subprocess.run(["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "--offload-arch=gfx950",
                        "-I", os.path.join(HERE, "..", "..", "include"), "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o", asm],
                       check=True, capture_output=True)
        lines = open(asm).read().split("\n")
    name = f"_ZN4nlpf19screen_level_kernelIffLi{rad}ELi{nw}ELi{r}EEEvNS_10ScreenArgsIT0_EE:"
    start = next(i for i, l in enumerate(lines) if l.startswith(name))
    blocks, cur = [], None
    for l in lines[start + 1:]:
        t = l.strip()
        if t.startswith(".Lfunc_end"):
            break
        if re.match(r"^\.LBB\d+_\d+:", t):
            cur = [t.split(":")[0], collections.Counter()]
            blocks.append(cur)
            continue
        if not t or t.startswith((".", ";")) or cur is None:
            continue
        cur[1][t.split()[0]] += 1
    for label, c in blocks:
        ex = sum(v for k, v in c.items() if k.startswith("v_exp"))
        if ex < 8:
            continue
        valu = sum(v for k, v in c.items() if k.startswith("v_"))
        print(f"{label}: pairs {ex}  VALU {valu}  VALU/pair {valu / ex:.2f}  issue slots/pair {(valu + ex) / ex:.2f}"
              f"  LDS reads {sum(v for k, v in c.items() if k.startswith('ds_read'))}")


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:4]))
