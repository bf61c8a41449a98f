"""Builds the CPU emulation of the HyperBall library (test infrastructure).

The library's own sources are compiled with g++ against the HIP emulation in
tests/emu/hip/hip_runtime.h, with AddressSanitizer and UBSan.  Three textual
rewrites of hb_kernels.hip make it host C++:
  - `kernel<<<grid, block, shmem, stream>>>(args)` becomes
    `emu::launch(kernel, grid, block, shmem, stream, args)`;
  - the wave ballot inside group_ballot<G> becomes the G-lane emulated ballot;
  - the wavefront barrier inside lane_sync<W> becomes a W-lane barrier (the
    lanes of a group are OS threads here, not lockstep lanes).
hb_transpose.hip (rocPRIM, device-only) is replaced by emu_transpose.cpp.
Executables: one per driver source tests/emu/*_check.cpp, in tests/emu/build/.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(REPO, "webgraph-algo-rs_amd", "csrc")
OUT = os.path.join(HERE, "build")
FLAGS = ["-std=c++20", "-g", "-O1", "-fno-omit-frame-pointer", "-pthread",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         "-ffp-contract=off", "-DHB_EMULATION=1", "-DHB_SEED_ITEMS_MAX_NODES=1000", "-w"]


def _matching_paren(text, i):
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "(":
            depth += 1
        elif text[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced launch arguments")


def rewrite_kernels(text):
    out, pos, launches = [], 0, 0
    for m in re.finditer(r"([A-Za-z_]\w*(?:<[^<>;()]*>)?)<<<", text):
        if m.start() < pos:
            continue
        cfg_end = text.index(">>>", m.end())
        paren = cfg_end + 3
        assert text[paren] == "(", text[m.start():paren + 20]
        close = _matching_paren(text, paren)
        args = text[paren + 1:close].strip()
        cfg = text[m.end():cfg_end]
        out.append(text[pos:m.start()])
        out.append(f"emu::launch({m.group(1)}, {cfg}" + (f", {args})" if args else ")"))
        pos = close + 1
        launches += 1
    out.append(text[pos:])
    text = "".join(out)
    n_ballot = text.count("__ballot(pred)")
    n_barrier = text.count("__builtin_amdgcn_wave_barrier()")
    assert n_ballot == 1 and n_barrier == 1, (n_ballot, n_barrier)
    text = text.replace("__ballot(pred)", "emu::ballot<G>(pred)")
    text = text.replace("__builtin_amdgcn_wave_barrier()", "emu::group_barrier<W>()")
    return text, launches


def build(drivers=None, verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(CSRC, "hb_kernels.hip")) as f:
        text, launches = rewrite_kernels(f.read())
    kern = os.path.join(OUT, "hb_kernels_emu.cpp")
    with open(kern, "w") as f:
        f.write(f'#line 1 "{os.path.join(CSRC, "hb_kernels.hip")}"\n' + text)
    inc = ["-I", HERE, "-I", CSRC, "-I", os.path.join(REPO, "include")]
    objs = []
    for src, lang in ((kern, "c++"), (os.path.join(CSRC, "hb_engine.hip"), "c++"),
                      (os.path.join(CSRC, "hb_graph.cpp"), "c++"),
                      (os.path.join(HERE, "emu_runtime.cpp"), "c++"),
                      (os.path.join(HERE, "emu_transpose.cpp"), "c++")):
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        cmd = ["g++", *FLAGS, *inc, "-x", lang, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    if drivers is None:
        drivers = sorted(f for f in os.listdir(HERE) if f.endswith("_check.cpp"))
    # the oracle's C restatement, for the drivers that check against it
    oracle_src = os.path.join(REPO, "oracle", "hb_oracle.c")
    oracle_obj = os.path.join(OUT, "hb_oracle.o")
    subprocess.run(["gcc", "-std=c11", "-g", "-O1", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-ffp-contract=off", "-w", "-fopenmp",
                    "-c", oracle_src, "-o", oracle_obj], check=True)
    exes = []
    for d in drivers:
        exe = os.path.join(OUT, d[:-4])
        extra = (["-I", os.path.join(REPO, "oracle"), oracle_obj, "-fopenmp"]
                 if d.startswith("oracle") else [])
        subprocess.run(["g++", *FLAGS, *inc, os.path.join(HERE, d), *objs, *extra, "-lm", "-o", exe],
                       check=True)
        exes.append(exe)
    return exes, launches


if __name__ == "__main__":
    exes, launches = build(verbose="-v" in sys.argv)
    print(f"{launches} launch sites rewritten; built {', '.join(exes)}")
