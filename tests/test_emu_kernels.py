"""The library's own HIP kernels and engine, run on the CPU through the HIP
emulation in tests/emu (g++ with AddressSanitizer and UBSan; lanes are
threads, cross-lane operations are barrier-synchronised exchanges within the
lane group).  An out-of-bounds device access, a lane-group operation that
some lanes of the group skip, or a batch result that differs from the
single-instance result fails here, without a GPU.

The batch mixes all six batched shapes (log2m 6, 8, 10 with 5- and 6-bit
registers), so wavefronts of one workgroup run different shapes at once;
rows up to 600 successors exercise the workgroup-per-node kernel."""
import os
import re
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def emu_exes():
    import build_emu
    exes, launches = build_emu.build()
    assert launches >= 19
    return {os.path.basename(e): e for e in exes}


@pytest.mark.parametrize("graph,heavy,tune", [(1, 0, 1), (0, 64, 0)])
def test_emulated_batch_matches_single_instances(emu_exes, graph, heavy, tune):
    """(graph replay, batch autotuner) and (direct launches, heavy rows above 64)."""
    args = [emu_exes["batch_check"], "6", str(graph), "7", "150", "400", "600", str(heavy),
            str(tune)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-3000:] + r.stderr[-6000:]
    assert r.stdout.count("identical") == 6
    # the log2m-6 instances run their dense iterations as batched window iterations
    assert any(int(w) > 0 for w in re.findall(r"(\d+) window\)", r.stdout))


@pytest.mark.parametrize("log2m,r,n,dmax,seed,heavy,delta", [
    (6, 5, 500, 400, 4, 64, 1),     # 2 slabs, G = 2, heavy rows, delta gathering forced;
                                    # 10-dword counters at a 16-dword stride
    (8, 6, 400, 300, 5, 64, 1),     # 8 slabs, G = 8
    (10, 6, 300, 250, 3, 0, 0),     # 32 slabs: one change-mask bit each; the default rule
    (10, 5, 300, 250, 8, 64, 2),    # no delta gathering
    (12, 5, 200, 150, 6, 0, 1),     # 128 slabs: four per change-mask bit
    (5, 6, 300, 250, 7, 64, 1),     # one slab, G = 1 (6 dwords, stride 8)
    (8, 5, 300, 250, 9, 64, 0),     # 40 dwords, stride 48
    (7, 5, 300, 250, 10, 0, 1),     # 20 dwords, stride 32
    (13, 5, 200, 150, 6, 0, 1),     # two 64-lane segments per counter, delta gathering forced
    (8, 6, 500, 300, 5, 0, "0 2 128 1"),   # streamed counters: 4 tiles, blocks of 32 nodes
    (13, 5, 200, 150, 6, 0, "0 2 128 1"),  # streamed, two segments per counter
    (6, 5, 2000, 300, 9, 0, "0 2 1024"),   # 2 tiles, streamed by the per-iteration rule
])
def test_emulated_engine_matches_oracle(emu_exes, log2m, r, n, dmax, seed, heavy, delta):
    """Every iteration (standard, systolic, local, pre-local) of the emulated
    kernels, with and without delta gathering, equals the oracle's
    restatement of the reference engine: registers, modified bits, mode; then
    sums of distances."""
    args = [emu_exes["oracle_check"], str(log2m), str(r), str(n), str(dmax), str(seed), str(heavy),
            *str(delta).split()]
    res = subprocess.run(args, capture_output=True, text=True, timeout=900)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout[-3000:] + res.stderr[-6000:]
    assert "DIFFERENT" not in res.stdout


@pytest.mark.parametrize("r,n,dmax,seed,window", [
    (5, 3000, 300, 9, "0 0 0 0"),        # one block, the whole graph in the tile
    (6, 2000, 600, 11, "0 64 40 100"),   # 64-node blocks, most successors outside the tile,
                                         # rows above 100 on the whole workgroup
    (5, 2500, 400, 4, "0 128 0 4294967295"),  # the largest tile around 128-node blocks, every
                                              # long row on a wavefront
    (6, 2000, 600, 11, "1 64 40 100"),   # every iteration it can run: systolic and local
                                         # worklists too (bitmaps, visited arcs per arc)
])
def test_emulated_window_kernel_matches_oracle(emu_exes, r, n, dmax, seed, window):
    """The window kernel (log2m 6, delta gathering off) against the oracle
    every iteration, visited arcs included, under AddressSanitizer: the
    iterations that check every node, and in mode 1 the systolic and local
    worklist iterations as well."""
    args = [emu_exes["oracle_check"], "6", str(r), str(n), str(dmax), str(seed), "64", "2"]
    env = dict(os.environ, HB_CHECK_WINDOW=window)
    res = subprocess.run(args, capture_output=True, text=True, timeout=900, env=env)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout[-3000:] + res.stderr[-6000:]
    assert "DIFFERENT" not in res.stdout
    assert res.stdout.count(" window 1 ") >= 5
    if window.startswith("1"):
        modes = re.findall(r"mode (\d+) .* window 1 ", res.stdout)
        assert {"1", "5", "7"} <= set(modes), modes


def test_emulated_count_high_registers_match_oracle(emu_exes):
    """count() on the device path (count_all_kernel, every shape) equals the
    oracle's or_count on crafted counters whose registers sit in the high
    classes random graphs never reach: 28..31 for 5-bit registers (a 31
    register adds -2^-31, the reference's i32 shift, hyper_log_log_logic.rs:
    178) and 28..63 for 6-bit registers, mixed with zeros and uniform values."""
    res = subprocess.run([emu_exes["oracle_count_check"], "36"], capture_output=True, text=True,
                         timeout=900)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout[-3000:] + res.stderr[-6000:]
    assert res.stdout.count("identical") == 23
    assert "registers up to 31: identical" in res.stdout
    assert "registers up to 63: identical" in res.stdout
