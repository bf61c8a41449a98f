"""The library's own HIP kernels and engine, run on the CPU through the HIP
emulation in tests/emu (g++ with AddressSanitizer and UBSan; lanes are
threads, cross-lane operations are barrier-synchronised exchanges within the
lane group).  An out-of-bounds device access, a lane-group operation that
some lanes of the group skip, or a batch result that differs from the
single-instance result fails here, without a GPU.

The batch mixes all six batched shapes (log2m 6, 8, 10 with 5- and 6-bit
registers), so wavefronts of one workgroup run different shapes at once;
rows up to 900 successors exercise the workgroup-per-node kernel."""
import os
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
    args = [emu_exes["batch_check"], "6", str(graph), "7", "200", "600", "900", str(heavy),
            str(tune)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-3000:] + r.stderr[-6000:]
    assert r.stdout.count("identical") == 6
