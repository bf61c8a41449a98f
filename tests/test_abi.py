"""The C ABI library: loads, exports every declared symbol, and reports the
reference's errors without a GPU (validation happens before any device
call)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(REPO, "include", "hyperball.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hb_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(hbm):
    lib = hbm.lib()
    declared = _declared_symbols()
    assert len(declared) >= 35
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", hbm.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (hb_\w+)", out))
    assert set(declared) <= exported


def test_python_mirror_binds_every_symbol(hbm):
    bound = {name for name, _, _ in hbm._SIGNATURES}
    assert set(_declared_symbols()) == bound


def test_hll_statics_match_oracle(hbm, orc):
    for n in [0, 1, 1000, 10**8, 2**40]:
        assert hbm.register_size(n) == orc.register_size(n) == 5
    for k in range(4, 16):
        assert hbm.rel_std(k) == orc.rel_std(k)
    for rsd in [0.3, 0.1, 0.05, 0.01]:
        assert hbm.log_2_num_of_registers(rsd) == orc.log2_num_registers(rsd)


def _small(hbm):
    return hbm.Graph.from_arcs(5, [(0, 1), (1, 2), (2, 3), (3, 4)])


def test_builder_log2m_error(hbm):
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(_small(hbm), None, 3).build()
    assert e.value.code == hbm.HB_E_INVALID_ARG
    assert "at least 4; got 3" in e.value.message


def test_builder_word_alignment_error(hbm):
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(_small(hbm), None, 4).build()
    assert e.value.code == hbm.HB_E_INVALID_ARG
    assert "Use u16 or smaller" in e.value.message


def test_builder_unknown_counter_memory(hbm):
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(_small(hbm), None, 6).counter_memory(3).build()
    assert e.value.code == hbm.HB_E_INVALID_ARG
    assert "unknown counter memory" in e.value.message


def test_builder_transpose_mismatch(hbm):
    g = _small(hbm)
    t = hbm.Graph.from_arcs(6, [(1, 0)])
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_transpose(g, t, 6).build()
    assert "the transpose should have same number of nodes" in e.value.message
    t2 = hbm.Graph.from_arcs(5, [(1, 0)])
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_transpose(g, t2, 6).build()
    assert e.value.code == hbm.HB_E_INVALID_ARG


def test_builder_weights_length(hbm):
    with pytest.raises(hbm.HyperBallError):
        hbm.HyperBallBuilder.with_hyper_log_log(_small(hbm), None, 6, weights=np.ones(4, np.uint64))


def test_builder_bad_graph(hbm):
    bad = hbm.Graph(np.array([0, 2, 1, 2], np.uint64), np.array([0, 1], np.uint32))
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(bad, None, 6).build()
    assert "monotone" in e.value.message
    bad2 = hbm.Graph(np.array([0, 1, 2], np.uint64), np.array([0, 7], np.uint32))
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(bad2, None, 6).build()
    assert "out of range" in e.value.message


def test_unsupported_shape(hbm):
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(_small(hbm), None, 17).build()
    assert "no kernel" in e.value.message


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_reported(hbm):
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(_small(hbm), None, 6).build()
    assert e.value.code == hbm.HB_E_NO_DEVICE


def test_null_arguments(hbm):
    lib = hbm.lib()
    assert lib.hb_create(None, None, None, None) == hbm.HB_E_INVALID_ARG
    assert lib.hb_run_until_done(None) == hbm.HB_E_INVALID_ARG
    assert b"null" in lib.hb_last_error()
    assert lib.hb_destroy(None) == hbm.HB_OK


def test_struct_layouts_match_the_header(hbm, tmp_path):
    """The ctypes mirrors of hb_csr, hb_graph, hb_config, hb_params and
    hb_iteration_stats have the C layout (sizes and field offsets), as
    compiled by gcc from include/hyperball.h."""
    import ctypes
    structs = {"hb_csr": hbm._Csr, "hb_graph": hbm._Graph, "hb_config": hbm._Config,
               "hb_params": hbm._Params, "hb_iteration_stats": hbm.IterationStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "hyperball.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
