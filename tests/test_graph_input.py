"""Host-side graph input: the BVGraph decoder against the files' own
metadata, the transpose, and the synthetic generators."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cnr-2000")


def _props(name):
    out = {}
    for line in open(os.path.join(GOLD, name + ".properties")):
        if "=" in line and not line.startswith("#"):
            k, v = line.strip().split("=", 1)
            out[k] = v
    return out


@pytest.mark.parametrize("name", ["cnr-2000", "cnr-2000-t"])
def test_bvgraph_decode_matches_metadata(hbm, name):
    g = hbm.Graph.load_bvgraph(os.path.join(GOLD, name))
    p = _props(name)
    assert g.num_nodes == int(p["nodes"]) == 325557
    assert g.num_arcs == int(p["arcs"]) == 3216152
    # rows strictly increasing, ids in range
    d = np.diff(g.successors.astype(np.int64))
    starts = np.zeros(g.num_arcs, bool)
    starts[g.offsets[:-1][g.offsets[:-1] < g.num_arcs].astype(np.int64)] = True
    assert np.all((d > 0) | starts[1:])
    assert g.successors.max() < g.num_nodes


@pytest.mark.parametrize("name", ["cnr-2000", "cnr-2000-t"])
def test_bvgraph_bit_offsets_match_offsets_file(hbm, name):
    import ctypes
    base = os.path.join(GOLD, name).encode()
    n = int(_props(name)["nodes"])
    recorded = np.zeros(n + 1, np.uint64)
    seen = np.zeros(n + 1, np.uint64)
    hbm._check(hbm.lib().hb_bvgraph_load_offsets(base, n, recorded.ctypes.data))
    hbm._check(hbm.lib().hb_bvgraph_decoder_offsets(base, seen.ctypes.data, n + 1))
    assert np.array_equal(recorded, seen)


def test_cnr_transpose_consistency(hbm, cnr):
    g, gt = cnr
    t = g.transpose()
    assert np.array_equal(t.offsets, gt.offsets)
    assert np.array_equal(t.successors, gt.successors)
    tt = gt.transpose()
    assert np.array_equal(tt.offsets, g.offsets)
    assert np.array_equal(tt.successors, g.successors)


def test_transpose_brute_force(hbm):
    rng = np.random.default_rng(3)
    arcs = rng.integers(0, 50, size=(400, 2))
    g = hbm.Graph.from_arcs(50, arcs)
    t = g.transpose()
    fwd = {(u, int(v)) for u in range(50) for v in g.successors[g.offsets[u]:g.offsets[u + 1]]}
    bwd = {(int(v), u) for u in range(50) for v in t.successors[t.offsets[u]:t.offsets[u + 1]]}
    assert fwd == bwd


def test_uniform_generator(hbm):
    g = hbm.Graph.uniform(10000, 8, seed=1)
    assert g.num_arcs == 80000
    assert np.all(g.outdegrees() == 8)
    src = np.repeat(np.arange(10000), 8)
    assert not np.any(g.successors == src)  # no self loops
    rows = g.successors.reshape(10000, 8)
    assert np.all(np.diff(rows.astype(np.int64), axis=1) > 0)  # distinct, sorted
    g2 = hbm.Graph.uniform(10000, 8, seed=1)
    assert np.array_equal(g.successors, g2.successors)  # deterministic
    # roughly uniform in-degree
    indeg = np.bincount(g.successors, minlength=10000)
    assert abs(indeg.mean() - 8) < 1e-9 and indeg.std() < 4


def test_web_generator(hbm):
    n = 200000
    g = hbm.Graph.web(n, 2.1, 3, 10000, 1000, 0.7, 2)
    deg = g.outdegrees().astype(np.int64)
    assert deg.max() <= 10000
    src = np.repeat(np.arange(n), deg)
    assert not np.any(g.successors == src)
    d = np.diff(g.successors.astype(np.int64))
    starts = np.zeros(g.num_arcs, bool)
    starts[g.offsets[:-1][g.offsets[:-1] < g.num_arcs].astype(np.int64)] = True
    assert np.all((d > 0) | starts[1:])
    # about 70% of arcs fall in the window (wrapping distance <= 1000)
    dist = np.abs(g.successors.astype(np.int64) - src)
    dist = np.minimum(dist, n - dist)
    frac = np.mean(dist <= 1000)
    assert 0.65 < frac < 0.8
    # preferential attachment makes in-degree heavy tailed
    indeg = np.bincount(g.successors, minlength=n)
    assert indeg.max() > 10 * indeg.mean()  # copy model with 30% copying: exponent ~ 1 + 1/0.3
    # mean out-degree of Zipf(2.1) on [3, 10000] is about 15.7 before dedup
    assert 12 < g.num_arcs / n < 16.5
    g2 = hbm.Graph.web(n, 2.1, 3, 10000, 1000, 0.7, 2)
    assert np.array_equal(g.successors, g2.successors)


@pytest.mark.parametrize("name", ["cnr-2000", "cnr-2000-t"])
def test_bvgraph_parallel_decode_equals_sequential(hbm, name):
    """Parallel decode through .offsets (ranges balanced by bit length, about
    8 per thread, so copy references cross range boundaries many times)
    yields exactly the sequential decode."""
    base = os.path.join(GOLD, name)
    seq = hbm.Graph.load_bvgraph(base, threads=1)
    for t in (2, 3, 8, 64):
        par = hbm.Graph.load_bvgraph(base, threads=t)
        assert np.array_equal(seq.offsets, par.offsets), t
        assert np.array_equal(seq.successors, par.successors), t


def test_bvgraph_parallel_decode_rejects_wrong_offsets(hbm, tmp_path):
    """An .offsets index that does not describe the .graph stream is an error,
    not a silent misdecode."""
    import shutil
    for ext in (".graph", ".properties"):
        shutil.copy(os.path.join(GOLD, "cnr-2000" + ext), tmp_path / ("g" + ext))
    shutil.copy(os.path.join(GOLD, "cnr-2000-t.offsets"), tmp_path / "g.offsets")
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.Graph.load_bvgraph(str(tmp_path / "g"), threads=4)
    assert "BVGraph" in e.value.message
    # one thread ignores the index
    g = hbm.Graph.load_bvgraph(str(tmp_path / "g"), threads=1)
    assert g.num_arcs == 3216152


def test_bvgraph_writer_reproduces_cnr_2000(hbm, tmp_path):
    """The writer with the Java compressor's defaults re-encodes the decoded
    cnr-2000 into the reference's own files, byte for byte: the same
    reference choices, copy blocks, intervals and residual codes, and the
    same offsets stream."""
    g = hbm.Graph.load_bvgraph(os.path.join(GOLD, "cnr-2000"))
    base = str(tmp_path / "cnr-2000")
    g.store_bvgraph(base)
    for ext in (".graph", ".offsets"):
        with open(os.path.join(GOLD, "cnr-2000" + ext), "rb") as f, open(base + ext, "rb") as h:
            assert f.read() == h.read(), ext


def _random_graph(hbm, rng, n, density):
    """Rows mixing random successors, runs of consecutive ids (intervals),
    copies of earlier rows (references), self loops and empty rows."""
    rows = []
    for v in range(n):
        kind = rng.integers(5)
        if kind == 0 or n == 1:
            row = set()
        elif kind == 1 and rows:
            src = rows[max(0, v - int(rng.integers(1, 9)))]
            row = {x for x in src if rng.uniform() < 0.8} | set(rng.integers(0, n, 3).tolist())
        elif kind == 2:
            a = int(rng.integers(0, n))
            row = set(range(a, min(n, a + int(rng.integers(1, 12)))))
        else:
            row = set(rng.integers(0, n, int(rng.integers(1, density))).tolist())
        if rng.uniform() < 0.1:
            row.add(v)
        rows.append(sorted(row))
    return hbm.Graph.from_arcs(n, [(v, s) for v, r in enumerate(rows) for s in r])


BV_CASES = [
    {},
    {"window_size": 0},
    {"min_interval_length": 0},
    {"window_size": 1, "max_ref_count": 1},
    {"window_size": 16, "max_ref_count": 1000},
    {"max_ref_count": 0},
    {"min_interval_length": 2, "zeta_k": 1},
    {"zeta_k": 5, "residuals": 3},
    {"outdegrees": 1, "references": 0, "blocks": 1, "block_count": 3, "residuals": 1,
     "offsets": 1},
    {"outdegrees": 3, "references": 3, "blocks": 3, "block_count": 1, "residuals": 0,
     "offsets": 3, "zeta_k": 2},
    {"outdegrees": 2, "blocks": 2, "block_count": 2, "residuals": 2, "offsets": 0},
]


@pytest.mark.parametrize("case", range(len(BV_CASES)))
def test_bvgraph_round_trip(hbm, tmp_path, case):
    """decode(encode(G)) == G for every code of every component, windows and
    reference depths from none to deep, intervals on and off: sequential and
    parallel decodes agree, the .offsets stream is the decoder's own record
    of bit positions, and .properties carries the parameters."""
    import ctypes
    params = BV_CASES[case]
    rng = np.random.default_rng(100 + case)
    for n in (1, 2, 37, 3000):
        g = _random_graph(hbm, rng, n, 30)
        base = str(tmp_path / f"g{n}")
        g.store_bvgraph(base, **params)
        for threads in (1, 4):
            h = hbm.Graph.load_bvgraph(base, threads=threads)
            assert np.array_equal(h.offsets, g.offsets), (n, threads)
            assert np.array_equal(h.successors, g.successors), (n, threads)
        recorded = np.zeros(n + 1, np.uint64)
        seen = np.zeros(n + 1, np.uint64)
        hbm._check(hbm.lib().hb_bvgraph_load_offsets(base.encode(), n, recorded.ctypes.data))
        hbm._check(hbm.lib().hb_bvgraph_decoder_offsets(base.encode(), seen.ctypes.data, n + 1))
        assert np.array_equal(recorded, seen)
        props = dict(l.strip().split("=", 1) for l in open(base + ".properties")
                     if "=" in l and not l.startswith("#"))
        assert int(props["nodes"]) == n and int(props["arcs"]) == g.num_arcs
        assert int(props["windowsize"]) == params.get("window_size", 7)


def test_bvgraph_round_trip_transpose(hbm, cnr, tmp_path):
    """cnr-2000-t was compressed by another writer: its decode survives a
    round trip through this writer, whose stream is no longer."""
    _, gt = cnr
    base = str(tmp_path / "t")
    gt.store_bvgraph(base)
    h = hbm.Graph.load_bvgraph(base)
    assert np.array_equal(h.offsets, gt.offsets) and np.array_equal(h.successors, gt.successors)
    assert os.path.getsize(base + ".graph") <= os.path.getsize(os.path.join(GOLD, "cnr-2000-t.graph"))


def test_bvgraph_writer_rejects(hbm, tmp_path):
    g = hbm.Graph.from_arcs(3, [(0, 1), (0, 2)])
    with pytest.raises(hbm.HyperBallError) as e:
        g.store_bvgraph(str(tmp_path / "x"), residuals=7)
    assert e.value.code == hbm.HB_E_INVALID_ARG
    with pytest.raises(TypeError):
        g.store_bvgraph(str(tmp_path / "x"), bogus=1)
    # a row with a repeated successor cannot be written
    dup = hbm.Graph(np.array([0, 2, 2, 2], np.uint64), np.array([1, 1], np.uint32))
    with pytest.raises(hbm.HyperBallError) as e:
        dup.store_bvgraph(str(tmp_path / "y"))
    assert "strictly increasing" in e.value.message


# ---- Elias-Fano node indices and cumulative degrees (webgraph's .ef, .dcf) --

GOLD_IN = os.path.join(os.path.dirname(GOLD), "in-2004")


@pytest.mark.parametrize("gold,name", [(GOLD, "cnr-2000"), (GOLD, "cnr-2000-t"),
                                       (GOLD_IN, "in-2004")])
def test_elias_fano_index_equals_offsets_file(hbm, gold, name):
    # The reference's .ef and the Java .offsets are two encodings of the same
    # node-record bit offsets (in-2004's .graph is not in the fixtures, but
    # both of its indices are).
    base = os.path.join(gold, name)
    n = 0
    for line in open(base + ".properties"):
        if line.startswith("nodes="):
            n = int(line.split("=")[1])
    ef = hbm.load_elias_fano(base + ".ef")
    recorded = np.zeros(n + 1, np.uint64)
    hbm._check(hbm.lib().hb_bvgraph_load_offsets(base.encode(), n, recorded.ctypes.data))
    assert ef.shape == (n + 1,)
    assert np.array_equal(ef, recorded)
    assert ef[0] == 0 and np.all(np.diff(ef.astype(np.int64)) > 0)


def test_elias_fano_index_equals_decoder_offsets(hbm):
    base = os.path.join(GOLD, "cnr-2000")
    seen = np.zeros(325558, np.uint64)
    hbm._check(hbm.lib().hb_bvgraph_decoder_offsets(base.encode(), seen.ctypes.data, seen.size))
    assert np.array_equal(hbm.load_elias_fano(base + ".ef"), seen)


def test_dcf_equals_decoded_cumulative_degrees(hbm, cnr):
    g, _ = cnr
    dcf = hbm.load_elias_fano(os.path.join(GOLD, "cnr-2000.dcf"))
    assert np.array_equal(dcf, g.offsets)


def test_parallel_decode_through_elias_fano_index(hbm, cnr, tmp_path):
    # .graph, .properties and .ef only: the parallel decoder takes its node
    # ranges from the Elias-Fano index
    import shutil
    g, _ = cnr
    for ext in (".graph", ".properties", ".ef"):
        shutil.copy(os.path.join(GOLD, "cnr-2000" + ext), tmp_path / ("cnr-2000" + ext))
    base = str(tmp_path / "cnr-2000")
    for threads in (1, 4, 16):
        h = hbm.Graph.load_bvgraph(base, threads=threads)
        assert np.array_equal(h.offsets, g.offsets)
        assert np.array_equal(h.successors, g.successors)
    # an index that does not belong to the graph is caught: the transpose's
    # .ef has the same node count but other record offsets
    shutil.copy(os.path.join(GOLD, "cnr-2000-t.ef"), tmp_path / "cnr-2000.ef")
    with pytest.raises(hbm.HyperBallError, match="node index"):
        hbm.Graph.load_bvgraph(base, threads=4)


def test_elias_fano_rejects(hbm, tmp_path):
    import ctypes
    src = open(os.path.join(GOLD, "cnr-2000.ef"), "rb").read()
    cases = {
        "magic": b"epserdX " + src[8:],
        "truncated": src[: len(src) // 2],
        "header": src[:20],
        "type": src[:37] + src[37:37 + 10].replace(b"sux", b"sox") + src[47:],
    }
    for what, data in cases.items():
        p = tmp_path / (what + ".ef")
        p.write_bytes(data)
        with pytest.raises(hbm.HyperBallError, match="Elias-Fano"):
            hbm.load_elias_fano(str(p))
    with pytest.raises(hbm.HyperBallError, match="cannot open"):
        hbm.load_elias_fano(str(tmp_path / "missing.ef"))
    cnt = ctypes.c_uint64()
    small = np.zeros(10, np.uint64)
    with pytest.raises(hbm.HyperBallError, match="too small"):
        hbm._check(hbm.lib().hb_elias_fano_load(os.path.join(GOLD, "cnr-2000.ef").encode(),
                                                small.ctypes.data, small.size, ctypes.byref(cnt)))
    assert cnt.value == 325558
