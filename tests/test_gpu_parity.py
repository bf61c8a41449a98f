"""GPU parity: the HIP engine against the CPU oracle, through the C ABI.

Registers (in the reference layout) and modified bits must be bit-exact with
the sequential Jacobi oracle after every iteration
(src/algo/hyperball/hyperball_impl.rs:1302-1319 in the reference); f64
accumulators must be bit-exact with the oracle's restatement of the reference
engine; the neighbourhood function agrees within 1e-12 relative (the
reference's own value depends on thread order).
"""
import numpy as np
import pytest

from golden_io import assert_array_equal_rel, cnr_golden

pytestmark = pytest.mark.gpu


def _tiny_graphs(hbm):
    n = 40
    path = hbm.Graph.from_arcs(n, [(i, i + 1) for i in range(n - 1)])
    cycle = hbm.Graph.from_arcs(n, [(i, (i + 1) % n) for i in range(n)])
    star = hbm.Graph.from_arcs(n, [(0, i) for i in range(1, n)] + [(i, 0) for i in range(1, 5)])
    clique = hbm.Graph.from_arcs(12, [(i, j) for i in range(12) for j in range(12) if i != j])
    empty = hbm.Graph.from_arcs(10, [])
    loops = hbm.Graph.from_arcs(10, [(i, i) for i in range(10)] + [(1, 2), (2, 2)])
    return {"path": path, "cycle": cycle, "star": star, "clique": clique, "empty": empty,
            "self_loops": loops}


def _run_lockstep(hbm, orc, g, gt, log2m, reg_bits=0, jenkins=None, max_iter=10_000, heavy=None,
                  delta=None, counter_memory=None, tile_nodes=0, streaming=None, window=None):
    b = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, log2m).register_bits(reg_bits)
    if jenkins is not None:
        b = b.jenkins(jenkins)
    if counter_memory is not None:
        b = b.counter_memory(counter_memory, tile_nodes)
    hb = b.build()
    if counter_memory is not None:
        assert hb.params()["counter_memory"] == counter_memory
    if heavy is not None:
        hb.set_tuning(heavy)
    if delta is not None:
        hb.set_delta_gathering(delta)
    if streaming is not None:
        hb.set_streaming(streaming)
    if window is not None:
        hb.set_window(*window)
    hll = orc.Hll(log2m, reg_bits, g.num_nodes, orc.HASH_JENKINS if jenkins is not None else orc.HASH_SIPHASH13,
                  jenkins or 0)
    ref = orc.Jacobi(hll, g.offsets, g.successors)
    hb.init()
    assert np.array_equal(hb.counters(0), ref.curr), "seeded registers differ"
    assert not hb.counters(1).any(), "the other buffer is not zero before the first iteration"
    it = 0
    while it < max_iter:
        hb.iterate()
        k = ref.iterate()
        it += 1
        assert np.array_equal(hb.counters(0), ref.curr), f"registers differ after iteration {it}"
        assert np.array_equal(hb.counters(1), ref.next), f"previous registers differ after iteration {it}"
        assert np.array_equal(hb.modified_bits(), ref.modified), f"modified bits differ after iteration {it}"
        if k == 0:
            break
    stats = hb.iteration_stats()
    hb.close()
    return it, stats


@pytest.mark.parametrize("name", ["path", "cycle", "star", "clique", "empty", "self_loops"])
@pytest.mark.parametrize("transpose", [False, True])
def test_tiny_graphs_bit_exact(hbm, orc, name, transpose):
    g = _tiny_graphs(hbm)[name]
    gt = g.transpose() if transpose else None
    _run_lockstep(hbm, orc, g, gt, 6, jenkins=42)


@pytest.mark.parametrize("log2m,reg_bits", [(5, 6), (6, 5), (7, 5), (8, 5), (9, 6), (10, 5), (10, 6),
                                            (11, 5), (12, 5), (12, 6), (13, 5), (14, 6), (15, 5),
                                            (16, 5), (16, 6)])
def test_uniform_all_shapes_bit_exact(hbm, orc, log2m, reg_bits):
    g = hbm.Graph.uniform(3000, 3, seed=11)
    _run_lockstep(hbm, orc, g, g.transpose(), log2m, reg_bits, jenkins=7)


@pytest.mark.parametrize("hash_", ["siphash", "jenkins"])
def test_web_systolic_local_bit_exact(hbm, orc, hash_):
    g = hbm.Graph.web(20000, dmin=1, dmax=200, window=100, seed=5)
    gt = g.transpose()
    it, stats = _run_lockstep(hbm, orc, g, gt, 8, jenkins=(None if hash_ == "siphash" else 42))
    modes = {(s["systolic"], s["local"], s["pre_local"]) for s in stats}
    assert any(m[0] for m in modes), "no systolic iteration was exercised"
    first = next(i for i, s in enumerate(stats) if s["systolic"])
    assert stats[first]["checked"] < g.num_nodes, "the first systolic iteration checked every node"


@pytest.mark.parametrize("log2m,reg_bits", [(8, 5), (6, 6), (10, 6), (12, 6), (11, 5), (13, 6),
                                            (14, 5)])
def test_engine_outputs_match_reference_engine(hbm, orc, log2m, reg_bits):
    g = hbm.Graph.web(30000, dmin=1, dmax=300, window=200, seed=9)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, log2m).register_bits(reg_bits).jenkins(3)
          .sum_of_distances(True).sum_of_inverse_distances(True).build())
    hb.run_until_done()
    ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=log2m,
                              reg_bits=reg_bits, hash=orc.HASH_JENKINS, seed=3, sum_of_distances=True,
                              sum_of_inverse_distances=True)
    ref.run()
    assert np.array_equal(hb.counters(0), ref.counters(0))
    # per-node f64 accumulators are deterministic in the reference: bit-exact
    assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())
    assert np.array_equal(hb.harmonic_centralities(), ref.harmonic_centralities())
    assert np.array_equal(hb.lin_centrality(), ref.lin_centrality())
    assert np.array_equal(hb.closeness_centrality(), ref.closeness_centrality())
    assert np.array_equal(hb.nieminen_centrality(), ref.nieminen_centrality())
    assert np.array_equal(hb.reachable_nodes(), ref.reachable_nodes())
    nf, nf_ref = hb.neighborhood_function(), ref.neighborhood_function()
    assert len(nf) == len(nf_ref)
    np.testing.assert_allclose(nf, nf_ref, rtol=1e-12, atol=0)


def test_cnr_2000_java_goldens(hbm, cnr):
    """tests/test_hyperball.rs:221-302 of the reference: Jenkins seed 42,
    log2m 8, sum of distances and harmonic, run until done, 1e-6 relative."""
    g, gt = cnr
    hb = (hbm.HyperBallBuilder.with_transpose(g, gt, 8).jenkins(42)
          .sum_of_distances(True).sum_of_inverse_distances(True).build())
    hb.run_until_done()
    threshold = 1e-6
    assert_array_equal_rel(cnr_golden("sum_of_distances"), hb.sum_of_distances(), threshold, "sum of distances")
    assert_array_equal_rel(cnr_golden("harmonic_centrality"), hb.harmonic_centralities(), threshold,
                           "harmonic centralities")
    assert_array_equal_rel(cnr_golden("lin_centrality"), hb.lin_centrality(), threshold, "lin centralities")
    assert_array_equal_rel(cnr_golden("closeness_centrality"), hb.closeness_centrality(), threshold,
                           "closeness centralities")
    assert_array_equal_rel(cnr_golden("nieminen_centrality"), hb.nieminen_centrality(), threshold,
                           "nieminen centralities")


def test_cnr_2000_par_equals_seq_every_iteration(hbm, orc, cnr):
    """The reference's white-box test (hyperball_impl.rs:1263-1322): log2m 6,
    DefaultHasher, with transpose; every iteration equals sequential Jacobi."""
    g, gt = cnr
    it, stats = _run_lockstep(hbm, orc, g, gt, 6)
    assert it > 10
    assert any(s["systolic"] for s in stats) and any(s["local"] for s in stats)
    # the first systolic iteration ran over the marked predecessors, or, on the
    # window kernel, checked every node as the reference does
    first = next(i for i, s in enumerate(stats) if s["systolic"])
    assert not stats[first - 1]["systolic"]
    if stats[first]["window"]:
        assert stats[first]["checked"] == g.num_nodes
    else:
        assert stats[first]["checked"] < g.num_nodes


@pytest.mark.parametrize("heavy", [64, 0, 2**32 - 1])
@pytest.mark.parametrize("log2m,reg_bits", [(6, 5), (8, 5), (10, 6), (12, 5)])
def test_heavy_rows_bit_exact(hbm, orc, heavy, log2m, reg_bits):
    """Rows above the threshold take the workgroup-per-node kernel; results
    must not depend on the threshold."""
    rng = np.random.default_rng(log2m)
    n = 6000
    arcs = [(0, int(v)) for v in rng.choice(np.arange(1, n), 3000, replace=False)]
    arcs += [(7, int(v)) for v in rng.choice(np.arange(n), 700, replace=False) if v != 7]
    arcs += [(int(u), int((u * 7 + 13) % n)) for u in range(n)]
    arcs += [(int(u), int((u + 1) % n)) for u in range(0, n, 3)]
    g = hbm.Graph.from_arcs(n, arcs)
    _run_lockstep(hbm, orc, g, g.transpose(), log2m, reg_bits, jenkins=5, heavy=heavy)


@pytest.mark.parametrize("delta", [1, 2])
@pytest.mark.parametrize("log2m,reg_bits", [(5, 6), (6, 5), (8, 6), (10, 6), (12, 5)])
def test_delta_gathering_bit_exact(hbm, orc, delta, log2m, reg_bits):
    """Delta gathering forced on (1) and off (2): every iteration bit-exact
    against the Jacobi oracle, heavy rows included; with it on, only a part
    of the successor slabs is read."""
    g = hbm.Graph.web(20000, 2.1, 1, 2000, 300, 0.7, 13)
    it, stats = _run_lockstep(hbm, orc, g, g.transpose(), log2m, reg_bits, jenkins=3, heavy=64,
                              delta=delta)
    assert it > 5
    assert all(s["delta_gathering"] == (delta == 1) for s in stats)
    slabs = max(1, (1 << log2m) // 32)
    # iteration 0 computes its successors' seeded registers from their hashes
    assert stats[0]["gathered_slabs"] == 0 and stats[0]["visited_arcs"] > 0
    gathered = sum(s["gathered_slabs"] for s in stats[1:])
    total = sum(s["visited_arcs"] for s in stats[1:]) * slabs
    if delta == 2 or slabs == 1:
        assert gathered == total
    else:
        assert gathered < total


@pytest.mark.parametrize("window", [(0, 64, 40, 100), (0, 256, 200, 64), (0, 128, 0, 2**32 - 1),
                                    (0, 0, 0, 0), (2, 0, 0, 0)])
@pytest.mark.parametrize("reg_bits", [5, 6])
def test_window_kernel_bit_exact(hbm, orc, window, reg_bits):
    """The window kernel (64-register counters, iterations that check every
    node): every iteration bit-exact against the Jacobi oracle, with blocks
    and tiles small enough that most successors lie outside the tile, rows
    taking a wavefront or a workgroup, and switched off (mode 2).  Without a
    transpose every iteration checks every node, so every one but the seeded
    first runs the window kernel."""
    g = hbm.Graph.web(20000, 2.1, 1, 2000, 300, 0.7, 17)
    it, stats = _run_lockstep(hbm, orc, g, None, 6, reg_bits, jenkins=3, window=window)
    assert it > 5
    assert [s["window"] for s in stats] == [int(i > 0 and window[0] != 2) for i in range(it)]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("reg_bits", [5, 6])
@pytest.mark.parametrize("acc", ["ACC_F64", "ACC_F32_KAHAN", "ACC_F32"])
def test_window_kernel_engine_outputs(hbm, orc, reg_bits, acc, mode):
    """Window iterations against the oracle's restatement of the reference
    engine, with a transpose (dense iterations, then systolic and local
    worklists: large ones on the window kernel by the rule in mode 0, all of
    them in mode 1): registers, modified bits, visited arcs (from in-degrees
    or counted per arc) and modified counts every iteration; sums of
    distances and harmonic centralities bit-exact in f64 and within 1e-6
    relative with f32 accumulators (Kahan-compensated or not)."""
    g = hbm.Graph.web(30000, 2.1, 1, 3000, 400, 0.7, 21)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 6).register_bits(reg_bits).jenkins(11)
          .sum_of_distances(True).sum_of_inverse_distances(True)
          .accumulator_precision(getattr(hbm, acc)).build())
    hb.set_window(mode, 512, 300, 200)
    ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=6,
                              reg_bits=reg_bits, hash=orc.HASH_JENKINS, seed=11,
                              sum_of_distances=True, sum_of_inverse_distances=True)
    hb.init()
    ref.init()
    windows, systolic_windows = 0, 0
    for it in range(1000):
        hb.iterate()
        ref.iterate()
        s = hb.iteration_stats()[-1]
        windows += s["window"]
        systolic_windows += s["window"] and s["systolic"]
        assert np.array_equal(hb.counters(0), ref.counters(0)), f"registers differ at {it}"
        assert np.array_equal(hb.modified_bits(), ref.modified_bits()), f"modified bits differ at {it}"
        assert s["visited_arcs"] == ref.visited_arcs, f"visited arcs differ at {it}"
        assert s["modified"] == ref.modified
        if ref.modified == 0:
            break
    assert windows >= 3
    assert systolic_windows >= 1
    if acc == "ACC_F64":
        assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())
        assert np.array_equal(hb.harmonic_centralities(), ref.harmonic_centralities())
    else:
        np.testing.assert_allclose(hb.sum_of_distances(), ref.sum_of_distances(), rtol=1e-6, atol=0)
        np.testing.assert_allclose(hb.harmonic_centralities(), ref.harmonic_centralities(),
                                   rtol=1e-6, atol=0)
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(), rtol=1e-12,
                               atol=0)
    hb.close()


@pytest.mark.parametrize("log2m,reg_bits", [(6, 5), (8, 6), (10, 6), (12, 5)])
def test_host_counter_memory_bit_exact(hbm, orc, log2m, reg_bits):
    """Counter arrays in pinned host memory (the kernels read and write them
    over the host link): every iteration bit-exact against the Jacobi
    oracle, systolic and heavy rows included."""
    g = hbm.Graph.web(6000, 2.1, 1, 1500, 100, 0.7, 3 + log2m)
    it, stats = _run_lockstep(hbm, orc, g, g.transpose(), log2m, reg_bits, jenkins=2, heavy=128,
                              counter_memory=hbm.COUNTERS_HOST)
    assert it > 5 and any(s["systolic"] for s in stats)


@pytest.mark.parametrize("log2m,reg_bits,tile", [(6, 5, 512), (8, 6, 1024), (10, 6, 640),
                                                 (12, 5, 6016), (13, 5, 768), (8, 5, 0)])
def test_streamed_counter_memory_bit_exact(hbm, orc, log2m, reg_bits, tile):
    """Counter arrays in pinned host memory streamed through device memory in
    tiles (uploaded source blocks merged pass by pass): every iteration
    bit-exact against the Jacobi oracle, with systolic, local and pre-local
    iterations, many tiles and blocks, one tile (6016 >= n), counters of two
    segments, and the automatic tile size."""
    g = hbm.Graph.web(6000, 2.1, 1, 1500, 100, 0.7, 3 + log2m)
    it, stats = _run_lockstep(hbm, orc, g, g.transpose(), log2m, reg_bits, jenkins=2,
                              counter_memory=hbm.COUNTERS_STREAM, tile_nodes=tile, streaming=1)
    assert it > 5 and any(s["systolic"] for s in stats)
    assert any(s["local"] for s in stats)
    assert all(s["streamed"] for s in stats)


def test_streamed_counter_memory_rule_bit_exact(hbm, orc):
    """The per-iteration rule of streamed counters: the seeded first
    iteration and the sparse late ones read the host arrays in place, the
    dense ones stream them (2 tiles); every iteration bit-exact."""
    g = hbm.Graph.web(20000, 2.1, 3, 1500, 200, 0.7, 8)
    it, stats = _run_lockstep(hbm, orc, g, g.transpose(), 8, 5, jenkins=3,
                              counter_memory=hbm.COUNTERS_STREAM, tile_nodes=10016)
    streamed = [s["streamed"] for s in stats]
    assert streamed[0] == 0 and streamed[-1] == 0 and any(streamed), streamed
