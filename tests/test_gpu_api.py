"""GPU tests of the rest of the API surface: error paths, run variants,
weights, discount functions, accumulator precision, and full-size runs
checked through size-independent properties."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _web(hbm, n=20000, seed=9):
    g = hbm.Graph.web(n, dmin=1, dmax=300, window=200, seed=seed)
    return g, g.transpose()


def _ref(orc, g, gt, log2m=8, **kw):
    return orc.OracleHyperBall(g.offsets, g.successors,
                               None if gt is None else gt.offsets,
                               None if gt is None else gt.successors, log2m=log2m, **kw)


def test_accessors_before_run(hbm):
    g, gt = _web(hbm, 500)
    hb = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 6).sum_of_distances(True).build()
    for f in (hb.neighborhood_function, hb.sum_of_distances, hb.closeness_centrality,
              hb.reachable_nodes, lambda: hb.reachable_nodes_from(0)):
        with pytest.raises(hbm.HyperBallError) as e:
            f()
        assert e.value.code == hbm.HB_E_NOT_RUN
        assert "HyperBall was not run" in e.value.message
    # upper_bound 0: no iteration, accessors still fail (hyperball_impl.rs:624-630)
    hb.run(0)
    with pytest.raises(hbm.HyperBallError):
        hb.neighborhood_function()


def test_outputs_not_requested(hbm):
    g, gt = _web(hbm, 500)
    hb = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 6).build()
    hb.run_until_done()
    for f, text in ((hb.sum_of_distances, "Sum of distances were not requested"),
                    (hb.harmonic_centralities, "Sum of inverse distances were not requested"),
                    (hb.closeness_centrality, "closeness centrality"),
                    (hb.lin_centrality, "lin centrality"),
                    (hb.nieminen_centrality, "lin centrality"),
                    (lambda: hb.discounted_centrality(0), "Discount centrality of index 0")):
        with pytest.raises(hbm.HyperBallError) as e:
            f()
        assert e.value.code == hbm.HB_E_NOT_REQUESTED
        assert text in e.value.message
    with pytest.raises(hbm.HyperBallError) as e:
        hb.reachable_nodes_from(500)
    assert e.value.code == hbm.HB_E_INVALID_ARG
    assert hb.reachable_nodes_from(7) == hb.reachable_nodes()[7]


@pytest.mark.parametrize("ub,threshold", [(1, None), (3, None), (100, 0.05), (100, 0.001), (2**64 - 1, 0.0)])
def test_upper_bound_and_threshold(hbm, orc, ub, threshold):
    """run(upper_bound, threshold): stop rules of hyperball_impl.rs:553-578."""
    g, gt = _web(hbm)
    hb = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(11).build()
    hb.run(ub, threshold)
    ref = _ref(orc, g, gt, hash=orc.HASH_JENKINS, seed=11)
    ref.run(ub, threshold)
    nf, nf_ref = hb.neighborhood_function(), ref.neighborhood_function()
    assert len(nf) == len(nf_ref) == ref.iteration + 1
    np.testing.assert_allclose(nf, nf_ref, rtol=1e-12)
    assert np.array_equal(hb.counters(0), ref.counters(0))


def test_no_transpose_standard_iterations(hbm, orc):
    g, _ = _web(hbm)
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, None, 8).sum_of_distances(True)
          .sum_of_inverse_distances(True).build())
    hb.run_until_done()
    ref = _ref(orc, g, None, sum_of_distances=True, sum_of_inverse_distances=True)
    ref.run()
    assert not any(s["systolic"] for s in hb.iteration_stats())
    assert np.array_equal(hb.counters(0), ref.counters(0))
    assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())
    assert np.array_equal(hb.harmonic_centralities(), ref.harmonic_centralities())
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(), rtol=1e-12)


def test_default_hasher_matches_oracle(hbm, orc):
    """SipHash-1-3 with zero keys on device and in the oracle (the value
    itself is pinned only through the published SipHash round function)."""
    g, gt = _web(hbm, 5000)
    hb = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 7).register_bits(6).build()
    hb.init()
    hll = orc.Hll(7, 6, 5000)
    ref = orc.Jacobi(hll, g.offsets, g.successors)
    assert np.array_equal(hb.counters(0), ref.curr)


def test_weighted_init(hbm, orc):
    """Weighted nodes add weights[v] items each (hyperball_impl.rs:1167-1174),
    drawn from the seeded stand-in for rand::random."""
    g, gt = _web(hbm, 8000)
    w = (np.arange(8000) % 5).astype(np.uint64)
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8, weights=w).weight_seed(123)
          .sum_of_distances(True).build())
    hb.run_until_done()
    ref = _ref(orc, g, gt, weights=w, weight_seed=123, sum_of_distances=True)
    ref.run()
    assert np.array_equal(hb.counters(0), ref.counters(0))
    assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(), rtol=1e-12)


def test_discount_functions(hbm, orc):
    """discount_function (hyperball_impl.rs:291-297, :1084-1091)."""
    g, gt = _web(hbm)
    fns = (lambda d: 1.0 / (d * d), lambda d: math.exp(-d), lambda d: 1.0 / math.log2(d + 1))
    b = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(5)
    for f in fns:
        b = b.discount_function(f)
    hb = b.build()
    hb.run_until_done()
    ref = _ref(orc, g, gt, hash=orc.HASH_JENKINS, seed=5, discount_functions=fns)
    ref.run()
    for k in range(len(fns)):
        assert np.array_equal(hb.discounted_centrality(k), ref.discounted_centrality(k))
    with pytest.raises(hbm.HyperBallError):
        hb.discounted_centrality(3)


def test_no_discount_function_clears(hbm):
    g, gt = _web(hbm, 500)
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 6).discount_function(lambda d: 1.0)
          .no_discount_function().build())
    hb.run_until_done()
    with pytest.raises(hbm.HyperBallError):
        hb.discounted_centrality(0)


@pytest.mark.parametrize("mode,bound", [("f32_kahan", 1e-6), ("f32", 1e-3)])
def test_accumulator_precision(hbm, orc, mode, bound):
    """FP32 accumulators against the f64 oracle: Kahan-compensated sums stay
    within 1e-6 relative; plain f32 within 1e-3 on this graph."""
    g, gt = _web(hbm, 30000)
    acc = hbm.ACC_F32_KAHAN if mode == "f32_kahan" else hbm.ACC_F32
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(2).sum_of_distances(True)
          .sum_of_inverse_distances(True).accumulator_precision(acc).build())
    hb.run_until_done()
    ref = _ref(orc, g, gt, hash=orc.HASH_JENKINS, seed=2, sum_of_distances=True,
               sum_of_inverse_distances=True)
    ref.run()
    for got, exp in ((hb.sum_of_distances(), ref.sum_of_distances()),
                     (hb.harmonic_centralities(), ref.harmonic_centralities())):
        rel = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-300)
        rel[exp == 0] = np.abs(got[exp == 0])
        assert rel.max() < bound, rel.max()


def test_uniform_graph_full_size_bit_exact(hbm, orc):
    """n = 10^5 nodes, 8 uniform distinct successors each, log2m 8: every
    iteration bit-exact against the Jacobi oracle, outputs against the
    engine restatement."""
    g = hbm.Graph.uniform(100_000, 8, seed=1)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(42).sum_of_distances(True)
          .sum_of_inverse_distances(True).build())
    J = orc.Jacobi(orc.Hll(8, 0, 100_000, orc.HASH_JENKINS, 42), g.offsets, g.successors)
    hb.init()
    while True:
        hb.iterate()
        k = J.iterate()
        assert np.array_equal(hb.counters(0), J.curr)
        assert np.array_equal(hb.modified_bits(), J.modified)
        if k == 0:
            break
    ref = _ref(orc, g, gt, hash=orc.HASH_JENKINS, seed=42, sum_of_distances=True,
               sum_of_inverse_distances=True)
    ref.run()
    assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())
    assert np.array_equal(hb.harmonic_centralities(), ref.harmonic_centralities())
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(), rtol=1e-12)


@pytest.mark.parametrize("reg_bits", [5, 6])
def test_web_graph_full_size_properties(hbm, orc, reg_bits):
    """The web-scale graph at full size (n = 10^7): properties that hold at
    any size.  The fixed point is closed under successors (every counter
    dominates its successors' counters), one more iteration changes nothing,
    the neighbourhood function starts at n and is non-decreasing, and a
    second run reproduces every value bit for bit."""
    g = hbm.Graph.web(10_000_000, 2.1, 3, 10000, 1000, 0.7, 2)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 10).register_bits(reg_bits).jenkins(42)
          .sum_of_distances(True).build())
    hb.run_until_done()
    nf = hb.neighborhood_function()
    assert nf[0] == 10_000_000 and np.all(np.diff(nf) >= 0)
    sd1 = hb.sum_of_distances()
    hb.iterate()
    assert hb.iteration_stats()[-1]["modified"] == 0
    cnt = hb.counters(0)
    hll = orc.Hll(10, reg_bits, 10_000_000, orc.HASH_JENKINS, 42)
    rng = np.random.default_rng(0)
    for v in rng.choice(10_000_000, 300, replace=False):
        x = cnt[v].copy()
        for s in g.successors[g.offsets[v]:g.offsets[v + 1]][:50]:
            y = x.copy()
            hll.merge_max(y, cnt[s])
            assert np.array_equal(x, y), f"counter {v} does not dominate successor {s}"
    # reachable counts from the cache equal a fresh count of the counters
    sample = rng.choice(10_000_000, 200, replace=False)
    reach = hb.reachable_nodes()
    for v in sample:
        assert reach[v] == hll.count(cnt[v])
    hb.run_until_done()
    assert np.array_equal(hb.sum_of_distances(), sd1)
    assert np.array_equal(hb.neighborhood_function(), nf)


def test_autotune_keeps_results(hbm, orc):
    """hb_autotune picks one of the thresholds it timed, resets the run
    state, and the run after it equals the oracle."""
    g = hbm.Graph.web(30000, 2.1, 1, 3000, 300, 0.7, 4)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(3).sum_of_distances(True)
          .build())
    hb.run_until_done()
    chosen, table = hb.autotune(reps=2)
    assert len(table) >= 3 and table[-1][0] == 2**32 - 1
    assert chosen in [t for t, _ in table]
    assert all(ms > 0 for _, ms in table)
    with pytest.raises(hbm.HyperBallError):
        hb.sum_of_distances()  # run state was reset
    hb.run_until_done()
    ref = _ref(orc, g, gt, hash=orc.HASH_JENKINS, seed=3, sum_of_distances=True)
    ref.run()
    assert np.array_equal(hb.counters(0), ref.counters(0))
    assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())


def _transpose_cases(hbm, cnr):
    rng = np.random.default_rng(11)
    yield "empty", hbm.Graph.from_arcs(0, [])
    yield "one node", hbm.Graph.from_arcs(1, [])
    yield "self loops", hbm.Graph.from_arcs(5, [(0, 0), (1, 1), (1, 0), (4, 2), (2, 4), (3, 3)])
    yield "no arcs", hbm.Graph.from_arcs(1000, [])
    yield "star in", hbm.Graph.from_arcs(3000, [(int(v), 0) for v in range(1, 3000)])
    yield "uniform", hbm.Graph.uniform(50_000, 8, seed=int(rng.integers(1 << 30)))
    yield "web", hbm.Graph.web(200_000, 2.1, 1, 5000, 500, 0.7, 3)
    yield "cnr-2000", cnr[0]


def test_device_transpose_matches_host(hbm, cnr):
    """hb_graph_transpose_device (in-degree atomics, scan, stable radix sort)
    returns exactly the host counting-sort transpose: same offsets, same
    rows in the same (sorted) order."""
    for name, g in _transpose_cases(hbm, cnr):
        th, td = g.transpose(), g.transpose(device=0)
        assert np.array_equal(th.offsets, td.offsets), name
        assert np.array_equal(th.successors, td.successors), name
        # and it is the transpose: (u, v) in G <=> (v, u) in G^T
        if g.num_arcs:
            src = np.repeat(np.arange(g.num_nodes, dtype=np.uint64), np.diff(g.offsets.astype(np.int64)))
            tsrc = np.repeat(np.arange(g.num_nodes, dtype=np.uint64), np.diff(td.offsets.astype(np.int64)))
            a = np.sort(src * g.num_nodes + g.successors.astype(np.uint64))
            b = np.sort(td.successors.astype(np.uint64) * g.num_nodes + tsrc)
            assert np.array_equal(a, b), name
    # the reference's own transposed file of cnr-2000 (BVGraph cnr-2000-t)
    td = cnr[0].transpose(device=0)
    assert np.array_equal(td.offsets, cnr[1].offsets)
    assert np.array_equal(td.successors, cnr[1].successors)


def test_instance_with_device_transpose(hbm):
    """transpose_on_device(): systolic iterations from a transpose built in
    place equal those from a host transpose."""
    g, gt = _web(hbm, 40000, seed=21)
    runs = []
    for b in (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8),
              hbm.HyperBallBuilder.with_hyper_log_log(g, None, 8).transpose_on_device()):
        hb = b.jenkins(8).sum_of_distances(True).sum_of_inverse_distances(True).build()
        hb.run_until_done()
        runs.append((hb.counters(0), hb.sum_of_distances(), hb.harmonic_centralities(),
                     hb.neighborhood_function(), hb.iteration_stats()))
    (c1, s1, h1, n1, st1), (c2, s2, h2, n2, st2) = runs
    assert any(s["systolic"] for s in st2)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2) and np.array_equal(h1, h2)
    assert np.array_equal(n1, n2)
    for x, y in zip(st1, st2):
        for k in ("systolic", "local", "pre_local", "checked", "visited_arcs", "transpose_arcs",
                  "modified"):
            assert x[k] == y[k], k


def test_progress_callback_and_caller_stream(hbm):
    """hb_set_progress reports every iteration as it ends; hb_set_stream runs
    the instance on a caller's stream with the same results, and None gives
    it a private stream again.  The caller's stream comes from the HIP runtime
    the library itself uses."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    g, gt = _web(hbm, 20000, seed=4)
    hb = hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(1).sum_of_distances(True).build()
    seen = []
    hb.set_progress(lambda st: seen.append(st))
    hb.run_until_done()
    assert [s["iteration"] for s in seen] == list(range(len(hb.iteration_stats())))
    assert [s["modified"] for s in seen] == [s["modified"] for s in hb.iteration_stats()]
    sd_private = hb.sum_of_distances()
    hb.set_progress(None)
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    hb.set_stream(stream.value)
    hb.run_until_done()
    assert hip.hipStreamSynchronize(stream) == 0
    assert np.array_equal(hb.sum_of_distances(), sd_private)
    hb.set_stream(None)
    assert hip.hipStreamDestroy(stream) == 0
    hb.run_until_done()
    assert np.array_equal(hb.sum_of_distances(), sd_private)
    assert len(seen) == len(hb.iteration_stats())  # no calls after None


def test_host_counter_memory_outputs(hbm):
    """The same run with HBM counter arrays, pinned host ones, and host ones
    streamed through device memory (several tiles, one tile) gives identical
    outputs, statistics included; an unknown memory kind is rejected."""
    g, gt = _web(hbm, 20000, seed=12)
    runs = []
    for where, tile in ((hbm.COUNTERS_DEVICE, 0), (hbm.COUNTERS_HOST, 0),
                        (hbm.COUNTERS_STREAM, 2048), (hbm.COUNTERS_STREAM, 0)):
        hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).jenkins(4).sum_of_distances(True)
              .sum_of_inverse_distances(True).counter_memory(where, tile).build())
        assert hb.params()["counter_memory"] == where
        hb.run_until_done()
        runs.append((hb.counters(0), hb.counters(1), hb.sum_of_distances(),
                     hb.harmonic_centralities(), hb.neighborhood_function(),
                     [(s["checked"], s["modified"], s["visited_arcs"], s["copied"])
                      for s in hb.iteration_stats()]))
        hb.close()
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a, b)
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).counter_memory(3).build()
    assert e.value.code == hbm.HB_E_INVALID_ARG
