"""Batched execution: every instance of a batch must produce exactly what the
same configuration produces alone (registers, modified bits, accumulators and
the neighbourhood function, whose fixed-point sums do not depend on the
schedule)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _builders(hbm, k, seed0=100, with_transpose=True):
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(k):
        n = int(10 ** rng.uniform(3, 4.7))
        dmax = 3000 if i % 3 == 0 else 300  # rows above the heavy threshold
        g = hbm.Graph.web(n, 2.1, 1, dmax, 100, 0.7, seed0 + i)
        gt = g.transpose() if (with_transpose and i % 4 != 3) else None
        log2m = [6, 8, 10][i % 3]
        r = 5 if i % 2 == 0 else 6
        b = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, log2m).register_bits(r).jenkins(i)
             .sum_of_distances(True).sum_of_inverse_distances(i % 2 == 0))
        out.append(b)
    return out


def _compare(hbm, b, hb_batch):
    single = b.build()
    single.run_until_done()
    assert np.array_equal(hb_batch.counters(0), single.counters(0))
    assert np.array_equal(hb_batch.counters(1), single.counters(1))
    assert np.array_equal(hb_batch.modified_bits(), single.modified_bits())
    assert np.array_equal(hb_batch.neighborhood_function(), single.neighborhood_function())
    assert np.array_equal(hb_batch.sum_of_distances(), single.sum_of_distances())
    assert np.array_equal(hb_batch.reachable_nodes(), single.reachable_nodes())
    if b._hc:
        assert np.array_equal(hb_batch.harmonic_centralities(), single.harmonic_centralities())
    s1, s2 = hb_batch.iteration_stats(), single.iteration_stats()
    assert len(s1) == len(s2)
    for x, y in zip(s1, s2):
        for key in ("systolic", "local", "pre_local", "checked", "visited_arcs", "scanned_arcs",
                    "transpose_arcs", "copied", "modified", "gathered_slabs", "delta_gathering",
                    "nf"):
            assert x[key] == y[key], key


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_runs(hbm, graph):
    builders = _builders(hbm, 12)
    batch = hbm.HyperBallBatch(builders)
    batch.set_graph(graph)
    batch.run_until_done()
    assert len(batch.step_times()) >= max(len(h.iteration_stats()) for h in batch.instances)
    for b, hb in zip(builders, batch.instances):
        _compare(hbm, b, hb)
    batch.close()


def test_batch_heavy_threshold_per_instance(hbm):
    """Instances keep their own heavy-row threshold inside a batch; results
    do not depend on it."""
    builders = _builders(hbm, 6, seed0=31)
    batch = hbm.HyperBallBatch(builders)
    for j, hb in enumerate(batch.instances):
        hb.set_tuning([64, 128, 4096][j % 3])
    batch.run_until_done()
    for b, hb in zip(builders, batch.instances):
        _compare(hbm, b, hb)


def test_batch_rerun_and_stop_rules(hbm):
    builders = _builders(hbm, 6, seed0=7)
    batch = hbm.HyperBallBatch(builders)
    batch.run(3)
    for b, hb in zip(builders, batch.instances):
        single = b.build()
        single.run(3)
        assert np.array_equal(hb.neighborhood_function(), single.neighborhood_function())
        assert np.array_equal(hb.counters(0), single.counters(0))
    batch.run(100, 0.01)
    for b, hb in zip(builders, batch.instances):
        single = b.build()
        single.run(100, 0.01)
        assert np.array_equal(hb.neighborhood_function(), single.neighborhood_function())
    batch.run_until_done()
    for b, hb in zip(builders, batch.instances):
        _compare(hbm, b, hb)


def test_batch_instance_is_borrowed(hbm):
    builders = _builders(hbm, 2)
    batch = hbm.HyperBallBatch(builders)
    with pytest.raises(hbm.HyperBallError):
        hbm._check(hbm.lib().hb_destroy(batch.instances[0]._h))


def test_batch_rejects_unsupported_shape(hbm):
    g = hbm.Graph.uniform(100, 3, 1)
    with pytest.raises(hbm.HyperBallError) as e:
        hbm.HyperBallBatch([hbm.HyperBallBuilder.with_hyper_log_log(g, None, 7)])
    assert "log2m 6, 8 or 10" in e.value.message


def test_batch_autotune_keeps_results(hbm):
    builders = _builders(hbm, 4, seed0=55)
    batch = hbm.HyperBallBatch(builders)
    chosen, table = batch.autotune()
    assert chosen in [t for t, _ in table] and table[-1][0] == 2**32 - 1
    batch.run_until_done()
    for b, hb in zip(builders, batch.instances):
        _compare(hbm, b, hb)


def test_batch_degenerate_instances(hbm):
    """Instances with one node, no arcs, self loops only, a path and a clique
    run in one batch beside a normal graph, each equal to its own run."""
    n = 40
    graphs = [hbm.Graph.from_arcs(1, []),
              hbm.Graph.from_arcs(10, []),
              hbm.Graph.from_arcs(10, [(i, i) for i in range(10)]),
              hbm.Graph.from_arcs(n, [(i, i + 1) for i in range(n - 1)]),
              hbm.Graph.from_arcs(12, [(i, j) for i in range(12) for j in range(12)]),
              hbm.Graph.from_arcs(n, [(i, (i + 1) % n) for i in range(n)]),
              hbm.Graph.web(5000, 2.1, 1, 300, 100, 0.7, 3)]
    builders = []
    for i, g in enumerate(graphs):
        gt = g.transpose() if i != 5 else None
        builders.append(hbm.HyperBallBuilder.with_hyper_log_log(g, gt, (6, 8, 10)[i % 3])
                        .register_bits(5 + i % 2).jenkins(i).sum_of_distances(True)
                        .sum_of_inverse_distances(True))
    batch = hbm.HyperBallBatch(builders)
    batch.run_until_done()
    for b, hb in zip(builders, batch.instances):
        _compare(hbm, b, hb)
