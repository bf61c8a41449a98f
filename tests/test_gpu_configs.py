"""GPU parity of the benchmark configurations against the oracle.

- Registers of value 31 and above: Jenkins seeds found by exhaustive search
  make one node seed a register of 31 (5-bit) or 32-33 (6-bit), so the
  reference's i32 quirk of count() (a 31 register adds -2^-31,
  hyper_log_log_logic.rs:178) and the high register classes of the device
  count run in a lockstep test.
- Batched instances (the batched configuration, scaled down): every instance
  of a mixed batch against the oracle's restatement of the reference engine
  and against the sequential Jacobi oracle (hyperball_impl.rs:1230-1261,
  :1302-1319).
- Accumulator precision at log2m 12 (FP32 Kahan and plain FP32 against the
  f64 oracle, hyperball_impl.rs:1072-1093).
- The large-instance shape (log2m 6, 5-bit registers, +-10,000 window,
  FP32 Kahan accumulators) against the oracle at small n, and at full size
  (n = 10^8, 2.1e9 arcs) through properties that hold at any size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle(orc, g, gt, log2m, reg_bits, seed, **kw):
    return orc.OracleHyperBall(g.offsets, g.successors,
                               None if gt is None else gt.offsets,
                               None if gt is None else gt.successors,
                               log2m=log2m, reg_bits=reg_bits, hash=orc.HASH_JENKINS, seed=seed,
                               **kw)


def _registers(counters, m, r):
    """Register values of counters in the reference layout (register j at bit
    j * r, LSB first, across little-endian u64 words)."""
    c = np.ascontiguousarray(counters, dtype="<u8")
    bits = np.unpackbits(c.view(np.uint8), axis=1, bitorder="little")
    bits = bits[:, :m * r].reshape(len(c), m, r).astype(np.uint32)
    return bits @ (np.uint32(1) << np.arange(r, dtype=np.uint32))


def _rel_err(got, exp):
    err = np.abs(np.asarray(got, np.float64) - exp)
    nz = exp != 0
    err[nz] /= np.abs(exp[nz])
    return err


# (log2m, register bits, Jenkins seed, node, register value the node seeds):
# jenkins(node, seed) >> log2m has at least 30 trailing zeros.
HIGH_REGISTERS = [
    (8, 5, 668684, 1403, 31),
    (8, 6, 668684, 1403, 32),
    (7, 6, 668684, 1403, 33),
    (9, 5, 668684, 1403, 31),
    (6, 5, 246070, 303, 31),
    (10, 5, 400129, 822, 31),
    (12, 5, 985052, 423, 31),
]


@pytest.mark.parametrize("log2m,reg_bits,seed,node,value", HIGH_REGISTERS)
def test_high_registers_bit_exact(hbm, orc, log2m, reg_bits, seed, node, value):
    """A node seeds a register of `value` (checked through the oracle); the
    value spreads to every node that reaches it.  Registers and modified bits
    are bit-exact every iteration, the f64 outputs (which go through count()
    and so through the quirk for a 31 register) bit-exact at the end."""
    g = hbm.Graph.web(2000, 2.1, 3, 200, 100, 0.7, 17)
    gt = g.transpose()
    hll = orc.Hll(log2m, reg_bits, g.num_nodes, orc.HASH_JENKINS, seed)
    J = orc.Jacobi(hll, g.offsets, g.successors)
    assert hll.registers(J.curr[node]).max() == value
    assert np.array_equal(_registers(J.curr[node:node + 1], 1 << log2m, reg_bits)[0],
                          hll.registers(J.curr[node]))
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, log2m).register_bits(reg_bits)
          .jenkins(seed).sum_of_distances(True).sum_of_inverse_distances(True).build())
    hb.init()
    assert np.array_equal(hb.counters(0), J.curr)
    while True:
        hb.iterate()
        k = J.iterate()
        assert np.array_equal(hb.counters(0), J.curr)
        assert np.array_equal(hb.modified_bits(), J.modified)
        if k == 0:
            break
    reached = int((_registers(J.curr, 1 << log2m, reg_bits).max(axis=1) == value).sum())
    assert reached > 1, "the high register reached no other node"
    ref = _oracle(orc, g, gt, log2m, reg_bits, seed, sum_of_distances=True,
                  sum_of_inverse_distances=True)
    ref.run()
    assert np.array_equal(hb.counters(0), ref.counters(0))
    for got, exp in ((hb.sum_of_distances(), ref.sum_of_distances()),
                     (hb.harmonic_centralities(), ref.harmonic_centralities()),
                     (hb.reachable_nodes(), ref.reachable_nodes()),
                     (hb.lin_centrality(), ref.lin_centrality()),
                     (hb.nieminen_centrality(), ref.nieminen_centrality())):
        assert np.array_equal(got, exp)
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(),
                               rtol=1e-12, atol=0)
    print(f"log2m {log2m} r {reg_bits}: register {value} reached {reached} nodes")


def _batch_mix(hbm, k, seed0, nmin, nmax, shapes_seed):
    """k instances from the web generator, sizes log-uniform on [nmin, nmax],
    log2m drawn from {6, 8, 10}, 5- and 6-bit registers, a transpose on three
    of four, long rows (up to 3000 successors) on every third."""
    rng = np.random.default_rng(shapes_seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        dmax = 3000 if i % 3 == 0 else 10000
        g = hbm.Graph.web(n, 2.1, 3 if i % 2 else 1, dmax, 1000, 0.7, seed0 + i)
        gt = g.transpose() if i % 4 != 3 else None
        log2m = int(rng.choice([6, 8, 10]))
        r = 5 if rng.uniform() < 0.5 else 6
        out.append((g, gt, log2m, r, 42 + i))
    return out


def _builder(hbm, g, gt, log2m, r, seed):
    return (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, log2m).register_bits(r).jenkins(seed)
            .sum_of_distances(True).sum_of_inverse_distances(True))


def _check_instance(orc, hb, g, gt, log2m, r, seed, ub=None):
    ref = _oracle(orc, g, gt, log2m, r, seed, sum_of_distances=True,
                  sum_of_inverse_distances=True)
    if ub is None:
        ref.run()
    else:
        ref.run(ub)
    assert np.array_equal(hb.counters(0), ref.counters(0))
    assert np.array_equal(hb.counters(1), ref.counters(1))
    assert np.array_equal(hb.modified_bits(), ref.modified_bits())
    assert np.array_equal(hb.sum_of_distances(), ref.sum_of_distances())
    assert np.array_equal(hb.harmonic_centralities(), ref.harmonic_centralities())
    assert np.array_equal(hb.reachable_nodes(), ref.reachable_nodes())
    nf, nf_ref = hb.neighborhood_function(), ref.neighborhood_function()
    assert len(nf) == len(nf_ref)
    np.testing.assert_allclose(nf, nf_ref, rtol=1e-12, atol=0)


def test_batch_instances_match_oracle(hbm, orc):
    """Every instance of a 16-instance mixed batch (log2m 6/8/10, r 5/6,
    with and without transpose, heavy rows above 64 and 128 successors)
    equals the oracle's engine after 1, 2 and 3 iterations and at the end,
    and its final registers equal the Jacobi fixed point."""
    mix = _batch_mix(hbm, 16, 500, 1000, 50_000, 8)
    batch = hbm.HyperBallBatch([_builder(hbm, *x) for x in mix])
    for j, hb in enumerate(batch.instances):
        hb.set_tuning([64, 128, 512, 4096][j % 4])
    for ub in (1, 2, 3):
        batch.run(ub)
        for hb, x in zip(batch.instances, mix):
            _check_instance(orc, hb, *x, ub=ub)
    batch.run_until_done()
    for hb, (g, gt, log2m, r, seed) in zip(batch.instances, mix):
        _check_instance(orc, hb, g, gt, log2m, r, seed)
        J = orc.Jacobi(orc.Hll(log2m, r, g.num_nodes, orc.HASH_JENKINS, seed), g.offsets,
                       g.successors)
        while J.iterate():
            pass
        assert np.array_equal(hb.counters(0), J.curr)
    batch.close()


def test_batched_configuration_scaled_matches_oracle(hbm, orc):
    """The batched configuration scaled down: 64 instances of the web-scale
    generator (Zipf(2.1) out-degree on [3, 10000], +-1000 window, 70%
    window / 30% preferential attachment, own seed each), sizes log-uniform
    on [2000, 60000], log2m drawn from {6, 8, 10}, 6-bit registers, graph
    replay and the batch autotuner; every instance against the oracle."""
    rng = np.random.default_rng(3)
    ns = np.exp(rng.uniform(np.log(2000), np.log(60000), 64)).astype(np.int64)
    log2ms = rng.choice([6, 8, 10], 64)
    mix = []
    for i, n in enumerate(ns):
        g = hbm.Graph.web(int(n), 2.1, 3, 10000, 1000, 0.7, 1000 + i)
        mix.append((g, g.transpose(), int(log2ms[i]), 6, 42 + i))
    batch = hbm.HyperBallBatch([_builder(hbm, *x) for x in mix])
    batch.autotune(reps=1)
    batch.run_until_done()
    for hb, x in zip(batch.instances, mix):
        _check_instance(orc, hb, *x)
    batch.close()


@pytest.mark.parametrize("acc,bound", [("f32_kahan", 1e-6), ("f32", 2e-6)])
def test_precision_log2m12(hbm, orc, acc, bound):
    """The precision configuration at n = 3*10^4: the web-scale generator,
    log2m 12, 5-bit registers, FP32 Kahan-compensated or plain FP32 per-node
    accumulators against the oracle's f64 (max relative error per output;
    absolute where the f64 value is 0).  Registers stay exact integers and the
    reachable counts come from integer register sums: both bit-exact."""
    g = hbm.Graph.web(30_000, 2.1, 3, 10000, 1000, 0.7, 2)
    gt = g.transpose()
    mode = hbm.ACC_F32_KAHAN if acc == "f32_kahan" else hbm.ACC_F32
    hb = (_builder(hbm, g, gt, 12, 5, 42).accumulator_precision(mode).build())
    hb.run_until_done()
    ref = _oracle(orc, g, gt, 12, 5, 42, sum_of_distances=True, sum_of_inverse_distances=True)
    ref.run()
    assert np.array_equal(hb.counters(0), ref.counters(0))
    assert np.array_equal(hb.reachable_nodes(), ref.reachable_nodes())
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(),
                               rtol=1e-12, atol=0)
    for name, got, exp in (("sum_of_distances", hb.sum_of_distances(), ref.sum_of_distances()),
                           ("harmonic", hb.harmonic_centralities(), ref.harmonic_centralities()),
                           ("lin", hb.lin_centrality(), ref.lin_centrality()),
                           ("closeness", hb.closeness_centrality(), ref.closeness_centrality())):
        err = _rel_err(got, exp)
        print(f"{acc} {name}: max {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}")
        assert err.max() <= bound, (name, err.max())


def test_large_shape_small_n(hbm, orc):
    """The large-instance shape at n = 5*10^4: out-degree Zipf(2.1) on
    [4, 10000], +-10,000 window, log2m 6, 5-bit registers, FP32 Kahan
    accumulators.  Registers and modified bits bit-exact every iteration
    against Jacobi; accumulators within 1e-6 of the f64 oracle."""
    g = hbm.Graph.web(50_000, 2.1, 4, 10000, 10000, 0.7, 5)
    gt = g.transpose()
    hb = (_builder(hbm, g, gt, 6, 5, 42).accumulator_precision(hbm.ACC_F32_KAHAN).build())
    J = orc.Jacobi(orc.Hll(6, 5, g.num_nodes, orc.HASH_JENKINS, 42), g.offsets, g.successors)
    hb.init()
    while True:
        hb.iterate()
        k = J.iterate()
        assert np.array_equal(hb.counters(0), J.curr)
        assert np.array_equal(hb.modified_bits(), J.modified)
        if k == 0:
            break
    assert any(s["systolic"] for s in hb.iteration_stats())
    ref = _oracle(orc, g, gt, 6, 5, 42, sum_of_distances=True, sum_of_inverse_distances=True)
    ref.run()
    assert np.array_equal(hb.reachable_nodes(), ref.reachable_nodes())
    np.testing.assert_allclose(hb.neighborhood_function(), ref.neighborhood_function(),
                               rtol=1e-12, atol=0)
    for got, exp in ((hb.sum_of_distances(), ref.sum_of_distances()),
                     (hb.harmonic_centralities(), ref.harmonic_centralities())):
        assert _rel_err(got, exp).max() <= 1e-6


@pytest.mark.slow
def test_large_instance_full_size_properties(hbm, orc):
    """The large instance at full size: n = 10^8, about 2.1e9 arcs, transpose
    built on the device, log2m 6, 5-bit registers, FP32 Kahan accumulators.
    Iteration 0 equals the oracle's seeding on sampled nodes.  At the fixed
    point every counter dominates its successors' (sampled),
    one more iteration modifies nothing, the neighbourhood function starts
    at n and never decreases, reachable counts equal a fresh count() of the
    counters, and a second run reproduces every value bit for bit."""
    n = 100_000_000
    g = hbm.Graph.web(n, 2.1, 4, 10000, 10000, 0.7, 5)
    print(f"graph n={g.num_nodes} arcs={g.num_arcs}", flush=True)
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, None, 6).transpose_on_device()
          .register_bits(5).jenkins(42).sum_of_distances(True).sum_of_inverse_distances(True)
          .accumulator_precision(hbm.ACC_F32_KAHAN).build())
    hll = orc.Hll(6, 5, n, orc.HASH_JENKINS, 42)
    rng = np.random.default_rng(1)
    # Iteration 0 above 2^25 nodes hashes the successor ids again (no array of
    # seeded registers): sampled counters equal the oracle's seeding of the
    # node and its successors.
    hb.init()
    hb.iterate()
    cnt = hb.counters(0)
    for v in rng.choice(n, 300, replace=False):
        want = hll.new_counter()
        for item in [v, *g.successors[g.offsets[v]:g.offsets[v + 1]]]:
            hll.add(want, int(item))
        assert np.array_equal(cnt[v], want), f"iteration 0: counter {v}"
    del cnt
    hb.run_until_done()
    nf = hb.neighborhood_function()
    sd, hc = hb.sum_of_distances(), hb.harmonic_centralities()
    assert nf[0] == n and np.all(np.diff(nf) >= 0)
    assert np.all(sd >= 0) and np.all(hc >= 0)
    hb.iterate()
    assert hb.iteration_stats()[-1]["modified"] == 0
    cnt = hb.counters(0)
    for v in rng.choice(n, 300, replace=False):
        x = cnt[v].copy()
        for s in g.successors[g.offsets[v]:g.offsets[v + 1]][:50]:
            y = x.copy()
            hll.merge_max(y, cnt[s])
            assert np.array_equal(x, y), f"counter {v} does not dominate successor {s}"
    reach = hb.reachable_nodes()
    for v in rng.choice(n, 300, replace=False):
        assert reach[v] == hll.count(cnt[v])
    del cnt
    hb.run_until_done()
    assert np.array_equal(hb.neighborhood_function(), nf)
    assert np.array_equal(hb.sum_of_distances(), sd)
    assert np.array_equal(hb.harmonic_centralities(), hc)
    hb.close()
