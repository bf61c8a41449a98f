"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so that a plausible mistake (wrong shift, dropped term,
transposed operand) fails it: closed forms of the reference's formulas,
published test vectors, brute force on tiny inputs, invariants, and the
reference's own golden vectors on cnr-2000.
"""
import math
import os

import numpy as np
import pytest

from golden_io import assert_array_equal_rel, cnr_golden


# ---- HyperLogLogBuilder statics (hyper_log_log_logic.rs:275-303) -----------

def test_register_size_is_five_for_every_n(orc):
    # max(5, ceil(ln((ln n / ln 2) / ln 2))): the parenthesisation keeps the
    # value at 5 below 2^64 (ln(64/ln 2) = 4.53), weight sums included.
    for n in [0, 1, 2, 16, 10**4, 10**5, 325557, 10**7, 10**8, 2**32, 2**40, 2**63, 2**64 - 1]:
        assert orc.register_size(n) == 5


def test_rel_std_table(orc):
    assert orc.rel_std(4) == pytest.approx(1.106 / 4)
    assert orc.rel_std(5) == pytest.approx(1.070 / math.sqrt(32))
    assert orc.rel_std(6) == pytest.approx(1.054 / 8)
    assert orc.rel_std(7) == pytest.approx(1.046 / math.sqrt(128))
    assert orc.rel_std(10) == pytest.approx(1.04 / 32)


def test_log2_num_registers(orc):
    # ceil(log2((1.106 / rsd)^2))
    assert orc.log2_num_registers(1.106 / 4) == 4
    assert orc.log2_num_registers(0.0325) == 11
    assert orc.log2_num_registers(0.1) == 7


def test_hll_parameters(orc):
    h = orc.Hll(8, 0, 325557)
    assert (h.m, h.r, h.W) == (256, 5, 20)
    assert h.h.sentinel == 1 << 30
    assert h.h.alpha_m_m == pytest.approx(0.7213 / (1 + 1.079 / 256) * 256 * 256, rel=1e-15)
    assert orc.Hll(4, 4).h.alpha_m_m == pytest.approx(0.673 * 256)
    assert orc.Hll(6, 6).h.sentinel == 1 << 62
    # masks: H has bit r-1 of every register, L bit 0
    msb = np.array(h.h.msb[:h.W], dtype=np.uint64)
    lsb = np.array(h.h.lsb[:h.W], dtype=np.uint64)
    bits_h = {w * 64 + b for w in range(h.W) for b in range(64) if (int(msb[w]) >> b) & 1}
    bits_l = {w * 64 + b for w in range(h.W) for b in range(64) if (int(lsb[w]) >> b) & 1}
    assert bits_h == {5 * i + 4 for i in range(256)}
    assert bits_l == {5 * i for i in range(256)}


def test_hll_builder_errors(orc):
    with pytest.raises(ValueError):
        orc.Hll(3, 5)  # log2m < 4 (hyper_log_log_logic.rs:379-383)
    with pytest.raises(ValueError):
        orc.Hll(5, 5)  # 160 bits is not a multiple of 64 (:399-403)


# ---- hashes ---------------------------------------------------------------

def test_siphash_published_vector(orc):
    # SipHash-2-4 test vector: key 00..0f, message 00..0e (Aumasson & Bernstein
    # 2012, appendix A).
    key = bytes(range(16))
    k0 = int.from_bytes(key[:8], "little")
    k1 = int.from_bytes(key[8:], "little")
    assert orc.siphash(bytes(range(15)), k0, k1, 2, 4) == 0xA129CA6149BE45E5
    # the empty-message vector of the reference implementation's table
    assert orc.siphash(b"", k0, k1, 2, 4) == 0x726FDB47DD0E0E31


def test_siphash_zero_key_u64_messages_match_cpython(orc):
    """The path the reference's DefaultHasher takes for a usize: zero keys,
    one 8-byte little-endian message, the length byte in the last block.
    CPython 3.10 hashes bytes with SipHash-2-4 under its hash secret, which
    PYTHONHASHSEED=0 sets to zero (hash -1 is reported as -2), so it is an
    independent implementation of SipHash-2-4 with zero keys; the 1-3 variant
    differs only in the round counts."""
    import subprocess
    import sys
    if sys.version_info[:2] != (3, 10) or sys.hash_info.algorithm != "siphash24":
        pytest.skip("needs CPython 3.10 with SipHash-2-4 bytes hashing")
    xs = [0, 1, 2, 7, 255, 256, 65535, 2**32 - 1, 2**32, 10**7, 2**63, 2**64 - 1,
          0x0123456789ABCDEF, 0xFEDCBA9876543210]
    code = ("xs = " + repr(xs) + "\n"
            "print(' '.join(str(hash(x.to_bytes(8, 'little')) & (2**64 - 1)) for x in xs))")
    env = dict(os.environ, PYTHONHASHSEED="0")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         check=True).stdout.split()
    for x, h in zip(xs, out):
        mine = orc.siphash(x.to_bytes(8, "little"), 0, 0, 2, 4)
        if mine == 2**64 - 1:  # CPython reports hash -1 as -2
            mine = 2**64 - 2
        assert int(h) == mine, hex(x)


def test_siphash13_is_the_round_parameterised_function(orc):
    for x in [0, 1, 42, 2**63 + 5]:
        assert orc.siphash13_u64(x) == orc.siphash(x.to_bytes(8, "little"), 0, 0, 1, 3)


def test_jenkins_structure(orc):
    # deterministic, seed-sensitive, and an actual mix (bits spread)
    a = [orc.jenkins(i, 42) for i in range(1000)]
    assert len(set(a)) == 1000
    assert orc.jenkins(5, 42) != orc.jenkins(5, 43)
    ones = sum(bin(x).count("1") for x in a) / 1000
    assert 28 < ones < 36


# ---- counter logic --------------------------------------------------------

@pytest.mark.parametrize("log2m,r", [(6, 5), (8, 5), (5, 6), (8, 6), (4, 4), (10, 5)])
def test_register_roundtrip(orc, log2m, r):
    h = orc.Hll(log2m, r)
    c = h.new_counter()
    rng = np.random.default_rng(log2m * 10 + r)
    vals = rng.integers(0, 1 << r, size=h.m)
    for i, v in enumerate(vals):
        h.set(c, i, int(v))
    assert np.array_equal(h.registers(c), vals.astype(np.uint64))
    # the packed words are exactly the LSB-first concatenation
    bits = 0
    for i, v in enumerate(vals):
        bits |= int(v) << (i * r)
    words = [(bits >> (64 * k)) & (2**64 - 1) for k in range(h.W)]
    assert [int(w) for w in c] == words


def test_add_brute_force(orc):
    h = orc.Hll(8, 5, hash=orc.HASH_JENKINS, seed=42)
    c = h.new_counter()
    expect = np.zeros(256, dtype=np.uint64)
    for item in range(5000):
        x = orc.jenkins(item, 42)
        j = x & 255
        w = (x >> 8) | (1 << 30)
        rho = (w & -w).bit_length() - 1
        expect[j] = max(int(expect[j]), rho + 1)
        h.add(c, item)
    assert np.array_equal(h.registers(c), expect)


@pytest.mark.parametrize("log2m,r", [(4, 4), (6, 5), (8, 5), (10, 5), (5, 6), (8, 6), (10, 6)])
def test_broadword_merge_is_register_max(orc, log2m, r):
    h = orc.Hll(log2m, r)
    rng = np.random.default_rng(1000 + log2m * 7 + r)
    for trial in range(60):
        x = h.new_counter()
        y = h.new_counter()
        hi = (1 << r) if trial % 3 else 3  # many equal registers in some trials
        xv = rng.integers(0, hi, size=h.m)
        yv = rng.integers(0, hi, size=h.m)
        for i in range(h.m):
            h.set(x, i, int(xv[i]))
            h.set(y, i, int(yv[i]))
        a = x.copy()
        b = x.copy()
        h.merge_broadword(a, y)
        h.merge_max(b, y)
        assert np.array_equal(a, b)
        assert np.array_equal(h.registers(a), np.maximum(xv, yv).astype(np.uint64))


def test_count_closed_forms(orc):
    h = orc.Hll(8, 5)
    c = h.new_counter()
    assert h.count(c) == 0.0  # zeros = m: m ln(m/m)
    h.set(c, 17, 3)
    assert h.count(c) == 256 * math.log(256 / 255)  # linear counting of a singleton
    # golden minimum of the nieminen array: (256 ln(256/255))^2 (F9)
    assert (256 * math.log(256 / 255)) ** 2 == pytest.approx(1.0039203, rel=1e-7)
    # no zeros: the harmonic estimate alpha m^2 / sum 2^-v
    for i in range(256):
        h.set(c, i, 1 + (i % 7))
    hm = sum(2.0 ** -(1 + (i % 7)) for i in range(256))
    assert h.count(c) == h.h.alpha_m_m / hm


def test_count_reference_i32_shift_for_31(orc):
    # hyper_log_log_logic.rs:178 computes 1.0 / (1 << value) as f64 with an
    # i32 literal: a register at 31 contributes -2^-31.
    h = orc.Hll(6, 5)
    c = h.new_counter()
    for i in range(64):
        h.set(c, i, 31 if i < 3 else 20)
    hm = 61 * 2.0 ** -20 - 3 * 2.0 ** -31
    assert h.count(c) == h.h.alpha_m_m / hm


def test_count_r6_exact(orc):
    h = orc.Hll(6, 6)
    c = h.new_counter()
    for i in range(64):
        h.set(c, i, 40 + (i % 24))
    hm = math.fsum(2.0 ** -(40 + (i % 24)) for i in range(64))
    assert h.count(c) == pytest.approx(h.h.alpha_m_m / hm, rel=1e-15)


def test_hll_estimate_accuracy(orc):
    # statistical check in the spirit of tests/test_hyperloglog_counter.rs:
    # most trials land within 2 relative standard deviations
    for log2m in [6, 8, 10]:
        ok = 0
        for seed in range(40):
            h = orc.Hll(log2m, 5, hash=orc.HASH_JENKINS, seed=seed)
            c = h.new_counter()
            for item in range(20000):
                h.add(c, item)
            if abs(h.count(c) - 20000) / 20000 < 2 * orc.rel_std(log2m):
                ok += 1
        assert ok >= 34, (log2m, ok)


# ---- engines --------------------------------------------------------------

def _random_graph(hbm, n, k, seed):
    return hbm.Graph.uniform(n, k, seed)


def test_jacobi_matches_brute_force_reachability(hbm, orc):
    # With m = 4096 registers and 12 nodes there are no collisions that matter:
    # each counter after t iterations is the union of the seeded singletons of
    # the nodes within distance t, so counter equality follows reachability.
    g = hbm.Graph.from_arcs(12, [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (6, 6), (7, 8), (8, 9),
                                 (9, 10), (10, 11)])
    hll = orc.Hll(12, 5, hash=orc.HASH_JENKINS, seed=1)
    J = orc.Jacobi(hll, g.offsets, g.successors)
    singles = J.curr.copy()
    for t in range(1, 8):
        J.iterate()
        for v in range(12):
            # brute-force ball of radius t
            ball = {v}
            for _ in range(t):
                ball |= {int(s) for u in ball for s in g.successors[g.offsets[u]:g.offsets[u + 1]]}
            expect = singles[v].copy()
            for u in ball:
                hll.merge_max(expect, singles[u])
            assert np.array_equal(J.curr[v], expect)


def test_reference_engine_equals_jacobi_every_iteration(hbm, orc):
    """The reference's white-box test (hyperball_impl.rs:1263-1322) on the
    restatement: parallel, systolic and local modes equal Jacobi."""
    g = hbm.Graph.web(5000, dmin=1, dmax=100, window=50, seed=4)
    gt = g.transpose()
    ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=6, threads=4)
    J = orc.Jacobi(orc.Hll(6, 0, 5000), g.offsets, g.successors)
    ref.init()
    modes = set()
    while True:
        ref.iterate()
        k = J.iterate()
        modes.add(ref.mode)
        assert np.array_equal(ref.counters(0), J.curr)
        assert np.array_equal(ref.counters(1), J.next)
        assert np.array_equal(ref.modified_bits(), J.modified)
        if k == 0:
            break
    assert any(m & 1 for m in modes) and any(m & 2 for m in modes)


def test_reference_engine_nf_is_sum_of_counts(hbm, orc):
    g = hbm.Graph.web(4000, dmin=1, dmax=60, window=40, seed=8)
    gt = g.transpose()
    ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=8, threads=3)
    J = orc.Jacobi(orc.Hll(8, 0, 4000), g.offsets, g.successors)
    ref.run()
    nf = ref.neighborhood_function()
    exact = [4000.0]
    while len(exact) < len(nf):
        J.iterate()
        exact.append(max(math.fsum(J.hll.counts(J.curr)), exact[-1]))
    np.testing.assert_allclose(nf, exact, rtol=1e-12)


@pytest.mark.slow
def test_oracle_reproduces_java_goldens(orc, cnr):
    """tests/test_hyperball.rs:221-302 of the reference, on the oracle."""
    g, gt = cnr
    ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=8,
                              hash=orc.HASH_JENKINS, seed=42, sum_of_distances=True,
                              sum_of_inverse_distances=True)
    ref.run()
    t = 1e-6
    assert_array_equal_rel(cnr_golden("sum_of_distances"), ref.sum_of_distances(), t, "sd")
    assert_array_equal_rel(cnr_golden("harmonic_centrality"), ref.harmonic_centralities(), t, "hc")
    assert_array_equal_rel(cnr_golden("lin_centrality"), ref.lin_centrality(), t, "lin")
    assert_array_equal_rel(cnr_golden("closeness_centrality"), ref.closeness_centrality(), t, "closeness")
    assert_array_equal_rel(cnr_golden("nieminen_centrality"), ref.nieminen_centrality(), t, "nieminen")


def test_golden_fixture_planted_error_is_caught(orc):
    """A wrong hash seed moves the oracle's outputs far past 1e-6 on a small
    graph: the golden comparison is sensitive to the Jenkins restatement."""
    import hyperball as hbm
    g = hbm.Graph.uniform(2000, 4, 3)
    gt = g.transpose()
    out = []
    for seed in (42, 43):
        ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=8,
                                  hash=orc.HASH_JENKINS, seed=seed, sum_of_distances=True)
        ref.run()
        out.append(ref.sum_of_distances())
    rel = np.abs(out[0] - out[1]) / np.maximum(out[0], 1)
    assert rel.max() > 1e-3
