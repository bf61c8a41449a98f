"""hb_get_counters_noalloc against hb_get_counters through the C ABI: the
variant that stages through the library's per-GPU device array (no device
allocation) returns the same words, on consecutive calls with a large and
then a smaller instance (the array keeps the previous call's contents), for
both counter buffers, before and after iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instance(hbm, n, log2m, r, seed):
    g = hbm.Graph.web(n, 2.1, 3, 1000, 500, 0.7, seed)
    return (hbm.HyperBallBuilder.with_hyper_log_log(g, g.transpose(), log2m).register_bits(r)
            .jenkins(seed).build())


def test_noalloc_counters_match(hbm):
    # 40,000 counters of 768 bytes: 30.7 MB, more than the 16 MiB staging
    # array, so the copy runs in two chunks; then 3,000 counters of 160 bytes
    big = _instance(hbm, 40_000, 10, 6, 1)
    small = _instance(hbm, 3_000, 8, 5, 2)
    for hb in (big, small):
        hb.init()
        assert np.array_equal(hb.counters_noalloc(0), hb.counters(0))
        assert not hb.counters_noalloc(1).any()
        hb.iterate()
        hb.iterate()
        for which in (0, 1):
            a = hb.counters_noalloc(which)
            assert a.any()
            assert np.array_equal(a, hb.counters(which))
    # the large instance again after the small one overwrote the array's start
    assert np.array_equal(big.counters_noalloc(0), big.counters(0))
    big.run_until_done()
    small.run_until_done()
    assert np.array_equal(small.counters_noalloc(0), small.counters(0))
    assert np.array_equal(big.counters_noalloc(0), big.counters(0))
    big.close()
    small.close()


def test_noalloc_counters_errors(hbm):
    hb = _instance(hbm, 1000, 8, 5, 3)
    with pytest.raises(hbm.HyperBallError) as e:
        hb.counters_noalloc(0)  # before init
    assert e.value.code == hbm.HB_E_NOT_RUN
    hb.init()
    out = np.zeros(10, dtype=np.uint64)
    with pytest.raises(hbm.HyperBallError) as e:
        hbm._check(hbm.lib().hb_get_counters_noalloc(hb._h, 0, hbm._ptr(out), out.size))
    assert e.value.code == hbm.HB_E_INVALID_ARG and "too small" in e.value.message
    hb.close()
