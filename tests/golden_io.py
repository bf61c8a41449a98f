"""Readers for the committed fixtures under tests/golden/."""
import gzip
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cnr_golden(name: str) -> np.ndarray:
    """Java HyperBall results on cnr-2000 (log2m 8, Jenkins seed 42), raw
    big-endian f32 as in tests/hyperball_results/ of the reference
    (reader: tests/test_hyperball.rs:174-197)."""
    path = os.path.join(GOLDEN, "cnr-2000", f"cnr-2000_{name}.be-f32.gz")
    return np.frombuffer(gzip.open(path).read(), dtype=">f4").astype(np.float64)


def assert_array_equal_rel(expected, actual, threshold, name):
    """Relative comparison of tests/test_hyperball.rs:199-219: relative error
    where the expected value is non-zero, absolute otherwise."""
    expected = np.asarray(expected, dtype=np.float64)
    actual = np.asarray(actual, dtype=np.float64)
    assert expected.shape == actual.shape, name
    diff = np.abs(expected - actual)
    nz = expected != 0
    rel = np.where(nz, diff / np.where(nz, np.abs(expected), 1.0), diff)
    bad = np.nonzero(~(rel < threshold))[0]
    assert len(bad) == 0, (f"{name}: {len(bad)} elements off; first {bad[0]}: "
                           f"expected {expected[bad[0]]} actual {actual[bad[0]]} rel {rel[bad[0]]}")
    return float(rel.max()) if len(rel) else 0.0
