"""GPU results against the committed golden fixtures (made by the oracle,
tests/golden/make_golden.py): lengths, expanded counts and witnesses bit-exact."""
import os

import numpy as np
import pytest

import ssw

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("mode", ["batch", "plain"])
def test_plain_n10_fixture(ctx, mode):
    g = np.load(os.path.join(GOLDEN, "plain_n10.npz"))
    A, B = g["a"].astype(np.int32), g["b"].astype(np.int32)
    if mode == "plain":
        A, B = A[:1500], B[:1500]
    r = ctx.solve_uniform(10, A, B, witness=True, mode=mode)
    k = len(A)
    np.testing.assert_array_equal(r.length, g["length"][:k])
    np.testing.assert_array_equal(r.expanded, g["expanded"][:k])
    np.testing.assert_array_equal(r.sync, (g["length"][:k] >= 0).astype(np.uint8))
    stride = g["witness"].shape[1]
    np.testing.assert_array_equal(r.witness[:, :stride], g["witness"][:k])


def test_uniform_n16_fixture(ctx):
    g = np.load(os.path.join(GOLDEN, "uniform_n16_slice.npz"))
    A, B = g["a"].astype(np.int32), g["b"].astype(np.int32)
    r = ctx.compute(16, A, B)
    np.testing.assert_array_equal(r.length, g["length"])
    np.testing.assert_array_equal(r.expanded, g["expanded"])
    rw = ctx.compute(16, A[:500], B[:500], witness=True)
    np.testing.assert_array_equal(rw.length, g["length"][:500])
    np.testing.assert_array_equal(rw.witness[:, :g["witness500"].shape[1]], g["witness500"])


def test_large_path_on_fixture(ctx):
    g = np.load(os.path.join(GOLDEN, "uniform_n16_slice.npz"))
    idx = np.r_[0:40, g["cerny_positions"][:4]]
    A, B = g["a"][idx].astype(np.int32), g["b"][idx].astype(np.int32)
    r = ctx.solve_uniform(16, A, B, mode="large")
    np.testing.assert_array_equal(r.length, g["length"][idx])
    np.testing.assert_array_equal(r.expanded, g["expanded"][idx])
