"""The C-ABI library and the host-side plumbing -- CPU only (no compute calls).

* libssw.so and its bounds-checked build libssw_checked.so load and export
  every function declared in include/ssw.h;
* the ctypes structures match the C layout (sizeof/offsetof from gcc);
* generators: the counter-based uniform generator against a numpy twin,
  the two Cerny conventions (Problem.cs:13-29), the System.Random twin;
* the workload definitions (sizes, Cerny share, size distribution).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import ssw
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ssw.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(ssw_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("path", [ssw.LIB_PATH, ssw.CHECKED_LIB_PATH], ids=["libssw", "libssw_checked"])
def test_library_exports_every_declared_symbol(path):
    if path == ssw.CHECKED_LIB_PATH and not os.path.exists(path):
        pytest.skip("libssw_checked.so is not built (build.py --checked)")
    lib = ctypes.CDLL(path)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(ssw.EXPORTS) | {"ssw_copy_on_device"}


def test_abi_version():
    assert ssw.lib().ssw_abi_version() == 2


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "ssw.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(ssw_batch), offsetof(ssw_batch, delta), offsetof(ssw_batch, bound),
         sizeof(ssw_tuning), sizeof(ssw_result), offsetof(ssw_result, witness));
  printf("%zu %zu %zu\n", sizeof(ssw_server_config), sizeof(ssw_server_stats), sizeof(ssw_progress));
  return 0;
}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [ctypes.sizeof(ssw.Batch), ssw.Batch.delta.offset, ssw.Batch.bound.offset, ctypes.sizeof(ssw.Tuning),
            ctypes.sizeof(ssw.Result), ssw.Result.witness.offset, ctypes.sizeof(ssw.ServerConfig),
            ctypes.sizeof(ssw.ServerStats), ctypes.sizeof(ssw.Progress)]
    assert got == want


def test_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(ssw.SswError) as e:
        ssw.Context(0)
    assert e.value.code == -2


def _np_uniform(seed, n, count, first=0):
    key = W.splitmix64(np.uint64(seed))
    i = np.arange(first, first + count, dtype=np.uint64)[:, None, None]
    q = np.arange(n, dtype=np.uint64)[None, :, None]
    x = np.arange(2, dtype=np.uint64)[None, None, :]
    h = W.splitmix64(key ^ ((i << np.uint64(6)) | (q << np.uint64(1)) | x))
    v = ((h >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)
    return v[..., 0].astype(np.int32), v[..., 1].astype(np.int32)


@pytest.mark.parametrize("n", [1, 2, 10, 16, 30])
def test_uniform_generator_matches_numpy_twin(n):
    A, B = ssw.generate_uniform(1234, n, 500, first=77)
    A2, B2 = _np_uniform(1234, n, 500, first=77)
    np.testing.assert_array_equal(A, A2)
    np.testing.assert_array_equal(B, B2)
    assert A.min() >= 0 and A.max() < n


def test_uniform_generator_is_counter_based():
    A, B = ssw.generate_uniform(9, 12, 100)
    A2, B2 = ssw.generate_uniform(9, 12, 10, first=50)
    np.testing.assert_array_equal(A[50:60], A2)
    np.testing.assert_array_equal(B[50:60], B2)
    # roughly uniform
    A, _ = ssw.generate_uniform(5, 30, 20000)
    counts = np.bincount(A.ravel(), minlength=30)
    assert counts.min() > 0.9 * counts.mean() and counts.max() < 1.1 * counts.mean()


def test_cerny_conventions():
    a, b = ssw.generate_cerny(5, 0)   # Problem.GenerateWorstCase: A merges 0->1, B cycles
    assert list(a) == [1, 1, 2, 3, 4] and list(b) == [1, 2, 3, 4, 0]
    a, b = ssw.generate_cerny(5, 1)
    assert list(a) == [1, 2, 3, 4, 0] and list(b) == [1, 1, 2, 3, 4]
    with pytest.raises(ssw.SswError):
        ssw.generate_cerny(1, 0)       # GenerateWorstCase throws for n < 3; n = 2 is allowed here


def test_dotnet_generator_shape():
    # System.Random twin of FillArrayOfProblems (Problem.cs:32-58): deterministic,
    # in range; byte parity with .NET is unpinned (no .NET here, no golden vectors)
    A, B = ssw.generate_dotnet(1234567, 12, 50)
    A2, B2 = ssw.generate_dotnet(1234567, 12, 50)
    np.testing.assert_array_equal(A, A2)
    assert A.min() >= 0 and A.max() < 12 and B.min() >= 0 and B.max() < 12
    A3, _ = ssw.generate_dotnet(1234568, 12, 50)
    assert not np.array_equal(A, A3)


def test_workload_shapes():
    n, delta, pos = W.uniform_n16(count=20_000)
    assert delta.shape == (20_000, 32) and len(pos) == 200
    assert np.all(np.diff(pos // 100) == 1)  # one Cerny per block of 100
    ca, cb = ssw.generate_cerny(16, 1)
    assert np.all(delta[pos, :16] == ca) and np.all(delta[pos, 16:] == cb)
    sizes = W.uneven_sizes(1003, 200_000)
    counts = np.bincount(sizes, minlength=25)[12:]
    assert sizes.min() >= 12 and sizes.max() <= 24
    ratio = counts[:-2] / counts[2:]
    assert np.all(np.abs(ratio - 2.0) < 0.25)  # P(n) halves every two sizes
    npr, offs, d = W.uneven_12_24(count=5000)
    assert len(d) == int(2 * npr.astype(int).sum())
    reqs = W.serving(count=200)
    assert all(8 <= r[1] <= 30 for r in reqs) and sorted(r[0] for r in reqs) == [r[0] for r in reqs]
