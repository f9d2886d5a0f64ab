"""ctypes wrapper of the CPU oracle (oracle/ssw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- never by the product.

What each function restates (file:line in the reference):
  slim()          SlimCPUUntweaked.Compute/ComputeMany
                  (GPGPU/GPGPU/CPU/SlimCPUUntweaked.cs:19-222), n <= 16
  wide()          the same loop widened to uint32 masks, n <= 30
  witness()       CPU.ComputeOne back-tracking (GPGPU/GPGPU/CPU/CPU.cs:93-283)
  verify_word()   Verify.VerifyValidityOfSynchronizingWord
                  (GPGPU/GPGPU/Result veryfier/Verify.cs:12-53)
Independent pins: brute_force() (word enumeration) and pair_sync() (Cerny's
pair criterion).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libssw_oracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("oracle_slim_parallel", "oracle_wide_parallel"):
            f = getattr(L, name)
            f.restype = ctypes.c_int64
            f.argtypes = [ctypes.c_int, ctypes.c_int64, _i32p, _i32p, _i32p, _u8p, _i64p,
                          ctypes.c_int32, ctypes.c_int]
        L.oracle_witness.restype = ctypes.c_int
        L.oracle_witness.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _u8p, ctypes.c_int32]
        L.oracle_verify_word.restype = ctypes.c_int
        L.oracle_verify_word.argtypes = [ctypes.c_int, _i32p, _i32p, _u8p, ctypes.c_int32]
        L.oracle_brute_force.restype = ctypes.c_int
        L.oracle_brute_force.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int32, _i32p, _u8p]
        L.oracle_pair_sync.restype = ctypes.c_int
        L.oracle_pair_sync.argtypes = [ctypes.c_int, _i32p, _i32p]
        _lib = L
    return _lib


def _p(arr, t):
    return arr.ctypes.data_as(t)


def _as_i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _run(fn, n, A, B, bound, threads):
    A = _as_i32(A).reshape(-1, n)
    B = _as_i32(B).reshape(-1, n)
    count = A.shape[0]
    length = np.empty(count, np.int32)
    sync = np.empty(count, np.uint8)
    expanded = np.empty(count, np.int64)
    viol = fn(n, count, _p(A, _i32p), _p(B, _i32p), _p(length, _i32p), _p(sync, _u8p),
              _p(expanded, _i64p), bound, threads)
    if viol == -2:
        raise MemoryError("oracle: out of memory")
    return length, sync, expanded, int(viol)


def slim(n, A, B, bound=-1, threads=1):
    """Faithful SlimCPUUntweaked (uint16 masks, epoch stamps); n <= 16.

    Returns (length[-1 if not synchronizing], sync, expanded, violation)."""
    if not 1 <= n <= 16:
        raise ValueError("slim oracle needs 1 <= n <= 16")
    return _run(lib().oracle_slim_parallel, n, A, B, bound, threads)


def wide(n, A, B, bound=-1, threads=1):
    """The same search with uint32 masks; n <= 30."""
    if not 1 <= n <= 30:
        raise ValueError("wide oracle needs 1 <= n <= 30")
    return _run(lib().oracle_wide_parallel, n, A, B, bound, threads)


def solve(n, A, B, bound=-1, threads=1):
    return slim(n, A, B, bound, threads) if n <= 16 else wide(n, A, B, bound, threads)


def witness(n, a, b):
    """(L, word) with word a uint8 array (1 = letter b); L = -1 if not sync."""
    a = _as_i32(a)
    b = _as_i32(b)
    cap = max(1, (n - 1) * (n - 1) + 1)
    word = np.zeros(cap, np.uint8)
    L = np.zeros(1, np.int32)
    rc = lib().oracle_witness(n, _p(a, _i32p), _p(b, _i32p), _p(L, _i32p), _p(word, _u8p), cap)
    if rc != 0:
        raise RuntimeError(f"oracle_witness failed rc={rc}")
    L = int(L[0])
    return L, (word[:L].copy() if L >= 0 else np.zeros(0, np.uint8))


def verify_word(n, a, b, word) -> bool:
    a = _as_i32(a)
    b = _as_i32(b)
    w = np.ascontiguousarray(word, dtype=np.uint8)
    return bool(lib().oracle_verify_word(n, _p(a, _i32p), _p(b, _i32p), _p(w, _u8p), len(w)))


def brute_force(n, a, b, maxlen):
    a = _as_i32(a)
    b = _as_i32(b)
    word = np.zeros(max(1, maxlen), np.uint8)
    L = np.zeros(1, np.int32)
    lib().oracle_brute_force(n, _p(a, _i32p), _p(b, _i32p), maxlen, _p(L, _i32p), _p(word, _u8p))
    L = int(L[0])
    return L, (word[:L].copy() if L >= 0 else np.zeros(0, np.uint8))


def pair_sync(n, a, b) -> bool:
    a = _as_i32(a)
    b = _as_i32(b)
    return bool(lib().oracle_pair_sync(n, _p(a, _i32p), _p(b, _i32p)))
