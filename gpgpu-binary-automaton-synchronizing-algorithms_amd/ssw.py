"""Python binding of libssw.so (ctypes over the C ABI in include/ssw.h).

Used by the tests and bench.py.  Everything here is plumbing: every solve runs
in the HIP kernels of libssw.so, and loading fails loudly when the library is
missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libssw.so")

SSW_OK = 0
SSW_PENDING = 1
SSW_ERR_ARG = -1
SSW_ERR_DEVICE = -2
SSW_ERR_NOMEM = -3
SSW_ERR_BUDGET = -4
SSW_ERR_STATE = -5
SSW_ERR_INTERNAL = -6
SSW_ERR_TIMEOUT = -7
SSW_ERR_NOTFOUND = -8
SSW_ERR_CANCELLED = -9
SSW_HOST = 0
SSW_DEVICE = 1
LEN_NOT_SYNC = -1
LEN_INVALID = -3

EXPORTS = [
    "ssw_abi_version", "ssw_create", "ssw_destroy", "ssw_last_error", "ssw_compute", "ssw_verify",
    "ssw_best_parallelism", "ssw_solve_batch", "ssw_solve_plain", "ssw_solve_large",
    "ssw_get_tuning", "ssw_set_tuning", "ssw_autotune", "ssw_autotune_batch", "ssw_load_tuning",
    "ssw_get_progress",
    "ssw_set_progress_callback", "ssw_device_alloc", "ssw_device_free", "ssw_copy_to_device",
    "ssw_copy_to_host", "ssw_copy_on_device", "ssw_synchronize", "ssw_last_kernel_time", "ssw_generate_uniform",
    "ssw_generate_cerny", "ssw_generate_dotnet", "ssw_server_create", "ssw_server_submit",
    "ssw_server_poll", "ssw_server_wait", "ssw_server_stats_get", "ssw_server_destroy",
    "ssw_measure_atomics", "ssw_solve_batch_noalloc", "ssw_last_launch_stats", "ssw_set_verify_prefilter",
    "ssw_server_last_error", "ssw_server_cancel",
]


class SswError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ssw error {code}: {msg}")
        self.code = code


class Batch(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int64),
        ("uniform_n", ctypes.c_int32),
        ("witness_stride", ctypes.c_int32),
        ("n_per", ctypes.c_void_p),
        ("offset", ctypes.c_void_p),
        ("delta", ctypes.c_void_p),
        ("length", ctypes.c_void_p),
        ("sync", ctypes.c_void_p),
        ("expanded", ctypes.c_void_p),
        ("witness", ctypes.c_void_p),
        ("bound", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("group_lanes", ctypes.c_int32), ("block_threads", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("arena_lds", ctypes.c_int32), ("lane_entries", ctypes.c_int32),
                ("spill_arena_lds", ctypes.c_int32), ("spill_lane_entries", ctypes.c_int32),
                ("spill_group_lanes", ctypes.c_int32), ("narrow_bits", ctypes.c_int32)]


class Progress(ctypes.Structure):
    _fields_ = [("automata_done", ctypes.c_uint64), ("subsets_expanded", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("deferred", ctypes.c_uint64), ("in_flight", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("mem_budget", ctypes.c_uint64), ("small_n_max", ctypes.c_int32),
                ("levels_per_tick", ctypes.c_int32), ("max_large", ctypes.c_int32),
                ("want_witness", ctypes.c_int32), ("small_workers", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("length", ctypes.c_int32), ("n", ctypes.c_int32), ("expanded", ctypes.c_int64),
                ("sync", ctypes.c_uint8), ("witness_valid", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 6),
                ("latency_us", ctypes.c_double), ("witness", ctypes.c_uint8 * 112)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("automata_per_s", ctypes.c_double),
                ("subsets_expanded", ctypes.c_uint64)]


class LaunchStat(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("variant", ctypes.c_int32), ("group_lanes", ctypes.c_int32),
                ("block_threads", ctypes.c_int32), ("grid", ctypes.c_int32), ("lds_bytes", ctypes.c_int32),
                ("ring_entries", ctypes.c_int32), ("lane_entries", ctypes.c_int32), ("bitmap_bits", ctypes.c_int32),
                ("pass_", ctypes.c_int32), ("ms", ctypes.c_double), ("automata", ctypes.c_uint64),
                ("subsets", ctypes.c_uint64), ("expanded", ctypes.c_uint64), ("kernel", ctypes.c_char * 64)]


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(Progress), ctypes.c_void_p)

CHECKED_LIB_PATH = os.path.join(HERE, "libssw_checked.so")

_libs = {}


def lib(path=None):
    """Loads libssw.so (or another build of it, e.g. the bounds-checked
    libssw_checked.so); raises when it has not been built."""
    path = path or LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(path)
        vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
        sig = {
            "ssw_abi_version": ([], ctypes.c_int),
            "ssw_create": ([ctypes.c_int, u64, ctypes.c_uint32, ctypes.POINTER(vp)], ctypes.c_int),
            "ssw_destroy": ([vp], ctypes.c_int),
            "ssw_last_error": ([vp], ctypes.c_char_p),
            "ssw_compute": ([vp, i32, i64, vp, vp, vp, vp, vp, i32, vp, i32, vp], ctypes.c_int),
            "ssw_verify": ([vp, i32, i64, vp, vp, i32, vp], ctypes.c_int),
            "ssw_best_parallelism": ([vp], ctypes.c_int),
            "ssw_set_verify_prefilter": ([vp, i32], ctypes.c_int),
            "ssw_solve_batch": ([vp, ctypes.POINTER(Batch), i32, vp], ctypes.c_int),
            "ssw_solve_batch_noalloc": ([vp, ctypes.POINTER(Batch), vp], ctypes.c_int),
            "ssw_solve_plain": ([vp, ctypes.POINTER(Batch), i32, vp], ctypes.c_int),
            "ssw_solve_large": ([vp, ctypes.POINTER(Batch), i32, vp], ctypes.c_int),
            "ssw_get_tuning": ([vp, i32, ctypes.POINTER(Tuning)], ctypes.c_int),
            "ssw_set_tuning": ([vp, i32, ctypes.POINTER(Tuning)], ctypes.c_int),
            "ssw_autotune": ([vp, i32, i32, i64, u64, ctypes.c_char_p], ctypes.c_int),
            "ssw_load_tuning": ([vp, ctypes.c_char_p], ctypes.c_int),
            "ssw_autotune_batch": ([vp, ctypes.POINTER(Batch), i32], ctypes.c_int),
            "ssw_get_progress": ([vp, ctypes.POINTER(Progress)], ctypes.c_int),
            "ssw_set_progress_callback": ([vp, PROGRESS_FN, vp, i32], ctypes.c_int),
            "ssw_device_alloc": ([vp, u64, ctypes.POINTER(vp)], ctypes.c_int),
            "ssw_device_free": ([vp, vp], ctypes.c_int),
            "ssw_copy_to_device": ([vp, vp, vp, u64], ctypes.c_int),
            "ssw_copy_to_host": ([vp, vp, vp, u64], ctypes.c_int),
            "ssw_copy_on_device": ([vp, vp, vp, u64], ctypes.c_int),
            "ssw_synchronize": ([vp], ctypes.c_int),
            "ssw_last_kernel_time": ([vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)], ctypes.c_int),
            "ssw_generate_uniform": ([u64, i32, i64, i64, vp, vp], ctypes.c_int),
            "ssw_generate_cerny": ([i32, i32, vp, vp], ctypes.c_int),
            "ssw_generate_dotnet": ([i32, i32, i64, vp, vp], ctypes.c_int),
            "ssw_server_create": ([vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)], ctypes.c_int),
            "ssw_server_submit": ([vp, i32, vp, vp, ctypes.POINTER(u64)], ctypes.c_int),
            "ssw_server_poll": ([vp, u64, ctypes.POINTER(Result)], ctypes.c_int),
            "ssw_server_wait": ([vp, u64, i32, ctypes.POINTER(Result)], ctypes.c_int),
            "ssw_server_stats_get": ([vp, ctypes.POINTER(ServerStats)], ctypes.c_int),
            "ssw_server_destroy": ([vp], ctypes.c_int),
            "ssw_server_last_error": ([vp], ctypes.c_char_p),
            "ssw_server_cancel": ([vp, ctypes.c_uint64], ctypes.c_int),
            "ssw_measure_atomics": ([vp, i32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "ssw_last_launch_stats": ([vp, ctypes.POINTER(LaunchStat), i32, ctypes.POINTER(i32)], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _libs[path] = L
    return _libs[path]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _progress_dict(p):
    return dict(automata_done=p.automata_done, subsets_expanded=p.subsets_expanded, launches=p.launches,
                deferred=p.deferred, in_flight=p.in_flight)


# ---------------------------------------------------------------- generators

def generate_uniform(seed: int, n: int, count: int, first: int = 0):
    a = np.empty((count, n), np.int32)
    b = np.empty((count, n), np.int32)
    rc = lib().ssw_generate_uniform(seed, n, first, count, _ptr(a), _ptr(b))
    if rc:
        raise SswError(rc, "generate_uniform")
    return a, b


def generate_cerny(n: int, convention: int = 1):
    """convention 0: the reference's (a merges 0->1, b cycles); 1: swapped."""
    a = np.empty(n, np.int32)
    b = np.empty(n, np.int32)
    rc = lib().ssw_generate_cerny(n, convention, _ptr(a), _ptr(b))
    if rc:
        raise SswError(rc, "generate_cerny")
    return a, b


def generate_dotnet(seed: int, n: int, count: int):
    a = np.empty((count, n), np.int32)
    b = np.empty((count, n), np.int32)
    rc = lib().ssw_generate_dotnet(seed, n, count, _ptr(a), _ptr(b))
    if rc:
        raise SswError(rc, "generate_dotnet")
    return a, b


def pack(n_per, A_list, B_list):
    """Packs automata of any sizes into (n_per u8, offset u64, delta u8)."""
    n_per = np.asarray(n_per, np.uint8)
    offs = np.zeros(len(n_per), np.uint64)
    total = int(2 * n_per.astype(np.int64).sum())
    delta = np.empty(total, np.uint8)
    o = 0
    for i, (a, b) in enumerate(zip(A_list, B_list)):
        n = int(n_per[i])
        offs[i] = o
        delta[o:o + n] = a
        delta[o + n:o + 2 * n] = b
        o += 2 * n
    return n_per, offs, delta


def pack_uniform(A, B):
    A = np.asarray(A)
    B = np.asarray(B)
    return np.ascontiguousarray(np.concatenate([A, B], axis=1).astype(np.uint8))


def witness_bits_to_word(row, L):
    """Unpacks a bit-packed witness row into a uint8 array of letters (1 = b)."""
    if L <= 0:
        return np.zeros(0, np.uint8)
    if isinstance(row, (bytes, bytearray)):
        row = np.frombuffer(row, np.uint8)
    bits = np.unpackbits(np.asarray(row, np.uint8), bitorder="little")
    return bits[:L].astype(np.uint8)


def witness_stride(nmax: int) -> int:
    return max(1, ((nmax - 1) * (nmax - 1) + 7) // 8)


@dataclass
class Solved:
    length: np.ndarray
    sync: np.ndarray
    expanded: np.ndarray
    witness: np.ndarray | None
    violation: int


class Context:
    """One solver context on one HIP device (a stream, resident scratch)."""

    def __init__(self, device: int = 0, budget: int = 0, library: str = None):
        self._L = lib(library)
        h = ctypes.c_void_p()
        rc = self._L.ssw_create(device, budget, 0, ctypes.byref(h))
        if rc:
            raise SswError(rc, self._L.ssw_last_error(None).decode())
        self.h = h
        self._cb = None

    def close(self):
        if self.h:
            self._L.ssw_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != SSW_OK:
            raise SswError(rc, f"{what}: {self._L.ssw_last_error(self.h).decode()}")

    # -- IComputable-shaped entry points (host int32 arrays, one n) --------
    def compute(self, n, A, B, witness=False, bound=-1, want_sync=True, want_expanded=True):
        A = np.ascontiguousarray(A, np.int32).reshape(-1, n)
        B = np.ascontiguousarray(B, np.int32).reshape(-1, n)
        count = A.shape[0]
        length = np.empty(count, np.int32)
        sync = np.empty(count, np.uint8) if want_sync else None
        expanded = np.empty(count, np.int64) if want_expanded else None
        stride = witness_stride(n)
        wit = np.zeros((count, stride), np.uint8) if witness else None
        viol = ctypes.c_int64(-1)
        rc = self._L.ssw_compute(self.h, n, count, _ptr(A), _ptr(B), _ptr(length), _ptr(sync), _ptr(wit),
                                 stride, _ptr(expanded), bound, ctypes.byref(viol))
        self._check(rc, "ssw_compute")
        return Solved(length, sync, expanded, wit, viol.value)

    def verify(self, n, A, B, bound=-1):
        A = np.ascontiguousarray(A, np.int32).reshape(-1, n)
        B = np.ascontiguousarray(B, np.int32).reshape(-1, n)
        viol = ctypes.c_int64(-1)
        rc = self._L.ssw_verify(self.h, n, A.shape[0], _ptr(A), _ptr(B), bound, ctypes.byref(viol))
        self._check(rc, "ssw_verify")
        return viol.value

    def set_verify_prefilter(self, word_length):
        """6: ssw_verify tries the 64 words of length 6 before each search; 0: off."""
        self._check(self._L.ssw_set_verify_prefilter(self.h, word_length), "set_verify_prefilter")

    def best_parallelism(self):
        return self._L.ssw_best_parallelism(self.h)

    # -- packed batches on the host -----------------------------------------
    def solve_packed(self, delta, count, uniform_n=0, n_per=None, offset=None, witness=False,
                     bound=-1, mode="batch"):
        length = np.empty(count, np.int32)
        sync = np.empty(count, np.uint8)
        expanded = np.empty(count, np.int64)
        nmax = uniform_n if uniform_n > 0 else int(np.max(n_per)) if count else 1
        stride = witness_stride(nmax)
        wit = np.zeros((count, stride), np.uint8) if witness else None
        b = Batch(count=count, uniform_n=uniform_n, witness_stride=stride,
                  n_per=_ptr(None if n_per is None else np.ascontiguousarray(n_per, np.uint8)),
                  offset=_ptr(None if offset is None else np.ascontiguousarray(offset, np.uint64)),
                  delta=_ptr(np.ascontiguousarray(delta, np.uint8)), length=_ptr(length), sync=_ptr(sync),
                  expanded=_ptr(expanded), witness=_ptr(wit), bound=bound, reserved=0)
        # keep converted arrays alive for the call
        self._keep = (n_per, offset, delta)
        viol = ctypes.c_int64(-1)
        fn = {"batch": self._L.ssw_solve_batch, "plain": self._L.ssw_solve_plain,
              "large": self._L.ssw_solve_large}[mode]
        rc = fn(self.h, ctypes.byref(b), SSW_HOST, ctypes.byref(viol))
        self._check(rc, f"ssw_solve_{mode}")
        return Solved(length, sync, expanded, wit, viol.value)

    def solve_mixed(self, n_per, A_list, B_list, witness=False, bound=-1, mode="batch"):
        n_per, offs, delta = pack(n_per, A_list, B_list)
        return self.solve_packed(delta, len(n_per), 0, n_per, offs, witness, bound, mode)

    def solve_uniform(self, n, A, B, witness=False, bound=-1, mode="batch"):
        delta = pack_uniform(np.asarray(A).reshape(-1, n), np.asarray(B).reshape(-1, n))
        return self.solve_packed(delta.reshape(-1), delta.shape[0], n, None, None, witness, bound, mode)

    # -- device buffers ---------------------------------------------------------
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._check(self._L.ssw_device_alloc(self.h, nbytes, ctypes.byref(p)), "alloc")
        return p.value

    def free(self, p):
        self._L.ssw_device_free(self.h, p)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(max(1, arr.nbytes))
        self._check(self._L.ssw_copy_to_device(self.h, p, _ptr(arr), arr.nbytes), "copy")
        return p

    def to_host(self, p, dtype, count):
        out = np.empty(count, dtype)
        self._check(self._L.ssw_copy_to_host(self.h, _ptr(out), p, out.nbytes), "copy")
        return out

    def copy_on_device(self, dst, src, nbytes):
        self._check(self._L.ssw_copy_on_device(self.h, dst, src, nbytes), "copy on device")

    def solve_device(self, batch: Batch, mode="batch"):
        viol = ctypes.c_int64(-1)
        fn = {"batch": self._L.ssw_solve_batch, "plain": self._L.ssw_solve_plain,
              "large": self._L.ssw_solve_large}[mode]
        self._check(fn(self.h, ctypes.byref(batch), SSW_DEVICE, ctypes.byref(viol)), f"solve_{mode}")
        return viol.value

    def solve_device_noalloc(self, batch: Batch):
        """ssw_solve_batch_noalloc: device buffers, one n <= 16, no witness,
        nothing allocated during the call."""
        viol = ctypes.c_int64(-1)
        self._check(self._L.ssw_solve_batch_noalloc(self.h, ctypes.byref(batch), ctypes.byref(viol)),
                    "solve_batch_noalloc")
        return viol.value

    def synchronize(self):
        self._check(self._L.ssw_synchronize(self.h), "synchronize")

    def measure_atomics(self, kind):
        """Test-and-set ceiling of the device in operations/s: kind "lds" (8 KiB
        bitmap per wavefront) or "hbm" (128 MiB bitmap, device scope)."""
        ops = ctypes.c_double(0.0)
        self._check(self._L.ssw_measure_atomics(self.h, {"lds": 0, "hbm": 1}[kind], ctypes.byref(ops)),
                    "measure_atomics")
        return ops.value

    def last_launch_stats(self):
        """Per launch of the last batched solve: shape, HIP-event duration on
        its stream, automata finished, their E and subsets expanded."""
        buf = (LaunchStat * 32)()
        cnt = ctypes.c_int32(0)
        self._check(self._L.ssw_last_launch_stats(self.h, buf, 32, ctypes.byref(cnt)), "launch stats")
        out = []
        for k in range(min(cnt.value, 32)):
            x = buf[k]
            out.append({f: (getattr(x, f).decode() if f == "kernel" else getattr(x, f))
                        for f, _ in LaunchStat._fields_ if f != "pass_"} | {"pass": x.pass_})
        return out

    def last_kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._L.ssw_last_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)), "kernel time")
        return ms.value, n.value

    # -- tuning / progress -----------------------------------------------------
    def get_tuning(self, n):
        t = Tuning()
        self._check(self._L.ssw_get_tuning(self.h, n, ctypes.byref(t)), "get_tuning")
        return dict(group_lanes=t.group_lanes, block_threads=t.block_threads, chunk=t.chunk,
                    arena_lds=t.arena_lds, lane_entries=t.lane_entries, spill_arena_lds=t.spill_arena_lds,
                    spill_lane_entries=t.spill_lane_entries, spill_group_lanes=t.spill_group_lanes,
                    narrow_bits=t.narrow_bits)

    def set_tuning(self, n, group_lanes, block_threads, chunk, arena_lds, lane_entries=1, spill_arena_lds=0,
                   spill_lane_entries=0, spill_group_lanes=0, narrow_bits=0):
        t = Tuning(group_lanes, block_threads, chunk, arena_lds, lane_entries, spill_arena_lds, spill_lane_entries,
                   spill_group_lanes, narrow_bits)
        self._check(self._L.ssw_set_tuning(self.h, n, ctypes.byref(t)), "set_tuning")

    def autotune(self, n_lo, n_hi, samples=20000, seed=7, out_path=None):
        self._check(self._L.ssw_autotune(self.h, n_lo, n_hi, samples, seed,
                                         out_path.encode() if out_path else None), "autotune")

    def autotune_batch(self, batch: Batch):
        self._check(self._L.ssw_autotune_batch(self.h, ctypes.byref(batch), SSW_DEVICE), "autotune_batch")

    def load_tuning(self, path):
        self._check(self._L.ssw_load_tuning(self.h, path.encode()), "load_tuning")

    def progress(self):
        p = Progress()
        self._check(self._L.ssw_get_progress(self.h, ctypes.byref(p)), "progress")
        return _progress_dict(p)

    def set_progress_callback(self, fn, interval_ms=50):
        """Calls fn(progress dict) from the library's thread (None: stop).

        The library joins its previous thread before it returns, so the old
        trampoline is released only after that call: the thread may be inside
        it until then."""
        if fn is None:
            self._check(self._L.ssw_set_progress_callback(self.h, PROGRESS_FN(), None, 0), "progress cb")
            self._cb = None
            return

        def tramp(p, _user):
            fn(_progress_dict(p.contents))
        cb = PROGRESS_FN(tramp)
        self._check(self._L.ssw_set_progress_callback(self.h, cb, None, interval_ms), "progress cb")
        self._cb = cb


class Server:
    """Continuous-batching request loop on a context (ssw_server_*)."""

    def __init__(self, ctx: Context, mem_budget=0, small_n_max=20, levels_per_tick=32, max_large=0,
                 want_witness=False, small_workers=0):
        self._L = ctx._L
        self.ctx = ctx
        cfg = ServerConfig(mem_budget, small_n_max, levels_per_tick, max_large, 1 if want_witness else 0,
                           small_workers, 0)
        h = ctypes.c_void_p()
        rc = self._L.ssw_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise SswError(rc, self._L.ssw_last_error(ctx.h).decode())
        self.h = h

    def submit(self, n, a, b):
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        t = ctypes.c_uint64()
        rc = self._L.ssw_server_submit(self.h, n, _ptr(a), _ptr(b), ctypes.byref(t))
        if rc:
            raise SswError(rc, self._why("submit"))
        return t.value

    @staticmethod
    def _res(r: Result):
        L = r.length
        return dict(length=L, n=r.n, expanded=r.expanded, sync=bool(r.sync), latency_us=r.latency_us,
                    witness=(witness_bits_to_word(bytes(r.witness), L) if r.witness_valid else None))

    def poll(self, ticket):
        r = Result()
        rc = self._L.ssw_server_poll(self.h, ticket, ctypes.byref(r))
        if rc == SSW_PENDING:
            return None
        if rc:
            raise SswError(rc, self._why("poll"))
        return self._res(r)

    def wait(self, ticket, timeout_ms=60000):
        r = Result()
        rc = self._L.ssw_server_wait(self.h, ticket, timeout_ms, ctypes.byref(r))
        if rc:
            raise SswError(rc, self._why("wait"))
        return self._res(r)

    def cancel(self, ticket):
        """True when the request was still queued and is cancelled (its
        ticket released), False when it is already being solved or done."""
        rc = self._L.ssw_server_cancel(self.h, ticket)
        if rc == SSW_ERR_STATE:
            return False
        if rc:
            raise SswError(rc, self._why("cancel"))
        return True

    def _why(self, what):
        msg = self._L.ssw_server_last_error(self.h).decode()
        return f"{what}: {msg}" if msg else what

    def stats(self):
        s = ServerStats()
        rc = self._L.ssw_server_stats_get(self.h, ctypes.byref(s))
        if rc:
            raise SswError(rc, self._why("stats"))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            self._L.ssw_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
