"""The bounds-checked build (libssw_checked.so, -DSSW_CHECKED) on the launch
configurations of every route: its kernels check each LDS address the n <= 16
solver forms and each index into the HBM rings, arenas, visited and witness
bitmaps, recorded history and level tables, and a solve that formed one out of
bounds fails with SSW_ERR_INTERNAL naming the source line.  Each test below
passes only if no check fired and the results equal the CPU oracle's
(restatement of SlimCPUUntweaked.cs:51-222) bit for bit, or, for automata the
oracle cannot search in seconds, those of libssw.so."""
import os

import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cctx():
    if not os.path.exists(ssw.CHECKED_LIB_PATH):
        pytest.skip("libssw_checked.so is not built (build.py --checked)")
    c = ssw.Context(0, library=ssw.CHECKED_LIB_PATH)
    yield c
    c.close()


def _same_as_oracle(r, n, A, B, tag):
    L, S, E, _ = O.solve(n, A, B, threads=8)
    np.testing.assert_array_equal(r.length, L, err_msg=f"{tag}: length")
    np.testing.assert_array_equal(r.sync, S, err_msg=f"{tag}: sync")
    np.testing.assert_array_equal(r.expanded, E, err_msg=f"{tag}: expanded")
    return L


def _n16_batch():
    # ordinary random automata, the widest of a larger sample (frontiers that
    # outgrow the first launch's ring) and Cerny C_16 in both conventions
    n, delta, _ = W.uniform_n16(seed=91, count=20_000, cerny_every=0)
    A = delta[:, :n].astype(np.int32)
    B = delta[:, n:].astype(np.int32)
    _, _, E, _ = O.slim(n, A, B, threads=8)
    wide = np.argsort(E)[-200:]
    rows, cols = [A[:2000], A[wide]], [B[:2000], B[wide]]
    for conv in (0, 1):
        a, b = ssw.generate_cerny(n, conv)
        rows.append(np.tile(a, (8, 1)))
        cols.append(np.tile(b, (8, 1)))
    perm = np.random.default_rng(3).permutation(sum(len(r) for r in rows))
    return n, np.concatenate(rows)[perm], np.concatenate(cols)[perm]


def test_checked_n16_default_launch_chain(cctx):
    """The committed n = 16 tuning: relabelled 2^15-bit first launch, full
    bitmap second launch, 2048-entry ring and HBM ring in the third."""
    n, A, B = _n16_batch()
    print(f"checked n=16 default chain: {len(A)} automata", flush=True)
    r = cctx.compute(n, A, B)
    L = _same_as_oracle(r, n, A, B, "n=16 default")
    assert (L == (n - 1) ** 2).sum() >= 16


# (group_lanes, block_threads, chunk, arena_lds, lane_entries, spill_arena_lds, spill_lane_entries,
#  spill_group_lanes[, narrow_bits])
CONFIGS = [
    (64, 128, 16, 512, 1, 1024, 4, 256),    # wide levels move to the HBM ring and back
    (256, 256, 2, 1024, 4, 0, 0, 0),        # no spill: level-split groups with HBM relocation
    (64, 0, 4, 256, 1, 2048, 8, 0),         # eight entries per lane, ring too small for C_16
    (32, 0, 8, 512, 1, 4096, 4, 0, 14),     # two automata per wavefront, 2^14-bit bitmap
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=[",".join(map(str, c)) for c in CONFIGS])
def test_checked_n16_configurations(cctx, cfg):
    n, A, B = _n16_batch()
    saved = cctx.get_tuning(n)
    try:
        cctx.set_tuning(n, *cfg)
        print(f"checked config {cfg}", flush=True)
        r = cctx.compute(n, A, B)
    finally:
        cctx.set_tuning(n, **saved)
    _same_as_oracle(r, n, A, B, f"config {cfg}")


def test_checked_n16_witness(cctx):
    """Witness words from the recorded history (per-level good-subset lists,
    HBM history for C_16): the lexicographically smallest shortest word."""
    n, A, B = _n16_batch()
    A, B = A[:600], B[:600]
    print("checked n=16 witness", flush=True)
    r = cctx.compute(n, A, B, witness=True)
    L = _same_as_oracle(r, n, A, B, "n=16 witness")
    for i in range(0, len(A), 7):
        Lo, wo = O.witness(n, A[i], B[i])
        assert L[i] == Lo
        if Lo >= 0:
            np.testing.assert_array_equal(ssw.witness_bits_to_word(r.witness[i], Lo), wo, err_msg=f"automaton {i}")


def _mixed_17_24(rng):
    """n = 17..24: images of Q inside few states (the n <= 16 solver's two
    families), 16..19 states (the general solver's 18-state families), plain
    random ones and Cerny C_n (HBM bitmap)."""
    ns, As, Bs = [], [], []
    for n in range(17, 25):
        for k in range(24):
            kind = k % 6
            if kind == 0:
                a = rng.integers(0, 15, size=n)
                b = rng.integers(n - 15, n, size=n)
            elif kind == 1:
                a = rng.integers(0, 12, size=n)
                b = rng.integers(0, 12, size=n)
            elif kind == 2:
                m = min(n, 16 + (k // 6) % 4)
                a = np.concatenate([np.arange(m), rng.integers(0, m, size=n - m)])
                b = rng.integers(0, n, size=n)
            elif kind == 3 and k < 12:
                a, b = ssw.generate_cerny(n, k % 2)
            else:
                a = rng.integers(0, n, size=n)
                b = rng.integers(0, n, size=n)
            perm = rng.permutation(n)
            inv = np.argsort(perm)
            ns.append(n)
            As.append(np.asarray(perm[np.asarray(a)[inv]], np.int32))
            Bs.append(np.asarray(perm[np.asarray(b)[inv]], np.int32))
    return np.asarray(ns, np.uint8), As, Bs


@pytest.mark.parametrize("witness", [False, True])
def test_checked_n17_24_families(cctx, witness):
    ns, As, Bs = _mixed_17_24(np.random.default_rng(57))
    print(f"checked n=17..24 families witness={witness}: {len(ns)} automata", flush=True)
    r = cctx.solve_mixed(ns, As, Bs, witness=witness)
    for n in range(17, 25):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([As[i] for i in idx])
        B = np.stack([Bs[i] for i in idx])
        L, S, E, _ = O.solve(n, A, B, threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.sync[idx], S, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")
        if witness:
            for j, i in enumerate(idx[::5]):
                if r.length[i] >= 0:
                    w = ssw.witness_bits_to_word(r.witness[i], int(r.length[i]))
                    assert O.verify_word(n, As[i], Bs[i], w), f"n={n} automaton {i}"


def test_checked_uneven_sizes(cctx):
    """Uneven sizes 12..24 in one batch: every launch class at once."""
    n_per, offs, delta = W.uneven_12_24(count=20_000)
    print("checked uneven 12..24", flush=True)
    r = cctx.solve_packed(delta, len(n_per), n_per=n_per, offset=offs)
    for n in np.unique(n_per):
        idx = np.nonzero(n_per == n)[0][:1500]
        rows = offs[idx].astype(np.int64)
        A = delta[rows[:, None] + np.arange(n)[None, :]].astype(np.int32)
        B = delta[rows[:, None] + int(n) + np.arange(n)[None, :]].astype(np.int32)
        L, S, E, _ = O.solve(int(n), A, B, threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")


def test_checked_device_wide_engine(cctx, ctx):
    """The device-wide level loop (n = 25..26, witness on): Cerny C_n has
    length (n-1)^2; random automata agree with libssw.so."""
    rng = np.random.default_rng(8)
    ns, As, Bs = [], [], []
    for n in (25, 26):
        a, b = ssw.generate_cerny(n, 1)
        ns.append(n)
        As.append(np.asarray(a, np.int32))
        Bs.append(np.asarray(b, np.int32))
        for _ in range(3):
            ns.append(n)
            As.append(rng.integers(0, n, size=n).astype(np.int32))
            Bs.append(rng.integers(0, n, size=n).astype(np.int32))
    ns = np.asarray(ns, np.uint8)
    print("checked device-wide engine n=25..26", flush=True)
    r = cctx.solve_mixed(ns, As, Bs, witness=True, mode="large")
    want = ctx.solve_mixed(ns, As, Bs, witness=True, mode="large")
    np.testing.assert_array_equal(r.length, want.length)
    np.testing.assert_array_equal(r.expanded, want.expanded)
    np.testing.assert_array_equal(r.witness, want.witness)
    assert r.length[0] == 24 ** 2 and r.length[4] == 25 ** 2
    for i in (0, 4):
        w = ssw.witness_bits_to_word(r.witness[i], int(r.length[i]))
        assert O.verify_word(int(ns[i]), As[i], Bs[i], w)


def test_checked_allocation_free_entry(cctx):
    """ssw_solve_batch_noalloc on the static scratch: the same results as the
    allocating entry point."""
    n, delta, pos = W.uniform_n16(seed=5, count=30_000)
    count = len(delta)
    d = cctx.to_device(np.ascontiguousarray(delta).reshape(-1))
    bufs = [d]
    out = []
    try:
        for fn in (cctx.solve_device, cctx.solve_device_noalloc):
            L, S, E = cctx.alloc(4 * count), cctx.alloc(count), cctx.alloc(8 * count)
            bufs += [L, S, E]
            b = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d, length=L,
                          sync=S, expanded=E, witness=None, bound=-1, reserved=0)
            print(f"checked {fn.__name__}", flush=True)
            fn(b)
            out.append((cctx.to_host(L, np.int32, count), cctx.to_host(E, np.int64, count)))
    finally:
        for p in bufs:
            cctx.free(p)
    np.testing.assert_array_equal(out[1][0], out[0][0])
    np.testing.assert_array_equal(out[1][1], out[0][1])
    assert np.all(out[0][0][pos] == (n - 1) ** 2)
