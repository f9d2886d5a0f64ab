"""GPU parity of the batched persistent kernel and the plain one-per-launch path
against the CPU oracle (restatement of SlimCPUUntweaked.cs:51-222 / CPU.cs:93-283).

Bar: lengths, synchronizing flags and canonical expanded counts bit-exact;
witnesses equal to the oracle's (the lexicographically smallest shortest reset
word) and re-verified on Q as Verify.cs:12-53 does.
"""
import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu


def _check(n, A, B, res, witness_rows=None, tag=""):
    L, S, E, _ = O.solve(n, A, B, threads=8)
    bad = np.nonzero(res.length != L)[0]
    assert len(bad) == 0, f"{tag}: length mismatch at {bad[:10]} gpu={res.length[bad[:10]]} cpu={L[bad[:10]]}"
    np.testing.assert_array_equal(res.sync, S, err_msg=f"{tag}: sync")
    np.testing.assert_array_equal(res.expanded, E, err_msg=f"{tag}: expanded")
    if witness_rows is not None:
        for i in witness_rows:
            if L[i] < 0:
                continue
            w = ssw.witness_bits_to_word(res.witness[i], int(L[i]))
            assert O.verify_word(n, A[i], B[i], w), f"{tag}: witness {i} does not synchronize"
            if n <= 24:
                Lw, wo = O.witness(n, A[i], B[i])
                assert Lw == L[i]
                np.testing.assert_array_equal(w, wo, err_msg=f"{tag}: witness {i} differs from oracle")


@pytest.mark.parametrize("conv", [0, 1])
def test_cerny_known_answers(ctx, conv):
    # SimpleAutomataTests.cs:14-35 / IntegrityTest.cs:44-62: C_n has length (n-1)^2
    for n in range(2, 17):
        a, b = ssw.generate_cerny(n, conv)
        A = np.tile(a, (3, 1))
        B = np.tile(b, (3, 1))
        r = ctx.compute(n, A, B, witness=True)
        assert list(r.length) == [(n - 1) ** 2] * 3
        assert list(r.expanded) == [2 ** n - n - 1] * 3
        _check(n, A, B, r, witness_rows=range(3), tag=f"cerny{n}")
        # closed form of the unique shortest word (b a^(n-1))^(n-2) b, letters swapped for conv 0
        word = ([1] + [0] * (n - 1)) * (n - 2) + [1]
        if conv == 0:
            word = [1 - x for x in word]
        np.testing.assert_array_equal(ssw.witness_bits_to_word(r.witness[0], (n - 1) ** 2), word)


def test_plain_n10_full_with_witness(ctx):
    n, A, B = W.plain_n10()
    r = ctx.compute(n, A, B, witness=True)
    _check(n, A, B, r, witness_rows=range(len(A)), tag="n10")


def test_uniform_n16_slice(ctx):
    n, delta, pos = W.uniform_n16(count=20_000)
    A = delta[:, :n].astype(np.int32)
    B = delta[:, n:].astype(np.int32)
    r = ctx.solve_packed(delta.reshape(-1), len(delta), n)
    _check(n, A, B, r, tag="n16")
    assert np.all(r.length[pos] == 225)


def test_uniform_n16_witness(ctx):
    n, delta, pos = W.uniform_n16(count=3_000)
    A = delta[:, :n].astype(np.int32)
    B = delta[:, n:].astype(np.int32)
    r = ctx.solve_packed(delta.reshape(-1), len(delta), n, witness=True)
    _check(n, A, B, r, witness_rows=range(len(A)), tag="n16w")


def test_host_witness_calls_through_pinned_memory(ctx):
    """Calls with witness words of 2^18 automata or more from host memory go
    through pinned memory both ways (rows copied and checked by the host
    pool into the staging buffer, results copied back into pinned memory and
    out by the pool): lengths, flags and counts against the oracle for every
    automaton, words against the oracle's back-tracking on a sample (with the
    Cerny automata), both entry points, and a bad row reported by index."""
    n = 16
    count = (1 << 18) + 4_321
    A, B = ssw.generate_uniform(1621, n, count)
    ca, cb = ssw.generate_cerny(n)
    cern = (0, 131_072, count - 1)
    for i in cern:
        A[i], B[i] = ca, cb
    rows = ssw.pack_uniform(A, B).reshape(-1)
    print("witness through pinned memory: solve_packed", flush=True)
    r = ctx.solve_packed(rows, count, n, witness=True)
    sample = list(cern) + list(np.random.default_rng(1622).choice(count, 400, replace=False))
    _check(n, A, B, r, witness_rows=sample, tag="pinned witness rows")
    print("witness through pinned memory: compute", flush=True)
    r2 = ctx.compute(n, A, B, witness=True)
    np.testing.assert_array_equal(r2.length, r.length)
    np.testing.assert_array_equal(r2.expanded, r.expanded)
    np.testing.assert_array_equal(r2.witness, r.witness)
    bad = rows.copy().reshape(count, 2 * n)
    bad[200_003, 5] = n
    with pytest.raises(ssw.SswError, match="200003"):
        ctx.solve_packed(bad.reshape(-1), count, n, witness=True)


def test_host_mixed_sizes_back_through_pinned_memory(ctx):
    """A host batch of 2^18 automata or more of mixed sizes (rows at offsets,
    n = 12..20, sizes interleaved) returns its results through pinned memory
    and the host pool: every automaton against the oracle, per size."""
    rng = np.random.default_rng(1623)
    count = (1 << 18) + 777
    ns = rng.integers(12, 21, size=count)
    As, Bs = [], []
    for n in range(12, 21):
        k = int((ns == n).sum())
        a, b = ssw.generate_uniform(1700 + n, n, k)
        As.append(a)
        Bs.append(b)
    nxt = [0] * 21
    A_list, B_list = [], []
    for n in ns:
        A_list.append(As[n - 12][nxt[n]])
        B_list.append(Bs[n - 12][nxt[n]])
        nxt[n] += 1
    print("mixed sizes through pinned memory: solve", flush=True)
    r = ctx.solve_mixed(ns, A_list, B_list)
    for n in range(12, 21):
        idx = np.nonzero(ns == n)[0]
        L, S, E, _ = O.solve(n, As[n - 12], Bs[n - 12], threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.sync[idx], S, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")


def test_mixed_sizes_small(ctx):
    rng = np.random.default_rng(5)
    ns = rng.integers(1, 17, size=3000).astype(np.uint8)
    As, Bs = [], []
    for i, n in enumerate(ns):
        a, b = ssw.generate_uniform(77, int(n), 1, first=i)
        As.append(a[0])
        Bs.append(b[0])
    r = ctx.solve_mixed(ns, As, Bs, witness=True)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([As[i] for i in idx])
        B = np.stack([Bs[i] for i in idx])
        sub = ssw.Solved(r.length[idx], r.sync[idx], r.expanded[idx], r.witness[idx], -1)
        _check(int(n), A, B, sub, witness_rows=range(len(idx)), tag=f"mixed n={n}")


@pytest.mark.parametrize("n", [17, 18, 20])
def test_hbm_bitmap(ctx, n):
    A, B = ssw.generate_uniform(900 + n, n, 2000)
    r = ctx.compute(n, A, B, witness=True)
    _check(n, A, B, r, witness_rows=range(200), tag=f"hbm n={n}")


def test_plain_path(ctx):
    n, delta, pos = W.uniform_n16(count=300, cerny_every=100)
    A = delta[:, :n].astype(np.int32)
    B = delta[:, n:].astype(np.int32)
    r = ctx.solve_packed(delta.reshape(-1), len(delta), n, witness=True, mode="plain")
    _check(n, A, B, r, witness_rows=range(len(A)), tag="plain")


def test_violation_index(ctx):
    # a lowered bound exercises the Cerny-violation path (SlimCPUUntweaked.cs:189-213)
    n, A, B = W.plain_n10(count=500)
    L, _, _, v = O.slim(n, A, B, bound=7)
    r = ctx.compute(n, A, B, bound=7)
    assert r.violation == v == int(np.nonzero(L > 7)[0][0])
    assert ctx.verify(n, A, B, bound=7) == v
    assert ctx.verify(n, A, B) == -1


def test_bad_transition_rejected(ctx):
    A = np.zeros((2, 5), np.int32)
    B = np.zeros((2, 5), np.int32)
    B[1, 3] = 5
    with pytest.raises(ssw.SswError):
        ctx.compute(5, A, B)


def test_repeatable(ctx):
    n, delta, pos = W.uniform_n16(count=5_000)
    r1 = ctx.solve_packed(delta.reshape(-1), len(delta), n, witness=True)
    r2 = ctx.solve_packed(delta.reshape(-1), len(delta), n, witness=True)
    np.testing.assert_array_equal(r1.length, r2.length)
    np.testing.assert_array_equal(r1.witness, r2.witness)


def test_host_arrays_chunked(ctx):
    """Host int32 arrays large enough to be packed and solved in chunks (the
    packing of chunk k + 1 overlaps chunk k's transfer and kernels): every
    output equals the oracle and the violating index is the smallest over all
    chunks.  The same arrays with witness words (solved whole) agree."""
    n = 12
    count = 600_000  # two chunks of 300,000
    A, B = ssw.generate_uniform(4242, n, count)
    L, S, E, v = O.slim(n, A, B, bound=9, threads=8)
    r = ctx.compute(n, A, B, bound=9)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    assert r.violation == v == int(np.nonzero(L > 9)[0][0])
    r = ctx.compute(n, A, B, witness=True, bound=9)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.expanded, E)
    assert r.violation == v
    for i in range(0, count, 9973):
        if L[i] > 0:
            w = ssw.witness_bits_to_word(r.witness[i], int(L[i]))
            assert O.verify_word(n, A[i], B[i], w)
    # the first violation only in the last chunk
    late = np.nonzero(L > 9)[0]
    late = int(late[late > count - 20_000][0])
    A2, B2 = A.copy(), B.copy()
    for i in np.nonzero(L > 9)[0]:
        if i != late:
            A2[i] = 0  # a constant letter: length 1
    assert ctx.verify(n, A2, B2, bound=9) == late
    assert ctx.verify(n, A2, B2) == -1


def test_host_arrays_bad_transition_in_last_chunk(ctx):
    n = 10
    A, B = ssw.generate_uniform(77, n, 600_000)  # two chunks
    B[599_990, 4] = n
    with pytest.raises(ssw.SswError, match="automaton 599990 state 4"):
        ctx.compute(n, A, B)
    B[599_990, 4] = 0
    # a negative transition in the first chunk, after a valid run of the same range
    A[123_456, 7] = -1
    B[123_457, 0] = n + 3
    with pytest.raises(ssw.SswError, match="automaton 123456 state 7"):
        ctx.compute(n, A, B)
    # the context stays usable
    r = ctx.compute(n, A[:1000], B[:1000])
    L, _, _, _ = O.slim(n, A[:1000], B[:1000])
    np.testing.assert_array_equal(r.length, L)


def test_empty_batches(ctx):
    A = np.zeros((0, 7), np.int32)
    r = ctx.compute(7, A, A, witness=True)
    assert r.length.shape == (0,) and r.violation == -1
    assert ctx.verify(7, A, A) == -1
    for mode in ("batch", "plain", "large"):
        r = ctx.solve_packed(np.zeros(1, np.uint8), 0, 7, mode=mode)
        assert r.length.shape == (0,) and r.violation == -1


@pytest.mark.parametrize("mode", ["batch", "plain", "large"])
def test_device_batch_marks_invalid_automata(ctx, mode):
    # host batches are checked on the host (the call fails); a device batch is
    # checked by the kernels: an automaton with a transition >= n is skipped
    # with length SSW_LEN_INVALID and the others are solved
    n = 9
    A, B = ssw.generate_uniform(5, n, 64)
    A[3, 2] = n
    B[40, 8] = 200
    delta = ssw.pack_uniform(A, B).reshape(-1)
    with pytest.raises(ssw.SswError, match="automaton 3:"):
        ctx.solve_packed(delta, 64, n, mode=mode)
    count = 64
    d = ctx.to_device(delta)
    dl, ds, de = ctx.alloc(4 * count), ctx.alloc(count), ctx.alloc(8 * count)
    try:
        b = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d, length=dl,
                      sync=ds, expanded=de, witness=None, bound=-1, reserved=0)
        ctx.solve_device(b, mode=mode)
        L = ctx.to_host(dl, np.int32, count)
        S = ctx.to_host(ds, np.uint8, count)
        E = ctx.to_host(de, np.int64, count)
    finally:
        for p in (d, dl, ds, de):
            ctx.free(p)
    bad = [3, 40]
    assert np.all(L[bad] == ssw.LEN_INVALID) and np.all(S[bad] == 0) and np.all(E[bad] == 0)
    good = np.setdiff1d(np.arange(count), bad)
    Lo, So, Eo, _ = O.slim(n, A[good], B[good])
    np.testing.assert_array_equal(L[good], Lo)
    np.testing.assert_array_equal(S[good], So)
    np.testing.assert_array_equal(E[good], Eo)


@pytest.mark.parametrize("witness", [True, False])
@pytest.mark.parametrize("mode", ["batch", "plain", "large"])
def test_tiny_n_every_mode(ctx, mode, witness):
    # n = 1 (Q is a singleton: not synchronizing by the reference's rule, E = 1)
    # up to n = 5, in every entry point.  For such n a level's 2 * width images
    # can outnumber the 2^n subsets: the device-wide ring is sized to the
    # subsets, not to that bound.
    for n in (1, 2, 3, 4, 5):
        A, B = ssw.generate_uniform(11 + n, n, 200)
        if n >= 2:
            A[0], B[0] = ssw.generate_cerny(n, 1)
        r = ctx.solve_uniform(n, A, B, witness=witness, mode=mode)
        L, S, E, _ = O.slim(n, A, B)
        np.testing.assert_array_equal(r.length, L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.sync, S, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded, E, err_msg=f"n={n}")
        if witness:
            for i in np.nonzero(L >= 0)[0]:
                w = ssw.witness_bits_to_word(r.witness[i], int(L[i]))
                np.testing.assert_array_equal(w, O.witness(n, A[i], B[i])[1], err_msg=f"n={n} i={i}")


@pytest.mark.parametrize("mode", ["batch", "plain", "large"])
def test_device_batch_with_bad_size_rejected(ctx, mode):
    n_per = np.array([5, 0, 7], np.uint8)
    offs = np.array([0, 10, 10], np.uint64)
    delta = np.zeros(24, np.uint8)
    d = [ctx.to_device(x) for x in (delta, n_per, offs)]
    dl = ctx.alloc(12)
    try:
        b = ssw.Batch(count=3, uniform_n=0, witness_stride=0, n_per=d[1], offset=d[2], delta=d[0], length=dl,
                      sync=None, expanded=None, witness=None, bound=-1, reserved=0)
        with pytest.raises(ssw.SswError) as e:
            ctx.solve_device(b, mode=mode)
        assert e.value.code == ssw.SSW_ERR_ARG
    finally:
        for p in d + [dl]:
            ctx.free(p)


@pytest.mark.parametrize("mode", ["batch", "plain", "large"])
def test_violation_index_every_mode(ctx, mode):
    # a lowered bound: the first automaton whose length exceeds it
    # (SlimCPUUntweaked.cs:189-213), from every entry point
    n = 11
    A, B = ssw.generate_uniform(808, n, 300)
    L, _, _, v = O.slim(n, A, B, bound=8)
    assert v >= 0
    r = ctx.solve_uniform(n, A, B, bound=8, mode=mode)
    assert r.violation == v
    np.testing.assert_array_equal(r.length, L)


def test_offsets_in_any_order(ctx):
    # offsets need not follow the automaton order: the batch packed backwards
    rng = np.random.default_rng(21)
    ns = rng.integers(2, 17, size=500).astype(np.uint8)
    As, Bs = [], []
    for i, n in enumerate(ns):
        a, b = ssw.generate_uniform(99, int(n), 1, first=i)
        As.append(a[0])
        Bs.append(b[0])
    n_per, offs, delta = ssw.pack(ns, As, Bs)
    # rebuild delta with the automata stored in reverse order
    sizes = 2 * ns.astype(np.uint64)
    roffs = np.zeros(len(ns), np.uint64)
    pos = 0
    rdelta = np.empty_like(delta)
    for i in reversed(range(len(ns))):
        roffs[i] = pos
        rdelta[pos:pos + int(sizes[i])] = delta[int(offs[i]):int(offs[i]) + int(sizes[i])]
        pos += int(sizes[i])
    r1 = ctx.solve_packed(delta, len(ns), 0, n_per, offs, witness=True)
    r2 = ctx.solve_packed(rdelta, len(ns), 0, n_per, roffs, witness=True)
    np.testing.assert_array_equal(r1.length, r2.length)
    np.testing.assert_array_equal(r1.expanded, r2.expanded)
    np.testing.assert_array_equal(r1.witness, r2.witness)


@pytest.mark.parametrize("mode", ["batch", "plain", "large"])
def test_wide_witness_rows_zero_padded(ctx, mode):
    # rows wider than the word: the bytes past it are zero
    n = 10
    A, B = ssw.generate_uniform(31, n, 50)
    stride = 40
    delta = ssw.pack_uniform(A, B).reshape(-1)
    length = np.empty(50, np.int32)
    wit = np.full((50, stride), 0xAB, np.uint8)
    b = ssw.Batch(count=50, uniform_n=n, witness_stride=stride, n_per=None, offset=None,
                  delta=delta.ctypes.data, length=length.ctypes.data, sync=None, expanded=None,
                  witness=wit.ctypes.data, bound=-1, reserved=0)
    fn = {"batch": ctx._L.ssw_solve_batch, "plain": ctx._L.ssw_solve_plain, "large": ctx._L.ssw_solve_large}[mode]
    viol = ssw.ctypes.c_int64(-1)
    assert fn(ctx.h, ssw.ctypes.byref(b), ssw.SSW_HOST, ssw.ctypes.byref(viol)) == 0
    L, _, _, _ = O.slim(n, A, B)
    np.testing.assert_array_equal(length, L)
    for i in range(50):
        used = (int(L[i]) + 7) // 8 if L[i] > 0 else 0
        assert np.all(wit[i, used:] == 0), (i, L[i])
        if L[i] > 0:
            w = ssw.witness_bits_to_word(wit[i], int(L[i]))
            np.testing.assert_array_equal(w, O.witness(n, A[i], B[i])[1])


def test_packed_host_rows_first_bad_reported(ctx):
    """Packed host rows are checked on several host threads before anything is
    copied: the first bad automaton over all ranges is reported (not the first
    range's), and the next calls on the same context are solved exactly."""
    n = 12
    count = 200_000
    A, B = ssw.generate_uniform(31, n, count)
    delta = ssw.pack_uniform(A, B).reshape(-1)
    bad = delta.copy()
    bad[190_000 * 2 * n + 5] = 255
    bad[150_001 * 2 * n + n + 11] = n
    with pytest.raises(ssw.SswError, match="automaton 150001:"):
        ctx.solve_packed(bad, count, n)
    r = ctx.solve_packed(delta, count, n)
    L, S, E, _ = O.slim(n, A, B, threads=8)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    # a smaller batch after a larger one
    r = ctx.solve_packed(delta[:2 * n * 1000], 1000, n, witness=True)
    np.testing.assert_array_equal(r.length, L[:1000])


def test_host_arrays_chunked_n20(ctx):
    # the chunked host path for n > 16 (HBM bitmaps, hand-off of long
    # frontiers to the device-wide path) against the oracle
    n = 20
    count = 540_000  # two chunks
    A, B = ssw.generate_uniform(2020, n, count)
    L, S, E, _ = O.solve(n, A, B, threads=8)
    r = ctx.compute(n, A, B)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)


def test_host_arrays_pipelined_n16(ctx):
    """The pipelined host path (n <= 16, two or more chunks: the next chunk's
    transitions copied under the current chunk's launches, results copied
    back on their own stream) against the oracle, with Cerny C_16 stragglers
    in both chunks; lengths only as Verify asks for them; and the violation
    index of Verify when every violation lies in the second chunk."""
    n = 16
    count = 600_000  # two chunks of 300,000
    A, B = ssw.generate_uniform(1617, n, count)
    ca, cb = ssw.generate_cerny(n)
    for i in (7, 123_456, 299_999, 300_000, 450_001, count - 1):
        A[i], B[i] = ca, cb
    print("pipelined n=16: oracle", flush=True)
    L, S, E, _ = O.solve(n, A, B, threads=8)
    print("pipelined n=16: compute", flush=True)
    r = ctx.compute(n, A, B)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    assert r.violation == -1
    r2 = ctx.compute(n, A, B, want_sync=False, want_expanded=False)
    np.testing.assert_array_equal(r2.length, L)
    # a bound every automaton of the first chunk keeps: the first violation
    # (C_16 at 300,000 is the first automaton above the bound) is found in
    # the second chunk and reported at its index in the whole call
    bound = int(L[:300_000][L[:300_000] != 225].max())
    A2, B2 = A.copy(), B.copy()
    for i in (7, 123_456, 299_999):
        A2[i], B2[i] = A[0], B[0]
    L2 = L.copy()
    L2[[7, 123_456, 299_999]] = L[0]
    expected = int(np.nonzero(L2 > bound)[0][0])
    assert expected >= 300_000
    assert ctx.verify(n, A2, B2, bound=bound) == expected
    assert ctx.compute(n, A2, B2, bound=bound).violation == expected


def test_host_pipeline_every_automaton_to_the_last_launch(ctx):
    """The n = 16 host pipeline hands what the chunks' last launch takes to
    one list shared by the chunks (each segment of its queue sized for a 64th
    of the batch plus a slack of a few entries per chunk).  Here every
    automaton has a permutation letter, so the first launch hands each one to
    that list and the list is full to its last slot: results against the
    oracle through both host entry points, and the violation index of an
    automaton in the second chunk found by the last launch."""
    n = 16
    count = (1 << 19) + 12_345  # two chunks
    rng = np.random.default_rng(1620)
    A = np.argsort(rng.random((count, n)), axis=1).astype(np.int32)  # a permutes the states
    B = rng.integers(0, n, size=(count, n), dtype=np.int32)
    ca, cb = ssw.generate_cerny(n)
    for i in (0, 300_001, count - 1):
        A[i], B[i] = ca, cb
    print("pipeline all deferred: oracle", flush=True)
    L, S, E, _ = O.solve(n, A, B, threads=8)
    print("pipeline all deferred: compute", flush=True)
    r = ctx.compute(n, A, B)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    print("pipeline all deferred: packed rows", flush=True)
    rows = ssw.pack_uniform(A, B).reshape(-1)
    r2 = ctx.solve_packed(rows, count, n)
    np.testing.assert_array_equal(r2.length, L)
    np.testing.assert_array_equal(r2.expanded, E)
    # a bound below C_16's 225 that every other automaton keeps: the first
    # violation is C_16 at 0; without it, the one at 300,001 (second chunk)
    bound = int(L[L != 225].max())
    assert bound < 225
    assert ctx.compute(n, A, B, bound=bound).violation == 0
    A[0], B[0] = A[1], B[1]
    assert ctx.verify(n, A, B, bound=bound) == 300_001


def test_host_arrays_pipelined_bad_input_late(ctx):
    """A transition out of range in the second chunk, found while the first
    chunk is already on the device: the call fails naming that automaton,
    drains its work and leaves the context usable."""
    n = 13
    count = 560_000
    A, B = ssw.generate_uniform(1618, n, count)
    A[400_123, 5] = n
    with pytest.raises(ssw.SswError, match="400123"):
        ctx.compute(n, A, B)
    A[400_123, 5] = 0
    r = ctx.compute(n, A[:1000], B[:1000])
    L, _, _, _ = O.solve(n, A[:1000], B[:1000], threads=8)
    np.testing.assert_array_equal(r.length, L)


def test_host_rows_pipelined_n16(ctx):
    """Packed rows in host memory through ssw_solve_batch in pipelined chunks
    (rows copied into pinned staging and checked by host threads under the
    earlier chunks' transfers and kernels) against the oracle, the violation
    index with a lowered bound, and a bad row in the second chunk."""
    n = 16
    count = 560_000  # two chunks
    A, B = ssw.generate_uniform(1619, n, count)
    ca, cb = ssw.generate_cerny(n)
    for i in (3, 279_999, 280_000, count - 2):
        A[i], B[i] = ca, cb
    rows = ssw.pack_uniform(A, B).reshape(-1)
    print("pipelined rows n=16: oracle", flush=True)
    L, S, E, _ = O.solve(n, A, B, threads=8)
    print("pipelined rows n=16: solve", flush=True)
    r = ctx.solve_packed(rows, count, n)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    assert r.violation == -1
    assert ctx.solve_packed(rows, count, n, bound=100).violation == 3
    bad = rows.copy().reshape(count, 2 * n)
    bad[300_007, 2 * n - 1] = n
    with pytest.raises(ssw.SswError, match="300007"):
        ctx.solve_packed(bad.reshape(-1), count, n)
    r2 = ctx.solve_packed(rows[: 2 * n * 1000], 1000, n)
    np.testing.assert_array_equal(r2.length, L[:1000])
