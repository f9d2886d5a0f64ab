"""The allocation-free entry point (ssw_solve_batch_noalloc) against
ssw_solve_batch on the same device batches -- lengths, flags, expanded counts
and the violating index, bit for bit -- on calls in a row whose batches
shrink, so a call that reads what an earlier, larger one left in the shared
scratch would show; and the limits it refuses."""
import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu


def _perm_merge(seed, count, n=16):
    # a random permutation and a letter merging one pair of states: the
    # reachable family is nearly all 2^n subsets and the widest levels hold
    # 14k-19k of them, more than any LDS ring, so these searches reach the
    # last launch and its HBM ring
    rng = np.random.default_rng(seed)
    delta = np.empty((count, 2 * n), np.uint8)
    for i in range(count):
        a = rng.permutation(n)
        b = rng.permutation(n)
        p, q = rng.choice(n, 2, replace=False)
        b[p] = b[q]
        delta[i, :n] = a
        delta[i, n:] = b
    return delta


def _outputs(ctx, count):
    return ctx.alloc(4 * count), ctx.alloc(count), ctx.alloc(8 * count)


def _run_both(ctx, n, delta, bound=-1):
    """(batched results, allocation-free results), each (L, S, E, violation)."""
    count = len(delta)
    d = ctx.to_device(np.ascontiguousarray(delta).reshape(-1))
    bufs = [d]
    out = []
    try:
        for fn in (ctx.solve_device, ctx.solve_device_noalloc):
            L, S, E = _outputs(ctx, count)
            bufs += [L, S, E]
            b = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d, length=L,
                          sync=S, expanded=E, witness=None, bound=bound, reserved=0)
            v = fn(b)
            out.append((ctx.to_host(L, np.int32, count), ctx.to_host(S, np.uint8, count),
                        ctx.to_host(E, np.int64, count), v))
    finally:
        for p in bufs:
            ctx.free(p)
    return out


def _same(a, b, tag):
    for k, name in enumerate(("length", "sync", "expanded")):
        np.testing.assert_array_equal(b[k], a[k], err_msg=f"{tag}: {name}")
    assert a[3] == b[3], f"{tag}: violation {a[3]} vs {b[3]}"


def test_noalloc_equals_batched_two_calls_shrinking(ctx):
    # call 1: 300k n = 16 automata with Cerny C_16 stragglers and automata
    # whose levels outgrow every LDS ring
    n, delta, pos = W.uniform_n16(seed=77, count=300_000)
    wide = _perm_merge(5, 48)
    delta = np.concatenate([delta, wide])
    want, got = _run_both(ctx, n, delta)
    _same(want, got, "call 1")
    assert np.all(got[0][pos] == (n - 1) ** 2)
    # the wide ones against the oracle too
    Lo, So, Eo, _ = O.slim(n, wide[:, :n].astype(np.int32), wide[:, n:].astype(np.int32), threads=8)
    np.testing.assert_array_equal(got[0][-48:], Lo)
    np.testing.assert_array_equal(got[2][-48:], Eo)
    # call 2, smaller: n = 13 with C_13 stragglers
    n2, delta2, pos2 = W.uniform_n16(seed=78, count=20_000, n=13, cerny_every=100)
    want2, got2 = _run_both(ctx, n2, delta2)
    _same(want2, got2, "call 2")
    assert np.all(got2[0][pos2] == (n2 - 1) ** 2)


def test_noalloc_every_small_n(ctx):
    # n = 1 .. 16, one call each, batches shrinking from 4000 to 250
    for n in range(16, 0, -1):
        count = 250 * n
        A, B = ssw.generate_uniform(900 + n, n, count)
        delta = ssw.pack_uniform(A, B)
        if n >= 3:
            ca, cb = ssw.generate_cerny(n, 1)
            delta[::97, :n] = ca
            delta[::97, n:] = cb
        want, got = _run_both(ctx, n, delta)
        _same(want, got, f"n={n}")
        L, S, E, _ = O.slim(n, delta[:, :n].astype(np.int32), delta[:, n:].astype(np.int32), threads=8)
        np.testing.assert_array_equal(got[0], L, err_msg=f"n={n} vs oracle")
        np.testing.assert_array_equal(got[2], E, err_msg=f"n={n} vs oracle")


def test_noalloc_violation_index(ctx):
    n, delta, pos = W.uniform_n16(seed=79, count=50_000)
    want, got = _run_both(ctx, n, delta, bound=12)
    _same(want, got, "bound 12")
    assert got[3] == int(np.nonzero(got[0] > 12)[0][0])


def test_noalloc_refuses_what_it_cannot_hold(ctx):
    d = ctx.alloc(64)
    L = ctx.alloc(64)
    try:
        def call(**kw):
            f = dict(count=1, uniform_n=8, witness_stride=0, n_per=None, offset=None, delta=d, length=L, sync=None,
                     expanded=None, witness=None, bound=-1, reserved=0)
            f.update(kw)
            with pytest.raises(ssw.SswError) as e:
                ctx.solve_device_noalloc(ssw.Batch(**f))
            assert e.value.code == ssw.SSW_ERR_ARG
            return str(e.value)
        assert "at most 1048576" in call(count=(1 << 20) + 1, uniform_n=16)
        call(uniform_n=17)
        call(uniform_n=0)
        call(n_per=d)
        call(witness=L, witness_stride=32)
    finally:
        ctx.free(d)
        ctx.free(L)
