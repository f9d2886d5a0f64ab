"""Parity at the benchmark configurations' full sizes.

The 1,000,000-automaton n = 16 batch (1 % Cerny C_16), the 200,000-automaton
batch of uneven sizes (n = 12..24) and the large instances (C_24..C_30 plus
64 random n = 30 automata) are each solved whole on the GPU and compared,
every automaton, with the CPU oracle (the SlimCPUUntweaked restatement, and
its uint32 twin for n > 16): length, flag and expanded count, bit for bit.
Beside that, properties that do not depend on size:

* Cerny closed forms: C_n has length (n-1)^2 (SimpleAutomataTests.cs:14-35)
  and the search expands every subset but the n singletons and Q's last
  level, E = 2^n - n - 1;
* repeatability: a second solve returns identical arrays;
* the batched path against the plain one-automaton-per-launch path on a
  sample.
"""
import os

import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu


THREADS = max(1, min(16, os.cpu_count() or 1))


def _sample(count, k, seed):
    return np.sort(np.random.default_rng(seed).choice(count, size=min(k, count), replace=False))


@pytest.fixture(scope="module")
def n16(ctx):
    n, delta, pos = W.uniform_n16()
    r = ctx.solve_packed(delta.reshape(-1), len(delta), n)
    return n, delta, pos, r


def test_n16_cerny_closed_forms(n16):
    n, delta, pos, r = n16
    assert len(pos) == 10_000
    assert np.all(r.length[pos] == (n - 1) ** 2)
    assert np.all(r.sync[pos] == 1)
    assert np.all(r.expanded[pos] == 2 ** n - n - 1)


def test_n16_lengths_in_range(n16):
    n, delta, pos, r = n16
    L = r.length
    assert np.all((L == -1) | ((L >= 1) & (L <= (n - 1) ** 2)))
    np.testing.assert_array_equal(r.sync, (L >= 0).astype(np.uint8))
    assert np.all(r.expanded >= 1)


def test_n16_matches_oracle_whole(n16):
    # every one of the 1,000,000 automata, the 10,000 C_16 included
    n, delta, pos, r = n16
    A = np.ascontiguousarray(delta[:, :n]).astype(np.int32)
    B = np.ascontiguousarray(delta[:, n:]).astype(np.int32)
    L, S, E, _ = O.slim(n, A, B, threads=THREADS)
    bad = np.nonzero((r.length != L) | (r.sync != S) | (r.expanded != E))[0]
    assert bad.size == 0, f"{bad.size} automata differ, first {bad[:10].tolist()}"
    assert np.array_equal(r.length, L) and np.array_equal(r.sync, S) and np.array_equal(r.expanded, E)


def test_n16_repeatable(ctx, n16):
    n, delta, pos, r = n16
    r2 = ctx.solve_packed(delta.reshape(-1), len(delta), n)
    np.testing.assert_array_equal(r2.length, r.length)
    np.testing.assert_array_equal(r2.sync, r.sync)
    np.testing.assert_array_equal(r2.expanded, r.expanded)


def test_n16_batch_equals_plain_on_sample(ctx, n16):
    n, delta, pos, r = n16
    idx = np.union1d(_sample(len(delta), 600, 17), pos[:8])
    sub = np.ascontiguousarray(delta[idx])
    p = ctx.solve_packed(sub.reshape(-1), len(idx), n, mode="plain")
    np.testing.assert_array_equal(p.length, r.length[idx])
    np.testing.assert_array_equal(p.expanded, r.expanded[idx])


def test_uneven_matches_oracle_whole(ctx):
    # every one of the 200,000 automata of uneven sizes
    n_per, offs, delta = W.uneven_12_24()
    count = len(n_per)
    r = ctx.solve_packed(delta, count, 0, n_per, offs)
    r2 = ctx.solve_packed(delta, count, 0, n_per, offs)
    np.testing.assert_array_equal(r2.length, r.length)
    np.testing.assert_array_equal(r2.expanded, r.expanded)
    L = r.length
    assert np.all((L == -1) | ((L >= 1) & (L <= (n_per.astype(np.int32) - 1) ** 2)))
    Lo = np.empty(count, np.int32)
    So = np.empty(count, np.uint8)
    Eo = np.empty(count, np.int64)
    for n in np.unique(n_per):
        sel = np.nonzero(n_per == n)[0]
        rows = offs[sel].astype(np.int64)[:, None] + np.arange(int(n))[None, :]
        Lo[sel], So[sel], Eo[sel], _ = O.solve(int(n), delta[rows].astype(np.int32),
                                               delta[rows + int(n)].astype(np.int32), threads=THREADS)
    bad = np.nonzero((r.length != Lo) | (r.sync != So) | (r.expanded != Eo))[0]
    assert bad.size == 0, f"{bad.size} automata differ, first {bad[:10].tolist()} (n {n_per[bad[:10]].tolist()})"


def test_large_instances(ctx):
    items = W.large()
    n_per = np.array([it[0] for it in items], np.uint8)
    r = ctx.solve_mixed(n_per, [it[1] for it in items], [it[2] for it in items], mode="large")
    for i, (n, a, b, kind) in enumerate(items):
        if kind == "cerny":
            assert r.length[i] == (n - 1) ** 2, f"C_{n}"
            assert r.expanded[i] == 2 ** n - n - 1, f"C_{n}"
    rnd = [i for i, it in enumerate(items) if it[3] == "random"]
    n = items[rnd[0]][0]
    A = np.stack([items[i][1] for i in rnd]).astype(np.int32)
    B = np.stack([items[i][2] for i in rnd]).astype(np.int32)
    L, S, E, _ = O.solve(n, A, B, threads=8)
    np.testing.assert_array_equal(r.length[rnd], L)
    np.testing.assert_array_equal(r.sync[rnd], S)
    np.testing.assert_array_equal(r.expanded[rnd], E)
