"""GPU parity of the device-wide path (large automata, n up to 30) and of the
hand-off from the batched kernel to it, against the CPU oracle and the Cerny
closed forms (SimpleAutomataTests.cs:14-35: C_n has length (n-1)^2)."""
import numpy as np
import pytest

import oracle as O
import ssw

pytestmark = pytest.mark.gpu


def _cerny_word(n, conv):
    word = ([1] + [0] * (n - 1)) * (n - 2) + [1]
    return word if conv == 1 else [1 - x for x in word]


@pytest.mark.parametrize("n", [3, 8, 17, 20])
def test_large_cerny_with_witness(ctx, n):
    a, b = ssw.generate_cerny(n, 1)
    r = ctx.solve_uniform(n, a[None], b[None], witness=True, mode="large")
    assert r.length[0] == (n - 1) ** 2
    assert r.expanded[0] == 2 ** n - n - 1
    np.testing.assert_array_equal(ssw.witness_bits_to_word(r.witness[0], (n - 1) ** 2), _cerny_word(n, 1))


@pytest.mark.parametrize("n", [12, 20, 24])
def test_large_random_matches_oracle(ctx, n):
    A, B = ssw.generate_uniform(4000 + n, n, 48 if n < 24 else 12)
    r = ctx.solve_uniform(n, A, B, witness=True, mode="large")
    L, S, E, _ = O.solve(n, A, B, threads=8)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    for i in range(len(A)):
        if L[i] >= 0:
            w = ssw.witness_bits_to_word(r.witness[i], int(L[i]))
            assert O.verify_word(n, A[i], B[i], w)
            Lw, wo = O.witness(n, A[i], B[i])
            np.testing.assert_array_equal(w, wo)


def test_large_mixed_sizes_no_witness(ctx):
    ns, As, Bs = [], [], []
    for i, n in enumerate([9, 14, 21, 25, 18, 23, 26, 16]):
        a, b = ssw.generate_uniform(55, n, 1, first=i)
        ns.append(n)
        As.append(a[0])
        Bs.append(b[0])
    a, b = ssw.generate_cerny(22, 0)
    ns.append(22)
    As.append(a)
    Bs.append(b)
    r = ctx.solve_mixed(ns, As, Bs, mode="large")
    for i, n in enumerate(ns):
        L, S, E, _ = O.solve(n, As[i][None], Bs[i][None])
        assert (r.length[i], r.sync[i], r.expanded[i]) == (L[0], S[0], E[0]), f"automaton {i} n={n}"


def test_batch_defers_long_frontiers(ctx):
    # C_22 has 2^22 - 23 expanded subsets, beyond a group's HBM arena: the
    # batched launch hands it to the device-wide path mid-batch
    n = 22
    A, B = ssw.generate_uniform(99, n, 300)
    a, b = ssw.generate_cerny(n, 1)
    A[137], B[137] = a, b
    r = ctx.compute(n, A, B)
    assert r.length[137] == (n - 1) ** 2
    L, S, E, _ = O.solve(n, A, B, threads=8)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.expanded, E)
    assert ctx.progress()["deferred"] >= 1


def test_batch_defers_with_witness(ctx):
    n = 21
    A, B = ssw.generate_uniform(98, n, 64)
    a, b = ssw.generate_cerny(n, 1)
    A[5], B[5] = a, b
    r = ctx.compute(n, A, B, witness=True)
    assert r.length[5] == (n - 1) ** 2
    np.testing.assert_array_equal(ssw.witness_bits_to_word(r.witness[5], (n - 1) ** 2), _cerny_word(n, 1))
    L, S, E, _ = O.solve(n, A, B, threads=8)
    np.testing.assert_array_equal(r.length, L)
    for i in range(len(A)):
        if L[i] >= 0:
            assert O.verify_word(n, A[i], B[i], ssw.witness_bits_to_word(r.witness[i], int(L[i])))


def _two_closed_halves(h, seed):
    # two closed halves of h states each (C_h on one, a random automaton on
    # the other): never synchronizing, and the pairs of subsets one word
    # reaches in the two halves make a wide reachable family
    a1, b1 = ssw.generate_cerny(h, 1)
    A2, B2 = ssw.generate_uniform(seed, h, 1)
    a = np.concatenate([a1, A2[0] + h]).astype(np.int32)
    b = np.concatenate([b1, B2[0] + h]).astype(np.int32)
    return a, b


@pytest.mark.parametrize("mode", ["batch", "plain", "large"])
@pytest.mark.parametrize("witness", [False, True])
@pytest.mark.parametrize("h,seed,least", [(8, 7, 30_000), (10, 36, 100_000)])
def test_non_synchronizing_with_a_wide_family(ctx, mode, witness, h, seed, least):
    # n = 16 runs in the small-automaton kernels (ring spill chain), n = 20
    # in the HBM-bitmap ones and the device-wide path
    n = 2 * h
    a, b = _two_closed_halves(h, seed)
    c, d = ssw.generate_cerny(n, 1)
    A = np.stack([a, c, a])
    B = np.stack([b, d, b])
    L, S, E, _ = O.solve(n, A, B, threads=4)
    assert L[0] == -1 and E[0] > least
    r = ctx.solve_uniform(n, A, B, witness=witness, mode=mode)
    np.testing.assert_array_equal(r.length, L)
    np.testing.assert_array_equal(r.sync, S)
    np.testing.assert_array_equal(r.expanded, E)
    if witness:
        assert not r.witness[0].any() and not r.witness[2].any()


@pytest.mark.parametrize("n", [25, 26])
def test_large_cerny_witness_beyond_the_oracle(ctx, n):
    # n > 24 is past the oracle's witness restatement: the closed forms pin it
    # (length (n-1)^2, E = 2^n - n - 1, the word (b a^(n-1))^(n-2) b)
    a, b = ssw.generate_cerny(n, 1)
    r = ctx.solve_uniform(n, a[None], b[None], witness=True, mode="large")
    assert r.length[0] == (n - 1) ** 2
    assert r.expanded[0] == 2 ** n - n - 1
    w = ssw.witness_bits_to_word(r.witness[0], (n - 1) ** 2)
    np.testing.assert_array_equal(w, _cerny_word(n, 1))
    assert O.verify_word(n, a, b, w)
