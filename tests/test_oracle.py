"""Pins the CPU oracle (oracle/ssw_oracle.c) -- CPU only.

The reference holds one family of known answers: Cerny C_n is synchronizing
with length (n-1)^2 (AlgorithmCorrectnessTestProject/SimpleAutomataTests.cs:14-35,
IntegrityTest.cs:44-62).  Everything else is pinned here independently of the
restated search:
  * brute force over words in lexicographic order (length and witness),
  * Cerny's pair criterion (the synchronizing flag),
  * invariants: relabelling of states, swapping the letters, a constant
    letter, two permutations, the Cerny bound,
  * the faithful uint16 search (SlimCPUUntweaked) against the widened one,
  * the committed golden fixtures (regression).
"""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def cerny(n, conv):
    merge = [1 if q == 0 else q for q in range(n)]
    cycle = [(q + 1) % n for q in range(n)]
    return (merge, cycle) if conv == 0 else (cycle, merge)


def cerny_word(n, conv):
    w = ([1] + [0] * (n - 1)) * (n - 2) + [1]
    return w if conv == 1 else [1 - x for x in w]


def rand_automata(rng, n, count):
    return rng.integers(0, n, size=(count, n)), rng.integers(0, n, size=(count, n))


@pytest.mark.parametrize("conv", [0, 1])
def test_cerny_known_answers(conv):
    # SimpleAutomataTests.CheckWorstCase for n = 3..13, extended to n = 2..20
    for n in range(2, 21):
        a, b = cerny(n, conv)
        solver = O.slim if n <= 16 else O.wide
        L, S, E, _ = solver(n, np.array([a]), np.array([b]))
        assert (L[0], S[0]) == ((n - 1) ** 2, 1)
        assert E[0] == 2 ** n - n - 1  # every subset with >= 2 states lies closer than L
        if n <= 16:
            Lw, w = O.witness(n, a, b)
            assert Lw == (n - 1) ** 2
            assert list(w) == cerny_word(n, conv)  # the unique shortest reset word
            assert O.verify_word(n, a, b, w)


def test_brute_force_agrees_on_tiny_automata():
    rng = np.random.default_rng(11)
    for n in range(1, 7):
        A, B = rand_automata(rng, n, 300 if n < 6 else 120)
        L, S, E, _ = O.slim(n, A, B)
        for i in range(len(A)):
            maxlen = (n - 1) ** 2 if L[i] < 0 else L[i]
            Lb, wb = O.brute_force(n, A[i], B[i], maxlen)
            if L[i] < 0:
                # no word up to the Cerny bound synchronizes (n <= 6: the bound holds)
                assert Lb == -1 or n == 1
            else:
                assert Lb == L[i], (n, i, A[i], B[i])
                Lw, w = O.witness(n, A[i], B[i])
                assert Lw == L[i]
                np.testing.assert_array_equal(w, wb)  # lexicographically smallest


def test_pair_criterion_agrees_with_flags():
    rng = np.random.default_rng(12)
    for n in range(2, 15):
        A, B = rand_automata(rng, n, 200)
        L, S, E, _ = O.slim(n, A, B)
        for i in range(len(A)):
            assert bool(S[i]) == O.pair_sync(n, A[i], B[i]), (n, i)


def test_slim_equals_wide():
    rng = np.random.default_rng(13)
    for n in (3, 8, 12, 16):
        A, B = rand_automata(rng, n, 2000)
        r1 = O.slim(n, A, B)
        r2 = O.wide(n, A, B)
        for x, y in zip(r1[:3], r2[:3]):
            np.testing.assert_array_equal(x, y)


def test_parallel_chunks_equal_serial():
    # Parallel.For over contiguous chunks (SlimCPUUntweaked.cs:33-46)
    rng = np.random.default_rng(14)
    A, B = rand_automata(rng, 12, 3001)
    r1 = O.slim(12, A, B, threads=1)
    r7 = O.slim(12, A, B, threads=7)
    for x, y in zip(r1[:3], r7[:3]):
        np.testing.assert_array_equal(x, y)


def test_invariants():
    rng = np.random.default_rng(15)
    for n in (5, 9, 13):
        A, B = rand_automata(rng, n, 300)
        L, S, E, _ = O.slim(n, A, B)
        assert np.all(L <= (n - 1) ** 2)  # the Cerny bound (Verify.cs:60-62) holds for random automata
        # relabelling the states by a permutation leaves length and E unchanged
        perm = rng.permutation(n)
        inv = np.argsort(perm)
        Ap = perm[A[:, inv]]
        Bp = perm[B[:, inv]]
        Lp, Sp, Ep, _ = O.slim(n, Ap, Bp)
        np.testing.assert_array_equal(L, Lp)
        np.testing.assert_array_equal(E, Ep)
        # swapping the letters leaves length and E unchanged
        Ls, Ss, Es, _ = O.slim(n, B, A)
        np.testing.assert_array_equal(L, Ls)
        np.testing.assert_array_equal(E, Es)
    # a constant letter synchronizes in one step; two permutations never do
    for n in range(2, 12):
        const = np.zeros((1, n), int)
        ident = np.arange(n)[None]
        assert O.slim(n, const, ident)[0][0] == 1
        rot = np.roll(np.arange(n), 1)[None]
        assert O.slim(n, rot, ident)[0][0] == -1


def test_n1_matches_the_reference():
    # Q itself is the only singleton and is already visited: the reference
    # search (SlimCPUUntweaked.cs:149-151) reports "not synchronizing"
    L, S, E, _ = O.slim(1, np.zeros((1, 1), int), np.zeros((1, 1), int))
    assert (L[0], S[0], E[0]) == (-1, 0, 1)


def test_violation_index():
    # Verify returns the first automaton whose length exceeds the bound
    rng = np.random.default_rng(16)
    A, B = rand_automata(rng, 10, 400)
    L, S, E, v = O.slim(10, A, B, bound=8)
    assert v == int(np.nonzero(L > 8)[0][0])
    assert O.slim(10, A, B)[3] == -1
    assert O.slim(10, A, B, bound=8, threads=5)[3] == v


def test_verify_word_rejects_wrong_words():
    # a planted mistake in a witness is caught: flip each letter in turn
    a, b = cerny(6, 1)
    w = np.array(cerny_word(6, 1), np.uint8)
    assert O.verify_word(6, a, b, w)
    assert not O.verify_word(6, a, b, w[:-1])
    for k in range(len(w)):
        bad = w.copy()
        bad[k] ^= 1
        assert not O.verify_word(6, a, b, bad)


def test_golden_fixtures():
    g = np.load(os.path.join(GOLDEN, "plain_n10.npz"))
    A, B = g["a"].astype(np.int32), g["b"].astype(np.int32)
    L, S, E, _ = O.slim(10, A, B, threads=8)
    np.testing.assert_array_equal(L, g["length"])
    np.testing.assert_array_equal(E, g["expanded"])
    for i in range(0, len(A), 37):
        if L[i] >= 0:
            w = np.unpackbits(g["witness"][i], bitorder="little")[:L[i]]
            assert O.verify_word(10, A[i], B[i], w)
            np.testing.assert_array_equal(w, O.witness(10, A[i], B[i])[1])
    g = np.load(os.path.join(GOLDEN, "uniform_n16_slice.npz"))
    A, B = g["a"].astype(np.int32), g["b"].astype(np.int32)
    L, S, E, _ = O.slim(16, A, B, threads=8)
    np.testing.assert_array_equal(L, g["length"])
    np.testing.assert_array_equal(E, g["expanded"])
    assert np.all(L[g["cerny_positions"]] == 225)
