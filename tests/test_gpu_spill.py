"""GPU parity of the n <= 16 path's launch configurations against the CPU oracle
(restatement of SlimCPUUntweaked.cs:51-222): ring spill to a second launch,
groups of one or several wavefronts, frontiers that move between the LDS ring
and the HBM ring, lane-entry counts 1/2/4, nibble and byte image tables.

Every configuration must give the same lengths, flags and canonical expanded
counts, bit for bit.  The batch mixes Cerny C_16 in both letter conventions
(wide frontiers, 225 levels), the widest random automata of a larger sample
(frontiers of several hundred subsets) and ordinary random automata.
"""
import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu

N = 16


def _batch():
    _, delta, _ = W.uniform_n16(count=40_000, cerny_every=0)
    A = delta[:, :N].astype(np.int32)
    B = delta[:, N:].astype(np.int32)
    L, S, E, _ = O.solve(N, A, B, threads=8)
    wide = np.argsort(E)[-300:]             # the random automata with the most subsets
    plain = np.arange(3000)
    rows = [A[plain], A[wide]]
    cols = [B[plain], B[wide]]
    for conv in (0, 1):
        a, b = ssw.generate_cerny(N, conv)
        rows.append(np.tile(a, (40, 1)))
        cols.append(np.tile(b, (40, 1)))
    A2 = np.concatenate(rows)
    B2 = np.concatenate(cols)
    perm = np.random.default_rng(11).permutation(len(A2))
    return A2[perm], B2[perm]


@pytest.fixture(scope="module")
def case():
    A, B = _batch()
    L, S, E, _ = O.solve(N, A, B, threads=8)
    return A, B, L, S, E


# (group_lanes, block_threads, chunk, arena_lds, lane_entries, spill_arena_lds, spill_lane_entries,
#  spill_group_lanes[, narrow_bits])
CONFIGS = [
    (64, 128, 16, 512, 1, 4096, 4, 0),      # the committed n = 16 configuration
    (64, 0, 4, 256, 1, 4096, 2, 0),
    (64, 192, 8, 512, 2, 2048, 4, 0),       # spill ring too small for C_16: HBM ring in the second launch
    (64, 128, 16, 512, 1, 4096, 4, 128),    # two wavefronts per automaton in the second launch
    (64, 128, 16, 512, 1, 4096, 1, 256),
    (64, 128, 16, 512, 1, 1024, 4, 256),    # wide levels relocate to the HBM ring and back
    (32, 0, 8, 512, 1, 4096, 4, 0),         # two automata per wavefront in the first launch
    (256, 256, 2, 512, 2, 4096, 4, 0),      # four wavefronts per automaton, spilling from the first launch
    (256, 256, 2, 1024, 4, 0, 0, 0),        # no spill: level-split groups with HBM relocation
    (64, 0, 2, 2048, 4, 0, 0, 0),           # no spill: one wavefront, frontier moves to HBM
    (64, 128, 16, 512, 1, 4096, 4, 0, 15),  # relabelled states, 2^15-bit bitmap, three launches
    (64, 128, 16, 512, 1, 4096, 8, 0, 15),  # eight entries per lane in the last launch
    (64, 0, 4, 256, 1, 2048, 8, 0),         # eight entries per lane, ring too small for C_16's widest steps
    (64, 128, 16, 512, 1, 4096, 4, 0, 13),
    (32, 0, 8, 512, 1, 4096, 4, 0, 14),
    (256, 256, 2, 1024, 2, 4096, 4, 0, 14), # relabelled, four wavefronts per automaton
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=[",".join(map(str, c)) for c in CONFIGS])
def test_configuration_parity(ctx, case, cfg):
    A, B, L, S, E = case
    saved = ctx.get_tuning(N)
    try:
        ctx.set_tuning(N, *cfg)
        print(f"config {cfg}", flush=True)
        r = ctx.compute(N, A, B)
    finally:
        ctx.set_tuning(N, **saved)
    bad = np.nonzero(r.length != L)[0]
    assert len(bad) == 0, f"{cfg}: length mismatch at {bad[:10]} gpu={r.length[bad[:10]]} cpu={L[bad[:10]]}"
    np.testing.assert_array_equal(r.sync, S, err_msg=f"{cfg}: sync")
    np.testing.assert_array_equal(r.expanded, E, err_msg=f"{cfg}: expanded")


def test_spill_mixed_sizes(ctx):
    """Uneven sizes with spill on for n = 12..16 (relabelled states with an
    (n-1)-bit bitmap for odd n): every class spills into its own lists and
    later launches, concurrently with the others."""
    rng = np.random.default_rng(23)
    ns = rng.integers(12, 17, size=4000).astype(np.uint8)
    As, Bs = [], []
    for i, n in enumerate(ns):
        if i % 97 == 0:
            a, b = ssw.generate_cerny(int(n), i % 2)
        else:
            a = rng.integers(0, n, size=n)
            b = rng.integers(0, n, size=n)
        As.append(np.asarray(a, np.int32))
        Bs.append(np.asarray(b, np.int32))
    saved = {n: ctx.get_tuning(n) for n in range(12, 17)}
    try:
        for n in range(12, 17):
            ctx.set_tuning(n, 64, 0, 8, 256, 1, 4096, 4, 0, n - 1 if n % 2 else 0)
        r = ctx.solve_mixed(ns, As, Bs)
    finally:
        for n, t in saved.items():
            ctx.set_tuning(n, **t)
    for n in range(12, 17):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([As[i] for i in idx])
        B = np.stack([Bs[i] for i in idx])
        L, S, E, _ = O.solve(n, A, B, threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")


def test_relabelled_large_n(ctx):
    """n = 17..24 through the relabelled n <= 16 solver: automata with
    |a(Q) | b(Q)| <= 16 are solved there (root level expanded at start), the
    others (Cerny C_n, wide random ones) by the HBM-bitmap solver from the
    spill queue.  Checked against the oracle's uint32 search."""
    rng = np.random.default_rng(31)
    ns = rng.integers(17, 25, size=1500).astype(np.uint8)
    As, Bs = [], []
    for i, n in enumerate(ns):
        if i % 151 == 0:
            a, b = ssw.generate_cerny(int(n), i % 2)
        elif i % 7 == 0:
            # image of Q under a and b inside 12 states: |a(Q) | b(Q)| <= 12
            a = rng.integers(0, 12, size=n)
            b = rng.integers(0, 12, size=n)
        else:
            a = rng.integers(0, n, size=n)
            b = rng.integers(0, n, size=n)
        As.append(np.asarray(a, np.int32))
        Bs.append(np.asarray(b, np.int32))
    saved = {n: ctx.get_tuning(n) for n in range(16, 25)}
    try:
        for n in range(17, 25):
            t = dict(saved[n])
            t["narrow_bits"] = 16
            ctx.set_tuning(n, **t)
        r = ctx.solve_mixed(ns, As, Bs)
    finally:
        for n, t in saved.items():
            ctx.set_tuning(n, **t)
    for n in range(17, 25):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([As[i] for i in idx])
        B = np.stack([Bs[i] for i in idx])
        L, S, E, _ = O.solve(n, A, B, threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.sync[idx], S, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")


def _family_cases(rng):
    """Automata of n = 17..24 that probe the two-family encodings: a(Q) and
    b(Q) overlapping in few or many states, b(Q) inside a(Q), a constant
    letter whose state lies outside a(Q) (a family-B singleton at the root),
    |a(Q)| = 15 (the n <= 16 solver's 15-state families), 16..18 (handed on to
    the general solver's 18-state families), 19 and Cerny C_n for n > 18 (its
    HBM bitmap); Cerny C_17, C_18 have a permutation letter inside 18 states."""
    cases = []
    for n in range(17, 25):
        for k in range(40):
            kind = k % 8
            if kind == 0:     # a into 0..14, b into n-15..n-1: small overlap
                a = rng.integers(0, 15, size=n)
                b = rng.integers(n - 15, n, size=n)
            elif kind == 1:   # large overlap
                a = rng.integers(0, 15, size=n)
                b = rng.integers(2, 17, size=n)
            elif kind == 2:   # b(Q) inside a(Q)
                a = rng.integers(0, 15, size=n)
                b = rng.integers(0, 15, size=n) % 8
            elif kind == 3:   # b constant outside a(Q): length 1
                a = rng.integers(0, 10, size=n)
                b = np.full(n, n - 1)
            elif kind == 4:   # exactly 15 states in a(Q)
                a = np.concatenate([np.arange(15), rng.integers(0, 15, size=n - 15)])
                b = rng.integers(0, n, size=n)
            elif kind == 5:   # 16..19 states in a(Q): handed on (up to 18 encoded there)
                m = min(n, 16 + (k // 8) % 4)
                a = np.concatenate([np.arange(m), rng.integers(0, m, size=n - m)])
                b = rng.integers(0, n, size=n)
            elif kind == 6:
                a, b = ssw.generate_cerny(n, k % 2)
            else:             # plain random
                a = rng.integers(0, n, size=n)
                b = rng.integers(0, n, size=n)
            perm = rng.permutation(n)  # relabel the states: any label may hold any role
            inv = np.argsort(perm)
            a = perm[np.asarray(a)[inv]]
            b = perm[np.asarray(b)[inv]]
            cases.append((n, np.asarray(a, np.int32), np.asarray(b, np.int32)))
    return cases


def test_two_family_encoding(ctx):
    """n = 17..24 in the two-family encodings (subsets of a(Q) or of b(Q),
    a(Q) & b(Q) labelled alike in both): the n <= 16 solver's (15 states per
    family) and, for what it hands on, the general solver's (18 states, LDS
    bitmap; HBM bitmap past that).  Lengths, flags and expanded counts against
    the oracle's uint32 search, and most of the probe automata finished by
    the first encoded launch."""
    rng = np.random.default_rng(57)
    cases = _family_cases(rng)
    ns = np.array([c[0] for c in cases], np.uint8)
    As = [c[1] for c in cases]
    Bs = [c[2] for c in cases]
    saved = {n: ctx.get_tuning(n) for n in range(17, 25)}
    try:
        for n in range(17, 25):
            t = dict(saved[n])
            t["narrow_bits"] = 16
            ctx.set_tuning(n, **t)
        r = ctx.solve_mixed(ns, As, Bs)
        stats = ctx.last_launch_stats()
    finally:
        for n, t in saved.items():
            ctx.set_tuning(n, **t)
    for n in range(17, 25):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([As[i] for i in idx])
        B = np.stack([Bs[i] for i in idx])
        L, S, E, _ = O.solve(n, A, B, threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.sync[idx], S, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")
    fam = [x for x in stats if x["kernel"].endswith(",false,true>")]
    assert fam, [x["kernel"] for x in stats]
    # kinds 0..4 and 7 (6 of every 8) fit the encoding; a few may outgrow the ring
    assert sum(x["automata"] for x in fam) >= 0.6 * len(cases), [(x["n"], x["automata"]) for x in fam]


def test_general_pass_keeps_n20_on_the_batched_path(ctx):
    """What the two-family launch hands on runs in the general solver with
    four entries per lane; its HBM arena holds every subset of n <= 20 with
    that step's margin, so Cerny C_20 (2^20 - 21 subsets) finishes there and
    nothing is deferred to the device-wide path.  Lengths and expanded counts
    by the closed forms and the oracle."""
    rng = np.random.default_rng(77)
    ns, As, Bs = [], [], []
    for conv in (0, 1):
        a, b = ssw.generate_cerny(20, conv)
        ns.append(20)
        As.append(np.asarray(a, np.int32))
        Bs.append(np.asarray(b, np.int32))
    for n in (17, 18, 19, 20) * 30:
        ns.append(n)
        As.append(rng.integers(0, n, size=n).astype(np.int32))
        Bs.append(rng.integers(0, n, size=n).astype(np.int32))
    before = ctx.progress()["deferred"]
    r = ctx.solve_mixed(np.array(ns, np.uint8), As, Bs)
    assert ctx.progress()["deferred"] == before
    assert int(r.length[0]) == int(r.length[1]) == 19 * 19
    assert int(r.expanded[0]) == int(r.expanded[1]) == 2 ** 20 - 21
    for n in (17, 18, 19, 20):
        idx = [i for i in range(2, len(ns)) if ns[i] == n]
        L, S, E, _ = O.solve(n, np.stack([As[i] for i in idx]), np.stack([Bs[i] for i in idx]), threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")


WIT_CONFIGS = [
    (64, 128, 16, 512, 1, 4096, 4, 0),       # the committed spill chain
    (64, 0, 4, 256, 1, 2048, 4, 0),          # last launch's ring too small for C_16: HBM ring
    (64, 0, 2, 2048, 1, 0, 0, 0),            # no spill: frontiers move to the HBM ring
    (64, 128, 16, 512, 1, 4096, 4, 0, 15),   # relabelled states, three launches
    (64, 128, 16, 512, 1, 4096, 4, 0, 13),
    (64, 0, 4, 256, 1, 512, 4, 0),           # small last-launch ring: C_16's good lists outgrow it
]


@pytest.fixture(scope="module")
def words(case):
    A, B, L, S, E = case
    return [O.witness(N, A[i], B[i])[1] if L[i] >= 0 else None for i in range(len(A))]


@pytest.mark.parametrize("cfg", WIT_CONFIGS, ids=[",".join(map(str, c)) for c in WIT_CONFIGS])
def test_witness_words_every_configuration(ctx, case, words, cfg):
    """Witness words from the n <= 16 path (recorded levels, good marks in the
    LDS bitmap) through every route a search can take: first launch, spill
    chain, HBM ring, good-subset lists in LDS and, where they outgrow the
    ring area, the choice bitmap."""
    A, B, L, S, E = case
    saved = ctx.get_tuning(N)
    try:
        ctx.set_tuning(N, *cfg)
        r = ctx.compute(N, A, B, witness=True)
    finally:
        ctx.set_tuning(N, **saved)
    np.testing.assert_array_equal(r.length, L, err_msg=f"{cfg}: length")
    np.testing.assert_array_equal(r.expanded, E, err_msg=f"{cfg}: expanded")
    for i in range(len(A)):
        if L[i] >= 0:
            w = ssw.witness_bits_to_word(r.witness[i], int(L[i]))
            np.testing.assert_array_equal(w, words[i], err_msg=f"{cfg}: automaton {i}")
        else:
            assert not r.witness[i].any()


def test_witness_words_mixed_sizes(ctx):
    """One call over n = 10..20 with witness words: the n <= 16 classes run
    the small kernels' witness passes, the others the general kernel, on
    their own streams at the same time."""
    rng = np.random.default_rng(29)
    ns = rng.integers(10, 21, size=3000).astype(np.uint8)
    As, Bs = [], []
    for i, n in enumerate(ns):
        if i % 101 == 0:
            a, b = ssw.generate_cerny(int(n), 1)
        else:
            a = rng.integers(0, n, size=n)
            b = rng.integers(0, n, size=n)
        As.append(np.asarray(a, np.int32))
        Bs.append(np.asarray(b, np.int32))
    r = ctx.solve_mixed(ns, As, Bs, witness=True)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([As[i] for i in idx])
        B = np.stack([Bs[i] for i in idx])
        L, S, E, _ = O.solve(int(n), A, B, threads=8)
        np.testing.assert_array_equal(r.length[idx], L, err_msg=f"n={n}")
        np.testing.assert_array_equal(r.expanded[idx], E, err_msg=f"n={n}")
        for k, i in enumerate(idx):
            if L[k] >= 0:
                w = ssw.witness_bits_to_word(r.witness[i], int(L[k]))
                np.testing.assert_array_equal(w, O.witness(int(n), A[k], B[k])[1], err_msg=f"n={n} i={i}")
