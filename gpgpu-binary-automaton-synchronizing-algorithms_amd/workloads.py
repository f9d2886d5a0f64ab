"""Synthetic workloads of the benchmark configurations (all integer, seeded).

Transitions come from the library's counter-based splitmix64 generator
(ssw_generate_uniform), so any automaton can be regenerated from
(seed, index).  Cerny automata use the letter convention a = cycle
(q -> q+1 mod n), b = merge (0 -> 1), i.e. convention 1 of
ssw_generate_cerny; the reference's Problem.GenerateWorstCase
(GPGPU/SharedDefinitions/Problem.cs:13-29) is convention 0.

  plain_n10      10,000 random automata, n = 10
  uniform_n16    1,000,000 random automata, n = 16, 1 % (one per block of
                 100, at a seeded position) replaced by Cerny C_16
  uneven_12_24   200,000 random automata, n in 12..24 with
                 P(n) proportional to 2^(-(n-12)/2)
  large          Cerny C_24 .. C_30 plus 64 random automata with n = 30
  serving        Poisson arrivals, n uniform in 8..30, 80 % random / 20 % Cerny
"""
from __future__ import annotations

import numpy as np

import ssw

SEEDS = {"plain_n10": 1001, "uniform_n16": 1002, "uneven_12_24": 1003, "large": 1004, "serving": 1005}

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """Vectorised splitmix64 finaliser over uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def uniform01(seed, idx):
    """Deterministic doubles in [0, 1) for (seed, idx)."""
    h = splitmix64(splitmix64(np.uint64(seed)) ^ np.asarray(idx, np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def plain_n10(seed=SEEDS["plain_n10"], count=10_000):
    A, B = ssw.generate_uniform(seed, 10, count)
    return 10, A, B


def cerny_positions(seed, count, every=100):
    """One position per block of `every` automata, at a seeded offset."""
    blocks = count // every
    k = np.arange(blocks, dtype=np.uint64)
    off = (splitmix64(splitmix64(np.uint64(seed ^ 0xC3)) ^ k) % np.uint64(every)).astype(np.int64)
    return np.arange(blocks, dtype=np.int64) * every + off


def uniform_n16(seed=SEEDS["uniform_n16"], count=1_000_000, n=16, cerny_every=100):
    """Packed (count, 2n) uint8 transitions plus the Cerny positions."""
    A, B = ssw.generate_uniform(seed, n, count)
    delta = np.empty((count, 2 * n), np.uint8)
    delta[:, :n] = A
    delta[:, n:] = B
    pos = cerny_positions(seed, count, cerny_every) if cerny_every else np.zeros(0, np.int64)
    ca, cb = ssw.generate_cerny(n, 1)
    delta[pos, :n] = ca
    delta[pos, n:] = cb
    return n, delta, pos


def uneven_sizes(seed, count, n_lo=12, n_hi=24):
    ns = np.arange(n_lo, n_hi + 1)
    w = 2.0 ** (-(ns - n_lo) / 2.0)
    cdf = np.cumsum(w / w.sum())
    u = uniform01(seed ^ 0x5A, np.arange(count, dtype=np.uint64))
    return ns[np.minimum(np.searchsorted(cdf, u, side="right"), len(ns) - 1)].astype(np.uint8)


def uneven_12_24(seed=SEEDS["uneven_12_24"], count=200_000):
    """(n_per u8, offset u64, delta u8) for automata of uneven sizes."""
    n_per = uneven_sizes(seed, count)
    offs = np.zeros(count, np.uint64)
    offs[1:] = np.cumsum(2 * n_per[:-1].astype(np.uint64))
    delta = np.empty(int(2 * n_per.astype(np.int64).sum()), np.uint8)
    for n in np.unique(n_per):
        idx = np.nonzero(n_per == n)[0]
        # automaton index i draws from (seed, i) whatever its size
        for start in range(0, len(idx), 65536):
            sel = idx[start:start + 65536]
            A, B = _gen_rows(seed, int(n), sel)
            rows = offs[sel].astype(np.int64)
            cols = np.arange(int(n))
            delta[rows[:, None] + cols[None, :]] = A
            delta[rows[:, None] + int(n) + cols[None, :]] = B
    return n_per, offs, delta


def _gen_rows(seed, n, indices):
    """Transitions of automata `indices` (any order) of size n."""
    A = np.empty((len(indices), n), np.int32)
    B = np.empty((len(indices), n), np.int32)
    # contiguous runs are generated in one call each
    indices = np.asarray(indices, np.int64)
    if len(indices) == 0:
        return A, B
    breaks = np.nonzero(np.diff(indices) != 1)[0] + 1
    starts = np.concatenate([[0], breaks])
    ends = np.concatenate([breaks, [len(indices)]])
    for s, e in zip(starts, ends):
        a, b = ssw.generate_uniform(seed, n, int(e - s), first=int(indices[s]))
        A[s:e] = a
        B[s:e] = b
    return A, B


def large(seed=SEEDS["large"], n_random=64, n_random_size=30, cerny_ns=range(24, 31)):
    """List of (n, a, b, kind) for the large-instance configuration."""
    out = []
    for n in cerny_ns:
        a, b = ssw.generate_cerny(n, 1)
        out.append((n, a, b, "cerny"))
    A, B = ssw.generate_uniform(seed, n_random_size, n_random)
    for i in range(n_random):
        out.append((n_random_size, A[i], B[i], "random"))
    return out


def serving(seed=SEEDS["serving"], rate=200.0, count=1000, n_lo=8, n_hi=30, cerny_frac=0.2):
    """Poisson arrival stream: list of (t_arrival_s, n, a, b, kind)."""
    idx = np.arange(count, dtype=np.uint64)
    gaps = -np.log1p(-uniform01(seed ^ 0x11, idx)) / rate
    t = np.cumsum(gaps)
    ns = (n_lo + np.floor(uniform01(seed ^ 0x22, idx) * (n_hi - n_lo + 1))).astype(np.int64)
    ns = np.minimum(ns, n_hi)
    is_cerny = uniform01(seed ^ 0x33, idx) < cerny_frac
    reqs = []
    for i in range(count):
        n = int(ns[i])
        if is_cerny[i]:
            a, b = ssw.generate_cerny(n, 1)
            kind = "cerny"
        else:
            A, B = ssw.generate_uniform(seed, n, 1, first=i)
            a, b, kind = A[0], B[0], "random"
        reqs.append((float(t[i]), n, a, b, kind))
    return reqs
