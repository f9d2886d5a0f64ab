"""Writes the golden fixtures of tests/golden from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  plain_n10.npz     all 10,000 automata of the n = 10 workload (seed 1001):
                    transitions, length (-1 = not synchronizing), expanded
                    count, and the witness (lexicographically smallest
                    shortest reset word, bit-packed LSB first, 1 = b)
  uniform_n16_slice.npz
                    the first 10,000 automata of the 1M n = 16 workload
                    (seed 1002, 100 Cerny C_16 among them): transitions,
                    length, expanded count; witnesses for the first 500
The inputs are stored, so the fixtures do not depend on the generator.  The
oracle itself is pinned independently in tests/test_oracle.py (Cerny closed
forms, brute force, the pair criterion, invariants).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gpgpu-binary-automaton-synchronizing-algorithms_amd"))

import oracle  # noqa: E402
import workloads  # noqa: E402


def pack_witnesses(n, A, B, L, rows):
    stride = max(1, ((n - 1) ** 2 + 7) // 8)
    out = np.zeros((len(rows), stride), np.uint8)
    for k, i in enumerate(rows):
        if L[i] < 0:
            continue
        Lw, w = oracle.witness(n, A[i], B[i])
        assert Lw == L[i]
        out[k, :len(np.packbits(w, bitorder="little"))] = np.packbits(w, bitorder="little")
    return out


def main():
    n, A, B = workloads.plain_n10()
    L, S, E, _ = oracle.slim(n, A, B, threads=8)
    W = pack_witnesses(n, A, B, L, range(len(A)))
    np.savez_compressed(os.path.join(HERE, "plain_n10.npz"), a=A.astype(np.uint8), b=B.astype(np.uint8),
                        length=L, expanded=E, witness=W)
    n, delta, pos = workloads.uniform_n16(count=10_000)
    A = delta[:, :n].astype(np.int32)
    B = delta[:, n:].astype(np.int32)
    L, S, E, _ = oracle.slim(n, A, B, threads=8)
    W = pack_witnesses(n, A, B, L, range(500))
    np.savez_compressed(os.path.join(HERE, "uniform_n16_slice.npz"), a=A.astype(np.uint8), b=B.astype(np.uint8),
                        length=L, expanded=E, witness500=W, cerny_positions=pos)
    print("fixtures written")


if __name__ == "__main__":
    main()
