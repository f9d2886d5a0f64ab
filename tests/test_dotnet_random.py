"""The System.Random twin behind ssw_generate_dotnet (the reference's
Problem.FillArrayOfProblems / GenerateProblem, GPGPU/SharedDefinitions/Problem.cs:32-58).

The restatement below follows the published .NET Framework algorithm (Knuth's
subtractive generator: MSEED 161803398, 55-entry table, taps 21 apart, four
mixing rounds; Sample() = InternalSample() / int.MaxValue; Next(lo, hi) =
(int)(Sample() * (hi - lo)) + lo).  It is pinned by known answers of .NET
Framework's System.Random: new Random(0).Next() gives 1559595546, 1755192844,
1649316166, 1198642031, 442452829 and new Random(1).Next() gives 534011718
first.  The library's C++ generator is then compared with it on whole problem
sets, built the way Problem.cs builds them (one outer Random(seed) whose Next()
seeds each problem's own Random; per state, a then b).
"""
import numpy as np
import pytest

import ssw

MBIG = 0x7FFFFFFF
MSEED = 161803398


def _i32(x):
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= 1 << 31 else x


class DotnetRandom:
    def __init__(self, seed):
        sub = MBIG if seed == -(1 << 31) else abs(seed)
        mj = _i32(MSEED - sub)
        a = [0] * 56
        a[55] = mj
        mk = 1
        for i in range(1, 55):
            ii = (21 * i) % 55
            a[ii] = mk
            mk = _i32(mj - mk)
            if mk < 0:
                mk += MBIG
            mj = a[ii]
        for _ in range(4):
            for i in range(1, 56):
                a[i] = _i32(a[i] - a[1 + (i + 30) % 55])
                if a[i] < 0:
                    a[i] += MBIG
        self.a, self.inext, self.inextp = a, 0, 21

    def next(self):
        self.inext = 1 if self.inext + 1 >= 56 else self.inext + 1
        self.inextp = 1 if self.inextp + 1 >= 56 else self.inextp + 1
        r = _i32(self.a[self.inext] - self.a[self.inextp])
        if r == MBIG:
            r -= 1
        if r < 0:
            r += MBIG
        self.a[self.inext] = r
        return r

    def next_range(self, lo, hi):
        return int(self.next() * (1.0 / MBIG) * (hi - lo)) + lo


def test_known_answers():
    r = DotnetRandom(0)
    assert [r.next() for _ in range(5)] == [1559595546, 1755192844, 1649316166, 1198642031, 442452829]
    assert DotnetRandom(1).next() == 534011718


def _problem_set(seed, n, count):
    outer = DotnetRandom(seed)
    A = np.zeros((count, n), np.int32)
    B = np.zeros((count, n), np.int32)
    for i in range(count):
        r = DotnetRandom(outer.next())
        for q in range(n):
            A[i, q] = r.next_range(0, n)
            B[i, q] = r.next_range(0, n)
    return A, B


@pytest.mark.parametrize("seed,n", [(0, 3), (1234567, 13), (1234567 + 2097152, 16), (-42, 9), (2**31 - 1, 30)])
def test_generator_matches_restatement(seed, n):
    A, B = ssw.generate_dotnet(seed, n, 40)
    A2, B2 = _problem_set(seed, n, 40)
    np.testing.assert_array_equal(A, A2)
    np.testing.assert_array_equal(B, B2)
