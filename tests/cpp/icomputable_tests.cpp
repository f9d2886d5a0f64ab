// The reference's MSTest suites restated against IComputable in C++:
//   SimpleAutomataTests (AlgorithmCorrectnessTestProject/SimpleAutomataTests.cs:14-35)
//   IntegrityTest       (AlgorithmCorrectnessTestProject/IntegrityTest.cs:18-129)
// The cross-solver comparison uses the CPU oracle as the reference column (the
// restatement of SlimCPUUntweaked) and compares every length exactly (the
// reference only compared lengths above 9, IntegrityTest.cs:120-123).
#include "computable.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" int64_t oracle_slim_parallel(int n, int64_t count, const int32_t *A, const int32_t *B, int32_t *length,
                                        uint8_t *sync, int64_t *expanded, int32_t bound, int threads);

using namespace GPGPU;

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            ++failures;                                    \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                  \
            fprintf(stderr, "\n");                         \
        }                                                  \
    } while (0)

// the oracle as an IComputable (the reference's SlimCPUUntweaked column)
class OracleComputable : public IComputable {
public:
    int Compute(const std::vector<Problem> &p, int ri, std::vector<ComputationResult> *res, int wi, int count,
                int dop) override {
        if (count == 0) return -1;
        const int n = p[ri].size;
        std::vector<int32_t> a((size_t)count * n), b(a.size()), len(count);
        for (int i = 0; i < count; ++i)
            for (int q = 0; q < n; ++q) {
                a[(size_t)i * n + q] = p[ri + i].stateTransitioningMatrixA[q];
                b[(size_t)i * n + q] = p[ri + i].stateTransitioningMatrixB[q];
            }
        const int64_t v = oracle_slim_parallel(n, count, a.data(), b.data(), len.data(), nullptr, nullptr, -1, dop);
        if (res)
            for (int i = 0; i < count; ++i) {
                ComputationResult r;
                r.computationType = ComputationType::CPU_Parallel;
                r.size = n;
                r.algorithmName = "SlimCPUUntweaked";
                r.isSynchronizable = len[i] >= 0;
                r.shortestSynchronizingWordLength = len[i] >= 0 ? len[i] : 0;
                (*res)[wi + i] = r;
            }
        return v < 0 ? -1 : wi + (int)v;
    }
    int Verify(const std::vector<Problem> &p, int ri, int count, int dop) override {
        return Compute(p, ri, nullptr, -1, count, dop);
    }
    int GetBestParallelism() override { return 8; }
};

static void CheckWorstCase(IComputable &c, int n) {  // SimpleAutomataTests.cs:14-26
    std::vector<Problem> problems{Problem::GenerateWorstCase(n)};
    std::vector<ComputationResult> result(1);
    c.Compute(problems, 0, &result, 0, 1, c.GetBestParallelism());
    CHECK(result[0].isSynchronizable, "C_%d not synchronizable", n);
    CHECK(result[0].shortestSynchronizingWordLength == (n - 1) * (n - 1), "C_%d length %d", n,
          result[0].shortestSynchronizingWordLength);
    CHECK(Verify::VerifyValidityOfSynchronizingWord(problems[0], result[0], 1), "C_%d word invalid", n);
    CHECK(Verify::VerifyCernyConjecture(problems[0], result[0]), "C_%d bound", n);
}

static void AssertSameResults(const std::vector<std::vector<ComputationResult>> &results,
                              const std::vector<Problem> &problems) {  // IntegrityTest.cs:94-129
    for (size_t k = 1; k < results.size(); ++k) {
        CHECK(results[0].size() == results[k].size(), "incomplete results");
        for (size_t i = 0; i < results[0].size(); ++i) {
            const ComputationResult &r0 = results[0][i], &rk = results[k][i];
            CHECK(r0.isSynchronizable == rk.isSynchronizable, "problem %zu flag", i);
            CHECK(r0.shortestSynchronizingWordLength == rk.shortestSynchronizingWordLength,
                  "problem %zu length %d vs %d", i, r0.shortestSynchronizingWordLength,
                  rk.shortestSynchronizingWordLength);
            if (rk.isSynchronizable && !rk.shortestSynchronizingWord.empty())
                CHECK(Verify::VerifyValidityOfSynchronizingWord(problems[i], rk, 1), "problem %zu word", i);
        }
    }
}

static void AssertProblems(std::vector<IComputable *> &computables, const std::vector<Problem> &problems) {
    std::vector<std::vector<ComputationResult>> results;
    for (IComputable *c : computables) {
        results.emplace_back(problems.size());
        c->Compute(problems, 0, &results.back(), 0, (int)problems.size(), c->GetBestParallelism());
    }
    AssertSameResults(results, problems);
}

int main(int argc, char **argv) {
    const bool quick = argc > 1 && std::string(argv[1]) == "quick";
    HipComputable hip;
    OracleComputable cpu;
    std::vector<IComputable *> computables{&cpu, &hip};

    // SimpleAutomataTests.WorstCasesFrom3To13 / WorstCase13
    for (int n = 3; n < 14; ++n) CheckWorstCase(hip, n);
    CheckWorstCase(hip, 13);

    // IntegrityTest.WorstCases3To13ThreeTimes / WorstCase13ThreeTimes / WorstCase13SixteenTimes
    for (int n = 3; n < 14; ++n)
        AssertProblems(computables, {Problem::GenerateWorstCase(n), Problem::GenerateWorstCase(n),
                                     Problem::GenerateWorstCase(n)});
    AssertProblems(computables, std::vector<Problem>(16, Problem::GenerateWorstCase(13)));

    // IntegrityTest.LargeIntegrityCheck: 2^17 automata, n = 13, seeds 12456..12465
    const int count = quick ? 1 << 14 : 1 << 17;
    const int seeds = quick ? 2 : 10;
    std::vector<Problem> problems;
    for (int seed = 12456; seed < 12456 + seeds; ++seed) {
        Problem::FillArrayOfProblems(problems, count, 13, seed);
        AssertProblems(computables, problems);
    }

    // Compute / Verify index semantics: a violation at position 5 of a call
    // written at resultsWritingIndex 3 is reported as 8; Verify reports -1 + 5
    {
        std::vector<Problem> p;
        Problem::FillArrayOfProblems(p, 64, 9, 7);
        p[5] = Problem::GenerateWorstCase(9);  // length 64
        // lower the bound through the C ABI: the C++ adapter keeps (n-1)^2
        std::vector<int32_t> a, b, len(64);
        for (auto &q : p) {
            a.insert(a.end(), q.stateTransitioningMatrixA.begin(), q.stateTransitioningMatrixA.end());
            b.insert(b.end(), q.stateTransitioningMatrixB.begin(), q.stateTransitioningMatrixB.end());
        }
        int64_t v = -2;
        ssw_compute(hip.context(), 9, 64, a.data(), b.data(), len.data(), nullptr, nullptr, 0, nullptr, 63, &v);
        CHECK(v == 5, "violation index %lld", (long long)v);
        std::vector<ComputationResult> res(100);
        CHECK(hip.Compute(p, 0, &res, 3, 64, 1) == -1, "no violation of (n-1)^2");
        CHECK(res[3 + 5].shortestSynchronizingWordLength == 64, "C_9 at 8");
        CHECK(hip.Verify(p, 0, 64, 1) == -1, "verify");
        CHECK(hip.GetBestParallelism() == 1, "parallelism");
    }
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("icomputable tests passed\n");
    return 0;
}
