// computable.hpp -- the reference's solver plug-in contract in C++, over the
// C ABI of libssw.so.  Header-only.
//
// Mirrors, name for name:
//   Problem                    GPGPU/SharedDefinitions/Problem.cs:7-79
//   ComputationResult          GPGPU/GPGPU/Shared/ComputationResult.cs:9-20
//   BenchmarkResult            GPGPU/GPGPU/Shared/BenchmarkResult.cs:9-14
//   ComputationType            GPGPU/GPGPU/Shared/ComputationType.cs:9-12
//   IComputable                GPGPU/GPGPU/Interfaces/IComputable.cs:10-20
//   Verify                     GPGPU/GPGPU/Result veryfier/Verify.cs:12-62
// HipComputable is the MI355X solver behind IComputable (it takes the place of
// SlimGPUQueue, GPGPU/GPGPU/GPU/SlimGPUQueue.cs:16-208).  Semantics kept from
// the reference:
//   * all problems of one call share n, read from the first problem
//     (SlimGPUQueue.cs:49);
//   * a non-synchronizing automaton gets isSynchronizable = false and
//     shortestSynchronizingWordLength = 0 (the C# default);
//   * Compute returns -1, or resultsWritingIndex + i for the first automaton
//     whose length exceeds (n-1)^2 (SlimCPUUntweaked.cs:189-213); Verify is
//     Compute(..., null, -1, ...) (SlimCPUUntweaked.cs:225-226), so it
//     returns -1 + i exactly as the reference does;
//   * errors surface as exceptions (C#: Exception), here std::runtime_error.
#pragma once

#include "ssw.h"

#include <chrono>
#include <cstdint>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace GPGPU {

struct Problem {
    std::vector<int> stateTransitioningMatrixA;
    std::vector<int> stateTransitioningMatrixB;
    int size = 0;

    // Problem.cs:13-29 (A: 0 -> 1, else identity; B: q -> q+1 mod n)
    static Problem GenerateWorstCase(int n) {
        if (n < 3) throw std::runtime_error("wrong number!");
        Problem p;
        p.size = n;
        p.stateTransitioningMatrixA.resize(n);
        p.stateTransitioningMatrixB.resize(n);
        if (ssw_generate_cerny(n, 0, p.stateTransitioningMatrixA.data(), p.stateTransitioningMatrixB.data()) != SSW_OK)
            throw std::runtime_error("ssw_generate_cerny failed");
        return p;
    }

    // Problem.cs:32-58: System.Random-compatible draws (the library's twin).
    static void FillArrayOfProblems(std::vector<Problem> &problems, int problemCount, int problemSize, int seed) {
        std::vector<int32_t> a((size_t)problemCount * problemSize), b(a.size());
        if (ssw_generate_dotnet(seed, problemSize, problemCount, a.data(), b.data()) != SSW_OK)
            throw std::runtime_error("ssw_generate_dotnet failed");
        problems.resize(std::max<size_t>(problems.size(), (size_t)problemCount));
        for (int i = 0; i < problemCount; ++i) {
            Problem &p = problems[i];
            p.size = problemSize;
            p.stateTransitioningMatrixA.assign(a.begin() + (size_t)i * problemSize, a.begin() + (size_t)(i + 1) * problemSize);
            p.stateTransitioningMatrixB.assign(b.begin() + (size_t)i * problemSize, b.begin() + (size_t)(i + 1) * problemSize);
        }
    }

    std::string ToString() const {  // Problem.cs:60-78
        std::string s = "[[";
        for (int i = 0; i < size; ++i) s += std::to_string(stateTransitioningMatrixA[i]) + (i < size - 1 ? "," : "");
        s += "],[";
        for (int i = 0; i < size; ++i) s += std::to_string(stateTransitioningMatrixB[i]) + (i < size - 1 ? "," : "");
        return s + "]]";
    }
};

enum class ComputationType { CPU_Parallel, CPU_Serial, GPU, CPU_GPU_Combined };

struct BenchmarkResult {
    double benchmarkedTimeMs = 0;
    double totalTimeMs = 0;
};

struct ComputationResult {
    bool isSynchronizable = false;
    std::vector<bool> shortestSynchronizingWord;  // true = letter b
    int shortestSynchronizingWordLength = 0;
    BenchmarkResult benchmarkResult;
    int queueBreadth = 0;
    ComputationType computationType = ComputationType::CPU_Parallel;
    int size = 0;
    int discoveredVertices = 0;
    std::string algorithmName;
};

class IComputable {
public:
    virtual ~IComputable() = default;
    /// Is the Cerny conjecture false?  Index of the failing problem, else -1.
    virtual int Compute(const std::vector<Problem> &problemsToSolve, int problemsReadingIndex,
                        std::vector<ComputationResult> *computationResults, int resultsWritingIndex, int problemCount,
                        int degreeOfParallelism) = 0;
    virtual int Verify(const std::vector<Problem> &problemsToSolve, int problemsReadingIndex, int problemCount,
                       int degreeOfParallelism) = 0;
    virtual int GetBestParallelism() = 0;
};

class HipComputable : public IComputable {
public:
    explicit HipComputable(int device = 0, bool witnesses = true) : witnesses_(witnesses) {
        ssw_context *c = nullptr;
        const int rc = ssw_create(device, 0, 0, &c);
        if (rc != SSW_OK) throw std::runtime_error(std::string("ssw_create: ") + ssw_last_error(nullptr));
        ctx_.reset(c);
    }

    int Compute(const std::vector<Problem> &problems, int readIdx, std::vector<ComputationResult> *results, int writeIdx,
                int problemCount, int /*degreeOfParallelism: one stream per context*/) override {
        if (problemCount == 0) return -1;
        const auto t0 = std::chrono::steady_clock::now();
        const int n = problems[readIdx].size;
        std::vector<int32_t> a((size_t)problemCount * n), b(a.size());
        for (int i = 0; i < problemCount; ++i) {  // flattened as SlimGPUQueue.cs:108-126
            const Problem &p = problems[readIdx + i];
            if (p.size != n) throw std::runtime_error("all problems of one call must share n");
            std::copy(p.stateTransitioningMatrixA.begin(), p.stateTransitioningMatrixA.end(), a.begin() + (size_t)i * n);
            std::copy(p.stateTransitioningMatrixB.begin(), p.stateTransitioningMatrixB.end(), b.begin() + (size_t)i * n);
        }
        std::vector<int32_t> length(problemCount);
        std::vector<int64_t> expanded(problemCount);
        const bool want_words = results && witnesses_;
        const int stride = std::max(1, ((n - 1) * (n - 1) + 7) / 8);
        std::vector<uint8_t> words(want_words ? (size_t)problemCount * stride : 0);
        int64_t violation = -1;
        // Verify (no results wanted): only the violating index comes back
        const int rc = results ? ssw_compute(ctx_.get(), n, problemCount, a.data(), b.data(), length.data(), nullptr,
                                             want_words ? words.data() : nullptr, stride, expanded.data(), -1,
                                             &violation)
                               : ssw_verify(ctx_.get(), n, problemCount, a.data(), b.data(), -1, &violation);
        if (rc != SSW_OK)
            throw std::runtime_error(std::string(results ? "ssw_compute: " : "ssw_verify: ") + ssw_last_error(ctx_.get()));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (results) {
            for (int i = 0; i < problemCount; ++i) {
                ComputationResult r;
                r.computationType = ComputationType::GPU;
                r.size = n;
                r.algorithmName = "HipComputable";
                r.isSynchronizable = length[i] >= 0;
                r.shortestSynchronizingWordLength = length[i] >= 0 ? length[i] : 0;
                r.discoveredVertices = (int)expanded[i];
                r.benchmarkResult.totalTimeMs = ms;
                if (want_words && length[i] >= 0) {
                    r.shortestSynchronizingWord.resize(length[i]);
                    for (int k = 0; k < length[i]; ++k)
                        r.shortestSynchronizingWord[k] = (words[(size_t)i * stride + k / 8] >> (k % 8)) & 1;
                }
                (*results)[writeIdx + i] = std::move(r);
            }
        }
        return violation < 0 ? -1 : writeIdx + (int)violation;
    }

    int Verify(const std::vector<Problem> &problems, int readIdx, int problemCount, int degreeOfParallelism) override {
        return Compute(problems, readIdx, nullptr, -1, problemCount, degreeOfParallelism);
    }

    int GetBestParallelism() override { return ssw_best_parallelism(ctx_.get()); }

    ssw_context *context() const { return ctx_.get(); }

private:
    struct Closer {
        void operator()(ssw_context *c) const { ssw_destroy(c); }
    };
    std::unique_ptr<ssw_context, Closer> ctx_;
    bool witnesses_;
};

// Verify.cs:12-53 and :60-62
struct Verify {
    static bool VerifyValidityOfSynchronizingWord(const Problem &problem, const ComputationResult &result, int = 1) {
        if (!result.isSynchronizable) return true;
        std::set<int> finals;
        for (int q = 0; q < problem.size; ++q) {
            int s = q;
            for (bool isB : result.shortestSynchronizingWord)
                s = isB ? problem.stateTransitioningMatrixB[s] : problem.stateTransitioningMatrixA[s];
            finals.insert(s);
        }
        return finals.size() == 1;
    }
    static bool VerifyCernyConjecture(const Problem &problem, const ComputationResult &result) {
        return !result.isSynchronizable ||
               (int)result.shortestSynchronizingWord.size() <= (problem.size - 1) * (problem.size - 1);
    }
};

}  // namespace GPGPU
