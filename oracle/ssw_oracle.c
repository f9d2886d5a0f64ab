/*
 * ssw_oracle.c -- CPU oracle for the shortest-synchronizing-word search.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker /
 * the timed CPU baseline.  The product path (libssw.so) never links it.
 *
 * The reference is C# (.NET Framework 4.6.1 + Alea GPU) and cannot be built
 * in this image, so this file restates its CPU search in C:
 *
 *   oracle_slim_chunk / oracle_slim_parallel
 *       GPGPU/GPGPU/CPU/SlimCPUUntweaked.cs:19-49 (Parallel.For chunking) and
 *       :51-222 (ComputeMany): FIFO queue of uint16 subsets, uint32 epoch
 *       stamps per subset, the verticesUntilBump level counter, letter A
 *       examined before B, stop at the first singleton, n <= 16.
 *   oracle_wide_chunk / oracle_wide_parallel
 *       The same loop with uint32 masks for 17 <= n <= 30 (the reference's
 *       ushort types cap it at n = 16).  Visited stamps are replaced by a bit
 *       per subset cleared through the queue after each automaton, because a
 *       4-byte stamp per subset is 4 GiB at n = 30; the loop, level counter,
 *       A-before-B order and stop rule are unchanged.
 *   oracle_witness
 *       GPGPU/GPGPU/CPU/CPU.cs:93-283 (ComputeOne): BFS with parent pointers
 *       and back-tracking of the first singleton's BFS-tree path
 *       (CPU.cs:262-269).  The word found this way is the lexicographically
 *       smallest shortest reset word with a < b (FIFO order at each level is
 *       the lexicographic order of the tree words).
 *   oracle_verify_word
 *       GPGPU/GPGPU/Result veryfier/Verify.cs:12-53: apply the word to every
 *       state; valid iff exactly one state remains.
 *
 * Independent pins used by the tests (not restatements):
 *   oracle_brute_force  -- enumerate all words by length in lexicographic
 *                          order; the first synchronizing one is the
 *                          lexicographically smallest shortest reset word.
 *   oracle_pair_sync    -- Cerny's pair criterion: an automaton is
 *                          synchronizing iff every pair of states can be
 *                          merged (backward BFS on the pair graph).
 *
 * Output convention of every search function (the C-ABI's, not C#'s):
 *   length[i]   = shortest reset-word length, -1 if not synchronizing
 *   sync[i]     = 1/0
 *   expanded[i] = canonical expanded-subset count E: the number of distinct
 *                 subsets at BFS distance < L (all reachable subsets when
 *                 not synchronizing); independent of queue order.
 * Return value of the batch functions: index (0-based, into this batch) of
 * the first automaton whose length exceeds `bound` ((n-1)^2 when bound < 0),
 * or -1; the reference returns resultsWritingIndex + i
 * (SlimCPUUntweaked.cs:189-193,209-213).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#include "ssw_oracle.h"

static int32_t default_bound(int n, int32_t bound) {
    return bound >= 0 ? bound : (n - 1) * (n - 1);
}

/* ------------------------------------------------------------------------ */
/* SlimCPUUntweaked.ComputeMany, n <= 16 (SlimCPUUntweaked.cs:51-222).       */
/* ------------------------------------------------------------------------ */
int64_t oracle_slim_chunk(int n, int64_t count, const int32_t *A, const int32_t *B,
                          int32_t *length, uint8_t *sync, int64_t *expanded,
                          int32_t bound) {
    if (n < 1 || n > 16 || count <= 0) return -1;
    const uint32_t power = 1u << n;
    const uint16_t initial = (uint16_t)(power - 1);           /* :61 */
    const int32_t max_len = default_bound(n, bound);          /* :62 */
    uint32_t local_id = 1;                                     /* :66 */
    uint32_t *discovered = (uint32_t *)calloc(power, sizeof(uint32_t)); /* :67 */
    uint16_t *queue = (uint16_t *)malloc(power * sizeof(uint16_t));     /* :69 */
    uint16_t pa[16], pb[16];
    int64_t violation = -1;
    queue[0] = initial;                                        /* :71 */

    for (int64_t p = 0; p < count; p++, local_id++) {          /* :89 */
        const int32_t *a = A + p * n, *b = B + p * n;
        uint16_t reading = 0, writing = 1;                     /* :104-105 */
        discovered[initial] = local_id;                        /* :103 */
        for (int i = 0; i < n; i++) {                          /* :107-111 */
            pa[i] = (uint16_t)(1u << a[i]);
            pb[i] = (uint16_t)(1u << b[i]);
        }
        uint16_t next_distance = 1;                            /* :114 */
        uint16_t until_bump = 0xFFFF;                          /* :115 */
        int seeking_first = 1, found = 0;                      /* :116-117 */
        int64_t level_mark = 1;  /* queue length at the last bump = E */

        while (reading < writing) {                            /* :119 */
            uint16_t v = queue[reading++];                     /* :124 */
            if (--until_bump == 0) {                           /* :126-130 */
                ++next_distance;
                seeking_first = 1;
                level_mark = writing;
            }
            uint16_t ia = 0, ib = 0;                           /* :132 */
            uint16_t mask = 1;
            for (int i = 0; i < n; i++, mask <<= 1) {          /* :139-147 */
                if (mask & v) { ia |= pa[i]; ib |= pb[i]; }
            }
            if (discovered[ia] != local_id) {                  /* :149-165 */
                if ((ia & (uint16_t)(ia - 1)) == 0) { found = 1; break; }
                discovered[ia] = local_id;
                queue[writing++] = ia;
                if (seeking_first) {
                    seeking_first = 0;
                    until_bump = (uint16_t)(writing - reading);
                }
            }
            if (discovered[ib] != local_id) {                  /* :167-183 */
                if ((ib & (uint16_t)(ib - 1)) == 0) { found = 1; break; }
                discovered[ib] = local_id;
                queue[writing++] = ib;
                if (seeking_first) {
                    seeking_first = 0;
                    until_bump = (uint16_t)(writing - reading);
                }
            }
        }
        if (length) length[p] = found ? (int32_t)next_distance : -1;   /* :198-214 */
        if (sync) sync[p] = (uint8_t)found;
        if (expanded) expanded[p] = found ? level_mark : (int64_t)writing;
        if (found && next_distance > max_len && violation < 0) violation = p;
    }
    free(discovered);
    free(queue);
    return violation;
}

/* ------------------------------------------------------------------------ */
/* The same loop with uint32 masks for n <= 30.                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t *data;
    int64_t cap;
} u32vec;

static int vec_reserve(u32vec *v, int64_t need) {
    if (need <= v->cap) return 0;
    int64_t cap = v->cap ? v->cap : 1024;
    while (cap < need) cap *= 2;
    uint32_t *d = (uint32_t *)realloc(v->data, (size_t)cap * sizeof(uint32_t));
    if (!d) return -1;
    v->data = d;
    v->cap = cap;
    return 0;
}

int64_t oracle_wide_chunk(int n, int64_t count, const int32_t *A, const int32_t *B,
                          int32_t *length, uint8_t *sync, int64_t *expanded,
                          int32_t bound) {
    if (n < 1 || n > 30 || count <= 0) return -1;
    const uint64_t power = 1ull << n;
    const uint32_t initial = (uint32_t)(power - 1);
    const int32_t max_len = default_bound(n, bound);
    uint64_t *visited = (uint64_t *)calloc((power + 63) / 64, sizeof(uint64_t));
    u32vec queue = {0, 0};
    uint32_t pa[30], pb[30];
    int64_t violation = -1;
    if (!visited || vec_reserve(&queue, 1024)) { free(visited); return -2; }

    for (int64_t p = 0; p < count; p++) {
        const int32_t *a = A + p * n, *b = B + p * n;
        int64_t reading = 0, writing = 1;
        queue.data[0] = initial;
        visited[initial >> 6] |= 1ull << (initial & 63);
        for (int i = 0; i < n; i++) {
            pa[i] = 1u << a[i];
            pb[i] = 1u << b[i];
        }
        int32_t next_distance = 1;
        int64_t until_bump = -1;  /* "infinite" until the first discovery */
        int seeking_first = 1, found = 0;
        int64_t level_mark = 1;

        while (reading < writing) {
            uint32_t v = queue.data[reading++];
            if (--until_bump == 0) {
                ++next_distance;
                seeking_first = 1;
                level_mark = writing;
            }
            uint32_t ia = 0, ib = 0;
            for (int i = 0; i < n; i++)
                if ((v >> i) & 1u) { ia |= pa[i]; ib |= pb[i]; }
            uint32_t img[2] = {ia, ib};
            for (int x = 0; x < 2 && !found; x++) {
                uint32_t w = img[x];
                if (!((visited[w >> 6] >> (w & 63)) & 1ull)) {
                    if ((w & (w - 1)) == 0) { found = 1; break; }
                    visited[w >> 6] |= 1ull << (w & 63);
                    if (vec_reserve(&queue, writing + 1)) { free(visited); free(queue.data); return -2; }
                    queue.data[writing++] = w;
                    if (seeking_first) {
                        seeking_first = 0;
                        until_bump = writing - reading;
                    }
                }
            }
            if (found) break;
        }
        if (length) length[p] = found ? next_distance : -1;
        if (sync) sync[p] = (uint8_t)found;
        if (expanded) expanded[p] = found ? level_mark : writing;
        if (found && next_distance > max_len && violation < 0) violation = p;
        /* clear through the queue: every visited subset is in it */
        for (int64_t k = 0; k < writing; k++) visited[queue.data[k] >> 6] = 0;
    }
    free(visited);
    free(queue.data);
    return violation;
}

/* ------------------------------------------------------------------------ */
/* Parallel.For over P contiguous chunks (SlimCPUUntweaked.cs:19-49).        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, wide;
    int64_t count;
    const int32_t *A, *B;
    int32_t *length;
    uint8_t *sync;
    int64_t *expanded;
    int32_t bound;
    int64_t result;
} chunk_job;

static void *chunk_main(void *arg) {
    chunk_job *j = (chunk_job *)arg;
    if (j->count <= 0) { j->result = -1; return 0; }
    j->result = j->wide
        ? oracle_wide_chunk(j->n, j->count, j->A, j->B, j->length, j->sync, j->expanded, j->bound)
        : oracle_slim_chunk(j->n, j->count, j->A, j->B, j->length, j->sync, j->expanded, j->bound);
    return 0;
}

static int64_t run_parallel(int wide, int n, int64_t count, const int32_t *A, const int32_t *B,
                            int32_t *length, uint8_t *sync, int64_t *expanded,
                            int32_t bound, int threads) {
    if (count <= 0) return -1;                                 /* :21-22 */
    if (threads < 1) threads = 1;
    if (threads > count) threads = (int)count;                 /* :24-25 */
    int64_t part = (count + threads - 1) / threads;            /* :33 */
    chunk_job *jobs = (chunk_job *)calloc((size_t)threads, sizeof(chunk_job));
    pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        int64_t off = t * part;
        int64_t cnt = count - off < part ? count - off : part; /* :42 */
        chunk_job *j = &jobs[t];
        j->n = n; j->wide = wide; j->count = cnt; j->bound = bound;
        j->A = A + off * n; j->B = B + off * n;
        j->length = length ? length + off : 0;
        j->sync = sync ? sync + off : 0;
        j->expanded = expanded ? expanded + off : 0;
        if (threads == 1) chunk_main(j);
        else pthread_create(&tids[t], 0, chunk_main, j);
    }
    int64_t violation = -1;
    for (int t = 0; t < threads; t++) {
        if (threads > 1) pthread_join(tids[t], 0);
        /* the reference keeps whichever chunk reports first
           (Interlocked.CompareExchange, :44-45); the oracle reports the
           smallest violating index so the answer is deterministic */
        if (jobs[t].result == -2) violation = -2;
        if (jobs[t].result >= 0 && violation == -1) violation = t * part + jobs[t].result;
    }
    free(jobs);
    free(tids);
    return violation;
}

int64_t oracle_slim_parallel(int n, int64_t count, const int32_t *A, const int32_t *B,
                             int32_t *length, uint8_t *sync, int64_t *expanded,
                             int32_t bound, int threads) {
    return run_parallel(0, n, count, A, B, length, sync, expanded, bound, threads);
}

int64_t oracle_wide_parallel(int n, int64_t count, const int32_t *A, const int32_t *B,
                             int32_t *length, uint8_t *sync, int64_t *expanded,
                             int32_t bound, int threads) {
    return run_parallel(1, n, count, A, B, length, sync, expanded, bound, threads);
}

/* ------------------------------------------------------------------------ */
/* CPU.ComputeOne with parent pointers (CPU.cs:93-283), uint32 masks.        */
/* word[k] = 1 when letter k is b (ComputationResult.cs:12, true = b).       */
/* Returns 0 on success, -1 bad args, -2 out of memory, -3 cap too small.    */
/* ------------------------------------------------------------------------ */
int oracle_witness(int n, const int32_t *a, const int32_t *b,
                   int32_t *length, uint8_t *word, int32_t cap) {
    if (n < 1 || n > 24) return -1;
    const uint64_t power = 1ull << n;
    const uint32_t initial = (uint32_t)(power - 1);
    uint8_t *disc = (uint8_t *)calloc(power, 1);
    uint32_t *prev = (uint32_t *)malloc(power * sizeof(uint32_t));      /* :110,131 */
    uint8_t *prev_b = (uint8_t *)malloc(power);                         /* :111,132 */
    uint32_t *queue = (uint32_t *)malloc(power * sizeof(uint32_t));     /* :137 */
    uint32_t pa[24], pb[24];
    if (!disc || !prev || !prev_b || !queue) {
        free(disc); free(prev); free(prev_b); free(queue);
        return -2;
    }
    for (int i = 0; i < n; i++) { pa[i] = 1u << a[i]; pb[i] = 1u << b[i]; } /* :152-158 */
    int64_t head = 0, tail = 0;
    queue[tail++] = initial;
    disc[initial] = 1;                                                  /* :115 */
    int found = 0;
    uint32_t first_singleton = 0;
    int32_t first_distance = 0;
    uint32_t bump_vertex = 0;                                           /* :162 */
    int32_t next_distance = 1;                                          /* :163 */
    int seeking_first = 1;
    while (head < tail && !found) {                                     /* :165 */
        uint32_t v = queue[head++];
        if (v == bump_vertex) { ++next_distance; seeking_first = 1; }   /* :171-175 */
        uint32_t ia = 0, ib = 0;
        for (int i = 0; i < n; i++)                                     /* :184-192 */
            if ((v >> i) & 1u) { ia |= pa[i]; ib |= pb[i]; }
        uint32_t img[2] = {ia, ib};
        for (int x = 0; x < 2; x++) {                                   /* :195-234 */
            uint32_t w = img[x];
            if (disc[w]) continue;
            if (seeking_first) { seeking_first = 0; bump_vertex = w; }
            disc[w] = 1;
            prev[w] = v;
            prev_b[w] = (uint8_t)x;
            queue[tail++] = w;
            if ((w & (w - 1)) == 0) {
                found = 1;
                first_singleton = w;
                first_distance = next_distance;
                break;
            }
        }
    }
    int rc = 0;
    if (!found) {
        *length = -1;
    } else {
        *length = first_distance;
        if (first_distance > cap) {
            rc = -3;
        } else {                                                        /* :262-269 */
            uint32_t cur = first_singleton;
            for (int k = first_distance - 1; k >= 0; k--) {
                word[k] = prev_b[cur];
                cur = prev[cur];
            }
        }
    }
    free(disc); free(prev); free(prev_b); free(queue);
    return rc;
}

/* Verify.VerifyValidityOfSynchronizingWord (Verify.cs:12-53). */
int oracle_verify_word(int n, const int32_t *a, const int32_t *b,
                       const uint8_t *word, int32_t len) {
    uint32_t image = 0;
    for (int q = 0; q < n; q++) {
        int s = q;
        for (int32_t k = 0; k < len; k++) s = word[k] ? b[s] : a[s];
        image |= 1u << s;
    }
    return (image & (image - 1)) == 0;   /* exactly one distinct final state */
}

/* ------------------------------------------------------------------------ */
/* Independent pins.                                                          */
/* ------------------------------------------------------------------------ */
static int brute_dfs(int n, const uint32_t *pa, const uint32_t *pb, uint32_t set,
                     int depth, int target, uint8_t *word) {
    if (depth == target) return (set & (set - 1)) == 0;
    for (int x = 0; x < 2; x++) {
        const uint32_t *t = x ? pb : pa;
        uint32_t img = 0;
        for (int i = 0; i < n; i++) if ((set >> i) & 1u) img |= t[i];
        word[depth] = (uint8_t)x;
        if (brute_dfs(n, pa, pb, img, depth + 1, target, word)) return 1;
    }
    return 0;
}

/* Lexicographically smallest word of minimal length <= maxlen that maps Q to
   a singleton.  length = -1 when none exists up to maxlen. */
int oracle_brute_force(int n, const int32_t *a, const int32_t *b, int32_t maxlen,
                       int32_t *length, uint8_t *word) {
    if (n < 1 || n > 30) return -1;
    uint32_t pa[30], pb[30];
    for (int i = 0; i < n; i++) { pa[i] = 1u << a[i]; pb[i] = 1u << b[i]; }
    uint32_t q = (uint32_t)((1ull << n) - 1);
    for (int32_t L = 0; L <= maxlen; L++) {
        if (brute_dfs(n, pa, pb, q, 0, L, word)) { *length = L; return 0; }
    }
    *length = -1;
    return 0;
}

/* Cerny's criterion: synchronizing iff every pair {p,q} can be merged. */
int oracle_pair_sync(int n, const int32_t *a, const int32_t *b) {
    if (n <= 1) return 1;
    int np = n * n;
    uint8_t *good = (uint8_t *)calloc((size_t)np, 1);
    for (int p = 0; p < n; p++) good[p * n + p] = 1;
    /* backward closure: {p,q} good if {a(p),a(q)} or {b(p),b(q)} good */
    int changed = 1;
    while (changed) {
        changed = 0;
        for (int p = 0; p < n; p++)
            for (int q = 0; q < n; q++) {
                if (good[p * n + q]) continue;
                if (good[a[p] * n + a[q]] || good[b[p] * n + b[q]]) {
                    good[p * n + q] = 1;
                    changed = 1;
                }
            }
    }
    int all = 1;
    for (int i = 0; i < np; i++) all &= good[i];
    free(good);
    return all;
}
