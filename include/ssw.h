/*
 * ssw.h -- C ABI of the MI355X (gfx950) shortest-synchronizing-word solver.
 *
 * A binary automaton with n states (1 <= n <= 30) is given by two transition
 * arrays a[q], b[q] in [0, n).  The solver runs a breadth-first search over
 * the subsets of Q, starting from Q itself, and reports per automaton:
 *
 *   length    shortest reset-word length, -1 when not synchronizing
 *   sync      1 when synchronizing
 *   expanded  canonical expanded-subset count E: distinct subsets at BFS
 *             distance < length (all reachable subsets if not synchronizing)
 *   witness   (optional) the lexicographically smallest shortest reset word,
 *             a < b, bit-packed LSB first: bit k of byte k/8 is 1 when letter
 *             k (applied k-th, left to right from Q) is b.
 *
 * The entry points replace the reference's solver plug-in interface
 *   GPGPU/GPGPU/Interfaces/IComputable.cs:10-20
 * and its data contract
 *   GPGPU/SharedDefinitions/Problem.cs:7-79, GPGPU/GPGPU/Shared/ComputationResult.cs:9-20.
 * Conventions: every function returns an int status (0 ok, < 0 error; see
 * ssw_status); no C++ exception crosses the ABI; host buffers are owned by the
 * caller and only read/written during the call; device memory is owned by the
 * library (or by the caller for the *_DEVICE memory kind).  One HIP stream per
 * context; calls on one context are serialized by an internal mutex, and
 * contexts may be created from any thread.
 */
#ifndef SSW_H
#define SSW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSW_MAX_N 30
#define SSW_ABI_VERSION 2

enum ssw_status {
    SSW_OK = 0,
    SSW_PENDING = 1,           /* server: request not finished yet */
    SSW_ERR_ARG = -1,          /* bad argument (n out of range, delta >= n, NULL ...) */
    SSW_ERR_DEVICE = -2,       /* HIP runtime error */
    SSW_ERR_NOMEM = -3,        /* device allocation failed */
    SSW_ERR_BUDGET = -4,       /* request exceeds the memory budget */
    SSW_ERR_STATE = -5,        /* object in the wrong state (e.g. server stopped) */
    SSW_ERR_INTERNAL = -6,
    SSW_ERR_TIMEOUT = -7,
    SSW_ERR_NOTFOUND = -8,     /* unknown ticket */
    SSW_ERR_CANCELLED = -9     /* the request was cancelled (ssw_server_cancel) */
};

/* Per-automaton length codes written by the device besides L >= 0 / -1. */
#define SSW_LEN_NOT_SYNC (-1)
#define SSW_LEN_INVALID  (-3)  /* a transition >= n; the automaton was skipped */

/* Where the buffers of an ssw_batch live. */
enum ssw_memory { SSW_HOST = 0, SSW_DEVICE = 1 };

typedef struct ssw_context ssw_context;

/* ---- context ------------------------------------------------------------ */

/* Opens `device` (HIP ordinal).  mem_budget bytes of device memory may be
   used for solver scratch (0 = 80 % of the free memory).  flags: reserved, 0. */
int ssw_create(int device, uint64_t mem_budget, uint32_t flags, ssw_context **out);
int ssw_destroy(ssw_context *ctx);
/* Message of the last failed call on ctx (ctx may be NULL for ssw_create
   errors).  Valid until the next call on the same context. */
const char *ssw_last_error(const ssw_context *ctx);
int ssw_abi_version(void);

/* ---- IComputable-equivalent entry points (host buffers, one n per call) -- */

/* IComputable.Compute (IComputable.cs:13; e.g. SlimGPUQueue.cs:23-208).
   delta_a/delta_b are count*n int32 values laid out automaton after automaton,
   exactly as SlimGPUQueue.cs:108-126 flattens Problem[].  Outputs are count
   long; sync_out, witness_out and expanded_out may be NULL.  witness_stride is
   the byte stride between automata in witness_out and must hold
   ceil(((n-1)^2)/8) bytes.  bound < 0 means (n-1)^2 (Verify.cs:60-62).
   *violation_out (may be NULL) receives the smallest index i (0-based in this
   call) with a synchronizing length > bound, or -1.  The C# adapter returns
   resultsWritingIndex + i, as SlimCPUUntweaked.cs:189-213 does. */
int ssw_compute(ssw_context *ctx, int32_t n, int64_t count,
                const int32_t *delta_a, const int32_t *delta_b,
                int32_t *length_out, uint8_t *sync_out,
                uint8_t *witness_out, int32_t witness_stride,
                int64_t *expanded_out, int32_t bound, int64_t *violation_out);

/* IComputable.Verify (IComputable.cs:16; SlimCPUUntweaked.cs:225-226):
   lengths only, returns the first violating index through *violation_out. */
int ssw_verify(ssw_context *ctx, int32_t n, int64_t count,
               const int32_t *delta_a, const int32_t *delta_b,
               int32_t bound, int64_t *violation_out);

/* Verify-mode short-word pre-filter (SuperSlimGPUBreakthrough.cs:183-256,
   with SlimCPUSkipper.cs:85-86 skipping what it proves): word_length 6 makes
   ssw_verify on this context try, before each search of an automaton run with
   one wavefront, the 64 words of length 6 from Q, one per lane; a word that
   synchronizes proves a length <= 6, within any bound >= 6, and that search is
   skipped.  0 (the default) turns it off.  Only ssw_verify is affected (it
   returns the violating index alone); results are the same either way. */
int ssw_set_verify_prefilter(ssw_context *ctx, int32_t word_length);

/* IComputable.GetBestParallelism (IComputable.cs:18): the number of
   concurrent solver streams worth using; 1 for this library. */
int ssw_best_parallelism(ssw_context *ctx);

/* ---- batched / uneven-size entry points --------------------------------- */

/* Packed transition format: automaton i occupies 2*n_i bytes starting at
   delta[offset[i]]: a[0..n_i), then b[0..n_i).  When uniform_n > 0, n_per may
   be NULL (all n = uniform_n) and offset may be NULL (offset = 2*n*i). */
typedef struct ssw_batch {
    int64_t count;
    int32_t uniform_n;
    int32_t witness_stride;   /* bytes per automaton in witness (if witness != NULL) */
    const uint8_t *n_per;     /* [count] or NULL */
    const uint64_t *offset;   /* [count] or NULL */
    const uint8_t *delta;     /* packed transitions */
    int32_t *length;          /* [count] out */
    uint8_t *sync;            /* [count] out, nullable */
    int64_t *expanded;        /* [count] out, nullable */
    uint8_t *witness;         /* [count * witness_stride] out, nullable */
    int32_t bound;            /* < 0: (n-1)^2 per automaton */
    int32_t reserved;
} ssw_batch;

/* Batched persistent-kernel solve: every automaton of the batch (any mix of
   n) is fed to one persistent launch per thread-group width through a device
   work queue; automata whose frontier outgrows a group's arena are finished
   by the device-wide path.  memory = SSW_HOST or SSW_DEVICE (then every
   pointer in *b is a device pointer and the call returns after the results
   are written).  No counterpart in the reference beyond Compute; this is the
   "batched" entry point added to the SharedDefinitions contract. */
int ssw_solve_batch(ssw_context *ctx, const ssw_batch *b, int32_t memory,
                    int64_t *violation_out);

/* Allocation-free variant of ssw_solve_batch with device buffers.  It has no
   counterpart in the reference; it is for callers that cannot allocate device
   memory during a call (tight serving loops, harnesses that forbid hipMalloc).
   It calls no hipMalloc, hipFree or allocating library routine: every device
   buffer it uses is the context's (set up by ssw_create) or one static scratch
   array per GPU of SSW_NOALLOC_SCRATCH_BYTES (about 44 MiB, held while the
   process runs), for the work queues, spill lists and HBM frontier rings of
   the same kernels ssw_solve_batch runs.  Limits, refused with SSW_ERR_ARG:
   every pointer of *b is a device pointer, one uniform_n in [1,
   SSW_NOALLOC_MAX_N] (n_per NULL), no witness words (witness NULL), count
   <= SSW_NOALLOC_MAX_COUNT.  Results equal ssw_solve_batch's.  Calls on one
   GPU share its scratch; the library serialises them within a process, and a
   call leaves the scratch holding its own state, which the next call
   overwrites before reading. */
#define SSW_NOALLOC_MAX_N 16
#define SSW_NOALLOC_MAX_COUNT 1048576
#define SSW_NOALLOC_SCRATCH_BYTES 46202880
int ssw_solve_batch_noalloc(ssw_context *ctx, const ssw_batch *b, int64_t *violation_out);

/* One kernel launch per automaton, each a 256-thread workgroup running the
   level loop: the plain GPU BFS the batched form is measured against (the
   role SlimGPUQueue.cs:32-208 plays in the reference). */
int ssw_solve_plain(ssw_context *ctx, const ssw_batch *b, int32_t memory,
                    int64_t *violation_out);

/* Device-wide level loop (hipGraph of level steps), all automata of the batch
   concurrently, each with its bitmap (2^n/8 bytes) and frontier in HBM.
   For large instances (n >= 20). */
int ssw_solve_large(ssw_context *ctx, const ssw_batch *b, int32_t memory,
                    int64_t *violation_out);

/* ---- tuning ------------------------------------------------------------- */

typedef struct ssw_tuning {
    int32_t group_lanes;      /* threads per automaton: 16, 32, 64 or 256 */
    int32_t block_threads;    /* workgroup size (multiple of 64, <= 256); 0 = most groups per CU */
    int32_t chunk;            /* automata taken per work-queue fetch */
    int32_t arena_lds;        /* frontier entries kept in LDS per automaton */
    int32_t lane_entries;     /* frontier entries per lane per level step: 1, 2 or 4 */
    int32_t spill_arena_lds;  /* n <= 16: 0 = a frontier that outgrows the LDS ring moves to HBM;
                                 else such automata are re-solved by a second launch whose LDS
                                 ring holds this many entries */
    int32_t spill_lane_entries; /* frontier entries per lane in that second launch: 1, 2, 4, or 8
                                   (8 with one wavefront per automaton, else taken as 4) */
    int32_t spill_group_lanes;  /* threads per automaton in that second launch: 0 (= group_lanes),
                                   16, 32, 64, 128 or 256 */
    int32_t narrow_bits;        /* n <= 16 with spill on: 0, or 8..n-1 = the first launch relabels the
                                   states so that a(Q) | b(Q) takes the lowest labels and uses a
                                   2^narrow_bits-bit visited bitmap; automata with a wider a(Q) | b(Q)
                                   go through an extra launch with the full bitmap */
} ssw_tuning;

int ssw_get_tuning(ssw_context *ctx, int32_t n, ssw_tuning *out);
int ssw_set_tuning(ssw_context *ctx, int32_t n, const ssw_tuning *t);
/* Times the candidate configurations on `samples` random automata per n for
   n in [n_lo, n_hi] and keeps the fastest; writes the table as JSON to
   out_path when it is not NULL. */
int ssw_autotune(ssw_context *ctx, int32_t n_lo, int32_t n_hi, int64_t samples,
                 uint64_t seed, const char *out_path);
int ssw_load_tuning(ssw_context *ctx, const char *path);
/* Tunes the entry of n = b->uniform_n on the caller's own device batch (for
   instance a workload with long-tail automata); memory must be SSW_DEVICE. */
int ssw_autotune_batch(ssw_context *ctx, const ssw_batch *b, int32_t memory);

/* ---- progress (device counters) ----------------------------------------- */

typedef struct ssw_progress {
    uint64_t automata_done;
    uint64_t subsets_expanded;   /* sum of E over finished automata */
    uint64_t launches;
    uint64_t deferred;           /* automata handed to the device-wide path */
    uint64_t in_flight;          /* automata of the running solve not finished yet (0 between solves) */
} ssw_progress;

typedef void (*ssw_progress_fn)(const ssw_progress *p, void *user);

int ssw_get_progress(ssw_context *ctx, ssw_progress *out);
/* Calls fn from a library thread every interval_ms while a solve runs, and
   once more after every solve that ended since the thread last looked (also
   when set_progress_callback stops the thread), so even a solve shorter than
   the interval is reported with its final counts.  fn = NULL turns it off;
   the call returns after the thread has stopped, so fn is not called after
   that.  This is the hook a Notifications-style progress reporter
   (GPGPU/Notifications/INotify.cs:7-10) subscribes to. */
int ssw_set_progress_callback(ssw_context *ctx, ssw_progress_fn fn, void *user,
                              int32_t interval_ms);

/* ---- device memory helpers (callers without another HIP allocator) ------ */

int ssw_device_alloc(ssw_context *ctx, uint64_t bytes, void **out);
int ssw_device_free(ssw_context *ctx, void *p);
int ssw_copy_to_device(ssw_context *ctx, void *dst, const void *src, uint64_t bytes);
int ssw_copy_to_host(ssw_context *ctx, void *dst, const void *src, uint64_t bytes);
int ssw_copy_on_device(ssw_context *ctx, void *dst, const void *src, uint64_t bytes);
int ssw_synchronize(ssw_context *ctx);
/* Average duration (ms) of the last solve's main kernel, measured with HIP
   events on the context's stream, and its launch count. */
int ssw_last_kernel_time(ssw_context *ctx, double *ms, int64_t *launches);

/* ---- measurement --------------------------------------------------------- */

/* One kernel launch of the last batched solve on the context (no counterpart
   in the reference; the per-launch measurement behind the roofline report).
   Launches of one solve come in class order; within a class, a spill chain's
   launches in order. */
typedef struct ssw_launch_stat {
    int32_t n;                /* largest n of the launch's class */
    int32_t variant;          /* 0 = general kernel (HBM bitmaps, witnesses), > 0 = n <= 16 kernel */
    int32_t group_lanes;      /* threads per automaton */
    int32_t block_threads;    /* workgroup size */
    int32_t grid;             /* workgroups */
    int32_t lds_bytes;        /* LDS per workgroup */
    int32_t ring_entries;     /* LDS frontier ring (arena) entries per automaton */
    int32_t lane_entries;     /* frontier entries per lane per level step (at most) */
    int32_t bitmap_bits;      /* LDS visited bitmap per automaton (0: in HBM) */
    int32_t pass;             /* 0 the class's only launch, 1 first launch of a spill chain,
                                 2 a later launch of the chain */
    double ms;                /* time between HIP events recorded around the launch on its stream */
    uint64_t automata;        /* automata this launch finished */
    uint64_t subsets;         /* their canonical expanded counts E, summed */
    uint64_t expanded;        /* subsets this launch expanded -- two test-and-sets each --
                                 including the partial searches it handed to a later launch */
    char kernel[64];          /* e.g. "SmallSolver<WaveGroup<64>,4,true>" */
} ssw_launch_stat;

/* Fills out[0 .. min(cap, *count)) and sets *count to the number of launches
   of the last batched solve (ssw_solve_batch, ssw_compute, ssw_verify,
   ssw_solve_batch_noalloc; at most 32 are recorded; the plain and device-wide
   entry points record none). */
int ssw_last_launch_stats(ssw_context *ctx, ssw_launch_stat *out, int32_t cap, int32_t *count);

/* Test-and-set ceiling of the device (no counterpart in the reference; for the
   roofline report beside the solvers' own rates): every wavefront of the GPU
   issues atomicOr-with-return at random bits, either of an 8 KiB LDS bitmap
   per wavefront (the n <= 16 visited set) or of a 128 MiB HBM bitmap at device
   scope (the n = 30 visited set).  Writes operations per second. */
enum { SSW_ATOMICS_LDS = 0, SSW_ATOMICS_HBM = 1 };
int ssw_measure_atomics(ssw_context *ctx, int32_t kind, double *ops_per_s);

/* ---- generators (the SharedDefinitions Problem.* generators) ------------ */

/* Uniform i.i.d. transitions from a counter-based splitmix64 stream: value
   (seed, automaton index i, state q, letter x) -> [0, n).  Writes count*n
   values to a and b for automata first .. first+count-1.  Replaces
   Problem.GenerateProblem / FillArrayOfProblems (Problem.cs:32-58), whose
   System.Random stream is a .NET runtime detail. */
int ssw_generate_uniform(uint64_t seed, int32_t n, int64_t first, int64_t count,
                         int32_t *a, int32_t *b);
/* Cerny automaton C_n.  convention 0 (the reference, Problem.cs:13-29):
   a: 0->1 else identity, b: q -> q+1 mod n.  convention 1: letters swapped. */
int ssw_generate_cerny(int32_t n, int32_t convention, int32_t *a, int32_t *b);
/* .NET System.Random-compatible twin of Problem.FillArrayOfProblems
   (Problem.cs:32-58): count automata of size n from `seed` exactly as the
   reference's generator would draw them. */
int ssw_generate_dotnet(int32_t seed, int32_t n, int64_t count, int32_t *a, int32_t *b);

/* ---- serving (continuous batching) -------------------------------------- */

typedef struct ssw_server ssw_server;

typedef struct ssw_server_config {
    uint64_t mem_budget;        /* bytes of device memory for admitted large requests
                                   (0 = half of the owning context's budget) */
    int32_t small_n_max;        /* n <= this runs in the batched engine (at most 20) */
    int32_t levels_per_tick;    /* device-wide engine: levels per graph replay (0 = 16) */
    int32_t max_large;          /* device-wide engine: concurrent instances (0 = 256) */
    int32_t want_witness;       /* compute witnesses for every request */
    int32_t small_workers;      /* batched rounds that may run at once, each on its
                                   own streams (1..4, 0 = 2): a request arriving
                                   while a round runs starts in the next idle one */
    int32_t reserved;
} ssw_server_config;

typedef struct ssw_result {
    int32_t length;             /* -1 not synchronizing */
    int32_t n;
    int64_t expanded;
    uint8_t sync;
    uint8_t witness_valid;
    uint8_t pad[6];
    double latency_us;          /* submit -> result available */
    uint8_t witness[112];       /* ceil(29*29/8) = 106 bytes suffice */
} ssw_result;

typedef struct ssw_server_stats {
    uint64_t submitted, completed, rejected;
    double p50_us, p99_us, max_us, mean_us;
    double automata_per_s;      /* completed / (last completion - first submit) */
    uint64_t subsets_expanded;
} ssw_server_stats;

int ssw_server_create(ssw_context *ctx, const ssw_server_config *cfg, ssw_server **out);
/* Copies the request; returns at once with a ticket.  SSW_ERR_BUDGET when the
   request alone can never fit the budget (the same figure admission uses). */
int ssw_server_submit(ssw_server *s, int32_t n, const int32_t *a, const int32_t *b,
                      uint64_t *ticket);
/* SSW_OK and *out filled when done, SSW_PENDING otherwise.  A result is
   handed out once: the ticket is released by the call that returns it, and
   later calls with it return SSW_ERR_NOTFOUND. */
int ssw_server_poll(ssw_server *s, uint64_t ticket, ssw_result *out);
/* As poll, blocking up to timeout_ms (< 0: no limit); SSW_ERR_TIMEOUT when
   the result is not ready by then (the ticket stays valid).  When several
   callers wait on one ticket, one gets the result and the others
   SSW_ERR_NOTFOUND. */
int ssw_server_wait(ssw_server *s, uint64_t ticket, int32_t timeout_ms, ssw_result *out);
/* Cancels a request still waiting in its queue (not yet taken by an engine):
   SSW_OK and the ticket is released -- callers blocked in ssw_server_wait on
   it return SSW_ERR_CANCELLED, later calls with it SSW_ERR_NOTFOUND.  A
   request already being solved, or done, is not cancelled: SSW_ERR_STATE (its
   result stays available).  SSW_ERR_NOTFOUND for an unknown ticket. */
int ssw_server_cancel(ssw_server *s, uint64_t ticket);
/* Percentiles over the last 2^20 completed requests. */
int ssw_server_stats_get(ssw_server *s, ssw_server_stats *out);
/* Why the server stopped on an error (an engine's HIP failure: then every
   pending and later call returns that error), or "" while it has not.  The
   text stays valid until ssw_server_destroy. */
const char *ssw_server_last_error(ssw_server *s);
/* Stops taking submissions, finishes every request already submitted, wakes
   the callers blocked in ssw_server_wait (they return their result, or
   SSW_ERR_STATE) and returns once none is left inside the server's calls;
   results not collected are dropped.  No call may use s after this one. */
int ssw_server_destroy(ssw_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SSW_H */
