"""Benchmark: batched shortest-synchronizing-word search on one MI355X.

Headline workload (one "step" = one pass of the solver over one batch, inputs
already resident in HBM): 1,000,000 uniformly random binary automata with
n = 16, 1 % replaced by Cerny C_16 (length 225) at seeded positions; lengths,
flags and expanded-subset counts.  `value` is automata/s of that step.

Also reported on the same JSON line:
  * subsets expanded/s of the step, and the plain one-automaton-per-launch GPU
    BFS on a sample of the same batch (batched_vs_plain);
  * the roofline of the step's dominant launch (the launch with the longest
    duration, from HIP events recorded around each launch on its stream in
    every timed step): the kernels keep their visited bitmaps and frontiers in
    LDS, so the binding resource is the LDS test-and-set rate -- two per
    expanded subset, counted by the launch itself -- against the device's own
    ceiling for random LDS test-and-sets, measured in the same run;
    `traffic` is the launch's HBM bytes from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary;
  * e2e: the same batch from host memory to results in host memory, as the
    reference's driver times a whole call (packed rows through
    ssw_solve_batch, and int32 arrays through ssw_compute, lengths only);
  * cpu_baseline: the C restatement of the reference's SlimCPUUntweaked
    search (oracle/ssw_oracle.c) on a bounded sample, on every host core
    available to this process (core count and CPU model stated);
  * the other workloads, each run once: plain_n10 (10k automata, n = 10, with
    witnesses re-verified on Q), n16_witness (the headline batch with witness
    words, a sample re-verified on Q), uneven_12_24 (200k automata of uneven n
    in one call), verify_n13 (Verify of 2M automata from host arrays), large
    (Cerny C_24..C_30 + 64 random n = 30, device-wide path), serving (Poisson
    stream, n in 8..30, p50/p99 latency) and serving_sweep (offered load raised
    until the completion rate stops following it; served vs plain).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--no-extras | --extras name,name,...]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpgpu-binary-automaton-synchronizing-algorithms_amd")
sys.path.insert(0, PKG)

import ssw  # noqa: E402
import workloads  # noqa: E402

METRIC = "automata/s + subsets/s at n=16; p99 latency; batched/served vs plain GPU BFS"
HBM_PEAK_GBS = 8000.0
TUNING = os.path.join(PKG, "tuning_gfx950.json")
TRAFFIC = os.path.join(ROOT, "profiles", "batch_kernel_uniform_n16_traffic.txt")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(n, expanded):
    """Per automaton: 2n transition bytes; per expanded subset a 4-byte
    frontier read and two 8-byte test-and-set read-modify-writes; 4 bytes per
    discovered non-root subset written to the frontier; 8 result bytes.
    Total 24*E + 2n + 4 (what the search would move if its visited sets and
    frontiers lived in HBM)."""
    E = np.asarray(expanded, np.float64)
    n = np.asarray(n, np.float64)
    return float(np.sum(24.0 * E + 2.0 * n + 4.0))


def traffic_from_profile(key):
    """HBM bytes (`key`: per solve, or per launch k) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, scaled by a copy kernel."""
    if not os.path.exists(TRAFFIC):
        return None
    for line in open(TRAFFIC):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def host_cores():
    """Cores this process may run on: its CPU affinity, capped by a cgroup CPU
    quota when one is set; plus nproc and the CPU model for the record."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, p = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(p)
    except (OSError, ValueError):
        pass
    use = aff if quota is None else max(1, min(aff, int(math.ceil(quota))))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return use, {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota_cores": quota, "cpu_model": model}


def cpu_baseline(n, delta, sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    A = np.ascontiguousarray(delta[:sample, :n]).astype(np.int32)
    B = np.ascontiguousarray(delta[:sample, n:]).astype(np.int32)
    oracle.slim(n, A[:4000], B[:4000], threads=threads)  # warm-up
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        L, S, E, _ = oracle.slim(n, A, B, threads=threads)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    return sample / dt, L, S, E, dt


def verify_words(n, A, B, L, wit):
    """Re-applies each witness to every state (Verify.cs:12-53)."""
    ok = True
    idx = np.nonzero(L >= 0)[0]
    for i in idx:
        w = ssw.witness_bits_to_word(wit[i], int(L[i]))
        s = np.arange(n)
        for x in w:
            s = (B[i] if x else A[i])[s]
        ok &= bool(np.all(s == s[0]))
    return ok


def bench_e2e(ctx, n, delta, steps, warmup):
    """The headline batch from host memory to results in host memory: the
    whole call, transfers included, as Program.cs:137-143 times Verify."""
    count = len(delta)
    flat = np.ascontiguousarray(delta).reshape(-1)
    A = np.ascontiguousarray(delta[:, :n]).astype(np.int32)
    B = np.ascontiguousarray(delta[:, n:]).astype(np.int32)
    out = {}
    for name, fn in (("packed_rows_lengths_flags_E", lambda: ctx.solve_packed(flat, count, n)),
                     ("int32_compute_lengths", lambda: ctx.compute(n, A, B, want_sync=False,
                                                                   want_expanded=False))):
        for _ in range(warmup):
            fn()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        dt = (time.perf_counter() - t0) / steps
        out[name] = {"automata_per_s": count / dt, "ms_per_call": dt * 1e3}
    out["workload"] = (f"the {count:,} n={n} automata above from host memory to results in host memory "
                       f"(PCIe both ways), {warmup} warm-up + {steps} timed calls: packed u8 rows through "
                       "ssw_solve_batch (lengths, flags, E back), and int32 A/B arrays through ssw_compute "
                       "(the IComputable.Compute shape, lengths back)")
    return out


def bench_plain_n10(ctx):
    n, A, B = workloads.plain_n10()
    ctx.compute(n, A[:100], B[:100], witness=True)
    t0 = time.perf_counter()
    r = ctx.compute(n, A, B, witness=True)
    dt = time.perf_counter() - t0
    return {"workload": "10,000 random automata n=10 (seed 1001), lengths + witnesses, host buffers (PCIe included)",
            "automata_per_s": len(A) / dt, "ms": dt * 1e3,
            "non_synchronizing": int((r.length < 0).sum()), "max_length": int(r.length.max()),
            "witnesses_verified_on_Q": verify_words(n, A, B, r.length, r.witness)}


def bench_n16_witness(ctx, n, delta):
    """The headline batch with witness words, host buffers, PCIe included;
    median of 3 calls after one of the same size, which allocates the slot
    memory."""
    count = len(delta)
    flat = np.ascontiguousarray(delta).reshape(-1)
    ctx.solve_packed(flat, count, n, witness=True)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        r = ctx.solve_packed(flat, count, n, witness=True)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    kernel_ms = ctx.last_kernel_time()[0]
    idx = np.random.default_rng(19).choice(count, size=2000, replace=False)
    A = delta[idx, :n].astype(np.int32)
    B = delta[idx, n:].astype(np.int32)
    return {"workload": f"the {count:,} n={n} automata above with witness words (host buffers, PCIe included); "
                        "median of 3 calls after one of the same size",
            "automata_per_s": count / dt, "ms": dt * 1e3, "kernel_ms": kernel_ms,
            "launches": [{k: x[k] for k in ("kernel", "ms", "automata", "expanded")} for x in ctx.last_launch_stats()],
            "witnesses_verified_on_Q_sample": verify_words(n, A, B, r.length[idx], r.witness[idx]),
            "sample": 2000}


def bench_uneven(ctx):
    npr, offs, delta = workloads.uneven_12_24()
    count = len(npr)
    d = [ctx.to_device(x) for x in (delta, npr, offs)]
    d_len, d_exp = ctx.alloc(4 * count), ctx.alloc(8 * count)
    b = ssw.Batch(count=count, uniform_n=0, witness_stride=0, n_per=d[1], offset=d[2], delta=d[0],
                  length=d_len, sync=None, expanded=d_exp, witness=None, bound=-1, reserved=0)
    ctx.solve_device(b)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.solve_device(b)
        ts.append(time.perf_counter() - t0)
    E = ctx.to_host(d_exp, np.int64, count)
    L = ctx.to_host(d_len, np.int32, count)
    dt = min(ts)
    launches = ctx.last_launch_stats()
    for p in d + [d_len, d_exp]:
        ctx.free(p)
    return {"workload": "200,000 random automata, n in 12..24 with P(n) ~ 2^-(n-12)/2 (seed 1003), one call",
            "automata_per_s": count / dt, "subsets_per_s": float(E.sum()) / dt, "ms": dt * 1e3,
            "launches": [{k: x[k] for k in ("kernel", "n", "ms", "automata", "expanded")} for x in launches],
            "counts_per_n": np.bincount(npr)[12:].tolist(), "non_synchronizing": int((L < 0).sum())}


def bench_large(ctx):
    items = workloads.large()
    ns = [it[0] for it in items]
    # the first call allocates the instances' bitmaps and rings; the engine
    # keeps them for the next call (a resident engine's steady state)
    t0 = time.perf_counter()
    ctx.solve_mixed(ns, [it[1] for it in items], [it[2] for it in items], mode="large")
    first_dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = ctx.solve_mixed(ns, [it[1] for it in items], [it[2] for it in items], mode="large")
    dt = time.perf_counter() - t0
    cerny_ok = all(int(r.length[i]) == (n - 1) ** 2 for i, (n, a, b, k) in enumerate(items) if k == "cerny")
    hbm_ceiling = ctx.measure_atomics("hbm")
    tas = 2.0 * float(r.expanded.sum()) / dt
    return {"workload": "Cerny C_24..C_30 + 64 random n=30 (seed 1004), all concurrent, device-wide path; the second of two calls (the first allocates the instances' buffers, first_call_wall_s)",
            "wall_s": dt, "first_call_wall_s": first_dt, "subsets": int(r.expanded.sum()),
            "subsets_per_s": float(r.expanded.sum()) / dt,
            "atomics": {"where": "HBM", "achieved_per_s": tas, "ceiling_per_s": hbm_ceiling,
                        "frac": tas / hbm_ceiling,
                        "note": "two test-and-sets per expanded subset over the wall time (host loop included); "
                                "ceiling from ssw_measure_atomics (random device-scope atomicOr-with-return "
                                "into a 128 MiB bitmap)"},
            "cerny_lengths_equal_(n-1)^2": cerny_ok,
            "cerny_expanded_equal_2^n-n-1": all(int(r.expanded[i]) == 2 ** n - n - 1
                                               for i, (n, a, b, k) in enumerate(items) if k == "cerny"),
            "random_n30_lengths": sorted(set(int(x) for x, it in zip(r.length, items) if it[3] == "random"))}


def bench_verify_n13(ctx, count=1 << 21):
    """IComputable.Verify over 2^21 automata with n = 13 drawn by the
    System.Random twin (seed 1234567 + count, Program.cs:59,94), timed end to
    end from host arrays as Program.cs:137-143 times Verify; with and without
    the length-6 word pre-filter (SuperSlimGPUBreakthrough.cs:183-256)."""
    A, B = ssw.generate_dotnet(1234567 + count, 13, count)
    res = {}
    for pf in (0, 6, 0, 6):
        ctx.set_verify_prefilter(pf)
        ctx.verify(13, A, B)
        ts, ks = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            bad = ctx.verify(13, A, B)
            ts.append(time.perf_counter() - t0)
            ks.append(ctx.last_kernel_time()[0])
        r = res.setdefault(pf, {"ms": [], "kernel_ms": [], "violation": bad})
        r["ms"].append(float(np.median(ts)) * 1e3)
        r["kernel_ms"].append(float(np.median(ks)))
    ctx.set_verify_prefilter(0)
    off, on = res[0], res[6]
    dt = min(off["ms"]) * 1e-3
    bad = off["violation"]
    return {"workload": f"Verify of {count} random automata n=13 (System.Random twin, seed 1234567+count), "
                        "host arrays in, violating index out (PCIe included); median of 3 calls after one "
                        "of the same size, the better of two rounds",
            "automata_verified_per_s": count / dt, "ms": dt * 1e3,
            "violation": bad,
            "prefilter_length6_words": {"ms": min(on["ms"]), "kernel_ms_last_chunk": min(on["kernel_ms"]),
                                        "automata_verified_per_s": count / (min(on["ms"]) * 1e-3),
                                        "violation": on["violation"],
                                        "without_kernel_ms_last_chunk": min(off["kernel_ms"]),
                                        "note": "ssw_set_verify_prefilter(6): the 64 words of length 6 from Q, "
                                                "one per lane, before each search; both settings timed in this run"},
            "reference_published_n13": "SlimGPUQueue 1,953,926 automata/s at 2,097,152 problems "
                                       "(GTX 1060, PDF p.18); not the headline metric"}


def run_stream(ctx, reqs, small_workers=0):
    """Replays a request stream against a fresh server (arrival times kept);
    returns per-request results, the server's statistics and the span from
    the first submission to the last completion."""
    with ssw.Server(ctx, small_n_max=20, levels_per_tick=16, small_workers=small_workers) as srv:
        t0 = time.perf_counter()
        tickets = []
        for (t, n, a, b, kind) in reqs:
            now = time.perf_counter() - t0
            if now < t:
                time.sleep(t - now)
            tickets.append(srv.submit(n, a, b))
        res = [srv.wait(tk, 600000) for tk in tickets]
        st = srv.stats()
    return res, st


def stream_summary(reqs, res, st):
    lat = np.array([r["latency_us"] for r in res]) / 1e3
    cer = np.array([k == "cerny" for (_, _, _, _, k) in reqs])
    ns = np.array([n for (_, n, _, _, _) in reqs])
    small = ns <= 20
    return {"completed": st["completed"], "automata_per_s": st["automata_per_s"],
            "p50_ms": st["p50_us"] / 1e3, "p99_ms": st["p99_us"] / 1e3, "max_ms": st["max_us"] / 1e3,
            "random_p99_ms": float(np.percentile(lat[~cer], 99)) if (~cer).any() else None,
            "small_n_p99_ms": float(np.percentile(lat[small], 99)) if small.any() else None,
            "cerny_lengths_ok": all(r["length"] == (n - 1) ** 2 for r, (t, n, a, b, k) in zip(res, reqs)
                                    if k == "cerny")}


def bench_serving(ctx, rate, count):
    reqs = workloads.serving(rate=rate, count=count)
    with ssw.Server(ctx, small_n_max=20, levels_per_tick=16) as srv:
        # warm both engines
        a, b = ssw.generate_cerny(12, 1)
        srv.wait(srv.submit(12, a, b))
        a, b = ssw.generate_cerny(22, 1)
        srv.wait(srv.submit(22, a, b))
    res, st = run_stream(ctx, reqs)
    out = {"workload": f"Poisson stream {rate:g} req/s, {count} requests, n uniform in 8..30, 20% Cerny (seed 1005)"}
    out.update(stream_summary(reqs, res, st))
    out["subsets_per_s"] = st["subsets_expanded"] / max(1e-9, (reqs[-1][0] + st["max_us"] / 1e6))
    return out


def sweep_count(rate):
    """Requests per offered load: about 1.5 s of arrivals, 300..900 requests."""
    return int(min(900, max(300, 1.5 * rate)))


def bench_serving_sweep(ctx, rates):
    """Offered load raised until the completion rate stops following it; the
    plain one-automaton-per-launch path on the same requests; at the highest
    load, one small-request worker against two."""
    rows = []
    for rate in rates:
        reqs = workloads.serving(rate=rate, count=sweep_count(rate))
        res, st = run_stream(ctx, reqs)
        row = {"offered_per_s": rate}
        row.update(stream_summary(reqs, res, st))
        rows.append(row)
        log(f"serving {rate:g} req/s: {row['automata_per_s']:.1f} done/s, p99 {row['p99_ms']:.1f} ms")
    best = max(rows, key=lambda r: r["automata_per_s"])
    top = workloads.serving(rate=rates[-1], count=sweep_count(rates[-1]))
    res1, st1 = run_stream(ctx, top, small_workers=1)
    one = stream_summary(top, res1, st1)
    # the plain path: one launch per automaton, sequentially, same requests
    ns = [n for (_, n, _, _, _) in top]
    A = [a for (_, _, a, _, _) in top]
    B = [b for (_, _, _, b, _) in top]
    t0 = time.perf_counter()
    pr = ctx.solve_mixed(ns, A, B, mode="plain")
    plain_dt = time.perf_counter() - t0
    plain_rate = len(top) / plain_dt
    return {"workload": "Poisson streams of about 1.5 s of arrivals (300..900 requests), n uniform in 8..30, "
                        f"20% Cerny (seed 1005), offered at {', '.join(f'{r:g}' for r in rates)} req/s; "
                        "plain = the highest load's requests one launch per automaton on one stream",
            "loads": rows, "saturated_automata_per_s": best["automata_per_s"],
            "plain_automata_per_s": plain_rate, "plain_s": plain_dt,
            "plain_equals_served": bool(np.array_equal(pr.length, [r["length"] for r in res1])),
            "served_vs_plain": best["automata_per_s"] / plain_rate,
            "at_highest_load_one_small_worker": {k: one[k] for k in ("automata_per_s", "p99_ms", "random_p99_ms",
                                                                     "small_n_p99_ms")},
            "at_highest_load_two_small_workers": {k: rows[-1][k] for k in ("automata_per_s", "p99_ms", "random_p99_ms",
                                                                             "small_n_p99_ms")}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--count", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=500_000)
    ap.add_argument("--plain-sample", type=int, default=5_000)
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-count", type=int, default=600)
    ap.add_argument("--sweep-rates", default="200,400,800,1600")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="e2e,plain_n10,n16_witness,uneven_12_24,verify_n13,large,serving,serving_sweep",
                    help="comma list of the other workloads to run after the headline one")
    ap.add_argument("--tune", action="store_true",
                    help="autotune n=16 on this batch before timing (default: the committed table)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also time a device-to-device copy of this size (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the automata are independent, so on N GPUs each would solve "
              "its own shard of the batch with no communication; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    n, delta, pos = workloads.uniform_n16(count=args.count)
    count = len(delta)
    ctx = ssw.Context(0)
    if os.path.exists(TUNING):
        ctx.load_tuning(TUNING)
    d_delta = ctx.to_device(delta.reshape(-1))
    d_len = ctx.alloc(4 * count)
    d_sync = ctx.alloc(count)
    d_exp = ctx.alloc(8 * count)
    batch = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d_delta,
                      length=d_len, sync=d_sync, expanded=d_exp, witness=None, bound=-1, reserved=0)
    if args.tune:
        t0 = time.perf_counter()
        ctx.autotune_batch(batch)
        log(f"autotuned on the batch in {time.perf_counter() - t0:.2f} s: {ctx.get_tuning(n)}")

    for _ in range(args.warmup):
        ctx.solve_device(batch)
    ctx.synchronize()
    kms = []
    per_launch = []
    # each solve returns once its results are on the device (it waits on its
    # stream for the counters), so the steps' times add up to the K steps;
    # the statistics are read between steps, outside the timed work
    dt = 0.0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.solve_device(batch)
        dt += time.perf_counter() - t0
        kms.append(ctx.last_kernel_time()[0])
        per_launch.append(ctx.last_launch_stats())
    ctx.synchronize()
    ms_per_step = 1e3 * dt / args.steps
    launches = ctx.last_kernel_time()[1]
    length = ctx.to_host(d_len, np.int32, count)
    sync = ctx.to_host(d_sync, np.uint8, count)
    expanded = ctx.to_host(d_exp, np.int64, count)
    subsets = float(expanded.sum())
    kernel_ms = float(np.mean(kms))
    log(f"step {ms_per_step:.3f} ms, kernel {kernel_ms:.3f} ms, {count / (ms_per_step * 1e-3):.3e} automata/s")

    if args.copy_bytes:
        src, dst = ctx.alloc(args.copy_bytes), ctx.alloc(args.copy_bytes)
        ctx.copy_on_device(dst, src, args.copy_bytes)
        ctx.free(src)
        ctx.free(dst)

    # plain one-automaton-per-launch GPU BFS on a sample with the same Cerny share
    ps = min(args.plain_sample, count)
    pdelta = np.ascontiguousarray(delta[:ps])
    ctx.solve_packed(pdelta[:50].reshape(-1), 50, n, mode="plain")
    t1 = time.perf_counter()
    pr = ctx.solve_packed(pdelta.reshape(-1), ps, n, mode="plain")
    plain_dt = time.perf_counter() - t1
    plain_rate = ps / plain_dt
    assert np.array_equal(pr.length, length[:ps]), "plain and batched paths disagree"

    # per-launch table over the timed steps: mean HIP-event duration, counters
    # of one step (identical in every step: same batch, same kernels)
    nl = len(per_launch[0])
    table = []
    for k in range(nl):
        rec = dict(per_launch[-1][k])
        rec["ms"] = float(np.mean([step[k]["ms"] for step in per_launch]))
        rec["ms_min"] = float(np.min([step[k]["ms"] for step in per_launch]))
        rec["ms_max"] = float(np.max([step[k]["ms"] for step in per_launch]))
        rec["test_and_sets_per_s"] = 2.0 * rec["expanded"] / (rec["ms"] * 1e-3) if rec["ms"] > 0 else 0.0
        table.append(rec)
    dom = max(range(nl), key=lambda k: table[k]["ms"])
    lds_ceiling = ctx.measure_atomics("lds")
    for rec in table:
        rec["frac_of_lds_ceiling"] = rec["test_and_sets_per_s"] / lds_ceiling
    achieved = table[dom]["test_and_sets_per_s"]
    log(f"dominant launch {dom + 1} {table[dom]['kernel']}: {table[dom]['ms']:.3f} ms, "
        f"{achieved:.3e} test-and-sets/s of a {lds_ceiling:.3e}/s ceiling")
    traffic_solve = traffic_from_profile("traffic_bytes_per_solve")
    traffic_dom = traffic_from_profile(f"traffic_bytes_launch_{dom + 1}")
    abytes = algorithmic_bytes(np.full(count, n), expanded)
    roofline = {"bound": "lds_atomic", "achieved": achieved / 1e9, "peak": lds_ceiling / 1e9,
                "unit": "G test-and-sets/s", "frac": achieved / lds_ceiling, "traffic": traffic_dom,
                "kernel": table[dom]["kernel"], "launch": dom + 1, "launch_ms": table[dom]["ms"],
                "test_and_sets_per_launch": 2 * table[dom]["expanded"],
                "note": "two LDS test-and-sets (visited bitmap) per subset the launch expanded, counted by the "
                        "launch, over its mean duration from HIP events on its stream in the timed steps; peak = "
                        "ssw_measure_atomics('lds') in this run (every wavefront issuing random atomicOr-with-"
                        "return into an 8 KiB LDS bitmap); traffic = the launch's HBM bytes from "
                        "profiles/batch_kernel_uniform_n16_traffic.txt",
                "hbm": {"traffic_bytes_per_solve": traffic_solve,
                        "GB_per_s": traffic_solve / (kernel_ms * 1e-3) / 1e9 if traffic_solve else None,
                        "frac_of_8TB_per_s": (traffic_solve / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                              if traffic_solve else None),
                        "model_bytes_24E": abytes,
                        "note": "counter-measured HBM bytes per solve over the kernel time; the 24*E model is what "
                                "the search would move with its visited sets in HBM -- here they are in LDS"}}

    cpu = None
    parity = None
    if not args.no_cpu:
        threads, hostinfo = host_cores()
        sample = min(args.cpu_sample, count)
        rate, L, S, E, cdt = cpu_baseline(n, delta, sample, threads)
        parity = bool(np.array_equal(L, length[:sample]) and np.array_equal(S, sync[:sample])
                      and np.array_equal(E, expanded[:sample]))
        cpu = {"value": rate, "unit": "automata/s", "cores": threads, "kind": "port",
               "sample": f"first {sample} automata of the same batch ({int((pos < sample).sum())} Cerny C_16), "
                         f"oracle/ssw_oracle.c restatement of SlimCPUUntweaked on {threads} threads (every core "
                         f"available to the process), median of 3 after a warm-up, {cdt:.2f} s wall",
               **hostinfo}
        log(f"cpu baseline {rate:.3e} automata/s on {threads} threads ({hostinfo}), parity {parity}")

    value = count / (ms_per_step * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "automata/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "uniform_n16: 1,000,000 random binary automata n=16 (splitmix64 seed 1002), "
                               "1% replaced by Cerny C_16; lengths + flags + expanded counts, inputs in HBM",
                   "count": count, "n": n, "cerny": int(len(pos)), "launches_per_step": int(launches),
                   "tuning": ctx.get_tuning(n)},
        "subsets_per_s": subsets / (ms_per_step * 1e-3),
        "subsets_per_step": subsets,
        "kernel_ms": kernel_ms,
        "plain_gpu_automata_per_s": plain_rate,
        "plain_sample": ps,
        "batched_vs_plain": value / plain_rate,
        "parity_vs_oracle_on_cpu_sample": parity,
        "roofline": roofline,
        "launches": table,
        "cpu_baseline": cpu,
    }
    if not args.no_extras:
        wanted = set(x for x in args.extras.split(",") if x)
        rates = [float(x) for x in args.sweep_rates.split(",") if x]
        for name, fn in (("e2e", lambda: bench_e2e(ctx, n, delta, args.steps, args.warmup)),
                         ("plain_n10", lambda: bench_plain_n10(ctx)),
                         ("n16_witness", lambda: bench_n16_witness(ctx, n, delta)),
                         ("uneven_12_24", lambda: bench_uneven(ctx)),
                         ("verify_n13", lambda: bench_verify_n13(ctx)),
                         ("large", lambda: bench_large(ctx)),
                         ("serving", lambda: bench_serving(ctx, args.serving_rate, args.serving_count)),
                         ("serving_sweep", lambda: bench_serving_sweep(ctx, rates))):
            if name not in wanted:
                continue
            t0 = time.perf_counter()
            out[name] = fn()
            log(f"{name} done in {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
