"""Benchmark: batched shortest-synchronizing-word search on one MI355X.

Workload (one "step" = one pass of the solver over one batch, inputs already
resident in HBM): 1,000,000 uniformly random binary automata with n = 16,
1 % replaced by Cerny C_16 (length 225) at seeded positions.  Reports
automata/s (value), subsets expanded/s, the plain one-automaton-per-launch GPU
BFS on a sample of the same workload, the HBM roofline of the batched kernel,
and the CPU oracle (a C restatement of the reference's SlimCPUUntweaked)
timed on a bounded sample on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
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


def algorithmic_bytes(n, expanded):
    """Bytes the search must move per automaton: 2n transition bytes, per
    expanded subset a 4-byte frontier read and two 8-byte test-and-set RMWs,
    a 4-byte frontier write per discovered non-root subset, 8 result bytes:
    24*E + 2n + 4."""
    E = expanded.astype(np.float64)
    return float(np.sum(24.0 * E + 2.0 * n + 4.0))


def traffic_from_profile():
    """HBM bytes per launch measured by rocprofv3 PMC runs (committed summary)."""
    path = os.path.join(ROOT, "profiles", "batch_kernel_uniform_n16_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(n, delta, sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    A = delta[:sample, :n].astype(np.int32)
    B = delta[:sample, n:].astype(np.int32)
    oracle.slim(n, A[:1000], B[:1000], threads=threads)  # warm-up
    t0 = time.perf_counter()
    L, S, E, _ = oracle.slim(n, A, B, threads=threads)
    dt = time.perf_counter() - t0
    return sample / dt, L, S, E, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--count", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=250_000)
    ap.add_argument("--plain-sample", type=int, default=5_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every automaton is independent, so N GPUs would each "
              "solve their own shard of the batch; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    n, delta, pos = workloads.uniform_n16(count=args.count)
    count = len(delta)
    ctx = ssw.Context(0)
    d_delta = ctx.to_device(delta.reshape(-1))
    d_len = ctx.alloc(4 * count)
    d_sync = ctx.alloc(count)
    d_exp = ctx.alloc(8 * count)
    batch = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d_delta,
                      length=d_len, sync=d_sync, expanded=d_exp, witness=None, bound=-1, reserved=0)

    for _ in range(args.warmup):
        ctx.solve_device(batch)
    ctx.solve_device(batch)
    kernel_ms, launches = ctx.last_kernel_time()
    ctx.synchronize()
    t0 = time.perf_counter()
    kms = []
    for _ in range(args.steps):
        ctx.solve_device(batch)
        kms.append(ctx.last_kernel_time()[0])
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    length = ctx.to_host(d_len, np.int32, count)
    sync = ctx.to_host(d_sync, np.uint8, count)
    expanded = ctx.to_host(d_exp, np.int64, count)
    subsets = float(expanded.sum())
    kernel_ms = float(np.mean(kms))

    # plain one-automaton-per-launch GPU BFS on a sample with the same Cerny share
    ps = min(args.plain_sample, count)
    pdelta = np.ascontiguousarray(delta[:ps])
    t1 = time.perf_counter()
    pr = ctx.solve_packed(pdelta.reshape(-1), ps, n, mode="plain")
    plain_dt = time.perf_counter() - t1
    plain_rate = ps / plain_dt
    assert np.array_equal(pr.length, length[:ps]), "plain and batched paths disagree"

    abytes = algorithmic_bytes(n, expanded)
    achieved = abytes / (kernel_ms * 1e-3) / 1e9
    traffic = traffic_from_profile()

    cpu = None
    parity = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        sample = min(args.cpu_sample, count)
        rate, L, S, E, cdt = cpu_baseline(n, delta, sample, threads)
        parity = bool(np.array_equal(L, length[:sample]) and np.array_equal(S, sync[:sample])
                      and np.array_equal(E, expanded[:sample]))
        cpu = {"value": rate, "unit": "automata/s", "cores": threads, "kind": "port",
               "sample": f"first {sample} automata of the same batch (1% Cerny C_16), "
                         f"oracle/ssw_oracle.c restatement of SlimCPUUntweaked, {cdt:.2f} s wall"}

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
                               "1% replaced by Cerny C_16; lengths+flags+expanded counts, no witness",
                   "count": count, "n": n, "cerny": int(len(pos)), "launches_per_step": int(launches)},
        "subsets_per_s": subsets / (ms_per_step * 1e-3),
        "subsets_per_step": subsets,
        "kernel_ms": kernel_ms,
        "plain_gpu_automata_per_s": plain_rate,
        "plain_sample": ps,
        "batched_vs_plain": value / plain_rate,
        "parity_vs_oracle_on_cpu_sample": parity,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
