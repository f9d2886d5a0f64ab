"""Runs the library's per-n autotuner on device 0 and writes the tuning table.

    python autotune.py [--lo 4] [--hi 24] [--samples 200000] [--out tuning_gfx950.json]

bench.py and the tests load the committed table (tuning_gfx950.json) when it
exists; without it the library's built-in defaults apply.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import ssw  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lo", type=int, default=4)
    ap.add_argument("--hi", type=int, default=24)
    ap.add_argument("--samples", type=int, default=200_000)
    ap.add_argument("--seed", type=int, default=77)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_gfx950.json"))
    ap.add_argument("--n16-on-workload", action="store_true",
                    help="tune n=16 on the 1M-automaton n=16 workload with its 1%% Cerny share")
    ap.add_argument("--start", default=None,
                    help="tuning table to start from (default: the library's built-in defaults)")
    ap.add_argument("--skip-random", action="store_true",
                    help="keep the starting table for every n; only the workload tuning runs")
    args = ap.parse_args()
    with ssw.Context(0) as ctx:
        if args.start:
            ctx.load_tuning(args.start)
        for n in range(args.lo, args.hi + 1):
            if args.skip_random:
                break
            t0 = time.perf_counter()
            ctx.autotune(n, n, args.samples, args.seed, None)
            print(f"n={n}: {ctx.get_tuning(n)} ({time.perf_counter() - t0:.1f} s)", flush=True)
        if args.n16_on_workload:
            import workloads
            n, delta, _ = workloads.uniform_n16()
            count = len(delta)
            d_delta, d_len = ctx.to_device(delta.reshape(-1)), ctx.alloc(4 * count)
            b = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d_delta,
                          length=d_len, sync=None, expanded=None, witness=None, bound=-1, reserved=0)
            for _ in range(2):  # two passes of the coordinate search
                ctx.autotune_batch(b)
            print(f"n=16 on the workload: {ctx.get_tuning(16)}", flush=True)
        import json
        table = {"device": "gfx950", "samples": args.samples,
                 "tuning": {str(n): ctx.get_tuning(n) for n in range(args.lo, args.hi + 1)}}
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
        print("wrote", args.out)


if __name__ == "__main__":
    main()
