"""Benchmark driver in the shape of the reference's Program.Main
(GPGPU/GPGPU/Program.cs:21-227): for growing problem counts, draw random
automata with the System.Random-compatible generator (seed 1234567 + count,
Program.cs:59,94), time IComputable.Verify end to end (host arrays in, the
violating index out, Program.cs:111,137-143), append one CSV row
`problemsize,problemcount,<solver>...` of round(count / seconds)
(Program.cs:64-68,153), and on a Cerny counterexample append the automaton to
cernyFailed.csv (Program.cs:112-125).  Progress of the running solve comes
from the device counters through ssw_set_progress_callback -- the hook a
Notifications listener (GPGPU/Notifications/INotify.cs:7-10) subscribes to.

    python driver.py [--n 13] [--start 2097152] [--max 33554432] [--factor 2]
                     [--solvers batched,plain] [--csv gpgpu.csv] [--progress]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import ssw  # noqa: E402


def problem_string(n, a, b):
    """Problem.ToString (Problem.cs:60-78)."""
    return "[[" + ",".join(str(x) for x in a) + "],[" + ",".join(str(x) for x in b) + "]]"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=13)
    ap.add_argument("--start", type=int, default=1 << 21)
    ap.add_argument("--max", type=int, default=1 << 25)
    ap.add_argument("--factor", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=1234567)
    ap.add_argument("--solvers", default="batched")
    ap.add_argument("--csv", default="gpgpu.csv")
    ap.add_argument("--failed", default="cernyFailed.csv")
    ap.add_argument("--progress", action="store_true")
    args = ap.parse_args(argv)
    solvers = [s.strip() for s in args.solvers.split(",") if s.strip()]
    names = {"batched": "HipComputable", "plain": "HipPlainComputable"}
    ctx = ssw.Context(0)
    tuning = os.path.join(HERE, "tuning_gfx950.json")
    if os.path.exists(tuning):
        ctx.load_tuning(tuning)
    if args.progress:
        ctx.set_progress_callback(lambda p: print(f"  progress: {p['automata_done']} automata, "
                                                  f"{p['subsets_expanded']} subsets", flush=True), 200)
    rows = ["problemsize,problemcount," + ",".join(names[s] for s in solvers)]
    count = args.start
    while count < args.max:
        print(f"\n{count} problems", flush=True)
        A, B = ssw.generate_dotnet(args.seed + count, args.n, count)
        row = [str(args.n), str(count)]
        for s in solvers:
            t0 = time.perf_counter()
            if s == "batched":
                bad = ctx.verify(args.n, A, B)
            else:
                r = ctx.solve_uniform(args.n, A, B, mode="plain")
                L = r.length
                over = np.nonzero(L > (args.n - 1) ** 2)[0]
                bad = int(over[0]) if len(over) else -1
            dt = time.perf_counter() - t0
            if bad >= 0:
                with open(args.failed, "a") as f:
                    f.write(problem_string(args.n, A[bad], B[bad]))
                print(f"Cerny Conjecture is false.\nProblem description: {problem_string(args.n, A[bad], B[bad])}")
                return 1
            rate = count / dt
            row.append(str(round(rate)))
            print(f"{names[s]}: {rate:,.0f} problems per second, {dt * 1e3 / count * 1e3:.4f} us per problem",
                  flush=True)
        rows.append(",".join(row))
        with open(args.csv, "w") as f:
            f.write("\n".join(rows) + "\n")
        nxt = int(round(count * args.factor))
        count = nxt if nxt > count else count + 1
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
