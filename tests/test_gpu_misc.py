"""Progress counters and callback, the Program.cs-style driver, tuning table
round trip, and contexts used from several host threads."""
import json
import os
import threading
import time

import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "gpgpu-binary-automaton-synchronizing-algorithms_amd")


def test_progress_counters_and_callback(ctx):
    seen = []
    before = ctx.progress()
    assert before["in_flight"] == 0
    ctx.set_progress_callback(seen.append, 5)
    n, delta, pos = W.uniform_n16(count=200_000)
    r = ctx.solve_packed(delta.reshape(-1), len(delta), n)
    after = ctx.progress()
    time.sleep(0.05)  # ten intervals after the solve ended
    ctx.set_progress_callback(None)  # returns after the library's thread stopped
    assert after["automata_done"] - before["automata_done"] == len(delta)
    assert after["subsets_expanded"] - before["subsets_expanded"] == int(r.expanded.sum())
    assert after["in_flight"] == 0
    # the solve is reported, however short against the interval, and the last
    # report carries its final counts
    assert seen
    done = [p["automata_done"] for p in seen]
    assert done == sorted(done)
    assert done[-1] == after["automata_done"]
    assert seen[-1]["in_flight"] == 0
    assert all(0 <= p["in_flight"] <= len(delta) for p in seen)


def test_progress_of_a_pipelined_host_call(ctx):
    # int32 host arrays of 2^19+ automata run as several solves in one call:
    # in flight counts the call's automata not handed back yet, never more
    # than the call, and is 0 once the call returned
    seen = []
    n, delta, pos = W.uniform_n16(count=600_000)
    A = np.ascontiguousarray(delta[:, :n]).astype(np.int32)
    B = np.ascontiguousarray(delta[:, n:]).astype(np.int32)
    before = ctx.progress()
    ctx.set_progress_callback(seen.append, 1)
    ctx.compute(n, A, B)
    after = ctx.progress()
    time.sleep(0.02)
    ctx.set_progress_callback(None)
    assert after["automata_done"] - before["automata_done"] == len(delta)
    assert after["in_flight"] == 0
    assert seen and seen[-1]["in_flight"] == 0
    assert all(0 <= p["in_flight"] <= len(delta) for p in seen)
    done = [p["automata_done"] for p in seen]
    assert done == sorted(done)


def test_progress_reports_a_solve_shorter_than_the_interval(ctx):
    # a solve that starts and ends between two looks of the library's thread
    # (interval 1 s, a few hundred microseconds of work) is still reported:
    # at the next look, or by the last look made when the callback is turned off
    seen = []
    ctx.set_progress_callback(seen.append, 1000)
    A, B = ssw.generate_uniform(12, 10, 1000)
    ctx.compute(10, A, B)
    final = ctx.progress()["automata_done"]
    ctx.set_progress_callback(None)
    assert seen and seen[-1]["automata_done"] == final and seen[-1]["in_flight"] == 0


def test_driver_writes_csv(tmp_path):
    import driver
    csv = tmp_path / "gpgpu.csv"
    rc = driver.main(["--n", "12", "--start", "4096", "--max", "8193", "--csv", str(csv),
                      "--failed", str(tmp_path / "cernyFailed.csv"), "--solvers", "batched,plain"])
    assert rc == 0
    lines = csv.read_text().strip().split("\n")
    assert lines[0] == "problemsize,problemcount,HipComputable,HipPlainComputable"
    assert [l.split(",")[1] for l in lines[1:]] == ["4096", "8192"]
    assert all(int(x) > 0 for l in lines[1:] for x in l.split(",")[2:])


def test_tuning_round_trip(ctx, tmp_path):
    path = str(tmp_path / "t.json")
    saved = {n: ctx.get_tuning(n) for n in (9, 10)}
    ctx.autotune(9, 10, 20_000, 5, path)
    table = json.load(open(path))
    assert set(table["tuning"]) == {"9", "10"}
    ctx.set_tuning(9, 64, 0, 1, 4096, 1)
    ctx.load_tuning(path)
    assert ctx.get_tuning(9)["group_lanes"] == table["tuning"]["9"]["group_lanes"]
    A, B = ssw.generate_uniform(1, 10, 5000)
    r = ctx.compute(10, A, B)
    np.testing.assert_array_equal(r.length, O.slim(10, A, B)[0])
    for n, t in saved.items():
        ctx.set_tuning(n, **t)


def test_contexts_on_threads():
    A, B = ssw.generate_uniform(2, 14, 20_000)
    want = O.slim(14, A, B, threads=8)[0]
    errs = []

    def work(k):
        try:
            with ssw.Context(0) as c:
                for _ in range(3):
                    r = c.compute(14, A[k::3], B[k::3])
                    if not np.array_equal(r.length, want[k::3]):
                        errs.append(k)
        except Exception as e:  # noqa: BLE001
            errs.append(repr(e))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs


def test_autotune_on_a_caller_batch(ctx):
    # ssw_autotune_batch times the candidate configurations on the caller's
    # device batch (here n = 14 with Cerny stragglers) and keeps the fastest;
    # results stay exact under the chosen entry
    saved = ctx.get_tuning(14)
    n, delta, pos = W.uniform_n16(seed=4242, count=20_000, n=14, cerny_every=200)
    count = len(delta)
    d = ctx.to_device(delta.reshape(-1))
    dl, de = ctx.alloc(4 * count), ctx.alloc(8 * count)
    try:
        b = ssw.Batch(count=count, uniform_n=n, witness_stride=0, n_per=None, offset=None, delta=d, length=dl,
                      sync=None, expanded=de, witness=None, bound=-1, reserved=0)
        ctx.autotune_batch(b)
        ctx.solve_device(b)
        L = ctx.to_host(dl, np.int32, count)
        E = ctx.to_host(de, np.int64, count)
    finally:
        for p in (d, dl, de):
            ctx.free(p)
        chosen = ctx.get_tuning(14)
        ctx.set_tuning(14, **saved)
    assert chosen["group_lanes"] in (16, 32, 64, 256)
    assert np.all(L[pos] == (n - 1) ** 2) and np.all(E[pos] == 2 ** n - n - 1)
    idx = np.union1d(np.arange(0, count, 37), pos[:5])
    Lo, _, Eo, _ = O.slim(n, delta[idx, :n].astype(np.int32), delta[idx, n:].astype(np.int32))
    np.testing.assert_array_equal(L[idx], Lo)
    np.testing.assert_array_equal(E[idx], Eo)


def test_one_context_shared_by_threads(ctx):
    # calls on one context from several threads are serialised by the library
    A, B = ssw.generate_uniform(3, 13, 30_000)
    want = O.slim(13, A, B, threads=8)[0]
    errs = []

    def work(k):
        try:
            for _ in range(3):
                r = ctx.compute(13, A[k::4], B[k::4])
                if not np.array_equal(r.length, want[k::4]):
                    errs.append(k)
        except Exception as e:  # noqa: BLE001
            errs.append(repr(e))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs


def test_small_budget_rejects_what_cannot_fit():
    # a context with a 32 MiB budget cannot hold C_28's 32 MiB bitmap plus its
    # ring: the device-wide path says so instead of failing an allocation
    with ssw.Context(0, budget=32 << 20) as c:
        a, b = ssw.generate_cerny(28, 1)
        with pytest.raises(ssw.SswError) as e:
            c.solve_uniform(28, a[None], b[None], mode="large")
        assert e.value.code == ssw.SSW_ERR_BUDGET
        # small automata still solve under the same budget
        A, B = ssw.generate_uniform(5, 12, 2000)
        np.testing.assert_array_equal(c.compute(12, A, B).length, O.slim(12, A, B)[0])
        # witness words need the largest slots (recorded levels per group):
        # the grid shrinks to the budget instead of failing
        for n, count in ((12, 300), (16, 40)):
            A, B = ssw.generate_uniform(6, n, count)
            ca, cb = ssw.generate_cerny(n, 1)
            A = np.concatenate([A, ca[None]]).astype(np.int32)
            B = np.concatenate([B, cb[None]]).astype(np.int32)
            r = c.compute(n, A, B, witness=True)
            L = O.slim(n, A, B)[0]
            np.testing.assert_array_equal(r.length, L, err_msg=f"n={n}")
            for i in range(len(A)):
                if L[i] >= 0:
                    w = ssw.witness_bits_to_word(r.witness[i], int(L[i]))
                    np.testing.assert_array_equal(w, O.witness(n, A[i], B[i])[1],
                                                  err_msg=f"n={n} automaton {i}")


def test_verify_bound_zero():
    # with bound 0 every synchronizing automaton violates: Verify returns the
    # first one (IComputable.Verify, SlimCPUUntweaked.cs:189-213,225-226)
    with ssw.Context(0) as c:
        A, B = ssw.generate_uniform(8, 9, 500)
        A[:3] = np.arange(9)  # a permutation...
        B[:3] = np.arange(9)  # ... and the identity: not synchronizing
        L, _, _, v = O.slim(9, A, B, bound=0)
        assert v == int(np.nonzero(L >= 1)[0][0]) >= 3
        assert c.verify(9, A, B, bound=0) == v


def test_tuning_validation_and_malformed_tables(ctx, tmp_path):
    saved = ctx.get_tuning(12)
    bad = [dict(group_lanes=48, block_threads=128, chunk=4, arena_lds=512),
           dict(group_lanes=64, block_threads=96, chunk=4, arena_lds=512),
           dict(group_lanes=64, block_threads=128, chunk=0, arena_lds=512),
           dict(group_lanes=64, block_threads=128, chunk=4, arena_lds=512, lane_entries=3),
           dict(group_lanes=64, block_threads=128, chunk=4, arena_lds=512, narrow_bits=20)]
    for t in bad:
        with pytest.raises(ssw.SswError) as e:
            ctx.set_tuning(12, **t)
        assert e.value.code == ssw.SSW_ERR_ARG, t
    assert ctx.get_tuning(12) == saved
    # a table that cannot be read or holds no tuning object is refused
    p = tmp_path / "broken.json"
    p.write_text("{ not json")
    with pytest.raises(ssw.SswError):
        ctx.load_tuning(str(p))
    with pytest.raises(ssw.SswError):
        ctx.load_tuning(str(tmp_path / "missing.json"))
    assert ctx.get_tuning(12) == saved
    A, B = ssw.generate_uniform(9, 12, 3000)
    np.testing.assert_array_equal(ctx.compute(12, A, B).length, O.slim(12, A, B)[0])


def test_server_tickets_and_timeouts(ctx):
    with ssw.Server(ctx, small_n_max=16) as srv:
        with pytest.raises(ssw.SswError) as e:
            srv.poll(123456789)
        assert e.value.code == ssw.SSW_ERR_NOTFOUND
        a, b = ssw.generate_cerny(30, 1)  # about 0.14 s of work
        tk = srv.submit(30, a, b)
        with pytest.raises(ssw.SswError) as e:
            srv.wait(tk, 1)
        assert e.value.code == ssw.SSW_ERR_TIMEOUT
        r = srv.wait(tk, 120000)  # the ticket survives a timed-out wait
        assert r["length"] == 29 * 29
        with pytest.raises(ssw.SswError):
            srv.submit(31, np.zeros(31, np.int32), np.zeros(31, np.int32))
        with pytest.raises(ssw.SswError):
            srv.submit(5, np.array([0, 1, 2, 3, 5], np.int32), np.zeros(5, np.int32))


def test_last_launch_stats(ctx):
    # the launches of one solve account for every automaton once and for the
    # batch's expanded counts; each has an event-timed duration.  The shared
    # context runs the default tuning (one launch for n = 16); a context with
    # the committed table runs the tuned three-launch spill chain.
    n, delta, pos = W.uniform_n16(seed=99, count=100_000)
    with ssw.Context(0) as tuned:
        tuned.load_tuning(os.path.join(PKG, "tuning_gfx950.json"))
        for c, chain in ((ctx, False), (tuned, True)):
            r = c.solve_packed(delta.reshape(-1), len(delta), n)
            st = c.last_launch_stats()
            assert len(st) >= (3 if chain else 1)
            assert sum(x["automata"] for x in st) == len(delta)
            assert sum(x["subsets"] for x in st) == int(r.expanded.sum())
            assert all(x["ms"] > 0 and x["grid"] > 0 and x["kernel"].startswith("SmallSolver") for x in st)
            # subsets actually expanded: the finished searches' E less the
            # unexpanded tails of their last levels, plus the partial
            # searches handed to a later launch
            assert 0.5 * int(r.expanded.sum()) < sum(x["expanded"] for x in st) < 2 * int(r.expanded.sum())
            if chain:
                # the Cerny automata (a permutation letter) run in the last launch
                assert st[-1]["automata"] >= len(pos) and st[-1]["subsets"] >= len(pos) * (2 ** n - n - 1)
                assert [x["pass"] for x in st] == [1] + [2] * (len(st) - 1)
        # the plain path records none
        tuned.solve_packed(delta[:10].reshape(-1), 10, n, mode="plain")
        assert tuned.last_launch_stats() == []


@pytest.mark.parametrize("n", [6, 9, 13, 16])
def test_verify_prefilter_same_violation(n):
    # ssw_verify with the length-6 word pre-filter returns what it returns
    # without it, and what the oracle says, for bounds above, at and below 6
    # (below 6 a word of length <= 6 proves nothing and the filter stays off)
    A, B = ssw.generate_uniform(70 + n, n, 20_000)
    if n >= 3:
        ca, cb = ssw.generate_cerny(n, 1)
        A[::501] = ca
        B[::501] = cb
    L = O.slim(n, A, B, threads=8)[0]
    with ssw.Context(0) as c:
        c.load_tuning(os.path.join(PKG, "tuning_gfx950.json"))
        for bound in (-1, 12, 7, 6, 5, 3, 1):
            want = O.slim(n, A, B, bound=bound, threads=8)[3]
            c.set_verify_prefilter(0)
            off = c.verify(n, A, B, bound=bound)
            c.set_verify_prefilter(6)
            on = c.verify(n, A, B, bound=bound)
            assert off == want and on == want, (n, bound, off, on, want)
        # the filter runs: the searches it skips expand nothing (about a third
        # of random automata have a reset word of length <= 6 for n >= 9)
        spent = {}
        for pf in (0, 6):
            c.set_verify_prefilter(pf)
            p0 = c.progress()
            c.verify(n, A, B)
            p1 = c.progress()
            assert p1["automata_done"] - p0["automata_done"] == len(A)
            spent[pf] = p1["subsets_expanded"] - p0["subsets_expanded"]
        if c.get_tuning(n)["group_lanes"] == 64:  # the filter needs one wavefront per automaton
            assert spent[6] < spent[0]
        else:
            assert spent[6] == spent[0]
        c.set_verify_prefilter(6)
        # Compute is unaffected by the option: lengths of every automaton
        np.testing.assert_array_equal(c.compute(n, A, B).length, L)
        with pytest.raises(ssw.SswError):
            c.set_verify_prefilter(5)
