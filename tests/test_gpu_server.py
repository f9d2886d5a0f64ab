"""Serving loop (continuous batching): results in a mixed stream equal the
results of the same requests solved alone, and the oracle's."""
import time

import numpy as np
import pytest

import oracle as O
import ssw
import workloads as W

pytestmark = pytest.mark.gpu


def _stream(count, n_hi, seed=7):
    reqs = W.serving(seed=seed, rate=2000.0, count=count, n_lo=6, n_hi=n_hi, cerny_frac=0.2)
    return reqs


def test_mixed_stream_equals_alone(ctx):
    reqs = _stream(160, 24)
    with ssw.Server(ctx, small_n_max=16, levels_per_tick=8, want_witness=True) as srv:
        t0 = time.perf_counter()
        tickets = []
        for (t, n, a, b, kind) in reqs:
            # replay the Poisson arrival times so requests join while others run
            while time.perf_counter() - t0 < t:
                time.sleep(0.0002)
            tickets.append(srv.submit(n, a, b))
        mixed = [srv.wait(tk, 120000) for tk in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["p99_us"] >= st["p50_us"] > 0
    # alone: a fresh server, one request at a time
    with ssw.Server(ctx, small_n_max=16, levels_per_tick=8, want_witness=True) as srv:
        for k in range(0, len(reqs), 7):
            _, n, a, b, _ = reqs[k]
            alone = srv.wait(srv.submit(n, a, b), 120000)
            m = mixed[k]
            assert (alone["length"], alone["expanded"]) == (m["length"], m["expanded"]), f"request {k}"
            if m["length"] >= 0:
                np.testing.assert_array_equal(alone["witness"], m["witness"])
    # and the oracle
    for k, (_, n, a, b, kind) in enumerate(reqs):
        L, S, E, _ = O.solve(n, np.asarray(a)[None], np.asarray(b)[None])
        assert (mixed[k]["length"], mixed[k]["expanded"]) == (L[0], E[0]), f"request {k} n={n} {kind}"
        if L[0] >= 0:
            assert O.verify_word(n, a, b, mixed[k]["witness"])
            if n <= 20:
                np.testing.assert_array_equal(mixed[k]["witness"], O.witness(n, a, b)[1])


def test_admission_rejects_oversized(ctx):
    with ssw.Server(ctx, mem_budget=1 << 20, small_n_max=16) as srv:
        a, b = ssw.generate_cerny(30, 1)
        with pytest.raises(ssw.SswError) as e:
            srv.submit(30, a, b)
        assert e.value.code == ssw.SSW_ERR_BUDGET
        a, b = ssw.generate_cerny(12, 1)
        tk = srv.submit(12, a, b)
        r = srv.wait(tk)
        assert r["length"] == 121
        assert srv.stats()["rejected"] == 1
        # a result is handed out once; the ticket is then released
        with pytest.raises(ssw.SswError) as e:
            srv.poll(tk)
        assert e.value.code == ssw.SSW_ERR_NOTFOUND


def test_large_requests_overlap(ctx):
    # many large requests in flight at once all finish, each exact
    reqs = []
    for i in range(12):
        n = 18 + (i % 5)
        A, B = ssw.generate_uniform(300 + i, n, 1)
        reqs.append((n, A[0], B[0]))
    a, b = ssw.generate_cerny(21, 1)
    reqs.append((21, a, b))
    with ssw.Server(ctx, small_n_max=12, levels_per_tick=4, max_large=16) as srv:
        tickets = [srv.submit(n, a, b) for (n, a, b) in reqs]
        res = [srv.wait(t, 120000) for t in tickets]
    for (n, a, b), r in zip(reqs, res):
        L, S, E, _ = O.solve(n, np.asarray(a)[None], np.asarray(b)[None])
        assert (r["length"], r["expanded"]) == (L[0], E[0])


@pytest.mark.parametrize("want_witness", [False, True])
def test_everything_in_the_device_wide_engine(ctx, want_witness):
    # small_n_max = 0 routes every request, n = 1 included, to the device-wide
    # engine; results equal the oracle's
    with ssw.Server(ctx, small_n_max=0, want_witness=want_witness) as srv:
        reqs = []
        for n in range(1, 9):
            A, B = ssw.generate_uniform(600 + n, n, 6)
            for i in range(len(A)):
                reqs.append((n, A[i], B[i], srv.submit(n, A[i], B[i])))
        for n, a, b, tk in reqs:
            r = srv.wait(tk, 60000)
            L, S, E, _ = O.slim(n, a[None], b[None])
            assert r["length"] == L[0] and r["expanded"] == E[0] and r["sync"] == bool(S[0]), (n, r)
            if want_witness and L[0] >= 0:
                np.testing.assert_array_equal(r["witness"], O.witness(n, a, b)[1])


def test_destroy_with_requests_in_flight(ctx):
    # destroying the server while requests are queued or running drains them
    # and returns; nothing leaks into the next server
    srv = ssw.Server(ctx, small_n_max=16)
    for req in _stream(200, 24, seed=11):
        _, n, a, b, _ = req
        srv.submit(n, a, b)
    t0 = time.perf_counter()
    srv.close()
    assert time.perf_counter() - t0 < 60
    with ssw.Server(ctx, small_n_max=16) as srv2:
        a, b = ssw.generate_cerny(9, 1)
        assert srv2.wait(srv2.submit(9, a, b))["length"] == 64
        assert srv2.stats()["completed"] == 1


def test_cancel_queued_requests(ctx):
    # one large instance at a time: behind a long C_26 the other large
    # requests wait in the queue, so cancels of them win (their tickets are
    # then gone); the running one cannot be cancelled and its result arrives;
    # what is not cancelled equals the oracle
    rng = np.random.default_rng(23)
    with ssw.Server(ctx, small_n_max=16, max_large=1) as srv:
        a, b = ssw.generate_cerny(26, 0)
        first = srv.submit(26, a, b)
        reqs, tickets = [], []
        for i in range(20):
            n = 21 + i % 2
            a = rng.integers(0, n, size=n).astype(np.int32)
            b = rng.integers(0, n, size=n).astype(np.int32)
            reqs.append((n, a, b))
            tickets.append(srv.submit(n, a, b))
        time.sleep(0.002)  # the engine has taken the first request
        cancelled = {i for i in range(0, 20, 2) if srv.cancel(tickets[i])}
        assert len(cancelled) >= 8, cancelled
        assert srv.cancel(first) is False
        assert srv.wait(first)["length"] == 25 * 25
        for i in range(20):
            if i in cancelled:
                with pytest.raises(ssw.SswError) as e:
                    srv.poll(tickets[i])
                assert e.value.code == ssw.SSW_ERR_NOTFOUND
                continue
            n, a, b = reqs[i]
            r = srv.wait(tickets[i])
            L, S, E, _ = O.solve(n, a[None, :], b[None, :])
            assert r["length"] == L[0] and r["expanded"] == E[0], (i, n, r)
        assert srv.stats()["completed"] == 21 - len(cancelled)
