"""GPU: queries that outgrow a batch's segments, and results larger than
the in-step sort.

A batch gives each query a result segment (the batch's result pool split
evenly) and, when filtered on the device, a candidate segment.  A query that
outgrows either is run again: with segments sized to the count the step
measured, or through the host filter.  Its result is then merged with the
others, which come from the step itself.  RK_BATCH_POOL_IDS and RK_DFILTER_SLOTS shrink the segments
so that small datasets reach both paths; every result must equal the plain
form's (pinned to the oracle by the parity tests) and, on a few, the
oracle's."""
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

NU, NF, NQ = 120_000, 1_500, 64


@pytest.fixture(scope="module")
def data(rk, ctx):
    seed = rk.default_seed(4)
    arrs = rk.generate(4, NU, NF, seed=seed, nq_ties=16, seed_ties=seed)
    ds = rk.Dataset(ctx, *arrs)
    qs = [int(q) for q in rk.generate_queries(NF, NQ, seed)]
    ks = [1 + (7 * i) % 64 for i in range(NQ)]
    plain = [ds.query(q, k) for q, k in zip(qs, ks)]
    yield arrs, ds, qs, ks, plain
    ds.close()


def check(arrs, qs, ks, got, plain):
    sizes = [len(g) for g in got]
    for q, k, g, p in zip(qs, ks, got, plain):
        assert np.array_equal(g, p), (q, k)
    ux, uy, uid, fx, fy, fid = arrs
    O.set_threads(16)
    for i in (0, 9, 31):
        assert np.array_equal(got[i], O.rknn(ux, uy, uid, fx, fy, qs[i], ks[i])), (qs[i], ks[i])
    return sizes


@pytest.mark.parametrize("filt", ["device", "host"])
def test_results_beyond_their_segment(rk, data, filt, monkeypatch):
    arrs, ds, qs, ks, plain = data
    monkeypatch.setenv("RK_FILTER", filt)
    monkeypatch.setenv("RK_BATCH_POOL_IDS", str(NQ * 200))  # 200 ids per query
    got = ds.query_batch(list(zip(qs, ks)))
    sizes = check(arrs, qs, ks, got, plain)
    over = sum(s > 200 for s in sizes)
    print(f"{filt}: {over} of {NQ} results beyond a 200-id segment, largest {max(sizes)}", flush=True)
    assert over >= NQ // 2  # the path under test was taken


def test_candidate_lists_beyond_their_segment(rk, data, monkeypatch):
    arrs, ds, qs, ks, plain = data
    monkeypatch.setenv("RK_FILTER", "device")
    monkeypatch.setenv("RK_DFILTER_SLOTS", "24")
    lists = ds.filter_device(qs, ks)
    over = sum(len(c) > 24 for c in lists)
    print(f"{over} of {NQ} device candidate lists beyond 24 slots", flush=True)
    assert over >= NQ // 4
    got = ds.query_batch(list(zip(qs, ks)))
    check(arrs, qs, ks, got, plain)


def test_served_beyond_both_segments(rk, ctx, data, monkeypatch):
    """The serving loop with both segments shrunk: every request's result
    equals the plain form's."""
    arrs, ds, qs, ks, plain = data
    monkeypatch.setenv("RK_BATCH_POOL_IDS", str(32 * 200))
    monkeypatch.setenv("RK_DFILTER_SLOTS", "24")
    r = np.random.default_rng(5)
    order = [int(i) for i in r.integers(0, NQ, 200)]
    srv = rk.Server(ctx, max_batch=32, max_wait_s=0.002)
    try:
        tickets = []
        for n, i in enumerate(order):
            tickets.append(srv.submit(ds, qs[i], ks[i]))
            if n % 40 == 39:
                time.sleep(0.002)
        results = {}
        deadline = time.time() + 90
        while len(results) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t not in results:
                    got = srv.poll(t, NU)
                    if got is not None:
                        results[t] = got
            time.sleep(0.001)
        assert len(results) == len(tickets)
    finally:
        srv.stop()
    for t, i in zip(tickets, order):
        assert np.array_equal(results[t][0], plain[i]), (qs[i], ks[i])


@pytest.mark.parametrize("negative_ids", [False, True])
@pytest.mark.parametrize("filt", ["device", "host"])
def test_large_results_sorted_on_the_device(rk, ctx, filt, negative_ids, monkeypatch):
    """Results beyond the traversal block's in-LDS sort (4,096 ids) are
    radix-sorted by the same block through its segment and the device
    landing area: 1,000 users per facility, k up to 64, ids positive (26
    bits) or spread over the whole int32 range with signs."""
    monkeypatch.setenv("RK_FILTER", filt)
    seed = rk.default_seed(2)
    ux, uy, uid, fx, fy, fid = rk.generate(2, 400_000, 400, seed=seed)
    if negative_ids:
        r = np.random.default_rng(11)
        uid = r.choice(np.arange(-(1 << 31), (1 << 31) - 1, 4099, dtype=np.int64), len(uid), replace=False)
        uid = uid.astype(np.int32)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    try:
        # 64 queries: over 2M result ids, so the host copy of the fetch is split over threads
        qs = [int(q) for q in rk.generate_queries(400, 64, seed)]
        ks = [8 + (i % 16) * 4 for i in range(64)]
        got = ds.query_batch(list(zip(qs, ks)))
        if filt == "host":  # a prepared batch: fetch and the in-place results agree with the one call
            b = ds.prepare_batch([(q, k, ds.filter(q, k)) for q, k in zip(qs, ks)])
            try:
                b.run()
                for g, f, v in zip(got, b.fetch(), b.results()):
                    assert np.array_equal(g, f) and np.array_equal(g, v)
            finally:
                b.close()
        sizes = [len(g) for g in got]
        print(f"{filt} negative={negative_ids}: sizes {min(sizes)}..{max(sizes)}, total {sum(sizes)}, "
              f"{sum(s > 8192 for s in sizes)} beyond 8192", flush=True)
        assert sum(s > 8192 for s in sizes) >= 32
        assert sum(sizes) > 2 << 20
        for q, k, g in zip(qs, ks, got):
            assert np.all(g[1:] > g[:-1])
            assert np.array_equal(g, ds.query(q, k)), (q, k)
        O.set_threads(16)
        for i in (3, 15, 63):
            assert np.array_equal(got[i], O.rknn(ux, uy, uid, fx, fy, qs[i], ks[i]))
    finally:
        ds.close()


def test_fetch_refuses_a_candidate_overflow(rk, data, monkeypatch):
    """A device-filtered batch whose candidate list outgrew its segment: the
    run reports RK_ERR_CAPACITY and a plain fetch of that run is refused (its
    truncated lists would give too many results); a later run that fits is
    fetched normally."""
    arrs, ds, qs, ks, plain = data
    monkeypatch.setenv("RK_DFILTER_SLOTS", "24")
    b = ds.prepare_batch(list(zip(qs, ks)), device_filter=True)
    try:
        with pytest.raises(rk.RkError) as e:
            b.run()
        assert e.value.status == 4
        with pytest.raises(rk.RkError) as e:
            b.fetch()
        assert e.value.status == 4
    finally:
        b.close()
    monkeypatch.delenv("RK_DFILTER_SLOTS")
    idx = list(range(8))
    b = ds.prepare_batch([(qs[i], ks[i]) for i in idx], device_filter=True)
    try:
        b.run()
        for i, g in zip(idx, b.fetch()):
            assert np.array_equal(g, plain[i]), (qs[i], ks[i])
    finally:
        b.close()
