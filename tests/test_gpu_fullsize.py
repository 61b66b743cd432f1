"""GPU parity at the sizes the configurations run: the rknn-inf configuration
in full (1M users, 10k facilities, 100 queries, k=10), the clustered
configuration with ties at its full size (4M users, 50k facilities), the
batched configuration's dataset (10M users, 100k facilities) with its k mix,
the serving configuration's three datasets (1M, 10M and 50M users) served
together, and a dataset deep
enough for a 4-level tile tree (more than 64^4 = 16,777,216 users), as the
50M-user serving dataset has.  Every form -- plain, graph-captured batch,
device-filtered batch, one call, served -- against the FP64 oracle over the
host filter's candidate lists (the lists are sufficient: the filter tests
pin that against brute force)."""
import ctypes
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def oracle_sets(arrs, qs, ks, cands):
    ux, uy, uid, fx, fy, fid = arrs
    O.set_threads(16)
    return [O.rknn(ux, uy, uid, fx, fy, q, k, cand=c) for q, k, c in zip(qs, ks, cands)]


def served(rk, ctx, ds, qs, ks):
    srv = rk.Server(ctx, max_batch=64, max_wait_s=0.002)
    try:
        tickets = [srv.submit(ds, q, k) for q, k in zip(qs, ks)]
        res = {}
        deadline = time.time() + 90
        while len(res) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t not in res:
                    got = srv.poll(t, ds.nu)
                    if got is not None:
                        res[t] = got[0]
            time.sleep(0.001)
        assert len(res) == len(tickets)
        return [res[t] for t in tickets]
    finally:
        srv.stop()


def test_config2_full_size(rk, ctx):
    """rknn-inf form: 1,000,000 users, 10,000 facilities, 100 queries, k=10."""
    seed = rk.default_seed(2)
    arrs = rk.generate(2, 1_000_000, 10_000, seed=seed)
    qs = [int(q) for q in rk.generate_queries(10_000, 100, seed)]
    ks = [10] * 100
    ds = rk.Dataset(ctx, *arrs)
    try:
        cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
        exp = oracle_sets(arrs, qs, ks, cands)
        print("config 2 full size: |R| mean", np.mean([len(e) for e in exp]), flush=True)
        for i, (q, k, c) in enumerate(zip(qs, ks, cands)):
            assert np.array_equal(ds.query_plain(q, k, c), exp[i]), (q, k)
        b = ds.prepare_batch(list(zip(qs, ks, cands)))
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), qs[i]
        finally:
            b.close()
        for i, g in enumerate(ds.query_batch(list(zip(qs, ks)))):
            assert np.array_equal(g, exp[i]), qs[i]
    finally:
        ds.close()


def test_config3_full_size(rk, ctx):
    """Clustered form with near ties: 4,000,000 users and 50,000 facilities
    (Gaussian mixture and polylines on a 2^-7 grid, users placed exactly on
    the queries' bisectors or duplicated), 16 of its 200 queries at each
    k in {1, 5, 10, 25}: plain, graph-captured batch (host filter's lists)
    and device-filtered batch, against the oracle; the exact ties reach the
    binary64 re-check."""
    seed = rk.default_seed(3)
    nu, nf = 4_000_000, 50_000
    qs_all = [int(q) for q in rk.generate_queries(nf, 200, seed)]
    arrs = rk.generate(3, nu, nf, seed=seed, nq_ties=200, seed_ties=seed)
    qs = [q for q in qs_all[:16] for _ in range(4)]
    ks = [1, 5, 10, 25] * 16
    ds = rk.Dataset(ctx, *arrs)
    try:
        cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
        exp = oracle_sets(arrs, qs, ks, cands)
        print("config 3 full size: |C| mean", np.mean([len(c) for c in cands]), "|R| mean",
              np.mean([len(e) for e in exp]), flush=True)
        rech = 0
        for i, (q, k, c) in enumerate(zip(qs, ks, cands)):
            got, t = ds.query_plain(q, k, c, timing=True)
            assert np.array_equal(got, exp[i]), (q, k)
            rech += t.n_fp64_rechecks
        assert rech > 0  # the ties were settled in binary64
        b = ds.prepare_batch(list(zip(qs, ks, cands)))
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (qs[i], ks[i])
        finally:
            b.close()
        b = ds.prepare_batch(list(zip(qs, ks)), device_filter=True)
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (qs[i], ks[i])
        finally:
            b.close()
    finally:
        ds.close()


def test_config4_full_size(rk, ctx):
    """Batched form at its full size: 10,000,000 users and 100,000
    facilities (half uniform, half clustered with ties), 32 queries with k
    drawn uniformly from 1..50 as the configuration's batch has them, through
    the graph batch (host filter's lists), the device-filtered batch and the
    server, against the oracle."""
    seed = rk.default_seed(4)
    nu, nf = 10_000_000, 100_000
    arrs = rk.generate(4, nu, nf, seed=seed, nq_ties=32, seed_ties=seed)
    qs = [int(q) for q in rk.generate_queries(nf, 32, seed)]
    ks = [int(k) for k in np.random.default_rng(seed).integers(1, 51, size=32)]
    ds = rk.Dataset(ctx, *arrs)
    try:
        cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
        exp = oracle_sets(arrs, qs, ks, cands)
        print("config 4 full size: |C| mean", np.mean([len(c) for c in cands]), "|R| mean",
              np.mean([len(e) for e in exp]), flush=True)
        b = ds.prepare_batch(list(zip(qs, ks, cands)))
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (qs[i], ks[i])
        finally:
            b.close()
        b = ds.prepare_batch(list(zip(qs, ks)), device_filter=True)
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (qs[i], ks[i])
        finally:
            b.close()
        for i, g in enumerate(served(rk, ctx, ds, qs, ks)):
            assert np.array_equal(g, exp[i]), (qs[i], ks[i])
    finally:
        ds.close()


def test_config5_full_size(rk, ctx):
    """Serving configuration at its full size: three datasets resident
    together (1M users / 10k facilities, 10M / 100k, 50M / 200k; the
    config-4 generator), one server, requests of all three interleaved, k
    drawn from 1..50; every result against the oracle over the host filter's
    list (the server filters on the device)."""
    sizes = [(1_000_000, 10_000), (10_000_000, 100_000), (50_000_000, 200_000)]
    rng = np.random.default_rng(5)
    sets = []
    try:
        for j, (nu, nf) in enumerate(sizes):
            seed = rk.default_seed(4) + 1000 * (j + 1)
            arrs = rk.generate(4, nu, nf, seed=seed)
            ds = rk.Dataset(ctx, *arrs)
            qs = [int(q) for q in rk.generate_queries(nf, 4, seed)]
            ks = [int(k) for k in rng.integers(1, 51, size=4)]
            cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
            exp = oracle_sets(arrs, qs, ks, cands)
            del arrs
            sets.append((ds, qs, ks, exp))
        ds, qs, ks, exp = sets[2]  # 50M users: a 4-level tile tree
        cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
        b = ds.prepare_batch(list(zip(qs, ks, cands)))
        try:
            b.run()
            items, _, _ = b.stats()
            assert len(items) == 4, items
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (qs[i], ks[i])
        finally:
            b.close()
        srv = rk.Server(ctx, max_batch=64, max_wait_s=0.002)
        try:
            tickets = []
            for i in range(4):  # interleaved over the datasets
                for ds, qs, ks, exp in sets:
                    tickets.append((srv.submit(ds, qs[i], ks[i]), ds, exp[i], qs[i]))
            deadline = time.time() + 90
            left = list(tickets)
            while left and time.time() < deadline:
                nxt = []
                for t, ds, e, q in left:
                    got = srv.poll(t, ds.nu)
                    if got is None:
                        nxt.append((t, ds, e, q))
                    else:
                        assert np.array_equal(got[0], e), (ds.nu, q)
                left = nxt
                time.sleep(0.001)
            assert not left
        finally:
            srv.stop()
    finally:
        for ds, *_ in sets:
            ds.close()


@pytest.fixture(scope="module")
def deep(rk, ctx):
    """17M users over 20k facilities (config-4 generator: half uniform, half
    clustered with exact ties): a 4-level tile tree, results of ~850*k ids."""
    nu, nf = 17_000_000, 20_000
    seed = rk.default_seed(4)
    t0 = time.time()
    arrs = rk.generate(4, nu, nf, seed=seed, nq_ties=12, seed_ties=seed)
    qs = [int(q) for q in rk.generate_queries(nf, 12, seed)]
    ks = [1, 10, 50] * 4
    ds = rk.Dataset(ctx, *arrs)
    cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
    exp = oracle_sets(arrs, qs, ks, cands)
    print(f"deep dataset: {nu} users, {nf} facilities, |C| {[len(c) for c in cands]}, "
          f"|R| {[len(e) for e in exp]} ({time.time() - t0:.1f}s)", flush=True)
    yield arrs, ds, qs, ks, cands, exp
    ds.close()


def test_deep_tree_has_four_levels(deep):
    arrs, ds, qs, ks, cands, exp = deep
    b = ds.prepare_batch(list(zip(qs, ks, cands)))
    try:
        b.run()
        items, res, _ = b.stats()
        assert len(items) == 4, items
        assert all(n > 0 for n in items), items  # every level was descended
        assert res == sum(len(e) for e in exp)
        for i, g in enumerate(b.fetch()):
            assert np.array_equal(g, exp[i]), (qs[i], ks[i])
    finally:
        b.close()


def test_deep_plain(deep):
    arrs, ds, qs, ks, cands, exp = deep
    for q, k, c, e in zip(qs, ks, cands, exp):
        assert np.array_equal(ds.query_plain(q, k, c), e), (q, k)


def test_deep_device_filtered_batch(deep):
    arrs, ds, qs, ks, cands, exp = deep
    b = ds.prepare_batch(list(zip(qs, ks)), device_filter=True)
    try:
        b.run()
        for i, g in enumerate(b.fetch()):
            assert np.array_equal(g, exp[i]), (qs[i], ks[i])
    finally:
        b.close()


@pytest.mark.parametrize("filt", ["device", "host"])
def test_deep_one_call(deep, filt, monkeypatch):
    arrs, ds, qs, ks, cands, exp = deep
    monkeypatch.setenv("RK_FILTER", filt)
    for i, g in enumerate(ds.query_batch(list(zip(qs, ks)))):
        assert np.array_equal(g, exp[i]), (qs[i], ks[i])


def test_deep_served(rk, ctx, deep):
    arrs, ds, qs, ks, cands, exp = deep
    for i, g in enumerate(served(rk, ctx, ds, qs, ks)):
        assert np.array_equal(g, exp[i]), (qs[i], ks[i])


def test_deep_count_all(deep):
    """Count-all at 17M users: the counts cross PCIe through the pinned chunk
    ring (a pageable destination of 68 MB), in caller order and in the
    dataset's order, and the users with fewer than k closer candidates are
    exactly the result set."""
    arrs, ds, qs, ks, cands, exp = deep
    ux, uy, uid, fx, fy, fid = arrs
    order = ds.order()
    out = np.full(ds.nu, 0xFFFFFFFF, dtype=np.uint32)
    for i in (0, 1, 5):
        want = O.counts(ux, uy, fx, fy, qs[i], cand=cands[i])
        got = ds.count_all(qs[i], cands[i])
        assert np.array_equal(got, want), qs[i]
        assert np.array_equal(np.sort(uid[got < ks[i]]), exp[i]), qs[i]
        ds.count_all(qs[i], cands[i], sorted_order=True, out=out)
        assert np.array_equal(out, want[order]), qs[i]
    # a pinned destination (hipHostMalloc) takes one DMA instead of the ring
    hip = ctypes.CDLL("libamdhip64.so")
    ptr = ctypes.c_void_p()
    assert hip.hipHostMalloc(ctypes.byref(ptr), ctypes.c_size_t(4 * ds.nu), ctypes.c_uint(0)) == 0
    try:
        pinned = np.ctypeslib.as_array((ctypes.c_uint32 * ds.nu).from_address(ptr.value))
        pinned[:] = 0xFFFFFFFF
        want = O.counts(ux, uy, fx, fy, qs[2], cand=cands[2])
        ds.count_all(qs[2], cands[2], out=pinned)
        assert np.array_equal(pinned, want)
        ds.count_all(qs[2], cands[2], sorted_order=True, out=pinned)
        assert np.array_equal(pinned, want[order])
        del pinned
    finally:
        hip.hipHostFree(ptr)


def test_deep_count_all_many(deep):
    """Count-all over five queries in one call at 17M users: in the dataset's
    order each query's 32-bit counts take 68 MB of device memory, so the
    call runs two rounds (three queries, then two) and no launch counts more
    queries than a round holds; in caller order the 8-bit counts of all five
    fit one round.  Every row equals the oracle's counts."""
    arrs, ds, qs, ks, cands, exp = deep
    ux, uy, uid, fx, fy, fid = arrs
    order = ds.order()
    sel = [0, 1, 2, 3, 4]
    want = [O.counts(ux, uy, fx, fy, qs[i], cand=cands[i]) for i in sel]
    got, t = ds.count_all_many([(qs[i], cands[i]) for i in sel], timing=True, sorted_order=True)
    assert t.n_candidates == sum(len(cands[i]) for i in sel)
    assert t.n_launches >= 2
    for r, i in enumerate(sel):
        assert np.array_equal(got[r], want[r][order]), qs[i]
    del got
    got = ds.count_all_many([(qs[i], cands[i]) for i in sel])
    for r, i in enumerate(sel):
        assert np.array_equal(got[r], want[r]), qs[i]
