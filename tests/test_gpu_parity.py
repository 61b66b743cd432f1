"""GPU parity: every query form of the gfx950 engine against the FP64 oracle
and the committed golden fixtures.  Integer/index work: bit-exact sets."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(name):
    d = np.load(os.path.join(HERE, "golden", name))
    res = [d["ids"][d["offs"][i]: d["offs"][i + 1]] for i in range(len(d["qk"]))]
    return d, [(int(q), int(k)) for q, k in d["qk"]], res


@pytest.fixture(scope="module")
def g1(rk, ctx):
    d, qk, res = golden("config1.npz")
    ds = rk.Dataset(ctx, d["ux"], d["uy"], d["uid"], d["fx"], d["fy"], d["fid"])
    yield d, ds, qk, res
    ds.close()


@pytest.fixture(scope="module")
def gties(rk, ctx):
    d, qk, res = golden("ties.npz")
    ds = rk.Dataset(ctx, d["ux"], d["uy"], d["uid"], d["fx"], d["fy"], d["fid"])
    yield d, ds, qk, res
    ds.close()


def test_config1_bruteforce_golden(g1):
    d, ds, qk, res = g1
    for (q, k), exp in zip(qk, res):
        print("config1 bruteforce q", q, "k", k, flush=True)
        assert np.array_equal(ds.query_bruteforce(q, k), exp)


def test_config1_filtered_golden(g1):
    d, ds, qk, res = g1
    for (q, k), exp in zip(qk, res):
        cand = ds.filter(q, k)
        assert np.array_equal(ds.query_plain(q, k, cand), exp)
        assert np.array_equal(ds.query(q, k), exp)


def test_config1_batch_golden(g1):
    d, ds, qk, res = g1
    got = ds.query_batch([(q, k) for q, k in qk])
    for g, exp in zip(got, res):
        assert np.array_equal(g, exp)
    # explicit candidate lists, brute-force lists, in one batch
    mixed = [(q, k, np.arange(ds.nf, dtype=np.uint32)) for q, k in qk] + [(q, k) for q, k in qk]
    got = ds.query_batch(mixed)
    for i, g in enumerate(got):
        assert np.array_equal(g, res[i % len(res)])


def test_batch_enqueue_wait_golden(rk, g1, gties):
    """rk_batch_enqueue / rk_batch_wait: batches of two datasets, graph-
    captured and device-filtered, four runs in flight at once, each on its
    batch's own stream; every fetch equals the golden sets.  Misuse is an
    error, not a hang."""
    bs = []
    for d, ds, qk, res in (g1, gties):
        keep = [(q, k) for q, k in qk if k <= 64]
        exp = [r for (q, k), r in zip(qk, res) if k <= 64]
        bs.append((ds.prepare_batch([(q, k, ds.filter(q, k)) for q, k in keep]), exp))
        bs.append((ds.prepare_batch(keep, device_filter=True), exp))
    for rnd in range(3):
        for b, _ in bs:
            b.enqueue()
        for b, _ in bs:
            assert b.wait() > 0.0
        for b, exp in bs:
            for g, e in zip(b.fetch(), exp):
                assert np.array_equal(g, e), rnd
            for g, e in zip(b.results(), exp):  # in place, in the pinned landing area
                assert np.array_equal(g, e), rnd
    b = bs[0][0]
    with pytest.raises(rk.RkError) as e:
        b.wait()  # nothing enqueued
    assert e.value.status == 2
    b.enqueue()
    with pytest.raises(rk.RkError) as e:
        b.enqueue()  # one run in flight per batch
    assert e.value.status == 2
    with pytest.raises(rk.RkError) as e:
        b.fetch()  # not waited for
    assert e.value.status == 2
    b.wait()
    for g, e in zip(b.fetch(), bs[0][1]):
        assert np.array_equal(g, e)
    for b, _ in bs:
        b.close()


def test_ties_golden_every_form(gties):
    d, ds, qk, res = gties
    rechecks = 0
    for (q, k), exp in zip(qk, res):
        print("ties q", q, "k", k, flush=True)
        ids, t = ds.query(q, k, timing=True)
        rechecks += t.n_fp64_rechecks
        assert np.array_equal(ids, exp)
        assert np.array_equal(ds.query_bruteforce(q, k), exp)
    assert rechecks > 0  # users on bisectors reached the binary64 path
    got = ds.query_batch(qk)
    for g, exp in zip(got, res):
        assert np.array_equal(g, exp)


@pytest.mark.parametrize("kernel", ["valu4", "valu6", "valu8", "tuned"])
def test_count_all_matches_oracle(g1, gties, kernel, monkeypatch):
    """Dense per-user counts, every count-all kernel (4 or 8 users per
    lane) and the autotuned choice, against the oracle's counts."""
    if kernel != "tuned":
        monkeypatch.setenv("RK_COUNT_KERNEL", kernel)
    for d, ds, qk, _ in (g1, gties):
        for q, k in qk[:6]:
            cand = ds.filter(q, k)
            exp = O.counts(d["ux"], d["uy"], d["fx"], d["fy"], q, cand=cand)
            assert np.array_equal(ds.count_all(q, cand), exp)
            allf = np.arange(ds.nf, dtype=np.uint32)
            exp = O.counts(d["ux"], d["uy"], d["fx"], d["fy"], q, cand=allf)
            assert np.array_equal(ds.count_all(q, allf), exp)
            # the dataset-order form: the same counts, permuted by order()
            assert np.array_equal(ds.count_all(q, allf, sorted_order=True), exp[ds.order()])


@pytest.mark.parametrize("kernel", ["valu4", "valu6", "valu8"])
def test_count_all_wide_counts(rk, ctx, kernel, monkeypatch):
    """65,536 candidates or more: counts no longer fit 16 bits, so the kernel
    writes 32-bit ones (and the dataset-order form always does)."""
    monkeypatch.setenv("RK_COUNT_KERNEL", kernel)
    r = np.random.default_rng(17)
    nu, nf = 3000, 70_000
    ux, uy = (r.random(nu) * 1000).astype(np.float32), (r.random(nu) * 1000).astype(np.float32)
    fx, fy = (r.random(nf) * 1000).astype(np.float32), (r.random(nf) * 1000).astype(np.float32)
    fx[0], fy[0] = 5000.0, 5000.0  # far outside the users' box: nearly every facility is closer
    uid = np.arange(nf, nf + nu, dtype=np.int32)
    fid = np.arange(nf, dtype=np.int32)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    try:
        allf = np.arange(nf, dtype=np.uint32)
        for q in (0, 12345):
            exp = O.counts(ux, uy, fx, fy, q, cand=allf)
            if q == 0:
                assert exp.min() >= 65536  # every count needs more than 16 bits
            assert np.array_equal(ds.count_all(q, allf), exp)
            assert np.array_equal(ds.count_all(q, allf, sorted_order=True), exp[ds.order()])
    finally:
        ds.close()


def test_dataset_order_is_a_permutation(g1):
    d, ds, _, _ = g1
    order = ds.order()
    assert len(order) == ds.nu
    assert np.array_equal(np.sort(order), np.arange(ds.nu, dtype=np.uint32))


def test_hand_cases_on_gpu(rk, ctx):
    from test_oracle import hand_cases

    for U, F, q, k, exp in hand_cases():
        U = np.array(U, np.float32)
        F = np.array(F, np.float32)
        uid = np.arange(100, 100 + len(U), dtype=np.int32)
        ds = rk.Dataset(ctx, U[:, 0], U[:, 1], uid, F[:, 0], F[:, 1], np.arange(len(F), dtype=np.int32))
        try:
            assert list(ds.query_bruteforce(q, k)) == exp
            assert list(ds.query(q, k)) == exp
            assert list(ds.query_batch([(q, k)])[0]) == exp
        finally:
            ds.close()


def test_edge_cases(rk, ctx):
    r = np.random.default_rng(5)
    nu, nf = 1000, 50
    ux, uy = r.random(nu, np.float32) * 100, r.random(nu, np.float32) * 100
    uid = np.arange(nf, nf + nu, dtype=np.int32)
    fx, fy = r.random(nf, np.float32) * 100, r.random(nf, np.float32) * 100
    fid = np.arange(nf, dtype=np.int32)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    try:
        # zero candidates: every user qualifies (rknn-inf-gpu returns none here)
        assert np.array_equal(ds.query_plain(3, 1, np.zeros(0, np.uint32)), np.sort(uid))
        assert np.array_equal(ds.query_batch([(3, 1, np.zeros(0, np.uint32))])[0], np.sort(uid))
        # k >= N_f: everyone
        assert np.array_equal(ds.query(3, nf), np.sort(uid))
        assert np.array_equal(ds.query_bruteforce(3, 10 ** 6), np.sort(uid))
        # candidate list holding q itself and duplicates of q are ignored
        exp = O.rknn(ux, uy, uid, fx, fy, 3, 2, cand=np.array([3, 3, 7, 9], np.uint32))
        assert np.array_equal(ds.query_plain(3, 2, np.array([3, 3, 7, 9], np.uint32)), exp)
        # invalid arguments
        for bad in [(nf, 1), (0, 0)]:
            with pytest.raises(rk.RkError) as e:
                ds.query(*bad)
            assert e.value.status == 2
        with pytest.raises(rk.RkError):
            ds.query_plain(0, 1, np.array([nf + 5], np.uint32))
        with pytest.raises(rk.RkError):
            ds.query_batch([(nf, 1)])
    finally:
        ds.close()
    # one user; all users at one point; facilities co-located with q
    for U, F in [([(5.0, 5.0)], [(0.0, 0.0), (10.0, 10.0)]),
                 ([(1.0, 1.0)] * 130, [(0.0, 0.0), (2.0, 2.0), (1.0, 0.0)]),
                 ([(3.0, 4.0), (0.5, 0.5), (7.0, 1.0)] * 50, [(1.0, 1.0), (1.0, 1.0), (1.0, 1.0), (5.0, 5.0)])]:
        U, F = np.array(U, np.float32), np.array(F, np.float32)
        uid = np.arange(len(U), dtype=np.int32) * 3 + 7
        ds = rk.Dataset(ctx, U[:, 0], U[:, 1], uid, F[:, 0], F[:, 1], np.arange(len(F), dtype=np.int32))
        try:
            for q in range(len(F)):
                for k in (1, 2, 3):
                    exp = O.rknn(U[:, 0], U[:, 1], uid, F[:, 0], F[:, 1], q, k)
                    assert np.array_equal(ds.query(q, k), exp)
                    assert np.array_equal(ds.query_bruteforce(q, k), exp)
                    assert np.array_equal(ds.query_batch([(q, k)])[0], exp)
        finally:
            ds.close()


@pytest.mark.parametrize("config", [2, 3, 4])
def test_medium_random_vs_oracle(rk, ctx, config):
    nu, nf = 200_000, 5_000
    seed = rk.default_seed(config)
    ux, uy, uid, fx, fy, fid = rk.generate(config, nu, nf, seed=seed, nq_ties=16, seed_ties=seed)
    qs = rk.generate_queries(nf, 16, seed)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    try:
        ks = [1, 5, 10, 25, 50, 3, 7, 10, 10, 2, 40, 10, 15, 20, 30, 10]
        batch = ds.query_batch([(int(q), k) for q, k in zip(qs, ks)])
        for i, (q, k) in enumerate(zip(qs, ks)):
            print("config", config, "q", int(q), "k", k, flush=True)
            cand = ds.filter(int(q), k)
            exp = O.rknn(ux, uy, uid, fx, fy, int(q), k, cand=cand)
            got = ds.query_plain(int(q), k, cand)
            assert np.array_equal(got, exp)
            assert np.array_equal(batch[i], exp)
            if i < 4:  # the filter is sufficient: same set as brute force
                assert np.array_equal(O.rknn(ux, uy, uid, fx, fy, int(q), k), exp)
    finally:
        ds.close()


@pytest.mark.parametrize("form", ["one-call", "4-wave"])
def test_batch_sort_paths(rk, ctx, form, monkeypatch):
    """Result segments of every size class and id layout against the oracle:
    up to 4,096 ids (2,048 in the 4-wave shape) bucket-sorted in LDS by the
    traversal block, or radix-sorted by it when they bunch unevenly in their
    range; beyond that radix-sorted by the block through its segment; random,
    spatially coherent with far outliers, and negative ids.  The one-call
    form (device filter, 8 waves per query) and a prepared batch with host
    candidate lists in the 4-wave shape."""
    r = np.random.default_rng(11)
    nu, nf = 300_000, 200
    ux = (r.random(nu) * 1000).astype(np.float32)
    uy = (r.random(nu) * 1000).astype(np.float32)
    fx = (r.random(nf) * 1000).astype(np.float32)
    fy = (r.random(nf) * 1000).astype(np.float32)
    fid = np.arange(nf, dtype=np.int32)
    xrank = np.empty(nu, np.int64)
    xrank[np.argsort(ux, kind="stable")] = np.arange(nu)
    layouts = {
        "random": r.permutation(nu),
        "coherent+outliers": np.where(r.random(nu) < 0.01, xrank + (1 << 28), xrank),
        "negative": -xrank - 1,
    }
    qs = r.choice(nf, 8, replace=False)
    ks = [1, 2, 3, 5, 8, 3, 5, 1]
    O.set_threads(16)
    sizes = []
    for name, ids in layouts.items():
        uid = ids.astype(np.int32)
        ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
        try:
            print("sort paths:", form, "ids", name, flush=True)
            if form == "one-call":
                got = ds.query_batch([(int(q), k) for q, k in zip(qs, ks)])
            else:
                monkeypatch.setenv("RK_AUTOTUNE", "4")
                b = ds.prepare_batch([(int(q), k, ds.filter(int(q), k)) for q, k in zip(qs, ks)])
                try:
                    assert b.waves_per_query() == 4
                    b.run()
                    got = b.fetch()
                finally:
                    b.close()
            for q, k, g in zip(qs, ks, got):
                exp = O.rknn(ux, uy, uid, fx, fy, int(q), k)
                sizes.append(len(exp))
                assert np.array_equal(g, exp), (name, int(q), k)
        finally:
            ds.close()
    sizes = np.array(sizes)
    assert (sizes <= 2048).any() and ((sizes > 2048) & (sizes <= 4096)).any()
    assert ((sizes > 4096) & (sizes <= 8192)).any() and (sizes > 8192).any()


def test_autotuned_shapes_agree(rk, ctx, tmp_path, monkeypatch):
    """Every block shape of the traversal kernel returns the oracle's sets;
    the autotuner times them on the batch, records its choice in the JSON
    cache and reuses it."""
    seed = rk.default_seed(4)
    nu, nf, nq = 200_000, 5_000, 96
    ux, uy, uid, fx, fy, fid = rk.generate(4, nu, nf, seed=seed, nq_ties=nq, seed_ties=seed)
    qs = [int(q) for q in rk.generate_queries(nf, nq, seed)]
    ks = [1 + (7 * i) % 50 for i in range(nq)]
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    try:
        exp = [O.rknn(ux, uy, uid, fx, fy, q, k, cand=ds.filter(q, k)) for q, k in zip(qs[:12], ks)]
        for shape in ("8", "4"):
            monkeypatch.setenv("RK_AUTOTUNE", shape)
            b = ds.prepare_batch([(q, k) for q, k in zip(qs, ks)])
            assert b.waves_per_query() == int(shape)
            b.run()
            got = b.fetch()
            b.close()
            for g, e in zip(got, exp):
                assert np.array_equal(g, e), shape
        monkeypatch.delenv("RK_AUTOTUNE")
        cache = tmp_path / "autotune.json"
        monkeypatch.setenv("RK_TUNE_FILE", str(cache))
        b = ds.prepare_batch([(q, k) for q, k in zip(qs, ks)])
        w = b.waves_per_query()
        assert w in (4, 8)
        b.run()
        for g, e in zip(b.fetch(), exp):
            assert np.array_equal(g, e)
        b.close()
        text = cache.read_text()
        assert '"variant": "%d waves per query"' % w in text and "/batch/" in text and "/q128/" in text
    finally:
        ds.close()


@pytest.mark.parametrize("kernel", ["valu4", "valu6", "valu8", "tuned"])
def test_count_all_many_matches_oracle(g1, gties, kernel, monkeypatch):
    """Many queries' counts in one call (count kernels back to back): lists
    from the filter (8-bit counts), every facility (16-bit), a query left to
    the host filter (cand None) and an empty list, in both orders, each row
    against the oracle's counts; and the call's argument checks."""
    if kernel != "tuned":
        monkeypatch.setenv("RK_COUNT_KERNEL", kernel)
    for d, ds, qk, _ in (g1, gties):
        allf = np.arange(ds.nf, dtype=np.uint32)
        queries, want = [], []
        for j, (q, k) in enumerate(qk[:5]):
            if j == 2:
                queries.append((q, k, None))
                cand = ds.filter(q, k)
            else:
                cand = ds.filter(q, k) if j % 2 == 0 else allf
                queries.append((q, cand))
            want.append(O.counts(d["ux"], d["uy"], d["fx"], d["fy"], q, cand=cand))
        q0 = qk[0][0]
        queries.append((q0, np.zeros(0, dtype=np.uint32)))
        want.append(np.zeros(ds.nu, dtype=np.uint32))
        got, t = ds.count_all_many(queries, timing=True)
        assert got.shape == (len(queries), ds.nu)
        for r in range(len(queries)):
            assert np.array_equal(got[r], want[r]), (kernel, r)
        order = ds.order()
        out = np.full((len(queries), ds.nu + 3), 0xFFFFFFFF, dtype=np.uint32)  # stride beyond nu
        ds.count_all_many(queries, sorted_order=True, out=out)
        for r in range(len(queries)):
            assert np.array_equal(out[r, : ds.nu], want[r][order]), (kernel, r)
            assert (out[r, ds.nu:] == 0xFFFFFFFF).all()
    d, ds, qk, _ = g1
    lib, h = ds._lib, ds.h
    descs, keep = ds._descs([(qk[0][0], 5)])
    buf = np.zeros(ds.nu, dtype=np.uint32)
    import ctypes as C
    p = buf.ctypes.data_as(C.POINTER(C.c_uint32))
    assert lib.rk_count_all_many(h, descs, 1, p, ds.nu - 1, 0, None) != 0  # stride below nu
    assert lib.rk_count_all_many(h, descs, 1, None, ds.nu, 0, None) != 0  # NULL output
    assert lib.rk_count_all_many(h, None, 1, p, ds.nu, 0, None) != 0  # NULL queries
    assert lib.rk_count_all_many(h, descs, 0, None, 0, 0, None) == 0  # nothing to do
    bad, keep2 = ds._descs([(ds.nf, 5)])
    assert lib.rk_count_all_many(h, bad, 1, p, ds.nu, 0, None) != 0  # q out of range


@pytest.mark.parametrize("kernel", ["valu4", "valu6", "valu8"])
def test_count_all_count_widths(rk, ctx, kernel, monkeypatch):
    """The count kernel stores 8-bit counts below 256 candidates and 16-bit
    ones from 256: lists of 255 and 256 candidates whose every count is the
    list's length (q far outside the users' box, so every candidate is
    closer), and a 255-candidate list over uniform data, in both orders."""
    monkeypatch.setenv("RK_COUNT_KERNEL", kernel)
    r = np.random.default_rng(23)
    nu, nf = 5000, 600
    ux, uy = (r.random(nu) * 1000).astype(np.float32), (r.random(nu) * 1000).astype(np.float32)
    fx, fy = (r.random(nf) * 1000).astype(np.float32), (r.random(nf) * 1000).astype(np.float32)
    fx[0], fy[0] = 9000.0, 9000.0
    uid = np.arange(nf, nf + nu, dtype=np.int32)
    fid = np.arange(nf, dtype=np.int32)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    try:
        order = ds.order()
        for q, cand in ((0, np.arange(1, 256, dtype=np.uint32)), (0, np.arange(1, 257, dtype=np.uint32)),
                        (7, np.arange(300, 555, dtype=np.uint32))):
            exp = O.counts(ux, uy, fx, fy, q, cand=cand)
            if q == 0:
                assert exp.min() == len(cand)  # the widest count the width holds
            got, t = ds.count_all(q, cand, timing=True)
            assert t.n_candidates == len(cand)
            assert np.array_equal(got, exp), (kernel, q, len(cand))
            assert np.array_equal(ds.count_all(q, cand, sorted_order=True), exp[order]), (kernel, q, len(cand))
    finally:
        ds.close()


@pytest.mark.parametrize("kernel", ["valu4", "valu6", "valu8"])
def test_count_all_large_vs_oracle(rk, ctx, kernel, monkeypatch):
    """Count-all on clustered data with exact bisector ties (config 3) and
    thousands of candidates: every per-user count equals the oracle's."""
    monkeypatch.setenv("RK_COUNT_KERNEL", kernel)
    seed = rk.default_seed(3)
    nu, nf = 60_000, 3_000
    ux, uy, uid, fx, fy, fid = rk.generate(3, nu, nf, seed=seed, nq_ties=4, seed_ties=seed)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    try:
        for q in rk.generate_queries(nf, 4, seed)[:3]:
            q = int(q)
            order = ds.order()
            for cand in (ds.filter(q, 25), np.arange(nf, dtype=np.uint32)):
                exp = O.counts(ux, uy, fx, fy, q, cand=cand)
                got, t = ds.count_all(q, cand, timing=True)
                assert np.array_equal(got, exp), (kernel, q, len(cand))
                got = ds.count_all(q, cand, sorted_order=True)
                assert np.array_equal(got, exp[order]), (kernel, q, len(cand))
    finally:
        ds.close()


@pytest.mark.parametrize("kernel", ["valu4", "valu6", "valu8"])
def test_count_all_query_sets(rk, ctx, kernel, monkeypatch):
    """rk_count_all_many counts consecutive queries whose lists fit the LDS
    staging together (at most 16 queries and 1,024 records) in one launch,
    all of them at the set's widest count type; a longer list runs alone.
    Clustered data with exact bisector ties (the banded path and its binary64
    re-check): a full set of 16 lists, a set mixing an 8-bit (255) and a
    16-bit (256) list, a 1,100-candidate list, an empty list, and lists that
    overflow the staging together.  Every row equals the oracle's counts in
    both orders at set sizes 1, 3 and 16, and the launches are the sets':
    23, 9 and 5."""
    monkeypatch.setenv("RK_COUNT_KERNEL", kernel)
    seed = rk.default_seed(3)
    nu, nf = 40_000, 2_000
    ux, uy, uid, fx, fy, fid = rk.generate(3, nu, nf, seed=seed, nq_ties=4, seed_ties=seed)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    r = np.random.default_rng(5)
    qs = [int(q) for q in rk.generate_queries(nf, 24, seed)]

    def pick(q, n):  # n distinct candidates other than q
        return r.choice(np.setdiff1d(np.arange(nf, dtype=np.uint32), [q]), n, replace=False).astype(np.uint32)

    sizes = [50] * 16 + [255, 256, 1100, 0, 600, 500]
    queries = [(qs[i], pick(qs[i], n)) for i, n in enumerate(sizes)]
    queries.append((qs[22], ds.filter(qs[22], 25)))
    assert len(queries[-1][1]) <= 424  # joins the 500-list's set
    try:
        want = [O.counts(ux, uy, fx, fy, q, cand=c) for q, c in queries]
        order = ds.order()
        for set_max, launches in ((1, 23), (3, 9), (16, 5)):
            monkeypatch.setenv("RK_COUNT_SET", str(set_max))
            got, t = ds.count_all_many(queries, timing=True)
            assert t.n_launches == launches, (kernel, set_max)
            assert t.n_candidates == sum(len(c) for _, c in queries)
            for i in range(len(queries)):
                assert np.array_equal(got[i], want[i]), (kernel, set_max, i)
            got = ds.count_all_many(queries, sorted_order=True)
            for i in range(len(queries)):
                assert np.array_equal(got[i], want[i][order]), (kernel, set_max, i)
    finally:
        ds.close()


@pytest.mark.parametrize("scale,shift", [(1.0, -500.0), (1.0, 1.0e6), (1.0e-3, 0.0), (37.5, -2.0e4)])
def test_coordinate_ranges(rk, ctx, scale, shift):
    """The certified fp32 band and its binary64 re-check hold wherever the
    coordinates lie: negative, offset by 10^6 (fp32 steps of 1/16, so many
    exact ties), scaled to 10^-3, scaled and shifted.  The rknn-rt mesh
    builder assumes the origin at 0 (rt.cpp:268-269); here every form --
    plain, graph batch, device-filtered batch, count-all -- equals the oracle
    on the transformed fp32 coordinates."""
    seed = rk.default_seed(2)
    ux, uy, uid, fx, fy, fid = rk.generate(4, 50_000, 2_000, seed=seed, nq_ties=8, seed_ties=seed)
    t = lambda v: (v.astype(np.float64) * scale + shift).astype(np.float32)
    ux, uy, fx, fy = t(ux), t(uy), t(fx), t(fy)
    qs = [int(q) for q in rk.generate_queries(2_000, 8, seed)]
    ks = [1, 10, 25, 50] * 2
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    try:
        cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
        exp = [O.rknn(ux, uy, uid, fx, fy, q, k, cand=c) for q, k, c in zip(qs, ks, cands)]
        for q, k, c, e in zip(qs, ks, cands, exp):
            assert np.array_equal(ds.query_plain(q, k, c), e), (scale, shift, q, k)
        b = ds.prepare_batch(list(zip(qs, ks, cands)))
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (scale, shift, qs[i], ks[i])
        finally:
            b.close()
        b = ds.prepare_batch(list(zip(qs, ks)), device_filter=True)
        try:
            b.run()
            for i, g in enumerate(b.fetch()):
                assert np.array_equal(g, exp[i]), (scale, shift, qs[i], ks[i])
        finally:
            b.close()
        got = ds.count_all_many([(q, c) for q, c in zip(qs, cands)])
        for i, (q, c) in enumerate(zip(qs, cands)):
            assert np.array_equal(got[i], O.counts(ux, uy, fx, fy, q, cand=c)), (scale, shift, q)
    finally:
        ds.close()


def test_count_all_into_device_memory(rk, ctx):
    """rk_count_all, rk_count_all_sorted and rk_count_all_many write device
    destinations (hipMalloc'd memory of the context's device, through the
    HIP runtime the library uses) without a host copy: the same counts as
    the oracle's, in both orders, and nothing past each row's nu counts."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")

    def dev_alloc(n):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), C.c_size_t(4 * n)) == 0
        assert hip.hipMemset(p, 0x7F, C.c_size_t(4 * n)) == 0
        return p

    def dev_read(p, n):
        out = np.zeros(n, dtype=np.uint32)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(4 * n), 2) == 0  # device to host
        return out

    seed = rk.default_seed(3)
    ux, uy, uid, fx, fy, fid = rk.generate(3, 30_000, 1_500, seed=seed, nq_ties=4, seed_ties=seed)
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    bufs = []
    try:
        qs = [int(q) for q in rk.generate_queries(1_500, 4, seed)]
        cands = [ds.filter(q, 10) for q in qs]
        order = ds.order()
        lib, h = ds._lib, ds.h
        nu = ds.nu
        wants = [O.counts(ux, uy, fx, fy, q, cand=c) for q, c in zip(qs, cands)]
        for q, c, want in zip(qs, cands, wants):
            cp = c.ctypes.data_as(C.POINTER(C.c_uint32))
            for fn, exp in ((lib.rk_count_all, want), (lib.rk_count_all_sorted, want[order])):
                d = dev_alloc(nu)
                bufs.append(d)
                assert fn(h, q, cp, len(c), C.cast(d, C.POINTER(C.c_uint32)), None) == 0
                assert np.array_equal(dev_read(d, nu), exp), (fn, q)
        descs, keep = ds._descs([(q, 1, c) for q, c in zip(qs, cands)])
        for sorted_order in (0, 1):
            stride = nu + 5
            d = dev_alloc(len(qs) * stride)
            bufs.append(d)
            assert lib.rk_count_all_many(h, descs, len(qs), C.cast(d, C.POINTER(C.c_uint32)), stride, sorted_order,
                                         None) == 0
            got = dev_read(d, len(qs) * stride).reshape(len(qs), stride)
            for i, want in enumerate(wants):
                assert np.array_equal(got[i, :nu], want[order] if sorted_order else want), (sorted_order, i)
                assert (got[i, nu:] == 0x7F7F7F7F).all()
    finally:
        for d in bufs:
            hip.hipFree(d)
        ds.close()
