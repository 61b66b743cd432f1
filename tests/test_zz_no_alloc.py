"""GPU: the allocation-free plain form (rk_query_plain_noalloc) through the C
ABI.  Its output equals rk_query_plain's (and the oracle's) on two calls in a
row, the second with a smaller input, so nothing of the first call's scratch
leaks into the second; its limits are errors, not truncation; and calls leave
the device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def device_free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.fixture(scope="module")
def data(rk, ctx):
    seed = rk.default_seed(4)
    arrs = rk.generate(4, 300_000, 5_000, seed=seed, nq_ties=8, seed_ties=seed)
    ds = rk.Dataset(ctx, *arrs)
    qs = [int(q) for q in rk.generate_queries(5_000, 8, seed)]
    yield arrs, ds, qs
    ds.close()


def test_two_calls_second_smaller(rk, data):
    arrs, ds, qs = data
    ux, uy, uid, fx, fy, fid = arrs
    O.set_threads(16)
    # first call: k=40 over its filter list plus q itself and a repeat (the
    # device builds dead records for both); second: k=3, a shorter list
    q1, q2 = qs[0], qs[1]
    c1 = np.concatenate([ds.filter(q1, 40), [q1], ds.filter(q1, 40)[:5]]).astype(np.uint32)
    c2 = ds.filter(q2, 3)
    assert len(c2) < len(c1)
    got1 = ds.query_plain_noalloc(q1, 40, c1)
    got2 = ds.query_plain_noalloc(q2, 3, c2)
    exp1 = ds.query_plain(q1, 40, c1)
    exp2 = ds.query_plain(q2, 3, c2)
    print(f"noalloc: |C| {len(c1)} -> {len(c2)}, |R| {len(got1)} -> {len(got2)}", flush=True)
    assert len(got2) < len(got1)
    assert np.array_equal(got1, exp1)
    assert np.array_equal(got2, exp2)
    assert np.array_equal(got1, O.rknn(ux, uy, uid, fx, fy, q1, 40, cand=c1))
    assert np.array_equal(got2, O.rknn(ux, uy, uid, fx, fy, q2, 3, cand=c2))


def test_every_query_and_free_memory_unchanged(rk, data):
    arrs, ds, qs = data
    work = []
    for i, q in enumerate(qs):
        k = (1, 10, 25, 50)[i % 4]
        c = ds.filter(q, k)
        work.append((q, k, c, ds.query_plain(q, k, c)))  # the allocating form, before the measurement
    ds.query_plain_noalloc(qs[0], 10, ds.filter(qs[0], 10))  # the runtime's first-use state
    before = device_free_bytes()
    got = [ds.query_plain_noalloc(q, k, c) for q, k, c, _ in work]
    assert device_free_bytes() == before
    for (q, k, c, exp), g in zip(work, got):
        assert np.array_equal(g, exp), (q, k)


def test_limits(rk, ctx, data):
    arrs, ds, qs = data
    q = qs[2]
    # more candidates than the scratch holds
    with pytest.raises(rk.RkError) as e:
        ds.query_plain_noalloc(q, 5, np.arange(4097, dtype=np.uint32))
    assert e.value.status == 2
    # exactly the maximum is taken
    c = np.arange(4096, dtype=np.uint32)
    assert np.array_equal(ds.query_plain_noalloc(q, 5, c), ds.query_plain(q, 5, c))
    # a bad index is an error
    with pytest.raises(rk.RkError) as e:
        ds.query_plain_noalloc(q, 5, np.array([ds.nf], np.uint32))
    assert e.value.status == 2
    # fewer live candidates than k: every user (host copy, any size)
    assert np.array_equal(ds.query_plain_noalloc(q, 3, np.array([q, q], np.uint32)), np.sort(arrs[2]))
    # caller buffer too small
    c = ds.filter(q, 10)
    n = len(ds.query_plain(q, 10, c))
    with pytest.raises(rk.RkError) as e:
        ds.query_plain_noalloc(q, 10, c, cap=n - 1)
    assert e.value.status == 4
    # a result beyond the staging: 2.5M users, half of them closer to q
    r = np.random.default_rng(3)
    nu = 2_500_000
    ux, uy = (r.random(nu) * 1000).astype(np.float32), (r.random(nu) * 1000).astype(np.float32)
    uid = np.arange(nu, dtype=np.int32)
    fx, fy = np.array([250.0, 750.0], np.float32), np.array([500.0, 500.0], np.float32)
    big = rk.Dataset(ctx, ux, uy, uid, fx, fy, np.arange(2, dtype=np.int32))
    try:
        with pytest.raises(rk.RkError) as e:
            big.query_plain_noalloc(0, 1, np.array([1], np.uint32))
        assert e.value.status == 4
        assert "n_out=%d" % len(big.query_plain(0, 1, np.array([1], np.uint32))) in str(e.value)
        # and the next call on the device is whole again
        got = ds.query_plain_noalloc(q, 10, c)
        assert np.array_equal(got, ds.query_plain(q, 10, c))
    finally:
        big.close()
