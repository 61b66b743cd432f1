"""Serving loop (rk_serve_*): every result is identical alone and in the mix,
admission by memory, several resident datasets."""
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def datasets(rk, ctx):
    out = []
    for cfg, nu, nf in [(2, 100_000, 2_000), (4, 150_000, 3_000)]:
        seed = rk.default_seed(cfg)
        arrs = rk.generate(cfg, nu, nf, seed=seed, nq_ties=32, seed_ties=seed)
        out.append((arrs, rk.Dataset(ctx, *arrs), rk.generate_queries(nf, 32, seed)))
    yield out
    for _, ds, _ in out:
        ds.close()


@pytest.mark.parametrize("filt", ["device", "host"])
def test_alone_equals_in_the_mix(rk, ctx, datasets, filt, monkeypatch):
    """k up to 80: with the device filter the requests with k <= 64 are
    filtered inside the device steps and the others on the host, in separate
    steps; RK_FILTER=host filters all of them on the host."""
    monkeypatch.setenv("RK_FILTER", filt)
    r = np.random.default_rng(1)
    reqs = []
    for i in range(300):
        d = int(r.integers(0, len(datasets)))
        q = int(datasets[d][2][r.integers(0, 32)])
        k = int(r.integers(1, 81))
        reqs.append((d, q, k))
    srv = rk.Server(ctx, max_batch=64, max_wait_s=0.002)
    try:
        tickets = []
        for d, q, k in reqs:  # bursty arrivals while earlier steps run
            tickets.append(srv.submit(datasets[d][1], q, k))
            if len(tickets) % 50 == 0:
                time.sleep(0.003)
        results = {}
        deadline = time.time() + 90
        while len(results) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t not in results:
                    got = srv.poll(t, datasets[0][1].nu + datasets[1][1].nu)
                    if got is not None:
                        results[t] = got
            time.sleep(0.001)
        assert len(results) == len(tickets)
    finally:
        srv.stop()
    lat = []
    for t, (d, q, k) in zip(tickets, reqs):
        ids, latency = results[t]
        lat.append(latency)
        assert latency > 0
        alone = datasets[d][1].query(q, k)
        assert np.array_equal(ids, alone), (d, q, k)
    # spot-check against the oracle
    for t, (d, q, k) in list(zip(tickets, reqs))[:6]:
        ux, uy, uid, fx, fy, fid = datasets[d][0]
        O.set_threads(16)
        assert np.array_equal(results[t][0], O.rknn(ux, uy, uid, fx, fy, q, k))
    print("serve latency p50 %.3f ms p99 %.3f ms" % (np.percentile(lat, 50) * 1e3, np.percentile(lat, 99) * 1e3))


def test_admission_rejects_what_never_fits(rk, ctx, datasets):
    srv = rk.Server(ctx, max_batch=8, memory_budget=1 << 12)
    try:
        with pytest.raises(rk.RkError) as e:
            srv.submit(datasets[0][1], int(datasets[0][2][0]), 5)
        assert e.value.status == 7
    finally:
        srv.stop()


def test_admission_throttles_but_completes(rk, ctx, datasets):
    ds = datasets[0][1]
    q = int(datasets[0][2][1])
    # a budget that admits only a few requests at a time
    from rknn_amd import _lib  # noqa: F401

    srv = rk.Server(ctx, max_batch=256, memory_budget=3 << 20, max_wait_s=0.0)
    try:
        ts = [srv.submit(ds, q, 10) for _ in range(40)]
        exp = ds.query(q, 10)
        done = 0
        deadline = time.time() + 60
        pending = set(ts)
        while pending and time.time() < deadline:
            for t in list(pending):
                got = srv.poll(t, ds.nu)
                if got is not None:
                    assert np.array_equal(got[0], exp)
                    pending.discard(t)
                    done += 1
            time.sleep(0.001)
        assert done == 40
    finally:
        srv.stop()


def device_free_bytes():
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_budget_bounds_device_memory(rk, ctx, datasets):
    """A tight memory budget: the server lays its buffers out within it (steps
    of fewer queries), every request -- both datasets, both filter sides --
    still completes with its result alone, and the device's free memory
    drops by no more than the budget plus what the HIP runtime takes for the
    server's streams and events (allowed 16 MiB)."""
    budget = 16 << 20
    r = np.random.default_rng(7)
    reqs = [(int(r.integers(0, 2)), int(r.integers(0, 32)), int(r.integers(1, 81))) for _ in range(200)]
    reqs = [(d, int(datasets[d][2][i]), k) for d, i, k in reqs]
    alone = [datasets[d][1].query(q, k) for d, q, k in reqs]  # before the measurement: this allocates
    free0 = device_free_bytes()
    srv = rk.Server(ctx, max_batch=256, memory_budget=budget, max_wait_s=0.001)
    try:
        tickets = []
        for i, (d, q, k) in enumerate(reqs):
            tickets.append(srv.submit(datasets[d][1], q, k))
            if i % 50 == 49:
                time.sleep(0.002)
        res = {}
        deadline = time.time() + 90
        while len(res) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t not in res:
                    got = srv.poll(t, datasets[0][1].nu + datasets[1][1].nu)
                    if got is not None:
                        res[t] = got[0]
            time.sleep(0.001)
        assert len(res) == len(tickets)
        held, step = srv.info()
        grew = free0 - device_free_bytes()
        print(f"budget {budget >> 20} MiB: buffers {held / 2**20:.1f} MiB, steps of {step} queries, "
              f"device free memory down {grew / 2**20:.1f} MiB", flush=True)
        assert held <= budget
        assert step < 256  # the budget, not max_batch, sized the steps
        assert grew <= budget + (16 << 20)
    finally:
        srv.stop()
    for t, exp, req in zip(tickets, alone, reqs):
        assert np.array_equal(res[t], exp), req
