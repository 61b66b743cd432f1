"""GPU: the device influence-zone filter (rk_filter_device).

Its candidate lists may differ from the host filter's (another visiting
order), so the checks are on what the lists are for: every result computed
over them equals the oracle's (through the golden fixtures, the oracle
itself, or the host-filtered result, which the other GPU tests pin to the
oracle), and the lists are valid (indices in range, no duplicates, never q or
a facility at q's position)."""
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


def check_list(ds, fx, fy, q, cand):
    assert len(np.unique(cand)) == len(cand)
    assert np.all(cand < ds.nf)
    assert q not in set(cand.tolist())
    same = (fx[cand] == fx[q]) & (fy[cand] == fy[q])
    assert not same.any()


@pytest.mark.parametrize("name", ["config1.npz", "ties.npz"])
def test_device_filter_golden(rk, ctx, name):
    d, qk, res = golden(name)
    ds = rk.Dataset(ctx, d["ux"], d["uy"], d["uid"], d["fx"], d["fy"], d["fid"])
    try:
        cands = ds.filter_device([q for q, _ in qk], [min(k, 64) for _, k in qk])
        for (q, k), exp, cand in zip(qk, res, cands):
            if k > 64:
                continue
            print(name, "q", q, "k", k, "device |C|", len(cand), "host |C|", len(ds.filter(q, k)), flush=True)
            check_list(ds, d["fx"], d["fy"], q, cand)
            assert np.array_equal(ds.query_plain(q, k, cand), exp)
    finally:
        ds.close()


def test_device_filter_hand_cases(rk, ctx):
    from test_oracle import hand_cases

    for U, F, q, k, exp in hand_cases():
        U = np.array(U, np.float32)
        F = np.array(F, np.float32)
        uid = np.arange(100, 100 + len(U), dtype=np.int32)
        ds = rk.Dataset(ctx, U[:, 0], U[:, 1], uid, F[:, 0], F[:, 1], np.arange(len(F), dtype=np.int32))
        try:
            cand = ds.filter_device([q], [k])[0]
            check_list(ds, F[:, 0], F[:, 1], q, cand)
            assert list(ds.query_plain(q, k, cand)) == exp
        finally:
            ds.close()


@pytest.mark.parametrize("config,nu,nf", [(2, 200_000, 10_000), (3, 60_000, 3_000), (4, 100_000, 5_000)])
def test_device_filter_matches_host_results(rk, ctx, config, nu, nf):
    """Uniform, clustered-with-ties and mixed data, k from 1 to 64, many
    queries in one call: the result over the device list equals the result
    over the host filter's list, and a few equal the oracle's directly."""
    seed = rk.default_seed(config)
    ux, uy, uid, fx, fy, fid = rk.generate(config, nu, nf, seed=seed, nq_ties=8, seed_ties=seed)
    qs = [int(q) for q in rk.generate_queries(nf, 48, seed)]
    ks = [(1, 5, 10, 25, 50, 64)[i % 6] for i in range(len(qs))]
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    O.set_threads(16)
    try:
        cands = ds.filter_device(qs, ks)
        dev_total = host_total = 0
        for i, (q, k, cand) in enumerate(zip(qs, ks, cands)):
            check_list(ds, fx, fy, q, cand)
            host = ds.filter(q, k)
            dev_total += len(cand)
            host_total += len(host)
            got = ds.query_plain(q, k, cand)
            assert np.array_equal(got, ds.query_plain(q, k, host)), (config, q, k)
            if i < 6:
                assert np.array_equal(got, O.rknn(ux, uy, uid, fx, fy, q, k)), (config, q, k)
        print(f"config {config}: device |C| {dev_total / len(qs):.1f} host |C| {host_total / len(qs):.1f}",
              flush=True)
    finally:
        ds.close()


def test_device_filter_arguments(rk, ctx):
    r = np.random.default_rng(9)
    nu, nf = 500, 40
    ux, uy = r.random(nu, np.float32) * 10, r.random(nu, np.float32) * 10
    fx, fy = r.random(nf, np.float32) * 10, r.random(nf, np.float32) * 10
    ds = rk.Dataset(ctx, ux, uy, np.arange(nu, dtype=np.int32), fx, fy, np.arange(nf, dtype=np.int32))
    try:
        with pytest.raises(rk.RkError) as e:
            ds.filter_device([0], [65])  # beyond the sector heaps
        assert e.value.status == 2
        with pytest.raises(rk.RkError) as e:
            ds.filter_device([nf], [1])
        assert e.value.status == 2
        with pytest.raises(rk.RkError) as e:
            ds.filter_device([0], [64], cap=1)  # k >= N_f keeps every facility: too many for cap
        assert e.value.status == 4
        # k beyond the facility count: every user qualifies over any list
        cand = ds.filter_device([3], [64])[0]
        assert np.array_equal(ds.query_plain(3, 64, cand), ds.query_bruteforce(3, 64))
        assert len(ds.query_bruteforce(3, 64)) == nu
    finally:
        ds.close()


@pytest.mark.parametrize("name", ["config1.npz", "ties.npz"])
def test_device_batch_golden(rk, ctx, name, monkeypatch):
    """Device-filtered batches (filter, traversal and sort in one graph),
    replayed, and the one-call form with and without RK_FILTER=host: every
    result equals the golden fixtures."""
    d, qk, res = golden(name)
    ds = rk.Dataset(ctx, d["ux"], d["uy"], d["uid"], d["fx"], d["fy"], d["fid"])
    try:
        keep = [(i, q, k) for i, (q, k) in enumerate(qk) if k <= 64]
        b = ds.prepare_batch([(q, k) for _, q, k in keep], device_filter=True)
        for _ in range(3):
            b.run()
            for (i, q, k), g in zip(keep, b.fetch()):
                assert np.array_equal(g, res[i]), (name, q, k)
        b.close()
        for filt in ("device", "host"):
            monkeypatch.setenv("RK_FILTER", filt)
            got = ds.query_batch(qk)
            for g, exp in zip(got, res):
                assert np.array_equal(g, exp)
    finally:
        ds.close()


def test_device_batch_mixed_k_matches_host(rk, ctx):
    """Clustered data with exact ties, 256 queries with k from 1 to 64 in one
    device-filtered batch: each result equals the host-filtered batch's."""
    seed = rk.default_seed(4)
    nu, nf = 300_000, 20_000
    ux, uy, uid, fx, fy, fid = rk.generate(4, nu, nf, seed=seed, nq_ties=256, seed_ties=seed)
    qs = [int(q) for q in rk.generate_queries(nf, 256, seed)]
    r = np.random.default_rng(3)
    ks = [int(x) for x in r.integers(1, 65, len(qs))]
    ds = rk.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    try:
        b = ds.prepare_batch(list(zip(qs, ks)), device_filter=True)
        b.run()
        got = b.fetch()
        b.close()
        host = ds.query_batch([(q, k, ds.filter(q, k)) for q, k in zip(qs, ks)])
        for q, k, g, h in zip(qs, ks, got, host):
            assert np.array_equal(g, h), (q, k)
    finally:
        ds.close()
