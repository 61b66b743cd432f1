"""Host influence-zone filter (rk_filter_points): the candidate set must be
sufficient -- the oracle restricted to it gives the brute-force set -- on
random, clustered, tie-laden and adversarial inputs.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import oracle as O


def filt(rk, fx, fy, ux, uy, q, k):
    lib = rk.load()
    fx, fy = np.ascontiguousarray(fx, np.float32), np.ascontiguousarray(fy, np.float32)
    ux, uy = np.ascontiguousarray(ux, np.float32), np.ascontiguousarray(uy, np.float32)
    cap = max(1, len(fx))
    out = np.zeros(cap, np.uint32)
    n = C.c_size_t()
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    st = lib.rk_filter_points(p(fx, C.c_float), p(fy, C.c_float), len(fx), p(ux, C.c_float), p(uy, C.c_float),
                              len(ux), q, k, p(out, C.c_uint32), cap, C.byref(n))
    assert st == 0
    return out[: n.value].copy()


def check(rk, ux, uy, fx, fy, qs, ks):
    uid = np.arange(len(ux), dtype=np.int32)
    O.set_threads(8)
    for q in qs:
        for k in ks:
            c = filt(rk, fx, fy, ux, uy, int(q), k)
            assert len(set(c.tolist())) == len(c)  # unique
            assert int(q) not in c.tolist()
            d = (fx[c].astype(np.float64) - fx[q]) ** 2 + (fy[c].astype(np.float64) - fy[q]) ** 2
            assert np.all(np.diff(d) >= 0)  # nearest first
            exp = O.rknn(ux, uy, uid, fx, fy, int(q), k)
            got = O.rknn(ux, uy, uid, fx, fy, int(q), k, cand=c)
            assert np.array_equal(got, exp), (int(q), k, len(c))


@pytest.mark.parametrize("config", [1, 3, 4])
def test_sufficient_on_generated(rk, config):
    seed = rk.default_seed(config)
    ux, uy, uid, fx, fy, fid = rk.generate(config, 20000, 1500, seed=seed, nq_ties=12, seed_ties=seed)
    qs = rk.generate_queries(1500, 12, seed)
    check(rk, ux, uy, fx, fy, qs, [1, 2, 5, 10, 25, 50])


def test_candidate_counts_reasonable(rk):
    ux, uy, uid, fx, fy, fid = rk.generate(2, 50000, 5000)
    qs = rk.generate_queries(5000, 20, 9)
    for k, lim in [(1, 40), (10, 150), (50, 500)]:
        sizes = [len(filt(rk, fx, fy, ux, uy, int(q), k)) for q in qs]
        assert np.mean(sizes) < lim, (k, np.mean(sizes))
        assert min(sizes) >= min(k, 1)


def test_adversarial_geometries(rk):
    r = np.random.default_rng(0)
    # facilities on a circle around q=0: all equidistant from q
    ang = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    fx = np.concatenate([[0.0], 50 * np.cos(ang)]).astype(np.float32)
    fy = np.concatenate([[0.0], 50 * np.sin(ang)]).astype(np.float32)
    ux, uy = (r.random(5000) * 200 - 100).astype(np.float32), (r.random(5000) * 200 - 100).astype(np.float32)
    check(rk, ux, uy, fx, fy, [0, 1, 17], [1, 3, 8, 64, 70])
    # collinear facilities, q at the end and in the middle
    fx = np.arange(40, dtype=np.float32) * 7.0
    fy = np.zeros(40, np.float32)
    ux, uy = (r.random(3000) * 300).astype(np.float32), (r.random(3000) * 40 - 20).astype(np.float32)
    check(rk, ux, uy, fx, fy, [0, 20, 39], [1, 2, 5, 39])
    # duplicated facilities and a facility co-located with q
    fx = np.array([10, 10, 10, 20, 20, 0, 30], np.float32)
    fy = np.array([10, 10, 10, 20, 0, 20, 30], np.float32)
    ux, uy = (r.random(2000) * 40).astype(np.float32), (r.random(2000) * 40).astype(np.float32)
    check(rk, ux, uy, fx, fy, [0, 3, 6], [1, 2, 3, 6])
    # q in a corner of the users' box and outside it
    fx = (r.random(300) * 100).astype(np.float32)
    fy = (r.random(300) * 100).astype(np.float32)
    fx[0], fy[0] = 0.0, 0.0
    fx[1], fy[1] = -500.0, 50.0
    ux, uy = (r.random(4000) * 100).astype(np.float32), (r.random(4000) * 100).astype(np.float32)
    check(rk, ux, uy, fx, fy, [0, 1, 2], [1, 4, 16])


@pytest.mark.parametrize("scale", [1e-3, 1.0, 1e5])
def test_scales_and_near_coincident(rk, scale):
    r = np.random.default_rng(int(scale * 7) % 1000 + 1)
    fx = (r.random(400) * scale).astype(np.float32)
    fy = (r.random(400) * scale).astype(np.float32)
    # facilities within a few ulps of q
    fx[1] = np.nextafter(fx[0], np.float32(np.inf))
    fy[1] = fy[0]
    fx[2] = fx[0]
    fy[2] = np.nextafter(fy[0], np.float32(-np.inf))
    ux = (r.random(4000) * scale).astype(np.float32)
    uy = (r.random(4000) * scale).astype(np.float32)
    ux[:50] = fx[0] + (r.random(50).astype(np.float32) - 0.5) * np.float32(scale * 1e-6)
    uy[:50] = fy[0]
    check(rk, ux, uy, fx, fy, [0, 1, 5], [1, 2, 3, 10])


def test_k_larger_than_facilities(rk):
    r = np.random.default_rng(4)
    fx, fy = r.random(10, np.float32), r.random(10, np.float32)
    ux, uy = r.random(100, np.float32), r.random(100, np.float32)
    c = filt(rk, fx, fy, ux, uy, 3, 50)
    uid = np.arange(100, dtype=np.int32)
    assert np.array_equal(O.rknn(ux, uy, uid, fx, fy, 3, 50, cand=c), np.arange(100))
