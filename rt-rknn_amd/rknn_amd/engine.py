"""Python mirror of the engine's C ABI (include/rknn.h) for tests and bench.

The host program of the engine is the C++ CLI (bin/rknn-hip); this module
only moves numpy buffers across the C ABI.  Every call runs the gfx950
kernels; errors raise RkError carrying the rk_status code.
"""
import ctypes as C
import threading

import numpy as np

from . import _lib as L


class RkError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"rk status {status}: {message}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count():
    return L.load().rk_device_count()


class Context:
    """One HIP device and stream (rk_init / rk_shutdown)."""

    def __init__(self, device=0):
        self._lib = L.load()
        h = C.c_void_p()
        st = self._lib.rk_init(device, C.byref(h))
        if st != L.RK_OK:
            raise RkError(st, self._lib.rk_status_string(st).decode())
        self.h = h

    def check(self, st):
        if st != L.RK_OK:
            raise RkError(st, self._lib.rk_last_error(self.h).decode())

    def close(self):
        if self.h:
            self._lib.rk_shutdown(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# rk_query_desc (include/rknn.h): uint32 q, uint32 k, const uint32* cand, size_t nc
_DESC = np.dtype([("q", "<u4"), ("k", "<u4"), ("cand", "<u8"), ("nc", "<u8")])
assert _DESC.itemsize == C.sizeof(L.rk_query_desc)


class Dataset:
    """Device-resident users and facilities (rk_dataset_create)."""

    def __init__(self, ctx, ux, uy, uid, fx, fy, fid):
        self.ctx = ctx
        self._lib = ctx._lib
        self.ux, self.uy, self.uid = _f32(ux), _f32(uy), _i32(uid)
        self.fx, self.fy, self.fid = _f32(fx), _f32(fy), _i32(fid)
        h = C.c_void_p()
        st = self._lib.rk_dataset_create(
            ctx.h, _ptr(self.ux, C.c_float), _ptr(self.uy, C.c_float), _ptr(self.uid, C.c_int32), len(self.ux),
            _ptr(self.fx, C.c_float), _ptr(self.fy, C.c_float), _ptr(self.fid, C.c_int32), len(self.fx),
            C.byref(h))
        ctx.check(st)
        self.h = h
        self.nu, self.nf = len(self.ux), len(self.fx)
        self._batch_out = None  # query_batch's output buffer, reused while not in use elsewhere
        self._batch_lock = threading.Lock()

    def close(self):
        if self.h:
            self._lib.rk_dataset_destroy(self.h)
            self.h = None

    def device_bytes(self):
        b = C.c_uint64()
        self.ctx.check(self._lib.rk_dataset_info(self.h, None, None, C.byref(b)))
        return b.value

    def filter(self, q, k):
        n = C.c_size_t()
        st = self._lib.rk_filter(self.h, q, k, None, 0, C.byref(n))
        if st not in (L.RK_OK, L.RK_ERR_CAPACITY):
            self.ctx.check(st)
        out = np.zeros(max(1, n.value), dtype=np.uint32)
        self.ctx.check(self._lib.rk_filter(self.h, q, k, _ptr(out, C.c_uint32), len(out), C.byref(n)))
        return out[: n.value].copy()

    def filter_device(self, qs, ks, cap=None):
        """Candidate lists of the device filter (rk_filter_device), one
        uint32 array per query."""
        qs, ks = _u32(qs), _u32(ks)
        n = len(qs)
        nc = np.zeros(max(1, n), dtype=np.uint64)
        cap = cap if cap is not None else min(max(1, self.nf), 16384)
        out = np.zeros(max(1, n * cap), dtype=np.uint32)
        st = self._lib.rk_filter_device(self.h, _ptr(qs, C.c_uint32), _ptr(ks, C.c_uint32), n,
                                        _ptr(out, C.c_uint32), cap, _ptr(nc, C.c_uint64))
        self.ctx.check(st)
        return [out[i * cap: i * cap + int(nc[i])].copy() for i in range(n)]

    def _ids_call(self, fn, *args):
        cap = max(1, self.nu)
        out = np.empty(cap, dtype=np.int32)
        n = C.c_size_t()
        t = L.rk_timing()
        st = fn(self.h, *args, _ptr(out, C.c_int32), cap, C.byref(n), C.byref(t))
        self.ctx.check(st)
        return out[: n.value].copy(), t

    def query_plain(self, q, k, cand, timing=False):
        cand = _u32(cand)
        ids, t = self._ids_call(self._lib.rk_query_plain, q, k, _ptr(cand, C.c_uint32), len(cand))
        return (ids, t) if timing else ids

    def query_plain_noalloc(self, q, k, cand, cap=None):
        """rk_query_plain_noalloc: the plain form on the library's fixed
        per-device scratch (no device allocation).  cap bounds the output."""
        cand = _u32(cand)
        cap = max(1, self.nu) if cap is None else cap
        out = np.empty(max(1, cap), dtype=np.int32)
        n = C.c_size_t()
        st = self._lib.rk_query_plain_noalloc(self.h, q, k, _ptr(cand, C.c_uint32), len(cand),
                                              _ptr(out, C.c_int32), cap, C.byref(n))
        if st != L.RK_OK:
            raise RkError(st, self._lib.rk_last_error(self.ctx.h).decode() + f" (n_out={n.value})")
        return out[: n.value].copy()

    def query_bruteforce(self, q, k, timing=False):
        ids, t = self._ids_call(self._lib.rk_query_bruteforce, q, k)
        return (ids, t) if timing else ids

    def query(self, q, k, timing=False):
        ids, t = self._ids_call(self._lib.rk_query, q, k)
        return (ids, t) if timing else ids

    def count_all(self, q, cand, timing=False, sorted_order=False, out=None):
        """Per-user counts in caller order, or with sorted_order=True in the
        dataset's own order (counts[j] belongs to user order()[j]); out: a
        uint32 array of at least nu entries to write them into."""
        cand = _u32(cand)
        if out is not None:
            if out.dtype != np.uint32 or not out.flags.c_contiguous or len(out) < max(1, self.nu):
                raise ValueError("out must be a contiguous uint32 array of at least nu entries")
            counts = out
        else:
            counts = np.zeros(max(1, self.nu), dtype=np.uint32)
        t = L.rk_timing()
        fn = self._lib.rk_count_all_sorted if sorted_order else self._lib.rk_count_all
        self.ctx.check(fn(self.h, q, _ptr(cand, C.c_uint32), len(cand), _ptr(counts, C.c_uint32), C.byref(t)))
        counts = counts[: self.nu]
        return (counts, t) if timing else counts

    def count_all_many(self, queries, timing=False, sorted_order=False, out=None):
        """Count-all for many queries in one call: up to 16 consecutive
        queries whose lists fit 1,024 records together share a launch, and
        the launches run back to back on the device.  queries: (q, cand)
        pairs, or (q, k, None) to count over the host filter's list at k.
        Returns an (nq, nu) uint32 array (rows in caller order, or the
        dataset's with sorted_order), with the rk_timing when timing
        (n_launches: the launches made); out: a C-contiguous (nq, >= nu)
        uint32 array to write into."""
        qk = [(x[0], 1, x[1]) if len(x) == 2 else tuple(x) for x in queries]
        descs, keep = self._descs(qk)
        nq = len(qk)
        if out is not None:
            if (out.dtype != np.uint32 or out.ndim != 2 or not out.flags.c_contiguous or out.shape[0] < nq
                    or out.shape[1] < max(1, self.nu)):
                raise ValueError("out must be a C-contiguous (nq, >= nu) uint32 array")
            counts = out
        else:
            counts = np.zeros((max(1, nq), max(1, self.nu)), dtype=np.uint32)
        t = L.rk_timing()
        self.ctx.check(self._lib.rk_count_all_many(self.h, descs, nq, _ptr(counts, C.c_uint32), counts.shape[1],
                                                   1 if sorted_order else 0, C.byref(t)))
        counts = counts[:nq, : self.nu]
        return (counts, t) if timing else counts

    def order(self):
        """order[j]: index in the constructor's user arrays of the dataset's
        j-th user."""
        out = np.zeros(max(1, self.nu), dtype=np.uint32)
        self.ctx.check(self._lib.rk_dataset_order(self.h, _ptr(out, C.c_uint32)))
        return out[: self.nu]

    @staticmethod
    def _descs(queries):
        """rk_query_desc[] as a numpy record array (the struct's layout), the
        (q, k) columns filled at once; keep holds the candidate arrays."""
        n = len(queries)
        d = np.zeros(max(1, n), dtype=_DESC)
        keep = []
        if n:
            d["q"][:n] = [int(x[0]) for x in queries]
            d["k"][:n] = [int(x[1]) for x in queries]
            for i, x in enumerate(queries):
                if len(x) > 2 and x[2] is not None:
                    c = _u32(x[2])
                    keep.append(c)
                    d["cand"][i] = c.ctypes.data
                    d["nc"][i] = len(c)
        keep.append(d)
        return d.ctypes.data_as(C.POINTER(L.rk_query_desc)), keep

    def query_batch(self, queries, timing=False):
        """queries: iterable of (q, k) or (q, k, cand); cand None runs the filter."""
        queries = list(queries)
        descs, keep = self._descs(queries)
        nq = len(queries)
        offs = np.zeros(nq + 1, dtype=np.uint64)
        t = L.rk_timing()
        # the output buffer is kept between calls (its pages stay mapped); a
        # concurrent call on the same dataset takes a buffer of its own
        shared = self._batch_lock.acquire(blocking=False)
        try:
            out = self._batch_out if shared else None
            while True:
                if out is None:
                    out = np.empty(max(1, self.nu), dtype=np.int32)
                st = self._lib.rk_query_batch(self.h, descs, nq, _ptr(out, C.c_int32), len(out),
                                              _ptr(offs, C.c_uint64), C.byref(t))
                if st == L.RK_ERR_CAPACITY and int(offs[nq]) > len(out):
                    out = np.empty(int(offs[nq]), dtype=np.int32)
                    continue
                self.ctx.check(st)
                break
            total = int(offs[nq])
            res = np.split(out[:total].copy(), offs[1:nq].astype(np.int64)) if nq else []
            if shared:
                self._batch_out = out
        finally:
            if shared:
                self._batch_lock.release()
        return (res, t) if timing else res

    def prepare_batch(self, queries, device_filter=False):
        """queries: (q, k) or (q, k, cand); with device_filter=True only
        (q, k) pairs (k <= 64), filtered on the device inside each run."""
        return Batch(self, list(queries), device_filter)


class Batch:
    """Prepared graph-captured batch (rk_batch_prepare[_device] / run / fetch)."""

    def __init__(self, ds, queries, device_filter=False):
        self.ds = ds
        self._lib = ds._lib
        self.nq = len(queries)
        h = C.c_void_p()
        if device_filter:
            self._q = _u32([int(qd[0]) for qd in queries])
            self._k = _u32([int(qd[1]) for qd in queries])
            ds.ctx.check(self._lib.rk_batch_prepare_device(ds.h, _ptr(self._q, C.c_uint32), _ptr(self._k, C.c_uint32),
                                                           self.nq, C.byref(h)))
        else:
            self._descs, self._keep = Dataset._descs(queries)
            ds.ctx.check(self._lib.rk_batch_prepare(ds.h, self._descs, self.nq, C.byref(h)))
        self.h = h

    def run(self):
        s = C.c_double()
        self.ds.ctx.check(self._lib.rk_batch_run(self.h, C.byref(s)))
        return s.value

    def enqueue(self):
        """Puts one run on the batch's own stream and returns (rk_batch_enqueue):
        runs of different batches overlap on the device."""
        self.ds.ctx.check(self._lib.rk_batch_enqueue(self.h))

    def wait(self):
        """Waits for the enqueued run; its device seconds (rk_batch_wait)."""
        s = C.c_double()
        self.ds.ctx.check(self._lib.rk_batch_wait(self.h, C.byref(s)))
        return s.value

    def fetch_flat(self):
        """rk_batch_fetch into a buffer the batch keeps (grown when short):
        (ids, offsets) -- the ids of query i are ids[offsets[i]:offsets[i+1]];
        overwritten by the next fetch."""
        if not hasattr(self, "_fo"):
            self._fo = np.zeros(self.nq + 1, dtype=np.uint64)
            self._fbuf = np.empty(1 << 16, dtype=np.int32)
        while True:
            out = self._fbuf
            st = self._lib.rk_batch_fetch(self.h, _ptr(out, C.c_int32), len(out), _ptr(self._fo, C.c_uint64))
            if st == L.RK_ERR_CAPACITY and int(self._fo[self.nq]) > len(out):
                self._fbuf = np.empty(int(self._fo[self.nq]), dtype=np.int32)
                continue
            self.ds.ctx.check(st)
            return out, self._fo

    def fetch(self):
        out, offs = self.fetch_flat()
        return [out[int(offs[i]): int(offs[i + 1])].copy() for i in range(self.nq)]

    def results_raw(self):
        """rk_batch_results: per query the address and length of its sorted
        ids in the batch's pinned landing area (no copy), as two uint64
        arrays; valid until the batch runs again."""
        if not hasattr(self, "_rptr"):
            self._rptr = np.zeros(max(1, self.nq), dtype=np.uint64)
            self._rn = np.zeros(max(1, self.nq), dtype=np.uint64)
        self.ds.ctx.check(self._lib.rk_batch_results(self.h, self._rptr.ctypes.data_as(C.POINTER(C.c_void_p)),
                                                     _ptr(self._rn, C.c_uint64)))
        return self._rptr[: self.nq], self._rn[: self.nq]

    def results(self):
        """The last run's results as numpy views of the landing area (no
        copy); valid until the batch runs again."""
        ptrs, ns = self.results_raw()
        out = []
        for p, n in zip(ptrs.tolist(), ns.tolist()):
            if n == 0:
                out.append(np.empty(0, np.int32))
            else:
                out.append(np.ctypeslib.as_array(C.cast(C.c_void_p(p), C.POINTER(C.c_int32)), shape=(n,)))
        return out

    def stats(self):
        """(items per tree level, leaf first; results; binary64 re-checks) of the last run."""
        items = np.zeros(16, np.uint64)
        nl, res, rech = C.c_size_t(), C.c_uint64(), C.c_uint64()
        self.ds.ctx.check(self._lib.rk_batch_stats(self.h, _ptr(items, C.c_uint64), 16, C.byref(nl), C.byref(res),
                                                   C.byref(rech)))
        return items[: nl.value].tolist(), res.value, rech.value

    def waves_per_query(self):
        """Block shape the autotuner chose for the traversal kernel."""
        w = C.c_uint32()
        self.ds.ctx.check(self._lib.rk_batch_config(self.h, C.byref(w)))
        return w.value

    def close(self):
        if self.h:
            self._lib.rk_batch_destroy(self.h)
            self.h = None


class Server:
    """Continuous-batching request loop (rk_serve_*)."""

    def __init__(self, ctx, max_batch=256, memory_budget=4 << 30, max_wait_s=0.0005):
        self.ctx = ctx
        self._lib = ctx._lib
        cfg = L.rk_serve_config(max_batch, memory_budget, max_wait_s)
        h = C.c_void_p()
        ctx.check(self._lib.rk_serve_start(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, ds, q, k):
        t = C.c_uint64()
        st = self._lib.rk_serve_submit(self.h, ds.h, q, k, C.byref(t))
        self.ctx.check(st)
        return t.value

    def poll(self, ticket, cap):
        """(ids, latency_s) when done, None while in flight."""
        out = np.empty(max(1, cap), dtype=np.int32)
        n = C.c_size_t()
        lat = C.c_double()
        st = self._lib.rk_serve_poll(self.h, ticket, _ptr(out, C.c_int32), len(out), C.byref(n), C.byref(lat))
        if st == L.RK_ERR_NOT_READY:
            return None
        if st == L.RK_ERR_CAPACITY:
            return self.poll(ticket, n.value)
        self.ctx.check(st)
        return out[: n.value].copy(), lat.value

    def info(self):
        """(device bytes the server's buffers hold, queries a step takes at most)."""
        b, n = C.c_uint64(), C.c_uint32()
        self.ctx.check(self._lib.rk_serve_info(self.h, C.byref(b), C.byref(n)))
        return b.value, n.value

    def stop(self):
        if self.h:
            self._lib.rk_serve_stop(self.h)
            self.h = None


# ---------------------------------------------------------------- data sets

def default_seed(config):
    return L.load().rk_default_seed(config)


def generate(config, nu, nf, seed=None, nq_ties=0, seed_ties=None):
    """Seeded synthetic dataset (rk_generate): (ux, uy, uid, fx, fy, fid)."""
    lib = L.load()
    seed = default_seed(config) if seed is None else seed
    seed_ties = seed if seed_ties is None else seed_ties
    ux = np.empty(nu, np.float32)
    uy = np.empty(nu, np.float32)
    uid = np.empty(nu, np.int32)
    fx = np.empty(nf, np.float32)
    fy = np.empty(nf, np.float32)
    fid = np.empty(nf, np.int32)
    st = lib.rk_generate(config, nu, nf, seed, nq_ties, seed_ties, _ptr(ux, C.c_float), _ptr(uy, C.c_float),
                         _ptr(uid, C.c_int32), _ptr(fx, C.c_float), _ptr(fy, C.c_float), _ptr(fid, C.c_int32))
    if st != L.RK_OK:
        raise RkError(st, "rk_generate")
    return ux, uy, uid, fx, fy, fid


def generate_queries(nf, nq, seed):
    lib = L.load()
    out = np.empty(max(1, nq), np.uint32)
    st = lib.rk_generate_queries(nf, nq, seed, _ptr(out, C.c_uint32))
    if st != L.RK_OK:
        raise RkError(st, "rk_generate_queries")
    return out[:nq]


def load_text(path):
    lib = L.load()
    fx, fy, ux, uy = (C.POINTER(C.c_float)() for _ in range(4))
    fid, uid = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    nf, nu = C.c_uint64(), C.c_uint64()
    st = lib.rk_load_text(path.encode(), C.byref(fx), C.byref(fy), C.byref(fid), C.byref(nf), C.byref(ux),
                          C.byref(uy), C.byref(uid), C.byref(nu))
    if st != L.RK_OK:
        raise RkError(st, f"rk_load_text({path})")
    try:
        arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy() if n else np.empty(0, dt)
        res = (arr(ux, nu.value, np.float32), arr(uy, nu.value, np.float32), arr(uid, nu.value, np.int32),
               arr(fx, nf.value, np.float32), arr(fy, nf.value, np.float32), arr(fid, nf.value, np.int32))
    finally:
        for p in (fx, fy, fid, ux, uy, uid):
            lib.rk_free(C.cast(p, C.c_void_p))
    return res


def write_text(path, ux, uy, uid, fx, fy, fid):
    lib = L.load()
    ux, uy, uid, fx, fy, fid = _f32(ux), _f32(uy), _i32(uid), _f32(fx), _f32(fy), _i32(fid)
    st = lib.rk_write_text(path.encode(), _ptr(fx, C.c_float), _ptr(fy, C.c_float), _ptr(fid, C.c_int32), len(fx),
                           _ptr(ux, C.c_float), _ptr(uy, C.c_float), _ptr(uid, C.c_int32), len(ux))
    if st != L.RK_OK:
        raise RkError(st, f"rk_write_text({path})")
