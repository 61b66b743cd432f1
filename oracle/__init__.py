"""CPU oracle of the RkNN engine -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/liboracle.so (oracle.c, built by oracle/Makefile).
Only tests/, __graft_entry__.smoke() (as the checker) and bench.py's
cpu_baseline leg import this package; the engine never does.  See oracle.h
for the restated semantics (rknn-naive/rknn.cpp:169-198 in binary64 with the
strict tie rule of rknn-inf/inf.h:77-93) and for what pins each function.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(mutant=0):
    """The oracle library, or the planted-mutation build liboracle_mut<n>.so."""
    if mutant in _libs:
        return _libs[mutant]
    name = "liboracle.so" if mutant == 0 else f"liboracle_mut{mutant}.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    f32p, i32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)
    L.oracle_rknn.restype = C.c_int64
    L.oracle_rknn.argtypes = [f32p, f32p, i32p, C.c_uint64, f32p, f32p, C.c_uint64, C.c_uint32, C.c_uint32,
                              u32p, C.c_uint64, i32p]
    L.oracle_counts.restype = C.c_int
    L.oracle_counts.argtypes = [f32p, f32p, C.c_uint64, f32p, f32p, C.c_uint64, C.c_uint32, u32p, C.c_uint64,
                                C.c_uint32, u32p]
    L.oracle_rknn_naive_int.restype = C.c_int64
    L.oracle_rknn_naive_int.argtypes = [i32p, i32p, i32p, C.c_uint64, i32p, i32p, i32p, C.c_uint64, C.c_uint32,
                                        C.c_uint32, i32p]
    L.oracle_set_threads.restype = None
    L.oracle_set_threads.argtypes = [C.c_int]
    u64p = C.POINTER(C.c_uint64)
    L.oracle_tree_build.restype = C.c_void_p
    L.oracle_tree_build.argtypes = [f32p, f32p, i32p, C.c_uint64]
    L.oracle_tree_free.restype = None
    L.oracle_tree_free.argtypes = [C.c_void_p]
    L.oracle_tree_rknn.restype = C.c_int64
    L.oracle_tree_rknn.argtypes = [C.c_void_p, f32p, f32p, C.c_uint64, C.c_uint32, C.c_uint32, u32p, C.c_uint64,
                                   i32p]
    L.oracle_tree_rknn_many.restype = C.c_int
    L.oracle_tree_rknn_many.argtypes = [C.c_void_p, f32p, f32p, C.c_uint64, C.c_uint64, u32p, u32p, u64p, u32p,
                                        C.POINTER(C.c_int64)]
    _libs[mutant] = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def set_threads(n):
    lib().oracle_set_threads(int(n))


def rknn(ux, uy, uid, fx, fy, q, k, cand=None, mutant=0):
    """Sorted user ids of R(q,k) over C = cand (None: every facility but q)."""
    ux, uy = np.ascontiguousarray(ux, np.float32), np.ascontiguousarray(uy, np.float32)
    uid = np.ascontiguousarray(uid, np.int32)
    fx, fy = np.ascontiguousarray(fx, np.float32), np.ascontiguousarray(fy, np.float32)
    out = np.empty(max(1, len(ux)), np.int32)
    if cand is None:
        cp, nc = None, 0
    else:
        c = np.ascontiguousarray(cand, np.uint32)
        cp, nc = _p(c, C.c_uint32), len(c)
    n = lib(mutant).oracle_rknn(_p(ux, C.c_float), _p(uy, C.c_float), _p(uid, C.c_int32), len(ux),
                                _p(fx, C.c_float), _p(fy, C.c_float), len(fx), int(q), int(k), cp, nc,
                                _p(out, C.c_int32))
    if n < 0:
        raise ValueError("oracle_rknn: invalid arguments")
    return out[:n].copy()


def counts(ux, uy, fx, fy, q, cand=None, cap=0, mutant=0):
    """Per-user count of candidates strictly closer than facility q (cap 0 = all)."""
    ux, uy = np.ascontiguousarray(ux, np.float32), np.ascontiguousarray(uy, np.float32)
    fx, fy = np.ascontiguousarray(fx, np.float32), np.ascontiguousarray(fy, np.float32)
    out = np.zeros(max(1, len(ux)), np.uint32)
    if cand is None:
        cp, nc = None, 0
    else:
        c = np.ascontiguousarray(cand, np.uint32)
        cp, nc = _p(c, C.c_uint32), len(c)
    r = lib(mutant).oracle_counts(_p(ux, C.c_float), _p(uy, C.c_float), len(ux), _p(fx, C.c_float),
                                  _p(fy, C.c_float), len(fx), int(q), cp, nc, int(cap), _p(out, C.c_uint32))
    if r != 0:
        raise ValueError("oracle_counts: invalid arguments")
    return out[: len(ux)].copy()


def rknn_naive_int(ux, uy, uid, fx, fy, fid, q, k):
    """Literal rknn-naive on integer coordinates (ids in user order)."""
    a = [np.ascontiguousarray(v, np.int32) for v in (ux, uy, uid, fx, fy, fid)]
    out = np.empty(max(1, len(a[0])), np.int32)
    n = lib().oracle_rknn_naive_int(_p(a[0], C.c_int32), _p(a[1], C.c_int32), _p(a[2], C.c_int32), len(a[0]),
                                    _p(a[3], C.c_int32), _p(a[4], C.c_int32), _p(a[5], C.c_int32), len(a[3]),
                                    int(q), int(k), _p(out, C.c_int32))
    if n < 0:
        raise ValueError("oracle_rknn_naive_int: invalid arguments")
    return out[:n].copy()


class Tree:
    """The reference's tree-culled verification (oracle_tree.c, rknn-inf/inf.h:194-249)
    over one user set; build time is kept apart from query time, as the reference reports it."""

    def __init__(self, ux, uy, uid):
        self.ux = np.ascontiguousarray(ux, np.float32)
        self.uy = np.ascontiguousarray(uy, np.float32)
        self.uid = np.ascontiguousarray(uid, np.int32)
        self.h = lib().oracle_tree_build(_p(self.ux, C.c_float), _p(self.uy, C.c_float), _p(self.uid, C.c_int32),
                                         len(self.ux))
        if not self.h:
            raise MemoryError("oracle_tree_build")

    def rknn(self, fx, fy, q, k, cand=None):
        fx, fy = np.ascontiguousarray(fx, np.float32), np.ascontiguousarray(fy, np.float32)
        out = np.empty(max(1, len(self.ux)), np.int32)
        if cand is None:
            cp, nc = None, 0
        else:
            c = np.ascontiguousarray(cand, np.uint32)
            cp, nc = _p(c, C.c_uint32), len(c)
        n = lib().oracle_tree_rknn(self.h, _p(fx, C.c_float), _p(fy, C.c_float), len(fx), int(q), int(k), cp, nc,
                                   _p(out, C.c_int32))
        if n < 0:
            raise ValueError("oracle_tree_rknn: invalid arguments")
        return out[:n].copy()

    def counts_many(self, fx, fy, qs, ks, cands=None):
        """|R(q_i, k_i)| for many queries on the worker threads (set_threads)."""
        fx, fy = np.ascontiguousarray(fx, np.float32), np.ascontiguousarray(fy, np.float32)
        q = np.ascontiguousarray(qs, np.uint32)
        k = np.ascontiguousarray(ks, np.uint32)
        out = np.zeros(max(1, len(q)), np.int64)
        if cands is None:
            op = cp = None
        else:
            off = np.zeros(len(cands) + 1, np.uint64)
            off[1:] = np.cumsum([len(c) for c in cands])
            flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.uint32) for c in cands])
                                        if len(cands) else np.zeros(1, np.uint32), np.uint32)
            op, cp = _p(off, C.c_uint64), _p(flat, C.c_uint32)
        r = lib().oracle_tree_rknn_many(self.h, _p(fx, C.c_float), _p(fy, C.c_float), len(fx), len(q),
                                        _p(q, C.c_uint32), _p(k, C.c_uint32), op, cp, _p(out, C.c_int64))
        if r != 0:
            raise ValueError("oracle_tree_rknn_many: invalid arguments")
        return out[: len(q)].copy()

    def close(self):
        if self.h:
            lib().oracle_tree_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def ref_binary(name):
    """Path of a reference CPU program built from source into oracle/_ref, or None."""
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None
