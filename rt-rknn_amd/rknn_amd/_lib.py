"""ctypes binding of lib/librknn.so (include/rknn.h, include/rknn_io.h).

The library is the only compute path: importing this module when the shared
library has not been built raises, and every query goes through the gfx950
kernels.  There is no CPU fallback.
"""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "librknn.so")

RK_OK = 0
RK_ERR_NO_DEVICE = 1
RK_ERR_INVALID_ARG = 2
RK_ERR_OUT_OF_MEMORY = 3
RK_ERR_CAPACITY = 4
RK_ERR_HIP = 5
RK_ERR_IO = 6
RK_ERR_ADMISSION = 7
RK_ERR_NOT_READY = 8
RK_ERR_INTERNAL = 9

# limits of the allocation-free plain form (include/rknn.h)
RK_NOALLOC_MAX_CANDIDATES = 4096
RK_NOALLOC_MAX_IDS = 1048576


class rk_timing(C.Structure):
    _fields_ = [
        ("filter_s", C.c_double),
        ("upload_s", C.c_double),
        ("kernel_s", C.c_double),
        ("download_s", C.c_double),
        ("total_s", C.c_double),
        ("n_candidates", C.c_uint64),
        ("n_fp64_rechecks", C.c_uint64),
        ("n_launches", C.c_uint64),
    ]


class rk_query_desc(C.Structure):
    _fields_ = [
        ("q", C.c_uint32),
        ("k", C.c_uint32),
        ("cand", C.POINTER(C.c_uint32)),
        ("nc", C.c_uint64),
    ]


class rk_serve_config(C.Structure):
    _fields_ = [
        ("max_batch", C.c_uint32),
        ("memory_budget", C.c_uint64),
        ("max_wait_s", C.c_double),
    ]


# every symbol of include/rknn.h and include/rknn_io.h with its signature
P = C.c_void_p
F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int32)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
SZP = C.POINTER(C.c_size_t)
TP = C.POINTER(rk_timing)
ST = C.c_int

SIGNATURES = {
    # rknn.h
    "rk_init": (ST, [C.c_int, C.POINTER(P)]),
    "rk_shutdown": (ST, [P]),
    "rk_device_count": (C.c_int, []),
    "rk_status_string": (C.c_char_p, [ST]),
    "rk_last_error": (C.c_char_p, [P]),
    "rk_dataset_create": (ST, [P, F32P, F32P, I32P, C.c_size_t, F32P, F32P, I32P, C.c_size_t, C.POINTER(P)]),
    "rk_dataset_destroy": (ST, [P]),
    "rk_dataset_info": (ST, [P, U64P, U64P, U64P]),
    "rk_filter": (ST, [P, C.c_uint32, C.c_uint32, U32P, C.c_size_t, SZP]),
    "rk_filter_points": (ST, [F32P, F32P, C.c_size_t, F32P, F32P, C.c_size_t, C.c_uint32, C.c_uint32, U32P,
                              C.c_size_t, SZP]),
    "rk_filter_device": (ST, [P, U32P, U32P, C.c_size_t, U32P, C.c_size_t, U64P]),
    "rk_query_plain": (ST, [P, C.c_uint32, C.c_uint32, U32P, C.c_size_t, I32P, C.c_size_t, SZP, TP]),
    "rk_query_plain_noalloc": (ST, [P, C.c_uint32, C.c_uint32, U32P, C.c_size_t, I32P, C.c_size_t, SZP]),
    "rk_query_bruteforce": (ST, [P, C.c_uint32, C.c_uint32, I32P, C.c_size_t, SZP, TP]),
    "rk_query": (ST, [P, C.c_uint32, C.c_uint32, I32P, C.c_size_t, SZP, TP]),
    "rk_count_all": (ST, [P, C.c_uint32, U32P, C.c_size_t, U32P, TP]),
    "rk_count_all_sorted": (ST, [P, C.c_uint32, U32P, C.c_size_t, U32P, TP]),
    "rk_count_all_many": (ST, [P, C.POINTER(rk_query_desc), C.c_size_t, U32P, C.c_size_t, C.c_int, TP]),
    "rk_dataset_order": (ST, [P, U32P]),
    "rk_query_batch": (ST, [P, C.POINTER(rk_query_desc), C.c_size_t, I32P, C.c_size_t, U64P, TP]),
    "rk_batch_prepare": (ST, [P, C.POINTER(rk_query_desc), C.c_size_t, C.POINTER(P)]),
    "rk_batch_prepare_device": (ST, [P, U32P, U32P, C.c_size_t, C.POINTER(P)]),
    "rk_batch_run": (ST, [P, C.POINTER(C.c_double)]),
    "rk_batch_enqueue": (ST, [P]),
    "rk_batch_wait": (ST, [P, C.POINTER(C.c_double)]),
    "rk_batch_fetch": (ST, [P, I32P, C.c_size_t, U64P]),
    "rk_batch_results": (ST, [P, C.POINTER(C.c_void_p), U64P]),
    "rk_batch_destroy": (ST, [P]),
    "rk_batch_stats": (ST, [P, U64P, C.c_size_t, SZP, U64P, U64P]),
    "rk_batch_config": (ST, [P, U32P]),
    "rk_serve_start": (ST, [P, C.POINTER(rk_serve_config), C.POINTER(P)]),
    "rk_serve_submit": (ST, [P, P, C.c_uint32, C.c_uint32, U64P]),
    "rk_serve_poll": (ST, [P, C.c_uint64, I32P, C.c_size_t, SZP, C.POINTER(C.c_double)]),
    "rk_serve_info": (ST, [P, U64P, U32P]),
    "rk_serve_stop": (ST, [P]),
    # rknn_io.h
    "rk_load_text": (ST, [C.c_char_p, C.POINTER(F32P), C.POINTER(F32P), C.POINTER(I32P), U64P,
                          C.POINTER(F32P), C.POINTER(F32P), C.POINTER(I32P), U64P]),
    "rk_write_text": (ST, [C.c_char_p, F32P, F32P, I32P, C.c_uint64, F32P, F32P, I32P, C.c_uint64]),
    "rk_free": (None, [P]),
    "rk_generate_queries": (ST, [C.c_uint64, C.c_uint64, C.c_uint64, U32P]),
    "rk_generate": (ST, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                         F32P, F32P, I32P, F32P, F32P, I32P]),
    "rk_default_seed": (C.c_uint64, [C.c_int]),
}

_lib = None


def load():
    """Load the engine library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP engine first "
            "(python -c 'import __graft_entry__ as g; g.build()' or make -C rt-rknn_amd)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
