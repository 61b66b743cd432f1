"""C ABI boundary checks that need no GPU: the library loads, exports every
function include/*.h declares, and fails cleanly without a device."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("rknn.h", "rknn_io.h")]
NO_DEVICE = not os.path.exists("/dev/kfd")


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rk_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_headers_declare_the_abi():
    names = declared_functions()
    for must in ["rk_init", "rk_shutdown", "rk_dataset_create", "rk_filter", "rk_query_plain", "rk_query_batch",
                 "rk_batch_prepare", "rk_serve_start", "rk_serve_submit", "rk_serve_poll", "rk_serve_stop",
                 "rk_status_string", "rk_count_all", "rk_query_bruteforce", "rk_load_text", "rk_generate"]:
        assert must in names


def test_library_exports_every_declared_symbol(rk):
    lib_path = rk.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (rk_[a-z_0-9]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert missing == []
    # and the Python mirror binds exactly those
    from rknn_amd import _lib

    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_status_strings(rk):
    lib = rk.load()
    for s in range(10):
        assert lib.rk_status_string(s)
    assert lib.rk_status_string(99) == b"unknown status"


@pytest.mark.skipif(not NO_DEVICE, reason="a GPU is present")
def test_no_device_is_an_error_not_a_fallback(rk):
    import ctypes as C

    lib = rk.load()
    assert lib.rk_device_count() == 0
    h = C.c_void_p()
    assert lib.rk_init(0, C.byref(h)) == 1  # RK_ERR_NO_DEVICE
    with pytest.raises(rk.RkError) as e:
        rk.Context(0)
    assert e.value.status == 1


def test_missing_library_raises(monkeypatch):
    from rknn_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/librknn.so")
    with pytest.raises(RuntimeError):
        _lib.load()


def test_null_arguments_rejected(rk):
    import ctypes as C

    lib = rk.load()
    n = C.c_size_t()
    assert lib.rk_filter_points(None, None, 5, None, None, 0, 0, 1, None, 0, C.byref(n)) == 2
    assert lib.rk_init(0, None) == 2
    assert lib.rk_dataset_destroy(None) == 2
    assert lib.rk_batch_run(None, None) == 2
