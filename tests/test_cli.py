"""rknn-hip: the reference's CLI flags, error behaviour and stdout layout
(rknn-inf/inf.cpp:38-119, rknn-inf-gpu/inf.cpp:164-658)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "rt-rknn_amd", "bin", "rknn-hip")
NO_DEVICE = not os.path.exists("/dev/kfd")

need_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


@pytest.fixture
def small_file(rk, tmp_path):
    arrs = rk.generate(1, 3000, 200)
    p = str(tmp_path / "d.txt")
    rk.write_text(p, *arrs)
    return p, arrs


@need_cli
def test_unknown_option():
    r = run("-zz")
    assert r.returncode == 1 and "Unknown option '-zz'" in r.stderr and "Usage" in r.stderr


@need_cli
def test_missing_file():
    r = run("-if", "/nonexistent.txt")
    assert r.returncode == 1 and "Error opening file: /nonexistent.txt" in r.stderr


@need_cli
def test_bad_k(small_file):
    r = run("-if", small_file[0], "-k", "0")
    assert r.returncode == 1 and "Invalid value for k" in r.stderr


@need_cli
def test_bad_q(small_file):
    r = run("-if", small_file[0], "-q", "5000")
    assert r.returncode == 1 and "Invalid value for q" in r.stderr


@need_cli
@pytest.mark.skipif(not NO_DEVICE, reason="a GPU is present")
def test_no_device_is_an_error(small_file):
    r = run("-if", small_file[0], "-k", "3", "-q", "1")
    assert r.returncode == 1 and "no HIP device" in r.stderr
    assert "Loaded 200 facilities and 3000 users." in r.stdout


@need_cli
@pytest.mark.gpu
@pytest.mark.parametrize("brute", [False, True])
def test_log_layout_and_ids(small_file, tmp_path, brute):
    path, (ux, uy, uid, fx, fy, fid) = small_file
    ids_path = str(tmp_path / "ids.txt")
    args = ["-if", path, "-k", "4", "-q", "7", "-al", "ignored", "--ids", ids_path]
    if brute:
        args.append("--bruteforce")
    r = run(*args)
    assert r.returncode == 0, r.stderr
    exp = O.rknn(ux, uy, uid, fx, fy, 7, 4)
    lines = [line for line in r.stdout.splitlines() if line]
    order = ["Loading data...", "Loaded 200 facilities and 3000 users.", "building R*-trees...",
             "R*-tree is built in", "Filtering...", "Filtering time:", "Total bisectors created:",
             "Verifying...", "Using GPU acceleration for verification", f"Found {len(exp)} RkNN results.",
             "Verification time (GPU):", "Total time:"]
    pos = 0
    for want in order:
        while pos < len(lines) and not lines[pos].startswith(want):
            pos += 1
        assert pos < len(lines), (want, r.stdout)
    got = np.loadtxt(ids_path, dtype=np.int32, ndmin=1) if len(exp) else np.zeros(0, np.int32)
    assert np.array_equal(got, exp)


# ---- rknn-naive-hip: rknn-naive-gpu's command line (rknn-naive-gpu/rknn.cpp:262-352)

NAIVE = os.path.join(ROOT, "rt-rknn_amd", "bin", "rknn-naive-hip")
need_naive = pytest.mark.skipif(not os.path.exists(NAIVE), reason="CLI not built")


def run_naive(*args):
    return subprocess.run([NAIVE, *args], capture_output=True, text=True, timeout=120)


@need_naive
def test_naive_flags_and_errors(small_file):
    r = run_naive("-zz")
    assert r.returncode == 1 and "Unknown option '-zz'" in r.stderr and "--gpu" in r.stderr
    r = run_naive("-if", "/nonexistent.txt", "--gpu")
    assert r.returncode == 1 and "Error opening file: /nonexistent.txt" in r.stderr
    assert r.stdout.startswith("GPU acceleration enabled\nLoading data...")
    r = run_naive("-if", small_file[0], "-k", "0")
    assert r.returncode == 1 and "Invalid value for k" in r.stderr
    r = run_naive("-if", small_file[0], "-q", "200")  # 200 facilities: indices 0..199
    assert r.returncode == 1 and "Invalid value for q" in r.stderr


@need_naive
@pytest.mark.skipif(not NO_DEVICE, reason="a GPU is present")
def test_naive_no_device_is_an_error(small_file):
    r = run_naive("-if", small_file[0], "-k", "3", "-q", "1", "--gpu")
    assert r.returncode == 1 and "no HIP device" in r.stderr
    assert "Loaded 200 facilities and 3000 users." in r.stdout


@need_naive
@pytest.mark.gpu
def test_naive_log_and_ids_vs_oracle(rk, small_file, tmp_path):
    path, (ux, uy, uid, fx, fy, fid) = small_file
    ids_path = str(tmp_path / "ids.txt")
    r = run_naive("-if", path, "-k", "5", "-q", "7", "--gpu", "-al", "x", "--ids", ids_path)
    assert r.returncode == 0, r.stderr
    exp = O.rknn(ux, uy, uid, fx, fy, 7, 5)  # every other facility a candidate
    lines = [line for line in r.stdout.splitlines() if line]
    assert lines[:4] == ["GPU acceleration enabled", "Loading data...", "Loaded 200 facilities and 3000 users.",
                         "Using GPU: HIP device 0"]
    assert lines[4].startswith("RkNN search (GPU) completed in ") and lines[4].endswith("[s].")
    assert lines[5] == f"Found {len(exp)} RkNN results"
    got = np.loadtxt(ids_path, dtype=np.int32, ndmin=1) if len(exp) else np.zeros(0, np.int32)
    assert np.array_equal(got, exp)


@need_naive
@pytest.mark.gpu
def test_naive_query_by_facility_id(rk, tmp_path):
    """The reference kernel finds the query as the first facility carrying
    facility q's id (rknn-naive-gpu/rknn_kernel.cu:135-141): with that id
    duplicated at a lower index, the query is that facility."""
    ux, uy, uid, fx, fy, fid = rk.generate(1, 3000, 200)
    fid = fid.copy()
    fid[2] = fid[7]
    p = str(tmp_path / "dup.txt")
    rk.write_text(p, ux, uy, uid, fx, fy, fid)
    ids_path = str(tmp_path / "ids.txt")
    r = run_naive("-if", p, "-k", "4", "-q", "7", "--ids", ids_path)
    assert r.returncode == 0, r.stderr
    exp = O.rknn(ux, uy, uid, fx, fy, 2, 4)
    assert f"Found {len(exp)} RkNN results" in r.stdout
    got = np.loadtxt(ids_path, dtype=np.int32, ndmin=1) if len(exp) else np.zeros(0, np.int32)
    assert np.array_equal(got, exp)
