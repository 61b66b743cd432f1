"""Headline benchmark: RkNN queries/s at 10M users, 100k facilities, k=10,
batched (one graph-captured launch sequence per step) versus plain (one
counting launch, sync and readback per query), bit-exact sets.

A step = one pass of the batched form over one batch of NQ queries whose
candidate lists (from the host influence-zone filter) are resident in HBM;
the users are Hilbert-sorted and resident.  The step ends with every query's
sorted result ids in pinned host memory: its last kernel writes them there.
`value` counts queries/s over exactly --steps timed steps that rotate over
--batches distinct batches (8 by default: 8,192 distinct queries), three in
flight, each step's results taken in place (rk_batch_results) after its wait;
`value_fetch` is the same with every step's ids copied into a caller buffer
(rk_batch_fetch).  The JSON line also carries:
  roofline      dense counting kernel (count-all: every user x every candidate,
                4 fp32 flops per test): the first 32 queries in one
                rk_count_all_many call, which counts up to 16 queries per
                launch (k_count_set), HIP events on the kernel's stream around
                the launches, divided by the launch count; beside it the same
                queries at one launch each (k_count_all) and, as kmix_tflops,
                the k ~ U{1..50} lists
  count_all_call_ms  the whole count-all call, counts in caller order and in
                the dataset's order, with the copy of 10M counts to the host,
                and the same calls into device memory (no PCIe copy)
  cpu_baseline  the reference's tree-culled verification (rknn-inf) restated
                in C, exact in binary64, on the step's queries and candidate
                lists, one host thread (cpu_baseline_mt: every host core;
                cpu_bruteforce: rknn-naive semantics on two queries)
  cpu_reference the reference's own CPU programs, built from its sources into
                oracle/_ref: rknn-inf on the rknn-inf configuration (1M users,
                10k facilities), rknn-naive on the 10k/1k configuration; host
                core count and CPU model
  plain_*       the one-launch-per-query form on batch 0's queries
  mixed_k       the step with k drawn uniformly from 1..50 per query
  device_filtered_step  the step with the influence-zone filter inside it
                (queries in, results in host memory)
  end_to_end    one rk_query_batch call: filter (device, or host cores) + step +
                the copy into the caller's buffer, on the library's own clock,
                at k=10 and at k ~ U{1..50}
  serve         the serving loop under Poisson arrivals (rknn-serve-bench), k=10
  serve_datasets  the same with 1M/10k, 10M/100k and 50M/200k resident together, k ~ U{1..50}
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "rt-rknn_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# HBM bytes per count-all launch of 16 queries on the 10M-user, 100k-facility
# k=10 workload, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by a copy
# kernel of known size (profiles/k_count_all_10M_100k.md); the kernel reads
# every user once and writes one 8-bit count per user and query (the k=10
# lists hold fewer than 256 candidates), so the figure depends on the number
# of queries in the launch, not on which they are
MEASURED_TRAFFIC = {"valu4": 244.9e6, "valu6": 246.1e6, "valu8": 245.9e6}
MEASURED_TRAFFIC_QUERIES = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def overlapped(batches, steps, warmup):
    """Seconds per step over batch objects of the same queries, each enqueued
    while the others run (each batch runs on its own stream); the batches'
    results must agree, and the first one's are returned."""
    n = len(batches)
    for _ in range(warmup):
        for b in batches:
            b.enqueue()
        for b in batches:
            b.wait()
    pending = []
    t0 = time.perf_counter()
    for s in range(steps):
        b = batches[s % n]
        if len(pending) == n:
            pending.pop(0).wait()
        b.enqueue()
        pending.append(b)
    while pending:
        pending.pop(0).wait()
    per_step = (time.perf_counter() - t0) / steps
    res = batches[0].fetch()
    for b in batches[1:]:
        if not all(np.array_equal(a, c) for a, c in zip(res, b.fetch())):
            log("ERROR: batch objects of the same queries differ")
            sys.exit(1)
    return per_step, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--users", type=int, default=10_000_000)
    ap.add_argument("--facilities", type=int, default=100_000)
    ap.add_argument("--queries", type=int, default=1024)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--plain", type=int, default=1024, help="queries timed in the plain form")
    ap.add_argument("--roofline-queries", type=int, default=32)
    ap.add_argument("--cpu-queries", type=int, default=2, help="queries of the brute-force CPU context run")
    ap.add_argument("--inflight", type=int, default=3,
                    help="batches in flight in the timed steps, each on its own stream")
    ap.add_argument("--batches", type=int, default=8,
                    help="distinct batches of --queries queries rotated through the timed steps")
    ap.add_argument("--ref-queries", type=int, default=2,
                    help="queries per reference CPU program (cpu_reference; one process each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-qps", default="100000,300000,500000,1000000")
    ap.add_argument("--serve-seconds", type=float, default=0.5)
    ap.add_argument("--serve-datasets", default="1000000:10000,10000000:100000,50000000:200000",
                    help="users:facilities of the datasets resident together for the serve_datasets line")
    ap.add_argument("--serve-datasets-qps", default="100000,200000,300000")
    ap.add_argument("--table", action="store_true",
                    help="instead of the headline: one JSON line per config 1-3 (plain and batched forms, the CPU "
                         "tree restatement and, where built, the reference's own rknn-inf)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (queries shard with no "
              "communication; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import rknn_amd as R

    if args.table:
        return table(R)

    t0 = time.time()
    seed = R.default_seed(args.config)
    ux, uy, uid, fx, fy, fid = R.generate(args.config, args.users, args.facilities, seed=seed,
                                          nq_ties=args.queries, seed_ties=seed)
    qs = [int(q) for q in R.generate_queries(args.facilities, args.queries, seed)]
    log(f"generated config {args.config}: {args.users} users, {args.facilities} facilities "
        f"in {time.time() - t0:.1f}s")

    ctx = R.Context(0)
    t0 = time.time()
    ds = R.Dataset(ctx, ux, uy, uid, fx, fy, fid)
    log(f"dataset resident in {time.time() - t0:.2f}s ({ds.device_bytes() / 2**20:.0f} MiB)")

    # host influence-zone filter for every query of the rotation (outside the
    # timed region; reported)
    nb = max(1, args.batches)
    nq = args.queries
    qs_all = [int(q) for q in R.generate_queries(args.facilities, nb * nq, seed)]
    t0 = time.time()
    cands_all = [ds.filter(q, args.k) for q in qs_all]
    filter_s = time.time() - t0
    ncs = np.array([len(c) for c in cands_all])
    log(f"filter: {nb * nq} queries in {filter_s:.2f}s, |C| mean {ncs.mean():.1f} "
        f"min {ncs.min()} max {ncs.max()}")
    qs, cands = qs_all[:nq], cands_all[:nq]  # batch 0: the other sections' queries

    # ---- batched form: nb batches of distinct queries, each prepared once
    # (candidate lists uploaded, graph captured) and rotated step after step,
    # --inflight of them in flight (rk_batch_enqueue / rk_batch_wait, each batch on its own
    # stream); a step is done when its results are in host memory.  Their
    # first run sizes each batch's landing area (and captures its graph again).
    batches = [ds.prepare_batch([(q, args.k, c) for q, c in zip(qs_all[j * nq:(j + 1) * nq],
                                                                 cands_all[j * nq:(j + 1) * nq])])
               for j in range(nb)]
    for b in batches:
        b.run()
    batch = batches[0]
    for _ in range(args.warmup):
        batch.run()
    t0 = time.perf_counter()
    dev_sync = [batch.run() for _ in range(args.steps)]
    sync_ms = (time.perf_counter() - t0) / args.steps * 1e3
    n_inflight = max(1, args.inflight)

    def rotate(steps, take):
        """Seconds per step over the rotation; take(b) after each wait.  A
        batch holds one run at a time, so at most nb are in flight."""
        depth = min(n_inflight, nb)
        pending = []
        for s in range(max(steps, 0)):
            b = batches[s % nb]
            if len(pending) == depth:
                c = pending.pop(0)
                c.wait()
                take(c)
            b.enqueue()
            pending.append(b)
        while pending:
            c = pending.pop(0)
            c.wait()
            take(c)

    rotate(args.warmup * nb, lambda b: b.results_raw())
    t0 = time.perf_counter()
    rotate(args.steps, lambda b: b.results_raw())
    wall = time.perf_counter() - t0
    rotate(nb, lambda b: b.fetch_flat())
    t0 = time.perf_counter()
    rotate(args.steps, lambda b: b.fetch_flat())
    wall_fetch = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    value = nq * args.steps / wall
    value_fetch = nq * args.steps / wall_fetch
    # every batch's results, in place and fetched, must agree (checked after the timing)
    batched_all = []
    for b in batches:
        b.run()
        got = b.fetch()
        view = b.results()
        if not all(np.array_equal(a, v) for a, v in zip(got, view)):
            log("ERROR: a batch's in-place results differ from its fetched ones")
            sys.exit(1)
        batched_all.append(got)
    batched = batched_all[0]
    items, n_res, n_rech = batch.stats()
    log(f"batched work per step: items per level (leaf first) {items}, results {n_res}, fp64 re-checks {n_rech}; "
        f"autotuned block: {batch.waves_per_query()} waves per query")
    log(f"batched, {nb} batches rotating, results in host memory: {ms_step:.3f} ms/step, {value:,.0f} queries/s "
        f"(ids also copied to a caller buffer: {nq * args.steps / wall_fetch:,.0f}); one run at a time "
        f"{sync_ms:.3f} ms, device {np.mean(dev_sync) * 1e3:.3f} ms")

    # ---- mixed k (1..50 per query, the batched workload's k mix), same step
    rng = np.random.default_rng(seed)
    ks = rng.integers(1, 51, size=args.queries)
    mcands = [ds.filter(q, int(k)) for q, k in zip(qs, ks)]
    # (timed as the k=10 step: batches alternating, each on its own stream;
    # one run at a time beside it)
    mbs = [ds.prepare_batch([(q, int(k), c) for q, k, c in zip(qs, ks, mcands)]) for _ in range(n_inflight)]
    for _ in range(args.warmup):
        mbs[0].run()
    t0 = time.perf_counter()
    mdev = [mbs[0].run() for _ in range(args.steps)]
    msync = (time.perf_counter() - t0) / args.steps
    mstep, mixed_res = overlapped(mbs, args.steps, args.warmup)
    for b in mbs:
        b.close()
    mixed = {"qps": round(args.queries / mstep, 1), "ms_per_step": round(mstep * 1e3, 4),
             "sync_ms_per_step": round(msync * 1e3, 4), "device_ms": round(float(np.mean(mdev)) * 1e3, 4),
             "candidates_mean": round(float(np.mean([len(c) for c in mcands])), 2)}
    for i in range(min(32, args.queries)):  # spot check against the plain form
        if not np.array_equal(ds.query_plain(qs[i], int(ks[i]), mcands[i]), mixed_res[i]):
            log("ERROR: mixed-k batched and plain result sets differ")
            sys.exit(1)
    log(f"mixed k 1..50: {mixed['ms_per_step']:.3f} ms/step (one run at a time {mixed['sync_ms_per_step']:.3f} ms, "
        f"device {mixed['device_ms']:.3f} ms), "
        f"{mixed['qps']:,.0f} queries/s, |C| mean {mixed['candidates_mean']}")

    # ---- device-filtered step: (q, k) in HBM, the influence-zone filter inside
    # the step's graph (filter, per-query setup, traversal, sort); same queries
    dfilter = {}
    for name, kk, ref in (("k10", [args.k] * args.queries, batched), ("mixed_k", [int(k) for k in ks], mixed_res)):
        dbs = [ds.prepare_batch(list(zip(qs, kk)), device_filter=True) for _ in range(n_inflight)]
        for _ in range(args.warmup):
            dbs[0].run()
        t0 = time.perf_counter()
        ddev = [dbs[0].run() for _ in range(args.steps)]
        dsync = (time.perf_counter() - t0) / args.steps
        dstep, dres = overlapped(dbs, args.steps, args.warmup)
        for b in dbs:
            b.close()
        same = all(np.array_equal(a, b) for a, b in zip(dres, ref))
        dfilter[name] = {"qps": round(args.queries / dstep, 1), "ms_per_step": round(dstep * 1e3, 4),
                         "sync_ms_per_step": round(dsync * 1e3, 4),
                         "device_ms": round(float(np.mean(ddev)) * 1e3, 4), "same_sets": bool(same)}
        log(f"device-filtered step ({name}): {dstep * 1e3:.3f} ms/step (one run at a time {dsync * 1e3:.3f} ms, "
            f"device {np.mean(ddev) * 1e3:.3f} ms), {args.queries / dstep:,.0f} queries/s; same sets: {same}")
        if not same:
            log("ERROR: device-filtered and host-filtered batched result sets differ")
            sys.exit(1)

    # ---- end to end, one call (rk_query_batch): filter + step + download;
    # the filter on the device (default for k <= 64) and on the host's cores
    # (the library's own clock around the call, rk_timing.total_s, is the
    # C-ABI figure; the wall time of the Python wrapper is reported beside it)
    end_to_end = {}
    for name, filt, kk, ref in (("device_filter", "device", [args.k] * nq, batched),
                                ("host_filter", "host", [args.k] * nq, batched),
                                ("mixed_k_device_filter", "device", [int(k) for k in ks], mixed_res),
                                ("mixed_k_host_filter", "host", [int(k) for k in ks], mixed_res)):
        os.environ["RK_FILTER"] = filt
        ds.query_batch(list(zip(qs, kk)))  # warm: the dataset's batch buffers grow once
        calls, e2e_ok = [], True
        for _ in range(3):
            t0 = time.perf_counter()
            e2e_res, tm = ds.query_batch(list(zip(qs, kk)), timing=True)
            calls.append((tm.total_s, time.perf_counter() - t0, tm.filter_s, tm.kernel_s, tm.download_s,
                          tm.upload_s))
            e2e_ok = e2e_ok and all(np.array_equal(a, b) for a, b in zip(e2e_res, ref))
        lib_s, py_s, f_s, k_s, d_s, u_s = (float(np.median([c[j] for c in calls])) for j in range(6))
        end_to_end[name] = {"qps": round(nq / lib_s, 1), "ms": round(lib_s * 1e3, 3),
                            "filter_ms": round(f_s * 1e3, 3), "fill_ms": round(u_s * 1e3, 3),
                            "device_ms": round(k_s * 1e3, 3), "download_ms": round(d_s * 1e3, 3),
                            "python_wall_ms": round(py_s * 1e3, 3), "matches": bool(e2e_ok)}
        log(f"end to end, {name} (filter + step + copy to the caller): {nq / lib_s:,.0f} queries/s "
            f"(library clock, median of 3: filter {f_s * 1e3:.3f}, fill {u_s * 1e3:.3f}, device {k_s * 1e3:.3f}, "
            f"copy {d_s * 1e3:.3f} ms; {py_s * 1e3:.2f} ms through the Python wrapper); same sets: {e2e_ok}")
        if not e2e_ok:
            break
    del os.environ["RK_FILTER"]
    if not all(v["matches"] for v in end_to_end.values()):
        log("ERROR: end-to-end results differ from the batched step")
        sys.exit(1)

    # ---- plain form: one counting launch + sync + readback per query
    n_plain = min(args.plain, args.queries)
    lat = []
    for i in range(min(8, n_plain)):  # warm
        ds.query_plain(qs[i], args.k, cands[i])
    mismatches = 0
    t0 = time.perf_counter()
    for i in range(n_plain):
        s = time.perf_counter()
        got = ds.query_plain(qs[i], args.k, cands[i])
        lat.append(time.perf_counter() - s)
        mismatches += int(not np.array_equal(got, batched[i]))
    plain_wall = time.perf_counter() - t0
    plain_qps = n_plain / plain_wall
    lat_ms = np.array(lat) * 1e3
    # the rotation's other batches: 32 queries each against the plain form
    checked = n_plain
    for j in range(1, nb):
        for i in range(min(32, nq)):
            q = qs_all[j * nq + i]
            mismatches += int(not np.array_equal(ds.query_plain(q, args.k, cands_all[j * nq + i]),
                                                 batched_all[j][i]))
            checked += 1
    log(f"plain: {plain_qps:,.0f} queries/s, p50 {np.percentile(lat_ms, 50):.3f} ms, "
        f"p99 {np.percentile(lat_ms, 99):.3f} ms; batched == plain on {checked - mismatches}/{checked}")
    if mismatches:
        log("ERROR: batched and plain result sets differ")
        sys.exit(1)

    # ---- roofline: dense counting kernel (count-all), HIP events on its stream;
    # every variant (4, 6 or 8 users per lane; 3 VALU instructions per clean
    # test).  rk_count_all_many counts consecutive queries whose lists fit the
    # kernel's LDS staging together in one launch (k_count_set: up to 16
    # queries, 1,024 records); a launch's algorithmic flops are its queries'
    # 4 * nu * |C| summed
    variants = {}
    # (8-bit counts: the k=10 lists hold fewer than 256 candidates)
    kernel_names = {"valu4": "k_count_set<4,uint8_t>", "valu6": "k_count_set<6,uint8_t>",
                    "valu8": "k_count_set<8,uint8_t>"}
    kts = {kern: [] for kern in kernel_names}
    flops = {kern: 0.0 for kern in kernel_names}
    n_roof = min(args.roofline_queries, args.queries)
    roof_queries = [(qs[i], cands[i]) for i in range(n_roof)]
    roof_out = np.zeros((n_roof, args.users), dtype=np.uint32)
    launches = 0
    for rnd in range(6):  # the variants interleaved: one untimed round, then five timed
        for kern in kernel_names:
            os.environ["RK_COUNT_KERNEL"] = kern
            # the queries' count launches back to back on the kernel's stream,
            # HIP events around them: kernel_s / n_launches is the average launch
            _, t = ds.count_all_many(roof_queries, timing=True, out=roof_out)
            if rnd == 0:
                continue
            launches = t.n_launches
            kts[kern].append(t.kernel_s / t.n_launches)
            flops[kern] += 4.0 * args.users * t.n_candidates / t.n_launches
    for i in range(n_roof):  # the call's counts are the single-query call's
        if not np.array_equal(roof_out[i], ds.count_all(qs[i], cands[i])):
            log("ERROR: count-all over many queries differs from the one-query call")
            sys.exit(1)
    # the same queries with one launch each (k_count_all), for comparison
    per_query = {}
    os.environ["RK_COUNT_SET"] = "1"
    for kern in kernel_names:
        os.environ["RK_COUNT_KERNEL"] = kern
        best_q = None
        for _ in range(3):
            _, t = ds.count_all_many(roof_queries, timing=True, out=roof_out)
            tf = 4.0 * args.users * t.n_candidates / t.kernel_s / 1e12
            best_q = tf if best_q is None else max(best_q, tf)
        per_query[kern] = round(best_q, 3)
    del os.environ["RK_COUNT_SET"]
    del os.environ["RK_COUNT_KERNEL"]
    # the whole count-all call, counts in caller order (a scatter pass after
    # the kernel) and in the dataset's order (the kernel's output as it is),
    # each including its device-to-host copy of nu counts
    call_ms = {}
    counts_out = np.zeros(args.users, dtype=np.uint32)
    for name, so in (("caller_order", False), ("dataset_order", True)):
        tt, kk, dd = [], [], []
        for i in range(min(8, args.queries)):
            t0 = time.perf_counter()
            _, t = ds.count_all(qs[i], cands[i], timing=True, sorted_order=so, out=counts_out)
            tt.append(time.perf_counter() - t0)
            kk.append(t.kernel_s)
            dd.append(t.download_s)
        call_ms[name] = {"call": round(float(np.median(tt)) * 1e3, 3),
                         "kernel": round(float(np.median(kk)) * 1e3, 4),
                         "order_and_copy": round(float(np.median(dd)) * 1e3, 3)}
        log(f"count-all call, {name}: {call_ms[name]['call']} ms (kernel {call_ms[name]['kernel']} ms, "
            f"order + copy of {args.users} counts {call_ms[name]['order_and_copy']} ms)")
    # the same calls into device memory (the caller keeps the counts on the
    # device, as rknn-rt's result buffer is): no PCIe copy; library clock
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    dptr = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(dptr), ctypes.c_size_t(4 * args.users)) == 0:
        lib = R.load()
        out_p = ctypes.cast(dptr, ctypes.POINTER(ctypes.c_uint32))
        for name, fn in (("caller_order_device", lib.rk_count_all), ("dataset_order_device", lib.rk_count_all_sorted)):
            tt, kk = [], []
            for i in range(min(8, args.queries)):
                c = np.ascontiguousarray(cands[i], dtype=np.uint32)
                t = R._lib.rk_timing()
                st = fn(ds.h, qs[i], c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(c), out_p,
                        ctypes.byref(t))
                if st != 0:
                    log(f"ERROR: count-all into device memory failed ({st})")
                    sys.exit(1)
                tt.append(t.total_s)
                kk.append(t.kernel_s)
            call_ms[name] = {"call": round(float(np.median(tt)) * 1e3, 3),
                             "kernel": round(float(np.median(kk)) * 1e3, 4)}
            log(f"count-all call, {name}: {call_ms[name]['call']} ms on the library's clock "
                f"(kernel {call_ms[name]['kernel']} ms)")
        hip.hipFree(dptr)
    # the roofline's queries in one rk_count_all_many call, in the dataset's
    # order into device memory (the count kernels write the caller's rows)
    mptr = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(mptr), ctypes.c_size_t(4 * args.users * n_roof)) == 0:
        lib = R.load()
        descs, keep = ds._descs([(q, 1, c) for q, c in roof_queries])
        tt, kk = [], []
        for _ in range(3):
            t = R._lib.rk_timing()
            st = lib.rk_count_all_many(ds.h, descs, n_roof, ctypes.cast(mptr, ctypes.POINTER(ctypes.c_uint32)),
                                       args.users, 1, ctypes.byref(t))
            if st != 0:
                log(f"ERROR: count-all over many queries into device memory failed ({st})")
                sys.exit(1)
            tt.append(t.total_s)
            kk.append(t.kernel_s)
        call_ms["many_dataset_order_device"] = {"queries": n_roof, "call": round(float(np.median(tt)) * 1e3, 3),
                                                "kernel": round(float(np.median(kk)) * 1e3, 4)}
        log(f"count-all over {n_roof} queries into device memory: {call_ms['many_dataset_order_device']['call']} ms "
            f"on the library's clock (kernels {call_ms['many_dataset_order_device']['kernel']} ms)")
        hip.hipFree(mptr)
    for kern, kt in kts.items():
        variants[kern] = (flops[kern] / sum(kt) / 1e12, float(np.mean(kt)), flops[kern] / len(kt))
        log(f"count-all kernel ({kern}): {variants[kern][0]:.2f} TFLOP/s, {n_roof} queries in {launches} launches "
            f"(avg {variants[kern][1] * 1e3:.4f} ms); one launch per query {per_query[kern]:.2f} TFLOP/s")
    best = max(variants, key=lambda v: variants[v][0])
    achieved, avg_s, flops_launch = variants[best]
    roofline = {"bound": "valu", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": (MEASURED_TRAFFIC.get(best) if (args.users, args.facilities, args.k, args.config)
                            == (10_000_000, 100_000, 10, 4) and n_roof == MEASURED_TRAFFIC_QUERIES * launches
                            else None),
                "traffic_source": "profiles/k_count_all_10M_100k.md (calibrated PMC, bytes per launch)",
                "kernel": kernel_names[best],
                "avg_launch_ms": round(avg_s * 1e3, 4), "flops_per_launch": flops_launch,
                "queries": n_roof, "launches": launches,
                "variants_tflops": {k: round(v[0], 3) for k, v in variants.items()},
                "one_launch_per_query_tflops": per_query}
    # the same kernel over the batched workload's own k mix (k ~ U{1..50}: the
    # mixed-k step's candidate lists), flop-weighted over its queries
    os.environ["RK_COUNT_KERNEL"] = best
    n_mix = min(args.roofline_queries, args.queries)
    _, t = ds.count_all_many([(qs[i], mcands[i]) for i in range(n_mix)], timing=True)
    kmix_s, kmix_f = t.kernel_s, 4.0 * args.users * t.n_candidates
    del os.environ["RK_COUNT_KERNEL"]
    roofline["kmix_tflops"] = round(kmix_f / kmix_s / 1e12, 3)
    roofline["kmix_frac"] = round(kmix_f / kmix_s / 1e12 / PEAK_FP32_TFLOPS, 4)
    log(f"count-all kernel ({best}) over the k ~ U{{1..50}} lists: {roofline['kmix_tflops']:.2f} TFLOP/s")

    # ---- CPU baseline: the reference's tree-culled verification restated in C
    # (oracle/oracle_tree.c, rknn-inf/inf.h:194-249) on the same queries and
    # candidate lists as the timed step; tree build reported apart, as the
    # reference does (inf.cpp:163-179)
    cpu = None
    cpu_mt = None
    cpu_bf = None
    if not args.no_cpu:
        import oracle

        t0 = time.perf_counter()
        tree = oracle.Tree(ux, uy, uid)
        build_s = time.perf_counter() - t0
        ks_all = [args.k] * args.queries
        oracle.set_threads(1)
        t0 = time.perf_counter()
        cnt1 = tree.counts_many(fx, fy, qs, ks_all, cands)
        cpu_s = time.perf_counter() - t0
        ok = all(int(cnt1[i]) == len(batched[i]) for i in range(args.queries))
        ok &= all(np.array_equal(tree.rknn(fx, fy, qs[i], args.k, cands[i]), batched[i]) for i in range(8))
        cpu = {"value": round(args.queries / cpu_s, 2), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"all {args.queries} queries of the timed step with the same candidate lists: tree-culled "
                         f"verification (rknn-inf inf.h:194-249 restated, binary64, exact), one host thread; "
                         f"user tree built once in {build_s:.1f}s, outside the rate as in the reference",
               "matches_gpu": bool(ok)}
        log(f"cpu tree verification (1 thread): {args.queries / cpu_s:,.1f} queries/s "
            f"(tree build {build_s:.1f}s); GPU sets equal: {ok}")
        if not ok:
            sys.exit(1)
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        cntn = tree.counts_many(fx, fy, qs, ks_all, cands)
        mt_s = time.perf_counter() - t0
        ok_mt = bool(np.array_equal(cntn, cnt1))
        cpu_mt = {"value": round(args.queries / mt_s, 2), "unit": "queries/s", "cores": threads, "kind": "port",
                  "sample": f"the same {args.queries} queries, one query per thread", "matches_gpu": ok_mt}
        log(f"cpu tree verification ({threads} threads): {args.queries / mt_s:,.1f} queries/s; equal: {ok_mt}")
        if not ok_mt:
            sys.exit(1)
        # brute force (rknn-naive semantics) for context: two queries, one thread
        oracle.set_threads(1)
        n_bf = min(args.cpu_queries, args.queries)
        t0 = time.perf_counter()
        ok_bf = all(np.array_equal(oracle.rknn(ux, uy, uid, fx, fy, qs[i], args.k), batched[i]) for i in range(n_bf))
        bf_s = time.perf_counter() - t0
        cpu_bf = {"value": round(n_bf / bf_s, 4), "unit": "queries/s", "cores": 1, "kind": "port",
                  "sample": f"{n_bf} queries, every facility tested for every user (rknn-naive semantics)",
                  "matches_gpu": bool(ok_bf)}
        log(f"cpu brute force (1 thread): {n_bf / bf_s:.3f} queries/s; equal: {ok_bf}")
        if not ok_bf:
            sys.exit(1)
        tree.close()

    cpu_reference = None if args.no_cpu else reference_programs(R, args.ref_queries)

    # ---- serving loop under Poisson arrivals (native load generator)
    serve = serve_datasets = None
    sb = os.path.join(ROOT, "rt-rknn_amd", "bin", "rknn-serve-bench")
    if not args.no_serve and os.path.exists(sb):
        batch.close()  # free the step's buffers for the server's own
        cmd = [sb, "-c", str(args.config), "-u", str(args.users), "-f", str(args.facilities), "-k", str(args.k),
               "--fixed-k", "-s", str(args.serve_seconds), "--qps", args.serve_qps]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            log("serve bench failed:", r.stderr[-2000:])
            sys.exit(1)
        serve = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
        for pt in serve:
            log(f"serve k={args.k}: offered {pt['offered_qps']:,.0f}/s -> {pt['achieved_qps']:,.0f}/s, "
                f"p50 {pt['p50_ms']:.3f} ms, p99 {pt['p99_ms']:.3f} ms")
        batch = None
        # several datasets resident at once, requests (dataset, q, k ~ U{1..50})
        cmd = [sb, "-c", str(args.config), "-k", "50", "-s", str(args.serve_seconds), "--qps",
               args.serve_datasets_qps, "--datasets", args.serve_datasets]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            log("serve bench (datasets) failed:", r.stderr[-2000:])
            sys.exit(1)
        pts = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
        serve_datasets = {"datasets": args.serve_datasets, "k": "U{1..50}", "points": pts}
        for pt in pts:
            log(f"serve {args.serve_datasets}: offered {pt['offered_qps']:,.0f}/s -> "
                f"{pt['achieved_qps']:,.0f}/s, p50 {pt['p50_ms']:.3f} ms, p99 {pt['p99_ms']:.3f} ms")

    out = {
        "metric": "RkNN queries/s @10M users,k=10 batched vs plain; p99 latency; bit-exact sets",
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 generator, config %d)" % args.config,
        "config": {"workload": f"batched RkNN, {args.users} users, {args.facilities} facilities, "
                               f"{args.queries} queries per step, k={args.k}, results in host memory",
                   "users": args.users, "facilities": args.facilities, "queries_per_step": args.queries,
                   "k": args.k, "candidates_mean": round(float(ncs.mean()), 2), "generator_config": args.config,
                   "distinct_batches": nb, "batches_in_flight": min(n_inflight, nb)},
        "value_fetch": round(value_fetch, 1),
        "plain_qps": round(plain_qps, 1),
        "batched_vs_plain": round(value / plain_qps, 2),
        "batched_vs_plain_basis": "both with their sorted result ids in host memory: each step's last kernel "
                                  "writes them to pinned memory; each plain call reads its ids back",
        "plain_latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4),
                             "p99": round(float(np.percentile(lat_ms, 99)), 4)},
        "batched_device_ms": round(float(np.mean(dev_sync)) * 1e3, 4),
        "batched_sync_ms_per_step": round(sync_ms, 4),
        "filter_qps": round(nb * nq / filter_s, 1),  # the host filter, one query per call
        "roofline": roofline,
        "count_all_call_ms": call_ms,
        "cpu_baseline": cpu,
        "cpu_baseline_mt": cpu_mt,
        "cpu_bruteforce": cpu_bf,
        "cpu_reference": cpu_reference,
        "mixed_k": mixed,
        "device_filtered_step": dfilter,
        "end_to_end": end_to_end,
        "serve": serve,
        "serve_datasets": serve_datasets,
    }
    print(json.dumps(out), flush=True)
    if batch is not None:
        batch.close()
    ds.close()
    ctx.close()


def host_cpu():
    """Core count and model of this host (the CPU baselines run here)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "cpu_model": model}


def _ref_times(binary, path, q, k):
    """One run of a reference CPU program: its printed phase times and result count."""
    out = subprocess.run([binary, "-if", path, "-k", str(k), "-q", str(q)], capture_output=True, text=True,
                         timeout=600).stdout
    res = {}
    for line in out.splitlines():
        for key, tag in (("tree_build_s", "R*-tree is built in "), ("filter_s", "Filtering time: "),
                         ("verify_s", "Verification time: "), ("total_s", "Total time: "),
                         ("search_s", "RkNN search completed in ")):
            if line.startswith(tag):
                res[key] = float(line[len(tag):].split("[")[0])
        if line.startswith("Found "):
            res["found"] = int(line.split()[1])
    return res


def reference_programs(R, n):
    """The reference's own CPU programs, built from its sources into
    oracle/_ref, on this host (one process per query, one thread each, as
    they are written): rknn-inf on the rknn-inf configuration (1M users, 10k
    facilities, k=10; its R*-tree build reported apart, as it prints it,
    rknn-inf/inf.cpp:179,584-601) and rknn-naive on the 10k/1k configuration
    with integer coordinates (it parses integers, rknn-naive/rknn.cpp:306-328;
    its search time, :347-352).  The engine's counts on the same files beside
    them."""
    import tempfile

    import oracle

    out = {"host": host_cpu(), "threads_per_program": 1}
    inf, naive = oracle.ref_binary("rknn-inf"), oracle.ref_binary("rknn-naive")
    if not inf and not naive:
        out["absent"] = "oracle/_ref holds no reference programs on this host (they are built where the " \
                        "reference sources are, by oracle/Makefile ref)"
        return out
    ctx = R.Context(0)
    try:
        with tempfile.TemporaryDirectory() as td:
            if inf:
                seed = R.default_seed(2)
                arrs = R.generate(2, 1_000_000, 10_000, seed=seed)
                path = os.path.join(td, "config2.txt")
                R.write_text(path, *arrs)
                qs = [int(q) for q in R.generate_queries(10_000, n, seed)]
                runs = [_ref_times(inf, path, q, 10) for q in qs]
                ds = R.Dataset(ctx, *arrs)
                ours = [len(ds.query(q, 10)) for q in qs]
                ds.close()
                tot = float(np.mean([r["total_s"] for r in runs]))
                out["rknn_inf"] = {"workload": "1,000,000 users, 10,000 facilities, uniform, k=10", "queries": len(qs),
                                   "tree_build_s": round(float(np.mean([r["tree_build_s"] for r in runs])), 3),
                                   "filter_s": round(float(np.mean([r["filter_s"] for r in runs])), 6),
                                   "verify_s": round(float(np.mean([r["verify_s"] for r in runs])), 6),
                                   "total_s": round(tot, 6), "qps": round(1.0 / tot, 1),
                                   "found": [r.get("found") for r in runs], "engine_found": ours}
                log(f"reference rknn-inf (1 thread): {1.0 / tot:,.1f} queries/s over {len(qs)} queries "
                    f"(filter + verification; R*-tree build {out['rknn_inf']['tree_build_s']}s apart); "
                    f"found {out['rknn_inf']['found']}, engine {ours}")
            if naive:
                seed = R.default_seed(1)
                ux, uy, uid, fx, fy, fid = R.generate(1, 10_000, 1_000, seed=seed)
                ux, uy, fx, fy = (np.round(v).astype(np.float32) for v in (ux, uy, fx, fy))
                path = os.path.join(td, "config1_int.txt")
                R.write_text(path, ux, uy, uid, fx, fy, fid)
                qs = [int(q) for q in R.generate_queries(1_000, n, seed)]
                runs = [_ref_times(naive, path, q, 5) for q in qs]
                ds = R.Dataset(ctx, ux, uy, uid, fx, fy, fid)
                ours = [len(ds.query_bruteforce(q, 5)) for q in qs]
                ds.close()
                sec = float(np.mean([r["search_s"] for r in runs]))
                out["rknn_naive"] = {"workload": "10,000 users, 1,000 facilities, integer coordinates in [0,1000]^2, "
                                                 "k=5", "queries": len(qs), "search_s": round(sec, 6),
                                     "qps": round(1.0 / sec, 3), "found": [r.get("found") for r in runs],
                                     "engine_found": ours}
                log(f"reference rknn-naive (1 thread): {1.0 / sec:,.3f} queries/s over {len(qs)} queries; "
                    f"found {out['rknn_naive']['found']}, engine {ours}")
    finally:
        ctx.close()
    return out


# configs 1-3: (users, facilities, queries, k values)
TABLE = {1: (10_000, 1_000, 10, [5]), 2: (1_000_000, 10_000, 100, [10]), 3: (4_000_000, 50_000, 200, [1, 5, 10, 25])}


def table(R):
    """One JSON line per config 1-3: the plain form (one launch per query, with its host filter), the batched
    form over the config's queries, the CPU tree restatement on one thread, and the reference's own rknn-inf
    binary (oracle/_ref, built from its sources here) on a few queries of configs 1-2, with its printed times."""
    import tempfile

    import oracle

    for cfg, (nu, nf, nq, kset) in TABLE.items():
        seed = R.default_seed(cfg)
        ux, uy, uid, fx, fy, fid = R.generate(cfg, nu, nf, seed=seed, nq_ties=nq, seed_ties=seed)
        qs = [int(q) for q in R.generate_queries(nf, nq, seed)]
        ks = [kset[i % len(kset)] for i in range(nq)]
        ctx = R.Context(0)
        ds = R.Dataset(ctx, ux, uy, uid, fx, fy, fid)
        cands = [ds.filter(q, k) for q, k in zip(qs, ks)]
        for q, k, c in zip(qs[:4], ks, cands):  # warm
            ds.query_plain(q, k, c)
        lat, plain = [], []
        for q, k, c in zip(qs, ks, cands):
            t0 = time.perf_counter()
            plain.append(ds.query_plain(q, k, c))
            lat.append(time.perf_counter() - t0)
        b = ds.prepare_batch([(q, k, c) for q, k, c in zip(qs, ks, cands)])
        for _ in range(3):
            b.run()
        dev = [b.run() for _ in range(10)]
        batched = b.fetch()
        b.close()
        same = all(np.array_equal(a, c) for a, c in zip(plain, batched))
        tree = oracle.Tree(ux, uy, uid)
        oracle.set_threads(1)
        t0 = time.perf_counter()
        cnt = tree.counts_many(fx, fy, qs, ks, cands)
        cpu_s = time.perf_counter() - t0
        same &= all(int(n) == len(r) for n, r in zip(cnt, plain))
        tree.close()
        row = {"config": cfg, "users": nu, "facilities": nf, "queries": nq, "k": sorted(set(ks)),
               "candidates_mean": round(float(np.mean([len(c) for c in cands])), 2),
               "plain_qps": round(nq / sum(lat), 1),
               "plain_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
               "plain_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
               "batched_qps": round(nq / float(np.mean(dev)), 1), "batched_device_ms": round(float(np.mean(dev)) * 1e3, 4),
               "cpu_tree_qps_1thread": round(nq / cpu_s, 1), "sets_equal": bool(same)}
        ref = oracle.ref_binary("rknn-inf")
        if ref and cfg in (1, 2):
            with tempfile.TemporaryDirectory() as td:
                path = os.path.join(td, "data.txt")
                R.write_text(path, ux, uy, uid, fx, fy, fid)
                tot, found, build = [], [], []
                for q, k in list(zip(qs, ks))[:3]:
                    out = subprocess.run([ref, "-if", path, "-k", str(k), "-q", str(q)], capture_output=True,
                                         text=True, timeout=600).stdout
                    for line in out.splitlines():
                        if line.startswith("Total time:"):
                            tot.append(float(line.split()[2].split("[")[0]))
                        elif line.startswith("Found "):
                            found.append(int(line.split()[1]))
                        elif line.startswith("R*-tree is built in"):
                            build.append(float(line.split()[4].split("[")[0]))
                row["ref_rknn_inf"] = {"queries": len(tot), "total_s_mean": round(float(np.mean(tot)), 6) if tot else None,
                                       "tree_build_s": round(float(np.mean(build)), 3) if build else None,
                                       "found": found, "ours": [len(r) for r in plain[: len(found)]]}
        ds.close()
        ctx.close()
        print(json.dumps(row), flush=True)


if __name__ == "__main__":
    main()
