"""Headline benchmark: RkNN queries/s at 10M users, 100k facilities, k=10,
batched (one graph-captured launch sequence per step) versus plain (one
counting launch, sync and readback per query), bit-exact sets.

A step = one pass of the batched form over one batch of NQ queries whose
candidate lists (from the host influence-zone filter) are already resident
in HBM; the users are Hilbert-sorted and resident.  `value` counts queries/s
over exactly --steps timed steps, two batch objects of the same queries
alternating so that one is enqueued while the other runs
(batched_sync_ms_per_step: one run at a time).  The JSON line also carries:
  roofline      dense counting kernel (count-all: every user x every candidate,
                4 fp32 flops per test) timed with HIP events on its stream
  cpu_baseline  the reference's tree-culled verification (rknn-inf) restated
                in C, exact in binary64, on the step's queries and candidate
                lists, one host thread (cpu_baseline_mt: every host core;
                cpu_bruteforce: rknn-naive semantics on two queries)
  plain_*       the one-launch-per-query form on the same queries
  mixed_k       the same step with k drawn uniformly from 1..50 per query
  device_filtered_step  the same step with the influence-zone filter inside it
                (queries in, results in HBM)
  end_to_end    one rk_query_batch call: filter (device, or host cores) + step +
                download, on the library's own clock (Python wrapper time beside)
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

# HBM bytes per count-all launch on the 10M-user, 100k-facility k=10 workload,
# from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by a copy kernel of known
# size (profiles/k_count_all_10M_100k.md); the kernel reads every user once and
# writes one 16-bit count per user, so the figure does not depend on the query
MEASURED_TRAFFIC = {"valu4": 101.4e6, "valu6": 101.1e6, "valu8": 100.8e6}


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
    ap.add_argument("--roofline-queries", type=int, default=8)
    ap.add_argument("--cpu-queries", type=int, default=2, help="queries of the brute-force CPU context run")
    ap.add_argument("--inflight", type=int, default=2,
                    help="batch objects of the same queries alternating in the timed step, each on its own stream")
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

    # host influence-zone filter (outside the timed region; reported)
    t0 = time.time()
    cands = [ds.filter(q, args.k) for q in qs]
    filter_s = time.time() - t0
    ncs = np.array([len(c) for c in cands])
    log(f"filter: {args.queries} queries in {filter_s:.2f}s, |C| mean {ncs.mean():.1f} "
        f"min {ncs.min()} max {ncs.max()}")

    # ---- batched form: prepared once, replayed per step.  Batch objects of
    # the same queries (--inflight, 2 by default) alternate, each enqueued
    # while the others run (rk_batch_enqueue / rk_batch_wait).  Each batch
    # runs on a stream of its own, so the runs overlap on the device: the next
    # run's query blocks take the SIMDs that the previous run's finished
    # blocks free.  Every
    # step is a full pass over the 1,024 queries, its counts checked on the
    # host.  The synchronous form (one run at a time, rk_batch_run) is timed
    # beside it as the per-step latency, and its device time (HIP events
    # around the graph) is the step's device time alone.
    n_inflight = max(1, args.inflight)
    batches = [ds.prepare_batch([(q, args.k, c) for q, c in zip(qs, cands)]) for _ in range(n_inflight)]
    batch = batches[0]
    for _ in range(args.warmup):
        batch.run()
    t0 = time.perf_counter()
    dev_sync = [batch.run() for _ in range(args.steps)]
    sync_ms = (time.perf_counter() - t0) / args.steps * 1e3
    for _ in range(args.warmup):
        for b in batches:
            b.enqueue()
        for b in batches:
            b.wait()
    dev, pending = [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        b = batches[s % n_inflight]
        if len(pending) == n_inflight:
            dev.append(pending.pop(0).wait())
        b.enqueue()
        pending.append(b)
    while pending:
        dev.append(pending.pop(0).wait())
    wall = time.perf_counter() - t0
    batched = batch.fetch()
    for b in batches[1:]:
        if not all(np.array_equal(a, c) for a, c in zip(batched, b.fetch())):
            log("ERROR: batch objects of the same queries differ")
            sys.exit(1)
        b.close()
    items, n_res, n_rech = batch.stats()
    log(f"batched work per step: items per level (leaf first) {items}, results {n_res}, fp64 re-checks {n_rech}; "
        f"autotuned block: {batch.waves_per_query()} waves per query")
    ms_step = wall / args.steps * 1e3
    value = args.queries * args.steps / wall
    log(f"batched: {ms_step:.3f} ms/step (one run at a time {sync_ms:.3f} ms, device {np.mean(dev_sync) * 1e3:.3f} ms; "
        f"{np.mean(dev) * 1e3:.3f} ms from start to end of a run overlapping another), "
        f"{value:,.0f} queries/s")

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
    for filt in ("device", "host"):
        os.environ["RK_FILTER"] = filt
        ds.query_batch([(q, args.k) for q in qs])  # warm: the dataset's batch buffers grow once
        calls, e2e_ok = [], True
        for _ in range(3):
            t0 = time.perf_counter()
            e2e_res, tm = ds.query_batch([(q, args.k) for q in qs], timing=True)
            calls.append((tm.total_s, time.perf_counter() - t0, tm.filter_s, tm.kernel_s, tm.download_s))
            e2e_ok = e2e_ok and all(np.array_equal(a, b) for a, b in zip(e2e_res, batched))
        lib_s, py_s, f_s, k_s, d_s = (float(np.median([c[j] for c in calls])) for j in range(5))
        end_to_end[f"{filt}_filter"] = {"qps": round(args.queries / lib_s, 1), "ms": round(lib_s * 1e3, 3),
                                        "filter_ms": round(f_s * 1e3, 3), "device_ms": round(k_s * 1e3, 3),
                                        "download_ms": round(d_s * 1e3, 3),
                                        "python_wall_ms": round(py_s * 1e3, 3), "matches": bool(e2e_ok)}
        log(f"end to end, {filt} filter (filter + step + download): {args.queries / lib_s:,.0f} queries/s "
            f"(library clock, median of 3; {py_s * 1e3:.2f} ms through the Python wrapper); same sets: {e2e_ok}")
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
    log(f"plain: {plain_qps:,.0f} queries/s, p50 {np.percentile(lat_ms, 50):.3f} ms, "
        f"p99 {np.percentile(lat_ms, 99):.3f} ms; batched == plain on {n_plain - mismatches}/{n_plain}")
    if mismatches:
        log("ERROR: batched and plain result sets differ")
        sys.exit(1)

    # ---- roofline: dense counting kernel (count-all), HIP events on its stream;
    # every variant (4, 6 or 8 users per lane; 3 VALU
    # instructions per clean test)
    variants = {}
    kernel_names = {"valu4": "k_count_all<4,1>", "valu6": "k_count_all<6,1>", "valu8": "k_count_all<8,1>"}
    kts = {kern: [] for kern in kernel_names}
    flops = {kern: 0.0 for kern in kernel_names}
    for _ in range(3):  # the variants interleaved, three rounds each
        for kern in kernel_names:
            os.environ["RK_COUNT_KERNEL"] = kern
            for i in range(min(args.roofline_queries, args.queries)):
                _, t = ds.count_all(qs[i], cands[i], timing=True)
                kts[kern].append(t.kernel_s)
                flops[kern] += 4.0 * args.users * t.n_candidates
    del os.environ["RK_COUNT_KERNEL"]
    for kern, kt in kts.items():
        variants[kern] = (flops[kern] / sum(kt) / 1e12, float(np.mean(kt)), flops[kern] / len(kt))
        log(f"count-all kernel ({kern}): {variants[kern][0]:.2f} TFLOP/s over {len(kt)} launches, "
            f"avg {variants[kern][1] * 1e3:.4f} ms")
    best = max(variants, key=lambda v: variants[v][0])
    achieved, avg_s, flops_launch = variants[best]
    roofline = {"bound": "valu", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": (MEASURED_TRAFFIC.get(best) if (args.users, args.facilities, args.k, args.config)
                            == (10_000_000, 100_000, 10, 4) else None),
                "traffic_source": "profiles/k_count_all_10M_100k.md (calibrated PMC, bytes per launch)",
                "kernel": kernel_names[best],
                "avg_launch_ms": round(avg_s * 1e3, 4), "flops_per_launch": flops_launch,
                "variants_tflops": {k: round(v[0], 3) for k, v in variants.items()}}

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
                               f"{args.queries} queries per step, k={args.k}",
                   "users": args.users, "facilities": args.facilities, "queries_per_step": args.queries,
                   "k": args.k, "candidates_mean": round(float(ncs.mean()), 2), "generator_config": args.config,
                   "batches_in_flight": n_inflight},
        "plain_qps": round(plain_qps, 1),
        "batched_vs_plain": round(value / plain_qps, 2),
        "plain_latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4),
                             "p99": round(float(np.percentile(lat_ms, 99)), 4)},
        "batched_device_ms": round(float(np.mean(dev_sync)) * 1e3, 4),
        "batched_overlapped_run_ms": round(float(np.mean(dev)) * 1e3, 4),
        "batched_sync_ms_per_step": round(sync_ms, 4),
        "filter_qps": round(args.queries / filter_s, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_mt": cpu_mt,
        "cpu_bruteforce": cpu_bf,
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
