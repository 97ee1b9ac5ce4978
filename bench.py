"""Benchmark: the filtered PK-FK join + aggregation (dim 1.5M x fact 6M) on
one MI355X, through the C ABI, with tables resident in HBM.

A step is one full execution of the plan: the workspace init kernel, the dim
build kernel, the fact probe/aggregate kernel and the group-table scan
kernel, enqueued on one stream.  ``value`` is fact (probe-side) rows per
second of the whole job.  Prints one JSON line.  Other workloads (--workload)
time their own plans the same way; the star join c3 times the plan captured
into one hipGraph.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c0|c2_*|c3|cal]
                  [--rows-scale S] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "probe rows/s + HBM GB/s vs peak; batched vs serial queries/s and p99 at fixed QPS"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per row of the dominant (scan/probe) kernel: every column
# the plan references on that table, read once
WORKLOADS = {
    "c1": {"table": "fact", "kernel": "krnl_fact3", "bytes_per_row": 4 + 8 + 8 + 4,
           "desc": "filtered PK-FK hash join + group-by: dim 1.5M rows (d_cat == '2'), fact 6M rows "
                   "(1-year date window), group by d_key, SUM f_price*(1-f_disc); int32 keys, f64 sums"},
    "c1_refill": {"table": "fact", "kernel": "krnl_fact3", "bytes_per_row": 4 + 8 + 8 + 4,
                  "desc": "c1 with lane refill after the fact date filter (the 14% of rows that pass are "
                          "gathered into full wavefronts before the hash probe)"},
    "c0": {"table": "cfg0", "kernel": "krnl_cfg01", "bytes_per_row": 1 + 1 + 4 + 8 + 8 + 8 + 4,
           "desc": "scan-filter-groupby (Q1 shape): date <= cutoff, group by (g1, g2), "
                   "5 f64 sums + count"},
    "c3": {"table": "sales", "kernel": "krnl_sales5", "bytes_per_row": 4 + 4 + 4 + 4 + 8, "graph": True,
           "desc": "4-way star join + LIKE: sales 60M rows probe item 2M (LIKE '%green%' on 55-byte text, 5.3%), "
                   "cust 200k, nation 25; star join materialised once, grouped by (n_key, c_year) and by i_key "
                   "(SUM qty, SUM price, COUNT); whole plan replayed as one hipGraph"},
    "c3_refill": {"table": "sales", "kernel": "krnl_sales5", "bytes_per_row": 4 + 4 + 4 + 4 + 8, "graph": True,
                  "desc": "c3 with lane refill after the item join (survivors of several loop trips are "
                          "gathered into full wavefronts before the cust/nation probes)"},
    "cal": {"table": "k", "kernel": "krnl_k1", "bytes_per_row": 8 + 8,
            "desc": "copy of one int64 column (8 B read + 8 B written per row): counter calibration"},
}
for _v, _d in (("none", "per-lane match loops"), ("refill", "lane refill after the filter"),
               ("pushdown", "push-down match expansion"), ("both", "lane refill + push-down")):
    WORKLOADS["c2_" + _v] = {
        "table": "probe", "kernel": "krnl_probe2", "bytes_per_row": 4 + 4 + 8,
        "desc": "divergent expanding join (%s): build 1M Zipf(1.0) keys over 100k values, probe 16M, "
                "filter p_f < 10, COUNT + SUM(int64 b_pay * p_pay)" % _d}
# lane refill with the register buffer (one entry per lane, ds_permute /
# ds_bpermute) in place of the LDS queue
for _v in ("c1_refill", "c2_refill", "c2_both", "c3_refill"):
    WORKLOADS[_v + "_reg"] = dict(WORKLOADS[_v], desc=WORKLOADS[_v]["desc"] + "; register refill buffer")
# lane refill across a kernel boundary: a first launch queues the survivors in
# device memory, a second runs the rest of the pipeline over the queue
for _v in ("c1_refill", "c2_refill", "c3_refill"):
    WORKLOADS[_v + "_g"] = dict(WORKLOADS[_v], desc=WORKLOADS[_v]["desc"] + "; refill across two launches")


SERVING = {"desc": "serving: stream of c0-c3 queries (row scale log-uniform 0.1x-10x) over HBM-resident tables "
                   "at 10x scale, admission by workspace bytes, one HIP stream per slot"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c1_refill", choices=sorted(WORKLOADS) + ["c4", "q10", "tpch"])
    ap.add_argument("--sf", type=float, default=1.0, help="q10: TPC-H scale factor of the synthetic data")
    ap.add_argument("--scale", type=float, default=10.0, help="c4: resident table scale")
    ap.add_argument("--slots", type=int, default=16, help="c4: concurrent queries")
    ap.add_argument("--arena-gb", type=float, default=48.0, help="c4: device memory for query workspaces")
    ap.add_argument("--max-heavy", type=int, default=1, help="c4: heavy queries running at once (0 = no limit)")
    ap.add_argument("--max-batch", type=int, default=16,
                    help="c4: waiting queries of one config run as one batched plan, up to this many (<= 1: off)")
    ap.add_argument("--batch-ratio", type=int, default=8,
                    help="c4: queries share a batch only if their driving rows are within this factor (0 = any)")
    ap.add_argument("--batch-min-share", type=float, default=2.0,
                    help="c4: a batch's largest query runs alone when the other members' driving rows sum to "
                         "less than this share of its own (0 = never)")
    ap.add_argument("--heavy-mb", type=float, default=64.0, help="c4: workspace size of a heavy query")
    ap.add_argument("--rows-scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--untuned", action="store_true", help="ignore the autotuner's cached configuration")
    return ap.parse_args()


def cpuBaseline(workload, g, seconds):
    """The C oracle (oracle/cpu_oracle.c) on the full workload: repeated for
    about `seconds` on the host threads (at most 16, one GPU's share of a
    GPU box's cores), then for about `seconds` / 2 on one thread; median
    time -> rows/s."""
    from oracle import coracle
    from dogqc_amd import workloads
    cols = {t: {k: v[1] for k, v in g[t].items()} for t in g}
    if workload.startswith("c1"):
        fn = lambda th: coracle.c1(cols["dim"], cols["fact"], threads=th)
        rows = len(cols["fact"]["f_fk"])
    elif workload.startswith("c2"):
        fn = lambda th: coracle.c2(cols["build"], cols["probe"], threads=th)
        rows = len(cols["probe"]["p_key"])
    elif workload.startswith("c3"):
        fn = lambda th: coracle.c3(cols, workloads.C3_WORD, threads=th)
        rows = len(cols["sales"]["s_item"])
    else:
        fn = lambda th: coracle.c0(cols["cfg0"], threads=th)
        rows = len(cols["cfg0"]["qty"])

    def timed(threads, budget):
        fn(threads)
        times = []
        t_end = time.perf_counter() + budget
        while time.perf_counter() < t_end or len(times) < 3:
            t0 = time.perf_counter()
            fn(threads)
            times.append(time.perf_counter() - t0)
        return statistics.median(times), times

    threads = max(1, min(16, os.cpu_count() or 1))
    med, times = timed(threads, seconds)
    med1, times1 = timed(1, seconds / 2.0)
    return {"value": rows / med, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": "full workload (%d probe rows), %d repetitions in %.1f s, median %.2f ms"
                      % (rows, len(times), sum(times), med * 1e3),
            "single_thread": {"value": rows / med1, "unit": "rows/s", "cores": 1,
                              "sample": "same workload, %d repetitions in %.1f s, median %.2f ms"
                                        % (len(times1), sum(times1), med1 * 1e3)}}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        return benchServing(args)
    if args.workload == "q10":
        return benchQ10(args)
    if args.workload == "tpch":
        return benchTpch(args)
    import numpy as np
    from dogqc_amd import planlib, workloads, runtime as rt, runner
    from dogqc_amd.dbio import MemoryAccessor

    spec = WORKLOADS[args.workload]
    kind = planlib.PLANS[args.workload][0]
    rows = {t: max(1, int(n * args.rows_scale)) for t, n in workloads.FULL_ROWS[kind].items()}
    g = workloads.generate(kind, rows)
    acc = MemoryAccessor(g)
    from dogqc_amd import autotune
    rec = None if args.untuned else autotune.loadCached(args.workload, acc)
    if rec is not None:
        comp = autotune.translateVariant(args.workload, rec["block"], rec["rows_per_thread"])
        mod = rt.PlanModule(comp.compile(autotune.variantName(args.workload, rec["block"], rec["rows_per_thread"])))
        tuning = autotune.tuningFromDict(rec["tuning"])
        tuned = {"block": rec["block"], "rows_per_thread": rec["rows_per_thread"], "tuning": rec["tuning"]}
    else:
        comp = planlib.translate(args.workload)
        mod = rt.PlanModule(comp.compile(args.workload))
        tuning = rt.newTuning()
        tuned = None
    inputs = rt.BoundInputs(mod, acc)

    # first run sizes the tables (retries update `tuning`) and checks parity
    ex = rt.run(mod, inputs, tuning)
    got = runner.resultColumns(mod, ex, acc)
    temps = {t["name"]: runner.tempColumns(mod, ex, acc, t["name"]) for t in mod.desc["temps"]}
    eagerCount = ex.count
    eagerTemps = [int(ex.res.status.temp_count[i]) for i in range(len(mod.desc["temps"]))]
    ex.release()
    parity = checkParity(args.workload, g, got, temps)
    if args.workload == "cal":
        args.no_cpu = True

    stream = rt.Stream()
    lay = rt.layout(mod, inputs, tuning)
    ws = rt.DeviceBuffer(lay.ws_bytes)
    steps = mod.steps()
    for _ in range(args.warmup):
        rt.enqueue(mod, inputs, tuning, ws, stream)
    stream.sync()

    events = [[rt.Event() for _ in range(len(steps) + 1)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        rt.enqueue(mod, inputs, tuning, ws, stream, events[k])
    stream.sync()
    t1 = time.perf_counter()
    eager_ms = (t1 - t0) * 1e3 / args.steps
    per = {name: [] for name in steps}
    for k in range(args.steps):
        for i, name in enumerate(steps):
            per[name].append(events[k][i].elapsed(events[k][i + 1]))
    kernel_ms = {name: statistics.mean(v) for name, v in per.items()}

    # the same plan captured once into a hipGraph and replayed
    graph_ms, gstatus = graphReplay(mod, inputs, tuning, ws, stream, args.steps)
    # every replay redoes the whole plan (workspace clear included): its
    # status must match the eager run's (no overflow, same row counts)
    graph_ok = gstatus.flags == 0 and gstatus.result_count == eagerCount and \
        all(gstatus.temp_count[i] == n for i, n in enumerate(eagerTemps))
    if not graph_ok:
        raise RuntimeError("hipGraph replay disagrees with the eager run (flags %#x, %d rows vs %d)"
                           % (gstatus.flags, gstatus.result_count, eagerCount))
    # both timings run the whole plan per step; the faster is reported
    ms_per_step = min(graph_ms, eager_ms)

    probe_rows = rows[spec["table"]]
    dom = spec["kernel"]
    # a kernel launched in two phases (long-list split) counts both launches
    dom_ms = kernel_ms[dom] if dom in kernel_ms else sum(v for k, v in kernel_ms.items() if k.startswith(dom + "<"))
    alg_bytes = spec["bytes_per_row"] * probe_rows
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    # traffic summaries of other sizes than the workload's own carry the
    # probe rows in their name (c0 at 64M rows: krnl_cfg01_c0_64M_traffic.json)
    tag = args.workload if args.rows_scale == 1.0 else "%s_%dM" % (args.workload, probe_rows // 1000000)
    traffic = loadTraffic(tag, dom)

    out = {
        "metric": METRIC,
        "value": probe_rows / (ms_per_step * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 generator, dogqc_amd/workloads.py)",
        "config": {"workload": spec["desc"], "rows": rows, "seed": workloads.SEEDS[kind],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": ("profiles/%s_%s_traffic.json: rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B and "
                                        "TCC_EA0_WRREQ_WRITE_DRAM_32B of this kernel (L2-to-memory bytes in 32-byte "
                                        "units); read from the committed file, not measured in this run"
                                        % (dom, tag)) if traffic else None,
                     "kernel": dom,
                     "kernel_ms": dom_ms, "algorithmic_bytes": alg_bytes},
        "kernels_ms": kernel_ms,
        "graph_ms_per_step": graph_ms,
        "eager_ms_per_step": eager_ms,
        "timed": ("hipGraph replay (whole plan captured once; status checked against the eager run)"
                  if graph_ms <= eager_ms else "eager enqueue"),
        "parity": parity,
        "autotuned": tuned,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpuBaseline(args.workload, g, args.cpu_seconds)
    print(json.dumps(out))


def benchQ10(args):
    """TPC-H Q10 on synthetic data at --sf (SF1: 6,001,215 lineitem rows),
    the query of the reference's only published measurements (sample
    profiler reports: krnl_lineitem7 15.2 ms divergent, 5.1 ms with lane
    refill at opId 9; RTX 2080, FP32 decimals, grid 100 x 32).  The refill
    runs with the LDS queue and with the register buffer, and once more with
    the group-by's string keys hashed by the whole wavefront.  All plans
    carry the lane-activity profile after the lineitem join (opId 9);
    kernel times are taken with the profile off, the histogram in one more
    run with it on.  Every plan's result is checked against the oracle."""
    import numpy as np
    from dogqc_amd import planlib, runtime as rt, runner, tpch
    from dogqc_amd.dbio import MemoryAccessor
    t0 = time.perf_counter()
    acc = MemoryAccessor(tpch.generate(args.sf, seed=10))
    gen_s = time.perf_counter() - t0
    li = acc.schema["lineitem"]["size"]
    store = {}
    out = {"metric": METRIC, "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic TPC-H (dogqc_amd/tpch.py generate, seed 10; %.1f s on the host)" % gen_s,
           "config": {"workload": "TPC-H Q10, SF %g (%d lineitem rows), lineitem pipeline with and without "
                                  "lane refill at opId 9" % (args.sf, li), "parallelism": "single GPU"},
           "reference_sample": {"krnl_lineitem7_ms": {"divergent": 15.2, "lane_refill_op9": 5.1},
                                "total_ms": {"divergent": 16.2, "lane_refill_op9": 6.2},
                                "hardware": "RTX 2080 (README dev env), FP32 decimals, grid 100 x 32 lanes",
                                "source": "sample/profiling_report_{divergent,balanced}.pdf p.1"},
           "plans": {}}
    results = {}
    stream = rt.Stream()
    for name in ("tpch10_p9", "tpch10_lr9_p9", "tpch10_lr9reg_p9", "tpch10_ws_lr9_p9", "tpch10_lrg9_p9"):
        mod = rt.PlanModule(planlib.translate(name).compile(name))
        inputs = rt.BoundInputs(mod, acc, store=store)
        tun = rt.newTuning()
        ex = rt.run(mod, inputs, tun, stream)
        results[name] = runner.resultColumns(mod, ex, acc)
        lay = ex.res.layout
        ex.release()
        ws = rt.DeviceBuffer(lay.ws_bytes)
        steps = mod.steps()
        for _ in range(args.warmup):
            rt.enqueue(mod, inputs, tun, ws, stream)
        ev = [[rt.Event() for _ in range(len(steps) + 1)] for _ in range(args.steps)]
        stream.sync()
        for k in range(args.steps):
            rt.enqueue(mod, inputs, tun, ws, stream, ev[k])
        stream.sync()
        per = {s: statistics.median(ev[k][i].elapsed(ev[k][i + 1]) for k in range(args.steps))
               for i, s in enumerate(steps)}
        ws.free()
        tun.profile = 1
        exp = rt.run(mod, inputs, tun, stream)
        hist = exp.profile()[9]
        exp.release()
        iters = int(hist.sum())
        mean = float((np.arange(len(hist)) * hist).sum()) / max(1, iters)
        # the lineitem pipeline, both launches when its refill splits it
        lim = sum(v for k, v in per.items() if k == "krnl_lineitem7" or k.startswith("krnl_lineitem7<"))
        out["plans"][name] = {
            "kernels_ms": per, "total_ms": sum(per.values()), "krnl_lineitem7_ms": lim,
            "lineitem_rows_per_s": li / (lim * 1e-3),
            "p9_histogram": {"lanes": [int(i) for i in np.nonzero(hist)[0]],
                             "iterations": [int(hist[i]) for i in np.nonzero(hist)[0]],
                             "wave_iterations": iters, "mean_active_lanes_of_64": mean}}
        print("[q10] %s: lineitem7 %.3f ms, total %.3f ms, mean active lanes %.1f of 64 over %d wave trips"
              % (name, lim, sum(per.values()), mean, iters), file=sys.stderr, flush=True)
    a, b, c = results["tpch10_p9"], results["tpch10_lr9_p9"], results["tpch10_lr9reg_p9"]
    d, e = results["tpch10_ws_lr9_p9"], results["tpch10_lrg9_p9"]
    from oracle.planeval import evaluate
    alg, p, cfg, _ = planlib.build("tpch10", acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    want = evaluate(plan, acc)
    out["parity"] = {"oracle": "planeval", "groups": int(len(want["c_custkey"])),
                     "plain_ok": sameGroups(a, want, ["c_custkey"], ["revenue"]),
                     "refill_ok": sameGroups(b, want, ["c_custkey"], ["revenue"]),
                     "register_refill_ok": sameGroups(c, want, ["c_custkey"], ["revenue"]),
                     "wave_string_hash_ok": sameGroups(d, want, ["c_custkey"], ["revenue"]),
                     "kernel_split_refill_ok": sameGroups(e, want, ["c_custkey"], ["revenue"])}
    best = min(out["plans"].values(), key=lambda r: r["total_ms"])
    out["value"] = li / (best["total_ms"] * 1e-3)
    out["ms_per_step"] = best["total_ms"]
    out["roofline"] = None
    print(json.dumps(out))


def servingCheck():
    """The checker of the served sample: the C oracle over host copies of
    each sampled query's row prefixes, called after the timed runs."""
    from oracle import servecheck
    return servecheck.check


def benchServing(args):
    """c4: queries/s of the serving loop (all queries arriving at once) with
    `slots` concurrent queries against one at a time, and p50/p99 latency of
    both under a Poisson stream at 70% of the one-at-a-time capacity."""
    # one hardware queue per slot, so kernels of concurrent queries are not
    # serialised behind each other in HIP's default four queues
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.slots)))
    from dogqc_amd import serving
    nq = max(8, args.steps) if args.steps != 50 else 200
    log = lambda m: print("[c4] " + m, file=sys.stderr, flush=True)
    r = serving.benchmark(scale=args.scale, nqueries=nq, slots=args.slots, arena_gb=args.arena_gb, log=log,
                          max_heavy=args.max_heavy, heavy_mb=args.heavy_mb, max_batch=args.max_batch,
                          batch_ratio=args.batch_ratio, batch_min_share=args.batch_min_share,
                          check=servingCheck())
    out = {
        "metric": METRIC,
        "value": r["batched"]["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": nq,
        "warmup": 4,
        "ms_per_step": 1e3 / r["batched"]["qps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated on the device, dogqc_amd/datagen.py)",
        "config": {"workload": SERVING["desc"], "resident_gb": r["resident_gb"], "scale": r["scale"],
                   "queries": nq, "slots": r["slots"], "arena_gb": r["arena_gb"], "max_heavy": r["max_heavy"],
                   "heavy_mb": r["heavy_mb"], "max_batch": r["max_batch"], "batch_ratio": r["batch_ratio"],
                   "batch_min_share": r["batch_min_share"], "plans": r["plans"],
                   "parallelism": "single GPU"},
        "roofline": None,
        "serial_qps": r["serial"]["qps"],
        "batched_qps": r["batched"]["qps"],
        "speedup_vs_serial": r["speedup"],
        "offered_qps": r["offered_qps"],
        "latency_at_offered_qps": {"serial": {k: r["serial_at_rate"][k] for k in ("p50_ms", "p99_ms", "mean_ms")},
                                   "batched": {k: r["batched_at_rate"][k] for k in ("p50_ms", "p99_ms", "mean_ms")}},
        "capacity_runs": {"serial": r["serial"], "batched": r["batched"]},
        "batching": {"capacity_run": {k: r["batched"][k] for k in ("in_batches", "mean_batch")},
                     "at_offered_qps": {k: r["batched_at_rate"][k] for k in ("in_batches", "mean_batch")}},
        "by_config": {"serial": r["serial_by_config"], "batched": r["batched_by_config"]},
        "latency_at_offered_qps_by_config": {"serial": r["serial_at_rate_by_config"],
                                             "batched": r["batched_at_rate_by_config"]},
        "batched_at_high_rate": r["batched_at_high_rate"],
        "batched_at_high_rate_by_config": r["batched_at_high_rate_by_config"],
        "serving_target_check": {
            "batched_offered_qps": r["batched_at_high_rate"]["offered_qps"],
            "batched_x_serial_capacity": r["batched_at_high_rate"]["offered_x_serial_capacity"],
            "batched_p99_ms": r["batched_at_high_rate"]["p99_ms"],
            "serial_offered_qps": r["offered_qps"], "serial_p99_ms": r["serial_at_rate"]["p99_ms"],
            "batched_p99_le_serial_p99": r["batched_at_high_rate"]["p99_ms"] <= r["serial_at_rate"]["p99_ms"]},
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        "tuned_buckets": r.get("tuned_buckets"),
        "parity": r["verify"],
    }
    print(json.dumps(out))


def graphReplay(mod, inputs, tuning, ws, stream, steps):
    """Capture one enqueue of the plan into a hipGraph and time `steps`
    replays; returns (ms per replay, the status block after the last)."""
    import ctypes as C
    import numpy as np
    from dogqc_amd import runtime as rt
    L = rt.lib()
    gx = C.c_void_p()
    rt.check(L.dqc_graph_create(mod.handle, C.byref(inputs.struct), C.byref(tuning), ws.ptr, ws.nbytes,
                                stream.handle, C.byref(gx)), "dqc_graph_create")
    for _ in range(3):
        rt.check(L.dqc_graph_launch(gx, stream.handle), "dqc_graph_launch")
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        rt.check(L.dqc_graph_launch(gx, stream.handle), "dqc_graph_launch")
    stream.sync()
    t1 = time.perf_counter()
    rt.check(L.dqc_graph_destroy(gx), "dqc_graph_destroy")
    # the status block heads the workspace: the last replay's flags and counts
    raw = ws.toArray(np.uint8, C.sizeof(rt.Status))
    status = rt.Status.from_buffer_copy(raw.tobytes())
    return (t1 - t0) * 1e3 / steps, status


def benchTpch(args):
    """TPC-H Q1-Q22 on synthetic data at --sf, each plan as `dogqc_amd.tpch`
    writes it, with the autotuner's record of its size bucket when there is
    a current one (--untuned: never): per query the median time of --steps
    enqueues (after --warmup), its kernels, retries of the first run, and
    its result against the numpy plan oracle (whose time is the CPU
    column)."""
    from dogqc_amd import autotune, planlib, runtime as rt, runner, tpch
    from dogqc_amd.dbio import MemoryAccessor
    from oracle.planeval import evaluate
    t0 = time.perf_counter()
    acc = MemoryAccessor(tpch.generate(args.sf, seed=10))
    gen_s = time.perf_counter() - t0
    store = {}
    stream = rt.Stream()
    queries = {}
    li = acc.schema["lineitem"]["size"]
    for q in range(1, 23):
        name = "tpch%d" % q
        # the record of the plan's own size bucket: its largest input table
        plain = planlib.translate(name)
        driving = max(acc.schema[t]["size"] for t in plain.program.tables)
        rec = None if args.untuned else autotune.nearest(autotune.records(name), driving)
        if rec is not None:
            mod = rt.PlanModule(autotune.variantPath(name, rec["block"], rec["rows_per_thread"]))
            tun = autotune.tuningFromDict(rec["tuning"])
        else:
            mod = rt.PlanModule(plain.compile(name))
            tun = rt.newTuning()
        inputs = rt.BoundInputs(mod, acc, store=store)
        ex = rt.run(mod, inputs, tun, stream)       # retries grow tun's capacities in place
        retries = int(ex.res.retries)
        got = runner.resultColumns(mod, ex, acc)
        lay = ex.res.layout
        ex.release()
        ws = rt.DeviceBuffer(lay.ws_bytes)
        steps = mod.steps()
        for _ in range(args.warmup):
            rt.enqueue(mod, inputs, tun, ws, stream)
        ev = [[rt.Event() for _ in range(len(steps) + 1)] for _ in range(args.steps)]
        stream.sync()
        for k in range(args.steps):
            rt.enqueue(mod, inputs, tun, ws, stream, ev[k])
        stream.sync()
        totals = [ev[k][0].elapsed(ev[k][len(steps)]) for k in range(args.steps)]
        per = {st: statistics.median(ev[k][i].elapsed(ev[k][i + 1]) for k in range(args.steps))
               for i, st in enumerate(steps)}
        ws.free()
        alg, p, cfg, _ = planlib.build(name, acc)
        plan = alg.resolveAlgebraPlan(p, cfg)
        c0 = time.perf_counter()
        want = evaluate(plan, acc)
        cpu_s = time.perf_counter() - c0
        rows = len(next(iter(want.values()))) if want else 0
        queries[name] = {"ms": statistics.median(totals), "kernels": len(steps), "retries": retries,
                         "autotuned": None if rec is None else {"block": rec["block"],
                                                                "rows_per_thread": rec["rows_per_thread"]},
                         "result_rows": rows, "oracle_s": cpu_s, "parity_ok": sameResult(got, want),
                         "slowest_kernel": max(per, key=per.get), "slowest_kernel_ms": max(per.values())}
        print("[tpch] Q%d: %.3f ms (%d kernels, %d retries), %d rows, oracle %.2f s, parity %s"
              % (q, queries[name]["ms"], len(steps), retries, rows, cpu_s, queries[name]["parity_ok"]),
              file=sys.stderr, flush=True)
    total_ms = sum(r["ms"] for r in queries.values())
    out = {"metric": METRIC, "value": 22 / (total_ms * 1e-3), "unit": "queries/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": total_ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic TPC-H (dogqc_amd/tpch.py generate, seed 10; %.1f s on the host)" % gen_s,
           "config": {"workload": "TPC-H Q1-Q22 one after another, SF %g (%d lineitem rows), %d of them with "
                                  "autotuner records" % (args.sf, li, sum(1 for r in queries.values() if r["autotuned"])),
                      "parallelism": "single GPU"},
           "queries": queries,
           "parity": {"oracle": "planeval", "ok": all(r["parity_ok"] for r in queries.values())},
           "cpu_baseline": {"value": 22 / sum(r["oracle_s"] for r in queries.values()), "unit": "queries/s",
                            "cores": 1, "kind": "port",
                            "sample": "the numpy plan oracle on the same 22 queries and data, once each"}}
    print(json.dumps(out))


def sameResult(got, want):
    """Multiset equality of two results: rows ordered by their exact
    (non-float) columns, then by their float columns; exact columns equal,
    float columns within 1e-9 relative."""
    import numpy as np
    if set(got) != set(want):
        return False
    cols = list(want)
    n = len(want[cols[0]]) if cols else 0
    if any(len(got[c]) != n for c in cols) or any(len(want[c]) != n for c in cols):
        return False
    isf = {c: np.asarray(want[c]).dtype.kind == "f" for c in cols}
    exact = [c for c in cols if not isf[c]]
    floats = [c for c in cols if isf[c]]

    def order(res):
        keys = [[v if not isinstance(v, bytes) else v.decode("latin-1") for v in np.asarray(res[c]).tolist()]
                for c in exact + floats]
        return sorted(range(n), key=lambda i: tuple(k[i] for k in keys))
    og, ow = order(got), order(want)
    for c in cols:
        a, b = np.asarray(got[c])[og], np.asarray(want[c])[ow]
        if isf[c]:
            if not np.allclose(a.astype(np.float64), b.astype(np.float64), rtol=1e-9, atol=1e-9):
                return False
        elif list(a.tolist()) != list(b.tolist()):
            return False
    return True


def sameGroups(got, want, keyCols, tolCols):
    """Multiset equality of grouped results: keys and other exact columns
    bit-exact, tolCols within 1e-9 relative."""
    import numpy as np
    if set(got) != set(want) or len(got[keyCols[0]]) != len(want[keyCols[0]]):
        return False
    og = np.lexsort([np.asarray(got[k]) for k in reversed(keyCols)])
    ow = np.lexsort([np.asarray(want[k]) for k in reversed(keyCols)])
    for k in want:
        a, b = np.asarray(got[k])[og], np.asarray(want[k])[ow]
        if k in tolCols:
            if not np.allclose(a, b, rtol=1e-9, atol=0):
                return False
        elif not np.array_equal(a, b.astype(a.dtype)):
            return False
    return True


def checkParity(workload, g, got, temps):
    """Bit-exact keys/counts, f64 sums within 1e-9 relative, against the C oracle."""
    import numpy as np
    from oracle import coracle
    from dogqc_amd import workloads
    cols = {t: {k: v[1] for k, v in g[t].items()} for t in g}
    if workload.startswith("c3"):
        byItem, byNy = coracle.c3(cols, workloads.C3_WORD)
        ok = sameGroups(got, byItem, ["i_key"], {"revenue"}) and \
            sameGroups(temps["c3_by_nation_year"], byNy, ["n_key", "c_year"], {"revenue"})
        return {"oracle": "C", "groups": [int(len(byItem["i_key"])), int(len(byNy["n_key"]))],
                "joined_rows": int(byItem["cnt"].sum()), "ok": bool(ok)}
    if workload.startswith("c2"):
        cnt, s = coracle.c2(cols["build"], cols["probe"])
        ok = int(got["cnt"][0]) == cnt and int(got["sum_prod"][0]) == s
        return {"oracle": "C", "matches": int(cnt), "ok": bool(ok)}
    if workload == "cal":
        ok = np.array_equal(np.sort(got["k_v"]), cols["k"]["k_v"])
        return {"oracle": "identity", "rows": int(len(got["k_v"])), "ok": bool(ok)}
    if workload.startswith("c1"):
        keys, sums = coracle.c1(cols["dim"], cols["fact"])
        o = np.argsort(keys)
        gk = np.asarray(got["d_key"])
        go = np.argsort(gk)
        ok = np.array_equal(gk[go], keys[o]) and np.allclose(got["revenue"][go], sums[o], rtol=1e-9, atol=0)
        return {"oracle": "C", "groups": int(len(keys)), "ok": bool(ok)}
    ref = coracle.c0(cols["cfg0"])
    ok = len(ref) == len(got["g1"])
    for i in range(len(got["g1"])):
        r = ref.get((int(got["g1"][i]), int(got["g2"][i])))
        ok = ok and r is not None and int(got["count_order"][i]) == r[5] and \
            np.allclose([got["sum_qty"][i], got["sum_base_price"][i], got["sum_disc_price"][i],
                         got["sum_charge"][i]], r[:4], rtol=1e-9, atol=0)
    return {"oracle": "C", "groups": len(ref), "ok": bool(ok)}


def loadTraffic(workload, kernel):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 PMC summary (profiles/<kernel>_<workload>_traffic.json)."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


if __name__ == "__main__":
    main()
