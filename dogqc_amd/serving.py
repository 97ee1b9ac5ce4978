"""Serving workload: a stream of queries over HBM-resident tables, run
through the native serving loop of libdqcrt.so (dqc_serve_*).

Each query is one of the four benchmark plans (c0 scan-aggregate, c1 PK-FK
join, c2 expanding join, c3 star join) over a row prefix of the resident
tables: a row scale s, log-uniform in [0.1, 10], selects the first
s x (base rows) rows of every scaled table, the row-range predicate of the
workload.  The loop admits arrived queries while a slot (HIP stream) is
free and the query's workspace fits the arena, so queries of uneven sizes
overlap on the device; with one slot it is the one-query-at-a-time
baseline on the same code path.

``benchmark()`` measures both: capacity (all queries arrive at once,
queries per second over the makespan) and latency under a Poisson stream at
a fixed rate (p50/p99 of arrival-to-completion), and checks a sample of the
results between the alone and mixed runs (and, through a caller-supplied
check, against an independent result).
"""
import ctypes as C
import json
import math
import statistics
import time
from collections import OrderedDict

import numpy as np

from dogqc_amd import planlib, runtime as rt, runner, workloads

CONFIG_PLANS = OrderedDict([("c0", "c0"), ("c1", "c1_refill"), ("c2", "c2_pushdown"), ("c3", "c3_refill")])
SCALE_RANGE = (0.1, 10.0)


class Server(object):
    """ctypes wrapper of one dqc_serve server."""

    def __init__(self, slots, arena_bytes, max_retries=8, bypass_ms=50, size_feedback=True, max_heavy=0,
                 heavy_bytes=0, max_batch=0, batch_ratio=0, batch_min_share=0.0):
        cfg = rt.ServeConfig()
        cfg.slots, cfg.max_retries, cfg.arena_bytes, cfg.bypass_ms = slots, max_retries, int(arena_bytes), bypass_ms
        cfg.size_feedback = 1 if size_feedback else 0
        cfg.max_heavy, cfg.heavy_bytes = int(max_heavy), int(heavy_bytes)
        cfg.max_batch, cfg.batch_ratio = int(max_batch), int(batch_ratio)
        cfg.batch_min_share = float(batch_min_share)
        self.handle = C.c_void_p()
        self._alive = []      # host arrays the server reads after submit
        rt.check(rt.lib().dqc_serve_create(C.byref(cfg), C.byref(self.handle)), "dqc_serve_create")

    def submit(self, module, rows, cols, tuning=None, arrival_ns=0, keep=False, batch=None):
        """Queue a query; `batch` = (batched module, its BatchBinding) lets
        the server run it together with other queries of the same plan."""
        req = rt.ServeRequest()
        req.module = module.handle
        req.inputs = rt.Inputs(len(rows), len(cols), rows.ctypes.data_as(C.POINTER(C.c_int64)),
                               C.cast(cols, C.POINTER(C.c_void_p)))
        req.tuning = C.pointer(tuning) if tuning is not None else None
        req.arrival_ns = int(arrival_ns)
        req.keep_result = 1 if keep else 0
        if batch is not None:
            bmod, bb = batch
            req.batch_module = bmod.handle
            req.batch_ncols, req.batch_ntables = len(bb.prefix), len(bb.tableMap)
            req.batch_cols = C.cast(bb.cols, C.POINTER(C.c_void_p))
            req.batch_col_prefix = C.cast(bb.prefix, C.POINTER(C.c_int32))
            req.batch_table_map = C.cast(bb.tableMap, C.POINTER(C.c_int32))
            self._alive.append(bb)
        qid = C.c_int64()
        rt.check(rt.lib().dqc_serve_submit(self.handle, C.byref(req), C.byref(qid)), "dqc_serve_submit")
        return qid.value

    def start(self):
        rt.check(rt.lib().dqc_serve_start(self.handle), "dqc_serve_start")

    def wait(self, timeout_ms=-1):
        rt.check(rt.lib().dqc_serve_wait(self.handle, int(timeout_ms)), "dqc_serve_wait")

    def record(self, qid):
        rec = rt.ServeRecord()
        rt.check(rt.lib().dqc_serve_record_of(self.handle, qid, C.byref(rec)), "dqc_serve_record_of")
        return rec

    def release(self, qid):
        rt.check(rt.lib().dqc_serve_release(self.handle, qid), "dqc_serve_release")

    def destroy(self):
        if self.handle:
            rt.check(rt.lib().dqc_serve_destroy(self.handle), "dqc_serve_destroy")
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PlanSet(object):
    """The plan modules and runtime tunings of every config.  With
    `tuned`, each current autotuner record of a config's plan (one per input
    size bucket) gives a module variant and tuning, and a query takes the
    record of the bucket nearest its driving rows (``pick``); queries with no
    record within two buckets run the untuned plan.  ``modules`` /
    ``tunings`` hold the base-size choice, used where one module per config
    is needed (result decoding: every variant has the same result layout)."""

    def __init__(self, configs=tuple(CONFIG_PLANS), tuned=True):
        from dogqc_amd import autotune
        self.modules, self.tunings, self.names, self.batchModules = {}, {}, {}, {}
        self.variants = {}
        for cfg in configs:
            name = CONFIG_PLANS[cfg]
            plain = (rt.PlanModule(planlib.translate(name).compile(name)), rt.newTuning())
            byBucket = {}
            if tuned:
                for bucket, rec in autotune.records(name).items():
                    path = autotune.variantPath(name, rec["block"], rec["rows_per_thread"])
                    byBucket[bucket] = (rt.PlanModule(path), autotune.tuningFromDict(rec["tuning"]), rec)
            self.variants[cfg] = (plain, byBucket)
            base = max(workloads.FULL_ROWS[cfg].values())
            self.modules[cfg], self.tunings[cfg] = self.pick(cfg, base)
            self.names[cfg] = name
            bname = "%s_batch" % cfg
            self.batchModules[cfg] = rt.PlanModule(planlib.translate(bname).compile(bname))

    def pick(self, cfg, drivingRows):
        """(module, tuning) for a query of `cfg` whose largest table
        prefix has `drivingRows` rows."""
        from dogqc_amd import autotune
        plain, byBucket = self.variants[cfg]
        got = autotune.nearest({b: v for b, v in byBucket.items()}, drivingRows)
        return (got[0], got[1]) if got is not None else plain

    def tunedBuckets(self, cfg):
        return sorted(self.variants[cfg][1])


class Query(object):
    def __init__(self, config, scale, arrival_ns=0):
        self.config, self.scale, self.arrival_ns = config, scale, arrival_ns


def prefixRows(tables, config, scale):
    """Rows of every table of a config a query at `scale` reads: the first
    scale x base rows (the 25 nations are always read whole)."""
    rows = OrderedDict()
    for t, n in workloads.FULL_ROWS[config].items():
        if config == "c3" and t == "nation":
            rows[t] = tables.rows[t]
        else:
            rows[t] = max(1, min(tables.rows[t], int(round(n * scale))))
    return rows


def binding(module, tables, rows):
    """(table_rows, column pointers) of a module over resident tables."""
    desc = module.desc
    r = np.array([rows[t] for t in desc["tables"]], dtype=np.int64)
    cols = (C.c_void_p * max(1, len(desc["columns"])))()
    for c in desc["columns"]:
        cols[c["index"]] = tables.store[(c["table"], c["attr"], c["part"])].ptr
    return r, cols


class BatchBinding(object):
    """How a config's batched module reads the resident tables
    (dqc_serve_request.batch_*): column pointers, which columns hold the
    members' prefixes (of which table), and which of the plain module's
    tables each batched-module table takes its row count from."""

    def __init__(self, plain, batched, tables):
        pdesc, bdesc = plain.desc, batched.desc
        ncols = max(1, len(bdesc["columns"]))
        self.cols = (C.c_void_p * ncols)()
        self.prefix = (C.c_int32 * ncols)(*([-1] * ncols))
        for c in bdesc["columns"]:
            if c["table"] == "__batch":
                self.prefix[c["index"]] = bdesc["tables"].index(c["attr"])
            else:
                self.cols[c["index"]] = tables.store[(c["table"], c["attr"], c["part"])].ptr
        self.tableMap = (C.c_int32 * len(bdesc["tables"]))(
            *[-1 if t == "__batch" else pdesc["tables"].index(t) for t in bdesc["tables"]])


def workspaceBytes(tables, plans, config, scale):
    """Workspace bytes of one query (its admission footprint)."""
    rows = prefixRows(tables, config, scale)
    mod, tun = plans.pick(config, max(rows.values()))
    r, cols = binding(mod, tables, rows)
    inputs = rt.Inputs(len(r), len(cols), r.ctypes.data_as(C.POINTER(C.c_int64)), C.cast(cols, C.POINTER(C.c_void_p)))
    lay = rt.Layout()
    rt.check(rt.lib().dqc_query_layout(mod.handle, C.byref(inputs), C.byref(tun), C.byref(lay)), "dqc_query_layout")
    return int(lay.ws_bytes)


def makeMix(n, seed=4, configs=tuple(CONFIG_PLANS), scaleRange=SCALE_RANGE):
    """n queries: config uniform, row scale log-uniform over scaleRange."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = math.log(scaleRange[0]), math.log(scaleRange[1])
    cfg = rng.integers(0, len(configs), size=n)
    sc = np.exp(rng.uniform(lo, hi, size=n))
    return [Query(configs[c], float(s)) for c, s in zip(cfg, sc)]


def poissonArrivals(queries, qps, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / qps, size=len(queries))
    t = np.cumsum(gaps) - gaps[0]
    for q, a in zip(queries, t):
        q.arrival_ns = int(a * 1e9)


def run(tables, plans, queries, slots, arena_bytes, keep=(), timeout_s=600, log=None, max_heavy=0,
        heavy_bytes=0, max_batch=0, batch_ratio=0, batch_min_share=0.0):
    """Run queries through a server with `slots` slots; returns (records,
    server) -- the server stays alive so kept results can be read.  With
    `max_batch` > 1, waiting queries of one config run in batches of up to
    that many (batched plans, dogqc_amd/batching.py) whose driving row
    counts are within `batch_ratio` of each other (0: any); a batch's largest
    member runs on its own when the others hold less than `batch_min_share`
    of its driving rows (0: never).  With `log`, reports
    every 5 s how many queries have finished."""
    srv = Server(slots, arena_bytes, max_heavy=max_heavy, heavy_bytes=heavy_bytes, max_batch=max_batch,
                 batch_ratio=batch_ratio, batch_min_share=batch_min_share)
    ids = []
    bindings = {}
    for i, q in enumerate(queries):
        rows = prefixRows(tables, q.config, q.scale)
        mod, tun = plans.pick(q.config, max(rows.values()))
        r, cols = binding(mod, tables, rows)
        batch = None
        if max_batch > 1:
            if q.config not in bindings:
                bindings[q.config] = BatchBinding(plans.modules[q.config], plans.batchModules[q.config], tables)
            batch = (plans.batchModules[q.config], bindings[q.config])
        ids.append(srv.submit(mod, r, cols, tun, q.arrival_ns, keep=i in keep, batch=batch))
    srv.start()
    deadline = time.perf_counter() + timeout_s
    while True:
        rc = rt.lib().dqc_serve_wait(srv.handle, 5000)
        if rc == 0:
            break
        if rc != -1 or time.perf_counter() > deadline:    # -1: still running after 5 s
            rt.check(rc, "dqc_serve_wait")
            raise rt.DqcError("serving run did not finish within %d s" % timeout_s)
        if log:
            states = [srv.record(i).state for i in ids]
            running = [i for i, st in enumerate(states) if st == 1]
            log("%d of %d queries finished (%d slots); running: %s" % (
                sum(1 for st in states if st >= 2), len(ids), slots,
                ", ".join("%s x%.2f" % (queries[i].config, queries[i].scale) for i in running[:4])))
    recs = [srv.record(i) for i in ids]
    for i, rec in enumerate(recs):
        if rec.state != 2:
            raise rt.DqcError("query %d (%s x%.2f) failed with code %d" % (i, queries[i].config, queries[i].scale,
                                                                           rec.error))
    return recs, srv


def perConfig(queries, recs):
    """Per config: queries, mean scale, mean and total service time
    (admission to completion, ms) and retries."""
    out = OrderedDict()
    for c in CONFIG_PLANS:
        idx = [i for i, q in enumerate(queries) if q.config == c]
        if not idx:
            continue
        svc = [(recs[i].done_ns - recs[i].admit_ns) * 1e-6 for i in idx]
        out[c] = {"queries": len(idx), "mean_scale": float(np.mean([queries[i].scale for i in idx])),
                  "mean_service_ms": float(np.mean(svc)), "total_service_ms": float(np.sum(svc)),
                  "retries": int(sum(recs[i].retries for i in idx))}
    return out


def slowest(queries, recs, n=6):
    """The n queries of longest latency: config, scale, wait and run time
    (ms), batch size."""
    order = sorted(range(len(recs)), key=lambda i: recs[i].arrival_ns - recs[i].done_ns)[:n]
    return ", ".join("%s x%.2f wait %.1f run %.1f b%d" % (
        queries[i].config, queries[i].scale, (recs[i].admit_ns - recs[i].arrival_ns) * 1e-6,
        (recs[i].done_ns - recs[i].admit_ns) * 1e-6, recs[i].batch_size) for i in order)


def latencyByConfig(queries, recs):
    """p50/p99 latency (arrival to completion, ms) and mean scale per config."""
    out = OrderedDict()
    for c in CONFIG_PLANS:
        idx = [i for i, q in enumerate(queries) if q.config == c]
        if idx:
            lat = np.array([(recs[i].done_ns - recs[i].arrival_ns) * 1e-6 for i in idx])
            out[c] = {"queries": len(idx), "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max())}
    return out


def latencyStats(recs):
    lat = np.array([(r.done_ns - r.arrival_ns) * 1e-6 for r in recs])
    first = min(r.arrival_ns for r in recs)
    last = max(r.done_ns for r in recs)
    return {"queries": len(recs), "makespan_ms": (last - first) * 1e-6,
            "qps": len(recs) / ((last - first) * 1e-9), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "retries": int(sum(r.retries for r in recs))}


def resultOf(plans, config, rec):
    """Result columns (and c3's second aggregation) of a kept record; a
    query that ran in a batch gets its member's combination of the batch's
    segments."""
    res = rt.QueryResult()
    res.ws, res.ws_bytes, res.layout, res.status = rec.ws, rec.ws_bytes, rec.layout, rec.status
    batched = rec.batch_size > 1
    mod = plans.batchModules[config] if batched else plans.modules[config]
    ex = rt.Execution(mod, None, None, res)
    ex.owned = False
    out = runner.resultColumns(mod, ex, None)
    temps = {t["name"]: runner.tempColumns(mod, ex, None, t["name"]) for t in mod.desc["temps"]
             if t["name"] == "c3_by_nation_year"}
    if batched:
        from dogqc_amd import batching
        out, temps = batching.combine(config, out, temps, rec.batch_index)
    if config == "c3":
        out = (out, temps["c3_by_nation_year"])
    return out


def sameResult(config, a, b, rtol=1e-9):
    """Multiset equality of two results of one config: keys and counts
    bit-exact, float64 sums within rtol."""
    if config == "c3":
        return sameResult("c3a", a[0], b[0], rtol) and sameResult("c3b", a[1], b[1], rtol)
    keys = {"c0": ["g1", "g2"], "c1": ["d_key"], "c2": [], "c3a": ["i_key"], "c3b": ["n_key", "c_year"]}[config]
    tol = {"c0": {"sum_base_price", "sum_disc_price", "sum_charge", "avg_price", "avg_disc", "avg_qty", "sum_qty"},
           "c1": {"revenue"}, "c2": set(), "c3a": {"revenue"}, "c3b": {"revenue"}}[config]
    if list(a) != list(b):
        return False
    n = len(next(iter(a.values())))
    if any(len(v) != n for v in list(a.values()) + list(b.values())):
        return False
    oa = np.lexsort([np.asarray(a[k]) for k in reversed(keys)]) if keys else np.arange(n)
    ob = np.lexsort([np.asarray(b[k]) for k in reversed(keys)]) if keys else np.arange(n)
    for k in a:
        x, y = np.asarray(a[k])[oa], np.asarray(b[k])[ob]
        if k in tol:
            if not np.allclose(x, y, rtol=rtol, atol=0):
                return False
        elif not np.array_equal(x, y):
            return False
    return True


def benchmark(scale=10.0, nqueries=200, slots=16, arena_gb=48.0, load=0.7, seed=4, verify=24, log=None,
              max_heavy=1, heavy_mb=64.0, max_batch=16, batch_ratio=8, batch_min_share=2.0, check=None,
              rate_queries=1000, high_load=3.0, high_queries=2000):
    """Serving benchmark on resident tables at `scale` x the base sizes.
    Returns a dict of capacity (serial / batched QPS, all queries waiting at
    once), latency of both under a Poisson stream of `rate_queries` queries
    at `load` x the serial capacity, latency of the batched loop under a
    Poisson stream of `high_queries` at `high_load` x the serial capacity
    (a rate the one-at-a-time loop cannot sustain), and verification
    results.  `check(tables, config, scale, result) -> bool`, when given,
    is called after the timed runs on each result of the verified sample."""
    from dogqc_amd import datagen
    say = log or (lambda *a: None)
    t0 = time.perf_counter()
    tables = datagen.ResidentTables(scale=scale)
    say("resident tables: %.1f GB in %.1f s" % (tables.nbytes() / 1e9, time.perf_counter() - t0))
    plans = PlanSet()
    arena = int(arena_gb * (1 << 30))
    # warm up every module once (code objects load at their first launch):
    # the tuned variant of each size bucket (queries at 0.1x, 1x and 10x),
    # and the batched plans (four equal queries per config wait together, so
    # each config's batched plan runs once)
    run(tables, plans, [Query(c, s) for c in CONFIG_PLANS for s in (0.1, 1.0, 10.0)], 1, arena,
        log=say)[1].destroy()
    hvWarm = {"max_heavy": max_heavy, "heavy_bytes": int(heavy_mb * (1 << 20)), "max_batch": max(max_batch, 4),
              "batch_ratio": batch_ratio, "batch_min_share": batch_min_share}
    warm = [Query(c, 0.1) for c in CONFIG_PLANS for _ in range(4)]
    recsW, srvW = run(tables, plans, warm, slots, arena, log=say, **hvWarm)
    srvW.destroy()
    say("modules loaded and warmed up (%d of %d warm-up queries in batches)"
        % (sum(1 for r in recsW if r.batch_size > 1), len(warm)))
    queries = makeMix(nqueries, seed)
    # the verified sample: the smallest queries of each config, so the CPU
    # check over host copies of their prefixes stays within seconds
    order = sorted(range(nqueries), key=lambda i: queries[i].scale)
    sample = []
    for c in CONFIG_PLANS:
        sample += [i for i in order if queries[i].config == c][:max(1, verify // len(CONFIG_PLANS))]
    out = {"tuned_buckets": {c: plans.tunedBuckets(c) for c in CONFIG_PLANS},
           "scale": scale, "resident_gb": tables.nbytes() / 1e9, "queries": nqueries, "slots": slots,
           "arena_gb": arena_gb, "max_heavy": max_heavy, "heavy_mb": heavy_mb, "max_batch": max_batch,
           "batch_ratio": batch_ratio, "batch_min_share": batch_min_share, "plans": dict(plans.names)}
    # capacity: every query arrives at t = 0
    for q in queries:
        q.arrival_ns = 0
    recsS, srvS = run(tables, plans, queries, 1, arena, keep=set(sample), log=say)
    say("slowest alone: " + ", ".join("%s x%.2f %.1f ms" % (queries[i].config, queries[i].scale,
                                                             (recsS[i].done_ns - recsS[i].admit_ns) * 1e-6)
                                       for i in sorted(range(nqueries), key=lambda i: recsS[i].admit_ns -
                                                       recsS[i].done_ns)[:6]))
    out["serial"] = latencyStats(recsS)
    out["serial_by_config"] = perConfig(queries, recsS)
    say("serial by config: %s" % json.dumps(out["serial_by_config"]))
    for i in sorted(range(nqueries), key=lambda i: -recsS[i].retries)[:5]:
        L = recsS[i].layout
        say("retried %s x%.2f %d times: ht %s mm %s result %d temp %s split %s" % (
            queries[i].config, queries[i].scale, recsS[i].retries, list(L.ht_capacity[:L.nhts]),
            list(L.mm_entries[:L.nhts]), L.result_capacity, list(L.temp_capacity[:L.ntemps]),
            list(L.split_capacity[:L.nhts])))
    say("serial capacity: %.1f q/s" % out["serial"]["qps"])
    hv = {"max_heavy": max_heavy, "heavy_bytes": int(heavy_mb * (1 << 20)), "max_batch": max_batch,
          "batch_ratio": batch_ratio, "batch_min_share": batch_min_share}
    recsB, srvB = run(tables, plans, queries, slots, arena, keep=set(sample), log=say, **hv)
    out["batched"] = latencyStats(recsB)
    out["batched_by_config"] = perConfig(queries, recsB)
    for i in sorted(range(nqueries), key=lambda i: -recsB[i].retries)[:4]:
        if recsB[i].retries:
            L = recsB[i].layout
            say("batched run retried: %s x%.2f (batch of %d) %d times: ht %s temp %s result %d" % (
                queries[i].config, queries[i].scale, recsB[i].batch_size, recsB[i].retries,
                list(L.ht_capacity[:L.nhts]), list(L.temp_capacity[:L.ntemps]), L.result_capacity))
    sizes = [r.batch_size for r in recsB if r.batch_size > 1]
    out["batched"]["in_batches"] = len(sizes)
    out["batched"]["mean_batch"] = float(np.mean(sizes)) if sizes else 1.0
    say("batched capacity: %.1f q/s" % out["batched"]["qps"])
    # alone vs mixed, and the caller's check, on the sample
    alone_same, checked_ok = 0, 0
    for i in sample:
        q = queries[i]
        a = resultOf(plans, q.config, recsS[i])
        b = resultOf(plans, q.config, recsB[i])
        alone_same += sameResult(q.config, a, b)
        if check is not None:
            checked_ok += bool(check(tables, q.config, q.scale, b))
    out["verify"] = {"sample": len(sample), "alone_equals_mixed": alone_same,
                     "checked_ok": checked_ok if check is not None else None,
                     "ok": alone_same == len(sample) and (check is None or checked_ok == len(sample))}
    srvS.destroy()
    srvB.destroy()
    say("verified %d queries: alone == mixed %d, checked %s" % (len(sample), alone_same,
                                                                 checked_ok if check is not None else "-"))
    # latency at a fixed offered rate, on a longer stream of the same mix
    rate = load * out["serial"]["qps"]
    queries = makeMix(rate_queries, seed + 2)
    poissonArrivals(queries, rate, seed + 1)
    recs, srv = run(tables, plans, queries, 1, arena, log=say)
    srv.destroy()
    say("slowest at %.1f q/s, one slot: %s" % (rate, slowest(queries, recs)))
    out["serial_at_rate"] = latencyStats(recs)
    out["serial_at_rate_by_config"] = latencyByConfig(queries, recs)
    recs, srv = run(tables, plans, queries, slots, arena, log=say, **hv)
    srv.destroy()
    say("slowest at %.1f q/s: %s" % (rate, slowest(queries, recs)))
    out["batched_at_rate"] = latencyStats(recs)
    out["batched_at_rate_by_config"] = latencyByConfig(queries, recs)
    sizes = [r.batch_size for r in recs if r.batch_size > 1]
    out["batched_at_rate"]["in_batches"] = len(sizes)
    out["batched_at_rate"]["mean_batch"] = float(np.mean(sizes)) if sizes else 1.0
    out["offered_qps"] = rate
    out["speedup"] = out["batched"]["qps"] / out["serial"]["qps"]
    # the batched loop at a rate the one-at-a-time loop cannot sustain
    high = high_load * out["serial"]["qps"]
    queries = makeMix(high_queries, seed + 4)
    poissonArrivals(queries, high, seed + 3)
    recs, srv = run(tables, plans, queries, slots, arena, log=say, **hv)
    srv.destroy()
    say("slowest at %.1f q/s: %s" % (high, slowest(queries, recs)))
    st = latencyStats(recs)
    sizes = [r.batch_size for r in recs if r.batch_size > 1]
    st["in_batches"] = len(sizes)
    st["mean_batch"] = float(np.mean(sizes)) if sizes else 1.0
    st["offered_qps"] = high
    st["offered_x_serial_capacity"] = high_load
    # the stream kept up when completions span about the arrivals' span
    span = (max(q.arrival_ns for q in queries) - min(q.arrival_ns for q in queries)) * 1e-6
    st["arrival_span_ms"] = span
    out["batched_at_high_rate"] = st
    out["batched_at_high_rate_by_config"] = latencyByConfig(queries, recs)
    say("at %.1f q/s (%.1fx the serial capacity): batched p50 %.2f ms, p99 %.2f ms, %d of %d in batches; "
        "serial p99 at %.1f q/s: %.2f ms" % (high, high_load, st["p50_ms"], st["p99_ms"], st["in_batches"],
                                             len(recs), rate, out["serial_at_rate"]["p99_ms"]))
    say("at %.1f q/s: p99 serial %.2f ms, batched %.2f ms" % (rate, out["serial_at_rate"]["p99_ms"],
                                                            out["batched_at_rate"]["p99_ms"]))
    return out
