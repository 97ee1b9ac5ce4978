"""Autotuner for compiled plans.

Searches, per pipeline, the knobs the reference fixes globally
(KernelCall.defaultGridSize/defaultBlockSize, kernel.py:10-11; the refill
threshold, cudaDivergenceBuffer.py:15; the cfg htSizeFactor,
relationalAlgebra.py:453-455) plus the MI355X-specific ones:

  compile time (one module per variant, built ahead of time):
    workgroup size, rows per thread
  run time (dqc_plan_tuning, no recompilation):
    grid size per pipeline, refill threshold, long-list split threshold
    (pipelines annotated SPLITn), LDS aggregation slots and
    replicas (pipelines annotated LDSn), hash-table size factor, and the
    size of the join filter (LDS sizes together with the probe pipeline's
    grid, MB sizes kept in device memory)

A plan whose estimates are too small for the inputs (a fixed group count, a
selectivity) is first run with retries, and the capacities the retries grew
to become part of the search's starting point and of the record.

Each candidate is timed with HIP events around every launch (median of a few
runs), so one pipeline's knob is judged by that pipeline's kernel time; the
table size factor is judged by the whole plan (it moves cost between the
build, the probe and the init kernel).  Coordinate descent, one knob at a
time, in a fixed order.  Results are cached as JSON per plan and input-size
bucket (log2 of the largest input table's rows); a record holds a hash of
the generated plan source of the variant it was tuned on, so a record whose
plan has changed since is ignored (its per-pipeline grid, refill and table
knobs are indexed by pipeline and table number and could land on other
kernels).  The device headers and compiler flags are not part of that key:
they change how fast a kernel runs, not which knob belongs to which kernel.
``nearest`` picks the valid record of the bucket closest to a query's own
size.
"""
import hashlib
import json
import math
import os
import statistics

from dogqc_amd import planlib, runtime as rt

PKG = os.path.dirname(os.path.abspath(__file__))
CACHE_DIR = os.path.join(PKG, "_tuning")

VARIANTS = [(256, 4), (256, 2), (256, 1), (512, 4), (128, 4), (256, 8), (512, 1), (1024, 1)]
# the TPC-H plans are tuned over fewer compile variants (22 plans)
TPCH_VARIANTS = [(256, 4), (512, 1), (1024, 1)]


# dqc_plan_tuning fields a retry grows (dqc_runtime.hip adjustForRetry)
CAPACITY_FIELDS = ("ht_capacity", "mm_entries", "result_capacity", "temp_capacity", "split_capacity")


def variantName(name, block, rows):
    return "%s@b%du%d" % (name, block, rows)


def translateVariant(name, block, rows, acc=None):
    comp = planlib.translate(name, acc, block=block, rowsPerThread=rows)
    return comp


def planKey(source):
    """Hash of a generated plan source: what a tuning record is valid for."""
    return hashlib.sha256(source.encode()).hexdigest()[:20]


def tuningToDict(t):
    d = {}
    for f, _ in rt.Tuning._fields_:
        v = getattr(t, f)
        if hasattr(v, "__len__"):
            v = list(v)
            if any(v):
                d[f] = v
        elif v:
            d[f] = v
    return d


def tuningFromDict(d):
    t = rt.Tuning()
    for f, v in d.items():
        if isinstance(v, list):
            arr = getattr(t, f)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(t, f, v)
    return t


class Measurer(object):
    """Times one module on bound inputs: per-step kernel times (ms)."""

    def __init__(self, module, inputs, reps=5):
        self.module, self.inputs, self.reps = module, inputs, reps
        self.stream = rt.Stream()
        self.steps = module.steps()
        self.events = [[rt.Event() for _ in range(len(self.steps) + 1)] for _ in range(reps)]
        self.ws = None

    def measure(self, tuning):
        """Median kernel times and whether the tables fit without a retry
        (the caller only accepts tunings that fit)."""
        trial = tuningFromDict(tuningToDict(tuning))  # retries must not leak into `tuning`
        ex = rt.run(self.module, self.inputs, trial, self.stream, max_retries=6)
        ok = ex.res.retries == 0
        # the table, temp and queue capacities the retries grew to
        self.grown = {f: list(getattr(trial, f)) if hasattr(getattr(trial, f), "__len__") else getattr(trial, f)
                      for f in CAPACITY_FIELDS}
        need = ex.res.layout.ws_bytes
        ex.release()
        if self.ws is None or self.ws.nbytes < need:
            self.ws = rt.DeviceBuffer(need)
        rt.enqueue(self.module, self.inputs, trial, self.ws, self.stream)
        for r in range(self.reps):
            rt.enqueue(self.module, self.inputs, trial, self.ws, self.stream, self.events[r])
        self.stream.sync()
        per = {}
        for i, name in enumerate(self.steps):
            per[name] = statistics.median(self.events[r][i].elapsed(self.events[r][i + 1])
                                          for r in range(self.reps))
        per["_total"] = sum(v for k, v in per.items() if k != "_total")
        return per, ok


class Autotuner(object):

    def __init__(self, name, accessor, variants=None, reps=5, verbose=False):
        self.name = name
        self.acc = accessor
        self.variants = variants or VARIANTS
        self.reps = reps
        self.verbose = verbose
        self.log = []

    def _pipelines(self, module):
        return module.desc["pipelines"]

    def tuneModule(self, module, inputs, tuning):
        cus = rt.device()[0]
        m = Measurer(module, inputs, self.reps)
        best, ok = m.measure(tuning)
        if not ok:
            # the plan's estimates are too small for these inputs (a fixed
            # group count, a selectivity): start from the capacities its
            # retries reached, so that every candidate runs without a retry
            # and the record carries them for queries of this size
            for f, v in m.grown.items():
                if isinstance(v, list):
                    arr = getattr(tuning, f)
                    for i, x in enumerate(v):
                        arr[i] = x
                else:
                    setattr(tuning, f, v)
            best, ok = m.measure(tuning)
        steps = m.steps

        def kernelOf(pipe):
            names = [s for s in steps if s.split("<")[0] == pipe["name"]]
            return names

        def score(times, pipe=None):
            if pipe is None:
                return times["_total"]
            return sum(times[k] for k in kernelOf(pipe))

        def trySet(arrName, index, values, pipe):
            nonlocal best
            arr = getattr(tuning, arrName)
            cur = arr[index]
            bestVal, bestScore = cur, score(best, pipe)
            for v in values:
                if v == cur:
                    continue
                arr[index] = v
                times, fits = m.measure(tuning)
                self.log.append((arrName, index, v, score(times, pipe)))
                if fits and score(times, pipe) < bestScore * 0.98:
                    bestVal, bestScore, best = v, score(times, pipe), times
            arr[index] = bestVal

        for pipe in self._pipelines(module):
            pid = pipe["id"]
            trySet("grid", pid, [cus * k for k in (1, 2, 4, 8, 16, 32)], pipe)
            if any(a.startswith("LR") for a in pipe["annotations"]):
                trySet("refill_threshold", pid, [8, 16, 32, 48, 64], pipe)
            if any(a.startswith("SPLIT") for a in pipe["annotations"]):
                # list length above which a probe row's matches go to the
                # second launch (judged by the whole plan: the split scan and
                # both launches move with it)
                trySet("split_threshold", pid, [128, 256, 512, 2048], None)
            if any(a.startswith("LDS") for a in pipe["annotations"]):
                # LDS aggregation width: table slots (-1: no LDS tier), then
                # accumulator replicas
                trySet("lds_slots", pid, [-1, 256, 1024, 4096], pipe)
                if tuning.lds_slots[pid] >= 0:
                    trySet("lds_replicas", pid, [1, 4, 16, 64], pipe)
        if any(a.startswith("PART") for p in self._pipelines(module) for a in p["annotations"]):
            # partitions of the two-phase group-by (0: from the group count,
            # -1: one phase), judged by the whole plan
            keep, bestScore = tuning.part_agg, score(best)
            for v in (-1, 16, 64, 256, 1024):
                if v == keep:
                    continue
                tuning.part_agg = v
                times, fits = m.measure(tuning)
                self.log.append(("part_agg", v, score(times)))
                if fits and score(times) < bestScore * 0.98:
                    keep, bestScore, best = v, score(times), times
            tuning.part_agg = keep
        for ht in module.desc["hts"]:
            if tuning.ht_capacity[ht["index"]] > 0:
                continue        # an explicit capacity overrides the factor
            if ht["kind"] in ("unique", "agg"):
                trySet("ht_factor", ht["index"], [0.5, 0.75, 1.5, 2.0, 4.0], None)
            elif ht["kind"] == "multi":
                # a multimap has a bucket per build row by default; keys that
                # repeat leave most buckets empty, and a smaller directory
                # stays in L2
                trySet("ht_factor", ht["index"], [0.125, 0.25, 0.5, 0.75, 1.5], None)
        # the LDS join filter (pipelines annotated BFn) is copied into LDS by
        # every workgroup of its probe pipeline, so it can only pay with few,
        # large workgroups: its size and those pipelines' grids are searched
        # together, judged by the whole plan (the build sets the filter)
        bfPipes = [p["id"] for p in self._pipelines(module) if any(a.startswith("BF") for a in p["annotations"])]
        if bfPipes:
            keep = (tuning.join_filter, [tuning.grid[p] for p in bfPipes])
            bestScore = score(best)
            tuned = tuning.grid[bfPipes[0]]
            # LDS sizes with few workgroups; MB sizes stay in device memory
            # (L2) and keep the pipeline's own grid
            trials = [(jf, g) for jf in (32, 64, 96, 128) for g in (cus, 2 * cus, 4 * cus)] + \
                [(jf, tuned) for jf in (1024, 2048, 4096)]
            for jf, g in trials:
                tuning.join_filter = jf
                for p in bfPipes:
                    tuning.grid[p] = g
                times, fits = m.measure(tuning)
                self.log.append(("join_filter", jf, g, score(times)))
                if fits and score(times) < bestScore * 0.98:
                    keep, bestScore, best = (jf, [g] * len(bfPipes)), score(times), times
            tuning.join_filter = keep[0]
            for p, g in zip(bfPipes, keep[1]):
                tuning.grid[p] = g
        return best

    def run(self, store=None, rows=None):
        """Tune every compile-time variant; returns the best record.  With
        `store` (a pre-filled column store, e.g. resident tables) the plan
        reads the first rows[t] rows of each table from it."""
        results = []
        for block, rpt in self.variants:
            path = variantPath(self.name, block, rpt)
            module = rt.PlanModule(path)
            inputs = rt.BoundInputs(module, self.acc, store=store, rows=rows)
            tuning = rt.Tuning()
            times = self.tuneModule(module, inputs, tuning)
            tabRows = {t: int(n) for t, n in zip(module.desc["tables"], inputs.rows)}
            rec = {"plan": self.name, "block": block, "rows_per_thread": rpt, "module": os.path.basename(path),
                   "plan_source": planKey(translateVariant(self.name, block, rpt).source),
                   "bucket": bucketOf(max(tabRows.values()) if tabRows else 1), "rows": tabRows,
                   "tuning": tuningToDict(tuning), "kernels_ms": times, "total_ms": times["_total"]}
            if self.verbose:
                print("variant b%d u%d: %.4f ms %s" % (block, rpt, times["_total"],
                                                     {k: round(v, 4) for k, v in times.items()}), flush=True)
            results.append(rec)
        results.sort(key=lambda r: r["total_ms"])
        return results[0], results


def bucketOf(rows):
    return int(math.log2(max(1, int(rows))))


def sizeBucket(acc):
    return bucketOf(max(t["size"] for t in acc.schema.values()) if acc.schema else 1)


def variantPath(name, block, rows):
    """The compiled module of a variant (built ahead of time for the tuned
    plans; compiled here otherwise)."""
    return translateVariant(name, block, rows).compile(variantName(name, block, rows))


def _recordPath(name, bucket):
    return os.path.join(CACHE_DIR, "%s_n2e%d.json" % (name, bucket))


def cachePath(name, acc):
    return _recordPath(name, sizeBucket(acc))


def isCurrent(rec):
    """True when the record was tuned on the plan source the current
    translator generates for its variant."""
    try:
        return planKey(translateVariant(rec["plan"], rec["block"], rec["rows_per_thread"]).source) == \
            rec["plan_source"]
    except Exception:
        return False


def _load(path):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec if isCurrent(rec) else None


def loadCached(name, acc):
    """The current record of the input-size bucket of `acc`, or None."""
    return _load(cachePath(name, acc))


def records(name):
    """Every current record of a plan, by bucket."""
    out = {}
    if os.path.isdir(CACHE_DIR):
        prefix = name + "_n2e"
        for f in sorted(os.listdir(CACHE_DIR)):
            if f.startswith(prefix) and f.endswith(".json") and f[len(prefix):-5].isdigit():
                rec = _load(os.path.join(CACHE_DIR, f))
                if rec is not None:
                    out[int(f[len(prefix):-5])] = rec
    return out


def nearest(recs, rows, within=2):
    """The record whose bucket is closest to that of `rows` driving rows
    (ties: the larger bucket), if within `within` buckets; else None."""
    if not recs:
        return None
    b = bucketOf(rows)
    best = min(recs, key=lambda k: (abs(k - b), -k))
    return recs[best] if abs(best - b) <= within else None


def saveCached(name, acc, rec):
    os.makedirs(CACHE_DIR, exist_ok=True)
    with open(cachePath(name, acc), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)


def saveRecord(rec, directory=None):
    d = directory or CACHE_DIR
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, os.path.basename(_recordPath(rec["plan"], rec["bucket"])))
    with open(path, "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
    return path


def tuneServing(plans, scales=(0.1, 1.0, 10.0), tableScale=10.0, reps=5, out=None, log=print, variants=None):
    """Tune plans on HBM-resident tables at `tableScale` x the base sizes
    (generated on the device), once per row scale: a query at scale s reads
    the first s x base rows of every table, as the serving workload's do."""
    from dogqc_amd import datagen, serving, workloads
    tables = datagen.ResidentTables(scale=tableScale)
    acc = tables.schemaAccessor()
    saved = []
    for name in plans:
        config = planlib.PLANS[name][0]
        for sc in scales:
            rows = serving.prefixRows(tables, config, sc)
            tuner = Autotuner(name, acc, variants=variants, reps=reps, verbose=True)
            best, allRecs = tuner.run(store=tables.store, rows=rows)
            best["scale"] = sc
            path = saveRecord(best, out)
            saved.append(path)
            log("%s x%g (bucket %d): b%d u%d %.4f ms (variants %s) -> %s" % (
                name, sc, best["bucket"], best["block"], best["rows_per_thread"], best["total_ms"],
                " ".join("b%du%d:%.4f" % (r["block"], r["rows_per_thread"], r["total_ms"]) for r in allRecs),
                os.path.basename(path)))
    return saved


def main(argv=None):
    import argparse
    from dogqc_amd import workloads
    from dogqc_amd.dbio import MemoryAccessor
    ap = argparse.ArgumentParser(description="autotune a named plan on its synthetic workload")
    ap.add_argument("plan", nargs="+")
    ap.add_argument("--rows-scale", type=float, default=1.0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None, help="also write the record to this path")
    ap.add_argument("--serving", action="store_true",
                    help="tune at 0.1x, 1x and 10x on device-resident tables (records to --out-dir)")
    ap.add_argument("--scales", default="0.1,1,10")
    ap.add_argument("--out-dir", default=None, help="write records here instead of the package's record folder")
    ap.add_argument("--variants", default=None,
                    help="compile-time variants to search, as BLOCKxROWS[,...] (default: all eight)")
    args = ap.parse_args(argv)
    variants = None
    if args.variants:
        variants = [tuple(int(x) for x in v.split("x")) for v in args.variants.split(",")]
        for v in variants:
            if v not in VARIANTS:
                ap.error("variant %dx%d is not built (built: %s)" % (v + (", ".join("%dx%d" % w for w in VARIANTS),)))
        if any(planlib.PLANS[n][0] == "tpch" for n in args.plan) and any(v not in TPCH_VARIANTS for v in variants):
            ap.error("TPC-H plans are built in the variants %s" % ", ".join("%dx%d" % w for w in TPCH_VARIANTS))
    if args.serving:
        tuneServing(args.plan, [float(x) for x in args.scales.split(",")], reps=args.reps, out=args.out_dir,
                    log=lambda m: print(m, flush=True), variants=variants)
        return
    kinds = set(planlib.PLANS[n][0] for n in args.plan)
    if kinds == {"tpch"}:
        # TPC-H plans: synthetic data at scale factor --rows-scale, once for
        # all of them, and their own compile variants
        from dogqc_amd import tpch
        tacc = MemoryAccessor(tpch.generate(args.rows_scale, seed=10))
        variants = variants or TPCH_VARIANTS
    elif len(args.plan) > 1:
        ap.error("several plans at once only for TPC-H plans")
    for name in args.plan:
        kind = planlib.PLANS[name][0]
        if kind == "tpch":
            acc = tacc
        else:
            rows = {t: max(1, int(n * args.rows_scale)) for t, n in workloads.FULL_ROWS[kind].items()}
            acc = MemoryAccessor(workloads.generate(kind, rows))
        tuner = Autotuner(name, acc, variants=variants, reps=args.reps, verbose=True)
        best, allRecs = tuner.run()
        best["scale"] = args.rows_scale
        if args.out_dir:
            saveRecord(best, args.out_dir)
        else:
            saveCached(name, acc, best)
        if args.out:
            with open(args.out, "w") as f:
                json.dump({"best": best, "all": allRecs}, f, indent=1, sort_keys=True)
        print(json.dumps(best), flush=True)


if __name__ == "__main__":
    main()
