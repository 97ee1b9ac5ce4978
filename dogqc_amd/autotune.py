"""Autotuner for compiled plans.

Searches, per pipeline, the knobs the reference fixes globally
(KernelCall.defaultGridSize/defaultBlockSize, kernel.py:10-11; the refill
threshold, cudaDivergenceBuffer.py:15; the cfg htSizeFactor,
relationalAlgebra.py:453-455) plus the MI355X-specific ones:

  compile time (one module per variant, built ahead of time):
    workgroup size, rows per thread
  run time (dqc_plan_tuning, no recompilation):
    grid size per pipeline, refill threshold, LDS aggregation slots and
    replicas, hash-table size factor

Each candidate is timed with HIP events around every launch (median of a few
runs), so one pipeline's knob is judged by that pipeline's kernel time; the
table size factor is judged by the whole plan (it moves cost between the
build, the probe and the init kernel).  Coordinate descent, one knob at a
time, in a fixed order.  Results are cached as JSON per plan, variant and
input-size bucket.
"""
import json
import math
import os
import statistics

from dogqc_amd import planlib, runtime as rt

PKG = os.path.dirname(os.path.abspath(__file__))
CACHE_DIR = os.path.join(PKG, "_tuning")

VARIANTS = [(256, 4), (256, 2), (256, 1), (512, 4), (128, 4), (256, 8)]


def variantName(name, block, rows):
    return "%s@b%du%d" % (name, block, rows)


def translateVariant(name, block, rows, acc=None):
    comp = planlib.translate(name, acc, block=block, rowsPerThread=rows)
    return comp


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
        for ht in module.desc["hts"]:
            if ht["kind"] in ("unique", "multi", "agg"):
                trySet("ht_factor", ht["index"], [0.5, 0.75, 1.5, 2.0, 4.0], None)
        return best

    def run(self):
        """Tune every compile-time variant; returns the best record."""
        results = []
        for block, rows in self.variants:
            comp = translateVariant(self.name, block, rows)
            path = comp.compile(variantName(self.name, block, rows))
            module = rt.PlanModule(path)
            inputs = rt.BoundInputs(module, self.acc)
            tuning = rt.Tuning()
            times = self.tuneModule(module, inputs, tuning)
            rec = {"plan": self.name, "block": block, "rows_per_thread": rows, "module": os.path.basename(path),
                   "tuning": tuningToDict(tuning), "kernels_ms": times, "total_ms": times["_total"]}
            if self.verbose:
                print("variant b%d u%d: %.4f ms %s" % (block, rows, times["_total"],
                                                     {k: round(v, 4) for k, v in times.items()}), flush=True)
            results.append(rec)
        results.sort(key=lambda r: r["total_ms"])
        return results[0], results


def sizeBucket(acc):
    n = max(t["size"] for t in acc.schema.values()) if acc.schema else 1
    return int(math.log2(max(1, n)))


def cachePath(name, acc):
    return os.path.join(CACHE_DIR, "%s_n2e%d.json" % (name, sizeBucket(acc)))


def loadCached(name, acc):
    p = cachePath(name, acc)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def saveCached(name, acc, rec):
    os.makedirs(CACHE_DIR, exist_ok=True)
    with open(cachePath(name, acc), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)


def main(argv=None):
    import argparse
    from dogqc_amd import workloads
    from dogqc_amd.dbio import MemoryAccessor
    ap = argparse.ArgumentParser(description="autotune a named plan on its synthetic workload")
    ap.add_argument("plan")
    ap.add_argument("--rows-scale", type=float, default=1.0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None, help="also write the record to this path")
    args = ap.parse_args(argv)
    kind = planlib.PLANS[args.plan][0]
    rows = {t: max(1, int(n * args.rows_scale)) for t, n in workloads.FULL_ROWS[kind].items()}
    acc = MemoryAccessor(workloads.generate(kind, rows))
    tuner = Autotuner(args.plan, acc, reps=args.reps, verbose=True)
    best, allRecs = tuner.run()
    saveCached(args.plan, acc, best)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"best": best, "all": allRecs}, f, indent=1, sort_keys=True)
    print(json.dumps(best))


if __name__ == "__main__":
    main()
