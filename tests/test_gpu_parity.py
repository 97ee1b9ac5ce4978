"""GPU parity: every named plan, run through the C ABI on the HIP path,
against the CPU oracle (oracle/planeval.py) on the same seeded inputs.

Integer, string and per-row floating-point columns are bit-exact; FP64
SUM/AVG aggregates within 1e-9 relative (tests/helpers.py).
"""
import numpy as np
import pytest

from dogqc_amd import planlib, workloads, runtime as rt, runner
from dogqc_amd.dbio import MemoryAccessor
from oracle.planeval import evaluate
from tests.helpers import assertSameResult, toleranceColumns

pytestmark = pytest.mark.gpu

_MODULES = {}
_STORE = {}


def module(name):
    if name not in _MODULES:
        comp = planlib.translate(name)
        _MODULES[name] = rt.PlanModule(comp.compile(name))
    return _MODULES[name]


def runPlan(name, acc, tuning=None, storeKey=None):
    """GPU result (decoded) and the resolved plan used to compute it."""
    mod = module(name)
    store = _STORE.setdefault(storeKey, {}) if storeKey else {}
    inputs = rt.BoundInputs(mod, acc, store=store)
    ex = rt.run(mod, inputs, tuning or rt.newTuning())
    res = runner.resultColumns(mod, ex, acc)
    retries = ex.res.retries
    ex.release()
    return res, retries


def oracleFor(name, acc):
    alg, p, cfg, opts = planlib.build(name, acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    return evaluate(plan, acc), plan


def check(name, acc, tuning=None, storeKey=None, label=None):
    print("gpu case: %s %s" % (name, label or ""), flush=True)
    got, retries = runPlan(name, acc, tuning, storeKey)
    want, plan = oracleFor(name, acc)
    assertSameResult(got, want, toleranceColumns(plan), label="%s %s" % (name, label or ""))
    return got, retries


def workloadAcc(config, rows=None, seed=None):
    return MemoryAccessor(workloads.generate(config, rows, seed))


# ------------------------------------------------------------ workloads

@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 100003])
def test_c0_ragged(n):
    check("c0", workloadAcc("c0", {"cfg0": n}, seed=11 + n), label="n=%d" % n)


def test_c0_full():
    check("c0", workloadAcc("c0"), storeKey="c0full", label="1M rows")


@pytest.mark.parametrize("lds", [(-1, 0), (2, 0), (16, 1), (16, 8), (4096, 0)])
def test_c0_lds_tiers(lds):
    t = rt.newTuning()
    for i in range(rt.MAX_PIPES):
        t.lds_slots[i], t.lds_replicas[i] = lds
    check("c0", workloadAcc("c0", {"cfg0": 200000}, seed=5), tuning=t, label="lds=%r" % (lds,))


@pytest.mark.parametrize("grid", [1, 7, 4096])
def test_c0_grid(grid):
    t = rt.newTuning()
    for i in range(rt.MAX_PIPES):
        t.grid[i] = grid
    check("c0", workloadAcc("c0", {"cfg0": 300000}, seed=6), tuning=t, label="grid=%d" % grid)


@pytest.mark.parametrize("plan", ["c1", "c1_refill", "c1_refill_reg", "c1_refill_g"])
def test_c1_full(plan):
    check(plan, workloadAcc("c1"), storeKey="c1full", label="1.5M x 6M")


@pytest.mark.parametrize("plan", ["c1", "c1_refill", "c1_refill_reg", "c1_refill_g"])
@pytest.mark.parametrize("rows", [{"dim": 0, "fact": 1000}, {"dim": 1000, "fact": 0}, {"dim": 1, "fact": 1},
                                  {"dim": 3000, "fact": 20000}])
def test_c1_small(rows, plan):
    check(plan, workloadAcc("c1", rows, seed=21), label=repr(rows))


def test_c1_undersized_tables_retry():
    t = rt.newTuning()
    t.ht_capacity[0] = 16   # join table
    t.ht_capacity[1] = 8    # group table
    _, retries = check("c1", workloadAcc("c1", {"dim": 5000, "fact": 40000}, seed=22), tuning=t,
                       label="tiny tables")
    assert retries > 0


C2_VARIANTS = ["c2_none", "c2_refill", "c2_pushdown", "c2_both", "c2_refill_reg", "c2_both_reg", "c2_refill_g"]


@pytest.mark.parametrize("variant", C2_VARIANTS)
def test_c2_small(variant):
    check(variant, workloadAcc("c2", {"build": 2000, "probe": 50000}, seed=31), label="small")


@pytest.mark.parametrize("variant", C2_VARIANTS)
def test_c2_empty(variant):
    check(variant, workloadAcc("c2", {"build": 0, "probe": 100}, seed=32), label="empty build")
    check(variant, workloadAcc("c2", {"build": 100, "probe": 0}, seed=33), label="empty probe")


@pytest.mark.parametrize("plan", ["c2_both", "c2_both_reg", "c2_refill_reg"])
@pytest.mark.parametrize("thr", [1, 17, 63, 64])
def test_c2_refill_thresholds(thr, plan):
    t = rt.newTuning()
    for i in range(rt.MAX_PIPES):
        t.refill_threshold[i] = thr
    check(plan, workloadAcc("c2", {"build": 5000, "probe": 100000}, seed=34), tuning=t, label="thr=%d" % thr)


@pytest.mark.parametrize("thr", [1, 7, 100000])
def test_c2_split_thresholds(thr):
    """Long match lists queued for the second launch at any threshold (1:
    nearly every list, 100000: none) give the same result."""
    t = rt.newTuning()
    for i in range(rt.MAX_PIPES):
        t.split_threshold[i] = thr
    check("c2_pushdown", workloadAcc("c2", {"build": 20000, "probe": 200000}, seed=35), tuning=t,
          label="split threshold %d" % thr)


def test_c2_split_queue_overflow_retry():
    t = rt.newTuning()
    for i in range(rt.MAX_PIPES):
        t.split_threshold[i] = 2
    t.split_capacity[0] = 3
    _, retries = check("c2_pushdown", workloadAcc("c2", {"build": 20000, "probe": 200000}, seed=36), tuning=t,
                       label="tiny split queue")
    assert retries > 0


def c2ClosedForm(acc):
    """COUNT and SUM(b_pay * p_pay) of c2 without enumerating pairs:
    per key k, count = nb(k) * np(k) and sum = Sb(k) * Sp(k)."""
    bk = np.asarray(acc.column("build", "b_key")).astype(np.int64)
    bp = np.asarray(acc.column("build", "b_pay"))
    pk = np.asarray(acc.column("probe", "p_key")).astype(np.int64)
    pf = np.asarray(acc.column("probe", "p_f"))
    pp = np.asarray(acc.column("probe", "p_pay"))
    m = pf < 10
    pk, pp = pk[m], pp[m]
    K = int(max(bk.max(initial=0), pk.max(initial=0))) + 1
    nb = np.bincount(bk, minlength=K).astype(object)
    npr = np.bincount(pk, minlength=K).astype(object)
    sb = np.zeros(K, dtype=object)
    sp = np.zeros(K, dtype=object)
    np.add.at(sb, bk, bp.astype(object))
    np.add.at(sp, pk, pp.astype(object))
    cnt = int(np.dot(nb, npr))
    s = int(np.dot(sb, sp))
    return cnt, (s + 2 ** 63) % 2 ** 64 - 2 ** 63


@pytest.mark.parametrize("variant", C2_VARIANTS)
def test_c2_full_closed_form(variant):
    acc = _STORE.setdefault("c2acc", workloadAcc("c2"))
    print("gpu case: %s full 1M x 16M" % variant, flush=True)
    got, _ = runPlan(variant, acc, storeKey="c2full")
    cnt, s = c2ClosedForm(acc)
    assert int(got["cnt"][0]) == cnt
    assert int(got["sum_prod"][0]) == s


# ------------------------------------------------------------ c3 star join

C3_SMALL = {"sales": 200000, "item": 20000, "cust": 2000, "nation": 25}


def c3Run(acc, tuning=None, storeKey=None, plan="c3"):
    mod = module(plan)
    store = _STORE.setdefault(storeKey, {}) if storeKey else {}
    inputs = rt.BoundInputs(mod, acc, store=store)
    ex = rt.run(mod, inputs, tuning or rt.newTuning())
    got = runner.resultColumns(mod, ex, acc)
    ny = runner.tempColumns(mod, ex, acc, "c3_by_nation_year")
    retries = ex.res.retries
    ex.release()
    return got, ny, retries


def c3Oracle(acc):
    alg, p, cfg, _ = planlib.build("c3", acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    from oracle.planeval import Evaluator
    ev = Evaluator(acc)
    res = ev.run(plan)
    return res, ev.temps["c3_by_nation_year"][0]


@pytest.mark.parametrize("plan", ["c3", "c3_refill", "c3_refill_reg", "c3_refill_g"])
@pytest.mark.parametrize("rows", [C3_SMALL, {"sales": 1, "item": 1, "cust": 1, "nation": 1},
                                  {"sales": 0, "item": 100, "cust": 10, "nation": 25},
                                  {"sales": 5000, "item": 0, "cust": 10, "nation": 25},
                                  {"sales": 70001, "item": 3000, "cust": 300, "nation": 5}])
def test_c3_small(rows, plan):
    print("gpu case: %s %r" % (plan, rows), flush=True)
    acc = workloadAcc("c3", rows, seed=41)
    got, ny, _ = c3Run(acc, plan=plan)
    want, wantNy = c3Oracle(acc)
    assertSameResult(got, want, {"revenue"}, label="c3 by item")
    assertSameResult(ny, wantNy, {"revenue"}, label="c3 by nation, year")


def test_c3_temp_overflow_retry():
    print("gpu case: c3 undersized temp tables", flush=True)
    acc = workloadAcc("c3", C3_SMALL, seed=42)
    t = rt.newTuning()
    t.temp_capacity[0] = 10
    t.temp_capacity[1] = 1
    got, ny, retries = c3Run(acc, tuning=t)
    assert retries > 0
    want, wantNy = c3Oracle(acc)
    assertSameResult(got, want, {"revenue"}, label="c3 retry by item")
    assertSameResult(ny, wantNy, {"revenue"}, label="c3 retry by nation, year")


@pytest.mark.parametrize("plan", ["c3", "c3_refill_reg", "c3_refill_g"])
def test_c3_full_vs_c_oracle(plan):
    from oracle import coracle
    print("gpu case: %s full 60M x 2M x 200k x 25" % plan, flush=True)
    data = workloads.generate("c3")
    acc = MemoryAccessor(data)
    got, ny, _ = c3Run(acc, plan=plan, storeKey="c3full")
    tables = {t: {a: v[1] for a, v in cols.items()} for t, cols in data.items()}
    wantItem, wantNy = coracle.c3(tables, workloads.C3_WORD)
    assert len(wantItem["i_key"]) > 90000 and len(wantNy["n_key"]) == 175
    assertSameResult(got, wantItem, {"revenue"}, label="c3 full by item")
    assertSameResult(ny, wantNy, {"revenue"}, label="c3 full by nation, year")


def test_c3_graph_replays_match_eager():
    """The whole c3 plan captured into one hipGraph, replayed three times
    with host reads of the workspace in between: every replay equals the
    eager run."""
    print("gpu case: c3 graph replays", flush=True)
    acc = workloadAcc("c3", C3_SMALL, seed=43)
    mod = module("c3")
    inputs = rt.BoundInputs(mod, acc)
    t = rt.newTuning()
    ex = rt.run(mod, inputs, t)       # sizes the tables (tuning updated on retries)
    eager = runner.resultColumns(mod, ex, acc)
    eagerNy = runner.tempColumns(mod, ex, acc, "c3_by_nation_year")
    ex.release()
    lay = rt.layout(mod, inputs, t)
    ws = rt.DeviceBuffer(lay.ws_bytes)
    stream = rt.Stream()
    g = rt.Graph(mod, inputs, t, ws, stream)
    for k in range(3):
        g.launch(stream)
        stream.sync()
        v = rt.view(mod, inputs, t, ws)
        assertSameResult(runner.resultColumns(mod, v, acc), eager, {"revenue"}, rtol=1e-12,
                         label="replay %d" % k)
        assertSameResult(runner.tempColumns(mod, v, acc, "c3_by_nation_year"), eagerNy, {"revenue"}, rtol=1e-12,
                         label="replay %d temp" % k)
    g.destroy()


# ------------------------------------------------------------ features

@pytest.mark.parametrize("name", [n for n in planlib.PLANS if planlib.PLANS[n][0] == "feature"])
@pytest.mark.parametrize("size", [(5000, 700), (200, 50), (1, 1), (0, 0)])
def test_feature_plans(name, size):
    acc = planlib.featureData(*size)
    check(name, acc, label="t=%d u=%d" % size)


@pytest.mark.parametrize("name", planlib.STRING_PLANS)
@pytest.mark.parametrize("rows", [None, 1, 0])
def test_string_plans(name, rows):
    """LIKE (the reference's semantics, util.cuh:148-199), SUBSTRING, IN and
    string equality on every short string against 115 patterns, per row."""
    acc = planlib.stringData()
    if rows is not None:
        acc = planlib.stringRows(acc, rows)
    check(name, acc, label="rows=%s" % rows)


def test_feature_retry_all_buffers():
    acc = planlib.featureData()
    t = rt.newTuning()
    t.result_capacity = 1
    t.mm_entries[0] = 1
    _, retries = check("f_join_multi", acc, tuning=t, label="tiny result + entries")
    assert retries > 0
    t = rt.newTuning()
    t.ht_capacity[0] = 2
    _, retries = check("f_groupby_wide", acc, tuning=t, label="tiny group table")
    assert retries > 0


# ------------------------------------------------------------ TPC-H

_TPCH = {}


def tpchAcc(sf):
    if sf not in _TPCH:
        from dogqc_amd import tpch
        _TPCH[sf] = MemoryAccessor(tpch.generate(sf, seed=3))
    return _TPCH[sf]


@pytest.mark.parametrize("q", range(1, 23))
@pytest.mark.parametrize("sf", [0.05, 0.002])
def test_tpch(q, sf):
    """TPC-H Q1-Q22 on synthetic TPC-H data (300k and 12k lineitem rows),
    every operator of the plan API on the GPU."""
    check("tpch%d" % q, tpchAcc(sf), storeKey=("tpch", sf), label="sf=%g" % sf)


@pytest.mark.parametrize("name", ["tpch10_p9", "tpch10_lr9_p9", "tpch10_lr9reg_p9", "tpch10_ws_lr9_p9",
                                  "tpch10_lrg9_p9", "tpch3_profiled", "tpch3_profiled_reg", "tpch3_profiled_g"]
                         + ["tpch%d_ws" % q for q in planlib.WAVE_STRING_TPCH])
@pytest.mark.parametrize("sf", [0.05, 0.002])
def test_tpch_refill_variants(name, sf):
    """Q10 and Q3 with lane refill (LDS queue and register buffer) and
    lane-activity profiles in the plan, and the queries with string keys
    hashed by the whole wavefront, against the oracle."""
    check(name, tpchAcc(sf), storeKey=("tpch", sf), label="sf=%g" % sf)



_IMPORTED = {}


def importedTpch(root):
    """Synthetic TPC-H (12k lineitem rows) written as dbgen .tbl files and
    imported through dbio.dbAccess (libdqcimport.so) into an mmdb directory;
    plans read the column files through the runtime's C loader."""
    if "acc" not in _IMPORTED:
        from dogqc_amd import tpch, dbio
        tpch.writeTbl(tpch.generate(0.002, seed=8), str(root / "tbl"))
        _IMPORTED["acc"] = dbio.dbAccess(tpch.schema(), str(root / "mmdb"), str(root / "tbl"))
    return _IMPORTED["acc"]


@pytest.mark.parametrize("q", range(1, 23))
def test_tpch_on_imported_tbl(q, tmp_path_factory):
    """The reference's data path end to end (dbio.py:13-27, 117-253):
    .tbl -> importer -> mmdb column files -> HBM (dqc_mmdb_load) -> HIP
    plan, against the oracle over the same mmdb columns."""
    from dogqc_amd import dbio
    acc = importedTpch(tmp_path_factory.getbasetemp())
    assert isinstance(acc, dbio.Accessor) and acc.schema["lineitem"]["size"] > 10000
    check("tpch%d" % q, acc, storeKey=("imported",), label="imported .tbl")


def test_plan_file_flow_on_gpu(tmp_path, monkeypatch, tmp_path_factory):
    """A plan file and cfg file in the reference's format
    (tests/golden/plans), evaluated as query/tpch.py:46-58 does, compiled and
    executed through HipCompiler on imported data."""
    import os
    import dogqc_amd.scalarAlgebra as scal
    from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction
    from dogqc_amd.types import Type
    from dogqc_amd.util import loadScript
    from dogqc_amd.hipTranslator import HipCompiler
    from dogqc_amd.hiplang import CType
    print("gpu case: plan file tpch10 on imported data", flush=True)
    acc = importedTpch(tmp_path_factory.getbasetemp())
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans")
    monkeypatch.chdir(tmp_path)
    alg = RelationalAlgebra(acc)
    scope = {"alg": alg, "scal": scal, "Reduction": Reduction, "Type": Type}
    plan = eval(loadScript(os.path.join(here, "tpch10.py")), scope)
    cfg = eval(loadScript(os.path.join(here, "tpch10cfg.py")), scope)
    plan = alg.resolveAlgebraPlan(plan, cfg)
    comp = HipCompiler(algebraContext=alg, smArchitecture="sm_75", decimalRepr=CType.FP64, debug=False)
    comp.gencode(alg.translateToCompilerPlan(plan, comp))
    comp.compile("tpch10")
    out = comp.execute()
    want, rplan = oracleFor("tpch10", acc)
    assertSameResult(comp.result, want, toleranceColumns(rplan), label="plan file tpch10")
    assert "Result: %d tuples" % len(want["c_custkey"]) in out


def test_reference_driver_flow_with_profiler(tmp_path, monkeypatch):
    """The reference's driver sequence (query/tpch.py:41-69, profiler.py:
    27-35): resolve, HipCompiler with profilers and a lane-refill buffer,
    gencode, compile, execute(); the printed protocol is kept in <name>.log
    with a <pN> histogram whose iterations add up to the profiled wave trips."""
    import re
    from dogqc_amd import tpch
    from dogqc_amd.hipTranslator import HipCompiler
    from dogqc_amd.hiplang import CType
    print("gpu case: driver flow tpch3 with profiler", flush=True)
    assert planlib.PLANS["tpch3_profiled"][1]["profilers"] == [7]
    monkeypatch.chdir(tmp_path)
    acc = tpchAcc(0.002)
    alg, plan, cfg = tpch.q3(acc)
    plan = alg.resolveAlgebraPlan(plan, cfg)
    sel = [n.opId for n in alg.selectionNodes]
    comp = HipCompiler(algebraContext=alg, smArchitecture="sm_75", decimalRepr=CType.FP64, debug=False)
    comp.setProfilers(sel[-1:])
    comp.setBuffers(sel[:1])
    comp.gencode(alg.translateToCompilerPlan(plan, comp))
    comp.compile("tpch3_profiled")          # the prebuilt module of this plan
    out = comp.execute()
    log = (tmp_path / "tpch3_profiled.log").read_text()
    assert log == out
    block = re.search(r"<p%d>\n(.*?)</p%d>" % (sel[-1], sel[-1]), log, re.S).group(1)
    hist = [int(l.split(",")[1]) for l in block.strip().splitlines()]
    assert len(hist) == 65 and sum(hist) > 0
    assert "totalKernelTime" in log and "LR%d" % sel[0] in log and "P%d" % sel[-1] in log
    want, rplan = oracleFor("tpch3", acc)
    assertSameResult(comp.result, want, toleranceColumns(rplan), label="driver flow tpch3")
    assert "Result: %d tuples" % len(want["l_orderkey"]) in log


@pytest.mark.parametrize("reps", [1, 2, 8])
@pytest.mark.parametrize("name", ["f_groupby_hashed", "f_groupby_wide", "f_groupby_64", "f_outer", "f_scan_minmax"])
def test_group_accumulator_copies(name, reps):
    """Global group-by accumulators kept in 1, 2 or 8 copies (workgroup b
    updates copy b % R; the output pass combines them: sums added, MIN/MAX
    folded, AVG from the combined sum and count)."""
    t = rt.newTuning()
    t.xcd_replicas = reps
    t.lds_slots[0] = -1     # no LDS tier in the first pipeline: all rows take the global path
    t.lds_slots[1] = -1
    check(name, planlib.featureData(20000, 700, seed=5), tuning=t, label="copies=%d" % reps)


@pytest.mark.parametrize("reps", [1, 8])
def test_c3_group_copies(reps):
    t = rt.newTuning()
    t.xcd_replicas = reps
    check("c3_refill", workloadAcc("c3", {"sales": 200000, "item": 20000, "cust": 2000, "nation": 25}, seed=31),
          tuning=t, label="copies=%d" % reps)


# ------------------------------------------------------------ cardinality feedback

def test_census_counts_table_keys():
    """tuning.census: ht_fill holds the keys each table really holds (the
    dim rows passing c1's filter; the groups of a hashed group-by)."""
    print("gpu case: census", flush=True)
    acc = workloadAcc("c1", {"dim": 30000, "fact": 100000}, seed=41)
    mod = module("c1")
    t = rt.newTuning()
    t.census = 1
    ex = rt.run(mod, rt.BoundInputs(mod, acc), t)
    dim = acc.column("dim", "d_cat")
    joinIdx = [h["index"] for h in mod.desc["hts"] if h["kind"] == "unique"][0]
    assert ex.res.status.ht_fill[joinIdx] == int(np.sum(dim == ord("2")))
    ex.release()
    facc = planlib.featureData(20000, 700, seed=5)
    fmod = module("f_groupby_hashed")
    ex = rt.run(fmod, rt.BoundInputs(fmod, facc), t)
    aggIdx = [h["index"] for h in fmod.desc["hts"] if h["kind"] == "agg"][0]
    assert ex.res.status.ht_fill[aggIdx] == ex.count
    ex.release()


@pytest.mark.parametrize("name,acc", [("c1", lambda: workloadAcc("c1", {"dim": 30000, "fact": 200000}, seed=42)),
                                      ("f_groupby_wide", lambda: planlib.featureData(20000, 700, seed=6)),
                                      ("tpch18", lambda: tpchAcc(0.002)), ("tpch13", lambda: tpchAcc(0.002))])
def test_feedback_sizes_the_next_run(name, acc):
    """A first run whose tables are far too small retries; the sizes it
    measured let the second run (given the same bad sizes) finish without a
    retry, with the same result."""
    print("gpu case: feedback %s" % name, flush=True)
    a = acc()
    mod = module(name)
    fb = rt.Feedback()

    def tiny():
        t = rt.newTuning()
        for i in range(len(mod.desc["hts"])):
            t.ht_capacity[i] = 16
        t.result_capacity = 1
        return t
    ex1 = rt.run(mod, rt.BoundInputs(mod, a), tiny(), feedback=fb)
    first = runner.resultColumns(mod, ex1, a)
    assert ex1.res.retries > 0
    ex1.release()
    ex2 = rt.run(mod, rt.BoundInputs(mod, a), tiny(), feedback=fb)
    assert ex2.res.retries == 0
    second = runner.resultColumns(mod, ex2, a)
    ex2.release()
    want, plan = oracleFor(name, a)
    assertSameResult(first, want, toleranceColumns(plan), label="%s first" % name)
    assertSameResult(second, want, toleranceColumns(plan), label="%s second" % name)


# ------------------------------------------------------------ batched plans

@pytest.mark.parametrize("config", ["c0", "c1", "c2", "c3"])
def test_batched_plan_on_gpu(config):
    """One GPU run of a batched plan over five members' prefixes, combined
    per member on the host, equals each member's own query on the oracle."""
    from dogqc_amd import batching
    from tests.test_batching import SIZES, SCALES, prefixes, oracleMember
    print("gpu case: batched %s" % config, flush=True)
    sizes = {t: n * 10 for t, n in SIZES[config].items()}
    if config == "c3":
        sizes["nation"] = 25
    acc = MemoryAccessor(workloads.generate(config, sizes, seed=9))
    pref = {t: [max(1, int(round(sizes[t] * s))) for s in SCALES] for t in batching.SCALED[config]}
    bacc = batching.batchAccessor(acc, pref)
    mod = module("%s_batch" % config)
    ex = rt.run(mod, rt.BoundInputs(mod, bacc), rt.newTuning())
    got = runner.resultColumns(mod, ex, bacc)
    temps = {t["name"]: runner.tempColumns(mod, ex, bacc, t["name"]) for t in mod.desc["temps"]}
    ex.release()
    for k in range(len(SCALES)):
        want, wantTemps, plan = oracleMember(config, acc, pref, k)
        res, tmp = batching.combine(config, got, temps, k)
        assertSameResult(res, want, toleranceColumns(plan), label="%s member %d" % (config, k))
        for name, cols in tmp.items():
            w = {c: wantTemps[name][c] for c in cols}
            assertSameResult(cols, w, {"revenue"}, label="%s member %d %s" % (config, k, name))


# ------------------------------------------------------------ partitioned group-by

# plans with a group-by the translator can launch partitioned (exact keys, a
# pipeline loop uniform over the workgroup)
PART_FEATURES = ["f_semi", "f_anti_multi", "f_anti_neg", "f_cross", "f_like_patterns"]
PART_TPCH = ["1_unhinted", 2, "8_plain", 11, 15, 18, 20]


def partTuning(parts):
    t = rt.newTuning()
    t.part_agg = parts
    return t


def assertPartitioned(name):
    ann = [a for p in module(name).desc["pipelines"] for a in p["annotations"]]
    assert any(a.startswith("PART") for a in ann), name


@pytest.mark.parametrize("parts", [2, 64])
@pytest.mark.parametrize("name", PART_FEATURES)
def test_partitioned_groupby_features(name, parts):
    """Group-bys forced through the partitioned tier (tuning part_agg = n
    partitions): rows scattered to partitions, each partition aggregated by
    one workgroup in LDS."""
    assertPartitioned(name)
    for size in [(5000, 700), (200, 50), (1, 1)]:
        check(name, planlib.featureData(*size), tuning=partTuning(parts),
              label="parts=%d t=%d u=%d" % ((parts,) + size))


@pytest.mark.parametrize("parts", [2, 64, 1024])
@pytest.mark.parametrize("q", PART_TPCH)
def test_partitioned_groupby_tpch(q, parts):
    """TPC-H group-bys through the partitioned tier.  Q1's 4 groups (its plan
    without the group-count hint) land in at most 4 partitions, far more rows
    than a partition holds: the rows beyond its capacity update the global
    table directly."""
    name = "tpch%s" % q
    assertPartitioned(name)
    check(name, tpchAcc(0.05), tuning=partTuning(parts), storeKey=("tpch", 0.05), label="parts=%d" % parts)


@pytest.mark.parametrize("parts", [-1, 0, 16, 64])
def test_partitioned_groupby_c3(parts):
    """c3's by-item group-by (numgroups 120000: partitioned by default, 64
    partitions), off (-1) and forced to 16 / 64 partitions."""
    print("gpu case: c3_refill part_agg=%d" % parts, flush=True)
    acc = workloadAcc("c3", {"sales": 200000, "item": 20000, "cust": 2000, "nation": 25}, seed=33)
    got, ny, _ = c3Run(acc, tuning=partTuning(parts), plan="c3_refill")
    want, wantNy = c3Oracle(acc)
    assertSameResult(got, want, {"revenue"}, label="c3 by item parts=%d" % parts)
    assertSameResult(ny, wantNy, {"revenue"}, label="c3 by nation, year parts=%d" % parts)


# ------------------------------------------------------------ join filters

def filterTuning(kb):
    t = rt.newTuning()
    t.join_filter = kb
    return t


@pytest.mark.parametrize("kb", [-1, 1, 4, 64, 2048])
@pytest.mark.parametrize("plan", ["c1", "c1_refill"])
def test_join_filter_c1(plan, kb):
    """The LDS filter of c1's selective dim join (tuning join_filter): off
    (-1), 1 KB (many false positives), 4 KB and 64 KB; rows the filter rules
    out skip the probe."""
    for rows in [{"dim": 30000, "fact": 200000}, {"dim": 1, "fact": 1000}, {"dim": 1000, "fact": 0}]:
        check(plan, workloadAcc("c1", rows, seed=51), tuning=filterTuning(kb), label="kb=%d %r" % (kb, rows))


@pytest.mark.parametrize("kb", [-1, 1, 32, 96, 1024])
def test_join_filter_c3(kb):
    print("gpu case: c3_refill join_filter=%d" % kb, flush=True)
    acc = workloadAcc("c3", {"sales": 200000, "item": 20000, "cust": 2000, "nation": 25}, seed=52)
    got, ny, _ = c3Run(acc, tuning=filterTuning(kb), plan="c3_refill")
    want, wantNy = c3Oracle(acc)
    assertSameResult(got, want, {"revenue"}, label="c3 by item kb=%d" % kb)
    assertSameResult(ny, wantNy, {"revenue"}, label="c3 by nation, year kb=%d" % kb)


@pytest.mark.parametrize("kb", [-1, 1, 64])
def test_join_filter_tpch10(kb):
    check("tpch10", tpchAcc(0.05), tuning=filterTuning(kb), storeKey=("tpch", 0.05), label="kb=%d" % kb)


# ------------------------------------------------------------ partitioned build

PB_FEATURES = planlib.PB_FEATURES
PB_TPCH = planlib.PB_TPCH


def buildTuning(parts):
    t = rt.newTuning()
    t.part_build = parts
    return t


@pytest.mark.parametrize("parts", [-1, 2, 64, 1024])
@pytest.mark.parametrize("name", PB_FEATURES)
def test_partitioned_build_features(name, parts):
    """Unique builds forced through the partitioned build (tuning part_build
    = n partitions): records scattered to slot-range partitions, each range
    built in LDS and written with plain stores.  With 1024 partitions a small
    table's ranges are a slot or two wide, so most keys spill into the CAS
    pass; f_join_unique_lowest sizes the segments for ~1 build row, so most
    rows overflow them and are inserted directly."""
    name += "_pb"
    ann = [a for p in module(name).desc["pipelines"] for a in p["annotations"]]
    assert any(a.startswith("PB") for a in ann), name
    for size in [(5000, 700), (200, 50), (1, 1), (0, 0)]:
        check(name, planlib.featureData(*size), tuning=buildTuning(parts),
              label="parts=%d t=%d u=%d" % ((parts,) + size))


@pytest.mark.parametrize("parts", [2, 1024])
@pytest.mark.parametrize("q", PB_TPCH)
def test_partitioned_build_tpch(q, parts):
    check("tpch%d_pb" % q, tpchAcc(0.05), tuning=buildTuning(parts), storeKey=("tpch", 0.05), label="parts=%d" % parts)


@pytest.mark.parametrize("parts", [0, 16, 1024])
@pytest.mark.parametrize("plan", ["c1_pb", "c1_refill_pb"])
def test_partitioned_build_c1(plan, parts):
    for rows in [{"dim": 300000, "fact": 400000}, {"dim": 1000, "fact": 5000}]:
        check(plan, workloadAcc("c1", rows, seed=61), tuning=buildTuning(parts), label="parts=%d %r" % (parts, rows))


@pytest.mark.parametrize("parts", [0, 64])
def test_partitioned_build_c3(parts):
    print("gpu case: c3_refill part_build=%d" % parts, flush=True)
    acc = workloadAcc("c3", {"sales": 200000, "item": 200000, "cust": 100000, "nation": 25}, seed=62)
    got, ny, _ = c3Run(acc, tuning=buildTuning(parts), plan="c3_refill_pb")
    want, wantNy = c3Oracle(acc)
    assertSameResult(got, want, {"revenue"}, label="c3 by item parts=%d" % parts)
    assertSameResult(ny, wantNy, {"revenue"}, label="c3 by nation, year parts=%d" % parts)


@pytest.mark.parametrize("plan", ["c3_refill_g", "c1_refill_g"])
def test_global_refill_queue_overflow_retry(plan):
    """A kernel-split refill queue sized far below the survivors overflows:
    the status reports the count and the run repeats with that capacity."""
    mod = module(plan)
    q = [h["index"] for h in mod.desc["hts"] if h["kind"] == "queue"]
    assert len(q) == 1
    t = rt.newTuning()
    t.split_capacity[q[0]] = 16
    config = "c3" if plan.startswith("c3") else "c1"
    rows = C3_SMALL if config == "c3" else {"dim": 5000, "fact": 60000}
    acc = workloadAcc(config, rows, seed=55)
    if config == "c3":
        print("gpu case: %s undersized queue" % plan, flush=True)
        got, ny, retries = c3Run(acc, tuning=t, plan=plan)
        want, wantNy = c3Oracle(acc)
        assertSameResult(got, want, {"revenue"}, label="%s retry by item" % plan)
        assertSameResult(ny, wantNy, {"revenue"}, label="%s retry by nation, year" % plan)
    else:
        _, retries = check(plan, acc, tuning=t, label="undersized queue")
    assert retries > 0
