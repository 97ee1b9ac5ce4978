"""dqc_query_run_noalloc, the plan run in the runtime's static workspace,
against dqc_query_run on the same inputs, through the C ABI.

Each plan runs twice in a row through the allocation-free entry point, the
second time on a smaller input, so the second run starts from a workspace
holding the first run's tables.  Both results must equal dqc_query_run's
(keys, counts and strings exact; FP64 sums within 1e-9 relative, since the
two runs sum in different orders), and the plans' own oracle results.
"""
import ctypes as C
from collections import OrderedDict

import numpy as np
import pytest

from dogqc_amd import planlib, workloads, runtime as rt, runner, tpch
from dogqc_amd.dbio import MemoryAccessor
from oracle.planeval import evaluate
from tests.helpers import assertSameResult, toleranceColumns

pytestmark = pytest.mark.gpu

_MODULES = {}


def module(name):
    if name not in _MODULES:
        _MODULES[name] = rt.PlanModule(planlib.translate(name).compile(name))
    return _MODULES[name]


def noallocColumns(mod, inputs, acc, outRows, tuning=None):
    """Result of dqc_query_run_noalloc, decoded like runner.resultColumns."""
    dts = [rt.CTYPE_DTYPE[c["ctype"]] for c in mod.desc["result"]]
    outs = [rt.DeviceBuffer(max(1, outRows) * np.dtype(dt).itemsize) for dt in dts]
    res = rt.runNoAlloc(mod, inputs, outs, outRows, tuning or rt.newTuning())
    assert res.ws is None, "the allocation-free run reports no workspace to release"
    n = int(res.status.result_count)
    byIndex = {c["index"]: c for c in mod.desc["columns"]}
    got = OrderedDict()
    for c, dt, buf in zip(mod.desc["result"], dts, outs):
        arr = buf.toArray(dt, n)
        if c["ctype"] == "dqc::str_offs":
            src = byIndex[c["char_col"]]
            arr = runner.decodeStrings(arr, acc.column(src["table"], src["attr"])[1])
        got[c["name"]] = arr
    return got, res


def allocatingRun(mod, inputs, acc):
    ex = rt.run(mod, inputs, rt.newTuning())
    want = runner.resultColumns(mod, ex, acc)
    ex.release()
    return want


def oracleOf(name, acc):
    alg, p, cfg, _ = planlib.build(name, acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    return evaluate(plan, acc), toleranceColumns(plan)


WORKLOAD_CASES = [
    ("c1_refill", "c1", {"dim": 40000, "fact": 160000}, {"dim": 7000, "fact": 25000}),
    ("c2_both", "c2", {"build": 60000, "probe": 500000}, {"build": 9000, "probe": 70000}),
    ("c0", "c0", {"cfg0": 300000}, {"cfg0": 1000}),
    ("c3_refill", "c3", {"sales": 200000, "item": 20000, "cust": 2000, "nation": 25},
     {"sales": 30000, "item": 3000, "cust": 500, "nation": 25}),
]


@pytest.mark.parametrize("name,config,big,small", WORKLOAD_CASES, ids=[c[0] for c in WORKLOAD_CASES])
def test_noalloc_matches_allocating_run(name, config, big, small):
    mod = module(name)
    for label, rows in (("first", big), ("second, smaller", small)):
        print("gpu case: noalloc %s %s %s" % (name, label, rows), flush=True)
        acc = MemoryAccessor(workloads.generate(config, rows))
        inputs = rt.BoundInputs(mod, acc)
        lay = rt.layout(mod, inputs, rt.newTuning())
        assert lay.ws_bytes <= rt.NOALLOC_WS_BYTES
        want = allocatingRun(mod, inputs, acc)
        n = len(next(iter(want.values())))
        got, res = noallocColumns(mod, inputs, acc, n)
        oracle, tol = oracleOf(name, acc)
        assertSameResult(got, want, tol, label="%s %s vs dqc_query_run" % (name, label))
        assertSameResult(got, oracle, tol, label="%s %s vs oracle" % (name, label))


def test_noalloc_strings_two_calls():
    """TPC-H Q10 (string result columns, 16-byte offsets) at two sizes."""
    mod = module("tpch10")
    for sf in (0.01, 0.002):
        print("gpu case: noalloc tpch10 sf %g" % sf, flush=True)
        acc = MemoryAccessor(tpch.generate(sf, seed=7))
        inputs = rt.BoundInputs(mod, acc)
        want = allocatingRun(mod, inputs, acc)
        n = len(next(iter(want.values())))
        got, _ = noallocColumns(mod, inputs, acc, n + 5)
        oracle, tol = oracleOf("tpch10", acc)
        assertSameResult(got, want, tol, label="tpch10 sf %g vs dqc_query_run" % sf)
        assertSameResult(got, oracle, tol, label="tpch10 sf %g vs oracle" % sf)


def test_noalloc_rejects_oversized_workspace_and_short_outputs():
    """A workspace beyond the static scratch is refused before anything is
    enqueued; a result longer than the output columns is refused before the
    copy.  A run afterwards still gives the right result."""
    mod = module("c2_both")
    desc = mod.desc
    print("gpu case: noalloc c2_both oversized layout (host check only)", flush=True)
    rows = np.array([2000000, 32000000], dtype=np.int64)  # build, probe: a ~84 MB workspace
    assert desc["tables"] == ["build", "probe"]
    ptrs = (C.c_void_p * len(desc["columns"]))()
    inputs = rt.Inputs(2, len(desc["columns"]), rows.ctypes.data_as(C.POINTER(C.c_int64)),
                       C.cast(ptrs, C.POINTER(C.c_void_p)))

    class Bound(object):
        struct = inputs
    lay = rt.Layout()
    rt.check(rt.lib().dqc_query_layout(mod.handle, C.byref(inputs), C.byref(rt.newTuning()), C.byref(lay)), "layout")
    assert lay.ws_bytes > rt.NOALLOC_WS_BYTES
    outs = [rt.DeviceBuffer(64) for _ in desc["result"]]
    with pytest.raises(rt.DqcError, match="static scratch"):
        rt.runNoAlloc(mod, Bound(), outs, 1)

    print("gpu case: noalloc c1 short output columns", flush=True)
    mod1 = module("c1")
    acc = MemoryAccessor(workloads.generate("c1", {"dim": 20000, "fact": 80000}))
    inputs1 = rt.BoundInputs(mod1, acc)
    want = allocatingRun(mod1, inputs1, acc)
    n = len(want["d_key"])
    assert n > 1
    with pytest.raises(rt.DqcError, match="output columns hold"):
        noallocColumns(mod1, inputs1, acc, n - 1)
    got, _ = noallocColumns(mod1, inputs1, acc, n)
    oracle, tol = oracleOf("c1", acc)
    assertSameResult(got, want, tol, label="c1 after the refusals")
    assertSameResult(got, oracle, tol, label="c1 after the refusals vs oracle")
