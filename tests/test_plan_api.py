"""Plan API (drop-in boundary): opId numbering, attribute resolution, pruning,
cfg semantics, and the hash-table / result sizes of TPC-H Q10 against the
numbers of the reference's generated sample program (sample/tpch10.cu:
jht6 = 50, jht5 = 150000 buckets + 150000 entries, jht9 = 300000,
aht11 = 138748, result 69374) and profiler reports (edge estimates
25 / 1.5 M / 75 K / 150 K / 2.0 M / 69.4 K)."""
import ctypes
import json

import numpy as np
import pytest

import dogqc_amd.scalarAlgebra as scal
from dogqc_amd import tpch, workloads, runtime as rt
from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction, Join
from dogqc_amd.types import Type
from dogqc_amd.hipTranslator import HipCompiler
from dogqc_amd.hiplang import CType
from dogqc_amd.util import loadScript


def q10Resolved():
    acc = tpch.SchemaAccessor(tpch.schema())
    alg, plan, cfg = tpch.q10(acc)
    return alg, alg.resolveAlgebraPlan(plan, cfg)


def nodes(root):
    out = {}

    def walk(n):
        out[n.opId] = n
        for c in n.children():
            walk(c)
    walk(root)
    return out


def test_q10_opids_and_estimates():
    alg, plan = q10Resolved()
    byId = nodes(plan[-1])
    assert sorted(byId) == list(range(1, 13))
    kinds = {i: type(n).__name__ for i, n in byId.items()}
    assert kinds == {1: "Scan", 2: "Scan", 3: "Selection", 4: "Scan", 5: "EquiJoin", 6: "EquiJoin", 7: "Scan",
                     8: "Selection", 9: "EquiJoin", 10: "Map", 11: "Aggregation", 12: "Materialize"}
    est = {i: n.tupleNum for i, n in byId.items()}
    assert est[1] == 25 and est[2] == 1500000 and est[3] == 75000 and est[5] == 150000
    assert est[6] == 150000 and est[8] == 2040413 and est[9] == 69374 and est[11] == 69374 and est[12] == 69374
    assert byId[5].multimatch is True and byId[6].multimatch is False and byId[9].multimatch is False
    assert byId[5].htSizeFactor == 2.0


def test_q10_generated_table_sizes():
    alg, plan = q10Resolved()
    comp = HipCompiler(alg, "sm_75", CType.FP64)
    comp.gencode(alg.translateToCompilerPlan(plan, comp))
    lib = ctypes.CDLL(comp.compile("tpch10"))
    lib.dqc_plan_describe.restype = ctypes.c_char_p
    desc = json.loads(lib.dqc_plan_describe().decode())
    rows = np.array([tpch.SF1_ROWS[t] for t in desc["tables"]], dtype=np.int64)
    cols = (ctypes.c_void_p * len(desc["columns"]))()
    inp = rt.Inputs(len(rows), len(desc["columns"]), rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                    ctypes.cast(cols, ctypes.POINTER(ctypes.c_void_p)))
    lay = rt.Layout()
    assert lib.dqc_plan_layout_of(ctypes.byref(inp), ctypes.byref(rt.Tuning()), ctypes.byref(lay)) == 0
    sizes = {h["label"]: (lay.ht_capacity[h["index"]], lay.mm_entries[h["index"]], h["kind"]) for h in desc["hts"]}
    assert sizes["jht6"] == (50, 0, "unique")
    assert sizes["jht5"] == (150000, 150000, "multi")
    assert sizes["jht9"] == (300000, 0, "unique")
    assert sizes["aht11"][0] == 138748
    assert lay.result_capacity == 69374
    assert desc["steps"][0] == "dqc_init_kernel"
    # kernel decomposition: nation build, orders count/scan/insert, customer,
    # lineitem, aggregation scan (the nation and customer builds store
    # strings, so neither is a partitioned build)
    assert [s for s in desc["steps"] if s.startswith("krnl")] == [
        "krnl_nation1", "krnl_orders2<0>", "krnl_orders2<1>", "krnl_customer4", "krnl_lineitem7",
        "krnl_aggregation11"]
    assert "dqc_spill_insert" not in desc["steps"]


def planFileResolved():
    """The reference's driver flow over a plan file and its cfg file
    (query/tpch.py:46-58): eval(loadScript(path)) with alg, scal,
    Reduction and Type in scope, then resolveAlgebraPlan."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans")
    acc = tpch.SchemaAccessor(tpch.schema())
    alg = RelationalAlgebra(acc)
    scope = {"alg": alg, "scal": scal, "Reduction": Reduction, "Type": Type}
    plan = eval(loadScript(os.path.join(here, "tpch10.py")), scope)
    cfg = eval(loadScript(os.path.join(here, "tpch10cfg.py")), scope)
    return alg, alg.resolveAlgebraPlan(plan, cfg)


def test_plan_file_flow_hits_q10_fixture():
    """A plan file and a cfg file in the reference's format evaluate to the
    Q10 plan of the reference's sample run: opIds 1-12, its edge estimates,
    and the generated module's table sizes (sample/tpch10.cu: jht6 = 50,
    jht5 = 150000 + 150000, jht9 = 300000, aht11 = 138748, result 69374)."""
    alg, plan = planFileResolved()
    byId = nodes(plan[-1])
    assert sorted(byId) == list(range(1, 13))
    est = {i: n.tupleNum for i, n in byId.items()}
    assert (est[1], est[2], est[3], est[5], est[6], est[8], est[9], est[12]) == (
        25, 1500000, 75000, 150000, 150000, 2040413, 69374, 69374)
    comp = HipCompiler(algebraContext=alg, smArchitecture="sm_75", decimalRepr=CType.FP64, debug=False)
    comp.gencode(alg.translateToCompilerPlan(plan, comp))
    lib = ctypes.CDLL(comp.compile("tpch10"))
    lib.dqc_plan_describe.restype = ctypes.c_char_p
    desc = json.loads(lib.dqc_plan_describe().decode())
    rows = np.array([tpch.SF1_ROWS[t] for t in desc["tables"]], dtype=np.int64)
    cols = (ctypes.c_void_p * len(desc["columns"]))()
    inp = rt.Inputs(len(rows), len(desc["columns"]), rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                    ctypes.cast(cols, ctypes.POINTER(ctypes.c_void_p)))
    lay = rt.Layout()
    assert lib.dqc_plan_layout_of(ctypes.byref(inp), ctypes.byref(rt.Tuning()), ctypes.byref(lay)) == 0
    sizes = {h["label"]: (lay.ht_capacity[h["index"]], lay.mm_entries[h["index"]]) for h in desc["hts"]}
    assert (sizes["jht6"], sizes["jht5"], sizes["jht9"], sizes["aht11"][0], lay.result_capacity) == (
        (50, 0), (150000, 150000), (300000, 0), 138748, 69374)


def test_opids_follow_construction_order():
    acc = workloads.emptyAccessor("c0")
    alg, p, cfg = workloads.plan("c0", acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    byId = nodes(plan[-1])
    assert [type(byId[i]).__name__ for i in range(1, 7)] == ["Scan", "Selection", "Map", "Map", "Aggregation",
                                                             "Materialize"]


def test_avg_adds_count_and_types():
    acc = workloads.emptyAccessor("c0")
    alg = RelationalAlgebra(acc)
    aggs = [(Reduction.AVG, "qty", "a")]
    node = alg.aggregation(["g1"], aggs, alg.scan("cfg0"))
    assert [a[2] for a in node.aggregates] == ["a", "count_agg"]
    types = sorted((a.name, a.dataType) for a in node.aggregateAttributes.values())
    assert types == [("a", Type.DOUBLE), ("count_agg", Type.INT)]


def test_identifier_errors():
    acc = tpch.SchemaAccessor(tpch.schema())
    alg = RelationalAlgebra(acc)
    j = alg.join(("o_custkey", "c_custkey"), alg.scan("orders"), alg.scan("customer"))
    bad = alg.selection(scal.EqualsExpr(scal.AttrExpr("no_such"), scal.ConstExpr("1", Type.INT)), j)
    with pytest.raises(NameError):
        alg.resolveAlgebraPlan(bad, {})
    alg = RelationalAlgebra(acc)
    self_join = alg.join(("o_orderkey", "orders.o_orderkey"), alg.scan("orders"), alg.scan("orders", "o2"))
    with pytest.raises(SyntaxError):
        alg.resolveAlgebraPlan(self_join, {})


def test_qualified_and_alias_identifiers():
    acc = tpch.SchemaAccessor(tpch.schema())
    alg = RelationalAlgebra(acc)
    j = alg.join(("o1.o_orderkey", "o2.o_custkey"), alg.scan("orders", "o1"), alg.scan("orders", "o2"))
    plan = alg.resolveAlgebraPlan(alg.projection(["o2.o_orderkey"], j), {})
    assert [a.name for a in plan[-1].outRelation.values()] == ["o_orderkey"]


def test_equals_type_check():
    acc = workloads.emptyAccessor("c1")
    alg = RelationalAlgebra(acc)
    bad = alg.selection(scal.EqualsExpr(scal.AttrExpr("f_price"), scal.ConstExpr("1", Type.INT)), alg.scan("fact"))
    with pytest.raises(SyntaxError):
        alg.resolveAlgebraPlan(bad, {})


def test_pruning_keeps_only_used_columns():
    acc = workloads.emptyAccessor("c1")
    alg, p, cfg = workloads.plan("c1", acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    byId = nodes(plan[-1])
    assert sorted(a.name for a in byId[1].outRelation.values()) == ["d_cat", "d_key"]   # d_seg unused
    assert sorted(a.name for a in byId[3].outRelation.values()) == ["f_date", "f_disc", "f_fk", "f_price"]
    assert sorted(a.name for a in byId[5].outRelation.values()) == ["d_key", "f_disc", "f_price"]


def test_cfg_defaults():
    acc = workloads.emptyAccessor("c2")
    alg, p, _ = workloads.plan("c2", acc)
    plan = alg.resolveAlgebraPlan(p, {})
    byId = nodes(plan[-1])
    assert byId[3].selectivity == 1.0
    assert byId[4].multimatch is True and byId[4].selectivity == 1.0 and byId[4].htSizeFactor == 2.0
    assert byId[6].numgroups == 1 and byId[6].doGroup is False


def test_loadscript_strips_comments(tmp_path):
    f = tmp_path / "plan.py"
    f.write_text("# header\n\nalg.scan ( 'x' )\n   # indented comment\n")
    assert loadScript(str(f)).strip() == "alg.scan ( 'x' )"
    g = tmp_path / "empty.py"
    g.write_text("# nothing\n")
    assert loadScript(str(g)) == "{}"


def test_compiler_signature_accepts_reference_arguments():
    acc = workloads.emptyAccessor("c0")
    alg, p, cfg = workloads.plan("c0", acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    comp = HipCompiler(algebraContext=alg, smArchitecture=None, decimalRepr=CType.FP64, debug=False)
    comp.setBuffers([2])
    comp.setProfilers([2])
    src = comp.gencode(alg.translateToCompilerPlan(plan, comp))
    assert "krnl_cfg01" in src
    with pytest.raises(ValueError):
        HipCompiler(alg, "gfx942")


def test_generated_source_is_deterministic():
    from dogqc_amd import planlib
    a = planlib.translate("c2_both").source
    b = planlib.translate("c2_both").source
    assert a == b
