"""Pins the CPU oracle (oracle/planeval.py) before it is trusted as the
checker of the HIP path:

  * against a row-at-a-time brute-force evaluator (nested-loop joins,
    dictionary group-by, exact Fraction sums) on tiny inputs of every plan;
  * its scalar arithmetic against gcc compiling the same C expressions
    (the reference's generated code is C, so C decides the semantics);
  * against closed forms: per-group counts and integer sums by bincount, the
    expanding join's COUNT and SUM(b*p) factorised per key, the star join's
    two aggregations by inverse-permutation indexing;
  * against the independent multithreaded C oracle at full workload size;
  * a mistake planted in the oracle (floor instead of truncating division)
    must fail the gcc check.
"""
import os
import subprocess
import tempfile
from collections import OrderedDict

import numpy as np
import pytest

import dogqc_amd.scalarAlgebra as scal
from dogqc_amd import planlib, workloads
from dogqc_amd.dbio import MemoryAccessor
from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction
from dogqc_amd.types import Type
from oracle import planeval, bruteforce, coracle
from tests.helpers import assertSameResult, toleranceColumns

SMALL = {"c0": {"cfg0": 400}, "c1": {"dim": 120, "fact": 900}, "c2": {"build": 150, "probe": 3000},
         "c3": {"sales": 1500, "item": 300, "cust": 40, "nation": 5}}


_TPCH = []


def tpchAccessor():
    """Synthetic TPC-H at scale factor 0.001 (6000 lineitem rows)."""
    if not _TPCH:
        from dogqc_amd import tpch
        _TPCH.append(MemoryAccessor(tpch.generate(0.001, seed=3)))
    return _TPCH[0]


def resolved(name, acc):
    alg, p, cfg, _ = planlib.build(name, acc)
    return alg.resolveAlgebraPlan(p, cfg)


@pytest.mark.parametrize("name", [n for n in planlib.PLANS if planlib.PLANS[n][0] != "cal"])
def test_planeval_matches_bruteforce(name):
    kind = planlib.PLANS[name][0]
    if kind == "feature":
        acc = planlib.featureData(160, 40, seed=3)
    elif kind == "strings":
        acc = planlib.stringRows(planlib.stringData(1), 120)
    elif kind == "tpch":
        acc = tpchAccessor()
    elif kind == "batch":
        from dogqc_amd import batching
        config = planlib.PLANS[name][1]["config"]
        full = MemoryAccessor(workloads.generate(config, SMALL[config], seed=5))
        pref = {t: [max(1, SMALL[config][t] // 3), SMALL[config][t]] for t in batching.SCALED[config]}
        acc = batching.batchAccessor(full, pref)
    else:
        acc = MemoryAccessor(workloads.generate(kind, SMALL[kind], seed=5))
    plan = resolved(name, acc)
    want = bruteforce.evaluate(plan, acc)
    plan = resolved(name, acc)
    got = planeval.evaluate(plan, acc)
    # brute force sums exactly (Fractions) and rounds once, like math.fsum
    assertSameResult(got, want, toleranceColumns(plan), rtol=0, atol=0, label=name)


# ---------------------------------------------------------------- C semantics

def _scalarTable():
    t = {"i": (Type.INT, np.array([7, -7, 2147483647, -9, 0], dtype=np.int32)),
         "j": (Type.INT, np.array([2, 2, 1, 4, 3], dtype=np.int32)),
         "d": (Type.DATE, np.array([19950315, 19981231, 19920101, 20000229, 19960704], dtype=np.uint32)),
         "c": (Type.CHAR, np.array([82, 65, -1, 0, 127], dtype=np.int8)),
         "f": (Type.FLOAT, np.array([1.1, -2.5, 3.25, 1e-3, 7.0], dtype=np.float32)),
         "x": (Type.DOUBLE, np.array([0.1, 2.675, -1e10, 3.0, 1.0 / 3.0])),
         "b": (Type.INT64, np.array([3, -5, 1 << 40, -(1 << 41), 9], dtype=np.int64))}
    from collections import OrderedDict
    return MemoryAccessor(OrderedDict([("s", OrderedDict(t.items()))]))


A, K = scal.AttrExpr, scal.ConstExpr
CASES = [
    ("div_int", scal.DivExpr(A("i"), A("j"))),
    ("add_wrap", scal.AddExpr(A("i"), K("1", Type.INT))),
    ("mix_int_double", scal.MulExpr(A("i"), A("x"))),
    ("float_ops", scal.MulExpr(A("f"), scal.SubExpr(K("1.1f", Type.FLOAT), A("f")))),
    ("float_double", scal.AddExpr(A("f"), A("x"))),
    ("date_year", scal.ExtractExpr(A("d"), scal.ExtractType.YEAR)),
    ("date_month", scal.ExtractExpr(A("d"), scal.ExtractType.MONTH)),
    ("date_day", scal.ExtractExpr(A("d"), scal.ExtractType.DAY)),
    ("int64_mul", scal.MulExpr(A("b"), A("b"))),
    ("int64_div", scal.DivExpr(A("b"), K("7", Type.INT64))),
    ("date_cmp", scal.SmallerEqualExpr(A("d"), K("19960101", Type.DATE))),
    ("char_eq", scal.EqualsExpr(A("c"), K("R", Type.CHAR))),
    ("case", scal.CaseExpr([(scal.LargerExpr(A("x"), K("1.0", Type.DOUBLE)), A("i"))], A("f"))),
    ("neg_div_double", scal.DivExpr(A("x"), A("i"))),
    ("in_int", scal.InExpr(A("i"), [K("7", Type.INT), K("0", Type.INT), K("-9", Type.INT)])),
    ("in_date", scal.InExpr(A("d"), [K("19950315", Type.DATE), K("20000229", Type.DATE)])),
    ("not_or", scal.NotExpr(scal.OrExpr(scal.SmallerExpr(A("j"), K("2", Type.INT)),
                                        scal.EqualsExpr(A("c"), K("A", Type.CHAR))))),
    ("and_not", scal.AndExpr(scal.LargerEqualExpr(A("b"), K("3", Type.INT64)),
                             scal.NotExpr(scal.EqualsExpr(A("i"), A("j"))))),
]

_CTYPE = {"int32": "int32_t", "uint32": "uint32_t", "int8": "int8_t", "int64": "int64_t", "float32": "float",
          "float64": "double", "bool": "int32_t"}


def _cExpr(e):
    """C text of a scalar expression in the reference's literal spelling
    (scalarAlgebra.py:142-154); attributes are register variables."""
    if isinstance(e, scal.AttrExpr):
        return e.identifier + "_r"
    if isinstance(e, scal.ConstExpr):
        return {Type.FLOAT: "((float)%s)" % e.token, Type.CHAR: "'%s'" % e.token}.get(e.type, "(%s)" % e.token)
    ops = {scal.MulExpr: "*", scal.AddExpr: "+", scal.SubExpr: "-", scal.DivExpr: "/", scal.SmallerEqualExpr: "<=",
           scal.LargerExpr: ">", scal.EqualsExpr: "==", scal.SmallerExpr: "<", scal.LargerEqualExpr: ">=",
           scal.AndExpr: "&&", scal.OrExpr: "||"}
    if isinstance(e, scal.NotExpr):
        return "(!%s)" % _cExpr(e.child)
    if type(e) in ops:
        return "(%s %s %s)" % (_cExpr(e.leftChild), ops[type(e)], _cExpr(e.rightChild))
    if isinstance(e, scal.ExtractExpr):
        c = _cExpr(e.child)
        return {scal.ExtractType.YEAR: "(%s / 10000)" % c, scal.ExtractType.MONTH: "((%s / 100) %% 100)" % c,
                scal.ExtractType.DAY: "(%s %% 100)" % c}[e.extract]
    if isinstance(e, scal.CaseExpr):
        (w, t), = e.exprListWhenThen
        return "(%s ? (double)%s : (double)%s)" % (_cExpr(w), _cExpr(t), _cExpr(e.exprElse))
    raise NotImplementedError(type(e).__name__)


def _gccValues(acc, cases):
    """Compile the expressions with gcc (-ffp-contract=off, -fwrapv) and read
    back their values as bit patterns, per row."""
    s = acc.schema["s"]
    n = s["size"]
    lines = ["#include <stdio.h>", "#include <stdint.h>", "#include <string.h>", "int main(void) {"]
    for a, typ in s["attributes"].items():
        d = acc.column("s", a)
        ct = {Type.INT: "int32_t", Type.DATE: "uint32_t", Type.CHAR: "int8_t", Type.FLOAT: "float",
              Type.DOUBLE: "double", Type.INT64: "int64_t"}[typ]
        reg = "double" if typ in (Type.FLOAT, Type.DOUBLE) else ct
        vals = ", ".join(repr(float(v)) if typ in (Type.FLOAT, Type.DOUBLE) else str(int(v)) for v in d)
        if typ == Type.FLOAT:
            vals = ", ".join("%sf" % float.hex(float(v)) for v in d)
        if typ == Type.DOUBLE:
            vals = ", ".join(float.hex(float(v)) for v in d)
        if typ == Type.INT64:
            vals = ", ".join("%dLL" % int(v) for v in d)
        lines.append("  static const %s %s_col[%d] = {%s};" % (ct, a, n, vals))
        lines.append("  %s %s_r;" % (reg, a))
    lines.append("  for (int r = 0; r < %d; ++r) {" % n)
    for a in s["attributes"]:
        lines.append("    %s_r = %s_col[r];" % (a, a))
    for name, e, ct in cases:
        lines.append("    { %s v = (%s)%s; unsigned long long bits = 0; memcpy(&bits, &v, sizeof(v));"
                     " printf(\"%s %%d %%llu\\n\", r, bits); }" % (_CTYPE[ct], _CTYPE[ct], _cExpr(e), name))
    lines.append("  }")
    lines.append("  return 0;")
    lines.append("}")
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "x.c")
        exe = os.path.join(td, "x")
        with open(src, "w") as f:
            f.write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-O0", "-std=c11", "-ffp-contract=off", "-fwrapv", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout
    vals = {}
    for line in out.splitlines():
        name, r, bits = line.split()
        vals[(name, int(r))] = int(bits)
    return vals


def _oracleValues(acc, cases, evaluator=None):
    alg = RelationalAlgebra(acc)
    scan = alg.scan("s")
    node = scan
    for name, e in cases:
        node = alg.map(name, e, node)
    plan = alg.resolveAlgebraPlan(node, {})
    rel, n = (evaluator or planeval.Evaluator(acc)).node(plan[-1])
    out, types = {}, {}
    m = plan[-1].child
    while hasattr(m, "mapAttr"):
        arr, ct = (evaluator or planeval.Evaluator(acc)).expr(m.expression, rel, n)
        types[m.mapAttr.name] = ct
        arr = np.broadcast_to(np.asarray(arr), (n,))
        for r in range(n):
            v = np.asarray(arr[r]).astype(ct if ct != "bool" else "int32")
            out[(m.mapAttr.name, r)] = int.from_bytes(v.tobytes(), "little")
        m = m.child
    return out, types


def test_scalar_semantics_match_gcc():
    acc = _scalarTable()
    got, types = _oracleValues(acc, CASES)
    want = _gccValues(acc, [(n, e, types[n]) for n, e in CASES])
    assert got == want


def test_planted_division_mistake_is_caught(monkeypatch):
    """An oracle that floors integer division (Python's //) must disagree
    with C's truncation on negative operands."""
    acc = _scalarTable()
    cases = [("div_int", scal.DivExpr(A("i"), A("j")))]

    class Floor(planeval.Evaluator):
        def expr(self, e, rel, n):
            if isinstance(e, scal.DivExpr):
                a, at = self.expr(e.leftChild, rel, n)
                b, bt = self.expr(e.rightChild, rel, n)
                return (np.asarray(a).astype(np.int64) // np.asarray(b).astype(np.int64)).astype(np.int32), "int32"
            return planeval.Evaluator.expr(self, e, rel, n)

    got, types = _oracleValues(acc, cases, Floor(acc))
    want = _gccValues(acc, [(n, e, types[n]) for n, e in cases])
    assert got != want


# --------------------------------------------------------------- closed forms

def test_c0_counts_and_integer_sums_closed_form():
    g = workloads.generate("c0", {"cfg0": 50000}, seed=8)
    acc = MemoryAccessor(g)
    res = planeval.evaluate(resolved("c0", acc), acc)
    cols = {k: v[1] for k, v in g["cfg0"].items()}
    keep = cols["date"] <= 19981110
    key = cols["g1"][keep].astype(np.int64) * 256 + cols["g2"][keep].astype(np.int64)
    cnt = np.bincount(key, minlength=3 * 256)
    qty = np.bincount(key, weights=cols["qty"][keep].astype(np.float64), minlength=3 * 256)
    for i in range(len(res["g1"])):
        k = int(res["g1"][i]) * 256 + int(res["g2"][i])
        assert res["count_order"][i] == cnt[k]
        assert res["sum_qty"][i] == qty[k]          # integers below 2^53: exact in any order
        assert res["avg_qty"][i] == qty[k] / cnt[k]
    assert len(res["g1"]) == int((cnt > 0).sum()) == 6


def test_c2_closed_form():
    g = workloads.generate("c2", {"build": 3000, "probe": 60000}, seed=9)
    acc = MemoryAccessor(g)
    res = planeval.evaluate(resolved("c2_none", acc), acc)
    from tests.test_gpu_parity import c2ClosedForm
    cnt, s = c2ClosedForm(acc)
    assert int(res["cnt"][0]) == cnt and int(res["sum_prod"][0]) == s


def c3ClosedForm(g):
    """Both c3 aggregations by direct indexing (every dimension key is a
    permutation of 0..N-1, so key -> row is an inverse permutation) and
    bincount; the LIKE test as a Python substring test."""
    it, cu, na, sa = (g[t] for t in ("item", "cust", "nation", "sales"))
    offs, chars = it["i_text"][1]
    b = chars.tobytes()
    word = workloads.C3_WORD.encode()
    passing = np.array([word in b[int(offs[i]):int(offs[i + 1])] for i in range(len(offs) - 1)], dtype=bool)
    passByKey = np.zeros(len(passing), dtype=bool)
    passByKey[it["i_key"][1]] = passing
    yearByKey = np.zeros(len(cu["c_key"][1]), dtype=np.int64)
    yearByKey[cu["c_key"][1]] = cu["c_year"][1]
    si, sc, sn = (sa[k][1].astype(np.int64) for k in ("s_item", "s_cust", "s_nation"))
    keep = passByKey[si]
    qty, price = sa["s_qty"][1][keep].astype(np.float64), sa["s_price"][1][keep]
    si, sc, sn = si[keep], sc[keep], sn[keep]
    cntItem = np.bincount(si, minlength=len(passing))
    qtyItem = np.bincount(si, weights=qty, minlength=len(passing))
    ny = sn * 10000 + yearByKey[sc]
    keys, inv = np.unique(ny, return_inverse=True)
    return (passing, cntItem, qtyItem, keys, np.bincount(inv), np.bincount(inv, weights=qty),
            np.bincount(si, weights=price, minlength=len(passing)))


def test_c3_closed_form():
    g = workloads.generate("c3", {"sales": 40000, "item": 4000, "cust": 700, "nation": 25}, seed=10)
    acc = MemoryAccessor(g)
    passing, cntItem, qtyItem, nyKeys, nyCnt, nyQty, revItem = c3ClosedForm(g)
    assert 0.03 < passing.mean() < 0.08      # 1 - (91/92)^5 = 5.3 %
    alg, p, cfg, _ = planlib.build("c3", acc)
    ev = planeval.Evaluator(acc)
    res = ev.run(alg.resolveAlgebraPlan(p, cfg))
    k = res["i_key"].astype(np.int64)
    assert np.array_equal(np.sort(k), np.nonzero(cntItem)[0])
    assert np.array_equal(res["cnt"], cntItem[k]) and np.array_equal(res["sum_qty"], qtyItem[k])
    assert np.allclose(res["revenue"], revItem[k], rtol=1e-12, atol=0)
    ny, _ = ev.temps["c3_by_nation_year"]
    got = ny["n_key"].astype(np.int64) * 10000 + ny["c_year"]
    o = np.argsort(got)
    assert np.array_equal(got[o], nyKeys)
    assert np.array_equal(ny["cnt"][o], nyCnt) and np.array_equal(ny["sum_qty"][o], nyQty)
    byItem, byNy = coracle.c3({t: {a: v[1] for a, v in c.items()} for t, c in g.items()}, workloads.C3_WORD)
    assertSameResult(byItem, res, {"revenue"}, rtol=1e-12, label="C oracle by item")
    assertSameResult(byNy, OrderedDict((c, ny[c]) for c in byNy), {"revenue"}, rtol=1e-12, label="C oracle ny")


def test_c3_text_column_layout():
    assert len(workloads.VOCAB) == len(set(workloads.VOCAB)) == 92
    assert not [(a, b) for a in workloads.VOCAB for b in workloads.VOCAB if a != b and a in b]
    rng = np.random.Generator(np.random.PCG64(3))
    idx = rng.integers(0, 92, size=(500, 5))
    offs, chars = workloads.fixedWidthText(idx)
    b = chars.tobytes()
    assert np.array_equal(offs, np.arange(501, dtype=np.uint64) * 55) and len(b) == 500 * 55
    for i in range(500):
        row = b[55 * i:55 * (i + 1)].decode()
        assert row.rstrip(" ").split(" ") == [workloads.VOCAB[j] for j in idx[i]]


def test_empty_no_group_aggregate_yields_one_row():
    g = workloads.generate("c2", {"build": 10, "probe": 0}, seed=1)
    acc = MemoryAccessor(g)
    res = planeval.evaluate(resolved("c2_none", acc), acc)
    assert int(res["cnt"][0]) == 0 and int(res["sum_prod"][0]) == 0


# ------------------------------------------------------- C oracle agreement

def test_c_oracle_agrees_at_full_size():
    g = workloads.generate("c1")
    acc = MemoryAccessor(g)
    want = planeval.evaluate(resolved("c1", acc), acc)
    keys, sums = coracle.c1({k: v[1] for k, v in g["dim"].items()}, {k: v[1] for k, v in g["fact"].items()})
    o = np.argsort(keys)
    w = np.argsort(want["d_key"])
    assert np.array_equal(keys[o], want["d_key"][w])
    assert np.allclose(sums[o], want["revenue"][w], rtol=1e-12, atol=0)

    g2 = workloads.generate("c2", {"build": 200000, "probe": 2000000})
    cnt, s = coracle.c2({k: v[1] for k, v in g2["build"].items()}, {k: v[1] for k, v in g2["probe"].items()})
    from tests.test_gpu_parity import c2ClosedForm
    assert (cnt, s) == c2ClosedForm(MemoryAccessor(g2))

    g0 = workloads.generate("c0", {"cfg0": 300000})
    acc0 = MemoryAccessor(g0)
    res = planeval.evaluate(resolved("c0", acc0), acc0)
    ref = coracle.c0({k: v[1] for k, v in g0["cfg0"].items()})
    for i in range(len(res["g1"])):
        r = ref[(int(res["g1"][i]), int(res["g2"][i]))]
        assert int(res["count_order"][i]) == r[5]
        np.testing.assert_allclose([res["sum_qty"][i], res["sum_base_price"][i], res["sum_disc_price"][i],
                                    res["sum_charge"][i]], r[:4], rtol=1e-12)


@pytest.mark.parametrize("sf", [0.0005, 0.002, 0.05])
def test_tpch_generator_keys(sf):
    """The synthetic TPC-H keeps the keys the plans mark unique unique, and
    the foreign keys inside their referenced tables."""
    from dogqc_amd import tpch
    T = tpch.generate(sf, seed=3)
    col = lambda t, a: T[t][a][1]
    for t, k in (("customer", "c_custkey"), ("orders", "o_orderkey"), ("part", "p_partkey"),
                 ("supplier", "s_suppkey"), ("nation", "n_nationkey"), ("region", "r_regionkey")):
        assert len(np.unique(col(t, k))) == len(col(t, k)), t
    pairs = col("partsupp", "ps_partkey").astype(np.int64) << 32 | col("partsupp", "ps_suppkey")
    assert len(np.unique(pairs)) == len(pairs)
    lpairs = col("lineitem", "l_partkey").astype(np.int64) << 32 | col("lineitem", "l_suppkey")
    assert np.isin(lpairs, pairs).all()
    assert np.isin(col("lineitem", "l_orderkey"), col("orders", "o_orderkey")).all()
    assert np.isin(col("orders", "o_custkey"), col("customer", "c_custkey")).all()
    assert (col("orders", "o_custkey") % 3 != 0).all()
    assert (col("lineitem", "l_receiptdate") > col("lineitem", "l_shipdate")).all()

