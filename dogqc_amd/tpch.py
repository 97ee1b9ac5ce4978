"""TPC-H schema in the plan API's table-dict format, and query 10 written
with the plan API.

The schema follows the TPC-H specification's column names and the type
choices of the reference's schema module (query/schema/tpch.py:23-122):
keys and counts INT, money FLOAT, flags CHAR, dates DATE as yyyymmdd,
text STRING.  Sizes are the scale-factor-1 row counts.
"""
from collections import OrderedDict

import numpy as np

import dogqc_amd.scalarAlgebra as scal
from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction
from dogqc_amd.types import Type

SF1_ROWS = {"lineitem": 6001215, "orders": 1500000, "customer": 150000, "part": 200000, "partsupp": 800000,
            "supplier": 10000, "nation": 25, "region": 5}

_T = Type
_COLUMNS = {
    "lineitem": [("l_orderkey", _T.INT), ("l_partkey", _T.INT), ("l_suppkey", _T.INT), ("l_linenumber", _T.INT),
                 ("l_quantity", _T.INT), ("l_extendedprice", _T.FLOAT), ("l_discount", _T.FLOAT),
                 ("l_tax", _T.FLOAT), ("l_returnflag", _T.CHAR), ("l_linestatus", _T.CHAR),
                 ("l_shipdate", _T.DATE), ("l_commitdate", _T.DATE), ("l_receiptdate", _T.DATE),
                 ("l_shipinstruct", _T.STRING), ("l_shipmode", _T.STRING), ("l_comment", _T.STRING)],
    "customer": [("c_custkey", _T.INT), ("c_name", _T.STRING), ("c_address", _T.STRING), ("c_nationkey", _T.INT),
                 ("c_phone", _T.STRING), ("c_acctbal", _T.FLOAT), ("c_mktsegment", _T.STRING),
                 ("c_comment", _T.STRING)],
    "orders": [("o_orderkey", _T.INT), ("o_custkey", _T.INT), ("o_orderstatus", _T.CHAR),
               ("o_totalprice", _T.FLOAT), ("o_orderdate", _T.DATE), ("o_orderpriority", _T.STRING),
               ("o_clerk", _T.STRING), ("o_shippriority", _T.INT), ("o_comment", _T.STRING)],
    "partsupp": [("ps_partkey", _T.INT), ("ps_suppkey", _T.INT), ("ps_availqty", _T.INT),
                 ("ps_supplycost", _T.FLOAT), ("ps_comment", _T.STRING)],
    "part": [("p_partkey", _T.INT), ("p_name", _T.STRING), ("p_mfgr", _T.STRING), ("p_brand", _T.STRING),
             ("p_type", _T.STRING), ("p_size", _T.INT), ("p_container", _T.STRING), ("p_retailprice", _T.FLOAT),
             ("p_comment", _T.STRING)],
    "supplier": [("s_suppkey", _T.INT), ("s_name", _T.STRING), ("s_address", _T.STRING), ("s_nationkey", _T.INT),
                 ("s_phone", _T.STRING), ("s_acctbal", _T.FLOAT), ("s_comment", _T.STRING)],
    "nation": [("n_nationkey", _T.INT), ("n_name", _T.STRING), ("n_regionkey", _T.INT), ("n_comment", _T.STRING)],
    "region": [("r_regionkey", _T.INT), ("r_name", _T.STRING), ("r_comment", _T.STRING)],
}


def schema(rows=None):
    rows = SF1_ROWS if rows is None else rows
    out = {}
    for name, cols in _COLUMNS.items():
        out[name] = {"name": name, "size": rows.get(name, 0), "numColumns": len(cols),
                     "attributes": OrderedDict(cols)}
    return out


class SchemaAccessor(object):
    """An accessor carrying only a schema (for translation and sizing)."""

    def __init__(self, schemaDict):
        self.schema = schemaDict


def q10(acc):
    """TPC-H Q10 (returned item reporting) with the cfg of the reference's
    sample run: opIds 1..11 in construction order, Result 12."""
    alg = RelationalAlgebra(acc)
    A, C = scal.AttrExpr, scal.ConstExpr
    plan = alg.aggregation(
        ["c_custkey", "c_name", "c_acctbal", "c_phone", "n_name", "c_address", "c_comment"],
        [(Reduction.SUM, "rev", "revenue")],
        alg.map("rev", scal.MulExpr(A("l_extendedprice"), scal.SubExpr(C("1.0f", Type.FLOAT), A("l_discount"))),
                alg.join(("l_orderkey", "o_orderkey"),
                         alg.join(("c_nationkey", "n_nationkey"),
                                  alg.scan("nation"),
                                  alg.join(("o_custkey", "c_custkey"),
                                           alg.selection(scal.AndExpr(
                                               scal.LargerEqualExpr(A("o_orderdate"), C("19931001", Type.DATE)),
                                               scal.SmallerExpr(A("o_orderdate"), C("19940101", Type.DATE))),
                                               alg.scan("orders")),
                                           alg.scan("customer"))),
                         alg.selection(scal.EqualsExpr(A("l_returnflag"), C("R", Type.CHAR)),
                                       alg.scan("lineitem")))))
    cfg = {3: {"selectivity": 0.05}, 8: {"selectivity": 0.34}, 6: {"multimatch": False},
           9: {"selectivity": 0.034, "multimatch": False}}
    return alg, plan, cfg



# ------------------------------------------------------------------ queries
#
# TPC-H Q1-Q22 written with the plan API from the specification's SQL
# (validation parameters), in the form of the reference's plan files
# (query/plans/tpch{1..22}.py: one expression over `alg`, `scal`,
# `Reduction`, `Type`).  As there, ORDER BY and LIMIT are not part of the
# plans (results are row multisets), correlated subqueries are decorrelated
# into joins/aggregations, and money is FLOAT.  Joins whose build side is a
# primary key are marked unique ("multimatch": False).


class _Q(object):
    """Builder with short names for the plan API; collects the cfg."""

    def __init__(self, acc):
        self.alg = RelationalAlgebra(acc)
        self.cfg = {}

    def pk(self, node):
        """Mark a join as unique on its build side."""
        self.cfg.setdefault(node.opId, {})["multimatch"] = False
        return node

    def hint(self, node, **kw):
        self.cfg.setdefault(node.opId, {}).update(kw)
        return node


A = scal.AttrExpr


def _c(tok, typ):
    return scal.ConstExpr(tok, typ)


def _d(yyyymmdd):
    return _c(str(yyyymmdd), Type.DATE)


def _s(text):
    return _c(text, Type.STRING)


def _i(v):
    return _c(str(v), Type.INT)


def _f(v):
    """A FLOAT constant (compared with FLOAT money columns exactly)."""
    return _c(v + "f", Type.FLOAT)


def _and(*es):
    out = es[-1]
    for e in reversed(es[:-1]):
        out = scal.AndExpr(e, out)
    return out


def _or(*es):
    out = es[-1]
    for e in reversed(es[:-1]):
        out = scal.OrExpr(e, out)
    return out


def _between(a, lo, hi):
    return scal.AndExpr(scal.LargerEqualExpr(a, lo), scal.SmallerEqualExpr(a, hi))


def _in(a, values):
    return scal.InExpr(a, values)


def _disc(price="l_extendedprice", disc="l_discount"):
    return scal.MulExpr(A(price), scal.SubExpr(_f("1.0"), A(disc)))


def _ne(a, b):
    return scal.NotExpr(scal.EqualsExpr(a, b))


def q1(acc):
    q = _Q(acc)
    alg = q.alg
    plan = alg.aggregation(
        ["l_returnflag", "l_linestatus"],
        [(Reduction.SUM, "l_quantity", "sum_qty"), (Reduction.SUM, "l_extendedprice", "sum_base_price"),
         (Reduction.SUM, "disc_price", "sum_disc_price"), (Reduction.SUM, "charge", "sum_charge"),
         (Reduction.AVG, "l_quantity", "avg_qty"), (Reduction.AVG, "l_extendedprice", "avg_price"),
         (Reduction.AVG, "l_discount", "avg_disc"), (Reduction.COUNT, "", "count_order")],
        alg.map("charge", scal.MulExpr(A("disc_price"), scal.AddExpr(_f("1.0"), A("l_tax"))),
                alg.map("disc_price", _disc(),
                        alg.selection(scal.SmallerEqualExpr(A("l_shipdate"), _d(19980902)),
                                      alg.scan("lineitem")))))
    # group-count hint of the reference's cfg (query/plans/tpch1cfg.py)
    q.hint(plan, numgroups=100)
    return alg, plan, q.cfg


def _regionNations(q, region, alias=None):
    """nations of a region: n_nationkey, n_name (prefixed by alias)."""
    alg = q.alg
    n = (alias + ".") if alias else ""
    return q.pk(alg.join(("r_regionkey", n + "n_regionkey"),
                         alg.selection(scal.EqualsExpr(A("r_name"), _s(region)), alg.scan("region")),
                         alg.scan("nation", alias)))


def q2(acc):
    q = _Q(acc)
    alg = q.alg

    def europeSuppliers(sAlias, nAlias, rAlias):
        n, s_ = nAlias + ".", sAlias + "."
        nations = q.pk(alg.join((rAlias + ".r_regionkey", n + "n_regionkey"),
                                alg.selection(scal.EqualsExpr(A(rAlias + ".r_name"), _s("EUROPE")),
                                              alg.scan("region", rAlias)),
                                alg.scan("nation", nAlias)))
        return q.pk(alg.join((n + "n_nationkey", s_ + "s_nationkey"), nations, alg.scan("supplier", sAlias)))

    minCost = alg.aggregation(
        ["ps1.ps_partkey"], [(Reduction.MIN, "ps1.ps_supplycost", "min_cost")],
        q.pk(alg.join(("s1.s_suppkey", "ps1.ps_suppkey"), europeSuppliers("s1", "n1", "r1"),
                      alg.scan("partsupp", "ps1"))))
    parts = alg.selection(_and(scal.EqualsExpr(A("p_size"), _i(15)),
                               scal.LikeExpr(A("p_type"), _s("%BRASS"))), alg.scan("part"))
    candidates = q.pk(alg.join(("p_partkey", "ps2.ps_partkey"), parts, alg.scan("partsupp", "ps2")))
    withSupp = q.pk(alg.join(("s2.s_suppkey", "ps2.ps_suppkey"), europeSuppliers("s2", "n2", "r2"), candidates))
    best = q.pk(alg.innerjoin([("ps1.ps_partkey", "ps2.ps_partkey")],
                              scal.EqualsExpr(A("ps2.ps_supplycost"), A("min_cost")), minCost, withSupp))
    plan = alg.projection(["s2.s_acctbal", "s2.s_name", "n2.n_name", "p_partkey", "p_mfgr", "s2.s_address",
                           "s2.s_phone", "s2.s_comment"], best)
    return alg, plan, q.cfg


def q3(acc):
    q = _Q(acc)
    alg = q.alg
    co = q.pk(alg.join(("c_custkey", "o_custkey"),
                       alg.selection(scal.EqualsExpr(A("c_mktsegment"), _s("BUILDING")), alg.scan("customer")),
                       alg.selection(scal.SmallerExpr(A("o_orderdate"), _d(19950315)), alg.scan("orders"))))
    col = q.pk(alg.join(("o_orderkey", "l_orderkey"), co,
                        alg.selection(scal.LargerExpr(A("l_shipdate"), _d(19950315)), alg.scan("lineitem"))))
    plan = alg.aggregation(["l_orderkey", "o_orderdate", "o_shippriority"], [(Reduction.SUM, "rev", "revenue")],
                           alg.map("rev", _disc(), col))
    return alg, plan, q.cfg


def q4(acc):
    q = _Q(acc)
    alg = q.alg
    plan = alg.aggregation(
        ["o_orderpriority"], [(Reduction.COUNT, "", "order_count")],
        alg.semijoin([("l_orderkey", "o_orderkey")], None,
                     alg.selection(scal.SmallerExpr(A("l_commitdate"), A("l_receiptdate")), alg.scan("lineitem")),
                     alg.selection(_and(scal.LargerEqualExpr(A("o_orderdate"), _d(19930701)),
                                        scal.SmallerExpr(A("o_orderdate"), _d(19931001))), alg.scan("orders"))))
    q.hint(plan, numgroups=100)      # query/plans/tpch4cfg.py
    return alg, plan, q.cfg


def q5(acc):
    q = _Q(acc)
    alg = q.alg
    supp = q.pk(alg.join(("n_nationkey", "s_nationkey"), _regionNations(q, "ASIA"), alg.scan("supplier")))
    co = q.pk(alg.join(("c_custkey", "o_custkey"), alg.scan("customer"),
                       alg.selection(_and(scal.LargerEqualExpr(A("o_orderdate"), _d(19940101)),
                                          scal.SmallerExpr(A("o_orderdate"), _d(19950101))), alg.scan("orders"))))
    col = q.pk(alg.join(("o_orderkey", "l_orderkey"), co, alg.scan("lineitem")))
    all_ = q.pk(alg.innerjoin([("s_suppkey", "l_suppkey"), ("s_nationkey", "c_nationkey")], None, supp, col))
    plan = alg.aggregation(["n_name"], [(Reduction.SUM, "rev", "revenue")], alg.map("rev", _disc(), all_))
    return alg, plan, q.cfg


def q6(acc):
    q = _Q(acc)
    alg = q.alg
    plan = alg.aggregation(
        [], [(Reduction.SUM, "rev", "revenue")],
        alg.map("rev", scal.MulExpr(A("l_extendedprice"), A("l_discount")),
                alg.selection(_and(scal.LargerEqualExpr(A("l_shipdate"), _d(19940101)),
                                   scal.SmallerExpr(A("l_shipdate"), _d(19950101)),
                                   _between(A("l_discount"), _f("0.05"), _f("0.07")),
                                   scal.SmallerExpr(A("l_quantity"), _i(24))), alg.scan("lineitem"))))
    return alg, plan, q.cfg


def q7(acc):
    q = _Q(acc)
    alg = q.alg
    pair = [_s("FRANCE"), _s("GERMANY")]
    supp = q.pk(alg.join(("n1.n_nationkey", "s_nationkey"),
                         alg.selection(_in(A("n1.n_name"), pair), alg.scan("nation", "n1")), alg.scan("supplier")))
    cust = q.pk(alg.join(("n2.n_nationkey", "c_nationkey"),
                         alg.selection(_in(A("n2.n_name"), pair), alg.scan("nation", "n2")), alg.scan("customer")))
    co = q.pk(alg.join(("c_custkey", "o_custkey"), cust, alg.scan("orders")))
    col = q.pk(alg.join(("o_orderkey", "l_orderkey"), co,
                        alg.selection(_between(A("l_shipdate"), _d(19950101), _d(19961231)), alg.scan("lineitem"))))
    all_ = q.pk(alg.join(("s_suppkey", "l_suppkey"), supp, col))
    cross = alg.selection(_or(_and(scal.EqualsExpr(A("n1.n_name"), _s("FRANCE")),
                                   scal.EqualsExpr(A("n2.n_name"), _s("GERMANY"))),
                              _and(scal.EqualsExpr(A("n1.n_name"), _s("GERMANY")),
                                   scal.EqualsExpr(A("n2.n_name"), _s("FRANCE")))), all_)
    named = alg.map("cust_nation", A("n2.n_name"), alg.map("supp_nation", A("n1.n_name"), cross))
    plan = alg.aggregation(["supp_nation", "cust_nation", "l_year"], [(Reduction.SUM, "volume", "revenue")],
                           alg.map("l_year", scal.ExtractExpr(A("l_shipdate"), scal.ExtractType.YEAR),
                                   alg.map("volume", _disc(), named)))
    return alg, plan, q.cfg


def q8(acc):
    q = _Q(acc)
    alg = q.alg
    cust = q.pk(alg.join(("n1.n_nationkey", "c_nationkey"), _regionNations(q, "AMERICA", "n1"),
                         alg.scan("customer")))
    co = q.pk(alg.join(("c_custkey", "o_custkey"), cust,
                       alg.selection(_between(A("o_orderdate"), _d(19950101), _d(19961231)), alg.scan("orders"))))
    pl = q.pk(alg.join(("p_partkey", "l_partkey"),
                       alg.selection(scal.EqualsExpr(A("p_type"), _s("ECONOMY ANODIZED STEEL")), alg.scan("part")),
                       alg.scan("lineitem")))
    col = q.pk(alg.join(("o_orderkey", "l_orderkey"), co, pl))
    supp = q.pk(alg.join(("n2.n_nationkey", "s_nationkey"), alg.scan("nation", "n2"), alg.scan("supplier")))
    all_ = q.pk(alg.join(("s_suppkey", "l_suppkey"), supp, col))
    vol = alg.map("brazil", scal.CaseExpr([(scal.EqualsExpr(A("n2.n_name"), _s("BRAZIL")), A("volume"))],
                                          _f("0.0")),
                  alg.map("volume", _disc(),
                          alg.map("o_year", scal.ExtractExpr(A("o_orderdate"), scal.ExtractType.YEAR), all_)))
    agg = alg.aggregation(["o_year"], [(Reduction.SUM, "brazil", "brazil_volume"),
                                       (Reduction.SUM, "volume", "total_volume")], vol)
    plan = alg.projection(["o_year", "mkt_share"],
                          alg.map("mkt_share", scal.DivExpr(A("brazil_volume"), A("total_volume")), agg))
    return alg, plan, q.cfg


def q9(acc):
    q = _Q(acc)
    alg = q.alg
    psg = q.pk(alg.join(("p_partkey", "ps_partkey"),
                        alg.selection(scal.LikeExpr(A("p_name"), _s("%green%")), alg.scan("part")),
                        alg.scan("partsupp")))
    li = q.pk(alg.innerjoin([("ps_partkey", "l_partkey"), ("ps_suppkey", "l_suppkey")], None, psg,
                            alg.scan("lineitem")))
    sn = q.pk(alg.join(("n_nationkey", "s_nationkey"), alg.scan("nation"), alg.scan("supplier")))
    lis = q.pk(alg.join(("s_suppkey", "l_suppkey"), sn, li))
    lo = q.pk(alg.join(("o_orderkey", "l_orderkey"), alg.scan("orders"), lis))
    amount = scal.SubExpr(_disc(), scal.MulExpr(A("ps_supplycost"), A("l_quantity")))
    plan = alg.aggregation(["n_name", "o_year"], [(Reduction.SUM, "amount", "sum_profit")],
                           alg.map("amount", amount,
                                   alg.map("o_year", scal.ExtractExpr(A("o_orderdate"), scal.ExtractType.YEAR), lo)))
    q.hint(plan, numgroups=1000)     # query/plans/tpch9cfg.py
    return alg, plan, q.cfg


def q11(acc, fraction="0.0001"):
    q = _Q(acc)
    alg = q.alg

    def germanStock(k):
        n, s_, ps = "n%d." % k, "s%d." % k, "ps%d." % k
        supp = q.pk(alg.join((n + "n_nationkey", s_ + "s_nationkey"),
                             alg.selection(scal.EqualsExpr(A(n + "n_name"), _s("GERMANY")),
                                           alg.scan("nation", "n%d" % k)),
                             alg.scan("supplier", "s%d" % k)))
        stock = q.pk(alg.join((s_ + "s_suppkey", ps + "ps_suppkey"), supp, alg.scan("partsupp", "ps%d" % k)))
        return alg.map("v%d" % k, scal.MulExpr(A(ps + "ps_supplycost"), A(ps + "ps_availqty")), stock)

    total = alg.map("threshold", scal.MulExpr(A("total"), _c(fraction, Type.DOUBLE)),
                    alg.aggregation([], [(Reduction.SUM, "v1", "total")], germanStock(1)))
    perPart = alg.aggregation(["ps2.ps_partkey"], [(Reduction.SUM, "v2", "value")], germanStock(2))
    plan = alg.projection(["ps2.ps_partkey", "value"],
                          alg.crossjoin(scal.LargerExpr(A("value"), A("threshold")), total, perPart))
    return alg, plan, q.cfg


def q12(acc):
    q = _Q(acc)
    alg = q.alg
    high = _or(scal.EqualsExpr(A("o_orderpriority"), _s("1-URGENT")),
               scal.EqualsExpr(A("o_orderpriority"), _s("2-HIGH")))
    li = alg.selection(_and(_in(A("l_shipmode"), [_s("MAIL"), _s("SHIP")]),
                            scal.SmallerExpr(A("l_commitdate"), A("l_receiptdate")),
                            scal.SmallerExpr(A("l_shipdate"), A("l_commitdate")),
                            scal.LargerEqualExpr(A("l_receiptdate"), _d(19940101)),
                            scal.SmallerExpr(A("l_receiptdate"), _d(19950101))), alg.scan("lineitem"))
    ol = q.pk(alg.join(("o_orderkey", "l_orderkey"), alg.scan("orders"), li))
    plan = alg.aggregation(
        ["l_shipmode"], [(Reduction.SUM, "high", "high_line_count"), (Reduction.SUM, "low", "low_line_count")],
        alg.map("low", scal.CaseExpr([(high, _i(0))], _i(1)),
                alg.map("high", scal.CaseExpr([(high, _i(1))], _i(0)), ol)))
    q.hint(plan, numgroups=7)        # query/plans/tpch12cfg.py
    return alg, plan, q.cfg


def q13(acc):
    q = _Q(acc)
    alg = q.alg
    orders = alg.selection(scal.NotExpr(scal.LikeExpr(A("o_comment"), _s("%special%requests%"))),
                           alg.scan("orders"))
    perCust = alg.aggregation(["c_custkey"], [(Reduction.COUNT, "o_orderkey", "c_count")],
                              alg.outerjoin([("o_custkey", "c_custkey")], None, orders, alg.scan("customer")))
    plan = alg.aggregation(["c_count"], [(Reduction.COUNT, "", "custdist")], perCust)
    q.hint(plan, numgroups=64)
    return alg, plan, q.cfg


def q14(acc):
    q = _Q(acc)
    alg = q.alg
    pl = q.pk(alg.join(("p_partkey", "l_partkey"), alg.scan("part"),
                       alg.selection(_and(scal.LargerEqualExpr(A("l_shipdate"), _d(19950901)),
                                          scal.SmallerExpr(A("l_shipdate"), _d(19951001))), alg.scan("lineitem"))))
    rev = alg.map("promo", scal.CaseExpr([(scal.LikeExpr(A("p_type"), _s("PROMO%")), A("rev"))], _f("0.0")),
                  alg.map("rev", _disc(), pl))
    agg = alg.aggregation([], [(Reduction.SUM, "promo", "promo_sum"), (Reduction.SUM, "rev", "rev_sum")], rev)
    plan = alg.projection(["promo_revenue"],
                          alg.map("promo_revenue",
                                  scal.DivExpr(scal.MulExpr(_c("100.0", Type.DOUBLE), A("promo_sum")), A("rev_sum")),
                                  agg))
    return alg, plan, q.cfg


def q15(acc):
    q = _Q(acc)
    alg = q.alg
    view = alg.createtemp(
        "revenue0",
        alg.aggregation(["l_suppkey"], [(Reduction.SUM, "rev", "total_revenue")],
                        alg.map("rev", _disc(),
                                alg.selection(_and(scal.LargerEqualExpr(A("l_shipdate"), _d(19960101)),
                                                   scal.SmallerExpr(A("l_shipdate"), _d(19960401))),
                                              alg.scan("lineitem")))))
    top = alg.aggregation([], [(Reduction.MAX, "r1.total_revenue", "max_revenue")], alg.scan("revenue0", "r1"))
    best = alg.crossjoin(scal.EqualsExpr(A("r2.total_revenue"), A("max_revenue")), top, alg.scan("revenue0", "r2"))
    plan = alg.projection(["s_suppkey", "s_name", "s_address", "s_phone", "r2.total_revenue"],
                          q.pk(alg.join(("s_suppkey", "r2.l_suppkey"), alg.scan("supplier"), best)))
    return alg, [view, plan], q.cfg


def q16(acc):
    q = _Q(acc)
    alg = q.alg
    parts = alg.selection(_and(_ne(A("p_brand"), _s("Brand#45")),
                               scal.NotExpr(scal.LikeExpr(A("p_type"), _s("MEDIUM POLISHED%"))),
                               _in(A("p_size"), [_i(v) for v in (49, 14, 23, 45, 19, 3, 36, 9)])),
                          alg.scan("part"))
    pps = q.pk(alg.join(("p_partkey", "ps_partkey"), parts, alg.scan("partsupp")))
    clean = alg.antijoin([("s_suppkey", "ps_suppkey")], None,
                         alg.selection(scal.LikeExpr(A("s_comment"), _s("%Customer%Complaints%")),
                                       alg.scan("supplier")), pps)
    distinct = alg.aggregation(["p_brand", "p_type", "p_size", "ps_suppkey"], [(Reduction.COUNT, "", "pairs")],
                               clean)
    plan = alg.aggregation(["p_brand", "p_type", "p_size"], [(Reduction.COUNT, "", "supplier_cnt")], distinct)
    return alg, plan, q.cfg


def q17(acc):
    q = _Q(acc)
    alg = q.alg
    parts = alg.selection(_and(scal.EqualsExpr(A("p_brand"), _s("Brand#23")),
                               scal.EqualsExpr(A("p_container"), _s("MED BOX"))), alg.scan("part"))
    avgq = alg.map("limit", scal.MulExpr(_c("0.2", Type.DOUBLE), A("avg_qty")),
                   alg.aggregation(["l1.l_partkey"], [(Reduction.AVG, "l1.l_quantity", "avg_qty")],
                                   q.pk(alg.join(("p_partkey", "l1.l_partkey"), parts, alg.scan("lineitem", "l1")))))
    small = alg.selection(scal.SmallerExpr(A("l2.l_quantity"), A("limit")),
                          q.pk(alg.join(("l1.l_partkey", "l2.l_partkey"), avgq, alg.scan("lineitem", "l2"))))
    plan = alg.projection(["avg_yearly"],
                          alg.map("avg_yearly", scal.DivExpr(A("total"), _c("7.0", Type.DOUBLE)),
                                  alg.aggregation([], [(Reduction.SUM, "l2.l_extendedprice", "total")], small)))
    return alg, plan, q.cfg


def q18(acc, quantity=300):
    q = _Q(acc)
    alg = q.alg
    big = alg.selection(scal.LargerExpr(A("sum_qty"), _i(quantity)),
                        alg.aggregation(["l1.l_orderkey"], [(Reduction.SUM, "l1.l_quantity", "sum_qty")],
                                        alg.scan("lineitem", "l1")))
    bo = q.pk(alg.join(("l1.l_orderkey", "o_orderkey"), big, alg.scan("orders")))
    cbo = q.pk(alg.join(("c_custkey", "o_custkey"), alg.scan("customer"), bo))
    all_ = q.pk(alg.join(("o_orderkey", "l2.l_orderkey"), cbo, alg.scan("lineitem", "l2")))
    plan = alg.aggregation(["c_name", "c_custkey", "o_orderkey", "o_orderdate", "o_totalprice"],
                           [(Reduction.SUM, "l2.l_quantity", "quantity")], all_)
    return alg, plan, q.cfg


def q19(acc):
    q = _Q(acc)
    alg = q.alg

    def branch(brand, containers, lo, hi, size):
        return _and(scal.EqualsExpr(A("p_brand"), _s(brand)),
                    _in(A("p_container"), [_s(c) for c in containers]),
                    _between(A("l_quantity"), _i(lo), _i(hi)),
                    _between(A("p_size"), _i(1), _i(size)))

    li = alg.selection(_and(_in(A("l_shipmode"), [_s("AIR"), _s("AIR REG")]),
                            scal.EqualsExpr(A("l_shipinstruct"), _s("DELIVER IN PERSON"))), alg.scan("lineitem"))
    pl = q.pk(alg.join(("p_partkey", "l_partkey"), alg.scan("part"), li))
    sel = alg.selection(_or(branch("Brand#12", ["SM CASE", "SM BOX", "SM PACK", "SM PKG"], 1, 11, 5),
                            branch("Brand#23", ["MED BAG", "MED BOX", "MED PKG", "MED PACK"], 10, 20, 10),
                            branch("Brand#34", ["LG CASE", "LG BOX", "LG PACK", "LG PKG"], 20, 30, 15)), pl)
    plan = alg.aggregation([], [(Reduction.SUM, "rev", "revenue")], alg.map("rev", _disc(), sel))
    return alg, plan, q.cfg


def q20(acc):
    q = _Q(acc)
    alg = q.alg
    forest = alg.semijoin([("p_partkey", "ps_partkey")], None,
                          alg.selection(scal.LikeExpr(A("p_name"), _s("forest%")), alg.scan("part")),
                          alg.scan("partsupp"))
    shipped = alg.aggregation(["l_partkey", "l_suppkey"], [(Reduction.SUM, "l_quantity", "sum_qty")],
                              alg.selection(_and(scal.LargerEqualExpr(A("l_shipdate"), _d(19940101)),
                                                 scal.SmallerExpr(A("l_shipdate"), _d(19950101))),
                                            alg.scan("lineitem")))
    excess = alg.selection(scal.LargerExpr(A("ps_availqty"), A("half")),
                           alg.map("half", scal.MulExpr(_c("0.5", Type.DOUBLE), A("sum_qty")),
                                   q.pk(alg.innerjoin([("l_partkey", "ps_partkey"), ("l_suppkey", "ps_suppkey")],
                                                      None, shipped, forest))))
    canada = q.pk(alg.join(("n_nationkey", "s_nationkey"),
                           alg.selection(scal.EqualsExpr(A("n_name"), _s("CANADA")), alg.scan("nation")),
                           alg.scan("supplier")))
    plan = alg.projection(["s_name", "s_address"],
                          alg.semijoin([("ps_suppkey", "s_suppkey")], None, excess, canada))
    return alg, plan, q.cfg


def q21(acc):
    q = _Q(acc)
    alg = q.alg
    late = alg.selection(scal.LargerExpr(A("l1.l_receiptdate"), A("l1.l_commitdate")), alg.scan("lineitem", "l1"))
    supp = q.pk(alg.join(("n_nationkey", "s_nationkey"),
                         alg.selection(scal.EqualsExpr(A("n_name"), _s("SAUDI ARABIA")), alg.scan("nation")),
                         alg.scan("supplier")))
    sl = q.pk(alg.join(("s_suppkey", "l1.l_suppkey"), supp, late))
    osl = q.pk(alg.join(("o_orderkey", "l1.l_orderkey"),
                        alg.selection(scal.EqualsExpr(A("o_orderstatus"), _c("F", Type.CHAR)), alg.scan("orders")),
                        sl))
    others = alg.semijoin([("l2.l_orderkey", "l1.l_orderkey")], _ne(A("l2.l_suppkey"), A("l1.l_suppkey")),
                          alg.scan("lineitem", "l2"), osl)
    lateOthers = alg.antijoin([("l3.l_orderkey", "l1.l_orderkey")], _ne(A("l3.l_suppkey"), A("l1.l_suppkey")),
                              alg.selection(scal.LargerExpr(A("l3.l_receiptdate"), A("l3.l_commitdate")),
                                            alg.scan("lineitem", "l3")), others)
    plan = alg.aggregation(["s_name"], [(Reduction.COUNT, "", "numwait")], lateOthers)
    return alg, plan, q.cfg


def q22(acc):
    q = _Q(acc)
    alg = q.alg
    codes = [_s(c) for c in ("13", "31", "23", "29", "30", "18", "17")]

    def code(alias):
        return scal.SubstringExpr(A(alias + ".c_phone"), _i(1), _i(2))

    avg = alg.aggregation([], [(Reduction.AVG, "c1.c_acctbal", "avg_bal")],
                          alg.selection(_and(scal.LargerExpr(A("c1.c_acctbal"), _f("0.00")), _in(code("c1"), codes)),
                                        alg.scan("customer", "c1")))
    rich = alg.crossjoin(scal.LargerExpr(A("c2.c_acctbal"), A("avg_bal")), avg,
                         alg.selection(_in(code("c2"), codes), alg.scan("customer", "c2")))
    noOrders = alg.antijoin([("o_custkey", "c2.c_custkey")], None, alg.scan("orders"), rich)
    plan = alg.aggregation(["cntrycode"], [(Reduction.COUNT, "", "numcust"), (Reduction.SUM, "c2.c_acctbal",
                                                                             "totacctbal")],
                           alg.map("cntrycode", code("c2"), noOrders))
    return alg, plan, q.cfg


def q10plan(acc):
    return q10(acc)


QUERIES = OrderedDict((n, f) for n, f in [
    (1, q1), (2, q2), (3, q3), (4, q4), (5, q5), (6, q6), (7, q7), (8, q8), (9, q9), (10, q10plan), (11, q11),
    (12, q12), (13, q13), (14, q14), (15, q15), (16, q16), (17, q17), (18, q18), (19, q19), (20, q20), (21, q21),
    (22, q22)])


# ------------------------------------------------------------------ data
#
# A synthetic TPC-H database: the specification's value domains and key
# relationships (dbgen itself is not in this image), at any scale factor.
# What the 22 queries depend on is kept: orderkeys sparse in runs of 8 per
# 32, a third of the customers without orders (custkey % 3 == 0), lineitem
# suppliers drawn from the part's four partsupp rows, ship/commit/receipt
# dates relative to the order date, return flag and line status from the
# 1995-06-17 cut-off, order status from the line statuses, text from the
# specification's word lists (colours, type and container syllables,
# segments, priorities, ship modes and instructions), phone country codes
# nationkey + 10, and a few comments matching the Q13/Q16 patterns.

NATIONS = [("ALGERIA", 0), ("ARGENTINA", 1), ("BRAZIL", 1), ("CANADA", 1), ("EGYPT", 4), ("ETHIOPIA", 0),
           ("FRANCE", 3), ("GERMANY", 3), ("INDIA", 2), ("INDONESIA", 2), ("IRAN", 4), ("IRAQ", 4), ("JAPAN", 2),
           ("JORDAN", 4), ("KENYA", 0), ("MOROCCO", 0), ("MOZAMBIQUE", 0), ("PERU", 1), ("CHINA", 2),
           ("ROMANIA", 3), ("SAUDI ARABIA", 4), ("VIETNAM", 2), ("RUSSIA", 3), ("UNITED KINGDOM", 3),
           ("UNITED STATES", 1)]
REGIONS = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
COLORS = ("almond antique aquamarine azure beige bisque black blanched blue blush brown burlywood burnished "
          "chartreuse chiffon chocolate coral cornflower cornsilk cream cyan dark deep dim dodger drab firebrick "
          "floral forest frosted gainsboro ghost goldenrod green grey honeydew hot indian ivory khaki lace "
          "lavender lawn lemon light lime linen magenta maroon medium metallic midnight mint misty moccasin "
          "navajo navy olive orange orchid pale papaya peach peru pink plum powder puff purple red rose rosy "
          "royal saddle salmon sandy seashell sienna sky slate smoke snow spring steel tan thistle tomato "
          "turquoise violet wheat white yellow").split()
TYPE_S1 = ["STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO"]
TYPE_S2 = ["ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED"]
TYPE_S3 = ["TIN", "NICKEL", "BRASS", "STEEL", "COPPER"]
CONT_S1 = ["SM", "LG", "MED", "JUMBO", "WRAP"]
CONT_S2 = ["CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM"]
SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "MACHINERY", "HOUSEHOLD"]
PRIORITIES = ["1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW"]
SHIPMODES = ["REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"]
INSTRUCTIONS = ["DELIVER IN PERSON", "COLLECT COD", "NONE", "TAKE BACK RETURN"]
WORDS = ("furiously sly careful blithe quick fluffy slow quiet ruthless thin close dogged daring brave stealthy "
         "permanent enticing idle busy regular final ironic even bold silent foxes ideas theodolites pinto beans "
         "instructions dependencies excuses platelets asymptotes courts dolphins multipliers sauternes warthogs "
         "frets dinos attainments somas Tiresias patterns forges braids hockey players frays warhorses dugouts "
         "notornis epitaphs pearls tithes waters orbits gifts sheaves depths sentiments decoys realms pains "
         "grouches escapades accounts deposits packages requests sleep wake are cajole haggle nag use boost "
         "affix detect integrate maintain nod was lose sublate solve thrash promise engage hinder print x-ray "
         "breach eat grow impress mold poach serve run dazzle snooze doze unwind kindle play hang believe doubt "
         "above across against along among around at atop before behind beneath beside besides between beyond "
         "by despite during except for from in inside instead of into near of on outside over past since "
         "through throughout to toward under until up upon without with within quickly carefully "
         "slyly blithely furiously").split()

CUTOFF = 19950617          # lines shipped after it are open ('O'), others final ('F')


def _strings(values):
    """(offsets, chars) of a STRING column from a list of bytes."""
    lens = np.fromiter((len(v) for v in values), dtype=np.uint64, count=len(values))
    offs = np.zeros(len(values) + 1, dtype=np.uint64)
    np.cumsum(lens, out=offs[1:])
    return offs, np.frombuffer(b"".join(values), dtype=np.uint8).copy() if values else np.zeros(0, np.uint8)


def _text(rng, n, lo, hi, plant=None, every=0):
    """n comments of lo..hi random words; every `every`-th row (offset by a
    random phase) gets the words of `plant` inserted in order."""
    words = [w.encode() for w in WORDS]
    counts = rng.integers(lo, hi + 1, size=n)
    picks = rng.integers(0, len(words), size=int(counts.sum()))
    out, k = [], 0
    for i, c in enumerate(counts):
        row = [words[j] for j in picks[k:k + c]]
        k += c
        if plant and every and i % every == every // 2:
            pos = sorted(rng.integers(0, len(row) + 1, size=len(plant)))
            for p, w in sorted(zip(pos, plant), reverse=True):
                row.insert(p, w.encode())
        out.append(b" ".join(row))
    return out


def _days(yyyymmdd):
    s = np.array(["%04d-%02d-%02d" % (d // 10000, d // 100 % 100, d % 100) for d in np.atleast_1d(yyyymmdd)],
                 dtype="datetime64[D]")
    return s


def _yyyymmdd(days):
    """datetime64[D] array to uint32 yyyymmdd."""
    y = days.astype("datetime64[Y]")
    m = days.astype("datetime64[M]")
    year = y.astype(np.int64) + 1970
    month = (m - y).astype(np.int64) + 1
    day = (days - m).astype(np.int64) + 1
    return (year * 10000 + month * 100 + day).astype(np.uint32)


def _phones(rng, nationkeys):
    n = len(nationkeys)
    a, b, c = rng.integers(100, 1000, n), rng.integers(100, 1000, n), rng.integers(1000, 10000, n)
    return [b"%02d-%03d-%03d-%04d" % (k + 10, x, y, z) for k, x, y, z in zip(nationkeys, a, b, c)]


def _money(rng, n, lo, hi):
    """Uniform cents in [lo, hi] as FLOAT (the schema's money type)."""
    return (rng.integers(int(round(lo * 100)), int(round(hi * 100)) + 1, size=n) / 100.0).astype(np.float32)


def _tblField(typ, data):
    """One column as dbgen prints it: DATE yyyy-mm-dd, money with 2
    decimals, CHAR one character, text verbatim."""
    if typ == Type.STRING:
        if isinstance(data, tuple):
            offs, chars = data
            raw = np.asarray(chars, dtype=np.uint8).tobytes()
            return [raw[int(offs[i]):int(offs[i + 1])].decode("latin-1") for i in range(len(offs) - 1)]
        return [x.decode("latin-1") for x in data]
    a = np.asarray(data)
    if typ == Type.DATE:
        return ["%04d-%02d-%02d" % (d // 10000, d // 100 % 100, d % 100) for d in a.astype(np.int64)]
    if typ in (Type.FLOAT, Type.DOUBLE):
        return ["%.2f" % float(v) for v in a]
    if typ == Type.CHAR:
        return [chr(int(v) & 0xFF) for v in a]
    return [str(int(v)) for v in a]


def writeTbl(tables, directory):
    """Write tables (as returned by ``generate``) as dbgen-format .tbl files
    (one line per row, every field followed by '|'), the input of the
    importer (dbio.dbAccess, dbio.py:117-253)."""
    import os
    os.makedirs(directory, exist_ok=True)
    for name, cols in tables.items():
        fields = [_tblField(typ, data) for typ, data in cols.values()]
        with open(os.path.join(directory, name + ".tbl"), "w", encoding="latin-1", newline="\n") as f:
            for row in zip(*fields):
                f.write("|".join(row) + "|\n")


def generate(sf=0.01, seed=1):
    """The eight TPC-H tables at scale factor `sf`, as plan-API column dicts
    (table -> OrderedDict attr -> (Type, data)); see dbio.MemoryAccessor."""
    rng = np.random.default_rng(seed)
    nS = max(4, int(round(10000 * sf)))
    nP = max(8, int(round(200000 * sf)))
    nC = max(6, int(round(150000 * sf)))
    nO = max(8, int(round(1500000 * sf)))
    T = OrderedDict()

    T["region"] = OrderedDict([("r_regionkey", (Type.INT, np.arange(5, dtype=np.int32))),
                               ("r_name", (Type.STRING, _strings([r.encode() for r in REGIONS]))),
                               ("r_comment", (Type.STRING, _strings(_text(rng, 5, 4, 10))))])
    T["nation"] = OrderedDict([("n_nationkey", (Type.INT, np.arange(25, dtype=np.int32))),
                               ("n_name", (Type.STRING, _strings([n.encode() for n, _ in NATIONS]))),
                               ("n_regionkey", (Type.INT, np.array([r for _, r in NATIONS], dtype=np.int32))),
                               ("n_comment", (Type.STRING, _strings(_text(rng, 25, 4, 12))))])

    snat = rng.integers(0, 25, size=nS).astype(np.int32)
    T["supplier"] = OrderedDict([
        ("s_suppkey", (Type.INT, np.arange(1, nS + 1, dtype=np.int32))),
        ("s_name", (Type.STRING, _strings([b"Supplier#%09d" % k for k in range(1, nS + 1)]))),
        ("s_address", (Type.STRING, _strings(_text(rng, nS, 1, 3)))),
        ("s_nationkey", (Type.INT, snat)),
        ("s_phone", (Type.STRING, _strings(_phones(rng, snat)))),
        ("s_acctbal", (Type.FLOAT, _money(rng, nS, -999.99, 9999.99))),
        ("s_comment", (Type.STRING, _strings(_text(rng, nS, 5, 14, ("Customer", "Complaints"), 7))))])

    pk = np.arange(1, nP + 1, dtype=np.int64)
    colors = [c.encode() for c in COLORS]
    names = rng.integers(0, len(colors), size=(nP, 5))
    mfgr = rng.integers(1, 6, size=nP)
    brand = mfgr * 10 + rng.integers(1, 6, size=nP)
    ptype = [b"%s %s %s" % (TYPE_S1[a].encode(), TYPE_S2[b].encode(), TYPE_S3[c].encode())
             for a, b, c in zip(rng.integers(0, 6, nP), rng.integers(0, 5, nP), rng.integers(0, 5, nP))]
    cont = [b"%s %s" % (CONT_S1[a].encode(), CONT_S2[b].encode())
            for a, b in zip(rng.integers(0, 5, nP), rng.integers(0, 8, nP))]
    retail = ((90000 + (pk // 10) % 20001 + 100 * (pk % 1000)) / 100.0).astype(np.float32)
    T["part"] = OrderedDict([
        ("p_partkey", (Type.INT, pk.astype(np.int32))),
        ("p_name", (Type.STRING, _strings([b" ".join(colors[j] for j in row) for row in names]))),
        ("p_mfgr", (Type.STRING, _strings([b"Manufacturer#%d" % m for m in mfgr]))),
        ("p_brand", (Type.STRING, _strings([b"Brand#%d" % b for b in brand]))),
        ("p_type", (Type.STRING, _strings(ptype))),
        ("p_size", (Type.INT, rng.integers(1, 51, size=nP).astype(np.int32))),
        ("p_container", (Type.STRING, _strings(cont))),
        ("p_retailprice", (Type.FLOAT, retail)),
        ("p_comment", (Type.STRING, _strings(_text(rng, nP, 2, 6))))])

    # four distinct suppliers per part, a quarter of the suppliers apart
    # (the specification's spread formula without its (partkey-1)/S term,
    # which repeats a supplier when S is small)
    step = nS // 4
    psPart = np.repeat(pk, 4)
    psIdx = np.tile(np.arange(4, dtype=np.int64), nP)
    psSupp = (psPart + psIdx * step) % nS + 1
    nPS = len(psPart)
    T["partsupp"] = OrderedDict([
        ("ps_partkey", (Type.INT, psPart.astype(np.int32))),
        ("ps_suppkey", (Type.INT, psSupp.astype(np.int32))),
        ("ps_availqty", (Type.INT, rng.integers(1, 10000, size=nPS).astype(np.int32))),
        ("ps_supplycost", (Type.FLOAT, _money(rng, nPS, 1.0, 1000.0))),
        ("ps_comment", (Type.STRING, _strings(_text(rng, nPS, 4, 12))))])

    cnat = rng.integers(0, 25, size=nC).astype(np.int32)
    T["customer"] = OrderedDict([
        ("c_custkey", (Type.INT, np.arange(1, nC + 1, dtype=np.int32))),
        ("c_name", (Type.STRING, _strings([b"Customer#%09d" % k for k in range(1, nC + 1)]))),
        ("c_address", (Type.STRING, _strings(_text(rng, nC, 1, 3)))),
        ("c_nationkey", (Type.INT, cnat)),
        ("c_phone", (Type.STRING, _strings(_phones(rng, cnat)))),
        ("c_acctbal", (Type.FLOAT, _money(rng, nC, -999.99, 9999.99))),
        ("c_mktsegment", (Type.STRING, _strings([SEGMENTS[i].encode() for i in rng.integers(0, 5, nC)]))),
        ("c_comment", (Type.STRING, _strings(_text(rng, nC, 4, 12))))])

    # orders: sparse keys, customers not divisible by three
    i = np.arange(nO, dtype=np.int64)
    okey = (i // 8) * 32 + i % 8 + 1
    withOrders = np.array([k for k in range(1, nC + 1) if k % 3 != 0], dtype=np.int64)
    ocust = withOrders[rng.integers(0, len(withOrders), size=nO)]
    start = np.datetime64("1992-01-01")
    odays = start + rng.integers(0, (np.datetime64("1998-08-02") - start).astype(int) + 1, size=nO)
    odate = _yyyymmdd(odays)

    # lineitem: 1..7 lines per order
    nl = rng.integers(1, 8, size=nO)
    nL = int(nl.sum())
    lo = np.repeat(np.arange(nO), nl)
    lnum = (np.arange(nL) - np.repeat(np.cumsum(nl) - nl, nl) + 1).astype(np.int32)
    lpart = rng.integers(1, nP + 1, size=nL)
    lsupp = (lpart + rng.integers(0, 4, size=nL) * step) % nS + 1
    qty = rng.integers(1, 51, size=nL).astype(np.int32)
    price = (qty * retail[lpart - 1].astype(np.float64)).astype(np.float32)
    disc = (rng.integers(0, 11, size=nL) / 100.0).astype(np.float32)
    tax = (rng.integers(0, 9, size=nL) / 100.0).astype(np.float32)
    ship = odays[lo] + rng.integers(1, 122, size=nL)
    commit = odays[lo] + rng.integers(30, 91, size=nL)
    receipt = ship + rng.integers(1, 31, size=nL)
    shipD, commitD, receiptD = _yyyymmdd(ship), _yyyymmdd(commit), _yyyymmdd(receipt)
    rflag = np.where(receiptD <= CUTOFF, np.where(rng.integers(0, 2, size=nL) == 0, ord("R"), ord("A")),
                     ord("N")).astype(np.int8)
    lstatus = np.where(shipD > CUTOFF, ord("O"), ord("F")).astype(np.int8)

    openLines = np.bincount(lo, weights=(lstatus == ord("O")), minlength=nO)
    ostatus = np.where(openLines == nl, ord("O"), np.where(openLines == 0, ord("F"), ord("P"))).astype(np.int8)
    charge = price.astype(np.float64) * (1 + tax.astype(np.float64)) * (1 - disc.astype(np.float64))
    total = np.bincount(lo, weights=charge, minlength=nO).astype(np.float32)
    T["orders"] = OrderedDict([
        ("o_orderkey", (Type.INT, okey.astype(np.int32))),
        ("o_custkey", (Type.INT, ocust.astype(np.int32))),
        ("o_orderstatus", (Type.CHAR, ostatus)),
        ("o_totalprice", (Type.FLOAT, total)),
        ("o_orderdate", (Type.DATE, odate)),
        ("o_orderpriority", (Type.STRING, _strings([PRIORITIES[k].encode() for k in rng.integers(0, 5, nO)]))),
        ("o_clerk", (Type.STRING, _strings([b"Clerk#%09d" % k for k in rng.integers(1, max(2, nO // 1500) + 1,
                                                                                     nO)]))),
        ("o_shippriority", (Type.INT, np.zeros(nO, dtype=np.int32))),
        ("o_comment", (Type.STRING, _strings(_text(rng, nO, 3, 10, ("special", "requests"), 50))))])
    T["lineitem"] = OrderedDict([
        ("l_orderkey", (Type.INT, okey[lo].astype(np.int32))),
        ("l_partkey", (Type.INT, lpart.astype(np.int32))),
        ("l_suppkey", (Type.INT, lsupp.astype(np.int32))),
        ("l_linenumber", (Type.INT, lnum)),
        ("l_quantity", (Type.INT, qty)),
        ("l_extendedprice", (Type.FLOAT, price)),
        ("l_discount", (Type.FLOAT, disc)),
        ("l_tax", (Type.FLOAT, tax)),
        ("l_returnflag", (Type.CHAR, rflag)),
        ("l_linestatus", (Type.CHAR, lstatus)),
        ("l_shipdate", (Type.DATE, shipD)),
        ("l_commitdate", (Type.DATE, commitD)),
        ("l_receiptdate", (Type.DATE, receiptD)),
        ("l_shipinstruct", (Type.STRING, _strings([INSTRUCTIONS[k].encode() for k in rng.integers(0, 4, nL)]))),
        ("l_shipmode", (Type.STRING, _strings([SHIPMODES[k].encode() for k in rng.integers(0, 7, nL)]))),
        ("l_comment", (Type.STRING, _strings(_text(rng, nL, 2, 6))))])
    order = ["lineitem", "customer", "orders", "partsupp", "part", "supplier", "nation", "region"]
    return OrderedDict((t, T[t]) for t in order)
