"""TPC-H schema in the plan API's table-dict format, and query 10 written
with the plan API.

The schema follows the TPC-H specification's column names and the type
choices of the reference's schema module (query/schema/tpch.py:23-122):
keys and counts INT, money FLOAT, flags CHAR, dates DATE as yyyymmdd,
text STRING.  Sizes are the scale-factor-1 row counts.
"""
from collections import OrderedDict

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
