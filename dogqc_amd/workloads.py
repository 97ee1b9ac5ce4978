"""Synthetic workloads: schemas, seeded column generators and plans written
in the reference plan API.

  c0  scan-filter-groupby (TPC-H Q1 shape, query/plans/tpch1.py) over one
      table: g1 CHAR {0,1,2}, g2 CHAR {0,1}, qty INT U[1,50],
      price DOUBLE = cents/100 with cents U[90000, 10500000],
      disc DOUBLE k/100 k U{0..10}, tax DOUBLE k/100 k U{0..8},
      date DATE uniform over 1992-01-01..1998-12-31; filter date <= 19981110.
  c1  filtered PK-FK join + aggregation: dim (d_key INT permutation of
      0..N-1, d_cat CHAR '0'..'4', d_seg INT U[0,1e6)), fact (f_fk INT
      U[0, Ndim), f_price, f_disc, f_date); dim filter d_cat == '2', fact
      filter 19950101 <= f_date < 19960101; group by d_key,
      SUM f_price * (1 - f_disc).
  c2  expanding multimatch join: build (b_key INT Zipf(s=1) over 100k
      values, b_pay INT64 U[1, 2^20)), probe (p_key INT U[0, 200k),
      p_f INT U[0,100), p_pay INT64 U[1, 2^20)); filter p_f < 10;
      COUNT and SUM b_pay * p_pay.
  cal a one-column INT64 table copied to the result (8 bytes read and 8
      written per row): calibrates memory-traffic counters.

Random numbers come from numpy's PCG64 generator, so a (config, rows, seed)
triple always yields the same columns.
"""
from collections import OrderedDict

import numpy as np

import dogqc_amd.scalarAlgebra as scal
from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction
from dogqc_amd.types import Type
from dogqc_amd.dbio import MemoryAccessor

SEEDS = {"c0": 1000, "c1": 1001, "c2": 1002, "c3": 1003, "cal": 0}
FULL_ROWS = {"c0": {"cfg0": 1000000},
             "c1": {"dim": 1500000, "fact": 6000000},
             "c2": {"build": 1000000, "probe": 16000000},
             "cal": {"k": 64 * 1024 * 1024}}


def dateTable():
    """yyyymmdd of every day 1992-01-01 .. 1998-12-31 (2557 days)."""
    days = np.arange(np.datetime64("1992-01-01"), np.datetime64("1999-01-01"))
    ymd = days.astype("datetime64[D]").astype(str)
    return np.array([int(s.replace("-", "")) for s in ymd], dtype=np.uint32)


_DATES = None


def _dates():
    global _DATES
    if _DATES is None:
        _DATES = dateTable()
    return _DATES


def _money(rng, n):
    return rng.integers(90000, 10500001, size=n, dtype=np.int64).astype(np.float64) / 100.0


def _hundredths(rng, n, hi):
    return rng.integers(0, hi + 1, size=n, dtype=np.int64).astype(np.float64) / 100.0


def _randDates(rng, n):
    d = _dates()
    return d[rng.integers(0, len(d), size=n)]


# ----------------------------------------------------------------- schemas

def schema(config):
    if config == "c0":
        return {"cfg0": OrderedDict([("g1", Type.CHAR), ("g2", Type.CHAR), ("qty", Type.INT),
                                     ("price", Type.DOUBLE), ("disc", Type.DOUBLE), ("tax", Type.DOUBLE),
                                     ("date", Type.DATE)])}
    if config == "c1":
        return {"dim": OrderedDict([("d_key", Type.INT), ("d_cat", Type.CHAR), ("d_seg", Type.INT)]),
                "fact": OrderedDict([("f_fk", Type.INT), ("f_price", Type.DOUBLE), ("f_disc", Type.DOUBLE),
                                     ("f_date", Type.DATE)])}
    if config == "cal":
        return {"k": OrderedDict([("k_v", Type.INT64)])}
    if config == "c2":
        return {"build": OrderedDict([("b_key", Type.INT), ("b_pay", Type.INT64)]),
                "probe": OrderedDict([("p_key", Type.INT), ("p_f", Type.INT), ("p_pay", Type.INT64)])}
    raise KeyError(config)


# -------------------------------------------------------------- generators

def generate(config, rows=None, seed=None):
    """Columns of a workload as an OrderedDict table -> OrderedDict attr ->
    (Type, ndarray).  ``rows`` maps table -> row count (default: full size)."""
    rows = dict(FULL_ROWS[config], **(rows or {}))
    seed = SEEDS[config] if seed is None else seed
    rng = np.random.Generator(np.random.PCG64(seed))
    sch = schema(config)
    out = OrderedDict()
    if config == "c0":
        n = rows["cfg0"]
        cols = OrderedDict()
        cols["g1"] = rng.integers(0, 3, size=n).astype(np.int8)
        cols["g2"] = rng.integers(0, 2, size=n).astype(np.int8)
        cols["qty"] = rng.integers(1, 51, size=n).astype(np.int32)
        cols["price"] = _money(rng, n)
        cols["disc"] = _hundredths(rng, n, 10)
        cols["tax"] = _hundredths(rng, n, 8)
        cols["date"] = _randDates(rng, n)
        out["cfg0"] = cols
    elif config == "c1":
        nd, nf = rows["dim"], rows["fact"]
        dim = OrderedDict()
        dim["d_key"] = rng.permutation(nd).astype(np.int32)
        dim["d_cat"] = (rng.integers(0, 5, size=nd) + ord("0")).astype(np.int8)
        dim["d_seg"] = rng.integers(0, 1000000, size=nd).astype(np.int32)
        fact = OrderedDict()
        fact["f_fk"] = rng.integers(0, max(1, nd), size=nf).astype(np.int32)
        fact["f_price"] = _money(rng, nf)
        fact["f_disc"] = _hundredths(rng, nf, 10)
        fact["f_date"] = _randDates(rng, nf)
        out["dim"], out["fact"] = dim, fact
    elif config == "c2":
        nb, np_ = rows["build"], rows["probe"]
        nvals = min(100000, max(1, nb))
        ranks = np.arange(1, nvals + 1, dtype=np.float64)
        cdf = np.cumsum(1.0 / ranks)
        cdf /= cdf[-1]
        perm = rng.permutation(nvals).astype(np.int32)
        r = np.searchsorted(cdf, rng.random(nb), side="right")
        r = np.minimum(r, nvals - 1)
        build = OrderedDict()
        build["b_key"] = perm[r]
        build["b_pay"] = rng.integers(1, 1 << 20, size=nb, dtype=np.int64)
        probe = OrderedDict()
        probe["p_key"] = rng.integers(0, 2 * nvals, size=np_).astype(np.int32)
        probe["p_f"] = rng.integers(0, 100, size=np_).astype(np.int32)
        probe["p_pay"] = rng.integers(1, 1 << 20, size=np_, dtype=np.int64)
        out["build"], out["probe"] = build, probe
    elif config == "cal":
        out["k"] = OrderedDict([("k_v", np.arange(rows["k"], dtype=np.int64))])
    else:
        raise KeyError(config)
    typed = OrderedDict()
    for t, cols in out.items():
        typed[t] = OrderedDict((a, (sch[t][a], cols[a])) for a in sch[t])
    return typed


def emptyAccessor(config):
    """Schema-only accessor (zero rows) for code generation."""
    tables = OrderedDict()
    for t, attrs in schema(config).items():
        tables[t] = OrderedDict((a, (typ, np.zeros(0, dtype=np.int8))) for a, typ in attrs.items())
    return MemoryAccessor(tables)


# ------------------------------------------------------------------- plans

def plan(config, acc):
    """(alg, plan, cfg) for a workload, built with the reference plan API."""
    alg = RelationalAlgebra(acc)
    if config == "c0":
        p = alg.aggregation(
            ["g1", "g2"],
            [(Reduction.SUM, "qty", "sum_qty"),
             (Reduction.SUM, "price", "sum_base_price"),
             (Reduction.SUM, "disc_price", "sum_disc_price"),
             (Reduction.SUM, "charge", "sum_charge"),
             (Reduction.AVG, "qty", "avg_qty"),
             (Reduction.AVG, "price", "avg_price"),
             (Reduction.AVG, "disc", "avg_disc"),
             (Reduction.COUNT, "", "count_order")],
            alg.map("disc_price",
                    scal.MulExpr(scal.AttrExpr("price"),
                                 scal.SubExpr(scal.ConstExpr("1.0f", Type.FLOAT), scal.AttrExpr("disc"))),
                    alg.map("charge",
                            scal.MulExpr(
                                scal.MulExpr(scal.AttrExpr("price"),
                                             scal.SubExpr(scal.ConstExpr("1.0f", Type.FLOAT),
                                                          scal.AttrExpr("disc"))),
                                scal.AddExpr(scal.ConstExpr("1.0f", Type.FLOAT), scal.AttrExpr("tax"))),
                            alg.selection(
                                scal.SmallerEqualExpr(scal.AttrExpr("date"), scal.ConstExpr("19981110", Type.DATE)),
                                alg.scan("cfg0")))))
        cfg = {5: {"numgroups": 6}}
    elif config == "c1":
        p = alg.aggregation(
            ["d_key"],
            [(Reduction.SUM, "rev", "revenue")],
            alg.map("rev",
                    scal.MulExpr(scal.AttrExpr("f_price"),
                                 scal.SubExpr(scal.ConstExpr("1.0f", Type.FLOAT), scal.AttrExpr("f_disc"))),
                    alg.join(("d_key", "f_fk"),
                             alg.selection(scal.EqualsExpr(scal.AttrExpr("d_cat"), scal.ConstExpr("2", Type.CHAR)),
                                           alg.scan("dim")),
                             alg.selection(scal.AndExpr(
                                 scal.LargerEqualExpr(scal.AttrExpr("f_date"), scal.ConstExpr("19950101", Type.DATE)),
                                 scal.SmallerExpr(scal.AttrExpr("f_date"), scal.ConstExpr("19960101", Type.DATE))),
                                 alg.scan("fact")))))
        cfg = {2: {"selectivity": 0.2}, 4: {"selectivity": 0.143},
               5: {"multimatch": False, "selectivity": 0.2}, 7: {"numgroups": 150000}}
    elif config == "c2":
        p = alg.aggregation(
            [],
            [(Reduction.COUNT, "", "cnt"), (Reduction.SUM, "prod", "sum_prod")],
            alg.map("prod", scal.MulExpr(scal.AttrExpr("b_pay"), scal.AttrExpr("p_pay")),
                    alg.join(("b_key", "p_key"),
                             alg.scan("build"),
                             alg.selection(scal.SmallerExpr(scal.AttrExpr("p_f"), scal.ConstExpr("10", Type.INT)),
                                           alg.scan("probe")))))
        cfg = {3: {"selectivity": 0.1}, 4: {"multimatch": True, "selectivity": 0.5}}
    elif config == "cal":
        # copy of one 8-byte column: reads and writes exactly 8 bytes per row
        p = alg.projection(["k_v"], alg.scan("k"))
        cfg = {}
    else:
        raise KeyError(config)
    return alg, p, cfg
