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
  c3  4-way star join + string predicate: sales (s_item INT U[0, 2M),
      s_cust INT U[0, 200k), s_nation INT U[0, 25), s_qty INT U[1,50],
      s_price DOUBLE as c0's price), item (i_key INT permutation of 0..N-1,
      i_text STRING of 55 bytes: 5 words drawn with replacement from the 92
      TPC-H colour words, space-joined and space-padded; no word is a
      substring of another, so LIKE '%green%' is a word test passing
      1-(91/92)^5 = 5.3%), cust (c_key permutation, c_year INT U{1992..1998}),
      nation (n_key permutation of 0..24).  The star join is materialised
      once as a temp table; one aggregation groups it by (n_key, c_year)
      (~175 groups) into a second temp table, the result groups it by i_key
      (~106k groups).
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

# the 92 colour words of the TPC-H specification (P_NAME); none is a
# substring of another
VOCAB = ("almond antique aquamarine azure beige bisque black blanched blue blush brown burlywood burnished "
         "chartreuse chiffon chocolate coral cornflower cornsilk cream cyan dark deep dim dodger drab firebrick "
         "floral forest frosted gainsboro ghost goldenrod green grey honeydew hot indian ivory khaki lace lavender "
         "lawn lemon light lime linen magenta maroon medium metallic midnight mint misty moccasin navajo navy "
         "olive orange orchid pale papaya peach peru pink plum powder puff purple red rose rosy royal saddle "
         "salmon sandy seashell sienna sky slate smoke snow spring steel tan thistle tomato turquoise violet "
         "wheat white yellow").split()
TEXT_WIDTH = 55
C3_WORD = "green"
FULL_ROWS = {"c0": {"cfg0": 1000000},
             "c1": {"dim": 1500000, "fact": 6000000},
             "c2": {"build": 1000000, "probe": 16000000},
             "c3": {"sales": 60000000, "item": 2000000, "cust": 200000, "nation": 25},
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
    if config == "c3":
        return OrderedDict([
            ("sales", OrderedDict([("s_item", Type.INT), ("s_cust", Type.INT), ("s_nation", Type.INT),
                                   ("s_qty", Type.INT), ("s_price", Type.DOUBLE)])),
            ("item", OrderedDict([("i_key", Type.INT), ("i_text", Type.STRING)])),
            ("cust", OrderedDict([("c_key", Type.INT), ("c_year", Type.INT)])),
            ("nation", OrderedDict([("n_key", Type.INT)]))])
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
    elif config == "c3":
        ns, ni, nc, nn = rows["sales"], rows["item"], rows["cust"], rows["nation"]
        item = OrderedDict()
        item["i_key"] = rng.permutation(ni).astype(np.int32)
        item["i_text"] = fixedWidthText(rng.integers(0, len(VOCAB), size=(ni, 5)))
        cust = OrderedDict()
        cust["c_key"] = rng.permutation(nc).astype(np.int32)
        cust["c_year"] = rng.integers(1992, 1999, size=nc).astype(np.int32)
        nation = OrderedDict()
        nation["n_key"] = rng.permutation(nn).astype(np.int32)
        sales = OrderedDict()
        sales["s_item"] = rng.integers(0, max(1, ni), size=ns).astype(np.int32)
        sales["s_cust"] = rng.integers(0, max(1, nc), size=ns).astype(np.int32)
        sales["s_nation"] = rng.integers(0, max(1, nn), size=ns).astype(np.int32)
        sales["s_qty"] = rng.integers(1, 51, size=ns).astype(np.int32)
        sales["s_price"] = _money(rng, ns)
        out["sales"], out["item"], out["cust"], out["nation"] = sales, item, cust, nation
    elif config == "cal":
        out["k"] = OrderedDict([("k_v", np.arange(rows["k"], dtype=np.int64))])
    else:
        raise KeyError(config)
    typed = OrderedDict()
    for t, cols in out.items():
        typed[t] = OrderedDict((a, (sch[t][a], cols[a])) for a in sch[t])
    return typed


def fixedWidthText(words):
    """STRING column (offsets, chars) of TEXT_WIDTH-byte rows: the words of
    each row of ``words`` (indices into VOCAB) joined by single spaces and
    padded with spaces."""
    n, k = words.shape
    lens = np.array([len(w) for w in VOCAB], dtype=np.int64)
    table = np.full((len(VOCAB), max(lens)), ord(" "), dtype=np.uint8)
    for i, w in enumerate(VOCAB):
        table[i, :len(w)] = np.frombuffer(w.encode(), dtype=np.uint8)
    out = np.full((n, TEXT_WIDTH), ord(" "), dtype=np.uint8)
    pos = np.zeros(n, dtype=np.int64)
    rowsIdx = np.arange(n)
    for j in range(k):
        wl = lens[words[:, j]]
        if int((pos + wl).max(initial=0)) > TEXT_WIDTH:
            raise ValueError("words do not fit the text width")
        for c in range(table.shape[1]):
            m = c < wl
            out[rowsIdx[m], pos[m] + c] = table[words[m, j], c]
        pos += wl + 1
    offs = np.arange(n + 1, dtype=np.uint64) * np.uint64(TEXT_WIDTH)
    return (offs, out.reshape(-1) if n else np.zeros(1, dtype=np.uint8))


def emptyAccessor(config):
    """Schema-only accessor (zero rows) for code generation."""
    tables = OrderedDict()
    for t, attrs in schema(config).items():
        tables[t] = OrderedDict(
            (a, (typ, (np.zeros(1, dtype=np.uint64), np.zeros(1, dtype=np.uint8)) if typ == Type.STRING
                 else np.zeros(0, dtype=np.int8))) for a, typ in attrs.items())
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
    elif config == "c3":
        like = scal.LikeExpr(scal.AttrExpr("i_text"), scal.ConstExpr("%" + C3_WORD + "%", Type.STRING))
        star = alg.join(("n_key", "s_nation"), alg.scan("nation"),
                        alg.join(("c_key", "s_cust"), alg.scan("cust"),
                                 alg.join(("i_key", "s_item"), alg.selection(like, alg.scan("item")),
                                          alg.scan("sales"))))
        aggs = [(Reduction.SUM, "s_qty", "sum_qty"), (Reduction.SUM, "s_price", "revenue"),
                (Reduction.COUNT, "", "cnt")]
        p = [alg.createtemp("c3_star", alg.projection(["i_key", "n_key", "c_year", "s_qty", "s_price"], star)),
             alg.createtemp("c3_by_nation_year", alg.aggregation(["n_key", "c_year"], aggs, alg.scan("c3_star"))),
             alg.aggregation(["i_key"], aggs, alg.scan("c3_star"))]
        cfg = {4: {"selectivity": 0.06}, 6: {"multimatch": False, "selectivity": 0.06},
               7: {"multimatch": False}, 8: {"multimatch": False}, 12: {"numgroups": 175},
               15: {"numgroups": 120000}}
    elif config == "cal":
        # copy of one 8-byte column: reads and writes exactly 8 bytes per row
        p = alg.projection(["k_v"], alg.scan("k"))
        cfg = {}
    else:
        raise KeyError(config)
    return alg, p, cfg
