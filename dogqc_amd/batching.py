"""Batched plans: one run of a workload plan for several queries that read
prefixes of the same tables (shared scans for the serving loop).

The members of a batch are ordered by growing prefixes; every table's prefix
row counts are then nondecreasing over the members (all tables of a query
scale together).  The batched plan scans the largest prefix once.  Each
scanned row gets its segment, the first member whose prefix holds it
(scal.SegmentExpr over the `__batch.<table>` input column), a joined row the
larger of its two sides' segments (scal.MaxExpr), and every aggregation
groups by the segment too.  Member k's result is then the combination of
segments 0..k: sums and counts add, MIN/MAX fold, averages are sums over
counts.  No reference counterpart (the reference runs one query per process,
codegen.py:209-218).
"""
from collections import OrderedDict

import numpy as np

import dogqc_amd.scalarAlgebra as scal
from dogqc_amd import workloads
from dogqc_amd.dbio import MemoryAccessor
from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction
from dogqc_amd.types import Type

A, C = scal.AttrExpr, scal.ConstExpr

# group-count hints sized for batches of up to this many members: the
# segment multiplies a query's groups, and a hint that is too small costs
# retries while one too large only costs workspace (small group-bys stay on
# the LDS tier; see AggregationTranslator)
MEMBERS = 16

# tables whose prefix grows with a query's scale, per config (c3's 25 nations
# are always read whole)
SCALED = {"c0": ["cfg0"], "c1": ["dim", "fact"], "c2": ["build", "probe"], "c3": ["item", "cust", "sales"]}


def _seg(alg, name, table, node):
    return alg.map(name, scal.SegmentExpr(table), node)


def plan(config, acc):
    """(alg, plan, cfg, options) of the batched plan of a workload config;
    options carry the serving variant's knobs (lane-refill positions,
    push-down), as planlib's named variants do."""
    alg = RelationalAlgebra(acc)
    one = C("1.0f", Type.FLOAT)
    if config == "c0":
        sel = alg.selection(scal.SmallerEqualExpr(A("date"), C("19981110", Type.DATE)),
                            _seg(alg, "seg", "cfg0", alg.scan("cfg0")))
        charge = alg.map("charge", scal.MulExpr(scal.MulExpr(A("price"), scal.SubExpr(one, A("disc"))),
                                                scal.AddExpr(one, A("tax"))), sel)
        dp = alg.map("disc_price", scal.MulExpr(A("price"), scal.SubExpr(one, A("disc"))), charge)
        p = alg.aggregation(["seg", "g1", "g2"],
                            [(Reduction.SUM, "qty", "sum_qty"), (Reduction.SUM, "price", "sum_base_price"),
                             (Reduction.SUM, "disc_price", "sum_disc_price"), (Reduction.SUM, "charge", "sum_charge"),
                             (Reduction.SUM, "disc", "sum_disc"), (Reduction.COUNT, "", "count_order")], dp)
        return alg, p, {p.opId: {"numgroups": 6 * MEMBERS}}, {}
    if config == "c1":
        dim = alg.selection(scal.EqualsExpr(A("d_cat"), C("2", Type.CHAR)), _seg(alg, "seg_d", "dim", alg.scan("dim")))
        fact = alg.selection(scal.AndExpr(scal.LargerEqualExpr(A("f_date"), C("19950101", Type.DATE)),
                                          scal.SmallerExpr(A("f_date"), C("19960101", Type.DATE))),
                             _seg(alg, "seg_f", "fact", alg.scan("fact")))
        j = alg.join(("d_key", "f_fk"), dim, fact)
        cfg = {j.opId: {"multimatch": False, "selectivity": 0.2}}
        rev = alg.map("rev", scal.MulExpr(A("f_price"), scal.SubExpr(one, A("f_disc"))),
                      alg.map("seg", scal.MaxExpr(A("seg_d"), A("seg_f")), j))
        return (alg, alg.aggregation(["seg", "d_key"], [(Reduction.SUM, "rev", "revenue")], rev), cfg,
                {"buffers": [fact.opId]})
    if config == "c2":
        build = _seg(alg, "seg_b", "build", alg.scan("build"))
        probe = alg.selection(scal.SmallerExpr(A("p_f"), C("10", Type.INT)), _seg(alg, "seg_p", "probe", alg.scan("probe")))
        j = alg.join(("b_key", "p_key"), build, probe)
        cfg = {j.opId: {"multimatch": True, "selectivity": 0.5}}
        prod = alg.map("prod", scal.MulExpr(A("b_pay"), A("p_pay")),
                       alg.map("seg", scal.MaxExpr(A("seg_b"), A("seg_p")), j))
        p = alg.aggregation(["seg"], [(Reduction.COUNT, "", "cnt"), (Reduction.SUM, "prod", "sum_prod")], prod)
        cfg[p.opId] = {"numgroups": MEMBERS}
        return alg, p, cfg, {"pushdown": True}
    if config == "c3":
        like = scal.LikeExpr(A("i_text"), C("%" + workloads.C3_WORD + "%", Type.STRING))
        item = alg.selection(like, _seg(alg, "seg_i", "item", alg.scan("item")))
        sales = _seg(alg, "seg_s", "sales", alg.scan("sales"))
        j1 = alg.join(("i_key", "s_item"), item, sales)
        cust = _seg(alg, "seg_c", "cust", alg.scan("cust"))
        j2 = alg.join(("c_key", "s_cust"), cust, alg.map("seg_is", scal.MaxExpr(A("seg_i"), A("seg_s")), j1))
        j3 = alg.join(("n_key", "s_nation"), alg.scan("nation"),
                      alg.map("seg", scal.MaxExpr(A("seg_is"), A("seg_c")), j2))
        star = alg.createtemp("c3_star", alg.projection(["seg", "i_key", "n_key", "c_year", "s_qty", "s_price"], j3))
        aggs = [(Reduction.SUM, "s_qty", "sum_qty"), (Reduction.SUM, "s_price", "revenue"), (Reduction.COUNT, "", "cnt")]
        # (segment, year) packed into one INT so that the group key is 64
        # bits (exact tags, LDS tier): year * 64 + segment
        yseg = alg.map("yseg", scal.AddExpr(scal.MulExpr(A("c_year"), C("64", Type.INT)), A("seg")),
                       alg.scan("c3_star"))
        ny = alg.createtemp("c3_by_nation_year", alg.aggregation(["yseg", "n_key"], aggs, yseg))
        byItem = alg.aggregation(["seg", "i_key"], aggs, alg.scan("c3_star"))
        # (nation, year) groups: 175 per member, on the LDS tier (4096 slots);
        # (segment, item) groups are bounded by the star rows (no hint: sized
        # from the temp table's row estimate, so large batches do not retry)
        # and number ~106k per member at 1x: a partitioned group-by of 1024
        # partitions instead of one global atomic per star row
        cfg = {j1.opId: {"multimatch": False, "selectivity": 0.06}, j2.opId: {"multimatch": False},
               j3.opId: {"multimatch": False}, ny.child.opId: {"numgroups": 176 * MEMBERS},
               byItem.opId: {"partitions": 1024}}
        # a 2 MB filter of the item keys in device memory (the batched plan
        # runs without a tuning record): measured on the serving mix, batched
        # capacity 2925 -> 4007 q/s, the c3 queries alone 45.4 -> 28.6 ms
        return alg, [star, ny, byItem], cfg, {"buffers": [j1.opId], "nontemporal": True, "join_filter": 2048}
    raise KeyError(config)


def batchColumns(prefixes):
    """The `__batch` input table: per scaled table, its members' prefix row
    counts (member order), as an accessor table."""
    return OrderedDict((t, (Type.INT64, np.asarray(v, dtype=np.int64))) for t, v in prefixes.items())


def prefixAccessor(acc, rows):
    """An accessor over the first rows[t] rows of `acc`'s tables (tables not
    in `rows` whole)."""
    tables = OrderedDict()
    for t in acc.schema:
        if t == "__batch":
            continue
        n = rows.get(t, acc.schema[t]["size"])
        cols = OrderedDict()
        for a, typ in acc.schema[t]["attributes"].items():
            d = acc.column(t, a)
            if typ == Type.STRING:
                offs, chars = d
                cols[a] = (typ, (np.asarray(offs)[:n + 1], chars))
            else:
                cols[a] = (typ, np.asarray(d)[:n])
        tables[t] = cols
    return MemoryAccessor(tables)


def batchAccessor(acc, prefixes):
    """The input of a batched run: `acc`'s tables cut to the largest
    prefix, plus the `__batch` table of the members' prefixes."""
    return withBatch(prefixAccessor(acc, {t: int(v[-1]) for t, v in prefixes.items()}), prefixes)


def withBatch(acc, prefixes):
    """An accessor over `acc`'s tables plus the `__batch` table."""
    tables = OrderedDict()
    for t in acc.schema:
        tables[t] = OrderedDict((a, (typ, acc.column(t, a))) for a, typ in acc.schema[t]["attributes"].items())
    tables["__batch"] = batchColumns(prefixes)
    return MemoryAccessor(tables)


# -------------------------------------------------------------- combine

def _fold(cols, member, keys, sums=(), mins=(), maxs=()):
    """Rows of segments <= member, grouped by `keys`: sums added, MIN/MAX
    folded; returns OrderedDict of columns (keys first, in key order)."""
    seg = np.asarray(cols["seg"])
    m = seg <= member
    sub = {k: np.asarray(v)[m] for k, v in cols.items()}
    n = int(m.sum())
    if keys:
        kt = [sub[k] for k in keys]
        order = np.lexsort(kt[::-1]) if n else np.zeros(0, dtype=np.int64)
        kt = [k[order] for k in kt]
        brk = np.ones(n, dtype=bool)
        if n:
            brk[1:] = np.zeros(n - 1, dtype=bool)
            for k in kt:
                brk[1:] |= k[1:] != k[:-1]
        starts = np.flatnonzero(brk)
        out = OrderedDict((k, v[starts]) for k, v in zip(keys, kt))
    else:
        order = np.arange(n)
        starts = np.zeros(1 if n else 0, dtype=np.int64)
        out = OrderedDict()
    for c in sums:
        v = sub[c][order]
        out[c] = np.add.reduceat(v, starts) if n else v[:0]
    for c in mins:
        v = sub[c][order]
        out[c] = np.minimum.reduceat(v, starts) if n else v[:0]
    for c in maxs:
        v = sub[c][order]
        out[c] = np.maximum.reduceat(v, starts) if n else v[:0]
    return out


def combine(config, result, temps, member):
    """Member `member`'s result, in the columns of the unbatched plan, from
    a batched run's result (and temp tables, for c3)."""
    if config == "c0":
        f = _fold(result, member, ["g1", "g2"],
                  sums=["sum_qty", "sum_base_price", "sum_disc_price", "sum_charge", "sum_disc", "count_order"])
        cnt = f["count_order"].astype(np.float64)
        out = OrderedDict((k, f[k]) for k in ("g1", "g2", "sum_qty", "sum_base_price", "sum_disc_price", "sum_charge"))
        out["avg_qty"] = f["sum_qty"] / cnt
        out["avg_price"] = f["sum_base_price"] / cnt
        out["avg_disc"] = f["sum_disc"] / cnt
        out["count_order"] = f["count_order"]
        return out, {}
    if config == "c1":
        return _fold(result, member, ["d_key"], sums=["revenue"]), {}
    if config == "c2":
        f = _fold(result, member, [], sums=["cnt", "sum_prod"])
        if len(f["cnt"]) == 0:      # an ungrouped aggregate has one row
            f = OrderedDict([("cnt", np.zeros(1, dtype=np.int32)), ("sum_prod", np.zeros(1, dtype=np.int64))])
        return f, {}
    if config == "c3":
        sums = ["sum_qty", "revenue", "cnt"]
        ny = OrderedDict(temps["c3_by_nation_year"])
        yseg = np.asarray(ny.pop("yseg")).astype(np.int64)
        ny["seg"], ny["c_year"] = (yseg % 64).astype(np.int32), (yseg // 64).astype(np.int32)
        return (_fold(result, member, ["i_key"], sums=sums),
                {"c3_by_nation_year": _fold(ny, member, ["n_key", "c_year"], sums=sums)})
    raise KeyError(config)
