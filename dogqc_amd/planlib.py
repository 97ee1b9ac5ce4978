"""Named plans: the workloads and the feature plans the GPU tests run.

Each entry builds a plan with the reference plan API and carries the
translator knobs it is compiled with.  ``translate(name, acc)`` resolves and
translates a plan; ``build()`` in the repository's entry module compiles every
entry ahead of time so no compilation happens on the GPU machine.
"""
from collections import OrderedDict
import contextlib
import itertools

import numpy as np

import dogqc_amd.scalarAlgebra as scal
from dogqc_amd import workloads
from dogqc_amd.relationalAlgebra import RelationalAlgebra, Reduction
from dogqc_amd.types import Type
from dogqc_amd.dbio import MemoryAccessor
from dogqc_amd.hiplang import CType
from dogqc_amd.hipTranslator import HipCompiler, StringHash, ScanTranslator
from dogqc_amd.hashJoins import EquiJoinTranslator
from dogqc_amd.hipDivergenceBuffer import LaneRefill, BufferType
from dogqc_amd.kernel import KernelCall


@contextlib.contextmanager
def knobs(pushdown=False, block=None, threshold=None, rowsPerThread=None, partBuild=False, bufferType=None,
          waveStrings=False, nontemporal=False, joinFilter=0):
    old = (EquiJoinTranslator.usePushDownJoin, KernelCall.defaultBlockSize, LaneRefill.threshold,
           KernelCall.defaultRowsPerThread, EquiJoinTranslator.usePartitionedBuild, LaneRefill.bufferType,
           StringHash.wave, ScanTranslator.nontemporal, EquiJoinTranslator.defaultJoinFilter)
    ScanTranslator.nontemporal = nontemporal
    EquiJoinTranslator.defaultJoinFilter = joinFilter
    EquiJoinTranslator.usePushDownJoin = pushdown
    EquiJoinTranslator.usePartitionedBuild = partBuild
    LaneRefill.bufferType = bufferType or BufferType.SMEM
    StringHash.wave = bool(waveStrings)
    if block is not None:
        KernelCall.defaultBlockSize = block
    if threshold is not None:
        LaneRefill.threshold = threshold
    if rowsPerThread is not None:
        KernelCall.defaultRowsPerThread = rowsPerThread
    try:
        yield
    finally:
        (EquiJoinTranslator.usePushDownJoin, KernelCall.defaultBlockSize, LaneRefill.threshold,
         KernelCall.defaultRowsPerThread, EquiJoinTranslator.usePartitionedBuild, LaneRefill.bufferType,
         StringHash.wave, ScanTranslator.nontemporal, EquiJoinTranslator.defaultJoinFilter) = old


# ------------------------------------------------------------- feature plans
# A small multi-type table "t" and a dimension "u" for joins, LIKE, MIN/MAX,
# semi/anti joins and hashed (non-packable) group keys.

FEATURE_SCHEMA = OrderedDict([
    ("t", OrderedDict([("t_id", Type.INT), ("t_grp", Type.INT), ("t_cat", Type.CHAR), ("t_val", Type.DOUBLE),
                       ("t_f", Type.FLOAT), ("t_big", Type.INT64), ("t_date", Type.DATE), ("t_txt", Type.STRING)])),
    ("u", OrderedDict([("u_id", Type.INT), ("u_grp", Type.INT), ("u_w", Type.DOUBLE), ("u_name", Type.STRING)])),
])

WORDS = ["alpha", "bravo", "charlie", "delta", "echo", "foxtrot", "golf", "hotel", "india", "juliet"]


def featureData(nt=5000, nu=700, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    t = OrderedDict()
    t["t_id"] = (Type.INT, np.arange(nt, dtype=np.int32))
    t["t_grp"] = (Type.INT, rng.integers(-50, 50, size=nt).astype(np.int32))
    t["t_cat"] = (Type.CHAR, rng.integers(ord("a"), ord("e"), size=nt).astype(np.int8))
    t["t_val"] = (Type.DOUBLE, np.round(rng.uniform(-100, 100, size=nt), 2))
    t["t_f"] = (Type.FLOAT, np.round(rng.uniform(0, 10, size=nt), 1).astype(np.float32))
    t["t_big"] = (Type.INT64, rng.integers(-(1 << 40), 1 << 40, size=nt, dtype=np.int64))
    t["t_date"] = (Type.DATE, workloads.dateTable()[rng.integers(0, 2557, size=nt)])
    txt = [(" ".join(WORDS[j] for j in rng.integers(0, len(WORDS), size=rng.integers(0, 4)))).encode()
           for _ in range(nt)]
    t["t_txt"] = (Type.STRING, txt)
    u = OrderedDict()
    u["u_id"] = (Type.INT, rng.permutation(nu).astype(np.int32) * 3)
    u["u_grp"] = (Type.INT, rng.integers(-50, 50, size=nu).astype(np.int32))
    u["u_w"] = (Type.DOUBLE, np.round(rng.uniform(0, 5, size=nu), 3))
    u["u_name"] = (Type.STRING, [WORDS[j].encode() for j in rng.integers(0, len(WORDS), size=nu)])
    return _asAccessor(OrderedDict([("t", t), ("u", u)]))


def _asAccessor(tables):
    conv = OrderedDict()
    for name, cols in tables.items():
        conv[name] = OrderedDict()
        for a, (typ, data) in cols.items():
            if typ == Type.STRING and not isinstance(data, tuple):
                lens = np.fromiter((len(s) for s in data), dtype=np.uint64, count=len(data))
                offs = np.zeros(len(data) + 1, dtype=np.uint64)
                np.cumsum(lens, out=offs[1:])
                chars = np.frombuffer(b"".join(data) or b"\0", dtype=np.uint8)
                data = (offs, chars)
            conv[name][a] = (typ, data)
    return MemoryAccessor(conv)


def featureEmpty():
    return featureData(0, 0)


def _A(n):
    return scal.AttrExpr(n)


def _C(tok, typ):
    return scal.ConstExpr(tok, typ)


def _featurePlan(name, alg):
    if name == "f_scan_minmax":
        # no-group MIN/MAX/SUM/COUNT/AVG over a filtered scan with INT64 and FLOAT inputs
        p = alg.aggregation([], [(Reduction.MIN, "t_val", "mn"), (Reduction.MAX, "t_val", "mx"),
                                 (Reduction.SUM, "t_big", "sb"), (Reduction.MAX, "t_big", "mb"),
                                 (Reduction.AVG, "t_f", "af"), (Reduction.COUNT, "", "c")],
                            alg.selection(scal.LargerExpr(_A("t_date"), _C("19950101", Type.DATE)), alg.scan("t")))
        return p, {}
    if name == "f_groupby_hashed":
        # INT64 + CHAR group key does not pack into 63 bits -> hashed global table
        p = alg.aggregation(["t_big2", "t_cat"], [(Reduction.SUM, "t_val", "s"), (Reduction.COUNT, "", "c")],
                            alg.map("t_big2", scal.DivExpr(_A("t_big"), _C("1099511627776", Type.INT64)),
                                    alg.scan("t")))
        return p, {3: {"numgroups": 8}}
    if name == "f_groupby_wide":
        # many groups: (t_grp, t_cat) with an under-estimated numgroups -> overflow + retry
        p = alg.aggregation(["t_grp", "t_cat"], [(Reduction.SUM, "t_val", "s"), (Reduction.MIN, "t_id", "mi"),
                                                  (Reduction.MAX, "t_date", "md"), (Reduction.COUNT, "", "c")],
                            alg.scan("t"))
        return p, {2: {"numgroups": 4}}
    if name == "f_like_project":
        # LIKE filter + string materialisation + map arithmetic
        p = alg.projection(["t_id", "t_txt", "v2"],
                           alg.map("v2", scal.MulExpr(_A("t_val"), _C("2", Type.INT)),
                                   alg.selection(scal.LikeExpr(_A("t_txt"), _C("%echo%", Type.STRING)),
                                                 alg.scan("t"))))
        return p, {2: {"selectivity": 0.01}}
    if name == "f_join_unique":
        # PK join on u_id (unique), build payload + string, grouped by a build attribute
        p = alg.aggregation(["u_name"], [(Reduction.SUM, "w", "sw"), (Reduction.COUNT, "", "c")],
                            alg.map("w", scal.MulExpr(_A("u_w"), _A("t_val")),
                                    alg.join(("u_id", "t_id"), alg.scan("u"),
                                             alg.selection(scal.EqualsExpr(_A("t_cat"), _C("b", Type.CHAR)),
                                                           alg.scan("t")))))
        return p, {4: {"multimatch": False}, 6: {"numgroups": 16}}
    if name == "f_join_unique_lowest":
        # a PK join with the build side's selectivity set far too low: the
        # tables are sized for ~1 build row, so every structure that is sized
        # from it (a partitioned build's segments) overflows
        p = alg.aggregation(["u_grp"], [(Reduction.SUM, "w", "sw"), (Reduction.COUNT, "", "c")],
                            alg.map("w", scal.MulExpr(_A("u_w"), _A("t_val")),
                                    alg.join(("u_id", "t_id"),
                                             alg.selection(scal.LargerExpr(_A("u_grp"), _C("-1000", Type.INT)),
                                                           alg.scan("u")),
                                             alg.scan("t"))))
        return p, {2: {"selectivity": 0.001}, 4: {"multimatch": False}, 6: {"numgroups": 128}}
    if name == "f_groupjoin":
        # group-by on the key of a unique join: accumulates at the join slot
        p = alg.aggregation(["u_id"], [(Reduction.SUM, "t_val", "s"), (Reduction.MIN, "t_val", "mn"),
                                        (Reduction.AVG, "t_f", "af"), (Reduction.COUNT, "", "c")],
                            alg.join(("u_id", "t_id"), alg.scan("u"),
                                     alg.selection(scal.LargerExpr(_A("t_val"), _C("-50.0", Type.DOUBLE)),
                                                   alg.scan("t"))))
        return p, {4: {"multimatch": False}}
    if name == "f_join_multi":
        # multimatch join on grp (many duplicates both sides), result rows materialised
        p = alg.projection(["t_id", "u_id", "t_grp"],
                           alg.join(("u_grp", "t_grp"), alg.scan("u"),
                                    alg.selection(scal.SmallerExpr(_A("t_id"), _C("300", Type.INT)), alg.scan("t"))))
        return p, {4: {"multimatch": True}}
    if name == "f_join_string":
        # multimatch join on a string key (single-word t_txt = u_name, about
        # 70 u rows per word), grouped by t_cat
        p = alg.aggregation(["t_cat"], [(Reduction.COUNT, "", "c"), (Reduction.SUM, "u_w", "s")],
                            alg.join(("u_name", "t_txt"), alg.scan("u"), alg.scan("t")))
        return p, {3: {"multimatch": True}}
    if name == "f_semi":
        p = alg.aggregation(["t_cat"], [(Reduction.COUNT, "", "c")],
                            alg.semijoin(("u_grp", "t_grp"), None, alg.scan("u"), alg.scan("t")))
        return p, {3: {"multimatch": False}}
    if name == "f_anti_multi":
        p = alg.aggregation(["t_cat"], [(Reduction.COUNT, "", "c"), (Reduction.SUM, "t_val", "s")],
                            alg.antijoin(("u_grp", "t_grp"),
                                         scal.LargerExpr(_A("u_w"), _C("2.5", Type.DOUBLE)),
                                         alg.scan("u"), alg.scan("t")))
        return p, {3: {"multimatch": True}}
    if name == "f_groupby_64":
        # two INT keys = 64-bit exact tag; (-1, -1) is the key stored in the
        # table's extra slot
        p = alg.aggregation(["t_grp", "g2"], [(Reduction.SUM, "t_val", "s"), (Reduction.COUNT, "", "c"),
                                               (Reduction.MIN, "t_id", "mi")],
                            alg.map("g2", scal.MulExpr(_A("t_grp"), _C("1", Type.INT)), alg.scan("t")))
        return p, {3: {"numgroups": 128}}
    if name == "f_groupby_int64":
        # one INT64 key (64-bit exact tag), about 2k groups
        p = alg.aggregation(["t_bq"], [(Reduction.SUM, "t_f", "sf"), (Reduction.COUNT, "", "c")],
                            alg.map("t_bq", scal.DivExpr(_A("t_big"), _C("1073741824", Type.INT64)), alg.scan("t")))
        return p, {3: {"numgroups": 2100}}
    if name == "f_join_neg":
        # unique join (32-bit tags) on a key that includes -1, the key stored
        # in the table's extra slot; grouped by the build key (group-join)
        p = alg.aggregation(["uk"], [(Reduction.SUM, "t_val", "s"), (Reduction.COUNT, "", "c")],
                            alg.join(("uk", "t_grp"),
                                     alg.map("uk", scal.SubExpr(scal.DivExpr(_A("u_id"), _C("3", Type.INT)),
                                                                _C("1", Type.INT)), alg.scan("u")),
                                     alg.scan("t")))
        return p, {4: {"multimatch": False}}
    if name == "f_join_packed":
        # unique join on a key that includes -1 (the extra slot) carrying one
        # 4-byte build column: tag and column share one 64-bit slot
        p = alg.aggregation(["t_cat"], [(Reduction.SUM, "u_grp", "s"), (Reduction.COUNT, "", "c"),
                                        (Reduction.MIN, "u_grp", "lo"), (Reduction.MAX, "u_grp", "hi")],
                            alg.join(("uk", "t_grp"),
                                     alg.map("uk", scal.SubExpr(scal.DivExpr(_A("u_id"), _C("3", Type.INT)),
                                                                _C("1", Type.INT)), alg.scan("u")),
                                     alg.scan("t")))
        return p, {4: {"multimatch": False}}
    if name == "f_anti_neg":
        p = alg.aggregation(["t_cat"], [(Reduction.COUNT, "", "c")],
                            alg.antijoin(("uk", "t_grp"), None,
                                         alg.map("uk", scal.SubExpr(scal.DivExpr(_A("u_id"), _C("3", Type.INT)),
                                                                    _C("1", Type.INT)), alg.scan("u")),
                                         alg.scan("t")))
        return p, {4: {"multimatch": False}}
    if name == "f_cross":
        # TPC-H Q11 shape (query/plans/tpch11.py): groups whose sum exceeds
        # a fraction of the total, via a nested-loop join with a one-row side
        p = alg.projection(["t_grp", "sv"],
                           alg.crossjoin(scal.LargerExpr(_A("sv"), _A("lim")),
                                         alg.map("lim", scal.MulExpr(_A("tot"), _C("0.02", Type.DOUBLE)),
                                                 alg.aggregation([], [(Reduction.SUM, "t_f", "tot")],
                                                                 alg.scan("t"))),
                                         alg.aggregation(["t_grp"], [(Reduction.SUM, "t_f", "sv")],
                                                         alg.scan("t"))))
        return p, {6: {"numgroups": 128}}
    if name == "f_cross_pairs":
        # a ten-row inner side, string condition, counted per probe category
        p = alg.aggregation(["t_cat"], [(Reduction.COUNT, "", "c"), (Reduction.SUM, "u_w", "sw")],
                            alg.crossjoin(scal.AndExpr(scal.LargerExpr(_A("u_w"), _A("t_f")),
                                                       scal.EqualsExpr(_A("u_name"), _C("echo", Type.STRING))),
                                          alg.selection(scal.SmallerExpr(_A("u_id"), _C("90", Type.INT)),
                                                        alg.scan("u")),
                                          alg.scan("t")))
        return p, {2: {"selectivity": 0.05}}
    if name == "f_like_patterns":
        # LIKE patterns of every shape the matcher generator distinguishes:
        # inner literals (short, across words, longer than 8 bytes, several
        # in order), prefix, suffix, prefix + suffix, exact, '%' alone, the
        # empty pattern, and one with '_' (general matcher); counted per category
        pats = ["%ha%", "%lpha g%", "%charlie delta%", "alpha%", "%echo", "golf", "%a%a%a%", "a%o",
                "%", "", "%_lf%", "%ec%ho%", "%juliet juliet%", "%o%o"]
        node = alg.scan("t")
        aggs = []
        for i, pat in enumerate(pats):
            node = alg.map("m%d" % i, scal.CaseExpr([(scal.LikeExpr(_A("t_txt"), _C(pat, Type.STRING)),
                                                      _C("1", Type.INT))], _C("0", Type.INT)), node)
            aggs.append((Reduction.SUM, "m%d" % i, "n%d" % i))
        return alg.aggregation(["t_cat"], aggs, node), {}
    if name == "f_outer":
        # TPC-H Q13 shape (query/plans/tpch13.py): probe rows without a
        # match count 0 through COUNT(build attribute)
        p = alg.aggregation(["t_grp"], [(Reduction.COUNT, "u_id", "c"), (Reduction.SUM, "u_w", "sw"),
                                        (Reduction.MAX, "u_w", "mw"), (Reduction.COUNT, "", "n")],
                            alg.outerjoin([("u_grp", "t_grp")], None,
                                          alg.selection(scal.LargerExpr(_A("u_w"), _C("2.5", Type.DOUBLE)),
                                                        alg.scan("u")),
                                          alg.scan("t")))
        return p, {5: {"numgroups": 128}}
    if name == "f_outer_project":
        # outer join with an extra condition, null build attributes materialised as 0 / ''
        p = alg.projection(["t_id", "u_id", "u_w", "u_name"],
                           alg.outerjoin([("u_grp", "t_grp")], scal.LargerExpr(_A("u_w"), _C("4.0", Type.DOUBLE)),
                                         alg.scan("u"),
                                         alg.selection(scal.SmallerExpr(_A("t_id"), _C("400", Type.INT)),
                                                       alg.scan("t"))))
        return p, {}
    if name == "f_temp":
        # TPC-H Q15 shape (query/plans/tpch15.py): a temp table of per-group
        # sums, the maximum over one scan of it joined back to a second scan,
        # then to the dimension
        p = [alg.createtemp("tsum", alg.aggregation(["t_grp"], [(Reduction.SUM, "t_val", "sv"),
                                                               (Reduction.COUNT, "", "c")], alg.scan("t"))),
             alg.projection(["u_id", "u_name", "t_grp", "c", "best"],
                            alg.join(("t_grp", "u_grp"),
                                     alg.join(("best", "sv"),
                                              alg.aggregation([], [(Reduction.MAX, "sv", "best")],
                                                              alg.scan("tsum", "r1")),
                                              alg.scan("tsum", "r2")),
                                     alg.scan("u")))]
        return p, {2: {"numgroups": 100}, 7: {"multimatch": False}, 9: {"multimatch": False}}
    if name == "f_temp_strings":
        # strings and a filter through a temp table, grouped by the string
        p = [alg.createtemp("tlike", alg.projection(["t_id", "t_txt", "t_val"],
                                                    alg.selection(scal.LikeExpr(_A("t_txt"), _C("%golf%", Type.STRING)),
                                                                  alg.scan("t")))),
             alg.aggregation(["t_txt"], [(Reduction.SUM, "t_val", "s"), (Reduction.COUNT, "", "c")],
                             alg.selection(scal.LargerExpr(_A("t_id"), _C("10", Type.INT)), alg.scan("tlike")))]
        return p, {2: {"selectivity": 0.05}, 7: {"numgroups": 64}}
    raise KeyError(name)


# ------------------------------------------------------------- string plans
# Every string over {a, b, c} of at most 5 bytes, and every string over
# {a, %, _} of at most 3 bytes, each 4 times with SUBSTRING arguments in
# [-2, 7]; LIKE patterns over {a, b, %, _} of at most 3 bytes and longer ones
# of the shapes the matcher generator distinguishes.

STRING_SCHEMA = OrderedDict([
    ("w", OrderedDict([("w_id", Type.INT), ("w_s", Type.STRING), ("w_f", Type.INT), ("w_l", Type.INT)])),
])


def _allStrings(alphabet, maxLen):
    out = [b""]
    for n in range(1, maxLen + 1):
        out += [bytes(t) for t in itertools.product(alphabet, repeat=n)]
    return out


LIKE_PATTERNS = [p.decode() for p in _allStrings(b"ab%_", 3)[1:]] + [
    "", "ab%b", "a%%c", "a%%", "%%c", "%%", "%%%", "abab", "%abc", "abc%", "a%b%c", "%a%b%", "%ab%ba%",
    "_%_", "%_a_%", "ab%_b", "a_%_c", "%c%", "cab%", "%bca", "a%a%a", "%aaa%", "b%%a", "%a%%b%", "_b_",
    "abcab", "%bcab", "ab%ab", "%%a%", "c%c%", "%%ab%%"]


def stringData(copies=4, seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    strs = (_allStrings(b"abc", 5) + _allStrings(b"a%_", 3)[1:]) * copies
    n = len(strs)
    w = OrderedDict()
    w["w_id"] = (Type.INT, np.arange(n, dtype=np.int32))
    w["w_s"] = (Type.STRING, strs)
    w["w_f"] = (Type.INT, rng.integers(-2, 8, size=n).astype(np.int32))
    w["w_l"] = (Type.INT, rng.integers(-2, 8, size=n).astype(np.int32))
    return _asAccessor(OrderedDict([("w", w)]))


def stringRows(acc, n):
    """The first n rows of a string table (an empty or one-row case)."""
    offs, chars = acc.column("w", "w_s")
    strs = [bytes(chars[offs[i]:offs[i + 1]]) for i in range(n)]
    w = OrderedDict()
    w["w_id"] = (Type.INT, np.asarray(acc.column("w", "w_id"))[:n])
    w["w_s"] = (Type.STRING, strs)
    w["w_f"] = (Type.INT, np.asarray(acc.column("w", "w_f"))[:n])
    w["w_l"] = (Type.INT, np.asarray(acc.column("w", "w_l"))[:n])
    return _asAccessor(OrderedDict([("w", w)]))


def _stringPlan(name, alg):
    if name == "s_like":
        # bit k of column b<j> = w_s LIKE LIKE_PATTERNS[16 j + k]
        node = alg.scan("w")
        cols = ["w_id"]
        for j in range(0, len(LIKE_PATTERNS), 16):
            expr = None
            for k, pat in enumerate(LIKE_PATTERNS[j:j + 16]):
                term = scal.CaseExpr([(scal.LikeExpr(_A("w_s"), _C(pat, Type.STRING)), _C(str(1 << k), Type.INT))],
                                     _C("0", Type.INT))
                expr = term if expr is None else scal.AddExpr(expr, term)
            node = alg.map("b%d" % (j // 16), expr, node)
            cols.append("b%d" % (j // 16))
        return alg.projection(cols, node), {}
    if name == "s_substr":
        # per-row SUBSTRING arguments; equality and IN on the substring
        node = alg.map("sub", scal.SubstringExpr(_A("w_s"), _A("w_f"), _A("w_l")), alg.scan("w"))
        node = alg.map("eq", scal.CaseExpr([(scal.EqualsExpr(_A("sub"), _C("ab", Type.STRING)), _C("1", Type.INT))],
                                           _C("0", Type.INT)), node)
        node = alg.map("inn", scal.CaseExpr([(scal.InExpr(_A("sub"), [_C(t, Type.STRING) for t in
                                                                       ("", "a", "%_", "cab")]),
                                              _C("1", Type.INT))], _C("0", Type.INT)), node)
        return alg.projection(["w_id", "sub", "eq", "inn"], node), {}
    if name == "s_filter":
        # string predicates as a filter: IN, NOT, OR, equality of two strings
        cond = scal.OrExpr(scal.InExpr(_A("w_s"), [_C(t, Type.STRING) for t in ("abc", "", "a%")]),
                           scal.AndExpr(scal.NotExpr(scal.LikeExpr(_A("w_s"), _C("%b%", Type.STRING))),
                                        scal.EqualsExpr(scal.SubstringExpr(_A("w_s"), _C("1", Type.INT),
                                                                           _C("2", Type.INT)),
                                                        scal.SubstringExpr(_A("w_s"), _C("2", Type.INT),
                                                                           _C("2", Type.INT)))))
        return alg.projection(["w_id", "w_s"], alg.selection(cond, alg.scan("w"))), {2: {"selectivity": 0.2}}
    raise KeyError(name)


STRING_PLANS = ["s_like", "s_substr", "s_filter"]

FEATURE_PLANS = ["f_scan_minmax", "f_groupby_hashed", "f_groupby_wide", "f_like_project", "f_join_unique",
                 "f_join_unique_lowest",
                 "f_groupjoin", "f_join_multi", "f_semi", "f_anti_multi", "f_temp", "f_temp_strings", "f_groupby_64",
                 "f_groupby_int64", "f_outer", "f_outer_project", "f_join_neg", "f_anti_neg",
                 "f_join_packed", "f_cross", "f_cross_pairs", "f_like_patterns", "f_join_string"]

# name -> (workload or "feature", options)
PLANS = OrderedDict()
PLANS["c0"] = ("c0", {})
PLANS["c1"] = ("c1", {})
PLANS["c2_none"] = ("c2", {})
PLANS["c2_refill"] = ("c2", {"buffers": [3]})
PLANS["c2_pushdown"] = ("c2", {"pushdown": True})
PLANS["c2_both"] = ("c2", {"buffers": [3], "pushdown": True})
PLANS["c2_refill_reg"] = ("c2", {"buffers": [3], "buffer_type": "reg"})
PLANS["c2_both_reg"] = ("c2", {"buffers": [3], "pushdown": True, "buffer_type": "reg"})
PLANS["c1_refill"] = ("c1", {"buffers": [4]})
# the c3 star join reads its base columns with the nontemporal hint: measured
# at 60M sales rows, the plan 0.805 -> 0.766 ms (the sales kernel 0.514 ->
# 0.500 ms); c1 and c2 ran slower with it (c1_refill's fact kernel 0.031 ->
# 0.035 ms, c2's probe 0.217 -> 0.229 ms), and hinting only the loads at the
# top of a pipeline gained nothing on c3
C3 = {"nontemporal": True}
PLANS["c3"] = ("c3", dict(C3))
PLANS["c3_refill"] = ("c3", dict(C3, buffers=[6]))
PLANS["cal"] = ("cal", {})
for _n in FEATURE_PLANS:
    PLANS[_n] = ("feature", {})
for _n in STRING_PLANS:
    PLANS[_n] = ("strings", {})
PLANS["f_join_multi_pushdown"] = ("feature", {"pushdown": True, "base": "f_join_multi"})
PLANS["f_join_string_pushdown"] = ("feature", {"pushdown": True, "base": "f_join_string"})
# string keys hashed and compared by the whole wavefront (StringHash.wave)
PLANS["f_join_string_ws"] = ("feature", {"wave_strings": True, "base": "f_join_string"})
PLANS["f_join_string_pushdown_ws"] = ("feature", {"pushdown": True, "wave_strings": True, "base": "f_join_string"})
PLANS["f_like_project_refill"] = ("feature", {"buffers": [2], "base": "f_like_project"})
PLANS["f_like_project_refill_reg"] = ("feature", {"buffers": [2], "base": "f_like_project", "buffer_type": "reg"})


# TPC-H Q1-Q22 (dogqc_amd/tpch.py)
# Lane refill (LDS queue) at one operator of some queries. Q10 refills after
# its lineitem join (opId 9), the placement of the reference's balanced Q10
# (sample/tpch10_balanced.cu:356-457): tuned at SF1 on MI355X, 0.565 ms
# without it, 0.237 ms with it (0.242 ms with the queue across two launches).
# The others come from timing a refill at every selection and join with
# >= 100 K input rows of every query, SF1, 512x1, untuned, results equal to
# the plan without it, and keeping those 8% or more faster: Q3 0.165 ->
# 0.143 ms, Q5 0.200 -> 0.183, Q7 0.197 -> 0.179, Q8 0.204 -> 0.174,
# Q9 0.389 -> 0.342, Q14 0.081 -> 0.073, Q19 0.240 -> 0.179.
TPCH_REFILL = {3: 8, 5: 13, 7: 13, 8: 13, 9: 6, 10: 9, 14: 3, 19: 2}
for _q in range(1, 23):
    PLANS["tpch%d" % _q] = ("tpch", {"query": _q, "buffers": [TPCH_REFILL[_q]]} if _q in TPCH_REFILL
                            else {"query": _q})
# the same queries without the refill (a pipeline with a refill point keeps
# its group-by out of the partitioned tier)
for _q in TPCH_REFILL:
    PLANS["tpch%d_plain" % _q] = ("tpch", {"query": _q})

# plans translated with the partitioned unique build (tuning part_build)
PB_FEATURES = ["f_groupjoin", "f_join_unique_lowest", "f_semi", "f_temp", "f_join_neg", "f_anti_neg", "f_join_packed"]
PB_TPCH = [3, 5, 9, 18, 21]
for _n in PB_FEATURES:
    PLANS[_n + "_pb"] = ("feature", {"part_build": True, "base": _n})
for _q in PB_TPCH:
    PLANS["tpch%d_pb" % _q] = ("tpch", {"query": _q, "part_build": True})
PLANS["c1_pb"] = ("c1", {"part_build": True})
PLANS["c1_refill_pb"] = ("c1", {"buffers": [4], "part_build": True})
PLANS["c3_refill_pb"] = ("c3", dict(C3, buffers=[6], part_build=True))
# batched serving plans (dogqc_amd/batching.py): one run for several queries
for _c in ("c0", "c1", "c2", "c3"):
    PLANS["%s_batch" % _c] = ("batch", {"config": _c})

# Q3 with a lane-refill buffer after the customer selection and a
# lane-activity profile after the lineitem selection (profiler.py usage)
PLANS["tpch3_profiled"] = ("tpch", {"query": 3, "buffers": [2], "profilers": [7]})
# TPC-H Q10 as the reference's sample profiler reports ran it: a lane-activity
# profile after the lineitem join (opId 9), without and with lane refill there
# Q1 without its group-count hint: its four groups then run through the
# partitioned group-by, whose partitions overflow into the global table
PLANS["tpch1_unhinted"] = ("tpch", {"query": 1, "no_group_hints": True})
PLANS["tpch10_p9"] = ("tpch", {"query": 10, "profilers": [9]})
PLANS["tpch10_lr9_p9"] = ("tpch", {"query": 10, "buffers": [9], "profilers": [9]})
# TPC-H queries with string join or group keys, their strings hashed by the
# whole wavefront (StringHash.wave) instead of lane by lane
WAVE_STRING_TPCH = [4, 5, 7, 9, 10, 12, 16, 18, 21, 22]
for _q in WAVE_STRING_TPCH:
    PLANS["tpch%d_ws" % _q] = ("tpch", {"query": _q, "wave_strings": True})
PLANS["tpch10_ws_lr9_p9"] = ("tpch", {"query": 10, "buffers": [9], "profilers": [9], "wave_strings": True})
# lane refill with the register buffer (BufferType.REG) in place of the LDS queue
PLANS["tpch10_lr9reg_p9"] = ("tpch", {"query": 10, "buffers": [9], "profilers": [9], "buffer_type": "reg"})
PLANS["tpch3_profiled_reg"] = ("tpch", {"query": 3, "buffers": [2], "profilers": [7], "buffer_type": "reg"})
PLANS["c1_refill_reg"] = ("c1", {"buffers": [4], "buffer_type": "reg"})
PLANS["c3_refill_reg"] = ("c3", dict(C3, buffers=[6], buffer_type="reg"))
# lane refill across a kernel boundary (BufferType.GLOBAL): survivors queued
# in device memory by a first launch, the rest of the pipeline run over the
# queue by a second
PLANS["c1_refill_g"] = ("c1", {"buffers": [4], "buffer_type": "global"})
PLANS["c2_refill_g"] = ("c2", {"buffers": [3], "buffer_type": "global"})
PLANS["c3_refill_g"] = ("c3", dict(C3, buffers=[6], buffer_type="global"))
PLANS["tpch10_lrg9_p9"] = ("tpch", {"query": 10, "buffers": [9], "profilers": [9], "buffer_type": "global"})
PLANS["tpch3_profiled_g"] = ("tpch", {"query": 3, "buffers": [2], "profilers": [7], "buffer_type": "global"})


def build(name, acc):
    """(alg, plan, cfg, options) for a named plan over an accessor."""
    kind, opts = PLANS[name]
    if kind == "tpch":
        from dogqc_amd import tpch
        alg, p, cfg = tpch.QUERIES[opts["query"]](acc)
        if opts.get("no_group_hints"):
            # group counts estimated from the inputs, as for a plan whose cfg
            # gives none
            cfg = {k: {kk: vv for kk, vv in v.items() if kk != "numgroups"} for k, v in cfg.items()}
        return alg, p, cfg, opts
    if kind == "batch":
        from dogqc_amd import batching
        alg, p, cfg, knobsOf = batching.plan(opts["config"], acc)
        merged = dict(knobsOf)
        merged.update(opts)
        return alg, p, cfg, merged
    if kind == "feature":
        alg = RelationalAlgebra(acc)
        p, cfg = _featurePlan(opts.get("base", name), alg)
        return alg, p, cfg, opts
    if kind == "strings":
        alg = RelationalAlgebra(acc)
        p, cfg = _stringPlan(name, alg)
        return alg, p, cfg, opts
    alg, p, cfg = workloads.plan(kind, acc)
    return alg, p, cfg, opts


def emptyAccessor(name):
    kind, _ = PLANS[name]
    if kind == "tpch":
        from dogqc_amd import tpch
        return tpch.SchemaAccessor(tpch.schema())
    if kind == "batch":
        from dogqc_amd import batching
        cfgName = PLANS[name][1]["config"]
        return batching.withBatch(workloads.emptyAccessor(cfgName),
                                  {t: np.zeros(1, dtype=np.int64) for t in batching.SCALED[cfgName]})
    if kind == "strings":
        return stringRows(stringData(1), 0)
    return featureEmpty() if kind == "feature" else workloads.emptyAccessor(kind)


# plans the autotuner explores compile-time variants of (built ahead of time)
TUNE_PLANS = ["c0", "c1", "c1_refill", "c2_none", "c2_refill", "c2_pushdown", "c2_both", "c3", "c3_refill",
              "c1_refill_reg", "c2_refill_reg", "c2_both_reg", "c3_refill_reg", "c1_refill_g", "c3_refill_g"]


def translate(name, acc=None, decimalRepr=CType.FP64, block=None, rowsPerThread=None):
    """Resolve and translate a named plan; returns the HipCompiler with
    generated source (call .compile() to get the module).  ``block`` and
    ``rowsPerThread`` override the plan's launch-shape knobs."""
    acc = acc or emptyAccessor(name)
    alg, p, cfg, opts = build(name, acc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    with knobs(pushdown=opts.get("pushdown", False), block=block or opts.get("block"),
               threshold=opts.get("threshold"), rowsPerThread=rowsPerThread,
               partBuild=opts.get("part_build", False),
               bufferType={"reg": BufferType.REG, "global": BufferType.GLOBAL}.get(opts.get("buffer_type"),
                                                                                    BufferType.SMEM),
               waveStrings=opts.get("wave_strings", False),
               nontemporal=opts.get("nontemporal", False),
               joinFilter=opts.get("join_filter", 0)):
        comp = HipCompiler(alg, "gfx950", decimalRepr)
        comp.setBuffers(opts.get("buffers", []))
        comp.setProfilers(opts.get("profilers", []))
        tp = alg.translateToCompilerPlan(plan, comp)
        comp.gencode(tp)
    comp.planName = name
    comp.resolvedPlan = plan
    return comp
