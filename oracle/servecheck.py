"""Oracle results of serving queries (c0-c3 over row prefixes of resident
tables) and their comparison with a served result.

TEST INFRASTRUCTURE ONLY: used by tests/test_serving.py and by the serving
leg of bench.py to check a sample of served queries after the timed runs,
never by the product path.
"""
from collections import OrderedDict

import numpy as np

from oracle import coracle


def oracleResult(tables, config, scale):
    """The C oracle's result of one query, computed over host copies of the
    row prefixes the query scans."""
    from dogqc_amd import serving, workloads
    rows = serving.prefixRows(tables, config, scale)
    acc = tables.hostAccessor(rows)
    cols = {t: {a: acc.column(t, a) for a in acc.schema[t]["attributes"]} for t in rows}
    if config == "c0":
        ref = coracle.c0(cols["cfg0"])
        out = OrderedDict((k, []) for k in ("g1", "g2", "sum_qty", "sum_base_price", "sum_disc_price",
                                             "sum_charge", "count_order"))
        for (g1, g2), v in sorted(ref.items()):
            for k, x in zip(out, (g1, g2) + tuple(v[:4]) + (v[5],)):
                out[k].append(x)
        return OrderedDict((k, np.array(v)) for k, v in out.items())
    if config == "c1":
        keys, sums = coracle.c1(cols["dim"], cols["fact"])
        return OrderedDict([("d_key", keys), ("revenue", sums)])
    if config == "c2":
        cnt, s = coracle.c2(cols["build"], cols["probe"])
        return OrderedDict([("cnt", np.array([cnt], dtype=np.int32)), ("sum_prod", np.array([s], dtype=np.int64))])
    return coracle.c3(cols, workloads.C3_WORD)


def matchesOracle(config, got, want):
    """A served result against the oracle's: keys and counts bit-exact,
    float64 sums within 1e-9 relative.  The oracle's c0 tuple carries fewer
    columns than the plan, so c0 compares the oracle's columns only."""
    from dogqc_amd import serving
    if config == "c0":
        sub = OrderedDict((k, np.asarray(got[k])) for k in want)
        sub["count_order"] = sub["count_order"].astype(np.int64)
        want = OrderedDict(want)
        want["count_order"] = want["count_order"].astype(np.int64)
        want["g1"], want["g2"] = want["g1"].astype(np.int8), want["g2"].astype(np.int8)
        return serving.sameResult("c0", sub, want)
    if config == "c2":
        return int(got["cnt"][0]) == int(want["cnt"][0]) and int(got["sum_prod"][0]) == int(want["sum_prod"][0])
    return serving.sameResult(config, got, want)


def check(tables, config, scale, got):
    """True when `got` (a served result) equals the oracle's."""
    return bool(matchesOracle(config, got, oracleResult(tables, config, scale)))
