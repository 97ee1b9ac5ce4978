"""Batched plans (dogqc_amd/batching.py) on the CPU oracle: one evaluation
of the batched plan over the largest prefix, combined per member, equals
the unbatched plan evaluated on each member's own prefix."""
from collections import OrderedDict

import numpy as np
import pytest

from dogqc_amd import batching, workloads
from dogqc_amd.dbio import MemoryAccessor
from oracle import planeval
from tests.helpers import assertSameResult, toleranceColumns

SIZES = {"c0": {"cfg0": 4000}, "c1": {"dim": 600, "fact": 5000}, "c2": {"build": 300, "probe": 6000},
         "c3": {"sales": 6000, "item": 900, "cust": 120, "nation": 25}}
SCALES = [0.1, 0.35, 0.35, 0.8, 1.0]


def prefixes(config, scales=SCALES):
    return OrderedDict((t, [max(1, int(round(SIZES[config][t] * s))) for s in scales])
                       for t in batching.SCALED[config])


def oracleMember(config, acc, pref, k):
    """The unbatched plan on member k's prefixes: (result, temps, plan)."""
    pacc = batching.prefixAccessor(acc, {t: v[k] for t, v in pref.items()})
    alg, p, cfg = workloads.plan(config, pacc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    ev = planeval.Evaluator(pacc)
    res = ev.run(plan)
    temps = {name: cols for name, (cols, n) in ev.temps.items()}
    return res, temps, plan


def batchedOracle(config, acc, pref):
    bacc = batching.batchAccessor(acc, pref)
    alg, p, cfg, _ = batching.plan(config, bacc)
    plan = alg.resolveAlgebraPlan(p, cfg)
    ev = planeval.Evaluator(bacc)
    res = ev.run(plan)
    return res, {name: cols for name, (cols, n) in ev.temps.items()}


@pytest.mark.parametrize("config", ["c0", "c1", "c2", "c3"])
def test_batched_plan_combines_to_each_member(config):
    acc = MemoryAccessor(workloads.generate(config, SIZES[config], seed=9))
    pref = prefixes(config)
    got, gotTemps = batchedOracle(config, acc, pref)
    assert int(np.max(got["seg"])) <= len(SCALES) - 1
    for k in range(len(SCALES)):
        want, wantTemps, plan = oracleMember(config, acc, pref, k)
        res, temps = batching.combine(config, got, gotTemps, k)
        assertSameResult(res, want, toleranceColumns(plan), rtol=1e-12, label="%s member %d" % (config, k))
        for name, cols in temps.items():
            w = OrderedDict((c, wantTemps[name][c]) for c in cols)
            assertSameResult(cols, w, {"revenue"}, rtol=1e-12, label="%s member %d %s" % (config, k, name))


def test_single_member_batch_is_the_query():
    acc = MemoryAccessor(workloads.generate("c1", SIZES["c1"], seed=3))
    pref = prefixes("c1", [0.6])
    got, _ = batchedOracle("c1", acc, pref)
    assert set(np.unique(got["seg"])) <= {0}
    want, _, plan = oracleMember("c1", acc, pref, 0)
    res, _ = batching.combine("c1", got, {}, 0)
    assertSameResult(res, want, toleranceColumns(plan), rtol=1e-12)
