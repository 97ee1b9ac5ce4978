"""Serving loop (dqc_serve_*) and device-side table generation, on the GPU.

  * the generator kernels against a numpy restatement of their formulas;
  * every query of a mixed stream gives the same result alone (one slot)
    and in the mix (eight slots), and the CPU oracle's result;
  * admission never runs more queries at once than their workspaces fit in
    the arena, a query larger than the arena fails cleanly, and a query
    whose tables overflow is re-admitted with larger ones.
"""
import ctypes as C

import numpy as np
import pytest

from dogqc_amd import datagen, runtime as rt, serving, workloads
from oracle import servecheck

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def rnd(seed, column, rows):
    """The generator's random word of (seed, column, row) for a row array."""
    rows = rows.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) ^ (np.uint64(column) * np.uint64(0x9E3779B97F4A7C15))
        z = z + (rows + np.uint64(1)) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(r, lo, hi):
    rng = np.uint64(hi - lo + 1)
    return lo + (((r >> np.uint64(32)) * rng) >> np.uint64(32)).astype(np.int64)


def prefixKeys(r, n, nkeys):
    """dqc_gen_prefix_key: the key of row floor(u * ceil(nkeys * (i+1) / n))
    of a table keyed by the affine permutation of permParams(nkeys)."""
    i = np.arange(n, dtype=np.int64)
    rows = np.clip((nkeys * (i + 1) + n - 1) // n, 1, nkeys).astype(np.uint64)
    row = ((r >> np.uint64(32)) * rows) >> np.uint64(32)
    a, c = datagen.permParams(nkeys)
    return ((np.uint64(a) * row + np.uint64(c)) % np.uint64(nkeys)).astype(np.int32)


_TABLES = {}


def tables(scale):
    if scale not in _TABLES:
        _TABLES[scale] = datagen.ResidentTables(scale=scale, seed=77)
    return _TABLES[scale]


def test_datagen_matches_restatement():
    print("gpu case: datagen restatement", flush=True)
    T = tables(0.02)
    seed = T.seed
    cid = T.columnIds
    n = T.rows["cfg0"]
    rows = np.arange(n)
    assert cid[("cfg0", "g1")] == 16 and cid[("cfg0", "qty")] == 48
    g1 = T.store[("cfg0", "g1", None)].toArray(np.int8, n)
    assert np.array_equal(g1, uniform(rnd(seed, 16, rows), 0, 2).astype(np.int8))
    qty = T.store[("cfg0", "qty", None)].toArray(np.int32, n)
    assert np.array_equal(qty, uniform(rnd(seed, 48, rows), 1, 50).astype(np.int32))
    price = T.store[("cfg0", "price", None)].toArray(np.float64, n)
    assert np.array_equal(price, uniform(rnd(seed, cid[("cfg0", "price")], rows), 90000, 10500000)
                          .astype(np.float64) / 100.0)
    date = T.store[("cfg0", "date", None)].toArray(np.uint32, n)
    assert np.array_equal(date, workloads.dateTable()[uniform(rnd(seed, cid[("cfg0", "date")], rows), 0, 2556)])
    nf = T.rows["fact"]
    fk = T.store[("fact", "f_fk", None)].toArray(np.int32, nf)
    assert np.array_equal(fk, prefixKeys(rnd(seed, cid[("fact", "f_fk")], np.arange(nf)), nf, T.rows["dim"]))
    # every prefix of a referencing table stays inside the same fraction of
    # the referenced one: a query at any scale finds all its keys
    for ref, t, a in (("dim", "fact", "f_fk"), ("item", "sales", "s_item"), ("cust", "sales", "s_cust")):
        keys = T.store[(ref, {"dim": "d_key", "item": "i_key", "cust": "c_key"}[ref], None)].toArray(
            np.int32, T.rows[ref])
        fks = T.store[(t, a, None)].toArray(np.int32, T.rows[t])
        for f in (0.01, 0.1, 0.37, 1.0):
            nr, nt = max(1, int(round(T.rows[ref] * f))), max(1, int(round(T.rows[t] * f)))
            assert np.isin(fks[:nt], keys[:nr]).all(), (t, a, f)
    # permutations are permutations
    for t, a in (("dim", "d_key"), ("item", "i_key"), ("cust", "c_key"), ("nation", "n_key")):
        k = T.store[(t, a, None)].toArray(np.int32, T.rows[t])
        assert np.array_equal(np.sort(k), np.arange(T.rows[t], dtype=np.int32)), t
    # Zipf keys: inverse CDF of 1/rank over 100k values
    nb = T.rows["build"]
    bk = T.store[("build", "b_key", None)].toArray(np.int32, nb)
    ranks = np.arange(1, 100001, dtype=np.float64)
    cdf = np.cumsum(1.0 / ranks)
    cdf /= cdf[-1]
    a, c = datagen.permParams(100000)
    values = ((a * np.arange(100000, dtype=np.int64) + c) % 100000).astype(np.int32)
    u = (rnd(seed, cid[("build", "b_key")], np.arange(nb)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.minimum(np.searchsorted(cdf, u, side="right"), 99999)
    assert np.array_equal(bk, values[r])
    # text rows: 5 vocabulary words, single spaces, space padded to 55 bytes
    ni = T.rows["item"]
    offs = T.store[("item", "i_text", "offset")].toArray(np.uint64, ni + 1)
    chars = T.store[("item", "i_text", "char")].toArray(np.uint8, ni * 55).tobytes()
    assert np.array_equal(offs, np.arange(ni + 1, dtype=np.uint64) * 55)
    tc = cid[("item", "i_text")]
    idx = np.stack([uniform(rnd(seed, tc + j, np.arange(ni)), 0, 91) for j in range(5)], axis=1)
    for i in range(0, ni, max(1, ni // 500)):
        words = chars[55 * i:55 * (i + 1)].decode().rstrip(" ").split(" ")
        assert words == [workloads.VOCAB[k] for k in idx[i]]


def test_alone_equals_mixed_and_oracle():
    print("gpu case: serving alone vs mixed", flush=True)
    T = tables(0.02)
    plans = serving.PlanSet()
    queries = serving.makeMix(40, seed=11, scaleRange=(0.001, 0.02))
    arena = 4 << 30
    keep = set(range(len(queries)))
    recsA, srvA = serving.run(T, plans, queries, 1, arena, keep=keep)
    recsM, srvM = serving.run(T, plans, queries, 8, arena, keep=keep)
    overlap = 0
    for i, q in enumerate(queries):
        a = serving.resultOf(plans, q.config, recsA[i])
        m = serving.resultOf(plans, q.config, recsM[i])
        assert serving.sameResult(q.config, a, m), "query %d %s x%.4f alone != mixed" % (i, q.config, q.scale)
        want = servecheck.oracleResult(T, q.config, q.scale)
        assert servecheck.matchesOracle(q.config, m, want), "query %d %s x%.4f != oracle" % (i, q.config, q.scale)
        overlap += sum(1 for j, r in enumerate(recsM) if j != i and r.admit_ns < recsM[i].done_ns
                       and recsM[i].admit_ns < r.done_ns)
    assert overlap > 0          # the mixed run did run queries concurrently
    srvA.destroy()
    srvM.destroy()


def test_admission_respects_arena_and_rejects_oversized():
    print("gpu case: serving admission", flush=True)
    T = tables(0.02)
    plans = serving.PlanSet()
    queries = [serving.Query("c1", 0.02) for _ in range(12)]
    need = serving.workspaceBytes(T, plans, "c1", 0.02)
    rows = serving.prefixRows(T, "c1", 0.02)
    mod, tun = plans.pick("c1", max(rows.values()))
    r, cols = serving.binding(mod, T, rows)
    arena = int(need * 2.5)     # room for two at a time
    recs, srv = serving.run(T, plans, queries, 8, arena)
    events = sorted([(rc.admit_ns, 1) for rc in recs] + [(rc.done_ns, -1) for rc in recs])
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert peak <= 2
    srv.destroy()
    small = serving.Server(4, need // 2)
    qid = small.submit(mod, r, cols, tun)
    small.start()
    small.wait(60000)
    rec = small.record(qid)
    assert rec.state == 3 and rec.error == -5      # DQC_SERVE_FAILED, DQC_ENOMEM
    small.destroy()


def test_overflow_is_readmitted():
    print("gpu case: serving retry", flush=True)
    T = tables(0.02)
    plans = serving.PlanSet(tuned=False)
    t = rt.newTuning()
    t.ht_capacity[0] = 16
    t.ht_capacity[1] = 8
    plans.variants["c1"] = ((plans.modules["c1"], t), {})
    q = serving.Query("c1", 0.02)
    recs, srv = serving.run(T, plans, [q], 2, 1 << 30, keep={0})
    assert recs[0].retries > 0
    got = serving.resultOf(plans, "c1", recs[0])
    assert servecheck.matchesOracle("c1", got, servecheck.oracleResult(T, "c1", 0.02))
    srv.destroy()


def test_batched_queries_equal_alone_and_oracle():
    """With max_batch, waiting queries of one config run as one batched plan
    (all arrive at once here, so batches form); each member's combined result
    equals its run alone and the CPU oracle, and kept batch workspaces are
    released through their members."""
    print("gpu case: serving batches", flush=True)
    T = tables(0.02)
    plans = serving.PlanSet()
    queries = serving.makeMix(32, seed=12, scaleRange=(0.001, 0.02))
    keep = set(range(len(queries)))
    recsA, srvA = serving.run(T, plans, queries, 1, 4 << 30, keep=keep)
    recsB, srvB = serving.run(T, plans, queries, 4, 4 << 30, keep=keep, max_batch=8)
    assert sum(1 for r in recsB if r.batch_size > 1) >= 8
    for i, q in enumerate(queries):
        a = serving.resultOf(plans, q.config, recsA[i])
        b = serving.resultOf(plans, q.config, recsB[i])
        assert serving.sameResult(q.config, a, b), "query %d %s x%.4f alone != batched (batch %d)" % (
            i, q.config, q.scale, recsB[i].batch_size)
        assert servecheck.matchesOracle(q.config, b, servecheck.oracleResult(T, q.config, q.scale)), \
            "query %d %s x%.4f batched != oracle" % (i, q.config, q.scale)
    for i in range(len(queries)):
        srvB.release(i)
    srvA.destroy()
    srvB.destroy()



def test_batch_min_share_runs_a_dominant_query_alone():
    """With batch_min_share, a batch runs only if its members other than the
    largest hold at least that share of the largest's driving rows; results
    stay the same."""
    print("gpu case: serving batch min share", flush=True)
    T = tables(0.02)
    plans = serving.PlanSet()
    queries = serving.makeMix(32, seed=17, scaleRange=(0.0005, 0.02))
    keep = set(range(len(queries)))
    recsA, srvA = serving.run(T, plans, queries, 1, 4 << 30, keep=keep)
    recsB, srvB = serving.run(T, plans, queries, 4, 4 << 30, keep=keep, max_batch=8, batch_min_share=1.0)
    driving = [max(serving.prefixRows(T, q.config, q.scale).values()) for q in queries]
    batches = {}
    for i, r in enumerate(recsB):
        if r.batch_size > 1:
            batches.setdefault((r.slot, r.admit_ns), []).append(i)
    assert batches, "no batch formed"
    for members in batches.values():
        rows = sorted(driving[i] for i in members)
        assert sum(rows[:-1]) >= rows[-1], [(queries[i].config, queries[i].scale) for i in members]
    for i, q in enumerate(queries):
        assert serving.sameResult(q.config, serving.resultOf(plans, q.config, recsA[i]),
                                  serving.resultOf(plans, q.config, recsB[i])), "query %d alone != batched" % i
    for i in range(len(queries)):
        srvB.release(i)
    srvA.destroy()
    srvB.destroy()


def test_batch_ratio_keeps_batches_of_similar_size():
    """With batch_ratio, a batch holds only queries whose driving row counts
    are within that factor of its first query's; results stay the same."""
    print("gpu case: serving batch ratio", flush=True)
    T = tables(0.02)
    plans = serving.PlanSet()
    queries = serving.makeMix(32, seed=13, scaleRange=(0.0005, 0.02))
    keep = set(range(len(queries)))
    recsA, srvA = serving.run(T, plans, queries, 1, 4 << 30, keep=keep)
    recsB, srvB = serving.run(T, plans, queries, 4, 4 << 30, keep=keep, max_batch=8, batch_ratio=2)
    driving = [max(serving.prefixRows(T, q.config, q.scale).values()) for q in queries]
    batches = {}
    for i, r in enumerate(recsB):
        if r.batch_size > 1:
            batches.setdefault((r.slot, r.admit_ns), []).append(i)
    assert batches, "no batch formed"
    for members in batches.values():
        assert len({queries[i].config for i in members}) == 1
        rows = [driving[i] for i in members]
        assert max(rows) <= 4 * min(rows), [(queries[i].config, queries[i].scale) for i in members]
    for i, q in enumerate(queries):
        assert serving.sameResult(q.config, serving.resultOf(plans, q.config, recsA[i]),
                                  serving.resultOf(plans, q.config, recsB[i])), "query %d alone != batched" % i
    for i in range(len(queries)):
        srvB.release(i)
    srvA.destroy()
    srvB.destroy()
