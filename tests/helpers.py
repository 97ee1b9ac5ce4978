"""Result comparison for parity tests: results are multisets of rows.

Integer, string and non-aggregated floating-point columns must match bit for
bit; floating-point SUM/AVG aggregates (whose GPU summation order is not
fixed) within ``rtol`` relative error (1e-9 by default) plus ``atol``.
"""
import numpy as np

import dogqc_amd.relationalAlgebra as ra
from dogqc_amd.relationalAlgebra import Reduction


def toleranceColumns(plan):
    """Names of the result columns that carry a floating-point SUM/AVG
    aggregate: the aggregate outputs themselves and anything derived from
    them (MIN/MAX over them, maps of them, their temp-table copies)."""
    import dogqc_amd.scalarAlgebra as scal
    tol = set()

    def uses(expr):
        if isinstance(expr, scal.AttrExpr):
            return expr.attr.id in tol
        return any(uses(c) for c in expr.children())

    def walk(node):
        for c in node.children():
            walk(c)
        if isinstance(node, ra.Scan) and node.isTempScan:
            for id, a in node.outRelation.items():
                if getattr(a, "sourceId", None) in tol:
                    tol.add(id)
        elif isinstance(node, ra.Map) and uses(node.expression):
            tol.add(node.mapAttr.id)
        elif isinstance(node, ra.Aggregation):
            for id, (inId, red) in node.aggregateTuples.items():
                if red in (Reduction.SUM, Reduction.AVG) or (red in (Reduction.MIN, Reduction.MAX) and inId in tol):
                    tol.add(id)
    for root in plan:
        walk(root)
    return {a.name for id, a in plan[-1].outRelation.items() if id in tol}


def _sortKey(cols, names):
    keys = []
    for n in reversed(names):
        c = cols[n]
        if c.dtype == object:
            _, inv = np.unique(c.astype(bytes), return_inverse=True) if len(c) else (None, np.zeros(0, int))
            keys.append(inv)
        else:
            keys.append(c)
    return np.lexsort(keys) if keys else np.arange(0)


def _widen(g, w):
    """Both arrays in one type wide enough for either's values: the wider
    float, or the wider integer (int64 for a signed/unsigned mix, which the
    result columns' value ranges fit)."""
    if g.dtype == w.dtype:
        return g, w
    if g.dtype.kind == "f" or w.dtype.kind == "f":
        t = np.float64 if max(g.itemsize, w.itemsize) == 8 else np.float32
        return g.astype(t), w.astype(t)
    t = np.result_type(g.dtype, w.dtype)
    if t.kind == "f":
        t = np.int64
    return g.astype(t), w.astype(t)


def assertSameResult(got, want, tolCols=(), rtol=1e-9, atol=0.0, label=""):
    assert list(got) == list(want), "%s: columns %s != %s" % (label, list(got), list(want))
    n = len(next(iter(want.values()))) if want else 0
    for k in got:
        assert len(got[k]) == n, "%s: %s has %d rows, expected %d" % (label, k, len(got[k]), n)
    exact = [k for k in want if k not in tolCols]
    og, ow = _sortKey(got, exact), _sortKey(want, exact)
    for k in want:
        g, w = np.asarray(got[k])[og], np.asarray(want[k])[ow]
        if k in tolCols:
            g, w = g.astype(np.float64), w.astype(np.float64)
            bad = ~(np.isclose(g, w, rtol=rtol, atol=atol) | (np.isnan(g) & np.isnan(w)))
            assert not bad.any(), "%s: column %s differs beyond rtol=%g at %d rows, e.g. %r vs %r" % (
                label, k, rtol, int(bad.sum()), g[bad][:3], w[bad][:3])
        elif w.dtype == object:
            assert list(g) == list(w), "%s: column %s differs" % (label, k)
        else:
            # exact columns compare at the wider of the two types, so a GPU
            # column that wrapped or rounded into a narrower type fails
            assert g.dtype.itemsize >= w.dtype.itemsize, "%s: column %s is %s on the GPU, %s in the oracle" % (
                label, k, g.dtype, w.dtype)
            g, w = _widen(g, w)
            if g.dtype.kind == "f":
                bits = np.uint64 if g.itemsize == 8 else np.uint32
                same = (g.view(bits) == w.view(bits)) | (np.isnan(g) & np.isnan(w))
                assert same.all(), "%s: column %s not bit-identical at %d rows, e.g. %r vs %r" % (
                    label, k, int((~same).sum()), g[~same][:3], w[~same][:3])
            else:
                assert np.array_equal(g, w), "%s: column %s differs at %d rows" % (label, k, int((g != w).sum()))
