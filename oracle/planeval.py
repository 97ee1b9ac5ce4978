"""CPU oracle: evaluates a resolved plan tree of the plan API with numpy.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP backend:
only tests, the smoke check and the CPU-baseline leg of the benchmark may use
it, never the product path.

It restates the reference's operator semantics operator-at-a-time:
  * selection keeps rows whose predicate is true (cudaTranslator.py:358-366);
  * map assigns an expression to a register of the attribute's type
    (cudaTranslator.py:410-422) with C arithmetic: usual arithmetic
    conversions, truncating integer division, float/double promotion, literal
    spellings of scalarAlgebra.py:142-154 and the register type map of
    codegen.py:37-45 (DATE unsigned, FLOAT/DOUBLE -> decimal type);
  * equi-joins (hashJoins.py): INNER yields every matching (build, probe)
    pair, SEMI/ANTI keep probe rows with/without a match (extra join
    conditions evaluated per pair), OUTER adds every probe row without a
    match once with null build attributes (hashJoins.py:473-530; a null
    value reads as 0 / empty string, COUNT(attr), SUM, MIN and MAX skip it,
    cudaTranslator.py:475-476), with sort-based matching instead of hash
    tables;
  * aggregation (cudaTranslator.py:425-525): groups by the key tuple; SUM
    and AVG of doubles use math.fsum (exactly rounded, so the oracle's own
    error is below one ulp), integer sums are exact; COUNT is int32; AVG =
    SUM / COUNT; without grouping exactly one row is produced; MIN starts at
    the type's maximum and MAX at its lowest value;
  * a temp table (createtemp, relationalAlgebra.py Materialize.temp) holds
    its child's rows; a later root's scan of it reads them by attribute name;
  * LIKE, SUBSTRING and string equality are strfuncs.py's restatements of
    util.cuh:90-102, 132-199 (the reference's LIKE, not textbook LIKE).
The result is an OrderedDict attribute name -> numpy array (strings as
object arrays of bytes), in the plan's output attribute order.
"""
import math
import re
from collections import OrderedDict

import numpy as np

import dogqc_amd.relationalAlgebra as ra
import dogqc_amd.scalarAlgebra as scal
from dogqc_amd.relationalAlgebra import Reduction, Join
from dogqc_amd.types import Type
from oracle import strfuncs

FP64 = "double"

# C types used by the oracle, as numpy dtypes ("str" = object array of bytes)
_REG = {Type.INT: "int32", Type.DATE: "uint32", Type.CHAR: "int8", Type.INT64: "int64", Type.BOOLEAN: "int32",
        Type.STRING: "str"}


def regType(t, decimal):
    if t in (Type.FLOAT, Type.DOUBLE):
        return "float64" if decimal == FP64 else "float32"
    return _REG[t]


_RANK = {"bool": 0, "int8": 1, "int32": 2, "uint32": 3, "int64": 4, "float32": 5, "float64": 6}


def _arith(a, b):
    """Usual arithmetic conversions for a binary operator on C types."""
    if "float64" in (a, b):
        return "float64"
    if "float32" in (a, b):
        return "float32"
    prom = [("int32" if t in ("bool", "int8") else t) for t in (a, b)]
    if "int64" in prom:
        return "int64"
    if "uint32" in prom:
        return "uint32"
    return "int32"


def _literal(e, decimal):
    """(value array of length 1, ctype) of a ConstExpr with its C spelling."""
    tok, t = e.token, e.type
    if t == Type.STRING:
        return e.token.encode("latin-1"), "str"
    if t == Type.CHAR:
        return np.int8(ord(tok)), "int8"
    if t == Type.FLOAT:
        v = np.float32(float(tok.rstrip("fF")))
        return v, "float32"
    if t in (Type.INT, Type.DATE, Type.INT64, Type.DOUBLE):
        s = tok.strip()
        if re.fullmatch(r"[+-]?\d+", s):
            v = int(s)
            if -2 ** 31 <= v < 2 ** 31:
                return np.int32(v), "int32"
            return np.int64(v), "int64"
        if s[-1:] in "fF":
            return np.float32(float(s[:-1])), "float32"
        return np.float64(float(s)), "float64"
    raise NotImplementedError("literal of type %s" % t)


def _cast(v, ct):
    if ct == "str":
        return v
    if isinstance(v, np.ndarray):
        if v.dtype == object:
            raise TypeError("string used as number")
        with np.errstate(over="ignore", invalid="ignore"):
            if np.dtype(ct).kind in "iu" and v.dtype.kind == "f":
                # C float -> integer conversion truncates toward zero
                return np.trunc(v).astype(np.int64).astype(ct)
            return v.astype(ct)
    return np.array(v).astype(ct)[()]


class Evaluator(object):

    def __init__(self, accessor, decimal=FP64):
        self.acc = accessor
        self.decimal = decimal
        self.temps = {}      # temp table name -> {attribute name: column}

    # ------------------------------------------------------------ scalars
    def expr(self, e, rel, n):
        """(array or scalar, ctype)."""
        if isinstance(e, scal.AttrExpr):
            return rel[e.attr.id], self.ctype(e.attr)
        if isinstance(e, scal.ConstExpr):
            return _literal(e, self.decimal)
        if isinstance(e, (scal.AndExpr, scal.OrExpr)):
            a, _ = self.expr(e.leftChild, rel, n)
            b, _ = self.expr(e.rightChild, rel, n)
            a, b = np.asarray(a) != 0, np.asarray(b) != 0
            return (a & b if isinstance(e, scal.AndExpr) else a | b), "bool"
        if isinstance(e, scal.NotExpr):
            a, _ = self.expr(e.child, rel, n)
            return np.asarray(a) == 0, "bool"
        if isinstance(e, scal.LikeExpr):
            a, _ = self.expr(e.leftChild, rel, n)
            p, _ = self.expr(e.rightChild, rel, n)
            if isinstance(p, bytes):
                return strfuncs.likeColumn(a, p), "bool"
            return np.fromiter((strfuncs.like(x, y) for x, y in zip(a, p)), dtype=bool, count=len(a)), "bool"
        if isinstance(e, (scal.EqualsExpr, scal.SmallerExpr, scal.SmallerEqualExpr, scal.LargerExpr,
                          scal.LargerEqualExpr)):
            a, at = self.expr(e.leftChild, rel, n)
            b, bt = self.expr(e.rightChild, rel, n)
            if at == "str" or bt == "str":
                if not isinstance(e, scal.EqualsExpr):
                    raise NotImplementedError("string ordering")
                if isinstance(a, bytes):
                    a, b = b, a
                if isinstance(b, bytes):
                    return np.fromiter((strfuncs.equals(x, b) for x in a), dtype=bool, count=len(a)), "bool"
                return np.fromiter((strfuncs.equals(x, y) for x, y in zip(a, b)), dtype=bool, count=len(a)), "bool"
            ct = _arith(at, bt)
            a, b = _cast(a, ct), _cast(b, ct)
            op = {scal.EqualsExpr: np.equal, scal.SmallerExpr: np.less, scal.SmallerEqualExpr: np.less_equal,
                  scal.LargerExpr: np.greater, scal.LargerEqualExpr: np.greater_equal}[type(e)]
            return op(a, b), "bool"
        if isinstance(e, (scal.MulExpr, scal.AddExpr, scal.SubExpr, scal.DivExpr)):
            a, at = self.expr(e.leftChild, rel, n)
            b, bt = self.expr(e.rightChild, rel, n)
            ct = _arith(at, bt)
            a, b = _cast(a, ct), _cast(b, ct)
            with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
                if isinstance(e, scal.MulExpr):
                    r = np.multiply(a, b, dtype=ct)
                elif isinstance(e, scal.AddExpr):
                    r = np.add(a, b, dtype=ct)
                elif isinstance(e, scal.SubExpr):
                    r = np.subtract(a, b, dtype=ct)
                elif np.dtype(ct).kind == "f":
                    r = np.divide(a, b, dtype=ct)
                else:
                    # C integer division truncates toward zero
                    q = np.floor_divide(np.abs(a.astype(np.int64)), np.abs(b.astype(np.int64)))
                    r = (q * np.sign(a.astype(np.int64)) * np.sign(b.astype(np.int64))).astype(ct)
            return r, ct
        if isinstance(e, scal.AbsExpr):
            a, at = self.expr(e.child, rel, n)
            return np.abs(a), at
        if isinstance(e, scal.ExtractExpr):
            a, at = self.expr(e.child, rel, n)
            a = np.asarray(a).astype(np.int64)
            if e.extract == scal.ExtractType.DAY:
                r = a % 100
            elif e.extract == scal.ExtractType.MONTH:
                r = (a // 100) % 100
            else:
                r = a // 10000
            ct = _arith(at, "int32")
            return r.astype(ct), ct
        if isinstance(e, scal.SegmentExpr):
            # first batch member whose prefix of the table holds the row
            rows = rel[("__row", e.table)]
            prefix = np.asarray(self.acc.column("__batch", e.table), dtype=np.int64)
            return np.searchsorted(prefix, rows, side="right").astype(np.int32), "int32_t"
        if isinstance(e, scal.MaxExpr):
            a, _ = self.expr(e.leftChild, rel, n)
            b, _ = self.expr(e.rightChild, rel, n)
            return np.maximum(np.asarray(a, dtype=np.int32), np.asarray(b, dtype=np.int32)), "int32_t"
        if isinstance(e, scal.SubstringExpr):
            s, _ = self.expr(e.exprString, rel, n)
            f, _ = self.expr(e.exprFrom, rel, n)
            ln, _ = self.expr(e.exprFor, rel, n)
            f = np.broadcast_to(np.asarray(f, dtype=np.int64), (len(s),))
            ln = np.broadcast_to(np.asarray(ln, dtype=np.int64), (len(s),))
            out = np.empty(len(s), dtype=object)
            for i, x in enumerate(s):
                out[i] = strfuncs.substring(x, int(f[i]), int(ln[i]))
            return out, "str"
        if isinstance(e, scal.CaseExpr):
            ct = regType(Type.FLOAT, self.decimal)
            if e.exprElse is not None:
                v, vt = self.expr(e.exprElse, rel, n)
                out = np.broadcast_to(_cast(v, ct), (n,)).copy()
            else:
                out = np.zeros(n, dtype=ct)
            for w, t in reversed(e.exprListWhenThen):
                cond, _ = self.expr(w, rel, n)
                v, _ = self.expr(t, rel, n)
                v = np.broadcast_to(_cast(v, ct), (n,))
                cond = np.broadcast_to(np.asarray(cond) != 0, (n,))
                out = np.where(cond, v, out)
            return out, ct
        raise NotImplementedError(type(e).__name__)

    def ctype(self, attr):
        return regType(attr.dataType, self.decimal)

    # ---------------------------------------------------------- relations
    def run(self, plan):
        """Evaluate the plan (list of roots, the last being the Result)."""
        out = None
        for node in plan:
            rel, n = self.node(node)
            out = (node, rel, n)
        node, rel, n = out
        res = OrderedDict()
        for id, a in node.outRelation.items():
            res[a.name] = rel[id]
        return res

    def node(self, node):
        m = getattr(self, "_" + type(node).__name__)
        return m(node)

    def _Scan(self, node):
        tname = node.table["name"]
        rel = {}
        n = None
        if node.isTempScan:
            cols, n = self.temps[tname]
            for id, a in node.outRelation.items():
                rel[id] = cols[a.name]
            return rel, n
        for id, a in node.outRelation.items():
            data = self.acc.column(tname, a.name)
            if a.dataType == Type.STRING:
                offs, chars = data
                offs = np.asarray(offs, dtype=np.int64)
                b = np.asarray(chars, dtype=np.uint8).tobytes()
                vals = np.empty(len(offs) - 1, dtype=object)
                for i in range(len(offs) - 1):
                    vals[i] = b[offs[i]:offs[i + 1]]
                rel[id] = vals
            else:
                rel[id] = _cast(np.asarray(data), self.ctype(a))
            n = len(rel[id])
        if n is None:
            n = node.table["size"]
        rel[("__row", tname)] = np.arange(n, dtype=np.int64)   # for SegmentExpr
        return rel, n

    def _Selection(self, node):
        rel, n = self.node(node.child)
        mask, _ = self.expr(node.condition, rel, n)
        mask = np.broadcast_to(np.asarray(mask) != 0, (n,))
        return {k: v[mask] for k, v in rel.items()}, int(mask.sum())

    def _Map(self, node):
        rel, n = self.node(node.child)
        v, vt = self.expr(node.expression, rel, n)
        ct = self.ctype(node.mapAttr)
        if ct == "str":
            v = np.broadcast_to(np.asarray(v, dtype=object), (n,))
        else:
            v = np.broadcast_to(_cast(v, ct), (n,))
        rel = dict(rel)
        rel[node.mapAttr.id] = np.array(v)
        return rel, n

    def _Projection(self, node):
        rel, n = self.node(node.child)
        return rel, n

    def _Materialize(self, node):
        """Result, or a temp table (relationalAlgebra.py createtemp): its
        columns become scannable by name for the plan's later roots."""
        rel, n = self.node(node.child)
        if node.matType == ra.MaterializationType.TEMPTABLE:
            self.temps[node.table["name"]] = ({a.name: rel[id] for id, a in node.outRelation.items()}, n)
        return rel, n

    @staticmethod
    def _codes(cols):
        """Dense integer code per row of a tuple of key columns (shared
        numbering across all rows given)."""
        code = np.zeros(len(cols[0]) if cols else 0, dtype=np.int64)
        for c in cols:
            if c.dtype == object:
                _, inv = np.unique(c.astype(bytes) if len(c) else c, return_inverse=True)
            else:
                _, inv = np.unique(c, return_inverse=True)
            k = int(inv.max()) + 1 if len(inv) else 1
            code = code * k + inv.astype(np.int64)
            _, code = np.unique(code, return_inverse=True)
            code = code.astype(np.int64)
        return code

    def _EquiJoin(self, node):
        L, nl = self.node(node.leftChild)
        R, nr = self.node(node.rightChild)
        bk = list(node.buildKeyAttributes.values())
        pk = list(node.probeKeyAttributes.values())
        cols = []
        for b, p in zip(bk, pk):
            bc, pc = L[b.id], R[p.id]
            if bc.dtype != object:
                ct = _arith(self.ctype(b), self.ctype(p))
                bc, pc = _cast(bc, ct), _cast(pc, ct)
            cols.append(np.concatenate([bc, pc]))
        code = self._codes(cols)
        cb, cp = code[:nl], code[nl:]
        order = np.argsort(cb, kind="stable")
        sb = cb[order]
        lo = np.searchsorted(sb, cp, side="left")
        hi = np.searchsorted(sb, cp, side="right")
        cnt = (hi - lo).astype(np.int64)
        total = int(cnt.sum())
        pi = np.repeat(np.arange(nr, dtype=np.int64), cnt)
        within = np.arange(total, dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        bi = order[np.repeat(lo, cnt) + within] if total else np.zeros(0, dtype=np.int64)
        pairs = {k: v[bi] for k, v in L.items()}
        pairs.update({k: v[pi] for k, v in R.items()})
        if node.conditions is not None and len(pi):
            m, _ = self.expr(node.conditions, pairs, len(pi))
            keep = np.broadcast_to(np.asarray(m) != 0, (len(pi),))
            pi, bi = pi[keep], bi[keep]
            pairs = {k: v[keep] for k, v in pairs.items()}
        if node.joinType == Join.INNER:
            return pairs, len(pi)
        if node.joinType in (Join.SEMI, Join.ANTI):
            hit = np.zeros(nr, dtype=bool)
            hit[pi] = True
            keep = hit if node.joinType == Join.SEMI else ~hit
            return {k: v[keep] for k, v in R.items()}, int(keep.sum())
        if node.joinType == Join.OUTER:
            # probe rows without a match appear once, build attributes null
            hit = np.zeros(nr, dtype=bool)
            hit[pi] = True
            miss = np.nonzero(~hit)[0]
            out = {}
            for k, v in pairs.items():
                if k in R:
                    out[k] = np.concatenate([v, R[k][miss]])
                elif isinstance(k, tuple):
                    out[k] = np.concatenate([v, np.ones(len(miss), dtype=bool)])
                else:
                    out[k] = np.concatenate([v, np.zeros(len(miss), dtype=v.dtype) if v.dtype != object
                                             else np.array([b""] * len(miss), dtype=object)])
            for id in node.leftChild.outRelation:
                if id in out and ("isnull", id) not in out:
                    out[("isnull", id)] = np.concatenate([np.zeros(len(pi), dtype=bool),
                                                          np.ones(len(miss), dtype=bool)])
            return out, len(pi) + len(miss)
        raise NotImplementedError(str(node.joinType))

    def _CrossJoin(self, node):
        """Every (left, right) pair that satisfies the condition
        (cudaTranslator.py:369-397)."""
        L, nl = self.node(node.leftChild)
        R, nr = self.node(node.rightChild)
        li = np.tile(np.arange(nl, dtype=np.int64), nr)
        ri = np.repeat(np.arange(nr, dtype=np.int64), nl)
        pairs = {k: v[li] for k, v in L.items()}
        pairs.update({k: v[ri] for k, v in R.items()})
        n = nl * nr
        if n:
            m, _ = self.expr(node.condition, pairs, n)
            keep = np.broadcast_to(np.asarray(m) != 0, (n,))
            pairs = {k: v[keep] for k, v in pairs.items()}
            n = int(keep.sum())
        return pairs, n

    def _Aggregation(self, node):
        rel, n = self.node(node.child)
        groups = list(node.groupAttributes.values())
        if groups:
            code = self._codes([rel[g.id] for g in groups]) if n else np.zeros(0, dtype=np.int64)
            ng = int(code.max()) + 1 if n else 0
        else:
            code = np.zeros(n, dtype=np.int64)
            ng = 1
        order = np.argsort(code, kind="stable")
        sc = code[order]
        bounds = np.searchsorted(sc, np.arange(ng + 1), side="left")
        out = {}
        first = order[bounds[:-1]] if ng and n else np.zeros(0, dtype=np.int64)
        for g in groups:
            out[g.id] = rel[g.id][first]
        counts = (bounds[1:] - bounds[:-1]).astype(np.int64)
        sums = {}
        for id, (inId, red) in node.aggregateTuples.items():
            attr = node.aggregateAttributes[id]
            ct = self.ctype(attr)
            null = rel.get(("isnull", inId)) if inId is not None else None
            null = null[order] if null is not None else None
            if red == Reduction.COUNT:
                if null is None:
                    out[id] = counts.astype(np.int32)
                else:   # COUNT(attr) skips null values
                    out[id] = np.array([int((~null[bounds[gi]:bounds[gi + 1]]).sum()) for gi in range(ng)],
                                       dtype=np.int32)
                continue
            vals = rel[inId][order]
            inCt = self.ctype(node.aggregateInAttributes[inId])
            res = np.empty(ng, dtype=ct)
            for gi in range(ng):
                seg = vals[bounds[gi]:bounds[gi + 1]]
                n = len(seg)
                if null is not None:    # SUM/MIN/MAX/AVG skip null values (AVG still divides by COUNT(*))
                    seg = seg[~null[bounds[gi]:bounds[gi + 1]]]
                res[gi] = self._reduce(red, seg, inCt, ct, n)
            out[id] = res
        return out, ng

    @staticmethod
    def _reduce(red, seg, inCt, ct, rows=None):
        kind = np.dtype(ct).kind
        rows = len(seg) if rows is None else rows
        if red in (Reduction.SUM, Reduction.AVG):
            if kind == "f":
                if np.dtype(inCt).kind in "iu":
                    s = float(sum(int(x) for x in seg))
                else:
                    s = math.fsum(seg.astype(np.float64))
                if red == Reduction.AVG:
                    return s / float(rows) if rows else float("nan")
                return s
            # integer sum with two's-complement wrap-around
            s = int(np.sum(seg.astype(np.int64).astype(object))) if len(seg) else 0
            s = (s + 2 ** 63) % 2 ** 64 - 2 ** 63
            return s
        if red == Reduction.MIN:
            if len(seg) == 0:
                return np.finfo(ct).max if kind == "f" else np.iinfo(ct).max
            return _cast(seg.min(), ct)
        if red == Reduction.MAX:
            if len(seg) == 0:
                return -np.finfo(ct).max if kind == "f" else np.iinfo(ct).min
            return _cast(seg.max(), ct)
        raise NotImplementedError(str(red))


def evaluate(plan, accessor, decimal=FP64):
    return Evaluator(accessor, decimal).run(plan)
