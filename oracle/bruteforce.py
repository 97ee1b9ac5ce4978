"""Row-at-a-time brute-force evaluator for tiny inputs.

TEST INFRASTRUCTURE ONLY: used to pin ``planeval`` (the numpy oracle) on
inputs of a few hundred rows, with deliberately different algorithms:
tuples as Python dicts, nested-loop joins, dictionary group-by, per-row
scalar arithmetic in numpy scalar types (C semantics), sums accumulated with
Python's exact integer / ``fractions.Fraction`` arithmetic and rounded once.
Semantics as documented in ``planeval``.
"""
from collections import OrderedDict
from fractions import Fraction

import numpy as np

import dogqc_amd.relationalAlgebra as ra
from dogqc_amd.relationalAlgebra import Reduction, Join
from dogqc_amd.types import Type
from oracle.planeval import Evaluator, regType, FP64


class BruteForce(object):

    def __init__(self, accessor, decimal=FP64):
        self.acc = accessor
        self.decimal = decimal
        self.scalar = Evaluator(accessor, decimal)
        self.temps = {}      # temp table name -> list of {attribute name: value}

    def value(self, e, row):
        """Evaluate an expression on one tuple (a dict attr id -> value) by
        running the numpy expression code on length-1 arrays."""
        rel = {k: np.array([v], dtype=object if isinstance(v, bytes) else None) for k, v in row.items()}
        v, _ = self.scalar.expr(e, rel, 1)
        v = np.asarray(v)
        return v.reshape(-1)[0] if v.ndim else v[()]

    def rows(self, node):
        if isinstance(node, ra.Scan) and node.isTempScan:
            byName = {a.name: id for id, a in node.outRelation.items()}
            return [{byName[k]: v for k, v in r.items() if k in byName} for r in self.temps[node.table["name"]]]
        if isinstance(node, ra.Scan):
            n = node.table["size"]
            cols = {}
            for id, a in node.outRelation.items():
                d = self.acc.column(node.table["name"], a.name)
                if a.dataType == Type.STRING:
                    offs, chars = d
                    b = np.asarray(chars, dtype=np.uint8).tobytes()
                    cols[id] = [b[int(offs[i]):int(offs[i + 1])] for i in range(len(offs) - 1)]
                else:
                    cols[id] = list(np.asarray(d).astype(regType(a.dataType, self.decimal)))
                n = len(cols[id])
            rows = [{id: cols[id][i] for id in cols} for i in range(n)]
            for i, r in enumerate(rows):
                r[("__row", node.table["name"])] = i      # for SegmentExpr
            return rows
        if isinstance(node, ra.Selection):
            return [r for r in self.rows(node.child) if self.value(node.condition, r)]
        if isinstance(node, ra.Map):
            out = []
            ct = regType(node.mapAttr.dataType, self.decimal)
            for r in self.rows(node.child):
                v = self.value(node.expression, r)
                r = dict(r)
                r[node.mapAttr.id] = v if ct == "str" else np.array(v).astype(ct)[()]
                out.append(r)
            return out
        if isinstance(node, ra.Materialize) and node.matType == ra.MaterializationType.TEMPTABLE:
            rows = self.rows(node.child)
            self.temps[node.table["name"]] = [{a.name: r[id] for id, a in node.outRelation.items()} for r in rows]
            return rows
        if isinstance(node, (ra.Projection, ra.Materialize)):
            return self.rows(node.child)
        if isinstance(node, ra.EquiJoin):
            L = self.rows(node.leftChild)
            R = self.rows(node.rightChild)
            bk = list(node.buildKeyAttributes)
            pk = list(node.probeKeyAttributes)
            out = []
            for p in R:
                matched = False
                for b in L:
                    if all(b[x] == p[y] for x, y in zip(bk, pk)):
                        pair = dict(b)
                        pair.update(p)
                        if node.conditions is not None and not self.value(node.conditions, pair):
                            continue
                        matched = True
                        if node.joinType in (Join.INNER, Join.OUTER):
                            out.append(pair)
                if node.joinType == Join.SEMI and matched:
                    out.append(p)
                if node.joinType == Join.ANTI and not matched:
                    out.append(p)
                if node.joinType == Join.OUTER and not matched:
                    # build attributes null: value 0 / empty string plus a null marker
                    row = dict(p)
                    for id, att in node.leftChild.outRelation.items():
                        ct = regType(att.dataType, self.decimal)
                        row[id] = b"" if ct == "str" else np.zeros(1, dtype=ct)[0]
                        row[("isnull", id)] = True
                    out.append(row)
            return out
        if isinstance(node, ra.CrossJoin):
            out = []
            for p in self.rows(node.rightChild):
                for b in self.rows(node.leftChild):
                    pair = dict(b)
                    pair.update(p)
                    if self.value(node.condition, pair):
                        out.append(pair)
            return out
        if isinstance(node, ra.Aggregation):
            rows = self.rows(node.child)
            groups = OrderedDict()
            gids = list(node.groupAttributes)
            for r in rows:
                groups.setdefault(tuple(r[g] for g in gids), []).append(r)
            if not gids and not groups:
                groups[()] = []
            out = []
            for key, members in groups.items():
                o = dict(zip(gids, key))
                for id, (inId, red) in node.aggregateTuples.items():
                    ct = regType(node.aggregateAttributes[id].dataType, self.decimal)
                    # null values (outer join) are skipped
                    vals = [m[inId] for m in members if not m.get(("isnull", inId))] if inId is not None else []
                    if red == Reduction.COUNT:
                        o[id] = np.int32(len(members) if inId is None else len(vals))
                    elif red in (Reduction.SUM, Reduction.AVG):
                        if np.dtype(ct).kind == "f":
                            s = sum((Fraction(float(v)) for v in vals), Fraction(0))
                            if red == Reduction.AVG:
                                # C: the rounded sum divided by the group's row count
                                o[id] = float(s) / float(len(members)) if members else float("nan")
                            else:
                                o[id] = float(s)
                        else:
                            s = sum(int(v) for v in vals)
                            o[id] = np.int64((s + 2 ** 63) % 2 ** 64 - 2 ** 63)
                    elif red == Reduction.MIN:
                        o[id] = np.array(min(vals)).astype(ct)[()] if vals else (
                            np.finfo(ct).max if np.dtype(ct).kind == "f" else np.iinfo(ct).max)
                    elif red == Reduction.MAX:
                        o[id] = np.array(max(vals)).astype(ct)[()] if vals else (
                            -np.finfo(ct).max if np.dtype(ct).kind == "f" else np.iinfo(ct).min)
                out.append(o)
            return out
        raise NotImplementedError(type(node).__name__)

    def run(self, plan):
        rows = None
        for node in plan:
            rows = self.rows(node)
        node = plan[-1]
        res = OrderedDict()
        for id, a in node.outRelation.items():
            vals = [r[id] for r in rows]
            if a.dataType == Type.STRING:
                res[a.name] = np.array(vals, dtype=object)
            else:
                # an empty result keeps its column's register type
                res[a.name] = np.array(vals) if vals else np.zeros(0, dtype=regType(a.dataType, self.decimal))
        return res


def evaluate(plan, accessor, decimal=FP64):
    return BruteForce(accessor, decimal).run(plan)
