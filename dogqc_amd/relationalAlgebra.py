"""Relational algebra plan API: operator trees, attribute resolution, pruning
and cardinality configuration.

This is the drop-in boundary for plans: names, constructor arguments, opId
numbering, attribute identifiers and cfg-dict semantics match the reference
(dogqc/relationalAlgebra.py), so plan files such as ``query/plans/tpch10.py``
and their ``*cfg.py`` dicts (keyed by opId) evaluate unchanged.

  * opIds are handed out in construction order (relationalAlgebra.py:30-32,
    152); Python evaluates arguments first, so children precede parents.
  * identifiers are ``name``, ``table.name`` and ``alias.name``
    (relationalAlgebra.py:301-314); ambiguity raises SyntaxError, absence
    NameError (176-186).
  * an operator keeps in its outRelation only attributes touched by a later
    (higher opId) operator (199-204).
  * configure() estimates cardinalities from cfg: selection ``selectivity``
    (342-349), join ``multimatch``/``selectivity``/``htSizeFactor``
    (441-458), cross join (488-500), aggregation ``numgroups`` (564-574).
"""
from collections import OrderedDict
from enum import Enum

from dogqc_amd.types import Type
from dogqc_amd.util import listWrap


class Attribute(object):

    def __init__(self, name, id, opId):
        self.name = name
        self.id = id
        self.creation = opId
        self.numElements = None
        self.lastUse = 0
        self.dataType = None
        self.identifiers = [name]


class Context(object):

    def __init__(self, database):
        self.database = database
        self.schema = dict(database.schema)
        self.opIdNum = 0
        self.attributeIdNum = 0
        self.tableVersions = dict()

    def opId(self):
        self.opIdNum += 1
        return self.opIdNum

    def createTable(self, identifier):
        table = {"name": identifier, "attributes": OrderedDict()}
        self.schema[identifier] = table
        return table

    def resolveTable(self, identifier):
        version = self.tableVersions.get(identifier, 0) + 1
        self.tableVersions[identifier] = version
        return version, self.schema[identifier]

    def createAttribute(self, attributeName, tableName=None):
        self.attributeIdNum += 1
        attr = Attribute(attributeName, self.attributeIdNum, self.opIdNum)
        return attr.id, attr


class RelationalAlgebra(object):

    def __init__(self, acc):
        self.ctxt = Context(acc)
        self.selectionNodes = []
        self.joinNodes = []
        self.semiJoinNodes = []
        self.antiJoinNodes = []

    def scan(self, table, alias=None):
        return Scan(self.ctxt, table, alias)

    def selection(self, condition, child):
        node = Selection(self.ctxt, condition, child)
        self.selectionNodes.append(node)
        return node

    def projection(self, attributes, child):
        return Projection(self.ctxt, attributes, child)

    def map(self, attrName, expression, child):
        return Map(self.ctxt, attrName, expression, child)

    def createtemp(self, identifier, child):
        return Materialize.temp(self.ctxt, identifier, child)

    def result(self, child):
        return Materialize.result(self.ctxt, child)

    def join(self, condition, leftChild, rightChild):
        return self.innerjoin(condition, None, leftChild, rightChild)

    def crossjoin(self, condition, leftChild, rightChild):
        return CrossJoin(self.ctxt, Join.CROSS, condition, leftChild, rightChild)

    def innerjoin(self, equalityConditions, otherConditions, leftChild, rightChild):
        node = EquiJoin(self.ctxt, Join.INNER, equalityConditions, otherConditions, leftChild, rightChild)
        self.joinNodes.append(node)
        return node

    def semijoin(self, equalityConditions, otherConditions, leftChild, rightChild):
        node = EquiJoin(self.ctxt, Join.SEMI, equalityConditions, otherConditions, leftChild, rightChild)
        self.semiJoinNodes.append(node)
        return node

    def antijoin(self, equalityConditions, otherConditions, leftChild, rightChild):
        node = EquiJoin(self.ctxt, Join.ANTI, equalityConditions, otherConditions, leftChild, rightChild)
        self.antiJoinNodes.append(node)
        return node

    def outerjoin(self, equalityConditions, otherConditions, leftChild, rightChild):
        return EquiJoin(self.ctxt, Join.OUTER, equalityConditions, otherConditions, leftChild, rightChild)

    def aggregation(self, groupAttributes, aggregates, child):
        return Aggregation(self.ctxt, groupAttributes, aggregates, child)

    def showGraph(self, plan):
        """Write the plan as Graphviz DOT text to ``query.dot`` (no rendering)."""
        lines = ["digraph {", "  rankdir=BT;"]
        for node in listWrap(plan):
            node.toDOT(lines)
        lines.append("}")
        with open("query.dot", "w") as f:
            f.write("\n".join(lines) + "\n")
        return "\n".join(lines)

    def resolveAlgebraPlan(self, plan, cfg):
        plan = listWrap(plan)
        plan[-1] = self.result(plan[-1])
        for node in plan:
            node.resolve(self.ctxt)
            node.prune()
            node.configure(cfg, self.ctxt)
        return plan

    def translateToCompilerPlan(self, plan, translator):
        return [node.translate(translator, self.ctxt) for node in plan]


def _cfgValue(cfg, opId, key, default):
    try:
        return cfg[opId][key]
    except (KeyError, TypeError, IndexError):
        return default


class AlgExpr(object):

    def __init__(self, ctxt):
        self.opId = ctxt.opId()
        self.inRelation = dict()
        self.outRelation = dict()

    def resolve(self, ctxt):
        pass

    def prune(self):
        pass

    def configure(self, cfg, ctxt):
        pass

    def translate(self, translator, ctxt):
        pass

    def touchAttribute(self, id):
        self.inRelation[id].lastUse = self.opId

    def buildIdentifiers(self):
        identifiers, ambiguous = dict(), dict()
        for id, att in self.inRelation.items():
            for ident in att.identifiers:
                if ident in identifiers:
                    ambiguous[ident] = True
                else:
                    identifiers[ident] = id
        return identifiers, ambiguous

    def resolveAttribute(self, identifier):
        idents, ambiguous = self.buildIdentifiers()
        if identifier in ambiguous:
            raise SyntaxError("Identifier %s is ambiguous in operator %s (%d)"
                              % (identifier, self.DOTcaption(), self.opId))
        if identifier not in idents:
            raise NameError("Attribute %s not found in %s operator (%d)"
                            % (identifier, self.DOTcaption(), self.opId))
        id = idents[identifier]
        self.touchAttribute(id)
        return self.inRelation[id]

    def pruneOutRelation(self):
        self.outRelation = {id: a for id, a in self.outRelation.items() if a.lastUse > self.opId}

    def children(self):
        return []

    def DOTcaption(self):
        return self.__class__.__name__

    def toDOT(self, lines):
        for c in self.children():
            c.toDOT(lines)
            lines.append('  "%d" -> "%d";' % (c.opId, self.opId))
        lines.append('  "%d" [label="%s (%d)"];' % (self.opId, self.DOTcaption(), self.opId))


class LiteralAlgExpr(AlgExpr):

    def prune(self):
        self.pruneOutRelation()

    def translate(self, translator, ctxt):
        return translator.translate(self)


class UnaryAlgExpr(AlgExpr):

    def __init__(self, ctxt, child):
        AlgExpr.__init__(self, ctxt)
        self.child = child

    def children(self):
        return [self.child]

    def prune(self):
        self.child.prune()
        self.pruneOutRelation()

    def translate(self, translator, ctxt):
        return translator.translate(self, self.child.translate(translator, ctxt))


class BinaryAlgExpr(AlgExpr):

    def __init__(self, ctxt, leftChild, rightChild):
        AlgExpr.__init__(self, ctxt)
        self.leftChild = leftChild
        self.rightChild = rightChild

    def children(self):
        return [self.leftChild, self.rightChild]

    def prune(self):
        self.leftChild.prune()
        self.rightChild.prune()
        self.pruneOutRelation()

    def translate(self, translator, ctxt):
        left = self.leftChild.translate(translator, ctxt)
        right = self.rightChild.translate(translator, ctxt)
        return translator.translate(self, left, right)


class Scan(LiteralAlgExpr):

    def __init__(self, ctxt, tableName, alias=None):
        LiteralAlgExpr.__init__(self, ctxt)
        self.scanAttributes = dict()
        self.scanTableId, self.table = ctxt.resolveTable(tableName)
        self.isTempScan = "isTempTable" in self.table
        self.tableAlias = alias

    def resolve(self, ctxt):
        for attrName, typ in self.table["attributes"].items():
            id, attr = ctxt.createAttribute(attrName, self.table["name"])
            attr.dataType = typ
            attr.numElements = self.table["size"]
            attr.identifiers.append(self.table["name"] + "." + attrName)
            if self.tableAlias is not None:
                attr.identifiers.append(self.tableAlias + "." + attrName)
            if self.isTempScan:
                attr.sourceId = self.table["sourceAttributes"][attrName]
            self.scanAttributes[id] = attr
        self.outRelation = self.scanAttributes
        return self.outRelation

    def configure(self, cfg, ctxt):
        self.tupleNum = self.table["size"]
        return self.tupleNum

    def DOTcaption(self):
        return "Tempscan" if self.isTempScan else "Scan"


class Selection(UnaryAlgExpr):

    def __init__(self, ctxt, condition, child):
        UnaryAlgExpr.__init__(self, ctxt, child)
        self.condition = condition

    def resolve(self, ctxt):
        self.inRelation = self.child.resolve(ctxt)
        self.conditionAttributes = self.condition.resolve(self)
        self.outRelation = self.inRelation
        return self.outRelation

    def configure(self, cfg, ctxt):
        childTupleNum = self.child.configure(cfg, ctxt)
        self.selectivity = _cfgValue(cfg, self.opId, "selectivity", 1.0)
        self.tupleNum = int(childTupleNum * self.selectivity)
        return self.tupleNum


class Map(UnaryAlgExpr):

    def __init__(self, ctxt, attrName, expression, child):
        UnaryAlgExpr.__init__(self, ctxt, child)
        self.expression = expression
        self.attrName = attrName

    def resolve(self, ctxt):
        self.inRelation = self.child.resolve(ctxt)
        mapped = self.expression.resolve(self)
        id, attr = ctxt.createAttribute(self.attrName)
        attr.dataType = self.expression.type
        self.mapAttr = attr
        self.mapStringAttributes = [a for a in mapped.values() if a.dataType == Type.STRING]
        out = dict(self.inRelation)
        out[attr.id] = attr
        self.outRelation = out
        return out

    def configure(self, cfg, ctxt):
        self.tupleNum = self.child.configure(cfg, ctxt)
        return self.tupleNum


class Join(Enum):
    CROSS = 1
    INNER = 2
    SEMI = 3
    ANTI = 4
    OUTER = 5


class EquiJoin(BinaryAlgExpr):
    """Hash equi-join.  The left child is the build side, the right child the
    probe side (hashJoins.py:24-33); SEMI/ANTI output only the probe side
    (relationalAlgebra.py:435-438)."""

    def __init__(self, ctxt, joinType, equalityConditions, otherConditions, leftChild, rightChild):
        BinaryAlgExpr.__init__(self, ctxt, leftChild, rightChild)
        self.equalities = listWrap(equalityConditions)
        self.conditions = otherConditions
        self.joinType = joinType

    def resolve(self, ctxt):
        inLeft = self.leftChild.resolve(ctxt)
        self.inRelation.update(inLeft)
        inRight = self.rightChild.resolve(ctxt)
        self.inRelation.update(inRight)

        condAttrs = self.conditions.resolve(self) if self.conditions is not None else dict()
        self.conditionProbeAttributes = {id: a for id, a in condAttrs.items() if id in inRight}
        self.conditionBuildAttributes = {id: a for id, a in condAttrs.items() if id in inLeft}

        self.buildKeyAttributes = OrderedDict()
        self.probeKeyAttributes = OrderedDict()
        for ident1, ident2 in self.equalities:
            a1 = self.resolveAttribute(ident1)
            a2 = self.resolveAttribute(ident2)
            if a1.id in inLeft and a2.id in inRight:
                self.buildKeyAttributes[a1.id] = a1
                self.probeKeyAttributes[a2.id] = a2
            elif a1.id in inRight and a2.id in inLeft:
                self.buildKeyAttributes[a2.id] = a2
                self.probeKeyAttributes[a1.id] = a1
            else:
                raise NameError("Attributes " + a1.name + " and " + a2.name + " come from the same relation.")
        if self.joinType in (Join.INNER, Join.OUTER):
            self.outRelation = self.inRelation
        else:
            self.outRelation = inRight
        return self.outRelation

    def configure(self, cfg, ctxt):
        left = self.leftChild.configure(cfg, ctxt)
        right = self.rightChild.configure(cfg, ctxt)
        self.multimatch = _cfgValue(cfg, self.opId, "multimatch", True)
        self.selectivity = _cfgValue(cfg, self.opId, "selectivity", 1.0)
        self.htSizeFactor = _cfgValue(cfg, self.opId, "htSizeFactor", 2.0)
        self.tupleNum = int(max(left, right) * self.selectivity)
        return self.tupleNum

    def DOTcaption(self):
        s = str(self.joinType) + ("*" if self.conditions is not None else "")
        return s[0] + s[1:].lower()


class CrossJoin(BinaryAlgExpr):

    def __init__(self, ctxt, joinType, condition, leftChild, rightChild):
        BinaryAlgExpr.__init__(self, ctxt, leftChild, rightChild)
        self.condition = condition
        self.joinType = joinType

    def resolve(self, ctxt):
        inLeft = self.leftChild.resolve(ctxt)
        self.inRelation.update(inLeft)
        inRight = self.rightChild.resolve(ctxt)
        self.inRelation.update(inRight)
        self.conditionAttributes = self.condition.resolve(self)
        self.outRelation = self.inRelation
        return self.outRelation

    def configure(self, cfg, ctxt):
        left = self.leftChild.configure(cfg, ctxt)
        right = self.rightChild.configure(cfg, ctxt)
        self.multimatch = _cfgValue(cfg, self.opId, "multimatch", True)
        self.selectivity = _cfgValue(cfg, self.opId, "selectivity", 1.0)
        self.tupleNum = int(left * right * self.selectivity)
        return self.tupleNum


class Reduction(Enum):
    SUM = 1
    COUNT = 2
    AVG = 3
    MIN = 4
    MAX = 5


class Aggregation(UnaryAlgExpr):
    """Group-by aggregation.  Aggregate attributes are created at construction
    (so they take attribute ids before any scan attribute), typed DOUBLE except
    COUNT, which is INT; an AVG without a COUNT gets a ``count_agg`` COUNT
    appended (relationalAlgebra.py:528-543).  Extension: SUM/MIN/MAX over an
    INT64 input are INT64 (exact 64-bit integer arithmetic)."""

    def __init__(self, ctxt, groupingIdentifiers, aggregates, child):
        UnaryAlgExpr.__init__(self, ctxt, child)
        self.groupingIdentifiers = listWrap(groupingIdentifiers)
        self.aggregates = listWrap(aggregates)
        self.aggregateTuplesCreated = OrderedDict()
        self.aggregateAttributes = dict()
        self.avgAggregates = dict()
        if any(a[0] == Reduction.AVG for a in self.aggregates):
            if not any(a[0] == Reduction.COUNT for a in self.aggregates):
                self.aggregates.append((Reduction.COUNT, "", "count_agg"))
        for reductionType, inputIdentifier, aliasName in self.aggregates:
            id, aggAttr = ctxt.createAttribute(aliasName)
            aggAttr.dataType = Type.DOUBLE
            if reductionType == Reduction.COUNT:
                aggAttr.dataType = Type.INT
                self.countAttr = aggAttr
            self.aggregateAttributes[id] = aggAttr
            if reductionType == Reduction.AVG:
                self.avgAggregates[id] = aggAttr
            self.aggregateTuplesCreated[id] = (aggAttr, inputIdentifier, reductionType)

    def resolve(self, ctxt):
        self.inRelation = self.child.resolve(ctxt)
        self.groupAttributes = OrderedDict()
        for ident in self.groupingIdentifiers:
            att = self.resolveAttribute(ident)
            self.groupAttributes[att.id] = att
        self.aggregateInAttributes = OrderedDict()
        self.aggregateTuples = dict()
        for id, (aggAttr, inputIdentifier, reductionType) in self.aggregateTuplesCreated.items():
            inId = None
            if inputIdentifier != "":
                inAtt = self.resolveAttribute(inputIdentifier)
                self.aggregateInAttributes[inAtt.id] = inAtt
                inId = inAtt.id
                if inAtt.dataType == Type.INT64 and reductionType in (Reduction.SUM, Reduction.MIN, Reduction.MAX):
                    aggAttr.dataType = Type.INT64
            self.aggregateTuples[id] = (inId, reductionType)
        self.outRelation.update(self.groupAttributes)
        self.outRelation.update(self.aggregateAttributes)
        return self.outRelation

    def configure(self, cfg, ctxt):
        childTupleNum = self.child.configure(cfg, ctxt)
        self.doGroup = len(self.groupAttributes) > 0
        self.numgroups = _cfgValue(cfg, self.opId, "numgroups", None)
        self.numgroupsFromCfg = self.numgroups is not None
        # partitions of a partitioned group-by when the tuning leaves it to
        # the plan (an extension: the reference has no such key)
        self.partitions = _cfgValue(cfg, self.opId, "partitions", None)
        if self.numgroups is None:
            self.numgroups = childTupleNum
        if not self.doGroup:
            self.numgroups = 1
        self.tupleNum = self.numgroups
        return self.tupleNum


class Projection(UnaryAlgExpr):

    def __init__(self, ctxt, identifiers, child):
        UnaryAlgExpr.__init__(self, ctxt, child)
        self.identifiers = listWrap(identifiers)

    def resolve(self, ctxt):
        self.inRelation = self.child.resolve(ctxt)
        self.projectionAttributes = dict()
        for ident in self.identifiers:
            att = self.resolveAttribute(ident)
            self.projectionAttributes[att.id] = att
        self.outRelation = self.projectionAttributes
        return self.outRelation

    def configure(self, cfg, ctxt):
        self.tupleNum = self.child.configure(cfg, ctxt)
        return self.tupleNum


class MaterializationType(Enum):
    RESULT = 1
    TEMPTABLE = 2


class Materialize(UnaryAlgExpr):

    @staticmethod
    def result(ctxt, child):
        mat = Materialize(ctxt, child)
        mat.matType = MaterializationType.RESULT
        return mat

    @staticmethod
    def temp(ctxt, identifier, child):
        mat = Materialize(ctxt, child)
        mat.table = ctxt.createTable(identifier)
        mat.table["isTempTable"] = True
        mat.matType = MaterializationType.TEMPTABLE
        return mat

    def resolve(self, ctxt):
        self.inRelation = self.child.resolve(ctxt)
        for id in self.inRelation:
            self.touchAttribute(id)
        if self.matType == MaterializationType.TEMPTABLE:
            self.table["sourceAttributes"] = dict()
            for id, att in self.inRelation.items():
                self.table["attributes"][att.name] = att.dataType
                self.table["sourceAttributes"][att.name] = att.id
            self.table["numColumns"] = len(self.inRelation)
        self.outRelation = self.inRelation
        return self.outRelation

    def prune(self):
        self.child.prune()

    def configure(self, cfg, ctxt):
        self.tupleNum = self.child.configure(cfg, ctxt)
        if self.matType == MaterializationType.TEMPTABLE:
            self.table["size"] = self.tupleNum
        return self.tupleNum

    def DOTcaption(self):
        return "Result" if self.matType == MaterializationType.RESULT else "Temptable"
