"""Scalar expression trees of the plan API (predicates, arithmetic, strings).

Class names, constructor signatures and typing rules follow the reference
(dogqc/scalarAlgebra.py).  Unlike the reference, an expression here is pure
data: it only resolves its attributes and types.  Code emission lives in the
HIP translator (``hipExpr.py``) and evaluation in the CPU oracle, so one tree
feeds both without either depending on the other.

Typing rules kept from the reference:
  * binary expressions: equal operand types keep the type, otherwise DOUBLE
    (scalarAlgebra.py:81-88) -- this includes comparisons and AND/OR;
  * unary expressions (NOT, ABS, EXTRACT) take the child's type
    (scalarAlgebra.py:54-57);
  * CASE is FLOAT (259), SUBSTRING is STRING (216);
  * EqualsExpr raises on mismatched non-FLOAT operand types (173-178),
    LikeExpr requires STRING operands (196-202).
"""
from enum import Enum

from dogqc_amd.types import Type


class ScalarExpr(object):
    # expression ids are only used for plan visualisation (scalarAlgebra.py:11-16)
    idSource = 1000

    def __init__(self):
        self.type = None
        self.exprId = ScalarExpr.idSource
        ScalarExpr.idSource += 1

    def children(self):
        return []

    def resolve(self, algExpr):
        attrs = {}
        for c in self.children():
            attrs.update(c.resolve(algExpr))
        self.resolveType()
        return attrs

    def resolveType(self):
        pass

    def DOTstr(self):
        return "%s (%d)" % (self.__class__.__name__, self.exprId)


class AttrExpr(ScalarExpr):

    def __init__(self, identifier):
        ScalarExpr.__init__(self)
        self.identifier = identifier

    def resolve(self, algExpr):
        self.attr = algExpr.resolveAttribute(self.identifier)
        self.type = self.attr.dataType
        return {self.attr.id: self.attr}


class ConstExpr(ScalarExpr):

    def __init__(self, token, type):
        ScalarExpr.__init__(self)
        self.token = token
        self.type = type

    def resolve(self, algExpr):
        return {}


class BoolExpr(ScalarExpr):

    def __init__(self):
        ScalarExpr.__init__(self)
        self.type = Type.BOOLEAN


class _Binary(ScalarExpr):

    def __init__(self, leftChild, rightChild):
        ScalarExpr.__init__(self)
        self.leftChild = leftChild
        self.rightChild = rightChild

    def children(self):
        return [self.leftChild, self.rightChild]

    def resolveType(self):
        lt, rt = self.leftChild.type, self.rightChild.type
        self.type = lt if lt == rt else Type.DOUBLE
        self.checkOperands()

    def checkOperands(self):
        pass


class _Unary(ScalarExpr):

    def __init__(self, child):
        ScalarExpr.__init__(self)
        self.child = child

    def children(self):
        return [self.child]

    def resolveType(self):
        self.type = self.child.type


def _mismatch(expr):
    return SyntaxError("Comparing unmatched types: %s and %s in scalar expr %d"
                       % (expr.leftChild.type, expr.rightChild.type, expr.exprId))


class EqualsExpr(_Binary):

    def checkOperands(self):
        lt, rt = self.leftChild.type, self.rightChild.type
        if Type.FLOAT not in (lt, rt) and lt != rt:
            raise _mismatch(self)
        self.comparisonType = lt


class LikeExpr(_Binary):

    def checkOperands(self):
        if self.leftChild.type != self.rightChild.type:
            raise _mismatch(self)
        if self.leftChild.type != Type.STRING:
            raise SyntaxError("Prefix check only for strings.")
        self.comparisonType = self.leftChild.type


class SmallerExpr(_Binary):
    pass


class SmallerEqualExpr(_Binary):
    pass


class LargerExpr(_Binary):
    pass


class LargerEqualExpr(_Binary):
    pass


class AndExpr(_Binary):
    pass


class OrExpr(_Binary):
    pass


class NotExpr(_Unary):
    pass


def InExpr(v, exprList):
    """``v IN (e0, e1, ...)`` as an OR tree of equalities (scalarAlgebra.py:330-334)."""
    tree = EqualsExpr(v, exprList[0])
    for e in exprList[1:]:
        tree = OrExpr(EqualsExpr(v, e), tree)
    return tree


class ArithExpr(ScalarExpr):

    def __init__(self):
        ScalarExpr.__init__(self)
        self.type = Type.INT


class BinaryNumericExpr(_Binary):
    pass


class MulExpr(BinaryNumericExpr):
    pass


class DivExpr(BinaryNumericExpr):
    pass


class AddExpr(BinaryNumericExpr):
    pass


class SubExpr(BinaryNumericExpr):
    pass


class AbsExpr(_Unary):
    pass


class ExtractType(Enum):
    DAY = 1
    MONTH = 2
    YEAR = 3


class ExtractExpr(_Unary):
    """Date part of a yyyymmdd DATE value; keeps the child's type (DATE)."""

    def __init__(self, child, extractType):
        _Unary.__init__(self, child)
        self.extract = extractType


class SubstringExpr(ScalarExpr):
    """SQL SUBSTRING(str FROM from FOR len), 1-based (util.cuh:132-142)."""

    def __init__(self, exprString, exprFrom, exprFor):
        ScalarExpr.__init__(self)
        self.exprString = exprString
        self.exprFrom = exprFrom
        self.exprFor = exprFor
        self.type = Type.STRING

    def children(self):
        return [self.exprString, self.exprFrom, self.exprFor]

    def resolveType(self):
        if self.exprString.type != Type.STRING:
            raise SyntaxError("Need STRING type, type is: " + str(self.exprString.type))
        if self.exprFrom.type != Type.INT:
            raise SyntaxError("Need INT type, type is: " + str(self.exprFrom.type))
        if self.exprFor.type != Type.INT:
            raise SyntaxError("Need INT type, type is: " + str(self.exprFor.type))
        self.type = Type.STRING


class CaseExpr(ScalarExpr):

    def __init__(self, exprListWhenThen, exprElse=None):
        ScalarExpr.__init__(self)
        if len(exprListWhenThen) == 0:
            raise SyntaxError("The case expression needs at least one when clause.")
        self.exprListWhenThen = exprListWhenThen
        self.exprElse = exprElse
        self.type = Type.FLOAT

    def children(self):
        out = []
        for w, t in self.exprListWhenThen:
            out += [w, t]
        if self.exprElse is not None:
            out.append(self.exprElse)
        return out

    def resolveType(self):
        self.type = Type.FLOAT


# ---------------------------------------------------------------- batching
# Extensions without a reference counterpart, used by batched plans (one run
# of a plan for several queries that read prefixes of the same tables).

class SegmentExpr(ScalarExpr):
    """The batch segment of the current row of a scanned table: the index of
    the first batch member whose prefix of `table` holds the row (members are
    ordered by growing prefixes, given by the plan's `__batch` input).  Must
    be mapped directly above the scan of `table`."""

    def __init__(self, table):
        ScalarExpr.__init__(self)
        self.table = table
        self.type = Type.INT

    def resolve(self, algExpr):
        return {}


class MaxExpr(_Binary):
    """The larger of two INT values (the segment of a joined row)."""

    def resolveType(self):
        if self.leftChild.type != Type.INT or self.rightChild.type != Type.INT:
            raise SyntaxError("MaxExpr takes INT operands")
        self.type = Type.INT
