"""Produce/consume translator tree (contract of dogqc/translatorBase.py:6-65).

A translator wraps one algebra operator.  ``produce`` opens pipelines from the
leaves upwards; ``consume`` is called by the child with the code position of
the current pipeline and emits the operator's code into it, then calls the
parent's ``consume``.  Binary operators receive two consume calls, the first
from the left (build) pipeline, the second from the right (probe) pipeline.
"""


class Translator(object):

    def __init__(self, algExpr):
        self.algExpr = algExpr
        self.parent = None

    def produce(self, ctxt):
        pass

    def consume(self, ctxt):
        pass

    def opId(self):
        return self.algExpr.opId


class LiteralTranslator(Translator):
    pass


class UnaryTranslator(Translator):

    def __init__(self, algExpr, child):
        Translator.__init__(self, algExpr)
        self.child = child
        child.parent = self


class BinaryTranslator(Translator):

    def __init__(self, algExpr, leftChild, rightChild):
        Translator.__init__(self, algExpr)
        self.leftChild = leftChild
        self.rightChild = rightChild
        leftChild.parent = self
        rightChild.parent = self
