"""Scalar expression -> HIP C++ expression text.

Literal spelling follows the reference so that arithmetic has identical C
semantics (scalarAlgebra.py:142-154): INT/DATE/DOUBLE tokens verbatim, FLOAT
tokens as ``(float)token``, CHAR as a character literal, STRING as a pointer
range into a device-side constant array.
"""
import dogqc_amd.scalarAlgebra as scal
from dogqc_amd.types import Type
from dogqc_amd.hiplang import langType, cLiteralString


class ExprEmitter(object):

    def __init__(self, program, decimalRepr, access, pipe=None):
        self.program = program
        self.decimalRepr = decimalRepr
        self.access = access  # attr -> C expression of its register
        self.pipe = pipe

    def emit(self, e):
        m = getattr(self, "_" + e.__class__.__name__, None)
        if m is None:
            raise NotImplementedError("scalar expression " + e.__class__.__name__)
        return m(e)

    def _AttrExpr(self, e):
        return self.access(e.attr)

    def _ConstExpr(self, e):
        t = e.type
        if t in (Type.INT, Type.DATE, Type.DOUBLE, Type.INT64):
            return "(" + e.token + ")"
        if t == Type.FLOAT:
            return "((float)" + e.token + ")"
        if t == Type.CHAR:
            if len(e.token) != 1:
                raise SyntaxError("CHAR constant must be one character: %r" % e.token)
            return "((int8_t)%d)" % ord(e.token)
        if t == Type.STRING:
            return self.program.stringConstant(e.token)
        raise NotImplementedError("constant of type " + str(t))

    def _bin(self, e, op):
        return "(%s %s %s)" % (self.emit(e.leftChild), op, self.emit(e.rightChild))

    def _EqualsExpr(self, e):
        if e.leftChild.type == Type.STRING:
            return "dqc::string_equals(%s, %s)" % (self.emit(e.leftChild), self.emit(e.rightChild))
        return self._bin(e, "==")

    def _LikeExpr(self, e):
        import dogqc_amd.scalarAlgebra as scal
        r = e.rightChild
        if isinstance(r, scal.ConstExpr) and r.type == Type.STRING and "_" not in r.token:
            fn = self.program.likeMatcher(r.token)
            return "%s(%s)" % (fn, self.emit(e.leftChild))
        return "dqc::string_like(%s, %s)" % (self.emit(e.leftChild), self.emit(r))

    def _SmallerExpr(self, e):
        return self._bin(e, "<")

    def _SmallerEqualExpr(self, e):
        return self._bin(e, "<=")

    def _LargerExpr(self, e):
        return self._bin(e, ">")

    def _LargerEqualExpr(self, e):
        return self._bin(e, ">=")

    def _AndExpr(self, e):
        return self._bin(e, "&&")

    def _OrExpr(self, e):
        return self._bin(e, "||")

    def _NotExpr(self, e):
        return "(!%s)" % self.emit(e.child)

    def _MulExpr(self, e):
        return self._bin(e, "*")

    def _DivExpr(self, e):
        return self._bin(e, "/")

    def _AddExpr(self, e):
        return self._bin(e, "+")

    def _SubExpr(self, e):
        return self._bin(e, "-")

    def _AbsExpr(self, e):
        fn = "fabs" if e.child.type in (Type.FLOAT, Type.DOUBLE) else "abs"
        return "%s(%s)" % (fn, self.emit(e.child))

    def _ExtractExpr(self, e):
        c = self.emit(e.child)
        if e.extract == scal.ExtractType.DAY:
            return "(%s %% 100)" % c
        if e.extract == scal.ExtractType.MONTH:
            return "((%s / 100) %% 100)" % c
        return "(%s / 10000)" % c

    def _SegmentExpr(self, e):
        """Batch segment of the scanned row: binary search of its index in
        the members' prefix row counts of the table (batch input column
        `__batch.<table>`, one INT64 per member, nondecreasing)."""
        pipe = self.pipe
        if pipe is None or getattr(pipe, "scanTable", None) != e.table:
            raise ValueError("SegmentExpr(%r) must be mapped directly above the scan of that table" % e.table)
        prog = self.program
        bt = prog.table("__batch")
        ci = prog.column("__batch", e.table, None, "int64_t")
        f = pipe.param("c%d" % ci, "const int64_t*", "C_%d" % ci)
        nb = pipe.param("bmembers", "int64_t", "L.rows[%d]" % bt)
        if pipe.loopOps:
            # rows were reassigned to lanes: no common tile, one search per row
            return "dqc::segment_of(%s, %s, %s[u])" % (f, nb, pipe.tid)
        return "dqc::segment_of(%s, %s, %s[u], base, %d)" % (f, nb, pipe.tid, pipe.bs * pipe.U)

    def _MaxExpr(self, e):
        return "std::max<int32_t>(%s, %s)" % (self.emit(e.leftChild), self.emit(e.rightChild))

    def _SubstringExpr(self, e):
        return "dqc::string_substring(%s, %s, %s)" % (self.emit(e.exprString), self.emit(e.exprFrom),
                                                      self.emit(e.exprFor))

    def _CaseExpr(self, e):
        ct = langType(Type.FLOAT, self.decimalRepr)
        out = "((%s)%s)" % (ct, self.emit(e.exprElse)) if e.exprElse is not None else "((%s)0)" % ct
        for w, t in reversed(e.exprListWhenThen):
            out = "(%s ? (%s)%s : %s)" % (self.emit(w), ct, self.emit(t), out)
        return out


def likeMatcherDecl(name, pattern, lit):
    """Device function matching LIKE `pattern` (literals and '%' only) with
    dqc::string_like's semantics (util.cuh:148-199).  When the pattern holds
    a '%' and no "%%", a string of at least P + S bytes (P before the first
    '%', S after the last) has the textbook answer: the first segment is a
    prefix, the last a suffix (unless the pattern starts / ends with '%'),
    the others are found in order in between.  Shorter strings, and other
    patterns, call dqc::string_like.  `lit(text)` names the device constant
    holding a literal."""
    segs = pattern.split("%")
    body = ["static __device__ __forceinline__ bool %s(dqc::str_t s) {" % name]
    if len(segs) == 1 or "%%" in pattern:
        body.append("    return dqc::string_like(s, %s);" % lit(pattern)[0])
        body.append("}")
        return "\n".join(body)
    pre, mid, suf = segs[0], segs[1:-1], segs[-1]
    n = lambda t: len(t.encode("latin-1"))
    body.append("    const char* pos = s.start;")
    body.append("    const char* lim = s.end;")
    if pre or suf:
        body.append("    if (s.end - s.start < %d) return dqc::string_like(s, %s);"
                    % (n(pre) + n(suf), lit(pattern)[0]))
    if suf:
        body.append("    lim = s.end - %d;" % n(suf))
        body.append("    if (!dqc::bytes_equal(lim, %s, %du)) return false;" % (lit(suf)[1], n(suf)))
    if pre:
        body.append("    if (!dqc::bytes_equal(pos, %s, %du)) return false;" % (lit(pre)[1], n(pre)))
        body.append("    pos += %d;" % n(pre))
    for m in mid:
        b = m.encode("latin-1")
        H = min(len(b), 8)
        head = 0
        for k in range(H):
            head = (head << 8) | b[k]
        hmask = (1 << (8 * H)) - 1
        body.append("    pos = dqc::string_find(pos, lim, %s, %du, 0x%xull, 0x%xull, %du);"
                    % (lit(m)[1], len(b), head, hmask, H))
        body.append("    if (!pos) return false;")
    body.append("    return true;")
    body.append("}")
    return "\n".join(body)


def stringConstantDecl(name, token):
    return "static __device__ const char %s[] = %s;" % (name, cLiteralString(token))
