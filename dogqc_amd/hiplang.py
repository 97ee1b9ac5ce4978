"""Emission helpers for generated HIP C++: C types and an indenting code buffer.

``CType`` keeps the member names the reference's drivers pass to the compiler
(``CType.FP32`` / ``CType.FP64`` as ``decimalRepr``, cudalang.py:2-13) but the
values are the fixed-width C++ spellings used by the generated HIP code.
"""
from dogqc_amd.types import Type


class CType(object):
    INT = "int32_t"
    UINT = "uint32_t"
    UINT64 = "uint64_t"
    INT64 = "int64_t"
    FP32 = "float"
    FP64 = "double"
    CHAR = "int8_t"
    SIZE = "uint64_t"
    STR_TYPE = "dqc::str_t"
    STR_OFFS = "dqc::str_offs"
    BOOL = "int32_t"

    # limits used to initialise MIN/MAX aggregates (MAX starts at the lowest
    # finite value, not at DBL_MIN as in cudalang.py:33-40)
    maxValue = {INT: "INT32_MAX", UINT: "UINT32_MAX", INT64: "INT64_MAX", FP32: "FLT_MAX", FP64: "DBL_MAX",
                CHAR: "INT8_MAX"}
    lowestValue = {INT: "INT32_MIN", UINT: "0u", INT64: "INT64_MIN", FP32: "(-FLT_MAX)", FP64: "(-DBL_MAX)",
                   CHAR: "INT8_MIN"}
    size = {INT: 4, UINT: 4, UINT64: 8, INT64: 8, FP32: 4, FP64: 8, CHAR: 1, STR_TYPE: 16, STR_OFFS: 16, BOOL: 4}


def langType(relType, decimalRepr):
    """Register type of a relational type (codegen.py:37-45, plus INT64)."""
    if relType in (Type.FLOAT, Type.DOUBLE):
        return decimalRepr
    return {Type.INT: CType.INT, Type.DATE: CType.UINT, Type.CHAR: CType.CHAR, Type.STRING: CType.STR_TYPE,
            Type.INT64: CType.INT64, Type.BOOLEAN: CType.BOOL}[relType]


def columnType(relType):
    """Element type of a base-table column on device."""
    return {Type.INT: CType.INT, Type.DATE: CType.UINT, Type.CHAR: CType.CHAR, Type.FLOAT: CType.FP32,
            Type.DOUBLE: CType.FP64, Type.INT64: CType.INT64, Type.BOOLEAN: CType.CHAR}[relType]


# bit width of a key column when packed into an exact 63-bit tag; None means
# the key cannot be packed (the table then stores a hash tag and the keys)
PACK_BITS = {Type.INT: 32, Type.DATE: 32, Type.CHAR: 8, Type.BOOLEAN: 8}


def ident(name):
    """A C identifier fragment from an attribute name."""
    out = "".join(c if c.isalnum() else "_" for c in name)
    return out[:16]


class Code(object):
    """Text buffer with block nesting."""

    def __init__(self, indent=0):
        self.lines = []
        self.depth = indent

    def emit(self, line):
        for ln in line.split("\n"):
            self.lines.append("    " * self.depth + ln if ln else "")

    def open(self, header):
        self.emit(header + " {")
        self.depth += 1

    def close(self, trailer=""):
        self.depth -= 1
        self.emit("}" + trailer)

    def extend(self, other):
        for ln in other.lines:
            self.lines.append("    " * self.depth + ln if ln else "")

    def text(self):
        return "\n".join(self.lines) + ("\n" if self.lines else "")

    def mark(self):
        """Placeholder line index, to be replaced later with fill()."""
        self.lines.append(None)
        return len(self.lines) - 1, self.depth

    def fill(self, mark, text):
        idx, depth = mark
        self.lines[idx] = "\n".join("    " * depth + ln for ln in text.split("\n"))


def cLiteralString(s):
    out = []
    for ch in s:
        if ch in "\\\"":
            out.append("\\" + ch)
        elif 32 <= ord(ch) < 127:
            out.append(ch)
        else:
            out.append("\\x%02x" % (ord(ch) & 0xFF))
    return '"' + "".join(out) + '"'
