"""CPU restatement of the reference's .tbl parsing (test infrastructure only).

The reference imports TPC-H ``.tbl`` text with a generated C++ program
(dbio.py:172-244) around csv.h's ``io::CSVReader<ncols + 1,
trim_chars<' '>, no_quote_escape<'|'>>``: one column more than the table for
the terminal '|', spaces trimmed on both sides of every column, no quoting,
integer overflow an error.  Each column is parsed by csv.h's ``parse()`` for
the C type the importer declares for it (codegen.py:37-45):

  * INT / INT64: ``parse_signed_integer`` (csv.h:946-995): optional sign,
    decimal digits only, out of range an error, the empty string 0;
  * FLOAT: ``parse_float<float>`` (csv.h:1009-1066), every step in float
    arithmetic (so not correctly rounded); DOUBLE: the same steps in double
    (the reference parses DOUBLE as float too; 8-byte DOUBLE columns are this
    build's extension);
  * CHAR: exactly one character (csv.h:921-928);
  * DATE: the schema's toDate (query/schema/tpch.py:9-20), digits at the fixed
    positions of YYYY-MM-DD, no validation;
  * STRING: the trimmed bytes.

Lines end at '\\n', a '\\r' before it is dropped (csv.h:478-480), and an empty
line is a row with too few columns.  Pinned against a driver around the
reference's own csv.h (oracle/ref_tbl_driver.cpp) on the fixtures of
tests/golden/tbl_*.
"""
import numpy as np

INT, INT64, FLOAT, DOUBLE, CHAR, DATE, STRING = "INT", "INT64", "FLOAT", "DOUBLE", "CHAR", "DATE", "STRING"

# error kinds, the order of include/dogqc_amd_import.h
COLUMNS, DIGIT, OVERFLOW, SINGLE_CHAR, SHORT_DATE = 1, 2, 3, 4, 5

_LIMITS = {INT: (-(1 << 31), (1 << 31) - 1), INT64: (-(1 << 63), (1 << 63) - 1)}


class TblError(Exception):
    def __init__(self, line, column, kind):
        Exception.__init__(self, "line %d column %d: error %d" % (line, column, kind))
        self.line, self.column, self.kind = line, column, kind


def _int(s, typ):
    """csv.h parse_signed_integer: the value, or DIGIT / OVERFLOW."""
    neg = s[:1] == b"-"
    body = s[1:] if s[:1] in (b"-", b"+") else s
    x = 0
    lo, hi = _LIMITS[typ]
    for ch in body:
        if not 48 <= ch <= 57:
            raise ValueError(DIGIT)
        x = 10 * x + (ch - 48)
        if (-x if neg else x) < lo or (-x if neg else x) > hi:
            raise ValueError(OVERFLOW)
    return -x if neg else x


def _float(s, ftype):
    """csv.h parse_float<T>, each operation rounded to T (numpy scalars)."""
    with np.errstate(over="ignore", under="ignore"):
        return _floatSteps(s, ftype)


def _floatSteps(s, ftype):
    T = ftype
    i = 0
    n = len(s)
    neg = False
    if i < n and s[i] == 45:
        neg = True
        i += 1
    elif i < n and s[i] == 43:
        i += 1
    x = T(0)
    ten = T(10)
    while i < n and 48 <= s[i] <= 57:
        x = x * ten
        x = x + T(s[i] - 48)
        i += 1
    if i < n and s[i] in (46, 44):      # '.' or ','
        i += 1
        pos = T(1)
        while i < n and 48 <= s[i] <= 57:
            pos = pos / ten
            x = x + T(s[i] - 48) * pos
            i += 1
    if i < n and s[i] in (101, 69):     # 'e' or 'E'
        e = _exponent(s[i + 1:])
        if e != 0:
            if e < 0:
                base = T(0.1)
                e = -e
            else:
                base = T(10)
            while e != 1:
                if e & 1 == 0:
                    base = base * base
                    e >>= 1
                else:
                    x = x * base
                    e -= 1
            x = x * base
    elif i < n:
        raise ValueError(DIGIT)
    return -x if neg else x


def _exponent(s):
    """parse_signed_integer<int, set_to_max_on_overflow> of the exponent."""
    neg = s[:1] == b"-"
    body = s[1:] if s[:1] in (b"-", b"+") else s
    x = 0
    for ch in body:
        if not 48 <= ch <= 57:
            raise ValueError(DIGIT)
        x = 10 * x + (ch - 48)
        if x > (1 << 31) - 1 + (1 if neg else 0):
            return -(1 << 31) if neg else (1 << 31) - 1
    return -x if neg else x


def _date(s):
    if len(s) < 10:
        raise ValueError(SHORT_DATE)
    c = [b - 256 if b > 127 else b for b in s[:10]]    # char is signed
    d = 0
    for p, w in zip((0, 1, 2, 3, 5, 6, 8, 9), (10000000, 1000000, 100000, 10000, 1000, 100, 10, 1)):
        d += (c[p] - 48) * w
    return d & 0xFFFFFFFF


def _value(s, typ):
    if typ in (INT, INT64):
        return _int(s, typ)
    if typ == FLOAT:
        return _float(s, np.float32)
    if typ == DOUBLE:
        return _float(s, np.float64)
    if typ == CHAR:
        if len(s) != 1:
            raise ValueError(SINGLE_CHAR)
        return s[0] - 256 if s[0] > 127 else s[0]
    if typ == DATE:
        return _date(s)
    return bytes(s)


def parse(data, types):
    """Rows of a .tbl file (bytes) as a list of lists of values: ints, numpy
    float32/float64 scalars, bytes.  Raises TblError at the first bad row."""
    rows = []
    lines = data.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    for ln, line in enumerate(lines, 1):
        if line.endswith(b"\r"):
            line = line[:-1]
        cols = line.split(b"|")
        if len(cols) != len(types) + 1:
            raise TblError(ln, -1, COLUMNS)
        row = []
        for k, (c, typ) in enumerate(zip(cols, types)):
            try:
                row.append(_value(c.strip(b" "), typ))
            except ValueError as e:
                raise TblError(ln, k, e.args[0])
        rows.append(row)
    return rows


def columns(data, types):
    """parse() as numpy columns of the mmdb dtypes; STRING as (offsets, chars)."""
    rows = parse(data, types)
    out = []
    dt = {INT: np.int32, INT64: np.int64, FLOAT: np.float32, DOUBLE: np.float64, CHAR: np.int8, DATE: np.uint32}
    for k, typ in enumerate(types):
        vals = [r[k] for r in rows]
        if typ == STRING:
            offs = np.zeros(len(vals) + 1, dtype=np.uint64)
            np.cumsum([len(v) for v in vals], out=offs[1:])
            out.append((offs, np.frombuffer(b"".join(vals), dtype=np.uint8)))
        else:
            out.append(np.array(vals, dtype=dt[typ]))
    return out
