"""String functions of the plan API, restated on the CPU: LIKE, SUBSTRING
and string equality with the reference's semantics.

TEST INFRASTRUCTURE ONLY: the oracle's scalar evaluator (planeval.py) calls
these; the product path never does.

LIKE follows `stringLikeCheck` (dogqc/include/util.cuh:148-199) step by step,
which is not textbook SQL LIKE:

  * the literal prefix (before the first '%') and the literal suffix (after
    the last '%') are compared independently, each only as far as the string
    reaches, so they may overlap or run past the string's ends
    ('ab' LIKE 'ab%b' and 'a' LIKE 'a_' hold);
  * a pattern without '%' holds for every string that starts and ends with
    it ('abab' LIKE 'ab');
  * the segments between the first and last '%' are found left to right in
    the bytes between prefix and suffix, and an empty segment ('%%' inside
    the pattern) only matches a literal '%' byte of the string.

Reads past either operand: the reference reads the string byte at the end of
the searched region (`*sTrace` before `sTrace < sInEnd`, util.cuh:185),
which cannot change the result (the conjunction is false either way), and
pattern bytes past the end only after an empty segment matched a '%' of the
string; pattern constants are NUL-terminated C literals there
(codegen.py:47-50), so this restatement reads 0 past the pattern's end.
The empty pattern holds for every string whatever byte precedes it.
"""
import re

import numpy as np


def _cmp(c, l):
    """cmpLike (util.cuh:144-146): equal bytes, or '_' in the pattern."""
    return c == l or l == 0x5F


def like(s, p):
    """`s LIKE p` as util.cuh:148-199 computes it (s, p: bytes)."""
    n, m = len(s), len(p)
    pc = lambda i: p[i] if 0 <= i < m else 0         # NUL past the literal
    PCT = 0x25
    lPos = sPos = None
    lInStart, lInEnd = 0, m
    sInStart, sInEnd = 0, n
    # prefix (util.cuh:155-165)
    if pc(0) != PCT:
        sPos, lPos = 0, 0
        while lPos < m and sPos < n and pc(lPos) != PCT:
            if not _cmp(s[sPos], pc(lPos)):
                return False
            lPos += 1
            sPos += 1
        lInStart, sInStart = lPos, sPos
    # suffix (util.cuh:167-177); the empty pattern reads the byte before
    # the literal here, and both outcomes lead to a match below
    if m > 0 and pc(m - 1) != PCT:
        sPos, lPos = n - 1, m - 1
        while lPos >= 0 and sPos >= 0 and pc(lPos) != PCT:
            if not _cmp(s[sPos], pc(lPos)):
                return False
            lPos -= 1
            sPos -= 1
        lInEnd, sInEnd = lPos, sPos + 1
    elif m == 0:
        lInEnd, lPos = -1, -1
    # infixes (util.cuh:180-197)
    if lInStart < lInEnd:
        lPos = lInStart + 1
        sPos = sInStart
        while sPos < sInEnd and lPos < lInEnd:
            lTrace, sTrace = lPos, sPos
            while sTrace < sInEnd and _cmp(s[sTrace], pc(lTrace)):
                lTrace += 1
                if pc(lTrace) == PCT:
                    lTrace += 1
                    lPos = lTrace
                    sPos = sTrace
                    break
                sTrace += 1
            sPos += 1
    return lPos >= lInEnd


def _textbookRegex(p):
    out = []
    for ch in p.decode("latin-1"):
        out.append(".*" if ch == "%" else "." if ch == "_" else re.escape(ch))
    return re.compile("".join(out), re.DOTALL)


def likeColumn(values, p):
    """LIKE over a column of byte strings.  For a pattern that holds at
    least one '%' and no "%%", a string of at least P + S bytes (P bytes
    before the first '%', S after the last) gets the textbook answer: the
    prefix and suffix are then compared whole and apart, and the left-to-right
    search of non-empty segments between them is the textbook match.  Those
    rows take a regular expression; every other row takes `like`."""
    out = np.empty(len(values), dtype=bool)
    fast = b"%" in p and b"%%" not in p
    if fast:
        P, S = p.index(b"%"), len(p) - 1 - p.rindex(b"%")
        rx = _textbookRegex(p)
    for i, s in enumerate(values):
        if fast and len(s) >= P + S:
            out[i] = rx.fullmatch(s.decode("latin-1")) is not None
        else:
            out[i] = like(s, p)
    return out


def substring(s, frm, ln):
    """SUBSTRING(s FROM frm FOR ln) (util.cuh:132-142): start = frm - 1,
    end = start + ln, then start clipped up to 0 and end down to len(s).
    Where that leaves end < start the reference yields a string of negative
    length; here it is the empty string."""
    st = frm - 1
    en = st + ln
    st, en = max(st, 0), min(en, len(s))
    return s[st:en] if st < en else b""


def equals(a, b):
    """stringEquals (util.cuh:90-102): equal lengths and bytes."""
    return a == b
