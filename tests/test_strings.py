"""The oracle's string functions (oracle/strfuncs.py), pinned without the
oracle's own evaluator:

  * LIKE against answers traced by hand through util.cuh:148-199, and
    against a closed form of that code derived by case analysis (below),
    over every string of at most 5 bytes on {a, b, c} and every pattern of
    at most 4 bytes on {a, b, %, _};
  * the column evaluator's textbook fast path against the step-by-step
    restatement;
  * SUBSTRING against Python slicing; equality against bytes equality;
  * planted mistakes in the restatements fail these checks.
"""
import inspect
import itertools
import random
import re

import numpy as np
import pytest

from oracle import strfuncs

# (string, pattern, util.cuh:148-199's answer), each traced through the C
# code by hand
KNOWN = [
    (b"ab", b"ab%b", True),     # prefix and suffix overlap
    (b"a", b"a_", True),        # the prefix loop stops at the string's end
    (b"ab", b"ab_", False),
    (b"abab", b"ab", True),     # no '%': starts and ends with the pattern
    (b"abcab", b"ab", True),
    (b"abc", b"ab", False),
    (b"", b"ab", True),         # 2n + 2 >= m with nothing compared
    (b"", b"abc", False),
    (b"c", b"cbc", True),
    (b"bc", b"%abc", True),     # the suffix loop stops at the string's start
    (b"c", b"%abc", False),
    (b"", b"%a", True),
    (b"", b"a%", False),
    (b"a", b"ab%", False),
    (b"a", b"ab%a", True),
    (b"abc", b"a%%c", True),    # nothing between the first and last '%'
    (b"abc", b"a%%%c", False),  # an empty segment needs a literal '%'
    (b"a%c", b"a%%%c", True),
    (b"abc", b"a%%", False),
    (b"abc", b"%%c", True),
    (b"", b"%%", False),
    (b"", b"%", True),
    (b"xyz", b"%", True),
    (b"", b"", True),           # the empty pattern matches every string
    (b"xyz", b"", True),
    (b"green", b"%green%", True),
    (b"gree", b"%green%", False),
    (b"a green b", b"%green%", True),
    (b"axbxc", b"a%b%c", True),
    (b"acb", b"a%b%c", False),
    (b"PROMO BRUSHED", b"PROMO%", True),
    (b"LARGE BRASS", b"%BRASS", True),
    (b"BRASS", b"%BRASS", True),
    (b"RASS", b"%BRASS", True),  # a suffix one byte longer than the string
    (b"ASS", b"%BRASS", False),
]


def _pm(x, y):
    """Bytes x match pattern bytes y ('_' any byte), equal lengths."""
    return len(x) == len(y) and all(c == l or l == 0x5F for c, l in zip(x, y))


def closedForm(s, p):
    """util.cuh:148-199 by case analysis, for strings without '%' bytes.

    P / S = bytes before the first / after the last '%' (both m without
    '%').  The prefix loop compares min(n, P) bytes and leaves
    lInStart = min(n, P); the suffix loop compares min(n, S) bytes and leaves
    lInEnd = m - 1 - min(n, S) (m when S = 0).  The infix scan runs only when
    n > P + S and the pattern has a '%'; otherwise the answer is
    lInEnd <= lInStart + 1.  When it runs it walks the segments of M, the
    bytes between the first and last '%' split at '%', that start before
    lInEnd: with S > 0 lInEnd is the last '%', so a trailing empty segment
    of M (M ending in '%') is never reached; with S = 0 lInEnd = m and every
    segment is walked.  The segments are searched left to right in
    s[P : n - S]; an empty one never matches a string without '%' bytes, and
    no segments at all (one '%', or S > 0 and M empty) is a match."""
    n, m = len(s), len(p)
    if m == 0:
        return True
    if b"%" in p:
        P, S = p.index(b"%"), m - 1 - p.rindex(b"%")
    else:
        P = S = m
    k1, k2 = min(n, P), min(n, S)
    if not _pm(s[:k1], p[:k1]) or not _pm(s[n - k2:], p[m - k2:]):
        return False
    if b"%" not in p or n <= P + S:
        lInEnd = m if S == 0 else m - 1 - k2
        return lInEnd <= k1 + 1
    first, last = P, m - 1 - S
    if first == last:
        return True
    segs = p[first + 1:last].split(b"%")
    if S > 0 and segs[-1] == b"":
        segs = segs[:-1]
    if any(x == b"" for x in segs):
        return False
    mid, pos = s[P:n - S], 0
    for x in segs:
        hit = next((j for j in range(pos, len(mid) - len(x) + 1) if _pm(mid[j:j + len(x)], x)), None)
        if hit is None:
            return False
        pos = hit + len(x)
    return True


def _all(alphabet, maxLen):
    out = [b""]
    for n in range(1, maxLen + 1):
        out += [bytes(t) for t in itertools.product(alphabet, repeat=n)]
    return out


STRINGS = _all(b"abc", 5)
PATTERNS = _all(b"ab%_", 4)


def _safe(like):
    """A mutant that raises counts as giving a wrong answer."""
    def f(s, p):
        try:
            return like(s, p)
        except IndexError:
            return None
    return f


def _checkKnown(like):
    return [(s, p, want) for s, p, want in KNOWN if like(s, p) != want]


def _checkClosed(like, strings=STRINGS, patterns=PATTERNS):
    bad = []
    for p in patterns:
        for s in strings:
            if like(s, p) != closedForm(s, p):
                bad.append((s, p))
                if len(bad) > 5:
                    return bad
    return bad


def test_like_known_answers():
    assert _checkKnown(strfuncs.like) == []


def test_closed_form_known_answers():
    """The closed form itself agrees with the hand-traced answers (those on
    strings without '%' bytes)."""
    assert [(s, p) for s, p, want in KNOWN if b"%" not in s and closedForm(s, p) != want] == []


def test_like_equals_closed_form_exhaustive():
    assert _checkClosed(strfuncs.like) == []


def test_like_equals_closed_form_longer_random():
    rng = random.Random(5)
    strings = [bytes(rng.choice(b"abc") for _ in range(rng.randint(0, 12))) for _ in range(400)]
    patterns = [bytes(rng.choice(b"abc%_") for _ in range(rng.randint(1, 9))) for _ in range(400)]
    assert _checkClosed(strfuncs.like, strings, patterns) == []


def test_like_textbook_where_the_reference_is():
    """A pattern with a '%' and no "%%" gives the textbook answer on strings
    of at least P + S bytes (the fast path's premise)."""
    for p in PATTERNS:
        if b"%" not in p or b"%%" in p:
            continue
        P, S = p.index(b"%"), len(p) - 1 - p.rindex(b"%")
        rx = re.compile(p.decode().replace("%", ".*").replace("_", "."), re.DOTALL)
        for s in STRINGS:
            if len(s) >= P + S:
                assert strfuncs.like(s, p) == (rx.fullmatch(s.decode()) is not None), (s, p)


def test_like_column_fast_path():
    strings = np.array(_all(b"ab%_", 4), dtype=object)
    for p in PATTERNS + [b"", b"%%ab%%", b"a%b%%", b"%_%_%"]:
        want = [strfuncs.like(s, p) for s in strings]
        assert list(strfuncs.likeColumn(strings, p)) == want, p


def _mutant(src, old, new):
    assert old in src, old
    ns = dict(strfuncs.__dict__)
    exec(compile(src.replace(old, new, 1), "<mutant>", "exec"), ns)
    return ns[src.split("(")[0].replace("def ", "").strip()]


LIKE_MUTANTS = [
    ("return lPos >= lInEnd", "return lPos > lInEnd"),
    ("lInEnd, sInEnd = lPos, sPos + 1", "lInEnd, sInEnd = lPos, sPos"),
    ("lPos = lInStart + 1", "lPos = lInStart"),
    ("lTrace += 1\n                    lPos = lTrace", "lPos = lTrace"),
    ("while lPos < m and sPos < n and", "while lPos < m and"),
    ("return c == l or l == 0x5F", "return c == l"),
]


@pytest.mark.parametrize("old,new", LIKE_MUTANTS)
def test_planted_like_mistakes_fail(old, new):
    src = inspect.getsource(strfuncs.like)
    if "return c == l" in old:
        src = inspect.getsource(strfuncs._cmp)
        cmp = _mutant(src, old, new)
        mutant = _mutant(inspect.getsource(strfuncs.like).replace("_cmp(", "cmp_("), "cmp_(", "cmp_(")
        mutant.__globals__["cmp_"] = cmp
    else:
        mutant = _mutant(src, old, new)
    mutant = _safe(mutant)
    assert _checkKnown(mutant) or _checkClosed(mutant), "planted mistake %r not caught" % new


def test_planted_textbook_like_fails():
    textbook = lambda s, p: re.fullmatch(re.escape(p.decode()).replace("%", ".*").replace("_", "."),
                                         s.decode(), re.DOTALL) is not None
    assert _checkKnown(textbook)


def sliceForm(s, frm, ln):
    """SUBSTRING by Python slicing: bytes [frm - 1, frm - 1 + ln) of s
    clipped to s, empty when the range is empty."""
    lo, hi = frm - 1, frm - 1 + ln
    return s[max(lo, 0):max(min(hi, len(s)), 0)] if hi > max(lo, 0) else b""


def _checkSubstr(f):
    bad = []
    for s in _all(b"ab", 4) + [b"abcdefgh"]:
        for frm in range(-3, 12):
            for ln in range(-3, 12):
                if f(s, frm, ln) != sliceForm(s, frm, ln):
                    bad.append((s, frm, ln))
    return bad


def test_substring_is_slicing():
    assert _checkSubstr(strfuncs.substring) == []
    assert strfuncs.substring(b"13-555", 1, 2) == b"13"        # TPC-H Q22's country code


@pytest.mark.parametrize("old,new", [("st = frm - 1", "st = frm"), ("en = st + ln", "en = st + ln - 1"),
                                     ("max(st, 0)", "st")])
def test_planted_substring_mistakes_fail(old, new):
    assert _checkSubstr(_mutant(inspect.getsource(strfuncs.substring), old, new))


def test_equals_is_bytes_equality():
    xs = _all(b"ab", 3)
    for a in xs:
        for b in xs:
            assert strfuncs.equals(a, b) == (a == b)
