"""Writes tests/golden/tbl_golden.json: .tbl fixtures and what the reference's
own CSV reader (csv.h, through oracle/_ref/tbl_ref, built by `make -C oracle
ref`) makes of them.

Each case holds the input bytes (latin-1), the column type letters
(i INT, l INT64, f FLOAT, d DOUBLE, c CHAR, t DATE, s STRING) and either the
rows as the driver prints them (integers in decimal, float/double bit
patterns in hex, strings as 's' + hex bytes) or the error (line, kind).

Run from the repository root:  python3 tests/golden/make_tbl_golden.py
"""
import json
import os
import random
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "tbl_ref")
OUT = os.path.join(ROOT, "tests", "golden", "tbl_golden.json")

LINEITEM = "iiiiifffccttt" + "sss"
WORDS = ["furiously", "regular", "accounts", "deposits", "sleep", "carefully", "final", "ironic", "packages",
         "blithely", "express", "pending", "requests", "quickly", "bold", "silent", "theodolites", "the"]
MODES = ["TRUCK", "MAIL", "REG AIR", "AIR", "FOB", "RAIL", "SHIP"]
INSTRUCT = ["DELIVER IN PERSON", "COLLECT COD", "NONE", "TAKE BACK RETURN"]


def date(r):
    return "%04d-%02d-%02d" % (r.randint(1992, 1998), r.randint(1, 12), r.randint(1, 28))


def lineitemRows(n, seed):
    """dbgen-like lineitem rows: 2-digit decimals, dates, words."""
    r = random.Random(seed)
    out = []
    for i in range(n):
        qty = r.randint(1, 50)
        price = r.randint(90000, 10500000)
        out.append("|".join([
            str(i // 4 + 1), str(r.randint(1, 200000)), str(r.randint(1, 10000)), str(i % 4 + 1), str(qty),
            "%d.%02d" % (price // 100, price % 100), "0.%02d" % r.randint(0, 10), "0.%02d" % r.randint(0, 8),
            r.choice("ANR"), r.choice("OF"), date(r), date(r), date(r), r.choice(INSTRUCT), r.choice(MODES),
            " ".join(r.choice(WORDS) for _ in range(r.randint(1, 6)))]) + "|")
    return "\n".join(out) + "\n"


def floatRows(n, seed):
    """Decimals with many digits, signs, exponents: where float arithmetic
    step by step differs from correct rounding."""
    r = random.Random(seed)
    out = []
    for _ in range(n):
        k = r.randint(0, 9)
        if k == 0:
            s = "%d.%0*d" % (r.randint(0, 10 ** 7), r.randint(1, 12), r.randint(0, 10 ** 9))
        elif k == 1:
            s = "%se%d" % (repr(r.uniform(-10, 10)), r.randint(-40, 40))
        elif k == 2:
            s = "%+.*f" % (r.randint(0, 15), r.uniform(-1e6, 1e6))
        elif k == 3:
            s = "%d,%d" % (r.randint(0, 999), r.randint(0, 999))
        elif k == 4:
            s = "%.3E" % r.uniform(-1e30, 1e30)
        else:
            s = "%d.%02d" % (r.randint(0, 10 ** 6), r.randint(0, 99))
        out.append("%s|%s|" % (s, s))
    return "\n".join(out) + "\n"


CASES = [
    ("lineitem", LINEITEM, lineitemRows(300, 1)),
    ("floats", "fd", floatRows(600, 2)),
    ("float_edges", "fd", "0.1|0.1|\n.5|.5|\n5.|5.|\n-|-|\n||\n+3.25|+3.25|\n-0.0|-0.0|\n1e0|1e0|\n1e-3|1e-3|\n"
                          "2.5e+2|2.5e+2|\n1E39|1E309|\n1e-50|1e-400|\n  7.125  |  7.125  |\n123456789.123|123456789.123|\n"
                          "1e99999999999|1e99999999999|\n0.30000000000000004|0.30000000000000004|\n"),
    ("int_edges", "il", "2147483647|9223372036854775807|\n-2147483648|-9223372036854775808|\n+5|+5|\n-|-|\n||\n"
                        " 42 | -42 |\n007|007|\n"),
    ("chars_dates", "ctc", "A|1995-03-15|x|\n Z |2000-02-29|  q|\n#|1992-01-01|||\n".replace("|||", "|%|")),
    ("strings", "ss", "hello world|  padded  |\n||\na,b;c|\"quoted\"|\n\xe9t\xe9|tab\there|\n"),
    ("crlf", "is", "1|one|\r\n2|two|\r\n"),
    ("no_final_newline", "is", "1|one|\n2|two|"),
    ("empty", "is", ""),
    ("err_too_many", "ii", "1|2|\n1|2|3|\n"),
    ("err_too_few", "ii", "1|2|\n3|\n"),
    ("err_no_terminal", "ii", "1|2\n"),
    ("err_empty_line", "ii", "1|2|\n\n3|4|\n"),
    ("err_digit", "ii", "1|2|\n3|4x|\n"),
    ("err_overflow", "i", "2147483648|\n"),
    ("err_underflow", "i", "-2147483649|\n"),
    ("err_overflow64", "l", "9223372036854775808|\n"),
    ("err_float", "f", "1.5|\n1.2.3|\n"),
    ("err_float_exp", "f", "1e5x|\n"),
    ("err_char_two", "c", "A|\nAB|\n"),
    ("err_char_empty", "c", "A|\n|\n"),
]

KINDS = [("columns", 1), ("invalid digit", 2), ("overflows", 3), ("underflows", 3), ("single character", 4)]


def run(text, types):
    with tempfile.NamedTemporaryFile("wb", suffix=".tbl", delete=False) as f:
        f.write(text.encode("latin-1"))
        path = f.name
    try:
        p = subprocess.run([DRIVER, path, types], stdout=subprocess.PIPE, check=False)
    finally:
        os.unlink(path)
    lines = p.stdout.decode("latin-1").splitlines()
    if lines and lines[-1].startswith("ERROR "):
        _, line, what = lines[-1].split(" ", 2)
        kind = [k for s, k in KINDS if s in what.lower()]
        if not kind:
            raise RuntimeError("unknown error text: " + what)
        return {"error": {"line": int(line), "kind": kind[0]}}
    if p.returncode != 0:
        raise RuntimeError("driver failed (%d)" % p.returncode)
    return {"rows": lines}


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the driver first: make -C oracle ref")
    out = []
    for name, types, text in CASES:
        res = run(text, types)
        res.update({"name": name, "types": types, "input": text})
        out.append(res)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote %d cases to %s" % (len(out), OUT))


if __name__ == "__main__":
    main()
