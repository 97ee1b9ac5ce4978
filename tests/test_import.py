"""The .tbl importer (libdqcimport.so, host code) and its oracle.

  * the oracle's restatement of the reference's CSV parsing (oracle/tblparse.py)
    against the golden vectors of the reference's own reader
    (tests/golden/tbl_golden.json, written by tests/golden/make_tbl_golden.py
    from csv.h) and, where oracle/_ref/tbl_ref is built, against that reader
    live on freshly generated fixtures;
  * the native importer against the golden vectors, bit for bit, and against
    the oracle on inputs large enough for many parallel chunks;
  * the first error in line order is the one reported, with no files left;
  * dbAccess(schema, dbpath, csvpath) imports once and reuses the database.
"""
import ctypes
import json
import os
import random
import struct
import subprocess
import sys

import numpy as np
import pytest

from dogqc_amd import build, dbio
from dogqc_amd.types import Type
from oracle import tblparse as tp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "tbl_golden.json")
REF = os.path.join(ROOT, "oracle", "_ref", "tbl_ref")
HEADER = os.path.join(ROOT, "include", "dogqc_amd_import.h")

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_tbl_golden as mk  # noqa: E402

LETTER_TYPE = {"i": Type.INT, "l": Type.INT64, "f": Type.FLOAT, "d": Type.DOUBLE, "c": Type.CHAR, "t": Type.DATE,
               "s": Type.STRING}
LETTER_ORACLE = {"i": tp.INT, "l": tp.INT64, "f": tp.FLOAT, "d": tp.DOUBLE, "c": tp.CHAR, "t": tp.DATE,
                 "s": tp.STRING}


def cases():
    with open(GOLDEN) as f:
        return json.load(f)


def fmt(v, t):
    """A value as the reference reader's driver prints it."""
    if t == "f":
        return "%08x" % struct.unpack("<I", np.float32(v).tobytes())[0]
    if t == "d":
        return "%016x" % struct.unpack("<Q", np.float64(v).tobytes())[0]
    if t == "s":
        return "s" + v.hex()
    return str(v)


def oracleOutcome(text, letters):
    try:
        rows = tp.parse(text, [LETTER_ORACLE[x] for x in letters])
    except tp.TblError as e:
        return {"error": {"line": e.line, "kind": e.kind}}
    return {"rows": ["\t".join(fmt(v, t) for v, t in zip(r, letters)) for r in rows]}


@pytest.fixture(scope="module")
def importer():
    if not os.path.exists(dbio.IMPORT_SO):
        build.buildImport()
    return dbio.importLib()


def importFile(tmp_path, text, letters, threads=4):
    tbl = tmp_path / "x.tbl"
    tbl.write_bytes(text)
    table = {"name": "x", "attributes": {"a%d" % k: LETTER_TYPE[t] for k, t in enumerate(letters)}}
    rows, chars = dbio.importTable(str(tmp_path), table, str(tbl), threads=threads)
    acc = dbio.Accessor(str(tmp_path), {"x": dict(table, size=rows)})
    return rows, [acc.column("x", "a%d" % k) for k in range(len(letters))], chars


def assertColumnsEqual(got, want):
    for k, (g, w) in enumerate(zip(got, want)):
        if isinstance(w, tuple):
            assert np.array_equal(g[0], w[0]) and np.array_equal(g[1], w[1]), "column %d" % k
        else:
            assert g.dtype == w.dtype and np.array_equal(g.view(np.uint8), w.view(np.uint8)), "column %d" % k


# ---------------------------------------------------------------- oracle

@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_oracle_matches_reference_reader_golden(case):
    want = {k: case[k] for k in ("rows", "error") if k in case}
    assert oracleOutcome(case["input"].encode("latin-1"), case["types"]) == want


def test_golden_floats_are_not_correctly_rounded():
    """The fixtures tell step-by-step float parsing from strtof: a correctly
    rounding parser fails many of them."""
    case = [c for c in cases() if c["name"] == "floats"][0]
    differs = 0
    for line, row in zip(case["input"].splitlines(), case["rows"]):
        s = line.split("|")[0].replace(",", ".")
        if fmt(np.float32(float(s)), "f") != row.split("\t")[0]:
            differs += 1
    assert differs > 50


@pytest.mark.skipif(not os.path.exists(REF), reason="reference reader not built (make -C oracle ref)")
@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_matches_reference_reader_live(seed):
    for name, letters, text in [("lineitem", mk.LINEITEM, mk.lineitemRows(400, seed)),
                                ("floats", "fd", mk.floatRows(800, seed))]:
        assert oracleOutcome(text.encode("latin-1"), letters) == mk.run(text, letters), name


def test_planted_float_mistake_is_caught(monkeypatch):
    """A correctly rounded parse in place of the step-by-step one fails the
    golden vectors."""
    monkeypatch.setattr(tp, "_float", lambda s, T: T(float(s.replace(b",", b".") or b"0")))
    case = [c for c in cases() if c["name"] == "floats"][0]
    assert oracleOutcome(case["input"].encode("latin-1"), case["types"])["rows"] != case["rows"]


# ---------------------------------------------------------------- importer

def test_header_and_exports(importer):
    import re
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(dqc_[a-z0-9_]+)\s*\(", text)))
    assert names == ["dqc_import_tbl", "dqc_tbl_rows"]
    for n in names:
        assert hasattr(importer, n)
    assert ctypes.sizeof(dbio.ImportReport) == 8 + 8 + 4 + 4 + 8 * 64 + 256


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_importer_matches_golden(importer, tmp_path, case):
    text = case["input"].encode("latin-1")
    if "error" in case:
        with pytest.raises(dbio.ImportError_) as ei:
            importFile(tmp_path, text, case["types"])
        assert (ei.value.line, ei.value.kind) == (case["error"]["line"], case["error"]["kind"])
        assert sorted(os.listdir(tmp_path)) == ["x.tbl"]        # nothing left behind
        return
    rows, got, _ = importFile(tmp_path, text, case["types"])
    assert rows == len(case["rows"])
    printed = ["\t".join(fmt(got[k][i], t) if t != "s" else
                         "s" + got[k][1][int(got[k][0][i]):int(got[k][0][i + 1])].tobytes().hex()
                         for k, t in enumerate(case["types"])) for i in range(rows)]
    assert printed == case["rows"]
    assert dbio.tblRows(str(tmp_path / "x.tbl")) == rows


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_importer_many_chunks_matches_oracle(importer, tmp_path, threads):
    text = mk.lineitemRows(24000, 5).encode()          # about 3 MB: several 1 MB chunks
    assert len(text) > 2 * (1 << 20)
    rows, got, chars = importFile(tmp_path, text, mk.LINEITEM, threads)
    want = tp.columns(text, [LETTER_ORACLE[x] for x in mk.LINEITEM])
    assert rows == 24000
    assertColumnsEqual(got, want)
    assert chars == {"a13": len(want[13][1]), "a14": len(want[14][1]), "a15": len(want[15][1])}


def test_importer_reports_first_error_in_line_order(importer, tmp_path):
    lines = mk.lineitemRows(24000, 6).encode().split(b"\n")
    bad = list(lines)
    bad[20000] = bad[20000].replace(b"|", b"||", 1)      # wrong field count, a late chunk
    cols = bad[9000].split(b"|")
    cols[4] = b"1x"                                       # bad digit, an earlier chunk
    bad[9000] = b"|".join(cols)
    with pytest.raises(dbio.ImportError_) as ei:
        importFile(tmp_path, b"\n".join(bad), mk.LINEITEM, 8)
    assert (ei.value.line, ei.value.column, ei.value.kind) == (9001, 4, tp.DIGIT)
    bad = list(lines)
    bad[20000] = bad[20000].replace(b"|", b"||", 1)
    bad[20001] = bad[20001].replace(b"A|", b"AB|", 1)
    with pytest.raises(dbio.ImportError_) as ei:
        importFile(tmp_path, b"\n".join(bad), mk.LINEITEM, 8)
    assert (ei.value.line, ei.value.column, ei.value.kind) == (20001, -1, tp.COLUMNS)
    assert sorted(os.listdir(tmp_path)) == ["x.tbl"]


def test_short_date_is_an_error(importer, tmp_path):
    with pytest.raises(dbio.ImportError_) as ei:
        importFile(tmp_path, b"1995-03-15|\n1995-3-1|\n", "t")
    assert (ei.value.line, ei.value.kind) == (2, tp.SHORT_DATE)


def test_dbaccess_imports_once(importer, tmp_path):
    csv = tmp_path / "tbl"
    csv.mkdir()
    (csv / "lineitem.tbl").write_bytes(mk.lineitemRows(500, 7).encode())
    (csv / "nation.tbl").write_bytes(b"0|ALGERIA|0| haggle. carefully final deposits|\n"
                                     b"1|ARGENTINA|1|al foxes promise slyly|\n")
    names = ["l_orderkey", "l_partkey", "l_suppkey", "l_linenumber", "l_quantity", "l_extendedprice", "l_discount",
             "l_tax", "l_returnflag", "l_linestatus", "l_shipdate", "l_commitdate", "l_receiptdate", "l_shipinstruct",
             "l_shipmode", "l_comment"]
    schema = {"dateformat": "fixed YYYY-MM-DD",
              "lineitem": {"name": "lineitem", "size": 0, "numColumns": 16,
                           "attributes": dict(zip(names, [LETTER_TYPE[x] for x in mk.LINEITEM]))},
              "nation": {"name": "nation", "size": 0, "numColumns": 4,
                         "attributes": {"n_nationkey": Type.INT, "n_name": Type.STRING, "n_regionkey": Type.INT,
                                        "n_comment": Type.STRING}}}
    db = str(tmp_path / "mmdb")
    acc = dbio.dbAccess(schema, db, str(csv))
    assert acc.schema["lineitem"]["size"] == 500 and acc.schema["nation"]["size"] == 2
    assert acc.schema["nation"]["charSizes"]["n_name"] == len("ALGERIAARGENTINA")
    offs, chars = acc.column("nation", "n_name")
    assert chars[offs[1]:offs[2]].tobytes() == b"ARGENTINA"
    want = tp.columns(mk.lineitemRows(500, 7).encode(), [LETTER_ORACLE[x] for x in mk.LINEITEM])
    assertColumnsEqual([acc.column("lineitem", n) for n in names], want)
    stamp = os.stat(dbio.columnPath(db, "lineitem", "l_quantity")).st_mtime_ns
    acc2 = dbio.dbAccess(schema, db, str(csv))               # same sizes: no import
    assert os.stat(dbio.columnPath(db, "lineitem", "l_quantity")).st_mtime_ns == stamp
    assert acc2.schema["nation"]["charSizes"] == acc.schema["nation"]["charSizes"]
    acc3 = dbio.dbAccess(schema, db)                         # no csvpath: the cached schema
    assert acc3.schema["lineitem"]["size"] == 500
    (csv / "nation.tbl").write_bytes(b"0|ALGERIA|0|x|\n")    # changed size: imported again
    acc4 = dbio.dbAccess(schema, db, str(csv))
    assert acc4.schema["nation"]["size"] == 1
