"""mmdb column store: reader, writer and the database accessor of the plan API.

File format (dogqc/include/mappedmalloc.h:11-82, dbio.py:207-251):
  * one file per column, ``<dbpath>/<table>_<attribute>``;
  * STRING columns are two files, ``..._offset`` (uint64 ``offs[N+1]``,
    ``offs[0] = 0``) and ``..._char`` (the concatenated bytes);
  * each file is an 8-byte little-endian header holding
    ``filesize = 8 + payload_bytes + 1``, the payload at byte 8, and one
    trailing NUL byte.
The reference writer stores only the low 4 header bytes (mappedmalloc.h:26),
which breaks files of 4 GiB and more; this writer stores all 8, and a file
written by the reference reads identically (its upper 4 bytes are zero).

``dbAccess`` keeps the reference signature (dbio.py:13-27): given a
``csvpath`` it imports ``<csvpath>/<table>.tbl`` text tables (dbgen output)
with the native importer ``libdqcimport.so`` (include/dogqc_amd_import.h)
unless ``<dbpath>`` already holds them at the same sizes.  Synthetic tables are
written column by column with ``writeTable``.  The schema is cached as JSON in
``<dbpath>/dbschema.json`` (the reference pickles it, dbio.py:62-64).
"""
import ctypes as C
import json
import os
from collections import OrderedDict

import numpy as np

from dogqc_amd.types import Type, COLUMN_DTYPE

HEADER_BYTES = 8


def columnPath(dbpath, table, attr, part=None):
    name = "%s_%s" % (table, attr)
    if part is not None:
        name += "_" + part
    return os.path.join(dbpath, name)


def writeMmdb(path, data):
    """Write one mmdb file from a numpy array or bytes."""
    buf = data.tobytes() if isinstance(data, np.ndarray) else bytes(data)
    filesize = HEADER_BYTES + len(buf) + 1
    with open(path, "wb") as f:
        f.write(int(filesize).to_bytes(8, "little"))
        f.write(buf)
        f.write(b"\0")


def payloadBytes(path):
    with open(path, "rb") as f:
        filesize = int.from_bytes(f.read(8), "little")
    if filesize < HEADER_BYTES + 1:
        raise IOError("corrupt mmdb header in " + path)
    return filesize - HEADER_BYTES - 1


def readMmdb(path, dtype):
    """Memory-map the payload of an mmdb file as a read-only numpy array."""
    nbytes = payloadBytes(path)
    itemsize = np.dtype(dtype).itemsize
    if nbytes % itemsize:
        raise IOError("payload of %s is not a multiple of %d bytes" % (path, itemsize))
    if nbytes == 0:
        return np.zeros(0, dtype=dtype)
    return np.memmap(path, dtype=dtype, mode="r", offset=HEADER_BYTES, shape=(nbytes // itemsize,))


def _typeName(t):
    return t.name


def _typeFromName(n):
    return Type[n]


class Accessor(object):
    """Schema plus column access for one mmdb directory (dbio.py:30-86)."""

    def __init__(self, dbpath, schema):
        self.dbpath = dbpath
        self.schema = schema

    def file(self, table, attr, part=None):
        return columnPath(self.dbpath, table, attr, part)

    def column(self, table, attr):
        """numpy view of a column; STRING returns (offsets, chars)."""
        typ = self.schema[table]["attributes"][attr]
        if typ == Type.STRING:
            return (readMmdb(self.file(table, attr, "offset"), np.uint64),
                    readMmdb(self.file(table, attr, "char"), np.uint8))
        return readMmdb(self.file(table, attr), COLUMN_DTYPE[typ])

    def writeSchema(self):
        out = {}
        for name, t in self.schema.items():
            if not isinstance(t, dict):
                continue
            out[name] = {"name": t["name"], "size": int(t["size"]), "numColumns": t["numColumns"],
                         "attributes": [[a, _typeName(ty)] for a, ty in t["attributes"].items()],
                         "charSizes": dict(t.get("charSizes", {}))}
        with open(os.path.join(self.dbpath, "dbschema.json"), "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)

    @staticmethod
    def loadSchema(dbpath):
        path = os.path.join(dbpath, "dbschema.json")
        if not os.path.exists(path):
            return None
        with open(path) as f:
            raw = json.load(f)
        schema = {}
        for name, t in raw.items():
            schema[name] = {"name": t["name"], "size": t["size"], "numColumns": t["numColumns"],
                            "attributes": OrderedDict((a, _typeFromName(ty)) for a, ty in t["attributes"]),
                            "charSizes": t.get("charSizes", {})}
        return schema


def tableSizeFromFiles(dbpath, table):
    """Row count of a table from its first column file."""
    attr, typ = next(iter(table["attributes"].items()))
    if typ == Type.STRING:
        return payloadBytes(columnPath(dbpath, table["name"], attr, "offset")) // 8 - 1
    return payloadBytes(columnPath(dbpath, table["name"], attr)) // np.dtype(COLUMN_DTYPE[typ]).itemsize


IMPORT_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdqcimport.so")

# include/dogqc_amd_import.h DQC_COL_*
IMPORT_TYPE = {Type.INT: 0, Type.INT64: 1, Type.FLOAT: 2, Type.DOUBLE: 3, Type.CHAR: 4, Type.DATE: 5,
               Type.STRING: 6}


class ImportReport(C.Structure):
    _fields_ = [("rows", C.c_int64), ("error_line", C.c_int64), ("error_column", C.c_int32),
                ("error_kind", C.c_int32), ("char_bytes", C.c_int64 * 64), ("message", C.c_char * 256)]


class ImportError_(IOError):
    """A .tbl file the importer rejects; ``line``/``column``/``kind`` as in
    dqc_import_report."""

    def __init__(self, report):
        IOError.__init__(self, report.message.decode(errors="replace"))
        self.line, self.column, self.kind = report.error_line, report.error_column, report.error_kind


_importLib = None


def importLib():
    global _importLib
    if _importLib is None:
        if not os.path.exists(IMPORT_SO):
            raise RuntimeError("libdqcimport.so is not built (run python3 -m dogqc_amd.build)")
        L = C.CDLL(IMPORT_SO)
        L.dqc_tbl_rows.argtypes = [C.c_char_p, C.POINTER(C.c_int64)]
        L.dqc_import_tbl.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_char_p), C.c_int32,
                                     C.POINTER(ImportReport)]
        _importLib = L
    return _importLib


def tblRows(path):
    """Rows of a .tbl file (its lines; Importer.retrieveTableSizes,
    dbio.py:124-138)."""
    n = C.c_int64()
    rc = importLib().dqc_tbl_rows(path.encode(), C.byref(n))
    if rc != 0:
        raise IOError("cannot read %s (%d)" % (path, rc))
    return n.value


def importTable(dbpath, table, tblPath, threads=0):
    """Import one .tbl file into the mmdb column files of ``table`` (a schema
    dict); returns (rows, {STRING attr: char bytes}).  Replaces the
    reference's generated importer (dbio.py:172-244)."""
    attrs = list(table["attributes"].items())
    types = []
    for a, t in attrs:
        if t not in IMPORT_TYPE:
            raise ValueError("cannot import %s.%s of type %s" % (table["name"], a, t))
        types.append(IMPORT_TYPE[t])
    paths = [columnPath(dbpath, table["name"], a).encode() for a, _ in attrs]
    ctypes_ = (C.c_int32 * len(types))(*types)
    cpaths = (C.c_char_p * len(paths))(*paths)
    rep = ImportReport()
    rc = importLib().dqc_import_tbl(tblPath.encode(), len(types), ctypes_, cpaths, threads, C.byref(rep))
    if rc != 0:
        raise ImportError_(rep)
    chars = {a: int(rep.char_bytes[k]) for k, (a, t) in enumerate(attrs) if t == Type.STRING}
    return int(rep.rows), chars


def _sameTables(cached, sized):
    if cached is None:
        return False
    for name, t in sized.items():
        c = cached.get(name)
        if c is None or c["size"] != t["size"] or list(c["attributes"].items()) != list(t["attributes"].items()):
            return False
    return True


def dbAccess(schema, dbpath, csvpath=None, doReload=False, waitEnter=False):
    """Open an mmdb directory (dbio.py:13-27).  With ``csvpath``, table sizes
    come from the .tbl files, and the tables are imported when the directory
    does not hold them at those sizes or ``doReload`` is set.  Without it,
    sizes come from the cached schema or, failing that, the column files."""
    os.makedirs(dbpath, exist_ok=True)
    cached = Accessor.loadSchema(dbpath)
    if csvpath is not None:
        sized = OrderedDict()
        for name, t in schema.items():
            if not isinstance(t, dict):
                continue
            t = dict(t)
            t["attributes"] = OrderedDict(t["attributes"])
            t["size"] = tblRows(os.path.join(csvpath, name + ".tbl"))
            sized[name] = t
        if _sameTables(cached, sized) and not doReload:
            for name, t in sized.items():
                t["charSizes"] = cached[name].get("charSizes", {})
        else:
            if waitEnter:
                input("Press Enter to OVERWRITE database...")
            for name, t in sized.items():
                rows, chars = importTable(dbpath, t, os.path.join(csvpath, name + ".tbl"))
                t["size"], t["charSizes"] = rows, chars
        acc = Accessor(dbpath, dict(cached or {}))
        acc.schema.update(sized)
        acc.writeSchema()
        acc.schema = sized
        return acc
    sized = {}
    for name, t in schema.items():
        if not isinstance(t, dict):
            continue
        t = dict(t)
        t["attributes"] = OrderedDict(t["attributes"])
        if cached is not None and name in cached:
            t["size"] = cached[name]["size"]
            t["charSizes"] = cached[name].get("charSizes", {})
        else:
            t["size"] = tableSizeFromFiles(dbpath, t)
        sized[name] = t
    return Accessor(dbpath, sized)


def writeTable(dbpath, name, columns):
    """Write a table.  ``columns`` is an OrderedDict attr -> (Type, data) where
    data is a numpy array, or for STRING a list/array of bytes values or an
    (offsets, chars) pair.  Returns the schema dict of the table."""
    os.makedirs(dbpath, exist_ok=True)
    attrs = OrderedDict()
    size = None
    charSizes = {}
    for attr, (typ, data) in columns.items():
        attrs[attr] = typ
        if typ == Type.STRING:
            if isinstance(data, tuple):
                offs, chars = data
                offs = np.ascontiguousarray(offs, dtype=np.uint64)
                chars = np.ascontiguousarray(chars, dtype=np.uint8)
            else:
                lens = np.fromiter((len(s) for s in data), dtype=np.uint64, count=len(data))
                offs = np.zeros(len(data) + 1, dtype=np.uint64)
                np.cumsum(lens, out=offs[1:])
                chars = np.frombuffer(b"".join(data), dtype=np.uint8)
            n = len(offs) - 1
            writeMmdb(columnPath(dbpath, name, attr, "offset"), offs)
            writeMmdb(columnPath(dbpath, name, attr, "char"), chars)
            charSizes[attr] = int(len(chars))
        else:
            arr = np.ascontiguousarray(data, dtype=COLUMN_DTYPE[typ])
            n = len(arr)
            writeMmdb(columnPath(dbpath, name, attr), arr)
        if size is None:
            size = n
        elif size != n:
            raise ValueError("column %s.%s has %d rows, expected %d" % (name, attr, n, size))
    table = {"name": name, "size": int(size or 0), "numColumns": len(attrs), "attributes": attrs,
             "charSizes": charSizes}
    acc = Accessor(dbpath, Accessor.loadSchema(dbpath) or {})
    acc.schema[name] = table
    acc.writeSchema()
    return table


class MemoryAccessor(object):
    """An accessor over in-memory numpy columns (no files), for generated data.
    ``tables`` maps table name -> OrderedDict attr -> (Type, data)."""

    def __init__(self, tables):
        self.schema = {}
        self.data = {}
        for name, cols in tables.items():
            attrs = OrderedDict((a, t) for a, (t, _) in cols.items())
            n = None
            for a, (t, d) in cols.items():
                rows = (len(d[0]) - 1) if t == Type.STRING else len(d)
                n = rows if n is None else n
                if rows != n:
                    raise ValueError("ragged table " + name)
                self.data[(name, a)] = d
            self.schema[name] = {"name": name, "size": int(n or 0), "numColumns": len(attrs),
                                 "attributes": attrs}

    def column(self, table, attr):
        return self.data[(table, attr)]
