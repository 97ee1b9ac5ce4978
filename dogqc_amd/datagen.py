"""HBM-resident workload tables generated on the device (libdqcgen.so).

The serving workload keeps the tables of the four benchmark workloads
resident in HBM at a multiple of their base sizes (10x: about 20 GB of
columns) and runs every query on a row prefix of them.  Generating that
much data on the host and copying it over would take minutes, so the
columns are filled by device kernels, each value a pure function of (seed,
column id, row) (include/dogqc_amd_datagen.h).  The distributions are those
of ``workloads.generate``, except for the foreign keys of c1 (f_fk) and c3
(s_item, s_cust): a query reads a prefix of every table, so a key drawn
uniformly over the whole referenced table would miss that table's prefix
(at 1x of 10x tables, nine keys in ten).  Those keys are drawn so that every
prefix of the referencing table stays inside the same fraction of the
referenced one (dqc_gen_prefix_key), which keeps each config's join
selectivities at every query scale.  The random streams differ
(counter-based here, PCG64 there), so the two are different draws.

``ResidentTables.store`` maps (table, attr, part) to a device buffer in the
form ``runtime.BoundInputs`` takes as a pre-filled column store, and
``hostAccessor(rows)`` copies row prefixes back for the CPU oracle.
"""
import ctypes as C
import math
import os
from collections import OrderedDict

import numpy as np

from dogqc_amd import runtime as rt, workloads
from dogqc_amd.dbio import MemoryAccessor
from dogqc_amd.types import Type

PKG = os.path.dirname(os.path.abspath(__file__))
GEN_SO = os.path.join(PKG, "libdqcgen.so")

I8, I32, U32, I64 = 0, 1, 2, 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(GEN_SO):
            raise RuntimeError("libdqcgen.so is not built (run python3 -m dogqc_amd.build)")
        rt.lib()
        L = C.CDLL(GEN_SO)
        vp, i64, u64, u32, i32 = C.c_void_p, C.c_int64, C.c_uint64, C.c_uint32, C.c_int32
        L.dqc_gen_uniform.argtypes = [vp, i32, i64, u64, u32, i64, i64, vp]
        L.dqc_gen_decimal.argtypes = [vp, i64, u64, u32, i64, i64, C.c_double, vp]
        L.dqc_gen_pick_u32.argtypes = [vp, i64, u64, u32, vp, i64, vp]
        L.dqc_gen_affine_perm.argtypes = [vp, i64, u64, u64, vp]
        L.dqc_gen_inverse_cdf.argtypes = [vp, i64, u64, u32, vp, vp, i32, vp]
        L.dqc_gen_text.argtypes = [vp, vp, i64, i32, u64, u32, vp, vp, i32, i32, vp]
        L.dqc_gen_prefix_key.argtypes = [vp, i64, u64, u32, i64, u64, u64, vp]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise rt.DqcError("%s failed (%d)" % (what, rc))


def permParams(n):
    """(a, b) of the affine permutation i -> (a*i + b) mod n."""
    if n <= 1:
        return 0, 0
    a = max(1, int(n * 0.6180339887498949))
    while math.gcd(a, n) != 1:
        a += 1
    a %= n
    if math.gcd(a, n) != 1 or a == 0:
        a = 1
    return a, int(n * 0.3141592653589793) % n


class ResidentTables(object):
    """The tables of configs c0-c3 at `scale` times their base sizes, in HBM."""

    def __init__(self, scale=10.0, seed=2024, configs=("c0", "c1", "c2", "c3")):
        rt.device()
        self.scale, self.seed = scale, seed
        self.store = {}
        self.columnIds = {}      # (table, attr) -> column id of the random stream
        self.rows = OrderedDict()
        self.schema = OrderedDict()
        self._col = 0
        self._aux = []
        for cfg in configs:
            sch = workloads.schema(cfg)
            for t, attrs in sch.items():
                n = max(1, int(round(workloads.FULL_ROWS[cfg][t] * scale)))
                if cfg == "c3" and t == "nation":
                    n = workloads.FULL_ROWS[cfg][t]   # the 25 nations do not scale
                self.rows[t] = n
                self.schema[t] = attrs
            getattr(self, "_gen_" + cfg)()
        rt.check(rt.lib().dqc_stream_sync(None), "generation")

    # ---------------------------------------------------------- columns
    def _new(self, table, attr, nbytes, part=None):
        b = rt.DeviceBuffer(max(1, nbytes))
        self.store[(table, attr, part)] = b
        return b

    def _colId(self, table, attr):
        self._col += 1
        self.columnIds[(table, attr)] = self._col * 16     # text columns use column + word index
        return self._col * 16

    def _uniform(self, table, attr, dtype, lo, hi):
        n = self.rows[table]
        size = {I8: 1, I32: 4, U32: 4, I64: 8}[dtype]
        b = self._new(table, attr, n * size)
        _check(lib().dqc_gen_uniform(b.ptr, dtype, n, self.seed, self._colId(table, attr), lo, hi, None), attr)

    def _decimal(self, table, attr, lo, hi, div):
        n = self.rows[table]
        b = self._new(table, attr, n * 8)
        _check(lib().dqc_gen_decimal(b.ptr, n, self.seed, self._colId(table, attr), lo, hi, float(div), None), attr)

    def _dates(self, table, attr):
        n = self.rows[table]
        tab = rt.DeviceBuffer.fromArray(workloads.dateTable())
        self._aux.append(tab)
        b = self._new(table, attr, n * 4)
        _check(lib().dqc_gen_pick_u32(b.ptr, n, self.seed, self._colId(table, attr), tab.ptr, 2557, None), attr)

    def _prefixKey(self, table, attr, ref):
        """Foreign keys into table `ref` (keys from _perm) that a query
        reading the first f x rows of both tables always finds: a query at
        any scale keeps the config's join selectivities."""
        n, nk = self.rows[table], self.rows[ref]
        b = self._new(table, attr, n * 4)
        a, c = permParams(nk)
        _check(lib().dqc_gen_prefix_key(b.ptr, n, self.seed, self._colId(table, attr), nk, a, c, None), attr)

    def _perm(self, table, attr):
        n = self.rows[table]
        b = self._new(table, attr, n * 4)
        a, c = permParams(n)
        _check(lib().dqc_gen_affine_perm(b.ptr, n, a, c, None), attr)

    # ---------------------------------------------------------- configs
    def _gen_c0(self):
        t = "cfg0"
        self._uniform(t, "g1", I8, 0, 2)
        self._uniform(t, "g2", I8, 0, 1)
        self._uniform(t, "qty", I32, 1, 50)
        self._decimal(t, "price", 90000, 10500000, 100)
        self._decimal(t, "disc", 0, 10, 100)
        self._decimal(t, "tax", 0, 8, 100)
        self._dates(t, "date")

    def _gen_c1(self):
        self._perm("dim", "d_key")
        self._uniform("dim", "d_cat", I8, ord("0"), ord("4"))
        self._uniform("dim", "d_seg", I32, 0, 999999)
        self._prefixKey("fact", "f_fk", "dim")
        self._decimal("fact", "f_price", 90000, 10500000, 100)
        self._decimal("fact", "f_disc", 0, 10, 100)
        self._dates("fact", "f_date")

    def _gen_c2(self):
        nvals = 100000
        ranks = np.arange(1, nvals + 1, dtype=np.float64)
        cdf = np.cumsum(1.0 / ranks)
        cdf /= cdf[-1]
        a, c = permParams(nvals)
        values = ((a * np.arange(nvals, dtype=np.int64) + c) % nvals).astype(np.int32)
        dcdf, dval = rt.DeviceBuffer.fromArray(cdf), rt.DeviceBuffer.fromArray(values)
        self._aux += [dcdf, dval]
        n = self.rows["build"]
        b = self._new("build", "b_key", n * 4)
        _check(lib().dqc_gen_inverse_cdf(b.ptr, n, self.seed, self._colId("build", "b_key"), dcdf.ptr, dval.ptr,
                                         nvals, None),
               "b_key")
        self._uniform("build", "b_pay", I64, 1, (1 << 20) - 1)
        self._uniform("probe", "p_key", I32, 0, 2 * nvals - 1)
        self._uniform("probe", "p_f", I32, 0, 99)
        self._uniform("probe", "p_pay", I64, 1, (1 << 20) - 1)

    def _gen_c3(self):
        self._perm("item", "i_key")
        n = self.rows["item"]
        w = workloads.TEXT_WIDTH
        words = [v.encode() for v in workloads.VOCAB]
        voffs = np.zeros(len(words) + 1, dtype=np.int32)
        np.cumsum([len(x) for x in words], out=voffs[1:])
        vchars = np.frombuffer(b"".join(words), dtype=np.uint8)
        dvc, dvo = rt.DeviceBuffer.fromArray(vchars), rt.DeviceBuffer.fromArray(voffs)
        self._aux += [dvc, dvo]
        offs = self._new("item", "i_text", (n + 1) * 8, "offset")
        chars = self._new("item", "i_text", n * w, "char")
        _check(lib().dqc_gen_text(offs.ptr, chars.ptr, n, w, self.seed, self._colId("item", "i_text"), dvc.ptr, dvo.ptr,
                                  len(words), 5, None), "i_text")
        self._perm("cust", "c_key")
        self._uniform("cust", "c_year", I32, 1992, 1998)
        self._perm("nation", "n_key")
        t = "sales"
        self._prefixKey(t, "s_item", "item")
        self._prefixKey(t, "s_cust", "cust")
        self._uniform(t, "s_nation", I32, 0, self.rows["nation"] - 1)
        self._uniform(t, "s_qty", I32, 1, 50)
        self._decimal(t, "s_price", 90000, 10500000, 100)

    # ---------------------------------------------------------- access
    def nbytes(self):
        return sum(b.nbytes for b in self.store.values())

    def schemaAccessor(self):
        """Schema-only accessor (table dicts with the resident row counts)."""
        class _Acc(object):
            pass
        acc = _Acc()
        acc.schema = OrderedDict()
        for t, attrs in self.schema.items():
            acc.schema[t] = {"name": t, "size": self.rows[t], "numColumns": len(attrs),
                             "attributes": OrderedDict(attrs)}
        return acc

    def hostAccessor(self, prefix):
        """MemoryAccessor over copies of the first prefix[t] rows of the
        given tables (for the CPU oracle)."""
        tables = OrderedDict()
        for t, n in prefix.items():
            cols = OrderedDict()
            for a, typ in self.schema[t].items():
                if typ == Type.STRING:
                    offs = self.store[(t, a, "offset")].toArray(np.uint64, n + 1)
                    chars = self.store[(t, a, "char")].toArray(np.uint8, max(1, int(offs[-1])))
                    cols[a] = (typ, (offs, chars))
                else:
                    dt = {Type.INT: np.int32, Type.DATE: np.uint32, Type.CHAR: np.int8, Type.INT64: np.int64,
                          Type.DOUBLE: np.float64, Type.FLOAT: np.float32}[typ]
                    cols[a] = (typ, self.store[(t, a, None)].toArray(dt, n))
            tables[t] = cols
        return MemoryAccessor(tables)
