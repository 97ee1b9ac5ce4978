"""HIP translator: relational plans -> one gfx950 HIP module per plan.

Replaces the reference's CudaCompiler (dogqc/cudaTranslator.py:31-101) and its
code emitter (cudalang.py, codegen.py, kernel.py, gpuio.py) with a produce /
consume translator that emits HIP C++ for CDNA4 directly:

  * every pipeline is one kernel with a block-uniform grid-stride loop over
    its source rows (cudaTranslator.py:284-297 uses a per-warp flush flag);
  * base-table columns are loaded lazily, at the first operator that needs
    them and only for lanes still active, instead of eagerly for every row
    (cudaTranslator.py:316-319), so filtered-out rows never fetch the columns
    only later operators read;
  * grouped aggregation goes through a per-workgroup LDS table with
    replicated accumulators before touching the global table
    (cudaTranslator.py:431-495 issues one global atomic per row and aggregate);
  * hash tables use the lock-free tag design of dqc_device.h;
  * the host part of the module exports the C ABI of include/dogqc_amd.h
    (sizes from runtime row counts, overflow status, no exit()).

Public API kept from the reference: ``HipCompiler(algebraContext, arch,
decimalRepr, debug)`` with ``translate``, ``gencode``, ``compile``,
``execute``, ``setBuffers``, ``setProfilers``, ``showGraph``.
"""
import json
import math
import struct
import sys
from collections import OrderedDict

import dogqc_amd.relationalAlgebra as ra
from dogqc_amd.relationalAlgebra import Reduction, MaterializationType
from dogqc_amd.types import Type, COLUMN_BYTES
from dogqc_amd.hiplang import CType, Code, langType, columnType, PACK_BITS, ident
from dogqc_amd.hipExpr import ExprEmitter, stringConstantDecl, likeMatcherDecl
from dogqc_amd.translatorBase import LiteralTranslator, UnaryTranslator, BinaryTranslator
from dogqc_amd.kernel import KernelCall
from dogqc_amd.util import listWrap

WAVE = 64
LDS_AGG_BUDGET = 48 * 1024


def _pow2ceil(x):
    return 1 << max(0, int(math.ceil(math.log2(max(1, x)))))


def _pow2floor(x):
    return 1 << max(0, int(math.floor(math.log2(max(1, x)))))


def _fillPattern(ctype, value):
    """64-bit fill pattern for an element type and a (Python) value."""
    if ctype == CType.FP64:
        b = struct.pack("<d", value)
    elif ctype == CType.FP32:
        b = struct.pack("<f", value) * 2
    elif ctype in (CType.INT, CType.BOOL):
        b = struct.pack("<i", value) * 2
    elif ctype == CType.UINT:
        b = struct.pack("<I", value) * 2
    elif ctype == CType.INT64:
        b = struct.pack("<q", value)
    elif ctype == CType.CHAR:
        b = struct.pack("<b", value) * 8
    else:
        raise ValueError(ctype)
    return "0x%016xull" % struct.unpack("<Q", b)[0]


def _identity(kind, ctype):
    """Initial accumulator value (Python) for an aggregate kind."""
    if kind in ("sum", "count"):
        return 0.0 if ctype in (CType.FP32, CType.FP64) else 0
    big = {CType.FP64: sys.float_info.max, CType.FP32: 3.4028234663852886e38, CType.INT: 2 ** 31 - 1,
           CType.UINT: 2 ** 32 - 1, CType.INT64: 2 ** 63 - 1, CType.CHAR: 127}
    low = {CType.FP64: -sys.float_info.max, CType.FP32: -3.4028234663852886e38, CType.INT: -2 ** 31,
           CType.UINT: 0, CType.INT64: -2 ** 63, CType.CHAR: -128}
    return big[ctype] if kind == "min" else low[ctype]


def _cIdentity(kind, ctype):
    if kind in ("sum", "count"):
        return "(%s)0" % ctype
    return "(%s)%s" % (ctype, CType.maxValue[ctype] if kind == "min" else CType.lowestValue[ctype])


class AttrState(object):
    __slots__ = ("attr", "reg", "ctype", "loaded", "loader", "deps")

    def __init__(self, attr, reg, ctype):
        self.attr = attr
        self.reg = reg
        self.ctype = ctype
        self.loaded = False
        self.loader = None   # C expression producing the value of row u
        self.deps = ()       # index registers the loader reads


class RowLoop(object):
    """Emits ``for (u = 0; u < U; ++u)`` around per-row code (nothing when a
    pipeline processes one row per thread and iteration)."""

    def __init__(self, pipe):
        self.pipe = pipe

    def __enter__(self):
        if self.pipe.U > 1:
            self.pipe.code.emit("#pragma unroll")
            self.pipe.code.open("for (int u = 0; u < %d; ++u)" % self.pipe.U)
        return self

    def __exit__(self, *exc):
        if self.pipe.U > 1:
            self.pipe.code.close()


class Pipeline(object):
    """One generated kernel.

    Each thread handles U rows per loop trip (U > 1 only for pipelines
    without inner loops): row u of trip ``base`` is ``base + u*B + tid``, so
    for every u a wavefront reads 64 consecutive rows.  Per-row registers are
    arrays ``name[U]`` indexed by ``u``; all per-row code sits inside
    ``rowLoop()`` blocks, so the U independent rows issue their loads
    together.  When U == 1 the index is the constant ``u = 0``."""

    def __init__(self, program, name, source, U=1):
        self.program = program
        self.name = name
        self.pipeId = len(program.pipes)
        self.bs = int(KernelCall.defaultBlockSize)
        if self.bs % WAVE or not (WAVE <= self.bs <= 1024):
            raise ValueError("block size must be a multiple of 64 in [64, 1024]")
        self.U = max(1, int(U))
        self.defaultGrid = int(KernelCall.defaultGridSize or 0)
        self.source = source
        self.params = OrderedDict()         # field -> (ctype, host expr)
        self.regs = OrderedDict()           # name -> (ctype, init, perRow)
        self.staticLds = []
        self.prologue = Code(1)
        self.code = Code(2)
        self.iterationEnd = Code(2)
        self.epilogue = Code(1)
        self.attrs = {}
        self.indexRegs = OrderedDict()
        self.phased = False
        self.phaseBetween = None     # launch steps between PHASE 0 and PHASE 1
        self.phaseRows = {}          # phase -> (host grid rows, device row count) overrides
        self.partAgg = None          # partitioned group-by: PHASE 1 gives workgroup b partition b
        self.ldsFilter = None        # join filter at the start of dynamic LDS: {"dev", "host"} byte counts
        self.dynLdsExpr = None
        self.refillPoints = []
        self.annotations = []
        self.innerDepth = 0
        self.probedJoins = []
        self.nullFlags = {}         # attribute id -> null flag (outer-join build attributes)
        self.loopOps = 0            # operators so far that reassign tuples to lanes
        self.uniqueCount = 0
        self.param("n", "int64_t", source["rows"])
        self.devRows = source.get("devRows")   # device-side row count (temp-table scans)
        self.tid = "tid_" + ident(source["tidName"])
        self.indexRegs[self.tid] = "int64_t"
        self.reg(self.tid, "int64_t")
        self.A = "active[u]"

    def rowLoop(self):
        return RowLoop(self)

    # ---- parameters and registers
    def param(self, field, ctype, hostExpr):
        if field in self.params and self.params[field] != (ctype, hostExpr):
            raise ValueError("conflicting kernel parameter " + field)
        self.params[field] = (ctype, hostExpr)
        return "p." + field

    def reg(self, name, ctype, init=None, perRow=True):
        """Declare a register (an array of U when perRow); returns the
        expression for the current row (``name[u]``) or the scalar name."""
        if name not in self.regs:
            self.regs[name] = (ctype, init, perRow)
        return name + "[u]" if self.regs[name][2] else name

    def emit(self, line):
        self.code.emit(line)

    # ---- attribute file
    def attrReg(self, attr):
        return "a%d_%s" % (attr.id, ident(attr.name))

    def declareAttr(self, attr, ctype=None):
        ctype = ctype or langType(attr.dataType, self.program.decimalRepr)
        st = AttrState(attr, self.attrReg(attr), ctype)
        self.attrs[attr.id] = st
        self.reg(st.reg, ctype)
        return st

    def setLazy(self, attr, loader, deps, ctype=None):
        st = self.declareAttr(attr, ctype)
        st.loader = loader
        st.deps = tuple(deps)
        return st

    def setComputed(self, attr, ctype=None):
        st = self.declareAttr(attr, ctype)
        st.loaded = True
        return st

    def has(self, attr):
        return attr.id in self.attrs

    def uniqueId(self):
        """A number not used before in this pipeline (names of locals)."""
        self.uniqueCount += 1
        return self.uniqueCount

    def access(self, attr):
        st = self.attrs.get(attr.id)
        if st is None:
            raise KeyError("attribute %s (%d) is not available in pipeline %s" % (attr.name, attr.id, self.name))
        if not st.loaded:
            raise RuntimeError("attribute %s accessed before load" % attr.name)
        return st.reg + "[u]"

    def need(self, attrs):
        """Emit loads for attributes not yet in registers (active rows only)."""
        pending = []
        seen = set()
        for a in attrs:
            st = self.attrs.get(a.id)
            if st is None:
                raise KeyError("attribute %s (%d) is not available in pipeline %s" % (a.name, a.id, self.name))
            if not st.loaded and a.id not in seen:
                pending.append(st)
                seen.add(a.id)
        if not pending:
            return
        with self.rowLoop():
            self.code.open("if (active[u])")
            for st in pending:
                # a callable loader is rendered now, for the lanes as they are here
                loader = st.loader() if callable(st.loader) else st.loader
                self.emit("%s[u] = %s;" % (st.reg, loader))
                st.loaded = True
            self.code.close()

    def liveState(self, attrs):
        """Registers that carry the given attributes across a lane
        reassignment: loaded registers plus index registers of lazy ones."""
        regs = OrderedDict()
        for a in attrs:
            st = self.attrs.get(a.id)
            if st is None:
                continue
            if st.loaded:
                regs[st.reg] = st.ctype
            else:
                for d in st.deps:
                    regs[d] = self.indexRegs.get(d) or self.regs[d][0]
            if a.id in self.nullFlags:
                r = self.nullFlags[a.id][:-3]     # the flag register, without [u]
                regs[r] = self.regs[r][0]
        return regs

    # ---- rendering
    def render(self):
        U, B = self.U, self.bs
        k = Code()
        k.open("struct KP_%s" % self.name)
        for f, (ct, _) in self.params.items():
            k.emit("%s %s;" % (ct, f))
        k.close(";")
        if self.phased:
            k.emit("template <int PHASE>")
        k.open("__global__ void __launch_bounds__(%d) %s(const KP_%s p)" % (B, self.name, self.name))
        for d in self.staticLds:
            k.emit(d)
        if self.dynLdsExpr is not None:
            k.emit("extern __shared__ __align__(16) unsigned char dqc_dyn_lds[];")
        k.emit("const uint32_t lane = dqc::lane_id();")
        k.emit("const uint32_t wave = threadIdx.x / 64u;")
        k.emit("(void)lane; (void)wave;")
        k.emit("dqc_plan_status* st = p.st;")
        k.emit("(void)st;")
        if self.devRows is not None:
            # rows appended by an earlier kernel of the plan, at most the capacity
            base = "std::min<int64_t>((int64_t)(%s), p.n)" % self.devRows
        else:
            base = "p.n"
        pa = self.partAgg
        if pa is not None:
            # PHASE 1: workgroup b reads partition b, rows 0..total-1 over the
            # segments the PHASE 0 workgroups filled (part_segments)
            k.emit("int64_t nrows = %s;" % base)
            k.open("if constexpr (PHASE == 1)")
            k.emit("nrows = blockIdx.x < %s ? (int64_t)dqc::part_segments(%s, blockIdx.x, %s, %s, %s, %s) : 0;"
                   % (pa["np"], pa["count"], pa["nseg"], pa["cap"], pa["pref"], pa["scan"]))
            k.close()
        elif 1 in self.phaseRows:
            k.emit("const int64_t nrows = PHASE == 1 ? (int64_t)(%s) : %s;" % (self.phaseRows[1][1], base))
        else:
            k.emit("const int64_t nrows = %s;" % base)
        if U == 1:
            k.emit("constexpr int u = 0;")
        for name, (ct, init, perRow) in self.regs.items():
            if perRow:
                k.emit("%s %s[%d];" % (ct, name, U))
                if init is not None:
                    k.emit("for (int i = 0; i < %d; ++i) %s[i] = %s;" % (U, name, init))
            else:
                k.emit("%s %s%s;" % (ct, name, "" if init is None else " = " + init))
        k.extend(self.prologue)
        if pa is not None:
            k.emit("const int64_t stride = PHASE == 1 ? (int64_t)%d : (int64_t)gridDim.x * %d;" % (B * U, B * U))
            k.open("for (int64_t base = PHASE == 1 ? 0 : (int64_t)blockIdx.x * %d;; base += stride)" % (B * U))
        else:
            k.emit("const int64_t stride = (int64_t)gridDim.x * %d;" % (B * U))
            k.open("for (int64_t base = (int64_t)blockIdx.x * %d;; base += stride)" % (B * U))
        k.emit("const bool more = base < nrows;")
        k.emit("int active[%d];" % U)
        if U > 1:
            k.emit("#pragma unroll")
            k.open("for (int u = 0; u < %d; ++u)" % U)
        k.emit("%s[u] = base + (int64_t)(u * %d) + (int64_t)threadIdx.x;" % (self.tid, B))
        k.emit("active[u] = more && %s[u] < nrows;" % self.tid)
        if U > 1:
            k.close()
        k.extend(self.code)
        k.extend(self.iterationEnd)
        k.emit("if (!more) break;")
        k.close()
        k.extend(self.epilogue)
        k.close()
        return k.text()


class Program(object):
    """Everything generated for one plan: kernels, buffers, launch steps."""

    def __init__(self, compiler):
        self.compiler = compiler
        self.decimalRepr = compiler.decimalRepr
        self.pipes = []
        self.steps = []
        self.buffers = OrderedDict()     # name -> dict(ctype, count, init)
        self.tables = OrderedDict()      # table name -> index
        self.columns = OrderedDict()     # (table, attr, part) -> dict(index, ctype)
        self.strings = OrderedDict()
        self.structDecls = []             # record types of multimap entries
        self.likes = OrderedDict()        # LIKE pattern -> matcher function name
        self.likeDecls = []
        self.hts = []
        self.estimates = OrderedDict()   # opId -> C++ expression
        self.hostVars = OrderedDict()    # extra host layout statements
        self.splitVars = OrderedDict()   # long-list queue capacities (after the zero fill of L.scap)
        self.result = None
        self.temps = OrderedDict()       # temp table name -> dict(index, columns, est)
        self.charSource = {}             # attr id -> column slot index of its char heap
        self.profiles = []
        self.planNodes = []

    # ---- inputs
    def table(self, name):
        if name not in self.tables:
            self.tables[name] = len(self.tables)
        return self.tables[name]

    def column(self, table, attr, part, ctype):
        key = (table, attr, part)
        if key not in self.columns:
            self.table(table)
            self.columns[key] = {"index": len(self.columns), "ctype": ctype}
        return self.columns[key]["index"]

    def stringConstant(self, token):
        if token not in self.strings:
            self.strings[token] = "dqc_lit%d" % len(self.strings)
        name = self.strings[token]
        return "dqc::str_t{%s, %s + %d}" % (name, name, len(token.encode("latin-1")))

    def likeMatcher(self, pattern):
        """Name of the generated device function matching LIKE `pattern`
        (literals and '%' only; see hipExpr.likeMatcherDecl)."""
        if pattern not in self.likes:
            self.likes[pattern] = "dqc_like%d" % len(self.likes)

            def lit(text):
                full = self.stringConstant(text)
                return full, self.strings[text]
            self.likeDecls.append((self.likes[pattern], pattern, lit))
        return self.likes[pattern]

    def newTemp(self, name, est):
        if name in self.temps:
            raise ValueError("temp table %s is created twice" % name)
        if len(self.temps) >= 8:
            raise ValueError("too many temp tables in one plan")
        rec = {"index": len(self.temps), "name": name, "columns": [], "est": est}
        self.temps[name] = rec
        return rec

    def tempColumns(self):
        return [c for t in self.temps.values() for c in t["columns"]]

    # ---- workspace
    def buffer(self, name, ctype, count, init=None):
        if name in self.buffers:
            raise ValueError("duplicate buffer " + name)
        self.buffers[name] = {"ctype": ctype, "count": count, "init": init}
        return name

    def newHt(self, kind, opId, label):
        rec = {"index": len(self.hts), "kind": kind, "opId": opId, "label": label}
        if rec["index"] >= 32:
            raise ValueError("too many hash tables in one plan")
        self.hts.append(rec)
        return rec

    def newPipeline(self, name, source, U=1):
        if len(self.pipes) >= 32:
            raise ValueError("too many pipelines in one plan")
        pipe = Pipeline(self, name, source, U)
        self.pipes.append(pipe)
        return pipe

    # ---- estimates (relationalAlgebra configure() formulas on runtime rows)
    def estimate(self, node):
        if node.opId in self.estimates:
            return "est_%d" % node.opId
        if isinstance(node, ra.Scan):
            if node.isTempScan:
                tmp = self.temps.get(node.table["name"])
                if tmp is None:
                    raise ValueError("temp table %s is scanned before it is created" % node.table["name"])
                i = tmp["index"]
                expr = "(t->temp_capacity[%d] > 0 ? t->temp_capacity[%d] : %s)" % (i, i, tmp["est"])
            else:
                expr = "rows_%d" % self.table(node.table["name"])
        elif isinstance(node, ra.Selection):
            expr = "(int64_t)((double)%s * %r)" % (self.estimate(node.child), float(node.selectivity))
        elif isinstance(node, (ra.Map, ra.Projection, ra.Materialize)):
            expr = self.estimate(node.child)
        elif isinstance(node, ra.EquiJoin):
            expr = "(int64_t)((double)std::max(%s, %s) * %r)" % (
                self.estimate(node.leftChild), self.estimate(node.rightChild), float(node.selectivity))
        elif isinstance(node, ra.CrossJoin):
            expr = "(int64_t)((double)%s * (double)%s * %r)" % (
                self.estimate(node.leftChild), self.estimate(node.rightChild), float(node.selectivity))
        elif isinstance(node, ra.Aggregation):
            self.estimate(node.child)
            if not node.doGroup:
                expr = "(int64_t)1"
            elif node.numgroups is not None and not _numgroupsFromChild(node):
                expr = "(int64_t)%d" % int(node.numgroups)
            else:
                expr = self.estimate(node.child)
                # keys of small domains (CHAR, BOOLEAN) bound the group count
                # whatever the input size: the table and the LDS tier are
                # sized for that many groups, not for the input rows
                dom = _keyDomain(node.groupAttributes.values())
                if dom is not None:
                    expr = "std::min<int64_t>(%s, (int64_t)%d)" % (expr, dom)
        else:
            raise NotImplementedError(type(node).__name__)
        self.estimates[node.opId] = expr
        return "est_%d" % node.opId

    # ---- rendering
    def render(self):
        out = Code()
        out.emit("// generated by dogqc_amd.hipTranslator for gfx950")
        out.emit("#include <hip/hip_runtime.h>")
        out.emit("#include <stdint.h>")
        out.emit("#include <float.h>")
        out.emit("#include <algorithm>")
        out.emit('#include "dogqc_amd.h"')
        out.emit('#include "dqc_device.h"')
        out.emit("")
        matchers = [likeMatcherDecl(name, pattern, lit) for name, pattern, lit in self.likeDecls]
        for token, name in self.strings.items():
            out.emit(stringConstantDecl(name, token))
        for decl in self.structDecls:
            out.emit(decl)
        out.emit("")
        for m in matchers:
            out.extend(_asCode(m))
            out.emit("")
        for pipe in self.pipes:
            out.extend(_asCode(pipe.render()))
            out.emit("")
        out.extend(_asCode(self.renderInitKernel()))
        out.extend(_asCode(self.renderHost()))
        return out.text()

    def renderInitKernel(self):
        k = Code()
        k.open("struct KP_dqc_init")
        nf = len(self.fills) + 1
        k.emit("uint64_t* ptr[%d];" % nf)
        k.emit("uint64_t words[%d];" % nf)
        k.emit("uint64_t pattern[%d];" % nf)
        k.close(";")
        k.open("__global__ void __launch_bounds__(256) dqc_init_kernel(const KP_dqc_init p)")
        k.emit("const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;")
        k.emit("const uint64_t gstride = (uint64_t)gridDim.x * 256u;")
        k.open("for (int f = 0; f < %d; ++f)" % (len(self.fills) + 1))
        k.emit("uint64_t* q = p.ptr[f];")
        k.emit("const uint64_t w = p.words[f], v = p.pattern[f];")
        k.emit("for (uint64_t i = gid; i < w; i += gstride) q[i] = v;")
        k.close()
        k.close()
        return k.text()

    def renderHost(self):
        h = Code()
        nt, nc = len(self.tables), len(self.columns)
        h.emit("namespace {")
        h.emit("struct Layout {")
        h.emit("    uint64_t off[%d];" % max(1, len(self.buffers)))
        h.emit("    uint64_t cnt[%d];" % max(1, len(self.buffers)))
        h.emit("    uint64_t bytes;")
        h.emit("    int64_t htcap[%d];" % max(1, len(self.hts)))
        h.emit("    int64_t mment[%d];" % max(1, len(self.hts)))
        h.emit("    int64_t rcap;")
        h.emit("    int64_t tcap[%d];" % max(1, len(self.temps)))
        h.emit("    int64_t scap[%d];" % max(1, len(self.hts)))
        h.emit("    int64_t rows[%d];" % max(1, nt))
        h.emit("    uint32_t ldsw[%d];" % max(1, len(self.pipes)))
        h.emit("    uint32_t ldsr[%d];" % max(1, len(self.pipes)))
        h.emit("    uint32_t arep[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t apart[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t apbits[%d];" % max(1, len(self.hts)))
        h.emit("    int64_t apcap[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t apseg[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t bfw[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t bspan[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t bspill[%d];" % max(1, len(self.hts)))
        h.emit("    uint32_t bfp[%d];" % max(1, len(self.pipes)))
        h.emit("    int64_t apgroups[%d];" % max(1, len(self.hts)))
        h.emit("};")
        h.emit("static inline uint64_t align256(uint64_t x) { return (x + 255u) & ~(uint64_t)255u; }")
        h.emit("static inline uint32_t pow2floor(uint64_t x) { uint32_t r = 1; while ((uint64_t)r * 2u <= x && r < (1u << 30)) r *= 2u; return r; }")
        h.emit("static inline uint32_t pow2ceil(uint64_t x) { uint32_t r = 1; while ((uint64_t)r < x && r < (1u << 30)) r *= 2u; return r; }")
        h.emit("static inline float fac(float f) { return f > 0.0f ? f : 1.0f; }")
        h.emit("static int g_lds = 0;")
        h.open("static uint64_t maxLds()")
        h.open("if (g_lds == 0)")
        h.emit("int dev = 0; (void)hipGetDevice(&dev);")
        h.emit("int b = 0;")
        h.emit("if (hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || b <= 0) b = 65536;")
        h.emit("g_lds = b;")
        h.close()
        h.emit("return (uint64_t)g_lds;")
        h.close()
        h.emit("static int g_cus = 0;")
        h.open("static int computeUnits()")
        h.open("if (g_cus == 0)")
        h.emit("int dev = 0; (void)hipGetDevice(&dev);")
        h.emit("int cus = 0;")
        h.emit("if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;")
        h.emit("g_cus = cus;")
        h.close()
        h.emit("return g_cus;")
        h.close()
        h.open("static unsigned gridFor(int32_t req, int64_t rows, int bs, int perCu)")
        h.emit("if (req > 0) return (unsigned)req;")
        h.emit("int64_t need = (rows + bs - 1) / bs;")
        h.emit("int64_t cap = (int64_t)computeUnits() * perCu;")
        h.emit("int64_t g = std::min<int64_t>(need, cap);")
        h.emit("return (unsigned)std::max<int64_t>(1, g);")
        h.close()
        h.open("static int computeLayout(const dqc_plan_inputs* in, const dqc_plan_tuning* t, Layout& L)")
        h.emit("if (!in || !t) return DQC_EINVAL;")
        h.emit("if (in->ntables != %d || in->ncols != %d) return DQC_EINVAL;" % (nt, nc))
        for name, ti in self.tables.items():
            h.emit("const int64_t rows_%d = in->table_rows[%d];  // %s" % (ti, ti, name))
            h.emit("if (rows_%d < 0) return DQC_EINVAL;" % ti)
            h.emit("L.rows[%d] = rows_%d;" % (ti, ti))
        for opId, expr in self.estimates.items():
            h.emit("const int64_t est_%d = std::max<int64_t>(0, %s);" % (opId, expr))
            h.emit("(void)est_%d;" % opId)
        h.emit("for (int i = 0; i < %d; ++i) L.bfp[i] = 0u;" % max(1, len(self.pipes)))
        h.emit("for (int i = 0; i < %d; ++i) { L.arep[i] = 1u; L.apart[i] = 0u; L.apbits[i] = 0u; L.apcap[i] = 0; "
               "L.apseg[i] = 0u; L.apgroups[i] = 0; L.bfw[i] = 0u; L.bspan[i] = 0u; L.bspill[i] = 0u; }"
               % max(1, len(self.hts)))
        for name, stmt in self.hostVars.items():
            # a callable is rendered now, when every pipeline is complete
            h.emit(stmt() if callable(stmt) else stmt)
        h.emit("for (int i = 0; i < %d; ++i) L.scap[i] = 0;" % max(1, len(self.hts)))
        for name, stmt in self.splitVars.items():
            h.emit(stmt)
        h.emit("uint64_t off = align256(sizeof(dqc_plan_status));")
        for i, (name, b) in enumerate(self.buffers.items()):
            h.emit("L.cnt[%d] = (uint64_t)(%s);  // %s" % (i, b["count"], name))
            h.emit("L.off[%d] = off; off = align256(off + L.cnt[%d] * sizeof(%s));" % (i, i, b["ctype"]))
        h.emit("L.bytes = off;")
        h.emit("return DQC_OK;")
        h.close()
        h.emit("}  // namespace")
        h.emit("")
        # describe
        h.emit('extern "C" const char* dqc_plan_describe(void) {')
        h.emit("    return R\"DQCJSON(%s)DQCJSON\";" % json.dumps(self.describe()))
        h.emit("}")
        # layout
        h.open('extern "C" int dqc_plan_layout_of(const dqc_plan_inputs* in, const dqc_plan_tuning* t, dqc_plan_layout* out)')
        h.emit("Layout L; int rc = computeLayout(in, t, L); if (rc) return rc;")
        h.emit("if (!out) return DQC_EINVAL;")
        h.emit("*out = dqc_plan_layout{};")
        h.emit("out->ws_bytes = L.bytes;")
        h.emit("out->nsteps = %d;" % len(self.steps))
        h.emit("out->nhts = %d;" % len(self.hts))
        h.emit("for (int i = 0; i < %d; ++i) { out->ht_capacity[i] = L.htcap[i]; out->mm_entries[i] = L.mment[i]; "
               "out->split_capacity[i] = L.scap[i]; }" % len(self.hts))
        h.emit("out->result_capacity = L.rcap;")
        res = self.result["columns"] if self.result else []
        h.emit("out->nresult = %d;" % len(res))
        for i, c in enumerate(res):
            h.emit("out->result_offset[%d] = L.off[%d];" % (i, list(self.buffers).index(c["buffer"])))
        h.emit("out->ntemps = %d;" % len(self.temps))
        h.emit("out->ntempcols = %d;" % len(self.tempColumns()))
        for t in self.temps.values():
            h.emit("out->temp_capacity[%d] = L.tcap[%d];" % (t["index"], t["index"]))
        for i, c in enumerate(self.tempColumns()):
            h.emit("out->temp_offset[%d] = L.off[%d];" % (i, list(self.buffers).index(c["buffer"])))
        h.emit("out->nprofile = %d;" % len(self.profiles))
        if self.profiles:
            h.emit("out->profile_offset = L.off[%d];" % list(self.buffers).index("profile"))
        h.emit("return DQC_OK;")
        h.close()
        # enqueue
        h.open('extern "C" int dqc_plan_enqueue(const dqc_plan_inputs* in, const dqc_plan_tuning* t, void* ws, uint64_t ws_bytes, void* stream, void* const* ev)')
        h.emit("Layout L; int rc = computeLayout(in, t, L); if (rc) return rc;")
        h.emit("if (!ws || ws_bytes < L.bytes) return DQC_ENOMEM;")
        h.emit("if (((uintptr_t)ws & 15u) != 0u) return DQC_EINVAL;   // 16-byte loads and stores of table regions")
        h.emit("hipStream_t s = (hipStream_t)stream;")
        h.emit("unsigned char* W = (unsigned char*)ws;")
        h.emit("dqc_plan_status* st = (dqc_plan_status*)W;")
        h.emit("(void)st;")
        for i, name in enumerate(self.buffers):
            h.emit("%s* B_%s = (%s*)(W + L.off[%d]);" % (self.buffers[name]["ctype"], name,
                                                          self.buffers[name]["ctype"], i))
            h.emit("(void)B_%s;" % name)
        for (table, attr, part), c in self.columns.items():
            h.emit("const %s* C_%d = (const %s*)in->cols[%d];  // %s.%s%s" % (
                c["ctype"], c["index"], c["ctype"], c["index"], table, attr, "" if part is None else "." + part))
            h.emit("(void)C_%d;" % c["index"])
        h.emit("(void)hipGetLastError();")
        for si, step in enumerate(self.steps):
            h.emit("if (ev) (void)hipEventRecord((hipEvent_t)ev[%d], s);" % si)
            h.open("")
            for line in step["host"](self):
                h.emit(line)
            h.close()
        h.emit("if (ev) (void)hipEventRecord((hipEvent_t)ev[%d], s);" % len(self.steps))
        census = []
        for ht in self.hts:
            for prefix in ("jht", "aht"):
                name = "%s%d_tags" % (prefix, ht["opId"])
                if name in self.buffers:
                    census.append((ht["index"], name, self.buffers[name]["ctype"], list(self.buffers).index(name)))
        if census:
            h.open("if (t->census)")
            for i, name, ct, bi in census:
                h.emit("hipLaunchKernelGGL(dqc::dqc_census<%s>, dim3(gridFor(0, (int64_t)L.cnt[%d], 256, 4)), "
                       "dim3(256), 0, s, (const %s*)B_%s, (uint64_t)L.cnt[%d], "
                       "(unsigned long long*)&st->ht_fill[%d]);" % (ct, bi, ct, name, bi, i))
            h.close()
        h.emit("return hipGetLastError() == hipSuccess ? DQC_OK : DQC_EHIP;")
        h.close()
        return h.text()

    def describe(self):
        cols = []
        for (table, attr, part), c in self.columns.items():
            cols.append({"table": table, "attr": attr, "part": part, "ctype": c["ctype"], "index": c["index"]})
        return {
            "tables": list(self.tables),
            "columns": cols,
            "result": self.result["columns"] if self.result else [],
            "temps": [{"name": t["name"], "index": t["index"], "columns": t["columns"]} for t in self.temps.values()],
            "steps": [s["name"] for s in self.steps],
            "pipelines": [{"name": p.name, "id": p.pipeId, "block": p.bs, "annotations": p.annotations}
                          for p in self.pipes],
            "hts": [{"index": h["index"], "kind": h["kind"], "opId": h["opId"], "label": h["label"]}
                    for h in self.hts],
            "profiles": self.profiles,
        }

    # ---- steps
    def addInitStep(self):
        self.fills = [(n, b) for n, b in self.buffers.items() if b["init"] is not None]

        def host(prog):
            lines = ["KP_dqc_init kp{};",
                     "kp.ptr[0] = (uint64_t*)st; kp.words[0] = sizeof(dqc_plan_status) / 8; kp.pattern[0] = 0;"]
            for i, (name, b) in enumerate(prog.fills):
                idx = list(prog.buffers).index(name)
                lines.append("kp.ptr[%d] = (uint64_t*)B_%s;" % (i + 1, name))
                lines.append("kp.words[%d] = (L.cnt[%d] * sizeof(%s) + 7) / 8;" % (i + 1, idx, b["ctype"]))
                lines.append("kp.pattern[%d] = %s;" % (i + 1, b["init"]))
            lines.append("hipLaunchKernelGGL(dqc_init_kernel, dim3(gridFor(0, %s, 256, 4)), dim3(256), 0, s, kp);"
                         % prog._maxFillWords())
            return lines

        self.steps.insert(0, {"name": "dqc_init_kernel", "host": host})

    def _maxFillWords(self):
        terms = ["(int64_t)(sizeof(dqc_plan_status) / 8)"]
        for name, b in self.fills:
            idx = list(self.buffers).index(name)
            terms.append("(int64_t)((L.cnt[%d] * sizeof(%s) + 7) / 8)" % (idx, b["ctype"]))
        expr = terms[0]
        for t in terms[1:]:
            expr = "std::max<int64_t>(%s, %s)" % (expr, t)
        return expr


def _asCode(text):
    c = Code()
    for ln in text.rstrip("\n").split("\n"):
        c.lines.append(ln)
    return c


def _numgroupsFromChild(node):
    return not getattr(node, "numgroupsFromCfg", False)


# distinct values a group key of this type can take, for the types whose
# domain is small enough to bound a table
_KEY_DOMAIN = {Type.CHAR: 256, Type.BOOLEAN: 2}


def _keyDomain(attrs):
    """Number of distinct key tuples the group attributes can form, or None
    when one of them has no small domain or the product exceeds 2^16."""
    attrs = list(attrs)
    if not attrs:
        return None
    d = 1
    for a in attrs:
        n = _KEY_DOMAIN.get(getattr(a, "dataType", None))
        if n is None:
            return None
        d *= n
    return d if d <= (1 << 16) else None


def pipelineLaunch(prog, pipe, phase=None, rowsExpr=None, perCu=8):
    """Host statements launching a pipeline kernel."""
    lines = ["KP_%s kp{};" % pipe.name]
    for f, (ct, expr) in pipe.params.items():
        lines.append("kp.%s = (%s)(%s);" % (f, ct, expr))
    rows = rowsExpr or pipe.params["n"][1]
    per = str(perCu)
    if "lw" in pipe.params:
        # an LDS group table is flushed once per workgroup: with many slots,
        # fewer workgroups (8 per CU at <= 512 slots, 1 per CU at >= 4096)
        per = "std::max<int>(1, (int)(%d * 512u / std::max<uint32_t>(512u, L.ldsw[%d])))" % (perCu, pipe.pipeId)
    lds = "0" if pipe.dynLdsExpr is None else "(size_t)(%s)" % pipe.dynLdsExpr
    pa = pipe.partAgg
    if pa is not None and phase == 0:
        # rows are scattered to partitions: no LDS group table in this pass
        lines.append("const bool part = %s > 0u;" % pa["hostNp"])
        if "lw" in pipe.params:
            lines.append("if (part) kp.lw = 0u;")
        keep = "(size_t)0" if pipe.ldsFilter is None else "(size_t)(%s)" % pipe.ldsFilter["host"]
        lds = "(part ? %s : %s)" % (keep, lds)
    if phase == 1 and getattr(pipe, "globalRefill", False):
        # the survivors' launch sizes its grid from the queue; the tuned grid
        # belongs to the first launch (with an LDS join filter, a few large
        # workgroups there)
        lines.append("const unsigned g = gridFor(0, %s, %d, %s);" % (rows, pipe.bs * pipe.U, per))
    else:
        lines.append("const unsigned g = gridFor(t->grid[%d] > 0 ? t->grid[%d] : %d, %s, %d, %s);"
                     % (pipe.pipeId, pipe.pipeId, pipe.defaultGrid, rows, pipe.bs * pipe.U, per))
    if pa is not None and phase == 0:
        # each workgroup fills its own segment of every partition
        lines[-1] = "const unsigned g = part ? (unsigned)%s : %s;" % (pa["hostNseg"], lines[-1][len("const unsigned g = "):-1])
    if pa is not None and phase == 1:
        # one workgroup per partition; nothing to do without partitions
        lines[-1] = "const unsigned g = (unsigned)%s;" % pa["hostNp"]
        lines.append("if (g > 0u) {")
    if phase == 1 and getattr(pipe, "globalRefill", False) and pipe.ldsFilter is not None:
        # the second launch of a kernel-split lane refill runs what follows
        # the filtered probe: no filter copied into LDS, none allocated
        lines.append("kp.%s = 0u;" % pipe.ldsFilter["param"])
        lds = lds.replace(pipe.ldsFilter["host"], "(uint64_t)0")
    kname = pipe.name if phase is None else "%s<%d>" % (pipe.name, phase)
    if pipe.dynLdsExpr is not None:
        # dynamic LDS beyond 64 KB (gfx950 has 160 KB per workgroup) is opted into per kernel
        lines.append("if (%s > 65536u) (void)hipFuncSetAttribute((const void*)&%s, "
                     "hipFuncAttributeMaxDynamicSharedMemorySize, (int)(%s));" % (lds, kname, lds))
    lines.append("hipLaunchKernelGGL((%s), dim3(g), dim3(%d), %s, s, kp);" % (kname, pipe.bs, lds))
    if pa is not None and phase == 1:
        lines.append("}")
    return lines


class HipCompiler(object):

    def __init__(self, algebraContext, arch="gfx950", decimalRepr=CType.FP32, debug=False, smArchitecture=None):
        if isinstance(arch, str) and arch.startswith("sm_"):
            arch = "gfx950"   # reference drivers pass a CUDA sm_XX target positionally
        if arch != "gfx950":
            raise ValueError("only gfx950 (MI355X) is supported, got " + str(arch))
        if decimalRepr not in (CType.FP32, CType.FP64):
            raise ValueError("decimalRepr must be CType.FP32 or CType.FP64")
        self.arch = arch
        self.decimalRepr = decimalRepr
        self.debug = debug
        self.algebraContext = algebraContext
        self.bufferPositions = []
        self.profilerPositions = []
        self.innerLoopCount = 0
        self.program = None
        self.pipe = None
        self.source = None
        self.module = None
        self.planName = None

    def setBuffers(self, positions):
        self.bufferPositions = list(positions)

    def setProfilers(self, positions):
        self.profilerPositions = list(positions)

    def translate(self, algebraExpr, child1=None, child2=None):
        from dogqc_amd.hashJoins import EquiJoinTranslator
        from dogqc_amd.hipDivergenceBuffer import LaneRefill
        from dogqc_amd.hipLaneActivityProfile import LaneActivityProfiler
        if isinstance(algebraExpr, ra.Scan):
            op = ScanTranslator(algebraExpr)
        elif isinstance(algebraExpr, ra.Selection):
            op = SelectionTranslator(algebraExpr, child1)
        elif isinstance(algebraExpr, ra.Map):
            op = MapTranslator(algebraExpr, child1)
        elif isinstance(algebraExpr, ra.EquiJoin):
            op = EquiJoinTranslator(algebraExpr, child1, child2)
        elif isinstance(algebraExpr, ra.CrossJoin):
            op = CrossJoinTranslator(algebraExpr, child1, child2)
        elif isinstance(algebraExpr, ra.Aggregation):
            op = AggregationTranslator(algebraExpr, child1)
        elif isinstance(algebraExpr, ra.Projection):
            op = ProjectionTranslator(algebraExpr, child1)
        elif isinstance(algebraExpr, ra.Materialize):
            op = MaterializeTranslator(algebraExpr, child1)
        else:
            raise NotImplementedError(type(algebraExpr).__name__)
        if algebraExpr.opId in self.bufferPositions:
            op = LaneRefill(algebraExpr, op)
        if algebraExpr.opId in self.profilerPositions:
            op = LaneActivityProfiler(algebraExpr, op)
        return op

    def gencode(self, translationPlan):
        self.program = Program(self)
        self.planRoots = listWrap(translationPlan)
        for node in self.planRoots:
            self.program.planNodes.append(node.algExpr)
            node.produce(self)
        self.program.addInitStep()
        self.source = self.program.render()
        return self.source

    def compile(self, filename):
        from dogqc_amd import plancache
        if self.source is None:
            raise RuntimeError("gencode() must run before compile()")
        self.module = plancache.compileSource(self.source, filename, debug=self.debug)
        if not getattr(self, "planName", None):
            self.planName = filename      # execute() logs to <filename>.log
        return self.module

    def execute(self):
        from dogqc_amd import runner
        if self.module is None:
            raise RuntimeError("compile() must run before execute()")
        out, self.result = runner.executeAndPrint(self)
        return out

    def showGraph(self, plan):
        lines = ["digraph {"]
        for node in listWrap(plan):
            node.algExpr.toDOT(lines)
        lines.append("}")
        with open("query.dot", "w") as f:
            f.write("\n".join(lines) + "\n")
        return "\n".join(lines)

    # helpers for translators
    def exprEmitter(self):
        return ExprEmitter(self.program, self.decimalRepr, self.pipe.access, self.pipe)


# ------------------------------------------------------------------ operators

def pipelineShape(translator):
    """Walk up from a pipeline's source translator to its sink and report
    (hasInnerLoop, passFraction): whether an operator of the pipeline opens a
    per-lane loop (multimatch probe, lane refill) and the product of the
    configured selectivities of the selections before the first probe and of
    that probe when it is a unique (single-match) join."""
    from dogqc_amd.hashJoins import EquiJoinTranslator
    from dogqc_amd.hipDivergenceBuffer import LaneRefill
    from dogqc_amd.relationalAlgebra import Join
    loops, frac, beforeJoin = False, 1.0, True
    t = translator
    while t.parent is not None:
        p = t.parent
        if isinstance(p, LaneRefill):
            loops = True
        elif isinstance(p, CrossJoinTranslator):
            if p.leftChild is t:
                break
            loops = True
            beforeJoin = False
        elif isinstance(p, EquiJoinTranslator):
            if p.leftChild is t:
                break
            if beforeJoin and not p.algExpr.multimatch and p.algExpr.joinType in (Join.INNER, Join.SEMI):
                frac *= min(1.0, float(getattr(p.algExpr, "selectivity", 1.0)))
            beforeJoin = False
            if p.algExpr.multimatch and p.algExpr.joinType in (Join.INNER, Join.OUTER):
                loops = True
        elif isinstance(p, SelectionTranslator) and beforeJoin:
            frac *= float(getattr(p.algExpr, "selectivity", 1.0))
        elif isinstance(p, (AggregationTranslator, MaterializeTranslator)):
            break
        t = p
    return loops, frac


_SHARED_DECL = None


def staticLdsBytes(pipe):
    """Bytes of a pipeline's static __shared__ arrays (its declarations)."""
    import re
    global _SHARED_DECL
    if _SHARED_DECL is None:
        _SHARED_DECL = re.compile(r"__shared__\s+(.+?)\s+\w+((?:\[\d+\])*)\s*;")
    sizes = {"uint8_t": 1, "int8_t": 1, "char": 1, "int16_t": 2, "uint16_t": 2, "int32_t": 4, "uint32_t": 4,
             "float": 4, "int64_t": 8, "uint64_t": 8, "unsigned long long": 8, "double": 8, "dqc::str_t": 16,
             "dqc::str_offs": 16}
    total = 0
    for d in pipe.staticLds:
        m = _SHARED_DECL.search(d)
        if not m:
            total += 1024
            continue
        n = 1
        for dim in re.findall(r"\[(\d+)\]", m.group(2)):
            n *= int(dim)
        total += n * sizes.get(m.group(1).strip(), 8)
    return total


def firstUseFractions(scan):
    """For each attribute a scan opens, the configured fraction of the
    scanned rows still active at the operator that first reads it: the
    product of the selectivities of the selections and unique-join probes
    before that operator.  Attributes first read after a per-lane loop
    (multimatch probe, lane refill, nested loop) are left out: they are
    read per match, not per row."""
    return useFractions(scan, set(scan.algExpr.outRelation), list(scan.algExpr.outRelation.values()))


def useFractions(start, mine, terminal):
    """First-use fractions of the attribute ids `mine`, walking the plan up
    from translator `start` (the rows there count as 1); a sink operator
    (aggregation, materialisation, projection) reads the attributes
    `terminal`."""
    from dogqc_amd.hashJoins import EquiJoinTranslator
    from dogqc_amd.hipDivergenceBuffer import LaneRefill
    from dogqc_amd.relationalAlgebra import Join
    out = {}

    def use(attrs, frac):
        for a in attrs:
            if a.id in mine and a.id not in out:
                out[a.id] = frac

    def exprAttrs(e):
        found = []
        if e is not None:
            _collectAttrs(e, found)
        return found

    frac = 1.0
    t = start
    while t.parent is not None:
        p = t.parent
        a = p.algExpr
        if isinstance(p, (LaneRefill, CrossJoinTranslator)):
            # rows reach what follows through a refill queue (or per inner
            # row): left out.  For a refill across launches, loading such a
            # column at the top of the first launch made that launch stream
            # every column and left no LDS for a join filter (c3: 0.51 ms for
            # the first launch alone, against 0.51 ms for the whole kernel
            # with the LDS queue).
            break
        if isinstance(p, SelectionTranslator):
            use(exprAttrs(a.condition), frac)
            frac *= float(getattr(a, "selectivity", 1.0))
        elif isinstance(p, EquiJoinTranslator):
            if p.leftChild is t:          # build side: everything is stored here
                use(terminal, frac)
                break
            use(list(a.probeKeyAttributes.values()), frac)
            if a.multimatch and a.joinType in (Join.INNER, Join.OUTER):
                break
            use(exprAttrs(a.conditions), frac)
            if a.joinType in (Join.INNER, Join.SEMI):
                frac *= min(1.0, float(getattr(a, "selectivity", 1.0)))
        elif isinstance(p, MapTranslator):
            use(exprAttrs(a.expression), frac)
        elif isinstance(p, (AggregationTranslator, MaterializeTranslator, ProjectionTranslator)):
            use(terminal, frac)
            break
        t = p
    return out


def lineTouchFraction(rowFraction, elemBytes, lineBytes=128):
    """Expected fraction of a column's cache lines read when a random
    `rowFraction` of its rows is read."""
    per = max(1, lineBytes // max(1, elemBytes))
    return 1.0 - (1.0 - min(1.0, max(0.0, rowFraction))) ** per


def rowsPerThread(translator):
    loops, _ = pipelineShape(translator)
    return 1 if loops else max(1, int(KernelCall.defaultRowsPerThread))


class ScanTranslator(LiteralTranslator):
    """Opens a pipeline over a base table.  Columns load lazily at first use,
    except those whose cache lines would nearly all be read anyway: a column
    first read where a fraction f of the rows is still active touches
    1 - (1 - f)^(128 / element bytes) of its 128-byte lines, and when that is
    at least EAGER_LINE_FRACTION the column loads at the top, with the
    first, so its load is not one more dependent memory round trip."""

    EAGER_LINE_FRACTION = 0.8
    # base-table column loads with the nontemporal hint (streamed through
    # the caches with low priority, so hash tables probed beside the scan
    # keep their lines)
    nontemporal = False

    def produce(self, ctxt):
        a = self.algExpr
        prog = ctxt.program
        tname = a.table["name"]
        name = "krnl_%s%d" % (ident(tname), a.opId)
        if a.isTempScan:
            self._produceTemp(ctxt, name)
            return
        ti = prog.table(tname)
        pipe = prog.newPipeline(name, {"rows": "L.rows[%d]" % ti, "tidName": "%s%d" % (tname, a.scanTableId)},
                                rowsPerThread(self))
        pipe.scanTable = tname
        ctxt.pipe = pipe
        tid = pipe.tid + "[u]"
        for id, attr in a.outRelation.items():
            if attr.dataType == Type.STRING:
                oi = prog.column(tname, attr.name, "offset", CType.UINT64)
                ci = prog.column(tname, attr.name, "char", "char")
                fo = pipe.param("c%d" % oi, "const uint64_t*", "C_%d" % oi)
                fc = pipe.param("c%d" % ci, "const char*", "C_%d" % ci)
                pipe.setLazy(attr, "dqc::string_scan(%s, %s, %s)" % (fo, fc, tid), [pipe.tid])
                prog.charSource[attr.id] = ci
            else:
                ct = columnType(attr.dataType)
                ci = prog.column(tname, attr.name, None, ct)
                f = pipe.param("c%d" % ci, "const %s*" % ct, "C_%d" % ci)
                lt = langType(attr.dataType, prog.decimalRepr)
                ref = "%s[%s]" % (f, tid)
                if ScanTranslator.nontemporal:
                    ref = "__builtin_nontemporal_load(&%s)" % ref
                pipe.setLazy(attr, "(%s)%s" % (lt, ref), [pipe.tid])
        pipe.param("st", "dqc_plan_status*", "st")
        self._finish(ctxt, pipe)

    def _finish(self, ctxt, pipe):
        fracs = firstUseFractions(self)
        eager = []
        for id, attr in self.algExpr.outRelation.items():
            if id not in fracs:
                continue
            size = 8 if attr.dataType == Type.STRING else COLUMN_BYTES.get(attr.dataType, 8)
            if lineTouchFraction(fracs[id], size) >= ScanTranslator.EAGER_LINE_FRACTION:
                eager.append(attr)
        if eager:
            pipe.need(eager)
        self.parent.consume(ctxt)
        closePipeline(ctxt, pipe)

    def _produceTemp(self, ctxt, name):
        """Scan of a temp table written by an earlier pipeline of the plan
        (the reference's tempscan over a materialised table,
        cudaTranslator.py:168-182): the columns are workspace buffers and the
        row count is read on the device from the temp table's append
        counter, so the launch sequence needs no host round trip."""
        a = self.algExpr
        prog = ctxt.program
        tmp = prog.temps.get(a.table["name"])
        if tmp is None:
            raise ValueError("temp table %s is scanned before it is created" % a.table["name"])
        ti = tmp["index"]
        pipe = prog.newPipeline(name, {"rows": "L.tcap[%d]" % ti, "devRows": "st->temp_count[%d]" % ti,
                                       "tidName": "%s%d" % (a.table["name"], a.scanTableId)}, rowsPerThread(self))
        ctxt.pipe = pipe
        tid = pipe.tid + "[u]"
        byName = {c["name"]: c for c in tmp["columns"]}
        for id, attr in a.outRelation.items():
            c = byName[attr.name]
            f = pipe.param("t%d_%d" % (ti, c["pos"]), "const %s*" % c["ctype"], "B_%s" % c["buffer"])
            pipe.setLazy(attr, "%s[%s]" % (f, tid), [pipe.tid], ctype=c["ctype"])
            if c["char_col"] is not None:
                prog.charSource[attr.id] = c["char_col"]
        pipe.param("st", "dqc_plan_status*", "st")
        self._finish(ctxt, pipe)


def closePipeline(ctxt, pipe):
    """Finish the kernel: resolve refill flush conditions and add the launch
    step(s).  Multimap builds launch twice (count, then insert) around the
    bucket scan."""
    from dogqc_amd.hipDivergenceBuffer import finishRefills
    finishRefills(pipe)
    prog = ctxt.program
    if pipe.phased:
        between = pipe.phaseBetween if pipe.phaseBetween is not None else mmScanSteps(pipe.phasedHt)
        prog.steps.append({"name": pipe.name + "<0>", "host": lambda p, pipe=pipe: pipelineLaunch(p, pipe, 0)})
        for s in between:
            prog.steps.append(s)
        rows1 = pipe.phaseRows[1][0] if 1 in pipe.phaseRows else None
        prog.steps.append({"name": pipe.name + "<1>",
                           "host": lambda p, pipe=pipe, r=rows1: pipelineLaunch(p, pipe, 1, rowsExpr=r)})
        for s in getattr(pipe, "phaseAfter", None) or []:
            prog.steps.append(s)
    else:
        prog.steps.append({"name": pipe.name, "host": lambda p, pipe=pipe: pipelineLaunch(p, pipe)})
    ctxt.pipe = None


def mmScanSteps(ht):
    """Bucket-count exclusive scan of a multimap (three kernels)."""
    dirb, sums, idx = ht["dirBuffer"], ht["sumsBuffer"], ht["index"]

    def reduce_(prog):
        return ["const uint64_t nb = (uint64_t)L.htcap[%d];" % idx,
                "const unsigned g = (unsigned)std::max<uint64_t>(1, (nb + dqc::SCAN_TILE - 1) / dqc::SCAN_TILE);",
                "hipLaunchKernelGGL(dqc::dqc_mm_scan_reduce, dim3(g), dim3(dqc::SCAN_BS), 0, s, "
                "(const uint32_t*)B_%s, nb, B_%s);" % (dirb, sums)]

    def blocks(prog):
        return ["const uint64_t nb = (uint64_t)L.htcap[%d];" % idx,
                "const uint64_t g = std::max<uint64_t>(1, (nb + dqc::SCAN_TILE - 1) / dqc::SCAN_TILE);",
                "hipLaunchKernelGGL(dqc::dqc_mm_scan_blocks, dim3(1), dim3(dqc::SCAN_BS), 0, s, B_%s, g, "
                "&st->mm_total[%d]);" % (sums, idx)]

    def apply_(prog):
        return ["const uint64_t nb = (uint64_t)L.htcap[%d];" % idx,
                "const unsigned g = (unsigned)std::max<uint64_t>(1, (nb + dqc::SCAN_TILE - 1) / dqc::SCAN_TILE);",
                "hipLaunchKernelGGL(dqc::dqc_mm_scan_apply, dim3(g), dim3(dqc::SCAN_BS), 0, s, "
                "(uint32_t*)B_%s, nb, (const uint64_t*)B_%s, %du);" % (dirb, sums, 1 if ht.get("ranked") else 0)]

    return [{"name": "dqc_mm_scan_reduce", "host": reduce_}, {"name": "dqc_mm_scan_blocks", "host": blocks},
            {"name": "dqc_mm_scan_apply", "host": apply_}]


class SelectionTranslator(UnaryTranslator):

    def produce(self, ctxt):
        self.child.produce(ctxt)

    def consume(self, ctxt):
        pipe = ctxt.pipe
        pipe.emit("// selection (opId %d)" % self.algExpr.opId)
        pipe.need(self.algExpr.conditionAttributes.values())
        cond = ctxt.exprEmitter().emit(self.algExpr.condition)
        with pipe.rowLoop():
            pipe.emit("if (active[u]) active[u] = (%s) ? 1 : 0;" % cond)
        self.parent.consume(ctxt)


class MapTranslator(UnaryTranslator):

    def produce(self, ctxt):
        self.child.produce(ctxt)

    def consume(self, ctxt):
        a = self.algExpr
        pipe = ctxt.pipe
        prog = ctxt.program
        pipe.emit("// map (opId %d)" % a.opId)
        import dogqc_amd.scalarAlgebra as scal
        if isinstance(a.expression, scal.SegmentExpr):
            # a batch segment depends on the row index only: computed where it
            # is first read, so rows filtered out before never search for it
            emitter = ctxt.exprEmitter()
            emitter.emit(a.expression)          # checks the scan and binds the prefix column
            pipe.setLazy(a.mapAttr, lambda e=a.expression: emitter.emit(e), [pipe.tid])
            self.parent.consume(ctxt)
            return
        used = []
        _collectAttrs(a.expression, used)
        pipe.need(used)
        if len(a.mapStringAttributes) == 1 and a.mapStringAttributes[0].id in prog.charSource:
            prog.charSource[a.mapAttr.id] = prog.charSource[a.mapStringAttributes[0].id]
        st = pipe.declareAttr(a.mapAttr)
        expr = ctxt.exprEmitter().emit(a.expression)
        with pipe.rowLoop():
            pipe.emit("if (active[u]) %s[u] = (%s)(%s);" % (st.reg, st.ctype, expr))
        st.loaded = True
        self.parent.consume(ctxt)


def _collectAttrs(expr, out):
    import dogqc_amd.scalarAlgebra as scal
    if isinstance(expr, scal.AttrExpr):
        out.append(expr.attr)
    for c in expr.children():
        _collectAttrs(c, out)


class ProjectionTranslator(UnaryTranslator):

    def produce(self, ctxt):
        self.child.produce(ctxt)

    def consume(self, ctxt):
        self.parent.consume(ctxt)


class MaterializeTranslator(UnaryTranslator):
    """Result materialisation (cudaTranslator.py:538-579, 599-610).

    Every wavefront appends its rows to its own staging area in LDS
    (positions from the ballot and mbcnt, the count kept wave-uniform in a
    register); when the next ballot would not fit, and once at the end of
    the kernel, the wavefront reserves one range of the output with a single
    global atomic and copies the stage out with coalesced stores.  No
    workgroup barrier is involved, so waves never wait for each other and
    the append works inside per-lane match loops.  The reference spends one
    global atomic per warp ballot on a single counter, which serialises in
    L2.  The count keeps growing past the capacity so the host can retry
    with the exact size."""

    STAGE_BYTES_PER_WAVE = 4096

    def produce(self, ctxt):
        self.child.produce(ctxt)

    def consume(self, ctxt):
        a = self.algExpr
        isResult = a.matType == MaterializationType.RESULT
        materializeRows(ctxt, list(a.outRelation.values()), isResult, None if isResult else a.table["name"],
                        ctxt.program.estimate(a), a.opId)


def materializeRows(ctxt, attrs, isResult, tableName, est, opId):
    """Append the active rows' `attrs` to the result (isResult) or to the
    temp table `tableName` (see MaterializeTranslator); returns the temp
    table's record."""
    names = [att.name for att in attrs]
    dup = sorted(set(n for n in names if names.count(n) > 1))
    if dup:
        raise ValueError("columns %s of %s %s have the same name; rename them with a map"
                         % (dup, "the result" if isResult else "temp table", tableName or ""))
    pipe = ctxt.pipe
    prog = ctxt.program
    pipe.need(attrs)
    tmp = None
    a = None
    if isResult:
        pipe.emit("// materialize result (opId %d)" % opId)
        prog.hostVars["rcap"] = ("L.rcap = t->result_capacity > 0 ? t->result_capacity : "
                                 "std::max<int64_t>(%s, 1);" % est)
        cap = pipe.param("rcap", "uint64_t", "L.rcap")
        counter = "st->result_count"
        bprefix, fprefix = "res", "r"
    else:
        # temp table: same append path into workspace columns; string
        # attributes keep their (start, end) pointers into the resident
        # char heap
        tmp = prog.newTemp(tableName, est)
        ti = tmp["index"]
        pipe.emit("// materialize temp table %s (opId %d)" % (tableName, opId))
        prog.hostVars["tcap%d" % ti] = ("L.tcap[%d] = std::max<int64_t>(1, t->temp_capacity[%d] > 0 ? "
                                       "t->temp_capacity[%d] : %s);" % (ti, ti, ti, est))
        cap = pipe.param("tcap%d" % ti, "uint64_t", "L.tcap[%d]" % ti)
        counter = "st->temp_count[%d]" % ti
        bprefix, fprefix = "tmp%d_" % ti, "w%d_" % ti
    cols = []
    rowBytes = 0
    for i, att in enumerate(attrs):
        st = pipe.attrs[att.id]
        ct = CType.STR_OFFS if (isResult and att.dataType == Type.STRING) else st.ctype
        rowBytes += CType.size.get(ct, 8) if ct != CType.STR_TYPE else 16
        bname = "%s%d" % (bprefix, i)
        prog.buffer(bname, ct, "L.rcap" if isResult else "L.tcap[%d]" % ti)
        f = pipe.param("%s%d" % (fprefix, i), "%s*" % ct, "B_%s" % bname)
        val = pipe.access(att)
        null = pipe.nullFlags.get(att.id)
        if ct == CType.STR_OFFS:
            ci = prog.charSource[att.id]
            fc = pipe.param("c%d" % ci, "const char*", "C_%d" % ci)
            if null:    # a null string is the empty string
                val = "(%s ? dqc::str_t{%s, %s} : %s)" % (null, fc, fc, val)
            val = "dqc::to_string_offset(%s, %s)" % (fc, val)
        elif null and ct != CType.STR_TYPE:
            val = "(%s ? (%s)0 : %s)" % (null, ct, val)
        cols.append({"name": att.name, "type": att.dataType.name, "ctype": ct, "buffer": bname,
                     "char_col": prog.charSource.get(att.id), "field": f, "val": val})
    public = [{k: v for k, v in c.items() if k not in ("field", "val")} for c in cols]
    if isResult:
        prog.result = {"columns": public}
    else:
        base = len(prog.tempColumns())
        for i, c in enumerate(public):
            c["pos"] = i
            c["slot"] = base + i
        tmp["columns"] = public
    NW = pipe.bs // WAVE
    SW = max(64, min(1024, (MaterializeTranslator.STAGE_BYTES_PER_WAVE // max(1, rowBytes)) // 64 * 64))
    for i, c in enumerate(cols):
        pipe.staticLds.append("__shared__ %s stg%d[%d][%d];" % (c["ctype"], i, NW, SW))
    wcnt = pipe.reg("stg_n", "uint32_t", "0u", perRow=False)

    def flush(code):
        # one global atomic per wavefront flush, coalesced copy-out by the
        # lanes currently executing (all 64 unless inside divergent code)
        code.emit("dqc::wave_lds_sync();")
        code.emit("const uint64_t fx = dqc::ballot(true);")
        code.emit("const int fl = dqc::first_lane(fx);")
        code.emit("unsigned long long wb = 0;")
        code.emit("if ((int)lane == fl) wb = atomicAdd((unsigned long long*)&%s, (unsigned long long)%s);"
                  % (counter, wcnt))
        code.emit("wb = dqc::uniform64(dqc::shfl(wb, fl));")
        code.open("for (uint32_t i = dqc::lanes_below(fx); i < %s; i += dqc::popc64(fx))" % wcnt)
        code.emit("const uint64_t wp = wb + i;")
        code.open("if (wp < %s)" % cap)
        for i, c in enumerate(cols):
            code.emit("%s[wp] = stg%d[wave][i];" % (c["field"], i))
        code.close()
        code.close()
        code.emit("dqc::wave_lds_sync();")
        code.emit("%s = 0u;" % wcnt)

    with pipe.rowLoop():
        pipe.emit("const uint64_t wm = dqc::ballot(active[u]);")
        pipe.emit("const uint32_t wn = dqc::popc64(wm);")
        pipe.code.open("if (%s + wn > %du)" % (wcnt, SW))
        flush(pipe.code)
        pipe.code.close()
        pipe.code.open("if (active[u])")
        pipe.emit("const uint32_t sp = %s + dqc::lanes_below(wm);" % wcnt)
        for i, c in enumerate(cols):
            pipe.emit("stg%d[wave][sp] = %s;" % (i, c["val"]))
        pipe.code.close()
        pipe.emit("%s += wn;" % wcnt)
    # final flush, once per workgroup: one global atomic reserves the rows of
    # all its wavefronts' stages (a per-wavefront atomic here serialises
    # thousands of wavefronts on one counter at the end of short kernels)
    pipe.staticLds.append("__shared__ uint32_t stg_blk[%d];" % NW)
    pipe.staticLds.append("__shared__ unsigned long long stg_base;")
    e = pipe.epilogue
    e.emit("if (lane == 0u) stg_blk[wave] = %s;" % wcnt)
    e.emit("__syncthreads();")
    e.open("if (threadIdx.x == 0u)")
    e.emit("uint32_t tot = 0u;")
    e.emit("for (uint32_t w = 0; w < %du; ++w) tot += stg_blk[w];" % NW)
    e.emit("stg_base = tot ? atomicAdd((unsigned long long*)&%s, (unsigned long long)tot) : 0ull;" % counter)
    e.close()
    e.emit("__syncthreads();")
    e.open("if (%s > 0u)" % wcnt)
    e.emit("uint64_t wb = stg_base;")
    e.emit("for (uint32_t w = 0; w < wave; ++w) wb += stg_blk[w];")
    e.open("for (uint32_t i = lane; i < %s; i += 64u)" % wcnt)
    e.emit("const uint64_t wp = wb + i;")
    e.open("if (wp < %s)" % cap)
    for i, c in enumerate(cols):
        e.emit("%s[wp] = stg%d[wave][i];" % (c["field"], i))
    e.close()
    e.close()
    e.close()
    return tmp


class CrossJoinTranslator(BinaryTranslator):
    """Nested-loop join (cudaTranslator.py:369-397, relationalAlgebra.py
    crossjoin): the left input is materialised into a workspace table, and
    every row of the right pipeline walks all of its rows, evaluating the
    join condition per pair.  The walk is wavefront-uniform: all lanes read
    the same inner row in a trip, one broadcast load per value.  Meant for
    the small inner sides of TPC-H Q7/Q11/Q22 (single aggregates, nation
    pairs)."""

    def __init__(self, algExpr, leftChild, rightChild):
        BinaryTranslator.__init__(self, algExpr, leftChild, rightChild)
        self.consumeCall = 0

    def produce(self, ctxt):
        self.leftChild.produce(ctxt)
        self.rightChild.produce(ctxt)

    def _innerAttrs(self):
        a = self.algExpr
        used = set(a.outRelation) | set(getattr(a, "conditionAttributes", {}))
        return [att for id, att in a.leftChild.outRelation.items() if id in used]

    def consume(self, ctxt):
        self.consumeCall += 1
        a = self.algExpr
        if self.consumeCall % 2 == 1:
            attrs = self._innerAttrs()
            ctxt.pipe.emit("// nested-loop join: materialise the inner side (opId %d)" % a.opId)
            self.inner = materializeRows(ctxt, attrs, False, "inner%d" % a.opId,
                                         ctxt.program.estimate(a.leftChild), a.opId)
            return
        pipe = ctxt.pipe
        J = a.opId
        tmp = self.inner
        ti = tmp["index"]
        pipe.emit("// nested-loop join (opId %d)" % J)
        fcap = pipe.param("xcap%d" % J, "uint64_t", "L.tcap[%d]" % ti)
        pos = pipe.reg("xpos%d" % J, "uint32_t", "0")
        pipe.indexRegs["xpos%d" % J] = "uint32_t"
        byName = {c["name"]: c for c in tmp["columns"]}
        for att in self._innerAttrs():
            c = byName[att.name]
            f = pipe.param("x%d_%d" % (J, c["pos"]), "const %s*" % c["ctype"], "B_%s" % c["buffer"])
            pipe.setLazy(att, "%s[%s]" % (f, pos), ["xpos%d" % J], c["ctype"])
            if c["char_col"] is not None:
                ctxt.program.charSource[att.id] = c["char_col"]
        condRight = [att for id, att in getattr(a, "conditionAttributes", {}).items()
                     if id in a.rightChild.outRelation]
        pipe.need(condRight)
        pipe.emit("const uint64_t xn%d = std::min<uint64_t>(st->temp_count[%d], %s);" % (J, ti, fcap))
        pipe.emit("const int xact%d = active[u];" % J)
        pipe.loopOps += 1
        pipe.code.open("for (uint64_t xj = 0; xj < xn%d; ++xj)" % J)
        pipe.emit("%s = (uint32_t)xj;" % pos)
        pipe.emit("active[u] = xact%d;" % J)
        pipe.need([att for att in self._innerAttrs() if att.id in getattr(a, "conditionAttributes", {})])
        cond = ctxt.exprEmitter().emit(a.condition)
        pipe.emit("if (active[u]) active[u] = (%s) ? 1 : 0;" % cond)
        ctxt.innerLoopCount += 1
        pipe.innerDepth += 1
        self.parent.consume(ctxt)
        pipe.innerDepth -= 1
        ctxt.innerLoopCount -= 1
        pipe.code.close()
        pipe.emit("active[u] = 0;")


class AggregationTranslator(UnaryTranslator):
    """Hash aggregation (cudaTranslator.py:425-525).

    Accumulators: SUM and AVG over the same input share one sum; COUNT is one
    counter; MIN/MAX their own.  Three tiers, chosen per row at run time:
      * no grouping: per-lane register accumulators, one wave reduction and
        one global atomic per wavefront at the end of the kernel;
      * grouping with keys that pack into an exact tag (63 bits with the
        top bit set, or exactly 64 bits stored as key + 1 with the single
        key whose tag would be 0 kept in an extra slot at index cap): a workgroup
        LDS table of W tag slots whose accumulators are replicated R times
        (lane i uses replica i mod R, so lanes of one wavefront rarely hit the
        same LDS word), flushed to the global table once per workgroup; rows
        whose group does not fit the LDS table go straight to the global
        table;
      * other keys: the global table in hashed mode.
    """

    def produce(self, ctxt):
        self.child.produce(ctxt)
        self.consumeHashTable(ctxt)

    def _accumulators(self, ctxt):
        """Distinct accumulators.  An input that can be null (the build side
        of an outer join) guards its accumulator: COUNT(attr), SUM, MIN and
        MAX skip null rows (cudaTranslator.py:475-476 does so for COUNT);
        AVG divides by COUNT(*) as in the reference."""
        a = self.algExpr
        dec = ctxt.decimalRepr
        nulls = ctxt.pipe.nullFlags if ctxt.pipe is not None else {}
        self.accs = []
        self.aggAcc = OrderedDict()
        index = {}
        for id, (inId, red) in a.aggregateTuples.items():
            aggAttr = a.aggregateAttributes[id]
            ct = langType(aggAttr.dataType, dec)
            kind = {Reduction.SUM: "sum", Reduction.AVG: "sum", Reduction.COUNT: "count",
                    Reduction.MIN: "min", Reduction.MAX: "max"}[red]
            inAttr = a.aggregateInAttributes[inId] if inId is not None else None
            null = nulls.get(inAttr.id) if inAttr is not None else None
            if kind == "count" and null is None:
                inAttr = None
            key = (kind, inAttr.id if inAttr is not None else None, ct)
            if key not in index:
                index[key] = len(self.accs)
                self.accs.append({"kind": kind, "in": inAttr, "ctype": ct, "i": len(self.accs), "null": null})
            self.aggAcc[id] = index[key]
        if a.avgAggregates and not any(c["kind"] == "count" and c["in"] is None for c in self.accs):
            # AVG's divisor: COUNT(*) even when the plan only counts a nullable attribute
            self.accs.append({"kind": "count", "in": None, "ctype": CType.INT, "i": len(self.accs), "null": None})

    def _fusable(self, pipe, groups):
        """A group-by whose keys are exactly the keys of a unique join probed
        earlier in this pipeline (exact tags, no lane reassignment since) can
        accumulate at that join's slot: equal keys always resolve to the same
        slot, so slots are groups.  Small configured group counts keep the
        LDS tier instead."""
        a = self.algExpr
        if getattr(a, "numgroupsFromCfg", False) and a.numgroups <= 4096:
            return None
        if pipe.innerDepth:
            return None
        gids = sorted(g.id for g in groups)
        for pj in reversed(pipe.probedJoins):
            if pj["loopOps"] != pipe.loopOps:
                continue
            if gids == sorted(pj["buildKeyIds"]) or gids == sorted(pj["probeKeyIds"]):
                return pj
        return None

    def consume(self, ctxt):
        a = self.algExpr
        pipe = ctxt.pipe
        prog = ctxt.program
        self._accumulators(ctxt)
        pipe.emit("// aggregation (opId %d)" % a.opId)
        ins = [acc["in"] for acc in self.accs if acc["in"] is not None and acc["kind"] != "count"]
        groups = list(a.groupAttributes.values())
        pipe.need(groups + ins)
        est = prog.estimate(a)
        self.doGroup = a.doGroup
        vals = []
        for acc in self.accs:
            if acc["kind"] == "count":
                vals.append("(%s)1" % acc["ctype"])
            else:
                vals.append("(%s)%s" % (acc["ctype"], pipe.access(acc["in"])))
        self.fused = self._fusable(pipe, groups) if a.doGroup else None
        if self.fused is not None:
            self._consumeFused(ctxt, pipe, prog, vals)
            return
        ht = prog.newHt("agg", a.opId, "aht%d" % a.opId)
        self.ht = ht
        idx = ht["index"]
        if not a.doGroup:
            prog.hostVars["htcap%d" % idx] = "L.htcap[%d] = 1; L.mment[%d] = 0;" % (idx, idx)
            self.accBuffers = []
            for acc in self.accs:
                b = prog.buffer("agg%d_%d" % (a.opId, acc["i"]), acc["ctype"], "1",
                                _fillPattern(acc["ctype"], _identity(acc["kind"], acc["ctype"])))
                self.accBuffers.append(b)
                f = pipe.param("acc%d_%d" % (a.opId, acc["i"]), "%s*" % acc["ctype"], "B_" + b)
                r = pipe.reg("racc%d_%d" % (a.opId, acc["i"]), acc["ctype"], _cIdentity(acc["kind"], acc["ctype"]),
                             perRow=False)
                acc["field"], acc["reg"] = f, r
            with pipe.rowLoop():
                pipe.code.open("if (active[u])")
                for acc, v in zip(self.accs, vals):
                    pipe.emit(_guarded(acc, _combine(acc["kind"], acc["reg"], v)))
                pipe.code.close()
            # end of kernel: wave reduction, the waves' partials combined in
            # LDS, one atomic per workgroup (thousands of atomics on one
            # address serialise at its memory channel)
            nw = pipe.bs // 64
            for acc in self.accs:
                red = {"sum": "wave_sum", "count": "wave_sum", "min": "wave_min", "max": "wave_max"}[acc["kind"]]
                pipe.staticLds.append("__shared__ %s wred%d_%d[%d];" % (acc["ctype"], a.opId, acc["i"], nw))
                pipe.epilogue.emit("%s = dqc::%s(%s);" % (acc["reg"], red, acc["reg"]))
            pipe.epilogue.open("if (lane == 0)")
            for acc in self.accs:
                pipe.epilogue.emit("wred%d_%d[wave] = %s;" % (a.opId, acc["i"], acc["reg"]))
            pipe.epilogue.close()
            pipe.epilogue.emit("__syncthreads();")
            pipe.epilogue.open("if (threadIdx.x == 0)")
            for acc in self.accs:
                w = "wred%d_%d" % (a.opId, acc["i"])
                pipe.epilogue.emit("for (uint32_t w = 1; w < %du; ++w) %s" % (nw, _combine(acc["kind"], w + "[0]",
                                                                                          w + "[w]")))
                pipe.epilogue.emit("%s;" % _atomic(acc["kind"], acc["field"] + "[0]", w + "[0]"))
            pipe.epilogue.close()
            return

        # grouped
        packed = _packKeys(groups)
        self.exact = packed is not None
        self.exact64 = self.exact and packed[2]
        extra = " + 1" if self.exact64 else ""
        factor = "fac(t->ht_factor[%d])" % idx
        prog.hostVars["htcap%d" % idx] = (
            "L.htcap[%d] = t->ht_capacity[%d] > 0 ? t->ht_capacity[%d] : "
            "std::max<int64_t>(64, (int64_t)((double)%s * 2.0 * %s)); L.mment[%d] = 0;\n"
            "if (L.htcap[%d] >= (int64_t)0xffffffffll) return DQC_ENOMEM;" % (idx, idx, idx, est, factor, idx, idx))
        tags = prog.buffer("aht%d_tags" % a.opId, CType.UINT64, "L.htcap[%d]%s" % (idx, extra), "0x0ull")
        ftags = pipe.param("aht%d" % a.opId, "uint64_t*", "B_" + tags)
        fcap = pipe.param("aht%d_cap" % a.opId, "uint32_t", "L.htcap[%d]" % idx)
        self.accBuffers = []
        for acc in self.accs:
            b = prog.buffer("agg%d_%d" % (a.opId, acc["i"]), acc["ctype"],
                            "(L.htcap[%d]%s) * (int64_t)L.arep[%d]" % (idx, extra, idx),
                            _fillPattern(acc["ctype"], _identity(acc["kind"], acc["ctype"])))
            self.accBuffers.append(b)
            acc["field"] = pipe.param("acc%d_%d" % (a.opId, acc["i"]), "%s*" % acc["ctype"], "B_" + b)
        # copies of the accumulators, one per group of workgroups: workgroup
        # b updates copy b % R, so with R = 8 each XCD (workgroups are dealt
        # to the 8 XCDs round robin) updates its own lines in its own L2
        # (tuning knob xcd_replicas; the output pass combines the copies)
        self.repStride = "(uint64_t)(L.htcap[%d]%s)" % (idx, extra)
        arep = pipe.param("arep%d" % a.opId, "uint32_t", "L.arep[%d]" % idx)
        astr = pipe.param("astr%d" % a.opId, "uint64_t", self.repStride)
        aoff = "aoff%d" % a.opId
        pipe.prologue.emit("const uint64_t %s = (uint64_t)(blockIdx.x & (%s - 1u)) * %s;" % (aoff, arep, astr))
        accBytes = sum(CType.size[acc["ctype"]] for acc in self.accs)

        def arepVar():
            # one copy unless the tuning asks for more: measured on c3's
            # 106k-group table, 8 copies saved 3% in the update kernel and
            # cost more in clearing and combining them
            prog.hostVars["arep%d" % idx] = (
                "{ const int32_t x = t->xcd_replicas; uint32_t R = 1u;\n"
                "  if (x > 0 && (uint64_t)(L.htcap[%d]%s) * %du * 8u <= (512ull << 20)) R = std::min<uint32_t>(8u, pow2floor((uint64_t)x));\n"
                "  L.arep[%d] = R; }" % (idx, extra, accBytes, idx))
        overflow = ("else { dqc::raise_status(&st->flags, DQC_ST_HT_OVERFLOW); atomicOr(&st->ht_overflow_mask, %du); }"
                    % (1 << idx))

        self.keyBuffers = []
        if self.exact:
            tagExpr, self.unpack, _ = packed
            G = a.opId
            upd = {"gtag": pipe.reg("gtag%d" % G, "uint64_t"), "gs0": pipe.reg("gs0_%d" % G, "uint32_t"),
                   "gt0": pipe.reg("gt0_%d" % G, "uint64_t"), "gneed": pipe.reg("gneed%d" % G, "int"),
                   "ftags": ftags, "fcap": fcap, "aoff": aoff, "overflow": overflow}
            if self._partitionable(pipe):
                estEff = self._partitioned(ctxt, pipe, prog, tagExpr(pipe), vals, upd, est)
                self._ldsTier(ctxt, pipe, prog, ftags, fcap, estEff, aoff)
            else:
                self._exactUpdate(pipe, tagExpr(pipe), vals, upd, "gh")
                self._ldsTier(ctxt, pipe, prog, ftags, fcap, est, aoff)
            arepVar()
        else:
            keyFields = []
            for k, g in enumerate(groups):
                ct = pipe.attrs[g.id].ctype
                b = prog.buffer("aht%d_k%d" % (a.opId, k), ct, "L.htcap[%d]" % idx)
                self.keyBuffers.append((b, ct))
                keyFields.append(pipe.param("aht%d_k%d" % (a.opId, k), "%s*" % ct, "B_" + b))
            self.keyFields = keyFields
            with pipe.rowLoop():
                pre = hoistStringHashes(pipe, groups)
                pipe.code.open("if (active[u])")
                pipe.emit("uint64_t gh = 0;")
                for g in groups:
                    pipe.emit("gh = dqc::hash_combine(gh, %s);" % _keyBits(g, pipe.access(g), pre))
                pipe.emit("const uint64_t gtag = dqc::TAG_TOP | (gh >> 1);")
                # keys written and compared during the kernel by lanes of
                # every XCD: coherent (agent-scope atomic) stores and loads
                writes = " ".join("dqc::coherent_store(&%s[s], %s);" % (f, pipe.access(g))
                                  for f, g in zip(keyFields, groups))
                eqs = " && ".join(_keyEq(g, "dqc::coherent_load(&%s[s])" % f, pipe.access(g))
                                  for f, g in zip(keyFields, groups))
                pipe.emit("const int64_t gs = dqc::agg_slot_hashed(%s, %s, gtag, gh, &st->flags, [&](uint32_t s) { %s }, "
                          "[&](uint32_t s) { return %s; });" % (ftags, fcap, writes, eqs))
                pipe.code.open("if (gs >= 0)")
                for acc, v in zip(self.accs, vals):
                    pipe.emit(_guarded(acc, "%s;" % _atomic(acc["kind"], "%s[%s + gs]" % (acc["field"], aoff), v)))
                pipe.code.close()
                pipe.emit(overflow)
                pipe.code.close()
            arepVar()

    def _exactUpdate(self, pipe, tagCode, vals, upd, ldsHash):
        """Exact-tag group update of the active rows: the LDS table when the
        group has (or gets) a slot there, else the global table.  `ldsHash`
        is the hash expression whose top bits pick the LDS slot."""
        gtag, gs0, gt0, gneed = upd["gtag"], upd["gs0"], upd["gt0"], upd["gneed"]
        ftags, fcap = upd["ftags"], upd["fcap"]
        # pass 1: LDS table, or the first global slot (loads of all rows of
        # the thread issued together)
        with pipe.rowLoop():
            pipe.emit("%s = 0;" % gneed)
            pipe.code.open("if (active[u])")
            pipe.emit("%s = %s;" % (gtag, tagCode))
            pipe.emit("const uint64_t gh = dqc::mix64(%s);" % gtag)
            pipe.emit("int ls = -1;")
            if self.exact64:
                # tag 0 (key ~0) bypasses the LDS table and uses slot cap
                pipe.emit("if (p.lw && %s != 0) ls = dqc::lds_slot_exact(ltags, p.lw, %s, %s);" % (gtag, gtag, ldsHash))
            else:
                pipe.emit("if (p.lw) ls = dqc::lds_slot_exact(ltags, p.lw, %s, %s);" % (gtag, ldsHash))
            pipe.code.open("if (ls >= 0)")
            pipe.emit("const uint32_t li = (uint32_t)ls * p.lr + (lane & (p.lr - 1u));")
            for acc, v in zip(self.accs, vals):
                pipe.emit(_guarded(acc, _ldsOp(acc["kind"], "&lacc%d[li]" % acc["i"], v)))
            pipe.code.close()
            pipe.code.open("else")
            pipe.emit("%s = 1;" % gneed)
            self._firstGlobalSlot(pipe, upd)
            pipe.code.close()
            pipe.code.close()
        self._globalUpdate(pipe, vals, upd)

    def _firstGlobalSlot(self, pipe, upd):
        gtag, gs0, gt0, ftags, fcap = upd["gtag"], upd["gs0"], upd["gt0"], upd["ftags"], upd["fcap"]
        if self.exact64:
            pipe.emit("%s = %s != 0 ? dqc::slot_of(gh, %s) : %s;" % (gs0, gtag, fcap, fcap))
            pipe.emit("%s = %s != 0 ? %s[%s] : 0ull;" % (gt0, gtag, ftags, gs0))
        else:
            pipe.emit("%s = dqc::slot_of(gh, %s);" % (gs0, fcap))
            pipe.emit("%s = %s[%s];" % (gt0, ftags, gs0))

    def _globalUpdate(self, pipe, vals, upd):
        """pass 2: resolve collisions / claim slots, update the group."""
        gtag, gs0, gt0, gneed = upd["gtag"], upd["gs0"], upd["gt0"], upd["gneed"]
        ftags, fcap, aoff = upd["ftags"], upd["fcap"], upd["aoff"]
        with pipe.rowLoop():
            pipe.code.open("if (%s)" % gneed)
            pipe.emit("const int64_t gs = %s == %s ? (int64_t)%s : dqc::agg_slot_exact_from(%s, %s, %s, %s, 0u, &st->flags);"
                      % (gt0, gtag, gs0, ftags, fcap, gtag, gs0))
            pipe.code.open("if (gs >= 0)")
            for acc, v in zip(self.accs, vals):
                pipe.emit(_guarded(acc, "%s;" % _atomic(acc["kind"], "%s[%s + gs]" % (acc["field"], aoff), v)))
            if self.exact64:
                pipe.emit("if (%s == 0) %s[%s] = 1ull;   // marks the extra slot used" % (gtag, ftags, fcap))
            pipe.code.close()
            pipe.emit(upd["overflow"])
            pipe.code.close()

    # partitioned group-by: at most this many partitions and first-pass
    # workgroups (segments per partition); auto mode above this many groups
    # (below, the LDS table holds a workgroup's groups), aiming at this many
    # groups per partition
    PART_MAX = 1024
    PART_SEG_MAX = 2048
    PART_MIN_GROUPS = 16384
    PART_GROUPS = 256

    def _partitionable(self, pipe):
        """A group-by with many groups updates the global table once per row,
        and those device-scope atomics are the bottleneck.  It can instead be
        launched in two phases when its update is not nested in a loop of the
        pipeline (no inner loops, lane reassignment or refill marks before it)
        and no accumulator skips null rows."""
        if pipe.innerDepth or pipe.loopOps or pipe.phased or pipe.refillPoints:
            return False
        if any(l is None for l in pipe.code.lines):
            return False
        if any(acc["null"] for acc in self.accs):
            return False
        a = self.algExpr
        if getattr(a, "partitions", None):
            return True
        if getattr(a, "numgroupsFromCfg", False) and a.numgroups <= self.PART_MIN_GROUPS:
            return False
        return True

    def _partitioned(self, ctxt, pipe, prog, tagCode, vals, upd, est):
        """Partitioned group-by, two launches of the pipeline kernel.

        PHASE 0 runs the pipeline; each row's group goes to partition
        (top bits of its hash).  Workgroup w of the launch owns segment w of
        every partition and appends (tag, accumulator inputs) records there
        through an LDS cursor per partition -- no global atomics, no barriers
        in the loop; rows that do not fit their segment (and the tag-0 key)
        update the global table directly.  At its end the workgroup stores
        its segment counts.  PHASE 1 gives workgroup b partition b: its rows
        are the records of the segments of partition b, aggregated in the
        LDS table (slots from the hash bits below the partition bits) and
        flushed once per group.  Partitions hold disjoint groups, so the
        global table receives one update per group and workgroup instead of
        one per row.  With no partitions (tuning part_agg, or a small or
        unknown group count) PHASE 0 is the plain exact-tag update and
        PHASE 1 is not launched."""
        a = self.algExpr
        G = a.opId
        idx = self.ht["index"]
        B, U = pipe.bs, pipe.U
        groupsKnown = "(int64_t)%d" % int(a.numgroups) if getattr(a, "numgroupsFromCfg", False) else "(int64_t)0"
        cfgParts = 0
        if getattr(a, "partitions", None):
            cfgParts = min(self.PART_MAX, max(2, _pow2floor(int(a.partitions))))
        prog.hostVars["apart%d" % idx] = (
            "{ const int32_t req = t->part_agg;\n"
            "  const int64_t known = t->ht_capacity[%d] > 0 ? t->ht_capacity[%d] / 2 : %s;\n"
            "  uint32_t np = 0u;\n"
            "  if (req > 0) np = std::min<uint32_t>(%du, std::max<uint32_t>(2u, pow2floor((uint64_t)req)));\n"
            "  else if (req == 0 && %d > 0) np = %du;\n"
            "  else if (req == 0 && known > %d) np = std::min<uint32_t>(%du, std::max<uint32_t>(2u, pow2ceil((uint64_t)(known / %d))));\n"
            "  uint32_t bits = 0u; while ((1u << bits) < np) ++bits;\n"
            "  L.apart[%d] = np; L.apbits[%d] = bits;\n"
            "  L.apgroups[%d] = known > 0 ? known : %s;\n"
            "  if (np) {\n"
            "    const uint32_t seg = std::min<uint32_t>(%du, gridFor(t->grid[%d] > 0 ? t->grid[%d] : %d, %s, %d, 1));\n"
            "    const double mu = (double)std::max<int64_t>(1, %s) / ((double)seg * np);\n"
            "    L.apseg[%d] = seg;\n"
            "    L.apcap[%d] = std::min<int64_t>(((int64_t)1 << 31) / seg, (int64_t)(mu * 1.25 + 4.0 * std::sqrt(mu)) + 16);\n"
            "  } }"
            % (idx, idx, groupsKnown, self.PART_MAX, cfgParts, cfgParts, self.PART_MIN_GROUPS, self.PART_MAX,
               self.PART_GROUPS,
               idx, idx, idx, est,
               self.PART_SEG_MAX, pipe.pipeId, pipe.pipeId, pipe.defaultGrid, pipe.params["n"][1], B * U,
               prog.estimate(a.child), idx, idx))
        # a record is 1 + (inputs) 8-byte words, stored row-wise so that a
        # partition's records in a segment share cache lines
        K = 1 + sum(1 for acc in self.accs if acc["kind"] != "count")
        recs = "(uint64_t)L.apart[%d] * L.apseg[%d] * (uint64_t)L.apcap[%d] * %d" % (idx, idx, idx, K)
        fnp = pipe.param("np%d" % G, "uint32_t", "L.apart[%d]" % idx)
        fbits = pipe.param("pbits%d" % G, "uint32_t", "L.apbits[%d]" % idx)
        fseg = pipe.param("pseg%d" % G, "uint32_t", "L.apseg[%d]" % idx)
        fpcap = pipe.param("pcap%d" % G, "uint32_t", "L.apcap[%d]" % idx)
        bcnt = prog.buffer("apc%d" % G, CType.UINT, "(uint64_t)L.apart[%d] * L.apseg[%d]" % (idx, idx))
        fcnt = pipe.param("apc%d" % G, "uint32_t*", "B_" + bcnt)
        brec = prog.buffer("aprec%d" % G, CType.UINT64, recs)
        frec = pipe.param("aprec%d" % G, "uint64_t*", "B_" + brec)
        words = {}      # accumulator -> word of the record
        for acc in self.accs:
            if acc["kind"] != "count":
                words[acc["i"]] = 1 + len(words)
        pipe.annotations.append("PART%d" % G)
        gtag, gneed = upd["gtag"], upd["gneed"]
        cur, pref, scan = "pcur%d" % G, "ppref%d" % G, "pscan%d" % G
        pipe.staticLds.append("__shared__ uint32_t %s[%d];" % (cur, self.PART_MAX))
        pipe.staticLds.append("__shared__ uint32_t %s[%d];" % (pref, self.PART_SEG_MAX + 1))
        pipe.staticLds.append("__shared__ uint32_t %s[%d];" % (scan, 1024 // WAVE + 1))
        # (uniform branch: no barrier when the group-by is not partitioned)
        pipe.prologue.open("if (%s > 0u)" % fnp)
        pipe.prologue.emit("for (uint32_t i = threadIdx.x; i < %s; i += %du) %s[i] = 0u;" % (fnp, B, cur))
        pipe.prologue.emit("__syncthreads();")
        pipe.prologue.close()
        # PHASE 0: everything before this point, then the update or the scatter
        pipe.code.lines.insert(0, "    " * (pipe.code.depth - 1) + "if constexpr (PHASE == 0) {")
        pipe.code.depth += 1
        pipe.code.open("if (%s == 0u)" % fnp)
        self._exactUpdate(pipe, tagCode, vals, upd, "gh")
        pipe.code.close()
        pipe.code.open("else")
        pipe.emit("// append to this workgroup's segment of the row's partition")
        with pipe.rowLoop():
            pipe.emit("%s = 0;" % gneed)
            pipe.code.open("if (active[u])")
            pipe.emit("%s = %s;" % (gtag, tagCode))
            pipe.emit("const uint64_t gh = dqc::mix64(%s);" % gtag)
            pipe.emit("const uint32_t pb = (uint32_t)(gh >> (64u - %s));" % fbits)
            pipe.emit("const uint32_t pos = %s != 0ull ? atomicAdd(&%s[pb], 1u) : 0xffffffffu;" % (gtag, cur))
            pipe.code.open("if (pos < %s)" % fpcap)
            pipe.emit("uint64_t* rec = %s + (((uint64_t)pb * %s + blockIdx.x) * %s + pos) * %du;" % (frec, fseg, fpcap, K))
            pipe.emit("rec[0] = %s;" % gtag)
            for acc, v in zip(self.accs, vals):
                if acc["i"] in words:
                    pipe.emit("rec[%d] = dqc::to_word<%s>(%s);" % (words[acc["i"]], acc["ctype"], v))
            pipe.code.close()
            pipe.code.open("else")
            pipe.emit("// segment full, or the tag-0 key: the global table")
            pipe.emit("%s = 1;" % gneed)
            self._firstGlobalSlot(pipe, upd)
            pipe.code.close()
            pipe.code.close()
        self._globalUpdate(pipe, vals, upd)
        pipe.code.close()   # end of the scatter
        pipe.code.close()   # end of PHASE 0
        # PHASE 1: the loop index is a record of this workgroup's partition
        pipe.code.open("else")
        rr = pipe.reg("prec%d" % G, "uint64_t")
        pipe.emit("dqc::part_records<%d>(%s, %s, blockIdx.x, %s, %s, active, %s);"
                  % (U, pref, fseg, fpcap, pipe.tid, rr[:-3]))
        # every row's record is loaded before any row probes the LDS table
        ptag = pipe.reg("ptag%d" % G, "uint64_t")
        pvals = {i: pipe.reg("pv%d_%d" % (G, i), "uint64_t") for i in words}
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            pipe.emit("const uint64_t* rec = %s + %s * %du;" % (frec, rr, K))
            pipe.emit("%s = rec[0];" % ptag)
            for i, w in words.items():
                pipe.emit("%s = rec[%d];" % (pvals[i], w))
            pipe.code.close()
        vals1 = []
        for acc, v in zip(self.accs, vals):
            vals1.append(v if acc["i"] not in words else "dqc::from_word<%s>(%s)" % (acc["ctype"], pvals[acc["i"]]))
        self._exactUpdate(pipe, ptag, vals1, upd, "(gh << %s)" % fbits)
        pipe.code.close()
        # PHASE 0 end: the segment counts
        pipe.epilogue.open("if (PHASE == 0 && %s > 0u)" % fnp)
        pipe.epilogue.emit("__syncthreads();")
        pipe.epilogue.emit("for (uint32_t i = threadIdx.x; i < %s; i += %du) %s[(uint64_t)i * %s + blockIdx.x] = "
                           "std::min<uint32_t>(%s[i], %s);" % (fnp, B, fcnt, fseg, cur, fpcap))
        pipe.epilogue.close()
        pipe.phased = True
        pipe.phaseBetween = []
        pipe.partAgg = {"np": fnp, "cap": fpcap, "count": fcnt, "nseg": fseg, "pref": pref, "scan": scan,
                        "hostNp": "L.apart[%d]" % idx, "hostNseg": "L.apseg[%d]" % idx}
        # the LDS table of PHASE 1: sized for 4x the groups of one partition
        # (short probes; measured on 234 groups per partition: 2048 slots
        # 89 us, 512 slots 110 us), at most 4096 slots
        return ("(L.apart[%d] ? std::min<int64_t>(2048, std::max<int64_t>(1, 4 * L.apgroups[%d] / (int64_t)L.apart[%d])) : %s)"
                % (idx, idx, idx, est))

    def _consumeFused(self, ctxt, pipe, prog, vals):
        """Accumulate at the slot of the probed unique join (see _fusable);
        a per-slot touched byte marks groups that received rows."""
        a = self.algExpr
        pj = self.fused
        pipe.annotations.append("GJ%d" % pj["opId"])
        cap = "L.htcap[%d]%s" % (pj["htIndex"], " + 1" if pj["tag32"] else "")
        self.accBuffers = []
        for acc in self.accs:
            b = prog.buffer("agg%d_%d" % (a.opId, acc["i"]), acc["ctype"], cap,
                            _fillPattern(acc["ctype"], _identity(acc["kind"], acc["ctype"])))
            self.accBuffers.append(b)
            acc["field"] = pipe.param("acc%d_%d" % (a.opId, acc["i"]), "%s*" % acc["ctype"], "B_" + b)
        self.touched = prog.buffer("agg%d_hit" % a.opId, CType.CHAR, cap, _fillPattern(CType.CHAR, 0))
        fhit = pipe.param("agg%d_hit" % a.opId, "int8_t*", "B_" + self.touched)
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            pipe.emit("const int64_t gs = %s[u];" % pj["slot"])
            for acc, v in zip(self.accs, vals):
                pipe.emit(_guarded(acc, "%s;" % _atomic(acc["kind"], "%s[gs]" % acc["field"], v)))
            pipe.emit("%s[gs] = 1;" % fhit)
            pipe.code.close()

    def _ldsTier(self, ctxt, pipe, prog, ftags, fcap, est, aoff):
        """Workgroup LDS table: W slots (runtime, power of two), accumulators
        replicated R times; flushed once per workgroup."""
        pid = pipe.pipeId
        pipe.annotations.append("LDS%d" % self.algExpr.opId)
        # an estimate beyond 8192 slots turns the tier off, except for keys of
        # a small domain (CHAR, BOOLEAN) without a group-count hint: their
        # groups are often few whatever the input size, so the tier stays
        # with 4096 slots (groups beyond them go to the global table)
        a = self.algExpr
        bounded = _keyDomain(a.groupAttributes.values()) is not None and not getattr(a, "numgroupsFromCfg", False)
        overflowW = "4096u" if bounded else "0"
        accBytes = sum(CType.size[acc["ctype"]] for acc in self.accs)
        # LDS in use besides the table: static arrays, a join filter in front
        used = "(uint64_t)%du" % (staticLdsBytes(pipe) + 1024)
        devBase, hostBase = "(uint64_t)0", None
        if pipe.ldsFilter is not None:
            used = "(%s + %s)" % (used, pipe.ldsFilter["host"])
            devBase, hostBase = pipe.ldsFilter["dev"], pipe.ldsFilter["host"]
        prog.hostVars["lds%d" % pid] = (
            "{ int64_t w = t->lds_slots[%d]; uint32_t W = 0, R = 1;\n"
            "  if (w > 0) W = pow2floor((uint64_t)w);\n"
            "  else if (w == 0) W = std::max<uint32_t>(16u, pow2ceil((uint64_t)std::max<int64_t>(1, %s) * 2u));\n"
            "  if (W > 8192u) W = %s;\n"
            "  W = std::min<uint32_t>(W, 4096u);\n"
            "  if (W) {\n"
            "    int64_t r = t->lds_replicas[%d];\n"
            "    if (r > 0) R = std::min<uint32_t>(64u, pow2floor((uint64_t)r));\n"
            "    else { R = 64u; while (R > 1u && (uint64_t)W * (8u + (uint64_t)R * %du) > %du) R /= 2u; }\n"
            "    while (W > 16u && (uint64_t)W * (8u + (uint64_t)R * %du) + %s > maxLds()) W /= 2u;\n"
            "    if ((uint64_t)W * (8u + (uint64_t)R * %du) + %s > maxLds()) W = 0;\n"
            "  }\n"
            "  L.ldsw[%d] = W; L.ldsr[%d] = W ? R : 1u; }"
            % (pid, est, overflowW, pid, accBytes, LDS_AGG_BUDGET, accBytes, used, accBytes, used, pid, pid))
        pipe.param("lw", "uint32_t", "L.ldsw[%d]" % pid)
        pipe.param("lr", "uint32_t", "L.ldsr[%d]" % pid)
        # dynamic LDS: tags, then one region per accumulator (8-byte aligned)
        parts = ["(uint64_t)L.ldsw[%d] * 8u" % pid]
        for acc in self.accs:
            parts.append("(((uint64_t)L.ldsw[%d] * L.ldsr[%d] * %d + 7u) & ~(uint64_t)7u)"
                         % (pid, pid, CType.size[acc["ctype"]]))
        if hostBase is not None:
            parts.insert(0, hostBase)
        pipe.dynLdsExpr = " + ".join(parts)
        pro = pipe.prologue
        pro.emit("uint64_t* ltags = (uint64_t*)(dqc_dyn_lds + %s);" % devBase)
        pro.emit("uint64_t loff = %s + (uint64_t)p.lw * 8u;" % devBase)
        for acc in self.accs:
            pro.emit("%s* lacc%d = (%s*)(dqc_dyn_lds + loff);" % (acc["ctype"], acc["i"], acc["ctype"]))
            pro.emit("loff += ((uint64_t)p.lw * p.lr * %d + 7u) & ~(uint64_t)7u;" % CType.size[acc["ctype"]])
        pro.emit("for (uint32_t i = threadIdx.x; i < p.lw; i += %du) ltags[i] = 0;" % pipe.bs)
        pro.open("for (uint32_t i = threadIdx.x; i < p.lw * p.lr; i += %du)" % pipe.bs)
        for acc in self.accs:
            pro.emit("lacc%d[i] = %s;" % (acc["i"], _cIdentity(acc["kind"], acc["ctype"])))
        pro.close()
        pro.emit("__syncthreads();")
        ep = pipe.epilogue
        ep.emit("__syncthreads();")
        # few replicas: one thread per slot folds them; many: one wavefront
        # per slot reduces them across lanes
        ep.open("if (p.lr < 16u)")
        ep.open("for (uint32_t s = threadIdx.x; s < p.lw; s += %du)" % pipe.bs)
        ep.emit("const uint64_t tg = ltags[s];")
        ep.emit("if (tg == 0) continue;")
        for acc in self.accs:
            ep.emit("%s f%d = lacc%d[s * p.lr];" % (acc["ctype"], acc["i"], acc["i"]))
            ep.emit("for (uint32_t r = 1; r < p.lr; ++r) { const %s v = lacc%d[s * p.lr + r]; %s }"
                    % (acc["ctype"], acc["i"], _combine(acc["kind"], "f%d" % acc["i"], "v")))
        ep.emit("const int64_t gs = dqc::agg_slot_exact(%s, %s, tg, dqc::mix64(tg), &st->flags);" % (ftags, fcap))
        ep.open("if (gs >= 0)")
        for acc in self.accs:
            ep.emit("%s;" % _atomic(acc["kind"], "%s[%s + gs]" % (acc["field"], aoff), "f%d" % acc["i"]))
        ep.close()
        ep.emit("else { dqc::raise_status(&st->flags, DQC_ST_HT_OVERFLOW); atomicOr(&st->ht_overflow_mask, %du); }"
                % (1 << self.ht["index"]))
        ep.close()
        ep.close()
        ep.open("else")
        ep.open("for (uint32_t s = wave; s < p.lw; s += %du)" % (pipe.bs // WAVE))
        ep.emit("const uint64_t tg = ltags[s];")
        ep.emit("if (tg == 0) continue;")
        for acc in self.accs:
            red = {"sum": "wave_sum", "count": "wave_sum", "min": "wave_min", "max": "wave_max"}[acc["kind"]]
            ep.emit("%s f%d = lane < p.lr ? lacc%d[s * p.lr + lane] : %s;" % (
                acc["ctype"], acc["i"], acc["i"], _cIdentity(acc["kind"], acc["ctype"])))
            ep.emit("f%d = dqc::%s(f%d);" % (acc["i"], red, acc["i"]))
        ep.open("if (lane == 0)")
        ep.emit("const int64_t gs = dqc::agg_slot_exact(%s, %s, tg, dqc::mix64(tg), &st->flags);" % (ftags, fcap))
        ep.open("if (gs >= 0)")
        for acc in self.accs:
            ep.emit("%s;" % _atomic(acc["kind"], "%s[%s + gs]" % (acc["field"], aoff), "f%d" % acc["i"]))
        ep.close()
        ep.emit("else { dqc::raise_status(&st->flags, DQC_ST_HT_OVERFLOW); atomicOr(&st->ht_overflow_mask, %du); }"
                % (1 << self.ht["index"]))
        ep.close()
        ep.close()
        ep.close()

    def consumeHashTable(self, ctxt):
        """Second pipeline: scan the aggregation table, emit one tuple per
        occupied slot, AVG = SUM / COUNT (cudaTranslator.py:497-525)."""
        a = self.algExpr
        prog = ctxt.program
        name = "krnl_aggregation%d" % a.opId
        rows = "L.htcap[%d]" % (self.fused["htIndex"] if self.fused else self.ht["index"])
        if (self.fused is None and self.doGroup and self.exact64) or (self.fused is not None and self.fused["tag32"]):
            rows += " + 1"     # the extra slot
        pipe = prog.newPipeline(name, {"rows": rows, "tidName": "agg%d" % a.opId}, rowsPerThread(self))
        ctxt.pipe = pipe
        tid = pipe.tid + "[u]"
        pipe.param("st", "dqc_plan_status*", "st")
        accFields = []
        for acc, b in zip(self.accs, self.accBuffers):
            accFields.append(pipe.param("acc%d_%d" % (a.opId, acc["i"]), "const %s*" % acc["ctype"], "B_" + b))
        groups = list(a.groupAttributes.values())
        if self.fused is not None:
            pj = self.fused
            J = pj["opId"]
            join = pj["join"]
            fhit = pipe.param("agg%d_hit" % a.opId, "const int8_t*", "B_" + self.touched)
            tt = "uint32_t" if pj["tag32"] else "uint64_t"
            ftags = pipe.param("jht%d" % J, "const %s*" % tt, "B_jht%d_tags" % J)
            gtag = pipe.reg("gtag", CType.UINT64, "0")
            keyIds = pj["buildKeyIds"] if sorted(g.id for g in groups) == sorted(pj["buildKeyIds"]) \
                else pj["probeKeyIds"]
            with pipe.rowLoop():
                pipe.emit("if (active[u]) active[u] = %s[%s] != 0;" % (fhit, tid))
                if pj["tag32"]:
                    # the packed key: tag - 1, all bits set in the extra slot
                    fjcap = pipe.param("jht%d_cap" % J, "uint32_t", "L.htcap[%d]" % pj["htIndex"])
                    pipe.emit("if (active[u]) %s = %s == (int64_t)%s ? 0xffffffffull : (uint64_t)(%s[%s] - 1u);"
                              % (gtag, tid, fjcap, ftags, tid))
                else:
                    pipe.emit("if (active[u]) %s = %s[%s];" % (gtag, ftags, tid))
                for g in groups:
                    k = keyIds.index(g.id)
                    shift = sum(PACK_BITS[b.dataType] for b in join.buildKeys[:k])
                    w = PACK_BITS[join.buildKeys[k].dataType]
                    src = "uint32_t" if w == 32 else "uint8_t"
                    st = pipe.setComputed(g)
                    pipe.emit("%s[u] = (%s)(%s)(%s >> %d);" % (st.reg, st.ctype, src, gtag, shift))
        elif a.doGroup:
            ftags = pipe.param("aht%d" % a.opId, "const uint64_t*", "B_aht%d_tags" % a.opId)
            gtag = pipe.reg("gtag", CType.UINT64, "0")
            with pipe.rowLoop():
                if self.exact64:
                    fcap = pipe.param("aht%d_cap" % a.opId, "uint32_t", "L.htcap[%d]" % self.ht["index"])
                    pipe.emit("if (active[u]) { %s = %s[%s]; active[u] = %s != 0 ? 1 : 0; }" % (gtag, ftags, tid, gtag))
                    pipe.emit("const uint64_t pk = %s == (int64_t)%s ? ~0ull : %s - 1ull;" % (tid, fcap, gtag))
                    src = "pk"
                else:
                    pipe.emit("if (active[u]) { %s = %s[%s]; active[u] = (%s >> 63) ? 1 : 0; }"
                              % (gtag, ftags, tid, gtag))
                    src = gtag
                if self.exact:
                    for g, unpack in zip(groups, self.unpack):
                        st = pipe.setComputed(g)
                        pipe.emit("%s[u] = %s;" % (st.reg, unpack(src)))
            if not self.exact:
                for g, (b, ct), k in zip(groups, self.keyBuffers, range(len(groups))):
                    f = pipe.param("aht%d_k%d" % (a.opId, k), "const %s*" % ct, "B_" + b)
                    pipe.setLazy(g, "%s[%s]" % (f, tid), [pipe.tid], ct)
        countAcc = None
        for acc in self.accs:
            if acc["kind"] == "count" and acc["in"] is None:
                countAcc = acc
        replicated = self.fused is None and a.doGroup
        if replicated:
            arep = pipe.param("arep%d" % a.opId, "uint32_t", "L.arep[%d]" % self.ht["index"])
            astr = pipe.param("astr%d" % a.opId, "uint64_t", self.repStride)

        def accValue(i):
            if not replicated:
                return "%s[%s]" % (accFields[i], tid)
            fn = {"sum": "rep_sum", "count": "rep_sum", "min": "rep_min", "max": "rep_max"}[self.accs[i]["kind"]]
            return "dqc::%s(%s, %s, %s, %s)" % (fn, accFields[i], tid, arep, astr)

        for id, accIdx in self.aggAcc.items():
            attr = a.aggregateAttributes[id]
            ct = langType(attr.dataType, ctxt.decimalRepr)
            loader = "(%s)%s" % (ct, accValue(accIdx))
            if id in a.avgAggregates:
                loader = "(%s)(%s / (%s)%s)" % (ct, accValue(accIdx), ct, accValue(countAcc["i"]))
            pipe.setLazy(attr, loader, [pipe.tid], ct)
        self.parent.consume(ctxt)
        closePipeline(ctxt, pipe)


def _guarded(acc, stmt):
    """An accumulator update, skipped for rows whose input is null."""
    if acc.get("null"):
        return "if (!(%s)) { %s }" % (acc["null"], stmt)
    return stmt


def _combine(kind, reg, v):
    if kind in ("sum", "count"):
        return "%s += %s;" % (reg, v)
    if kind == "min":
        return "%s = %s < %s ? %s : %s;" % (reg, v, reg, v, reg)
    return "%s = %s > %s ? %s : %s;" % (reg, v, reg, v, reg)


def _atomic(kind, lvalue, v):
    fn = {"sum": "atomic_add", "count": "atomic_add", "min": "atomic_min", "max": "atomic_max"}[kind]
    return "dqc::%s(&%s, %s)" % (fn, lvalue, v)


def _ldsOp(kind, ptr, v):
    fn = {"sum": "lds_add", "count": "lds_add", "min": "lds_min", "max": "lds_max"}[kind]
    return "dqc::%s(%s, %s);" % (fn, ptr, v)


class StringHash(object):
    """How string keys are hashed: by each lane for its own row (default),
    or, with ``wave``, by the whole wavefront one string at a time
    (dqc::string_hash_wave; the reference's stringHashPushDown,
    hashing.cuh:39-62, which hashTableUtil.py:20 leaves commented out).  Both
    give the same hash value, so tables built one way are probed the other."""
    wave = False


def hoistStringHashes(pipe, attrs):
    """With StringHash.wave, emit the wavefront string hash of every STRING
    attribute of `attrs` for the current row, at the current (wavefront-
    uniform) position; returns {attr id: hash register} for _keyBits."""
    pre = {}
    if not StringHash.wave:
        return pre
    for a in attrs:
        if a.dataType == Type.STRING and a.id not in pre:
            name = "wsh%d_%d" % (a.id, pipe.uniqueId())
            pipe.emit("const uint64_t %s = dqc::string_hash_wave(active[u], %s);" % (name, pipe.access(a)))
            pre[a.id] = name
    return pre


def _keyBits(attr, reg, pre=None):
    t = attr.dataType
    if t == Type.STRING:
        if pre and attr.id in pre:
            return pre[attr.id]
        return "dqc::string_hash(%s)" % reg
    if t == Type.DOUBLE or t == Type.FLOAT:
        return "(uint64_t)__double_as_longlong((%s) == 0 ? 0.0 : (double)(%s))" % (reg, reg)
    if t == Type.INT64:
        return "(uint64_t)(%s)" % reg
    return "(uint64_t)(uint32_t)(%s)" % reg


def _keyEq(attr, a, b):
    if attr.dataType == Type.STRING:
        return "dqc::string_equals(%s, %s)" % (a, b)
    return "(%s == %s)" % (a, b)


_KEY_MASK = {8: "0xffull", 32: "0xffffffffull", 64: "0xffffffffffffffffull"}
_KEY_SRC = {8: "uint8_t", 32: "uint32_t", 64: "uint64_t"}


def _packKeys(attrs):
    """Exact tag of a group-key tuple, or None when the keys need more than
    64 bits.  Up to 63 bits: TAG_TOP | packed (never 0).  Exactly 64 bits:
    packed + 1, which is 0 only for packed == ~0 (that key gets the extra
    slot at index cap).  Returns (tagExprFn(pipe), [unpackFn(packedVar)],
    is64); the unpack functions take the tag itself in the 63-bit form and
    the packed value in the 64-bit form."""
    widths = []
    for a in attrs:
        w = _GROUP_BITS.get(a.dataType)
        if w is None:
            return None
        widths.append(w)
    if sum(widths) > 64 or not widths:
        return None
    is64 = sum(widths) == 64

    def tagExpr(pipe):
        parts = []
        shift = 0
        for a, w in zip(attrs, widths):
            parts.append("(((uint64_t)(%s) & %s) << %d)" % (pipe.access(a), _KEY_MASK[w], shift))
            shift += w
        if is64:
            return "((%s) + 1ull)" % " | ".join(parts)
        return " | ".join(["dqc::TAG_TOP"] + parts)

    unpacks = []
    shift = 0
    for a, w in zip(attrs, widths):
        ct = langType(a.dataType, CType.FP64)
        unpacks.append(lambda tv, s=shift, src=_KEY_SRC[w], ct=ct: "(%s)(%s)((%s) >> %d)" % (ct, src, tv, s))
        shift += w
    return tagExpr, unpacks, is64


# bits of a group key in an exact aggregation tag (joins pack up to 63 bits
# with PACK_BITS; INT64 keys fit only the 64-bit aggregation form)
_GROUP_BITS = dict(PACK_BITS)
_GROUP_BITS[Type.INT64] = 64
