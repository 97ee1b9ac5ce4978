"""ctypes binding of libdqcrt.so and of plan modules (include/dogqc_amd.h).

This is the Python side of the C ABI: it binds numpy/mmdb columns to a plan
module's input slots, keeps them resident in device memory, runs the plan
with the runtime's retry driver and copies result columns back.  There is no
fallback: if the shared library or a GPU is missing, calls raise.
"""
import ctypes as C
import json
import math
import os
from collections import OrderedDict

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libdqcrt.so")

MAX_PIPES, MAX_HTS, MAX_RESULT_COLS, MAX_TEMPS, PROFILE_BUCKETS = 32, 32, 64, 8, 65

RUNTIME_SYMBOLS = [
    "dqc_last_error", "dqc_device_init", "dqc_malloc", "dqc_free", "dqc_memcpy_h2d", "dqc_memcpy_d2h",
    "dqc_memset", "dqc_stream_create", "dqc_stream_destroy", "dqc_stream_sync", "dqc_event_create",
    "dqc_event_destroy", "dqc_event_elapsed_ms", "dqc_event_record", "dqc_mmdb_load", "dqc_module_load",
    "dqc_module_unload", "dqc_module_describe", "dqc_query_layout", "dqc_query_enqueue", "dqc_query_run",
    "dqc_query_release", "dqc_graph_create", "dqc_graph_launch", "dqc_graph_destroy",
    "dqc_serve_create", "dqc_serve_submit", "dqc_serve_start", "dqc_serve_wait", "dqc_serve_record_of",
    "dqc_serve_release", "dqc_serve_destroy", "dqc_query_run_noalloc",
]
NOALLOC_WS_BYTES = 64 << 20  # DQC_NOALLOC_WS_BYTES
GEN_SYMBOLS = ["dqc_gen_uniform", "dqc_gen_decimal", "dqc_gen_pick_u32", "dqc_gen_affine_perm",
               "dqc_gen_inverse_cdf", "dqc_gen_text", "dqc_gen_prefix_key"]
PLAN_SYMBOLS = ["dqc_plan_describe", "dqc_plan_layout_of", "dqc_plan_enqueue"]


class Inputs(C.Structure):
    _fields_ = [("ntables", C.c_int32), ("ncols", C.c_int32), ("table_rows", C.POINTER(C.c_int64)),
                ("cols", C.POINTER(C.c_void_p))]


class Tuning(C.Structure):
    _fields_ = [("grid", C.c_int32 * MAX_PIPES), ("refill_threshold", C.c_int32 * MAX_PIPES),
                ("lds_slots", C.c_int32 * MAX_PIPES), ("lds_replicas", C.c_int32 * MAX_PIPES),
                ("ht_factor", C.c_float * MAX_HTS), ("ht_capacity", C.c_int64 * MAX_HTS),
                ("mm_entries", C.c_int64 * MAX_HTS), ("result_capacity", C.c_int64),
                ("temp_capacity", C.c_int64 * MAX_TEMPS), ("split_capacity", C.c_int64 * MAX_HTS),
                ("split_threshold", C.c_int32 * MAX_PIPES), ("profile", C.c_int32), ("xcd_replicas", C.c_int32),
                ("census", C.c_int32), ("part_agg", C.c_int32), ("join_filter", C.c_int32),
                ("part_build", C.c_int32), ("reserved", C.c_int32 * 2)]


class Status(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("ht_overflow_mask", C.c_uint32), ("mm_overflow_mask", C.c_uint32),
                ("reserved", C.c_uint32), ("result_count", C.c_uint64), ("mm_total", C.c_uint64 * MAX_HTS),
                ("temp_count", C.c_uint64 * MAX_TEMPS), ("split_count", C.c_uint64 * MAX_HTS),
                ("ht_fill", C.c_uint64 * MAX_HTS)]


class Layout(C.Structure):
    _fields_ = [("ws_bytes", C.c_uint64), ("nsteps", C.c_int32), ("nhts", C.c_int32),
                ("ht_capacity", C.c_int64 * MAX_HTS), ("mm_entries", C.c_int64 * MAX_HTS),
                ("result_capacity", C.c_int64), ("nresult", C.c_int32), ("nprofile", C.c_int32),
                ("result_offset", C.c_uint64 * MAX_RESULT_COLS), ("profile_offset", C.c_uint64),
                ("ntemps", C.c_int32), ("ntempcols", C.c_int32), ("temp_capacity", C.c_int64 * MAX_TEMPS),
                ("temp_offset", C.c_uint64 * MAX_RESULT_COLS), ("split_capacity", C.c_int64 * MAX_HTS)]


class QueryResult(C.Structure):
    _fields_ = [("ws", C.c_void_p), ("ws_bytes", C.c_uint64), ("layout", Layout), ("status", Status),
                ("retries", C.c_int32)]


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_retries", C.c_int32), ("arena_bytes", C.c_uint64),
                ("bypass_ms", C.c_int32), ("size_feedback", C.c_int32), ("max_heavy", C.c_int32),
                ("heavy_bytes", C.c_uint64), ("max_batch", C.c_int32), ("batch_ratio", C.c_int32),
                ("batch_min_share", C.c_double)]


class ServeRequest(C.Structure):
    _fields_ = [("module", C.c_void_p), ("inputs", Inputs), ("tuning", C.POINTER(Tuning)),
                ("arrival_ns", C.c_int64), ("keep_result", C.c_int32), ("batch_ntables", C.c_int32),
                ("batch_module", C.c_void_p), ("batch_cols", C.POINTER(C.c_void_p)),
                ("batch_col_prefix", C.POINTER(C.c_int32)), ("batch_table_map", C.POINTER(C.c_int32)),
                ("batch_ncols", C.c_int32), ("reserved", C.c_int32)]


class ServeRecord(C.Structure):
    _fields_ = [("arrival_ns", C.c_int64), ("admit_ns", C.c_int64), ("done_ns", C.c_int64),
                ("state", C.c_int32), ("retries", C.c_int32), ("error", C.c_int32), ("slot", C.c_int32),
                ("ws", C.c_void_p), ("ws_bytes", C.c_uint64), ("batch_size", C.c_int32),
                ("batch_index", C.c_int32), ("layout", Layout), ("status", Status)]


class DqcError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libdqcrt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise DqcError("libdqcrt.so is not built (run python3 -m dogqc_amd.build)")
        L = C.CDLL(LIBPATH)
        L.dqc_last_error.restype = C.c_char_p
        L.dqc_module_describe.restype = C.c_char_p
        vp = C.c_void_p
        L.dqc_device_init.argtypes = [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
        L.dqc_malloc.argtypes = [C.POINTER(vp), C.c_uint64]
        L.dqc_free.argtypes = [vp]
        L.dqc_memcpy_h2d.argtypes = [vp, vp, C.c_uint64]
        L.dqc_memcpy_d2h.argtypes = [vp, vp, C.c_uint64]
        L.dqc_memset.argtypes = [vp, C.c_int, C.c_uint64]
        L.dqc_stream_create.argtypes = [C.POINTER(vp)]
        L.dqc_stream_destroy.argtypes = [vp]
        L.dqc_stream_sync.argtypes = [vp]
        L.dqc_event_create.argtypes = [C.POINTER(vp)]
        L.dqc_event_destroy.argtypes = [vp]
        L.dqc_event_record.argtypes = [vp, vp]
        L.dqc_event_elapsed_ms.argtypes = [vp, vp, C.POINTER(C.c_float)]
        L.dqc_mmdb_load.argtypes = [C.c_char_p, C.POINTER(vp), C.POINTER(C.c_uint64)]
        L.dqc_module_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.dqc_module_unload.argtypes = [vp]
        L.dqc_module_describe.argtypes = [vp]
        L.dqc_query_layout.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Tuning), C.POINTER(Layout)]
        L.dqc_query_enqueue.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Tuning), vp, C.c_uint64, vp,
                                        C.POINTER(vp)]
        L.dqc_query_run.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Tuning), vp, C.c_int32, C.POINTER(QueryResult)]
        L.dqc_query_release.argtypes = [C.POINTER(QueryResult)]
        L.dqc_query_run_noalloc.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Tuning), vp, C.c_int32,
                                            C.POINTER(vp), C.c_int64, C.POINTER(QueryResult)]
        L.dqc_graph_create.argtypes = [vp, C.POINTER(Inputs), C.POINTER(Tuning), vp, C.c_uint64, vp, C.POINTER(vp)]
        L.dqc_graph_launch.argtypes = [vp, vp]
        L.dqc_graph_destroy.argtypes = [vp]
        L.dqc_serve_create.argtypes = [C.POINTER(ServeConfig), C.POINTER(vp)]
        L.dqc_serve_submit.argtypes = [vp, C.POINTER(ServeRequest), C.POINTER(C.c_int64)]
        L.dqc_serve_start.argtypes = [vp]
        L.dqc_serve_wait.argtypes = [vp, C.c_int64]
        L.dqc_serve_record_of.argtypes = [vp, C.c_int64, C.POINTER(ServeRecord)]
        L.dqc_serve_release.argtypes = [vp, C.c_int64]
        L.dqc_serve_destroy.argtypes = [vp]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().dqc_last_error().decode(errors="replace")
        raise DqcError("%s failed (%d): %s" % (what, rc, msg))


_device = None


def device(index=0):
    """Initialise the device once; returns (compute units, HBM bytes)."""
    global _device
    if _device is None:
        cus, hbm = C.c_int32(), C.c_uint64()
        check(lib().dqc_device_init(index, C.byref(cus), C.byref(hbm)), "dqc_device_init")
        _device = (cus.value, hbm.value)
    return _device


class DeviceBuffer(object):
    """Owned device allocation."""

    def __init__(self, nbytes):
        device()
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(lib().dqc_malloc(C.byref(self.ptr), self.nbytes), "dqc_malloc")

    @staticmethod
    def fromArray(arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(max(arr.nbytes, 1))
        if arr.nbytes:
            check(lib().dqc_memcpy_h2d(b.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes), "dqc_memcpy_h2d")
        b.dtype, b.count = arr.dtype, len(arr)
        return b

    @staticmethod
    def fromMmdb(path):
        """The payload of an mmdb column file, loaded by the runtime's C
        reader (dqc_mmdb_load) into a new device allocation."""
        device()
        b = DeviceBuffer.__new__(DeviceBuffer)
        b.ptr = C.c_void_p()
        n = C.c_uint64()
        check(lib().dqc_mmdb_load(path.encode(), C.byref(b.ptr), C.byref(n)), "dqc_mmdb_load")
        b.nbytes = int(n.value)
        return b

    def toArray(self, dtype, count, offset=0):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            src = C.c_void_p(self.ptr.value + offset)
            check(lib().dqc_memcpy_d2h(out.ctypes.data_as(C.c_void_p), src, out.nbytes), "dqc_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().dqc_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copyToHost(ptr, dtype, count):
    out = np.empty(count, dtype=dtype)
    if out.nbytes:
        check(lib().dqc_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes), "dqc_memcpy_d2h")
    return out


class Stream(object):

    def __init__(self):
        device()
        self.handle = C.c_void_p()
        check(lib().dqc_stream_create(C.byref(self.handle)), "dqc_stream_create")

    def sync(self):
        check(lib().dqc_stream_sync(self.handle), "dqc_stream_sync")

    def __del__(self):
        try:
            if self.handle:
                lib().dqc_stream_destroy(self.handle)
        except Exception:
            pass


class Event(object):

    def __init__(self):
        self.handle = C.c_void_p()
        check(lib().dqc_event_create(C.byref(self.handle)), "dqc_event_create")

    def record(self, stream):
        check(lib().dqc_event_record(self.handle, stream.handle), "dqc_event_record")

    def elapsed(self, other):
        ms = C.c_float()
        check(lib().dqc_event_elapsed_ms(self.handle, other.handle, C.byref(ms)), "dqc_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().dqc_event_destroy(self.handle)
        except Exception:
            pass


CTYPE_DTYPE = {"int32_t": np.int32, "uint32_t": np.uint32, "int64_t": np.int64, "uint64_t": np.uint64,
               "float": np.float32, "double": np.float64, "int8_t": np.int8, "char": np.uint8,
               "dqc::str_offs": np.dtype([("start", np.uint64), ("end", np.uint64)]),
               "dqc::str_t": np.dtype([("start", np.uint64), ("end", np.uint64)])}


class PlanModule(object):
    """A loaded plan module."""

    def __init__(self, path):
        device()
        self.path = path
        self.handle = C.c_void_p()
        check(lib().dqc_module_load(path.encode(), C.byref(self.handle)), "dqc_module_load")
        self.desc = json.loads(lib().dqc_module_describe(self.handle).decode())

    def steps(self):
        return self.desc["steps"]

    def __del__(self):
        try:
            if self.handle:
                lib().dqc_module_unload(self.handle)
        except Exception:
            pass


class BoundInputs(object):
    """Device-resident inputs of one plan: columns uploaded from an accessor
    (``column(table, attr)`` -> array or (offsets, chars)) or given device
    pointers.  Uploaded buffers are cached per (table, attr) in ``store``."""

    def __init__(self, module, accessor, store=None, rows=None):
        self.module = module
        self.store = {} if store is None else store
        desc = module.desc
        self.rows = np.zeros(max(1, len(desc["tables"])), dtype=np.int64)
        for i, t in enumerate(desc["tables"]):
            self.rows[i] = accessor.schema[t]["size"] if rows is None else rows[t]
        self.ptrs = (C.c_void_p * max(1, len(desc["columns"])))()
        for c in desc["columns"]:
            key = (c["table"], c["attr"], c["part"])
            if key not in self.store and hasattr(accessor, "file"):
                # an mmdb directory (dbio.Accessor): the C reader loads the file
                path = accessor.file(c["table"], c["attr"], c["part"])
                if os.path.exists(path):
                    self.store[key] = DeviceBuffer.fromMmdb(path)
            if key not in self.store:
                data = accessor.column(c["table"], c["attr"])
                if c["part"] == "offset":
                    data = data[0]
                elif c["part"] == "char":
                    data = data[1]
                self.store[key] = DeviceBuffer.fromArray(np.ascontiguousarray(data))
            self.ptrs[c["index"]] = self.store[key].ptr
        self.struct = Inputs(len(desc["tables"]), len(desc["columns"]),
                             self.rows.ctypes.data_as(C.POINTER(C.c_int64)),
                             C.cast(self.ptrs, C.POINTER(C.c_void_p)))


def newTuning():
    return Tuning()


class Execution(object):
    """Result of one plan run: owns the workspace until released."""

    def __init__(self, module, inputs, tuning, res):
        self.module, self.inputs, self.tuning, self.res = module, inputs, tuning, res
        self.owned = True

    @property
    def count(self):
        return int(self.res.status.result_count)

    def columns(self):
        """Result columns as an OrderedDict name -> numpy array."""
        out = OrderedDict()
        n = min(self.count, int(self.res.layout.result_capacity))
        for i, c in enumerate(self.module.desc["result"]):
            dt = CTYPE_DTYPE[c["ctype"]]
            out[c["name"]] = copyToHost(self.res.ws + self.res.layout.result_offset[i], dt, n)
        return out

    def tempCount(self, name):
        t = self._temp(name)
        return int(self.res.status.temp_count[t["index"]])

    def _temp(self, name):
        for t in self.module.desc["temps"]:
            if t["name"] == name:
                return t
        raise KeyError("no temp table %s in this plan" % name)

    def temp(self, name):
        """Columns of a temp table (createtemp) as an OrderedDict name ->
        numpy array.  String columns hold device (start, end) pointers into
        the input char heap."""
        t = self._temp(name)
        lay = self.res.layout
        n = min(self.tempCount(name), int(lay.temp_capacity[t["index"]]))
        out = OrderedDict()
        for c in t["columns"]:
            dt = CTYPE_DTYPE[c["ctype"]]
            out[c["name"]] = copyToHost(self.res.ws + lay.temp_offset[c["slot"]], dt, n)
        return out

    def profile(self):
        lay = self.res.layout
        if lay.nprofile == 0:
            return {}
        arr = copyToHost(self.res.ws + lay.profile_offset, np.uint64, lay.nprofile * PROFILE_BUCKETS)
        return {p: arr[i * PROFILE_BUCKETS:(i + 1) * PROFILE_BUCKETS]
                for i, p in enumerate(self.module.desc["profiles"])}

    def release(self):
        if self.res.ws and self.owned:
            lib().dqc_query_release(C.byref(self.res))
        self.res.ws = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Graph(object):
    """One plan enqueue captured into an executable hipGraph: every launch
    of the plan replays the whole kernel sequence with one call."""

    def __init__(self, module, inputs, tuning, ws, stream):
        self.handle = C.c_void_p()
        check(lib().dqc_graph_create(module.handle, C.byref(inputs.struct), C.byref(tuning), ws.ptr, ws.nbytes,
                                     stream.handle, C.byref(self.handle)), "dqc_graph_create")

    def launch(self, stream):
        check(lib().dqc_graph_launch(self.handle, stream.handle), "dqc_graph_launch")

    def destroy(self):
        if self.handle:
            lib().dqc_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def view(module, inputs, tuning, ws):
    """An Execution over a caller-owned workspace after an enqueue or graph
    launch has completed (the status is read from the workspace now; the
    workspace is not freed by the view)."""
    res = QueryResult()
    res.ws, res.ws_bytes = ws.ptr.value, ws.nbytes
    res.layout = layout(module, inputs, tuning)
    st = ws.toArray(np.uint8, C.sizeof(Status))
    C.memmove(C.addressof(res.status), st.ctypes.data, C.sizeof(Status))
    ex = Execution(module, inputs, tuning, res)
    ex.owned = False
    return ex


class Feedback(object):
    """Cardinality feedback: measured sizes replace the cfg estimates
    (selectivity, htSizeFactor, numgroups; relationalAlgebra.py:342-349,
    441-458, 564-574) on later runs of the same plan module.

    A run through run(..., feedback=fb) counts the keys every unique-join and
    group-by table really holds (tuning.census, dqc_plan_status.ht_fill);
    with the multimap entry totals, temp-table and result row counts and
    long-list queue lengths of the status block, learn() keeps them per
    module together with the size of the largest input table.  apply() sets
    explicit capacities for the next run, scaled by how much that table grew
    or shrank: LOAD slots per key for hash tables, the counts themselves for
    row buffers.  At the same input sizes the next run therefore needs no
    overflow retry and allocates what the data needs."""

    LOAD = 2.0

    def __init__(self):
        self.records = {}

    @staticmethod
    def key(module):
        return os.path.basename(module.path)

    def learn(self, module, inputs, execution):
        st = execution.res.status
        lay = execution.res.layout
        hts = module.desc.get("hts", [])
        self.records[self.key(module)] = {
            "rows": int(max(inputs.rows)) if len(inputs.rows) else 0,
            "kinds": [h["kind"] for h in hts],
            "fill": [int(st.ht_fill[h["index"]]) for h in hts],
            "mm": [int(st.mm_total[h["index"]]) for h in hts],
            "split": [int(st.split_count[h["index"]]) if lay.split_capacity[h["index"]] > 0 else 0 for h in hts],
            "temps": [int(st.temp_count[i]) for i in range(len(module.desc.get("temps", [])))],
            "result": int(st.result_count) if module.desc.get("result") else 0,
        }

    def apply(self, module, inputs, tuning):
        """Explicit capacities from the module's last learned run; returns
        False when there is none."""
        rec = self.records.get(self.key(module))
        if rec is None:
            return False
        now = int(max(inputs.rows)) if len(inputs.rows) else 0
        scale = now / float(rec["rows"]) if rec["rows"] > 0 else 1.0
        grow = lambda n: int(math.ceil(n * scale))
        for i, kind in enumerate(rec["kinds"]):
            if kind == "multi":
                # entries exactly; one bucket per entry bounds the bucket
                # lists by the keys' own duplicates
                if rec["mm"][i] > 0:
                    tuning.mm_entries[i] = max(1, grow(rec["mm"][i]))
                tuning.ht_capacity[i] = max(64, grow(rec["mm"][i]))
            else:
                tuning.ht_capacity[i] = max(64, int(math.ceil(rec["fill"][i] * scale * self.LOAD)))
            if rec["split"][i] > 0:
                tuning.split_capacity[i] = max(1, grow(rec["split"][i]))
        for i, n in enumerate(rec["temps"]):
            tuning.temp_capacity[i] = max(1, grow(n))
        if module.desc.get("result"):
            tuning.result_capacity = max(1, grow(rec["result"]))
        return True

    def save(self, path):
        with open(path, "w") as f:
            json.dump(self.records, f, indent=1, sort_keys=True)

    def load(self, path):
        with open(path) as f:
            self.records.update(json.load(f))


def run(module, inputs, tuning=None, stream=None, max_retries=8, feedback=None):
    """Run a plan to completion (retries on overflow update `tuning` in
    place).  With a Feedback, its capacities for this module are applied
    first and the run's measured sizes are learned afterwards."""
    tuning = tuning or Tuning()
    stream = stream or Stream()
    if feedback is not None:
        feedback.apply(module, inputs, tuning)
        tuning.census = 1
    res = QueryResult()
    check(lib().dqc_query_run(module.handle, C.byref(inputs.struct), C.byref(tuning), stream.handle,
                              max_retries, C.byref(res)), "dqc_query_run")
    ex = Execution(module, inputs, tuning, res)
    if feedback is not None:
        feedback.learn(module, inputs, ex)
    return ex


def runNoAlloc(module, inputs, outputs, out_rows, tuning=None, stream=None, max_retries=8):
    """Run a plan in the runtime's static workspace (dqc_query_run_noalloc):
    no device memory is allocated; the result rows are copied into
    `outputs`, one caller-owned DeviceBuffer per result column holding
    `out_rows` rows.  Returns the QueryResult (layout, status, retries)."""
    tuning = tuning or Tuning()
    stream = stream or Stream()
    cols = (C.c_void_p * max(1, len(outputs)))(*[b.ptr.value for b in outputs])
    res = QueryResult()
    check(lib().dqc_query_run_noalloc(module.handle, C.byref(inputs.struct), C.byref(tuning), stream.handle,
                                      max_retries, cols, out_rows, C.byref(res)), "dqc_query_run_noalloc")
    return res


def layout(module, inputs, tuning):
    lay = Layout()
    check(lib().dqc_query_layout(module.handle, C.byref(inputs.struct), C.byref(tuning), C.byref(lay)),
          "dqc_query_layout")
    return lay


def enqueue(module, inputs, tuning, ws, stream, events=None):
    evs = None
    if events is not None:
        evs = (C.c_void_p * len(events))(*[e.handle.value for e in events])
    check(lib().dqc_query_enqueue(module.handle, C.byref(inputs.struct), C.byref(tuning), ws.ptr, ws.nbytes,
                                  stream.handle, evs), "dqc_query_enqueue")
