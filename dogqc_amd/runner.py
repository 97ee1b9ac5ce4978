"""Execution of a compiled plan for ``HipCompiler.execute()``.

Binds the plan's input columns from the algebra context's database accessor,
runs the module through the runtime (device-resident inputs, retry on
overflow), and prints the reference's text protocol: ``Result: N tuples``,
up to ten sample rows and a ``<timing>`` block with one line per kernel
launch (codegen.py:153-156, cudaTranslator.py:607-643), plus ``<pN>``
lane-activity blocks when profilers are placed (cudaLaneActivityProfile.py),
and keeps it as ``<plan name>.log`` like the reference's execute().
"""
from collections import OrderedDict

import numpy as np

from dogqc_amd import runtime as rt


def decodeStrings(col, chars):
    out = np.empty(len(col), dtype=object)
    b = np.asarray(chars, dtype=np.uint8).tobytes()
    for i, (s, e) in enumerate(zip(col["start"], col["end"])):
        out[i] = b[int(s):int(e)]
    return out


def resultColumns(module, execution, accessor):
    """Result as OrderedDict name -> numpy array, strings decoded to bytes."""
    cols = execution.columns()
    desc = module.desc
    byIndex = {c["index"]: c for c in desc["columns"]}
    out = OrderedDict()
    for c, (name, arr) in zip(desc["result"], cols.items()):
        if c["ctype"] == "dqc::str_offs":
            src = byIndex[c["char_col"]]
            chars = accessor.column(src["table"], src["attr"])[1]
            out[name] = decodeStrings(arr, chars)
        else:
            out[name] = arr
    return out


def tempColumns(module, execution, accessor, name):
    """A temp table's columns, strings decoded to bytes (their device
    pointers are turned into offsets of the bound char heap)."""
    t = execution._temp(name)
    cols = execution.temp(name)
    byIndex = {c["index"]: c for c in module.desc["columns"]}
    out = OrderedDict()
    for c in t["columns"]:
        arr = cols[c["name"]]
        if c["ctype"] == "dqc::str_t":
            base = int(execution.inputs.ptrs[c["char_col"]])
            offs = np.zeros(len(arr), dtype=[("start", np.uint64), ("end", np.uint64)])
            offs["start"] = arr["start"] - np.uint64(base)
            offs["end"] = arr["end"] - np.uint64(base)
            src = byIndex[c["char_col"]]
            out[c["name"]] = decodeStrings(offs, accessor.column(src["table"], src["attr"])[1])
        else:
            out[c["name"]] = arr
    return out


# sizes measured by HipCompiler.execute() runs, applied to later executions
# of the same plan module in this process (runtime.Feedback)
FEEDBACK = rt.Feedback()


def timeSteps(module, inputs, tuning, stream, ws):
    """Per-launch milliseconds of one enqueue (events between launches)."""
    steps = module.steps()
    events = [rt.Event() for _ in range(len(steps) + 1)]
    rt.enqueue(module, inputs, tuning, ws, stream, events)
    stream.sync()
    return [(name, events[i].elapsed(events[i + 1])) for i, name in enumerate(steps)]


def stepLabels(module):
    """Timing-line labels of the launch steps: a pipeline kernel's name
    followed by its annotations (lane refill LRn, profile Pn, push-down PDn,
    ...), as the reference's Timestamp labels (code.py:21-26)."""
    notes = {p["name"]: p.get("annotations", []) for p in module.desc.get("pipelines", [])}
    return [" ".join([s] + list(notes.get(s, []))) for s in module.steps()]


def protocolText(count, rows, profiles, times, printRows=10):
    """The generated program's stdout in the reference's protocol:
    ``Result: N tuples`` and sample rows, one ``<pN>`` block per lane-activity
    profile ("lanes, iterations" lines; 0..64 lanes on a wave64), and the
    ``<timing>`` block of "%32s: %6.3f ms" lines with ``totalKernelTime``
    (codegen.py:98-105, 153-156; code.py:21-26; query/profiler.py reads the
    blocks back).  `rows` maps column -> array; `times` is [(label, ms)]."""
    lines = ["", "Result: %d tuples" % count]
    n = min(printRows, count)
    for i in range(n):
        parts = []
        for name, arr in rows.items():
            v = arr[i]
            if isinstance(v, bytes):
                s = v.decode("latin-1")
            elif isinstance(v, (float, np.floating)):
                s = "%15.2f" % v
            else:
                s = "%10d" % int(v)
            parts.append("%s: %s" % (name, s))
        lines.append("  ".join(parts))
    if count > printRows:
        lines.append("[...]")
    for p, hist in profiles.items():
        lines.append("<p%d>" % p)
        lines.extend("%d, %d" % (i, int(c)) for i, c in enumerate(hist))
        lines.append("</p%d>" % p)
        lines.append("")
    lines.append("<timing>")
    total = 0.0
    for name, ms in times:
        lines.append("%32s: %6.3f ms" % (name, ms))
        total += ms
    lines.append("%32s: %6.3f ms" % ("totalKernelTime", total))
    lines.append("</timing>")
    return "\n".join(lines) + "\n"


def executeAndPrint(compiler, printRows=10):
    acc = compiler.algebraContext.ctxt.database
    module = rt.PlanModule(compiler.module)
    inputs = rt.BoundInputs(module, acc)
    tuning = rt.newTuning()
    if module.desc["profiles"]:
        tuning.profile = 1
    stream = rt.Stream()
    ex = rt.run(module, inputs, tuning, stream, feedback=FEEDBACK)
    res = resultColumns(module, ex, acc)
    profiles = ex.profile()
    ws = rt.DeviceBuffer(ex.res.ws_bytes)
    times = timeSteps(module, inputs, tuning, stream, ws)
    labels = stepLabels(module)
    out = protocolText(ex.count, res, profiles, [(labels[i], ms) for i, (_, ms) in enumerate(times)], printRows)
    print(out)
    # the reference's execute() keeps the output as <plan>.log for
    # query/profiler.py (codegen.py:209-218)
    name = getattr(compiler, "planName", None) or "query"
    with open(name + ".log", "w") as f:
        f.write(out)
    ex.release()
    return out, res
