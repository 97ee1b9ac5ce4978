"""The printed protocol of HipCompiler.execute() (runner.protocolText), read
back the way the reference's profiler reads a query log (query/profiler.py:
112-244: the <pN> blocks as "lanes, iters" CSV without the 0-lane row, the
<timing> lines holding kernel names, annotations and totalKernelTime)."""
import io
import re

import numpy as np

from dogqc_amd import planlib, runner


class _Module(object):
    def __init__(self, desc):
        self.desc = desc

    def steps(self):
        return self.desc["steps"]


def _tag(text, name):
    m = re.search(r"<%s>\n(.*?)</%s>" % (name, name), text, re.S)
    return m.group(1) if m else None


def test_protocol_blocks_read_back():
    hist = np.arange(65, dtype=np.uint32) * 3
    rows = {"k": np.array([1, 2, 3], dtype=np.int32), "s": np.array([b"x", b"yy", b"z"], dtype=object),
            "v": np.array([0.5, 1.25, 2.0])}
    times = [("dqc_init_kernel", 0.01), ("krnl_lineitem1 LR3 P5", 1.5), ("dqc_mm_scan_up", 0.02)]
    out = runner.protocolText(3, rows, {5: hist}, times, printRows=2)
    assert "Result: 3 tuples" in out and "[...]" in out
    block = _tag(out, "p5")
    parsed = np.loadtxt(io.StringIO(block), delimiter=",", dtype=np.int64)
    assert parsed.shape == (65, 2)
    assert np.array_equal(parsed[1:, 0], np.arange(1, 65)) and np.array_equal(parsed[:, 1], hist)
    timing = _tag(out, "timing").splitlines()
    kernels = [l for l in timing if any(t in l for t in ("krnl_", "scanMulti", "totalKernel"))]
    assert kernels[0].strip().startswith("krnl_lineitem1 LR3 P5:")
    assert re.match(r"^\s*totalKernelTime:\s+1\.530 ms$", kernels[-1])
    assert all(len(l.split(":")[0]) >= 32 for l in timing)


def test_step_labels_carry_annotations():
    comp = planlib.translate("c2_both")
    from dogqc_amd import plancache  # noqa: F401
    import json
    src = comp.source
    desc = json.loads(re.search(r'R"DQCJSON\((.*)\)DQCJSON"', src).group(1))
    labels = runner.stepLabels(_Module(desc))
    annotated = [l for l in labels if " " in l]
    assert annotated, labels
    assert any("LR3" in l for l in annotated) and any("PD" in l for l in annotated)
