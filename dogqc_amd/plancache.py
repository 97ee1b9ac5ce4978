"""Compiled-plan cache: generated HIP source -> in-tree shared object.

The key is a hash of the generated source, the compiler flags and the device
runtime headers, so a module is rebuilt exactly when something it depends on
changes.  Modules live in ``dogqc_amd/_plans/`` next to the package (not in a
user cache directory), so modules built on a CPU-only machine travel with the
tree to the GPU machine.  Replaces the reference's write-<q>.cu-then-nvcc step
(codegen.py:170-192).
"""
import hashlib
import os
import subprocess
import sys
import tempfile
import time

PKG = os.path.dirname(os.path.abspath(__file__))
PLANDIR = os.path.join(PKG, "_plans")
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
HIPCC = os.environ.get("DQC_HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-add, so per-row arithmetic is
# bit-identical to the CPU oracle.  The packed-fp32 feature is disabled for
# all device code.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-function"]

_HEADERS = [os.path.join(CSRC, "dqc_device.h"), os.path.join(INCLUDE, "dogqc_amd.h")]


def _headerDigest():
    h = hashlib.sha256()
    for p in _HEADERS:
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def sourceKey(source, debug=False):
    h = hashlib.sha256()
    h.update(source.encode())
    h.update(" ".join(HIPCC_FLAGS + (["-g"] if debug else [])).encode())
    h.update(_headerDigest().encode())
    return h.hexdigest()[:20]


def _safeName(name):
    base = os.path.basename(str(name)) or "plan"
    return "".join(c if c.isalnum() or c in "_-" else "_" for c in base)[:48]


def modulePath(source, name, debug=False):
    return os.path.join(PLANDIR, "%s_%s.so" % (_safeName(name), sourceKey(source, debug)))


def compileSource(source, name, debug=False, verbose=False):
    """Return the path of the compiled module, building it if needed."""
    so = modulePath(source, name, debug)
    if os.path.exists(so):
        return so
    if os.environ.get("DQC_REQUIRE_PREBUILT") == "1":
        raise RuntimeError("plan module %s is not prebuilt (DQC_REQUIRE_PREBUILT=1)" % so)
    os.makedirs(PLANDIR, exist_ok=True)
    src = so[:-3] + ".hip"
    with open(src, "w") as f:
        f.write(source)
    fd, tmp = tempfile.mkstemp(suffix=".so", dir=PLANDIR)
    os.close(fd)
    cmd = [HIPCC] + HIPCC_FLAGS + (["-g"] if debug else []) + ["-I", CSRC, "-I", INCLUDE, src, "-o", tmp]
    t0 = time.time()
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if proc.returncode != 0:
        os.unlink(tmp)
        errs = "\n".join(l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, errs[-6000:]))
    os.replace(tmp, so)
    if verbose:
        print("compiled %s in %.1f s" % (os.path.basename(so), time.time() - t0), file=sys.stderr)
    return so
