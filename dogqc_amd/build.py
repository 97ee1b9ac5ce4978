"""Ahead-of-time build: the runtime library and every named plan module.

``buildAll()`` compiles ``csrc/dqc_runtime.hip``, ``csrc/dqc_serve.hip`` and
``csrc/dqc_noalloc.hip``
into ``libdqcrt.so``, ``csrc/dqc_gen.hip`` into ``libdqcgen.so``,
``csrc/dqc_import.cpp`` into ``libdqcimport.so`` (host code) and
every plan of ``planlib.PLANS`` into ``_plans/`` (in parallel), always from
source, for gfx950.
"""
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

from dogqc_amd import plancache

PKG = plancache.PKG
RUNTIME_SRCS = [os.path.join(PKG, "csrc", f) for f in ("dqc_runtime.hip", "dqc_serve.hip", "dqc_noalloc.hip")]
RUNTIME_SO = os.path.join(PKG, "libdqcrt.so")


def buildRuntime():
    tmp = RUNTIME_SO + ".tmp"
    cmd = [plancache.HIPCC] + plancache.HIPCC_FLAGS + ["-I", plancache.INCLUDE] + RUNTIME_SRCS + \
        ["-o", tmp, "-ldl", "-lpthread"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if proc.returncode != 0:
        raise RuntimeError("building libdqcrt.so failed:\n" + proc.stderr[-6000:])
    os.replace(tmp, RUNTIME_SO)
    return RUNTIME_SO


GEN_SRC = os.path.join(PKG, "csrc", "dqc_gen.hip")
GEN_SO = os.path.join(PKG, "libdqcgen.so")


def buildGen():
    """libdqcgen.so: device-side generation of the benchmark tables."""
    tmp = GEN_SO + ".tmp"
    cmd = [plancache.HIPCC] + plancache.HIPCC_FLAGS + ["-I", plancache.INCLUDE, GEN_SRC, "-o", tmp]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if proc.returncode != 0:
        raise RuntimeError("building libdqcgen.so failed:\n" + proc.stderr[-6000:])
    os.replace(tmp, GEN_SO)
    return GEN_SO


IMPORT_SRC = os.path.join(PKG, "csrc", "dqc_import.cpp")
IMPORT_SO = os.path.join(PKG, "libdqcimport.so")


def buildImport():
    """libdqcimport.so: the .tbl importer (host code only, built with g++;
    -ffp-contract=off keeps the decimal parser's float steps unfused)."""
    tmp = IMPORT_SO + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-Wall",
           "-I", plancache.INCLUDE, IMPORT_SRC, "-o", tmp]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if proc.returncode != 0:
        raise RuntimeError("building libdqcimport.so failed:\n" + proc.stderr[-6000:])
    os.replace(tmp, IMPORT_SO)
    return IMPORT_SO


def buildPlans(names=None, jobs=None, force=False):
    from dogqc_amd import planlib
    names = list(planlib.PLANS) if names is None else names
    comps = [planlib.translate(n) for n in names]
    if force:
        for c, n in zip(comps, names):
            p = plancache.modulePath(c.source, n)
            if os.path.exists(p):
                os.unlink(p)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        paths = list(ex.map(lambda cn: cn[0].compile(cn[1]), zip(comps, names)))
    return dict(zip(names, paths))


def buildVariants(jobs=None):
    """Compile-time variants of the tunable plans (autotune.VARIANTS)."""
    from dogqc_amd import planlib, autotune
    work = []
    for name in planlib.TUNE_PLANS:
        for block, rows in autotune.VARIANTS:
            comp = autotune.translateVariant(name, block, rows)
            work.append((comp, autotune.variantName(name, block, rows)))
    for q in range(1, 23):
        for block, rows in autotune.TPCH_VARIANTS:
            name = "tpch%d" % q
            comp = autotune.translateVariant(name, block, rows)
            work.append((comp, autotune.variantName(name, block, rows)))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(lambda cn: cn[0].compile(cn[1]), work))


def cleanStale(keep):
    """Remove plan modules and sources no current plan refers to."""
    keep = set(os.path.basename(p)[:-3] for p in keep)
    if not os.path.isdir(plancache.PLANDIR):
        return
    for f in os.listdir(plancache.PLANDIR):
        stem = f.rsplit(".", 1)[0]
        if stem not in keep:
            os.unlink(os.path.join(plancache.PLANDIR, f))


def buildAll(verbose=True):
    t0 = time.time()
    buildRuntime()
    buildGen()
    buildImport()
    paths = buildPlans(force=True)
    variants = buildVariants()
    cleanStale(list(paths.values()) + variants)
    if verbose:
        print("built libdqcrt.so, libdqcgen.so, libdqcimport.so and %d plan modules in %.1f s" % (len(paths) + len(variants),
                                                                                 time.time() - t0),
              file=sys.stderr)
    return paths


if __name__ == "__main__":
    buildAll()
