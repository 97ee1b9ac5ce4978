"""The C ABI: libdqcrt.so and plan modules load and export every function
include/dogqc_amd.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from dogqc_amd import runtime as rt, planlib, build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dogqc_amd.h")
GEN_HEADER = os.path.join(ROOT, "include", "dogqc_amd_datagen.h")


def declared(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dqc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def runtimeLib():
    if not os.path.exists(rt.LIBPATH):
        build.buildRuntime()
    return ctypes.CDLL(rt.LIBPATH)


def test_header_lists_runtime_and_plan_abi():
    names = declared()
    assert set(rt.PLAN_SYMBOLS) <= set(names)
    assert set(rt.RUNTIME_SYMBOLS) <= set(names)
    assert set(names) == set(rt.PLAN_SYMBOLS) | set(rt.RUNTIME_SYMBOLS)


def test_runtime_exports(runtimeLib):
    for s in rt.RUNTIME_SYMBOLS:
        assert hasattr(runtimeLib, s), s


def test_datagen_header_and_exports():
    from dogqc_amd import datagen
    assert declared(GEN_HEADER) == sorted(rt.GEN_SYMBOLS)
    if not os.path.exists(datagen.GEN_SO):
        build.buildGen()
    lib = ctypes.CDLL(datagen.GEN_SO)
    for s in rt.GEN_SYMBOLS:
        assert hasattr(lib, s), s


def test_serve_structs_match_header():
    # slots, max_retries, arena_bytes, bypass_ms, size_feedback, max_heavy, 4 bytes of
    # padding before heavy_bytes, max_batch, batch_ratio, batch_min_share
    assert ctypes.sizeof(rt.ServeConfig) == 4 + 4 + 8 + 4 + 4 + 4 + 4 + 8 + 4 * 2 + 8 == 56
    assert rt.ServeConfig.max_batch.offset == 40 and rt.ServeConfig.batch_ratio.offset == 44
    assert rt.ServeConfig.batch_min_share.offset == 48
    assert rt.ServeConfig.heavy_bytes.offset == 32
    assert ctypes.sizeof(rt.ServeRequest) == 8 + 24 + 8 + 8 + 4 + 4 + 8 * 4 + 4 + 4 == 96
    assert rt.ServeRequest.batch_module.offset == 56
    assert ctypes.sizeof(rt.ServeRecord) == 8 * 3 + 4 * 4 + 8 + 8 + 4 + 4 + ctypes.sizeof(rt.Layout) + \
        ctypes.sizeof(rt.Status)


def test_plan_module_exports_and_describe():
    comp = planlib.translate("c1")
    path = comp.compile("c1")
    lib = ctypes.CDLL(path)
    for s in rt.PLAN_SYMBOLS:
        assert hasattr(lib, s), s
    lib.dqc_plan_describe.restype = ctypes.c_char_p
    import json
    desc = json.loads(lib.dqc_plan_describe().decode())
    assert desc["tables"] == ["dim", "fact"]
    assert [c["attr"] for c in desc["columns"]] == ["d_key", "d_cat", "f_fk", "f_price", "f_disc", "f_date"]
    assert [c["name"] for c in desc["result"]] == ["d_key", "revenue"]


def test_struct_sizes_match_header():
    # the ctypes mirrors must have the C layout
    # (the same sizes are static_asserts in the runtime, which includes the header)
    assert ctypes.sizeof(rt.Tuning) == 4 * 32 * 4 + 4 * 32 + 8 * 32 * 2 + 8 + 8 * 8 + 8 * 32 + 4 * 32 + 4 + 4 * 7 \
        == 1640
    assert rt.Tuning.part_agg.offset == rt.Tuning.census.offset + 4
    assert rt.Tuning.join_filter.offset == rt.Tuning.part_agg.offset + 4
    assert rt.Tuning.part_build.offset == rt.Tuning.join_filter.offset + 4
    assert ctypes.sizeof(rt.Status) == 16 + 8 + 8 * 32 + 8 * 8 + 8 * 32 + 8 * 32 == 856
    assert ctypes.sizeof(rt.Layout) == 8 + 4 + 4 + 8 * 32 * 2 + 8 + 4 + 4 + 8 * 64 + 8 + 4 + 4 + 8 * 8 + 8 * 64 \
        + 8 * 32 == 1904
