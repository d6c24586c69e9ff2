"""CPU checks of the C-ABI boundary: the built libraries load and export every function that
include/*.h declares, struct layouts match, and host-side logic (planner, argument checks) runs
without a GPU. No compute call is made here."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
INC = ROOT / "include"
LIB = ROOT / "knn-toolbox_amd" / "lib"


def declared_functions(header: Path) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
    text = re.sub(r"#[^\n]*", "", text)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{}]*\)\s*;", text)
    return sorted(set(names))


def _lib(name):
    p = LIB / name
    if not p.exists():
        pytest.fail(f"{p} missing: run python knn-toolbox_amd/build.py (or __graft_entry__.build())")
    return ctypes.CDLL(str(p))


def test_header_parser_sees_the_api():
    f = declared_functions(INC / "cuknns.h")
    assert {"gpuknnsBitonic", "gpuknnsHeap", "knnsplan", "knnsexecute"} <= set(f)
    g = declared_functions(INC / "knns_native.h")
    assert {"knns_corpus_create", "knns_search", "knns_last_error"} <= set(g)


@pytest.mark.parametrize("lib", ["libcuknn.so", "libcuknn_double.so"])
def test_dropin_exports(lib):
    L = _lib(lib)
    for fn in declared_functions(INC / "cuknns.h"):
        assert hasattr(L, fn), f"{lib} does not export {fn}"


def test_native_exports():
    L = _lib("libknns_amd.so")
    for fn in declared_functions(INC / "knns_native.h"):
        assert hasattr(L, fn), f"libknns_amd.so does not export {fn}"


def test_dropin_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB / "libcuknn.so")],
                         capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for fn in ("gpuknnsBitonic", "gpuknnsHeap", "knnsplan", "knnsexecute"):
        assert fn in syms


def test_plan_struct_layout_matches_header():
    # knnplan = {fn ptr, 5 ints} (cuknns.h:34-41): 8 + 5*4 -> 32 bytes with tail padding
    class KnnPlan(ctypes.Structure):
        _fields_ = [("pt2Function", ctypes.c_void_p), ("objects", ctypes.c_int), ("numQueries", ctypes.c_int),
                    ("dimentions", ctypes.c_int), ("k", ctypes.c_int), ("numStreams", ctypes.c_int)]
    L = _lib("libcuknn.so")
    L.knnsplan.argtypes = [ctypes.POINTER(KnnPlan), ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long]
    L.knnsplan.restype = None
    p = KnnPlan()
    L.knnsplan(ctypes.byref(p), 16384, 1024, 128, 16)  # table-driven: no GPU work
    assert (p.objects, p.numQueries, p.dimentions, p.k, p.numStreams) == (16384, 1024, 128, 16, 1)
    heap = ctypes.cast(L.gpuknnsHeap, ctypes.c_void_p).value
    bit = ctypes.cast(L.gpuknnsBitonic, ctypes.c_void_p).value
    assert p.pt2Function in (heap, bit)


def test_native_params_layout(knns_mod):
    p = knns_mod.Params()
    knns_mod.native().knns_params_init(ctypes.byref(p), 1, 7)
    assert (p.precision, p.k, p.select, p.shortlist, p.flags) == (1, 7, 0, 0, 0)
    # layout of knns_params / knns_stats as gcc lays them out from include/knns_native.h
    assert ctypes.sizeof(knns_mod.Params) == 56 and knns_mod.Params.qb.offset == 48
    assert ctypes.sizeof(knns_mod.Stats) == 104 and knns_mod.Stats.shortlist.offset == 96


def test_planner_shapes(knns_mod, builtin_plans):
    s = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")
    assert s["qb"] == 3 and s["dc"] == 128 and s["dp"] == 128
    assert s["groups"] == 108  # ceil(10000 / 96) rounded up to whole workgroups of 4 groups
    assert s["rounds"] >= 3 and 1 <= s["slices_max"] <= 128
    s = knns_mod.plan_describe(16384, 43, 1024, 16, "fp32")
    assert s["qb"] == 1 and s["dc"] == 64 and s["dp"] == 64 and s["groups"] == 32
    s = knns_mod.plan_describe(1000, 2048, 10, 8, "fp32")
    assert s["dc"] == 128 and s["dp"] == 2048


def test_native_argument_errors_without_gpu(knns_mod):
    L = knns_mod.native()
    h = ctypes.c_void_p()
    import numpy as np
    x = np.zeros((4, 4), np.float32)
    # invalid shapes are rejected before any device work
    assert L.knns_corpus_create(ctypes.byref(h), x.ctypes.data, 0, 0, 4, 0) == knns_mod.KNNS_ERR_INVALID
    assert b"out of range" in L.knns_last_error()
    assert L.knns_corpus_create(ctypes.byref(h), x.ctypes.data, 0, 4, 0, 0) == knns_mod.KNNS_ERR_INVALID
    assert L.knns_corpus_create(ctypes.byref(h), x.ctypes.data, 7, 4, 4, 0) == knns_mod.KNNS_ERR_INVALID
    assert L.knns_corpus_create(ctypes.byref(h), None, 0, 4, 4, 0) == knns_mod.KNNS_ERR_INVALID
    s = knns_mod.Stats()
    assert L.knns_plan_describe(10, 4, 1, 2000, 0, ctypes.byref(s)) == knns_mod.KNNS_ERR_INVALID
    p = knns_mod.Params()
    L.knns_params_init(ctypes.byref(p), 0, 4)
    assert L.knns_search_rows(None, 0, 4, x.ctypes.data, 0, 4, ctypes.byref(p), x.ctypes.data,
                              x.ctypes.data) == knns_mod.KNNS_ERR_INVALID
    assert b"null corpus" in L.knns_last_error()
    # the one-shot call takes host buffers only; both checks run before any device work
    p.flags = knns_mod.KNNS_DEVICE_PTRS
    assert L.knns_search_once(x.ctypes.data, 0, 4, 4, x.ctypes.data, 4, ctypes.byref(p), x.ctypes.data,
                              x.ctypes.data) == knns_mod.KNNS_ERR_INVALID
    assert b"host buffers" in L.knns_last_error()
    assert L.knns_search_once(x.ctypes.data, 0, 4, 4, x.ctypes.data, 4, None, x.ctypes.data,
                              x.ctypes.data) == knns_mod.KNNS_ERR_INVALID


def test_dropin_rejects_bad_sizes_without_gpu(knns_mod):
    import numpy as np
    L = knns_mod.dropin(False)
    q = np.zeros((2, 4), np.float32)
    out = np.zeros((2, 4), np.float32)
    t = L.gpuknnsHeap(q.ctypes.data, q.ctypes.data, out.ctypes.data, out.ctypes.data, 0, 2, 4, 4, 1)
    assert t >= float(np.finfo(np.float32).max)
    assert b"invalid sizes" in L.knnsGetLastError()


def test_bucketed_search_rejects_bad_arguments_without_gpu(knns_mod):
    import numpy as np
    L = knns_mod.dropin(False)
    q = np.zeros((6, 4), np.float32)
    x = np.zeros((10, 4), np.float32)
    out = np.zeros((6, 3), np.float32)

    def call(bs, qo, qs, objects=10, nq=6, c=None):
        bs, qo, qs = (np.asarray(a, np.int32) for a in (bs, qo, qs))
        return L.gpuknnLSH(q.ctypes.data, x.ctypes.data, out.ctypes.data, out.ctypes.data, None, objects, nq, 4,
                           3, 1, bs.ctypes.data, qo.ctypes.data, len(bs) if c is None else c, qs.ctypes.data)

    flt_max = float(np.finfo(np.float32).max)
    assert call([5, 6], [0, 3], [3, 3]) >= flt_max                 # buckets hold 11 rows of 10
    assert b"buckets hold 11 rows" in L.knnsGetLastError()
    assert call([5, 5], [0, 4], [3, 3]) >= flt_max                 # cluster 1's queries run past 6
    assert b"cluster 1" in L.knnsGetLastError()
    assert call([5, 0], [0, 3], [3, 3]) >= flt_max                 # queries with an empty bucket
    assert b"empty bucket" in L.knnsGetLastError()
    assert call([5, -1], [0, 3], [3, 0]) >= flt_max
    assert call([5, 5], [0, 3], [3, 3], c=0) >= flt_max
    assert b"invalid sizes" in L.knnsGetLastError()
