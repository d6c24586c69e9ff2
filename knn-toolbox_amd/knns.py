"""Python binding of the in-tree HIP libraries (ctypes over the C ABI).

Two layers, mirroring the C headers:

* ``dropin`` / ``DropIn`` -- the reference's C API (include/cuknns.h): ``gpuknnsBitonic``,
  ``gpuknnsHeap``, ``gpuknnLSH``, ``knnsplan``, ``knnsexecute``, with knntype (float32, or float64 for the
  double library) row-major inputs and Q x k outputs holding indices as knntype values, exactly as
  the reference's C callers see them (src/knnTest.cu:93-100).
* ``Corpus`` -- the native engine (include/knns_native.h): resident corpus, int32 indices,
  precision modes, per-query k, all-kNN self exclusion, device pointers.

There is no CPU fallback: if the libraries are not built, or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_DIR = PKG / "lib"

(KNNS_OK, KNNS_ERR_INVALID, KNNS_ERR_NOMEM, KNNS_ERR_DEVICE, KNNS_ERR_UNSUPPORTED, KNNS_ERR_OVERFLOW,
 KNNS_ERR_TIMEOUT) = range(7)
KNNS_F32, KNNS_F64 = 0, 1
PREC = {"fp32": 0, "bf16": 1, "fp8": 2, "fp64": 3}
SELECT = {"auto": 0, "threshold": 1, "heap": 1, "bitonic": 2}
KNNS_CORPUS_DEVICE_PTR, KNNS_CORPUS_BF16, KNNS_CORPUS_FP8 = 0x1, 0x2, 0x4
KNNS_DEVICE_PTRS, KNNS_EXCLUDE_SELF, KNNS_DIST_F64, KNNS_KERNEL_TIMES, KNNS_NO_CERTIFY = 0x1, 0x2, 0x4, 0x8, 0x10

STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "NOMEM", 3: "DEVICE", 4: "UNSUPPORTED", 5: "OVERFLOW", 6: "TIMEOUT"}


class KnnsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"knns {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("select", ctypes.c_int), ("k", ctypes.c_int),
                ("shortlist", ctypes.c_int), ("flags", ctypes.c_uint), ("self_offset", ctypes.c_long),
                ("ratio", ctypes.c_int), ("slices", ctypes.c_int), ("stream", ctypes.c_void_p),
                ("qb", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("total_s", ctypes.c_double), ("scan_s", ctypes.c_double), ("select_s", ctypes.c_double),
                ("rounds", ctypes.c_long), ("slow_compactions", ctypes.c_long), ("candidates", ctypes.c_long),
                ("groups", ctypes.c_int), ("slices_max", ctypes.c_int), ("qb", ctypes.c_int),
                ("dc", ctypes.c_int), ("dp", ctypes.c_int), ("uncertified", ctypes.c_long),
                ("researched", ctypes.c_long), ("research_s", ctypes.c_double),
                ("shortlist", ctypes.c_long)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Times(ctypes.Structure):
    _fields_ = [("dst_time", ctypes.c_double), ("srch_time", ctypes.c_double), ("knn_time", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("max_k", ctypes.c_int), ("max_batch", ctypes.c_long),
                ("mem_budget", ctypes.c_long), ("lanes", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_long), ("queries", ctypes.c_long), ("batches", ctypes.c_long),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("mean_batch", ctypes.c_double), ("busy_s", ctypes.c_double), ("buckets", ctypes.c_long),
                ("max_batch", ctypes.c_long), ("workspace_bytes", ctypes.c_long), ("lanes", ctypes.c_long),
                ("overlapped", ctypes.c_long), ("device_s", ctypes.c_double), ("researched", ctypes.c_long)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_native = None
_dropin = {}


def lib_path(name: str) -> Path:
    return LIB_DIR / name


def native():
    """Load libknns_amd.so (raises if it was not built)."""
    global _native
    if _native is None:
        p = lib_path("libknns_amd.so")
        if not p.exists():
            raise KnnsError(KNNS_ERR_DEVICE, f"{p} not built; run knn-toolbox_amd/build.py")
        L = ctypes.CDLL(str(p))
        V, I, L_, U = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_uint
        L.knns_params_init.argtypes = [ctypes.POINTER(Params), I, I]
        L.knns_params_init.restype = None
        L.knns_corpus_create.argtypes = [ctypes.POINTER(V), V, I, L_, I, U]
        L.knns_corpus_create.restype = I
        L.knns_corpus_destroy.argtypes = [V]
        L.knns_corpus_destroy.restype = I
        L.knns_corpus_rows.argtypes = [V]
        L.knns_corpus_rows.restype = L_
        L.knns_corpus_dim.argtypes = [V]
        L.knns_corpus_dim.restype = I
        L.knns_search.argtypes = [V, V, I, L_, ctypes.POINTER(Params), V, V]
        L.knns_search.restype = I
        L.knns_search_multi.argtypes = [V, V, I, L_, V, ctypes.POINTER(Params), V, V]
        L.knns_search_multi.restype = I
        L.knns_search_batches.argtypes = [V, V, I, L_, V, V, V, ctypes.POINTER(Params), V, V]
        L.knns_search_batches.restype = I
        L.knns_search_noalloc.argtypes = [V, V, I, L_, ctypes.POINTER(Params), V, V]
        L.knns_search_noalloc.restype = I
        L.knns_search_rows.argtypes = [V, L_, L_, V, I, L_, ctypes.POINTER(Params), V, V]
        L.knns_search_rows.restype = I
        L.knns_search_once.argtypes = [V, I, L_, I, V, L_, ctypes.POINTER(Params), V, V]
        L.knns_search_once.restype = I
        L.knns_plan_describe.argtypes = [L_, I, L_, I, I, ctypes.POINTER(Stats)]
        L.knns_plan_describe.restype = I
        L.knns_last_error.argtypes = []
        L.knns_last_error.restype = ctypes.c_char_p
        L.knns_last_stats.argtypes = [ctypes.POINTER(Stats)]
        L.knns_last_stats.restype = I
        L.knns_device_count.argtypes = []
        L.knns_device_count.restype = I
        L.knns_release_workspace.argtypes = []
        L.knns_release_workspace.restype = None
        L.knns_set_tuning_table.argtypes = [ctypes.c_char_p]
        L.knns_set_tuning_table.restype = I
        SC, SS = ctypes.POINTER(ServerConfig), ctypes.POINTER(ServerStats)
        L.knns_server_create.argtypes = [ctypes.POINTER(V), V, SC]
        L.knns_server_create.restype = I
        L.knns_server_submit.argtypes = [V, V, I, L_, I, V, V, ctypes.POINTER(ctypes.c_ulonglong)]
        L.knns_server_submit.restype = I
        L.knns_server_wait.argtypes = [V, ctypes.c_ulonglong, ctypes.c_double]
        L.knns_server_wait.restype = I
        L.knns_server_stats_get.argtypes = [V, SS, I]
        L.knns_server_stats_get.restype = I
        L.knns_server_destroy.argtypes = [V]
        L.knns_server_destroy.restype = I
        _native = L
    return _native


def dropin(double: bool = False):
    """Load libcuknn.so (float knntype) or libcuknn_double.so."""
    if double not in _dropin:
        p = lib_path("libcuknn_double.so" if double else "libcuknn.so")
        if not p.exists():
            raise KnnsError(KNNS_ERR_DEVICE, f"{p} not built; run knn-toolbox_amd/build.py")
        L = ctypes.CDLL(str(p))
        kt = ctypes.c_void_p
        I = ctypes.c_int
        for fn in ("gpuknnsBitonic", "gpuknnsHeap"):
            f = getattr(L, fn)
            f.argtypes = [kt, kt, kt, kt, I, I, I, I, I]
            f.restype = ctypes.c_double
        ip = ctypes.c_void_p
        L.gpuknnLSH.argtypes = [kt, kt, kt, kt, kt, I, I, I, I, I, ip, ip, I, ip]
        L.gpuknnLSH.restype = ctypes.c_float
        L.knnsGetLastError.argtypes = []
        L.knnsGetLastError.restype = ctypes.c_char_p
        L.knnsGetLastTimes.argtypes = []
        L.knnsGetLastTimes.restype = Times
        L.knnsSetStageTimes.argtypes = [I]
        L.knnsSetStageTimes.restype = None
        _dropin[double] = L
    return _dropin[double]


def _err(status: int):
    raise KnnsError(status, native().knns_last_error().decode(errors="replace"))


def device_count() -> int:
    return native().knns_device_count()


def last_stats() -> dict:
    s = Stats()
    native().knns_last_stats(ctypes.byref(s))
    return s.as_dict()


def predicted_rank(shortlist: int, seen: float) -> int:
    """knns_predicted_rank: the rank that sets a certified search's threshold after the fraction
    `seen` of its rows (the shortlist itself at seen = 1)."""
    L = native()
    L.knns_predicted_rank.argtypes = [ctypes.c_int, ctypes.c_double]
    L.knns_predicted_rank.restype = ctypes.c_int
    return int(L.knns_predicted_rank(int(shortlist), float(seen)))


def plan_describe(m: int, d: int, n: int, k: int, precision: str = "fp32") -> dict:
    s = Stats()
    st = native().knns_plan_describe(m, d, n, k, PREC[precision], ctypes.byref(s))
    if st:
        _err(st)
    return s.as_dict()


def _ptr(a) -> int:
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor


class Corpus:
    """A corpus resident in device memory (native API)."""

    def __init__(self, data, bf16: bool = False, device_ptr: bool = False, m: int | None = None,
                 d: int | None = None, dtype=None, fp8: bool = False):
        L = native()
        if device_ptr:
            self.m, self.d = int(m), int(d)
            self.dtype = np.dtype(dtype)
            src = _ptr(data)
        else:
            data = np.ascontiguousarray(data)
            if data.dtype not in (np.float32, np.float64) or data.ndim != 2:
                raise KnnsError(KNNS_ERR_INVALID, "corpus must be a 2-D float32/float64 array")
            self.m, self.d = data.shape
            self.dtype = data.dtype
            src = data.ctypes.data
        flags = ((KNNS_CORPUS_BF16 if bf16 else 0) | (KNNS_CORPUS_FP8 if fp8 else 0) |
                 (KNNS_CORPUS_DEVICE_PTR if device_ptr else 0))
        h = ctypes.c_void_p()
        st = L.knns_corpus_create(ctypes.byref(h), src, KNNS_F64 if self.dtype == np.float64 else KNNS_F32,
                                  self.m, self.d, flags)
        if st:
            _err(st)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            native().knns_corpus_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _params(self, precision, k, select, shortlist, exclude_self, self_offset, ratio, slices, flags, qb=0):
        p = Params()
        native().knns_params_init(ctypes.byref(p), PREC[precision], int(k))
        p.select = SELECT[select]
        p.shortlist = int(shortlist)
        p.flags = flags | (KNNS_EXCLUDE_SELF if exclude_self else 0)
        p.self_offset = int(self_offset)
        p.ratio = int(ratio)
        p.slices = int(slices)
        p.qb = int(qb)
        return p

    def search(self, queries, k: int, precision: str = "fp32", select: str = "auto", shortlist: int = 0,
               exclude_self: bool = False, self_offset: int = 0, ratio: int = 0, slices: int = 0,
               dist_f64: bool = False, certify: bool = True):
        """Host arrays in, (dist (n,k), idx (n,k) int32) out, sorted by (distance, index).
        certify=False skips the shortlist certificate and its re-search (KNNS_NO_CERTIFY)."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        n = q.shape[0]
        out64 = self.dtype == np.float64 or dist_f64
        D = np.empty((n, k), dtype=np.float64 if out64 else np.float32)
        I = np.empty((n, k), dtype=np.int32)
        p = self._params(precision, k, select, shortlist, exclude_self, self_offset, ratio, slices,
                         (KNNS_DIST_F64 if dist_f64 else 0) | (0 if certify else KNNS_NO_CERTIFY))
        st = native().knns_search(self._h, q.ctypes.data, KNNS_F64 if self.dtype == np.float64 else KNNS_F32,
                                  n, ctypes.byref(p), D.ctypes.data, I.ctypes.data)
        if st:
            _err(st)
        return D, I

    def search_rows(self, queries, k: int, row_begin: int, row_end: int, precision: str = "fp32",
                    select: str = "auto", shortlist: int = 0):
        """search() over corpus rows [row_begin, row_end) only (knns_search_rows); indices are corpus
        rows."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        n = q.shape[0]
        out64 = self.dtype == np.float64
        D = np.empty((n, k), dtype=np.float64 if out64 else np.float32)
        I = np.empty((n, k), dtype=np.int32)
        p = self._params(precision, k, select, shortlist, False, 0, 0, 0, 0)
        st = native().knns_search_rows(self._h, row_begin, row_end, q.ctypes.data,
                                       KNNS_F64 if out64 else KNNS_F32, n, ctypes.byref(p), D.ctypes.data,
                                       I.ctypes.data)
        if st:
            _err(st)
        return D, I

    def search_device(self, q_ptr, n: int, k: int, dist_ptr, idx_ptr, precision: str = "fp32",
                      select: str = "auto", shortlist: int = 0, ratio: int = 0, slices: int = 0,
                      stream: int = 0, kernel_times: bool = False, qb: int = 0):
        """Device pointers (ints or torch tensors); synchronous. kernel_times: fill scan_s and
        select_s of last_stats() (one event per launch). qb: query blocks per scan wave."""
        p = self._params(precision, k, select, shortlist, False, 0, ratio, slices,
                         KNNS_DEVICE_PTRS | (KNNS_KERNEL_TIMES if kernel_times else 0), qb)
        p.stream = stream or None
        st = native().knns_search(self._h, _ptr(q_ptr) if not isinstance(q_ptr, int) else q_ptr,
                                  KNNS_F64 if self.dtype == np.float64 else KNNS_F32, n, ctypes.byref(p),
                                  _ptr(dist_ptr) if not isinstance(dist_ptr, int) else dist_ptr,
                                  _ptr(idx_ptr) if not isinstance(idx_ptr, int) else idx_ptr)
        if st:
            _err(st)

    def search_noalloc(self, q_ptr, n: int, k: int, dist_ptr, idx_ptr, precision: str = "fp32",
                       select: str = "auto", shortlist: int = 0, stream: int = 0):
        """knns_search_noalloc: device pointers (ints or torch tensors), no allocation; the
        workspace is the library's fixed device scratch array. Synchronous."""
        p = self._params(precision, k, select, shortlist, False, 0, 0, 0, KNNS_DEVICE_PTRS)
        p.stream = stream or None
        st = native().knns_search_noalloc(self._h, q_ptr if isinstance(q_ptr, int) else _ptr(q_ptr),
                                          KNNS_F64 if self.dtype == np.float64 else KNNS_F32, n, ctypes.byref(p),
                                          dist_ptr if isinstance(dist_ptr, int) else _ptr(dist_ptr),
                                          idx_ptr if isinstance(idx_ptr, int) else _ptr(idx_ptr))
        if st:
            _err(st)

    def search_multi(self, queries, kq, precision: str = "fp32", select: str = "auto", shortlist: int = 0):
        """Per-query k. Returns a list of (dist, idx) per query."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        kq = np.ascontiguousarray(kq, dtype=np.int32)
        n = q.shape[0]
        tot = int(kq.sum())
        out64 = self.dtype == np.float64
        D = np.empty(tot, dtype=np.float64 if out64 else np.float32)
        I = np.empty(tot, dtype=np.int32)
        p = self._params(precision, 0, select, shortlist, False, 0, 0, 0, 0)
        st = native().knns_search_multi(self._h, q.ctypes.data, KNNS_F64 if out64 else KNNS_F32, n,
                                        kq.ctypes.data, ctypes.byref(p), D.ctypes.data, I.ctypes.data)
        if st:
            _err(st)
        off = np.concatenate([[0], np.cumsum(kq)])
        return [(D[off[i]:off[i + 1]], I[off[i]:off[i + 1]]) for i in range(n)]


    def search_batches(self, queries, sizes, ks=None, self_rows=None, precision: str = "fp32",
                       select: str = "auto", shortlist: int = 0, k: int = 0):
        """knns_search_batches: the batches (sizes[b] consecutive rows of `queries` each, k = ks[b]
        or k, excluding corpus row self_rows[b] + j for query j when self_rows[b] >= 0) packed into
        one search. Returns (dist, idx) as flat arrays packed per query, and the offsets."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        nb = sizes.shape[0]
        if int(sizes.sum()) != q.shape[0]:
            raise KnnsError(KNNS_ERR_INVALID, f"batch sizes sum to {int(sizes.sum())}, not {q.shape[0]} queries")
        kk = np.full(nb, int(k), dtype=np.int32) if ks is None else np.ascontiguousarray(ks, dtype=np.int32)
        kq = np.repeat(kk, sizes)
        tot = int(kq.sum())
        out64 = self.dtype == np.float64
        D = np.empty(tot, dtype=np.float64 if out64 else np.float32)
        I = np.empty(tot, dtype=np.int32)
        sr = None if self_rows is None else np.ascontiguousarray(self_rows, dtype=np.int64)
        p = self._params(precision, int(k), select, shortlist, False, 0, 0, 0, 0)
        st = native().knns_search_batches(self._h, q.ctypes.data, KNNS_F64 if out64 else KNNS_F32, nb,
                                          sizes.ctypes.data, None if ks is None else kk.ctypes.data,
                                          None if sr is None else sr.ctypes.data, ctypes.byref(p),
                                          D.ctypes.data, I.ctypes.data)
        if st:
            _err(st)
        return D, I, np.concatenate([[0], np.cumsum(kq)])


# ------------------------------------------------------------------------------------------------
# drop-in API (the reference's C interface)
# ------------------------------------------------------------------------------------------------
def _dropin_call(fn_name: str, query, data, k: int, double: bool, num_streams: int = 1):
    L = dropin(double)
    kt = np.float64 if double else np.float32
    q = np.ascontiguousarray(query, dtype=kt)
    x = np.ascontiguousarray(data, dtype=kt)
    nq, d = q.shape
    m = x.shape[0]
    vals = np.empty((nq, k), dtype=kt)
    inds = np.empty((nq, k), dtype=kt)
    t = getattr(L, fn_name)(q.ctypes.data, x.ctypes.data, vals.ctypes.data, inds.ctypes.data,
                             m, nq, d, k, num_streams)
    if t >= float(np.finfo(np.float32).max):
        raise KnnsError(KNNS_ERR_DEVICE, L.knnsGetLastError().decode(errors="replace"))
    return vals, inds, t


def dropin_stage_times(enable: bool, double: bool = False) -> None:
    """knnsSetStageTimes: per-stage device times in knnsGetLastTimes (off by default)."""
    dropin(double).knnsSetStageTimes(1 if enable else 0)


def dropin_last_times(double: bool = False) -> dict:
    """knnsGetLastTimes of the drop-in library: {dst_time, srch_time, knn_time} in ms."""
    t = dropin(double).knnsGetLastTimes()
    return {"dst_time": t.dst_time, "srch_time": t.srch_time, "knn_time": t.knn_time}


def gpuknns_bitonic(query, data, k: int, double: bool = False, num_streams: int = 1):
    """gpuknnsBitonic(query, data, values, indices, objects, numQueries, attributes, k, numStreams)."""
    return _dropin_call("gpuknnsBitonic", query, data, k, double, num_streams)


def gpuknns_heap(query, data, k: int, double: bool = False, num_streams: int = 1):
    return _dropin_call("gpuknnsHeap", query, data, k, double, num_streams)


def gpuknn_lsh(query, data, k: int, bucket_sizes, query_offsets, query_sizes, double: bool = False,
               num_streams: int = 1):
    """gpuknnLSH(query, data, values, indices, dp, objects, numQueries, attributes, k, numStreams,
    bucketSize, query_offsets, c, query_sizes): the queries of cluster j (rows query_offsets[j] ..
    +query_sizes[j]) get their exact top-k within bucket j (the j-th run of bucket_sizes[j]
    consecutive corpus rows). Returns (values (nq, k), indices (nq, k) as knntype corpus rows,
    seconds); rows of queries in no cluster are left as zeros."""
    L = dropin(double)
    kt = np.float64 if double else np.float32
    q = np.ascontiguousarray(query, dtype=kt)
    x = np.ascontiguousarray(data, dtype=kt)
    bs = np.ascontiguousarray(bucket_sizes, dtype=np.int32)
    qo = np.ascontiguousarray(query_offsets, dtype=np.int32)
    qs = np.ascontiguousarray(query_sizes, dtype=np.int32)
    if not (bs.shape == qo.shape == qs.shape) or bs.ndim != 1:
        raise KnnsError(KNNS_ERR_INVALID, "bucket_sizes, query_offsets and query_sizes differ in length")
    nq, d = q.shape
    vals = np.zeros((nq, k), dtype=kt)
    inds = np.zeros((nq, k), dtype=kt)
    t = L.gpuknnLSH(q.ctypes.data, x.ctypes.data, vals.ctypes.data, inds.ctypes.data, None, x.shape[0], nq, d,
                    k, num_streams, bs.ctypes.data, qo.ctypes.data, len(bs), qs.ctypes.data)
    if t >= float(np.finfo(np.float32).max):
        raise KnnsError(KNNS_ERR_DEVICE, L.knnsGetLastError().decode(errors="replace"))
    return vals, inds, t


# ---- raw .bin files of the knnTest driver (src/knnTest.cu:17-48, demo/exportData.m,
# demo/importData.m) ----
def knntest_path(double: bool = False) -> Path:
    """The built knnTest (float knntype) or knnTestDouble executable."""
    return Path(__file__).resolve().parent / "bin" / ("knnTestDouble" if double else "knnTest")


def export_bin(path, rows, double: bool = False) -> None:
    """Write rows (n x d, row-major = MATLAB's d x n column-major) as raw knntype values."""
    np.ascontiguousarray(rows, dtype=np.float64 if double else np.float32).tofile(str(path))


def import_results(nq: int, k: int, double: bool = False, directory=".") -> tuple:
    """Read KNNdist.bin / KNNidx.bin written by knnTest: (dist (nq, k), idx (nq, k) as knntype)."""
    kt = np.float64 if double else np.float32
    d = Path(directory)
    dist = np.fromfile(str(d / "KNNdist.bin"), dtype=kt)
    idx = np.fromfile(str(d / "KNNidx.bin"), dtype=kt)
    if dist.size != nq * k or idx.size != nq * k:
        raise KnnsError(KNNS_ERR_INVALID, f"result files hold {dist.size}/{idx.size} values, expected {nq * k}")
    return dist.reshape(nq, k), idx.reshape(nq, k)


class Server:
    """Continuous-batching server over a resident corpus (knns_server_*): submit() returns a
    ticket at once, result() blocks until that request's (dist, idx) are ready."""

    def __init__(self, corpus: "Corpus", precision: str = "bf16", max_k: int = 100, max_batch: int = 4096,
                 mem_budget: int = 0, lanes: int = 0):
        self._corpus = corpus           # keeps the corpus alive
        self._dtype = corpus.dtype
        cfg = ServerConfig(PREC[precision], int(max_k), int(max_batch), int(mem_budget), int(lanes))
        h = ctypes.c_void_p()
        st = native().knns_server_create(ctypes.byref(h), corpus._h, ctypes.byref(cfg))
        if st:
            _err(st)
        self._h = h
        self._out = {}

    def submit(self, queries, k: int) -> int:
        q = np.ascontiguousarray(queries, dtype=self._dtype)
        n = q.shape[0]
        D = np.empty((n, k), dtype=self._dtype)
        I = np.empty((n, k), dtype=np.int32)
        t = ctypes.c_ulonglong()
        st = native().knns_server_submit(self._h, q.ctypes.data, KNNS_F64 if self._dtype == np.float64 else KNNS_F32,
                                         n, int(k), D.ctypes.data, I.ctypes.data, ctypes.byref(t))
        if st:
            _err(st)
        self._out[t.value] = (D, I)
        return t.value

    def result(self, ticket: int, timeout: float = -1.0):
        """Blocks until the request is done (timeout < 0: no limit). On a timeout the request
        stays live and its output arrays stay referenced here (the worker still writes them), so
        result() can be called again for the same ticket."""
        if ticket not in self._out:
            raise KnnsError(KNNS_ERR_INVALID, f"unknown ticket {ticket}")
        st = native().knns_server_wait(self._h, ticket, float(timeout))
        if st == KNNS_ERR_TIMEOUT:
            raise TimeoutError(f"request {ticket} not done after {timeout} s")
        D, I = self._out.pop(ticket)
        if st:
            _err(st)
        return D, I

    def stats(self, reset: bool = False) -> dict:
        s = ServerStats()
        st = native().knns_server_stats_get(self._h, ctypes.byref(s), 1 if reset else 0)
        if st:
            _err(st)
        return s.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            native().knns_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
