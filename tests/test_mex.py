"""MEX gateways (knn-toolbox_amd/mex/mexknns.cpp) and the toolbox semantics, without MATLAB.

The gateway source is built against tests/mex_shim/mex.h (a stand-in for the MEX API subset it
uses) into knn-toolbox_amd/lib/mex_test/*_shim.so; these tests build mxArrays through the shim
and call mexFunction. CPU tests cover the argument checks (all raised before any device work);
GPU tests compare the outputs of the four gateways with the FP64 oracle and apply the toolbox's
post-processing (knn-toolbox_amd/toolbox/gpuknn*.m: idx+1, stable column sort, first k rows)."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

LIBDIR = Path(__file__).resolve().parent.parent / "knn-toolbox_amd" / "lib" / "mex_test"
MX_DOUBLE, MX_SINGLE, MX_INT32 = 6, 7, 12
VARIANTS = {"mexknnsBitonic": np.float32, "mexknnsHeap": np.float32,
            "mexknnsDBitonic": np.float64, "mexknnsDHeap": np.float64}
_libs = {}


def gateway(name):
    if name not in _libs:
        p = LIBDIR / f"{name}_shim.so"
        if not p.exists():
            pytest.fail(f"{p} missing: run python knn-toolbox_amd/build.py")
        L = ctypes.CDLL(str(p))
        V = ctypes.c_void_p
        L.shim_matrix.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, V, ctypes.c_int]
        L.shim_matrix.restype = V
        L.shim_call.argtypes = [ctypes.c_int, ctypes.POINTER(V), ctypes.c_int, ctypes.POINTER(V)]
        L.shim_call.restype = ctypes.c_int
        L.shim_error_id.restype = ctypes.c_char_p
        L.shim_error_msg.restype = ctypes.c_char_p
        for f in ("mxGetM", "mxGetN"):
            getattr(L, f).argtypes = [V]
            getattr(L, f).restype = ctypes.c_size_t
        L.mxGetData.argtypes = [V]
        L.mxGetData.restype = V
        L.mxGetClassID.argtypes = [V]
        L.mxGetClassID.restype = ctypes.c_int
        L.shim_free_all.restype = None
        _libs[name] = L
    return _libs[name]


def mx(L, a, cls=None, complex_flag=0):
    """numpy rows (n x d, row-major) -> MATLAB d x n matrix (same bytes)."""
    a = np.ascontiguousarray(a)
    if a.ndim < 2:
        a = a.reshape(1, 1)
    if cls is None:
        cls = {np.dtype(np.float64): MX_DOUBLE, np.dtype(np.float32): MX_SINGLE, np.dtype(np.int32): MX_INT32}[a.dtype]
    return L.shim_matrix(cls, a.shape[1], a.shape[0], a.ctypes.data, complex_flag)


def call(name, nlhs, args):
    L = gateway(name)
    prhs = (ctypes.c_void_p * len(args))(*args)
    plhs = (ctypes.c_void_p * max(nlhs, 1))()
    st = L.shim_call(nlhs, plhs, len(args), prhs)
    if st:
        return L.shim_error_id().decode(), L.shim_error_msg().decode()
    outs = []
    for i in range(max(nlhs, 1)):
        a = plhs[i]
        m, n = L.mxGetM(a), L.mxGetN(a)
        dt = np.float64 if L.mxGetClassID(a) == MX_DOUBLE else np.float32
        buf = (ctypes.c_char * (m * n * np.dtype(dt).itemsize)).from_address(L.mxGetData(a)) if m * n else b""
        outs.append(np.frombuffer(bytes(buf), dtype=dt).reshape(n, m).copy())   # n columns of m
    return outs


def toolbox_finish(dist, idx, k):
    """gpuknnBitonic.m / gpuknnHeap.m after the MEX call (columns = queries; here rows)."""
    idx = idx + 1
    order = np.argsort(dist, axis=1, kind="stable")
    return np.take_along_axis(dist, order, axis=1)[:, :k], np.take_along_axis(idx, order, axis=1)[:, :k]


@pytest.mark.parametrize("name", list(VARIANTS))
def test_argument_errors(name):
    kt = VARIANTS[name]
    L = gateway(name)
    Q = np.zeros((3, 5), kt)
    X = np.zeros((10, 5), kt)
    k = np.array(4.0)
    s = np.array(1.0)
    assert call(name, 2, [mx(L, Q), mx(L, X), mx(L, k)])[0] == "knns:args"
    assert call(name, 6, [mx(L, Q), mx(L, X), mx(L, k), mx(L, s)])[0] == "knns:args"
    other = np.float64 if kt == np.float32 else np.float32
    assert call(name, 2, [mx(L, Q.astype(other)), mx(L, X), mx(L, k), mx(L, s)])[0] == "knns:class"
    assert call(name, 2, [mx(L, Q, complex_flag=1), mx(L, X), mx(L, k), mx(L, s)])[0] == "knns:class"
    assert call(name, 2, [mx(L, np.zeros((3, 4), kt)), mx(L, X), mx(L, k), mx(L, s)])[0] == "knns:dims"
    for bad in (0.0, 1025.0, 2.5):
        assert call(name, 2, [mx(L, Q), mx(L, X), mx(L, np.array(bad)), mx(L, s)])[0] == "knns:k"
    assert call(name, 2, [mx(L, Q), mx(L, X), mx(L, k), mx(L, np.array(0.0))])[0] == "knns:streams"
    L.shim_free_all()


@pytest.mark.parametrize("name", list(VARIANTS))
def test_no_queries_returns_empty_without_device_work(name):
    kt = VARIANTS[name]
    L = gateway(name)
    outs = call(name, 5, [mx(L, np.zeros((0, 5), kt)), mx(L, np.zeros((10, 5), kt)), mx(L, np.array(3.0)),
                          mx(L, np.array(1.0))])
    assert outs[0].shape == (0, 3) and outs[1].shape == (0, 3)
    assert [float(o.ravel()[0]) for o in outs[2:]] == [0.0, 0.0, 0.0]
    L.shim_free_all()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(VARIANTS))
def test_gateways_match_oracle(gpu, oracle_mod, name):
    kt = VARIANTS[name]
    L = gateway(name)
    m, n, d, k = 2000, 40, 24, 16
    print(f"\n[case] mex {name} m={m} n={n} d={d} k={k}", flush=True)
    R = oracle_mod.uniform(m, d, seed=51, dtype=kt)
    Q = oracle_mod.uniform(n, d, seed=52, dtype=kt)
    dist, idx, t, dt, st = call(name, 5, [mx(L, Q), mx(L, R), mx(L, np.array(float(k))), mx(L, np.array(4.0))])
    Do, Io = oracle_mod.exact_topk(Q, R, k)
    assert dist.shape == (n, k) and idx.shape == (n, k)    # k x q in MATLAB
    assert np.array_equal(idx.astype(np.int64), Io.astype(np.int64))
    assert np.array_equal(dist, Do.astype(kt))
    assert float(t[0, 0]) > 0 and float(dt[0, 0]) > 0 and float(st[0, 0]) > 0
    # the toolbox wrapper: 1-based indices, columns sorted (already sorted), first k rows
    D1, I1 = toolbox_finish(dist, idx, k)
    assert np.array_equal(I1.astype(np.int64), Io.astype(np.int64) + 1)
    assert np.array_equal(D1, Do.astype(kt))
    L.shim_free_all()
