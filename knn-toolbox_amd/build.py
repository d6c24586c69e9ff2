"""Build recipe for the in-tree HIP libraries (gfx950).

    python knn-toolbox_amd/build.py   # lib/libknns_amd.so, libcuknn.so, libcuknn_double.so, bin/knnTest{,Double}

Outputs go to knn-toolbox_amd/lib/ (git-ignored, shipped with the tree). Every HIP source is built
with -packed-fp32-ops disabled on the device side; the host pass warns that it does not know the
feature, which is harmless.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
INC = PKG.parent / "include"
ARCH = os.environ.get("KNNS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", str(INC), "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out and os.environ.get("KNNS_BUILD_VERBOSE"):
        print(out, file=sys.stderr)


def build(verbose: bool = False) -> list[Path]:
    LIB.mkdir(exist_ok=True)
    objs = []
    jobs = []
    for src in ["knns_kernels.hip", "knns_engine.hip", "knns_server.hip"]:
        obj = LIB / (Path(src).stem + ".o")
        # Kernels: NaN inputs are outside the supported domain, and without NaN semantics the
        # filter's max trees compile to bare v_max/v_max3 (no per-operand canonicalisation).
        extra = ["-fno-honor-nans"] if src == "knns_kernels.hip" else []
        jobs.append([HIPCC, *COMMON, *DEVICE_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    core = LIB / "libknns_amd.so"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(core)])
    outs = [core]
    for name, extra in [("libcuknn.so", []), ("libcuknn_double.so", ["-D__DOUBLE__"])]:
        so = LIB / name
        _run([HIPCC, *COMMON, *DEVICE_FLAGS, *extra, "-shared", str(CSRC / "cuknns_dropin.cpp"),
              "-L", str(LIB), "-lknns_amd", "-Wl,-rpath,$ORIGIN", "-o", str(so)])
        outs.append(so)
    # knnTest / knnTestDouble: the file-driven C driver (host code only) over the drop-in libraries
    BIN.mkdir(exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    for name, lib, extra in [("knnTest", "cuknn", []), ("knnTestDouble", "cuknn_double", ["-D__DOUBLE__"])]:
        exe = BIN / name
        _run([cxx, "-O2", "-std=c++17", "-Wall", "-I", str(INC), *extra, str(PKG / "cli" / "knnTest.cpp"),
              "-L", str(LIB), f"-l{lib}", "-Wl,-rpath,$ORIGIN/../lib", "-o", str(exe)])
        outs.append(exe)
    # MEX gateways (mex/mexknns.cpp): against the test stand-in of the MEX API always (driven by
    # tests/test_mex.py), and as real MEX files next to the toolbox .m files when MATLAB's mex or
    # Octave's mkoctfile is installed
    shim = PKG.parent / "tests" / "mex_shim"
    mexdir = LIB / "mex_test"
    mexdir.mkdir(exist_ok=True)
    variants = [("mexknnsBitonic", "cuknn", []), ("mexknnsHeap", "cuknn", ["-DKNNS_MEX_HEAP"]),
                ("mexknnsDBitonic", "cuknn_double", ["-D__DOUBLE__"]),
                ("mexknnsDHeap", "cuknn_double", ["-D__DOUBLE__", "-DKNNS_MEX_HEAP"])]
    if shim.exists():
        for name, lib, extra in variants:
            so = mexdir / f"{name}_shim.so"
            _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I", str(shim), "-I", str(INC), *extra,
                  str(PKG / "mex" / "mexknns.cpp"), str(shim / "mex_shim.cpp"), "-L", str(LIB), f"-l{lib}",
                  "-Wl,-rpath,$ORIGIN/..", "-o", str(so)])
            outs.append(so)
    mex = shutil.which("mex") or shutil.which("mkoctfile")
    if mex:
        for name, lib, extra in variants:
            octave = mex.endswith("mkoctfile")
            cmd = [mex] + (["--mex"] if octave else []) + [
                f"-I{INC}", *extra, str(PKG / "mex" / "mexknns.cpp"), f"-L{LIB}", f"-l{lib}",
                "-o" if octave else "-output", str(PKG / "toolbox" / name)]
            _run(cmd)
    for o in objs:
        o.unlink(missing_ok=True)
    if verbose:
        for o in outs:
            print(f"built {o}")
    return outs


if __name__ == "__main__":
    build(verbose=True)
