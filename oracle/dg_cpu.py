"""ctypes binding of ``oracle/dg_cpu.c`` (TEST INFRASTRUCTURE ONLY).

The OpenMP C restatement of the fused right-hand sides is the CPU baseline
``bench.py`` times and a second checker of the fused operator forms; only
``tests/``, ``bench.py``'s ``cpu_baseline`` leg and ``__graft_entry__`` touch
it.  It takes the arrays a HIP plan takes (``discr.ref`` operators,
``packed_geometry``, ``packed_ext_map``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdg_cpu.so")
EQUATIONS = {"wave": 0, "euler": 1, "maxwell": 2}


def build() -> str:
    """Compile libdg_cpu.so with gcc and OpenMP (oracle/Makefile's recipe)."""
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           os.path.join(HERE, "dg_cpu.c"), "-o", LIB + ".tmp", "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode:
        raise RuntimeError(f"gcc failed:\n{proc.stderr}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        pp = ctypes.POINTER(ctypes.c_void_p)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        lib.dg_cpu_rhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                   dp, dp, dp, ip, dp, ip, ctypes.c_double, pp, pp]
        lib.dg_cpu_rhs.restype = ctypes.c_int
        lib.dg_cpu_set_threads.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


class CpuOperator:
    """One operator on one tetrahedral discretization, arrays held for reuse."""

    def __init__(self, discr, equation: str, param: float, threads: int | None = None):
        self.lib = load()
        if threads:
            self.lib.dg_cpu_set_threads(int(threads))
        ref = discr.ref
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)   # noqa: E731
        self.arrays = [f64(ref.d_weak), f64(ref.d_strong), f64(ref.lift),
                       np.ascontiguousarray(ref.face_nodes, dtype=np.int32),
                       f64(discr.packed_geometry()),
                       np.ascontiguousarray(discr.packed_ext_map(), dtype=np.int32)]
        self.eq = EQUATIONS[equation]
        self.param = float(param)
        self.np, self.nfp, self.k = discr.np, discr.nfp, discr.nelements

    def __call__(self, q: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty_like(q) if out is None else out
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        a = self.arrays
        qp = (ctypes.c_void_p * len(q))(*[f.ctypes.data for f in q])
        op = (ctypes.c_void_p * len(out))(*[f.ctypes.data for f in out])
        rc = self.lib.dg_cpu_rhs(self.eq, self.np, self.nfp, self.k,
                                 a[0].ctypes.data_as(dp), a[1].ctypes.data_as(dp),
                                 a[2].ctypes.data_as(dp), a[3].ctypes.data_as(ip),
                                 a[4].ctypes.data_as(dp), a[5].ctypes.data_as(ip),
                                 self.param, qp, op)
        if rc:
            raise ValueError("dg_cpu_rhs: bad arguments")
        return out
