"""ctypes binding of the CPU oracle (liboracle.so) and of the reference-header
build (_ref/libppfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product never uses it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libppfref.so")

u32 = ctypes.c_uint32
f32 = ctypes.c_float
vp = ctypes.c_void_p


class OracleResult(ctypes.Structure):
    _fields_ = [("iter", u32), ("resid", f32), ("status", ctypes.c_int32),
                ("break_cause", ctypes.c_int32), ("break_iter", ctypes.c_int32),
                ("err0", f32)]


_lib = None
_ref = None


def load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB)
        lib.oracle_symm3x3.argtypes = [vp, vp, vp]
        lib.oracle_bj_invert.argtypes = [u32, vp, vp]
        lib.oracle_bj_invert.restype = u32
        lib.oracle_spmv.argtypes = [u32, vp, vp, vp, vp, vp, vp]
        lib.oracle_merged_diag.argtypes = [u32, vp, vp, vp, vp, vp]
        lib.oracle_pcg.argtypes = [u32, vp, vp, vp, vp, vp, vp, f32, u32, vp,
                                   ctypes.POINTER(OracleResult)]
        lib.oracle_pcg.restype = ctypes.c_int
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def load_ref():
    """The reference's own headers compiled host-side, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        lib = ctypes.CDLL(REF_LIB)
        lib.ref_symm3x3.argtypes = [vp, vp, vp]
        _ref = lib
    return _ref


def _p(a):
    return a.ctypes.data_as(vp)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def set_threads(n):
    load().oracle_set_threads(int(n))


def symm3x3(a9, lib=None):
    a9 = _f32(a9).reshape(-1, 9)
    val = np.zeros((len(a9), 3), np.float32)
    vec = np.zeros((len(a9), 9), np.float32)
    fn = (lib or load()).oracle_symm3x3 if lib is None else lib.ref_symm3x3
    for k in range(len(a9)):
        fn(_p(a9[k]), _p(val[k]), _p(vec[k]))
    return val, vec


def ref_symm3x3(a9):
    ref = load_ref()
    if ref is None:
        return None
    return symm3x3(a9, lib=ref)


def bj_invert(a9):
    a9 = _f32(a9).reshape(-1, 9)
    out = np.zeros_like(a9)
    bad = load().oracle_bj_invert(len(a9), _p(a9), _p(out))
    return out, int(bad)


def spmv(sys, x):
    x = _f32(x)
    y = np.zeros_like(x)
    ro, ci = _u32(sys["row_offset"]), _u32(sys["col_index"])
    val, cd = _f32(sys["value"]), _f32(sys["cdiag"])
    load().oracle_spmv(sys["n"], _p(ro), _p(ci), _p(val), _p(cd), _p(x), _p(y))
    return y


def pcg(sys, x0=None, tol=1e-3, max_iter=10000, hist=False):
    n = sys["n"]
    x = np.zeros(3 * n, np.float32) if x0 is None else _f32(x0).copy()
    ro, ci = _u32(sys["row_offset"]), _u32(sys["col_index"])
    val, cd, b = _f32(sys["value"]), _f32(sys["cdiag"]), _f32(sys["b"])
    h = np.zeros(max_iter, np.float32) if hist else None
    res = OracleResult()
    load().oracle_pcg(n, _p(ro), _p(ci), _p(val), _p(cd), _p(b), _p(x), tol,
                      max_iter, _p(h) if hist else None, ctypes.byref(res))
    out = dict(iter=res.iter, resid=res.resid, status=res.status,
               break_cause=res.break_cause, break_iter=res.break_iter,
               err0=res.err0, x=x)
    if hist:
        out["err_hist"] = h[:res.iter]
    return out
