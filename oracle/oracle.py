"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  See plsqr_oracle.c for what each
function restates (reference file:line) and what pins it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "plsqr3_ref_fp64")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
TRACE_WIDTH = 8


class OrcSys(C.Structure):
    _fields_ = [("m_k", C.c_int), ("m_d", C.c_int), ("n", C.c_int),
                ("k_col_ptr", ip), ("k_row_idx", ip), ("k_val", dp),
                ("d_row_ptr", ip), ("d_col_idx", ip), ("d_val", dp),
                ("dc_col_ptr", ip), ("dc_row_idx", ip), ("dc_val", dp)]


class OrcOut(C.Structure):
    _fields_ = [("istop", C.c_int), ("itn", C.c_int),
                ("anorm", C.c_double), ("acond", C.c_double), ("rnorm", C.c_double),
                ("arnorm", C.c_double), ("xnorm", C.c_double), ("bnorm", C.c_double),
                ("dxmax", C.c_double), ("maxdx", C.c_int), ("arnorm_true", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_spmv_csc.argtypes = [C.c_int, ip, ip, dp, dp, dp]
        lib.orc_spmv_csr_v2.argtypes = [C.c_int, ip, ip, dp, dp, C.c_int, dp]
        lib.orc_spmvT_csc_v2.argtypes = [C.c_int, ip, ip, dp, dp, C.c_int, dp]
        lib.orc_aprod.argtypes = [C.POINTER(OrcSys), C.c_int, dp, dp, dp]
        lib.orc_aprod.restype = C.c_int
        for f in ("orc_dnrm2", "orc_norm_v"):
            getattr(lib, f).argtypes = [C.c_int, dp]
            getattr(lib, f).restype = C.c_double
        lib.orc_norm_u.argtypes = [C.c_int, dp, C.c_int, dp]
        lib.orc_norm_u.restype = C.c_double
        lib.orc_d2norm.argtypes = [C.c_double, C.c_double]
        lib.orc_d2norm.restype = C.c_double
        lib.orc_dscal.argtypes = [C.c_int, C.c_double, dp]
        lib.orc_lsqr.argtypes = [C.POINTER(OrcSys), dp, dp, dp, dp, dp, dp, C.c_int, C.c_double,
                                 C.c_double, dp, C.POINTER(OrcOut)]
        lib.orc_lsqr.restype = C.c_int
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


class Oracle:
    """Holds contiguous copies of a system's arrays for the C oracle."""

    def __init__(self, system):
        self.lib = load()
        self.m_k, self.m_d, self.n = system.m_k, system.m_d, system.n
        self.arrs = [np.ascontiguousarray(a, dtype=t) for a, t in (
            (system.k_col_ptr, np.int32), (system.k_row_idx, np.int32), (system.k_val, np.float64),
            (system.d_row_ptr, np.int32), (system.d_col_idx, np.int32), (system.d_val, np.float64),
            (system.dc_col_ptr, np.int32), (system.dc_row_idx, np.int32), (system.dc_val, np.float64))]
        a = self.arrs
        s = OrcSys()
        s.m_k, s.m_d, s.n = self.m_k, self.m_d, self.n
        s.k_col_ptr, s.k_row_idx, s.k_val = _i(a[0]), _i(a[1]), _d(a[2])
        s.d_row_ptr, s.d_col_idx, s.d_val = _i(a[3]), _i(a[4]), _d(a[5])
        s.dc_col_ptr, s.dc_row_idx, s.dc_val = _i(a[6]), _i(a[7]), _d(a[8])
        self.s = s

    def aprod(self, mode, x, y):
        """In place on host arrays: mode 1 y += A x, mode 2 x += A^T y (y = [y_k; y_d])."""
        yk = np.ascontiguousarray(y[:self.m_k])
        yd = np.ascontiguousarray(y[self.m_k:])
        rc = self.lib.orc_aprod(C.byref(self.s), int(mode), _d(x), _d(yk), _d(yd))
        assert rc == 0
        y[:self.m_k] = yk
        y[self.m_k:] = yd

    def lsqr(self, b, itn, damp=0.0, tol=0.0, want_se=False, trace=True):
        yk = np.array(b, dtype=np.float64, copy=True)
        yd = np.zeros(self.m_d, np.float64)
        v = np.zeros(self.n)
        w = np.zeros(self.n)
        x = np.zeros(self.n)
        se = np.zeros(self.n) if want_se else None
        tr = np.full((itn + 1, TRACE_WIDTH), np.nan)
        out = OrcOut()
        rc = self.lib.orc_lsqr(C.byref(self.s), _d(yk), _d(yd), _d(v), _d(w), _d(x),
                               _d(se) if se is not None else None, int(itn), float(damp),
                               float(tol), _d(tr) if trace else None, C.byref(out))
        assert rc == 0
        res = dict(out.as_dict(), v=v, w=w, u=np.concatenate([yk, yd]))
        if se is not None:
            res["se"] = se
        return x, res, tr

    def dnrm2(self, v):
        v = np.ascontiguousarray(v, np.float64)
        return self.lib.orc_dnrm2(v.size, _d(v))

    def d2norm(self, a, b):
        return self.lib.orc_d2norm(float(a), float(b))
