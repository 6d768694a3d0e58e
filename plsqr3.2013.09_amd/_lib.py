"""ctypes binding of libplsqr_hip.so (include/plsqr_hip.h).

The Python side is a thin host mirror for tests and benchmarks: every
computation happens in the C/HIP library.  Arrays cross the boundary as
numpy views of library-owned memory or as pointers into caller-owned numpy
arrays; nothing here falls back to a CPU implementation.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplsqr_hip.so")

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)

PLSQR_OK = 0
ERR_NAMES = {0: "OK", 1: "ARG", 2: "BOUNDS", 3: "IO", 4: "HIP", 5: "NOMEM", 6: "STATE",
             7: "FORMAT", 8: "NONFINITE"}

GEOM_OPPOSITE_FACES = 0
GEOM_TOP_SOURCES = 1

OP_DEFAULT = 0
OP_EXACT_ORDER = 1
OP_WITH_FP32 = 2

PREC_FP64 = 0
PREC_MIXED = 1
PRECOND_NONE = 0
PRECOND_COLSCALE = 1

TRACE_WIDTH = 8
TRACE_FIELDS = ("alpha", "beta", "rnorm", "test1", "test2", "anorm", "acond", "xnorm")


class PlsqrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"plsqr error {ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class SystemC(C.Structure):
    _fields_ = [
        ("m_k", C.c_int), ("m_d", C.c_int), ("n", C.c_int), ("owned", C.c_int),
        ("nnz_k", C.c_longlong), ("nnz_d", C.c_longlong),
        ("k_col_ptr", c_int_p), ("k_row_idx", c_int_p), ("k_val", c_double_p),
        ("d_row_ptr", c_int_p), ("d_col_idx", c_int_p), ("d_val", c_double_p),
        ("dc_col_ptr", c_int_p), ("dc_row_idx", c_int_p), ("dc_val", c_double_p),
        ("b", c_double_p), ("x_true", c_double_p),
    ]


class GenParamsC(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("n_rays", C.c_longlong), ("geometry", C.c_int), ("seed", C.c_ulonglong),
        ("sigma", C.c_double), ("lambda_", C.c_double),
        ("w_xx", C.c_double), ("w_yy", C.c_double), ("w_zz", C.c_double),
        ("n_blobs", C.c_int), ("blob_width", C.c_double),
    ]


class DampingSpecC(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_param", C.c_int),
                ("w", C.c_double * 7), ("cross", C.c_int)]


class GridC(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)]


class SolveOptsC(C.Structure):
    _fields_ = [
        ("itn", C.c_int), ("tol", C.c_double), ("precision", C.c_int),
        ("refine_every", C.c_int), ("damp", C.c_double),
        ("atol", C.c_double), ("btol", C.c_double), ("conlim", C.c_double),
        ("itnlim", C.c_int), ("precond", C.c_int), ("anorm_ref", C.c_double),
        ("refine_adaptive", C.c_int), ("precond_power", C.c_double),
    ]


class SolveOutC(C.Structure):
    _fields_ = [
        ("istop", C.c_int), ("itn", C.c_int),
        ("anorm", C.c_double), ("acond", C.c_double), ("rnorm", C.c_double),
        ("arnorm", C.c_double), ("xnorm", C.c_double), ("bnorm", C.c_double),
        ("arnorm_true", C.c_double), ("elapsed_ms", C.c_double),
        ("refinements", C.c_int), ("inner_total", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchTuningC(C.Structure):
    _fields_ = [
        ("rows_per_wave", C.c_int), ("nnz_per_lane", C.c_int), ("block_rows", C.c_int),
        ("block_cols", C.c_int), ("ms_rows", C.c_double), ("ms_cols", C.c_double),
        ("candidates", C.c_int), ("gather_kib", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BATCH_MAX_RHS = 4
BATCH_NO_TUNE = 1
BATCH_TUNE_MIXED = 2

APROD_REF_T = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                          c_double_p, c_double_p, C.c_void_p, C.c_int, C.c_int)

# name -> (restype, argtypes); every symbol include/plsqr_hip.h declares
SIGNATURES = {
    "plsqr_system_free": (None, [C.POINTER(SystemC)]),
    "plsqr_gen_system": (C.c_int, [C.POINTER(GenParamsC), C.POINTER(SystemC)]),
    "plsqr_gen_rhs": (C.c_int, [C.POINTER(SystemC), C.POINTER(GenParamsC), c_double_p, c_double_p]),
    "plsqr_last_global_error": (C.c_char_p, []),
    "plsqr_write_reference_files": (C.c_int, [C.POINTER(SystemC), C.c_char_p] + [C.c_char_p] * 7),
    "plsqr_read_reference_files": (C.c_int, [C.c_char_p] * 8 + [C.c_int] * 3 + [C.POINTER(SystemC)]),
    "plsqr_write_x": (C.c_int, [C.c_char_p, c_double_p, C.c_int]),
    "plsqr_read_x": (C.c_int, [C.c_char_p, c_double_p, C.c_int]),
    "plsqr_op_create": (C.c_int, [C.POINTER(SystemC), C.c_int, C.POINTER(C.c_void_p)]),
    "plsqr_op_create_ex": (C.c_int, [C.POINTER(SystemC), C.c_int, C.POINTER(GridC),
                                     C.POINTER(C.c_void_p)]),
    "plsqr_op_destroy": (None, [C.c_void_p]),
    "plsqr_last_error": (C.c_char_p, [C.c_void_p]),
    "plsqr_clear_error": (None, [C.c_void_p]),
    "plsqr_op_dims": (C.c_int, [C.c_void_p, c_int_p, c_int_p, c_int_p,
                                C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "plsqr_op_device_bytes": (C.c_size_t, [C.c_void_p]),
    "plsqr_op_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "plsqr_op_get_stream": (C.c_void_p, [C.c_void_p]),
    "plsqr_op_set_vectors": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "plsqr_op_get_vectors": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "plsqr_aprod_ref": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               c_double_p, c_double_p, C.c_void_p, C.c_int, C.c_int]),
    "plsqr_aprod": (None, [C.c_int, C.c_int, C.c_int, c_double_p, c_double_p, C.c_void_p]),
    "plsqr_aprod_split": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, c_double_p]),
    "plsqr_aprod_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_aprod_f32": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p]),
    "plsqr_kernel_rows_to_csc": (C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_int_p,
                                           c_double_p]),
    "plsqr_ker2plsqr3": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_char_p]),
    "plsqr_ker_ascii2bin": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
    "plsqr_gen_damping": (C.c_int, [C.POINTER(DampingSpecC), C.POINTER(SystemC)]),
    "plsqr_write_damping_files": (C.c_int, [C.POINTER(DampingSpecC), C.c_char_p] + [C.c_char_p] * 4),
    "plsqr_lsqr_solve_se": (C.c_int, [C.c_void_p, c_double_p, C.POINTER(SolveOptsC), c_double_p, c_double_p,
                                      C.POINTER(SolveOutC)]),
    "plsqr_aprod_device_f32": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_op_to_internal": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_op_to_external": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_aprod_device_internal": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_aprod_device_internal_f32": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_lsqr_solve": (C.c_int, [C.c_void_p, c_double_p, C.POINTER(SolveOptsC), c_double_p,
                                   c_double_p, C.POINTER(SolveOutC), c_double_p]),
    "plsqr_lsqr_solve_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveOptsC),
                                          C.c_void_p, C.c_void_p, C.POINTER(SolveOutC),
                                          c_double_p]),
    "plsqr_residual_norms": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p,
                                       c_double_p, c_double_p]),
    "plsqr_lsqr_solve_x0": (C.c_int, [C.c_void_p, c_double_p, c_double_p, C.POINTER(SolveOptsC),
                                      c_double_p, C.POINTER(SolveOutC), c_int_p]),
    "plsqr_lsqr_solve_x0_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.POINTER(SolveOptsC), C.c_void_p,
                                             C.POINTER(SolveOutC), c_int_p]),
    "plsqr_gen_field": (C.c_int, [C.POINTER(GenParamsC), c_double_p]),
    "plsqr_cnn_param_count": (C.c_int, []),
    "plsqr_cnn_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong,
                                   C.POINTER(C.c_void_p)]),
    "plsqr_cnn_destroy": (None, [C.c_void_p]),
    "plsqr_cnn_last_error": (C.c_char_p, [C.c_void_p]),
    "plsqr_cnn_device_bytes": (C.c_size_t, [C.c_void_p]),
    "plsqr_cnn_get_params": (C.c_int, [C.c_void_p, c_float_p]),
    "plsqr_cnn_set_params": (C.c_int, [C.c_void_p, c_float_p]),
    "plsqr_cnn_set_data": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p]),
    "plsqr_cnn_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_ulonglong,
                                  c_double_p]),
    "plsqr_cnn_predict": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p]),
    "plsqr_cnn_gradient": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, c_float_p,
                                     c_double_p]),
    "plsqr_cnn_conv_forward": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_float_p, c_float_p,
                                         c_float_p, C.c_int, c_float_p]),
    "plsqr_cnn_conv_backward": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_float_p, c_float_p,
                                          c_float_p, c_float_p, c_float_p, c_float_p]),
    "plsqr_cnn_conv_forward_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "plsqr_batch_create": (C.c_int, [C.POINTER(C.POINTER(SystemC)), C.POINTER(GridC), C.c_int,
                                     C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "plsqr_batch_destroy": (None, [C.c_void_p]),
    "plsqr_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "plsqr_batch_clear_error": (None, [C.c_void_p]),
    "plsqr_batch_dims": (C.c_int, [C.c_void_p, c_int_p, c_int_p, C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "plsqr_batch_device_bytes": (C.c_size_t, [C.c_void_p]),
    "plsqr_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(BatchTuningC)]),
    "plsqr_batch_get_tuning": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BatchTuningC)]),
    "plsqr_batch_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BatchTuningC)]),
    "plsqr_batch_aprod": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, c_double_p]),
    "plsqr_batch_solve": (C.c_int, [C.c_void_p, c_double_p, C.POINTER(SolveOptsC), c_double_p,
                                    C.POINTER(SolveOutC), c_double_p]),
    "plsqr_batch_set_anorm_refs": (C.c_int, [C.c_void_p, c_double_p]),
    "plsqr_batch_solve_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveOptsC),
                                           C.c_void_p, C.POINTER(SolveOutC), c_double_p]),
}

_lib = None


def load(path=None):
    """Load the built library (fails loudly when it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc, op=None):
    if rc != PLSQR_OK:
        lib = load()
        msg = lib.plsqr_last_error(op) if op else lib.plsqr_last_global_error()
        raise PlsqrError(rc, (msg or b"").decode(errors="replace"))


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_float_p)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


def _view(ptr, n, dtype, owner=None):
    """numpy view of library memory; the view keeps `owner` (the System) alive."""
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    ct = C.c_int if dtype == np.int32 else C.c_double
    buf = (ct * int(n)).from_address(C.addressof(ptr.contents))
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype)


class System:
    """A system in the reference's loaded layout (kernel CSC, damping CSR and CSC, b)."""

    def __init__(self, c_struct, keep=None, grid=None):
        self.c = c_struct
        self._keep = keep  # numpy arrays backing a caller-built struct
        self.grid = grid   # (nx, ny, nz) of the voxel grid the columns come from, if known

    @classmethod
    def from_arrays(cls, m_k, m_d, n, k_col_ptr, k_row_idx, k_val, d_row_ptr, d_col_idx, d_val,
                    dc_col_ptr, dc_row_idx, dc_val, b, x_true=None, grid=None):
        # copies: the inputs may be views of another System's library-owned memory
        arrs = [np.array(a, dtype=t, copy=True, order="C") for a, t in (
            (k_col_ptr, np.int32), (k_row_idx, np.int32), (k_val, np.float64),
            (d_row_ptr, np.int32), (d_col_idx, np.int32), (d_val, np.float64),
            (dc_col_ptr, np.int32), (dc_row_idx, np.int32), (dc_val, np.float64),
            (b, np.float64))]
        xt = None if x_true is None else np.array(x_true, dtype=np.float64, copy=True)
        s = SystemC()
        s.m_k, s.m_d, s.n, s.owned = int(m_k), int(m_d), int(n), 0
        s.nnz_k, s.nnz_d = int(arrs[0][-1]), int(arrs[3][-1])
        s.k_col_ptr, s.k_row_idx, s.k_val = iptr(arrs[0]), iptr(arrs[1]), dptr(arrs[2])
        s.d_row_ptr, s.d_col_idx, s.d_val = iptr(arrs[3]), iptr(arrs[4]), dptr(arrs[5])
        s.dc_col_ptr, s.dc_row_idx, s.dc_val = iptr(arrs[6]), iptr(arrs[7]), dptr(arrs[8])
        s.b = dptr(arrs[9])
        s.x_true = dptr(xt) if xt is not None else None
        return cls(s, keep=(arrs, xt), grid=grid)

    def __del__(self):
        try:
            if self.c.owned:
                load().plsqr_system_free(C.byref(self.c))
        except Exception:
            pass

    # numpy views (no copies)
    m_k = property(lambda self: self.c.m_k)
    m_d = property(lambda self: self.c.m_d)
    n = property(lambda self: self.c.n)
    nnz_k = property(lambda self: self.c.nnz_k)
    nnz_d = property(lambda self: self.c.nnz_d)
    k_col_ptr = property(lambda self: _view(self.c.k_col_ptr, self.n + 1, np.int32, self))
    k_row_idx = property(lambda self: _view(self.c.k_row_idx, self.nnz_k, np.int32, self))
    k_val = property(lambda self: _view(self.c.k_val, self.nnz_k, np.float64, self))
    d_row_ptr = property(lambda self: _view(self.c.d_row_ptr, self.m_d + 1, np.int32, self))
    d_col_idx = property(lambda self: _view(self.c.d_col_idx, self.nnz_d, np.int32, self))
    d_val = property(lambda self: _view(self.c.d_val, self.nnz_d, np.float64, self))
    dc_col_ptr = property(lambda self: _view(self.c.dc_col_ptr, self.n + 1, np.int32, self))
    dc_row_idx = property(lambda self: _view(self.c.dc_row_idx, self.nnz_d, np.int32, self))
    dc_val = property(lambda self: _view(self.c.dc_val, self.nnz_d, np.float64, self))
    b = property(lambda self: _view(self.c.b, self.m_k, np.float64, self))
    x_true = property(lambda self: _view(self.c.x_true, self.n, np.float64, self)
                      if self.c.x_true else None)

    def write_reference_files(self, directory, prefix=""):
        """Write the seven files the reference CLI reads; returns their names."""
        names = dict(ker_f=prefix + "ker.bin", ker_i=prefix + "ker.info",
                     damp_f=prefix + "damp_row_data.bin", damp_i=prefix + "damp_row_info.bin",
                     damp_f_bycol=prefix + "damp_col_data.bin",
                     damp_i_bycol=prefix + "damp_col_info.bin", b_k=prefix + "b.list")
        lib = load()
        _check(lib.plsqr_write_reference_files(
            C.byref(self.c), directory.encode(), *[names[k].encode() for k in (
                "ker_f", "ker_i", "damp_f", "damp_i", "damp_f_bycol", "damp_i_bycol", "b_k")]))
        return names

    @classmethod
    def read_reference_files(cls, directory, names, row_k, row_d, col):
        lib = load()
        s = SystemC()
        _check(lib.plsqr_read_reference_files(
            directory.encode(), *[names[k].encode() for k in (
                "ker_f", "ker_i", "damp_f", "damp_i", "damp_f_bycol", "damp_i_bycol", "b_k")],
            int(row_k), int(row_d), int(col), C.byref(s)))
        return cls(s)


def gen_params(nx, ny, nz, n_rays, geometry, seed, sigma, lam, w=1.0, n_blobs=5, blob_width=3.0):
    p = GenParamsC()
    p.nx, p.ny, p.nz = int(nx), int(ny), int(nz)
    p.n_rays, p.geometry, p.seed = int(n_rays), int(geometry), int(seed)
    p.sigma, p.lambda_ = float(sigma), float(lam)
    p.w_xx = p.w_yy = p.w_zz = float(w)
    p.n_blobs, p.blob_width = int(n_blobs), float(blob_width)
    return p


def generate(nx, ny, nz, n_rays, geometry, seed, sigma, lam, w=1.0, n_blobs=5, blob_width=3.0):
    lib = load()
    s = SystemC()
    p = gen_params(nx, ny, nz, n_rays, geometry, seed, sigma, lam, w, n_blobs, blob_width)
    _check(lib.plsqr_gen_system(C.byref(p), C.byref(s)))
    return System(s, grid=(int(nx), int(ny), int(nz)))


def gen_field(nx, ny, nz, seed, n_blobs=5, blob_width=3.0):
    """The generator's x_true blob field alone."""
    p = gen_params(nx, ny, nz, 1, 0, seed, 0.0, 0.0, 1.0, n_blobs, blob_width)
    x = np.empty(int(nx) * int(ny) * int(nz), np.float64)
    _check(load().plsqr_gen_field(C.byref(p), dptr(x)))
    return x


def gen_rhs(system, nx, ny, nz, seed, sigma, n_blobs=5, blob_width=3.0):
    """Fresh (b, x_true) for an existing matrix."""
    lib = load()
    p = gen_params(nx, ny, nz, 1, 0, seed, sigma, 0.0, 1.0, n_blobs, blob_width)
    b = np.empty(system.m_k, np.float64)
    xt = np.empty(system.n, np.float64)
    _check(lib.plsqr_gen_rhs(C.byref(system.c), C.byref(p), dptr(b), dptr(xt)))
    return b, xt


# Workloads named by their shapes (the measured configurations of the build).
CONFIGS = {
    # 16^3 grid, 20,000 rays between opposite faces, lambda 0.1, sigma 0.01, 100 iterations
    "grid16_rays20k": dict(nx=16, ny=16, nz=16, n_rays=20000, geometry=GEOM_OPPOSITE_FACES,
                           seed=2013, sigma=0.01, lam=0.1, itn=100),
    # 67^3 grid, 2,000,000 rays from the top face, lambda 0.05, sigma 0.02, 100 iterations
    "grid67_rays2m": dict(nx=67, ny=67, nz=67, n_rays=2_000_000, geometry=GEOM_TOP_SOURCES,
                          seed=2014, sigma=0.02, lam=0.05, itn=100),
}


# Batches of independent problems: grids of edge U{lo..hi} (cubic), rays =
# rays_per_col x columns between opposite faces, lambda 0.1 damping, n_rhs
# right-hand sides per problem from their own blob fields and noise levels.
BATCH_CONFIGS = {
    "batch24_grid20to48_rhs4": dict(n_problems=24, edge_lo=20, edge_hi=48, rays_per_col=8, n_rhs=4,
                                    geometry=GEOM_OPPOSITE_FACES, seed=2015, sigma_lo=0.01,
                                    sigma_hi=0.05, lam=0.1),
}


def batch_specs(name=None, **override):
    """Per-problem generation parameters of a batch workload (deterministic)."""
    cfg = dict(BATCH_CONFIGS[name]) if name else {}
    cfg.update(override)
    rng = np.random.default_rng(cfg["seed"])
    P, K = cfg["n_problems"], cfg["n_rhs"]
    edges = rng.integers(cfg["edge_lo"], cfg["edge_hi"] + 1, size=P)
    sigmas = rng.uniform(cfg["sigma_lo"], cfg["sigma_hi"], size=(P, K))
    specs = []
    for p in range(P):
        e = int(edges[p])
        sys_seed = int(cfg["seed"]) * 1000 + p
        specs.append(dict(grid=(e, e, e), n_rays=int(cfg["rays_per_col"]) * e ** 3,
                          geometry=cfg["geometry"], seed=sys_seed, lam=cfg["lam"],
                          rhs=[(sys_seed * 16 + k + 1, float(sigmas[p, k])) for k in range(K)]))
    return specs


def generate_batch(name=None, specs=None, **override):
    """(systems, B, X_true): B[p] and X_true[p] are (n_rhs, m_k) and (n_rhs, n) arrays."""
    if specs is None:
        specs = batch_specs(name, **override)
    systems, B, XT = [], [], []
    for sp in specs:
        nx, ny, nz = sp["grid"]
        s = generate(nx, ny, nz, sp["n_rays"], sp["geometry"], sp["seed"], sp["rhs"][0][1], sp["lam"])
        bs, xs = [], []
        for seed, sigma in sp["rhs"]:
            b, xt = gen_rhs(s, nx, ny, nz, seed, sigma)
            bs.append(b)
            xs.append(xt)
        systems.append(s)
        B.append(np.stack(bs))
        XT.append(np.stack(xs))
    return systems, B, XT


def generate_config(name, **override):
    cfg = dict(CONFIGS[name])
    cfg.update(override)
    cfg.pop("itn", None)
    return generate(**cfg)


def kernel_rows_to_csc(m_k, n, row_ptr, col1, val):
    """Kernel rows (row_ptr from 0, 1-based columns in any order within a row)
    to the kernel CSC on the GPU: (col_ptr, row_idx 0-based, val)."""
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col1 = np.ascontiguousarray(col1, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    nnz = int(row_ptr[-1]) if row_ptr.size else 0
    cp = np.empty(n + 1, np.int32)
    ri = np.empty(max(nnz, 1), np.int32)
    vo = np.empty(max(nnz, 1), np.float64)
    _check(load().plsqr_kernel_rows_to_csc(int(m_k), int(n), iptr(row_ptr), iptr(col1), dptr(val), iptr(cp),
                                           iptr(ri), dptr(vo)))
    return cp, ri[:nnz], vo[:nnz]


def ker2plsqr3(directory, list_file, n_col, out_bin, out_info):
    """The Ker2PLSQR3 step on the GPU: row files listed in directory/list_file
    to the column-sorted kernel files directory/out_bin and out_info."""
    _check(load().plsqr_ker2plsqr3(directory.encode(), list_file.encode(), int(n_col), out_bin.encode(),
                                   out_info.encode()))


def ker_ascii2bin(list_path, in_dir, out_dir, out_list=None):
    """ASCII kernel rows ("col ix iy iz value") to the 16-byte row files."""
    _check(load().plsqr_ker_ascii2bin(list_path.encode(), in_dir.encode(), out_dir.encode(),
                                      out_list.encode() if out_list else None))


def damping_spec(nx, ny, nz, n_param=1, weights=(1.0,) * 7, cross=True):
    """Damping of damping_binary.py: weights = identity, xx, yy, zz, xy, xz, yz."""
    d = DampingSpecC()
    d.nx, d.ny, d.nz, d.n_param, d.cross = int(nx), int(ny), int(nz), int(n_param), int(bool(cross))
    w = [float(v) for v in weights]
    assert len(w) == 7
    for k in range(7):
        d.w[k] = w[k]
    return d


def gen_damping(nx, ny, nz, n_param=1, weights=(1.0,) * 7, cross=True):
    """A System holding only the damping part (m_k = 0)."""
    s = SystemC()
    _check(load().plsqr_gen_damping(C.byref(damping_spec(nx, ny, nz, n_param, weights, cross)), C.byref(s)))
    return System(s)


def write_damping_files(directory, nx, ny, nz, n_param=1, weights=(1.0,) * 7, cross=True, prefix="damp"):
    """The four damping files (names as the reference tool writes them for
    the prefix); returns the names."""
    names = dict(damp_f=prefix + "_row_data.bin", damp_i=prefix + "_row_info.bin",
                 damp_f_bycol=prefix + "_col_data.bin", damp_i_bycol=prefix + "_col_info.bin")
    _check(load().plsqr_write_damping_files(C.byref(damping_spec(nx, ny, nz, n_param, weights, cross)),
                                            directory.encode(), *[names[k].encode() for k in (
                                                "damp_f", "damp_i", "damp_f_bycol", "damp_i_bycol")]))
    return names


def read_x(path, n):
    x = np.empty(n, np.float64)
    _check(load().plsqr_read_x(path.encode(), dptr(x), int(n)))
    return x


def write_x(path, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    _check(load().plsqr_write_x(path.encode(), dptr(x), int(x.size)))
