"""MI355X-native LSQR hot path of PLSQR3 (host-side Python mirror).

The product is libplsqr_hip.so (C-ABI in include/plsqr_hip.h); this package
loads it with ctypes and mirrors the reference's operator interface:

    Operator.aprod(mode, x, y)   <- aprod(version, procRank, numProc, mode, m, n,
                                          x, y, UsrWrk, isFinalIteration, iteration)
                                    (source/aprod.h:10-13, source/lsqr.h:38-40)
    Operator.lsqr(b, itn, ...)   <- call_lsqr / lsqr (source/lsqr.c:34-184, 195-917)

Every call runs on the GPU through the library; when the library is missing,
load() raises instead of falling back to anything else.
"""
import ctypes as C

import numpy as np

from ._lib import (  # noqa: F401
    BATCH_CONFIGS, BATCH_MAX_RHS, BATCH_NO_TUNE, BATCH_TUNE_MIXED, BatchTuningC, GridC, CONFIGS,
    ERR_NAMES, GEOM_OPPOSITE_FACES, GEOM_TOP_SOURCES, LIB_PATH, OP_DEFAULT, OP_EXACT_ORDER,
    OP_WITH_FP32, PREC_FP64, PREC_MIXED, PRECOND_COLSCALE, PRECOND_NONE, SIGNATURES, TRACE_FIELDS,
    TRACE_WIDTH, PlsqrError,
    SolveOptsC, SolveOutC, SystemC, System, batch_specs, dptr, fptr, gen_field, gen_rhs, generate,
    kernel_rows_to_csc, ker2plsqr3, ker_ascii2bin, damping_spec, gen_damping, write_damping_files,
    DampingSpecC,
    generate_batch, generate_config, load, read_x, write_x, _check,
)


class Operator:
    """Device-resident operator A = [A_k ; D] of one system (plsqr_op)."""

    def __init__(self, system, flags=OP_DEFAULT, grid="auto"):
        """grid: (nx, ny, nz) of the voxel grid, None for none, "auto" = system.grid."""
        lib = load()
        h = C.c_void_p()
        if grid == "auto":
            grid = getattr(system, "grid", None)
        g = None
        if grid is not None:
            g = GridC()
            g.nx, g.ny, g.nz = (int(v) for v in grid)
        rc = lib.plsqr_op_create_ex(C.byref(system.c), int(flags), C.byref(g) if g else None,
                                    C.byref(h))
        if rc != 0:
            msg = lib.plsqr_last_error(h) if h.value else lib.plsqr_last_global_error()
            if h.value:
                lib.plsqr_op_destroy(h)
            raise PlsqrError(rc, (msg or b"").decode(errors="replace"))
        self.h = h
        self.lib = lib
        self.m_k, self.m_d, self.n = system.m_k, system.m_d, system.n
        self.m = self.m_k + self.m_d
        self.nnz_k, self.nnz_d = system.nnz_k, system.nnz_d
        self.flags = flags

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.plsqr_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self):
        return int(self.lib.plsqr_op_device_bytes(self.h))

    def last_error(self):
        return (self.lib.plsqr_last_error(self.h) or b"").decode(errors="replace")

    def _raise_if_error(self):
        msg = self.last_error()
        if msg:
            self.lib.plsqr_clear_error(self.h)
            raise PlsqrError(1, msg)

    # -- the callback boundary ------------------------------------------
    def aprod(self, mode, x, y):
        """mode 1: y += A x ; mode 2: x += A^T y   (in place, host float64 arrays)."""
        assert x.dtype == np.float64 and y.dtype == np.float64
        self.lib.plsqr_aprod_ref(0, 0, 1, int(mode), self.m, self.n, dptr(x), dptr(y),
                                 self.h, 0, 0)
        self._raise_if_error()

    def aprod_classic(self, mode, x, y):
        self.lib.plsqr_aprod(int(mode), self.m, self.n, dptr(x), dptr(y), self.h)
        self._raise_if_error()

    def aprod_split(self, mode, x, y_k, y_d):
        _check(self.lib.plsqr_aprod_split(self.h, int(mode), dptr(x), dptr(y_k), dptr(y_d)), self.h)

    def aprod_device(self, mode, x_ptr, y_ptr):
        """Device pointers (ints), enqueued on the handle's stream."""
        _check(self.lib.plsqr_aprod_device(self.h, int(mode), C.c_void_p(x_ptr), C.c_void_p(y_ptr)),
               self.h)

    def aprod_f32(self, mode, x, y):
        """The mixed-precision solver's operator: FP32 values and vectors with
        dot products accumulated in FP64; in place on host float32 arrays."""
        assert x.dtype == np.float32 and y.dtype == np.float32 and x.size == self.n and y.size == self.m
        _check(self.lib.plsqr_aprod_f32(self.h, int(mode), fptr(x), fptr(y)), self.h)

    def aprod_device_f32(self, mode, x_ptr, y_ptr):
        """FP32 device pointers (ints), enqueued on the handle's stream."""
        _check(self.lib.plsqr_aprod_device_f32(self.h, int(mode), C.c_void_p(x_ptr), C.c_void_p(y_ptr)),
               self.h)

    def to_internal(self, space, ext_ptr, int_ptr):
        """Device vector (ints: pointers) from the caller's order to the
        handle's; space 0: columns (n), 1: rows (m)."""
        _check(self.lib.plsqr_op_to_internal(self.h, int(space), C.c_void_p(ext_ptr), C.c_void_p(int_ptr)), self.h)

    def to_external(self, space, int_ptr, ext_ptr):
        _check(self.lib.plsqr_op_to_external(self.h, int(space), C.c_void_p(int_ptr), C.c_void_p(ext_ptr)), self.h)

    def aprod_device_internal(self, mode, x_ptr, y_ptr):
        """aprod on device vectors in the handle's order: the SpMV kernels only."""
        _check(self.lib.plsqr_aprod_device_internal(self.h, int(mode), C.c_void_p(x_ptr), C.c_void_p(y_ptr)),
               self.h)

    def aprod_device_internal_f32(self, mode, x_ptr, y_ptr):
        _check(self.lib.plsqr_aprod_device_internal_f32(self.h, int(mode), C.c_void_p(x_ptr), C.c_void_p(y_ptr)),
               self.h)

    def set_vectors(self, u=None, v=None):
        _check(self.lib.plsqr_op_set_vectors(self.h, dptr(u) if u is not None else None,
                                             dptr(v) if v is not None else None), self.h)

    def get_vectors(self):
        u = np.empty(self.m, np.float64)
        v = np.empty(self.n, np.float64)
        _check(self.lib.plsqr_op_get_vectors(self.h, dptr(u), dptr(v)), self.h)
        return u, v

    def aprod_resident(self, mode):
        """aprod with NULL x/y: acts on the handle's resident u and v (PLSQR3 convention)."""
        self.lib.plsqr_aprod_ref(0, 0, 1, int(mode), self.m, self.n, None, None, self.h, 0, 0)
        self._raise_if_error()

    # -- LSQR -------------------------------------------------------------
    @staticmethod
    def make_opts(itn, tol=0.0, precision=PREC_FP64, refine_every=25, damp=0.0, precond=PRECOND_NONE,
                  anorm_ref=0.0, refine_adaptive=False, precond_power=0.0):
        o = SolveOptsC()
        o.itn, o.tol, o.precision, o.refine_every, o.damp = int(itn), float(tol), int(precision), \
            int(refine_every), float(damp)
        o.atol = o.btol = o.conlim = 0.0
        o.itnlim = 0
        o.precond, o.anorm_ref, o.refine_adaptive = int(precond), float(anorm_ref), int(bool(refine_adaptive))
        o.precond_power = float(precond_power)
        return o

    def lsqr(self, b, itn, tol=0.0, precision=PREC_FP64, refine_every=25, damp=0.0,
             trace=False, want_w=False, precond=PRECOND_NONE, anorm_ref=0.0, refine_adaptive=False,
             precond_power=0.0):
        """Returns (x, info dict[, trace array (itn+1, 8)])."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.size == self.m_k
        x = np.empty(self.n, np.float64)
        w = np.empty(self.n, np.float64) if want_w else None
        out = SolveOutC()
        tr = np.full((itn + 1, TRACE_WIDTH), np.nan) if trace else None
        o = self.make_opts(itn, tol, precision, refine_every, damp, precond, anorm_ref, refine_adaptive,
                           precond_power)
        _check(self.lib.plsqr_lsqr_solve(self.h, dptr(b), C.byref(o), dptr(x),
                                         dptr(w) if w is not None else None, C.byref(out),
                                         tr.ctypes.data_as(C.POINTER(C.c_double))
                                         if tr is not None else None), self.h)
        info = out.as_dict()
        if want_w:
            info["w"] = w
        return (x, info, tr) if trace else (x, info)

    def lsqr_se(self, b, itn, tol=0.0, damp=0.0):
        """Plain FP64 solve with the standard-error estimates: (x, se, info)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.size == self.m_k
        x = np.empty(self.n, np.float64)
        se = np.empty(self.n, np.float64)
        out = SolveOutC()
        o = self.make_opts(itn, tol, PREC_FP64, 25, damp)
        _check(self.lib.plsqr_lsqr_solve_se(self.h, dptr(b), C.byref(o), dptr(x), dptr(se), C.byref(out)), self.h)
        return x, se, out.as_dict()

    def lsqr_device(self, b_ptr, x_ptr, itn, tol=0.0, precision=PREC_FP64, refine_every=25,
                    damp=0.0, w_ptr=None, precond=PRECOND_NONE, anorm_ref=0.0, refine_adaptive=False,
                    precond_power=0.0):
        out = SolveOutC()
        o = self.make_opts(itn, tol, precision, refine_every, damp, precond, anorm_ref, refine_adaptive,
                           precond_power)
        _check(self.lib.plsqr_lsqr_solve_device(self.h, C.c_void_p(b_ptr), C.byref(o),
                                                C.c_void_p(x_ptr),
                                                C.c_void_p(w_ptr) if w_ptr else None,
                                                C.byref(out), None), self.h)
        return out.as_dict()

    def lsqr_x0(self, b, x0, itn, tol=0.0, precision=PREC_FP64, refine_every=25, damp=0.0,
                precond=PRECOND_NONE, anorm_ref=0.0, refine_adaptive=False,
                precond_power=0.0):
        """Warm start from x0 (None = zero start); returns (x, info) with info["used_x0"]."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        x = np.empty(self.n, np.float64)
        out = SolveOutC()
        used = C.c_int(0)
        o = self.make_opts(itn, tol, precision, refine_every, damp, precond, anorm_ref, refine_adaptive,
                           precond_power)
        _check(self.lib.plsqr_lsqr_solve_x0(self.h, dptr(b), dptr(x0), C.byref(o), dptr(x), C.byref(out),
                                            C.byref(used)), self.h)
        info = out.as_dict()
        info["used_x0"] = bool(used.value)
        return x, info

    def lsqr_x0_device(self, b_ptr, x0_ptr, x_ptr, itn, tol=0.0, precision=PREC_FP64, refine_every=25,
                       precond=PRECOND_NONE, anorm_ref=0.0, refine_adaptive=False,
                       precond_power=0.0):
        out = SolveOutC()
        used = C.c_int(0)
        o = self.make_opts(itn, tol, precision, refine_every, 0.0, precond, anorm_ref, refine_adaptive,
                           precond_power)
        _check(self.lib.plsqr_lsqr_solve_x0_device(self.h, C.c_void_p(b_ptr),
                                                   C.c_void_p(x0_ptr) if x0_ptr else None, C.byref(o),
                                                   C.c_void_p(x_ptr), C.byref(out), C.byref(used)), self.h)
        info = out.as_dict()
        info["used_x0"] = bool(used.value)
        return info

    def residual_norms(self, b, x):
        rn, an, af = C.c_double(), C.c_double(), C.c_double()
        _check(self.lib.plsqr_residual_norms(self.h, dptr(np.ascontiguousarray(b, np.float64)),
                                             dptr(np.ascontiguousarray(x, np.float64)),
                                             C.byref(rn), C.byref(an), C.byref(af)), self.h)
        return rn.value, an.value, af.value


class Batch:
    """Independent systems solved together (plsqr_batch), n_rhs right-hand
    sides each.  Vectors cross as lists with one (n_rhs, length) array per
    system; the library packs them system after system."""

    def __init__(self, systems, n_rhs, grids="auto", flags=0):
        lib = load()
        self.lib = lib
        self.systems = list(systems)
        P = len(self.systems)
        arr = (C.POINTER(SystemC) * P)(*[C.pointer(s.c) for s in self.systems])
        if grids == "auto":
            grids = [getattr(s, "grid", None) for s in self.systems]
        g = None
        if grids is not None:
            g = (GridC * P)()
            for i, gr in enumerate(grids):
                if gr is not None:
                    g[i].nx, g[i].ny, g[i].nz = (int(v) for v in gr)
        h = C.c_void_p()
        rc = lib.plsqr_batch_create(arr, g, P, int(n_rhs), int(flags), C.byref(h))
        if rc != 0:
            raise PlsqrError(rc, (lib.plsqr_batch_last_error(None) or b"").decode(errors="replace"))
        self.h = h
        self.P, self.K = P, int(n_rhs)
        self.m_k = [s.m_k for s in self.systems]
        self.m = [s.m_k + s.m_d for s in self.systems]
        self.n = [s.n for s in self.systems]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.plsqr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            msg = (self.lib.plsqr_batch_last_error(self.h) or b"").decode(errors="replace")
            self.lib.plsqr_batch_clear_error(self.h)
            raise PlsqrError(rc, msg)

    @property
    def device_bytes(self):
        return int(self.lib.plsqr_batch_device_bytes(self.h))

    def dims(self):
        p, k = C.c_int(), C.c_int()
        r, c, z = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._check(self.lib.plsqr_batch_dims(self.h, C.byref(p), C.byref(k), C.byref(r), C.byref(c),
                                              C.byref(z)))
        return dict(n_problems=p.value, n_rhs=k.value, rows=r.value, cols=c.value, nnz=z.value)

    def _pack(self, arrs, lengths):
        out = []
        for a, ln in zip(arrs, lengths):
            a = np.asarray(a, dtype=np.float64).reshape(self.K, ln)
            out.append(a.ravel())
        return np.ascontiguousarray(np.concatenate(out))

    def _unpack(self, flat, lengths):
        res, off = [], 0
        for ln in lengths:
            res.append(flat[off:off + self.K * ln].reshape(self.K, ln).copy())
            off += self.K * ln
        return res

    def aprod(self, mode, X, Y, precision=PREC_FP64):
        """mode 1: Y[p] += A_p X[p]; mode 2: X[p] += A_p^T Y[p] (lists of (K, len) arrays);
        returns the updated (X, Y)."""
        x = self._pack(X, self.n)
        y = self._pack(Y, self.m)
        self._check(self.lib.plsqr_batch_aprod(self.h, int(mode), int(precision), dptr(x), dptr(y)))
        return self._unpack(x, self.n), self._unpack(y, self.m)

    def set_anorm_refs(self, refs):
        """Per-(problem, rhs) anorm of a column-scaled solve's FP64 test2 ((P, K)
        array, e.g. [[info["anorm"] ...] ...] of an FP64 batch solve), or None."""
        if refs is None:
            self._check(self.lib.plsqr_batch_set_anorm_refs(self.h, None))
            return
        r = np.ascontiguousarray(refs, dtype=np.float64).reshape(self.P * self.K)
        self._check(self.lib.plsqr_batch_set_anorm_refs(self.h, dptr(r)))

    def solve(self, B, itn, tol=0.0, precision=PREC_FP64, refine_every=25, damp=0.0, precond=PRECOND_NONE,
              precond_power=0.0, refine_adaptive=False):
        """B: list of (K, m_k) arrays.  Returns (X list of (K, n), infos [p][k] dicts, total_ms)."""
        b = self._pack(B, self.m_k)
        x = np.empty(sum(self.n) * self.K, np.float64)
        out = (SolveOutC * (self.P * self.K))()
        tot = C.c_double()
        o = Operator.make_opts(itn, tol, precision, refine_every, damp, precond, precond_power=precond_power,
                               refine_adaptive=refine_adaptive)
        self._check(self.lib.plsqr_batch_solve(self.h, dptr(b), C.byref(o), dptr(x), out, C.byref(tot)))
        infos = [[out[p * self.K + k].as_dict() for k in range(self.K)] for p in range(self.P)]
        return self._unpack(x, self.n), infos, tot.value

    def solve_device(self, b_ptr, x_ptr, itn, tol=0.0, precision=PREC_FP64, refine_every=25,
                     damp=0.0, precond=PRECOND_NONE, precond_power=0.0, refine_adaptive=False):
        """b_ptr, x_ptr: device arrays in the library's batch layout.  Returns (infos, total_ms)."""
        out = (SolveOutC * (self.P * self.K))()
        tot = C.c_double()
        o = Operator.make_opts(itn, tol, precision, refine_every, damp, precond, precond_power=precond_power,
                               refine_adaptive=refine_adaptive)
        self._check(self.lib.plsqr_batch_solve_device(self.h, C.c_void_p(b_ptr), C.byref(o),
                                                      C.c_void_p(x_ptr), out, C.byref(tot)))
        infos = [[out[p * self.K + k].as_dict() for k in range(self.K)] for p in range(self.P)]
        return infos, tot.value

    def pack_b(self, B):
        return self._pack(B, self.m_k)

    def autotune(self, precision=PREC_FP64, reps=5):
        t = BatchTuningC()
        self._check(self.lib.plsqr_batch_autotune(self.h, int(precision), int(reps), C.byref(t)))
        return t.as_dict()

    def get_tuning(self, precision=PREC_FP64):
        t = BatchTuningC()
        self._check(self.lib.plsqr_batch_get_tuning(self.h, int(precision), C.byref(t)))
        return t.as_dict()

    def set_tuning(self, precision=PREC_FP64, rows_per_wave=4, nnz_per_lane=4, block_rows=256,
                   block_cols=256, gather_kib=32):
        t = BatchTuningC()
        t.rows_per_wave, t.nnz_per_lane, t.block_rows, t.block_cols, t.gather_kib = (
            int(rows_per_wave), int(nnz_per_lane), int(block_rows), int(block_cols), int(gather_kib))
        self._check(self.lib.plsqr_batch_set_tuning(self.h, int(precision), C.byref(t)))


# launch shapes the batch autotuner chooses among
BATCH_SHAPES = [dict(rows_per_wave=r, nnz_per_lane=u, block_rows=b, block_cols=256)
                for r in (16, 8, 4, 2) for b in (128, 256) for u in (2, 4)] + \
               [dict(rows_per_wave=4, nnz_per_lane=4, block_rows=256, block_cols=c) for c in (128, 512)] + \
               [dict(rows_per_wave=4, nnz_per_lane=4, block_rows=256, block_cols=c, gather_kib=64)
                for c in (128, 256, 512)]


class WarmStartNet:
    """The warm-start network (plsqr_cnn): A^T b on the voxel grid -> x0."""

    def __init__(self, grid, max_batch=4, seed=1):
        lib = load()
        self.lib = lib
        h = C.c_void_p()
        nx, ny, nz = (int(v) for v in grid)
        rc = lib.plsqr_cnn_create(nx, ny, nz, int(max_batch), int(seed), C.byref(h))
        if rc != 0:
            raise PlsqrError(rc, (lib.plsqr_cnn_last_error(None) or b"").decode(errors="replace"))
        self.h = h
        self.grid = (nx, ny, nz)
        self.V = nx * ny * nz
        self.max_batch = int(max_batch)
        self.n_params = int(lib.plsqr_cnn_param_count())

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.plsqr_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise PlsqrError(rc, (self.lib.plsqr_cnn_last_error(self.h) or b"").decode(errors="replace"))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(self.lib.plsqr_cnn_get_params(self.h, fptr(p)))
        return p

    def set_params(self, p):
        self._check(self.lib.plsqr_cnn_set_params(self.h, fptr(np.ascontiguousarray(p, np.float32))))

    def set_data(self, atb, xstar):
        atb = np.ascontiguousarray(atb, np.float64).reshape(-1, self.V)
        xstar = np.ascontiguousarray(xstar, np.float64).reshape(-1, self.V)
        assert atb.shape == xstar.shape
        self._check(self.lib.plsqr_cnn_set_data(self.h, atb.shape[0], dptr(atb), dptr(xstar)))

    def train(self, steps, batch, lr=1e-3, seed=7):
        loss = np.empty(max(int(steps), 1), np.float64)
        self._check(self.lib.plsqr_cnn_train(self.h, int(steps), int(batch), float(lr), int(seed), dptr(loss)))
        return loss[:int(steps)]

    def predict(self, atb):
        atb = np.ascontiguousarray(atb, np.float64).reshape(-1, self.V)
        x0 = np.empty_like(atb)
        self._check(self.lib.plsqr_cnn_predict(self.h, atb.shape[0], dptr(atb), dptr(x0)))
        return x0

    def gradient(self, atb, xstar):
        atb = np.ascontiguousarray(atb, np.float64).reshape(-1, self.V)
        xstar = np.ascontiguousarray(xstar, np.float64).reshape(-1, self.V)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        self._check(self.lib.plsqr_cnn_gradient(self.h, atb.shape[0], dptr(atb), dptr(xstar), fptr(g),
                                                C.byref(loss)))
        return loss.value, g


def conv_forward(grid, inp, w, bias=None, relu=False):
    """One 32 -> 32 layer through the MFMA kernel: inp (B, V, 32), w (27, 32, 32) [t][n][k]."""
    lib = load()
    nx, ny, nz = grid
    inp = np.ascontiguousarray(inp, np.float32)
    B = inp.shape[0]
    out = np.empty_like(inp)
    rc = lib.plsqr_cnn_conv_forward(nx, ny, nz, B, fptr(inp), fptr(np.ascontiguousarray(w, np.float32)),
                                    fptr(None if bias is None else np.ascontiguousarray(bias, np.float32)),
                                    int(bool(relu)), fptr(out))
    if rc:
        raise PlsqrError(rc, (lib.plsqr_cnn_last_error(None) or b"").decode(errors="replace"))
    return out


def conv_backward(grid, inp, w, dout):
    lib = load()
    nx, ny, nz = grid
    inp = np.ascontiguousarray(inp, np.float32)
    dout = np.ascontiguousarray(dout, np.float32)
    B = inp.shape[0]
    din = np.empty_like(inp)
    dw = np.empty((27, 32, 32), np.float32)
    db = np.empty(32, np.float32)
    rc = lib.plsqr_cnn_conv_backward(nx, ny, nz, B, fptr(inp), fptr(np.ascontiguousarray(w, np.float32)),
                                     fptr(dout), fptr(din), fptr(dw), fptr(db))
    if rc:
        raise PlsqrError(rc, (lib.plsqr_cnn_last_error(None) or b"").decode(errors="replace"))
    return din, dw, db
