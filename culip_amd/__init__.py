"""culip_amd: MI355X-native BILDLt preconditioner + SQMR for KKT systems.

Python mirror of the reference's C++ interfaces (dthuerck/culip):
  BlockiLDLt   libs/la/preconditioner/block_ildlt.cuh:188-311
  CSRMatrixSPMV, NormalMatrixSPMV, AugmentedMatrixSPMV   libs/la/spmv.cuh:36-149
  SQMR         libs/la/iterative_solvers/sqmr.cuh:21-72
  ImportLDUPreconditioner   libs/la/preconditioner/import_ldu_preconditioner.cuh:36-125
  read_matrix_csr / read_dense_vector  libs/test/test.h
over the C-ABI of libculip_amd.so. Device vectors are torch CUDA (HIP)
float64 tensors; torch only provides device memory and streams.
"""
import atexit
import ctypes as C
import itertools
import os
import weakref

import numpy as np

from . import _lib
from ._lib import (CulipError, PREC_F32, PREC_F64, SOLVE_LEFT, SOLVE_MIDDLE, SOLVE_RIGHT,
                   ORDER_NATURAL, ORDER_BLOCK_COLOR, TRSV_LEVELS, TRSV_SINGLE, TRSV_JACOBI,
                   check, lib)

__all__ = ["CulipError", "Batch", "KKT", "Prep", "BlockiLDLt", "CSRMatrixSPMV", "AugmentedMatrixSPMV",
           "SQMR", "ImportLDUPreconditioner", "NormalMatrixSPMV", "Server", "read_matrix_csr", "read_dense_vector", "csr_struct", "PREC_F32",
           "PREC_F64", "ORDER_NATURAL", "ORDER_BLOCK_COLOR", "permute_scale",
           "weighted_matching", "structural_rank", "btf_rank_bound"]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _solve_algorithm(a):
    """'single' | 'levels' | 'jacobi' (or the CULIP_TRSV_* value)."""
    return {"levels": TRSV_LEVELS, "single": TRSV_SINGLE, "jacobi": TRSV_JACOBI}.get(a, a)


def csr_struct(row_ptr, col, val, n=None):
    """Builds a Csr view; returns (struct, keepalive)."""
    rp, ci, cv = _i32(row_ptr), _i32(col), _f64(val)
    m = len(rp) - 1
    s = _lib.Csr(m, m if n is None else n, len(ci), _ptr(rp, C.c_int32), _ptr(ci, C.c_int32),
                 _ptr(cv, C.c_double))
    return s, (rp, ci, cv)


def _copy_csr(c):
    m, nnz = c.m, c.nnz
    rp = np.ctypeslib.as_array(c.row_ptr, shape=(m + 1,)).copy()
    ci = np.ctypeslib.as_array(c.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    cv = np.ctypeslib.as_array(c.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
    return rp, ci, cv


def _dptr(t):
    """Device pointer of a torch tensor (float64, contiguous)."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise CulipError(1, "device vectors must be CUDA (HIP) torch tensors")
    if t.dtype != torch.float64 or not t.is_contiguous():
        raise CulipError(1, "device vectors must be contiguous float64")
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(int(stream))


# Live library objects, by creation order. At interpreter exit they are
# released from an atexit hook, while the module, torch and the HIP runtime
# are all still intact (servers first, then solvers, preconditioners,
# operators and host objects; newer before older within a kind); after that
# every __del__ does nothing. A __del__ that ran later, during module
# teardown, could find the library's globals already cleared.
_live = weakref.WeakValueDictionary()
_serial = itertools.count()
_closing = False


class _Handle:
    """An object that owns a library handle in self._h; subclasses free it
    in _release()."""

    def _track(self):
        _live[next(_serial)] = self

    def __del__(self):
        if _closing is not False:
            return
        self._release()


def _release_all():
    global _closing
    objs = [o for _, o in sorted(_live.items(), key=lambda kv: -kv[0])]
    for kind in (Server, SQMR, BlockiLDLt, ImportLDUPreconditioner, _Spmv, Prep, KKT):
        for o in objs:
            if isinstance(o, kind):
                try:
                    o._release()
                except Exception:   # exit goes on; the process frees the rest
                    pass
    _closing = True


atexit.register(_release_all)


class KKT(_Handle):
    """Synthetic KKT system [[H, A'], [A, -delta I]] (host)."""

    def __init__(self, n, m, h_offdiag_per_row=4, a_nnz_per_col=3, h_diag=(1.0, 10.0),
                 h_diag_log=False, h_offdiag_std=0.1, delta=1e-8, seed=0, barrier=None):
        bl, bh = barrier if barrier else (0.0, 0.0)
        p = _lib.KktParams(n, m, h_offdiag_per_row, a_nnz_per_col, h_diag[0], h_diag[1],
                           1 if h_diag_log else 0, h_offdiag_std, delta, seed, bl, bh)
        h = C.c_void_p()
        check(lib().culip_kkt_generate(C.byref(p), C.byref(h)))
        self._h = h
        self._track()
        self.n, self.m = n, m
        c = _lib.Csr()
        rhs = C.POINTER(C.c_double)()
        check(lib().culip_kkt_view(h, C.byref(c), C.byref(rhs)))
        self.row_ptr, self.col, self.val = _copy_csr(c)
        self.N = c.m
        self.rhs = np.ctypeslib.as_array(rhs, shape=(c.m,)).copy()

    def barrier_values(self, w_lo=1e-4, w_hi=1e4, seed=0):
        out = np.empty(len(self.val))
        check(lib().culip_kkt_barrier_values(self._h, w_lo, w_hi, seed, _ptr(out, C.c_double)))
        return out

    def _release(self):
        if getattr(self, "_h", None):
            lib().culip_kkt_destroy(self._h)
            self._h = None


class Prep(_Handle):
    """Ordering, contiguous blocking and static 1x1 pivot starts."""

    def __init__(self, row_ptr, col, val, n_primal, max_block=32, ordering=ORDER_NATURAL):
        s, keep = csr_struct(row_ptr, col, val)
        h = C.c_void_p()
        check(lib().culip_prep_create(C.byref(s), n_primal, max_block, ordering, C.byref(h)))
        self._h = h
        self._track()
        c = _lib.Csr()
        perm = C.POINTER(C.c_int32)()
        nbk = C.c_int32()
        bs = C.POINTER(C.c_int32)()
        npv = C.c_int32()
        ps = C.POINTER(C.c_int32)()
        check(lib().culip_prep_view(h, C.byref(c), C.byref(perm), C.byref(nbk), C.byref(bs),
                                    C.byref(npv), C.byref(ps)))
        self.row_ptr, self.col, self.val = _copy_csr(c)
        self.perm = np.ctypeslib.as_array(perm, shape=(c.m,)).copy()
        self.block_starts = np.ctypeslib.as_array(bs, shape=(nbk.value,)).copy()
        self.piv_starts = np.ctypeslib.as_array(ps, shape=(npv.value,)).copy()

    def permute_values(self, vals):
        v = _f64(vals)
        out = np.empty_like(v)
        check(lib().culip_prep_permute_values(self._h, _ptr(v, C.c_double),
                                              _ptr(out, C.c_double)))
        return out

    def _release(self):
        if getattr(self, "_h", None):
            lib().culip_prep_destroy(self._h)
            self._h = None


class BlockiLDLt(_Handle):
    """BlockiLDLt<T, false, false> (static pivots); prec 'f64' or 'f32' = T."""

    def __init__(self, row_ptr, col, val, piv_starts, block_starts, prec="f64"):
        s, keep = csr_struct(row_ptr, col, val)
        ps, bs = _i32(piv_starts), _i32(block_starts)
        h = C.c_void_p()
        pr = PREC_F32 if prec in ("f32", PREC_F32) else PREC_F64
        check(lib().culip_bildlt_create(C.byref(s), len(ps), _ptr(ps, C.c_int32), len(bs),
                                        _ptr(bs, C.c_int32), pr, C.byref(h)))
        self._h = h
        self._track()
        self.N = len(row_ptr) - 1
        self.prec = "f32" if pr == PREC_F32 else "f64"

    def n(self):
        return lib().culip_bildlt_n(self._h)

    def symbolic(self, fill_level=0, fill_factor=1.0):
        check(lib().culip_bildlt_symbolic(self._h, fill_level, fill_factor))

    def compute(self, fill_level=0, fill_factor=1.0, threshold=1e-4, stream=None):
        check(lib().culip_bildlt_compute(self._h, fill_level, fill_factor, threshold,
                                         _stream(stream)))
        return 0.0   # the reference's compute() returns 0 (block_ildlt.impl.cuh:145)

    def refactor(self, vals, threshold=1e-4, stream=None):
        import torch
        if isinstance(vals, torch.Tensor):
            check(lib().culip_bildlt_refactor(self._h, _dptr(vals), 1, threshold,
                                              _stream(stream)))
        else:
            v = _f64(vals)
            check(lib().culip_bildlt_refactor(self._h, v.ctypes.data_as(C.c_void_p), 0,
                                              threshold, _stream(stream)))

    def set_solve_algorithm(self, algorithm, num_jacobi_sweeps=0):
        """set_solve_algorithm: 'single' (default: one launch per solve),
        'levels' (one launch per level) or 'jacobi' (num_jacobi_sweeps
        block-Jacobi sweeps)."""
        check(lib().culip_bildlt_set_solve_algorithm(self._h, _solve_algorithm(algorithm),
                                                     num_jacobi_sweeps))

    def export_factorized(self, folder):
        """Writes out_U.mtx, out_D.mtx, out_P.mtx under folder; returns the
        achieved fill factor (2 nnz(U) + nnz(D)) / nnz(A)."""
        f = C.c_double()
        check(lib().culip_bildlt_export_factorized(self._h, os.fsencode(folder), C.byref(f)))
        return f.value

    def _solve(self, side, b, x, transpose, stream):
        check(lib().culip_bildlt_solve(self._h, side, 1 if transpose else 0, _dptr(b), _dptr(x),
                                       _stream(stream)))

    def solve_noalloc(self, side, b, x, transpose=False, stream=None):
        """solve_<side> without allocating (culip_bildlt_solve_noalloc):
        exact solves as level launches, block-Jacobi work in the library's
        static scratch; side 'left' | 'middle' | 'right'."""
        sd = {"left": SOLVE_LEFT, "middle": SOLVE_MIDDLE, "right": SOLVE_RIGHT}[side]
        check(lib().culip_bildlt_solve_noalloc(self._h, sd, 1 if transpose else 0, _dptr(b),
                                               _dptr(x), _stream(stream)))

    def solve_left(self, b, x, transpose=False, stream=None):
        self._solve(SOLVE_LEFT, b, x, transpose, stream)

    def solve_middle(self, b, x, transpose=False, stream=None):
        self._solve(SOLVE_MIDDLE, b, x, transpose, stream)

    def solve_right(self, b, x, transpose=False, stream=None):
        self._solve(SOLVE_RIGHT, b, x, transpose, stream)

    def left_product(self, A, src, x, stream=None):
        """x = U' \\ (P' (A src)) for a CSRMatrixSPMV A, folded into one
        launch where the solve schedule allows; returns whether it was."""
        f = C.c_int32()
        check(lib().culip_bildlt_solve_left_product(self._h, A._h, _dptr(src), _dptr(x),
                                                    C.byref(f), _stream(stream)))
        return bool(f.value)

    def get_int(self, name):
        p = C.POINTER(C.c_int32)()
        n = C.c_int64()
        check(lib().culip_bildlt_get_int(self._h, name.encode(), C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.int32)

    def get_real(self, name):
        p = C.POINTER(C.c_double)()
        n = C.c_int64()
        check(lib().culip_bildlt_get_real(self._h, name.encode(), C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0)

    def device_bytes(self):
        return lib().culip_bildlt_device_bytes(self._h)

    def set_pivoting(self, mode):
        """'static' (default), 'bk' (Bunch-Kaufman) or 'rook': numeric
        pivoting inside pivot blocks (BlockiLDLt<T, bk, rook>)."""
        m = {"static": 0, "bk": 1, "rook": 2}[mode]
        check(lib().culip_bildlt_set_pivoting(self._h, m))

    def set_pivot_precision(self, p):
        """FP32 factor: 'f64' (default) forms the 2x2 pivot steps and D
        solves in FP64, 'factor' in FP32 as BlockiLDLt<float> does."""
        check(lib().culip_bildlt_set_pivot_precision(self._h, {"f64": 0, "factor": 1}[p]))

    def set_graphs(self, enable=True):
        """Refactors on explicit streams replay a captured HIP graph."""
        check(lib().culip_bildlt_set_graphs(self._h, 1 if enable else 0))

    def _inject_solve_error(self):
        """Testing hook: records a timed-out single-launch solve from the host."""
        check(lib().culip_bildlt_inject_solve_error(self._h))

    def _release(self):
        if getattr(self, "_h", None):
            lib().culip_bildlt_destroy(self._h)
            self._h = None


class _Spmv(_Handle):
    def multiply(self, x, y, alpha=1.0, beta=0.0, transpose=False, stream=None):
        check(lib().culip_spmv_multiply(self._h, _dptr(x), _dptr(y), alpha, beta,
                                        1 if transpose else 0, _stream(stream)))

    def m(self):
        return lib().culip_spmv_m(self._h)

    def n(self):
        return lib().culip_spmv_n(self._h)

    def _release(self):
        if getattr(self, "_h", None):
            lib().culip_spmv_destroy(self._h)
            self._h = None


class CSRMatrixSPMV(_Spmv):
    def __init__(self, row_ptr, col, val, n=None):
        s, self._keep = csr_struct(row_ptr, col, val, n)
        h = C.c_void_p()
        check(lib().culip_spmv_create_csr(C.byref(s), C.byref(h)))
        self._h = h
        self._track()

    def equilibrate(self, sweeps, d, stream=None):
        """Values <- D A D by `sweeps` equilibration sweeps; d (device
        float64 tensor, n) receives diag(D)."""
        check(lib().culip_spmv_equilibrate(self._h, sweeps, _dptr(d), _stream(stream)))

    def update_values(self, vals, stream=None):
        v = _f64(vals)
        check(lib().culip_spmv_update_values(self._h, v.ctypes.data_as(C.c_void_p), 0,
                                             _stream(stream)))


class AugmentedMatrixSPMV(_Spmv):
    """K = [diag(d) A'; A delta*I(ix)] (spmv.cuh:116-149), dimension n + m;
    with H = (row_ptr, col, val) of an n x n symmetric CSR matrix in place
    of d, K = [H A'; A delta*I(ix)] (culip_spmv_create_kkt_csr)."""

    def __init__(self, a_row_ptr, a_col, a_val, a_ncols, d=None, delta=0.0, ix=None, H=None):
        s, self._keep = csr_struct(a_row_ptr, a_col, a_val, a_ncols)
        h = C.c_void_p()
        ixa = _i32(ix if ix is not None else [])
        ixp = _ptr(ixa, C.c_int32) if len(ixa) else None
        if H is not None:
            hs, self._hkeep = csr_struct(*H)
            check(lib().culip_spmv_create_kkt_csr(C.byref(hs), C.byref(s), delta, len(ixa), ixp,
                                                  C.byref(h)))
        else:
            dd = _f64(d)
            check(lib().culip_spmv_create_kkt(C.byref(s), _ptr(dd, C.c_double), delta,
                                              len(ixa), ixp, C.byref(h)))
        self._h = h
        self._track()


class NormalMatrixSPMV(_Spmv):
    """N = A diag(d)^-1 A' + delta*I(ix) (spmv.cuh:74-107), dimension m."""

    def __init__(self, a_row_ptr, a_col, a_val, a_ncols, d, delta=0.0, ix=None):
        s, self._keep = csr_struct(a_row_ptr, a_col, a_val, a_ncols)
        dd = _f64(d)
        h = C.c_void_p()
        ixa = _i32(ix if ix is not None else [])
        check(lib().culip_spmv_create_normal(C.byref(s), _ptr(dd, C.c_double), delta,
                                             len(ixa),
                                             _ptr(ixa, C.c_int32) if len(ixa) else None,
                                             C.byref(h)))
        self._h = h
        self._track()


class ImportLDUPreconditioner(_Handle):
    """ImportLDUPreconditioner<T> (import_ldu_preconditioner.cuh:36-125):
    M = inv(S_r) P_r' L D R P_c' inv(S_c) with R = L' (U None) or U.

    L, U: (row_ptr, col, val) host CSR; tri / is_piv: the tridiagonal D and
    its pivot-start flags (None: no D); exact_solve False applies 10 Jacobi
    sweeps instead of the substitution; prec 'f32' stores L and U in FP32."""

    def __init__(self, L, tri=None, is_piv=None, U=None, exact_solve=True, prec="f64"):
        ls, keep_l = csr_struct(*L)
        us, keep_u = csr_struct(*U) if U is not None else (None, None)
        self._keep = (keep_l, keep_u)
        t = _f64(tri) if tri is not None else None
        p = _i32(is_piv) if is_piv is not None else None
        pr = PREC_F32 if prec in ("f32", PREC_F32) else PREC_F64
        h = C.c_void_p()
        check(lib().culip_ldu_create(C.byref(ls), C.byref(us) if us is not None else None,
                                     _ptr(t, C.c_double) if t is not None else None,
                                     _ptr(p, C.c_int32) if p is not None else None,
                                     1 if exact_solve else 0, pr, C.byref(h)))
        self._h = h
        self._track()
        self.N = len(L[0]) - 1

    def n(self):
        return lib().culip_ldu_n(self._h)

    def is_left(self):
        return True

    def is_middle(self):
        return bool(lib().culip_ldu_is_middle(self._h))

    def is_right(self):
        return True

    def set_permutation(self, p_r, p_c=None):
        """Push convention: entry i moves to p[i]; None removes."""
        if p_r is None:
            check(lib().culip_ldu_set_permutation(self._h, None, None))
            return
        a, b = _i32(p_r), _i32(p_c if p_c is not None else p_r)
        check(lib().culip_ldu_set_permutation(self._h, _ptr(a, C.c_int32), _ptr(b, C.c_int32)))

    def set_scaling(self, s_r, s_c=None):
        if s_r is None:
            check(lib().culip_ldu_set_scaling(self._h, None, None))
            return
        a, b = _f64(s_r), _f64(s_c if s_c is not None else s_r)
        check(lib().culip_ldu_set_scaling(self._h, _ptr(a, C.c_double), _ptr(b, C.c_double)))

    def _solve(self, side, b, x, transpose, stream):
        check(lib().culip_ldu_solve(self._h, side, 1 if transpose else 0, _dptr(b), _dptr(x),
                                    _stream(stream)))

    def solve_noalloc(self, side, b, x, transpose=False, stream=None):
        """solve_<side> without allocating (culip_bildlt_solve_noalloc):
        exact solves as level launches, block-Jacobi work in the library's
        static scratch; side 'left' | 'middle' | 'right'."""
        sd = {"left": SOLVE_LEFT, "middle": SOLVE_MIDDLE, "right": SOLVE_RIGHT}[side]
        check(lib().culip_bildlt_solve_noalloc(self._h, sd, 1 if transpose else 0, _dptr(b),
                                               _dptr(x), _stream(stream)))

    def solve_left(self, b, x, transpose=False, stream=None):
        self._solve(SOLVE_LEFT, b, x, transpose, stream)

    def solve_middle(self, b, x, transpose=False, stream=None):
        self._solve(SOLVE_MIDDLE, b, x, transpose, stream)

    def solve_right(self, b, x, transpose=False, stream=None):
        self._solve(SOLVE_RIGHT, b, x, transpose, stream)

    def _release(self):
        if getattr(self, "_h", None):
            lib().culip_ldu_destroy(self._h)
            self._h = None


class SQMR(_Handle):
    """SQMR<double>(spmv, tol, max_it, prec); solve(b, x) -> (residual, iterations)."""

    def __init__(self, spmv, tol=1e-8, max_it=1000, prec=None, true_residual_period=0,
                 max_restarts=0, separate_product=False):
        self._spmv, self._prec = spmv, prec
        o = _lib.SqmrOpts(tol, max_it, true_residual_period, max_restarts,
                          1 if separate_product else 0)
        h = C.c_void_p()
        if isinstance(prec, ImportLDUPreconditioner):
            check(lib().culip_sqmr_create_ldu(spmv._h, prec._h, C.byref(o), C.byref(h)))
        else:
            check(lib().culip_sqmr_create(spmv._h, prec._h if prec is not None else None,
                                          C.byref(o), C.byref(h)))
        self._h = h
        self._track()
        self.result = None

    def set_residual_scale(self, s):
        """Measure true residuals as ||(b - A x) ./ s|| / ||b ./ s|| (s: device
        tensor kept alive here; None resets)."""
        self._res_scale = s
        check(lib().culip_sqmr_set_residual_scale(self._h, _dptr(s)))

    def solve(self, b, x, stream=None):
        r = _lib.SqmrResult()
        check(lib().culip_sqmr_solve(self._h, _dptr(b), _dptr(x), C.byref(r), _stream(stream)))
        self.result = {"residual": r.residual, "iterations": r.iterations,
                       "converged": bool(r.converged), "restarts": r.restarts,
                       "true_rel_residual": r.true_rel_residual}
        return r.residual, r.iterations

    def _release(self):
        if getattr(self, "_h", None):
            lib().culip_sqmr_destroy(self._h)
            self._h = None


class Server(_Handle):
    """Request loop for KKT solves over resident sparsity patterns.

    add_pattern(K, piv_starts, block_starts) -> pattern id (symbolic analysis
    and upload, once); submit(pattern, vals, rhs) -> request id (host arrays,
    copied); wait(request) -> (x, result, timing, slot) with timing =
    {queued, service, latency, submit} seconds."""

    def __init__(self, slots=4, prec="f32", hbm_budget=0, fill_level=0, fill_factor=1.0,
                 threshold=1e-4, tol=1e-8, max_it=1000, true_residual_period=20,
                 max_restarts=10, solve_algorithm="levels", jacobi_sweeps=0,
                 equilibrate_sweeps=0, graphs=False):
        pr = PREC_F32 if prec in ("f32", PREC_F32) else PREC_F64
        o = _lib.ServerOpts(slots, int(hbm_budget), pr, fill_level, fill_factor, threshold,
                            _lib.SqmrOpts(tol, max_it, true_residual_period, max_restarts),
                            _solve_algorithm(solve_algorithm), jacobi_sweeps,
                            equilibrate_sweeps, 1 if graphs else 0)
        h = C.c_void_p()
        check(lib().culip_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._track()
        self._n = {}
        self._req_n = {}

    def add_pattern(self, row_ptr, col, val, piv_starts, block_starts):
        s, keep = csr_struct(row_ptr, col, val)
        ps, bs = _i32(piv_starts), _i32(block_starts)
        p = C.c_int32()
        check(lib().culip_server_add_pattern(self._h, C.byref(s), len(ps), _ptr(ps, C.c_int32),
                                             len(bs), _ptr(bs, C.c_int32), C.byref(p)))
        self._n[p.value] = len(row_ptr) - 1
        return p.value

    def set_pattern_solve(self, pattern, algorithm, jacobi_sweeps=0):
        """Triangular-solve schedule of one pattern, before its first submit."""
        check(lib().culip_server_set_pattern_solve(self._h, pattern,
                                                   _solve_algorithm(algorithm), jacobi_sweeps))

    def submit(self, pattern, vals, rhs):
        v, b = _f64(vals), _f64(rhs)
        r = C.c_int64()
        check(lib().culip_server_submit(self._h, pattern, _ptr(v, C.c_double),
                                        _ptr(b, C.c_double), C.byref(r)))
        self._req_n[r.value] = b.size
        return r.value

    def wait(self, request):
        x = np.empty(self._req_n.pop(request))
        res = _lib.SqmrResult()
        t = np.empty(4)
        slot = C.c_int32()
        check(lib().culip_server_wait(self._h, request, _ptr(x, C.c_double), C.byref(res),
                                      _ptr(t, C.c_double), C.byref(slot)))
        result = {"residual": res.residual, "iterations": res.iterations,
                  "converged": bool(res.converged), "restarts": res.restarts,
                  "true_rel_residual": res.true_rel_residual}
        return x, result, {"queued": t[0], "service": t[1], "latency": t[2],
                           "submit": t[3]}, slot.value

    def stats(self):
        out = np.zeros(5, np.int64)
        check(lib().culip_server_stats(self._h, _ptr(out, C.c_int64)))
        return {"used": int(out[0]), "peak": int(out[1]), "budget": int(out[2]),
                "instances": int(out[3]), "completed": int(out[4])}

    def close(self):
        if getattr(self, "_h", None):
            lib().culip_server_destroy(self._h)
            self._h = None

    def _release(self):
        self.close()


class Batch(_Handle):
    """Independent systems refactored and solved together (culip_batch):
    each step of the refactor (a level of all systems) and of the SQMR
    recurrence is one launch over all systems taking part, with block-Jacobi
    triangular solves of `jacobi_sweeps` sweeps. add(K, piv_starts,
    block_starts) -> index (symbolic analysis and upload; thread-safe,
    before the first solve); solve(vals, rhs) -> (xs, results) for lists of
    arrays, one per system: numpy arrays (host), or float64 CUDA tensors
    (device; xs are then CUDA tensors). Each system's x and result equal its
    solve alone with the same options (BlockiLDLt with Jacobi sweeps +
    SQMR)."""

    def __init__(self, prec="f32", fill_level=0, fill_factor=1.0, threshold=1e-4, tol=1e-8,
                 max_it=1000, true_residual_period=20, max_restarts=10, jacobi_sweeps=2):
        pr = PREC_F32 if prec in ("f32", PREC_F32) else PREC_F64
        o = _lib.BatchOpts(pr, fill_level, fill_factor, threshold,
                           _lib.SqmrOpts(tol, max_it, true_residual_period, max_restarts),
                           jacobi_sweeps)
        h = C.c_void_p()
        check(lib().culip_batch_create(C.byref(o), C.byref(h)))
        self._h = h
        self._track()
        self._n = {}

    def add(self, row_ptr, col, val, piv_starts, block_starts):
        s, keep = csr_struct(row_ptr, col, val)
        ps, bs = _i32(piv_starts), _i32(block_starts)
        i = C.c_int32()
        check(lib().culip_batch_add(self._h, C.byref(s), len(ps), _ptr(ps, C.c_int32), len(bs),
                                    _ptr(bs, C.c_int32), C.byref(i)))
        self._n[i.value] = len(row_ptr) - 1
        return i.value

    def solve(self, vals, rhs, stream=None):
        S = len(self._n)
        if len(vals) != S or len(rhs) != S:
            raise ValueError(f"batch: {S} systems, got {len(vals)} value and {len(rhs)} "
                             "right-hand-side arrays")
        P = C.POINTER(C.c_double)
        device = S > 0 and hasattr(vals[0], "is_cuda")
        if device:
            import torch
            for a in list(vals) + list(rhs):
                if not (a.is_cuda and a.dtype == torch.float64 and a.is_contiguous()):
                    raise ValueError("batch: device arrays must be contiguous float64 CUDA "
                                     "tensors")
            xs = [torch.empty(self._n[i], dtype=torch.float64, device=vals[0].device)
                  for i in range(S)]
            addr = lambda a: C.cast(C.c_void_p(a.data_ptr()), P)
            keep = (vals, rhs)
        else:
            v = [_f64(a) for a in vals]
            b = [_f64(a) for a in rhs]
            xs = [np.empty(self._n[i]) for i in range(S)]
            addr = lambda a: _ptr(a, C.c_double)
            keep = (v, b)
        pv = (P * S)(*[addr(a) for a in keep[0]])
        pb = (P * S)(*[addr(a) for a in keep[1]])
        px = (P * S)(*[addr(a) for a in xs])
        res = (_lib.SqmrResult * S)()
        check(lib().culip_batch_solve(self._h, pv, pb, px, res, 1 if device else 0,
                                      _stream(stream)))
        out = [{"residual": r.residual, "iterations": r.iterations,
                "converged": bool(r.converged), "restarts": r.restarts,
                "true_rel_residual": r.true_rel_residual} for r in res]
        return xs, out

    def info(self):
        """systems, device bytes, and the last solve's seconds: refactors
        and SQMR"""
        c, b = C.c_int32(), C.c_int64()
        t = np.zeros(2)
        check(lib().culip_batch_info(self._h, C.byref(c), C.byref(b), _ptr(t, C.c_double)))
        return {"systems": c.value, "device_bytes": b.value, "factor_s": float(t[0]),
                "sqmr_s": float(t[1])}

    def close(self):
        if getattr(self, "_h", None):
            lib().culip_batch_destroy(self._h)
            self._h = None

    def _release(self):
        self.close()


def weighted_matching(row_ptr, col, val, n):
    """Maximum-product matching of a square matrix with its scaling
    (culip_weighted_matching): dict of row_match, u, v, row_scale,
    col_scale, sym_scale, objective, feasible."""
    s, keep = csr_struct(row_ptr, col, val, n)
    m = s.m
    rm = np.empty(m, np.int32)
    out = {k: np.empty(m) for k in ("u", "v", "row_scale", "col_scale", "sym_scale")}
    obj = C.c_double()
    feas = C.c_int32()
    check(lib().culip_weighted_matching(C.byref(s), _ptr(rm, C.c_int32),
                                        *[_ptr(out[k], C.c_double) for k in
                                          ("u", "v", "row_scale", "col_scale", "sym_scale")],
                                        C.byref(obj), C.byref(feas)))
    out.update(row_match=rm, objective=obj.value, feasible=bool(feas.value))
    return out


def _pattern_struct(row_ptr, col, m, n):
    rp, ci = _i32(row_ptr), _i32(col)
    s = _lib.Csr(m, n, len(ci), _ptr(rp, C.c_int32), _ptr(ci, C.c_int32), None)
    return s, (rp, ci)


def structural_rank(row_ptr, col, m, n):
    """Structural rank of an m x n pattern and a maximum matching
    (culip_structural_rank): (rank, row_match with -1 for unmatched rows)."""
    s, keep = _pattern_struct(row_ptr, col, m, n)
    r = C.c_int32()
    rm = np.empty(m, np.int32)
    check(lib().culip_structural_rank(C.byref(s), C.byref(r), _ptr(rm, C.c_int32)))
    return r.value, rm


def btf_rank_bound(row_ptr, col, m, n):
    """Rank lower bound by greedy triangular peeling (culip_btf_rank_bound):
    (bound, 1-based row order, 1-based column order)."""
    s, keep = _pattern_struct(row_ptr, col, m, n)
    b = C.c_int32()
    ro, co = np.empty(m, np.int32), np.empty(n, np.int32)
    check(lib().culip_btf_rank_bound(C.byref(s), C.byref(b), _ptr(ro, C.c_int32),
                                     _ptr(co, C.c_int32)))
    return b.value, ro, co


def read_matrix_csr(path):
    """Test<T>::read_matrix_csr (libs/test/test.impl.h:218-228)."""
    h = C.c_void_p()
    check(lib().culip_mm_read_csr(str(path).encode(), C.byref(h)))
    c = _lib.Csr()
    check(lib().culip_host_matrix_view(h, C.byref(c)))
    out = _copy_csr(c) + (c.n,)
    lib().culip_host_matrix_destroy(h)
    return out


def read_dense_vector(path):
    """Test<T>::read_dense_vector (libs/test/test.impl.h:236-300)."""
    n = C.c_int64()
    p = C.POINTER(C.c_double)()
    check(lib().culip_mm_read_vector(str(path).encode(), C.byref(n), C.byref(p)))
    v = np.ctypeslib.as_array(p, shape=(n.value,)).copy()
    lib().culip_host_free(p)
    return v


def permute_scale(row_ptr, col, val, row_perm, col_perm=None, perm_is_old_to_new=False,
                  row_scale=None, col_scale=None, scale_before_match=False):
    """T_matrix_permute_scale (libs/la/sparse_la.impl.cuh:600-657): with new ->
    old vectors B(i, j) = A(p_r[i], p_c[j]) / scale; returns (row_ptr, col, val)."""
    s, keep = csr_struct(row_ptr, col, val)
    pr = _i32(row_perm)
    pc = _i32(col_perm if col_perm is not None else row_perm)
    rs = _f64(row_scale) if row_scale is not None else None
    cs = _f64(col_scale) if col_scale is not None else None
    h = C.c_void_p()
    check(lib().culip_csr_permute_scale(C.byref(s), _ptr(pr, C.c_int32), _ptr(pc, C.c_int32),
                                        1 if perm_is_old_to_new else 0,
                                        _ptr(rs, C.c_double) if rs is not None else None,
                                        _ptr(cs, C.c_double) if cs is not None else None,
                                        1 if scale_before_match else 0, C.byref(h)))
    c = _lib.Csr()
    check(lib().culip_host_matrix_view(h, C.byref(c)))
    out = _copy_csr(c)
    lib().culip_host_matrix_destroy(h)
    return out
