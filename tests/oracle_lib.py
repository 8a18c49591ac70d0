"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: a plain-C restatement of the reference
algorithm used as the checker, never as the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


class SqmrRes(C.Structure):
    _fields_ = [("residual", C.c_double), ("iterations", C.c_int),
                ("flag_delta", C.c_int), ("flag_res", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        h = C.CDLL(ORACLE_SO)
        I, D, P = C.c_int, C.c_double, C.c_void_p
        pi, pd = C.POINTER(C.c_int), C.POINTER(C.c_double)
        for suf in ("d", "f"):
            f = getattr(h, f"or_bildlt_create_{suf}")
            f.restype, f.argtypes = P, [I, pi, pi, pd, I, pi, I, pi]
            f = getattr(h, f"or_bildlt_symbolic_{suf}")
            f.restype, f.argtypes = I, [P, I, D]
            f = getattr(h, f"or_bildlt_factor_{suf}")
            f.restype, f.argtypes = I, [P, D]
            f = getattr(h, f"or_bildlt_set_pivoting_{suf}")
            f.restype, f.argtypes = None, [P, I]
            f = getattr(h, f"or_bildlt_set_pivot_precision_{suf}")
            f.restype, f.argtypes = None, [P, I]
            for side in ("left", "middle", "right"):
                f = getattr(h, f"or_bildlt_solve_{side}_{suf}")
                f.restype, f.argtypes = None, [P, pd, pd]
            f = getattr(h, f"or_bildlt_get_{suf}")
            f.restype, f.argtypes = C.c_long, [P, C.c_char_p, C.POINTER(P)]
            f = getattr(h, f"or_bildlt_destroy_{suf}")
            f.restype, f.argtypes = None, [P]
            f = getattr(h, f"or_sqmr_bildlt_{suf}")
            f.restype, f.argtypes = I, [I, pi, pi, pd, P, pd, pd, D, I, C.POINTER(SqmrRes)]
        h.or_csr_spmv.restype, h.or_csr_spmv.argtypes = None, [I, pi, pi, pd, pd, pd]
        h.or_kkt_spmv.restype = None
        h.or_kkt_spmv.argtypes = [I, I, pi, pi, pd, pd, D, I, pi, pd, pd, D, D]
        h.or_sqmr_ldu_d.restype = I
        h.or_sqmr_ldu_d.argtypes = [I, pi, pi, pd, pi, pi, pd, pd, pi, pi, pd, pd, pd, D, I,
                                    C.POINTER(SqmrRes)]
        h.or_solve_D.restype, h.or_solve_D.argtypes = None, [I, pd, pi, pd]
        _lib = h
    return _lib


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# names holding reals in the oracle handle (REAL = T)
_REAL_NAMES = {"val_store", "fval_store", "tri", "row_norms"}


class OracleBildlt:
    """BlockiLDLt<T> restated on the CPU; T = 'd' (double) or 'f' (float)."""

    def __init__(self, row_ptr, col, val, piv_starts, block_starts, t="d"):
        self.t = t
        self._keep = (_i(row_ptr), _i(col), _d(val), _i(piv_starts), _i(block_starts))
        rp, ci, cv, ps, bs = self._keep
        self.n = len(rp) - 1
        self._h = getattr(lib(), f"or_bildlt_create_{t}")(self.n, _pi(rp), _pi(ci), _pd(cv),
                                                           len(ps), _pi(ps), len(bs), _pi(bs))

    def symbolic(self, fill_level=0, fill_factor=1.0):
        r = getattr(lib(), f"or_bildlt_symbolic_{self.t}")(self._h, fill_level, fill_factor)
        if r != 0:
            raise RuntimeError("oracle symbolic failed")

    def factor(self, threshold):
        getattr(lib(), f"or_bildlt_factor_{self.t}")(self._h, threshold)

    def set_pivoting(self, mode):
        """'static' (default), 'bk' (Bunch-Kaufman) or 'rook'."""
        m = {"static": 0, "bk": 1, "rook": 2}[mode]
        getattr(lib(), f"or_bildlt_set_pivoting_{self.t}")(self._h, m)

    def set_pivot_precision(self, p):
        """'factor' (default: 2x2 steps in the factor's type, as the
        reference) or 'f64' (in double, the library's FP32 default)."""
        getattr(lib(), f"or_bildlt_set_pivot_precision_{self.t}")(self._h, {"factor": 0, "f64": 1}[p])

    def get(self, name):
        p = C.c_void_p()
        n = getattr(lib(), f"or_bildlt_get_{self.t}")(self._h, name.encode(), C.byref(p))
        if n < 0:
            raise KeyError(name)
        if n == 0:
            return np.zeros(0)
        if name in _REAL_NAMES:
            ct = C.c_double if self.t == "d" else C.c_float
            arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,))
            return arr.astype(np.float64)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int)), shape=(n,)).copy()

    def solve(self, side, b):
        b = _d(b)
        x = np.zeros_like(b)
        getattr(lib(), f"or_bildlt_solve_{side}_{self.t}")(self._h, _pd(b), _pd(x))
        return x

    def __del__(self):
        if getattr(self, "_h", None):
            getattr(lib(), f"or_bildlt_destroy_{self.t}")(self._h)
            self._h = None


def csr_spmv(row_ptr, col, val, x):
    rp, ci, cv, xx = _i(row_ptr), _i(col), _d(val), _d(x)
    y = np.zeros(len(rp) - 1)
    lib().or_csr_spmv(len(rp) - 1, _pi(rp), _pi(ci), _pd(cv), _pd(xx), _pd(y))
    return y


def kkt_spmv(a_row, a_col, a_val, n, d, delta, ix, x, y=None, alpha=1.0, beta=0.0):
    ar, ac, av, dd, xx = _i(a_row), _i(a_col), _d(a_val), _d(d), _d(x)
    ixa = _i(ix if ix is not None else [])
    m = len(ar) - 1
    yy = np.zeros(n + m) if y is None else _d(y).copy()
    lib().or_kkt_spmv(m, n, _pi(ar), _pi(ac), _pd(av), _pd(dd), delta, len(ixa), _pi(ixa),
                      _pd(xx), _pd(yy), alpha, beta)
    return yy


def sqmr_bildlt(row_ptr, col, val, prec, b, tol=1e-8, max_it=1000):
    rp, ci, cv, bb = _i(row_ptr), _i(col), _d(val), _d(b)
    x = np.zeros_like(bb)
    r = SqmrRes()
    t = prec.t if prec is not None else "d"
    getattr(lib(), f"or_sqmr_bildlt_{t}")(len(rp) - 1, _pi(rp), _pi(ci), _pd(cv),
                                         prec._h if prec is not None else None, _pd(bb), _pd(x),
                                         tol, max_it, C.byref(r))
    return x, r


def sqmr_ldu(row_ptr, col, val, l_csr, tri, is_piv, perm, scale, b, tol, max_it):
    rp, ci, cv, bb = _i(row_ptr), _i(col), _d(val), _d(b)
    x = np.zeros_like(bb)
    r = SqmrRes()
    if l_csr is None:
        lib().or_sqmr_ldu_d(len(rp) - 1, _pi(rp), _pi(ci), _pd(cv), None, None, None, None,
                            None, None, None, _pd(bb), _pd(x), tol, max_it, C.byref(r))
        return x, r
    lr, lc, lv = _i(l_csr[0]), _i(l_csr[1]), _d(l_csr[2])
    tr, ip, pm, sc = _d(tri), _i(is_piv), _i(perm), _d(scale)
    lib().or_sqmr_ldu_d(len(rp) - 1, _pi(rp), _pi(ci), _pd(cv), _pi(lr), _pi(lc), _pd(lv),
                        _pd(tr), _pi(ip), _pi(pm), _pd(sc), _pd(bb), _pd(x), tol, max_it,
                        C.byref(r))
    return x, r


def solve_D(tri, is_piv, b):
    tr, ip, bb = _d(tri), _i(is_piv), _d(b).copy()
    lib().or_solve_D(len(bb), _pd(tr), _pi(ip), _pd(bb))
    return bb


def sqmr_bildlt_refined(row_ptr, col, val, prec, b, tol=1e-8, max_it=1000, max_rounds=20):
    """The reference SQMR (its quasi-residual stop, sqmr.impl.cuh:89-107)
    followed by rounds of iterative refinement, x += SQMR(b - K x), until
    the FP64 true relative residual ||b - K x|| / ||b|| meets tol: the
    stopping rule of the device path's true-residual gate. Returns
    (x, total iterations, refinement rounds, true relative residual)."""
    b = _d(b)
    x = np.zeros_like(b)
    nb = np.linalg.norm(b)
    its, rounds = 0, 0
    rel = 1.0
    for rounds in range(max_rounds + 1):
        r = b - csr_spmv(row_ptr, col, val, x)
        rel = np.linalg.norm(r) / nb if nb > 0 else 0.0
        if rel <= tol or rounds == max_rounds:
            break
        # the round's own stop is relative to ||r|| and to ||K|| ||x||, so it
        # aims at 1% of the remaining target tol ||b|| / ||r||
        dx, res = sqmr_bildlt(row_ptr, col, val, prec, r, min(1.0, 0.01 * tol / rel), max_it)
        its += res.iterations
        x += dx
    return x, its, rounds, rel
