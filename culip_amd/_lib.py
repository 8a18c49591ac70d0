"""ctypes binding of libculip_amd.so (the C-ABI declared in include/culip_amd.h).

The shared library is built in-tree (culip_amd/libculip_amd.so). Loading
fails loudly when it is missing: there is no fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libculip_amd.so")

OK = 0
STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "NOMEM", 4: "STATE", 5: "NUMERIC",
          6: "BUDGET", 7: "UNSUPPORTED", 8: "NO_DEVICE"}

PREC_F64, PREC_F32 = 0, 1
SOLVE_LEFT, SOLVE_MIDDLE, SOLVE_RIGHT = 0, 1, 2
ORDER_NATURAL, ORDER_BLOCK_COLOR = 0, 1
TRSV_LEVELS, TRSV_SINGLE, TRSV_JACOBI = 0, 1, 2


class CulipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"culip_amd {STATUS.get(status, status)}: {msg}")
        self.status = status


class Csr(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int32)), ("col", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double))]


class KktParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("h_offdiag_per_row", C.c_int32),
                ("a_nnz_per_col", C.c_int32), ("h_diag_lo", C.c_double),
                ("h_diag_hi", C.c_double), ("h_diag_log", C.c_int32),
                ("h_offdiag_std", C.c_double), ("delta", C.c_double), ("seed", C.c_uint64),
                ("barrier_lo", C.c_double), ("barrier_hi", C.c_double)]


class SqmrOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_it", C.c_int32),
                ("true_residual_period", C.c_int32), ("max_restarts", C.c_int32),
                ("separate_product", C.c_int32)]


class SqmrResult(C.Structure):
    _fields_ = [("residual", C.c_double), ("iterations", C.c_int32),
                ("converged", C.c_int32), ("restarts", C.c_int32),
                ("true_rel_residual", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("hbm_budget", C.c_int64), ("prec", C.c_int),
                ("fill_level", C.c_int32), ("fill_factor", C.c_double),
                ("threshold", C.c_double), ("sqmr", SqmrOpts),
                ("solve_algorithm", C.c_int), ("jacobi_sweeps", C.c_int32),
                ("equilibrate_sweeps", C.c_int32), ("graphs", C.c_int32)]


class BatchOpts(C.Structure):
    _fields_ = [("prec", C.c_int), ("fill_level", C.c_int32), ("fill_factor", C.c_double),
                ("threshold", C.c_double), ("sqmr", SqmrOpts), ("jacobi_sweeps", C.c_int32)]


_lib = None

# (name, restype, argtypes)
_SIGS = [
    ("culip_last_error", C.c_char_p, []),
    ("culip_version", C.c_char_p, []),
    ("culip_kkt_generate", C.c_int, [C.POINTER(KktParams), C.POINTER(C.c_void_p)]),
    ("culip_kkt_view", C.c_int, [C.c_void_p, C.POINTER(Csr), C.POINTER(C.POINTER(C.c_double))]),
    ("culip_kkt_barrier_values", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_uint64,
                                           C.POINTER(C.c_double)]),
    ("culip_kkt_destroy", None, [C.c_void_p]),
    ("culip_prep_create", C.c_int, [C.POINTER(Csr), C.c_int32, C.c_int32, C.c_int,
                                    C.POINTER(C.c_void_p)]),
    ("culip_prep_view", C.c_int, [C.c_void_p, C.POINTER(Csr), C.POINTER(C.POINTER(C.c_int32)),
                                  C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_int32)),
                                  C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_int32))]),
    ("culip_prep_permute_values", C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]),
    ("culip_prep_destroy", None, [C.c_void_p]),
    ("culip_mm_read_csr", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("culip_mm_read_vector", C.c_int, [C.c_char_p, C.POINTER(C.c_int64),
                                       C.POINTER(C.POINTER(C.c_double))]),
    ("culip_csr_permute_scale", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.c_int32,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                                          C.POINTER(C.c_void_p)]),
    ("culip_host_matrix_view", C.c_int, [C.c_void_p, C.POINTER(Csr)]),
    ("culip_host_matrix_destroy", None, [C.c_void_p]),
    ("culip_host_free", None, [C.c_void_p]),
    ("culip_bildlt_create", C.c_int, [C.POINTER(Csr), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                      C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_void_p)]),
    ("culip_bildlt_compute", C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double,
                                       C.c_void_p]),
    ("culip_bildlt_symbolic", C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    ("culip_bildlt_refactor", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                        C.c_void_p]),
    ("culip_bildlt_solve", C.c_int, [C.c_void_p, C.c_int, C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    ("culip_bildlt_n", C.c_int32, [C.c_void_p]),
    ("culip_bildlt_set_solve_algorithm", C.c_int, [C.c_void_p, C.c_int, C.c_int32]),
    ("culip_bildlt_export_factorized", C.c_int, [C.c_void_p, C.c_char_p,
                                                 C.POINTER(C.c_double)]),
    ("culip_bildlt_get_int", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.POINTER(C.c_int32)),
                                       C.POINTER(C.c_int64)]),
    ("culip_bildlt_get_real", C.c_int, [C.c_void_p, C.c_char_p,
                                        C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int64)]),
    ("culip_bildlt_device_bytes", C.c_int64, [C.c_void_p]),
    ("culip_bildlt_inject_solve_error", C.c_int, [C.c_void_p]),
    ("culip_bildlt_set_graphs", C.c_int, [C.c_void_p, C.c_int32]),
    ("culip_bildlt_set_pivoting", C.c_int, [C.c_void_p, C.c_int]),
    ("culip_bildlt_set_pivot_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("culip_bildlt_solve_left_product", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.POINTER(C.c_int32),
                                                  C.c_void_p]),
    ("culip_bildlt_solve_noalloc", C.c_int, [C.c_void_p, C.c_int, C.c_int32, C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
    ("culip_bildlt_destroy", None, [C.c_void_p]),
    ("culip_spmv_create_csr", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_void_p)]),
    ("culip_spmv_create_kkt", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_double), C.c_double,
                                        C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_void_p)]),
    ("culip_spmv_create_normal", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_double), C.c_double,
                                           C.c_int32, C.POINTER(C.c_int32),
                                           C.POINTER(C.c_void_p)]),
    ("culip_spmv_update_values", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    ("culip_spmv_multiply", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                      C.c_double, C.c_int32, C.c_void_p]),
    ("culip_spmv_m", C.c_int32, [C.c_void_p]),
    ("culip_spmv_n", C.c_int32, [C.c_void_p]),
    ("culip_spmv_destroy", None, [C.c_void_p]),
    ("culip_sqmr_create", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SqmrOpts),
                                    C.POINTER(C.c_void_p)]),
    ("culip_sqmr_solve", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(SqmrResult),
                                   C.c_void_p]),
    ("culip_sqmr_destroy", None, [C.c_void_p]),
    ("culip_ldu_create", C.c_int, [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int32), C.c_int32, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    ("culip_ldu_set_permutation", C.c_int, [C.c_void_p, C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32)]),
    ("culip_ldu_set_scaling", C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
    ("culip_ldu_solve", C.c_int, [C.c_void_p, C.c_int, C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    ("culip_ldu_n", C.c_int32, [C.c_void_p]),
    ("culip_ldu_is_middle", C.c_int32, [C.c_void_p]),
    ("culip_ldu_destroy", None, [C.c_void_p]),
    ("culip_sqmr_create_ldu", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SqmrOpts),
                                        C.POINTER(C.c_void_p)]),
    ("culip_server_create", C.c_int, [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    ("culip_server_add_pattern", C.c_int, [C.c_void_p, C.POINTER(Csr), C.c_int32,
                                           C.POINTER(C.c_int32), C.c_int32,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("culip_server_set_pattern_solve", C.c_int, [C.c_void_p, C.c_int32, C.c_int, C.c_int32]),
    ("culip_server_submit", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("culip_server_wait", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double),
                                    C.POINTER(SqmrResult), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int32)]),
    ("culip_server_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("culip_server_destroy", None, [C.c_void_p]),
    ("culip_batch_create", C.c_int, [C.POINTER(BatchOpts), C.POINTER(C.c_void_p)]),
    ("culip_batch_add", C.c_int, [C.c_void_p, C.POINTER(Csr), C.c_int32, C.POINTER(C.c_int32),
                                  C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("culip_batch_solve", C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_double)),
                                    C.POINTER(C.POINTER(C.c_double)),
                                    C.POINTER(C.POINTER(C.c_double)), C.POINTER(SqmrResult),
                                    C.c_int32, C.c_void_p]),
    ("culip_batch_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double)]),
    ("culip_batch_destroy", None, [C.c_void_p]),
    ("culip_spmv_create_kkt_csr", C.c_int, [C.POINTER(Csr), C.POINTER(Csr), C.c_double,
                                            C.c_int32, C.POINTER(C.c_int32),
                                            C.POINTER(C.c_void_p)]),
    ("culip_spmv_equilibrate", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("culip_sqmr_set_residual_scale", C.c_int, [C.c_void_p, C.c_void_p]),
    ("culip_weighted_matching", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int32)]),
    ("culip_structural_rank", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]),
    ("culip_btf_rank_bound", C.c_int, [C.POINTER(Csr), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib():
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CulipError(-1, f"{LIB_PATH} not built (run __graft_entry__.build())")
        h = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(status):
    if status != OK:
        msg = lib().culip_last_error().decode()
        raise CulipError(status, msg)
