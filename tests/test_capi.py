"""The C-ABI library loads on any host, exports every entry point declared
in include/culip_amd.h, reports errors as status codes, and its host-side
input preparation (generator, ordering, blocking) behaves as documented."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca
from culip_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "culip_amd.h")


def test_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    declared = set(re.findall(r"\b(culip_[a-z0-9_]+)\s*\(", text))
    assert len(declared) > 30
    h = _lib.lib()
    missing = [s for s in sorted(declared) if not hasattr(h, s)]
    assert not missing, missing
    assert set(_lib.EXPORTED) <= declared


def test_errors_are_status_codes():
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    cv = np.array([1.0, 2.0])
    # a block start that is not a pivot start: 2x2 pivot straddling blocks
    g = ca.BlockiLDLt(rp, ci, cv, [0], [0, 1])
    with pytest.raises(ca.CulipError) as e:
        g.symbolic(0, 1.0)
    assert e.value.status == 1 and "pivot" in str(e.value)
    with pytest.raises(ca.CulipError) as e:
        ca.BlockiLDLt(rp, ci, cv, [0, 1], [0, 1]).symbolic(0, 0.5)
    assert "fill factor" in str(e.value)
    with pytest.raises(ca.CulipError):
        ca.BlockiLDLt(rp, np.array([0, 5], np.int32), cv, [0, 1], [0, 1])
    with pytest.raises(ca.CulipError):
        ca.Prep(rp, ci, cv, 1, 64)


def test_generator_is_deterministic_and_symmetric():
    a = ca.KKT(500, 200, 4, 3, seed=7)
    b = ca.KKT(500, 200, 4, 3, seed=7)
    c = ca.KKT(500, 200, 4, 3, seed=8)
    assert np.array_equal(a.val, b.val) and np.array_equal(a.rhs, b.rhs)
    assert not np.array_equal(a.val, c.val)
    K = sp.csr_matrix((a.val, a.col, a.row_ptr))
    assert (abs(K - K.T) > 0).nnz == 0
    n, m = 500, 200
    H = K[:n, :n]
    A = K[n:, :n]
    assert np.allclose(K[n:, n:].toarray(), -1e-8 * np.eye(m))
    assert H.nnz == n + n * 4 // 2 * 2
    assert (np.diff(A.tocsc().indptr) == 3).all()
    d = H.diagonal()
    assert d.min() >= 1.0 and d.max() <= 10.0
    off = (H - sp.diags(d)).data
    assert abs(off.std() - 0.1) < 0.02


def test_barrier_values_change_h_diagonal_only():
    k = ca.KKT(300, 100, 4, 3, seed=1)
    v = k.barrier_values(1e-4, 1e4, seed=3)
    K0 = sp.csr_matrix((k.val, k.col, k.row_ptr))
    K1 = sp.csr_matrix((v, k.col, k.row_ptr))
    D = (K1 - K0)
    assert abs(D - sp.diags(D.diagonal())).max() == 0
    w = K1.diagonal()[:300] - K0.diagonal()[:300]
    assert w.min() >= 1e-4 * (1 - 1e-12) and w.max() <= 1e4 * (1 + 1e-12)
    # log-uniform: about half the weights below the geometric midpoint 1
    assert 0.4 < np.mean(w < 1.0) < 0.6
    assert np.array_equal(K1.diagonal()[300:], K0.diagonal()[300:])
    # a second draw replaces the barrier term instead of compounding it
    k2 = ca.KKT(300, 100, 4, 3, seed=1, barrier=(1e-6, 1e6))
    v2 = k2.barrier_values(1e-4, 1e4, seed=3)
    assert np.array_equal(v2, v)


def test_generator_barrier_term():
    k0 = ca.KKT(500, 200, 4, 3, seed=4, delta=1e-10)
    k1 = ca.KKT(500, 200, 4, 3, seed=4, delta=1e-10, barrier=(1e-6, 1e6))
    assert np.array_equal(k0.rhs, k1.rhs) and np.array_equal(k0.col, k1.col)
    K0 = sp.csr_matrix((k0.val, k0.col, k0.row_ptr))
    K1 = sp.csr_matrix((k1.val, k1.col, k1.row_ptr))
    D = K1 - K0
    w = D.diagonal()[:500]
    assert abs(D - sp.diags(D.diagonal())).max() == 0 and not D.diagonal()[500:].any()
    lw = np.log10(w)
    assert lw.min() >= -6 - 1e-9 and lw.max() <= 6 + 1e-9
    assert abs(lw.mean()) < 0.6 and 2.9 < lw.std() < 4.0   # U(-6, 6): std 3.46
    # H = H0 + diag(w) stays positive definite
    assert np.linalg.eigvalsh(K1[:500, :500].toarray()).min() > 0


@pytest.mark.parametrize("order", [ca.ORDER_NATURAL, ca.ORDER_BLOCK_COLOR])
def test_prep_permutes_and_blocks(order):
    k = ca.KKT(1000, 400, 4, 3, seed=2)
    p = ca.Prep(k.row_ptr, k.col, k.val, 1000, 16, order)
    K = sp.csr_matrix((k.val, k.col, k.row_ptr))
    P = sp.csr_matrix((p.val, p.col, p.row_ptr))
    assert sorted(p.perm) == list(range(1400))
    assert (abs(K[p.perm][:, p.perm] - P) > 0).nnz == 0
    assert np.array_equal(p.permute_values(k.val), p.val)
    bs = list(p.block_starts) + [1400]
    assert all(1 <= bs[i + 1] - bs[i] <= 16 for i in range(len(bs) - 1))
    assert np.array_equal(p.piv_starts, np.arange(1400))
    # primal rows first, dual rows after, blocks never mix them
    assert (p.perm[:1000] < 1000).all()
    assert 1000 in bs
    if order == ca.ORDER_BLOCK_COLOR:
        g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
        g.symbolic(0, 1.0)
        gn = ca.BlockiLDLt(k.row_ptr, k.col, k.val, p.piv_starts,
                           ca.Prep(k.row_ptr, k.col, k.val, 1000, 16).block_starts)
        gn.symbolic(0, 1.0)
        assert g.get_int("scalars")[6] < gn.get_int("scalars")[6]


def test_matrix_market_reader(golden):
    rp, ci, cv, n = ca.read_matrix_csr(f"{golden}/sqmr_small_A.mtx")
    assert rp.size - 1 == 62 and n == 62 and rp[-1] == 342
    A = sp.csr_matrix((cv, ci, rp))
    assert (abs(A - A.T) > 0).nnz == 0
    assert all((np.diff(ci[rp[i]:rp[i + 1]]) >= 0).all() for i in range(62))
    p = ca.read_dense_vector(f"{golden}/sqmr_small_prec_p.mtx")
    assert sorted(p.astype(int)) == list(range(62))


DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "kkt_driver")


def test_cpp_driver_builds_against_the_class_header():
    """tests/cpp/kkt_driver.cpp (the reference driver's call sequence over
    include/culip_amd.hpp) is built with the library and links to it; with no
    arguments it prints its usage without touching a device."""
    import subprocess
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_server_needs_a_device():
    """Without a HIP device the request loop refuses to start with a status
    code (no device work is attempted)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ca.CulipError) as e:
        ca.Server(slots=2)
    assert e.value.status == 8


def test_batch_needs_a_device_and_checks_options():
    """The batch refuses to start without a HIP device (NO_DEVICE), and its
    C entry points reject null arguments with INVALID before any device
    work."""
    import ctypes as C
    import torch
    from culip_amd import _lib
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ca.CulipError) as e:
        ca.Batch(prec="f32", jacobi_sweeps=2)
    assert e.value.status == 8
    L = _lib.lib()
    h = C.c_void_p()
    assert L.culip_batch_create(None, C.byref(h)) == 1
    assert L.culip_batch_solve(None, None, None, None, None, 0, None) == 1
    assert L.culip_batch_add(None, None, 0, None, 0, None, None) == 1
    assert L.culip_batch_info(None, None, None, None) == 1
    L.culip_batch_destroy(None)


def test_imported_factor_checks_inputs_before_the_device():
    """ImportLDUPreconditioner: malformed inputs are rejected on the host
    with INVALID; well-formed ones need a device (NO_DEVICE here)."""
    import torch
    L = (np.array([0, 1, 3], np.int32), np.array([0, 0, 1], np.int32), np.array([2.0, 1.0, 3.0]))
    with pytest.raises(ca.CulipError) as e:
        ca.ImportLDUPreconditioner(L, tri=np.zeros(6))          # D without pivot flags
    assert e.value.status == 1
    rect = (np.array([0, 1, 2], np.int32), np.array([0, 2], np.int32), np.array([1.0, 1.0]))
    with pytest.raises(ca.CulipError) as e:
        ca.ImportLDUPreconditioner((rect[0], rect[1], rect[2]), U=L)
    assert e.value.status == 1                                   # column out of range
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ca.CulipError) as e:
        ca.ImportLDUPreconditioner(L)
    assert e.value.status == 8


def test_permute_scale_matches_the_reference_convention():
    """culip_csr_permute_scale with a new -> old vector gives B(i, j) =
    A(p[i], p[j]) (host_permute with perm_is_old_to_new = false,
    block_ildlt_invariants.h:29-69), columns sorted; an old -> new vector
    is its inverse; scales divide entries by s_r s_c of the new (or, before
    matching, the old) indices. Checked against scipy."""
    rng = np.random.default_rng(2)
    n = 30
    A = sp.random(n, n, density=0.2, random_state=3, format="csr")
    A = (A + A.T).tocsr()
    A.sort_indices()
    p = rng.permutation(n).astype(np.int32)
    rp, ci, cv = ca.permute_scale(A.indptr, A.indices, A.data, p)
    B = sp.csr_matrix((cv, ci, rp), shape=(n, n)).toarray()
    np.testing.assert_array_equal(B, A.toarray()[np.ix_(p, p)])
    assert all((np.diff(ci[rp[i]:rp[i + 1]]) > 0).all() for i in range(n))
    inv = np.argsort(p).astype(np.int32)
    rp2, ci2, cv2 = ca.permute_scale(A.indptr, A.indices, A.data, inv, perm_is_old_to_new=True)
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(ci2, ci)
    np.testing.assert_array_equal(cv2, cv)
    s = rng.uniform(0.5, 2.0, n)
    rp3, ci3, cv3 = ca.permute_scale(A.indptr, A.indices, A.data, p, row_scale=s, col_scale=s)
    B3 = sp.csr_matrix((cv3, ci3, rp3), shape=(n, n)).toarray()
    np.testing.assert_allclose(B3, A.toarray()[np.ix_(p, p)] / np.outer(s, s), rtol=1e-15)
    rp4, ci4, cv4 = ca.permute_scale(A.indptr, A.indices, A.data, p, row_scale=s, col_scale=s,
                                     scale_before_match=True)
    B4 = sp.csr_matrix((cv4, ci4, rp4), shape=(n, n)).toarray()
    np.testing.assert_allclose(B4, (A.toarray() / np.outer(s, s))[np.ix_(p, p)], rtol=1e-15)
    with pytest.raises(ca.CulipError):
        ca.permute_scale(A.indptr, A.indices, A.data, np.zeros(n, np.int32))
