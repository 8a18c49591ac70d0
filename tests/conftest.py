import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def sym_problem(n_blocks=6, bmax=8, density=0.15, two_by_two=True, seed=1, diag_shift=4.0):
    """Random symmetric indefinite test matrix with a contiguous blocking and
    a static 1x1/2x2 pivot structure (pivots never straddle blocks)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(2, bmax + 1, size=n_blocks)
    bs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    n = int(sizes.sum())
    A = np.zeros((n, n))
    mask = rng.random((n, n)) < density
    vals = rng.normal(size=(n, n))
    A[mask] = vals[mask]
    A = np.triu(A, 1)
    A = A + A.T
    d = rng.normal(size=n) * diag_shift
    d[np.abs(d) < 0.5] += 2.0
    A[np.arange(n), np.arange(n)] = d
    is_piv = np.ones(n, dtype=np.int32)
    if two_by_two:
        for b, s in zip(bs, sizes):
            r = b
            while r < b + s:
                if r + 1 < b + s and rng.random() < 0.35:
                    is_piv[r + 1] = 0
                    A[r, r + 1] = A[r + 1, r] = rng.normal() * 3.0
                    r += 2
                else:
                    r += 1
    piv = np.nonzero(is_piv)[0].astype(np.int32)
    rp = [0]
    ci, cv = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci.extend(nz)
        cv.extend(A[i, nz])
        rp.append(len(ci))
    return (np.array(rp, np.int32), np.array(ci, np.int32), np.array(cv)), piv, bs, A


@pytest.fixture(scope="session")
def hip_device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a HIP device")
    return torch.device("cuda:0")


def _read(path):
    import culip_amd as ca
    return ca.read_matrix_csr(path)


def ldu_problem(golden, prefix):
    """A sqmr_{small,big} fixture: A, L, tridiagonal D + is_piv, p, s."""
    import culip_amd as ca
    ar, ac, av, an = _read(f"{golden}/{prefix}_A.mtx")
    lr, lc, lv, ln = _read(f"{golden}/{prefix}_prec_L.mtx")
    dr, dc, dv, dn = _read(f"{golden}/{prefix}_prec_D.mtx")
    p = ca.read_dense_vector(f"{golden}/{prefix}_prec_p.mtx").astype(np.int32)
    s = ca.read_dense_vector(f"{golden}/{prefix}_prec_s.mtx")
    n = an
    # D to tridiagonal + is_piv (test_sqmr.cc:103-152)
    tri = np.zeros(3 * n)
    is_piv = np.ones(n, np.int32)
    i = 0
    while i < n:
        ln_i = dr[i + 1] - dr[i]
        if ln_i == 1:
            if i > 0:
                tri[3 * (i - 1) + 2] = 0
            tri[3 * i + 1] = dv[dr[i]]
            if i < n - 1:
                tri[3 * (i + 1)] = 0
            i += 1
        else:
            iv, ipv = dv[dr[i]:dr[i + 1]], dv[dr[i + 1]:dr[i + 2]]
            tri[3 * i], tri[3 * i + 1], tri[3 * i + 2] = 0, iv[0], iv[1]
            tri[3 * (i + 1)], tri[3 * (i + 1) + 1], tri[3 * (i + 1) + 2] = ipv[0], ipv[1], 0
            is_piv[i + 1] = 0
            i += 2
    return (ar, ac, av), (lr, lc, lv), tri, is_piv, p, s, n


def _dense_csr(A):
    rp, ci, cv = [0], [], []
    for i in range(A.shape[0]):
        nz = np.nonzero(A[i])[0]
        ci.extend(nz)
        cv.extend(A[i, nz])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(cv, dtype=np.float64)


def edge_problem(name):
    """Small edge cases: (csr, piv_starts, block_starts, dense A).
    chain-b1: tridiagonal with one-row blocks, so the level count equals the
      block count (the reference's one-past-end case at
      block_ildlt.impl.cuh:775, 789-791);
    one-1x1 / one-2x2: a single row / a single 2x2 pivot;
    diagonal: D = diag(A), U = I;
    dense-32: one full 32 x 32 block with 2x2 pivots;
    zero-pivots: zero diagonal entries that hit the tiny-pivot rules."""
    rng = np.random.default_rng(sum(map(ord, name)))
    if name == "chain-b1":
        n = 40
        A = np.diag(rng.uniform(2, 4, n)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
        piv, bs = np.arange(n), np.arange(n)
    elif name == "one-1x1":
        A = np.array([[3.0]])
        piv, bs = np.array([0]), np.array([0])
    elif name == "one-2x2":
        A = np.array([[1e-3, 2.0], [2.0, 1e-3]])
        piv, bs = np.array([0]), np.array([0])
    elif name == "diagonal":
        n = 50
        A = np.diag(rng.uniform(-3, 3, n) + 0.5)
        piv, bs = np.arange(n), np.arange(0, n, 8)
    elif name == "dense-32":
        n = 32
        Q = rng.normal(size=(n, n))
        A = Q + Q.T + np.diag(np.linspace(-8, 8, n))
        piv = np.array([0, 2, 3, 5, 6] + list(range(8, 32)))
        bs = np.array([0])
    elif name == "zero-pivots":
        n = 48
        A = np.zeros((n, n))
        for i in range(n - 1):
            if rng.random() < 0.3:
                A[i, i + 1] = A[i + 1, i] = rng.normal()
        d = rng.normal(size=n) * 3
        d[::5] = 0.0
        A[np.arange(n), np.arange(n)] = d
        piv, bs = np.arange(n), np.arange(0, n, 12)
    else:
        raise KeyError(name)
    return _dense_csr(A), np.asarray(piv, np.int32), np.asarray(bs, np.int32), A


EDGE_CASES = ["chain-b1", "one-1x1", "one-2x2", "diagonal", "dense-32", "zero-pivots"]
