"""Exiting Python with live library objects: they are released by the
package's atexit hook while the library (and, on a GPU, the HIP runtime)
is intact, and no destructor runs during module teardown."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HOST_SCRIPT = r"""
import sys
sys.path.insert(0, {root!r})
import culip_amd as ca
k = ca.KKT(300, 120, 4, 3, seed=0)
p = ca.Prep(k.row_ptr, k.col, k.val, 300, 16, ca.ORDER_BLOCK_COLOR)
g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
g.symbolic(0, 1.0)
cycle = [k, p, g]
cycle.append(cycle)          # a reference cycle: only the final collection frees it
print("objects live at exit", flush=True)
"""

GPU_SCRIPT = r"""
import sys
sys.path.insert(0, {root!r})
import torch
import culip_amd as ca
k = ca.KKT(2000, 800, 4, 3, seed=0)
p = ca.Prep(k.row_ptr, k.col, k.val, 2000, 16, ca.ORDER_BLOCK_COLOR)
g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
g.compute(0, 1.0, 1e-4)
A = ca.CSRMatrixSPMV(p.row_ptr, p.col, p.val)
s = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
b = torch.tensor(k.rhs[p.perm], device="cuda:0")
x = torch.zeros_like(b)
s.solve(b, x)
srv = ca.Server(slots=2, prec="f32")
pid = srv.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
srv.wait(srv.submit(pid, p.val, k.rhs[p.perm]))
rid = srv.submit(pid, p.val, k.rhs[p.perm])   # left queued or running at exit
cycle = [k, p, g, A, s, srv, x]
cycle.append(cycle)
print("objects live at exit", flush=True)
"""


def _run(script):
    r = subprocess.run([sys.executable, "-c", script.format(root=ROOT)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "objects live at exit" in r.stdout
    for bad in ("Exception ignored", "Traceback", "Segmentation", "Aborted"):
        assert bad not in r.stderr, r.stderr[-2000:]


def test_exit_with_live_host_objects():
    _run(HOST_SCRIPT)


@pytest.mark.gpu
def test_exit_with_live_device_objects(hip_device):
    _run(GPU_SCRIPT)
