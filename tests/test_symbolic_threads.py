"""The host symbolic phase splits its per-block-row passes over threads
(CULIP_HOST_THREADS); every structure it produces, the Schur-update pair
lists included, is the same bit for bit for one thread and for several."""
import json
import os
import subprocess
import sys

import pytest

from test_symbolic import NAMES

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CODE = """
import hashlib, json, sys
import numpy as np
import culip_amd as ca
names = json.loads(sys.argv[1])
out = {}
for n, B, fl, ff, order in ((30000, 16, 0, 1.0, ca.ORDER_BLOCK_COLOR),
                            (6000, 16, 1, 1.5, ca.ORDER_NATURAL),
                            (9000, 32, 2, 1.2, ca.ORDER_BLOCK_COLOR)):
    k = ca.KKT(n, int(0.4 * n), 4, 3, seed=5)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, B, order)
    g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
    g.symbolic(fl, ff)
    h = hashlib.sha256()
    for nm in names + ["upd_off", "upd_ki", "upd_kj", "scalars"]:
        h.update(np.ascontiguousarray(g.get_int(nm)).tobytes())
    h.update(np.ascontiguousarray(g.get_real("val_store")).tobytes())
    out[f"{n}-{fl}"] = [h.hexdigest(), int(g.get_int("upd_off")[-1])]
print(json.dumps(out))
"""


def _digests(threads):
    env = dict(os.environ, CULIP_HOST_THREADS=str(threads), PYTHONPATH=REPO)
    r = subprocess.run([sys.executable, "-c", CODE, json.dumps(NAMES)], env=env,
                       capture_output=True, text=True, cwd=REPO, timeout=600)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_symbolic_same_for_any_thread_count():
    one, many = _digests(1), _digests(7)
    assert one == many
    assert all(v[1] > 0 for v in one.values())
