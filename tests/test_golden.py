"""The committed FP64 reference of the 16384-sphere workload
(tests/golden/cfg5_fp64_reference.npz, made on the GPU by
tests/golden/make_cfg5_fp64_reference.py): it loads without pickle, covers
every sphere with finite velocities, records a converged GMRES(30) solve, and
its apply count is what GMRES(30) with Belos' explicit residual test spends
on that many iterations: the right-hand side, one apply per iteration, one
explicit residual per restart cycle, and the final single layer."""
import json
import math
import os

import numpy as np

from conftest import GOLDEN


def test_cfg5_fp64_reference_fixture():
    r = np.load(os.path.join(GOLDEN, "cfg5_fp64_reference.npz"), allow_pickle=False)
    meta = json.loads(str(r["meta"]))
    v = r["vel"]
    assert v.shape == (6 * 16384,) and np.isfinite(v).all()
    assert meta["fp64_relres"] <= 1e-8
    it = meta["fp64_iters"]
    assert meta["fp64_applies"] == 1 + it + math.ceil(it / 30) + 1
    assert meta["mixed_rel_err_V"] < 1e-7 and meta["mixed_rel_err_Omega"] < 1e-7
    # rigid-body velocities of forced spheres: translational speeds of the
    # order of the isolated-sphere drag |F| / (6 pi R) with |F| ~ sqrt(3)
    V = v.reshape(-1, 6)[:, :3]
    s = np.linalg.norm(V, axis=1)
    assert 0.01 < np.median(s) * 6 * math.pi < 10
