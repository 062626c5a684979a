"""GPU parity of the steps beside the solve, through the C ABI, against the
oracle on identical inputs: Brownian velocities (SRC/SphereSystem.cpp:323-360)
to 1e-14 of their scale (the Philox words are integer-exact; log, sin and cos
differ from the host's by an ulp or so) and collision pairs
(Sphere/CollisionSphere.hpp:51-90) with the same pairs in the same order and
phi0 and normals to 4e-16 (the same FP64 operations in the same order)."""
import sys

import numpy as np
import pytest

import pyoracle as O
import spheresimulator_amd as S

pytestmark = pytest.mark.gpu


def say(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,seed,step", [(1, 0, 0), (1000, 7, 3), (100003, 2 ** 40 + 5, 2 ** 33 + 1)])
def test_brownian_matches_oracle(n, seed, step):
    say("brownian n=%d seed=%d step=%d" % (n, seed, step))
    rad = np.random.default_rng(n).uniform(0.5, 1.5, n)
    args = dict(dt=0.02, kBT=0.0041, viscosity=0.9, scale=1.1, seed=seed, step=step)
    v = S.system.brownian_velocity(rad, **args)
    w = O.brownian_velocity(rad, **args)
    f = np.sqrt(2 * 0.0041 / 0.02) * 1.1
    sc = np.repeat(np.stack([np.sqrt(1 / (6 * np.pi * 0.9 * rad)), np.sqrt(1 / (8 * np.pi * 0.9 * rad ** 3))], 1),
                   3, axis=1).ravel() * f
    assert np.abs(v - w).max() <= 1e-14 * sc.max()
    assert np.abs((v - w) / sc).max() < 1e-13


def test_brownian_empty_and_errors():
    say("brownian empty and invalid")
    assert S.system.brownian_velocity([], dt=0.1, kBT=1.0).size == 0
    with pytest.raises(S.SphstkError):
        S.system.brownian_velocity([1.0, -1.0], dt=0.1, kBT=1.0)
    with pytest.raises(S.SphstkError):
        S.system.brownian_velocity([1.0], dt=0.0, kBT=1.0)


def _check_pairs(pos, rc, cb, gid=None):
    got = S.system.collision_pairs(pos, rc, cb, gid=gid)
    ii, jj, phi0, nrm = O.collision_pairs(pos, rc, cb, gid=gid)
    assert np.array_equal(got["idx_i"], ii) and np.array_equal(got["idx_j"], jj)
    assert np.allclose(got["phi0"], phi0, rtol=0, atol=4e-16 * max(1.0, np.abs(phi0).max(initial=0)))
    assert np.allclose(got["norm_i"], nrm, rtol=0, atol=4e-16)
    return got


def test_collision_pairs_random_polydisperse():
    say("collision pairs 3000 polydisperse spheres, permuted gid")
    rng = np.random.default_rng(11)
    n = 3000
    pos = rng.uniform(0, 30, (n, 3))
    rc = rng.uniform(0.3, 1.2, n)
    got = _check_pairs(pos, rc, 0.3, gid=rng.permutation(n))
    assert len(got["idx_i"]) > n
    assert np.array_equal(got["gamma"], np.maximum(-got["phi0"], 0))


def test_collision_pairs_lattice_sparse_and_edge_cases():
    say("collision pairs lattice, sparse clusters, n<2, no pairs")
    k, d = 8, 2.25
    g = np.arange(k)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * d
    got = _check_pairs(pos, np.ones(len(pos)), 0.3)
    assert len(got["idx_i"]) == 3 * k * k * (k - 1)
    # two tight clusters 1e6 apart: the cell edge grows, the pairs stay exact
    rng = np.random.default_rng(5)
    a = rng.uniform(0, 6, (200, 3))
    pos = np.concatenate([a, a[:100] + 1e6])
    _check_pairs(pos, np.full(300, 0.5), 0.3)
    assert len(S.system.collision_pairs(np.zeros((1, 3)), [1.0], 0.3)["idx_i"]) == 0
    assert len(S.system.collision_pairs(np.zeros((0, 3)), [], 0.3)["idx_i"]) == 0
    far = np.array([[0.0, 0, 0], [10.0, 0, 0]])
    assert len(S.system.collision_pairs(far, [1.0, 1.0], 0.3)["idx_i"]) == 0
    # a capacity below the count returns the count and writes nothing (the binding then retries)
    import ctypes
    L = S.load()
    pos = np.ascontiguousarray(np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * d, dtype=float)
    rc = np.ones(len(pos))
    cnt = ctypes.c_int64()
    dp = ctypes.POINTER(ctypes.c_double)
    ii = np.full(4, -7, dtype=np.int32)
    rc_ = L.sphstk_collision_pairs(len(rc), pos.ctypes.data_as(dp), rc.ctypes.data_as(dp), None, 0.3, 4,
                                   ctypes.byref(cnt), ii.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None,
                                   None)
    assert rc_ == 0 and cnt.value == 3 * k * k * (k - 1) and (ii == -7).all()


def test_collision_pairs_cfg5_packing():
    """The 16384-sphere clustered packing (log-normal radii, vf 0.2): the
    collision radius = the hydrodynamic one."""
    say("collision pairs 16384-sphere clustered packing")
    c = S.configs.config5()
    _check_pairs(c["pos"], c["radius"], 0.3)


def test_collision_lcp_hydro_matches_oracle():
    """The collision LCP with the hydrodynamic mobility (one GPU mobility solve per
    BBPGD iteration, SphereSTKMobMat::apply inside the LCP) against the oracle's
    BBPGD over the oracle's mobility, and against the exact LCP solution from the
    dense matrix dt D^T M D (every active set tried).  A compressed chain of six
    spheres with gaps inside the collision buffer."""
    say("collision LCP, hydrodynamic mobility, 6-sphere chain")
    rng = np.random.default_rng(4)
    n = 6
    pos = np.array([[0, 0, 0], [1.72, 0, 0], [0.86, 1.5, 0], [3.44, 0.1, 0], [5.16, 0, 0.1], [6.9, 0, 0]], float)
    pos += rng.uniform(-0.02, 0.02, pos.shape)
    rad = np.full(n, 0.8)
    quat = np.tile([1.0, 0, 0, 0], (n, 1))
    order = np.full(n, 6, dtype=np.int32)
    F = np.zeros((n, 6))
    F[:, :3] = -(pos - pos.mean(0)) * 10.0  # squeezes the chain
    F = F.ravel()
    dt = 0.5
    mob = S.SphereSTKMobMat(pos, rad, quat, order, params=S.solver_params(tol=1e-10))
    vk, _ = mob.solveMob(F)
    pairs = S.system.collision_pairs(pos, rad, 0.3)
    m = len(pairs["idx_i"])
    assert m >= 5 and (pairs["phi0"] > 0).all()
    g, Fc, Uc, info = S.system.collision_lcp(pairs, vk, dt, mob=mob, tol=1e-9)
    assert info["converged"] == 1 and info["mobility_solves"] >= info["iters"]
    assert (g > 0).sum() >= 2
    ref = O.MobilityOperator(pos, rad, quat, order)
    vko, _ = ref.solve(F, tol=1e-10)
    assert np.abs(vk - vko).max() < 1e-9 * np.abs(vko).max()

    def Mref(f):
        return ref.solve(f, tol=1e-10)[0]

    go, Fo, Uo, io = O.lcp_bbpgd(Mref, n, pairs["idx_i"], pairs["idx_j"], pairs["phi0"], pairs["norm_i"], vko, dt,
                                 tol=1e-9)
    assert io["converged"] == 1
    assert np.abs(g - go).max() < 1e-6 * np.abs(go).max()
    assert np.abs(Uc - Uo).max() < 1e-6 * np.abs(Uo).max()
    # exact LCP solution of the dense system (positive definite, not symmetric at
    # this resolution: a P-matrix, so the solution is unique)
    ii, jj, nrm = pairs["idx_i"], pairs["idx_j"], pairs["norm_i"]
    A = np.zeros((m, m))
    for k in range(m):
        f = np.zeros((n, 6))
        f[ii[k], :3] += nrm[k]
        f[jj[k], :3] -= nrm[k]
        u = Mref(f.ravel()).reshape(n, 6)[:, :3]
        A[:, k] = dt * np.einsum("kc,kc->k", nrm, u[ii] - u[jj])
    v = vko.reshape(n, 6)[:, :3]
    b = pairs["phi0"] + dt * np.einsum("kc,kc->k", nrm, v[ii] - v[jj])
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    want = O.lcp_enumerate(A, b)
    assert np.abs(g - want).max() < 1e-6 * np.abs(want).max()


@pytest.mark.parametrize("hydro", [False, True])
def test_resolve_collision_step(hydro):
    """SphereSystem::resolveCollision end to end (known velocity + Brownian term,
    pairs, LCP): the returned velocities keep every linearised gap
    phi0 + dt D^T vel non-negative to the LCP tolerance, and pairs that push
    (gamma > 0) close to contact; then one Euler step."""
    say("resolve collision, hydro=%s" % hydro)
    rng = np.random.default_rng(8)
    n = 27
    g = np.arange(3)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * 1.72 + rng.uniform(-0.02, 0.02, (n, 3))
    rad = np.full(n, 0.8)
    F = np.zeros((n, 6))
    F[:, :3] = -(pos - pos.mean(0)) * 10.0
    F[:, 3:] = rng.standard_normal((n, 3))
    dt = 0.5
    mob = None
    if hydro:
        mob = S.SphereSTKMobMat(pos, rad, np.tile([1.0, 0, 0, 0], (n, 1)), np.full(n, 6, dtype=np.int32))
    vel, pairs, gamma, fcol, info = S.system.resolve_collision(pos, rad, F.ravel(), dt, 0.3, mob=mob, radius=rad,
                                                               kBT=0.0041, scale_brown=1.0, seed=3, step=1,
                                                               tol=1e-7)
    assert info["converged"] == 1 and (gamma > 0).any()
    v = vel.reshape(n, 6)[:, :3]
    ii, jj, nrm = pairs["idx_i"], pairs["idx_j"], pairs["norm_i"]
    phi = pairs["phi0"] + dt * np.einsum("kc,kc->k", nrm, v[ii] - v[jj])
    assert phi.min() > -1e-6
    assert np.abs(phi[gamma > 0]).max() < 1e-6
    p2, q2 = S.system.step_euler(pos, np.tile([1.0, 0, 0, 0], (n, 1)), vel, dt)
    assert np.isfinite(p2).all() and np.allclose(np.linalg.norm(q2, axis=1), 1)
