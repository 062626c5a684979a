"""Pins of the oracle's operator and mobility solve (CPU only).

Closed forms: isolated sphere (Stokes drag and rotation), Stimson & Jeffery's
exact two-sphere solution (Proc. R. Soc. A 111, 1926); invariances of the
physics: rigid motions, permutations, symmetry of the mobility matrix.
"""
import numpy as np
import pytest

import pyoracle as O
from conftest import rel

ID = np.array([1.0, 0.0, 0.0, 0.0])


def quat_of(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])


def rotmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def stimson_jeffery(d, R=1.0, terms=300):
    """U / U_Stokes for two equal spheres translating along their line of centres."""
    al = np.arccosh(d / (2 * R))
    s = 0.0
    for n in range(1, terms):
        if (2 * n + 1) * al > 600:  # remaining terms are below double precision
            break
        s += n * (n + 1) / ((2 * n - 1) * (2 * n + 3)) * (
            1 - (4 * np.sinh((n + .5) * al) ** 2 - (2 * n + 1) ** 2 * np.sinh(al) ** 2)
            / (2 * np.sinh((2 * n + 1) * al) + (2 * n + 1) * np.sinh(2 * al)))
    return 1.0 / (4.0 / 3.0 * np.sinh(al) * s)


@pytest.mark.parametrize("R,mu,p", [(1.0, 1.0, 6), (0.7, 2.5, 8)])
def test_isolated_sphere(R, mu, p):
    M = O.MobilityOperator([[0.3, -1.0, 2.0]], [R], [quat_of([1, 2, 3], 0.7)], [p], viscosity=mu)
    F = np.array([1.0, -2.0, 0.5, 0.3, 0.9, -0.4])
    v, info = M.solve(F)
    assert info["shortcut"] == 1  # b = -(I/2 + K') rho0 = 0 exactly (MobMat.cpp:216-219)
    assert rel(v[:3], F[:3] / (6 * np.pi * mu * R)) < 1e-13
    assert rel(v[3:], F[3:] / (8 * np.pi * mu * R ** 3)) < 1e-13


@pytest.mark.parametrize("d,p,tol", [(4.0, 12, 1e-10), (3.0, 12, 1e-10), (2.5, 16, 1e-9), (2.2, 16, 2e-7)])
def test_stimson_jeffery(d, p, tol):
    M = O.MobilityOperator([[0, 0, 0], [0, 0, d]], [1, 1], [ID, ID], [p, p])
    F = np.array([0, 0, 1.0, 0, 0, 0] * 2)
    v, info = M.solve(F, tol=1e-12)
    assert info["converged"] == 1
    exact = stimson_jeffery(d)
    assert abs(v[2] * 6 * np.pi - exact) / exact < tol
    assert abs(v[8] * 6 * np.pi - exact) / exact < tol


def _small_config(rng, n=4, p=8):
    pos = np.array([[0, 0, 0], [2.6, 0.4, 0.1], [-0.5, 2.7, 0.6], [3.0, 3.5, -1.0]])[:n]
    rad = np.array([1.0, 0.9, 1.1, 0.8])[:n]
    quat = np.array([quat_of(rng.standard_normal(3), rng.uniform(0, np.pi)) for _ in range(n)])
    return pos, rad, quat, [p] * n


def test_rigid_motion_covariance():
    rng = np.random.default_rng(5)
    pos, rad, quat, order = _small_config(rng)
    F = rng.standard_normal(24)
    M = O.MobilityOperator(pos, rad, quat, order)
    v, _ = M.solve(F, tol=1e-12)
    qg = quat_of([0.3, -0.8, 0.5], 1.1)
    Rg = rotmat(qg)
    shift = np.array([3.0, -7.0, 11.0])
    pos2 = pos @ Rg.T + shift
    quat2 = np.array([qmul(qg, q) for q in quat])
    F2 = (F.reshape(-1, 2, 3) @ Rg.T).ravel()
    M2 = O.MobilityOperator(pos2, rad, quat2, order)
    v2, _ = M2.solve(F2, tol=1e-12)
    assert rel(v2, (v.reshape(-1, 2, 3) @ Rg.T).ravel()) < 1e-10
    # permutation of the spheres permutes the answer
    perm = [2, 0, 3, 1]
    M3 = O.MobilityOperator(pos[perm], rad[perm], quat[perm], order)
    v3, _ = M3.solve(F.reshape(-1, 6)[perm].ravel(), tol=1e-12)
    assert rel(v3, v.reshape(-1, 6)[perm].ravel()) < 1e-10


def test_mobility_matrix_symmetric():
    """The grand mobility matrix is symmetric (Lorentz reciprocity) up to the
    discretisation error, which shrinks with p."""
    rng = np.random.default_rng(9)
    pos, rad, quat, _ = _small_config(rng, n=3)
    errs = []
    for p in (8, 12):
        M = O.MobilityOperator(pos, rad, quat, [p] * 3)
        cols = [M.solve(e, tol=1e-12)[0] for e in np.eye(18)]
        A = np.array(cols).T
        errs.append(np.abs(A - A.T).max() / np.abs(A).max())
    assert errs[1] < 1e-5
    assert errs[1] < errs[0]


def test_zero_force_shortcut_and_nonconvergence():
    rng = np.random.default_rng(2)
    pos, rad, quat, order = _small_config(rng)
    M = O.MobilityOperator(pos, rad, quat, order)
    v, info = M.solve(np.zeros(24))
    assert info["shortcut"] == 1 and np.all(v == 0)
    v, info = M.solve(rng.standard_normal(24), max_iters=2)
    assert info["converged"] == 0 and info["iters"] == 2


def test_operator_p2p_decomposition(pos_initial):
    """apply = p2p(cId, cSL, cTrac) + lop(cLOP) with Y = beta Y + alpha (...)."""
    pos, rad = pos_initial
    n = 20
    quat = np.tile(ID, (n, 1))
    op = O.Operator(pos[:n], rad[:n], quat, [6] * n, 0.5, 0.0, 1.0, 1.0)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(op.dof)
    y0 = rng.standard_normal(op.dof)
    ref = op.apply_p2p(x, cId=0.5, cSL=0.0, cTrac=1.0)
    ref = op.apply_lop(x, ref, 1.0)
    y = op.apply(x, y0.copy(), alpha=2.0, beta=-0.5)
    assert rel(y, 2.0 * ref - 0.5 * y0) < 1e-14


def test_operator_p2p_on_sphere_ranges(pos_initial):
    """The sphere-range apply (used at full problem sizes on sampled spheres)
    equals the full apply on the points of its spheres, bit for bit, and leaves
    every other entry untouched; empty and clipped ranges are no-ops or clipped."""
    pos, rad = pos_initial
    n = 30
    rng = np.random.default_rng(4)
    quat = np.array([quat_of(rng.standard_normal(3), rng.uniform(0, 3)) for _ in range(n)])
    op = O.Operator(pos[:n], rad[:n], quat, [5, 6, 7] * (n // 3))
    _, _, _, off = op.grid()
    x = rng.standard_normal(op.dof)
    for coeffs in ((0.5, 0.0, 1.0), (0.0, 1.0, 0.0), (0.3, 0.7, -0.4)):
        full = op.apply_p2p(x, None, *coeffs)
        for T0, T1 in ((0, 4), (13, 14), (25, 30), (0, n)):
            a, b = 3 * off[T0], 3 * off[T1]
            got = op.apply_p2p_spheres(x, T0, T1, None, *coeffs)
            assert np.array_equal(got[a:b], full[a:b])
            assert not got[:a].any() and not got[b:].any()
            seed = rng.standard_normal(op.dof)
            acc = op.apply_p2p_spheres(x, T0, T1, seed.copy(), *coeffs)
            assert np.abs(acc[a:b] - seed[a:b] - full[a:b]).max() <= 1e-13 * np.abs(full).max()
            assert np.array_equal(acc[:a], seed[:a]) and np.array_equal(acc[b:], seed[b:])
    assert np.array_equal(op.apply_p2p_spheres(x, 5, 5, np.ones(op.dof), 0.5, 0.0, 1.0), np.ones(op.dof))
    clip = op.apply_p2p_spheres(x, n - 2, n + 10, None, 0.5, 0.0, 1.0)
    assert np.array_equal(clip, op.apply_p2p_spheres(x, n - 2, n, None, 0.5, 0.0, 1.0))
