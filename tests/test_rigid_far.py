"""Closed form of the traction far field of a rigidly moving sphere's force
density (CPU).  rho0 = F/(4 pi R^2) + 3/(8 pi R^4) T x (R n) (the reference's
initial density, BIE/SphereSTKMobMat.cpp:174-192) generates outside its sphere
exactly the flow of a Stokeslet F, its degenerate quadrupole (R^2/6) lap G F
and a rotlet T; K'[rho0](x) is that flow's traction on a surface with normal n
at x.  The HIP path evaluates it at the sphere centre for distant pairs; here
it is checked against the brute-force grid sum of the traction kernel
-(3/4 pi) r (r.f)(r.n)/r^5 (Shexp_test.cpp:357-397) on an order-16 grid."""
import numpy as np

import pyoracle as O


def rigid_traction_far(x, n, c, R, F, T):
    r = x - c
    d = np.linalg.norm(r, axis=1)[:, None]
    rn = np.sum(r * n, 1)[:, None]
    rF = r @ F
    rF = rF[:, None]
    Fn = (n @ F)[:, None]
    Txr = np.cross(T, r)
    Txr_n = np.sum(Txr * n, 1)[:, None]
    tS = -3.0 / (4 * np.pi) * r * rF * rn / d ** 5
    tD = R * R / (12 * np.pi) * (-3.0 * (F * rn + r * Fn + n * rF) / d ** 5 + 15.0 * rF * rn * r / d ** 7)
    tR = -3.0 / (8 * np.pi) * (Txr * rn + r * Txr_n) / d ** 5
    return tS + tD + tR


def test_rigid_sphere_traction_far_field_closed_form():
    rng = np.random.default_rng(3)
    c = np.array([0.3, -0.2, 0.1])
    R = 1.3
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    F = rng.standard_normal(3)
    T = rng.standard_normal(3)
    op = O.Operator(c[None], [R], q[None], np.array([16], dtype=np.int32))
    y, ny, w, _ = op.grid()
    rho0 = F / (4 * np.pi * R * R) + 3.0 / (8 * np.pi * R ** 4) * np.cross(T, R * ny)
    f = w[:, None] * rho0
    dirs = rng.standard_normal((40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    x = c + dirs * rng.uniform(4 * R, 12 * R, 40)[:, None]
    n = rng.standard_normal((40, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    r = x[:, None, :] - y[None, :, :]
    d = np.linalg.norm(r, axis=2)
    s = np.einsum("tsk,sk->ts", r, f) * np.einsum("tsk,tk->ts", r, n) / d ** 5
    brute = -3.0 / (4 * np.pi) * np.einsum("ts,tsk->tk", s, r)
    closed = rigid_traction_far(x, n, c, R, F, T)
    err = np.abs(closed - brute).max() / np.abs(brute).max()
    assert err < 1e-12, err
    # each part matters: dropping the quadrupole or the rotlet fails
    assert np.abs(closed - brute).max() < 1e-3 * np.abs(rigid_traction_far(x, n, c, 0.0, F, T) - brute).max()
    assert np.abs(closed - brute).max() < 1e-3 * np.abs(rigid_traction_far(x, n, c, R, F, 0 * T) - brute).max()
