"""Pins of the oracle's spherical-harmonic layer (CPU only).

Each check is fixed by mathematics independent of the oracle's own formulas:
library quadrature, closed forms, finite differences, brute-force quadrature
of the Stokeslet/stresslet (the method of Sphere/Shexp_test.cpp:109-524).
"""
import math

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

import pyoracle as O
from conftest import relmax


def tri(n, m):
    return n * (n + 1) // 2 + m


@pytest.mark.parametrize("n", [1, 2, 5, 9, 13, 17])
def test_gauss_legendre_matches_library(n):
    x, w = O.gl(n)
    xr, wr = leggauss(n)
    assert np.abs(x - xr).max() < 1e-14
    assert np.abs(w - wr).max() < 1e-14


@pytest.mark.parametrize("p", [4, 8, 12])
def test_unit_grid_moments_and_order(p):
    eta, w = O.unit_grid(p)
    assert eta.shape == ((p + 1) * (2 * p + 2), 3)
    assert abs(w.sum() - 4 * np.pi) < 1e-12
    assert np.abs(w @ eta).max() < 1e-13
    assert np.abs(eta.T @ (w[:, None] * eta) - 4 * np.pi / 3 * np.eye(3)).max() < 1e-13
    # north -> south rings, phi_k = 2 pi k / (2p+2)   (Sphere/Shexp.cpp:231-243)
    rings = eta[:, 2].reshape(p + 1, 2 * p + 2)
    assert np.all(np.diff(rings[:, 0]) < 0)
    phi = np.arctan2(eta[: 2 * p + 2, 1], eta[: 2 * p + 2, 0]) % (2 * np.pi)
    assert np.abs(phi - 2 * np.pi * np.arange(2 * p + 2) / (2 * p + 2)).max() < 1e-12


def test_solid_harmonics_closed_forms():
    x = np.array([0.3, -0.7, 1.1])
    r = np.linalg.norm(x)
    D = O.solid_harmonics(4, x)
    w = x[0] + 1j * x[1]
    assert abs(D[tri(0, 0)] - 1 / r) < 1e-15
    assert abs(D[tri(1, 0)] + x[2] / r ** 3) < 1e-15
    assert abs(D[tri(1, 1)] + w / r ** 3) < 1e-15
    assert abs(D[tri(2, 0)] - (3 * x[2] ** 2 / r ** 5 - 1 / r ** 3)) < 1e-14
    assert abs(D[tri(2, 2)] - 3 * w ** 2 / r ** 5) < 1e-14
    assert abs(D[tri(2, 1)] - 3 * w * x[2] / r ** 5) < 1e-14


def test_solid_harmonics_ladder_and_laplace():
    P = 7
    x0 = np.array([0.6, 0.45, -1.2])
    h = 1e-4
    D = O.solid_harmonics(P + 2, x0)
    for n in range(P):
        for m in range(n + 1):
            f = lambda xx: O.solid_harmonics(P + 2, xx)[tri(n, m)]  # noqa: E731
            g = [(f(x0 + h * e) - f(x0 - h * e)) / (2 * h) for e in np.eye(3)]
            lap = sum((f(x0 + h * e) - 2 * f(x0) + f(x0 - h * e)) / h ** 2 for e in np.eye(3))
            scale = abs(D[tri(n + 1, m)]) + abs(D[tri(n + 1, min(m + 1, n + 1))]) + 1e-300
            if m >= 1:
                gx = 0.5 * (D[tri(n + 1, m + 1)] - D[tri(n + 1, m - 1)])
                gy = -0.5j * (D[tri(n + 1, m + 1)] + D[tri(n + 1, m - 1)])
            else:
                gx, gy = D[tri(n + 1, 1)].real, D[tri(n + 1, 1)].imag
            assert abs(g[0] - gx) / scale < 1e-6
            assert abs(g[1] - gy) / scale < 1e-6
            assert abs(g[2] - D[tri(n + 1, m)]) / scale < 1e-6
            assert abs(lap) / (abs(D[tri(n, m)]) * 1e2 + scale) < 1e-4


@pytest.mark.parametrize("p", [6, 12])
def test_norms_by_quadrature(p):
    eta, w = O.unit_grid(p)
    Ds = np.array([O.solid_harmonics(p, e) for e in eta])
    for n in range(p + 1):
        for m in range(n + 1):
            N = 4 * np.pi * math.factorial(n - m) * math.factorial(n + m) / (2 * n + 1)
            q = (w * np.abs(Ds[:, tri(n, m)]) ** 2).sum()
            assert abs(q / N - 1) < 1e-12


@pytest.mark.parametrize("p", [4, 8, 12, 16])
def test_projection_idempotent_and_exact(p):
    eta, w = O.unit_grid(p)
    rng = np.random.default_rng(p)
    v = rng.standard_normal((len(w), 3))
    v1 = O.sh_synthesis(p, O.sh_analysis(p, v))
    v2 = O.sh_synthesis(p, O.sh_analysis(p, v1))
    assert relmax(v2, v1) < 1e-13
    # self-adjoint in the quadrature inner product
    u = rng.standard_normal((len(w), 3))
    pu = O.sh_synthesis(p, O.sh_analysis(p, u))
    assert abs(np.sum(w[:, None] * pu * v) - np.sum(w[:, None] * u * v1)) < 1e-11 * len(w)
    # band-limited fields are reproduced exactly: constants, rigid rotations,
    # and the surface gradient of x^3 (vector-SH degree <= 3)
    for f in (np.tile([0.3, -1.2, 0.5], (len(w), 1)), np.cross([0.4, 0.1, -0.9], eta),
              np.stack([3 * eta[:, 0] ** 2, 0 * eta[:, 0], 0 * eta[:, 0]], axis=1)):
        assert relmax(O.sh_synthesis(p, O.sh_analysis(p, f)), f) < 1e-13


def _coef_of(p, f):
    return O.sh_analysis(p, f)


@pytest.mark.parametrize("p", [6, 12])
def test_self_spectra_closed_forms(p):
    """Translation/rotation modes of the sphere (BIE/SphereSTKMobMat.cpp:174-196):
    S[c] = 2R/3 c, S[c x n] = R/3 c x n, exterior traction of rho0 = -rho0,
    interior = 0, principal value = -rho0/2."""
    eta, w = O.unit_grid(p)
    c = np.array([0.7, -0.2, 1.3])
    for f, s_fac in ((np.tile(c, (len(w), 1)), 2 / 3), (np.cross(c, eta), 1 / 3)):
        cf = _coef_of(p, f)
        assert relmax(O.sh_synthesis(p, O.sh_self(p, 0, cf)), s_fac * f) < 1e-13
        assert relmax(O.sh_synthesis(p, O.sh_self(p, 1, cf)), -f) < 1e-13
        assert np.abs(O.sh_synthesis(p, O.sh_self(p, 2, cf))).max() < 1e-13
        assert relmax(O.sh_synthesis(p, O.sh_self(p, 3, cf)), -0.5 * f) < 1e-13
    # uniform normal density: single layer vanishes, exterior traction vanishes
    cf = _coef_of(p, eta.copy())
    assert np.abs(O.sh_synthesis(p, O.sh_self(p, 0, cf))).max() < 1e-13
    assert np.abs(O.sh_synthesis(p, O.sh_self(p, 1, cf))).max() < 1e-13


@pytest.mark.parametrize("p", [8, 12])
def test_self_spectra_agree_with_exterior_limit_and_symmetry(p):
    eta, w = O.unit_grid(p)
    rng = np.random.default_rng(7)
    f = O.sh_synthesis(p, O.sh_analysis(p, rng.standard_normal((len(w), 3))))
    g = O.sh_synthesis(p, O.sh_analysis(p, rng.standard_normal((len(w), 3))))
    cf = O.sh_analysis(p, f)
    u, t = O.sh_eval_exterior(p, cf, eta, eta)
    assert relmax(u, O.sh_synthesis(p, O.sh_self(p, 0, cf))) < 1e-12
    assert relmax(t, O.sh_synthesis(p, O.sh_self(p, 1, cf))) < 1e-12
    # the single layer is self-adjoint: <S f, g> = <f, S g>
    Sf = O.sh_synthesis(p, O.sh_self(p, 0, cf))
    Sg = O.sh_synthesis(p, O.sh_self(p, 0, O.sh_analysis(p, g)))
    assert abs(np.sum(w[:, None] * Sf * g) - np.sum(w[:, None] * f * Sg)) < 1e-12 * np.abs(Sf).max() * len(w)
    # K' (principal value) is the adjoint of the double layer: <K' f, 1> = -<f, 1>/2 (Gauss)
    Kf = O.sh_synthesis(p, O.sh_self(p, 3, cf))
    for e in np.eye(3):
        assert abs(np.sum(w * (Kf @ e)) + 0.5 * np.sum(w * (f @ e))) < 1e-11 * np.abs(f).max()


def _upsampled(p, coef, pf=48):
    Q, Qf = (p + 1) * (p + 2) // 2, (pf + 1) * (pf + 2) // 2
    cc = coef[0::2] + 1j * coef[1::2]
    cf = np.zeros(3 * Qf, complex)
    for k in range(3):
        cf[k * Qf:k * Qf + Q] = cc[k * Q:(k + 1) * Q]
    out = np.zeros(6 * Qf)
    out[0::2], out[1::2] = cf.real, cf.imag
    vals = O.sh_synthesis(pf, out)
    eta, w = O.unit_grid(pf)
    return vals, eta, w


@pytest.mark.parametrize("rt,tol", [(4.0, 1e-12), (1.6, 1e-11)])
def test_exterior_eval_vs_bruteforce_quadrature(rt, tol):
    """Off-surface SL velocity and traction against brute-force quadrature over an
    upsampled grid (Sphere/Shexp_test.cpp:109-200, :301-411 with targets at rt R)."""
    p = 12
    eta0, w0 = O.unit_grid(p)
    rng = np.random.default_rng(3)
    coef = O.sh_analysis(p, rng.standard_normal((len(w0), 3)))
    vf, etaf, wf = _upsampled(p, coef)
    xt = rng.standard_normal((25, 3))
    xt *= rt / np.linalg.norm(xt, axis=1)[:, None]
    nt = rng.standard_normal((25, 3))
    nt /= np.linalg.norm(nt, axis=1)[:, None]
    u, t = O.sh_eval_exterior(p, coef, xt, nt)
    f = vf * wf[:, None]
    r = xt[:, None, :] - etaf[None]
    rn = np.linalg.norm(r, axis=2)
    rf = np.einsum("ijk,jk->ij", r, f)
    ub = (f[None] / rn[..., None] + r * (rf / rn ** 3)[..., None]).sum(1) / (8 * np.pi)
    tb = -3 / (4 * np.pi) * (r * (rf * np.einsum("ijk,ik->ij", r, nt) / rn ** 5)[..., None]).sum(1)
    assert relmax(u, ub) < tol
    assert relmax(t, tb) < tol
    # target normal parallel to the position (testTracSelf, Shexp_test.cpp:413-524)
    ns = xt / np.linalg.norm(xt, axis=1)[:, None]
    _, ts = O.sh_eval_exterior(p, coef, xt, ns)
    tbs = -3 / (4 * np.pi) * (r * (rf * np.einsum("ijk,ik->ij", r, ns) / rn ** 5)[..., None]).sum(1)
    assert relmax(ts, tbs) < tol


def test_far_sums_match_direct_numpy():
    rng = np.random.default_rng(11)
    pos = np.array([[0, 0, 0], [5.0, 0.3, 0.2], [0.4, 6.0, -1.0], [-3.5, -3.0, 4.0]])
    op = O.Operator(pos, [1.0, 0.8, 1.2, 0.9], np.tile([1.0, 0, 0, 0], (4, 1)), [4, 5, 4, 6])
    c, nv, w, off = op.grid()
    x = rng.standard_normal(op.dof)
    ut = op.far_sums(x, cSL=1.0, cTrac=0.0)
    tt = op.far_sums(x, cSL=0.0, cTrac=1.0)
    sph = np.repeat(np.arange(4), np.diff(off))
    f = x.reshape(-1, 3) * w[:, None]
    ub = np.zeros_like(c)
    tb = np.zeros_like(c)
    for i in range(len(c)):
        m = sph != sph[i]
        r = c[i] - c[m]
        rn = np.linalg.norm(r, axis=1)
        rf = (r * f[m]).sum(1)
        ub[i] = (f[m] / rn[:, None] + r * (rf / rn ** 3)[:, None]).sum(0) / (8 * np.pi)
        tb[i] = -3 / (4 * np.pi) * (r * (rf * (r @ nv[i]) / rn ** 5)[:, None]).sum(0)
    assert relmax(ut, ub) < 1e-13
    assert relmax(tt, tb) < 1e-13
