"""The collision LCP around the mobility (SRC/SphereSystem.cpp:445-474; the
reference's solver lives in the un-vendored SimToolbox and is restated from the
published BBPGD method).  The oracle's BBPGD is pinned by what the mathematics
fixes: for the free-drain mobility and pairs that share no sphere, the closed
form gamma_k = max(-b_k / A_kk, 0); for coupled pairs, the unique solution of
the LCP with SPD matrix found by trying every active set.  The library's LCP
(host code, here with the free-drain mobility, so no GPU is needed) must match
the oracle step for step."""
import numpy as np
import pytest

import pyoracle as O
import spheresimulator_amd as S


def _dense_A(mob, n, pairs, dt):
    """A = dt D^T M D column by column."""
    ii, jj, nrm = pairs["idx_i"], pairs["idx_j"], pairs["norm_i"]
    m = len(ii)
    A = np.zeros((m, m))
    for k in range(m):
        F = np.zeros((n, 6))
        F[ii[k], :3] += nrm[k]
        F[jj[k], :3] -= nrm[k]
        u = mob(F.ravel()).reshape(n, 6)[:, :3]
        A[:, k] = dt * np.einsum("kc,kc->k", nrm, u[ii] - u[jj])
    return A


def _b(pairs, vk, dt, buffer=0.0):
    v = vk.reshape(-1, 6)[:, :3]
    return pairs["phi0"] + buffer + dt * np.einsum("kc,kc->k", pairs["norm_i"], v[pairs["idx_i"]] - v[pairs["idx_j"]])


def _pairs(pos, rc, cb):
    ii, jj, phi0, nrm = O.collision_pairs(pos, rc, cb)
    return dict(idx_i=ii, idx_j=jj, phi0=phi0, norm_i=nrm)


def test_oracle_bbpgd_disjoint_pairs_closed_form():
    """Three isolated overlapping or touching pairs, free drain: gamma_k =
    max(-b_k/A_kk, 0) with A_kk = dt (1/(6 pi mu R_I) + 1/(6 pi mu R_J))."""
    pos = np.array([[0, 0, 0], [1.8, 0.2, 0], [20, 0, 0], [20, 1.9, 0], [40, 0, 0], [40, 0, 1.5]], float)
    rad = np.array([1.0, 0.9, 1.0, 1.0, 0.8, 0.6])
    mu, dt = 0.7, 1e-3
    pairs = _pairs(pos, rad, 0.3)
    assert len(pairs["idx_i"]) == 3
    vk = np.random.default_rng(1).standard_normal(36)
    vk[12:24] = 0.0  # pair 2 (spheres 2, 3) overlaps and does not move: gamma > 0
    g, F, U, info = O.lcp_bbpgd(O.free_drain(rad, mu), 6, pairs["idx_i"], pairs["idx_j"], pairs["phi0"],
                                pairs["norm_i"], vk, dt, tol=1e-13)
    b = _b(pairs, vk, dt)
    Akk = dt * (1 / (6 * np.pi * mu * rad[pairs["idx_i"]]) + 1 / (6 * np.pi * mu * rad[pairs["idx_j"]]))
    want = np.maximum(-b / Akk, 0)
    assert info["converged"] == 1
    assert np.allclose(g, want, rtol=1e-10, atol=1e-12)
    assert (want > 0).sum() >= 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_bbpgd_coupled_pairs_match_enumeration(seed):
    """A cluster of overlapping spheres (pairs share spheres): BBPGD equals the
    active-set solution, and the KKT conditions hold."""
    rng = np.random.default_rng(seed)
    # a bent chain: each sphere overlaps or nearly touches the next, a 4-sphere
    # ring closes on itself at the start
    pos = np.array([[0, 0, 0], [1.6, 0, 0], [0.8, 1.35, 0], [3.2, 0, 0], [4.8, 0, 0], [6.4, 0, 0]], float)
    pos += rng.uniform(-0.08, 0.08, pos.shape)
    rad = rng.uniform(0.78, 0.86, 6)
    pairs = _pairs(pos, rad, 0.3)
    m = len(pairs["idx_i"])
    assert 5 <= m <= 8
    mu, dt = 1.0, 1e-2
    vk = rng.standard_normal(36)
    mob = O.free_drain(rad, mu)
    g, F, U, info = O.lcp_bbpgd(mob, 6, pairs["idx_i"], pairs["idx_j"], pairs["phi0"], pairs["norm_i"], vk, dt,
                                tol=1e-12, max_iters=20000)
    A = _dense_A(mob, 6, pairs, dt)
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > 1e-6 * ev.max()  # SPD: the LCP has one solution
    b = _b(pairs, vk, dt)
    want = O.lcp_enumerate(A, b)
    assert info["converged"] == 1
    assert np.abs(g - want).max() < 1e-8 * max(1.0, np.abs(want).max())
    phi = A @ g + b
    assert (g >= 0).all() and (phi >= -1e-11).all() and abs(g @ phi) < 1e-9
    assert np.allclose(U, mob(F), rtol=1e-12, atol=1e-14)


def test_library_lcp_free_drain_matches_oracle():
    """The library's BBPGD (host code; free-drain mobility) against the oracle's on
    the same inputs, iteration counts included; buffer and an initial guess."""
    rng = np.random.default_rng(7)
    n = 60
    pos = rng.uniform(0, 8, (n, 3))
    rad = rng.uniform(0.5, 0.9, n)
    pairs = _pairs(pos, rad, 0.3)
    m = len(pairs["idx_i"])
    assert m > 20
    vk = rng.standard_normal(6 * n)
    g0 = np.maximum(-pairs["phi0"], 0)
    for buf in (0.0, 0.05):
        g, F, U, info = S.system.collision_lcp(pairs, vk, 1e-2, radius=rad, viscosity=0.8, buffer=buf, tol=1e-10,
                                               gamma0=g0)
        go, Fo, Uo, io = O.lcp_bbpgd(O.free_drain(rad, 0.8), n, pairs["idx_i"], pairs["idx_j"], pairs["phi0"],
                                     pairs["norm_i"], vk, 1e-2, buffer=buf, tol=1e-10, gamma0=g0)
        assert info["converged"] == io["converged"] == 1
        assert abs(info["iters"] - io["iters"]) <= 2
        assert np.abs(g - go).max() < 1e-9 * np.abs(go).max()
        assert np.abs(F - Fo).max() < 1e-9 * np.abs(Fo).max()
        assert np.abs(U - Uo).max() < 1e-9 * np.abs(Uo).max()
        A = _dense_A(O.free_drain(rad, 0.8), n, pairs, 1e-2)
        phi = A @ g + _b(pairs, vk, 1e-2, buf)
        assert (g >= 0).all() and (phi >= -1e-10).all() and np.abs(np.minimum(g, phi)).max() <= 1e-10


def test_library_lcp_edge_cases():
    n = 3
    vk = np.zeros(18)
    empty = dict(idx_i=np.zeros(0, np.int32), idx_j=np.zeros(0, np.int32), phi0=np.zeros(0), norm_i=np.zeros((0, 3)))
    g, F, U, info = S.system.collision_lcp(empty, vk, 0.1, radius=np.ones(n))
    assert g.size == 0 and not F.any() and info["converged"] == 1
    # separating pair: gamma = 0 at once
    pairs = dict(idx_i=np.array([0], np.int32), idx_j=np.array([1], np.int32), phi0=np.array([0.1]),
                 norm_i=np.array([[1.0, 0, 0]]))
    g, F, U, info = S.system.collision_lcp(pairs, vk, 0.1, radius=np.ones(n))
    assert g[0] == 0.0 and info["iters"] == 0 and info["converged"] == 1
    bad = dict(pairs, idx_j=np.array([3], np.int32))
    with pytest.raises(S.SphstkError):
        S.system.collision_lcp(bad, vk, 0.1, radius=np.ones(n))
    with pytest.raises(S.SphstkError):
        S.system.collision_lcp(pairs, vk, 0.0, radius=np.ones(n))
