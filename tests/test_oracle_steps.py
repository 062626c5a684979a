"""The oracle's steps beside the solve, pinned without the reference (whose
random streams come from an un-vendored TRNG pool and whose collision search
lives in SimToolbox):

* Philox4x32-10 against the known-answer vectors published with the
  algorithm (Random123 kat_vectors, philox4x32 with 10 rounds);
* Brownian velocities (SRC/SphereSystem.cpp:323-360) through their law:
  zero mean, variance 2 kBT/dt scale^2 / (6 pi mu R) per translational and
  / (8 pi mu R^3) per rotational component, no correlation between components,
  spheres or steps, and reproducibility for a fixed (seed, step);
* collision pairs (Sphere/CollisionSphere.hpp:51-90) on geometries whose pairs
  are known in closed form: a cubic lattice, pairs placed just inside and just
  outside the buffer, the I side set by gid and the buffer scaled by the I
  sphere's radius (the test is not symmetric in I and J)."""
import numpy as np
import pytest

import pyoracle as O


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, want):
    assert list(O.philox4x32_10(ctr, key)) == want


def test_brownian_law():
    n = 200000
    rad = np.where(np.arange(n) % 2 == 0, 1.0, 0.5)
    dt, kBT, mu, scale = 0.01, 0.0041, 0.7, 1.3
    v = O.brownian_velocity(rad, dt, kBT, viscosity=mu, scale=scale, seed=12345, step=7).reshape(n, 6)
    f2 = 2 * kBT / dt * scale ** 2
    for R in (1.0, 0.5):
        sel = v[rad == R]
        m = len(sel)
        vt = f2 / (6 * np.pi * mu * R)
        vr = f2 / (8 * np.pi * mu * R ** 3)
        z = np.concatenate([sel[:, :3] / np.sqrt(vt), sel[:, 3:] / np.sqrt(vr)], axis=1)
        # standard normal components: mean within 5 sigma, variance within 5 sigma (sd of s^2 = sqrt(2/m))
        assert np.abs(z.mean(axis=0)).max() < 5 / np.sqrt(m)
        assert np.abs(z.var(axis=0) - 1).max() < 5 * np.sqrt(2 / m)
        c = np.corrcoef(z.T)
        assert np.abs(c - np.eye(6)).max() < 5 / np.sqrt(m)
        # fourth moment of a normal is 3 (Box-Muller, not a uniform law of the right variance)
        assert np.abs((z ** 4).mean(axis=0) - 3).max() < 5 * np.sqrt(96 / m)
    # neighbouring spheres and steps are uncorrelated; (seed, step) reproduces
    a = v[:, 0]
    assert abs(np.corrcoef(a[:-1], a[1:])[0, 1]) < 5 / np.sqrt(n)
    w = O.brownian_velocity(rad, dt, kBT, viscosity=mu, scale=scale, seed=12345, step=8).reshape(n, 6)
    assert abs(np.corrcoef(v[:, 0], w[:, 0])[0, 1]) < 5 / np.sqrt(n)
    u = O.brownian_velocity(rad, dt, kBT, viscosity=mu, scale=scale, seed=12346, step=7).reshape(n, 6)
    assert abs(np.corrcoef(v[:, 0], u[:, 0])[0, 1]) < 5 / np.sqrt(n)
    assert np.array_equal(v, O.brownian_velocity(rad, dt, kBT, viscosity=mu, scale=scale, seed=12345,
                                                 step=7).reshape(n, 6))
    assert not O.brownian_velocity(rad[:10], dt, 0.0).any()


def test_collision_cubic_lattice():
    """k^3 spheres of radius R on a cubic lattice of spacing d with d - 2R < colbuf R
    and d sqrt(2) - 2R > colbuf R: exactly the 3 k^2 (k-1) axis neighbours, with
    phi0 = d - 2R and normI = -(lattice step)."""
    k, d, R, cb = 5, 2.25, 1.0, 0.3
    g = np.arange(k)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * d
    ii, jj, phi0, nrm = O.collision_pairs(pos, np.full(len(pos), R), cb)
    assert len(ii) == 3 * k * k * (k - 1)
    assert np.all(ii < jj)
    assert np.all(np.lexsort((jj, ii)) == np.arange(len(ii)))
    step = pos[jj] - pos[ii]
    assert np.allclose(np.abs(step).sum(1), d, rtol=0, atol=1e-13)
    assert np.allclose(phi0, d - 2 * R, rtol=0, atol=1e-14)
    assert np.allclose(nrm, -step / d, rtol=0, atol=1e-15)


def test_collision_buffer_edges_and_orientation():
    """Separations colbuf*Rc_I*(1 -+ 1e-6) on either side of the buffer, the I side
    chosen by gid, and the asymmetric buffer colbuf*Rc_I (CollisionSphere.hpp:67)."""
    cb = 0.3
    for eps, want in ((-1e-6, 1), (1e-6, 0)):
        RA, RB = 1.0, 0.4
        sep = cb * RA * (1 + eps)
        pos = np.array([[0.0, 0.0, 0.0], [RA + RB + sep, 0.0, 0.0]])
        ii, jj, phi0, nrm = O.collision_pairs(pos, [RA, RB], cb)
        assert len(ii) == want
        if want:
            assert (ii[0], jj[0]) == (0, 1)
            assert abs(phi0[0] - sep) < 1e-15
            assert np.allclose(nrm[0], [-1.0, 0.0, 0.0], rtol=0, atol=2e-16)
    # sep between cb*0.4 and cb*1.0: a pair only when the large sphere is the I side
    pos = np.array([[0.0, 0.0, 0.0], [1.4 + 0.2, 0.0, 0.0]])
    assert len(O.collision_pairs(pos, [1.0, 0.4], cb, gid=[0, 1])[0]) == 1
    ii, jj, phi0, nrm = O.collision_pairs(pos, [1.0, 0.4], cb, gid=[5, 2])
    assert len(ii) == 0
    # overlapping spheres: phi0 < 0; I side = smaller gid, normal points from J to I
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 1.5, 0.0]])
    ii, jj, phi0, nrm = O.collision_pairs(pos, [1.0, 1.0], cb, gid=[9, 3])
    assert (ii[0], jj[0]) == (1, 0) and abs(phi0[0] + 0.5) < 1e-15
    assert np.allclose(nrm[0], [0.0, 1.0, 0.0], rtol=0, atol=2e-16)


def test_collision_pairs_are_all_close_pairs():
    """Random polydisperse spheres: the pair set equals the set of all index pairs
    with centre distance below Rc_I + Rc_J + cb Rc_I (from the pairwise distance
    matrix), each unordered pair once."""
    rng = np.random.default_rng(3)
    n = 300
    pos = rng.uniform(0, 12, (n, 3))
    rc = rng.uniform(0.3, 0.8, n)
    gid = rng.permutation(n)
    cb = 0.3
    ii, jj, phi0, nrm = O.collision_pairs(pos, rc, cb, gid=gid)
    D = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    near = (D < rc[:, None] * (1 + cb) + rc[None]) & (gid[:, None] < gid[None])
    want = set(zip(*np.nonzero(near)))
    assert set(zip(ii.tolist(), jj.tolist())) == want
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, rtol=0, atol=1e-15)
    assert np.allclose(phi0, D[ii, jj] - rc[ii] - rc[jj], rtol=0, atol=1e-14)
