"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle on
identical seeded inputs.  Tolerances: relative 1e-11 for single operator
applies (pure FP64 arithmetic in a different order), 1e-10 for full FP64
mobility solves (the north-star parity bar), 1e-7 for the mixed-precision
solve against the FP64 oracle."""
import sys

import numpy as np
import pytest

import pyoracle as O
import spheresimulator_amd as S
from conftest import rel, relmax

pytestmark = pytest.mark.gpu

ID = [1.0, 0.0, 0.0, 0.0]


def say(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def quat_of(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])


def small_geometry(seed=0, n=8, orders=(6, 8, 5, 7)):
    rng = np.random.default_rng(seed)
    pos = [np.zeros(3)]
    rad = [1.0]
    while len(pos) < n:
        r = rng.uniform(0.6, 1.3)
        x = rng.uniform(-5, 5, 3)
        if all(np.linalg.norm(x - q) > r + rq + 0.15 for q, rq in zip(pos, rad)):
            pos.append(x)
            rad.append(r)
    quat = np.array([quat_of(rng.standard_normal(3), rng.uniform(0, 3)) for _ in range(n)])
    order = np.array([orders[i % len(orders)] for i in range(n)], dtype=np.int32)
    return np.array(pos), np.array(rad), quat, order


@pytest.fixture(scope="module")
def geo():
    return small_geometry()


@pytest.mark.parametrize("coeffs", [(0.5, 0.0, 1.0), (0.0, 1.0, 0.0), (-0.5, 0.0, -1.0), (0.3, 0.7, -0.4)])
def test_p2p_parity(geo, coeffs):
    say("p2p parity coeffs=%s" % (coeffs,))
    pos, rad, quat, order = geo
    ref = O.Operator(pos, rad, quat, order)
    gpu = S.SphereSTKSHOperator(pos, rad, quat, order)
    assert gpu.n_dof == ref.dof
    assert gpu.n_near_pairs == len(ref.near_pairs()[0]) > 0
    x = np.random.default_rng(1).standard_normal(ref.dof)
    y0 = np.random.default_rng(2).standard_normal(ref.dof)
    want = ref.apply_p2p(x, y0.copy(), *coeffs)
    got = gpu.applyP2POP(x, y0.copy(), *coeffs)
    assert relmax(got, want) < 1e-11


def test_grid_and_near_list_match(geo):
    say("grid/near list")
    pos, rad, quat, order = geo
    ref = O.Operator(pos, rad, quat, order)
    gpu = S.SphereSTKSHOperator(pos, rad, quat, order)
    c, nv, w, off = ref.grid()
    g = gpu.grid()
    assert np.abs(g["coords"] - c.ravel()).max() < 1e-14
    assert np.abs(g["norms"] - nv.ravel()).max() < 1e-15
    assert np.abs(g["weights"] - w).max() < 1e-15
    assert np.array_equal(g["index"], off[:-1]) and np.array_equal(g["length"], np.diff(off))
    t, s = gpu.near_pairs()
    tr, sr = ref.near_pairs()
    assert np.array_equal(t, tr) and np.array_equal(s, sr)


def test_near_list_batched_and_overlap_rejected():
    """The device near-pair search over a batch whose suspensions occupy the same
    region of space: pairs only within a suspension, each suspension's list equal
    to the oracle's for that suspension alone (same order), and two overlapping
    spheres of one suspension rejected at creation."""
    say("near list: batched suspensions in one region, overlap rejected")
    c = S.configs.config4(n_susp=12)
    gpu = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], counts=c["counts"])
    t, s = gpu.near_pairs()
    off = np.concatenate([[0], np.cumsum(c["counts"])])
    want_t, want_s = [], []
    for b in range(len(c["counts"])):
        sl = slice(off[b], off[b + 1])
        tr, sr = O.Operator(c["pos"][sl], c["radius"][sl], c["quat"][sl], c["order"][sl]).near_pairs()
        want_t.append(tr + off[b])
        want_s.append(sr + off[b])
    assert len(t) > 100
    assert np.array_equal(t, np.concatenate(want_t)) and np.array_equal(s, np.concatenate(want_s))
    pos = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [6.9, 0.0, 0.0]])
    with pytest.raises(S.SphstkError, match="overlapping spheres 1 and 2"):
        S.SphereSTKSHOperator(pos, [1.0, 1.0, 1.0], None, 4)
    # the same three spheres split over two suspensions do not interact
    ok = S.SphereSTKSHOperator(pos, [1.0, 1.0, 1.0], None, 4, counts=[2, 1])
    assert ok.n_near_pairs == 0


def test_apply_lop_project_parity(geo):
    say("apply/lop/project")
    pos, rad, quat, order = geo
    ref = O.Operator(pos, rad, quat, order, 0.5, 0.0, 1.0, 1.0)
    gpu = S.SphereSTKSHOperator(pos, rad, quat, order, 0.5, 0.0, 1.0, 1.0)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(ref.dof)
    y0 = rng.standard_normal(ref.dof)
    assert relmax(gpu.apply(x, y0.copy(), 1.7, -0.3), ref.apply(x, y0.copy(), 1.7, -0.3)) < 1e-11
    assert relmax(gpu.applyLOP(x, y0.copy(), 0.9), ref.apply_lop(x, y0.copy(), 0.9)) < 1e-13
    px = gpu.projectNullSpace(x.copy())
    assert relmax(px, ref.project(x)) < 1e-12
    assert relmax(gpu.projectNullSpace(px.copy()), px) < 1e-12
    # two columns at once
    X = np.concatenate([x, y0])
    Y = gpu.apply(X)
    assert relmax(Y[:ref.dof], ref.apply(x)) < 1e-11 and relmax(Y[ref.dof:], ref.apply(y0)) < 1e-11


def test_pos_initial_operator_parity(pos_initial):
    """The reference's own input file posInitial.txt (100 spheres, 130 near pairs)."""
    say("posInitial.txt operator parity")
    pos, rad = pos_initial
    n = len(rad)
    quat = np.tile(ID, (n, 1))
    order = np.full(n, 6, dtype=np.int32)
    ref = O.Operator(pos, rad, quat, order)
    gpu = S.SphereSTKSHOperator(pos, rad, quat, order)
    x = np.random.default_rng(4).standard_normal(ref.dof)
    for cf in ((0.5, 0.0, 1.0), (0.0, 1.0, 0.0)):
        assert relmax(gpu.applyP2POP(x, np.zeros(ref.dof), *cf), ref.apply_p2p(x, None, *cf)) < 1e-11


def test_batched_operator_parity():
    say("batched operator parity")
    geos = [small_geometry(seed=s, n=k) for s, k in ((5, 3), (6, 6), (7, 1), (8, 4))]
    pos = np.concatenate([g[0] for g in geos])
    rad = np.concatenate([g[1] for g in geos])
    quat = np.concatenate([g[2] for g in geos])
    order = np.concatenate([g[3] for g in geos])
    counts = [len(g[1]) for g in geos]
    gpu = S.SphereSTKSHOperator(pos, rad, quat, order, 0.5, 0.0, 1.0, 1.0, counts=counts)
    x = np.random.default_rng(9).standard_normal(gpu.n_dof)
    y = gpu.apply(x)
    o = 0
    for g in geos:
        ref = O.Operator(*g, 0.5, 0.0, 1.0, 1.0)
        assert relmax(y[o:o + ref.dof], ref.apply(x[o:o + ref.dof])) < 1e-11
        o += ref.dof


def test_isolated_sphere_closed_form():
    say("isolated sphere")
    for prec in (S.PREC_FP64, S.PREC_MIXED):
        M = S.SphereSTKMobMat([[0.3, -1.0, 2.0]], [0.7], [quat_of([1, 2, 3], 0.7)], [8], viscosity=2.5,
                              precision=prec)
        F = np.array([1.0, -2.0, 0.5, 0.3, 0.9, -0.4])
        v, info = M.solveMob(F)
        assert info["shortcut"] == 1
        assert rel(v[:3], F[:3] / (6 * np.pi * 2.5 * 0.7)) < 1e-13
        assert rel(v[3:], F[3:] / (8 * np.pi * 2.5 * 0.7 ** 3)) < 1e-13


def test_mobility_cfg1_fp64_and_mixed():
    """64 spheres p=8 (the reference's CPU-runnable case): FP64 GPU matches the
    oracle to 1e-10; mixed precision to 1e-7."""
    say("cfg1 mobility FP64 + mixed")
    c = S.configs.config1()
    ref = O.MobilityOperator(c["pos"], c["radius"], c["quat"], c["order"])
    vr, ir = ref.solve(c["force"], tol=1e-8)
    assert ir["converged"] == 1
    gpu = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"])
    v, info = gpu.solveMob(c["force"])
    assert info["converged"] == 1
    V, W = v.reshape(-1, 6)[:, :3], v.reshape(-1, 6)[:, 3:]
    Vr, Wr = vr.reshape(-1, 6)[:, :3], vr.reshape(-1, 6)[:, 3:]
    assert rel(V, Vr) < 1e-10
    assert rel(W, Wr) < 1e-10
    assert abs(info["iters"] - ir["iters"]) <= 1
    mix = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_MIXED)
    vm, im = mix.solveMob(c["force"])
    assert im["converged"] == 1 and im["relres"] <= 1e-8
    assert rel(vm.reshape(-1, 6)[:, :3], Vr) < 1e-7
    assert rel(vm.reshape(-1, 6)[:, 3:], Wr) < 1e-7


def test_batched_mobility_matches_individual_solves():
    say("batched mobility")
    geos = [small_geometry(seed=s, n=k, orders=(o,)) for s, k, o in ((11, 5, 6), (12, 9, 8), (13, 2, 10))]
    pos = np.concatenate([g[0] for g in geos])
    rad = np.concatenate([g[1] for g in geos])
    quat = np.concatenate([g[2] for g in geos])
    order = np.concatenate([g[3] for g in geos])
    counts = [len(g[1]) for g in geos]
    F = np.random.default_rng(3).standard_normal(6 * len(rad))
    for prec, tol in ((S.PREC_FP64, 1e-10), (S.PREC_MIXED, 1e-7)):
        gpu = S.SphereSTKMobMat(pos, rad, quat, order, counts=counts, precision=prec)
        v, info = gpu.solveMob(F)
        assert info["converged"] == 1
        o = 0
        for g in geos:
            k = len(g[1])
            vr, _ = O.MobilityOperator(*g).solve(F[6 * o:6 * (o + k)], tol=1e-8)
            assert rel(v[6 * o:6 * (o + k)], vr) < tol
            o += k


def test_warm_start_and_density():
    say("warm start")
    pos, rad, quat, order = small_geometry(seed=21, n=5, orders=(6,))
    F = np.random.default_rng(5).standard_normal(30)
    gpu = S.SphereSTKMobMat(pos, rad, quat, order)
    v1, i1 = gpu.solveMob(F)
    v2, i2 = gpu.solveMob(F)  # x0 = xLast (MobMat.cpp:222, :245)
    assert i2["iters"] < i1["iters"]
    assert rel(v2, v1) < 1e-9
    rho = gpu.getDensitySolution()
    ref = O.MobilityOperator(pos, rad, quat, order)
    vr, _ = ref.solve(F, tol=1e-8)
    assert rel(v1, vr) < 1e-10
    assert np.isfinite(rho).all() and rho.size == gpu.op.n_dof


def test_set_density_warm_start():
    """The constructor's warm start (MobMat.cpp:67-93): a new handle seeded with the
    previous handle's density starts from x = projectNullSpace(rho) (rho0 is not
    subtracted: the reference's member rho is still zero there) and converges to
    the same velocities; a zero density is a cold start, bit for bit."""
    say("set_density warm start")
    pos, rad, quat, order = small_geometry(seed=33, n=6, orders=(6, 8))
    F = np.random.default_rng(9).standard_normal(36)
    a = S.SphereSTKMobMat(pos, rad, quat, order)
    v1, i1 = a.solveMob(F)
    rho = a.getDensitySolution()
    b = S.SphereSTKMobMat(pos, rad, quat, order)
    b.set_density(rho)
    want = b.op.projectNullSpace(rho.copy())
    assert np.abs(b.get_unknown() - want).max() <= 1e-14 * np.abs(want).max()
    v2, i2 = b.solveMob(F)
    assert i2["converged"] == 1 and rel(v2, v1) < 1e-9
    c = S.SphereSTKMobMat(pos, rad, quat, order)
    c.set_density(np.zeros_like(rho))
    v3, i3 = c.solveMob(F)
    assert i3["iters"] == i1["iters"] and np.array_equal(v3, v1)


@pytest.mark.parametrize("p,direct,tol", [(8, True, 2e-6), (8, False, 2e-6), (12, False, 1e-5)])
def test_inner_fp32_operator_accuracy(p, direct, tol):
    """The FP32 inner operator approximates the FP64 operator on a white band-
    limited density (the projection of a random field).  Direct FP32 far field:
    FP32 rounding only.  Multi-level far field (pairs beyond 3 (R_s + R_t) through
    coarser grids, order p/2+1 first): the truncated degrees decay like (1/6)^(p/2+2),
    about 4e-6 of the operator at p=12; two refinement steps against the FP64
    residual reach 1e-8 with an inner operator up to ~1e-5 off."""
    say("inner FP32 operator (p=%d direct=%s)" % (p, direct))
    c = S.configs.make_suspension(96 if p == 8 else 256, p, 31, vf=0.1, min_gap=0.2)
    gpu = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0,
                                fp32_direct=direct)
    x = gpu.projectNullSpace(np.random.default_rng(8).standard_normal(gpu.n_dof))
    y64 = gpu.apply(x)
    y32 = gpu.apply_inner(x)
    err = rel(y32, y64)
    print("inner rel err %.3e" % err)
    assert err < tol


def test_coarse_fp32_operator_accuracy():
    """The coarse FP32 operator of the mixed form's inner solves (coarse grids
    of order 5, 4, 3, 2 from 2, 3, 5, 8 (R_s + R_t)) approximates the FP64
    operator to about 1e-2 on a band-limited density: coarser than the FP32
    inner operator, fine enough that each refinement step gains about 1e-2.
    With the FP32 direct far field it is the FP32 inner operator itself."""
    say("coarse FP32 operator (256 spheres p=12)")
    c = S.configs.make_suspension(256, 12, 31, vf=0.1, min_gap=0.2)
    gpu = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    x = gpu.projectNullSpace(np.random.default_rng(8).standard_normal(gpu.n_dof))
    y64 = gpu.apply(x)
    e_c, e_32 = rel(gpu.apply_coarse(x), y64), rel(gpu.apply_inner(x), y64)
    print("coarse rel err %.3e, inner rel err %.3e" % (e_c, e_32))
    assert e_32 < e_c < 3e-2
    pairs_c, pairs_32 = gpu.kernel_stats(0, 3)["pairs"], gpu.kernel_stats(0, 1)["pairs"]
    print("pair evaluations per apply: coarse %.4g, inner %.4g" % (pairs_c, pairs_32))
    assert 0 < pairs_c < 0.6 * pairs_32
    direct = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0,
                                   fp32_direct=True)
    assert np.array_equal(direct.apply_coarse(x), direct.apply_inner(x))


def test_inner_operator_source_cells():
    """Source-cell aggregation of the FP32 inner operator's two coarsest levels
    (cells of 16 spheres, the leaves of a k-d split, as 200 equivalent point
    forces, used for targets beyond 4 cell radii): on a dense suspension where
    cells are in use it evaluates fewer pairs, and the operator stays as close
    to the FP64 direct-sum operator as without cells (band-limited and white
    densities), and to the oracle's direct sums on sampled spheres.  The
    operator's own error at this density (p=8, vf 0.2) is about 1.2e-5 with
    the default levels; the refinement takes its cheap second step for
    errors up to about 5e-5."""
    say("inner FP32 operator with source cells (4096 spheres p=8 vf 0.2)")
    c = S.configs.make_suspension(4096, 8, 37, vf=0.2, min_gap=0.1)
    args = (c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    with_cells = S.SphereSTKSHOperator(*args)
    without = S.SphereSTKSHOperator(*args, cells=False)
    white = np.random.default_rng(9).standard_normal(with_cells.n_dof)
    band = with_cells.projectNullSpace(white.copy())
    for name, x in (("band-limited", band), ("white", white)):
        y64 = with_cells.apply(x)
        e_cells, e_plain = rel(with_cells.apply_inner(x), y64), rel(without.apply_inner(x), y64)
        print("%s: inner rel err with cells %.3e, without %.3e" % (name, e_cells, e_plain))
        assert e_cells < 2e-5 and e_cells < 1.5 * e_plain + 2e-6
    p_cells, p_plain = with_cells.kernel_stats(0, 1)["pairs"], without.kernel_stats(0, 1)["pairs"]
    print("pairs per inner apply: with cells %.4g, without %.4g" % (p_cells, p_plain))
    assert p_cells < 0.97 * p_plain
    ref = O.Operator(*args)
    _, _, _, off = ref.grid()
    sample = (0, 1500, 4095)
    rows = np.concatenate([np.arange(3 * off[i], 3 * off[i + 1]) for i in sample])
    want = np.zeros(ref.dof)
    for i in sample:
        ref.apply_p2p_spheres(band, i, i + 1, want, 0.5, 0.0, 1.0)
    ref.apply_lop(band, want, 1.0)
    err = rel(with_cells.apply_inner(band)[rows], want[rows])
    print("inner with cells vs oracle on sampled spheres %.3e" % err)
    assert err < 2e-5


def test_inner_level_orders_from_environment(monkeypatch):
    """SPHSTK_LEVEL_ALPHA with explicit FP32 level orders: orders equal to the
    defaults (p=8: 8, 5, 4, 3; -1 keeps a default) give the default inner
    operator bit for bit, lower orders on the far levels change it slightly,
    and a malformed value makes creation fail with an error."""
    say("inner level orders from SPHSTK_LEVEL_ALPHA")
    c = S.configs.make_suspension(96, 8, 31, vf=0.02, min_gap=0.2)

    def inner(x):
        op = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
        return op.apply_inner(x)

    monkeypatch.delenv("SPHSTK_LEVEL_ALPHA", raising=False)
    op = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    x = np.random.default_rng(5).standard_normal(op.n_dof)
    ref, y64 = op.apply_inner(x), op.apply(x)
    del op
    for v in ("3,6,9,16:-1,-1,-1,-1", "3,6,9,16:8,5,4,3", "3,6,9,16"):
        monkeypatch.setenv("SPHSTK_LEVEL_ALPHA", v)
        assert np.array_equal(inner(x), ref), v
    monkeypatch.setenv("SPHSTK_LEVEL_ALPHA", "3,6,9,16:-1,-1,3,2")
    low = inner(x)
    print("lower far-level orders: change %.3e, error vs FP64 %.3e (default %.3e)" %
          (rel(low, ref), rel(low, y64), rel(ref, y64)))
    assert 0 < rel(low, ref) < 1e-3
    for bad in ("3,6,9", "3,6,9,16:1,2", "3,6,9,16:-1,-1,-1,99"):
        monkeypatch.setenv("SPHSTK_LEVEL_ALPHA", bad)
        with pytest.raises(S.SphstkError):
            inner(x)


@pytest.mark.parametrize("p", [10, 12])
def test_outer_fp64_operator_accuracy(p):
    """The FP64 multi-level outer operator of the mixed form (pairs beyond
    s' = 5.8 through grids of order min(p, 9/7/6/5)) matches the FP64 direct
    operator on a band-limited density and on a white one (no band limit), to
    far below the 1e-8 GMRES tolerance the refinement residual serves (about
    3e-10 here, p=12).  The
    dilute suspension (vf 0.02) makes the levels pay, so they are in use."""
    say("outer FP64 multi-level operator (p=%d)" % p)
    c = S.configs.make_suspension(256, p, 33, vf=0.02, min_gap=0.2)
    gpu = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    white = np.random.default_rng(9).standard_normal(gpu.n_dof)
    for name, x, tol in (("band-limited", gpu.projectNullSpace(white.copy()), 2e-9), ("white", white, 2e-9)):
        err = rel(gpu.apply_outer(x), gpu.apply(x))
        print("outer %s rel err %.3e" % (name, err))
        assert err < tol
    ml, direct = gpu.kernel_stats(0, 2)["pairs"], gpu.kernel_stats(0, 0)["pairs"]
    print("pair evaluations: multi-level %.4g, direct %.4g" % (ml, direct))
    assert 0 < ml < 0.75 * direct


def test_mixed_outer_levels_built_in_background():
    """The mixed form builds its FP64 outer level set on a host thread during
    the set-up and the first inner solve: a mobility destroyed right after
    creation joins that thread, and solves on a fresh mobility equal those on
    one whose set-up finished before the first solve (the standalone operator
    builds the set on first use)."""
    say("mixed form: FP64 outer level set built in the background")
    c = S.configs.config2()
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_MIXED)
    del M
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_MIXED)
    v1, info1 = M.solveMob(c["force"])
    M.set_unknown(np.zeros(M.n_dof))
    v2, info2 = M.solveMob(c["force"])
    assert info1["converged"] == 1 and np.array_equal(v1, v2)
    op = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    x = np.random.default_rng(17).standard_normal(op.n_dof)
    assert np.array_equal(op.apply_outer(x), M.op.apply_outer(x))


def test_mixed_matches_fp64_low_order_polydisperse():
    """Mixed form vs FP64 form on polydisperse suspensions with low orders (the
    first suspensions of the batched workload: p in {6..16}, R in [0.5, 1.5],
    gaps 0.1 (R_s + R_t)): final velocities agree to 1e-8, i.e. the sphere-
    centre (Faxen) sums are used only where the grid quadrature matches them."""
    say("mixed vs FP64, polydisperse low order")
    c = S.configs.config4(n_susp=6)
    v = {}
    for prec in (S.PREC_FP64, S.PREC_MIXED):
        M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], counts=c["counts"], precision=prec)
        v[prec], info = M.solveMob(c["force"])
        assert info["converged"] == 1
    a, b = v[S.PREC_MIXED].reshape(-1, 6), v[S.PREC_FP64].reshape(-1, 6)
    ev, ew = rel(a[:, :3], b[:, :3]), rel(a[:, 3:], b[:, 3:])
    print("rel err V %.3e Omega %.3e" % (ev, ew))
    assert ev < 1e-8 and ew < 1e-8


def test_autotuned_shapes_give_the_same_operator():
    """SPHSTK_FLAG_AUTOTUNE picks far-field tile shapes by timing candidates on
    the problem itself; whatever it picks, the operators are unchanged (up to
    summation order)."""
    say("autotuned far-field shapes")
    c = S.configs.make_suspension(256, 8, 35, vf=0.05, min_gap=0.2)
    args = (c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    ref = S.SphereSTKSHOperator(*args)
    tuned = S.SphereSTKSHOperator(*args, autotune=True)
    sh = tuned.far_shape()
    print("shape", sh, "default", ref.far_shape())
    assert sh["tpl"] in (1, 2, 4) and 1 <= sh["split"] <= 64
    x = np.random.default_rng(13).standard_normal(ref.n_dof)
    assert rel(tuned.apply(x), ref.apply(x)) < 1e-13
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_MIXED, autotune=True)
    assert M.op.far_shape()["inner_tpl"] in (1, 2, 4)
    v, info = M.solveMob(c["force"])
    vr, _ = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"]).solveMob(c["force"])
    assert info["converged"] == 1 and rel(v, vr) < 1e-8


def test_inner_apply_graph_replays_match_eager():
    """SPHSTK_FLAG_GRAPH: the FP32 inner operator's apply is captured once and
    replayed.  Every replay, with other eager work and synchronisations in
    between, equals the eager apply of the same input; the mixed solve with
    graphs equals the one without."""
    say("inner apply HIP graph replays")
    c = S.configs.make_suspension(128, 8, 37, vf=0.05, min_gap=0.2)
    args = (c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    eager = S.SphereSTKSHOperator(*args)
    gr = S.SphereSTKSHOperator(*args, graph=True)
    rng = np.random.default_rng(17)
    for k in range(5):
        x = rng.standard_normal(eager.n_dof)
        want = eager.apply_inner(x)
        got = gr.apply_inner(x)
        print("replay %d rel diff %.3e" % (k, rel(got, want)))
        assert rel(got, want) <= 1e-14
        gr.apply(x)  # eager FP64 work between the replays
    kw = dict(precision=S.PREC_MIXED)
    v0, i0 = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], **kw).solveMob(c["force"])
    v1, i1 = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], graph=True, **kw).solveMob(c["force"])
    assert i1["converged"] == 1 and i1["iters"] == i0["iters"] and rel(v1, v0) <= 1e-13


def test_pole_values_and_vtu_output(tmp_path):
    """Shexp::calcPoleValue and writeVTU: for a band-limited grid field (a
    constant, a rigid rotation c x n and a cubic in n) the expansion at the two
    poles of each rotated SH frame equals the field there; the VTU file holds
    one piece per sphere with the poles first and last, zero pole weights and
    the grid weights summing to 4 pi R^2."""
    import xml.etree.ElementTree as ET
    say("pole values and VTU output")
    pos, rad, quat, order = small_geometry(seed=51, n=4, orders=(8, 6, 10, 7))
    op = S.SphereSTKSHOperator(pos, rad, quat, order)
    g = op.grid()
    nv = g["norms"].reshape(-1, 3)
    a = np.array([0.3, -0.7, 1.1])
    c = np.array([0.4, 0.2, -0.5])

    def field(n):
        return a + np.cross(c, n) + np.stack([n[:, 0] * n[:, 1] * n[:, 2], n[:, 0] ** 2, n[:, 2] ** 3 - n[:, 1]], 1)

    x = field(nv).ravel()
    poles = op.pole_values(x)
    for s in range(len(rad)):
        w, u = quat[s, 0], quat[s, 1:]
        t = 2 * np.cross(u, [0.0, 0.0, 1.0])
        ez = np.array([0.0, 0.0, 1.0]) + w * t + np.cross(u, t)
        want = field(np.stack([ez, -ez]))
        assert np.abs(poles[s] - want).max() < 1e-12
    f = tmp_path / "density.vtu"
    S.system.write_vtu(f, op, pos, rad, quat, order, x)
    pieces = ET.parse(f).getroot().find("UnstructuredGrid").findall("Piece")
    assert len(pieces) == len(rad)
    arrays = S.system.read_vtu_arrays(f)
    for s, piece in enumerate(pieces):
        p = int(order[s])
        assert int(piece.get("NumberOfPoints")) == 2 * p * p + 4 * p + 4
        arr = arrays[s]
        vals = arr["stkmob"].reshape(-1, 3)
        assert np.allclose(vals[0], poles[s, 0], rtol=0, atol=1e-14)
        assert np.allclose(vals[-1], poles[s, 1], rtol=0, atol=1e-14)
        w = arr["weights"]
        assert w[0] == 0.0 and w[-1] == 0.0 and abs(w.sum() - 4 * np.pi * rad[s] ** 2) < 1e-12
        pts = arr["points"].reshape(-1, 3)
        assert np.allclose(np.linalg.norm(pts - pos[s], axis=1), rad[s], rtol=0, atol=1e-13)


@pytest.mark.parametrize("ncol", [2, 3, 5, 9])
def test_multi_rhs_solve_matches_single_solves(ncol):
    """Several right-hand sides on one geometry (two suspensions) in one lock-step
    FP64 solve: each column equals the oracle's solve of that column (the traction
    far field shared by blocks of up to 8 columns; 9 = 8 + 1), a zero column
    takes the zero-RHS shortcut, and the handle's own warm start is untouched."""
    say("multi-RHS solve ncol=%d" % ncol)
    geos = [small_geometry(seed=61, n=8, orders=(6, 8)), small_geometry(seed=62, n=6, orders=(7,))]
    pos = np.concatenate([g[0] + 40.0 * i for i, g in enumerate(geos)])
    rad = np.concatenate([g[1] for g in geos])
    quat = np.concatenate([g[2] for g in geos])
    order = np.concatenate([g[3] for g in geos])
    counts = [len(g[1]) for g in geos]
    n6 = 6 * len(rad)
    F = np.random.default_rng(70 + ncol).standard_normal((ncol, n6))
    if ncol == 3:
        F[1] = 0.0
    gpu = S.SphereSTKMobMat(pos, rad, quat, order, counts=counts)
    v0, _ = gpu.solveMob(F[0])
    x0 = gpu.get_unknown()
    V, info = gpu.solveMob_multi(F)
    assert V.shape == (ncol, n6)
    assert info["converged"] == 1 and info["relres"] <= 1e-8
    assert np.array_equal(gpu.get_unknown(), x0)
    for c in range(ncol):
        if not F[c].any():
            assert not V[c].any()
            continue
        o = 0
        for g, k in zip(geos, counts):
            pg = g[0] + 40.0 * (o > 0)
            vr, _ = O.MobilityOperator(pg, g[1], g[2], g[3]).solve(F[c, 6 * o:6 * (o + k)], tol=1e-8)
            assert rel(V[c, 6 * o:6 * (o + k)], vr) < 1e-10
            o += k
    V2, _ = gpu.solveMob_multi(F[::-1].copy())  # second call reuses the copies
    assert relmax(V2[::-1], V) < 1e-10
    gpu.set_unknown(np.zeros_like(x0))
    v1, _ = gpu.solveMob(F[0])
    assert rel(v1, v0) < 1e-12


def test_multi_rhs_mixed_form():
    """The mixed form takes several right-hand sides as a batch of copies: each
    column meets the mixed bar (1e-7) against its own FP64 solve."""
    say("multi-RHS solve, mixed form, 3 columns")
    pos, rad, quat, order = small_geometry(seed=71, n=8, orders=(6, 8))
    F = np.random.default_rng(72).standard_normal((3, 48))
    mix = S.SphereSTKMobMat(pos, rad, quat, order, precision=S.PREC_MIXED)
    V, info = mix.solveMob_multi(F)
    assert info["converged"] == 1 and info["applies_fp32"] > 0
    ref = S.SphereSTKMobMat(pos, rad, quat, order)
    for k in range(3):
        ref.set_unknown(np.zeros(ref.op.n_dof))
        v, _ = ref.solveMob(F[k])
        assert rel(V[k], v) < 1e-7


def test_multi_rhs_solve_large_matches_sequential():
    """The shared traction far field at a size with split source slices: eight
    columns in one solve against eight single FP64 solves on the GPU."""
    say("multi-RHS solve 256 spheres p=8, 8 columns")
    c = S.configs.make_suspension(256, 8, seed=9, vf=0.15, random_orientation=True)
    F = np.random.default_rng(80).standard_normal((8, 6 * len(c["radius"])))
    gpu = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"])
    V, info = gpu.solveMob_multi(F)
    assert info["converged"] == 1
    for k in range(8):
        gpu.set_unknown(np.zeros(gpu.op.n_dof))
        v, _ = gpu.solveMob(F[k])
        assert rel(V[k], v) < 1e-10


@pytest.mark.parametrize("R,axis,ang,p", [(1.0, (0, 0, 1), 0.0, 8), (1.7, (1, 2, 3), 0.7, 12), (0.6, (0, 1, 0), 2.0, 6)])
def test_self_inverse_is_exact_for_an_isolated_sphere(R, axis, ang, p):
    """The mixed form's inner preconditioner M^-1 inverts a sphere's self operator
    1/2 I + K'_PV P + L exactly (all degrees, the rigid-body modes through L, the
    part beyond degree p): for one sphere of any radius and orientation
    A M^-1 x = x to round-off, with A the GPU operator itself."""
    say("self inverse, isolated sphere R=%g p=%d" % (R, p))
    q = quat_of(axis, ang) if ang else np.array(ID)
    op = S.SphereSTKSHOperator([[0.3, -1.0, 2.0]], [R], [q], np.array([p], dtype=np.int32), 0.5, 0.0, 1.0, 1.0)
    x = np.random.default_rng(p).standard_normal(op.n_dof)
    y = op.apply_self_inverse(x)
    assert np.abs(op.apply(y) - x).max() < 1e-12 * np.abs(x).max()


def test_self_inverse_preconditioned_mixed_solve():
    """With the inner solve right-preconditioned by M^-1 (SPHSTK_FLAG_INNER_PRECOND)
    the mixed solve still meets the 1e-7 bar against the FP64 oracle, in fewer
    inner iterations."""
    say("mixed solve with the isolated-sphere preconditioner")
    pos, rad, quat, order = small_geometry(seed=41, n=8, orders=(6, 8, 7))
    F = np.random.default_rng(12).standard_normal(48)
    vr, _ = O.MobilityOperator(pos, rad, quat, order).solve(F, tol=1e-10)
    its = {}
    for flag in (False, True):
        M = S.SphereSTKMobMat(pos, rad, quat, order, precision=S.PREC_MIXED, inner_precond=flag)
        v, info = M.solveMob(F)
        assert info["converged"] == 1 and rel(v, vr) < 1e-7
        its[flag] = info["iters"]
    assert its[True] < its[False]


@pytest.mark.parametrize("prec,tol", [(S.PREC_FP64, 1e-10), (S.PREC_MIXED, 1e-7)])
def test_batched_with_empty_and_single_sphere_suspensions(prec, tol):
    """Batches may hold empty suspensions and one-sphere suspensions (no far or near
    pairs): every non-empty one matches its own oracle solve, the single sphere
    the closed form V = F/(6 pi mu R), Omega = T/(8 pi mu R^3)."""
    say("batched mobility with empty and single-sphere suspensions, prec=%d" % prec)
    g1 = small_geometry(seed=81, n=4, orders=(6,))
    g2 = (np.array([[5.0, 5.0, 5.0]]), np.array([0.7]), np.array([quat_of([1, 1, 0], 0.4)]),
          np.array([8], dtype=np.int32))
    g3 = small_geometry(seed=82, n=5, orders=(7, 6))
    parts = [g1, None, g2, g3, None]
    counts = [0 if g is None else len(g[1]) for g in parts]
    geos = [g for g in parts if g is not None]
    pos = np.concatenate([g[0] for g in geos])
    rad = np.concatenate([g[1] for g in geos])
    quat = np.concatenate([g[2] for g in geos])
    order = np.concatenate([g[3] for g in geos])
    F = np.random.default_rng(83).standard_normal(6 * len(rad))
    gpu = S.SphereSTKMobMat(pos, rad, quat, order, counts=counts, precision=prec)
    v, info = gpu.solveMob(F)
    assert info["converged"] == 1
    o = 0
    for g in geos:
        k = len(g[1])
        if k == 1:
            R = g[1][0]
            want = np.concatenate([F[6 * o:6 * o + 3] / (6 * np.pi * R), F[6 * o + 3:6 * o + 6] / (8 * np.pi * R ** 3)])
            assert rel(v[6 * o:6 * o + 6], want) < (1e-12 if prec == S.PREC_FP64 else tol)
        else:
            vr, _ = O.MobilityOperator(*g).solve(F[6 * o:6 * (o + k)], tol=1e-8)
            assert rel(v[6 * o:6 * (o + k)], vr) < tol
        o += k
