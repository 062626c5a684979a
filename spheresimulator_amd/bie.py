"""Python mirror of the reference's BIE layer over the C ABI.

SphereSTKSHOperator  <->  BIE/SphereSTKSHOperator.hpp:112-231
SphereSTKMobMat      <->  BIE/SphereSTKMobMat.hpp:26-79

Same names, argument meaning and error behaviour (size mismatches raise, the
way TEUCHOS_ASSERT throws in the reference).  Vectors are numpy float64 arrays
in the reference's DOF layout; every call runs on the GPU through libsphstk.so.
"""
import ctypes

import numpy as np

from ._lib import SolveInfo, SolverParams, check, load

_dp = ctypes.POINTER(ctypes.c_double)

PREC_FP64 = 0
PREC_MIXED = 1
FLAG_TIMING = 1
FLAG_FP32_DIRECT = 2
FLAG_AUTOTUNE = 4
FLAG_GRAPH = 8
FLAG_INNER_PRECOND = 16
FLAG_NO_CELLS = 32
FLAG_EXPLICIT_REFINE = 64
FLAG_NO_COARSE = 128


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a, n=None, name="array"):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError("%s has %d entries, expected %d" % (name, a.size, n))
    return a


def _out(a, n, name):
    """An output array the library writes n doubles into: float64, C-contiguous, n entries."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous or a.size != n:
        raise ValueError("%s must be a C-contiguous float64 array of %d entries" % (name, n))
    if not a.flags.writeable:
        raise ValueError("%s must be writeable" % name)
    return a


class _Spheres:
    def __init__(self, pos, radius, quat, order, counts):
        self.pos = _f64(pos, name="pos")
        self.n = self.pos.size // 3
        self.radius = _f64(radius, self.n, "radius")
        self.quat = _f64(quat if quat is not None else np.tile([1.0, 0, 0, 0], (self.n, 1)), 4 * self.n, "quat")
        self.order = np.ascontiguousarray(np.broadcast_to(np.asarray(order, dtype=np.int32), (self.n,)))
        self.counts = np.ascontiguousarray(np.asarray(counts if counts is not None else [self.n], dtype=np.int32))
        if int(self.counts.sum()) != self.n:
            raise ValueError("suspension sphere counts do not add up to the number of spheres")


class SphereSTKSHOperator:
    """Y := beta Y + alpha (cId I + cSL S + cTrac K' + cLOP L) X on the GPU."""

    def __init__(self, pos, radius, quat, order, cId=0.0, cSL=0.0, cTrac=0.0, cLOP=0.0, counts=None, timing=False,
                 fp32_direct=False, autotune=False, graph=False, cells=True, _handle=None, _owner=None):
        L = load()
        self._owner = _owner
        if _handle is not None:
            self._h = ctypes.c_void_p(_handle)
            self._own = False
        else:
            s = _Spheres(pos, radius, quat, order, counts)
            h = ctypes.c_void_p()
            check(L.sphstk_operator_create_batched(len(s.counts), s.counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                   _d(s.pos), _d(s.radius), _d(s.quat),
                                                   s.order.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                   cId, cSL, cTrac, cLOP,
                                                   (FLAG_TIMING if timing else 0) |
                                                   (FLAG_FP32_DIRECT if fp32_direct else 0) |
                                                   (FLAG_AUTOTUNE if autotune else 0) |
                                                   (FLAG_GRAPH if graph else 0) |
                                                   (0 if cells else FLAG_NO_CELLS), ctypes.byref(h)))
            self._h = h
            self._own = True
        ns, npt, nd, nn = (ctypes.c_int64() for _ in range(4))
        check(L.sphstk_operator_sizes(self._h, ctypes.byref(ns), ctypes.byref(npt), ctypes.byref(nd),
                                      ctypes.byref(nn)))
        self.n_spheres, self.n_points, self.n_dof, self.n_near_pairs = ns.value, npt.value, nd.value, nn.value

    def __del__(self):
        if getattr(self, "_own", False) and getattr(self, "_h", None):
            load().sphstk_operator_destroy(self._h)
            self._h = None

    # --- SphereSTKSHOperator interface -----------------------------------
    def apply(self, X, Y=None, alpha=1.0, beta=0.0):
        X = _f64(X)
        if X.size % self.n_dof:
            raise ValueError("X length is not a multiple of the DOF count")
        ncol = X.size // self.n_dof
        Y = np.zeros(X.size) if Y is None else _out(Y, X.size, "Y")
        check(load().sphstk_operator_apply(self._h, _d(X), _d(Y), ncol, alpha, beta))
        return Y

    def apply_inner(self, X, Y=None, alpha=1.0, beta=0.0):
        """The FP32 inner operator of the mixed-precision solve (one column)."""
        X = _f64(X, self.n_dof, "X")
        Y = np.zeros(self.n_dof) if Y is None else _out(Y, self.n_dof, "Y")
        check(load().sphstk_operator_apply_inner(self._h, _d(X), _d(Y), alpha, beta))
        return Y

    def apply_outer(self, X, Y=None, alpha=1.0, beta=0.0):
        """The FP64 multi-level outer operator of the mixed-precision solve (one column)."""
        X = _f64(X, self.n_dof, "X")
        Y = np.zeros(self.n_dof) if Y is None else _out(Y, self.n_dof, "Y")
        check(load().sphstk_operator_apply_outer(self._h, _d(X), _d(Y), alpha, beta))
        return Y

    def apply_coarse(self, X, Y=None, alpha=1.0, beta=0.0):
        """The coarse FP32 operator of the mixed-precision solve's inner solves (one column)."""
        X = _f64(X, self.n_dof, "X")
        Y = np.zeros(self.n_dof) if Y is None else _out(Y, self.n_dof, "Y")
        check(load().sphstk_operator_apply_coarse(self._h, _d(X), _d(Y), alpha, beta))
        return Y

    def applyP2POP(self, inp, out, cId, cSL, cTrac):
        """out += cId in + cSL S[in] + cTrac K'[in]  (accumulates, as the reference)."""
        inp = _f64(inp, self.n_dof, "in")
        _out(out, self.n_dof, "out")
        check(load().sphstk_operator_apply_p2p(self._h, _d(inp), _d(out), cId, cSL, cTrac))
        return out

    def applyLOP(self, inp, out, cLOP):
        inp = _f64(inp, self.n_dof, "in")
        _out(out, self.n_dof, "out")
        check(load().sphstk_operator_apply_lop(self._h, _d(inp), _d(out), cLOP))
        return out

    def apply_self_inverse(self, x):
        """M^-1 x with M each sphere's self operator 1/2 I + K'_self P + L (exact inverse per
        sphere and SH degree; the mixed form's inner preconditioner)."""
        x = _f64(x, self.n_dof, "x")
        y = np.zeros(self.n_dof)
        check(load().sphstk_operator_apply_self_inverse(self._h, _d(x), _d(y)))
        return y

    def projectNullSpace(self, inout):
        _out(inout, self.n_dof, "inout")
        check(load().sphstk_operator_project_nullspace(self._h, _d(inout)))
        return inout

    def grid(self):
        n, ns = self.n_points, self.n_spheres
        w = np.zeros(n)
        c = np.zeros(3 * n)
        nv = np.zeros(3 * n)
        idx = np.zeros(ns, dtype=np.int64)
        ln = np.zeros(ns, dtype=np.int64)
        i64 = ctypes.POINTER(ctypes.c_int64)
        check(load().sphstk_operator_get_grid(self._h, _d(w), _d(c), _d(nv), idx.ctypes.data_as(i64),
                                              ln.ctypes.data_as(i64)))
        return dict(weights=w, coords=c, norms=nv, index=idx, length=ln)

    def getGridWeights(self):
        return self.grid()["weights"]

    def getGridCoords(self):
        return self.grid()["coords"]

    def getGridNorms(self):
        return self.grid()["norms"]

    def getGridNumberIndex(self):
        return self.grid()["index"]

    def getGridNumberLength(self):
        return self.grid()["length"]

    def near_pairs(self):
        k = self.n_near_pairs
        t = np.zeros(k, dtype=np.int32)
        s = np.zeros(k, dtype=np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        check(load().sphstk_operator_get_near_pairs(self._h, t.ctypes.data_as(i32), s.ctypes.data_as(i32)))
        return t, s

    def kernel_stats(self, kind=0, prec=0):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        pairs = ctypes.c_double()
        check(load().sphstk_operator_kernel_stats(self._h, kind, prec, ctypes.byref(ms), ctypes.byref(n),
                                                  ctypes.byref(pairs)))
        return dict(ms=ms.value, launches=n.value, pairs=pairs.value)

    def reset_stats(self):
        check(load().sphstk_operator_reset_stats(self._h))

    def pole_values(self, x):
        """SH expansion of the grid field x at each sphere's two poles (lab frame), shape (N, 2, 3)."""
        x = _f64(x, self.n_dof, "x")
        out = np.zeros(6 * self.n_spheres)
        check(load().sphstk_operator_pole_values(self._h, _d(x), _d(out)))
        return out.reshape(-1, 2, 3)

    def far_shape(self):
        """Far-field work-list shapes in use: FP64 targets per lane and source slices, FP32 targets per lane."""
        t, s, i = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(load().sphstk_operator_far_shape(self._h, ctypes.byref(t), ctypes.byref(s), ctypes.byref(i)))
        return dict(tpl=t.value, split=s.value, inner_tpl=i.value)


def solver_params(tol=1e-8, restart=30, max_iters=1000, explicit_check=1, inner_tol=3e-5, max_refine=10):
    return SolverParams(tol, restart, max_iters, explicit_check, inner_tol, max_refine)


class SphereSTKMobMat:
    """Force/torque (6N) -> velocity/angular velocity (6N) through A = 0.5 I + K' + L."""

    def __init__(self, pos, radius, quat, order, viscosity=1.0, counts=None, precision=PREC_FP64, params=None,
                 timing=False, fp32_direct=False, autotune=False, graph=False, inner_precond=False, cells=True,
                 explicit_refine=False, coarse=True):
        L = load()
        s = _Spheres(pos, radius, quat, order, counts)
        self.n_spheres = s.n
        self.params = params or solver_params()
        h = ctypes.c_void_p()
        check(L.sphstk_mobility_create_batched(len(s.counts), s.counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                               _d(s.pos), _d(s.radius), _d(s.quat),
                                               s.order.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), viscosity,
                                               ctypes.byref(self.params), precision,
                                               (FLAG_TIMING if timing else 0) |
                                               (FLAG_FP32_DIRECT if fp32_direct else 0) |
                                               (FLAG_AUTOTUNE if autotune else 0) |
                                               (FLAG_GRAPH if graph else 0) |
                                               (FLAG_INNER_PRECOND if inner_precond else 0) |
                                               (FLAG_EXPLICIT_REFINE if explicit_refine else 0) |
                                               (0 if coarse else FLAG_NO_COARSE) |
                                               (0 if cells else FLAG_NO_CELLS), ctypes.byref(h)))
        self._h = h
        self.n_dof = self.op.n_dof

    @property
    def op(self):
        """The operator A = 0.5 I + K' + L owned by this mobility.  A view made on
        each access: it keeps the mobility alive while it is referenced, and the
        mobility holds no reference to it, so dropping the mobility frees its
        device memory at once."""
        return SphereSTKSHOperator(None, None, None, None, _handle=load().sphstk_mobility_operator(self._h),
                                   _owner=self)

    def __del__(self):
        if getattr(self, "_h", None):
            load().sphstk_mobility_destroy(self._h)
            self._h = None

    def set_params(self, params=None, precision=PREC_FP64):
        self.params = params or self.params
        check(load().sphstk_mobility_set_params(self._h, ctypes.byref(self.params), precision))

    def solveMob(self, force):
        force = _f64(force, 6 * self.n_spheres, "force")
        vel = np.zeros(6 * self.n_spheres)
        info = SolveInfo()
        check(load().sphstk_mobility_solve(self._h, _d(force), _d(vel), ctypes.byref(info)))
        return vel, info.as_dict()

    def solveMob_multi(self, force):
        """Several right-hand sides (ncol, 6N) in one FP64 lock-step solve from x0 = 0; returns (ncol, 6N)."""
        force = np.ascontiguousarray(np.asarray(force, dtype=np.float64).reshape(-1, 6 * self.n_spheres))
        vel = np.zeros_like(force)
        info = SolveInfo()
        check(load().sphstk_mobility_solve_multi(self._h, force.shape[0], _d(force), _d(vel), ctypes.byref(info)))
        return vel, info.as_dict()

    def apply(self, X, Y=None, alpha=1.0, beta=0.0):
        X = _f64(X)
        n6 = 6 * self.n_spheres
        if X.size % n6:
            raise ValueError("X length is not a multiple of 6 N")
        Y = np.zeros(X.size) if Y is None else _out(Y, X.size, "Y")
        check(load().sphstk_mobility_apply(self._h, _d(X), _d(Y), X.size // n6, alpha, beta))
        return Y

    def getDensitySolution(self):
        rho = np.zeros(self.n_dof)
        check(load().sphstk_mobility_get_density(self._h, _d(rho)))
        return rho

    def get_unknown(self):
        x = np.zeros(self.n_dof)
        check(load().sphstk_mobility_get_unknown(self._h, _d(x)))
        return x

    def set_unknown(self, x):
        x = _f64(x, self.n_dof, "x")
        check(load().sphstk_mobility_set_unknown(self._h, _d(x)))

    def set_density(self, rho):
        """Warm start from a stored density, x = projectNullSpace(rho), as the
        reference's constructor does with the layer's gridValue (MobMat.cpp:67-93)."""
        rho = _f64(rho, self.n_dof, "rho")
        check(load().sphstk_mobility_set_density(self._h, _d(rho)))
