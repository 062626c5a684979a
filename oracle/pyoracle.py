"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
CPU-baseline leg of bench.py.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


class GmresParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("restart", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("explicit_check", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("restarts", ctypes.c_int), ("converged", ctypes.c_int),
                ("shortcut", ctypes.c_int), ("relres_implicit", ctypes.c_double),
                ("relres_explicit", ctypes.c_double), ("applies", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_gl.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_solid_harmonics.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_unit_grid.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_sh_analysis.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_sh_synthesis.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_sh_self_spectral.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_sh_eval_exterior.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_create.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip] + [ctypes.c_double] * 4
        L.orc_create.restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_npts.argtypes = [ctypes.c_void_p]
        L.orc_npts.restype = ctypes.c_long
        L.orc_get_grid.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _lp]
        L.orc_num_near_pairs.argtypes = [ctypes.c_void_p]
        L.orc_num_near_pairs.restype = ctypes.c_long
        L.orc_get_near_pairs.argtypes = [ctypes.c_void_p, _ip, _ip]
        L.orc_apply_p2p.argtypes = [ctypes.c_void_p, _dp, _dp] + [ctypes.c_double] * 3
        L.orc_apply_p2p_spheres.argtypes = [ctypes.c_void_p, _dp, _dp] + [ctypes.c_double] * 3 + [ctypes.c_int] * 2
        L.orc_apply_lop.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double]
        L.orc_apply.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_project.argtypes = [ctypes.c_void_p, _dp]
        L.orc_far_sums.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_far_sums_range.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_long, ctypes.c_long]
        L.orc_gmres.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, GmresParams, ctypes.POINTER(SolveInfo)]
        L.orc_mobility_solve.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp, _dp, GmresParams,
                                         ctypes.POINTER(SolveInfo)]
        L.orc_num_threads.restype = ctypes.c_int
        _up = ctypes.POINTER(ctypes.c_uint)
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]
        L.orc_brownian_velocity.argtypes = [ctypes.c_int, _dp] + [ctypes.c_double] * 4 + [ctypes.c_ulonglong] * 2 + [_dp]
        L.orc_collision_pairs.argtypes = [ctypes.c_int, _dp, _dp, _ip, ctypes.c_double, ctypes.c_long, _ip, _ip, _dp,
                                          _dp]
        L.orc_collision_pairs.restype = ctypes.c_long
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def tri_size(P):
    return (P + 1) * (P + 2) // 2


def gl(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gl(n, _d(x), _d(w))
    return x, w


def solid_harmonics(P, x):
    x = _c(x)
    D = np.zeros(2 * tri_size(P))
    lib().orc_solid_harmonics(P, _d(x), _d(D))
    return D[0::2] + 1j * D[1::2]


def unit_grid(p):
    n = (p + 1) * (2 * p + 2)
    eta = np.zeros((n, 3))
    w = np.zeros(n)
    lib().orc_unit_grid(p, _d(eta), _d(w))
    return eta, w


def sh_analysis(p, vals):
    vals = _c(vals)
    coef = np.zeros(3 * (p + 1) * (p + 2))
    lib().orc_sh_analysis(p, _d(vals), _d(coef))
    return coef


def sh_synthesis(p, coef):
    coef = _c(coef)
    n = (p + 1) * (2 * p + 2)
    vals = np.zeros((n, 3))
    lib().orc_sh_synthesis(p, _d(coef), _d(vals))
    return vals


def sh_self(p, which, coef):
    coef = _c(coef)
    out = np.zeros_like(coef)
    lib().orc_sh_self_spectral(p, which, _d(coef), _d(out))
    return out


def sh_eval_exterior(p, coef, xi, nu=None):
    coef = _c(coef)
    xi = _c(xi).reshape(-1, 3)
    m = xi.shape[0]
    u = np.zeros((m, 3))
    t = np.zeros((m, 3))
    if nu is None:
        lib().orc_sh_eval_exterior(p, _d(coef), m, _d(xi), None, _d(u), None)
        return u
    nu = _c(nu).reshape(-1, 3)
    lib().orc_sh_eval_exterior(p, _d(coef), m, _d(xi), _d(nu), _d(u), _d(t))
    return u, t


class Operator:
    """FP64 CPU restatement of SphereSTKSHOperator (BIE/SphereSTKSHOperator.hpp:112-231)."""

    def __init__(self, pos, radius, quat, order, cId=0.0, cSL=0.0, cTrac=0.0, cLOP=0.0):
        self.pos = _c(pos).reshape(-1, 3)
        self.nsph = self.pos.shape[0]
        self.radius = _c(radius)
        self.quat = _c(quat).reshape(-1, 4)
        self.order = _c(order, np.int32)
        self.coeffs = (cId, cSL, cTrac, cLOP)
        self._h = lib().orc_create(self.nsph, _d(self.pos), _d(self.radius), _d(self.quat),
                                   self.order.ctypes.data_as(_ip), cId, cSL, cTrac, cLOP)
        self.npts = lib().orc_npts(self._h)
        self.dof = 3 * self.npts

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def grid(self):
        c = np.zeros((self.npts, 3))
        nv = np.zeros((self.npts, 3))
        w = np.zeros(self.npts)
        off = np.zeros(self.nsph + 1, dtype=np.int64)
        lib().orc_get_grid(self._h, _d(c), _d(nv), _d(w), off.ctypes.data_as(_lp))
        return c, nv, w, off

    def near_pairs(self):
        k = lib().orc_num_near_pairs(self._h)
        t = np.zeros(k, dtype=np.int32)
        s = np.zeros(k, dtype=np.int32)
        if k:
            lib().orc_get_near_pairs(self._h, t.ctypes.data_as(_ip), s.ctypes.data_as(_ip))
        return t, s

    def apply_p2p(self, x, out=None, cId=0.0, cSL=0.0, cTrac=0.0):
        x = _c(x).ravel()
        out = np.zeros(self.dof) if out is None else out
        lib().orc_apply_p2p(self._h, _d(x), _d(out), cId, cSL, cTrac)
        return out

    def apply_p2p_spheres(self, x, T0, T1, out=None, cId=0.0, cSL=0.0, cTrac=0.0):
        """apply_p2p evaluated on the points of target spheres [T0, T1) only;
        the other entries of out are left as they were."""
        x = _c(x).ravel()
        out = np.zeros(self.dof) if out is None else out
        lib().orc_apply_p2p_spheres(self._h, _d(x), _d(out), cId, cSL, cTrac, int(T0), int(T1))
        return out

    def apply_lop(self, x, out=None, cLOP=1.0):
        x = _c(x).ravel()
        out = np.zeros(self.dof) if out is None else out
        lib().orc_apply_lop(self._h, _d(x), _d(out), cLOP)
        return out

    def apply(self, x, y=None, alpha=1.0, beta=0.0):
        x = _c(x).ravel()
        y = np.zeros(self.dof) if y is None else y
        lib().orc_apply(self._h, _d(x), _d(y), alpha, beta)
        return y

    def project(self, x):
        x = _c(x).ravel().copy()
        lib().orc_project(self._h, _d(x))
        return x

    def far_sums(self, x, cSL=0.0, cTrac=0.0):
        x = _c(x).ravel()
        out = np.zeros(self.dof)
        lib().orc_far_sums(self._h, _d(x), _d(out), cSL, cTrac)
        return out

    def far_sums_range(self, x, t0, t1, cSL=0.0, cTrac=0.0):
        x = _c(x).ravel()
        out = np.zeros(self.dof)
        lib().orc_far_sums_range(self._h, _d(x), _d(out), cSL, cTrac, t0, t1)
        return out

    def gmres(self, b, x0=None, tol=1e-8, restart=30, max_iters=1000, explicit_check=1):
        b = _c(b).ravel()
        x = np.zeros(self.dof) if x0 is None else _c(x0).ravel().copy()
        info = SolveInfo()
        prm = GmresParams(tol, restart, max_iters, explicit_check)
        lib().orc_gmres(self._h, _d(b), _d(x), 1 if x0 is None else 0, prm, ctypes.byref(info))
        return x, info.as_dict()


class MobilityOperator(Operator):
    """FP64 CPU restatement of SphereSTKMobMat (BIE/SphereSTKMobMat.hpp:26-79)."""

    def __init__(self, pos, radius, quat, order, viscosity=1.0):
        super().__init__(pos, radius, quat, order, 0.5, 0.0, 1.0, 1.0)
        self.viscosity = viscosity
        self.x_last = None

    def solve(self, force, tol=1e-8, restart=30, max_iters=1000, explicit_check=1, warm=False):
        force = _c(force).ravel()
        vel = np.zeros(6 * self.nsph)
        info = SolveInfo()
        prm = GmresParams(tol, restart, max_iters, explicit_check)
        x = self.x_last.copy() if (warm and self.x_last is not None) else np.zeros(self.dof)
        lib().orc_mobility_solve(self._h, self.viscosity, _d(force), _d(vel), _d(x), prm, ctypes.byref(info))
        self.x_last = x
        return vel, info.as_dict()


def num_threads():
    return lib().orc_num_threads()


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block: four 32-bit words from counter (4 words) and key (2 words)."""
    c = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32))
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
    o = np.zeros(4, dtype=np.uint32)
    up = ctypes.POINTER(ctypes.c_uint)
    lib().orc_philox4x32_10(c.ctypes.data_as(up), k.ctypes.data_as(up), o.ctypes.data_as(up))
    return o


def brownian_velocity(radius, dt, kBT, viscosity=1.0, scale=1.0, seed=0, step=0):
    r = _c(np.asarray(radius, dtype=np.float64).ravel())
    v = np.zeros(6 * r.size)
    lib().orc_brownian_velocity(r.size, _d(r), viscosity, kBT, dt, scale, seed, step, _d(v))
    return v


def collision_pairs(pos, radius_col, colbuf, gid=None):
    """Brute-force collision pairs: (idx_i, idx_j, phi0, norm_i[m, 3]) ordered by I then J."""
    p = _c(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    rc = _c(np.asarray(radius_col, dtype=np.float64).ravel())
    n = rc.size
    ip = ctypes.POINTER(ctypes.c_int)
    g = None if gid is None else _c(np.asarray(gid, dtype=np.int32).ravel(), np.int32)
    gp = g.ctypes.data_as(ip) if g is not None else None
    m = lib().orc_collision_pairs(n, _d(p), _d(rc), gp, colbuf, 0, None, None, None, None)
    ii = np.zeros(m, dtype=np.int32)
    jj = np.zeros(m, dtype=np.int32)
    phi0 = np.zeros(m)
    nrm = np.zeros(3 * m)
    if m:
        lib().orc_collision_pairs(n, _d(p), _d(rc), gp, colbuf, m, ii.ctypes.data_as(ip), jj.ctypes.data_as(ip),
                                  _d(phi0), _d(nrm))
    return ii, jj, phi0, nrm.reshape(-1, 3)



def free_drain(radius, mu):
    """Diagonal mobility without hydrodynamic interaction (SRC/SphereSystem.cpp:237-268)."""
    r = np.asarray(radius, dtype=np.float64).ravel()
    d = np.repeat(np.stack([1 / (6 * np.pi * r * mu), 1 / (8 * np.pi * r ** 3 * mu)], 1), 3, axis=1).ravel()
    return lambda F: d * F


def lcp_bbpgd(mobility, n, ii, jj, phi0, nrm, vel_known, dt, buffer=0.0, tol=1e-5, max_iters=2000, gamma0=None):
    """Collision LCP by Barzilai-Borwein projected gradient descent (Barzilai and
    Borwein 1988; Yan et al., J. Comput. Phys. 2019), the solver SimToolbox's
    CollisionSolver applies around the mobility (SRC/SphereSystem.cpp:445-474):
    gamma >= 0, phi = phi0 + buffer + dt D^T (v_known + M D gamma) >= 0,
    gamma . phi = 0; residual max |min(gamma, phi)|.  Small cases only."""
    ii = np.asarray(ii)
    jj = np.asarray(jj)
    nrm = np.asarray(nrm, dtype=np.float64).reshape(-1, 3)
    m = len(ii)

    def D(g):
        F = np.zeros((n, 6))
        np.add.at(F[:, :3], ii, g[:, None] * nrm)
        np.add.at(F[:, :3], jj, -g[:, None] * nrm)
        return F.ravel()

    def Dt(v):
        v = v.reshape(n, 6)[:, :3]
        return np.einsum("kc,kc->k", nrm, v[ii] - v[jj])

    info = dict(iters=0, mobility_solves=0, converged=0, residual=0.0)
    if m == 0:
        info["converged"] = 1
        return np.zeros(0), np.zeros(6 * n), np.zeros(6 * n), info
    b = np.asarray(phi0, dtype=np.float64) + buffer + dt * Dt(np.asarray(vel_known, dtype=np.float64))

    def A(x):
        u = mobility(D(x))
        info["mobility_solves"] += 1
        return dt * Dt(u), u

    g = np.maximum(np.zeros(m) if gamma0 is None else np.asarray(gamma0, dtype=np.float64), 0.0)
    Ag, U = A(g)
    phi = Ag + b
    res = np.abs(np.minimum(g, phi)).max()
    if res > tol:
        Aphi, _ = A(phi)
        gAg = phi @ Aphi
        alpha = (phi @ phi) / gAg if gAg > 0 else 1.0
        for it in range(1, max_iters + 1):
            gn = np.maximum(g - alpha * phi, 0.0)
            Agn, Un = A(gn)
            phin = Agn + b
            info["iters"] = it
            res = np.abs(np.minimum(gn, phin)).max()
            s, y = gn - g, phin - phi
            g, phi, U = gn, phin, Un
            if res <= tol:
                break
            ss, sy, yy = s @ s, s @ y, y @ y
            if ss == 0.0:
                break
            a = ss / sy if it % 2 else sy / yy
            if np.isfinite(a) and a > 0:
                alpha = a
    info["residual"] = float(res)
    info["converged"] = int(res <= tol)
    return g, D(g), U, info


def lcp_enumerate(A, b):
    """Exact solution of the LCP (A positive definite, hence a P-matrix with one
    solution; small m) by trying every active set:
    gamma_S = -A_SS^-1 b_S >= 0 and phi = A gamma + b >= 0 off S."""
    import itertools
    m = len(b)
    best = None
    for k in range(m + 1):
        for S in itertools.combinations(range(m), k):
            S = list(S)
            g = np.zeros(m)
            if S:
                g[S] = np.linalg.solve(A[np.ix_(S, S)], -b[S])
            phi = A @ g + b
            if (g >= -1e-12).all() and (phi >= -1e-12).all():
                if best is None:
                    best = g
    return best
