"""ctypes binding of libsphstk.so (the C ABI in include/sphstk.h).

The library is built in-tree by spheresimulator_amd/csrc/Makefile.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphstk.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

EXPORTED = [
    "sphstk_last_error", "sphstk_version", "sphstk_device_count",
    "sphstk_operator_create", "sphstk_operator_create_batched", "sphstk_operator_destroy",
    "sphstk_operator_sizes", "sphstk_operator_apply", "sphstk_operator_apply_device", "sphstk_operator_apply_inner",
    "sphstk_operator_apply_outer", "sphstk_operator_apply_coarse",
    "sphstk_operator_apply_p2p", "sphstk_operator_apply_p2p_device", "sphstk_operator_apply_lop",
    "sphstk_operator_project_nullspace", "sphstk_operator_apply_self_inverse", "sphstk_operator_pole_values", "sphstk_operator_get_grid", "sphstk_operator_get_near_pairs",
    "sphstk_operator_far_shape", "sphstk_operator_kernel_stats", "sphstk_operator_reset_stats",
    "sphstk_mobility_create", "sphstk_mobility_create_batched", "sphstk_mobility_destroy",
    "sphstk_mobility_operator", "sphstk_mobility_set_params", "sphstk_mobility_solve", "sphstk_mobility_solve_multi", "sphstk_mobility_apply",
    "sphstk_mobility_get_density", "sphstk_mobility_get_unknown", "sphstk_mobility_set_unknown",
    "sphstk_mobility_set_density",
    "sphstk_brownian_velocity", "sphstk_collision_pairs", "sphstk_collision_pairs_noalloc", "sphstk_collision_lcp",
]


class LcpInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("mobility_solves", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolverParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("restart", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("explicit_check", ctypes.c_int), ("inner_tol", ctypes.c_double), ("max_refine", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("outer", ctypes.c_int), ("converged", ctypes.c_int),
                ("shortcut", ctypes.c_int), ("relres", ctypes.c_double), ("applies_fp64", ctypes.c_int),
                ("applies_fp32", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SphstkError(RuntimeError):
    pass


_LIB = None


def load():
    """Load libsphstk.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise SphstkError("libsphstk.so is not built (%s); run __graft_entry__.build() or make -C "
                          "spheresimulator_amd/csrc" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.sphstk_last_error.restype = ctypes.c_char_p
    L.sphstk_version.restype = ctypes.c_char_p
    L.sphstk_device_count.restype = ctypes.c_int
    L.sphstk_operator_create.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip] + [ctypes.c_double] * 4 + [
        ctypes.c_int, ctypes.POINTER(_vp)]
    L.sphstk_operator_create_batched.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp, _ip] + [ctypes.c_double] * 4 + [
        ctypes.c_int, ctypes.POINTER(_vp)]
    L.sphstk_operator_destroy.argtypes = [_vp]
    L.sphstk_operator_sizes.argtypes = [_vp, _i64p, _i64p, _i64p, _i64p]
    L.sphstk_operator_apply.argtypes = [_vp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.sphstk_operator_apply_device.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.sphstk_operator_apply_inner.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double]
    L.sphstk_operator_apply_outer.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double]
    L.sphstk_operator_apply_coarse.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double]
    L.sphstk_operator_pole_values.argtypes = [_vp, _dp, _dp]
    L.sphstk_operator_far_shape.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int)]
    L.sphstk_operator_apply_p2p.argtypes = [_vp, _dp, _dp] + [ctypes.c_double] * 3
    L.sphstk_operator_apply_p2p_device.argtypes = [_vp, _vp, _vp] + [ctypes.c_double] * 3
    L.sphstk_operator_apply_lop.argtypes = [_vp, _dp, _dp, ctypes.c_double]
    L.sphstk_operator_project_nullspace.argtypes = [_vp, _dp]
    L.sphstk_operator_apply_self_inverse.argtypes = [_vp, _dp, _dp]
    L.sphstk_operator_get_grid.argtypes = [_vp, _dp, _dp, _dp, _i64p, _i64p]
    L.sphstk_operator_get_near_pairs.argtypes = [_vp, _i32p, _i32p]
    L.sphstk_operator_kernel_stats.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _i64p, _dp]
    L.sphstk_operator_reset_stats.argtypes = [_vp]
    L.sphstk_mobility_create.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_double,
                                         ctypes.POINTER(SolverParams), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(_vp)]
    L.sphstk_mobility_create_batched.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp, _ip, ctypes.c_double,
                                                 ctypes.POINTER(SolverParams), ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(_vp)]
    L.sphstk_mobility_destroy.argtypes = [_vp]
    L.sphstk_mobility_operator.argtypes = [_vp]
    L.sphstk_mobility_operator.restype = _vp
    L.sphstk_mobility_set_params.argtypes = [_vp, ctypes.POINTER(SolverParams), ctypes.c_int]
    L.sphstk_mobility_solve.argtypes = [_vp, _dp, _dp, ctypes.POINTER(SolveInfo)]
    L.sphstk_mobility_solve_multi.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(SolveInfo)]
    L.sphstk_mobility_apply.argtypes = [_vp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.sphstk_mobility_get_density.argtypes = [_vp, _dp]
    L.sphstk_mobility_get_unknown.argtypes = [_vp, _dp]
    L.sphstk_mobility_set_unknown.argtypes = [_vp, _dp]
    L.sphstk_mobility_set_density.argtypes = [_vp, _dp]
    L.sphstk_brownian_velocity.argtypes = [ctypes.c_int, _dp] + [ctypes.c_double] * 4 + [ctypes.c_uint64] * 2 + [_dp]
    L.sphstk_collision_lcp.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int64, _ip, _ip, _dp, _dp,
                                       _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                       _dp, ctypes.POINTER(LcpInfo)]
    L.sphstk_collision_pairs.argtypes = [ctypes.c_int, _dp, _dp, _ip, ctypes.c_double, ctypes.c_int64, _i64p,
                                         _ip, _ip, _dp, _dp]
    L.sphstk_collision_pairs_noalloc.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int64, _i64p,
                                                 _vp, _vp, _vp, _vp, _vp]
    for name in EXPORTED:
        if name.startswith("sphstk_") and name not in ("sphstk_last_error", "sphstk_version",
                                                        "sphstk_device_count", "sphstk_mobility_operator"):
            getattr(L, name).restype = ctypes.c_int
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        msg = load().sphstk_last_error().decode()
        raise SphstkError("sphstk error %d: %s" % (rc, msg))
    return rc
