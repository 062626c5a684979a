"""The collision-pair search that allocates nothing
(sphstk_collision_pairs_noalloc: device arrays in and out, intermediate arrays
in the library's static device scratch) against the allocating entry point
(sphstk_collision_pairs, host arrays) on the same inputs: the same count, the
same pairs in the same order, phi0 and normals bit for bit (the same kernels).
Two calls in a row, the second with fewer spheres (the scratch keeps the
first call's contents, so everything the second reads must be cleared by it),
then the count-only and small-capacity modes, the index as id, and the
rejection of a search larger than the scratch.  Device arrays come from the
HIP runtime that libsphstk.so itself links (hipMalloc through ctypes; test
plumbing only)."""
import ctypes
import sys

import numpy as np
import pytest

import spheresimulator_amd as S

pytestmark = pytest.mark.gpu

def say(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def allocating(pos, rc, cb, gid):
    """The existing entry point, through its Python wrapper."""
    return S.system.collision_pairs(pos, rc, cb, gid=gid)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libsphstk.so links."""

    def __init__(self):
        S.load()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.ok(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)))
        if a.nbytes:
            self.ok(self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1))
        return p

    def get(self, p, n, dtype):
        a = np.empty(n, dtype=dtype)
        if a.nbytes:
            self.ok(self.L.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, a.nbytes, 2))
        return a

    def free(self, *ps):
        for p in ps:
            if p is not None:
                self.ok(self.L.hipFree(p))


def noalloc(pos, rc, cb, gid, capacity=None, outputs=True):
    H = _Hip()
    L = S.load()
    n = len(rc)
    dpos = H.put(np.asarray(pos, dtype=np.float64).ravel())
    drc = H.put(np.asarray(rc, dtype=np.float64))
    dgid = None if gid is None else H.put(np.asarray(gid, dtype=np.int32))
    cnt = ctypes.c_int64()
    try:
        # count first, then fill (as a caller with preallocated output would)
        S._lib.check(L.sphstk_collision_pairs_noalloc(n, dpos, drc, dgid, cb, 0, ctypes.byref(cnt), None, None, None,
                                                      None, None))
        m = cnt.value
        if not outputs:
            return m
        cap = m if capacity is None else capacity
        k = max(cap, 1)
        ii = H.put(np.full(k, -7, dtype=np.int32))
        jj = H.put(np.full(k, -7, dtype=np.int32))
        phi = H.put(np.full(k, -7.0))
        nrm = H.put(np.full(3 * k, -7.0))
        try:
            S._lib.check(L.sphstk_collision_pairs_noalloc(n, dpos, drc, dgid, cb, cap, ctypes.byref(cnt), ii, jj, phi,
                                                          nrm, None))
            assert cnt.value == m
            hi = H.get(ii, k, np.int32)
            return dict(m=m, idx_i=hi[:m], idx_j=H.get(jj, k, np.int32)[:m], phi0=H.get(phi, k, np.float64)[:m],
                        norm_i=H.get(nrm, 3 * k, np.float64)[:3 * m].reshape(-1, 3), untouched=bool((hi == -7).all()))
        finally:
            H.free(ii, jj, phi, nrm)
    finally:
        H.free(dpos, drc, dgid)


def same(a, b):
    assert a["m"] == len(b["idx_i"])
    assert np.array_equal(a["idx_i"], b["idx_i"]) and np.array_equal(a["idx_j"], b["idx_j"])
    assert np.array_equal(a["phi0"], b["phi0"]) and np.array_equal(a["norm_i"], b["norm_i"])


def test_noalloc_matches_allocating_search_twice_then_smaller():
    c = S.configs.config5()
    say("no-alloc collision pairs: 16384-sphere packing")
    first = noalloc(c["pos"], c["radius"], 0.3, None)
    same(first, allocating(c["pos"], c["radius"], 0.3, None))
    assert first["m"] > 10000
    rng = np.random.default_rng(11)
    n = 3000
    pos = rng.uniform(0, 30, (n, 3))
    rc = rng.uniform(0.3, 1.2, n)
    gid = rng.permutation(n).astype(np.int32)
    say("no-alloc collision pairs: then 3000 polydisperse spheres, permuted gid")
    second = noalloc(pos, rc, 0.3, gid)
    same(second, allocating(pos, rc, 0.3, gid))
    assert second["m"] > n


def test_noalloc_count_only_capacity_and_limits():
    say("no-alloc collision pairs: count only, small capacity, n < 2, limit")
    k, d = 8, 2.25
    g = np.arange(k)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * d
    rc = np.ones(len(pos))
    want = 3 * k * k * (k - 1)
    assert noalloc(pos, rc, 0.3, None, outputs=False) == want
    small = noalloc(pos, rc, 0.3, None, capacity=4)
    assert small["m"] == want and small["untouched"]
    same(noalloc(pos, rc, 0.3, None), allocating(pos, rc, 0.3, None))
    assert noalloc(np.zeros((1, 3)), [1.0], 0.3, None, outputs=False) == 0
    L = S.load()
    cnt = ctypes.c_int64()
    rc_ = L.sphstk_collision_pairs_noalloc(262145, None, None, None, 0.3, 0, ctypes.byref(cnt), None,
                                           None, None, None, None)
    assert rc_ == -1 and b"exceeds" in L.sphstk_last_error()
    with pytest.raises(S.SphstkError):
        noalloc(np.array([[0.0, 0, 0], [np.nan, 0, 0]]), [1.0, 1.0], 0.3, None, outputs=False)
