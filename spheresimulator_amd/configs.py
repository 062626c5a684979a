"""Seeded synthetic suspensions for the five benchmark/parity configurations.

All randomness comes from a counter-based SplitMix64 stream (one seed per
configuration), so the CPU oracle and the GPU library consume bit-identical
inputs.  Geometry follows the reference's input conventions: radius-R spheres,
one SH layer of order p per sphere with the identity (or a seeded uniform)
orientation quaternion (w, x, y, z) as in SRC/SphereSystem.cpp:54-57, forces
and torques as [fx, fy, fz, tx, ty, tz] per sphere (BIE/SphereSTKMobMat.cpp:176-177).
"""
import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """Counter-based SplitMix64: value i = mix(seed + (i + 1) * gamma)."""

    _BLOCK = 65536

    def __init__(self, seed):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.ctr = 0
        self._buf = np.zeros(0, dtype=np.uint64)  # values ctr .. ctr + len - 1, computed ahead

    def _mix(self, first, n):
        with np.errstate(over="ignore"):
            i = np.arange(first + 1, first + 1 + n, dtype=np.uint64)
            z = self.seed + i * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def u64(self, n):
        if n > len(self._buf):
            self._buf = self._mix(self.ctr, max(n, self._BLOCK))
        z = self._buf[:n]
        self._buf = self._buf[n:]
        self.ctr += n
        return z

    def uniform(self, n):
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def unit_vectors(self, n):
        z = 2.0 * self.uniform(n) - 1.0
        ph = 2.0 * np.pi * self.uniform(n)
        s = np.sqrt(1.0 - z * z)
        return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)

    def quaternions(self, n):
        """Uniform random unit quaternions (Shoemake)."""
        u = self.uniform(3 * n).reshape(n, 3)
        a = np.sqrt(1 - u[:, 0])
        b = np.sqrt(u[:, 0])
        return np.stack([b * np.cos(2 * np.pi * u[:, 2]), a * np.sin(2 * np.pi * u[:, 1]),
                         a * np.cos(2 * np.pi * u[:, 1]), b * np.sin(2 * np.pi * u[:, 2])], axis=1)


class _Grid:
    """Cell hash for non-overlap checks during sequential placement."""

    _OFF = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]

    def __init__(self, h):
        self.h = h
        self.cells = {}
        self.n = 0
        self._p = np.zeros((1024, 3))
        self._r = np.zeros(1024)

    @property
    def pos(self):
        return self._p[:self.n]

    @property
    def rad(self):
        return self._r[:self.n]

    def key(self, x):
        return (int(math.floor(x[0] / self.h)), int(math.floor(x[1] / self.h)), int(math.floor(x[2] / self.h)))

    def ok(self, x, r, gap_fn, skip=-1):
        k0, k1, k2 = self.key(x)
        get = self.cells.get
        idx = [j for dx, dy, dz in self._OFF for j in get((k0 + dx, k1 + dy, k2 + dz), ()) if j != skip]
        if not idx:
            return True
        d = x - self._p[idx]
        lim = r + self._r[idx] + gap_fn(r, self._r[idx])
        return not bool(np.any(np.einsum("ij,ij->i", d, d) < lim * lim))

    def add(self, x, r):
        if self.n == len(self._r):
            self._p = np.concatenate([self._p, np.zeros_like(self._p)])
            self._r = np.concatenate([self._r, np.zeros_like(self._r)])
        self.cells.setdefault(self.key(x), []).append(self.n)
        self._p[self.n] = x
        self._r[self.n] = r
        self.n += 1


def _rsa(st, radii, lo, hi, gap_fn, max_tries=2000000, wall=True):
    """Random sequential addition of spheres of the given radii with centres in
    the cube [lo, hi]^3 (whole spheres inside it when `wall`)."""
    n = len(radii)
    h = 2 * max(radii) + gap_fn(max(radii), max(radii)) + 1e-9
    g = _Grid(h)
    tries = 0
    batch = 4096
    buf = st.uniform(3 * batch).reshape(batch, 3)
    bi = 0
    for i in range(n):
        r = radii[i]
        while True:
            if bi == batch:
                buf = st.uniform(3 * batch).reshape(batch, 3)
                bi = 0
            x = lo + r + buf[bi] * (hi - lo - 2 * r) if wall else lo + buf[bi] * (hi - lo)
            bi += 1
            tries += 1
            if tries > max_tries:
                raise RuntimeError("random sequential addition did not converge")
            if g.ok(x, r, gap_fn):
                g.add(x, r)
                break
    return g.pos.copy()


def make_suspension(n, p, seed, vf=None, box=None, radius=1.0, min_gap=0.2, torque_std=1.0,
                    random_orientation=False):
    """Uniform non-overlapping spheres of one radius.

    Either `box` (edge of the cube [0, box]^3) or the volume fraction `vf` fixes
    the cube.  Surface gaps are at least `min_gap`."""
    st = Stream(seed)
    if box is None:
        box = (n * 4.0 / 3.0 * math.pi * radius ** 3 / vf) ** (1.0 / 3.0)
    radii = np.full(n, float(radius))
    pos = _rsa(st, radii, 0.0, box, lambda a, b: min_gap)
    quat = st.quaternions(n) if random_orientation else np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    force = np.zeros((n, 6))
    force[:, :3] = st.normal(3 * n).reshape(n, 3)
    if torque_std:
        force[:, 3:] = torque_std * st.normal(3 * n).reshape(n, 3)
    return dict(pos=pos, radius=radii, quat=quat, order=np.full(n, p, dtype=np.int32), force=force.ravel(),
                counts=np.array([n], dtype=np.int32), box=box)


def config1(seed=1):
    """64 spheres R=1 uniform in [0,20]^3 with surface gap >= 0.2, p=8, F~N(0,1), T=0."""
    c = make_suspension(64, 8, seed, box=20.0, min_gap=0.2, torque_std=0.0)
    c["name"] = "cfg1: 64 spheres p=8 (CPU reference case)"
    return c


def config2(seed=2):
    """1024 spheres R=1 at volume fraction 0.05, p=12, F,T~N(0,1)."""
    c = make_suspension(1024, 12, seed, vf=0.05, min_gap=0.2, torque_std=1.0)
    c["name"] = "cfg2: 1024 spheres p=12 vf=0.05"
    return c


def config3(seed=3):
    """4096 spheres R=1 at volume fraction 0.1, p=16, F~N(0,1), T~N(0, 0.5^2)."""
    c = make_suspension(4096, 16, seed, vf=0.1, min_gap=0.2, torque_std=0.5)
    c["name"] = "cfg3: 4096 spheres p=16 vf=0.1"
    return c


def config4(seed=4, n_susp=256):
    """Batched: n_susp suspensions, N_b = round(exp(U(ln 8, ln 512))), p_b from
    {6,8,10,12,16}, R ~ U[0.5,1.5], vf_b ~ U[0.05,0.15], F~N(0,1), T=0."""
    st = Stream(seed)
    ps = np.array([6, 8, 10, 12, 16], dtype=np.int32)
    pos, rad, order, counts, boxes = [], [], [], [], []
    for b in range(n_susp):
        n = int(round(math.exp(math.log(8) + st.uniform(1)[0] * (math.log(512) - math.log(8)))))
        p = int(ps[min(int(st.uniform(1)[0] * 5), 4)])
        vf = 0.05 + 0.10 * st.uniform(1)[0]
        r = 0.5 + st.uniform(n)
        box = (np.sum(4.0 / 3.0 * math.pi * r ** 3) / vf) ** (1.0 / 3.0)
        pb = _rsa(st, r, 0.0, box, lambda a, c: 0.1 * (a + c), wall=False)
        pos.append(pb)
        rad.append(r)
        order.append(np.full(n, p, dtype=np.int32))
        counts.append(n)
        boxes.append(box)
    n = int(sum(counts))
    force = np.zeros((n, 6))
    force[:, :3] = st.normal(3 * n).reshape(n, 3)
    return dict(pos=np.concatenate(pos), radius=np.concatenate(rad), quat=np.tile([1.0, 0.0, 0.0, 0.0], (n, 1)),
                order=np.concatenate(order), force=force.ravel(), counts=np.array(counts, dtype=np.int32),
                box=np.array(boxes), name="cfg4: %d batched suspensions" % n_susp)


def config5(seed=5, n=16384, p=12, vf=0.2, chain=8):
    """Dense near-field case: radii LogNormal(0, 0.2), volume fraction vf, chains
    of `chain` spheres placed by random walks whose bond gaps are
    U(0.02, 0.5) * min(R_i, R_j); other pairs keep a gap >= 0.01 min(R_i, R_j)."""
    st = Stream(seed)
    radii = np.exp(0.2 * st.normal(n))
    box = (np.sum(4.0 / 3.0 * math.pi * radii ** 3) / vf) ** (1.0 / 3.0)
    gap_fn = lambda a, b: 0.01 * np.minimum(a, b)  # noqa: E731
    g = _Grid(2 * radii.max() * 1.6)
    bonds = 0
    i = 0
    while i < n:
        # chain seed
        r = radii[i]
        while True:
            x = r + st.uniform(3) * (box - 2 * r)
            if g.ok(x, r, gap_fn):
                break
        g.add(x, r)
        prev = i
        i += 1
        for _ in range(chain - 1):
            if i >= n:
                break
            r = radii[i]
            placed = False
            for _t in range(64):
                d = st.unit_vectors(1)[0]
                gap = (0.02 + 0.48 * st.uniform(1)[0]) * min(r, radii[prev])
                x = g.pos[prev] + (radii[prev] + r + gap) * d
                if np.any(x < r) or np.any(x > box - r):
                    continue
                if g.ok(x, r, gap_fn):
                    placed = True
                    break
            if not placed:
                break
            g.add(x, r)
            bonds += 1
            prev = i
            i += 1
    pos = np.array(g.pos)
    force = st.normal(6 * n)
    return dict(pos=pos, radius=radii, quat=np.tile([1.0, 0.0, 0.0, 0.0], (n, 1)),
                order=np.full(n, p, dtype=np.int32), force=force, counts=np.array([n], dtype=np.int32), box=box,
                bonds=bonds, name="cfg5: %d spheres p=%d lognormal R vf=%.2f chains" % (n, p, vf))


CONFIGS = {1: config1, 2: config2, 3: config3, 4: config4, 5: config5}


def near_pair_stats(pos, radius):
    """Number of ordered near pairs (centre distance < 1.5 (R_s + R_t)) and the
    fraction of them with surface gap < 0.1 min(R_s, R_t)."""
    from scipy.spatial import cKDTree
    t = cKDTree(pos)
    rmax = radius.max()
    pairs = t.query_pairs(3.0 * rmax, output_type="ndarray")
    if len(pairs) == 0:
        return 0, 0.0
    d = np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1)
    rs, rt = radius[pairs[:, 0]], radius[pairs[:, 1]]
    near = d < 1.5 * (rs + rt)
    gap = d - rs - rt
    close = near & (gap < 0.1 * np.minimum(rs, rt))
    return 2 * int(near.sum()), float(close.sum()) / max(1, int(near.sum()))
