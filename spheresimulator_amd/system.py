"""The steps on either side of the mobility solve in the reference's time loop:
the XYZ position files (SphereSystem::readXYZ / writeXYZ,
SRC/SphereSystem.cpp:129-179) and the explicit Euler step of positions and
orientations (Sphere::stepEuler, Sphere/Sphere.cpp:367-373).  Host-side;
orientations are unit quaternions (w, x, y, z) as at the C ABI."""
import numpy as np


def read_xyz(path):
    """Sphere ids, radii and centres of an XYZ file (SRC/SphereSystem.cpp:129-156):
    two header lines, then one row `type gid radius x y z` per sphere.
    Returns (gid int64[N], radius[N], pos[N, 3])."""
    gid, rad, pos = [], [], []
    with open(path) as f:
        f.readline()
        f.readline()
        for line in f:
            t = line.split()
            if not t:
                continue
            if len(t) < 6:
                raise ValueError("XYZ row needs `type gid radius x y z`: %r" % line)
            gid.append(int(t[1]))
            rad.append(float(t[2]))
            pos.append((float(t[3]), float(t[4]), float(t[5])))
    return np.array(gid, dtype=np.int64), np.array(rad), np.array(pos).reshape(-1, 3)


def write_xyz(path, gid, radius, pos, time=0.0):
    """Writes the XYZ format of SRC/SphereSystem.cpp:158-179: the sphere count,
    `SphereNumber: N, time: t`, then `S gid R x y z` rows with 14 significant
    digits (tab separated), so that read_xyz returns the same doubles."""
    gid = np.asarray(gid, dtype=np.int64).ravel()
    radius = np.asarray(radius, dtype=np.float64).ravel()
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    if not (len(gid) == len(radius) == len(pos)):
        raise ValueError("gid, radius and pos must describe the same spheres")
    with open(path, "w") as f:
        f.write("%d\n" % len(gid))
        f.write("SphereNumber: %d, time: %f\n" % (len(gid), time))
        for g, r, p in zip(gid, radius, pos):
            f.write("S\t%d\t%.14g\t%.14g\t%.14g\t%.14g\n" % (g, r, p[0], p[1], p[2]))


def _qmul(a, b):
    w1, x1, y1, z1 = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    w2, x2, y2, z2 = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=-1)


def step_euler(pos, quat, vel, dt):
    """Sphere::stepEuler (Sphere/Sphere.cpp:367-373): pos += V dt, and every
    orientation is turned about Omega by |Omega| dt (quaternion q <- dq q with
    dq = (cos(|Omega| dt / 2), sin(|Omega| dt / 2) Omega / |Omega|), renormalised).
    `vel` holds [V, Omega] per sphere (6N), as solveMob returns it.
    Returns new (pos, quat)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    quat = np.asarray(quat, dtype=np.float64).reshape(-1, 4)
    vel = np.asarray(vel, dtype=np.float64).reshape(-1, 6)
    if not (len(pos) == len(quat) == len(vel)):
        raise ValueError("pos, quat and vel must describe the same spheres")
    om = vel[:, 3:]
    w = np.linalg.norm(om, axis=1)
    half = 0.5 * w * dt
    axis = np.divide(om, w[:, None], out=np.zeros_like(om), where=w[:, None] > 0)
    dq = np.concatenate([np.cos(half)[:, None], np.sin(half)[:, None] * axis], axis=1)
    q = _qmul(dq, quat)
    q /= np.linalg.norm(q, axis=1)[:, None]
    return pos + vel[:, :3] * dt, q
