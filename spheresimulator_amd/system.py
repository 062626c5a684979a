"""The steps on either side of the mobility solve in the reference's time loop:
the XYZ position files (SphereSystem::readXYZ / writeXYZ,
SRC/SphereSystem.cpp:129-179), the explicit Euler step of positions and
orientations (Sphere::stepEuler, Sphere/Sphere.cpp:367-373) and the density
VTU files (Shexp::writeVTU) on the host; Brownian velocities
(SRC/SphereSystem.cpp:323-360) and collision pairs
(Sphere/CollisionSphere.hpp:51-90) on the GPU through the C ABI.
Orientations are unit quaternions (w, x, y, z) as at the C ABI."""
import base64
import ctypes
import struct

import numpy as np

from ._lib import LcpInfo, check, load


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


def grid_with_poles_cells(p):
    """Cells of one sphere's grid with poles (Shexp::getGridWithPoleCellConnect,
    Sphere/Shexp.cpp:273-380): point 0 is the north pole, then the (p+1) rings
    of 2p+2 grid points from north to south, then the south pole; triangles
    (VTK type 5) fan around the poles, quads (type 9) join neighbouring rings.
    Returns (connectivity, offsets (end of each cell), types)."""
    nphi = 2 * p + 2
    south = 2 * p * p + 4 * p + 3
    conn, off, typ = [], [], []

    def cell(nodes, t):
        conn.extend(nodes)
        off.append(len(conn))
        typ.append(t)

    index = 0
    for _ in range(2 * p + 1):
        index += 1
        cell([0, index, index + 1], 5)
    index += 1
    cell([0, nphi, 1], 5)
    for _ in range(1, p + 1):
        for _ in range(2 * p + 1):
            index += 1
            cell([index, index + 1, index + 1 - nphi, index - nphi], 9)
        index += 1
        cell([index, index - (2 * p + 1), index - (2 * p + 1) - nphi, index - nphi], 9)
    cell([south, south - 1, south - 1 - (2 * p + 1)], 5)
    for _ in range(2 * p + 1):
        index -= 1
        cell([south, index + 1, index], 5)
    return np.array(conn, dtype=np.int32), np.array(off, dtype=np.int32), np.array(typ, dtype=np.uint8)


def _qrot(q, v):
    w, u = q[0], np.asarray(q[1:])
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


_VTK_TYPES = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}


def _array(f, name, a, ncomp, vtype):
    """One inline binary DataArray of VTK XML: base64 of the byte count (UInt32,
    the default header of file version 0.1) followed by base64 of the
    little-endian values, the two encoded separately as VTK writes and reads
    them (the reference's IOHelper::writeDataArrayBase64 output)."""
    raw = np.ascontiguousarray(np.asarray(a).ravel(), dtype=_VTK_TYPES[vtype]).tobytes()
    f.write('<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="binary">\n' % (vtype, name, ncomp))
    f.write(base64.b64encode(struct.pack("<I", len(raw))).decode("ascii"))
    f.write(base64.b64encode(raw).decode("ascii"))
    f.write("\n</DataArray>\n")


def read_vtu_arrays(path):
    """The DataArrays of each Piece of a VTU file written by write_vtu (inline
    base64 binary, or ascii): a list with one {name: numpy array} per Piece."""
    import xml.etree.ElementTree as ET
    out = []
    for piece in ET.parse(path).getroot().find("UnstructuredGrid").findall("Piece"):
        arrs = {}
        for d in piece.iter("DataArray"):
            dt = np.dtype(_VTK_TYPES[d.get("type")])
            text = (d.text or "").strip()
            if d.get("format") == "ascii":
                arrs[d.get("Name")] = np.array(text.split(), dtype=dt)
                continue
            nbytes = struct.unpack("<I", base64.b64decode(text[:8]))[0]
            data = base64.b64decode(text[8:])
            if len(data) != nbytes:
                raise ValueError("DataArray %s: %d bytes, header says %d" % (d.get("Name"), len(data), nbytes))
            arrs[d.get("Name")] = np.frombuffer(data, dtype=dt).copy()
        out.append(arrs)
    return out


def write_vtu(path, op, pos, radius, quat, order, density, name="stkmob"):
    """Grid density of every sphere as a VTK unstructured grid, one Piece per
    sphere (SphereSystem::writeVTK -> Shexp::writeVTU, Sphere/Shexp.cpp:96-204):
    the points with poles (north pole, the grid, south pole; Shexp.cpp:206-271),
    the density (3 components, the grid values as given and the poles from the
    SH expansion, `op.pole_values`), the quadrature weights (0 at the poles) and
    the triangle/quad cells of grid_with_poles_cells.  `op` is the
    SphereSTKSHOperator of these spheres.  Data arrays are inline base64
    binary as the reference writes them (Shexp.cpp:182-195)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    radius = np.asarray(radius, dtype=np.float64).ravel()
    quat = np.asarray(quat, dtype=np.float64).reshape(-1, 4)
    order = np.asarray(order).ravel()
    density = np.asarray(density, dtype=np.float64).ravel()
    g = op.grid()
    poles = op.pole_values(density)
    coords = g["coords"].reshape(-1, 3)
    weights = g["weights"]
    dens = density.reshape(-1, 3)
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        f.write("<UnstructuredGrid>\n")
        for s in range(len(radius)):
            p = int(order[s])
            a, n = int(g["index"][s]), int(g["length"][s])
            ez = _qrot(quat[s], np.array([0.0, 0.0, 1.0]))
            pts = np.vstack([pos[s] + radius[s] * ez, coords[a:a + n], pos[s] - radius[s] * ez])
            val = np.vstack([poles[s, 0], dens[a:a + n], poles[s, 1]])
            w = np.concatenate([[0.0], weights[a:a + n], [0.0]])
            conn, off, typ = grid_with_poles_cells(p)
            f.write('<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(pts), len(typ)))
            f.write('<PointData Scalars="scalars">\n')
            _array(f, name, val, 3, "Float64")
            _array(f, "weights", w, 1, "Float64")
            f.write("</PointData>\n<Points>\n")
            _array(f, "points", pts, 3, "Float64")
            f.write("</Points>\n<Cells>\n")
            _array(f, "connectivity", conn, 1, "Int32")
            _array(f, "offsets", off, 1, "Int32")
            _array(f, "types", typ, 1, "UInt8")
            f.write("</Cells>\n</Piece>\n")
        f.write("</UnstructuredGrid>\n</VTKFile>\n")


def brownian_velocity(radius, dt, kBT, viscosity=1.0, scale=1.0, seed=0, step=0):
    """SphereSystem::getVelocityBrown (SRC/SphereSystem.cpp:323-360) on the GPU:
    [V, Omega] per sphere (6N) with V ~ sqrt(2 kBT/dt/(6 pi mu R)) scale N(0, I) and
    Omega ~ sqrt(2 kBT/dt/(8 pi mu R^3)) scale N(0, I).  The normals come from
    Philox4x32-10 with counter (sphere, block, step) and key seed, so (seed, step)
    fixes them."""
    radius = np.ascontiguousarray(np.asarray(radius, dtype=np.float64).ravel())
    vel = np.zeros(6 * radius.size)
    check(load().sphstk_brownian_velocity(radius.size, radius.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          viscosity, kBT, dt, scale, seed, step,
                                          vel.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return vel


def collision_pairs(pos, radius_col, colbuf, gid=None):
    """Collision blocks of SphereSystem::resolveCollision / CollisionSphere::collide
    (SRC/SphereSystem.cpp:414-431, Sphere/CollisionSphere.hpp:51-90) on the GPU:
    pairs (I, J) with gid[I] < gid[J] and sep = |x_J - x_I| - Rc_I - Rc_J < colbuf Rc_I,
    ordered by I then J.  Returns a dict of idx_i, idx_j, phi0 (= sep), norm_i
    (= -(x_J - x_I)/|x_J - x_I|), norm_j = -norm_i, gamma (max(-sep, 0)) and the
    stress |x_J - x_I| norm_i norm_i^T per pair."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    rc = np.ascontiguousarray(np.asarray(radius_col, dtype=np.float64).ravel())
    n = len(rc)
    if len(pos) != n:
        raise ValueError("pos and radius_col must describe the same spheres")
    g = None if gid is None else np.ascontiguousarray(np.asarray(gid, dtype=np.int32).ravel())
    if g is not None and len(g) != n:
        raise ValueError("gid must have one entry per sphere")
    L = load()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    gp = g.ctypes.data_as(ip) if g is not None else None
    cap = 6 * n
    while True:
        ii = np.zeros(cap, dtype=np.int32)
        jj = np.zeros(cap, dtype=np.int32)
        phi0 = np.zeros(cap)
        nrm = np.zeros(3 * cap)
        cnt = ctypes.c_int64()
        check(L.sphstk_collision_pairs(n, pos.ctypes.data_as(dp), rc.ctypes.data_as(dp), gp, colbuf, cap,
                                       ctypes.byref(cnt), ii.ctypes.data_as(ip), jj.ctypes.data_as(ip),
                                       phi0.ctypes.data_as(dp), nrm.ctypes.data_as(dp)))
        if cnt.value <= cap:
            break
        cap = cnt.value
    m = cnt.value
    ii, jj, phi0, nrm = ii[:m], jj[:m], phi0[:m], nrm[:3 * m].reshape(-1, 3)
    r = np.linalg.norm(pos[jj] - pos[ii], axis=1)
    return dict(idx_i=ii, idx_j=jj, phi0=phi0, norm_i=nrm, norm_j=-nrm, gamma=np.maximum(-phi0, 0.0),
                stress=r[:, None, None] * nrm[:, :, None] * nrm[:, None, :])


def collision_lcp(pairs, vel_known, dt, mob=None, radius=None, viscosity=1.0, buffer=0.0, tol=1e-5, max_iters=2000,
                  gamma0=None):
    """The collision LCP of SphereSystem::resolveCollision (SRC/SphereSystem.cpp:445-474)
    by Barzilai-Borwein projected gradient descent in the library: pairs as returned
    by collision_pairs, vel_known the 6N velocities from the known forces (plus the
    Brownian term), mob a SphereSTKMobMat (one full mobility solve per iteration) or
    None for the free-drain mobility of `radius` and `viscosity`.  Returns
    (gamma, force_col, vel_col, info); the sphere velocities are vel_col + vel_known."""
    ii = np.ascontiguousarray(pairs["idx_i"], dtype=np.int32)
    jj = np.ascontiguousarray(pairs["idx_j"], dtype=np.int32)
    phi0 = np.ascontiguousarray(pairs["phi0"], dtype=np.float64)
    nrm = np.ascontiguousarray(np.asarray(pairs["norm_i"], dtype=np.float64).reshape(-1, 3))
    vk = np.ascontiguousarray(np.asarray(vel_known, dtype=np.float64).ravel())
    n = vk.size // 6
    if vk.size != 6 * n:
        raise ValueError("vel_known must hold 6 values per sphere")
    m = len(ii)
    g = np.zeros(m) if gamma0 is None else np.ascontiguousarray(np.asarray(gamma0, dtype=np.float64).ravel())
    if g.size != m:
        raise ValueError("gamma0 must have one entry per pair")
    g = g.copy()
    r = None if radius is None else np.ascontiguousarray(np.asarray(radius, dtype=np.float64).ravel())
    F = np.zeros(6 * n)
    U = np.zeros(6 * n)
    info = LcpInfo()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    check(load().sphstk_collision_lcp(mob._h if mob is not None else None, n,
                                      r.ctypes.data_as(dp) if r is not None else None, viscosity, m,
                                      ii.ctypes.data_as(ip), jj.ctypes.data_as(ip), phi0.ctypes.data_as(dp),
                                      nrm.ctypes.data_as(dp), vk.ctypes.data_as(dp), dt, buffer, tol, max_iters,
                                      g.ctypes.data_as(dp), F.ctypes.data_as(dp), U.ctypes.data_as(dp),
                                      ctypes.byref(info)))
    return g, F, U, info.as_dict()


def resolve_collision(pos, radius_col, force_known, dt, colbuf, mob=None, radius=None, viscosity=1.0, kBT=0.0,
                      scale_brown=0.0, seed=0, step=0, buffer=0.0, tol=1e-5, max_iters=2000, gid=None):
    """One SphereSystem::resolveCollision (SRC/SphereSystem.cpp:403-478): the known
    velocity M F_known (a mobility solve of `mob`, or the free drain of `radius`
    and `viscosity` when mob is None) plus the Brownian term when scale_brown > 0
    (getVelocityKnown, :280-292), the collision pairs, the LCP, and the sphere
    velocities vel_col + vel_known (:472-476).  Returns (vel 6N, pairs, gamma, force_col, info)."""
    F = np.asarray(force_known, dtype=np.float64).ravel()
    rc = np.asarray(radius_col, dtype=np.float64).ravel()
    if F.size != 6 * len(rc):
        raise ValueError("force_known must hold 6 values per sphere")
    if mob is not None:
        vk, _ = mob.solveMob(F)
        rh = rc if radius is None else np.asarray(radius, dtype=np.float64).ravel()
    else:
        if radius is None:
            raise ValueError("radius is needed for the free-drain mobility")
        rh = np.asarray(radius, dtype=np.float64).ravel()
        d = np.repeat(np.stack([1 / (6 * np.pi * rh * viscosity), 1 / (8 * np.pi * rh ** 3 * viscosity)], 1), 3,
                      axis=1).ravel()
        vk = d * F
    if scale_brown > 0:
        vk = vk + brownian_velocity(rh, dt, kBT, viscosity=viscosity, scale=scale_brown, seed=seed, step=step)
    pairs = collision_pairs(pos, rc, colbuf, gid=gid)
    gamma, fcol, vcol, info = collision_lcp(pairs, vk, dt, mob=mob, radius=rh, viscosity=viscosity, buffer=buffer,
                                            tol=tol, max_iters=max_iters, gamma0=pairs["gamma"])
    return vcol + vk, pairs, gamma, fcol, info

