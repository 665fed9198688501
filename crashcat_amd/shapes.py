"""Host-side shape records and mass properties, mirroring crashcat's shape API.

Shapes are created on the host (as in crashcat, where shape construction is
setup, not stepping) and uploaded to the device as plain descriptors.

* box.create       src/shapes/box.ts:55-69,   mass src/body/mass-properties.ts:36-56
* sphere.create    src/shapes/sphere.ts:56-70, mass :99-124
* capsule.create   src/shapes/capsule.ts:51-74, mass :115-158
* convexHull.create src/shapes/convex-hull.ts:106-447 (COM-relative points, covariance
  inertia, per-vertex supporting planes, convex-radius reduction) and the shrunk
  support points of ConvexHullNoConvexSupport (:732-786).

The hull's polygon faces come from a convex hull of the input points (Qhull via
scipy) whose coplanar triangles are merged; crashcat's own quickhull builder
(src/shapes/utils/convex-hull-builder.ts) is host setup and is not restated.
"""
import math

import numpy as np

from . import _abi

DEFAULT_CONVEX_RADIUS = 0.05  # src/collision/support.ts:5
DEFAULT_SHAPE_DENSITY = 1000.0  # src/shapes/shapes.ts:121

SPHERE = _abi.CC_SHAPE_SPHERE
BOX = _abi.CC_SHAPE_BOX
CAPSULE = _abi.CC_SHAPE_CAPSULE
CONVEX_HULL = _abi.CC_SHAPE_CONVEX_HULL


class Shape:
    """A convex shape record (type + parameters + local bounds + mass data)."""

    def __init__(self, type_, **kw):
        self.type = type_
        self.radius = kw.get("radius", 0.0)
        self.half_height = kw.get("half_height", 0.0)
        self.half_extents = tuple(kw.get("half_extents", (0.0, 0.0, 0.0)))
        self.convex_radius = kw.get("convex_radius", 0.0)
        self.density = kw.get("density", DEFAULT_SHAPE_DENSITY)
        self.center_of_mass = tuple(kw.get("center_of_mass", (0.0, 0.0, 0.0)))
        self.aabb_min = tuple(kw["aabb_min"])
        self.aabb_max = tuple(kw["aabb_max"])
        self.mass = kw["mass"]
        self.inertia = np.asarray(kw["inertia"], dtype=np.float64).reshape(3, 3)
        # hull data (COM-relative)
        self.points = kw.get("points")
        self.shrunk_points = kw.get("shrunk_points")
        self.faces = kw.get("faces")  # list of vertex-index lists (CCW about outward normal)
        self.plane_normals = kw.get("plane_normals")
        self.plane_constants = kw.get("plane_constants")
        self.volume = kw.get("volume", 0.0)

    def key(self):
        """Hashable identity used to deduplicate identical shapes at upload."""
        if self.type == CONVEX_HULL:
            return (self.type, id(self))
        return (self.type, self.radius, self.half_height, self.half_extents, self.convex_radius, self.density)


# --------------------------------------------------------------------------- box
def box(half_extents, convex_radius=DEFAULT_CONVEX_RADIUS, density=DEFAULT_SHAPE_DENSITY):
    """box.create({halfExtents, convexRadius, density}) — src/shapes/box.ts:55-69."""
    hx, hy, hz = (float(v) for v in half_extents)
    size = (2 * hx, 2 * hy, 2 * hz)
    mass = size[0] * size[1] * size[2] * density  # setMassAndInertiaOfSolidBox
    sq = [s * s for s in size]
    inertia = np.diag([(sq[1] + sq[2]) * (mass / 12.0), (sq[0] + sq[2]) * (mass / 12.0), (sq[0] + sq[1]) * (mass / 12.0)])
    return Shape(BOX, half_extents=(hx, hy, hz), convex_radius=float(convex_radius), density=density,
                 aabb_min=(-hx, -hy, -hz), aabb_max=(hx, hy, hz), mass=mass, inertia=inertia,
                 volume=8 * hx * hy * hz)


# ------------------------------------------------------------------------ sphere
def sphere(radius, density=DEFAULT_SHAPE_DENSITY):
    """sphere.create({radius, density}) — src/shapes/sphere.ts:56-70, mass :99-124."""
    r = float(radius)
    r2 = r * r
    mass = (4.0 / 3.0) * math.pi * r * r2 * density
    i = (2.0 / 5.0) * mass * r2
    return Shape(SPHERE, radius=r, density=density, aabb_min=(-r, -r, -r), aabb_max=(r, r, r), mass=mass,
                 inertia=np.diag([i, i, i]), volume=(4 / 3) * math.pi * r * r * r)


# ----------------------------------------------------------------------- capsule
def capsule(radius, half_height_of_cylinder, density=DEFAULT_SHAPE_DENSITY):
    """capsule.create({radius, halfHeightOfCylinder}) — src/shapes/capsule.ts:51-74, mass :115-158."""
    r = float(radius)
    h = float(half_height_of_cylinder)
    if r <= 0 or h <= 0:
        raise ValueError("Invalid radius or height, must be > 0")
    volume = math.pi * r * r * (2 * h) + (4.0 / 3.0) * math.pi * r * r * r
    mass = volume * density
    rsq = r * r
    height = 2.0 * h
    cyl = math.pi * height * rsq * density
    hemi = ((2.0 * math.pi) / 3.0) * rsq * r * density
    iy = rsq * cyl * 0.5
    ixz = cyl * ((1.0 / 12.0) * height * height + (1.0 / 4.0) * rsq)
    hxz = hemi * ((2.0 / 5.0) * rsq + h * h + (3.0 / 8.0) * h * r)
    hy = hemi * (2.0 / 5.0) * rsq
    t = h + r
    return Shape(CAPSULE, radius=r, half_height=h, density=density, aabb_min=(-r, -t, -r), aabb_max=(r, t, r),
                 mass=mass, inertia=np.diag([ixz + hxz, iy + hy, ixz + hxz]), volume=volume)


# ------------------------------------------------------------------- convex hull
def _hull_faces(pts, tol):
    """Polygon faces (CCW about the outward normal) of the convex hull of pts."""
    from scipy.spatial import ConvexHull

    hull = ConvexHull(pts)
    tris = []
    for simplex, eq in zip(hull.simplices, hull.equations):
        a, b, c = (int(i) for i in simplex)
        n = np.cross(pts[b] - pts[a], pts[c] - pts[a])
        if np.dot(n, eq[:3]) < 0:
            b, c = c, b
        tris.append((a, b, c, eq[:3] / np.linalg.norm(eq[:3]), float(eq[3])))
    # merge coplanar neighbours (normals within tolerance and all vertices on the plane)
    n_t = len(tris)
    parent = list(range(n_t))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    edge_owner = {}
    for ti, (a, b, c, _, _) in enumerate(tris):
        for e in ((a, b), (b, c), (c, a)):
            edge_owner[e] = ti
    for ti, (a, b, c, n, d) in enumerate(tris):
        for (u, v) in ((a, b), (b, c), (c, a)):
            tj = edge_owner.get((v, u))
            if tj is None:
                continue
            nj = tris[tj][3]
            if np.dot(n, nj) > 1.0 - 1e-9:
                others = [tris[tj][k] for k in range(3)]
                if all(abs(np.dot(n, pts[o]) + d) < tol for o in others):
                    parent[find(ti)] = find(tj)
    groups = {}
    for ti in range(n_t):
        groups.setdefault(find(ti), []).append(ti)
    faces = []
    for g in sorted(groups.values(), key=lambda x: min(x)):
        if len(g) == 1:
            a, b, c = tris[g[0]][:3]
            faces.append([a, b, c])
            continue
        # boundary edges of the merged group, walked into a loop
        inner = set()
        edges = []
        for ti in g:
            a, b, c = tris[ti][:3]
            edges += [(a, b), (b, c), (c, a)]
        es = set(edges)
        bnd = [e for e in edges if (e[1], e[0]) not in es]
        nxt = {u: v for (u, v) in bnd}
        start = bnd[0][0]
        loop = [start]
        cur = nxt[start]
        while cur != start:
            loop.append(cur)
            cur = nxt[cur]
        faces.append(loop)
        del inner
    return faces


def convex_hull(positions, convex_radius=DEFAULT_CONVEX_RADIUS, density=DEFAULT_SHAPE_DENSITY,
                hull_tolerance=1e-3, max_error_convex_radius=0.05):
    """convexHull.create({positions, convexRadius}) — src/shapes/convex-hull.ts:106-447."""
    pts_in = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    faces_idx = _hull_faces(pts_in, hull_tolerance)
    # centroid / normal per face (triangle fan)
    fc = []
    for f in faces_idx:
        v = pts_in[f]
        nrm = np.zeros(3)
        for k in range(1, len(f) - 1):
            nrm += np.cross(v[k] - v[0], v[k + 1] - v[0])
        fc.append((v.mean(axis=0), nrm))
    # getCenterOfMassAndVolume (convex-hull-builder.ts:492-550)
    v4 = np.zeros(3)
    for c, _ in fc:
        v4 += c
    v4 = v4 / len(fc)
    vol = 0.0
    com = np.zeros(3)
    for f in faces_idx:
        v1 = pts_in[f[0]]
        v2 = pts_in[f[1]]
        for k in range(2, len(f)):
            v3 = pts_in[f[k]]
            vt = float(np.dot(v1 - v4, np.cross(v2 - v4, v3 - v4)))
            vol += vt
            com += vt * (v1 + v2 + v3 + v4)
            v2 = v3
    com = com / (4.0 * vol)
    volume = vol / 6.0
    # covariance inertia (convex-hull.ts:137-193)
    canon = np.array([[1 / 60, 1 / 120, 1 / 120], [1 / 120, 1 / 60, 1 / 120], [1 / 120, 1 / 120, 1 / 60]])
    cov = np.zeros((3, 3))
    for f in faces_idx:
        t1 = pts_in[f[0]] - com
        t2 = pts_in[f[1]] - com
        for k in range(2, len(f)):
            t3 = pts_in[f[k]] - com
            A = np.column_stack([t1, t2, t3])
            cov += np.linalg.det(A) * (A @ canon @ A.T)
            t2 = t3
    inertia = np.eye(3) * np.trace(cov) - cov
    # runtime representation: points renumbered in face order, relative to COM
    vmap = {}
    points = []
    faces = []
    for f in faces_idx:
        loop = []
        for oi in f:
            if oi not in vmap:
                vmap[oi] = len(points)
                points.append(pts_in[oi] - com)
            loop.append(vmap[oi])
        faces.append(loop)
    points = np.array(points)
    normals = []
    consts = []
    for (c, nrm) in fc:
        nn = nrm / np.linalg.norm(nrm)
        normals.append(nn)
        consts.append(-float(np.dot(nn, c - com)))
    normals = np.array(normals)
    consts = np.array(consts)
    # per-point supporting planes (convex-hull.ts:272-330)
    pfaces = []
    cos1 = math.cos(math.pi / 180.0)
    for p in range(len(points)):
        nb = [fi for fi, f in enumerate(faces) if p in f]
        best3 = None
        biggest = 0.05
        best2 = None
        smallest = cos1
        for i1 in range(len(nb)):
            n1 = normals[nb[i1]]
            for i2 in range(i1 + 1, len(nb)):
                n2 = normals[nb[i2]]
                cr = np.cross(n1, n2)
                d = float(np.dot(n1, n2))
                if d < smallest:
                    smallest = d
                    best2 = (nb[i1], nb[i2])
                for i3 in range(i2 + 1, len(nb)):
                    v = abs(float(np.dot(cr, normals[nb[i3]])))
                    if v > biggest:
                        biggest = v
                        best3 = (nb[i1], nb[i2], nb[i3])
        if best3 is not None:
            pfaces.append(best3)
        elif best2 is not None:
            pfaces.append(best2)
        else:
            pfaces.append((nb[0],))
    # convex radius reduction (convex-hull.ts:333-429)
    cr = float(convex_radius)
    if cr > 0:
        min_size = float("inf")
        for n, c in zip(normals, consts):
            md = 0.0
            for pt in points:
                dist = -(float(np.dot(n, pt)) + c)
                md = max(md, dist)
            min_size = min(min_size, md)
        cr = min(cr, 0.5 * min_size)
    if cr > 0:
        for pi, pf in enumerate(pfaces):
            if len(pf) == 1:
                continue
            n1 = normals[pf[0]]
            n2 = normals[pf[1]]
            if len(pf) == 3:
                n3 = normals[pf[2]]
                mask = np.array([1.0, 1.0, 1.0])
            else:
                n3 = np.cross(n1, n2)
                n3 = n3 / np.linalg.norm(n3)
                mask = np.array([1.0, 1.0, 0.0])
            N = np.array([n1, n2, n3])  # rows
            det = float(np.linalg.det(N))
            if abs(det) < 1e-12:
                cr = 0.0
                break
            # adjoint(N) * mask with a perspective divide by det, then divided by det again
            adj = np.linalg.inv(N) * det
            shift = (adj @ mask) / det / det
            off = float(np.linalg.norm(shift))
            if off > 1:
                cr = min(cr, max_error_convex_radius / (off - 1))
    # shrunk support points (convex-hull.ts:732-786)
    shrunk = np.zeros_like(points)
    for pi, pf in enumerate(pfaces):
        pt = points[pi]
        if len(pf) == 1:
            shrunk[pi] = pt - normals[pf[0]] * cr
            continue
        pl = []
        for k in range(2):
            n = normals[pf[k]]
            pl.append((n, -float(np.dot(n, pt)) - (-cr)))  # fromNormalAndPoint then offset(-r)
        if len(pf) == 3:
            n = normals[pf[2]]
            pl.append((n, -float(np.dot(n, pt)) + cr))
        else:
            n = np.cross(pl[0][0], pl[1][0])
            n = n / np.linalg.norm(n)
            pl.append((n, -float(np.dot(n, pt))))
        (c1, w1), (c2, w2), (c3, w3) = pl
        den = float(np.dot(c1, np.cross(c2, c3)))
        if den == 0:
            shrunk[pi] = pt - c1 * cr
        else:
            shrunk[pi] = (-w1 * np.cross(c2, c3) - w2 * np.cross(c3, c1) - w3 * np.cross(c1, c2)) / den
    mn = points.min(axis=0)
    mx = points.max(axis=0)
    return Shape(CONVEX_HULL, convex_radius=cr, density=density, center_of_mass=tuple(com),
                 aabb_min=tuple(mn), aabb_max=tuple(mx), mass=volume * density, inertia=inertia * density,
                 points=points, shrunk_points=shrunk, faces=faces, plane_normals=normals, plane_constants=consts,
                 volume=volume)


# ------------------------------------------------------------------ mass -> motion
def _jacobi(inertia):
    """decomposePrincipalMomentsOfInertia — src/body/motion-properties.ts:199-414."""
    a = np.array(inertia, dtype=np.float64).copy()
    rot = np.eye(3)
    b = np.array([a[0, 0], a[1, 1], a[2, 2]])
    d = b.copy()
    z = np.zeros(3)
    converged = False
    tiny = 5e-324
    for sweep in range(50):
        sm = abs(a[0, 1]) + abs(a[0, 2]) + abs(a[1, 2])
        avg = sm / 9.0
        if avg < tiny:
            converged = True
            break
        thresh = 0.2 * avg if sweep < 4 else tiny
        for ip in range(2):
            for iq in range(ip + 1, 3):
                apq = a[ip, iq]
                g = 100.0 * abs(apq)
                if sweep > 4 and abs(d[ip]) + g == abs(d[ip]) and abs(d[iq]) + g == abs(d[iq]):
                    a[ip, iq] = 0
                elif abs(apq) > thresh:
                    h = d[iq] - d[ip]
                    if abs(h) + g == abs(h):
                        t = apq / h
                    else:
                        theta = 0.5 * h / apq
                        t = 1.0 / (abs(theta) + math.sqrt(1.0 + theta * theta))
                        if theta < 0:
                            t = -t
                    c = 1.0 / math.sqrt(1.0 + t * t)
                    s = t * c
                    tau = s / (1.0 + c)
                    hh = t * apq
                    a[ip, iq] = 0
                    z[ip] -= hh
                    z[iq] += hh
                    d[ip] -= hh
                    d[iq] += hh

                    def rot_(m, i, j, k, l):
                        gv = m[i, j]
                        hv = m[k, l]
                        m[i, j] = gv - s * (hv + gv * tau)
                        m[k, l] = hv + s * (gv - hv * tau)

                    for j in range(0, ip):
                        rot_(a, j, ip, j, iq)
                    for j in range(ip + 1, iq):
                        rot_(a, ip, j, j, iq)
                    for j in range(iq + 1, 3):
                        rot_(a, ip, j, iq, j)
                    for j in range(3):
                        rot_(rot, j, ip, j, iq)
        b += z
        d = b.copy()
        z[:] = 0
    idx = [0, 1, 2]
    if d[idx[0]] < d[idx[1]]:
        idx[0], idx[1] = idx[1], idx[0]
    if d[idx[1]] < d[idx[2]]:
        idx[1], idx[2] = idx[2], idx[1]
    if d[idx[0]] < d[idx[1]]:
        idx[0], idx[1] = idx[1], idx[0]
    d2 = np.array([d[i] for i in idx])
    r2 = np.column_stack([rot[:, i] for i in idx])
    if np.dot(np.cross(r2[:, 0], r2[:, 1]), r2[:, 2]) < 0:
        r2[:, 2] = -r2[:, 2]
    return converged, r2, d2


def _quat_from_mat3(r):
    """quat.fromMat3 (gl-matrix) on a column-major 3x3."""
    m = [r[0, 0], r[1, 0], r[2, 0], r[0, 1], r[1, 1], r[2, 1], r[0, 2], r[1, 2], r[2, 2]]
    tr = m[0] + m[4] + m[8]
    out = [0.0, 0.0, 0.0, 0.0]
    if tr > 0.0:
        root = math.sqrt(tr + 1.0)
        out[3] = 0.5 * root
        root = 0.5 / root
        out[0] = (m[5] - m[7]) * root
        out[1] = (m[6] - m[2]) * root
        out[2] = (m[1] - m[3]) * root
    else:
        i = 0
        if m[4] > m[0]:
            i = 1
        if m[8] > m[i * 3 + i]:
            i = 2
        j = (i + 1) % 3
        k = (i + 2) % 3
        root = math.sqrt(m[i * 3 + i] - m[j * 3 + j] - m[k * 3 + k] + 1.0)
        out[i] = 0.5 * root
        root = 0.5 / root
        out[3] = (m[j * 3 + k] - m[k * 3 + j]) * root
        out[j] = (m[j * 3 + i] + m[i * 3 + j]) * root
        out[k] = (m[k * 3 + i] + m[i * 3 + k]) * root
    return tuple(out)


def motion_mass(mass, inertia, allowed_dofs=0b111111):
    """setMassProperties — src/body/motion-properties.ts:419-481.

    Returns (inv_mass, inv_inertia_diagonal, inertia_rotation quaternion)."""
    trans = allowed_dofs & 7
    rot_axes = (allowed_dofs >> 3) & 7
    inv_mass = 0.0 if trans == 0 else 1.0 / mass
    if rot_axes == 0:
        return inv_mass, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)
    ok, r, d = _jacobi(inertia)
    if ok:
        if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > 1e-20:
            inv = tuple((1.0 / x) if x != 0 else 0.0 for x in d)
            return inv_mass, inv, _quat_from_mat3(r)
    s = 2.5 * inv_mass
    return inv_mass, (s, s, s), (0.0, 0.0, 0.0, 1.0)
