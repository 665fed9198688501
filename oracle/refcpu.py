"""ctypes binding of the FP64 CPU oracle (librefcpu.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the results the GPU path is compared with or as the timed
CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from crashcat_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.environ.get("CC_ORACLE_LIB") or os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.oc_create.restype = C.c_void_p
        L.oc_create.argtypes = [C.POINTER(_abi.Settings), C.POINTER(_abi.BodyDesc), C.c_int32,
                                C.POINTER(_abi.ShapeDesc), C.c_int32, C.POINTER(_abi.HullData),
                                C.POINTER(_abi.PointConstraintDesc), C.c_int32, C.POINTER(_abi.MeshData)]
        L.oc_destroy.argtypes = [C.c_void_p]
        L.oc_step.argtypes = [C.c_void_p, C.c_double]
        L.oc_get_bodies.argtypes = [C.c_void_p, C.POINTER(_abi.BodyState)]
        L.oc_set_bodies.argtypes = [C.c_void_p, C.POINTER(_abi.BodyState)]
        L.oc_get_pairs.restype = C.c_int32
        L.oc_get_pairs.argtypes = [C.c_void_p, ip, C.c_int32]
        L.oc_get_contacts.restype = C.c_int32
        L.oc_get_contacts.argtypes = [C.c_void_p, C.POINTER(_abi.ContactOut), C.c_int32]
        L.oc_get_contact_cache.restype = C.c_int32
        L.oc_get_contact_cache.argtypes = [C.c_void_p, C.POINTER(_abi.CachedContact), C.c_int32, dp]
        L.oc_get_sleep_state.argtypes = [C.c_void_p, C.POINTER(_abi.BodySleep)]
        L.oc_get_point_impulses.argtypes = [C.c_void_p, dp]
        L.oc_get_removed_contacts.restype = C.c_int32
        L.oc_get_removed_contacts.argtypes = [C.c_void_p, C.POINTER(_abi.ContactRemoved), C.c_int32]
        L.oc_get_ccd_bodies.restype = C.c_int32
        L.oc_get_ccd_bodies.argtypes = [C.c_void_p, C.POINTER(_abi.CcdBody), C.c_int32]
        L.oc_get_counters.argtypes = [C.c_void_p, C.POINTER(_abi.StepCounters)]
        L.oc_add_forces.argtypes = [C.c_void_p, dp, dp, C.c_int32]
        L.oc_set_settings.argtypes = [C.c_void_p, C.POINTER(_abi.Settings)]
        L.oc_sort_key.restype = C.c_uint32
        L.oc_sort_key.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.oc_contact_tangents.argtypes = [dp, dp, dp]
        L.oc_island_iterations.restype = C.c_int32
        L.oc_island_iterations.argtypes = [C.c_void_p, C.c_int32, ip, ip]
        L.oc_broadphase_aabbs.restype = C.c_int32
        L.oc_broadphase_aabbs.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), ip, ip, ip, C.c_int32,
                                          C.c_double, C.POINTER(_abi.Settings), ip, C.c_int32]
        vp = C.c_void_p
        for fn in ("oc_collide", "oc_collide_convex"):
            getattr(L, fn).argtypes = [vp, vp, C.c_int32, dp, dp, C.c_int32, dp, dp, C.c_double, vp, vp]
        L.oc_collide_mesh.restype = C.c_int32
        L.oc_collide_mesh.argtypes = [vp, vp, vp, C.c_int32, dp, dp, C.c_int32, dp, dp, C.c_double, dp, vp, vp,
                                      C.POINTER(C.c_uint32), C.c_int32]
        L.oc_closest_on_triangle.restype = C.c_int32
        L.oc_closest_on_triangle.argtypes = [dp, dp, dp, dp, dp]
        L.oc_manifold_from_hit.restype = C.c_int32
        L.oc_manifold_from_hit.argtypes = [vp, C.c_double, dp, dp, dp]
        L.oc_clip_poly_vs_plane.restype = C.c_int32
        L.oc_clip_poly_vs_plane.argtypes = [dp, C.c_int32, dp, dp, dp]
        L.oc_clip_poly_vs_poly.restype = C.c_int32
        L.oc_clip_poly_vs_poly.argtypes = [dp, C.c_int32, dp, C.c_int32, dp, dp]
        L.oc_clip_poly_vs_edge.restype = C.c_int32
        L.oc_clip_poly_vs_edge.argtypes = [dp, C.c_int32, dp, dp, dp, dp]
        L.oc_manifold_between_two_faces.restype = C.c_int32
        L.oc_manifold_between_two_faces.argtypes = [dp, dp, dp, C.c_double, dp, C.c_int32, dp, C.c_int32, dp, dp]
        L.oc_prune.restype = C.c_int32
        L.oc_prune.argtypes = [dp, dp, C.c_int32, dp]
        L.oc_barycentric2.argtypes = [dp, dp, C.c_double, dp]
        L.oc_barycentric3.argtypes = [dp, dp, dp, C.c_double, dp]
        L.oc_inverse_inertia.argtypes = [dp, dp, dp, C.c_int32, dp]
        L.oc_supporting_face.restype = C.c_int32
        L.oc_supporting_face.argtypes = [vp, vp, C.c_int32, dp, dp, dp, dp]
        _LIB = L
    return _LIB


BODY_STATE_DTYPE = np.dtype(_abi.BodyState)
CONTACT_DTYPE = np.dtype(_abi.ContactOut)


class OracleWorld:
    """One FP64 oracle world built from a Scene (one world of it)."""

    def __init__(self, scene, world=0, bodies=None):
        """bodies: the world's body descriptors to use instead of the scene's (same count)."""
        L = lib()
        offs = scene.offsets()
        b0, b1 = offs[world], offs[world + 1]
        self.b0 = b0
        bodies = scene.bodies()[b0:b1].copy() if bodies is None else np.array(bodies, copy=True)
        assert len(bodies) == b1 - b0
        self._bodies = bodies
        shapes, hulls, self._keep = scene.c_shapes()
        if ((bodies["motion_quality"] == _abi.CC_MOTION_QUALITY_LINEAR_CAST)
                & (bodies["motion_type"] == _abi.CC_MOTION_DYNAMIC)).any():
            convex = (_abi.CC_SHAPE_SPHERE, _abi.CC_SHAPE_BOX, _abi.CC_SHAPE_CAPSULE, _abi.CC_SHAPE_CONVEX_HULL,
                      _abi.CC_SHAPE_CYLINDER)
            lc = (bodies["motion_quality"] == _abi.CC_MOTION_QUALITY_LINEAR_CAST) & (bodies["motion_type"] == _abi.CC_MOTION_DYNAMIC)
            for si in np.unique(bodies["shape"][lc]):  # the cast is restated for convex leaves
                if shapes[si].type not in convex + (_abi.CC_SHAPE_COMPOUND,):
                    raise ValueError("LINEAR_CAST bodies must be convex shapes (decorated or not) or compounds")
            for si in np.unique(bodies["shape"]):  # ... against convex leaves, planes and compounds of them
                if shapes[si].type not in convex + (_abi.CC_SHAPE_PLANE, _abi.CC_SHAPE_COMPOUND):
                    raise ValueError("LINEAR_CAST bodies need a world of convex shapes, planes and compounds")
        pcs = [(a - b0, b - b0, la, lb, ov, op) for (a, b, la, lb, ov, op) in scene.point_constraints if b0 <= a < b1]
        pa = (_abi.PointConstraintDesc * max(len(pcs), 1))()
        for i, (a, b, la, lb, ov, op) in enumerate(pcs):
            pa[i].body_a, pa[i].body_b, pa[i].enabled = a, b, 1
            pa[i].local_point_a[:] = la
            pa[i].local_point_b[:] = lb
            pa[i].num_velocity_steps_override = ov
            pa[i].num_position_steps_override = op
        self._pcs = pa
        self.npc = len(pcs)
        self._settings = scene.settings
        self.n = b1 - b0
        self.h = L.oc_create(C.byref(scene.settings), bodies.ctypes.data_as(C.POINTER(_abi.BodyDesc)), self.n,
                             shapes, self._keep["num_shapes"], C.byref(hulls), pa, len(pcs), C.byref(self._keep["meshes"]))

    def __del__(self):
        try:
            lib().oc_destroy(self.h)
        except Exception:
            pass

    def step(self, dt=1.0 / 60.0, n=1):
        for _ in range(n):
            lib().oc_step(self.h, dt)

    def bodies(self):
        out = np.zeros(self.n, dtype=BODY_STATE_DTYPE)
        lib().oc_get_bodies(self.h, out.ctypes.data_as(C.POINTER(_abi.BodyState)))
        return out

    def set_bodies(self, states):
        st = np.ascontiguousarray(states, dtype=BODY_STATE_DTYPE)
        lib().oc_set_bodies(self.h, st.ctypes.data_as(C.POINTER(_abi.BodyState)))

    def add_forces(self, force=None, torque=None, wake=True):
        """addForce / addTorque on every body: (n, 3) arrays, either may be None."""
        dp = C.POINTER(C.c_double)
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64).reshape(self.n, 3)
        t = None if torque is None else np.ascontiguousarray(torque, dtype=np.float64).reshape(self.n, 3)
        lib().oc_add_forces(self.h, None if f is None else f.ctypes.data_as(dp), None if t is None else t.ctypes.data_as(dp),
                            int(bool(wake)))

    def set_settings(self, settings):
        """Replace the world settings between steps (a copy is kept)."""
        self._settings = settings
        lib().oc_set_settings(self.h, C.byref(settings))

    def island_iterations(self, body):
        """(velocity, position) iteration counts of the body's island after the last step, or None."""
        v, p = C.c_int32(), C.c_int32()
        if not lib().oc_island_iterations(self.h, body, C.byref(v), C.byref(p)):
            return None
        return v.value, p.value

    def pairs(self):
        n = lib().oc_get_pairs(self.h, None, 0)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        lib().oc_get_pairs(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return out[:n]

    def contacts(self):
        n = lib().oc_get_contacts(self.h, None, 0)
        out = np.zeros(max(n, 1), dtype=CONTACT_DTYPE)
        lib().oc_get_contacts(self.h, out.ctypes.data_as(C.POINTER(_abi.ContactOut)), n)
        return out[:n]

    def contact_cache(self):
        """(records, previous time step): the cache the next step matches, sorted by key."""
        pdt = C.c_double()
        n = lib().oc_get_contact_cache(self.h, None, 0, C.byref(pdt))
        out = np.zeros(max(n, 1), dtype=np.dtype(_abi.CachedContact))
        lib().oc_get_contact_cache(self.h, out.ctypes.data_as(C.POINTER(_abi.CachedContact)), n, None)
        return out[:n], pdt.value

    def sleep_state(self):
        out = np.zeros(self.n, dtype=np.dtype(_abi.BodySleep))
        lib().oc_get_sleep_state(self.h, out.ctypes.data_as(C.POINTER(_abi.BodySleep)))
        return out

    def point_impulses(self):
        out = np.zeros((max(len(self._pcs), 1), 3))
        lib().oc_get_point_impulses(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:self.npc]

    def removed_contacts(self):
        n = lib().oc_get_removed_contacts(self.h, None, 0)
        out = np.zeros(max(n, 1), dtype=np.dtype(_abi.ContactRemoved))
        lib().oc_get_removed_contacts(self.h, out.ctypes.data_as(C.POINTER(_abi.ContactRemoved)), n)
        return out[:n]

    def ccd_bodies(self):
        n = lib().oc_get_ccd_bodies(self.h, None, 0)
        out = np.zeros(max(n, 1), dtype=np.dtype(_abi.CcdBody))
        lib().oc_get_ccd_bodies(self.h, out.ctypes.data_as(C.POINTER(_abi.CcdBody)), n)
        return out[:n]

    def counters(self):
        c = _abi.StepCounters()
        lib().oc_get_counters(self.h, C.byref(c))
        return {k: getattr(c, k) for k, _ in _abi.StepCounters._fields_}


def broadphase_aabbs(mn, mx, motion, sleeping, layer, margin, settings):
    L = lib()
    mn = np.ascontiguousarray(mn, dtype=np.float32)
    mx = np.ascontiguousarray(mx, dtype=np.float32)
    ip = C.POINTER(C.c_int32)
    mo = np.ascontiguousarray(motion, dtype=np.int32)
    sl = np.ascontiguousarray(sleeping, dtype=np.int32)
    la = np.ascontiguousarray(layer, dtype=np.int32)
    n = len(mo)
    fp = C.POINTER(C.c_float)
    cnt = L.oc_broadphase_aabbs(mn.ctypes.data_as(fp), mx.ctypes.data_as(fp), mo.ctypes.data_as(ip),
                                sl.ctypes.data_as(ip), la.ctypes.data_as(ip), n, margin, C.byref(settings), None, 0)
    out = np.zeros((max(cnt, 1), 2), dtype=np.int32)
    L.oc_broadphase_aabbs(mn.ctypes.data_as(fp), mx.ctypes.data_as(fp), mo.ctypes.data_as(ip),
                          sl.ctypes.data_as(ip), la.ctypes.data_as(ip), n, margin, C.byref(settings),
                          out.ctypes.data_as(ip), cnt)
    return out[:cnt]


# ---------------------------------------------------------------- unit entry points

class Hit(C.Structure):
    _fields_ = [
        ("hit", C.c_int32),
        ("status", C.c_int32),
        ("point_a", C.c_double * 3),
        ("point_b", C.c_double * 3),
        ("axis", C.c_double * 3),
        ("penetration", C.c_double),
        ("face_a_n", C.c_int32),
        ("face_b_n", C.c_int32),
        ("face_a", C.c_double * 96),
        ("face_b", C.c_double * 96),
    ]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vec(v, n=3):
    a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    assert a.size == n
    return a


def _poly(v):
    a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    return a, len(a)


def collide(scene, sa, pa, qa, sb, pb, qb, max_separation=0.0, convex_path=False):
    """collideShapeVsShape for two shapes of `scene` at world poses.  Returns a dict or None."""
    L = lib()
    shapes, hulls, keep = scene.c_shapes()
    h = Hit()
    fn = L.oc_collide_convex if convex_path else L.oc_collide
    pa, qa, pb, qb = _vec(pa), _vec(qa, 4), _vec(pb), _vec(qb, 4)
    fn(shapes, C.byref(hulls), sa, _dp(pa), _dp(qa), sb, _dp(pb), _dp(qb), C.c_double(max_separation),
       C.byref(scene.settings), C.byref(h))
    if not h.hit:
        return None
    return {
        "hit": h, "status": h.status, "point_a": np.array(h.point_a[:]), "point_b": np.array(h.point_b[:]),
        "axis": np.array(h.axis[:]), "penetration": h.penetration,
        "face_a": np.array(h.face_a[:3 * h.face_a_n]).reshape(-1, 3),
        "face_b": np.array(h.face_b[:3 * h.face_b_n]).reshape(-1, 3),
    }


def manifold_from_hit(hit, max_contact_distance):
    L = lib()
    nrm = np.zeros(3)
    oa = np.zeros((64, 3))
    ob = np.zeros((64, 3))
    n = L.oc_manifold_from_hit(C.byref(hit["hit"]), C.c_double(max_contact_distance), _dp(nrm), _dp(oa), _dp(ob))
    return nrm, oa[:n], ob[:n]


def clip_poly_vs_plane(poly, origin, normal):
    p, n = _poly(poly)
    o, nn = _vec(origin), _vec(normal)
    out = np.zeros((64, 3))
    k = lib().oc_clip_poly_vs_plane(_dp(p), n, _dp(o), _dp(nn), _dp(out))
    return out[:k]


def clip_poly_vs_poly(poly, clip, normal):
    p, n = _poly(poly)
    c, m = _poly(clip)
    nn = _vec(normal)
    out = np.zeros((64, 3))
    k = lib().oc_clip_poly_vs_poly(_dp(p), n, _dp(c), m, _dp(nn), _dp(out))
    return out[:k]


def clip_poly_vs_edge(poly, e1, e2, normal):
    p, n = _poly(poly)
    a, b, nn = _vec(e1), _vec(e2), _vec(normal)
    out = np.zeros((64, 3))
    k = lib().oc_clip_poly_vs_edge(_dp(p), n, _dp(a), _dp(b), _dp(nn), _dp(out))
    return out[:k]


def manifold_between_two_faces(p1, p2, axis, max_dist, face1, face2):
    a, b, ax = _vec(p1), _vec(p2), _vec(axis)
    f1, n1 = _poly(face1) if len(face1) else (np.zeros((1, 3)), 0)
    f2, n2 = _poly(face2) if len(face2) else (np.zeros((1, 3)), 0)
    oa = np.zeros((64, 3))
    ob = np.zeros((64, 3))
    k = lib().oc_manifold_between_two_faces(_dp(a), _dp(b), _dp(ax), C.c_double(max_dist), _dp(f1), n1, _dp(f2),
                                            n2, _dp(oa), _dp(ob))
    return oa[:k], ob[:k]


def prune(rel_a, rel_b, axis):
    a = np.ascontiguousarray(rel_a, dtype=np.float64).reshape(-1, 3).copy()
    b = np.ascontiguousarray(rel_b, dtype=np.float64).reshape(-1, 3).copy()
    assert len(a) > 4, "pruning needs more than four points"
    ax = _vec(axis)
    k = lib().oc_prune(_dp(a), _dp(b), len(a), _dp(ax))
    return a[:k], b[:k]


def barycentric2(a, b, tol=1e-12):
    r = np.zeros(4)
    A, B = _vec(a), _vec(b)
    lib().oc_barycentric2(_dp(A), _dp(B), C.c_double(tol), _dp(r))
    return r[0], r[1], bool(r[3])


def barycentric3(a, b, c, tol=1e-12):
    r = np.zeros(4)
    A, B, Cc = _vec(a), _vec(b), _vec(c)
    lib().oc_barycentric3(_dp(A), _dp(B), _dp(Cc), C.c_double(tol), _dp(r))
    return r[0], r[1], r[2], bool(r[3])


def inverse_inertia(inv_diag, inertia_rot, body_quat, dofs=0b111111):
    """World-space inverse inertia as a 3x3 matrix (row, column)."""
    d, iq, bq = _vec(inv_diag), _vec(inertia_rot, 4), _vec(body_quat, 4)
    out = np.zeros(9)
    lib().oc_inverse_inertia(_dp(d), _dp(iq), _dp(bq), dofs, _dp(out))
    return out.reshape(3, 3).T  # stored column-major


def supporting_face(scene, s, direction, pos=(0, 0, 0), quat=(0, 0, 0, 1)):
    shapes, hulls, keep = scene.c_shapes()
    d, p, q = _vec(direction), _vec(pos), _vec(quat, 4)
    out = np.zeros((64, 3))
    k = lib().oc_supporting_face(shapes, C.byref(hulls), s, _dp(d), _dp(p), _dp(q), _dp(out))
    return out[:k]


def closest_on_triangle(a, b, c):
    """getClosestPointOnTriangle: (point, squared distance, feature bits) of triangle abc to the origin."""
    p = _vec((0, 0, 0))
    d2 = C.c_double()
    v = [_vec(a), _vec(b), _vec(c)]
    f = lib().oc_closest_on_triangle(_dp(v[0]), _dp(v[1]), _dp(v[2]), _dp(p), C.byref(d2))
    return np.array(p[:]), d2.value, f


def collide_mesh(scene, sa, pa, qa, sm, pm=(0, 0, 0), qm=(0, 0, 0, 1), max_separation=0.0, move_dir=(0, 0, 0),
                 backfaces=True, active_edges=True, cap=256):
    """Hits of convex shape sa against mesh shape sm (collideConvexVsTriangleMesh) with the
    default collide settings (createDefaultCollideShapeSettings, collide-shape-vs-shape.ts:203-213)
    unless given: list of (Hit, sub-shape id of the triangle)."""
    L = lib()
    shapes, hulls, keep = scene.c_shapes()
    st = _abi.Settings()
    C.pointer(st)[0] = scene.settings
    st.collide_with_backfaces = 1 if backfaces else 0
    st.collide_only_with_active_edges = 1 if active_edges else 0
    out = (Hit * cap)()
    sub = (C.c_uint32 * cap)()
    v = [_vec(pa), _vec(qa, 4), _vec(pm), _vec(qm, 4), _vec(move_dir)]
    n = L.oc_collide_mesh(C.cast(shapes, C.c_void_p), C.cast(C.pointer(hulls), C.c_void_p),
                          C.cast(C.pointer(keep["meshes"]), C.c_void_p), sa, _dp(v[0]), _dp(v[1]), sm, _dp(v[2]),
                          _dp(v[3]), max_separation, _dp(v[4]), C.cast(C.pointer(st), C.c_void_p), out, sub, cap)
    return [(out[i], sub[i]) for i in range(min(n, cap))]


def sort_key(a, b, sub_a, sub_b):
    """computeContactSortKey as the oracle computes it."""
    return int(lib().oc_sort_key(a, b, sub_a, sub_b))


def contact_tangents(n):
    """The oracle's tangent basis (t1, t2) of a contact normal."""
    nn = _vec(n)
    t1 = np.zeros(3)
    t2 = np.zeros(3)
    lib().oc_contact_tangents(_dp(nn), _dp(t1), _dp(t2))
    return t1, t2
