"""Scene description shared by the device path and its checkers, plus the seeded
scene generators of the benchmark configurations.

A Scene is the plain-data form of one or more crashcat worlds (bodies, shapes,
hull data, point constraints, settings) laid out exactly as the C ABI expects
(include/crashcat_amd.h).  Generators use splitmix64-seeded xoshiro256** so the
same seed gives the same initial state on every machine.
"""
import ctypes as C
import math

import numpy as np

from . import _abi
from . import shapes as S


# ---------------------------------------------------------------- settings / layers
def create_world_settings():
    """createWorldSettings — src/world-settings.ts:224-264 (defaults)."""
    s = _abi.Settings()
    s.gravity[:] = (0.0, -9.81, 0.0)
    s.speculative_contact_distance = 0.02
    s.manifold_tolerance = 1e-3
    s.collision_tolerance = 1e-4
    s.penetration_tolerance = 1e-4
    s.normal_cos_max_delta_rotation = math.cos(5 * math.pi / 180)
    s.contact_point_preserve_lambda_max_dist_sq = 0.01 * 0.01
    s.min_velocity_for_restitution = 1.0
    s.penetration_slop = 0.02
    s.baumgarte_factor = 0.2
    s.max_penetration_distance = 0.2
    s.time_before_sleep = 0.5
    s.point_velocity_sleep_threshold = 0.03
    s.velocity_iterations = 10
    s.position_iterations = 2
    s.gravity_enabled = 1
    s.warm_starting = 1
    s.allow_sleeping = 1
    s.num_object_layers = 0
    s.collide_with_backfaces = 0
    s.collide_only_with_active_edges = 1
    s.ccd_linear_cast_threshold = 0.05
    s.ccd_linear_cast_max_penetration = 0.25
    s.ccd_min_velocity_for_restitution = 1.0
    return s


def add_object_layer(settings):
    """addObjectLayer — src/layers.ts (broadphase layers follow the object-layer pair table)."""
    i = settings.num_object_layers
    if i >= _abi.CC_MAX_OBJECT_LAYERS:
        raise ValueError("too many object layers")
    settings.num_object_layers = i + 1
    return i


def enable_collision(settings, a, b):
    """enableCollision — src/layers.ts:88-112."""
    settings.object_layer_pairs[a * _abi.CC_MAX_OBJECT_LAYERS + b] = 1
    settings.object_layer_pairs[b * _abi.CC_MAX_OBJECT_LAYERS + a] = 1


def test_world_settings():
    """The layer setup of the reference's test helper (tst/helpers.ts:3-26):
    MOVING collides with MOVING and NOT_MOVING."""
    s = create_world_settings()
    moving = add_object_layer(s)
    not_moving = add_object_layer(s)
    enable_collision(s, moving, moving)
    enable_collision(s, moving, not_moving)
    return s, moving, not_moving


BODY_DTYPE = np.dtype(_abi.BodyDesc)


def _quat_yaw(theta):
    return (0.0, math.sin(theta / 2), 0.0, math.cos(theta / 2))


class Scene:
    """Bodies of one or more independent worlds sharing a shape table and settings."""

    def __init__(self, settings=None):
        if settings is None:
            settings, self.layer_moving, self.layer_static = test_world_settings()
        else:
            self.layer_moving, self.layer_static = 0, 1
        self.settings = settings
        self.shapes = []
        self._shape_index = {}
        self.rows = []  # list of dicts, converted lazily
        self.world_offsets = [0]
        self.point_constraints = []
        self._bodies = None

    # -------------------------------------------------------------- building
    def add_shape(self, shape):
        k = shape.key()
        if k in self._shape_index:
            return self._shape_index[k]
        self._shape_index[k] = len(self.shapes)
        self.shapes.append(shape)
        return len(self.shapes) - 1

    def add_body(self, shape, position, quaternion=(0.0, 0.0, 0.0, 1.0), motion_type=_abi.CC_MOTION_DYNAMIC,
                 linear_velocity=(0.0, 0.0, 0.0), angular_velocity=(0.0, 0.0, 0.0), object_layer=None,
                 friction=0.2, restitution=0.0, gravity_factor=1.0, linear_damping=0.05, angular_damping=0.05,
                 max_linear_velocity=500.0, max_angular_velocity=0.25 * math.pi * 60.0, allow_sleeping=True,
                 allowed_dofs=0b111111, friction_combine=4, restitution_combine=3, collision_groups=0xFFFFFFFF,
                 collision_mask=0xFFFFFFFF, force=(0.0, 0.0, 0.0), torque=(0.0, 0.0, 0.0),
                 num_velocity_steps_override=0, num_position_steps_override=0, sensor=False, motion_quality=0, sleeping=False):
        """rigidBody.create — src/body/rigid-body.ts:390-434, defaults :120-139; force / torque are
        the accumulators addForce / addTorque fill before the first step; `sleeping` starts a
        dynamic body asleep (rigidBody.sleep, src/body/sleep.ts:205-222)."""
        si = self.add_shape(shape)
        if object_layer is None:
            object_layer = self.layer_static if motion_type == _abi.CC_MOTION_STATIC else self.layer_moving
        if motion_type == _abi.CC_MOTION_STATIC:
            inv_mass, inv_i, irot = 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)
        else:
            inv_mass, inv_i, irot = S.motion_mass(shape.mass, shape.inertia, allowed_dofs)
        self.rows.append((si, motion_type, object_layer, collision_groups, collision_mask, int(allow_sleeping),
                          allowed_dofs, friction_combine, restitution_combine,
                          int(bool(sleeping) and motion_type != _abi.CC_MOTION_STATIC), tuple(position), tuple(quaternion),
                          tuple(linear_velocity), tuple(angular_velocity), inv_mass, inv_i, irot, friction,
                          restitution, gravity_factor, linear_damping, angular_damping, max_linear_velocity,
                          max_angular_velocity, tuple(force), tuple(torque), int(num_velocity_steps_override),
                          int(num_position_steps_override), int(bool(sensor)), int(motion_quality)))
        self._bodies = None
        return len(self.rows) - 1

    def end_world(self):
        """Close the current world; following bodies belong to a new world."""
        if self.world_offsets[-1] != len(self.rows):
            self.world_offsets.append(len(self.rows))

    def add_point_constraint(self, a, b, local_a, local_b, num_velocity_steps_override=0,
                             num_position_steps_override=0):
        self.point_constraints.append((a, b, tuple(local_a), tuple(local_b), int(num_velocity_steps_override),
                                       int(num_position_steps_override)))

    # -------------------------------------------------------------- C views
    @property
    def num_bodies(self):
        return len(self.rows)

    @property
    def num_worlds(self):
        offs = self.offsets()
        return len(offs) - 1

    def offsets(self):
        offs = list(self.world_offsets)
        if offs[-1] != len(self.rows):
            offs.append(len(self.rows))
        return offs

    def bodies(self):
        """numpy structured array with the layout of cc_body_desc."""
        if self._bodies is None:
            arr = np.zeros(len(self.rows), dtype=BODY_DTYPE)
            if self.rows:
                cols = list(zip(*self.rows))
                names = ["shape", "motion_type", "object_layer", "collision_groups", "collision_mask",
                         "allow_sleeping", "allowed_dofs", "friction_combine", "restitution_combine", "sleeping",
                         "position", "quaternion", "linear_velocity", "angular_velocity", "inv_mass",
                         "inv_inertia_diagonal", "inertia_rotation", "friction", "restitution", "gravity_factor",
                         "linear_damping", "angular_damping", "max_linear_velocity", "max_angular_velocity",
                         "force", "torque", "num_velocity_steps_override", "num_position_steps_override", "sensor",
                         "motion_quality"]
                for n, c in zip(names, cols):
                    arr[n] = np.asarray(c)
            self._bodies = arr
        return self._bodies

    def world(self, w):
        """World w as a one-world scene (same shapes and settings, constraints renumbered)."""
        offs = self.offsets()
        b0, b1 = offs[w], offs[w + 1]
        out = Scene(self.settings)
        out.layer_moving, out.layer_static = self.layer_moving, self.layer_static
        out.shapes, out._shape_index = self.shapes, self._shape_index
        out._bodies = self.bodies()[b0:b1].copy()
        out.rows = [None] * (b1 - b0)
        out.world_offsets = [0]
        out.point_constraints = [(a - b0, b - b0, la, lb, ov, op) for (a, b, la, lb, ov, op) in self.point_constraints
                                 if b0 <= a < b1]
        return out

    def c_bodies(self):
        b = self.bodies()
        return b.ctypes.data_as(C.POINTER(_abi.BodyDesc)), b

    def c_shapes(self):
        """(shape array, HullData, keepalive) in the C layout; built once per shape list (shapes
        are only ever appended)."""
        key = (id(self.shapes), len(self.shapes))
        cached = getattr(self, "_cshapes", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        # a compound's children follow the scene's shapes, each a leaf desc whose local transform
        # is the child's position and quaternion
        descs = list(self.shapes)
        child_range = {}
        for i, s in enumerate(self.shapes):
            if s.type == S.COMPOUND:
                child_range[i] = (len(descs), len(s.children))
                descs.extend((pos, q, ch) for pos, q, ch in s.children)
        n = len(descs)
        arr = (_abi.ShapeDesc * max(n, 1))()
        pts, shr, ffv, fnv, fvi, pn, pc = [], [], [], [], [], [], []
        np_, nf, nfv = 0, 0, 0
        mpos, mtri, mnode = [], [], []
        nmv, nmt, nmn = 0, 0, 0
        for i, s in enumerate(descs):
            d = arr[i]
            if isinstance(s, tuple):  # a compound's child
                pos, q, s = s
                d.local_position[:] = pos
                d.local_quaternion[:] = q
                d.decorated = s.decorated | _abi.CC_DECO_TRANSFORMED
            else:
                d.local_position[:] = s.local_position
                d.local_quaternion[:] = s.local_quaternion
                d.decorated = s.decorated
            if i in child_range:
                d.compound_first_child, d.compound_num_children = child_range[i]
            d.type = s.type
            d.radius = s.radius
            d.half_height = s.half_height
            d.half_extents[:] = s.half_extents
            d.convex_radius = s.convex_radius
            d.center_of_mass[:] = s.center_of_mass
            d.aabb_min[:] = s.aabb_min
            d.aabb_max[:] = s.aabb_max
            d.scale[:] = s.scale
            if s.type == S.CONVEX_HULL:
                d.hull_first_point = np_
                d.hull_num_points = len(s.points)
                d.hull_first_face = nf
                d.hull_num_faces = len(s.faces)
                pts.append(np.asarray(s.points, dtype=np.float64))
                shr.append(np.asarray(s.shrunk_points, dtype=np.float64))
                for f, nrm, c in zip(s.faces, s.plane_normals, s.plane_constants):
                    ffv.append(nfv)
                    fnv.append(len(f))
                    fvi.extend(f)
                    nfv += len(f)
                    pn.append(nrm)
                    pc.append(c)
                np_ += len(s.points)
                nf += len(s.faces)
            if s.type == S.TRIANGLE_MESH:
                d.mesh_first_triangle = nmt
                d.mesh_num_triangles = len(s.mesh_triangles)
                d.mesh_first_node = nmn
                d.mesh_num_nodes = len(s.mesh_nodes)
                t = np.array(s.mesh_triangles, dtype=np.int32)
                t[:, :3] += nmv
                mpos.append(np.asarray(s.mesh_positions, dtype=np.float64))
                mtri.append(t)
                mnode.append(np.asarray(s.mesh_nodes, dtype=np.float64))
                nmv += len(s.mesh_positions)
                nmt += len(t)
                nmn += len(s.mesh_nodes)
        keep = {}
        keep["pts"] = np.ascontiguousarray(np.concatenate(pts) if pts else np.zeros((1, 3))).astype(np.float64)
        keep["shr"] = np.ascontiguousarray(np.concatenate(shr) if shr else np.zeros((1, 3))).astype(np.float64)
        keep["ffv"] = np.asarray(ffv if ffv else [0], dtype=np.int32)
        keep["fnv"] = np.asarray(fnv if fnv else [0], dtype=np.int32)
        keep["fvi"] = np.asarray(fvi if fvi else [0], dtype=np.int32)
        keep["pn"] = np.ascontiguousarray(np.asarray(pn if pn else np.zeros((1, 3)), dtype=np.float64))
        keep["pc"] = np.asarray(pc if pc else [0.0], dtype=np.float64)
        h = _abi.HullData()
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        h.points = keep["pts"].ctypes.data_as(dp)
        h.shrunk_points = keep["shr"].ctypes.data_as(dp)
        h.face_first_vertex = keep["ffv"].ctypes.data_as(ip)
        h.face_num_vertices = keep["fnv"].ctypes.data_as(ip)
        h.face_vertex_indices = keep["fvi"].ctypes.data_as(ip)
        h.plane_normals = keep["pn"].ctypes.data_as(dp)
        h.plane_constants = keep["pc"].ctypes.data_as(dp)
        h.total_points = np_
        h.total_faces = nf
        h.total_face_vertices = nfv
        keep["mpos"] = np.ascontiguousarray(np.concatenate(mpos) if mpos else np.zeros((1, 3)), dtype=np.float64)
        keep["mtri"] = np.ascontiguousarray(np.concatenate(mtri) if mtri else np.zeros((1, 4)), dtype=np.int32)
        keep["mnode"] = np.ascontiguousarray(np.concatenate(mnode) if mnode else np.zeros((1, 8)), dtype=np.float64)
        m = _abi.MeshData()
        m.positions = keep["mpos"].ctypes.data_as(dp)
        m.triangles = keep["mtri"].ctypes.data_as(ip)
        m.nodes = keep["mnode"].ctypes.data_as(dp)
        m.total_vertices = nmv
        m.total_triangles = nmt
        m.total_nodes = nmn
        keep["meshes"] = m
        keep["arr"] = arr
        keep["num_shapes"] = n  # the scene's shapes and the compounds' children after them
        self._cshapes = (key, (arr, h, keep))
        return arr, h, keep

    def c_point_constraints(self):
        n = len(self.point_constraints)
        arr = (_abi.PointConstraintDesc * max(n, 1))()
        for i, (a, b, la, lb, ov, op) in enumerate(self.point_constraints):
            arr[i].body_a = a
            arr[i].body_b = b
            arr[i].enabled = 1
            arr[i].local_point_a[:] = la
            arr[i].local_point_b[:] = lb
            arr[i].num_velocity_steps_override = ov
            arr[i].num_position_steps_override = op
        return arr, n

    def c_offsets(self):
        offs = np.asarray(self.offsets(), dtype=np.int32)
        return offs

    def desc(self):
        """The scene as the JavaScript host's world description (crashcat_amd/node/index.js):
        plain numbers and lists keyed by the C ABI's field names, JSON-serialisable."""
        def struct_dict(st):
            out = {}
            for name, ctype in st._fields_:
                if name.endswith("_"):
                    continue
                v = getattr(st, name)
                out[name] = list(v) if hasattr(v, "__len__") else v
            return out
        arr, hull, keep = self.c_shapes()
        shapes = [struct_dict(arr[i]) for i in range(keep["num_shapes"])]
        bodies = []
        b = self.bodies()
        for row in b:
            d = {}
            for name in b.dtype.names:
                if name.endswith("_"):
                    continue
                v = row[name]
                d[name] = v.tolist() if hasattr(v, "tolist") else v
            bodies.append(d)
        pcs, npc = self.c_point_constraints()
        return {
            "settings": struct_dict(self.settings),
            "worldBodyOffsets": [int(x) for x in self.c_offsets()],
            "shapes": shapes,
            "hull": {
                "points": keep["pts"].reshape(-1)[: 3 * hull.total_points].tolist(),
                "shrunkPoints": keep["shr"].reshape(-1)[: 3 * hull.total_points].tolist(),
                "faceFirstVertex": keep["ffv"][: hull.total_faces].tolist(),
                "faceNumVertices": keep["fnv"][: hull.total_faces].tolist(),
                "faceVertexIndices": keep["fvi"][: hull.total_face_vertices].tolist(),
                "planeNormals": keep["pn"].reshape(-1)[: 3 * hull.total_faces].tolist(),
                "planeConstants": keep["pc"][: hull.total_faces].tolist(),
            },
            "mesh": {
                "positions": keep["mpos"].reshape(-1)[: 3 * keep["meshes"].total_vertices].tolist(),
                "triangles": keep["mtri"].reshape(-1)[: 4 * keep["meshes"].total_triangles].tolist(),
                "nodes": keep["mnode"].reshape(-1)[: 8 * keep["meshes"].total_nodes].tolist(),
            },
            "bodies": bodies,
            "pointConstraints": [struct_dict(pcs[i]) for i in range(npc)],
        }


    def crashcat_world(self):
        """The scene (one world) as crashcat's own objects, JSON-serialisable: a World's settings
        (src/world-settings.ts:4-200), bodies.pool of RigidBody records (src/body/rigid-body.ts:
        142-252) with motionProperties, shapes in crashcat's records (sphere.ts:41-55, box.ts:
        43-52, capsule.ts:32-49, convex-hull.ts:38-86, cylinder.ts:30-40, plane.ts:59-74,
        triangle-mesh.ts:105-112) and a point-constraint pool (point-constraint.ts:21-25).  A
        body's `shape` is an index into the returned `shapes` list (JSON has no shared objects;
        a loader puts the shape objects in place).  Bodies were created in index order, so body
        k has sequence k and id k + k * 2^32 (src/body/body-id.ts)."""
        assert self.num_worlds == 1, "one world"
        st = self.settings
        n = st.num_object_layers
        pairs = [int(st.object_layer_pairs[a * 8 + b]) for a in range(n) for b in range(n)]
        settings = {
            "layers": {"broadphaseLayers": 1, "objectLayers": n, "objectLayerToBroadphaseLayer": [0] * n,
                       "objectLayerPairs": pairs, "objectVsBroadphase": [1] * n, "broadphasePairs": [1]},
            "gravity": list(st.gravity), "gravityEnabled": bool(st.gravity_enabled),
            "contacts": {"contactPointPreserveLambdaMaxDistSq": st.contact_point_preserve_lambda_max_dist_sq},
            "narrowphase": {"collisionTolerance": st.collision_tolerance, "penetrationTolerance": st.penetration_tolerance,
                            "returnDeepestPoint": False,
                            "collideOnlyWithActiveEdges": bool(st.collide_only_with_active_edges),
                            "speculativeContactDistance": st.speculative_contact_distance,
                            "manifoldTolerance": st.manifold_tolerance,
                            "collideWithBackfaces": bool(st.collide_with_backfaces), "useManifoldReduction": True,
                            "normalCosMaxDeltaRotation": st.normal_cos_max_delta_rotation},
            "solver": {"minVelocityForRestitution": st.min_velocity_for_restitution,
                       "velocityIterations": st.velocity_iterations, "positionIterations": st.position_iterations,
                       "warmStarting": bool(st.warm_starting), "penetrationSlop": st.penetration_slop,
                       "baumgarteFactor": st.baumgarte_factor, "warmStartImpulseRatio": 1.0,
                       "maxPenetrationDistance": st.max_penetration_distance},
            "sleeping": {"allowSleeping": bool(st.allow_sleeping), "timeBeforeSleep": st.time_before_sleep,
                         "pointVelocitySleepThreshold": st.point_velocity_sleep_threshold},
            "ccd": {"linearCastThreshold": st.ccd_linear_cast_threshold,
                    "linearCastMaxPenetration": st.ccd_linear_cast_max_penetration,
                    "minVelocityForRestitution": st.ccd_min_velocity_for_restitution},
        }

        def aabb(sh):
            return [list(sh.aabb_min), list(sh.aabb_max)]

        def record(sh):
            base = {"aabb": aabb(sh), "centerOfMass": list(sh.center_of_mass), "density": sh.density,
                    "volume": sh.volume, "materialId": -1}
            if sh.type == S.COMPOUND:  # compound.ts:15-27
                base.pop("density")
                base.pop("materialId")
                base.update(type=5, children=[{"position": list(p), "quaternion": list(q), "shape": record(ch)}
                                              for p, q, ch in sh.children])
                return base
            if sh.wrap is not None:  # scaled.ts:31-38, transformed.ts:31-39, offset-center-of-mass.ts:40-47
                kind, par, inner = sh.wrap
                base.pop("density")
                base.pop("materialId")
                base["shape"] = record(inner)
                if kind == "scaled":
                    base.update(type=7, scale=list(par))
                elif kind == "transformed":
                    base.update(type=6, position=list(par[0]), quaternion=list(par[1]))
                else:
                    base.update(type=10, offset=list(par))
                return base
            if sh.type == S.SPHERE:
                base.update(type=0, radius=sh.radius)
            elif sh.type == S.BOX:
                base.update(type=1, halfExtents=list(sh.half_extents), convexRadius=sh.convex_radius)
            elif sh.type == S.CAPSULE:
                base.update(type=2, radius=sh.radius, halfHeightOfCylinder=sh.half_height)
            elif sh.type == S.CYLINDER:
                base.update(type=9, radius=sh.radius, halfHeight=sh.half_height, convexRadius=sh.convex_radius)
            elif sh.type == S.PLANE:
                base.update(type=11, plane={"normal": list(sh.half_extents), "constant": sh.radius},
                            halfExtent=sh.half_height)
            elif sh.type == S.CONVEX_HULL:
                vi, faces = [], []
                for f in sh.faces:
                    faces.append({"firstVertex": len(vi), "numVertices": len(f)})
                    vi.extend(int(v) for v in f)
                pts = []
                for p_, pf in zip(sh.points, sh.point_faces):
                    fs = [int(x) for x in pf] + [0] * (3 - len(pf))
                    pts.append({"position": [float(x) for x in p_], "numFaces": len(pf), "faces": fs})
                base.update(type=3, points=pts, faces=faces, vertexIndices=vi, convexRadius=sh.convex_radius,
                            planes=[{"normal": [float(x) for x in nrm], "constant": float(c)}
                                    for nrm, c in zip(sh.plane_normals, sh.plane_constants)])
            elif sh.type == S.TRIANGLE_MESH:
                pos = np.asarray(sh.mesh_positions, dtype=np.float64)
                tri = np.asarray(sh.mesh_triangles, dtype=np.int64)
                buf = []
                for a, b, c, act in tri:  # stride 8: a, b, c, normal, active edges, material
                    nrm = np.cross(pos[b] - pos[a], pos[c] - pos[a])
                    nrm = nrm / max(np.linalg.norm(nrm), 1e-300)
                    buf.extend([int(a), int(b), int(c), float(nrm[0]), float(nrm[1]), float(nrm[2]), int(act), -1])
                base.update(type=4, data={"positions": pos.reshape(-1).tolist(), "triangleBuffer": buf,
                                          "triangleCount": len(tri)},
                            bvh={"buffer": np.asarray(sh.mesh_nodes, dtype=np.float64).reshape(-1).tolist()})
            else:
                raise ValueError(f"shape type {sh.type}")
            return base

        shapes = [record(sh) for sh in self.shapes]
        b = self.bodies()
        pool = []
        for k, r in enumerate(b):
            pool.append({
                "_pooled": False, "id": k + k * 2 ** 32, "index": k, "sequence": k, "userData": None,
                "shape": int(r["shape"]), "position": r["position"].tolist(), "quaternion": r["quaternion"].tolist(),
                "centerOfMassPosition": [0.0, 0.0, 0.0], "motionType": int(r["motion_type"]),
                "motionProperties": {
                    "linearVelocity": r["linear_velocity"].tolist(), "angularVelocity": r["angular_velocity"].tolist(),
                    "invInertiaDiagonal": r["inv_inertia_diagonal"].tolist(),
                    "inertiaRotation": r["inertia_rotation"].tolist(),
                    "force": r["force"].tolist(), "torque": r["torque"].tolist(), "invMass": float(r["inv_mass"]),
                    "linearDamping": float(r["linear_damping"]), "angularDamping": float(r["angular_damping"]),
                    "maxLinearVelocity": float(r["max_linear_velocity"]),
                    "maxAngularVelocity": float(r["max_angular_velocity"]), "gravityFactor": float(r["gravity_factor"]),
                    "motionQuality": int(r["motion_quality"]), "allowedDegreesOfFreedom": int(r["allowed_dofs"]),
                    "numVelocityStepsOverride": int(r["num_velocity_steps_override"]),
                    "numPositionStepsOverride": int(r["num_position_steps_override"]),
                    "sleepTestSpheres": [], "allowSleeping": bool(r["allow_sleeping"]), "sleepTestTimer": 0},
                "friction": float(r["friction"]), "restitution": float(r["restitution"]),
                "frictionCombineMode": int(r["friction_combine"]), "restitutionCombineMode": int(r["restitution_combine"]),
                "collisionGroups": int(r["collision_groups"]), "collisionMask": int(r["collision_mask"]),
                "useManifoldReduction": True, "sensor": bool(r["sensor"]), "enhancedInternalEdgeRemoval": False,
                "collideKinematicVsNonDynamic": False, "sleeping": bool(r["sleeping"]),
                "aabb": [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]], "objectLayer": int(r["object_layer"]),
            })
        cons = []
        for k, (a, bb, la, lb, ov, op) in enumerate(self.point_constraints):
            cons.append({"_pooled": False, "_sleeping": False, "id": k, "index": k, "sequence": 0, "enabled": True,
                         "constraintPriority": 0, "numVelocityStepsOverride": ov, "numPositionStepsOverride": op,
                         "bodyIndexA": a, "bodyIndexB": bb, "localSpacePosition1": list(la),
                         "localSpacePosition2": list(lb)})
        return {"settings": settings, "shapes": shapes,
                "bodies": {"pool": pool, "freeIndices": [], "nextSequence": len(pool)},
                "constraints": {"pools": {"0": {"type": 0, "constraints": cons, "freeIndices": [],
                                                "nextSequence": len(cons)}} if cons else {}},
                "previousTimeStep": 0}


# ------------------------------------------------------------------------ PRNG
_M64 = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return x, z ^ (z >> 31)


class Rng:
    """xoshiro256** seeded by splitmix64."""

    def __init__(self, seed):
        x = seed & _M64
        self.s = []
        for _ in range(4):
            x, z = _splitmix64(x)
            self.s.append(z)

    def next_u64(self):
        s = self.s
        r = (((s[1] * 5) & _M64) << 7 | ((s[1] * 5) & _M64) >> 57) & _M64
        r = (r * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & _M64
        return r

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def normal(self, mean=0.0, std=1.0):
        u1 = self.uniform()
        u2 = self.uniform()
        if u1 < 1e-300:
            u1 = 1e-300
        return mean + std * math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)

    def quaternion(self):
        """Uniform random rotation (Shoemake)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a = math.sqrt(1 - u1)
        b = math.sqrt(u1)
        return (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3),
                b * math.cos(2 * math.pi * u3))


# ------------------------------------------------------------------ generators
def _floor(scene, half=(50.0, 0.5, 50.0), top=0.0, friction=0.2):
    return scene.add_body(S.box(half), (0.0, top - half[1], 0.0), motion_type=_abi.CC_MOTION_STATIC,
                          friction=friction)


def cube_heap(seed=1001, nx=10, ny=10, nz=10, scene=None, friction=0.5):
    """Config 1: 1,000 boxes (he 0.5, density 1000) on a 10x10x10 grid at 1.1 m pitch, the
    lowest layer 1 m above a static 50x0.5x50 floor, +-0.05 m jitter, random yaw, friction 0.5."""
    sc = scene if scene is not None else Scene()
    rng = Rng(seed)
    _floor(sc, friction=friction)
    shape = S.box((0.5, 0.5, 0.5))
    pitch = 1.1
    for j in range(ny):
        for i in range(nx):
            for k in range(nz):
                x = (i - (nx - 1) / 2) * pitch + rng.uniform(-0.05, 0.05)
                y = 1.0 + 0.5 + j * pitch + rng.uniform(-0.05, 0.05)
                z = (k - (nz - 1) / 2) * pitch + rng.uniform(-0.05, 0.05)
                q = _quat_yaw(rng.uniform(0, 2 * math.pi))
                sc.add_body(shape, (x, y, z), q, friction=friction)
    sc.end_world()
    return sc


def _unit_sphere_points(rng, n):
    pts = []
    for _ in range(n):
        z = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0, 2 * math.pi)
        r = math.sqrt(max(0.0, 1 - z * z))
        pts.append((r * math.cos(phi), r * math.sin(phi), z))
    return np.array(pts)


def mixed_pile(seed=1002, n=20000, extent=(60.0, 40.0, 60.0), scene=None, cylinders=False):
    """Config 2: n bodies drawn uniformly from sphere (r U[0.3,0.8]), box (he U[0.2,0.6]^3),
    capsule (r U[0.2,0.4], hh U[0.2,0.6]) and convex hull of 16 points on the unit sphere
    scaled U[0.4,1.0]; positions uniform in a 60x40x60 m box above a static floor;
    v ~ N(0,1) per axis, w ~ N(0,0.5) per axis.  With `cylinders`, a fifth kind: cylinders of
    r U[0.2,0.6], hh U[0.2,0.6] (convex radius 0.05)."""
    sc = scene if scene is not None else Scene()
    rng = Rng(seed)
    half = (max(50.0, extent[0] / 2 + 5), 0.5, max(50.0, extent[2] / 2 + 5))
    _floor(sc, half=half)
    kinds = 5 if cylinders else 4
    for _ in range(n):
        kind = min(kinds - 1, int(rng.uniform() * kinds))
        if kind == 4:
            shape = S.cylinder(rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6))
        elif kind == 0:
            shape = S.sphere(rng.uniform(0.3, 0.8))
        elif kind == 1:
            shape = S.box((rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6)))
        elif kind == 2:
            shape = S.capsule(rng.uniform(0.2, 0.4), rng.uniform(0.2, 0.6))
        else:
            pts = _unit_sphere_points(rng, 16) * rng.uniform(0.4, 1.0)
            shape = S.convex_hull(pts)
        x = rng.uniform(-extent[0] / 2, extent[0] / 2)
        y = rng.uniform(1.0, 1.0 + extent[1])
        z = rng.uniform(-extent[2] / 2, extent[2] / 2)
        q = rng.quaternion()
        v = (rng.normal(0, 1), rng.normal(0, 1), rng.normal(0, 1))
        w = (rng.normal(0, 0.5), rng.normal(0, 0.5), rng.normal(0, 0.5))
        sc.add_body(shape, (x, y, z), q, linear_velocity=v, angular_velocity=w)
    sc.end_world()
    return sc


def decorated_pile(seed=1009, n=300, extent=(12.0, 8.0, 12.0), scene=None):
    """A pile of decorated shapes over a static box floor, drawn uniformly from: scaled boxes
    (per-axis scale in [0.5, 2], each axis mirrored with probability 1/4, so some are inside out),
    uniformly scaled spheres, capsules and cylinders, scaled 16-point hulls (non-uniform), boxes,
    capsules, cylinders and hulls in a transformed decorator (offset up to 0.3 m, random rotation),
    and offset-centre-of-mass boxes and spheres; positions uniform in the extent above the floor,
    v ~ N(0, 1), w ~ N(0, 0.5)."""
    sc = scene if scene is not None else Scene()
    rng = Rng(seed)
    half = (max(50.0, extent[0] / 2 + 5), 0.5, max(50.0, extent[2] / 2 + 5))
    _floor(sc, half=half)

    def sgn():
        return -1.0 if rng.uniform() < 0.25 else 1.0

    for _ in range(n):
        kind = min(9, int(rng.uniform() * 10))
        if kind == 0:
            shape = S.scaled(S.box((rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5))),
                             (sgn() * rng.uniform(0.5, 2.0), sgn() * rng.uniform(0.5, 2.0), sgn() * rng.uniform(0.5, 2.0)))
        elif kind == 1:
            u = sgn() * rng.uniform(0.5, 2.0)
            shape = S.scaled(S.sphere(rng.uniform(0.2, 0.4)), (u, u, u))
        elif kind == 2:
            u = rng.uniform(0.5, 2.0)
            shape = S.scaled(S.capsule(rng.uniform(0.1, 0.25), rng.uniform(0.1, 0.3)), (u, u, u))
        elif kind == 3:
            u = rng.uniform(0.7, 1.5)
            shape = S.scaled(S.cylinder(rng.uniform(0.2, 0.4), rng.uniform(0.2, 0.4)), (u, u, u))
        elif kind == 4:
            pts = _unit_sphere_points(rng, 16) * rng.uniform(0.3, 0.6)
            shape = S.scaled(S.convex_hull(pts), (rng.uniform(0.6, 1.6), sgn() * rng.uniform(0.6, 1.6),
                                                  rng.uniform(0.6, 1.6)))
        elif kind in (5, 6, 7):
            inner = [S.box((rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5))),
                     S.capsule(rng.uniform(0.15, 0.3), rng.uniform(0.2, 0.5)),
                     S.cylinder(rng.uniform(0.2, 0.4), rng.uniform(0.2, 0.4))][kind - 5]
            if rng.uniform() < 0.5:
                inner = S.scaled(inner, (1.5, 1.5, 1.5))
            t = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3))
            shape = S.transformed(inner, t, rng.quaternion())
        elif kind == 8:
            pts = _unit_sphere_points(rng, 16) * rng.uniform(0.3, 0.6)
            shape = S.transformed(S.convex_hull(pts), (0.0, rng.uniform(-0.2, 0.2), 0.0), rng.quaternion())
        else:
            inner = S.box((rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5))) \
                if rng.uniform() < 0.5 else S.sphere(rng.uniform(0.2, 0.4))
            shape = S.offset_center_of_mass(inner, (rng.uniform(-0.1, 0.1), rng.uniform(-0.2, 0.0),
                                                    rng.uniform(-0.1, 0.1)))
        x = rng.uniform(-extent[0] / 2, extent[0] / 2)
        y = rng.uniform(1.0, 1.0 + extent[1])
        z = rng.uniform(-extent[2] / 2, extent[2] / 2)
        v = (rng.normal(0, 1), rng.normal(0, 1), rng.normal(0, 1))
        w = (rng.normal(0, 0.5), rng.normal(0, 0.5), rng.normal(0, 0.5))
        sc.add_body(shape, (x, y, z), rng.quaternion(), linear_velocity=v, angular_velocity=w)
    sc.end_world()
    return sc


def compound_pile(seed=1012, n=200, extent=(12.0, 8.0, 12.0), scene=None):
    """A pile over a static box floor: two thirds compound bodies of 2-4 children (boxes,
    spheres, capsules, cylinders and 12-point hulls, some scaled, at random positions within
    0.4 m and random orientations), one third plain boxes and spheres; positions uniform in the
    extent above the floor, v ~ N(0, 1), w ~ N(0, 0.5)."""
    sc = scene if scene is not None else Scene()
    rng = Rng(seed)
    half = (max(50.0, extent[0] / 2 + 5), 0.5, max(50.0, extent[2] / 2 + 5))
    _floor(sc, half=half)

    def leaf():
        k = min(4, int(rng.uniform() * 5))
        if k == 0:
            return S.box((rng.uniform(0.15, 0.35), rng.uniform(0.15, 0.35), rng.uniform(0.15, 0.35)))
        if k == 1:
            return S.sphere(rng.uniform(0.15, 0.3))
        if k == 2:
            return S.capsule(rng.uniform(0.1, 0.2), rng.uniform(0.1, 0.3))
        if k == 3:
            return S.cylinder(rng.uniform(0.15, 0.3), rng.uniform(0.15, 0.3))
        return S.convex_hull(_unit_sphere_points(rng, 12) * rng.uniform(0.2, 0.4))

    for _ in range(n):
        if rng.uniform() < 2.0 / 3.0:
            kids = []
            for _c in range(2 + min(2, int(rng.uniform() * 3))):
                ch = leaf()
                if rng.uniform() < 0.25:
                    ch = S.scaled(ch, (1.3, 1.3, 1.3) if ch.type in (S.SPHERE, S.CAPSULE, S.CYLINDER)
                                  else (rng.uniform(0.7, 1.4), rng.uniform(0.7, 1.4), rng.uniform(0.7, 1.4)))
                kids.append(((rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4)), rng.quaternion(), ch))
            shape = S.compound(kids)
        elif rng.uniform() < 0.5:
            shape = S.box((rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5)))
        else:
            shape = S.sphere(rng.uniform(0.2, 0.5))
        x = rng.uniform(-extent[0] / 2, extent[0] / 2)
        y = rng.uniform(1.0, 1.0 + extent[1])
        z = rng.uniform(-extent[2] / 2, extent[2] / 2)
        v = (rng.normal(0, 1), rng.normal(0, 1), rng.normal(0, 1))
        w = (rng.normal(0, 0.5), rng.normal(0, 0.5), rng.normal(0, 0.5))
        sc.add_body(shape, (x, y, z), rng.quaternion(), linear_velocity=v, angular_velocity=w)
    sc.end_world()
    return sc


def ccd_gallery(seed=1013, bullets=24, targets=60, scene=None, plane_floor=False, compound_targets=False,
                decorated_targets=False, decorated_bullets=False):
    """Fast LINEAR_CAST bodies (MotionQuality.LINEAR_CAST, src/update.ts:932-967) in a walled
    range: a static floor, four thin static walls (0.1 m) around a 16 m square, one tilted thin
    slab in the middle, and `targets` dynamic boxes, spheres, capsules, cylinders and 12-point hulls
    resting on the floor (a third of them asleep).  The bullets (the same five shapes, 0.1-0.3 m)
    start inside the walls at 30-150 m/s towards a random point on a wall or a target, with random
    spin and restitution in [0, 0.6); pairs of bullets fly at each other, one kinematic LINEAR_CAST
    paddle sweeps the range at 60 m/s, and a few slow LINEAR_CAST bodies stay below the cast
    threshold.  With `plane_floor` the floor is a static plane through the origin; with
    `compound_targets` every third target is a compound of two or three unscaled leaves; with
    `decorated_targets` every third is instead scaled, transformed, offset-centre-of-mass or a
    compound with a scaled child, in turn; with `decorated_bullets` every third bullet is scaled
    (non-uniformly for boxes and hulls), a compound of two leaves, in a transformed decorator or
    has an offset centre of mass."""
    sc = scene if scene is not None else Scene()
    rng = Rng(seed)
    if plane_floor:
        sc.add_body(S.plane((0.0, 1.0, 0.0)), (0.0, 0.0, 0.0), motion_type=_abi.CC_MOTION_STATIC)
    else:
        sc.add_body(S.box((12.0, 0.5, 12.0)), (0.0, -0.5, 0.0), motion_type=_abi.CC_MOTION_STATIC)
    for x, z, he in ((8.0, 0.0, (0.05, 3.0, 8.0)), (-8.0, 0.0, (0.05, 3.0, 8.0)),
                     (0.0, 8.0, (8.0, 3.0, 0.05)), (0.0, -8.0, (8.0, 3.0, 0.05))):
        sc.add_body(S.box(he), (x, 3.0, z), motion_type=_abi.CC_MOTION_STATIC)
    sc.add_body(S.box((2.0, 1.5, 0.05)), (0.0, 1.5, 0.0), _quat_yaw(0.6), motion_type=_abi.CC_MOTION_STATIC)

    def shape(lo, hi):
        k = min(4, int(rng.uniform() * 5))
        r = rng.uniform(lo, hi)
        if k == 0:
            return S.box((r, rng.uniform(lo, hi), rng.uniform(lo, hi)))
        if k == 1:
            return S.sphere(r)
        if k == 2:
            return S.capsule(rng.uniform(lo, hi), r * 0.6)
        if k == 3:
            return S.cylinder(rng.uniform(lo, hi), r)
        return S.convex_hull(_unit_sphere_points(rng, 12) * r)

    aims = []
    for _ in range(targets):
        p = (rng.uniform(-6.5, 6.5), rng.uniform(0.3, 0.6), rng.uniform(-6.5, 6.5))
        if abs(p[0]) < 2.5 and abs(p[2]) < 1.5:
            p = (p[0] + 5.0, p[1], p[2])
        tshape = shape(0.2, 0.4)
        if decorated_targets and len(aims) % 3 == 0:
            kind = (len(aims) // 3) % 4
            if kind == 0:
                tshape = S.scaled(S.box((0.2, 0.25, 0.3)), (1.5, 0.8, -1.2))
            elif kind == 1:
                tshape = S.transformed(shape(0.15, 0.3), (0.2, 0.1, -0.1), rng.quaternion())
            elif kind == 2:
                tshape = S.offset_center_of_mass(shape(0.2, 0.35), (0.05, -0.1, 0.08))
            else:
                tshape = S.compound([((0.25, 0.0, 0.0), rng.quaternion(), S.scaled(S.box((0.15, 0.15, 0.15)), (1.4, 1.4, 1.4))),
                                     ((-0.25, 0.0, 0.0), rng.quaternion(), shape(0.12, 0.2))])
        elif compound_targets and len(aims) % 3 == 0:
            tshape = S.compound([((rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2), rng.uniform(-0.3, 0.3)), rng.quaternion(),
                                  shape(0.12, 0.25)) for _ in range(2 + int(rng.uniform() * 2))])
        sc.add_body(tshape, p, rng.quaternion(), sleeping=rng.uniform() < 1.0 / 3.0)
        aims.append(p)

    nfired = [0]

    def bullet_shape():
        s = shape(0.1, 0.3)
        nfired[0] += 1
        if decorated_bullets and nfired[0] % 3 == 0:
            if nfired[0] % 2 == 0:
                u = rng.uniform(0.7, 1.4)
                s = S.scaled(s, (u, u, u) if s.type in (S.SPHERE, S.CAPSULE, S.CYLINDER)
                             else (rng.uniform(0.7, 1.4), rng.uniform(0.7, 1.4), rng.uniform(0.7, 1.4)))
            elif nfired[0] % 15 == 0:
                s = S.compound([((0.12, 0.0, 0.0), rng.quaternion(), s), ((-0.12, 0.0, 0.0), rng.quaternion(), shape(0.08, 0.15))])
            elif nfired[0] % 9 == 0:
                s = S.transformed(s, (rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)), rng.quaternion())
            else:
                s = S.offset_center_of_mass(s, (rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)))
        return s

    def fire(pos, aim, speed, **kw):
        d = np.array(aim, dtype=float) - np.array(pos, dtype=float)
        d /= np.linalg.norm(d)
        w = (rng.normal(0, 5), rng.normal(0, 5), rng.normal(0, 5))
        return sc.add_body(bullet_shape(), pos, rng.quaternion(), linear_velocity=tuple(speed * d),
                           angular_velocity=w, motion_quality=_abi.CC_MOTION_QUALITY_LINEAR_CAST,
                           restitution=rng.uniform(0.0, 0.6), **kw)

    for b in range(bullets):
        pos = (rng.uniform(-6.0, 6.0), rng.uniform(0.5, 4.0), rng.uniform(-6.0, 6.0))
        if rng.uniform() < 0.5:
            aim = aims[int(rng.uniform() * len(aims))]
        else:
            side = int(rng.uniform() * 4)
            t = rng.uniform(-6.0, 6.0)
            aim = [(8.0, 1.5, t), (-8.0, 1.5, t), (t, 1.5, 8.0), (t, 1.5, -8.0)][side]
        fire(pos, aim, rng.uniform(30.0, 150.0))
    for k in range(3):  # head-on pairs
        y = 3.0 + 0.4 * k
        z = rng.uniform(-5.0, 5.0)
        fire((-3.0, y, z), (3.0, y, z), 90.0, gravity_factor=0.0)
        fire((3.0, y, z), (-3.0, y, z), 90.0, gravity_factor=0.0)
    sc.add_body(S.box((0.1, 0.5, 1.0)), (-6.0, 4.5, 0.0), motion_type=_abi.CC_MOTION_KINEMATIC,
                linear_velocity=(60.0, 0.0, 0.0), motion_quality=_abi.CC_MOTION_QUALITY_LINEAR_CAST)
    for _ in range(4):  # below the threshold: 0.05 * inner radius per step
        sc.add_body(S.sphere(0.3), (rng.uniform(-5.0, 5.0), 1.0, rng.uniform(-5.0, 5.0)),
                    linear_velocity=(0.5, 0.0, 0.0), motion_quality=_abi.CC_MOTION_QUALITY_LINEAR_CAST)
    sc.end_world()
    return sc


def batched_worlds(seed=1003, num_worlds=4096, min_bodies=8, max_bodies=512, chains=2, chain_len=8, scene=None):
    """Config 3: independent worlds, body count per world log-uniform in [8, 512]; each world
    is a pile of boxes and spheres (he/r 0.5) on a 1.1 m grid above a floor plus `chains`
    chains of `chain_len` capsules (r 0.1, hh 0.3) joined end to end by point constraints."""
    sc = scene if scene is not None else Scene()
    box = S.box((0.5, 0.5, 0.5))
    ball = S.sphere(0.5)
    cap = S.capsule(0.1, 0.3)
    for w in range(num_worlds):
        rng = Rng((seed << 32) + w)
        n = int(math.floor(math.exp(rng.uniform(math.log(min_bodies), math.log(max_bodies)))))
        _floor(sc, friction=0.5)
        side = max(1, int(math.ceil(n ** (1 / 3))))
        for b in range(n):
            i, k, j = b % side, (b // side) % side, b // (side * side)
            x = (i - (side - 1) / 2) * 1.1 + rng.uniform(-0.05, 0.05)
            y = 1.5 + j * 1.1 + rng.uniform(-0.05, 0.05)
            z = (k - (side - 1) / 2) * 1.1 + rng.uniform(-0.05, 0.05)
            shape = box if rng.uniform() < 0.5 else ball
            sc.add_body(shape, (x, y, z), _quat_yaw(rng.uniform(0, 2 * math.pi)), friction=0.5)
        top = 1.5 + (n // (side * side) + 1) * 1.1 + 1.0
        for c in range(chains):
            z0 = (c - (chains - 1) / 2) * 2.0 + side * 1.1
            prev = None
            for l in range(chain_len):
                x = (l - (chain_len - 1) / 2) * 0.8
                # capsule axis along x: rotate 90 degrees about z
                q = (0.0, 0.0, math.sin(math.pi / 4), math.cos(math.pi / 4))
                idx = sc.add_body(cap, (x, top, z0), q, friction=0.5)
                if prev is not None:
                    # shared end point: +x end of the previous capsule (local -y after the 90 degree
                    # turn about z) and the -x end of this one (local +y)
                    sc.add_point_constraint(prev, idx, (0.0, -0.4, 0.0), (0.0, 0.4, 0.0))
                prev = idx
        sc.end_world()
    return sc


def throw_worlds(seed=1005, num_worlds=1024, friction=0.5, scene=None):
    """Config 5: independent worlds, each a 4x4x4 stack of boxes of half-extent 0.25 (density
    1000) resting on a static floor and a projectile sphere (r 0.4, 10 kg) launched at the stack
    from 4 m away with v0 ~ N((12, 2, 0), 1) m/s; one shared friction coefficient on every body
    (geometric-mean combine, so each contact's friction equals it); sleeping off.  The loss of
    world w is |centroid of its stack - target_w|^2 after the run, with a seeded target a little
    past the stack.  Returns the scene with `throw` = {stack: (W, 64) body indices, projectile:
    (W,) body indices, target: (W, 3)}."""
    sc = scene if scene is not None else Scene()
    sc.settings.allow_sleeping = 0
    box = S.box((0.25, 0.25, 0.25))
    r = 0.4
    ball = S.sphere(r, density=10.0 / (4.0 / 3.0 * math.pi * r ** 3))
    stack, proj, target = [], [], []
    for w in range(num_worlds):
        rng = Rng((seed << 32) + w)
        _floor(sc, friction=friction)
        ids = []
        for j in range(4):
            for i in range(4):
                for k in range(4):
                    x = (i - 1.5) * 0.505
                    y = 0.25 + 0.001 + j * 0.502
                    z = (k - 1.5) * 0.505
                    ids.append(sc.add_body(box, (x, y, z), friction=friction))
        v0 = (rng.normal(12.0, 1.0), rng.normal(2.0, 1.0), rng.normal(0.0, 1.0))
        proj.append(sc.add_body(ball, (-4.0, 1.0, 0.0), linear_velocity=v0, friction=friction))
        stack.append(ids)
        target.append((rng.uniform(0.2, 0.8), rng.uniform(0.8, 1.0), rng.uniform(-0.3, 0.3)))
        sc.end_world()
    sc.throw = {"stack": np.array(stack, dtype=np.int64), "projectile": np.array(proj, dtype=np.int64),
                "target": np.array(target)}
    return sc


def terrain_heights(rng, amps=(0.5, 3.0), wavelengths=(8.0, 64.0), terms=4):
    """height(x, z) = sum of `terms` seeded sinusoids: amplitude U[amps], wavelength U[wavelengths],
    direction uniform in the plane, phase U[0, 2 pi)."""
    par = []
    for _ in range(terms):
        a = rng.uniform(*amps)
        lam = rng.uniform(*wavelengths)
        th = rng.uniform(0, 2 * math.pi)
        ph = rng.uniform(0, 2 * math.pi)
        par.append((a, lam, math.cos(th), math.sin(th), ph))

    def h(x, z):
        y = 0.0 * x
        for a, lam, dx, dz, ph in par:
            y = y + a * np.sin(2 * math.pi * (dx * x + dz * z) / lam + ph)
        return y
    return h


def terrain_drop(seed=1004, vertices=512, spacing=0.5, n=50000, drop=(5.0, 30.0), hull_variants=64, scene=None,
                 spheres=False):
    """Config 4: a static triangle mesh over a `vertices` x `vertices` grid at `spacing` (511^2 * 2
    = 522,242 triangles at the full size), heights from terrain_heights; n boxes (he 0.5),
    capsules (r 0.25, hh 0.25) and convex hulls (16 points on a sphere of radius 0.5; one of
    `hull_variants` seeded hulls), a third each, dropped from U[drop] m above the terrain at
    positions uniform over it (1 m in from its edges), random orientation, at rest.  With
    `spheres`, every fourth body is a sphere of radius 0.5 instead (a quarter of each kind)."""
    sc = scene if scene is not None else Scene()
    rng = Rng(seed)
    h = terrain_heights(rng)
    ext = (vertices - 1) * spacing
    pos, idx = S.height_field_mesh(vertices, vertices, spacing, h, origin=(-ext / 2, -ext / 2))
    sc.add_body(S.triangle_mesh(pos, idx), (0.0, 0.0, 0.0), motion_type=_abi.CC_MOTION_STATIC, friction=0.5)
    box = S.box((0.5, 0.5, 0.5))
    cap = S.capsule(0.25, 0.25)
    hulls = [S.convex_hull(_unit_sphere_points(rng, 16) * 0.5) for _ in range(hull_variants)]
    ball = S.sphere(0.5)
    half = ext / 2 - 1.0
    for i in range(n):
        kind = i % (4 if spheres else 3)
        if kind == 3:
            shape = ball
        else:
            shape = box if kind == 0 else (cap if kind == 1 else hulls[int(rng.uniform(0, hull_variants)) % hull_variants])
        x = rng.uniform(-half, half)
        z = rng.uniform(-half, half)
        y = float(h(np.float64(x), np.float64(z))) + rng.uniform(*drop)
        sc.add_body(shape, (x, y, z), rng.quaternion(), friction=0.5)
    sc.end_world()
    return sc


def replicate(builder, copies, seed0, **kw):
    """A batch of `copies` independent worlds made by `builder` with seeds seed0, seed0+1, ..."""
    sc = Scene()
    for c in range(copies):
        builder(seed=seed0 + c, scene=sc, **kw)
    return sc


def as_fp32(bodies):
    """A copy of body descriptors or body states with the fields an FP32 context stores rounded to
    FP32 (position, orientation, linear and angular velocity).  An FP64 oracle started from the
    result and an FP32 context given the same array then hold one and the same state."""
    out = np.array(bodies, copy=True)
    for f in ("position", "quaternion", "linear_velocity", "angular_velocity"):
        out[f] = out[f].astype(np.float32).astype(np.float64)
    return out


def replicate_world(scene, copies, seed, extent=(60.0, 40.0, 60.0)):
    """A batch scene of `copies` worlds that reuse world 0's bodies and shapes with positions,
    orientations and velocities redrawn per world (config-2 sampling: positions uniform in the
    extent above the floor, v ~ N(0,1), w ~ N(0,0.5)); static bodies are copied unchanged."""
    base = scene.bodies()[: scene.offsets()[1]]
    out = Scene(scene.settings)
    out.shapes = scene.shapes
    out._shape_index = scene._shape_index
    out.point_constraints = []
    n = len(base)
    arrs = []
    for c in range(copies):
        b = base.copy()
        if c > 0:
            rng = np.random.Generator(np.random.PCG64(seed * 1000003 + c))
            dyn = b["motion_type"] != _abi.CC_MOTION_STATIC
            k = int(dyn.sum())
            b["position"][dyn, 0] = rng.uniform(-extent[0] / 2, extent[0] / 2, k)
            b["position"][dyn, 1] = rng.uniform(1.0, 1.0 + extent[1], k)
            b["position"][dyn, 2] = rng.uniform(-extent[2] / 2, extent[2] / 2, k)
            u = rng.uniform(size=(k, 3))
            a = np.sqrt(1 - u[:, 0])
            bb = np.sqrt(u[:, 0])
            b["quaternion"][dyn] = np.stack([a * np.sin(2 * np.pi * u[:, 1]), a * np.cos(2 * np.pi * u[:, 1]),
                                              bb * np.sin(2 * np.pi * u[:, 2]), bb * np.cos(2 * np.pi * u[:, 2])], 1)
            b["linear_velocity"][dyn] = rng.normal(0, 1, (k, 3))
            b["angular_velocity"][dyn] = rng.normal(0, 0.5, (k, 3))
        arrs.append(b)
    out._bodies = np.concatenate(arrs) if arrs else np.zeros(0, dtype=BODY_DTYPE)
    out.rows = [None] * len(out._bodies)
    out.world_offsets = [i * n for i in range(copies + 1)]
    return out
