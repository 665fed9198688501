"""ctypes mirrors of the plain-data structs declared in include/crashcat_amd.h."""
import ctypes as C

CC_SHAPE_SPHERE = 0
CC_SHAPE_BOX = 1
CC_SHAPE_CAPSULE = 2
CC_SHAPE_CONVEX_HULL = 3
CC_SHAPE_TRIANGLE_MESH = 4
CC_SHAPE_CYLINDER = 5
CC_SHAPE_PLANE = 6
CC_SHAPE_COMPOUND = 7
CC_DECO_SCALED = 1
CC_DECO_TRANSFORMED = 2
CC_DECO_OFFSET_COM = 4

CC_MOTION_STATIC = 0
CC_OK = 0
CC_E_ARG = -1
CC_E_HIP = -2
CC_E_CAPACITY = -3
CC_E_STATE = -4
CC_E_UNSUPPORTED = -5
CC_ORDER_REFERENCE = 0
CC_ORDER_COLORED = 1
CC_MOTION_KINEMATIC = 1
CC_MOTION_DYNAMIC = 2
CC_MOTION_QUALITY_DISCRETE = 0
CC_MOTION_QUALITY_LINEAR_CAST = 1

CC_PRECISION_FP32 = 0
CC_PRECISION_FP64 = 1

CC_MAX_OBJECT_LAYERS = 8

D3 = C.c_double * 3
D4 = C.c_double * 4


class Settings(C.Structure):
    _fields_ = [
        ("gravity", D3),
        ("speculative_contact_distance", C.c_double),
        ("manifold_tolerance", C.c_double),
        ("collision_tolerance", C.c_double),
        ("penetration_tolerance", C.c_double),
        ("normal_cos_max_delta_rotation", C.c_double),
        ("contact_point_preserve_lambda_max_dist_sq", C.c_double),
        ("min_velocity_for_restitution", C.c_double),
        ("penetration_slop", C.c_double),
        ("baumgarte_factor", C.c_double),
        ("max_penetration_distance", C.c_double),
        ("time_before_sleep", C.c_double),
        ("point_velocity_sleep_threshold", C.c_double),
        ("velocity_iterations", C.c_int32),
        ("position_iterations", C.c_int32),
        ("gravity_enabled", C.c_int32),
        ("warm_starting", C.c_int32),
        ("allow_sleeping", C.c_int32),
        ("num_object_layers", C.c_int32),
        ("object_layer_pairs", C.c_uint8 * (CC_MAX_OBJECT_LAYERS * CC_MAX_OBJECT_LAYERS)),
        ("collide_with_backfaces", C.c_int32),
        ("collide_only_with_active_edges", C.c_int32),
        ("ccd_linear_cast_threshold", C.c_double),
        ("ccd_linear_cast_max_penetration", C.c_double),
        ("ccd_min_velocity_for_restitution", C.c_double),
    ]


class ShapeDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("hull_first_point", C.c_int32),
        ("hull_num_points", C.c_int32),
        ("hull_first_face", C.c_int32),
        ("hull_num_faces", C.c_int32),
        ("pad_", C.c_int32),
        ("radius", C.c_double),
        ("half_height", C.c_double),
        ("half_extents", D3),
        ("convex_radius", C.c_double),
        ("center_of_mass", D3),
        ("aabb_min", D3),
        ("aabb_max", D3),
        ("mesh_first_triangle", C.c_int32),
        ("mesh_num_triangles", C.c_int32),
        ("mesh_first_node", C.c_int32),
        ("mesh_num_nodes", C.c_int32),
        ("decorated", C.c_int32),
        ("pad2_", C.c_int32),
        ("scale", D3),
        ("local_position", D3),
        ("local_quaternion", D4),
        ("compound_first_child", C.c_int32),
        ("compound_num_children", C.c_int32),
    ]


class MeshData(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_double)),
        ("triangles", C.POINTER(C.c_int32)),
        ("nodes", C.POINTER(C.c_double)),
        ("total_vertices", C.c_int32),
        ("total_triangles", C.c_int32),
        ("total_nodes", C.c_int32),
        ("pad_", C.c_int32),
    ]


class HullData(C.Structure):
    _fields_ = [
        ("points", C.POINTER(C.c_double)),
        ("shrunk_points", C.POINTER(C.c_double)),
        ("face_first_vertex", C.POINTER(C.c_int32)),
        ("face_num_vertices", C.POINTER(C.c_int32)),
        ("face_vertex_indices", C.POINTER(C.c_int32)),
        ("plane_normals", C.POINTER(C.c_double)),
        ("plane_constants", C.POINTER(C.c_double)),
        ("total_points", C.c_int32),
        ("total_faces", C.c_int32),
        ("total_face_vertices", C.c_int32),
        ("pad_", C.c_int32),
    ]


class BodyDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("motion_type", C.c_int32),
        ("object_layer", C.c_int32),
        ("collision_groups", C.c_uint32),
        ("collision_mask", C.c_uint32),
        ("allow_sleeping", C.c_int32),
        ("allowed_dofs", C.c_int32),
        ("friction_combine", C.c_int32),
        ("restitution_combine", C.c_int32),
        ("sleeping", C.c_int32),
        ("position", D3),
        ("quaternion", D4),
        ("linear_velocity", D3),
        ("angular_velocity", D3),
        ("inv_mass", C.c_double),
        ("inv_inertia_diagonal", D3),
        ("inertia_rotation", D4),
        ("friction", C.c_double),
        ("restitution", C.c_double),
        ("gravity_factor", C.c_double),
        ("linear_damping", C.c_double),
        ("angular_damping", C.c_double),
        ("max_linear_velocity", C.c_double),
        ("max_angular_velocity", C.c_double),
        ("force", D3),
        ("torque", D3),
        ("num_velocity_steps_override", C.c_int32),
        ("num_position_steps_override", C.c_int32),
        ("sensor", C.c_int32),
        ("motion_quality", C.c_int32),
    ]


class BodyState(C.Structure):
    _fields_ = [
        ("position", D3),
        ("quaternion", D4),
        ("center_of_mass_position", D3),
        ("linear_velocity", D3),
        ("angular_velocity", D3),
        ("aabb_min", D3),
        ("aabb_max", D3),
        ("sleeping", C.c_int32),
        ("island", C.c_int32),
    ]


class PointConstraintDesc(C.Structure):
    _fields_ = [
        ("body_a", C.c_int32),
        ("body_b", C.c_int32),
        ("enabled", C.c_int32),
        ("pad_", C.c_int32),
        ("local_point_a", D3),
        ("local_point_b", D3),
        ("num_velocity_steps_override", C.c_int32),
        ("num_position_steps_override", C.c_int32),
    ]


class ContactOut(C.Structure):
    _fields_ = [
        ("body_a", C.c_int32),
        ("body_b", C.c_int32),
        ("num_points", C.c_int32),
        ("persisted", C.c_int32),
        ("sub_shape_a", C.c_uint32),
        ("sub_shape_b", C.c_uint32),
        ("normal", D3),
        ("position_a", D3 * 4),
        ("position_b", D3 * 4),
        ("lambda_n", D4),
        ("lambda_t1", D4),
        ("lambda_t2", D4),
    ]


class CcdBody(C.Structure):
    _fields_ = [
        ("body_index", C.c_int32),
        ("hit_body_index", C.c_int32),
        ("delta_position", D3),
        ("contact_normal", D3),
        ("contact_point", D3),
        ("fraction", C.c_double),
        ("fraction_plus_slop", C.c_double),
        ("linear_cast_threshold_sq", C.c_double),
        ("max_penetration", C.c_double),
        ("combined_friction", C.c_double),
        ("combined_restitution", C.c_double),
        ("sub_shape_id_1", C.c_uint32),
        ("sub_shape_id_2", C.c_uint32),
    ]


class ContactRemoved(C.Structure):
    _fields_ = [
        ("body_a", C.c_int32),
        ("body_b", C.c_int32),
        ("sub_shape_a", C.c_uint32),
        ("sub_shape_b", C.c_uint32),
    ]


class StepCounters(C.Structure):
    _fields_ = [
        ("num_bodies", C.c_int64),
        ("num_awake", C.c_int64),
        ("num_pairs", C.c_int64),
        ("num_contacts", C.c_int64),
        ("num_points", C.c_int64),
        ("num_epa", C.c_int64),
        ("num_islands", C.c_int64),
        ("max_levels", C.c_int64),
        ("position_sweeps", C.c_int64),
        ("overflow", C.c_int64),
    ]


class CachedContact(C.Structure):
    _fields_ = [
        ("body_a", C.c_int32),
        ("body_b", C.c_int32),
        ("sub_shape_a", C.c_uint32),
        ("sub_shape_b", C.c_uint32),
        ("num_points", C.c_int32),
        ("pad_", C.c_int32),
        ("local_a", (C.c_double * 3) * 4),
        ("local_b", (C.c_double * 3) * 4),
        ("lambda_n", C.c_double * 4),
        ("lambda_t1", C.c_double * 4),
        ("lambda_t2", C.c_double * 4),
    ]


class BodySleep(C.Structure):
    _fields_ = [
        ("timer", C.c_double),
        ("center", (C.c_double * 3) * 3),
        ("radius", C.c_double * 3),
    ]


class Totals(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("pairs", C.c_int64),
        ("contacts", C.c_int64),
        ("points", C.c_int64),
        ("epa", C.c_int64),
        ("awake_body_steps", C.c_int64),
        ("solver_bytes", C.c_int64),
        ("gjk", C.c_int64),
        ("body_steps", C.c_int64),
        ("reserved", C.c_int64 * 7),
    ]


class Config(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("device", C.c_int32),
        ("max_pairs_per_body", C.c_int32),
        ("max_contacts_per_body", C.c_int32),
        ("block_solver", C.c_int32),
        ("block_bodies", C.c_int32),
        ("island_groups", C.c_int32),
        ("disable_graphs", C.c_int32),
        ("solver_order", C.c_int32),
        ("measure_residual", C.c_int32),
        ("narrowphase_fp32", C.c_int32),
    ]
