/*
 * Node N-API addon over the C ABI (include/crashcat_amd.h): the binding crashcat's JavaScript
 * host loads so that updateWorld (src/update.ts:45) can hand a world's step to the GPU.
 * Every buffer crosses as an ArrayBuffer or typed array owned by JavaScript; struct layouts
 * are taken from the C header at build time and exported by layout(), so the JavaScript side
 * packs fields by name and never hard-codes an offset.  Errors become thrown JS Errors carrying
 * cc_last_error(), as crashcat's assert() throws (src/utils/assert.ts:1-4).
 *
 * Build: gcc -O2 -shared -fPIC -I/usr/include/node -I<repo>/include cc_node.c
 *        -L<repo>/crashcat_amd -lcrashcat_amd -Wl,-rpath,'$ORIGIN/..' -o crashcat_amd.node
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include "crashcat_amd.h"

#define CHECK(env, call)                                                    \
    do {                                                                    \
        if ((call) != napi_ok) {                                            \
            napi_throw_error((env), NULL, "crashcat_amd: N-API call failed"); \
            return NULL;                                                    \
        }                                                                   \
    } while (0)

static napi_value throw_rc(napi_env env, cc_ctx* c, int rc) {
    char msg[512];
    snprintf(msg, sizeof msg, "crashcat_amd: %s (code %d)", c ? cc_last_error(c) : "no context", rc);
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static cc_ctx* get_ctx(napi_env env, napi_value v) {
    void* p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, NULL, "crashcat_amd: expected a context handle");
        return NULL;
    }
    return (cc_ctx*)p;
}

/* bytes of an ArrayBuffer, a typed array or a DataView */
static void* buffer_data(napi_env env, napi_value v, size_t* bytes) {
    bool is;
    void* data = NULL;
    *bytes = 0;
    if (napi_is_arraybuffer(env, v, &is) == napi_ok && is) {
        napi_get_arraybuffer_info(env, v, &data, bytes);
        return data;
    }
    if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
        napi_typedarray_type t;
        size_t len, off;
        napi_value ab;
        napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
        size_t es = 1;
        switch (t) {
            case napi_int16_array: case napi_uint16_array: es = 2; break;
            case napi_int32_array: case napi_uint32_array: case napi_float32_array: es = 4; break;
            case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: es = 8; break;
            default: es = 1;
        }
        *bytes = len * es;
        return data;
    }
    napi_throw_type_error(env, NULL, "crashcat_amd: expected an ArrayBuffer or a typed array");
    return NULL;
}

static void finalize_ctx(napi_env env, void* data, void* hint) {
    (void)env; (void)hint;
    if (data) cc_destroy((cc_ctx*)data);
}

/* ---- layout(): struct sizes and field offsets from the C header ---------------------- */
static void set_field(napi_env env, napi_value fields, const char* name, size_t off, const char* type, int count) {
    napi_value arr, v;
    napi_create_array_with_length(env, 3, &arr);
    napi_create_uint32(env, (uint32_t)off, &v); napi_set_element(env, arr, 0, v);
    napi_create_string_utf8(env, type, NAPI_AUTO_LENGTH, &v); napi_set_element(env, arr, 1, v);
    napi_create_uint32(env, (uint32_t)count, &v); napi_set_element(env, arr, 2, v);
    napi_set_named_property(env, fields, name, arr);
}
#define F(S, name, type, count) set_field(env, fields, #name, offsetof(S, name), type, count)
static napi_value add_struct(napi_env env, napi_value out, const char* sname, size_t size, napi_value fields) {
    napi_value o, v;
    napi_create_object(env, &o);
    napi_create_uint32(env, (uint32_t)size, &v);
    napi_set_named_property(env, o, "size", v);
    napi_set_named_property(env, o, "fields", fields);
    napi_set_named_property(env, out, sname, o);
    return o;
}
static napi_value layout(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value out, fields;
    CHECK(env, napi_create_object(env, &out));
    napi_create_object(env, &fields);
    F(cc_settings, gravity, "f64", 3); F(cc_settings, speculative_contact_distance, "f64", 1);
    F(cc_settings, manifold_tolerance, "f64", 1); F(cc_settings, collision_tolerance, "f64", 1);
    F(cc_settings, penetration_tolerance, "f64", 1); F(cc_settings, normal_cos_max_delta_rotation, "f64", 1);
    F(cc_settings, contact_point_preserve_lambda_max_dist_sq, "f64", 1);
    F(cc_settings, min_velocity_for_restitution, "f64", 1); F(cc_settings, penetration_slop, "f64", 1);
    F(cc_settings, baumgarte_factor, "f64", 1); F(cc_settings, max_penetration_distance, "f64", 1);
    F(cc_settings, time_before_sleep, "f64", 1); F(cc_settings, point_velocity_sleep_threshold, "f64", 1);
    F(cc_settings, velocity_iterations, "i32", 1); F(cc_settings, position_iterations, "i32", 1);
    F(cc_settings, gravity_enabled, "i32", 1); F(cc_settings, warm_starting, "i32", 1);
    F(cc_settings, allow_sleeping, "i32", 1); F(cc_settings, num_object_layers, "i32", 1);
    F(cc_settings, object_layer_pairs, "u8", CC_MAX_OBJECT_LAYERS * CC_MAX_OBJECT_LAYERS);
    F(cc_settings, collide_with_backfaces, "i32", 1); F(cc_settings, collide_only_with_active_edges, "i32", 1);
    F(cc_settings, ccd_linear_cast_threshold, "f64", 1); F(cc_settings, ccd_linear_cast_max_penetration, "f64", 1);
    F(cc_settings, ccd_min_velocity_for_restitution, "f64", 1);
    add_struct(env, out, "cc_settings", sizeof(cc_settings), fields);
    napi_create_object(env, &fields);
    F(cc_shape_desc, type, "i32", 1); F(cc_shape_desc, hull_first_point, "i32", 1);
    F(cc_shape_desc, hull_num_points, "i32", 1); F(cc_shape_desc, hull_first_face, "i32", 1);
    F(cc_shape_desc, hull_num_faces, "i32", 1); F(cc_shape_desc, radius, "f64", 1);
    F(cc_shape_desc, half_height, "f64", 1); F(cc_shape_desc, half_extents, "f64", 3);
    F(cc_shape_desc, convex_radius, "f64", 1); F(cc_shape_desc, center_of_mass, "f64", 3);
    F(cc_shape_desc, aabb_min, "f64", 3); F(cc_shape_desc, aabb_max, "f64", 3);
    F(cc_shape_desc, mesh_first_triangle, "i32", 1); F(cc_shape_desc, mesh_num_triangles, "i32", 1);
    F(cc_shape_desc, mesh_first_node, "i32", 1); F(cc_shape_desc, mesh_num_nodes, "i32", 1);
    F(cc_shape_desc, decorated, "i32", 1); F(cc_shape_desc, scale, "f64", 3);
    F(cc_shape_desc, local_position, "f64", 3); F(cc_shape_desc, local_quaternion, "f64", 4);
    F(cc_shape_desc, compound_first_child, "i32", 1); F(cc_shape_desc, compound_num_children, "i32", 1);
    add_struct(env, out, "cc_shape_desc", sizeof(cc_shape_desc), fields);
    napi_create_object(env, &fields);
    F(cc_body_desc, shape, "i32", 1); F(cc_body_desc, motion_type, "i32", 1);
    F(cc_body_desc, object_layer, "i32", 1); F(cc_body_desc, collision_groups, "u32", 1);
    F(cc_body_desc, collision_mask, "u32", 1); F(cc_body_desc, allow_sleeping, "i32", 1);
    F(cc_body_desc, allowed_dofs, "i32", 1); F(cc_body_desc, friction_combine, "i32", 1);
    F(cc_body_desc, restitution_combine, "i32", 1); F(cc_body_desc, sleeping, "i32", 1);
    F(cc_body_desc, position, "f64", 3); F(cc_body_desc, quaternion, "f64", 4);
    F(cc_body_desc, linear_velocity, "f64", 3); F(cc_body_desc, angular_velocity, "f64", 3);
    F(cc_body_desc, inv_mass, "f64", 1); F(cc_body_desc, inv_inertia_diagonal, "f64", 3);
    F(cc_body_desc, inertia_rotation, "f64", 4); F(cc_body_desc, friction, "f64", 1);
    F(cc_body_desc, restitution, "f64", 1); F(cc_body_desc, gravity_factor, "f64", 1);
    F(cc_body_desc, linear_damping, "f64", 1); F(cc_body_desc, angular_damping, "f64", 1);
    F(cc_body_desc, max_linear_velocity, "f64", 1); F(cc_body_desc, max_angular_velocity, "f64", 1);
    F(cc_body_desc, force, "f64", 3); F(cc_body_desc, torque, "f64", 3);
    F(cc_body_desc, num_velocity_steps_override, "i32", 1); F(cc_body_desc, num_position_steps_override, "i32", 1);
    F(cc_body_desc, sensor, "i32", 1); F(cc_body_desc, motion_quality, "i32", 1);
    add_struct(env, out, "cc_body_desc", sizeof(cc_body_desc), fields);
    napi_create_object(env, &fields);
    F(cc_body_state, position, "f64", 3); F(cc_body_state, quaternion, "f64", 4);
    F(cc_body_state, center_of_mass_position, "f64", 3); F(cc_body_state, linear_velocity, "f64", 3);
    F(cc_body_state, angular_velocity, "f64", 3); F(cc_body_state, aabb_min, "f64", 3);
    F(cc_body_state, aabb_max, "f64", 3); F(cc_body_state, sleeping, "i32", 1);
    F(cc_body_state, island, "i32", 1);
    add_struct(env, out, "cc_body_state", sizeof(cc_body_state), fields);
    napi_create_object(env, &fields);
    F(cc_point_constraint_desc, body_a, "i32", 1); F(cc_point_constraint_desc, body_b, "i32", 1);
    F(cc_point_constraint_desc, enabled, "i32", 1); F(cc_point_constraint_desc, local_point_a, "f64", 3);
    F(cc_point_constraint_desc, local_point_b, "f64", 3);
    F(cc_point_constraint_desc, num_velocity_steps_override, "i32", 1);
    F(cc_point_constraint_desc, num_position_steps_override, "i32", 1);
    add_struct(env, out, "cc_point_constraint_desc", sizeof(cc_point_constraint_desc), fields);
    napi_create_object(env, &fields);
    F(cc_step_counters, num_bodies, "i64", 1); F(cc_step_counters, num_awake, "i64", 1);
    F(cc_step_counters, num_pairs, "i64", 1); F(cc_step_counters, num_contacts, "i64", 1);
    F(cc_step_counters, num_points, "i64", 1); F(cc_step_counters, num_epa, "i64", 1);
    F(cc_step_counters, num_islands, "i64", 1); F(cc_step_counters, max_levels, "i64", 1);
    F(cc_step_counters, position_sweeps, "i64", 1); F(cc_step_counters, overflow, "i64", 1);
    add_struct(env, out, "cc_step_counters", sizeof(cc_step_counters), fields);
    napi_create_object(env, &fields);
    F(cc_contact_out, body_a, "i32", 1); F(cc_contact_out, body_b, "i32", 1);
    F(cc_contact_out, num_points, "i32", 1); F(cc_contact_out, persisted, "i32", 1);
    F(cc_contact_out, sub_shape_a, "u32", 1); F(cc_contact_out, sub_shape_b, "u32", 1);
    F(cc_contact_out, normal, "f64", 3); F(cc_contact_out, position_a, "f64", 12);
    F(cc_contact_out, position_b, "f64", 12); F(cc_contact_out, lambda_n, "f64", 4);
    F(cc_contact_out, lambda_t1, "f64", 4); F(cc_contact_out, lambda_t2, "f64", 4);
    add_struct(env, out, "cc_contact_out", sizeof(cc_contact_out), fields);
    napi_create_object(env, &fields);
    F(cc_contact_removed, body_a, "i32", 1); F(cc_contact_removed, body_b, "i32", 1);
    F(cc_contact_removed, sub_shape_a, "u32", 1); F(cc_contact_removed, sub_shape_b, "u32", 1);
    add_struct(env, out, "cc_contact_removed", sizeof(cc_contact_removed), fields);
    napi_create_object(env, &fields);
    F(cc_ccd_body, body_index, "i32", 1); F(cc_ccd_body, hit_body_index, "i32", 1);
    F(cc_ccd_body, delta_position, "f64", 3); F(cc_ccd_body, contact_normal, "f64", 3);
    F(cc_ccd_body, contact_point, "f64", 3); F(cc_ccd_body, fraction, "f64", 1);
    F(cc_ccd_body, fraction_plus_slop, "f64", 1); F(cc_ccd_body, linear_cast_threshold_sq, "f64", 1);
    F(cc_ccd_body, max_penetration, "f64", 1); F(cc_ccd_body, combined_friction, "f64", 1);
    F(cc_ccd_body, combined_restitution, "f64", 1);
    F(cc_ccd_body, sub_shape_id_1, "u32", 1); F(cc_ccd_body, sub_shape_id_2, "u32", 1);
    add_struct(env, out, "cc_ccd_body", sizeof(cc_ccd_body), fields);
    return out;
}
#undef F

/* ---- create(precision, maxPairsPerBody, maxContactsPerBody, device, solverOrder) -> handle - */
static napi_value create(napi_env env, napi_callback_info info) {
    size_t argc = 5;
    napi_value a[5];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    int32_t prec = CC_PRECISION_FP32, pairs = 24, cons = 8, dev = 0, order = CC_ORDER_REFERENCE;
    if (argc > 0) napi_get_value_int32(env, a[0], &prec);
    if (argc > 1) napi_get_value_int32(env, a[1], &pairs);
    if (argc > 2) napi_get_value_int32(env, a[2], &cons);
    if (argc > 3) napi_get_value_int32(env, a[3], &dev);
    if (argc > 4) napi_get_value_int32(env, a[4], &order);
    cc_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.precision = prec;
    cfg.device = dev;
    cfg.max_pairs_per_body = pairs;
    cfg.max_contacts_per_body = cons;
    cfg.solver_order = order;
    cc_ctx* c = NULL;
    int rc = cc_create(&c, &cfg);
    if (rc != CC_OK) {
        char msg[160];
        snprintf(msg, sizeof msg, "crashcat_amd: cc_create failed (code %d): %s", rc,
                 rc == CC_E_HIP ? "no usable HIP device" : "invalid configuration");
        if (c) cc_destroy(c);
        napi_throw_error(env, NULL, msg);
        return NULL;
    }
    napi_value h;
    CHECK(env, napi_create_external(env, c, finalize_ctx, NULL, &h));
    return h;
}

/* ---- upload(ctx, settings, offsets, bodies, shapes, hull, points) ----------------------- */
static napi_value upload(napi_env env, napi_callback_info info) {
    size_t argc = 7;
    napi_value a[7];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    if (argc < 7) { napi_throw_type_error(env, NULL, "crashcat_amd: upload needs 7 arguments"); return NULL; }
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t sb, ob, bb, shb, pb;
    const void* s = buffer_data(env, a[1], &sb);
    const int32_t* off = (const int32_t*)buffer_data(env, a[2], &ob);
    const void* bodies = buffer_data(env, a[3], &bb);
    const void* shapes = buffer_data(env, a[4], &shb);
    const void* pcs = buffer_data(env, a[6], &pb);
    if (!s || !off || (!bodies && bb) || (!shapes && shb)) return NULL;
    if (sb < sizeof(cc_settings) || ob < 2 * sizeof(int32_t) || bb % sizeof(cc_body_desc) || shb % sizeof(cc_shape_desc) ||
        pb % sizeof(cc_point_constraint_desc)) {
        napi_throw_range_error(env, NULL, "crashcat_amd: buffer sizes do not match the struct layouts");
        return NULL;
    }
    cc_hull_data hull;
    memset(&hull, 0, sizeof hull);
    napi_value h = a[5], v;
    bool has;
    /* bytes of each hull buffer (0 when absent): the totals below must fit in them */
    size_t bPts = 0, bShr = 0, bFfv = 0, bFnv = 0, bFvi = 0, bPn = 0, bPc = 0;
#define HULL(field, prop, bytes)                                                   \
    if (napi_has_named_property(env, h, prop, &has) == napi_ok && has) {           \
        napi_get_named_property(env, h, prop, &v);                                 \
        hull.field = buffer_data(env, v, &bytes);                                  \
        if (!hull.field) return NULL;                                              \
    }
    HULL(points, "points", bPts) HULL(shrunk_points, "shrunkPoints", bShr)
    HULL(face_first_vertex, "faceFirstVertex", bFfv) HULL(face_num_vertices, "faceNumVertices", bFnv)
    HULL(face_vertex_indices, "faceVertexIndices", bFvi) HULL(plane_normals, "planeNormals", bPn)
    HULL(plane_constants, "planeConstants", bPc)
#undef HULL
    if (napi_get_named_property(env, h, "totalPoints", &v) == napi_ok) napi_get_value_int32(env, v, &hull.total_points);
    if (napi_get_named_property(env, h, "totalFaces", &v) == napi_ok) napi_get_value_int32(env, v, &hull.total_faces);
    if (napi_get_named_property(env, h, "totalFaceVertices", &v) == napi_ok)
        napi_get_value_int32(env, v, &hull.total_face_vertices);
    {
        const size_t P = (size_t)(hull.total_points > 0 ? hull.total_points : 0);
        const size_t Fc = (size_t)(hull.total_faces > 0 ? hull.total_faces : 0);
        const size_t FV = (size_t)(hull.total_face_vertices > 0 ? hull.total_face_vertices : 0);
        if (hull.total_points < 0 || hull.total_faces < 0 || hull.total_face_vertices < 0 ||
            bPts < 3 * P * sizeof(double) || bShr < 3 * P * sizeof(double) || bFfv < Fc * sizeof(int32_t) ||
            bFnv < Fc * sizeof(int32_t) || bPn < 3 * Fc * sizeof(double) || bPc < Fc * sizeof(double) ||
            bFvi < FV * sizeof(int32_t)) {
            napi_throw_range_error(env, NULL, "crashcat_amd: hull totals exceed the hull buffers");
            return NULL;
        }
    }
    /* triangle meshes ride on the same object: meshPositions (Float64Array), meshTriangles
     * (Int32Array, 4 per triangle), meshNodes (Float64Array, the BVH node buffer) */
    cc_mesh_data mesh;
    memset(&mesh, 0, sizeof mesh);
    size_t mvb = 0, mtb = 0, mnb = 0;
    if (napi_has_named_property(env, h, "meshPositions", &has) == napi_ok && has) {
        napi_get_named_property(env, h, "meshPositions", &v);
        mesh.positions = (const double*)buffer_data(env, v, &mvb);
        napi_get_named_property(env, h, "meshTriangles", &v);
        mesh.triangles = (const int32_t*)buffer_data(env, v, &mtb);
        napi_get_named_property(env, h, "meshNodes", &v);
        mesh.nodes = (const double*)buffer_data(env, v, &mnb);
        if (!mesh.positions || !mesh.triangles || !mesh.nodes) return NULL;
        mesh.total_vertices = (int32_t)(mvb / (3 * sizeof(double)));
        mesh.total_triangles = (int32_t)(mtb / (4 * sizeof(int32_t)));
        mesh.total_nodes = (int32_t)(mnb / (8 * sizeof(double)));
    }
    int32_t nw = (int32_t)(ob / sizeof(int32_t)) - 1;
    int rc = cc_upload_scene(c, (const cc_settings*)s, nw, off, (const cc_body_desc*)bodies,
                             (int32_t)(bb / sizeof(cc_body_desc)), (const cc_shape_desc*)shapes,
                             (int32_t)(shb / sizeof(cc_shape_desc)), &hull, (const cc_point_constraint_desc*)pcs,
                             (int32_t)(pb / sizeof(cc_point_constraint_desc)), mesh.positions ? &mesh : NULL);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}

/* ---- step(ctx, dt, n): the updateWorld replacement; sync(ctx) --------------------------- */
static napi_value step(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value a[3];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    double dt = 1.0 / 60.0;
    int32_t n = 1;
    if (argc > 1) napi_get_value_double(env, a[1], &dt);
    if (argc > 2) napi_get_value_int32(env, a[2], &n);
    int rc = cc_step(c, dt, n);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}
static napi_value sync_(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value a[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    int rc = cc_sync(c);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}

/* ---- downloadBodies(ctx, out) / setBodies(ctx, in): cc_body_state arrays ------------------ */
static napi_value bodies_io(napi_env env, napi_callback_info info, int download) {
    size_t argc = 2;
    napi_value a[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t bytes;
    void* p = buffer_data(env, a[1], &bytes);
    if (!p) return NULL;
    if (bytes % sizeof(cc_body_state)) {
        napi_throw_range_error(env, NULL, "crashcat_amd: buffer is not a whole number of cc_body_state");
        return NULL;
    }
    int32_t n = (int32_t)(bytes / sizeof(cc_body_state));
    int rc = download ? cc_download_bodies(c, (cc_body_state*)p, n) : cc_set_bodies(c, (const cc_body_state*)p, n);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}
static napi_value download_bodies(napi_env env, napi_callback_info info) { return bodies_io(env, info, 1); }
static napi_value set_bodies(napi_env env, napi_callback_info info) { return bodies_io(env, info, 0); }

/* ---- counters(ctx, out) fills a cc_step_counters buffer ---------------------------------- */
static napi_value counters(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value a[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t bytes;
    void* p = buffer_data(env, a[1], &bytes);
    if (!p) return NULL;
    if (bytes < sizeof(cc_step_counters)) {
        napi_throw_range_error(env, NULL, "crashcat_amd: buffer smaller than cc_step_counters");
        return NULL;
    }
    int rc = cc_get_counters(c, (cc_step_counters*)p);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}

/* ---- downloadPairs(ctx) -> Int32Array of (a, b) ------------------------------------------ */
static napi_value download_pairs(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value a[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    int32_t count = 0;
    int rc = cc_download_pairs(c, NULL, 0, &count);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    void* data;
    napi_value ab, arr;
    CHECK(env, napi_create_arraybuffer(env, (size_t)count * 2 * sizeof(int32_t), &data, &ab));
    if (count > 0) {
        rc = cc_download_pairs(c, (int32_t*)data, count, &count);
        if (rc != CC_OK) return throw_rc(env, c, rc);
    }
    CHECK(env, napi_create_typedarray(env, napi_int32_array, (size_t)count * 2, ab, 0, &arr));
    return arr;
}

/* ---- downloadContacts(handle) / downloadRemovedContacts(handle) -> ArrayBuffer of records --- */
static napi_value download_records(napi_env env, napi_callback_info info, int removed) {
    size_t argc = 1;
    napi_value a[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    int32_t count = 0;
    int rc = removed ? cc_download_removed_contacts(c, NULL, 0, &count) : cc_download_contacts(c, NULL, 0, &count);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    const size_t rec = removed ? sizeof(cc_contact_removed) : sizeof(cc_contact_out);
    void* data;
    napi_value ab;
    CHECK(env, napi_create_arraybuffer(env, (size_t)count * rec, &data, &ab));
    if (count > 0) {
        int32_t got = 0;
        rc = removed ? cc_download_removed_contacts(c, (cc_contact_removed*)data, count, &got)
                     : cc_download_contacts(c, (cc_contact_out*)data, count, &got);
        if (rc != CC_OK) return throw_rc(env, c, rc);
    }
    return ab;
}
static napi_value download_contacts(napi_env env, napi_callback_info info) { return download_records(env, info, 0); }
static napi_value download_removed(napi_env env, napi_callback_info info) { return download_records(env, info, 1); }

/* ---- downloadCcdBodies(handle) -> ArrayBuffer of cc_ccd_body records (world.ccd.ccdBodies) --- */
static napi_value download_ccd(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value a[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    int32_t count = 0;
    int rc = cc_download_ccd_bodies(c, NULL, 0, &count);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    void* data;
    napi_value ab;
    CHECK(env, napi_create_arraybuffer(env, (size_t)count * sizeof(cc_ccd_body), &data, &ab));
    if (count > 0) {
        int32_t got = 0;
        rc = cc_download_ccd_bodies(c, (cc_ccd_body*)data, count, &got);
        if (rc != CC_OK) return throw_rc(env, c, rc);
    }
    return ab;
}

/* ---- addForces(ctx, force, torque, wake): Float64Array (3 per body) or null each ------- */
static napi_value add_forces(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value a[4];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    if (argc < 3) { napi_throw_type_error(env, NULL, "crashcat_amd: addForces needs 3 arguments"); return NULL; }
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    const double* fv[2] = {NULL, NULL};
    size_t by[2] = {0, 0};
    for (int k = 0; k < 2; k++) {
        napi_valuetype t;
        CHECK(env, napi_typeof(env, a[1 + k], &t));
        if (t == napi_null || t == napi_undefined) continue;
        fv[k] = (const double*)buffer_data(env, a[1 + k], &by[k]);
        if (!fv[k]) return NULL;
    }
    int32_t wake = 1;
    if (argc > 3) napi_get_value_int32(env, a[3], &wake);
    size_t n = fv[0] ? by[0] : by[1];
    if (n % (3 * sizeof(double)) || (fv[0] && fv[1] && by[0] != by[1])) {
        napi_throw_range_error(env, NULL, "crashcat_amd: force and torque need 3 doubles per body");
        return NULL;
    }
    int rc = cc_add_forces(c, fv[0], fv[1], (int32_t)(n / (3 * sizeof(double))), wake);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}

/* ---- fwd(ctx, dt, nsteps, checkpointEvery); bwd(ctx, dLdcom, dLdvel|null) -> ArrayBuffer --
 * bwd returns one ArrayBuffer of 10 * n doubles: dL/dv0 (3n), dL/dw0 (3n), dL/dcom0 (3n) and
 * dL/dfriction (n), the layout the JavaScript host slices */
static napi_value fwd(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value a[4];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    double dt = 1.0 / 60.0;
    int32_t n = 1, k = 10;
    if (argc > 1) napi_get_value_double(env, a[1], &dt);
    if (argc > 2) napi_get_value_int32(env, a[2], &n);
    if (argc > 3) napi_get_value_int32(env, a[3], &k);
    int rc = cc_fwd(c, dt, n, k);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}

static napi_value bwd(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value a[3];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    if (argc < 2) { napi_throw_type_error(env, NULL, "crashcat_amd: bwd needs a context and dL/dcom"); return NULL; }
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t gb = 0, vb = 0;
    const double* g = (const double*)buffer_data(env, a[1], &gb);
    if (!g) return NULL;
    const double* gv = NULL;
    if (argc > 2) {
        napi_valuetype t;
        CHECK(env, napi_typeof(env, a[2], &t));
        if (t != napi_null && t != napi_undefined) {
            gv = (const double*)buffer_data(env, a[2], &vb);
            if (!gv) return NULL;
        }
    }
    if (gb % (3 * sizeof(double)) || (gv && vb != gb)) {
        napi_throw_range_error(env, NULL, "crashcat_amd: dL/dcom and dL/dvel need 3 doubles per body");
        return NULL;
    }
    const size_t n = gb / (3 * sizeof(double));
    void* data;
    napi_value ab;
    CHECK(env, napi_create_arraybuffer(env, 10 * n * sizeof(double), &data, &ab));
    double* o = (double*)data;
    int rc = cc_bwd(c, g, gv, o, o + 3 * n, o + 6 * n, o + 9 * n);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return ab;
}

/* ---- snapshot(ctx) -> ArrayBuffer; restore(ctx, buffer) ---------------------------------- */
static napi_value snapshot(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value a[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    int64_t size = 0;
    int rc = cc_snapshot(c, NULL, 0, &size);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    void* data;
    napi_value ab;
    CHECK(env, napi_create_arraybuffer(env, (size_t)size, &data, &ab));
    rc = cc_snapshot(c, data, size, &size);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return ab;
}

static napi_value restore(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value a[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
    cc_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t bytes;
    void* p = buffer_data(env, a[1], &bytes);
    if (!p) return NULL;
    int rc = cc_restore(c, p, (int64_t)bytes);
    if (rc != CC_OK) return throw_rc(env, c, rc);
    return NULL;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor d[] = {
        {"layout", 0, layout, 0, 0, 0, napi_default, 0},
        {"create", 0, create, 0, 0, 0, napi_default, 0},
        {"upload", 0, upload, 0, 0, 0, napi_default, 0},
        {"step", 0, step, 0, 0, 0, napi_default, 0},
        {"sync", 0, sync_, 0, 0, 0, napi_default, 0},
        {"downloadBodies", 0, download_bodies, 0, 0, 0, napi_default, 0},
        {"setBodies", 0, set_bodies, 0, 0, 0, napi_default, 0},
        {"counters", 0, counters, 0, 0, 0, napi_default, 0},
        {"downloadPairs", 0, download_pairs, 0, 0, 0, napi_default, 0},
        {"downloadContacts", 0, download_contacts, 0, 0, 0, napi_default, 0},
        {"downloadRemovedContacts", 0, download_removed, 0, 0, 0, napi_default, 0},
        {"downloadCcdBodies", 0, download_ccd, 0, 0, 0, napi_default, 0},
        {"snapshot", 0, snapshot, 0, 0, 0, napi_default, 0},
        {"restore", 0, restore, 0, 0, 0, napi_default, 0},
        {"addForces", 0, add_forces, 0, 0, 0, napi_default, 0},
        {"fwd", 0, fwd, 0, 0, 0, napi_default, 0},
        {"bwd", 0, bwd, 0, 0, 0, napi_default, 0},
    };
    napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
