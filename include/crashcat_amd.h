/*
 * crashcat_amd — C ABI of the MI355X (gfx950) step hot path.
 *
 * The library replaces the body of crashcat's
 *     updateWorld(world, listener, timeStep)          src/update.ts:45
 * for worlds whose bodies use the convex shapes {sphere, box, capsule,
 * convex hull} and static triangle meshes (collided with boxes, capsules and
 * hulls), and whose listener has no callbacks.  Everything that
 * crosses this boundary is plain data: the host (JavaScript over N-API,
 * Python over ctypes, or C/C++) owns every buffer it passes in, calls copy
 * synchronously and keep no alias after they return; device memory is owned
 * by the context.  A context is used from one host thread at a time.
 *
 * All functions return 0 on success and a negative CC_E_* code on failure;
 * cc_last_error() then describes the failure (the host maps this to a
 * thrown Error, as crashcat's assert() does, src/utils/assert.ts:1-4).
 */
#ifndef CRASHCAT_AMD_H
#define CRASHCAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define CC_OK 0
#define CC_E_ARG (-1)       /* invalid argument / descriptor                   */
#define CC_E_HIP (-2)       /* HIP runtime failure                             */
#define CC_E_CAPACITY (-3)  /* a per-step capacity (pairs, contacts, EPA) ran out */
#define CC_E_STATE (-4)     /* call not valid in the current context state     */
#define CC_E_UNSUPPORTED (-5) /* shape/constraint type not handled on the GPU path */

/* ---- shape types: the dispatch matrix of src/shapes/shapes.ts:37-51 ----- */
#define CC_SHAPE_SPHERE 0
#define CC_SHAPE_BOX 1
#define CC_SHAPE_CAPSULE 2
#define CC_SHAPE_CONVEX_HULL 3
#define CC_SHAPE_TRIANGLE_MESH 4
#define CC_SHAPE_CYLINDER 5      /* src/shapes/cylinder.ts: radius, half_height, convex_radius; axis y */
#define CC_SHAPE_PLANE 6         /* src/shapes/plane.ts (static bodies only): half_extents = unit normal,
                                    radius = constant, half_height = half extent of the bounded form */
#define CC_SHAPE_COMPOUND 7      /* src/shapes/compound.ts: children are the shape descs
                                    [compound_first_child, + compound_num_children), each a convex
                                    leaf (scaled allowed) whose local transform (CC_DECO_TRANSFORMED)
                                    is the child's position and quaternion */

/* ---- motion types: src/body/motion-type.ts ------------------------------ */
#define CC_MOTION_STATIC 0
#define CC_MOTION_KINEMATIC 1
#define CC_MOTION_DYNAMIC 2

/* ---- motion quality: MotionQuality, src/body/motion-properties.ts:8-22 ---- */
#define CC_MOTION_QUALITY_DISCRETE 0
#define CC_MOTION_QUALITY_LINEAR_CAST 1

/* ---- material combine modes: src/constraints/combine-material.ts:1-8 ---- */
#define CC_COMBINE_AVERAGE 0
#define CC_COMBINE_MULTIPLY 1
#define CC_COMBINE_MIN 2
#define CC_COMBINE_MAX 3
#define CC_COMBINE_GEOMETRIC_MEAN 4

/* ---- precision of a context ------------------------------------------- */
#define CC_PRECISION_FP32 0 /* batched path: bulk FP32; GJK, EPA and inertia in FP64     */
#define CC_PRECISION_FP64 1 /* plain path: FP64 throughout                        */

#define CC_MAX_OBJECT_LAYERS 8

/* World settings; mirrors createWorldSettings() src/world-settings.ts:224-264. */
typedef struct cc_settings {
    double gravity[3];
    double speculative_contact_distance;    /* narrowphase.speculativeContactDistance (0.02) */
    double manifold_tolerance;              /* narrowphase.manifoldTolerance (1e-3)          */
    double collision_tolerance;             /* narrowphase.collisionTolerance (1e-4)         */
    double penetration_tolerance;           /* narrowphase.penetrationTolerance (1e-4)       */
    double normal_cos_max_delta_rotation;   /* cos(5 deg)                                     */
    double contact_point_preserve_lambda_max_dist_sq; /* contacts.* (1e-4)                     */
    double min_velocity_for_restitution;    /* solver.minVelocityForRestitution (1)           */
    double penetration_slop;                /* solver.penetrationSlop (0.02)                  */
    double baumgarte_factor;                /* solver.baumgarteFactor (0.2)                   */
    double max_penetration_distance;        /* solver.maxPenetrationDistance (0.2)            */
    double time_before_sleep;               /* sleeping.timeBeforeSleep (0.5)                 */
    double point_velocity_sleep_threshold;  /* sleeping.pointVelocitySleepThreshold (0.03)    */
    int32_t velocity_iterations;            /* solver.velocityIterations (10)                 */
    int32_t position_iterations;            /* solver.positionIterations (2)                  */
    int32_t gravity_enabled;
    int32_t warm_starting;
    int32_t allow_sleeping;
    int32_t num_object_layers;
    /* object_layer_pairs[a * CC_MAX_OBJECT_LAYERS + b] != 0 when object layers a and b
     * collide (src/layers.ts:88-112, enableCollision). */
    uint8_t object_layer_pairs[CC_MAX_OBJECT_LAYERS * CC_MAX_OBJECT_LAYERS];
    int32_t collide_with_backfaces;         /* narrowphase.collideWithBackfaces (false)       */
    int32_t collide_only_with_active_edges; /* narrowphase.collideOnlyWithActiveEdges (true)  */
    double ccd_linear_cast_threshold;       /* ccd.linearCastThreshold (0.05)                 */
    double ccd_linear_cast_max_penetration; /* ccd.linearCastMaxPenetration (0.25)            */
    double ccd_min_velocity_for_restitution; /* ccd.minVelocityForRestitution (1)             */
} cc_settings;

/* One shape; mirrors the fields of the Sphere/Box/Capsule/ConvexHull/TriangleMesh shape
 * records (src/shapes/sphere.ts:40-56, box.ts:42-52, capsule.ts:30-48, convex-hull.ts:37-62,
 * triangle-mesh.ts:88-95).  A triangle mesh's aabb is its BVH root's bounds and its centre
 * of mass is the origin (triangle-mesh.ts:114-147). */
#define CC_DECO_SCALED 1      /* the chain holds a scaled decorator: `scale` applies          */
#define CC_DECO_TRANSFORMED 2 /* ... a transformed decorator: the local transform applies     */
#define CC_DECO_OFFSET_COM 4  /* ... an offset centre of mass (already in center_of_mass)     */

typedef struct cc_shape_desc {
    int32_t type;               /* CC_SHAPE_* */
    int32_t hull_first_point;   /* hulls: offset into cc_hull_data.points / shrunk_points */
    int32_t hull_num_points;
    int32_t hull_first_face;    /* hulls: offset into face_first_vertex / face_num_vertices / planes */
    int32_t hull_num_faces;
    int32_t pad_;
    double radius;              /* sphere radius, capsule radius                    */
    double half_height;         /* capsule half height of the cylinder              */
    double half_extents[3];     /* box                                              */
    double convex_radius;       /* box, hull                                        */
    double center_of_mass[3];   /* shape.centerOfMass (hull: the hull's COM)        */
    double aabb_min[3];         /* shape.aabb (local bounds)                        */
    double aabb_max[3];
    int32_t mesh_first_triangle; /* meshes: offset into cc_mesh_data.triangles (in triangles) */
    int32_t mesh_num_triangles;
    int32_t mesh_first_node;     /* meshes: offset into cc_mesh_data.nodes (in nodes)          */
    int32_t mesh_num_nodes;
    /* Decorators flattened onto the leaf shape above (scaled.ts, transformed.ts,
     * offset-center-of-mass.ts; a decorator consumes no sub-shape id bits).  The leaf is collided
     * at the body pose composed with the local transform, p + rotate(q, local_position) and
     * q * local_quaternion (collideTransformedVsShape transformed.ts:370-420), with `scale`
     * applied to its support functions and supporting faces (sphere, capsule and cylinder use
     * |scale[0]|, scaled.ts:23-28; the product of nested scales).  The leaf's own fields above
     * stay unscaled, except that a scaled convex hull's shrunk_points are the scaled ones
     * (setConvexHullNoConvexSupportScaled convex-hull.ts:842-935).  center_of_mass, aabb_* and the
     * body's mass properties are the decorated shape's (an offset centre of mass moves only
     * those).  decorated: CC_DECO_* bits, 0 = scale 1 and no local transform.  Convex leaves
     * only, and not in a scene with a triangle mesh or a plane. */
    int32_t decorated;
    int32_t pad2_;
    double scale[3];
    double local_position[3];
    double local_quaternion[4];
    /* compounds (CC_SHAPE_COMPOUND): the children are shapes [first, first + num) of this array,
     * collided in index order with sub-shape ids pushIndex(child, num) (compound.ts:494-640);
     * center_of_mass and aabb_* are the compound's (compound.ts:60-150) */
    int32_t compound_first_child;
    int32_t compound_num_children;
} cc_shape_desc;

/* Flattened convex-hull data shared by all hull shapes.  Points are relative to the
 * hull's centre of mass (src/shapes/convex-hull.ts:228); shrunk_points are the
 * convex-radius-excluded support points (src/shapes/convex-hull.ts:732-786). */
typedef struct cc_hull_data {
    const double* points;              /* [3 * total_points]            */
    const double* shrunk_points;       /* [3 * total_points]            */
    const int32_t* face_first_vertex;  /* [total_faces], into face_vertex_indices (absolute) */
    const int32_t* face_num_vertices;  /* [total_faces]                 */
    const int32_t* face_vertex_indices;/* [total_face_vertices], indices relative to the hull's first point */
    const double* plane_normals;       /* [3 * total_faces]             */
    const double* plane_constants;     /* [total_faces]                 */
    int32_t total_points;
    int32_t total_faces;
    int32_t total_face_vertices;
    int32_t pad_;
} cc_hull_data;

/* Flattened triangle-mesh data shared by all mesh shapes: the TriangleMeshData and
 * TriangleMeshBVH a mesh shape holds after buildTriangleMesh
 * (src/shapes/utils/triangle-mesh-builder.ts:62-220).
 *   triangles: 4 ints per triangle — vertex a, b, c (absolute indices into positions) and
 *     the active-edge bits (bit 0 edge a-b, bit 1 b-c, bit 2 c-a; triangle-mesh-data.ts:4-14);
 *     a mesh's triangles are in BVH leaf order, so a leaf covers a contiguous range.
 *   nodes: 8 doubles per node, the reference's node buffer (triangle-mesh-bvh.ts:15-25):
 *     min xyz, max xyz, then for an internal node the right child's offset and the split
 *     axis, for a leaf the first triangle and -(count + 1).  Offsets count doubles from the
 *     mesh's first node (the left child of the node at offset o is at o + 8); triangle
 *     indices count from the mesh's first triangle. */
typedef struct cc_mesh_data {
    const double* positions;          /* [3 * total_vertices] */
    const int32_t* triangles;         /* [4 * total_triangles] */
    const double* nodes;              /* [8 * total_nodes]     */
    int32_t total_vertices;
    int32_t total_triangles;
    int32_t total_nodes;
    int32_t pad_;
} cc_mesh_data;

/* One body at upload; mirrors RigidBody + MotionProperties (src/body/rigid-body.ts:142-252,
 * src/body/motion-properties.ts:25-69).  Mass properties are already decomposed into
 * inv_mass / inv_inertia_diagonal / inertia_rotation (setMassProperties, :419-481). */
typedef struct cc_body_desc {
    int32_t shape;              /* index into the shape table                       */
    int32_t motion_type;        /* CC_MOTION_*                                      */
    int32_t object_layer;
    uint32_t collision_groups;
    uint32_t collision_mask;
    int32_t allow_sleeping;
    int32_t allowed_dofs;       /* 6-bit mask, bits 0-2 translation, 3-5 rotation   */
    int32_t friction_combine;   /* CC_COMBINE_*                                     */
    int32_t restitution_combine;
    int32_t sleeping;           /* initial sleep state                              */
    double position[3];
    double quaternion[4];       /* x, y, z, w                                       */
    double linear_velocity[3];
    double angular_velocity[3];
    double inv_mass;
    double inv_inertia_diagonal[3];
    double inertia_rotation[4];
    double friction;
    double restitution;
    double gravity_factor;
    double linear_damping;
    double angular_damping;
    double max_linear_velocity;
    double max_angular_velocity;
    /* accumulated force and torque (world space; motionProperties.force / .torque,
     * src/body/motion-properties.ts) that the next step integrates (src/update.ts:253-292) and
     * then clears (:234-238); cc_add_forces adds more between steps */
    double force[3];
    double torque[3];
    /* numVelocityStepsOverride / numPositionStepsOverride (src/body/motion-properties.ts:59):
     * 0 = the world default; an island solves max over its dynamic bodies and point constraints,
     * and at least the world default when any of them is 0 (src/islands.ts:344-386); at most 255 */
    int32_t num_velocity_steps_override;
    int32_t num_position_steps_override;
    /* body.sensor (src/body/rigid-body.ts:207-208) and motionProperties.motionQuality: sensors and
     * MotionQuality.LINEAR_CAST (CCD, src/update.ts:932-934) are not on the GPU path; a body with
     * either is rejected with CC_E_UNSUPPORTED rather than stepped differently */
    int32_t sensor;
    int32_t motion_quality;     /* CC_MOTION_QUALITY_* */
} cc_body_desc;

/* Body state read back after stepping. */
typedef struct cc_body_state {
    double position[3];
    double quaternion[4];
    double center_of_mass_position[3];
    double linear_velocity[3];
    double angular_velocity[3];
    double aabb_min[3];
    double aabb_max[3];
    int32_t sleeping;
    int32_t island;             /* island label (lowest body index of the island), -1 if none */
} cc_body_state;

/* Ball joint between two bodies (src/constraints/point-constraint.ts:40-80);
 * points are in each body's centre-of-mass space. */
typedef struct cc_point_constraint_desc {
    int32_t body_a;
    int32_t body_b;
    int32_t enabled;
    int32_t pad_;
    double local_point_a[3];
    double local_point_b[3];
    /* the constraint's iteration overrides (ConstraintBase numVelocityStepsOverride /
     * numPositionStepsOverride, read by getConstraintIterationOverrides, src/islands.ts:366-374);
     * 0 = world default; at most 255 */
    int32_t num_velocity_steps_override;
    int32_t num_position_steps_override;
} cc_point_constraint_desc;

/* One contact constraint of the last step, for parity taps
 * (ContactConstraint, src/constraints/contact-constraints.ts:31-95). */
typedef struct cc_contact_out {
    int32_t body_a;             /* lower id of the pair (src/constraints/contact-constraints.ts:448-453) */
    int32_t body_b;
    int32_t num_points;
    int32_t persisted;          /* 1: the step before had this (a, b, sub a, sub b) contact, the
                                   onContactPersisted case; 0: a new contact, onContactAdded
                                   (src/constraints/contact-constraints.ts:493-499) */
    uint32_t sub_shape_a;       /* sub-shape ids (src/body/sub-shape.ts); 0xffffffff when empty */
    uint32_t sub_shape_b;
    double normal[3];           /* world space, A -> B                              */
    double position_a[4][3];    /* world-space contact points on A and B at creation */
    double position_b[4][3];
    double lambda_n[4];         /* solved impulses (after the velocity iterations)  */
    double lambda_t1[4];
    double lambda_t2[4];
} cc_contact_out;

/* A contact of the step before the last that the last step did not make again: the
 * arguments of onContactRemoved (src/listener.ts:84-98), ids sorted body_a < body_b. */
typedef struct cc_contact_removed {
    int32_t body_a;
    int32_t body_b;
    uint32_t sub_shape_a;
    uint32_t sub_shape_b;
} cc_contact_removed;

/* One CCD record of the last step: CCDBody (src/ccd.ts:19-60) of a dynamic LINEAR_CAST body
 * whose displacement exceeded linearCastThreshold * innerRadius, as world.ccd.ccdBodies holds it
 * after updateWorld: in time-of-impact order (sortCCDBodies, src/update.ts:1607-1621) and after
 * resolveCCDContacts (a later collision invalidated by an earlier one has hit_body_index -1). */
typedef struct cc_ccd_body {
    int32_t body_index;
    int32_t hit_body_index;             /* -1: no hit, the body moved its full displacement */
    double delta_position[3];
    double contact_normal[3];
    double contact_point[3];
    double fraction;
    double fraction_plus_slop;
    double linear_cast_threshold_sq;
    double max_penetration;
    double combined_friction;
    double combined_restitution;
    uint32_t sub_shape_id_1;            /* the hit's sub-shape ids (EMPTY 0xffffffff on a leaf; */
    uint32_t sub_shape_id_2;            /* a compound child's pushed index); 0 without a hit   */
} cc_ccd_body;

/* One manifold of the contact cache the next step warm-starts from (CachedManifold in
 * world.contacts, src/contacts.ts:9-58; matched by (body pair, sub-shape ids), its points by
 * the lambda-preserve distance, contact-constraints.ts:651-686).  Points are in each body's
 * centre-of-mass frame (rotated into the body), body_a < body_b. */
typedef struct cc_cached_contact {
    int32_t body_a;
    int32_t body_b;
    uint32_t sub_shape_a;
    uint32_t sub_shape_b;
    int32_t num_points;         /* 1..4 */
    int32_t pad_;
    double local_a[4][3];
    double local_b[4][3];
    double lambda_n[4];
    double lambda_t1[4];
    double lambda_t2[4];
} cc_cached_contact;

/* A body's sleep test (motionProperties.sleepTestSpheres / sleepTestTimer,
 * src/body/motion-properties.ts:70-75, src/body/sleep.ts): three sphere centres and radii and the
 * time it has been at rest. */
typedef struct cc_body_sleep {
    double timer;
    double center[3][3];
    double radius[3];
} cc_body_sleep;

/* Per-step counters used by the byte model. */
typedef struct cc_step_counters {
    int64_t num_bodies;         /* dynamic bodies                                   */
    int64_t num_awake;
    int64_t num_pairs;
    int64_t num_contacts;       /* contact constraints (manifolds)                  */
    int64_t num_points;         /* contact points                                   */
    int64_t num_epa;            /* EPA invocations                                  */
    int64_t num_islands;
    int64_t max_levels;         /* deepest dependency-level schedule of any world   */
    int64_t position_sweeps;    /* position sweeps actually executed (max over worlds) */
    int64_t overflow;           /* non-zero if a capacity was exceeded              */
} cc_step_counters;

/* Running totals since upload (all worlds, all steps); feed the algorithmic byte model. */
typedef struct cc_totals {
    int64_t steps;
    int64_t pairs;              /* sum over steps of broadphase pairs                */
    int64_t contacts;           /* sum over steps of contact constraints             */
    int64_t points;             /* sum over steps of contact points                  */
    int64_t epa;                /* EPA invocations                                   */
    int64_t awake_body_steps;   /* sum over steps of awake dynamic bodies            */
    int64_t solver_bytes;       /* byte-model bytes of the velocity + position sweeps */
    int64_t gjk;                /* pairs handed to GJK (past the bounding-sphere filter) */
    int64_t body_steps;         /* sum over steps of all bodies, static ones included */
    int64_t reserved[7];
} cc_totals;

typedef struct cc_config {
    int32_t precision;          /* CC_PRECISION_*                                   */
    int32_t device;             /* HIP device ordinal                               */
    int32_t max_pairs_per_body; /* capacity factor for broadphase pairs (default 24) */
    int32_t max_contacts_per_body; /* capacity factor for contact constraints (default 12) */
    /* block sizes (64, 128, 256, 512 or 1024; 0 = autotuned: at the first cc_step after an
     * upload whose previous step made constraints, trial steps from a checkpoint of the state
     * time the candidates and the fastest is kept, cached per workload for the process) */
    int32_t block_solver;       /* threads per solver workgroup                     */
    int32_t block_bodies;       /* threads per workgroup of the one-thread-per-body kernels */
    int32_t island_groups;      /* most island groups a world is solved as (0 = by size, 1 = never split) */
    int32_t disable_graphs;     /* 0: each step phase is replayed as a captured hipGraph; 1: launched directly */
    /* CC_ORDER_REFERENCE: each island's constraints are solved in crashcat's order (bit-identical
     * to the FP64 oracle on the FP64 path); CC_ORDER_COLORED: in the colours of a graph colouring
     * (no two constraints of a colour share a dynamic body), far fewer sequential stages on large
     * islands, a different Gauss-Seidel order judged by its residual (cc_get_solver_residual) */
    int32_t solver_order;
    int32_t measure_residual;   /* 1: every step measures its velocity residual (one extra pass) */
    /* FP32 contexts: 0 (default) = GJK and EPA run in FP64 from the FP32 poses with FP64 shape
     * data, so their discrete choices (EPA's last face, GJK's closest points between parallel
     * features) are the FP64 path's; 1 = GJK and EPA in FP32 (EPA face normals, distances and
     * barycentrics still FP64).  Ignored by FP64 contexts. */
    int32_t narrowphase_fp32;
} cc_config;

#define CC_ORDER_REFERENCE 0
#define CC_ORDER_COLORED 1

typedef struct cc_ctx cc_ctx;

/* Lifecycle ---------------------------------------------------------------- */
/* replaces createWorld (src/world.ts:46-58) + device allocation */
int cc_create(cc_ctx** out, const cc_config* cfg);
int cc_destroy(cc_ctx* ctx);
const char* cc_last_error(const cc_ctx* ctx);

/* Upload a batch of independent worlds.  world_body_offsets has num_worlds+1
 * entries; bodies of world w are [offsets[w], offsets[w+1]) and keep creation order
 * (src/body/rigid-body.ts:398-407).  Point constraints refer to global body indices
 * and must join bodies of the same world.  Replaces rigidBody.create (src/body/rigid-body.ts:390-434)
 * + broadphase.addBody for all bodies at once.  meshes may be NULL when no shape is a
 * triangle mesh.  Bodies with a mesh shape must be static, and a scene with a mesh body
 * has no sphere bodies (the sphere-vs-mesh path, triangle-mesh.ts:1418-1780, is not on the
 * GPU): both are rejected with CC_E_UNSUPPORTED. */
int cc_upload_scene(cc_ctx* ctx, const cc_settings* settings,
                    int32_t num_worlds, const int32_t* world_body_offsets,
                    const cc_body_desc* bodies, int32_t num_bodies,
                    const cc_shape_desc* shapes, int32_t num_shapes,
                    const cc_hull_data* hulls,
                    const cc_point_constraint_desc* point_constraints, int32_t num_point_constraints,
                    const cc_mesh_data* meshes);

/* Build one triangle mesh's data from a triangle soup, on the host: weld vertices on a
 * 1e-8 grid, drop triangles whose edge cross product is shorter than degenerate_tolerance
 * and duplicates, mark active edges by the dihedral-angle cosine threshold, and build the
 * BVH with the CENTER split strategy and max_leaf_triangles per leaf; replaces
 * buildTriangleMesh (src/shapes/utils/triangle-mesh-builder.ts:62-220) with
 * triangleMeshBvh.build (src/shapes/utils/triangle-mesh-bvh.ts:368-576).  Outputs use the
 * cc_mesh_data layout with vertex indices from 0; capacities: out_positions 3 * num_vertices
 * doubles, out_triangles 4 * num_triangles ints, out_nodes 8 * (2 * num_triangles) doubles.
 * out_counts receives (vertices, triangles, nodes). */
int cc_build_triangle_mesh(const double* positions, int32_t num_vertices, const int32_t* indices,
                           int32_t num_triangles, double degenerate_tolerance, double active_edge_cos_threshold,
                           int32_t max_leaf_triangles, double* out_positions, int32_t* out_triangles,
                           double* out_nodes, int32_t* out_counts);

/* Advance every world by nsteps steps of dt; replaces updateWorld (src/update.ts:45-239).
 * Work is enqueued on the context's stream; call cc_sync to wait. */
int cc_step(cc_ctx* ctx, double dt, int32_t nsteps);
int cc_sync(cc_ctx* ctx);

/* Add force and torque to every dynamic body (world space, accumulated until the next step
 * integrates and clears them): addForce / addTorque (src/body/rigid-body.ts:840-869) for all
 * bodies at once.  force and torque are 3 * num_bodies doubles each (either may be NULL);
 * entries of non-dynamic bodies are ignored, as the reference ignores them.  With wake != 0 a
 * sleeping body that receives a non-zero force or torque is woken (sleep.ts:188-203). */
int cc_add_forces(cc_ctx* ctx, const double* force, const double* torque, int32_t num_bodies, int32_t wake);

/* Read back body state (all worlds, global body order). */
int cc_download_bodies(cc_ctx* ctx, cc_body_state* out, int32_t num_bodies);
/* Overwrite body poses/velocities (all worlds), e.g. to restart from an oracle state.  A
 * sleeping dynamic body whose `sleeping` field is 0 here is woken at once (sleep timer reset,
 * back in the active set), as wake (src/body/sleep.ts:188-203) does after addImpulse
 * (src/body/rigid-body.ts:887-892); `sleeping` = 1 never puts an awake body to sleep. */
int cc_set_bodies(cc_ctx* ctx, const cc_body_state* in, int32_t num_bodies);

/* Parity taps of the last step. */
/* Broadphase pairs as (a, b) global body indices, a < b, sorted ascending;
 * replaces broadphase.findCollidingPairs (src/broadphase/broadphase.ts:217-277). */
int cc_download_pairs(cc_ctx* ctx, int32_t* out_pairs, int32_t capacity, int32_t* out_count);
/* Contact constraints sorted by (body_a, body_b, sub_shape_a, sub_shape_b).  With `persisted`
 * they are the post-step form of onContactAdded / onContactPersisted (called from
 * addContactConstraint, src/constraints/contact-constraints.ts:493-499). */
int cc_download_contacts(cc_ctx* ctx, cc_contact_out* out, int32_t capacity, int32_t* out_count);
/* Contacts removed by the last step, sorted by (body_a, body_b, sub_shape_a, sub_shape_b):
 * the post-step form of onContactRemoved, which destroyContact (src/contacts.ts:297-307) calls
 * for every contact that destroyStaleContactsBetweenBodies (:404-430) and
 * destroyUnprocessedContacts (:531-555) find unprocessed.  *out_count is the number removed;
 * at most `capacity` are written.  Call cc_step with one step to see every step's events. */
int cc_download_removed_contacts(cc_ctx* ctx, cc_contact_removed* out, int32_t capacity, int32_t* out_count);
/* The CCD records of the last step (world.ccd.ccdBodies after updateWorld, src/update.ts:131-140),
 * world by world, each world's in time-of-impact order.  *out_count is their number; at most
 * `capacity` are written. */
int cc_download_ccd_bodies(cc_ctx* ctx, cc_ccd_body* out, int32_t capacity, int32_t* out_count);
/* cc_download_removed_contacts without any device allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory once they run, such as a
 * serving loop that reads every step's removed contacts (the plain entry point allocates and
 * frees a staging list per call, and a free synchronizes the whole device).  The staging list and
 * its counter live in one array per GPU, held until the process ends, of CC_NOALLOC_MAX_RECORDS
 * records: a previous step that kept more contacts is rejected with CC_E_CAPACITY.  Results are
 * those of cc_download_removed_contacts; with out != NULL and 0 < capacity < *out_count it fails
 * with CC_E_ARG (query the count with capacity 0 first).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (several host threads, several contexts) race. */
#define CC_NOALLOC_MAX_RECORDS (1 << 20)
int cc_download_removed_contacts_noalloc(cc_ctx* ctx, cc_contact_removed* out, int32_t capacity, int32_t* out_count);
int cc_get_counters(cc_ctx* ctx, cc_step_counters* out);

/* Broadphase pair set on caller-supplied AABBs (one world), without stepping; the
 * AABBs are used as given (float32 values widened to double on both sides). */
int cc_broadphase_aabbs(cc_ctx* ctx, const float* aabb_min, const float* aabb_max,
                        const int32_t* motion_type, const int32_t* sleeping, const int32_t* object_layer,
                        int32_t n, double margin, int32_t* out_pairs, int32_t capacity, int32_t* out_count);

/* Per-phase device time summed over every step of the last cc_step call (milliseconds,
 * HIP events on the context stream).  Phases, in order (12): acceleration integration
 * (k_accel), grid (cells and pairs with large bodies), cell pairs (k_bp_cells), pair layout,
 * pair filter and GJK list sort, GJK (k_np_gjk), EPA (k_epa), manifolds with faces of at most
 * 4 vertices (k_np_manifold, FP32: <float, 8>), the other manifolds with planes and triangle
 * meshes, islands, solver (k_world: levels, warm start, velocity and position sweeps,
 * integration, sleep), contact-cache hash. */
int cc_get_phase_times(cc_ctx* ctx, double* out_ms, int32_t capacity, int32_t* out_count);

/* World snapshot: the state one step hands to the next, as one binary blob in place of the
 * reference's JSON-serialisable world (README.md:2296-2298; warm-start state in world.contacts,
 * src/contacts.ts:9-58, and world.previousTimeStep, src/world.ts:38): every body's pose,
 * velocities, bounds and sleep state, the solver's body records, the point constraints with
 * their impulses, the last step's contact cache (keys, sub-shape ids, contact points and
 * impulses) and the previous time step.  cc_snapshot writes the blob when `capacity` is at
 * least *out_size (always set; call with capacity 0 for the size).  cc_restore loads one into
 * a context holding the same upload (same precision, bodies, shapes and point constraints;
 * CC_E_ARG otherwise): the steps that follow are bit-identical to those after the snapshot.
 * The parity taps (pairs, contacts, removed contacts) are empty until the next step. */
int cc_snapshot(cc_ctx* ctx, void* out, int64_t capacity, int64_t* out_size);

/* The state one step hands to the next, in plain records rather than the snapshot blob, so that a
 * host can bring it from elsewhere (another engine's world mid-run, the oracle): the contact
 * cache with the previous time step (world.contacts, world.previousTimeStep, src/world.ts:38;
 * it replaces the cache of the last step, which the next step reads), the bodies' sleep tests,
 * and the point constraints' accumulated impulses (3 per constraint, PointConstraintPart
 * totalLambda, src/constraints/point-constraint-part.ts).  Records are validated (CC_E_ARG:
 * bodies out of range or unordered, point counts outside 1..4, more records than the cache
 * holds, CC_E_CAPACITY).  cc_download_contact_cache returns the last step's cache, sorted by
 * (body_a, body_b, sub_shape_a, sub_shape_b), with the time step it was made with. */
int cc_set_contact_cache(cc_ctx* ctx, const cc_cached_contact* contacts, int32_t count, double previous_time_step);
int cc_download_contact_cache(cc_ctx* ctx, cc_cached_contact* out, int32_t capacity, int32_t* out_count,
                              double* out_previous_time_step);
int cc_set_sleep_state(cc_ctx* ctx, const cc_body_sleep* in, int32_t num_bodies);
int cc_set_point_impulses(cc_ctx* ctx, const double* lambda, int32_t num_point_constraints);

/* The velocity residual of the last step (needs cc_config.measure_residual): out[0] the sum over
 * every solved contact point of the impulse change a further velocity sweep would make (normal,
 * plus the friction pair after its disc projection), out[1] the sum of the normal impulses,
 * out[2] the largest single change, out[3] the number of points.  No counterpart in the
 * reference, which has no residual; it judges CC_ORDER_COLORED against CC_ORDER_REFERENCE. */
int cc_get_solver_residual(cc_ctx* ctx, double out[4]);

/* Differentiable stepping: a forward/backward pair over nsteps steps of dt, the adjoint of the
 * velocity-level contact solve and the integrator (reverse mode, per step in reverse order):
 * the friction-disc and normal-clamp projections with their active sets
 * (src/constraints/contact-constraints.ts:963-1053, axis-constraint-part.ts:578-707), the warm
 * start and the impulses it carries from step to step (:906-955, 651-686), the speculative
 * normal bias (:288-423), the acceleration and velocity integration (src/update.ts:245-319,
 * 896-976) and the combined friction (combine-material.ts:40-57).  Contact normals, lever arms,
 * effective masses and inverse inertias are held constant within a step, and the position
 * (Baumgarte) sweeps, the restitution bias and the velocity clamps are not differentiated.
 * cc_fwd steps like cc_step and keeps a checkpoint of the state every checkpoint_every steps;
 * cc_bwd recomputes each segment from its checkpoint with the solver recording, runs the
 * adjoint backwards through it, and leaves the context in the state cc_fwd left it in.  Worlds
 * must be solved unsplit (island_groups = 1, or at least 64 worlds) in crashcat's order
 * (CC_ORDER_REFERENCE), with sleeping off and no point constraints or iteration overrides
 * (CC_E_UNSUPPORTED otherwise).  There is no such
 * pair in the reference; it is the forward/backward of an autograd Function over updateWorld.
 * cc_bwd takes dL/d(centre of mass) of every body after the last step (3 * num_bodies doubles;
 * dl_dvel, the same for the final linear velocities, may be NULL) and writes dL/d(initial linear
 * and angular velocities), dL/d(initial centres of mass) and dL/d(friction coefficient of each
 * body) into the outputs that are not NULL.  It must follow cc_fwd with no cc_step, cc_set_bodies,
 * cc_add_forces or cc_restore in between (CC_E_STATE). */
int cc_fwd(cc_ctx* ctx, double dt, int32_t nsteps, int32_t checkpoint_every);
int cc_bwd(cc_ctx* ctx, const double* dl_dcom, const double* dl_dvel, double* dl_dv0, double* dl_dw0,
           double* dl_dcom0, double* dl_dfriction);
int cc_restore(cc_ctx* ctx, const void* in, int64_t size);
int cc_get_totals(cc_ctx* ctx, cc_totals* out);

#ifdef __cplusplus
}
#endif
#endif /* CRASHCAT_AMD_H */
