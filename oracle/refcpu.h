/*
 * refcpu — FP64 CPU restatement of crashcat's per-step collision-and-contact
 * hot path.  TEST INFRASTRUCTURE (the oracle): only tests/, the smoke check and
 * bench.py's CPU-baseline leg load it, and only as the results the GPU path is
 * compared with or as the timed CPU baseline; the product library never links or
 * calls it.
 *
 * It follows the reference files op for op (citations beside each function in
 * refcpu_collide.c / refcpu_step.c) in double precision, one world on one
 * thread (its scratch is thread-local, so separate worlds may step on separate
 * threads at once), with crashcat's own orderings: active list in wake/creation order, contact
 * constraints in pair order, warm start in creation order, islands numbered by
 * lowest active index and sorted by size, contacts key-sorted per island.
 * Deliberate differences, none of which changes a pair set or an island
 * partition: broadphase pairs are enumerated brute force (the pair set of the
 * dynamic AABB tree is tree-independent, src/broadphase/dbvt.ts:620-681), in
 * (querying body, other body) index order instead of tree order; body ids equal
 * body indices (no removals, so id order == index order).
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>
#include "../include/crashcat_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_world oc_world;

/* world / step */
oc_world* oc_create(const cc_settings* s, const cc_body_desc* bodies, int32_t nb,
                    const cc_shape_desc* shapes, int32_t ns, const cc_hull_data* hulls,
                    const cc_point_constraint_desc* pcs, int32_t npc, const cc_mesh_data* meshes);
void oc_destroy(oc_world* w);
void oc_step(oc_world* w, double dt);
int32_t oc_num_bodies(const oc_world* w);
void oc_get_bodies(const oc_world* w, cc_body_state* out);
void oc_set_bodies(oc_world* w, const cc_body_state* in);
/* addForce / addTorque for every body (3 doubles each; NULL for none); wake: wake a sleeping
 * body that receives a non-zero entry */
void oc_add_forces(oc_world* w, const double* force, const double* torque, int32_t wake);
/* replace the world settings between steps (tests isolate one solver stage this way) */
void oc_set_settings(oc_world* w, const cc_settings* s);
/* the island of a body after the last step: its velocity and position iteration counts
 * (island.numVelocitySteps / numPositionSteps, src/islands.ts:344-386); 0 when not in an island */
int32_t oc_island_iterations(const oc_world* w, int32_t body, int32_t* vel, int32_t* pos);
/* computeContactSortKey (src/constraints/contact-constraints.ts:1222-1231) */
uint32_t oc_sort_key(uint32_t a, uint32_t b, uint32_t sub_a, uint32_t sub_b);
/* the contact tangent basis of addContactConstraint (:536-556) */
void oc_contact_tangents(const double* normal, double* t1, double* t2);
/* pairs of the last step as (min,max) index pairs, sorted */
int32_t oc_get_pairs(const oc_world* w, int32_t* out, int32_t cap);
/* contact constraints of the last step, sorted by (a, b, sub-shape a, sub-shape b) */
int32_t oc_get_contacts(const oc_world* w, cc_contact_out* out, int32_t cap);
/* contacts the last step removed (destroyUnprocessedContacts), sorted like the contacts */
int32_t oc_get_removed_contacts(const oc_world* w, cc_contact_removed* out, int32_t cap);
/* the last step's CCD records in time-of-impact order (world.ccd.ccdBodies); returns their number */
int32_t oc_get_ccd_bodies(const oc_world* w, cc_ccd_body* out, int32_t cap);
/* the state the next step starts from, as the C ABI's plain records: the contact cache sorted by
 * (body_a, body_b, sub ids) (returns the count; with the previous time step), the bodies' sleep
 * tests, the point constraints' accumulated impulses (3 each) */
int32_t oc_get_contact_cache(const oc_world* w, cc_cached_contact* out, int32_t cap, double* prev_dt);
void oc_get_sleep_state(const oc_world* w, cc_body_sleep* out);
void oc_get_point_impulses(const oc_world* w, double* out);
void oc_get_counters(const oc_world* w, cc_step_counters* out);
/* broadphase only, on given AABBs (widened from float) */
int32_t oc_broadphase_aabbs(const float* mn, const float* mx, const int32_t* motion, const int32_t* sleeping,
                            const int32_t* layer, int32_t n, double margin, const cc_settings* s,
                            int32_t* out, int32_t cap);

/* unit entry points used by the known-answer tests */
typedef struct oc_hit {
    int32_t hit;
    int32_t status;          /* 0 GJK-colliding, 1 EPA, -1 none */
    double point_a[3], point_b[3], axis[3], penetration;
    int32_t face_a_n, face_b_n;
    double face_a[32 * 3], face_b[32 * 3];
} oc_hit;
/* collide two shapes at world poses (collideShapeVsShape, src/collision/narrowphase.ts:172) */
void oc_collide(const cc_shape_desc* shapes, const cc_hull_data* hulls, int32_t sa, const double* pa, const double* qa,
                int32_t sb, const double* pb, const double* qb, double max_separation, const cc_settings* s,
                oc_hit* out);
/* the convex GJK/EPA path for any pair (spheres included) */
void oc_collide_convex(const cc_shape_desc* shapes, const cc_hull_data* hulls, int32_t sa, const double* pa,
                       const double* qa, int32_t sb, const double* pb, const double* qb, double max_separation,
                       const cc_settings* s, oc_hit* out);
/* collideConvexVsTriangleMesh (src/shapes/triangle-mesh.ts:1026-1379), or collideSphereVsTriangleMesh
 * (:1418-1720) when sa is a sphere: every hit of shape sa against mesh shape sm, in traversal order, with the triangle's sub-shape id in sub_b[i];
 * s supplies the tolerances and the back-face / active-edge switches; returns the hit count */
int32_t oc_collide_mesh(const cc_shape_desc* shapes, const cc_hull_data* hulls, const cc_mesh_data* meshes, int32_t sa,
                        const double* pa, const double* qa, int32_t sm, const double* pm, const double* qm,
                        double max_separation, const double* move_dir, const cc_settings* s, oc_hit* out,
                        uint32_t* sub_b, int32_t cap);
/* getClosestPointOnTriangle (src/collision/triangle.ts:62-286): closest point of triangle abc to
 * the origin and its squared distance; returns the TriangleFeature bits */
int32_t oc_closest_on_triangle(const double* a, const double* b, const double* c, double* point, double* dist2);
/* manifold of one hit, pruned to <= 4 points (update.ts:385-563 for a single hit);
 * out_a/out_b are world points; returns the point count */
int32_t oc_manifold_from_hit(const oc_hit* hit, double max_contact_distance, double* normal_out,
                             double* out_a, double* out_b);
/* clipping primitives (src/manifold/clip.ts) */
int32_t oc_clip_poly_vs_poly(const double* poly, int32_t n, const double* clip, int32_t m, const double* normal,
                             double* out);
int32_t oc_clip_poly_vs_edge(const double* poly, int32_t n, const double* e1, const double* e2, const double* normal,
                             double* out);
int32_t oc_clip_poly_vs_plane(const double* poly, int32_t n, const double* origin, const double* normal, double* out);
int32_t oc_manifold_between_two_faces(const double* p1, const double* p2, const double* axis, double max_dist,
                                      const double* face1, int32_t n1, const double* face2, int32_t n2,
                                      double* out_a, double* out_b);
/* prune >4 points to <=4 (src/manifold/manifold.ts:192-313); points relative to base offset */
int32_t oc_prune(double* rel_a, double* rel_b, int32_t n, const double* axis);
/* barycentric coordinates (src/collision/closest-points.ts) */
void oc_barycentric2(const double* a, const double* b, double tol, double* uvw_valid);
void oc_barycentric3(const double* a, const double* b, const double* c, double tol, double* uvw_valid);
/* world inverse inertia (src/body/motion-properties.ts:494-621), column-major 3x3 */
void oc_inverse_inertia(const double* inv_diag, const double* inertia_rot, const double* body_quat, int32_t dofs,
                        double* out9);
/* supporting face of a shape in local direction, transformed to world (shapes.ts:489) */
int32_t oc_supporting_face(const cc_shape_desc* shapes, const cc_hull_data* hulls, int32_t s, const double* dir,
                           const double* pos, const double* quat, double* out);

#ifdef __cplusplus
}
#endif
#endif
