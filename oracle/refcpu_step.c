/*
 * refcpu_step.c — FP64 restatement of crashcat's updateWorld (src/update.ts:45-239)
 * around the narrowphase in refcpu_collide.c: acceleration integration,
 * broadphase pair predicate, manifold reduction collector, persistent contact
 * cache, contact-constraint setup, islands, warm start, sequential-impulse
 * velocity and position solves, integration and sleeping.  Oracle (test
 * infrastructure); citations are to the reference's files.
 */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include "refcpu_internal.h"

#define INACTIVE INT32_MAX /* INACTIVE_BODY_INDEX (src/body/sleep.ts:8) */
#define EMPTY_SUB 0xffffffffu

typedef struct {
    int shape, motion, layer;
    uint32_t groups, mask;
    int allowSleep, dof, fmode, rmode;
    V3 pos, com, v, w, force, torque;
    Q4 q;
    double invMass;
    V3 invI;
    Q4 inertiaRot;
    double friction, restitution, gravityFactor, linDamp, angDamp, maxLinV, maxAngV;
    int sleeping;
    int ovV, ovP; /* numVelocityStepsOverride / numPositionStepsOverride (motion-properties.ts:59) */
    int activeIndex;
    int islandIndex;
    V3 aabbMin, aabbMax;
    V3 sleepCenter[3];
    double sleepRadius[3];
    double sleepTimer;
    int quality;   /* motionProperties.motionQuality */
    int ccdIndex;  /* body.ccdBodyIndex: the body's CCD record this step, or -1 */
} OBody;

typedef struct { V3 p1, p2; double ln, lt1, lt2; } OCached;
typedef struct {
    int a, b;
    uint32_t sa, sb; /* sub-shape ids */
    int numPoints;
    V3 normalB;
    OCached pts[4];
    int processed;
    int persisted;
} OContact;

typedef struct { V3 r1x, r2x, iI1r1x, iI2r2x; double effMass, lambda, bias; } OAxis;
typedef struct { V3 posA, posB, localA, localB; OAxis n, t1, t2; } OCPoint;
typedef struct {
    int a, b, contact;
    uint32_t subA, subB;
    V3 normal, t1, t2;
    double friction, restitution, invMassA, invMassB;
    M3 invIA, invIB;
    int np;
    OCPoint pts[4];
    uint32_t sortKey;
    int persisted; /* the contact existed before this step (onContactPersisted, else onContactAdded) */
} OCC;

typedef struct {
    int a, b, enabled, sleeping;
    V3 localA, localB;
    V3 r1, r2;
    M3 iI1r1X, iI2r2X;
    double eff[16]; /* mat4 (3x3 part used, [15] marks active) */
    V3 lambda;
    int ovV, ovP; /* the constraint's iteration overrides */
} OPoint;

/* CCDBody (src/ccd.ts:19-60) */
typedef struct {
    int body, hit;
    uint32_t sub1, sub2; /* subShapeId1 / subShapeId2: the hit's sub-shape ids */
    V3 delta, normal, point;
    double fraction, fps, thrSq, maxPen, friction, restitution;
} OCcd;

typedef struct {
    int* bodies; int nb, cb;      /* active indices */
    int* contacts; int nc, cc;    /* contact constraint indices */
    int* cons; int nu, cu;        /* user constraint indices */
    int velSteps, posSteps;
} OIsland;

struct oc_world {
    cc_settings s;
    OShapes sh;
    cc_shape_desc* shapes;
    int ns;
    cc_hull_data hulls;
    double* hullBuf;
    int32_t* hullIdx;
    cc_mesh_data meshes;
    double* meshBuf;
    int32_t* meshIdx;
    OBody* b;
    int nb;
    int* active;
    int nactive;
    double prevDt;
    /* pairs */
    int* pairs; int npairs, cpairs;
    /* contacts cache */
    OContact* ct; int nct, cct;
    int* htab; int hcap;
    /* contacts removed by the last step: (a, b, subA, subB) each */
    cc_contact_removed* rem; int nrem, crem;
    /* constraints */
    OCC* cc; int ncc, ccc;
    /* user constraints */
    OPoint* pc; int npc;
    /* islands */
    int* bodyLinks; int* bodyIslands; int* contactLinks; int* consLinks; int* consIds; int nconsIds;
    OIsland* isl; int nisl, cisl;
    /* CCD: this step's records in body order, their time-of-impact order, and the two constraint
     * parts applyCCD reuses from one call to the next (module scope in update.ts:1453-1454) */
    OCcd* ccd; int nccd, cccd;
    int* ccdOrder;
    OAxis ccdN, ccdF;
    /* counters */
    cc_step_counters ctr;
};

/* qsort of n elements; a zero-length (possibly null) array is left alone, as the C library
 * requires a valid pointer even for n = 0 */
static void sort_some(void* base, size_t n, size_t size, int (*cmp)(const void*, const void*)) {
    if (n > 1) qsort(base, n, size, cmp);
}

static void* xrealloc(void* p, size_t n) {
    void* r = realloc(p, n ? n : 1);
    if (!r) { fprintf(stderr, "refcpu: out of memory\n"); abort(); }
    return r;
}

/* ---- body helpers (src/body/rigid-body.ts, motion-properties.ts, sleep.ts) ---- */
static const cc_shape_desc* body_shape(const oc_world* w, const OBody* b) { return &w->shapes[b->shape]; }

static void update_com(oc_world* w, OBody* b) { /* updateCenterOfMassPosition :527-536 */
    b->com = vadd(qrotate(b->q, vload(body_shape(w, b)->center_of_mass)), b->pos);
}
static void pos_from_com(oc_world* w, OBody* b) { /* updatePositionFromCenterOfMass :545-553 */
    b->pos = vsub(b->com, qrotate(b->q, vload(body_shape(w, b)->center_of_mass)));
}
static void update_aabb(oc_world* w, OBody* b) { /* updateAABB :561-674 */
    const cc_shape_desc* s = body_shape(w, b);
    M3 R = m3_from_quat(b->q);
    double mn[3] = {s->aabb_min[0], s->aabb_min[1], s->aabb_min[2]};
    double mx[3] = {s->aabb_max[0], s->aabb_max[1], s->aabb_max[2]};
    V3 lo = v3(0, 0, 0), hi = v3(0, 0, 0);
    for (int c = 0; c < 8; c++) {
        V3 p = v3((c & 1) ? mx[0] : mn[0], (c & 2) ? mx[1] : mn[1], (c & 4) ? mx[2] : mn[2]);
        V3 wp = vadd(m3_mul(&R, p), b->pos);
        if (c == 0) { lo = wp; hi = wp; }
        else {
            lo = v3(dmin(lo.x, wp.x), dmin(lo.y, wp.y), dmin(lo.z, wp.z));
            hi = v3(dmax(hi.x, wp.x), dmax(hi.y, wp.y), dmax(hi.z, wp.z));
        }
    }
    b->aabbMin = lo;
    b->aabbMax = hi;
}
static void set_transform(oc_world* w, OBody* b) { /* setTransform :758-774 (broadphase update is implicit) */
    update_com(w, b);
    update_aabb(w, b);
}

/* getInverseInertiaForRotation (src/body/motion-properties.ts:494-621) */
static M3 inv_inertia(V3 invDiag, Q4 iq, const M3* bodyR, int dof) {
    M3 I = m3_from_quat(iq);
    const double* b = bodyR->m;
    const double* i = I.m;
    double r[9];
    for (int c = 0; c < 3; c++)
        for (int row = 0; row < 3; row++)
            r[c * 3 + row] = b[row] * i[c * 3 + 0] + b[3 + row] * i[c * 3 + 1] + b[6 + row] * i[c * 3 + 2];
    double d[3] = {invDiag.x, invDiag.y, invDiag.z};
    double s[9];
    for (int c = 0; c < 3; c++)
        for (int row = 0; row < 3; row++) s[c * 3 + row] = d[c] * r[c * 3 + row];
    M3 o;
    /* out(col j, row k) = sum_c s[c][j] * r[c][k] */
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) o.m[j * 3 + k] = s[0 * 3 + j] * r[0 * 3 + k] + s[1 * 3 + j] * r[1 * 3 + k] + s[2 * 3 + j] * r[2 * 3 + k];
    int rot = (dof >> 3) & 7;
    if (rot != 7) {
        double m[3] = {(rot & 1) ? 1.0 : 0.0, (rot & 2) ? 1.0 : 0.0, (rot & 4) ? 1.0 : 0.0};
        for (int c = 0; c < 3; c++)
            for (int row = 0; row < 3; row++) o.m[c * 3 + row] *= m[row] * m[c];
    }
    return o;
}

void oc_inverse_inertia(const double* inv_diag, const double* inertia_rot, const double* body_quat, int32_t dofs,
                        double* out9) {
    V3 d = {inv_diag[0], inv_diag[1], inv_diag[2]};
    Q4 iq = {inertia_rot[0], inertia_rot[1], inertia_rot[2], inertia_rot[3]};
    Q4 bq = {body_quat[0], body_quat[1], body_quat[2], body_quat[3]};
    M3 R = m3_from_quat(bq);
    M3 o = inv_inertia(d, iq, &R, dofs);
    for (int k = 0; k < 9; k++) out9[k] = o.m[k];
}

static void sleep_points(oc_world* w, OBody* b, V3 out[3]) { /* getSleepTestPoints src/body/sleep.ts:22-76 */
    const cc_shape_desc* s = body_shape(w, b);
    out[0] = b->com;
    double x = (s->aabb_max[0] - s->aabb_min[0]) * 0.5, y = (s->aabb_max[1] - s->aabb_min[1]) * 0.5,
           z = (s->aabb_max[2] - s->aabb_min[2]) * 0.5;
    int low = (x <= y && x <= z) ? 0 : (y <= z ? 1 : 2);
    V3 ax = v3(1, 0, 0), ay = v3(0, 1, 0), az = v3(0, 0, 1);
    if (low == 0) { out[1] = vadd(b->com, vscale(qrotate(b->q, ay), y)); out[2] = vadd(b->com, vscale(qrotate(b->q, az), z)); }
    else if (low == 1) { out[1] = vadd(b->com, vscale(qrotate(b->q, ax), x)); out[2] = vadd(b->com, vscale(qrotate(b->q, az), z)); }
    else { out[1] = vadd(b->com, vscale(qrotate(b->q, ax), x)); out[2] = vadd(b->com, vscale(qrotate(b->q, ay), y)); }
}
static void reset_sleep_spheres(OBody* b, V3 p[3]) {
    for (int i = 0; i < 3; i++) { b->sleepCenter[i] = p[i]; b->sleepRadius[i] = 0; }
    b->sleepTimer = 0;
}
static void reset_sleep_timer(oc_world* w, OBody* b) { /* resetSleepTimer :120-127 */
    if (b->motion != CC_MOTION_DYNAMIC) return;
    V3 p[3];
    sleep_points(w, b, p);
    reset_sleep_spheres(b, p);
}
static int update_sleep_state(oc_world* w, OBody* b, double dt, double maxMove, double tBefore) { /* :88-117 */
    if (!b->allowSleep) return 0;
    V3 p[3];
    sleep_points(w, b, p);
    for (int i = 0; i < 3; i++) {
        double d = vlen(vsub(b->sleepCenter[i], p[i]));
        b->sleepRadius[i] = dmax(b->sleepRadius[i], d);
        if (b->sleepRadius[i] > maxMove) { reset_sleep_spheres(b, p); return 0; }
    }
    b->sleepTimer += dt;
    return b->sleepTimer >= tBefore;
}
static void add_active(oc_world* w, OBody* b, int idx) { /* addBodyToActiveBodies :132-142 */
    if (b->activeIndex != INACTIVE) return;
    b->activeIndex = w->nactive;
    w->active[w->nactive++] = idx;
}
static void remove_active(oc_world* w, OBody* b) { /* removeBodyFromActiveBodies :145-166 */
    if (b->activeIndex == INACTIVE) return;
    int last = w->nactive - 1;
    if (b->activeIndex != last) {
        int lastBody = w->active[last];
        w->active[b->activeIndex] = lastBody;
        w->b[lastBody].activeIndex = b->activeIndex;
    }
    b->activeIndex = INACTIVE;
    w->nactive--;
}
static void body_sleep(oc_world* w, OBody* b) { /* sleep :169-185 */
    if (b->motion == CC_MOTION_STATIC || b->sleeping) return;
    remove_active(w, b);
    b->sleeping = 1;
    b->v = v3(0, 0, 0);
    b->w = v3(0, 0, 0);
}
static void body_wake(oc_world* w, OBody* b, int idx) { /* wake :188-203 */
    if (b->motion == CC_MOTION_STATIC) return;
    reset_sleep_timer(w, b);
    if (!b->sleeping) return;
    b->sleeping = 0;
    add_active(w, b, idx);
}
static V3 dof_lin(V3 v, int dof) {
    if (!(dof & 1)) v.x = 0;
    if (!(dof & 2)) v.y = 0;
    if (!(dof & 4)) v.z = 0;
    return v;
}
static void clamp_lin(OBody* b) { /* clampLinearVelocity :624-630 */
    if (vlen2(b->v) > b->maxLinV * b->maxLinV) b->v = vscale(vnormalize(b->v), b->maxLinV);
}
static void clamp_ang(OBody* b) {
    if (vlen2(b->w) > b->maxAngV * b->maxAngV) b->w = vscale(vnormalize(b->w), b->maxAngV);
}
static void add_rot_step(OBody* b, V3 s, double sign) { /* add/subRotationStep src/body/rigid-body-step.ts:55-97 */
    double len = vlen(s);
    if (len > 1.0e-6) {
        V3 axis = vscale(s, 1 / len);
        Q4 r = qaxis_angle(axis, sign * len);
        b->q = qnormalize(qmul(r, b->q));
    }
}

/* ---- world creation -------------------------------------------------------- */
oc_world* oc_create(const cc_settings* s, const cc_body_desc* bodies, int32_t nb, const cc_shape_desc* shapes,
                    int32_t ns, const cc_hull_data* hulls, const cc_point_constraint_desc* pcs, int32_t npc,
                    const cc_mesh_data* meshes) {
    oc_world* w = (oc_world*)calloc(1, sizeof(oc_world));
    w->s = *s;
    w->shapes = (cc_shape_desc*)xrealloc(NULL, sizeof(cc_shape_desc) * (size_t)ns);
    memcpy(w->shapes, shapes, sizeof(cc_shape_desc) * (size_t)ns);
    w->ns = ns;
    if (hulls && hulls->total_points > 0) {
        int P = hulls->total_points, F = hulls->total_faces, FV = hulls->total_face_vertices;
        w->hullBuf = (double*)xrealloc(NULL, sizeof(double) * (size_t)(6 * P + 4 * F));
        double* p = w->hullBuf;
        memcpy(p, hulls->points, sizeof(double) * 3 * P); w->hulls.points = p; p += 3 * P;
        memcpy(p, hulls->shrunk_points, sizeof(double) * 3 * P); w->hulls.shrunk_points = p; p += 3 * P;
        memcpy(p, hulls->plane_normals, sizeof(double) * 3 * F); w->hulls.plane_normals = p; p += 3 * F;
        memcpy(p, hulls->plane_constants, sizeof(double) * F); w->hulls.plane_constants = p;
        w->hullIdx = (int32_t*)xrealloc(NULL, sizeof(int32_t) * (size_t)(2 * F + FV));
        int32_t* q = w->hullIdx;
        memcpy(q, hulls->face_first_vertex, sizeof(int32_t) * F); w->hulls.face_first_vertex = q; q += F;
        memcpy(q, hulls->face_num_vertices, sizeof(int32_t) * F); w->hulls.face_num_vertices = q; q += F;
        memcpy(q, hulls->face_vertex_indices, sizeof(int32_t) * FV); w->hulls.face_vertex_indices = q;
        w->hulls.total_points = P; w->hulls.total_faces = F; w->hulls.total_face_vertices = FV;
    }
    if (meshes && meshes->total_triangles > 0) {
        int V = meshes->total_vertices, T = meshes->total_triangles, N = meshes->total_nodes;
        w->meshBuf = (double*)xrealloc(NULL, sizeof(double) * (size_t)(3 * V + 8 * N));
        memcpy(w->meshBuf, meshes->positions, sizeof(double) * 3 * (size_t)V);
        memcpy(w->meshBuf + 3 * (size_t)V, meshes->nodes, sizeof(double) * 8 * (size_t)N);
        w->meshIdx = (int32_t*)xrealloc(NULL, sizeof(int32_t) * 4 * (size_t)T);
        memcpy(w->meshIdx, meshes->triangles, sizeof(int32_t) * 4 * (size_t)T);
        w->meshes.positions = w->meshBuf;
        w->meshes.nodes = w->meshBuf + 3 * (size_t)V;
        w->meshes.triangles = w->meshIdx;
        w->meshes.total_vertices = V; w->meshes.total_triangles = T; w->meshes.total_nodes = N;
    }
    w->sh.shapes = w->shapes;
    w->sh.hulls = &w->hulls;
    w->sh.meshes = &w->meshes;
    w->nb = nb;
    w->b = (OBody*)calloc((size_t)(nb ? nb : 1), sizeof(OBody));
    w->active = (int*)calloc((size_t)(nb ? nb : 1), sizeof(int));
    w->bodyLinks = (int*)calloc((size_t)(nb ? nb : 1), sizeof(int));
    w->bodyIslands = (int*)calloc((size_t)(nb ? nb : 1), sizeof(int));
    for (int i = 0; i < nb; i++) {
        const cc_body_desc* d = &bodies[i];
        OBody* b = &w->b[i];
        b->shape = d->shape; b->motion = d->motion_type; b->layer = d->object_layer;
        b->groups = d->collision_groups; b->mask = d->collision_mask;
        b->allowSleep = d->allow_sleeping; b->dof = d->allowed_dofs;
        b->fmode = d->friction_combine; b->rmode = d->restitution_combine;
        b->pos = vload(d->position);
        b->q.x = d->quaternion[0]; b->q.y = d->quaternion[1]; b->q.z = d->quaternion[2]; b->q.w = d->quaternion[3];
        b->v = vload(d->linear_velocity); b->w = vload(d->angular_velocity);
        b->invMass = d->inv_mass; b->invI = vload(d->inv_inertia_diagonal);
        b->inertiaRot.x = d->inertia_rotation[0]; b->inertiaRot.y = d->inertia_rotation[1];
        b->inertiaRot.z = d->inertia_rotation[2]; b->inertiaRot.w = d->inertia_rotation[3];
        b->friction = d->friction; b->restitution = d->restitution; b->gravityFactor = d->gravity_factor;
        b->linDamp = d->linear_damping; b->angDamp = d->angular_damping;
        b->maxLinV = d->max_linear_velocity; b->maxAngV = d->max_angular_velocity;
        b->sleeping = d->sleeping;
        b->ovV = d->num_velocity_steps_override;
        b->ovP = d->num_position_steps_override;
        b->quality = d->motion_quality;
        b->ccdIndex = -1;
        if (b->motion == CC_MOTION_DYNAMIC) { /* motionProperties.force / torque (dynamic bodies only) */
            b->force = vload(d->force);
            b->torque = vload(d->torque);
        }
        b->activeIndex = INACTIVE;
        b->islandIndex = -1;
        set_transform(w, b); /* updateShape :677-712 */
        if (b->motion == CC_MOTION_DYNAMIC) reset_sleep_timer(w, b);
        if (b->motion != CC_MOTION_STATIC && !b->sleeping) add_active(w, b, i);
    }
    w->npc = npc;
    w->pc = (OPoint*)calloc((size_t)(npc ? npc : 1), sizeof(OPoint));
    for (int i = 0; i < npc; i++) {
        w->pc[i].a = pcs[i].body_a; w->pc[i].b = pcs[i].body_b; w->pc[i].enabled = pcs[i].enabled;
        w->pc[i].localA = vload(pcs[i].local_point_a); w->pc[i].localB = vload(pcs[i].local_point_b);
        w->pc[i].ovV = pcs[i].num_velocity_steps_override; w->pc[i].ovP = pcs[i].num_position_steps_override;
    }
    w->consLinks = (int*)calloc((size_t)(npc ? npc : 1), sizeof(int));
    w->consIds = (int*)calloc((size_t)(npc ? npc : 1), sizeof(int));
    w->prevDt = 0;
    return w;
}

void oc_destroy(oc_world* w) {
    if (!w) return;
    free(w->shapes); free(w->hullBuf); free(w->hullIdx); free(w->meshBuf); free(w->meshIdx); free(w->b); free(w->active);
    free(w->pairs); free(w->ct); free(w->htab); free(w->rem); free(w->cc); free(w->pc);
    free(w->bodyLinks); free(w->bodyIslands); free(w->contactLinks); free(w->consLinks); free(w->consIds);
    for (int i = 0; i < w->cisl; i++) { free(w->isl[i].bodies); free(w->isl[i].contacts); free(w->isl[i].cons); }
    free(w->isl);
    free(w->ccd); free(w->ccdOrder);
    free(w);
}

/* ---- contact cache (src/contacts.ts) as a hash map keyed by (a, b, subA, subB) ---------- */
static uint64_t hkey(int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }
static uint32_t hhash(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
    return (uint32_t)k;
}
static uint32_t ckey(int a, int b, uint32_t sa, uint32_t sb) {
    return hhash(hkey(a, b)) ^ hhash(((uint64_t)sa << 32) | sb);
}
static void cache_rebuild(oc_world* w) {
    int need = 16;
    while (need < 2 * w->nct + 16) need <<= 1;
    if (need != w->hcap) { w->htab = (int*)xrealloc(w->htab, sizeof(int) * (size_t)need); w->hcap = need; }
    for (int i = 0; i < w->hcap; i++) w->htab[i] = -1;
    for (int i = 0; i < w->nct; i++) {
        uint32_t h = ckey(w->ct[i].a, w->ct[i].b, w->ct[i].sa, w->ct[i].sb) & (uint32_t)(w->hcap - 1);
        while (w->htab[h] >= 0) h = (h + 1) & (uint32_t)(w->hcap - 1);
        w->htab[h] = i;
    }
}
static int cache_find(oc_world* w, int a, int b, uint32_t sa, uint32_t sb) { /* findContact :481-513 */
    if (!w->hcap) return -1;
    uint32_t h = ckey(a, b, sa, sb) & (uint32_t)(w->hcap - 1);
    while (w->htab[h] >= 0) {
        OContact* c = &w->ct[w->htab[h]];
        if (c->a == a && c->b == b && c->sa == sa && c->sb == sb) return w->htab[h];
        h = (h + 1) & (uint32_t)(w->hcap - 1);
    }
    return -1;
}
static int cache_create(oc_world* w, int a, int b, uint32_t sa, uint32_t sb) { /* createContact :235-285 + setContact :141-170 */
    if (w->nct == w->cct) { w->cct = w->cct ? 2 * w->cct : 256; w->ct = (OContact*)xrealloc(w->ct, sizeof(OContact) * (size_t)w->cct); }
    OContact* c = &w->ct[w->nct];
    memset(c, 0, sizeof(*c));
    c->a = a; c->b = b; c->sa = sa; c->sb = sb;
    if (2 * (w->nct + 1) + 16 > w->hcap) { w->nct++; cache_rebuild(w); return w->nct - 1; }
    uint32_t h = ckey(a, b, sa, sb) & (uint32_t)(w->hcap - 1);
    while (w->htab[h] >= 0) h = (h + 1) & (uint32_t)(w->hcap - 1);
    w->htab[h] = w->nct;
    return w->nct++;
}
/* ---- constraint parts (src/constraints/constraint-part/axis-constraint-part.ts) ---- */
static void axis_props(OAxis* p, const OBody* A, const OBody* B, double invMA, double invMB, const M3* IA, const M3* IB,
                       V3 r1, V3 r2, V3 axis, double bias) { /* calculateConstraintProperties :169-201 */
    double inv = 0;
    if (A->motion == CC_MOTION_DYNAMIC) {
        p->r1x = vcross(r1, axis);
        p->iI1r1x = m3_mul(IA, p->r1x);
        inv += invMA + vdot(p->iI1r1x, p->r1x);
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        p->r2x = vcross(r2, axis);
        p->iI2r2x = m3_mul(IB, p->r2x);
        inv += invMB + vdot(p->iI2r2x, p->r2x);
    }
    if (inv == 0) { p->effMass = 0; p->lambda = 0; }
    else { p->effMass = 1 / inv; p->bias = bias; }
}
static int axis_active(const OAxis* p) { return p->effMass != 0; }
static double axis_total_lambda(const OAxis* p, const OBody* A, const OBody* B, V3 axis) { /* getTotalLambda :578-625 */
    int mA = A->motion != CC_MOTION_STATIC, mB = B->motion != CC_MOTION_STATIC;
    double jv;
    if (mA && mB) jv = vdot(axis, vsub(A->v, B->v));
    else if (mA) jv = vdot(axis, A->v);
    else if (mB) jv = -vdot(axis, B->v);
    else jv = 0;
    if (mA) jv += vdot(p->r1x, A->w);
    if (mB) jv -= vdot(p->r2x, B->w);
    double lam = p->effMass * (jv - p->bias);
    return p->lambda + lam;
}
static void axis_apply(OAxis* p, OBody* A, OBody* B, double invMA, double invMB, V3 axis, double total) { /* applyLambda :640-707 */
    double dl = total - p->lambda;
    p->lambda = total;
    if (dl == 0) return;
    if (A->motion == CC_MOTION_DYNAMIC) {
        double ls = dl * invMA;
        A->v.x -= axis.x * ls; A->v.y -= axis.y * ls; A->v.z -= axis.z * ls;
        A->v = dof_lin(A->v, A->dof);
        A->w.x -= p->iI1r1x.x * dl; A->w.y -= p->iI1r1x.y * dl; A->w.z -= p->iI1r1x.z * dl;
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        double ls = dl * invMB;
        B->v.x += axis.x * ls; B->v.y += axis.y * ls; B->v.z += axis.z * ls;
        B->v = dof_lin(B->v, B->dof);
        B->w.x += p->iI2r2x.x * dl; B->w.y += p->iI2r2x.y * dl; B->w.z += p->iI2r2x.z * dl;
    }
}
static void axis_warm(OAxis* p, OBody* A, OBody* B, double invMA, double invMB, V3 axis, double ratio) { /* warmStart :500-537 */
    p->lambda *= ratio;
    if (p->lambda == 0) return;
    if (A->motion == CC_MOTION_DYNAMIC) {
        A->v = dof_lin(vsub(A->v, vscale(axis, p->lambda * invMA)), A->dof);
        A->w = vsub(A->w, vscale(p->iI1r1x, p->lambda));
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        B->v = dof_lin(vadd(B->v, vscale(axis, p->lambda * invMB)), B->dof);
        B->w = vadd(B->w, vscale(p->iI2r2x, p->lambda));
    }
}
static int axis_solve_position(OAxis* p, OBody* A, OBody* B, double invMA, double invMB, V3 axis, double C, double beta) {
    /* solvePositionConstraintWithMassOverride :790-847 */
    if (C == 0 || !axis_active(p)) return 0;
    double lam = -p->effMass * beta * C;
    if (A->motion == CC_MOTION_DYNAMIC) {
        V3 st = vscale(axis, lam * invMA);
        A->com = dof_lin(vsub(A->com, st), A->dof); /* subPositionStep: applyTranslationDOFConstraint on the position */
        add_rot_step(A, vscale(p->iI1r1x, lam), -1.0);
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        V3 st = vscale(axis, lam * invMB);
        B->com = dof_lin(vadd(B->com, st), B->dof); /* addPositionStep */
        add_rot_step(B, vscale(p->iI2r2x, lam), 1.0);
    }
    return 1;
}

static double combine(double a, double b, int ma, int mb) { /* combineMaterial src/constraints/combine-material.ts:40-57 */
    static const int prio[5] = {0, 1, 3, 4, 2}; /* AVERAGE, MULTIPLY, MIN, MAX, GEOMETRIC_MEAN */
    int mode = prio[ma] >= prio[mb] ? ma : mb;
    switch (mode) {
    case CC_COMBINE_MULTIPLY: return a * b;
    case CC_COMBINE_MIN: return dmin(a, b);
    case CC_COMBINE_MAX: return dmax(a, b);
    case CC_COMBINE_GEOMETRIC_MEAN: return sqrt(a * b);
    default: return (a + b) / 2;
    }
}

static uint32_t sort_key(uint32_t a, uint32_t b, uint32_t sa, uint32_t sb) { /* computeContactSortKey :1222-1231 */
    uint32_t h = 0;
    h = h * 31u + a; h = h * 31u + b; h = h * 31u + sa; h = h * 31u + sb;
    return h;
}

/* tangent basis of a contact normal, addContactConstraint :536-556: t1 = (nz, 0, -nx)/|.| when
 * |nx| > |ny|, else (0, nz, -ny)/|.|; t2 = n x t1 */
static void contact_tangents(V3 n, V3* t1, V3* t2) {
    if (fabs(n.x) > fabs(n.y)) { double tl = sqrt(n.x * n.x + n.z * n.z); *t1 = v3(n.z / tl, 0, -n.x / tl); }
    else { double tl = sqrt(n.y * n.y + n.z * n.z); *t1 = v3(0, n.z / tl, -n.y / tl); }
    *t2 = vcross(n, *t1);
}

/* calculateNormalVelocityBias src/constraints/contact-constraints.ts:288-423 */
static double normal_bias(const oc_world* w, const OBody* A, const OBody* B, V3 pA, V3 pB, V3 n, V3 rA, V3 rB, double e,
                          double dt) {
    int mA = A->motion != CC_MOTION_STATIC, mB = B->motion != CC_MOTION_STATIC;
    int dA = A->motion == CC_MOTION_DYNAMIC, dB = B->motion == CC_MOTION_DYNAMIC;
    V3 rel;
    if (mA && mB) rel = vsub(vadd(B->v, vcross(B->w, rB)), vadd(A->v, vcross(A->w, rA)));
    else if (mA) rel = vneg(vadd(A->v, vcross(A->w, rA)));
    else if (mB) rel = vadd(B->v, vcross(B->w, rB));
    else return 0;
    double nv = vdot(rel, n);
    double pen = vdot(vsub(pA, pB), n);
    double spec = dmax(0, -pen / dt);
    V3 g = vload(w->s.gravity);
    if (e > 0 && nv < -w->s.min_velocity_for_restitution) {
        if (nv < -spec) {
            V3 acc = v3(0, 0, 0);
            if (mA && mB) acc = vadd(acc, vscale(g, B->gravityFactor - A->gravityFactor));
            else if (mA) acc = vadd(acc, vscale(g, -A->gravityFactor));
            else if (mB) acc = vadd(acc, vscale(g, B->gravityFactor));
            if (dA) acc = vadd(acc, vscale(A->force, -A->invMass));
            if (dB) acc = vadd(acc, vscale(B->force, B->invMass));
            double fdv = dmin(0, vdot(acc, n) * dt);
            return e * (nv - fdv);
        }
        return spec;
    }
    return spec;
}

/* addContactConstraint src/constraints/contact-constraints.ts:437-857 */
static int add_contact_constraint(oc_world* w, int ia, int ib, OManifold* m, uint32_t subA, uint32_t subB, double dt) {
    if (ia > ib) { /* swap so that id A < id B, manifold.swapShapes :86-107 */
        int t = ia; ia = ib; ib = t;
        uint32_t u = subA; subA = subB; subB = u;
        m->normal = vneg(m->normal);
        for (int i = 0; i < m->n; i++) { V3 x = m->relA[i]; m->relA[i] = m->relB[i]; m->relB[i] = x; }
    }
    OBody* A = &w->b[ia];
    OBody* B = &w->b[ib];
    int ci = cache_find(w, ia, ib, subA, subB);
    int existing = ci >= 0;
    if (!existing) ci = cache_create(w, ia, ib, subA, subB);
    OContact* c = &w->ct[ci];
    c->processed = 1;
    M3 rotA = m3_from_quat(A->q), rotB = m3_from_quat(B->q);
    c->normalB = vnormalize(m3_mul_t(&rotB, m->normal));
    c->numPoints = m->n;
    double fr = combine(A->friction, B->friction, A->fmode, B->fmode);
    double re = combine(A->restitution, B->restitution, A->rmode, B->rmode);
    if (existing) c->persisted = 1;
    if ((A->motion == CC_MOTION_DYNAMIC && A->invMass != 0) || (B->motion == CC_MOTION_DYNAMIC && B->invMass != 0)) {
        if (w->ncc == w->ccc) { w->ccc = w->ccc ? 2 * w->ccc : 256; w->cc = (OCC*)xrealloc(w->cc, sizeof(OCC) * (size_t)w->ccc); }
        OCC* k = &w->cc[w->ncc++];
        memset(k, 0, sizeof(*k));
        k->a = ia; k->b = ib; k->contact = ci;
        k->subA = subA; k->subB = subB;
        k->persisted = existing;
        k->sortKey = sort_key((uint32_t)ia, (uint32_t)ib, subA, subB);
        V3 n = m->normal;
        k->normal = n;
        contact_tangents(n, &k->t1, &k->t2);
        k->friction = fr; k->restitution = re;
        k->invMassA = A->invMass * 1.0; k->invMassB = B->invMass * 1.0;
        if (A->motion == CC_MOTION_DYNAMIC) k->invIA = inv_inertia(A->invI, A->inertiaRot, &rotA, A->dof);
        if (B->motion == CC_MOTION_DYNAMIC) k->invIB = inv_inertia(B->invI, B->inertiaRot, &rotB, B->dof);
        k->np = 0;
        for (int i = 0; i < m->n; i++) {
            OCPoint* cp = &k->pts[k->np++];
            cp->posA = vadd(m->base, m->relA[i]);
            cp->posB = vadd(m->base, m->relB[i]);
            cp->localA = m3_mul_t(&rotA, vsub(cp->posA, A->com));
            cp->localB = m3_mul_t(&rotB, vsub(cp->posB, B->com));
            int set = 0;
            if (existing) { /* in-place cache: slots < i already hold this step's points (reference behaviour) */
                for (int j = 0; j < c->numPoints; j++) {
                    OCached* pt = &c->pts[j];
                    double d1 = vlen2(vsub(cp->localA, pt->p1)), d2 = vlen2(vsub(cp->localB, pt->p2));
                    double th = w->s.contact_point_preserve_lambda_max_dist_sq;
                    if (d1 < th && d2 < th) {
                        cp->n.lambda = pt->ln; cp->t1.lambda = pt->lt1; cp->t2.lambda = pt->lt2;
                        set = 1;
                        break;
                    }
                }
            }
            if (!set) cp->n.lambda = cp->t1.lambda = cp->t2.lambda = 0;
            V3 mid = v3((cp->posA.x + cp->posB.x) * 0.5, (cp->posA.y + cp->posB.y) * 0.5, (cp->posA.z + cp->posB.z) * 0.5);
            V3 rA = vsub(mid, A->com), rB = vsub(mid, B->com);
            double bias = normal_bias(w, A, B, cp->posA, cp->posB, n, rA, rB, re, dt);
            axis_props(&cp->n, A, B, k->invMassA, k->invMassB, &k->invIA, &k->invIB, rA, rB, n, bias);
            if (fr > 0) {
                axis_props(&cp->t1, A, B, k->invMassA, k->invMassB, &k->invIA, &k->invIB, rA, rB, k->t1, 0);
                axis_props(&cp->t2, A, B, k->invMassA, k->invMassB, &k->invIA, &k->invIB, rA, rB, k->t2, 0);
            } else {
                cp->t1.effMass = 0; cp->t1.lambda = 0;
                cp->t2.effMass = 0; cp->t2.lambda = 0;
            }
            OCached* st = &c->pts[i];
            st->p1 = cp->localA; st->p2 = cp->localB;
            st->ln = cp->n.lambda; st->lt1 = cp->t1.lambda; st->lt2 = cp->t2.lambda;
        }
        return 1;
    }
    for (int i = 0; i < m->n; i++) {
        V3 pa = vadd(m->base, m->relA[i]), pb = vadd(m->base, m->relB[i]);
        OCached* st = &c->pts[i];
        st->p1 = m3_mul_t(&rotA, vsub(pa, A->com));
        st->p2 = m3_mul_t(&rotB, vsub(pb, B->com));
        st->ln = st->lt1 = st->lt2 = 0;
    }
    return 0;
}

/* ---- broadphase (src/broadphase/broadphase.ts:44-138, 217-277) --------------- */
static int queries(const OBody* b) { return b->motion != CC_MOTION_STATIC && !b->sleeping; }
static int64_t dedup(const OBody* b, int idx) { return queries(b) ? idx : (int64_t)INACTIVE; }
static int overlap(V3 amn, V3 amx, V3 bmn, V3 bmx) { /* box3.intersectsBox3, inclusive */
    return !(bmx.x < amn.x || bmn.x > amx.x || bmx.y < amn.y || bmn.y > amx.y || bmx.z < amn.z || bmn.z > amx.z);
}
static int pair_ok(const cc_settings* s, const OBody* a, int ia, const OBody* b, int ib) {
    if (ia == ib) return 0;
    if (!s->object_layer_pairs[a->layer * CC_MAX_OBJECT_LAYERS + b->layer]) return 0;
    if (!((a->groups & b->mask) != 0 && (b->groups & a->mask) != 0)) return 0;
    if (dedup(a, ia) >= dedup(b, ib)) return 0;
    if (a->motion == CC_MOTION_STATIC && b->motion == CC_MOTION_STATIC) return 0;
    if (a->motion != CC_MOTION_DYNAMIC && b->motion != CC_MOTION_DYNAMIC) return 0;
    return 1;
}
static int cmp_pair_qo(const void* x, const void* y) {
    const int* a = (const int*)x;
    const int* b = (const int*)y;
    if (a[0] != b[0]) return a[0] - b[0];
    return a[1] - b[1];
}
static void push_pair(oc_world* w, int a, int b) {
    if (w->npairs + 1 > w->cpairs) { w->cpairs = w->cpairs ? 2 * w->cpairs : 1024; w->pairs = (int*)xrealloc(w->pairs, sizeof(int) * 2 * (size_t)w->cpairs); }
    w->pairs[2 * w->npairs] = a;
    w->pairs[2 * w->npairs + 1] = b;
    w->npairs++;
}
/* Pair enumeration: sort-and-sweep on x over the expanded AABBs, then pairs ordered by
 * (querying body, other body) — the same list a brute-force double loop produces. */
static __thread double* g_sap_key;
static int cmp_sap(const void* x, const void* y) {
    int a = *(const int*)x, b = *(const int*)y;
    if (g_sap_key[a] != g_sap_key[b]) return g_sap_key[a] < g_sap_key[b] ? -1 : 1;
    return a - b;
}
static void find_pairs(oc_world* w) {
    double m = w->s.speculative_contact_distance;
    w->npairs = 0;
    int n = w->nb;
    int* ord = (int*)xrealloc(NULL, sizeof(int) * (size_t)(n ? n : 1));
    double* key = (double*)xrealloc(NULL, sizeof(double) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; i++) {
        ord[i] = i;
        key[i] = w->b[i].aabbMin.x - (queries(&w->b[i]) ? m : 0.0);
    }
    g_sap_key = key;
    sort_some(ord, (size_t)n, sizeof(int), cmp_sap);
    for (int oi = 0; oi < n; oi++) {
        int i = ord[oi];
        OBody* a = &w->b[i];
        double ahi = a->aabbMax.x + (queries(a) ? m : 0.0);
        for (int oj = oi + 1; oj < n; oj++) {
            int j = ord[oj];
            if (key[j] > ahi) break;
            OBody* b = &w->b[j];
            int q = -1, o = -1;
            if (pair_ok(&w->s, a, i, b, j)) { q = i; o = j; }
            else if (pair_ok(&w->s, b, j, a, i)) { q = j; o = i; }
            if (q < 0) continue;
            OBody* Q = &w->b[q];
            V3 mn = vsub(Q->aabbMin, v3(m, m, m)), mx = vadd(Q->aabbMax, v3(m, m, m));
            if (!overlap(mn, mx, w->b[o].aabbMin, w->b[o].aabbMax)) continue;
            push_pair(w, q, o);
        }
    }
    free(ord);
    free(key);
    sort_some(w->pairs, (size_t)w->npairs, 2 * sizeof(int), cmp_pair_qo);
}

/* ---- narrowphase of one pair with the reduction collector (src/update.ts:333-563, 714-850) ---- */
#define OC_MAX_MANIFOLDS 32
typedef struct {
    const oc_world* w;
    V3 base;                       /* bodyA.position */
    int reversed;                  /* the mesh is body A: hits come back inverted (collide-shape-vs-shape.ts:228-271) */
    int n;
    OManifold m[OC_MAX_MANIFOLDS];
    uint32_t subA[OC_MAX_MANIFOLDS], subB[OC_MAX_MANIFOLDS];
} OCollector;
static __thread OCollector g_col;
static __thread OManifold g_tmp;

/* addHit :385-520: merge by normal, else a new manifold (the shallowest one is replaced when
 * all 32 are in use), then append the hit's clipped points */
static void collector_add(OCollector* C, const OHit* h, uint32_t subA, uint32_t subB) {
    V3 normal = vnormalize(h->axis);
    int f = -1;
    for (int i = 0; i < C->n; i++)
        if (vdot(normal, C->m[i].normal) >= C->w->s.normal_cos_max_delta_rotation) { f = i; break; }
    if (f < 0) {
        if (C->n >= OC_MAX_MANIFOLDS) {
            int sh = 0;
            for (int i = 1; i < C->n; i++)
                if (C->m[i].depth < C->m[sh].depth) sh = i;
            if (h->penetration < C->m[sh].depth) return;
            f = sh;
        } else {
            f = C->n++;
        }
        OManifold* m = &C->m[f];
        m->n = 0;
        m->base = C->base;
        m->normal = normal;
        m->depth = h->penetration;
        C->subA[f] = subA;
        C->subB[f] = subB;
    } else {
        C->m[f].normal = vadd(C->m[f].normal, normal);
    }
    OManifold* m = &C->m[f];
    g_tmp.base = C->base;
    g_tmp.n = 0;
    double maxDist = C->w->s.speculative_contact_distance + C->w->s.manifold_tolerance;
    if (h->faceA.n >= 2 && h->faceB.n >= 3)
        ocx_manifold_between_two_faces(&g_tmp, h->pointA, h->pointB, h->axis, maxDist, &h->faceA, &h->faceB);
    else { g_tmp.relA[0] = vsub(h->pointA, g_tmp.base); g_tmp.relB[0] = vsub(h->pointB, g_tmp.base); g_tmp.n = 1; }
    for (int i = 0; i < g_tmp.n && m->n < 64; i++) { m->relA[m->n] = g_tmp.relA[i]; m->relB[m->n] = g_tmp.relB[i]; m->n++; }
    if (h->penetration > m->depth) m->depth = h->penetration;
    if (m->n > 32) ocx_prune(m, m->normal);
}
/* a convex-vs-triangle hit, through the inverted collector when the mesh is body A */
static void mesh_hit(void* ctx, const OHit* h, uint32_t subTri) {
    OCollector* C = (OCollector*)ctx;
    if (!C->reversed) { collector_add(C, h, EMPTY_SUB, subTri); return; }
    OHit r;
    r.hit = h->hit; r.status = h->status;
    r.pointA = h->pointB; r.pointB = h->pointA;
    r.axis = vneg(h->axis);
    r.penetration = h->penetration;
    r.faceA = h->faceB; r.faceB = h->faceA;
    collector_add(C, &r, subTri, EMPTY_SUB);
}
/* collideCompoundVsShape / collideShapeVsCompound (src/shapes/compound.ts:494-640): every child
 * of body A's compound, then of body B's, in index order, its sub-shape id pushed with
 * pushIndex(child, count) (src/body/sub-shape.ts:43-64); a child carries its position and
 * quaternion as a local transform, which ocx_collide composes with the body pose as the
 * compound does.  Each leaf pair's hit goes to the reduction collector. */
static uint32_t sub_push(uint32_t value, int bit, int id, int nbits) {
    const uint32_t mask = nbits ? (((1u << nbits) - 1u) << bit) : 0u;
    return (value & ~mask) | ((uint32_t)id << bit);
}
static void collide_compound(oc_world* w, OCollector* C, int sa, uint32_t subA, int bitA, V3 pa, Q4 qa, int sb,
                             uint32_t subB, int bitB, V3 pb, Q4 qb) {
    const cc_shape_desc* A = &w->shapes[sa];
    const cc_shape_desc* B = &w->shapes[sb];
    if (A->type == CC_SHAPE_COMPOUND || B->type == CC_SHAPE_COMPOUND) {
        const int onA = A->type == CC_SHAPE_COMPOUND;
        const cc_shape_desc* K = onA ? A : B;
        const int n = K->compound_num_children;
        const int nbits = n > 1 ? 32 - __builtin_clz((unsigned)(n - 1)) : 0;
        for (int i = 0; i < n; i++) {
            const int child = K->compound_first_child + i;
            if (onA) collide_compound(w, C, child, sub_push(subA, bitA, i, nbits), bitA + nbits, pa, qa, sb, subB, bitB, pb, qb);
            else collide_compound(w, C, sa, subA, bitA, pa, qa, child, sub_push(subB, bitB, i, nbits), bitB + nbits, pb, qb);
        }
        return;
    }
    OHit h;
    ocx_collide(&w->sh, sa, pa, qa, sb, pb, qb, w->s.speculative_contact_distance, w->s.collision_tolerance,
                w->s.penetration_tolerance, &h);
    if (h.hit) collector_add(C, &h, subA, subB);
}

static int narrowphase_pair(oc_world* w, int ia, int ib, double dt) {
    OBody* A = &w->b[ia];
    OBody* B = &w->b[ib];
    OCollector* C = &g_col;
    C->w = w;
    C->base = A->pos;
    C->n = 0;
    int ta = w->shapes[A->shape].type, tb = w->shapes[B->shape].type;
    if ((ta == CC_SHAPE_PLANE || tb == CC_SHAPE_PLANE) && (ta == CC_SHAPE_TRIANGLE_MESH || tb == CC_SHAPE_TRIANGLE_MESH))
        return 0; /* no plane-vs-mesh function */
    if (ta == CC_SHAPE_TRIANGLE_MESH || tb == CC_SHAPE_TRIANGLE_MESH) {
        if (ta == CC_SHAPE_TRIANGLE_MESH && tb == CC_SHAPE_TRIANGLE_MESH) return 0; /* no mesh-vs-mesh function */
        V3 move = vsub(A->v, B->v); /* activeEdgeMovementDirection :730-735 */
        double maxSep = w->s.speculative_contact_distance;
        if (ta == CC_SHAPE_SPHERE || tb == CC_SHAPE_SPHERE) { /* collideSphereVsTriangleMesh and its reverse */
            C->reversed = ta == CC_SHAPE_TRIANGLE_MESH;
            if (!C->reversed)
                ocx_collide_sphere_mesh(&w->sh, A->shape, A->pos, B->shape, B->pos, B->q, &w->s, maxSep, move, mesh_hit, C);
            else
                ocx_collide_sphere_mesh(&w->sh, B->shape, B->pos, A->shape, A->pos, A->q, &w->s, maxSep, move, mesh_hit, C);
        } else if (tb == CC_SHAPE_TRIANGLE_MESH) {
            C->reversed = 0;
            ocx_collide_mesh(&w->sh, A->shape, A->pos, A->q, B->shape, B->pos, B->q, &w->s, maxSep, move, mesh_hit, C);
        } else {
            C->reversed = 1;
            ocx_collide_mesh(&w->sh, B->shape, B->pos, B->q, A->shape, A->pos, A->q, &w->s, maxSep, move, mesh_hit, C);
        }
    } else if (ta == CC_SHAPE_COMPOUND || tb == CC_SHAPE_COMPOUND) {
        collide_compound(w, C, A->shape, EMPTY_SUB, 0, A->pos, A->q, B->shape, EMPTY_SUB, 0, B->pos, B->q);
    } else {
        OHit h;
        ocx_collide(&w->sh, A->shape, A->pos, A->q, B->shape, B->pos, B->q, w->s.speculative_contact_distance,
                    w->s.collision_tolerance, w->s.penetration_tolerance, &h);
        if (!h.hit) return 0;
        collector_add(C, &h, EMPTY_SUB, EMPTY_SUB);
    }
    /* finalizeAndCreateConstraints :522-555 */
    int made = 0;
    for (int i = 0; i < C->n; i++) {
        OManifold* m = &C->m[i];
        m->normal = vnormalize(m->normal);
        if (m->n > 4) ocx_prune(m, m->normal);
        if (m->n > 0) made |= add_contact_constraint(w, ia, ib, m, C->subA[i], C->subB[i], dt);
    }
    return made;
}

/* ---- islands (src/islands.ts) ----------------------------------------------- */
static int lowest(const oc_world* w, int i) { while (w->bodyLinks[i] != i) i = w->bodyLinks[i]; return i; }
static void link_bodies(oc_world* w, int a, int b) { /* linkBodies :103-140 */
    if (a >= w->nactive || b >= w->nactive) return;
    int f = lowest(w, a), s = lowest(w, b);
    if (f == s) return;
    if (f < s) w->bodyLinks[s] = f; else w->bodyLinks[f] = s;
    int lo = f < s ? f : s;
    if (w->bodyLinks[a] > lo) w->bodyLinks[a] = lo;
    if (w->bodyLinks[b] > lo) w->bodyLinks[b] = lo;
}
static void ipush(int** arr, int* n, int* cap, int v) {
    if (*n == *cap) { *cap = *cap ? 2 * *cap : 8; *arr = (int*)xrealloc(*arr, sizeof(int) * (size_t)*cap); }
    (*arr)[(*n)++] = v;
}
static int cmp_island(const void* x, const void* y) { /* compareIslandSize, stable via index */
    const OIsland* a = (const OIsland*)x;
    const OIsland* b = (const OIsland*)y;
    int na = a->nc + a->nu, nb = b->nc + b->nu;
    if (na != nb) return nb - na;
    return 0;
}
static void build_islands(oc_world* w) {
    int na = w->nactive;
    for (int i = 0; i < w->nb; i++) w->b[i].islandIndex = -1;
    for (int i = 0; i < na; i++) { w->bodyLinks[i] = i; w->bodyIslands[i] = 0; }
    w->contactLinks = (int*)xrealloc(w->contactLinks, sizeof(int) * (size_t)(w->ncc ? w->ncc : 1));
    for (int c = 0; c < w->ncc; c++) { /* linkContactConstraints :184-204 */
        OBody* A = &w->b[w->cc[c].a];
        OBody* B = &w->b[w->cc[c].b];
        if (A->activeIndex != INACTIVE && B->activeIndex != INACTIVE) link_bodies(w, A->activeIndex, B->activeIndex);
        w->contactLinks[c] = A->activeIndex < B->activeIndex ? A->activeIndex : B->activeIndex;
    }
    w->nconsIds = 0;
    for (int i = 0; i < w->npc; i++) { /* linkUserConstraints :210-226 */
        OPoint* p = &w->pc[i];
        if (!p->enabled) continue;
        OBody* A = &w->b[p->a];
        OBody* B = &w->b[p->b];
        if (A->activeIndex != INACTIVE || B->activeIndex != INACTIVE) {
            if (A->activeIndex != INACTIVE && B->activeIndex != INACTIVE) link_bodies(w, A->activeIndex, B->activeIndex);
            w->consIds[w->nconsIds] = i;
            w->consLinks[w->nconsIds] = A->activeIndex < B->activeIndex ? A->activeIndex : B->activeIndex;
            w->nconsIds++;
        }
    }
    int nisl = 0; /* finalize :240-389 */
    for (int i = 0; i < na; i++) {
        int l = w->bodyLinks[i];
        if (l != i) w->bodyIslands[i] = w->bodyIslands[l];
        else w->bodyIslands[i] = nisl++;
    }
    if (nisl > w->cisl) {
        w->isl = (OIsland*)xrealloc(w->isl, sizeof(OIsland) * (size_t)nisl);
        memset(w->isl + w->cisl, 0, sizeof(OIsland) * (size_t)(nisl - w->cisl));
        w->cisl = nisl;
    }
    w->nisl = nisl;
    for (int i = 0; i < nisl; i++) { w->isl[i].nb = w->isl[i].nc = w->isl[i].nu = 0; }
    for (int ai = 0; ai < na; ai++) {
        int bi = w->active[ai];
        int il = w->bodyIslands[ai];
        ipush(&w->isl[il].bodies, &w->isl[il].nb, &w->isl[il].cb, ai);
        w->b[bi].islandIndex = il;
    }
    for (int c = 0; c < w->ncc; c++) {
        int l = w->contactLinks[c];
        if (l >= INACTIVE) continue;
        int il = w->bodyIslands[l];
        ipush(&w->isl[il].contacts, &w->isl[il].nc, &w->isl[il].cc, c);
    }
    for (int k = 0; k < w->nconsIds; k++) {
        int l = w->consLinks[k];
        if (l >= INACTIVE) continue;
        int il = w->bodyIslands[l];
        ipush(&w->isl[il].cons, &w->isl[il].nu, &w->isl[il].cu, w->consIds[k]);
    }
    /* stable sort by constraint count, descending (Array.prototype.sort is stable) */
    for (int i = 1; i < nisl; i++) { /* insertion sort keeps stability */
        OIsland t = w->isl[i];
        int j = i - 1;
        while (j >= 0 && cmp_island(&w->isl[j], &t) > 0) { w->isl[j + 1] = w->isl[j]; j--; }
        w->isl[j + 1] = t;
    }
    for (int i = 0; i < nisl; i++) {
        OIsland* I = &w->isl[i];
        for (int k = 0; k < I->nb; k++) w->b[w->active[I->bodies[k]]].islandIndex = i;
        /* solver iterations of the island (islands.ts:344-386): the largest override of its dynamic
         * bodies and user constraints, and at least the world default if any of them uses 0 */
        int nv = 0, np = 0, defV = 0, defP = 0;
        for (int k = 0; k < I->nb; k++) {
            const OBody* b = &w->b[w->active[I->bodies[k]]];
            if (b->motion != CC_MOTION_DYNAMIC) continue;
            if (b->ovV > nv) nv = b->ovV;
            defV = defV || b->ovV == 0;
            if (b->ovP > np) np = b->ovP;
            defP = defP || b->ovP == 0;
        }
        for (int k = 0; k < I->nu; k++) {
            const OPoint* p = &w->pc[I->cons[k]];
            if (p->ovV > nv) nv = p->ovV;
            defV = defV || p->ovV == 0;
            if (p->ovP > np) np = p->ovP;
            defP = defP || p->ovP == 0;
        }
        if (defV && w->s.velocity_iterations > nv) nv = w->s.velocity_iterations;
        if (defP && w->s.position_iterations > np) np = w->s.position_iterations;
        I->velSteps = nv;
        I->posSteps = np;
    }
}

static __thread oc_world* g_sort_world;
static int cmp_contact(const void* x, const void* y) { /* sortContactIndices :1240-1264 + stability */
    const OCC* a = &g_sort_world->cc[*(const int*)x];
    const OCC* b = &g_sort_world->cc[*(const int*)y];
    if (a->sortKey != b->sortKey) return a->sortKey < b->sortKey ? -1 : 1;
    if (a->a != b->a) return a->a < b->a ? -1 : 1;
    if (a->b != b->b) return a->b < b->b ? -1 : 1;
    return *(const int*)x < *(const int*)y ? -1 : 1;
}
static int cmp_int(const void* x, const void* y) { return *(const int*)x - *(const int*)y; }

/* ---- point constraint (src/constraints/constraint-part/point-constraint-part.ts) ---- */
static M3 cross_mat(V3 r) { M3 m = {{0, r.z, -r.y, -r.z, 0, r.x, r.y, -r.x, 0}}; return m; }
static M3 m3_mm(const M3* a, const M3* b) {
    M3 o;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) o.m[c * 3 + r] = a->m[r] * b->m[c * 3] + a->m[3 + r] * b->m[c * 3 + 1] + a->m[6 + r] * b->m[c * 3 + 2];
    return o;
}
static M3 m3_mm_rt(const M3* a, const M3* b) { /* a * b^T */
    M3 o;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) o.m[c * 3 + r] = a->m[r] * b->m[c] + a->m[3 + r] * b->m[3 + c] + a->m[6 + r] * b->m[6 + c];
    return o;
}
static int m3_invert(const double* a, double* out) { /* mat4.invert3x3 on the 3x3 block, [15] = 1 when invertible */
    double a00 = a[0], a01 = a[1], a02 = a[2], a10 = a[4], a11 = a[5], a12 = a[6], a20 = a[8], a21 = a[9], a22 = a[10];
    double b01 = a22 * a11 - a12 * a21, b11 = -a22 * a10 + a12 * a20, b21 = a21 * a10 - a11 * a20;
    double det = a00 * b01 + a01 * b11 + a02 * b21;
    if (det == 0) return 0;
    det = 1.0 / det;
    memset(out, 0, 16 * sizeof(double));
    out[0] = b01 * det; out[1] = (-a22 * a01 + a02 * a21) * det; out[2] = (a12 * a01 - a02 * a11) * det;
    out[4] = b11 * det; out[5] = (a22 * a00 - a02 * a20) * det; out[6] = (-a12 * a00 + a02 * a10) * det;
    out[8] = b21 * det; out[9] = (-a21 * a00 + a01 * a20) * det; out[10] = (a11 * a00 - a01 * a10) * det;
    out[15] = 1;
    return 1;
}
static void point_props(OPoint* p, OBody* A, OBody* B) { /* calculateConstraintProperties :77-165 */
    M3 RA = m3_from_quat(A->q), RB = m3_from_quat(B->q);
    p->r1 = m3_mul(&RA, p->localA);
    p->r2 = m3_mul(&RB, p->localB);
    double K[16];
    memset(K, 0, sizeof(K));
    double sumInv = 0;
    memset(&p->iI1r1X, 0, sizeof(M3));
    memset(&p->iI2r2X, 0, sizeof(M3));
    if (A->motion == CC_MOTION_DYNAMIC) {
        sumInv += A->invMass;
        M3 I = inv_inertia(A->invI, A->inertiaRot, &RA, A->dof);
        M3 X = cross_mat(p->r1);
        p->iI1r1X = m3_mm(&I, &X);
        M3 t1 = m3_mm(&X, &I);
        M3 t2 = m3_mm_rt(&t1, &X);
        for (int c = 0; c < 3; c++) for (int r = 0; r < 3; r++) K[c * 4 + r] += t2.m[c * 3 + r];
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        sumInv += B->invMass;
        M3 I = inv_inertia(B->invI, B->inertiaRot, &RB, B->dof);
        M3 X = cross_mat(p->r2);
        p->iI2r2X = m3_mm(&I, &X);
        M3 t1 = m3_mm(&X, &I);
        M3 t2 = m3_mm_rt(&t1, &X);
        for (int c = 0; c < 3; c++) for (int r = 0; r < 3; r++) K[c * 4 + r] += t2.m[c * 3 + r];
    }
    K[0] += sumInv; K[5] += sumInv; K[10] += sumInv;
    if (!m3_invert(K, p->eff)) { memset(p->eff, 0, sizeof(p->eff)); p->lambda = v3(0, 0, 0); }
}
static V3 eff_mul(const double* e, V3 v) {
    return v3(e[0] * v.x + e[4] * v.y + e[8] * v.z, e[1] * v.x + e[5] * v.y + e[9] * v.z, e[2] * v.x + e[6] * v.y + e[10] * v.z);
}
static int point_apply(OPoint* p, OBody* A, OBody* B, V3 lam) { /* applyVelocityStep :232-262 */
    if (vlen2(lam) == 0) return 0;
    if (A->motion == CC_MOTION_DYNAMIC) {
        A->v = vadd(A->v, vscale(lam, -A->invMass));
        A->w = vsub(A->w, m3_mul(&p->iI1r1X, lam));
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        B->v = vadd(B->v, vscale(lam, B->invMass));
        B->w = vadd(B->w, m3_mul(&p->iI2r2X, lam));
    }
    return 1;
}
static void point_solve_velocity(OPoint* p, OBody* A, OBody* B) { /* solveVelocityConstraint :186-221 */
    if (p->eff[15] == 0) return;
    V3 jv = vsub(vsub(A->v, vcross(p->r1, A->w)), B->v);
    jv = vadd(jv, vcross(p->r2, B->w));
    V3 lam = eff_mul(p->eff, jv);
    p->lambda = vadd(p->lambda, lam);
    point_apply(p, A, B, lam);
}
static int point_solve_position(OPoint* p, OBody* A, OBody* B, double beta) { /* solvePositionConstraint :271-316 */
    point_props(p, A, B);
    if (p->eff[15] == 0) return 0;
    V3 sep = vsub(vsub(vadd(B->com, p->r2), A->com), p->r1);
    if (vlen2(sep) < 1e-12) return 0;
    V3 lam = vscale(eff_mul(p->eff, sep), -beta);
    if (A->motion == CC_MOTION_DYNAMIC) {
        A->com = dof_lin(vsub(A->com, vscale(lam, A->invMass)), A->dof);
        add_rot_step(A, m3_mul(&p->iI1r1X, lam), -1.0);
    }
    if (B->motion == CC_MOTION_DYNAMIC) {
        B->com = dof_lin(vadd(B->com, vscale(lam, B->invMass)), B->dof);
        add_rot_step(B, m3_mul(&p->iI2r2X, lam), 1.0);
    }
    return 1;
}

/* ---- contact solver (src/constraints/contact-constraints.ts:906-1216) ---------- */
static void warm_start_contacts(oc_world* w, double ratio) {
    for (int i = 0; i < w->ncc; i++) {
        OCC* k = &w->cc[i];
        OBody* A = &w->b[k->a];
        OBody* B = &w->b[k->b];
        for (int j = 0; j < k->np; j++) {
            OCPoint* cp = &k->pts[j];
            if (axis_active(&cp->t1) || axis_active(&cp->t2)) {
                axis_warm(&cp->t1, A, B, k->invMassA, k->invMassB, k->t1, ratio);
                axis_warm(&cp->t2, A, B, k->invMassA, k->invMassB, k->t2, ratio);
            }
            axis_warm(&cp->n, A, B, k->invMassA, k->invMassB, k->normal, ratio);
        }
    }
}
static void solve_velocity_contact(oc_world* w, OCC* k) {
    OBody* A = &w->b[k->a];
    OBody* B = &w->b[k->b];
    if (A->motion != CC_MOTION_DYNAMIC && B->motion != CC_MOTION_DYNAMIC) return;
    for (int j = 0; j < k->np; j++) {
        OCPoint* cp = &k->pts[j];
        if (axis_active(&cp->t1) || axis_active(&cp->t2)) {
            double l1 = axis_total_lambda(&cp->t1, A, B, k->t1);
            double l2 = axis_total_lambda(&cp->t2, A, B, k->t2);
            double mf = k->friction * cp->n.lambda;
            double fm2 = l1 * l1 + l2 * l2;
            if (fm2 > mf * mf) { double sc = mf / sqrt(fm2); l1 *= sc; l2 *= sc; }
            axis_apply(&cp->t1, A, B, k->invMassA, k->invMassB, k->t1, l1);
            axis_apply(&cp->t2, A, B, k->invMassA, k->invMassB, k->t2, l2);
        }
    }
    for (int j = 0; j < k->np; j++) {
        OCPoint* cp = &k->pts[j];
        double tl = axis_total_lambda(&cp->n, A, B, k->normal);
        axis_apply(&cp->n, A, B, k->invMassA, k->invMassB, k->normal, dmax(0, tl));
    }
}
static int solve_position_contact(oc_world* w, OCC* k) {
    OBody* A = &w->b[k->a];
    OBody* B = &w->b[k->b];
    if (A->motion != CC_MOTION_DYNAMIC && B->motion != CC_MOTION_DYNAMIC) return 0;
    int any = 0;
    M3 RA = m3_from_quat(A->q), RB = m3_from_quat(B->q);
    M3 IA, IB;
    memset(&IA, 0, sizeof(IA));
    memset(&IB, 0, sizeof(IB));
    if (A->motion == CC_MOTION_DYNAMIC) IA = inv_inertia(A->invI, A->inertiaRot, &RA, A->dof);
    if (B->motion == CC_MOTION_DYNAMIC) IB = inv_inertia(B->invI, B->inertiaRot, &RB, B->dof);
    for (int j = 0; j < k->np; j++) {
        OCPoint* cp = &k->pts[j];
        V3 pA = vadd(A->com, m3_mul(&RA, cp->localA));
        V3 pB = vadd(B->com, m3_mul(&RB, cp->localB));
        double sep = vdot(vsub(pB, pA), k->normal) + w->s.penetration_slop;
        sep = dmax(sep, -w->s.max_penetration_distance);
        if (sep >= 0) continue;
        V3 mid = v3((pA.x + pB.x) * 0.5, (pA.y + pB.y) * 0.5, (pA.z + pB.z) * 0.5);
        V3 rA = vsub(mid, A->com), rB = vsub(mid, B->com);
        axis_props(&cp->n, A, B, k->invMassA, k->invMassB, &IA, &IB, rA, rB, k->normal, 0);
        if (axis_solve_position(&cp->n, A, B, k->invMassA, k->invMassB, k->normal, sep, w->s.baumgarte_factor)) any = 1;
    }
    return any;
}

/* ---- continuous collision detection (src/update.ts:896-1677) ---------------- */
/* getShapeInnerRadius (shapes.ts:522-525): sphere.ts:147-149, box.ts:333-335, capsule.ts:254-256,
 * cylinder.ts:261-263, convex-hull.ts:550-558 (the nearest face plane), scaled.ts:127-130 */
static double inner_radius(const oc_world* w, int si) {
    const cc_shape_desc* d = &w->shapes[si];
    double r = 0;
    if (d->type == CC_SHAPE_COMPOUND) { /* the smallest child's (compound.ts:283-291) */
        r = DBL_MAX;
        for (int k = 0; k < d->compound_num_children; k++) r = dmin(r, inner_radius(w, d->compound_first_child + k));
        return r;
    }
    switch (d->type) {
    case CC_SHAPE_SPHERE: case CC_SHAPE_CAPSULE: r = d->radius; break;
    case CC_SHAPE_BOX: r = dmin(d->half_extents[0], dmin(d->half_extents[1], d->half_extents[2])); break;
    case CC_SHAPE_CYLINDER: r = dmin(d->half_height, d->radius); break;
    case CC_SHAPE_CONVEX_HULL: {
        r = DBL_MAX;
        for (int f = 0; f < d->hull_num_faces; f++) r = dmin(r, -w->hulls.plane_constants[d->hull_first_face + f]);
        r = dmax(0.0, r);
        break;
    }
    default: break;
    }
    if (d->decorated & CC_DECO_SCALED)
        r = dmin(fabs(d->scale[0]), dmin(fabs(d->scale[1]), fabs(d->scale[2]))) * r;
    return r;
}

/* the motion a LINEAR_CAST body that is dynamic or kinematic makes this step, zero for any other
 * (calculateBodyMotion :985-996) */
static V3 body_motion(const OBody* b, double dt) {
    if (b->quality == CC_MOTION_QUALITY_LINEAR_CAST && b->motion != CC_MOTION_STATIC) return vscale(b->v, dt);
    return v3(0, 0, 0);
}

/* findCCDContacts :1357-1437 with the visitor :1110-1240 and the collector :1003-1107.  The
 * reference visits the bodies that a swept-AABB cast through its tree reaches (dbvt.ts:869-975)
 * and whose bounds, widened by A's half size, a ray along the relative displacement meets
 * (:1194-1206); here a body is visited when A's bounds swept over the relative displacement,
 * widened by 1e-3 m and a relative 1e-5, overlap its bounds, in index order (the visit order
 * decides only between hits of equal fraction-plus-slop).  The manifold block of :1376-1432
 * never runs: the collector's reset after every visit (:1238) clears hasHit, so the contact
 * point stays the collector's. */
static int ccd_swept_overlap(const OBody* A, V3 rel, const OBody* B) {
    const double lo[3] = {A->aabbMin.x + dmin(rel.x, 0.0), A->aabbMin.y + dmin(rel.y, 0.0), A->aabbMin.z + dmin(rel.z, 0.0)};
    const double hi[3] = {A->aabbMax.x + dmax(rel.x, 0.0), A->aabbMax.y + dmax(rel.y, 0.0), A->aabbMax.z + dmax(rel.z, 0.0)};
    const double blo[3] = {B->aabbMin.x, B->aabbMin.y, B->aabbMin.z}, bhi[3] = {B->aabbMax.x, B->aabbMax.y, B->aabbMax.z};
    for (int k = 0; k < 3; k++) {
        const double m = 1e-3 + 1e-5 * (fabs(lo[k]) + fabs(hi[k]) + fabs(blo[k]) + fabs(bhi[k]));
        if (lo[k] - m > bhi[k] || blo[k] - m > hi[k]) return 0;
    }
    return 1;
}
/* one leaf of body j (shape sb at pb, qb, sub-shape id sub) through the cast and the collector's
 * addHit :1045-1098 */
static void ccd_leaf(oc_world* w, OCcd* c, int j, int sa, V3 pa, Q4 qa, V3 da, uint32_t subA, int sb, V3 pb, Q4 qb,
                     uint32_t sub, double dt) {
    const OBody* A = &w->b[c->body];
    const OBody* B = &w->b[j];
    double fraction;
    V3 pA, pen;
    const int hit = w->shapes[sb].type == CC_SHAPE_PLANE
                        ? ocx_cast_convex_plane(&w->sh, sa, pa, qa, da, sb, pb, qb, w->s.collide_with_backfaces,
                                                c->fps, &fraction, &pA, &pen)
                        : ocx_cast_convex(&w->sh, sa, pa, qa, da, sb, pb, qb, w->s.collide_with_backfaces,
                                          c->fps, &fraction, &pA, &pen);
    if (!hit) return;
    const V3 n = vnormalize(pen);
    const double den = vdot(n, c->delta);
    if (den <= c->maxPen) return;
    const double fps = fraction + c->maxPen / den;
    if (fps >= c->fps) return;
    c->hit = j;
    c->sub1 = subA;
    c->sub2 = sub;
    c->fraction = fraction;
    c->fps = fps;
    c->normal = n;
    V3 mb = body_motion(B, dt);
    if (vlen2(mb) > 1e-12) c->point = vadd(pA, vscale(mb, fraction));
    else c->point = pA;
    c->friction = combine(A->friction, B->friction, A->fmode, B->fmode);
    c->restitution = combine(A->restitution, B->restitution, A->rmode, B->rmode);
}

static void ccd_find(oc_world* w, OCcd* c, double dt) {
    const OBody* A = &w->b[c->body];
    for (int j = 0; j < w->nb; j++) {
        const OBody* B = &w->b[j];
        if (j == c->body) continue;
        if (!w->s.object_layer_pairs[A->layer * CC_MAX_OBJECT_LAYERS + B->layer]) continue;
        if (!((A->groups & B->mask) != 0 && (B->groups & A->mask) != 0)) continue;
        V3 rel = c->delta;
        if (B->ccdIndex >= 0) {
            const OCcd* cb = &w->ccd[B->ccdIndex];
            if (c->body > cb->body) continue;
            rel = vsub(rel, cb->delta);
        }
        if (vlen2(rel) < c->thrSq) continue;
        if (!ccd_swept_overlap(A, rel, B)) continue;
        /* A's leaves: castCompoundVsShape (compound.ts: each child at the composed pose, the
         * displacement turned by A's rotation, its id pushed) or castTransformedVsShape
         * (transformed.ts:537-633, the same for the one transform); scaled and
         * offset-centre-of-mass pass through */
        const cc_shape_desc* DA = &w->shapes[A->shape];
        const int cmpA = DA->type == CC_SHAPE_COMPOUND;
        const int na = cmpA ? DA->compound_num_children : 1;
        const int abits = na > 1 ? 32 - __builtin_clz((unsigned)(na - 1)) : 0;
        for (int ka = 0; ka < na; ka++) {
        const int sa = cmpA ? DA->compound_first_child + ka : A->shape;
        const cc_shape_desc* LA = &w->shapes[sa];
        V3 pa = A->pos, da = rel;
        Q4 qa = A->q;
        if (LA->decorated & CC_DECO_TRANSFORMED) {
            const Q4 lq = {LA->local_quaternion[0], LA->local_quaternion[1], LA->local_quaternion[2], LA->local_quaternion[3]};
            pa = vadd(A->pos, qrotate(A->q, vload(LA->local_position)));
            qa = qmul(A->q, lq);
            da = qrotate(A->q, rel);
        }
        const uint32_t subA = cmpA ? sub_push(EMPTY_SUB, 0, ka, abits) : EMPTY_SUB;
        const cc_shape_desc* D = &w->shapes[B->shape];
        if (D->type != CC_SHAPE_COMPOUND) {
            /* castShapeVsTransformed / castShapeVsScaled / castShapeVsOffsetCenterOfMass
             * (transformed.ts:634-700, scaled.ts:483-560, offset-center-of-mass.ts:458-530): the
             * leaf at the body pose composed with the local transform, its scale in the support */
            V3 pb = B->pos;
            Q4 qb = B->q;
            if (D->decorated & CC_DECO_TRANSFORMED) {
                const Q4 lq = {D->local_quaternion[0], D->local_quaternion[1], D->local_quaternion[2], D->local_quaternion[3]};
                pb = vadd(B->pos, qrotate(B->q, vload(D->local_position)));
                qb = qmul(B->q, lq);
            }
            ccd_leaf(w, c, j, sa, pa, qa, da, subA, B->shape, pb, qb, EMPTY_SUB, dt);
            continue;
        }
        /* castShapeVsCompound compound.ts: every child at the body pose composed with its
         * transform, its sub-shape id pushed */
        const int nk = D->compound_num_children;
        const int nbits = nk > 1 ? 32 - __builtin_clz((unsigned)(nk - 1)) : 0;
        for (int k = 0; k < nk; k++) {
            const int ci = D->compound_first_child + k;
            const cc_shape_desc* K = &w->shapes[ci];
            const Q4 lq = {K->local_quaternion[0], K->local_quaternion[1], K->local_quaternion[2], K->local_quaternion[3]};
            ccd_leaf(w, c, j, sa, pa, qa, da, subA, ci, vadd(B->pos, qrotate(B->q, vload(K->local_position))), qmul(B->q, lq),
                     sub_push(EMPTY_SUB, 0, k, nbits), dt);
        }
        }
    }
}

/* compareCCDBodies :1596-1604 */
static oc_world* g_ccd_world;
static int cmp_ccd(const void* x, const void* y) {
    const OCcd* a = &g_ccd_world->ccd[*(const int*)x];
    const OCcd* b = &g_ccd_world->ccd[*(const int*)y];
    if (a->fps != b->fps) return a->fps < b->fps ? -1 : 1;
    return a->body - b->body;
}

/* applyCCD :1457-1593: one normal iteration and one friction iteration at the contact, through
 * the two constraint parts kept from the previous call */
static void apply_ccd(oc_world* w, const OCcd* c, OBody* A, OBody* B) {
    if (A->motion != CC_MOTION_DYNAMIC) return;
    const V3 atHit = vadd(A->com, vscale(c->delta, c->fraction));
    const V3 r1 = vsub(c->point, atHit);
    const V3 r2 = vsub(c->point, B->com);
    const V3 v1 = vadd(A->v, vcross(A->w, r1));
    const V3 v2 = B->motion != CC_MOTION_STATIC ? vadd(B->v, vcross(B->w, r2)) : v3(0, 0, 0);
    const V3 rv = vsub(v2, v1);
    const double nv = vdot(rv, c->normal);
    if (nv >= 0) return;
    double bias = 0;
    if (c->restitution > 0 && nv < -w->s.ccd_min_velocity_for_restitution) bias = c->restitution * nv;
    const double invMA = 1.0 * A->invMass;
    const double invMB = B->motion == CC_MOTION_DYNAMIC ? 1.0 * B->invMass : 0;
    M3 IA, IB;
    memset(&IB, 0, sizeof(IB));
    M3 RA = m3_from_quat(A->q);
    IA = inv_inertia(A->invI, A->inertiaRot, &RA, A->dof);
    if (B->motion == CC_MOTION_DYNAMIC) {
        M3 RB = m3_from_quat(B->q);
        IB = inv_inertia(B->invI, B->inertiaRot, &RB, B->dof);
    }
    axis_props(&w->ccdN, A, B, invMA, invMB, &IA, &IB, r1, r2, c->normal, bias);
    double tot = axis_total_lambda(&w->ccdN, A, B, c->normal);
    axis_apply(&w->ccdN, A, B, invMA, invMB, c->normal, tot);
    if (c->friction > 0) {
        const V3 tv = vsub(rv, vscale(c->normal, nv));
        const double tl2 = vlen2(tv);
        if (tl2 > 1e-12) {
            const V3 dir = vscale(tv, 1.0 / sqrt(tl2));
            const double maxF = c->friction * w->ccdN.lambda;
            axis_props(&w->ccdF, A, B, invMA, invMB, &IA, &IB, r1, r2, dir, 0);
            tot = axis_total_lambda(&w->ccdF, A, B, dir);
            axis_apply(&w->ccdF, A, B, invMA, invMB, dir, dmax(-maxF, dmin(maxF, tot)));
        }
    }
    clamp_lin(A); clamp_ang(A);
    if (B->motion == CC_MOTION_DYNAMIC) { clamp_lin(B); clamp_ang(B); }
}

/* findCCDContacts, sortCCDBodies :1607-1621 and resolveCCDContacts :1626-1677.  The body's
 * ccdBodyIndex still names its place before the sort, and resolveCCDContacts looks it up in the
 * sorted array (:1647-1649); that lookup is kept as it is. */
static void ccd_step(oc_world* w, double dt) {
    for (int i = 0; i < w->nccd; i++) ccd_find(w, &w->ccd[i], dt);
    w->ccdOrder = (int*)xrealloc(w->ccdOrder, sizeof(int) * (size_t)w->nccd);
    for (int i = 0; i < w->nccd; i++) w->ccdOrder[i] = i;
    g_ccd_world = w;
    sort_some(w->ccdOrder, (size_t)w->nccd, sizeof(int), cmp_ccd);
    for (int i = 0; i < w->nccd; i++) {
        OCcd* c = &w->ccd[w->ccdOrder[i]];
        OBody* A = &w->b[c->body];
        if (c->hit == -1) {
            if (A->motion == CC_MOTION_DYNAMIC) {
                A->pos = vadd(A->pos, c->delta);
                set_transform(w, A);
            }
            continue;
        }
        OBody* B = &w->b[c->hit];
        if (B->ccdIndex >= 0) {
            OCcd* cb = &w->ccd[w->ccdOrder[B->ccdIndex]];
            if (cb->hit != -1 && cb->fraction > c->fraction) {
                cb->hit = -1;
                cb->fps = c->fraction;
            }
        }
        apply_ccd(w, c, A, B);
        if (A->motion == CC_MOTION_DYNAMIC) {
            A->pos = vadd(A->pos, vscale(c->delta, c->fps));
            set_transform(w, A);
        }
        if (B->sleeping) body_wake(w, B, c->hit);
    }
}

/* ---- the step (src/update.ts:45-239) ---------------------------------------- */
void oc_step(oc_world* w, double dt) {
    memset(&w->ctr, 0, sizeof(w->ctr));
    int64_t epa0 = ocx_epa_calls;
    w->ncc = 0;
    w->nccd = 0; /* ccd.clear src/ccd.ts:107-123 */
    for (int i = 0; i < w->nb; i++) w->b[i].ccdIndex = -1;
    for (int i = 0; i < w->nct; i++) w->ct[i].processed = 0;
    /* accelerationIntegrationUpdate :245-319 */
    for (int i = 0; i < w->nb; i++) {
        OBody* b = &w->b[i];
        if (b->motion != CC_MOTION_DYNAMIC || b->sleeping) continue;
        double gs = w->s.gravity_enabled ? b->gravityFactor : 0;
        V3 g = vload(w->s.gravity);
        b->v.x += (b->force.x * b->invMass + g.x * gs) * dt;
        b->v.y += (b->force.y * b->invMass + g.y * gs) * dt;
        b->v.z += (b->force.z * b->invMass + g.z * gs) * dt;
        double ld = dmax(0, 1 - b->linDamp * dt);
        b->v = vscale(b->v, ld);
        double s2 = vlen2(b->v);
        if (s2 > b->maxLinV * b->maxLinV) b->v = vscale(b->v, b->maxLinV / sqrt(s2));
        b->v = dof_lin(b->v, b->dof);
        M3 R = m3_from_quat(b->q);
        M3 I = inv_inertia(b->invI, b->inertiaRot, &R, b->dof);
        b->w = vadd(b->w, vscale(m3_mul(&I, b->torque), dt));
        double ad = dmax(0, 1 - b->angDamp * dt);
        b->w = vscale(b->w, ad);
        double a2 = vlen2(b->w);
        if (a2 > b->maxAngV * b->maxAngV) b->w = vscale(b->w, b->maxAngV / sqrt(a2));
        b->w = dof_lin(b->w, b->dof >> 3);
    }
    find_pairs(w);
    w->ctr.num_pairs = w->npairs;
    for (int p = 0; p < w->npairs; p++) {
        int ia = w->pairs[2 * p], ib = w->pairs[2 * p + 1];
        OBody* A = &w->b[ia];
        OBody* B = &w->b[ib];
        if (A->motion > B->motion || (A->motion == B->motion && ib < ia)) { int t = ia; ia = ib; ib = t; }
        int made = narrowphase_pair(w, ia, ib, dt);
        if (made) {
            if (w->b[ia].motion == CC_MOTION_DYNAMIC && w->b[ia].sleeping) body_wake(w, &w->b[ia], ia);
            if (w->b[ib].motion == CC_MOTION_DYNAMIC && w->b[ib].sleeping) body_wake(w, &w->b[ib], ib);
        }
    }
    /* destroyUnprocessedContacts: the constraint list refers to cache slots, remap them */
    {
        int* remap = (int*)xrealloc(NULL, sizeof(int) * (size_t)(w->nct ? w->nct : 1));
        int k = 0;
        w->nrem = 0;
        for (int i = 0; i < w->nct; i++) {
            if (w->ct[i].processed) { remap[i] = k; w->ct[k++] = w->ct[i]; }
            else { /* destroyContact :297-307 reports it to onContactRemoved */
                remap[i] = -1;
                if (w->nrem == w->crem) {
                    w->crem = w->crem ? 2 * w->crem : 256;
                    w->rem = (cc_contact_removed*)xrealloc(w->rem, sizeof(cc_contact_removed) * (size_t)w->crem);
                }
                cc_contact_removed* r = &w->rem[w->nrem++];
                r->body_a = w->ct[i].a; r->body_b = w->ct[i].b;
                r->sub_shape_a = w->ct[i].sa; r->sub_shape_b = w->ct[i].sb;
            }
        }
        w->nct = k;
        for (int i = 0; i < w->ncc; i++) w->cc[i].contact = remap[w->cc[i].contact];
        free(remap);
        cache_rebuild(w);
    }
    w->ctr.num_contacts = w->ncc;
    for (int i = 0; i < w->ncc; i++) w->ctr.num_points += w->cc[i].np;
    if (dt > 0) {
        for (int i = 0; i < w->npc; i++) { /* wakeBodiesInUserConstraints :863-890 */
            OPoint* p = &w->pc[i];
            if (!p->enabled) continue;
            OBody* A = &w->b[p->a];
            OBody* B = &w->b[p->b];
            int act = !A->sleeping || !B->sleeping;
            int dyn = A->motion == CC_MOTION_DYNAMIC || B->motion == CC_MOTION_DYNAMIC;
            if (act && dyn) {
                if (A->sleeping && A->motion == CC_MOTION_DYNAMIC) body_wake(w, A, p->a);
                if (B->sleeping && B->motion == CC_MOTION_DYNAMIC) body_wake(w, B, p->b);
            }
        }
        build_islands(w);
        for (int i = 0; i < w->npc; i++) w->pc[i].sleeping = w->b[w->pc[i].a].sleeping && w->b[w->pc[i].b].sleeping;
        for (int i = 0; i < w->npc; i++)
            if (w->pc[i].enabled && !w->pc[i].sleeping) point_props(&w->pc[i], &w->b[w->pc[i].a], &w->b[w->pc[i].b]);
        if (w->s.warm_starting) {
            double ratio = (w->s.warm_starting && w->prevDt > 0) ? dt / w->prevDt : 0.0;
            warm_start_contacts(w, ratio);
            for (int i = 0; i < w->npc; i++) {
                OPoint* p = &w->pc[i];
                if (!p->enabled || p->sleeping || p->eff[15] == 0) continue;
                p->lambda = vscale(p->lambda, ratio);
                point_apply(p, &w->b[p->a], &w->b[p->b], p->lambda);
            }
        }
        w->prevDt = dt;
        g_sort_world = w;
        for (int il = 0; il < w->nisl; il++) {
            OIsland* I = &w->isl[il];
            if (I->nc == 0 && I->nu == 0) continue;
            sort_some(I->cons, (size_t)I->nu, sizeof(int), cmp_int); /* sortConstraintIds: priority 0, index */
            sort_some(I->contacts, (size_t)I->nc, sizeof(int), cmp_contact);
            for (int it = 0; it < I->velSteps; it++) {
                for (int k = 0; k < I->nc; k++) solve_velocity_contact(w, &w->cc[I->contacts[k]]);
                for (int k = 0; k < I->nu; k++) {
                    OPoint* p = &w->pc[I->cons[k]];
                    point_solve_velocity(p, &w->b[p->a], &w->b[p->b]);
                }
            }
        }
        for (int i = 0; i < w->ncc; i++) { /* storeAppliedImpulses :1061-1076 */
            OCC* k = &w->cc[i];
            OContact* c = &w->ct[k->contact];
            for (int j = 0; j < k->np; j++) {
                c->pts[j].ln = k->pts[j].n.lambda;
                c->pts[j].lt1 = k->pts[j].t1.lambda;
                c->pts[j].lt2 = k->pts[j].t2.lambda;
            }
        }
        for (int i = 0; i < w->nb; i++) { /* velocityIntegrationUpdate :896-976 */
            OBody* b = &w->b[i];
            if (b->motion == CC_MOTION_STATIC || b->sleeping) continue;
            if (b->motion == CC_MOTION_DYNAMIC) { clamp_lin(b); clamp_ang(b); }
            V3 rv = vscale(b->w, dt);
            double ang = vlen(rv);
            if (ang > 1e-6) {
                V3 axis = vscale(rv, 1 / ang);
                b->q = qnormalize(qmul(qaxis_angle(axis, ang), b->q));
            }
            V3 disp = vscale(b->v, dt);
            if (b->quality == CC_MOTION_QUALITY_LINEAR_CAST && b->motion == CC_MOTION_DYNAMIC) { /* :932-967 */
                const double inner = inner_radius(w, b->shape);
                const double thr = w->s.ccd_linear_cast_threshold * inner;
                if (vlen2(disp) > thr * thr) {
                    if (w->nccd == w->cccd) {
                        w->cccd = w->cccd ? 2 * w->cccd : 16;
                        w->ccd = (OCcd*)xrealloc(w->ccd, sizeof(OCcd) * (size_t)w->cccd);
                    }
                    OCcd* c = &w->ccd[w->nccd];
                    memset(c, 0, sizeof(*c));
                    b->ccdIndex = w->nccd++;
                    c->body = i;
                    c->hit = -1;
                    c->delta = disp;
                    c->fraction = 1.0;
                    c->fps = 1.0;
                    c->thrSq = thr * thr;
                    c->maxPen = dmin(w->s.penetration_slop, w->s.ccd_linear_cast_max_penetration * inner);
                    set_transform(w, b); /* setQuaternion: the position waits for the CCD pass */
                    continue;
                }
            }
            b->com = vadd(b->com, disp);
            pos_from_com(w, b);
            set_transform(w, b);
        }
        if (w->nccd > 0) ccd_step(w, dt);
        int maxPos = 0;
        for (int il = 0; il < w->nisl; il++) {
            OIsland* I = &w->isl[il];
            if (I->nc == 0 && I->nu == 0) continue;
            for (int it = 0; it < I->posSteps; it++) {
                int applied = 0;
                if (I->nc > 0)
                    for (int k = 0; k < I->nc; k++)
                        if (solve_position_contact(w, &w->cc[I->contacts[k]])) applied = 1;
                if (I->nu > 0) {
                    int ua = 0;
                    for (int k = 0; k < I->nu; k++) {
                        OPoint* p = &w->pc[I->cons[k]];
                        if (point_solve_position(p, &w->b[p->a], &w->b[p->b], w->s.baumgarte_factor)) ua = 1;
                    }
                    applied = ua || applied;
                }
                if (it + 1 > maxPos) maxPos = it + 1;
                if (!applied) break;
            }
        }
        w->ctr.position_sweeps = maxPos;
        for (int i = 0; i < w->nb; i++) { /* updateBodyPositions :322-330 */
            OBody* b = &w->b[i];
            if (b->motion == CC_MOTION_STATIC || b->sleeping) continue;
            pos_from_com(w, b);
            set_transform(w, b);
        }
        if (w->s.allow_sleeping) { /* checkIslandSleep :392-423 */
            double maxMove = w->s.point_velocity_sleep_threshold * w->s.time_before_sleep;
            for (int il = 0; il < w->nisl; il++) {
                OIsland* I = &w->isl[il];
                int all = 1;
                for (int k = 0; k < I->nb; k++) {
                    OBody* b = &w->b[w->active[I->bodies[k]]];
                    if (b->motion != CC_MOTION_DYNAMIC) continue;
                    int can = update_sleep_state(w, b, dt, maxMove, w->s.time_before_sleep);
                    all = all && can;
                }
                if (all) {
                    for (int k = 0; k < I->nb; k++) {
                        int bi = w->active[I->bodies[k]];
                        OBody* b = &w->b[bi];
                        if (b->motion == CC_MOTION_DYNAMIC) body_sleep(w, b);
                    }
                }
            }
        }
    }
    for (int i = 0; i < w->nb; i++) { /* clear forces */
        OBody* b = &w->b[i];
        if (b->motion == CC_MOTION_STATIC || b->sleeping) continue;
        b->force = v3(0, 0, 0);
        b->torque = v3(0, 0, 0);
    }
    w->ctr.num_islands = w->nisl;
    w->ctr.num_epa = ocx_epa_calls - epa0;
    for (int i = 0; i < w->nb; i++) {
        if (w->b[i].motion == CC_MOTION_DYNAMIC) {
            w->ctr.num_bodies++;
            if (!w->b[i].sleeping) w->ctr.num_awake++;
        }
    }
}

/* ---- accessors ------------------------------------------------------------- */
int32_t oc_get_ccd_bodies(const oc_world* w, cc_ccd_body* out, int32_t cap) {
    for (int i = 0; i < w->nccd && i < cap; i++) {
        const OCcd* c = &w->ccd[w->ccdOrder[i]];
        cc_ccd_body* o = &out[i];
        o->body_index = c->body;
        o->hit_body_index = c->hit;
        vstore(o->delta_position, c->delta);
        vstore(o->contact_normal, c->normal);
        vstore(o->contact_point, c->point);
        o->fraction = c->fraction;
        o->fraction_plus_slop = c->fps;
        o->linear_cast_threshold_sq = c->thrSq;
        o->max_penetration = c->maxPen;
        o->combined_friction = c->friction;
        o->combined_restitution = c->restitution;
        o->sub_shape_id_1 = c->hit >= 0 ? c->sub1 : 0u;
        o->sub_shape_id_2 = c->hit >= 0 ? c->sub2 : 0u;
    }
    return w->nccd;
}
int32_t oc_num_bodies(const oc_world* w) { return w->nb; }

void oc_get_bodies(const oc_world* w, cc_body_state* out) {
    /* island label = lowest body index among the island's bodies */
    int* lab = (int*)calloc((size_t)(w->nisl ? w->nisl : 1), sizeof(int));
    for (int i = 0; i < w->nisl; i++) lab[i] = INT32_MAX;
    for (int i = 0; i < w->nb; i++) if (w->b[i].islandIndex >= 0 && i < lab[w->b[i].islandIndex]) lab[w->b[i].islandIndex] = i;
    for (int i = 0; i < w->nb; i++) {
        const OBody* b = &w->b[i];
        cc_body_state* o = &out[i];
        vstore(o->position, b->pos);
        o->quaternion[0] = b->q.x; o->quaternion[1] = b->q.y; o->quaternion[2] = b->q.z; o->quaternion[3] = b->q.w;
        vstore(o->center_of_mass_position, b->com);
        vstore(o->linear_velocity, b->v);
        vstore(o->angular_velocity, b->w);
        vstore(o->aabb_min, b->aabbMin);
        vstore(o->aabb_max, b->aabbMax);
        o->sleeping = b->sleeping;
        o->island = b->islandIndex >= 0 ? lab[b->islandIndex] : -1;
    }
    free(lab);
}

void oc_set_bodies(oc_world* w, const cc_body_state* in) {
    for (int i = 0; i < w->nb; i++) {
        OBody* b = &w->b[i];
        b->pos = vload(in[i].position);
        b->q.x = in[i].quaternion[0]; b->q.y = in[i].quaternion[1]; b->q.z = in[i].quaternion[2]; b->q.w = in[i].quaternion[3];
        b->v = vload(in[i].linear_velocity);
        b->w = vload(in[i].angular_velocity);
        set_transform(w, b);
        if (!in[i].sleeping && b->sleeping && b->motion == CC_MOTION_DYNAMIC) body_wake(w, b, i);
    }
}

void oc_add_forces(oc_world* w, const double* force, const double* torque, int32_t wake) { /* addForce/addTorque rigid-body.ts:840-869 */
    for (int i = 0; i < w->nb; i++) {
        OBody* b = &w->b[i];
        if (b->motion != CC_MOTION_DYNAMIC) continue;
        int nz = 0;
        if (force) {
            b->force = vadd(b->force, vload(force + 3 * i));
            nz = nz || force[3 * i] != 0 || force[3 * i + 1] != 0 || force[3 * i + 2] != 0;
        }
        if (torque) {
            b->torque = vadd(b->torque, vload(torque + 3 * i));
            nz = nz || torque[3 * i] != 0 || torque[3 * i + 1] != 0 || torque[3 * i + 2] != 0;
        }
        if (wake && nz && b->sleeping) body_wake(w, b, i);
    }
}

/* unit entry points of the known-answer tests */
void oc_set_settings(oc_world* w, const cc_settings* s) { w->s = *s; }
int32_t oc_island_iterations(const oc_world* w, int32_t body, int32_t* vel, int32_t* pos) {
    if (body < 0 || body >= w->nb || w->b[body].islandIndex < 0) return 0;
    const OIsland* I = &w->isl[w->b[body].islandIndex];
    *vel = I->velSteps;
    *pos = I->posSteps;
    return 1;
}
uint32_t oc_sort_key(uint32_t a, uint32_t b, uint32_t sa, uint32_t sb) { return sort_key(a, b, sa, sb); }
void oc_contact_tangents(const double* n, double* t1, double* t2) {
    V3 a, b;
    contact_tangents(vload(n), &a, &b);
    vstore(t1, a);
    vstore(t2, b);
}

static int cmp_pair(const void* x, const void* y) {
    const int* a = (const int*)x;
    const int* b = (const int*)y;
    if (a[0] != b[0]) return a[0] - b[0];
    return a[1] - b[1];
}
int32_t oc_get_pairs(const oc_world* w, int32_t* out, int32_t cap) {
    int n = w->npairs < cap ? w->npairs : cap;
    for (int i = 0; i < n; i++) {
        int a = w->pairs[2 * i], b = w->pairs[2 * i + 1];
        out[2 * i] = a < b ? a : b;
        out[2 * i + 1] = a < b ? b : a;
    }
    sort_some(out, (size_t)n, 2 * sizeof(int32_t), cmp_pair);
    return w->npairs;
}

static int cmp_cout(const void* x, const void* y) {
    const cc_contact_out* a = (const cc_contact_out*)x;
    const cc_contact_out* b = (const cc_contact_out*)y;
    if (a->body_a != b->body_a) return a->body_a - b->body_a;
    if (a->body_b != b->body_b) return a->body_b - b->body_b;
    if (a->sub_shape_a != b->sub_shape_a) return a->sub_shape_a < b->sub_shape_a ? -1 : 1;
    if (a->sub_shape_b != b->sub_shape_b) return a->sub_shape_b < b->sub_shape_b ? -1 : 1;
    return 0;
}
int32_t oc_get_contacts(const oc_world* w, cc_contact_out* out, int32_t cap) {
    int n = w->ncc < cap ? w->ncc : cap;
    for (int i = 0; i < n; i++) {
        const OCC* k = &w->cc[i];
        cc_contact_out* o = &out[i];
        memset(o, 0, sizeof(*o));
        o->body_a = k->a; o->body_b = k->b; o->num_points = k->np;
        o->persisted = k->persisted;
        o->sub_shape_a = k->subA; o->sub_shape_b = k->subB;
        vstore(o->normal, k->normal);
        for (int j = 0; j < k->np; j++) {
            vstore(o->position_a[j], k->pts[j].posA);
            vstore(o->position_b[j], k->pts[j].posB);
            o->lambda_n[j] = k->pts[j].n.lambda;
            o->lambda_t1[j] = k->pts[j].t1.lambda;
            o->lambda_t2[j] = k->pts[j].t2.lambda;
        }
    }
    sort_some(out, (size_t)n, sizeof(cc_contact_out), cmp_cout);
    return w->ncc;
}

static int cmp_removed(const void* x, const void* y) {
    const cc_contact_removed* a = (const cc_contact_removed*)x;
    const cc_contact_removed* b = (const cc_contact_removed*)y;
    if (a->body_a != b->body_a) return a->body_a < b->body_a ? -1 : 1;
    if (a->body_b != b->body_b) return a->body_b < b->body_b ? -1 : 1;
    if (a->sub_shape_a != b->sub_shape_a) return a->sub_shape_a < b->sub_shape_a ? -1 : 1;
    if (a->sub_shape_b != b->sub_shape_b) return a->sub_shape_b < b->sub_shape_b ? -1 : 1;
    return 0;
}

int32_t oc_get_removed_contacts(const oc_world* w, cc_contact_removed* out, int32_t cap) {
    int n = w->nrem < cap ? w->nrem : cap;
    if (n > 0) {
        cc_contact_removed* t = (cc_contact_removed*)xrealloc(NULL, sizeof(cc_contact_removed) * (size_t)w->nrem);
        memcpy(t, w->rem, sizeof(cc_contact_removed) * (size_t)w->nrem);
        sort_some(t, (size_t)w->nrem, sizeof(cc_contact_removed), cmp_removed);
        memcpy(out, t, sizeof(cc_contact_removed) * (size_t)n);
        free(t);
    }
    return w->nrem;
}

static int cmp_cached(const void* x, const void* y) {
    const cc_cached_contact* a = (const cc_cached_contact*)x;
    const cc_cached_contact* b = (const cc_cached_contact*)y;
    if (a->body_a != b->body_a) return a->body_a < b->body_a ? -1 : 1;
    if (a->body_b != b->body_b) return a->body_b < b->body_b ? -1 : 1;
    if (a->sub_shape_a != b->sub_shape_a) return a->sub_shape_a < b->sub_shape_a ? -1 : 1;
    if (a->sub_shape_b != b->sub_shape_b) return a->sub_shape_b < b->sub_shape_b ? -1 : 1;
    return 0;
}

/* world.contacts as plain records (src/contacts.ts:9-58): the cache the next step matches */
int32_t oc_get_contact_cache(const oc_world* w, cc_cached_contact* out, int32_t cap, double* prev_dt) {
    if (prev_dt) *prev_dt = w->prevDt;
    if (!out || cap <= 0) return w->nct;
    cc_cached_contact* t = (cc_cached_contact*)xrealloc(NULL, sizeof(cc_cached_contact) * (size_t)(w->nct ? w->nct : 1));
    for (int k = 0; k < w->nct; k++) {
        const OContact* c = &w->ct[k];
        cc_cached_contact* o = &t[k];
        memset(o, 0, sizeof(*o));
        o->body_a = c->a;
        o->body_b = c->b;
        o->sub_shape_a = c->sa;
        o->sub_shape_b = c->sb;
        o->num_points = c->numPoints;
        for (int j = 0; j < c->numPoints && j < 4; j++) {
            const OCached* p = &c->pts[j];
            o->local_a[j][0] = p->p1.x; o->local_a[j][1] = p->p1.y; o->local_a[j][2] = p->p1.z;
            o->local_b[j][0] = p->p2.x; o->local_b[j][1] = p->p2.y; o->local_b[j][2] = p->p2.z;
            o->lambda_n[j] = p->ln;
            o->lambda_t1[j] = p->lt1;
            o->lambda_t2[j] = p->lt2;
        }
    }
    sort_some(t, (size_t)w->nct, sizeof(cc_cached_contact), cmp_cached);
    const int n = w->nct < cap ? w->nct : cap;
    if (n > 0) memcpy(out, t, sizeof(cc_cached_contact) * (size_t)n);
    free(t);
    return w->nct;
}

/* motionProperties.sleepTestSpheres / sleepTestTimer of every body */
void oc_get_sleep_state(const oc_world* w, cc_body_sleep* out) {
    for (int i = 0; i < w->nb; i++) {
        const OBody* b = &w->b[i];
        out[i].timer = b->sleepTimer;
        for (int k = 0; k < 3; k++) {
            out[i].center[k][0] = b->sleepCenter[k].x;
            out[i].center[k][1] = b->sleepCenter[k].y;
            out[i].center[k][2] = b->sleepCenter[k].z;
            out[i].radius[k] = b->sleepRadius[k];
        }
    }
}

/* PointConstraintPart totalLambda of every point constraint */
void oc_get_point_impulses(const oc_world* w, double* out) {
    for (int k = 0; k < w->npc; k++) {
        out[3 * k] = w->pc[k].lambda.x;
        out[3 * k + 1] = w->pc[k].lambda.y;
        out[3 * k + 2] = w->pc[k].lambda.z;
    }
}

void oc_get_counters(const oc_world* w, cc_step_counters* out) { *out = w->ctr; }

int32_t oc_broadphase_aabbs(const float* mn, const float* mx, const int32_t* motion, const int32_t* sleeping,
                            const int32_t* layer, int32_t n, double margin, const cc_settings* s, int32_t* out,
                            int32_t cap) {
    OBody* b = (OBody*)calloc((size_t)(n ? n : 1), sizeof(OBody));
    for (int i = 0; i < n; i++) {
        b[i].aabbMin = v3(mn[3 * i], mn[3 * i + 1], mn[3 * i + 2]);
        b[i].aabbMax = v3(mx[3 * i], mx[3 * i + 1], mx[3 * i + 2]);
        b[i].motion = motion[i]; b[i].sleeping = sleeping[i]; b[i].layer = layer[i];
        b[i].groups = b[i].mask = 0xffffffffu;
    }
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        if (!queries(&b[i])) continue;
        V3 amn = vsub(b[i].aabbMin, v3(margin, margin, margin)), amx = vadd(b[i].aabbMax, v3(margin, margin, margin));
        for (int j = 0; j < n; j++) {
            if (!pair_ok(s, &b[i], i, &b[j], j)) continue;
            if (!overlap(amn, amx, b[j].aabbMin, b[j].aabbMax)) continue;
            if (cnt < cap) { out[2 * cnt] = i < j ? i : j; out[2 * cnt + 1] = i < j ? j : i; }
            cnt++;
        }
    }
    free(b);
    sort_some(out, (size_t)(cnt < cap ? cnt : cap), 2 * sizeof(int32_t), cmp_pair);
    return cnt;
}
