/*
 * refcpu_collide.c — FP64 restatement of crashcat's convex narrowphase:
 * support functions, GJK closest points, EPA, supporting faces, polygon
 * clipping, manifold generation and contact pruning.  Oracle (test
 * infrastructure); citations are to the reference's files.
 */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include "refcpu_internal.h"

int64_t ocx_epa_calls = 0;

/* ======================================================================== */
/* support functions (src/collision/support.ts and the shape files)       */
/* ======================================================================== */
enum { SK_ZERO, SK_SPHERE_R, SK_BOX, SK_CAP_NC, SK_CAP_R, SK_HULL, SK_TRI, SK_CYL };
typedef struct {
    int kind;
    V3 ta, tb, tc; /* SK_TRI: the triangle's vertices */
    double r, hh;
    V3 he;
    const double* pts;
    int np;
    double convexRadius;
    int transformed;
    M3 R;
    V3 t;
    int addRadius;
    double addR;
    int hscaled; /* a scaled hull's points, scaled as they are read (convexHullWithConvexScaledGetSupport) */
    V3 hs;
} OSup;

#define MODE_INCLUDE 0
#define MODE_EXCLUDE 1

/* The scale a decorated shape's leaf is collided with (scaled.ts:296-330 accumulates it down the
 * chain); 1 for a shape without a scaled decorator. */
static V3 shape_scale(const cc_shape_desc* d) {
    return (d->decorated & CC_DECO_SCALED) ? vload(d->scale) : v3(1, 1, 1);
}

/* getShapeSupportFunction with the shape's scale: sphere.ts:195-263 (uniform, |scale.x|),
 * support.ts:336-368 (box: |scale| per axis, convex radius by the smallest |scale| clamped to
 * the default), capsule.ts:305-404 (uniform), convex-hull.ts:979-1010 (scaled points, the
 * scaled shrunk points and radius), cylinder.ts:323-419 (uniform) */
static void sup_for_shape(OSup* s, const OShapes* sh, int si, int mode) {
    const cc_shape_desc* d = &sh->shapes[si];
    memset(s, 0, sizeof(*s));
    const V3 sc = shape_scale(d);
    const double ax = fabs(sc.x), ay = fabs(sc.y), az = fabs(sc.z);
    const double minScale = dmin(ax, dmin(ay, az));
    switch (d->type) {
    case CC_SHAPE_SPHERE:
        if (mode == MODE_INCLUDE) { s->kind = SK_SPHERE_R; s->r = ax * d->radius; s->convexRadius = 0; }
        else { s->kind = SK_ZERO; s->convexRadius = ax * d->radius; }
        break;
    case CC_SHAPE_BOX: {
        s->kind = SK_BOX;
        const V3 he = v3(ax * d->half_extents[0], ay * d->half_extents[1], az * d->half_extents[2]);
        if (mode == MODE_EXCLUDE) {
            double cr = dmin(d->convex_radius * minScale, 0.05); /* clamped to DEFAULT_CONVEX_RADIUS */
            s->he = v3(dmax(0, he.x - cr), dmax(0, he.y - cr), dmax(0, he.z - cr));
            s->convexRadius = cr;
        } else {
            s->he = he;
            s->convexRadius = 0;
        }
        break;
    }
    case CC_SHAPE_CAPSULE:
        s->hh = ax * d->half_height;
        if (mode == MODE_INCLUDE) { s->kind = SK_CAP_R; s->r = ax * d->radius; s->convexRadius = 0; }
        else { s->kind = SK_CAP_NC; s->convexRadius = ax * d->radius; }
        break;
    case CC_SHAPE_CYLINDER: /* setCylinderWithConvexSupport :378-397, setCylinderNoConvexSupport :323-340 */
        s->kind = SK_CYL;
        if (mode == MODE_INCLUDE) { s->hh = ax * d->half_height; s->r = ax * d->radius; s->convexRadius = 0; }
        else {
            const double cr = ax * d->convex_radius;
            s->hh = ax * d->half_height - cr;
            s->r = ax * d->radius - cr;
            s->convexRadius = cr;
        }
        break;
    case CC_SHAPE_CONVEX_HULL:
        s->kind = SK_HULL;
        s->np = d->hull_num_points;
        if (mode == MODE_INCLUDE || d->convex_radius == 0) {
            s->pts = sh->hulls->points + 3 * d->hull_first_point;
            s->convexRadius = 0;
            s->hscaled = sc.x != 1 || sc.y != 1 || sc.z != 1;
            s->hs = sc;
        } else {
            s->pts = sh->hulls->shrunk_points + 3 * d->hull_first_point;
            s->convexRadius = d->convex_radius * minScale;
        }
        break;
    default:
        fprintf(stderr, "refcpu: unsupported shape type %d\n", d->type);
        abort();
    }
}

static V3 sup_base(const OSup* s, V3 d) {
    switch (s->kind) {
    case SK_ZERO:
        return v3(0, 0, 0);
    case SK_SPHERE_R: { /* sphere.ts:214-231 */
        double l2 = vlen2(d);
        if (l2 > 0) { double sc = s->r / sqrt(l2); return vscale(d, sc); }
        return v3(0, 0, 0);
    }
    case SK_BOX: /* support.ts:319-323 */
        return v3(d.x >= 0 ? s->he.x : -s->he.x, d.y >= 0 ? s->he.y : -s->he.y, d.z >= 0 ? s->he.z : -s->he.z);
    case SK_CAP_NC: /* capsule.ts:291-297 */
        return v3(0, d.y > 0 ? s->hh : -s->hh, 0);
    case SK_CYL: { /* cylinderNoConvexGetSupport / cylinderWithConvexGetSupport cylinder.ts:295-311, 353-367 */
        double hl = sqrt(d.x * d.x + d.z * d.z);
        V3 o = v3(0, d.y >= 0 ? s->hh : -s->hh, 0);
        if (hl > 0) { double sc = s->r / hl; o.x = d.x * sc; o.z = d.z * sc; }
        return o;
    }
    case SK_CAP_R: { /* capsule.ts:330-345 */
        double l2 = vlen2(d);
        if (l2 > 0) {
            double sc = s->r / sqrt(l2);
            return v3(d.x * sc, d.y * sc + (d.y > 0 ? s->hh : -s->hh), d.z * sc);
        }
        return v3(0, s->hh, 0);
    }
    case SK_HULL: { /* convex-hull.ts:588-608, 639-658, 698-716 */
        double best = -INFINITY;
        V3 out = v3(0, 0, 0);
        for (int i = 0; i < s->np; i++) {
            V3 p = vload(s->pts + 3 * i);
            if (s->hscaled) p = v3(p.x * s->hs.x, p.y * s->hs.y, p.z * s->hs.z);
            double dt = vdot(p, d);
            if (dt > best) { best = dt; out = p; }
        }
        return out;
    }
    case SK_TRI: { /* triangleGetSupport support.ts:83-118 */
        double d1 = s->ta.x * d.x + s->ta.y * d.y + s->ta.z * d.z;
        double d2 = s->tb.x * d.x + s->tb.y * d.y + s->tb.z * d.z;
        double d3 = s->tc.x * d.x + s->tc.y * d.y + s->tc.z * d.z;
        if (d1 > d2) return d1 > d3 ? s->ta : s->tc;
        return d2 > d3 ? s->tb : s->tc;
    }
    }
    return v3(0, 0, 0);
}

/* TransformedSupport (support.ts:236-261) and AddConvexRadiusSupport (support.ts:288-303) */
static V3 sup_get(const OSup* s, V3 d) {
    V3 out;
    if (s->transformed) {
        V3 ld = m3_mul_t(&s->R, d);
        V3 p = sup_base(s, ld);
        out = vadd(m3_mul(&s->R, p), s->t);
    } else {
        out = sup_base(s, d);
    }
    if (s->addRadius) {
        double l2 = vlen2(d);
        if (l2 > 0) {
            double sc = s->addR / sqrt(l2);
            out = vadd(out, vscale(d, sc));
        }
    }
    return out;
}

/* ======================================================================== */
/* barycentric coordinates (src/collision/closest-points.ts:19-140)          */
/* ======================================================================== */
typedef struct { double u, v, w; int valid; } Bary;

static Bary bary2(V3 a, V3 b, double tol2) {
    Bary r = {0, 0, 0, 0};
    V3 ab = vsub(b, a);
    double den = vlen2(ab);
    if (den < tol2) {
        if (vlen2(a) < vlen2(b)) { r.u = 1; r.v = 0; }
        else { r.u = 0; r.v = 1; }
        r.valid = 0;
        return r;
    }
    r.v = -vdot(a, ab) / den;
    r.u = 1.0 - r.v;
    r.valid = 1;
    return r;
}

static Bary bary3(V3 a, V3 b, V3 c, double tol2) {
    Bary r = {0, 0, 0, 0};
    V3 ab = vsub(b, a), ac = vsub(c, a), bc = vsub(c, b);
    double d00 = vlen2(ab), d11 = vlen2(ac), d22 = vlen2(bc);
    if (d00 <= d22) {
        double d01 = vdot(ab, ac);
        double den = d00 * d11 - d01 * d01;
        if (fabs(den) < 1.0e-12) {
            if (d00 > d11) { Bary o = bary2(a, b, tol2); r.u = o.u; r.v = o.v; r.w = 0; }
            else { Bary o = bary2(a, c, tol2); r.u = o.u; r.w = o.v; r.v = 0; }
            r.valid = 0;
            return r;
        }
        double a0 = vdot(a, ab), a1 = vdot(a, ac);
        r.v = (d01 * a1 - d11 * a0) / den;
        r.w = (d01 * a0 - d00 * a1) / den;
        r.u = 1.0 - r.v - r.w;
    } else {
        double d12 = vdot(ac, bc);
        double den = d11 * d22 - d12 * d12;
        if (fabs(den) < 1.0e-12) {
            if (d11 > d22) { Bary o = bary2(a, c, tol2); r.u = o.u; r.w = o.v; r.v = 0; }
            else { Bary o = bary2(b, c, tol2); r.v = o.u; r.w = o.v; r.u = 0; }
            r.valid = 0;
            return r;
        }
        double c1 = vdot(c, ac), c2 = vdot(c, bc);
        r.u = (d22 * c1 - d12 * c2) / den;
        r.v = (d11 * c2 - d12 * c1) / den;
        r.w = 1.0 - r.u - r.v;
    }
    r.valid = 1;
    return r;
}

/* ======================================================================== */
/* GJK (src/collision/gjk.ts)                                               */
/* ======================================================================== */
typedef struct { V3 y, p, q; } SPt;
typedef struct { SPt pt[4]; int size; } Simplex;
typedef struct { V3 point; int set; } CPRes;

static double clamp01(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }

/* computeClosestPointOnLine gjk.ts:33-55 */
static CPRes cp_line(V3 a, V3 b, double tol2) {
    Bary bc = bary2(a, b, tol2);
    CPRes r;
    if (bc.v <= 0.0) { r.point = a; r.set = 1; }
    else if (bc.u <= 0.0) { r.point = b; r.set = 2; }
    else { r.point = vadd(vscale(a, bc.u), vscale(b, bc.v)); r.set = 3; }
    return r;
}

/* computeClosestPointOnTriangle gjk.ts:57-309 */
static CPRes cp_triangle(V3 inA, V3 inB, V3 inC, int mustIncludeC, double tol2) {
    CPRes r;
    V3 ac0 = vsub(inC, inA), bc0 = vsub(inC, inB);
    int swapAC = vlen2(bc0) < vlen2(ac0);
    V3 a = swapAC ? inC : inA;
    V3 c = swapAC ? inA : inC;
    V3 ab = vsub(inB, a), ac = vsub(c, a);
    V3 n = vcross(ab, ac);
    double n2 = vlen2(n);
    if (n2 < 1.0e-10) {
        int set = 4;
        V3 best = inC;
        double bestD = vlen2(inC);
        if (!mustIncludeC) {
            double a2 = vlen2(inA);
            if (a2 < bestD) { set = 1; best = inA; bestD = a2; }
            double b2 = vlen2(inB);
            if (b2 < bestD) { set = 2; best = inB; bestD = b2; }
        }
        V3 ac2 = vsub(c, a);
        double acl = vlen2(ac2);
        if (acl > tol2) {
            double v = clamp01(-vdot(a, ac2) / acl);
            V3 q = vadd(a, vscale(ac2, v));
            double d2 = vlen2(q);
            if (d2 < bestD) { set = 5; best = q; bestD = d2; }
        }
        V3 bc2 = vsub(inC, inB);
        double bcl = vlen2(bc2);
        if (bcl > tol2) {
            double v = clamp01(-vdot(inB, bc2) / bcl);
            V3 q = vadd(inB, vscale(bc2, v));
            double d2 = vlen2(q);
            if (d2 < bestD) { set = 6; best = q; bestD = d2; }
        }
        if (!mustIncludeC) {
            V3 ab2 = vsub(inB, inA);
            double abl = vlen2(ab2);
            if (abl > tol2) {
                double v = clamp01(-vdot(inA, ab2) / abl);
                V3 q = vadd(inA, vscale(ab2, v));
                double d2 = vlen2(q);
                if (d2 < bestD) { set = 3; best = q; }
            }
        }
        r.set = set;
        r.point = best;
        return r;
    }
    V3 ap = vneg(a);
    double d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { r.set = swapAC ? 4 : 1; r.point = a; return r; }
    V3 bp = vneg(inB);
    double d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { r.set = 2; r.point = inB; return r; }
    if (d1 * d4 <= d3 * d2 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        r.set = swapAC ? 6 : 3;
        r.point = vadd(a, vscale(ab, v));
        return r;
    }
    V3 cp = vneg(c);
    double d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { r.set = swapAC ? 1 : 4; r.point = c; return r; }
    if (d5 * d2 <= d1 * d6 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        r.set = 5;
        r.point = vadd(a, vscale(ac, w));
        return r;
    }
    double d43 = d4 - d3, d56 = d5 - d6;
    if (d3 * d6 <= d5 * d4 && d43 >= 0.0 && d56 >= 0.0) {
        double w = d43 / (d43 + d56);
        r.set = swapAC ? 3 : 6;
        r.point = vadd(inB, vscale(vsub(c, inB), w));
        return r;
    }
    r.set = 7;
    V3 sum = vadd(vadd(a, inB), c);
    double sc = vdot(sum, n) / (3 * n2);
    r.point = vscale(n, sc);
    return r;
}

/* isOriginOutsideOfTrianglePlanes gjk.ts:314-394 */
static void origin_outside(int out[4], V3 a, V3 b, V3 c, V3 d, double tol) {
    V3 ab = vsub(b, a), ac = vsub(c, a), ad = vsub(d, a), bd = vsub(d, b), bc = vsub(c, b);
    V3 abac = vcross(ab, ac), acad = vcross(ac, ad), adab = vcross(ad, ab), bdbc = vcross(bd, bc);
    double sp[4] = {vdot(a, abac), vdot(a, acad), vdot(a, adab), vdot(b, bdbc)};
    double sd[4] = {vdot(ad, abac), vdot(ab, acad), vdot(ac, adab), -vdot(ab, bdbc)};
    if (sd[0] > 0 && sd[1] > 0 && sd[2] > 0 && sd[3] > 0) {
        for (int i = 0; i < 4; i++) out[i] = sp[i] >= -tol;
        return;
    }
    if (sd[0] < 0 && sd[1] < 0 && sd[2] < 0 && sd[3] < 0) {
        for (int i = 0; i < 4; i++) out[i] = sp[i] <= tol;
        return;
    }
    out[0] = out[1] = out[2] = out[3] = 1;
}

/* computeClosestPointOnTetrahedron gjk.ts:399-487 */
static CPRes cp_tetra(V3 a, V3 b, V3 c, V3 d, int mustIncludeD, double tol) {
    double tol2 = tol * tol;
    CPRes r;
    r.set = 15;
    r.point = v3(0, 0, 0);
    double best = INFINITY;
    int o[4];
    origin_outside(o, a, b, c, d, tol);
    if (o[0]) {
        if (mustIncludeD) { r.set = 1; r.point = a; }
        else r = cp_triangle(a, b, c, 0, tol2);
        best = vlen2(r.point);
    }
    if (o[1]) {
        CPRes t = cp_triangle(a, c, d, mustIncludeD, tol2);
        double d2 = vlen2(t.point);
        if (d2 < best) { best = d2; r.point = t.point; r.set = (t.set & 1) + ((t.set & 6) << 1); }
    }
    if (o[2]) {
        CPRes t = cp_triangle(a, b, d, mustIncludeD, tol2);
        double d2 = vlen2(t.point);
        if (d2 < best) { best = d2; r.point = t.point; r.set = (t.set & 3) + ((t.set & 4) << 1); }
    }
    if (o[3]) {
        CPRes t = cp_triangle(b, c, d, mustIncludeD, tol2);
        double d2 = vlen2(t.point);
        if (d2 < best) { r.point = t.point; r.set = t.set << 1; }
    }
    return r;
}

typedef struct { V3 point; double dist2; int set; int found; } CPSRes;

/* computeClosestPointToSimplex gjk.ts:512-596 */
static void cp_simplex(CPSRes* res, double prevDist2, int lastPointPartOfClosest, const Simplex* s) {
    CPRes cp;
    switch (s->size) {
    case 1: cp.set = 1; cp.point = s->pt[0].y; break;
    case 2: cp = cp_line(s->pt[0].y, s->pt[1].y, 1e-10); break;
    case 3: cp = cp_triangle(s->pt[0].y, s->pt[1].y, s->pt[2].y, lastPointPartOfClosest, 1e-10); break;
    default: cp = cp_tetra(s->pt[0].y, s->pt[1].y, s->pt[2].y, s->pt[3].y, lastPointPartOfClosest, 1e-5); break;
    }
    double d2 = vlen2(cp.point);
    if (d2 < prevDist2) {
        res->point = cp.point;
        res->dist2 = d2;
        res->set = cp.set;
        res->found = 1;
        return;
    }
    res->found = 0;
}

#define GJK_TOLERANCE 1e-5
#define GJK_MAX_ITERATIONS 100

typedef struct { V3 pointA, pointB, axis; double dist2; Simplex simplex; } GjkOut;

/* gjkClosestPoints gjk.ts:1294-1546 */
static void gjk_closest_points(GjkOut* out, const OSup* A, const OSup* B, double tolerance, V3 direction,
                               double maxDist2) {
    double tol2 = tolerance * tolerance;
    Simplex s;
    s.size = 0;
    CPSRes c;
    c.point = direction;
    c.dist2 = vlen2(direction);
    c.set = 0;
    c.found = 1;
    double prevDist2 = DBL_MAX;
    int it = 0;
    while (it++ < GJK_MAX_ITERATIONS) {
        V3 p = sup_get(A, c.point);
        V3 q = sup_get(B, vneg(c.point));
        V3 w = vsub(p, q);
        double dot = vdot(c.point, w);
        if (dot < 0.0 && dot * dot > c.dist2 * maxDist2) {
            out->dist2 = DBL_MAX;
            out->axis = c.point;
            out->simplex.size = 0;
            out->pointA = out->pointB = v3(0, 0, 0);
            return;
        }
        s.pt[s.size].y = w;
        s.pt[s.size].p = p;
        s.pt[s.size].q = q;
        s.size++;
        cp_simplex(&c, prevDist2, 1, &s);
        if (!c.found) { s.size--; break; }
        if (c.set == 0xf) { c.point = v3(0, 0, 0); c.dist2 = 0.0; break; }
        int ns = 0;
        for (int i = 0; i < s.size; i++) {
            if (c.set & (1 << i)) {
                if (ns != i) s.pt[ns] = s.pt[i];
                ns++;
            }
        }
        s.size = ns;
        if (c.dist2 <= tol2) { c.point = v3(0, 0, 0); c.dist2 = 0.0; break; }
        double ymax = 0;
        for (int i = 0; i < s.size; i++) ymax = dmax(ymax, vlen2(s.pt[i].y));
        if (c.dist2 <= GJK_TOLERANCE * ymax) { c.point = v3(0, 0, 0); c.dist2 = 0.0; break; }
        c.point = vneg(c.point);
        if (prevDist2 - c.dist2 <= GJK_TOLERANCE * prevDist2) break;
        prevDist2 = c.dist2;
    }
    out->pointA = out->pointB = v3(0, 0, 0);
    out->simplex = s;
    if (s.size == 0) {
        out->dist2 = DBL_MAX;
        out->axis = v3(0, 0, 0);
        return;
    }
    if (s.size == 1) {
        out->pointA = s.pt[0].p;
        out->pointB = s.pt[0].q;
    } else if (s.size == 2) {
        Bary b = bary2(s.pt[0].y, s.pt[1].y, 1e-10);
        out->pointA = vadd(vscale(s.pt[0].p, b.u), vscale(s.pt[1].p, b.v));
        out->pointB = vadd(vscale(s.pt[0].q, b.u), vscale(s.pt[1].q, b.v));
    } else if (s.size == 3) {
        Bary b = bary3(s.pt[0].y, s.pt[1].y, s.pt[2].y, 1e-10);
        out->pointA = vadd(vadd(vscale(s.pt[0].p, b.u), vscale(s.pt[1].p, b.v)), vscale(s.pt[2].p, b.w));
        out->pointB = vadd(vadd(vscale(s.pt[0].q, b.u), vscale(s.pt[1].q, b.v)), vscale(s.pt[2].q, b.w));
    }
    out->axis = c.point;
    out->dist2 = c.dist2;
}

/* ======================================================================== */
/* penetration depth (src/collision/penetration.ts)                         */
/* ======================================================================== */
enum { PD_NOT_COLLIDING = 0, PD_COLLIDING = 1, PD_INDETERMINATE = 2 };
typedef struct { int status; V3 axis, pointA, pointB; } PenDepth;

/* penetrationDepthStepGJK penetration.ts:35-110 */
static void pd_step_gjk(PenDepth* out, Simplex* outSimplex, const OSup* A, const OSup* B, double crA, double crB,
                        V3 dir, double tol) {
    double cr = crA + crB, cr2 = cr * cr;
    GjkOut g;
    gjk_closest_points(&g, A, B, tol, dir, cr2);
    if (g.dist2 > cr2) {
        out->pointA = out->pointB = out->axis = v3(0, 0, 0);
        out->status = PD_NOT_COLLIDING;
        return;
    }
    out->pointA = g.pointA;
    out->pointB = g.pointB;
    out->axis = g.axis;
    *outSimplex = g.simplex;
    if (g.dist2 > 0.0) {
        double vl = sqrt(g.dist2);
        out->pointA = vadd(out->pointA, vscale(out->axis, crA / vl));
        out->pointB = vadd(out->pointB, vscale(out->axis, -(crB / vl)));
        out->status = PD_COLLIDING;
        return;
    }
    out->status = PD_INDETERMINATE;
}

/* ---- EPA convex hull builder (src/collision/epa-convex-hull-builder.ts) --- */
#define EPA_MAX_TRIANGLES 256
#define EPA_MAX_EDGE_LENGTH 128
#define EPA_MIN_TRIANGLE_AREA 1e-10
#define EPA_BARYCENTRIC_EPSILON 1e-3
#define EPA_MAX_POINTS_TO_INCLUDE_ORIGIN_IN_HULL 32
#define EPA_MAX_POINTS 128
#define EPA_POINT_CAP 256

typedef struct { int nbT, nbE, start; } EEdge;
typedef struct {
    EEdge e[3];
    V3 normal, centroid;
    double closestLengthSq;
    double lambda[2];
    int lambdaRel0, interior, removed, inQueue;
} ETri;
typedef struct { int tri, edge, iter; } EStack;

typedef struct {
    ETri tris[EPA_MAX_TRIANGLES];
    int freeList[EPA_MAX_TRIANGLES * 4];
    int nfree;
    int queue[EPA_MAX_TRIANGLES * 4];
    int nq;
    const V3* pos;
    EStack stack[EPA_MAX_EDGE_LENGTH];
    EEdge edges[EPA_MAX_EDGE_LENGTH];
    int nedges;
} EHull;

static int tri_facing(const ETri* t, V3 p) { return vdot(t->normal, vsub(p, t->centroid)) > 0.0; }

static void heap_push(EHull* h, int ti) { /* pushTriangleToQueue :95-116 */
    h->queue[h->nq++] = ti;
    h->tris[ti].inQueue = 1;
    int cur = h->nq - 1;
    while (cur > 0) {
        int par = (cur - 1) >> 1;
        if (h->tris[h->queue[par]].closestLengthSq > h->tris[h->queue[cur]].closestLengthSq) {
            int t = h->queue[par]; h->queue[par] = h->queue[cur]; h->queue[cur] = t;
            cur = par;
        } else break;
    }
}

static int heap_pop(EHull* h) { /* popClosestTriangleFromQueue :381-418 */
    if (h->nq == 0) return -1;
    int* q = h->queue;
    int tmp = q[h->nq - 1]; q[h->nq - 1] = q[0]; q[0] = tmp;
    int count = h->nq - 1;
    int largest = 0;
    for (;;) {
        int child = (largest << 1) + 1;
        if (child >= count) break;
        int prev = largest;
        if (h->tris[q[largest]].closestLengthSq > h->tris[q[child]].closestLengthSq) largest = child;
        ++child;
        if (child < count && h->tris[q[largest]].closestLengthSq > h->tris[q[child]].closestLengthSq) largest = child;
        if (prev == largest) break;
        int t = q[prev]; q[prev] = q[largest]; q[largest] = t;
    }
    return q[--h->nq];
}

static void free_tri(EHull* h, int ti) { h->freeList[h->nfree++] = ti; }

/* createTriangle :179-345 */
static int create_tri(EHull* h, int i1, int i2, int i3) {
    if (h->nfree == 0) return -1;
    int ti = h->freeList[--h->nfree];
    ETri* t = &h->tris[ti];
    t->closestLengthSq = INFINITY;
    t->lambda[0] = t->lambda[1] = 0;
    t->lambdaRel0 = 0; t->interior = 0; t->removed = 0; t->inQueue = 0;
    t->e[0].start = i1; t->e[1].start = i2; t->e[2].start = i3;
    t->e[0].nbT = t->e[1].nbT = t->e[2].nbT = -1;
    V3 y0 = h->pos[i1], y1 = h->pos[i2], y2 = h->pos[i3];
    t->centroid = v3((y0.x + y1.x + y2.x) / 3.0, (y0.y + y1.y + y2.y) / 3.0, (y0.z + y1.z + y2.z) / 3.0);
    V3 y10 = vsub(y1, y0), y20 = vsub(y2, y0), y21 = vsub(y2, y1);
    double y20d = vlen2(y20), y21d = vlen2(y21);
    if (y20d < y21d) {
        V3 n = vcross(y10, y20);
        t->normal = n;
        double nl = vlen2(n);
        if (nl > EPA_MIN_TRIANGLE_AREA) {
            double cdn = vdot(t->centroid, n);
            t->closestLengthSq = fabs(cdn) * cdn / nl;
            double y10d = vlen2(y10), y10y20 = vdot(y10, y20);
            double det = y10d * y20d - y10y20 * y10y20;
            if (det > 0.0) {
                double y0y10 = vdot(y0, y10), y0y20 = vdot(y0, y20);
                double l0 = (y10y20 * y0y20 - y20d * y0y10) / det;
                double l1 = (y10y20 * y0y10 - y10d * y0y20) / det;
                t->lambda[0] = l0; t->lambda[1] = l1; t->lambdaRel0 = 1;
                if (l0 > -EPA_BARYCENTRIC_EPSILON && l1 > -EPA_BARYCENTRIC_EPSILON && l0 + l1 < 1.0 + EPA_BARYCENTRIC_EPSILON)
                    t->interior = 1;
            }
        }
    } else {
        V3 n = vcross(y10, y21);
        t->normal = n;
        double nl = vlen2(n);
        if (nl > EPA_MIN_TRIANGLE_AREA) {
            double cdn = vdot(t->centroid, n);
            t->closestLengthSq = fabs(cdn) * cdn / nl;
            double y10d = vlen2(y10), y10y21 = vdot(y10, y21);
            double det = y10d * y21d - y10y21 * y10y21;
            if (det > 0.0) {
                double y1y10 = vdot(y1, y10), y1y21 = vdot(y1, y21);
                double l0 = (y21d * y1y10 - y10y21 * y1y21) / det;
                double l1 = (y10y21 * y1y10 - y10d * y1y21) / det;
                t->lambda[0] = l0; t->lambda[1] = l1; t->lambdaRel0 = 0;
                if (l0 > -EPA_BARYCENTRIC_EPSILON && l1 > -EPA_BARYCENTRIC_EPSILON && l0 + l1 < 1.0 + EPA_BARYCENTRIC_EPSILON)
                    t->interior = 1;
            }
        }
    }
    return ti;
}

static void link_tri(EHull* h, int t1, int e1, int t2, int e2) {
    h->tris[t1].e[e1].nbT = t2; h->tris[t1].e[e1].nbE = e2;
    h->tris[t2].e[e2].nbT = t1; h->tris[t2].e[e2].nbE = e1;
}

static void hull_initialize(EHull* h, int i1, int i2, int i3) { /* initialize :347-371 */
    h->nfree = EPA_MAX_TRIANGLES;
    for (int i = 0; i < EPA_MAX_TRIANGLES; i++) h->freeList[i] = EPA_MAX_TRIANGLES - 1 - i;
    h->nq = 0;
    int t1 = create_tri(h, i1, i2, i3);
    int t2 = create_tri(h, i1, i3, i2);
    link_tri(h, t1, 0, t2, 2);
    link_tri(h, t1, 1, t2, 1);
    link_tri(h, t1, 2, t2, 0);
    heap_push(h, t1);
    heap_push(h, t2);
}

static int find_facing(EHull* h, V3 p) { /* findFacingTriangle :420-445 */
    int best = -1;
    double bestD = 0.0;
    for (int i = 0; i < h->nq; i++) {
        ETri* t = &h->tris[h->queue[i]];
        if (t->removed) continue;
        double dot = vdot(t->normal, vsub(p, t->centroid));
        if (dot > 0.0) {
            double d2 = dot * dot / vlen2(t->normal);
            if (d2 > bestD) { best = h->queue[i]; bestD = d2; }
        }
    }
    return best;
}

static void unlink_tri(EHull* h, int ti) { /* unlinkTriangle :451-466 */
    ETri* t = &h->tris[ti];
    for (int i = 0; i < 3; i++) {
        if (t->e[i].nbT >= 0) {
            h->tris[t->e[i].nbT].e[t->e[i].nbE].nbT = -1;
            t->e[i].nbT = -1;
        }
    }
    if (!t->inQueue) free_tri(h, ti);
}

static int find_edge(EHull* h, int facing, V3 vtx) { /* findEdge :468-539 */
    h->nedges = 0;
    h->tris[facing].removed = 1;
    int sp = 0;
    h->stack[0].tri = facing; h->stack[0].edge = 0; h->stack[0].iter = -1;
    int nextExpected = -1;
    for (;;) {
        EStack* cur = &h->stack[sp];
        if (++cur->iter >= 3) {
            unlink_tri(h, cur->tri);
            if (--sp < 0) break;
        } else {
            EEdge e = h->tris[cur->tri].e[(cur->edge + cur->iter) % 3];
            int n = e.nbT;
            if (n >= 0 && !h->tris[n].removed) {
                if (tri_facing(&h->tris[n], vtx)) {
                    h->tris[n].removed = 1;
                    if (sp + 1 >= EPA_MAX_EDGE_LENGTH) return 0; /* the reference's fixed stack has 128 entries */
                    sp++;
                    h->stack[sp].tri = n; h->stack[sp].edge = e.nbE; h->stack[sp].iter = 0;
                } else {
                    if (e.start != nextExpected && nextExpected != -1) return 0;
                    nextExpected = h->tris[n].e[e.nbE].start;
                    if (h->nedges >= EPA_MAX_EDGE_LENGTH) return 0;
                    h->edges[h->nedges++] = e;
                }
            }
        }
    }
    if (h->nedges > 0 && h->edges[0].start != nextExpected) return 0;
    return h->nedges >= 3;
}

/* addPoint :541-593; new triangles returned in out[] */
static int add_point(EHull* h, int facing, int idx, double closestDist2, int* out, int* nout) {
    V3 p = h->pos[idx];
    *nout = 0;
    if (!find_edge(h, facing, p)) return 0;
    int ne = h->nedges;
    for (int i = 0; i < ne; i++) {
        int in = (i + 1) % ne;
        int nt = create_tri(h, h->edges[i].start, h->edges[in].start, idx);
        if (nt < 0) return 0;
        out[(*nout)++] = nt;
        ETri* t = &h->tris[nt];
        if ((t->interior && t->closestLengthSq < closestDist2) || t->closestLengthSq < 0.0) heap_push(h, nt);
    }
    for (int i = 0; i < ne; i++) {
        int in = (i + 1) % ne;
        link_tri(h, out[i], 0, h->edges[i].nbT, h->edges[i].nbE);
        link_tri(h, out[i], 1, out[in], 2);
    }
    return 1;
}

typedef struct { V3 y[EPA_POINT_CAP], p[EPA_POINT_CAP], q[EPA_POINT_CAP]; int n; } EPts;

static int epa_add(EPts* P, const OSup* A, const OSup* B, V3 d) { /* addEpaSupportPoint penetration.ts:155-184 */
    V3 p = sup_get(A, d);
    V3 q = sup_get(B, vneg(d));
    int i = P->n++;
    P->y[i] = vsub(p, q);
    P->p[i] = p;
    P->q[i] = q;
    return i;
}

/* penetrationDepthStepEPA penetration.ts:214-614 */
static int pd_step_epa(PenDepth* out, const OSup* A, const OSup* B, double tol, const Simplex* s) {
    static __thread EHull H; /* per thread: worlds may step on several threads */
    static __thread EPts P;
    ocx_epa_calls++;
    P.n = 0;
    for (int i = 0; i < s->size; i++) { P.y[i] = s->pt[i].y; P.p[i] = s->pt[i].p; P.q[i] = s->pt[i].q; }
    P.n = s->size;
    if (P.n == 1) {
        P.n--;
        epa_add(&P, A, B, v3(0, 1, 0));
        epa_add(&P, A, B, v3(-1, -1, -1));
        epa_add(&P, A, B, v3(1, -1, -1));
        epa_add(&P, A, B, v3(0, -1, 1));
    } else if (P.n == 2) {
        V3 axis = vsub(P.y[1], P.y[0]);
        double al = vlen(axis);
        V3 an = v3(axis.x / al, axis.y / al, axis.z / al);
        V3 d1;
        if (fabs(an.x) > fabs(an.y)) {
            double pl = sqrt(an.x * an.x + an.z * an.z);
            d1 = v3(an.z / pl, 0.0, -an.x / pl);
        } else {
            double pl = sqrt(an.y * an.y + an.z * an.z);
            d1 = v3(0.0, an.z / pl, -an.y / pl);
        }
        const double c = cos(120.0 * M_PI / 180.0), sn = sin(120.0 * M_PI / 180.0), omc = 1 - c;
        double r00 = c + an.x * an.x * omc, r01 = an.x * an.y * omc - an.z * sn, r02 = an.x * an.z * omc + an.y * sn;
        double r10 = an.y * an.x * omc + an.z * sn, r11 = c + an.y * an.y * omc, r12 = an.y * an.z * omc - an.x * sn;
        double r20 = an.z * an.x * omc - an.y * sn, r21 = an.z * an.y * omc + an.x * sn, r22 = c + an.z * an.z * omc;
        V3 d2 = v3(r00 * d1.x + r01 * d1.y + r02 * d1.z, r10 * d1.x + r11 * d1.y + r12 * d1.z, r20 * d1.x + r21 * d1.y + r22 * d1.z);
        V3 d3 = v3(r00 * d2.x + r01 * d2.y + r02 * d2.z, r10 * d2.x + r11 * d2.y + r12 * d2.z, r20 * d2.x + r21 * d2.y + r22 * d2.z);
        epa_add(&P, A, B, d1);
        epa_add(&P, A, B, d2);
        epa_add(&P, A, B, d3);
    }
    H.pos = P.y;
    hull_initialize(&H, 0, 1, 2);
    int newT[EPA_MAX_EDGE_LENGTH];
    int nnew;
    for (int i = 3; i < P.n; i++) {
        int t = find_facing(&H, P.y[i]);
        if (t >= 0) {
            if (!add_point(&H, t, i, DBL_MAX, newT, &nnew)) { out->status = PD_NOT_COLLIDING; return 0; }
        }
    }
    for (;;) {
        if (H.nq == 0) { out->status = PD_NOT_COLLIDING; return 0; }
        int t = H.queue[0];
        if (H.tris[t].removed) {
            heap_pop(&H);
            if (H.nq == 0) { out->status = PD_NOT_COLLIDING; return 0; }
            free_tri(&H, t);
            continue;
        }
        if (H.tris[t].closestLengthSq >= 0.0) break;
        heap_pop(&H);
        int ni = epa_add(&P, A, B, H.tris[t].normal);
        V3 w = P.y[ni];
        if (!tri_facing(&H.tris[t], w) || !add_point(&H, t, ni, DBL_MAX, newT, &nnew)) {
            out->status = PD_NOT_COLLIDING;
            return 0;
        }
        free_tri(&H, t);
        if (H.nq == 0 || P.n >= EPA_MAX_POINTS_TO_INCLUDE_ORIGIN_IN_HULL) { out->status = PD_NOT_COLLIDING; return 0; }
    }
    double closestDist2 = DBL_MAX;
    int last = -1;
    int flip = 0;
    do {
        int t = heap_pop(&H);
        if (t < 0) { out->status = PD_NOT_COLLIDING; return 0; }
        ETri* T = &H.tris[t];
        if (T->removed) { free_tri(&H, t); continue; }
        if (T->closestLengthSq >= closestDist2) break;
        if (last >= 0) free_tri(&H, last);
        last = t;
        int ni = epa_add(&P, A, B, T->normal);
        V3 w = P.y[ni];
        double dot = vdot(T->normal, w);
        if (dot < 0.0) { out->status = PD_NOT_COLLIDING; return 0; }
        double d2 = dot * dot / vlen2(T->normal);
        if (d2 - T->closestLengthSq < T->closestLengthSq * tol) break;
        closestDist2 = dmin(closestDist2, d2);
        if (!tri_facing(T, w)) break;
        if (!add_point(&H, t, ni, closestDist2, newT, &nnew)) break;
        int defect = 0;
        for (int i = 0; i < nnew; i++) {
            ETri* nt = &H.tris[newT[i]];
            if (vdot(nt->normal, nt->centroid) < 0.0) { defect = 1; break; }
        }
        if (defect) {
            V3 nn = vneg(T->normal);
            V3 p2 = sup_get(A, nn);
            V3 q2 = sup_get(B, T->normal);
            double dot2 = vdot(nn, vsub(p2, q2));
            if (dot2 < dot) flip = 1;
            break;
        }
    } while (H.nq > 0 && P.n < EPA_MAX_POINTS);
    if (last < 0) { out->status = PD_NOT_COLLIDING; return 0; }
    ETri* L = &H.tris[last];
    double nl2 = vlen2(L->normal);
    double cdn = vdot(L->centroid, L->normal);
    V3 pn = vscale(L->normal, cdn / nl2);
    if (vlen2(pn) < 1e-10) { out->status = PD_NOT_COLLIDING; return 0; }
    if (flip) pn = vneg(pn);
    int i0 = L->e[0].start, i1 = L->e[1].start, i2 = L->e[2].start;
    V3 ca, cb;
    double l0 = L->lambda[0], l1 = L->lambda[1];
    if (L->lambdaRel0) {
        ca = vadd(vadd(P.p[i0], vscale(vsub(P.p[i1], P.p[i0]), l0)), vscale(vsub(P.p[i2], P.p[i0]), l1));
        cb = vadd(vadd(P.q[i0], vscale(vsub(P.q[i1], P.q[i0]), l0)), vscale(vsub(P.q[i2], P.q[i0]), l1));
    } else {
        ca = vadd(vadd(P.p[i1], vscale(vsub(P.p[i0], P.p[i1]), l0)), vscale(vsub(P.p[i2], P.p[i1]), l1));
        cb = vadd(vadd(P.q[i1], vscale(vsub(P.q[i0], P.q[i1]), l0)), vscale(vsub(P.q[i2], P.q[i1]), l1));
    }
    out->axis = pn;
    out->pointA = ca;
    out->pointB = cb;
    out->status = PD_COLLIDING;
    return 1;
}

/* ======================================================================== */
/* supporting faces                                                          */
/* ======================================================================== */
/* transformFace (src/utils/face.ts:30-84): rotation times scale, then the translation */
static void face_transform(OFace* f, V3 pos, Q4 q, V3 sc) {
    M3 R = m3_from_quat(q);
    for (int k = 0; k < 3; k++) { R.m[k] *= sc.x; R.m[3 + k] *= sc.y; R.m[6 + k] *= sc.z; }
    for (int i = 0; i < f->n; i++) f->v[i] = vadd(m3_mul(&R, f->v[i]), pos);
}

/* isScaleInsideOut (src/utils/face.ts:87-89) */
static int inside_out(V3 sc) { return sc.x * sc.y * sc.z < 0; }

/* the vertices of a face in the opposite order */
static void face_reverse(OFace* f) {
    for (int i = 0, j = f->n - 1; i < j; i++, j--) { V3 t = f->v[i]; f->v[i] = f->v[j]; f->v[j] = t; }
}

static void face_box(OFace* f, V3 d, V3 he) { /* box.ts:129-332; an inside-out scale lists the same face reversed */
    double hx = he.x, hy = he.y, hz = he.z;
    double ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
    f->n = 4;
    if (ax >= ay && ax >= az) {
        if (d.x < 0) { f->v[0] = v3(hx, -hy, -hz); f->v[1] = v3(hx, hy, -hz); f->v[2] = v3(hx, hy, hz); f->v[3] = v3(hx, -hy, hz); }
        else { f->v[0] = v3(-hx, -hy, -hz); f->v[1] = v3(-hx, -hy, hz); f->v[2] = v3(-hx, hy, hz); f->v[3] = v3(-hx, hy, -hz); }
    } else if (ay >= ax && ay >= az) {
        if (d.y < 0) { f->v[0] = v3(-hx, hy, -hz); f->v[1] = v3(-hx, hy, hz); f->v[2] = v3(hx, hy, hz); f->v[3] = v3(hx, hy, -hz); }
        else { f->v[0] = v3(-hx, -hy, -hz); f->v[1] = v3(hx, -hy, -hz); f->v[2] = v3(hx, -hy, hz); f->v[3] = v3(-hx, -hy, hz); }
    } else {
        if (d.z < 0) { f->v[0] = v3(-hx, -hy, hz); f->v[1] = v3(hx, -hy, hz); f->v[2] = v3(hx, hy, hz); f->v[3] = v3(-hx, hy, hz); }
        else { f->v[0] = v3(-hx, -hy, -hz); f->v[1] = v3(-hx, hy, -hz); f->v[2] = v3(hx, hy, -hz); f->v[3] = v3(hx, -hy, -hz); }
    }
}

static void face_capsule(OFace* f, V3 d, double hh, double r) { /* capsule.ts:193-253 */
    double len = sqrt(d.x * d.x + d.z * d.z);
    if (len == 0.0) { f->n = 0; return; }
    double sx = (r / len) * d.x, sz = (r / len) * d.z;
    V3 top = v3(-sx, hh, -sz), bot = v3(-sx, -hh, -sz);
    double pt = vdot(top, d), pb = vdot(bot, d);
    double dl = vlen(d);
    if (fabs(pt - pb) < 0.02 * dl) { f->n = 2; f->v[0] = top; f->v[1] = bot; }
    else f->n = 0;
}

static void face_hull(OFace* f, const OShapes* sh, const cc_shape_desc* s, V3 d) { /* convex-hull.ts:479-548 */
    const cc_hull_data* H = sh->hulls;
    const V3 sc = shape_scale(s);
    const V3 inv = v3(1 / sc.x, 1 / sc.y, 1 / sc.z); /* normals transform by the inverse scale */
    int f0 = s->hull_first_face;
    V3 n0 = vload(H->plane_normals + 3 * f0);
    n0 = v3(inv.x * n0.x, inv.y * n0.y, inv.z * n0.z);
    double bestDot = vdot(n0, d) / vlen(n0);
    int best = 0;
    for (int i = 1; i < s->hull_num_faces; i++) {
        V3 n = vload(H->plane_normals + 3 * (f0 + i));
        n = v3(inv.x * n.x, inv.y * n.y, inv.z * n.z);
        double dt = vdot(n, d) / vlen(n);
        if (dt < bestDot) { bestDot = dt; best = i; }
    }
    int first = H->face_first_vertex[f0 + best], nv = H->face_num_vertices[f0 + best];
    int maxv = 32 / 2;
    int delta = (nv + maxv - 1) / maxv;
    f->n = 0;
    if (inside_out(sc)) {
        for (int i = nv - 1; i >= 0; i -= delta) {
            int vi = H->face_vertex_indices[first + i];
            f->v[f->n++] = vload(H->points + 3 * (s->hull_first_point + vi));
        }
    } else {
        for (int i = 0; i < nv; i += delta) {
            int vi = H->face_vertex_indices[first + i];
            f->v[f->n++] = vload(H->points + 3 * (s->hull_first_point + vi));
        }
    }
}

/* getShapeSupportingFace (src/shapes/shapes.ts:489-509) */
/* getSupportingFace of a cylinder (cylinder.ts:167-240, cap template :43-53): the vertical edge
 * of the side facing away from dir, or the cap's octagon, turned so that a vertex points along
 * dir's horizontal part when dir is more than 5 degrees from the axis */
static void face_cylinder(OFace* f, V3 d, double hh, double r, int insideOut) {
    static const double c45 = 0.707106769;
    static const double ox[8] = {0.0, c45, 1.0, c45, 0.0, -c45, -1.0, -c45};
    static const double oz[8] = {1.0, c45, 0.0, -c45, -1.0, -c45, 0.0, c45};
    double xz = d.x * d.x + d.z * d.z, y2 = d.y * d.y;
    if (xz > y2) {
        double hl = sqrt(xz);
        double ff = -r / hl;
        double vx = d.x * ff, vz = d.z * ff;
        f->n = 2;
        f->v[0] = v3(vx, hh, vz);
        f->v[1] = v3(vx, -hh, vz);
        return;
    }
    double sx = d.y < 0 ? r : -r, sy = d.y < 0 ? hh : -hh, sz = r;
    double rc = 1, rs = 0;
    if (xz > 0.00765427 * y2) {
        double hl = sqrt(xz);
        rc = d.x / hl;
        rs = d.z / hl;
    }
    f->n = 8;
    for (int i = 0; i < 8; i++) {
        const int t = insideOut ? 7 - i : i;
        double px = sx * ox[t], py = sy * 1.0, pz = sz * oz[t];
        f->v[i] = v3(px * rc - pz * rs, py, px * rs + pz * rc);
    }
}

static void supporting_face(OFace* f, const OShapes* sh, int si, V3 dir, V3 pos, Q4 q) {
    const cc_shape_desc* s = &sh->shapes[si];
    const V3 sc = shape_scale(s);
    switch (s->type) {
    case CC_SHAPE_SPHERE: f->n = 0; return;
    case CC_SHAPE_BOX:
        face_box(f, dir, vload(s->half_extents));
        if (inside_out(sc)) face_reverse(f);
        break;
    case CC_SHAPE_CAPSULE: face_capsule(f, dir, s->half_height, s->radius); if (f->n == 0) return; break;
    case CC_SHAPE_CONVEX_HULL: face_hull(f, sh, s, dir); break;
    case CC_SHAPE_CYLINDER:
        face_cylinder(f, dir, s->half_height, s->radius, inside_out(sc));
        break;
    default: f->n = 0; return;
    }
    face_transform(f, pos, q, sc);
}

/* ======================================================================== */
/* collide                                                                   */
/* ======================================================================== */
/* v' = v + 2 * cross(q.xyz, cross(q.xyz, v) + q.w * v)   (convex.ts inline form) */
static V3 qrot2(Q4 q, V3 v) {
    double uvx = q.y * v.z - q.z * v.y + q.w * v.x;
    double uvy = q.z * v.x - q.x * v.z + q.w * v.y;
    double uvz = q.x * v.y - q.y * v.x + q.w * v.z;
    double uuvx = q.y * uvz - q.z * uvy, uuvy = q.z * uvx - q.x * uvz, uuvz = q.x * uvy - q.y * uvx;
    return v3(v.x + 2 * uuvx, v.y + 2 * uuvy, v.z + 2 * uuvz);
}

/* The pose a decorated shape's leaf is collided at: a transformed decorator composes the body pose
 * with its own, p + rotate(q, t) and q * tq (collideTransformedVsShape transformed.ts:370-420);
 * scaled and offset-centre-of-mass decorators leave it unchanged. */
static void leaf_pose(const cc_shape_desc* d, V3* p, Q4* q) {
    if (!(d->decorated & CC_DECO_TRANSFORMED)) return;
    const V3 t = qrotate(*q, vload(d->local_position));
    const Q4 lq = {d->local_quaternion[0], d->local_quaternion[1], d->local_quaternion[2], d->local_quaternion[3]};
    *p = vadd(*p, t);
    *q = qmul(*q, lq);
}

/* collideSphereVsSphere src/shapes/sphere.ts:319-410 (radii times |scale.x|) */
static void collide_sphere_sphere(const OShapes* sh, int sa, V3 pa, int sb, V3 pb, OHit* out) {
    double ra = sh->shapes[sa].radius * fabs(shape_scale(&sh->shapes[sa]).x);
    double rb = sh->shapes[sb].radius * fabs(shape_scale(&sh->shapes[sb]).x);
    V3 d = vsub(pb, pa);
    double d2 = vlen2(d), rs = ra + rb;
    out->hit = 0;
    if (d2 >= rs * rs) return;
    double dist = sqrt(d2);
    double pen = rs - dist;
    V3 n;
    if (dist > 0) { double inv = 1 / dist; n = vscale(d, inv); }
    else n = v3(0, 1, 0);
    out->hit = 1;
    out->status = 0;
    out->pointA = vadd(pa, vscale(n, ra));
    out->pointB = vsub(pb, vscale(n, rb));
    out->axis = n;
    out->penetration = pen;
    out->faceA.n = 0;
    out->faceB.n = 0;
}

/* collideConvexVsConvex + collideConvexVsConvexLocal src/shapes/convex.ts:257-641 */
static void collide_convex_convex(const OShapes* sh, int sa, V3 pa, Q4 qa, int sb, V3 pb, Q4 qb, double maxSep,
                                  double colTol, double penTol, OHit* out) {
    out->hit = 0;
    Q4 ia = {-qa.x, -qa.y, -qa.z, qa.w};
    Q4 rel;
    rel.x = ia.w * qb.x + ia.x * qb.w + ia.y * qb.z - ia.z * qb.y;
    rel.y = ia.w * qb.y - ia.x * qb.z + ia.y * qb.w + ia.z * qb.x;
    rel.z = ia.w * qb.z + ia.x * qb.y - ia.y * qb.x + ia.z * qb.w;
    rel.w = ia.w * qb.w - ia.x * qb.x - ia.y * qb.y - ia.z * qb.z;
    V3 relPos = qrot2(ia, vsub(pb, pa));

    OSup A, B;
    sup_for_shape(&A, sh, sa, MODE_EXCLUDE);
    sup_for_shape(&B, sh, sb, MODE_EXCLUDE);
    B.transformed = 1;
    B.R = m3_from_quat(rel);
    B.t = relPos;
    V3 axis0 = relPos;
    if (vlen2(axis0) <= 1e-12) axis0 = v3(1, 0, 0);
    double maxSepUse = maxSep;
    PenDepth pd;
    Simplex simplex;
    simplex.size = 0;
    pd_step_gjk(&pd, &simplex, &A, &B, A.convexRadius + maxSepUse, B.convexRadius, axis0, colTol);
    int status = 0;
    if (pd.status == PD_NOT_COLLIDING) return;
    if (pd.status == PD_INDETERMINATE) {
        maxSepUse = dmin(maxSepUse, 1.0);
        OSup A2, B2;
        sup_for_shape(&A2, sh, sa, MODE_INCLUDE);
        sup_for_shape(&B2, sh, sb, MODE_INCLUDE);
        A2.addRadius = 1;
        A2.addR = maxSepUse;
        B2.transformed = 1;
        B2.R = B.R;
        B2.t = relPos;
        if (!pd_step_epa(&pd, &A2, &B2, penTol, &simplex)) return;
        status = 1;
    }
    V3 dd = vsub(pd.pointA, pd.pointB);
    double pen = sqrt(vlen2(dd)) - maxSepUse;
    double al = sqrt(vlen2(pd.axis));
    if (al > 0.0) {
        double corr = maxSepUse / al;
        pd.pointA = vadd(pd.pointA, vscale(pd.axis, -corr));
    }
    out->hit = 1;
    out->status = status;
    out->pointA = vadd(qrot2(qa, pd.pointA), pa);
    out->pointB = vadd(qrot2(qa, pd.pointB), pa);
    out->axis = qrot2(qa, pd.axis);
    out->penetration = pen;
    supporting_face(&out->faceA, sh, sa, vneg(pd.axis), pa, qa);
    Q4 crel = {-rel.x, -rel.y, -rel.z, rel.w};
    V3 dirB = qrot2(crel, pd.axis);
    V3 posB = vadd(qrot2(qa, relPos), pa);
    Q4 qB;
    qB.x = qa.w * rel.x + qa.x * rel.w + qa.y * rel.z - qa.z * rel.y;
    qB.y = qa.w * rel.y - qa.x * rel.z + qa.y * rel.w + qa.z * rel.x;
    qB.z = qa.w * rel.z + qa.x * rel.y - qa.y * rel.x + qa.z * rel.w;
    qB.w = qa.w * rel.w - qa.x * rel.x - qa.y * rel.y - qa.z * rel.z;
    supporting_face(&out->faceB, sh, sb, dirB, posB, qB);
}

/* fixNormal src/collision/active-edges.ts:74-153: the triangle normal replaces the GJK/EPA
 * normal when the contact point is on an inactive edge or inside the triangle */
static V3 fix_normal(V3 v0, V3 v1, V3 v2, V3 tn, int act, V3 p, V3 n, V3 md) {
    if (act == 7) return n;
    double nl = vlen(n), tl = vlen(tn);
    if (vdot(md, n) * tl < vdot(md, tn) * nl) return n;
    if (act == 0) return tn;
    if (vdot(tn, n) > 0.999848 * nl * tl) return n;
    Bary b = bary3(vsub(v0, p), vsub(v1, p), vsub(v2, p), 1e-10);
    int edge;
    if (b.u > 1.0 - 1.0e-4) edge = 5;
    else if (b.v > 1.0 - 1.0e-4) edge = 3;
    else if (b.w > 1.0 - 1.0e-4) edge = 6;
    else if (b.u < 1.0e-4) edge = 2;
    else if (b.v < 1.0e-4) edge = 4;
    else if (b.w < 1.0e-4) edge = 1;
    else return tn;
    return (act & edge) ? n : tn;
}

static __thread int g_mesh_calls;
int ocx_mesh_calls(void) { return g_mesh_calls; }

/* collideConvexVsTriangleMesh src/shapes/triangle-mesh.ts:1026-1379 at unit scale: the convex
 * shape A's bounds in mesh space cull the BVH (closer child first), each triangle is moved to
 * A's space and collided by GJK (EPA when the cores overlap) against A's support, the normal is
 * corrected at inactive edges, and each hit goes to fn in traversal order. */
void ocx_collide_mesh(const OShapes* sh, int sa, V3 pa, Q4 qa, int sm, V3 pm, Q4 qm, const cc_settings* s,
                      double maxSep0, V3 moveDir, ocx_hit_fn fn, void* ctx) {
    const cc_shape_desc* A = &sh->shapes[sa];
    const cc_shape_desc* M = &sh->shapes[sm];
    const cc_mesh_data* md = sh->meshes;
    g_mesh_calls++;
    Q4 iqa = {-qa.x, -qa.y, -qa.z, qa.w};
    Q4 t21q = qmul(iqa, qm);
    V3 t21p = qrotate(iqa, vsub(pm, pa));
    Q4 iqm = {-qm.x, -qm.y, -qm.z, qm.w};
    Q4 qAinB = qmul(iqm, qa);
    V3 pAinB = qrotate(iqm, vsub(pa, pm));
    V3 e = v3(maxSep0, maxSep0, maxSep0);
    V3 b1mn = vsub(vload(A->aabb_min), e), b1mx = vadd(vload(A->aabb_max), e);
    M3 R = m3_from_quat(qAinB);
    V3 q2mn = v3(0, 0, 0), q2mx = v3(0, 0, 0);
    for (int c = 0; c < 8; c++) { /* box3.transformMat4 of the local bounds */
        V3 p = v3((c & 1) ? A->aabb_max[0] : A->aabb_min[0], (c & 2) ? A->aabb_max[1] : A->aabb_min[1],
                  (c & 4) ? A->aabb_max[2] : A->aabb_min[2]);
        V3 wp = vadd(m3_mul(&R, p), pAinB);
        if (c == 0) { q2mn = wp; q2mx = wp; }
        else {
            q2mn = v3(dmin(q2mn.x, wp.x), dmin(q2mn.y, wp.y), dmin(q2mn.z, wp.z));
            q2mx = v3(dmax(q2mx.x, wp.x), dmax(q2mx.y, wp.y), dmax(q2mx.z, wp.z));
        }
    }
    q2mn = vsub(q2mn, e);
    q2mx = vadd(q2mx, e);
    OSup SA;
    sup_for_shape(&SA, sh, sa, MODE_EXCLUDE);
    if (M->mesh_num_nodes == 0) return;
    const double* nodes = md->nodes + 8 * (size_t)M->mesh_first_node;
    const int32_t* tris = md->triangles + 4 * (size_t)M->mesh_first_triangle;
    const int ntri = M->mesh_num_triangles;
    const int nbits = ntri > 1 ? 32 - __builtin_clz((unsigned)(ntri - 1)) : 0; /* pushIndex sub-shape.ts:60-64 */
    const uint32_t mask = nbits >= 32 ? 0xffffffffu : ((1u << nbits) - 1u);
    V3 qc = v3((q2mn.x + q2mx.x) * 0.5, (q2mn.y + q2mx.y) * 0.5, (q2mn.z + q2mx.z) * 0.5);
    V3 mdl = qrotate(iqa, moveDir);
    static __thread int* stack = NULL;
    static __thread int scap = 0;
    int ns = 0;
    if (!scap) { scap = 256; stack = (int*)malloc(sizeof(int) * (size_t)scap); }
    stack[ns++] = 0;
    while (ns > 0) {
        const int o = stack[--ns];
        const double* nd = nodes + o;
        if (!(nd[0] <= q2mx.x && nd[3] >= q2mn.x && nd[1] <= q2mx.y && nd[4] >= q2mn.y && nd[2] <= q2mx.z &&
              nd[5] >= q2mn.z))
            continue;
        if (nd[7] < 0) {
            const int t0 = (int)nd[6], cnt = -((int)nd[7] + 1);
            for (int i = 0; i < cnt; i++) {
                const int ti = t0 + i;
                const int32_t* tr = tris + 4 * (size_t)ti;
                V3 a = vadd(qrotate(t21q, vload(md->positions + 3 * (size_t)tr[0])), t21p);
                V3 b = vadd(qrotate(t21q, vload(md->positions + 3 * (size_t)tr[1])), t21p);
                V3 c = vadd(qrotate(t21q, vload(md->positions + 3 * (size_t)tr[2])), t21p);
                V3 tmn = v3(dmin(dmin(a.x, b.x), c.x), dmin(dmin(a.y, b.y), c.y), dmin(dmin(a.z, b.z), c.z));
                V3 tmx = v3(dmax(dmax(a.x, b.x), c.x), dmax(dmax(a.y, b.y), c.y), dmax(dmax(a.z, b.z), c.z));
                if (tmx.x < b1mn.x || tmn.x > b1mx.x || tmx.y < b1mn.y || tmn.y > b1mx.y || tmx.z < b1mn.z ||
                    tmn.z > b1mx.z)
                    continue;
                V3 nrm = vcross(vsub(b, a), vsub(c, a));
                int back = vdot(nrm, a) > 0.0;
                if (!s->collide_with_backfaces && back) continue;
                V3 axis = vneg(nrm);
                if (vlen2(axis) < 1e-10) axis = v3(1, 0, 0);
                else axis = vnormalize(axis);
                double maxSep = maxSep0;
                OSup T;
                memset(&T, 0, sizeof(T));
                T.kind = SK_TRI;
                T.ta = a; T.tb = b; T.tc = c;
                PenDepth pd;
                Simplex simplex;
                simplex.size = 0;
                pd_step_gjk(&pd, &simplex, &SA, &T, SA.convexRadius + maxSep, 0.0, axis, s->collision_tolerance);
                if (pd.status == PD_NOT_COLLIDING) continue;
                int status = 0;
                if (pd.status == PD_INDETERMINATE) {
                    maxSep = dmin(maxSep, 1.0);
                    OSup SA2;
                    sup_for_shape(&SA2, sh, sa, MODE_INCLUDE);
                    SA2.addRadius = 1;
                    SA2.addR = maxSep;
                    if (!pd_step_epa(&pd, &SA2, &T, s->penetration_tolerance, &simplex)) continue;
                    status = 1;
                }
                double pen = sqrt(vlen2(vsub(pd.pointA, pd.pointB))) - maxSep;
                double al = sqrt(vlen2(pd.axis));
                if (al > 0.0) {
                    double corr = maxSep / al;
                    pd.pointA = vadd(pd.pointA, vscale(pd.axis, -corr));
                }
                const int act = tr[3];
                if (s->collide_only_with_active_edges && act != 7)
                    pd.axis = fix_normal(a, b, c, back ? nrm : vneg(nrm), act, pd.pointB, pd.axis, mdl);
                OHit h;
                h.hit = 1;
                h.status = status;
                h.pointA = vadd(qrotate(qa, pd.pointA), pa);
                h.pointB = vadd(qrotate(qa, pd.pointB), pa);
                h.axis = qrotate(qa, pd.axis);
                h.penetration = pen;
                supporting_face(&h.faceA, sh, sa, vneg(pd.axis), pa, qa);
                h.faceB.n = 3;
                h.faceB.v[0] = a; h.faceB.v[1] = b; h.faceB.v[2] = c;
                face_transform(&h.faceB, pa, qa, v3(1, 1, 1));
                fn(ctx, &h, (0xffffffffu & ~mask) | (uint32_t)ti);
            }
        } else {
            const int l = o + 8, r = (int)nd[6];
            const double* ln = nodes + l;
            const double* rn = nodes + r;
            V3 lc = v3((ln[0] + ln[3]) * 0.5, (ln[1] + ln[4]) * 0.5, (ln[2] + ln[5]) * 0.5);
            V3 rc = v3((rn[0] + rn[3]) * 0.5, (rn[1] + rn[4]) * 0.5, (rn[2] + rn[5]) * 0.5);
            double ld = vlen2(vsub(lc, qc)), rd = vlen2(vsub(rc, qc));
            if (ns + 2 > scap) { scap *= 2; stack = (int*)realloc(stack, sizeof(int) * (size_t)scap); }
            if (ld <= rd) { stack[ns++] = r; stack[ns++] = l; }
            else { stack[ns++] = l; stack[ns++] = r; }
        }
    }
}

/* getClosestPointOnTriangle (src/collision/triangle.ts:62-198; degenerate triangles,
 * handleDegenerateTriangle :199-286): the point of triangle abc closest to the origin, its
 * squared distance, and the TriangleFeature bits (:9-24) of the feature it lies on. */
static int closest_on_triangle(V3 a, V3 b, V3 c, V3* p, double* dist2) {
    V3 ab = vsub(b, a), ac = vsub(c, a);
    V3 n = vcross(ab, ac);
    double nl = vlen2(n);
    if (nl < 1e-10) {
        int f = 4;
        *p = c;
        double best = vlen2(c);
        double al = vlen2(a);
        if (al < best) { f = 1; *p = a; best = al; }
        double bl = vlen2(b);
        if (bl < best) { f = 2; *p = b; best = bl; }
        const double eps2 = 1.1920929e-7 * 1.1920929e-7;
        double acl = vlen2(ac);
        if (acl > eps2) {
            double v = dmax(0.0, dmin(1.0, -vdot(a, ac) / acl));
            V3 q = v3(a.x + v * ac.x, a.y + v * ac.y, a.z + v * ac.z);
            double dq = vlen2(q);
            if (dq < best) { f = 5; *p = q; best = dq; }
        }
        V3 bc = vsub(c, b);
        double bcl = vlen2(bc);
        if (bcl > eps2) {
            double v = dmax(0.0, dmin(1.0, -vdot(b, bc) / bcl));
            V3 q = v3(b.x + v * bc.x, b.y + v * bc.y, b.z + v * bc.z);
            double dq = vlen2(q);
            if (dq < best) { f = 6; *p = q; best = dq; }
        }
        double abl = vlen2(ab);
        if (abl > eps2) {
            double v = dmax(0.0, dmin(1.0, -vdot(a, ab) / abl));
            V3 q = v3(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
            double dq = vlen2(q);
            if (dq < best) { f = 3; *p = q; best = dq; }
        }
        *dist2 = best;
        return f;
    }
    V3 ap = vneg(a);
    double d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { *p = a; *dist2 = vlen2(a); return 1; }
    V3 bp = vneg(b);
    double d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { *p = b; *dist2 = vlen2(b); return 2; }
    if (d1 * d4 <= d3 * d2 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        *p = v3(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
        *dist2 = vlen2(*p);
        return 3;
    }
    V3 cp = vneg(c);
    double d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { *p = c; *dist2 = vlen2(c); return 4; }
    if (d5 * d2 <= d1 * d6 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        *p = v3(a.x + w * ac.x, a.y + w * ac.y, a.z + w * ac.z);
        *dist2 = vlen2(*p);
        return 5;
    }
    double d43 = d4 - d3, d56 = d5 - d6;
    if (d3 * d6 <= d5 * d4 && d43 >= 0 && d56 >= 0) {
        double w = d43 / (d43 + d56);
        V3 bc = vsub(c, b);
        *p = v3(b.x + w * bc.x, b.y + w * bc.y, b.z + w * bc.z);
        *dist2 = vlen2(*p);
        return 6;
    }
    double cd = (a.x + b.x + c.x) * n.x + (a.y + b.y + c.y) * n.y + (a.z + b.z + c.z) * n.z;
    double sc = cd / (3 * nl);
    *p = v3(n.x * sc, n.y * sc, n.z * sc);
    *dist2 = vlen2(*p);
    return 7;
}

int32_t oc_closest_on_triangle(const double* a, const double* b, const double* c, double* point, double* dist2) {
    V3 p;
    int f = closest_on_triangle(vload(a), vload(b), vload(c), &p, dist2);
    vstore(point, p);
    return f;
}

/* FEATURE_TO_ACTIVE_EDGES src/collision/sphere-triangle.ts:218-227 */
static const int k_feature_edges[8] = {0, 5, 3, 1, 6, 4, 2, 0};

/* collideSphereVsTriangleMesh src/shapes/triangle-mesh.ts:1418-1720 (sphere A of radius r at
 * pa, mesh at pm/qm, unit scales): the hits against each triangle the BVH walk reaches, in the
 * walk's order, in world space. */
void ocx_collide_sphere_mesh(const OShapes* sh, int sa, V3 pa, int sm, V3 pm, Q4 qm, const cc_settings* s,
                             double maxSep, V3 moveDir, ocx_hit_fn fn, void* ctx) {
    const cc_shape_desc* A = &sh->shapes[sa];
    const cc_shape_desc* M = &sh->shapes[sm];
    const cc_mesh_data* md = sh->meshes;
    g_mesh_calls++;
    const double r = A->radius;
    const double ms2 = (r + maxSep) * (r + maxSep);
    Q4 iqm = {-qm.x, -qm.y, -qm.z, qm.w};
    V3 c = qrotate(iqm, vsub(pa, pm));
    const double er = r + maxSep;
    V3 bmn = v3(c.x - er, c.y - er, c.z - er), bmx = v3(c.x + er, c.y + er, c.z + er);
    if (M->mesh_num_nodes == 0) return;
    const double* nodes = md->nodes + 8 * (size_t)M->mesh_first_node;
    const int32_t* tris = md->triangles + 4 * (size_t)M->mesh_first_triangle;
    const int ntri = M->mesh_num_triangles;
    const int nbits = ntri > 1 ? 32 - __builtin_clz((unsigned)(ntri - 1)) : 0;
    const uint32_t mask = nbits >= 32 ? 0xffffffffu : ((1u << nbits) - 1u);
    V3 mdm = qrotate(iqm, moveDir);
    static __thread int* stack = NULL;
    static __thread int scap = 0;
    int ns = 0;
    if (!scap) { scap = 256; stack = (int*)malloc(sizeof(int) * (size_t)scap); }
    stack[ns++] = 0;
    while (ns > 0) {
        const int o = stack[--ns];
        const double* nd = nodes + o;
        if (!(nd[0] <= bmx.x && nd[3] >= bmn.x && nd[1] <= bmx.y && nd[4] >= bmn.y && nd[2] <= bmx.z && nd[5] >= bmn.z))
            continue;
        if (nd[7] < 0) {
            const int t0 = (int)nd[6], cnt = -((int)nd[7] + 1);
            for (int i = 0; i < cnt; i++) {
                const int ti = t0 + i;
                const int32_t* tr = tris + 4 * (size_t)ti;
                V3 v0 = vload(md->positions + 3 * (size_t)tr[0]);
                V3 v1 = vload(md->positions + 3 * (size_t)tr[1]);
                V3 v2 = vload(md->positions + 3 * (size_t)tr[2]);
                V3 a = vsub(v0, c), b = vsub(v1, c), cc = vsub(v2, c);
                V3 nrm = vcross(vsub(b, a), vsub(cc, a));
                int back = vdot(nrm, a) > 0.0;
                if (!s->collide_with_backfaces && back) continue;
                V3 p2;
                double d2;
                int f = closest_on_triangle(a, b, cc, &p2, &d2);
                if (d2 > ms2) continue;
                double pen = r - sqrt(d2);
                V3 axis = d2 > 0 ? vnormalize(p2) : v3(0, 1, 0);
                V3 p1 = v3(axis.x * r, axis.y * r, axis.z * r);
                const int act = tr[3];
                if (s->collide_only_with_active_edges && f != 7 && act != 7 && (act & k_feature_edges[f]) == 0) {
                    V3 na = back ? nrm : vneg(nrm);
                    if (vdot(mdm, axis) * vlen(na) >= vdot(mdm, na)) axis = na;
                }
                OHit h;
                memset(&h, 0, sizeof(h));
                h.hit = 1;
                h.status = 0;
                h.pointA = vadd(qrotate(qm, vadd(c, p1)), pm);
                h.pointB = vadd(qrotate(qm, vadd(c, p2)), pm);
                h.axis = qrotate(qm, axis);
                h.penetration = pen;
                h.faceA.n = 0; /* a sphere has no supporting face; the triangle's is not used then */
                h.faceB.n = 0;
                fn(ctx, &h, (0xffffffffu & ~mask) | (uint32_t)ti);
            }
        } else {
            const int l = o + 8, rr = (int)nd[6];
            const double* ln = nodes + l;
            const double* rn = nodes + rr;
            V3 lc = v3((ln[0] + ln[3]) * 0.5, (ln[1] + ln[4]) * 0.5, (ln[2] + ln[5]) * 0.5);
            V3 rc = v3((rn[0] + rn[3]) * 0.5, (rn[1] + rn[4]) * 0.5, (rn[2] + rn[5]) * 0.5);
            double ld = vlen2(vsub(lc, c)), rd = vlen2(vsub(rc, c));
            if (ns + 2 > scap) { scap *= 2; stack = (int*)realloc(stack, sizeof(int) * (size_t)scap); }
            if (ld <= rd) { stack[ns++] = rr; stack[ns++] = l; }
            else { stack[ns++] = l; stack[ns++] = rr; }
        }
    }
}

/* ---- planes (src/shapes/plane.ts), with mathcat's column-major mat4 ----------------------- */
typedef struct { double m[16]; } OM4;
static OM4 m4_from_rt(Q4 q, V3 v) { /* mat4.fromRotationTranslation */
    double x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    double xx = q.x * x2, xy = q.x * y2, xz = q.x * z2, yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    double wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    OM4 o = {{1 - (yy + zz), xy + wz, xz - wy, 0, xy - wz, 1 - (xx + zz), yz + wx, 0, xz + wy, yz - wx, 1 - (xx + yy), 0,
              v.x, v.y, v.z, 1}};
    return o;
}
static OM4 m4_invert(const OM4* A) { /* mat4.invert */
    const double* a = A->m;
    double a00 = a[0], a01 = a[1], a02 = a[2], a03 = a[3], a10 = a[4], a11 = a[5], a12 = a[6], a13 = a[7];
    double a20 = a[8], a21 = a[9], a22 = a[10], a23 = a[11], a30 = a[12], a31 = a[13], a32 = a[14], a33 = a[15];
    double b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
    double b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    double b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
    double b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
    det = 1.0 / det;
    OM4 o = {{(a11 * b11 - a12 * b10 + a13 * b09) * det, (a02 * b10 - a01 * b11 - a03 * b09) * det,
              (a31 * b05 - a32 * b04 + a33 * b03) * det, (a22 * b04 - a21 * b05 - a23 * b03) * det,
              (a12 * b08 - a10 * b11 - a13 * b07) * det, (a00 * b11 - a02 * b08 + a03 * b07) * det,
              (a32 * b02 - a30 * b05 - a33 * b01) * det, (a20 * b05 - a22 * b02 + a23 * b01) * det,
              (a10 * b10 - a11 * b08 + a13 * b06) * det, (a01 * b08 - a00 * b10 - a03 * b06) * det,
              (a30 * b04 - a31 * b02 + a33 * b00) * det, (a21 * b02 - a20 * b04 - a23 * b00) * det,
              (a11 * b07 - a10 * b09 - a12 * b06) * det, (a00 * b09 - a01 * b07 + a02 * b06) * det,
              (a31 * b01 - a30 * b03 - a32 * b00) * det, (a20 * b03 - a21 * b01 + a22 * b00) * det}};
    return o;
}
static OM4 m4_mul(const OM4* A, const OM4* B) { /* mat4.multiply: A * B */
    OM4 o;
    for (int j = 0; j < 4; j++) {
        double b0 = B->m[4 * j], b1 = B->m[4 * j + 1], b2 = B->m[4 * j + 2], b3 = B->m[4 * j + 3];
        for (int i = 0; i < 4; i++) o.m[4 * j + i] = b0 * A->m[i] + b1 * A->m[4 + i] + b2 * A->m[8 + i] + b3 * A->m[12 + i];
    }
    return o;
}
static V3 m4_point(const OM4* M, V3 p) { /* vec3.transformMat4 */
    const double* m = M->m;
    double w = m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15];
    if (w == 0) w = 1;
    return v3((m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12]) / w, (m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13]) / w,
              (m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]) / w);
}
/* scalePlane at unit scale (:184-194), then transformPlane (:196-208) */
static void plane_in(const cc_shape_desc* P, const OM4* M, V3* n, double* c) {
    V3 n0 = vload(P->half_extents);
    double l = vlen(n0);
    V3 sn = vscale(n0, 1 / l);
    double sc = P->radius / l;
    const double* m = M->m;
    *n = v3(m[0] * sn.x + m[4] * sn.y + m[8] * sn.z, m[1] * sn.x + m[5] * sn.y + m[9] * sn.z,
            m[2] * sn.x + m[6] * sn.y + m[10] * sn.z);
    *c = sc - (m[12] * n->x + m[13] * n->y + m[14] * n->z);
}
/* collideConvexVsPlane src/shapes/plane.ts:474-590, with the adaptive plane face :631-725 */
static void collide_convex_plane(const OShapes* sh, int sc, V3 pc, Q4 qc, int sp, V3 pp, Q4 qp, double maxSep,
                                 OHit* out) {
    OM4 TA = m4_from_rt(qc, pc), TB = m4_from_rt(qp, pp);
    OM4 iA = m4_invert(&TA);
    OM4 M = m4_mul(&iA, &TB);
    V3 n;
    double c;
    plane_in(&sh->shapes[sp], &M, &n, &c);
    int ty = sh->shapes[sc].type;
    OSup S;
    sup_for_shape(&S, sh, sc, (ty == CC_SHAPE_BOX || ty == CC_SHAPE_CYLINDER) ? MODE_INCLUDE : MODE_EXCLUDE);
    V3 s = sup_base(&S, vneg(n));
    double dist = vdot(n, s) + c;
    double cr = S.convexRadius;
    double pen = -dist + cr;
    out->hit = 0;
    out->faceA.n = out->faceB.n = 0;
    if (!(pen > -maxSep)) return;
    V3 p1 = vsub(s, v3(n.x * cr, n.y * cr, n.z * cr));
    V3 p2 = vsub(s, v3(n.x * dist, n.y * dist, n.z * dist));
    out->hit = 1;
    out->status = 0;
    out->pointA = m4_point(&TA, p1);
    out->pointB = m4_point(&TA, p2);
    out->axis = qrotate(qc, vneg(n));
    out->penetration = pen;
    supporting_face(&out->faceA, sh, sc, n, pc, qc);
    if (out->faceA.n > 0) {
        V3 wn;
        double wc;
        plane_in(&sh->shapes[sp], &TB, &wn, &wc);
        double dd = vdot(wn, pc) + wc; /* plane3.projectPoint */
        V3 ctr = vsub(pc, v3(wn.x * dd, wn.y * dd, wn.z * dd));
        V3 b1 = vcross(wn, v3(0, 1, 0)); /* getOrthogonalBasis :210-233 */
        double l2 = vlen2(b1);
        if (l2 < 1e-6) b1 = v3(1, 0, 0);
        else { double k = 1 / sqrt(l2); b1 = v3(b1.x * k, b1.y * k, b1.z * k); }
        V3 b2 = vnormalize(vcross(b1, wn));
        b1 = vnormalize(vcross(wn, b2));
        const cc_shape_desc* C = &sh->shapes[sc];
        double size = vlen(v3(C->aabb_max[0] - C->aabb_min[0], C->aabb_max[1] - C->aabb_min[1], C->aabb_max[2] - C->aabb_min[2]));
        b1 = v3(b1.x * size, b1.y * size, b1.z * size);
        b2 = v3(b2.x * size, b2.y * size, b2.z * size);
        out->faceB.n = 4;
        out->faceB.v[0] = vadd(ctr, vadd(b1, b2));
        out->faceB.v[1] = vadd(ctr, vsub(b1, b2));
        out->faceB.v[2] = vsub(ctr, vadd(b1, b2));
        out->faceB.v[3] = vadd(ctr, vadd(vneg(b1), b2));
    }
}

/* ======================================================================== */
/* shape cast for CCD                                                        */
/* ======================================================================== */
/* gjkCastShape gjk.ts:853-1267: shape A (transformed by posAInB / quatAInB) moves by `disp` against
 * B at rest, both without their convex radius; returns 0 for no hit, else the fraction, the contact
 * point on A and the separating axis, all in B's space. */
static int gjk_cast_shape(const OSup* A, const OSup* B, V3 disp, double tolerance, double crA, double crB,
                          double maxLambda, double* lambdaOut, V3* pointA, V3* sepAxis) {
    double tol2 = tolerance * tolerance;
    const double sumR = crA + crB;
    Simplex s;
    s.size = 0;
    double lambda = 0.0;
    V3 x = v3(0, 0, 0);
    V3 q = vneg(sup_get(B, v3(0, 0, 0)));
    V3 p = sup_get(A, v3(0, 0, 0));
    V3 v = vsub(q, p);
    double vLen2 = DBL_MAX;
    int allowRestart = 0;
    V3 prevV = v3(0, 0, 0);
    int it = 0;
    while (it < GJK_MAX_ITERATIONS) {
        it++;
        p = sup_get(A, vneg(v));
        q = sup_get(B, v);
        V3 w = vsub(x, vsub(q, p));
        const double vw = v.x * w.x + v.y * w.y + v.z * w.z - sumR * sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        if (vw > 0.0) {
            const double vr = v.x * disp.x + v.y * disp.y + v.z * disp.z;
            if (vr >= -1.0e-18) return 0;
            const double old = lambda;
            lambda -= vw / vr;
            if (old == lambda) break;
            if (lambda >= maxLambda) return 0;
            x = vscale(disp, lambda);
            vLen2 = DBL_MAX;
            tol2 = tolerance + sumR;
            tol2 = tol2 * tol2;
            allowRestart = 1;
        }
        s.pt[s.size].p = p;
        s.pt[s.size].q = q;
        s.size++;
        for (int i = 0; i < s.size; i++) s.pt[i].y = vadd(vsub(x, s.pt[i].q), s.pt[i].p);
        CPSRes c;
        cp_simplex(&c, vLen2, 0, &s);
        if (c.found) {
            vLen2 = c.dist2;
            v = c.point;
        } else {
            if (!allowRestart) break;
            allowRestart = 0;
            s.pt[0].p = p;
            s.pt[0].q = q;
            s.size = 1;
            v = vsub(x, q);
            vLen2 = DBL_MAX;
            continue;
        }
        if (c.set == 0xf) break;
        int ns = 0; /* updatePointSetPQ */
        for (int i = 0; i < s.size; i++)
            if (c.set & (1 << i)) {
                if (ns != i) s.pt[ns] = s.pt[i];
                ns++;
            }
        s.size = ns;
        if (vLen2 <= tol2) break;
        prevV = v;
    }
    for (int i = 0; i < s.size; i++) s.pt[i].y = vadd(vsub(x, s.pt[i].q), s.pt[i].p);
    const double vl = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    V3 nv = v3(0, 0, 0);
    if (vl > 0) nv = vscale(v, 1 / vl);
    V3 pa = v3(0, 0, 0), pb = v3(0, 0, 0);
    if (s.size == 1) {
        pb = vadd(s.pt[0].q, vscale(nv, crB));
        pa = lambda > 0.0 ? pb : vadd(s.pt[0].p, vscale(nv, -crA));
    } else if (s.size == 2) {
        Bary b = bary2(s.pt[0].y, s.pt[1].y, 1e-10);
        pb = vadd(vadd(vadd(pb, vscale(s.pt[0].q, b.u)), vscale(s.pt[1].q, b.v)), vscale(nv, crB));
        pa = lambda > 0.0 ? pb
                          : vadd(vadd(vadd(pa, vscale(s.pt[0].p, b.u)), vscale(s.pt[1].p, b.v)), vscale(nv, -crA));
    } else if (s.size >= 3) {
        Bary b = bary3(s.pt[0].y, s.pt[1].y, s.pt[2].y, 1e-10);
        pb = vadd(vadd(vadd(vadd(pb, vscale(s.pt[0].q, b.u)), vscale(s.pt[1].q, b.v)), vscale(s.pt[2].q, b.w)),
                  vscale(nv, crB));
        if (lambda > 0.0) pa = pb;
        else
            pa = vadd(vadd(vadd(vadd(pa, vscale(s.pt[0].p, b.u)), vscale(s.pt[1].p, b.v)), vscale(s.pt[2].p, b.w)),
                      vscale(nv, -crA));
    }
    *lambdaOut = lambda;
    *pointA = pa;
    *sepAxis = sumR > 0.0 ? vneg(v) : vneg(prevV);
    return 1;
}

/* castConvexVsConvex convex.ts:675-875 with penetrationCastShape penetration.ts:621-688 as the CCD
 * visitor calls them (update.ts:1218-1237: unit scales, returnDeepestPoint false, so no EPA, and
 * the supporting faces go unused).  A at (pa, qa) moves by `disp` (world space) against B at
 * (pb, qb).  Returns 1 for a hit accepted below `earlyOut`, with the fraction, A's contact point
 * and the (unnormalised) penetration axis in world space. */
int ocx_cast_convex(const OShapes* sh, int sa, V3 pa, Q4 qa, V3 disp, int sb, V3 pb, Q4 qb, int backfaces,
                    double earlyOut, double* fraction, V3* pointA, V3* penAxis) {
    const Q4 ib = {-qb.x, -qb.y, -qb.z, qb.w};
    const V3 posAInB = qrotate(ib, vsub(pa, pb));
    const Q4 quatAInB = qmul(ib, qa);
    const V3 dispInB = qrotate(ib, disp);
    OSup A, B;
    sup_for_shape(&A, sh, sa, MODE_EXCLUDE);
    sup_for_shape(&B, sh, sb, MODE_EXCLUDE);
    A.transformed = 1;
    A.R = m3_from_quat(quatAInB);
    A.t = posAInB;
    const double tolerance = 1e-4;
    double lambda;
    V3 pA, sep;
    if (!gjk_cast_shape(&A, &B, dispInB, tolerance, A.convexRadius, B.convexRadius, earlyOut, &lambda, &pA, &sep))
        return 0;
    if (vlen2(sep) < tolerance * tolerance) sep = dispInB; /* contactNormalInvalid */
    if (lambda > 1.0) return 0;
    if (!backfaces && vdot(sep, dispInB) <= 0) return 0;
    if (lambda >= earlyOut) return 0;
    *fraction = lambda;
    *pointA = vadd(qrotate(qb, pA), pb);
    *penAxis = qrotate(qb, sep);
    return 1;
}

/* castConvexVsPlane plane.ts:727-880 as the CCD visitor calls it (unit scales): the convex shape's
 * DEFAULT support point against the plane is taken to world space with A's transform and
 * measured against the plane in the plane shape's own frame; the contact point goes through the
 * plane body's transform (translated by the cast for a hit ahead) and the penetration axis is
 * -normal through that same matrix, as a point.  Returns 1 for a hit below `earlyOut`. */
int ocx_cast_convex_plane(const OShapes* sh, int sa, V3 pa, Q4 qa, V3 disp, int sp, V3 pb, Q4 qb, int backfaces,
                          double earlyOut, double* fraction, V3* pointA, V3* penAxis) {
    const cc_shape_desc* P = &sh->shapes[sp];
    const V3 sn0 = v3(P->half_extents[0] / 1.0, P->half_extents[1] / 1.0, P->half_extents[2] / 1.0);
    const double len = vlen(sn0);
    const V3 n = vscale(sn0, 1 / len);
    const double c = P->radius / len;
    const OM4 start = m4_from_rt(qa, pa);
    const int ty = sh->shapes[sa].type;
    OSup S;
    sup_for_shape(&S, sh, sa, (ty == CC_SHAPE_BOX || ty == CC_SHAPE_CYLINDER) ? MODE_INCLUDE : MODE_EXCLUDE);
    const Q4 iq = {-qa.x, -qa.y, -qa.z, qa.w};
    const V3 dirS = vneg(qrotate(iq, n));
    const V3 sw = m4_point(&start, sup_get(&S, dirS));
    const double sd = vdot(n, sw) + c;
    const double cr = S.convexRadius;
    const double pen = -sd + cr;
    const double dot = vdot(disp, n);
    OM4 hitM;
    double f;
    if (pen > 0.0) {
        if (!backfaces && dot > 0.0) return 0;
        if (pen <= -earlyOut) return 0;
        f = 0.0;
        hitM = m4_from_rt(qb, pb);
    } else if (dot < 0.0) {
        f = pen / dot;
        if (f >= earlyOut) return 0;
        const OM4 p2w = m4_from_rt(qb, pb);
        const V3 o = vscale(disp, f);
        hitM = p2w; /* mat4.translate */
        hitM.m[12] = p2w.m[0] * o.x + p2w.m[4] * o.y + p2w.m[8] * o.z + p2w.m[12];
        hitM.m[13] = p2w.m[1] * o.x + p2w.m[5] * o.y + p2w.m[9] * o.z + p2w.m[13];
        hitM.m[14] = p2w.m[2] * o.x + p2w.m[6] * o.y + p2w.m[10] * o.z + p2w.m[14];
        hitM.m[15] = p2w.m[3] * o.x + p2w.m[7] * o.y + p2w.m[11] * o.z + p2w.m[15];
    } else {
        return 0;
    }
    *fraction = f;
    *pointA = m4_point(&hitM, vsub(sw, vscale(n, cr)));
    *penAxis = m4_point(&hitM, vneg(n));
    return 1;
}

/* collideShapeVsShape dispatch (src/collision/narrowphase.ts:172-238, src/shapes/sphere.ts:154-167) */
void ocx_collide(const OShapes* sh, int sa, V3 pa, Q4 qa, int sb, V3 pb, Q4 qb, double maxSep, double colTol,
                 double penTol, OHit* out) {
    out->faceA.n = out->faceB.n = 0;
    int ta = sh->shapes[sa].type, tb = sh->shapes[sb].type;
    leaf_pose(&sh->shapes[sa], &pa, &qa); /* collideTransformedVsShape, then collideShapeVsTransformed */
    leaf_pose(&sh->shapes[sb], &pb, &qb);
    if (ta == CC_SHAPE_PLANE || tb == CC_SHAPE_PLANE) { /* plane.ts:297-310 */
        out->hit = 0;
        if (ta == tb) return;
        if (tb == CC_SHAPE_PLANE) { collide_convex_plane(sh, sa, pa, qa, sb, pb, qb, maxSep, out); return; }
        OHit h;
        collide_convex_plane(sh, sb, pb, qb, sa, pa, qa, maxSep, &h); /* reversedCollideShapeVsShape */
        *out = h;
        out->pointA = h.pointB; out->pointB = h.pointA;
        out->axis = vneg(h.axis);
        out->faceA = h.faceB; out->faceB = h.faceA;
        return;
    }
    if (ta == CC_SHAPE_SPHERE && tb == CC_SHAPE_SPHERE)
        collide_sphere_sphere(sh, sa, pa, sb, pb, out);
    else
        collide_convex_convex(sh, sa, pa, qa, sb, pb, qb, maxSep, colTol, penTol, out);
}

/* ======================================================================== */
/* clipping (src/manifold/clip.ts)                                           */
/* ======================================================================== */
static void clip_poly_vs_plane(OFace* out, const OFace* poly, V3 origin, V3 normal) { /* :13-92 */
    out->n = 0;
    if (poly->n < 2) return;
    V3 e1 = poly->v[poly->n - 1];
    double prevNum = vdot(vsub(origin, e1), normal);
    int prevIn = prevNum < 0.0;
    for (int j = 0; j < poly->n; j++) {
        V3 e2 = poly->v[j];
        double num = vdot(vsub(origin, e2), normal);
        int curIn = num < 0.0;
        if (curIn != prevIn) {
            V3 e12 = vsub(e2, e1);
            double den = vdot(e12, normal);
            if (den != 0.0) {
                double t = prevNum / den;
                if (out->n < OC_MAX_FACE) out->v[out->n++] = vadd(e1, vscale(e12, t));
            } else curIn = prevIn;
        }
        if (curIn && out->n < OC_MAX_FACE) out->v[out->n++] = e2;
        prevNum = num;
        prevIn = curIn;
        e1 = e2;
    }
}

static void clip_poly_vs_poly(OFace* out, const OFace* toClip, const OFace* clip, V3 normal) { /* :108-175 */
    static __thread OFace tmp[2];
    out->n = 0;
    if (toClip->n < 2 || clip->n < 3) return;
    int tmpIdx = 0;
    const OFace* src;
    OFace* dst;
    for (int i = 0; i < clip->n; i++) {
        V3 c1 = clip->v[i], c2 = clip->v[(i + 1) % clip->n];
        V3 cn = vcross(normal, vsub(c2, c1));
        src = (i == 0) ? toClip : &tmp[tmpIdx];
        tmpIdx ^= 1;
        dst = (i == clip->n - 1) ? out : &tmp[tmpIdx];
        clip_poly_vs_plane(dst, src, c1, cn);
        if (dst->n < 3) { out->n = 0; break; }
    }
}

static void clip_poly_vs_edge(OFace* out, const OFace* poly, V3 ev1, V3 ev2, V3 clipNormal) { /* :190-363 */
    out->n = 0;
    if (poly->n < 3) return;
    V3 edge = vsub(ev2, ev1);
    V3 en = vcross(clipNormal, edge);
    V3 p0 = poly->v[0], p1 = poly->v[1], p2 = poly->v[2];
    V3 pn = vcross(vsub(p2, p0), vsub(p1, p0));
    double pl2 = vlen2(pn);
    V3 v1 = vadd(ev1, vscale(pn, vdot(vsub(p0, ev1), pn) / pl2));
    V3 v2 = vadd(ev2, vscale(pn, vdot(vsub(p0, ev2), pn) / pl2));
    V3 v12 = vsub(v2, v1);
    double v12l = vlen2(v12);
    V3 e1 = poly->v[poly->n - 1];
    double prevNum = vdot(vsub(ev1, e1), en);
    int prevIn = prevNum < 0.0;
    for (int j = 0; j < poly->n; j++) {
        V3 e2 = poly->v[j];
        double num = vdot(vsub(ev1, e2), en);
        int curIn = num < 0.0;
        if (curIn != prevIn) {
            V3 e12 = vsub(e2, e1);
            double den = vdot(e12, en);
            V3 cl = (den != 0.0) ? vadd(e1, vscale(e12, prevNum / den)) : e1;
            double pr = vdot(vsub(cl, v1), v12);
            V3 o = pr < 0.0 ? v1 : (pr > v12l ? v2 : cl);
            if (out->n < OC_MAX_FACE) out->v[out->n++] = o;
        }
        prevNum = num;
        prevIn = curIn;
        e1 = e2;
    }
}

/* ======================================================================== */
/* manifold (src/manifold/manifold.ts)                                       */
/* ======================================================================== */
static void set_point(OManifold* m, int i, V3 w1, V3 w2) { /* setContactPoint :152-163 */
    m->relA[i] = vsub(w1, m->base);
    m->relB[i] = vsub(w2, m->base);
}

/* manifoldBetweenTwoFaces :345-446 */
void ocx_manifold_between_two_faces(OManifold* out, V3 p1, V3 p2, V3 axis, double maxDist, const OFace* f1,
                                    const OFace* f2) {
    out->n = 0;
    if (f1->n < 2 || f2->n < 3) {
        set_point(out, 0, p1, p2);
        out->n = 1;
        return;
    }
    OFace clipped;
    clipped.n = 0;
    if (f1->n >= 3) clip_poly_vs_poly(&clipped, f2, f1, axis);
    else if (f1->n == 2) clip_poly_vs_edge(&clipped, f2, f1->v[0], f1->v[1], axis);
    V3 origin = f1->v[0];
    V3 first = vsub(f1->v[1], origin);
    V3 pnrm;
    if (f1->n >= 3) pnrm = vcross(first, vsub(f1->v[2], origin));
    else pnrm = vcross(vcross(first, axis), first);
    double apn = vdot(axis, pnrm);
    if (apn != 0.0) {
        double al = vlen(axis);
        for (int i = 0; i < clipped.n; i++) {
            V3 q2 = clipped.v[i];
            double dist = vdot(vsub(q2, origin), pnrm) / apn;
            if (dist * al < maxDist) {
                V3 q1 = vadd(q2, vscale(axis, -dist));
                set_point(out, out->n, q1, q2);
                out->n++;
                if (out->n >= OC_MAX_MANIFOLD_POINTS) break;
            }
        }
    }
    if (out->n == 0) {
        set_point(out, 0, p1, p2);
        out->n = 1;
    }
}

/* pruneContactPoints :192-313 */
void ocx_prune(OManifold* m, V3 axis) {
    if (m->n <= 4) return;
    const double MIN_D2 = 1.0e-6;
    V3 proj[OC_MAX_MANIFOLD_POINTS];
    double pd2[OC_MAX_MANIFOLD_POINTS];
    for (int i = 0; i < m->n; i++) {
        V3 v1 = m->relA[i], v2 = m->relB[i];
        double dp = vdot(v1, axis);
        proj[i] = vadd(v1, vscale(axis, -dp));
        pd2[i] = dmax(MIN_D2, vlen2(vsub(v2, v1)));
    }
    int p1 = 0;
    double best = dmax(MIN_D2, vlen2(proj[0])) * pd2[0];
    for (int i = 1; i < m->n; i++) {
        double me = dmax(MIN_D2, vlen2(proj[i])) * pd2[i];
        if (me > best) { best = me; p1 = i; }
    }
    int p2 = -1;
    best = -INFINITY;
    for (int i = 0; i < m->n; i++) {
        if (i == p1) continue;
        double me = dmax(MIN_D2, vlen2(vsub(proj[i], proj[p1]))) * pd2[i];
        if (me > best) { best = me; p2 = i; }
    }
    int p3 = -1, p4 = -1;
    double minD = 0.0, maxD = 0.0;
    V3 perp = vcross(vsub(proj[p2], proj[p1]), axis);
    for (int i = 0; i < m->n; i++) {
        if (i == p1 || i == p2) continue;
        double d = vdot(vsub(proj[i], proj[p1]), perp);
        if (d < minD) { minD = d; p3 = i; }
        else if (d > maxD) { maxD = d; p4 = i; }
    }
    int order[4] = {p1, p3, p2, p4};
    V3 ra[4], rb[4];
    int n = 0;
    for (int k = 0; k < 4; k++) {
        if (order[k] != -1) { ra[n] = m->relA[order[k]]; rb[n] = m->relB[order[k]]; n++; }
    }
    for (int k = 0; k < n; k++) { m->relA[k] = ra[k]; m->relB[k] = rb[k]; }
    m->n = n;
}

/* ======================================================================== */
/* unit entry points                                                         */
/* ======================================================================== */
static OFace face_from(const double* v, int n) {
    OFace f;
    f.n = n;
    for (int i = 0; i < n; i++) f.v[i] = vload(v + 3 * i);
    return f;
}

void oc_collide(const cc_shape_desc* shapes, const cc_hull_data* hulls, int32_t sa, const double* pa, const double* qa,
                int32_t sb, const double* pb, const double* qb, double maxSep, const cc_settings* s, oc_hit* out) {
    OShapes sh = {shapes, hulls, NULL};
    OHit h;
    Q4 QA = {qa[0], qa[1], qa[2], qa[3]}, QB = {qb[0], qb[1], qb[2], qb[3]};
    ocx_collide(&sh, sa, vload(pa), QA, sb, vload(pb), QB, maxSep, s->collision_tolerance, s->penetration_tolerance, &h);
    memset(out, 0, sizeof(*out));
    out->hit = h.hit;
    out->status = h.hit ? h.status : -1;
    if (!h.hit) return;
    vstore(out->point_a, h.pointA);
    vstore(out->point_b, h.pointB);
    vstore(out->axis, h.axis);
    out->penetration = h.penetration;
    out->face_a_n = h.faceA.n;
    out->face_b_n = h.faceB.n;
    for (int i = 0; i < h.faceA.n; i++) vstore(out->face_a + 3 * i, h.faceA.v[i]);
    for (int i = 0; i < h.faceB.n; i++) vstore(out->face_b + 3 * i, h.faceB.v[i]);
}

/* the convex GJK/EPA path for any pair, spheres included (a sphere wrapped in a transformed
 * shape takes this path in src/collision/narrowphase.ts:172-238) */
void oc_collide_convex(const cc_shape_desc* shapes, const cc_hull_data* hulls, int32_t sa, const double* pa,
                       const double* qa, int32_t sb, const double* pb, const double* qb, double maxSep,
                       const cc_settings* s, oc_hit* out) {
    OShapes sh = {shapes, hulls, NULL};
    OHit h;
    Q4 QA = {qa[0], qa[1], qa[2], qa[3]}, QB = {qb[0], qb[1], qb[2], qb[3]};
    h.hit = 0;
    h.faceA.n = h.faceB.n = 0;
    collide_convex_convex(&sh, sa, vload(pa), QA, sb, vload(pb), QB, maxSep, s->collision_tolerance,
                          s->penetration_tolerance, &h);
    memset(out, 0, sizeof(*out));
    out->hit = h.hit;
    out->status = h.hit ? h.status : -1;
    if (!h.hit) return;
    vstore(out->point_a, h.pointA);
    vstore(out->point_b, h.pointB);
    vstore(out->axis, h.axis);
    out->penetration = h.penetration;
    out->face_a_n = h.faceA.n;
    out->face_b_n = h.faceB.n;
    for (int i = 0; i < h.faceA.n; i++) vstore(out->face_a + 3 * i, h.faceA.v[i]);
    for (int i = 0; i < h.faceB.n; i++) vstore(out->face_b + 3 * i, h.faceB.v[i]);
}

int32_t oc_manifold_from_hit(const oc_hit* hit, double maxDist, double* normal_out, double* out_a, double* out_b) {
    OManifold m;
    m.base = v3(0, 0, 0);
    m.n = 0;
    V3 axis = vload(hit->axis);
    V3 n = vnormalize(axis);
    OFace fa = face_from(hit->face_a, hit->face_a_n), fb = face_from(hit->face_b, hit->face_b_n);
    if (fa.n >= 2 && fb.n >= 3)
        ocx_manifold_between_two_faces(&m, vload(hit->point_a), vload(hit->point_b), axis, maxDist, &fa, &fb);
    else { set_point(&m, 0, vload(hit->point_a), vload(hit->point_b)); m.n = 1; }
    n = vnormalize(n);
    if (m.n > 4) ocx_prune(&m, n);
    vstore(normal_out, n);
    for (int i = 0; i < m.n; i++) { vstore(out_a + 3 * i, m.relA[i]); vstore(out_b + 3 * i, m.relB[i]); }
    return m.n;
}

int32_t oc_clip_poly_vs_poly(const double* poly, int32_t n, const double* clip, int32_t m, const double* normal,
                             double* out) {
    OFace a = face_from(poly, n), b = face_from(clip, m), o;
    clip_poly_vs_poly(&o, &a, &b, vload(normal));
    for (int i = 0; i < o.n; i++) vstore(out + 3 * i, o.v[i]);
    return o.n;
}

int32_t oc_clip_poly_vs_plane(const double* poly, int32_t n, const double* origin, const double* normal, double* out) {
    OFace a = face_from(poly, n), o;
    clip_poly_vs_plane(&o, &a, vload(origin), vload(normal));
    for (int i = 0; i < o.n; i++) vstore(out + 3 * i, o.v[i]);
    return o.n;
}

int32_t oc_clip_poly_vs_edge(const double* poly, int32_t n, const double* e1, const double* e2, const double* normal,
                             double* out) {
    OFace a = face_from(poly, n), o;
    clip_poly_vs_edge(&o, &a, vload(e1), vload(e2), vload(normal));
    for (int i = 0; i < o.n; i++) vstore(out + 3 * i, o.v[i]);
    return o.n;
}

int32_t oc_manifold_between_two_faces(const double* p1, const double* p2, const double* axis, double maxDist,
                                      const double* f1, int32_t n1, const double* f2, int32_t n2, double* out_a,
                                      double* out_b) {
    OManifold m;
    m.base = v3(0, 0, 0);
    OFace a = face_from(f1, n1), b = face_from(f2, n2);
    ocx_manifold_between_two_faces(&m, vload(p1), vload(p2), vload(axis), maxDist, &a, &b);
    for (int i = 0; i < m.n; i++) { vstore(out_a + 3 * i, m.relA[i]); vstore(out_b + 3 * i, m.relB[i]); }
    return m.n;
}

int32_t oc_prune(double* rel_a, double* rel_b, int32_t n, const double* axis) {
    OManifold m;
    m.base = v3(0, 0, 0);
    m.n = n;
    for (int i = 0; i < n; i++) { m.relA[i] = vload(rel_a + 3 * i); m.relB[i] = vload(rel_b + 3 * i); }
    ocx_prune(&m, vload(axis));
    for (int i = 0; i < m.n; i++) { vstore(rel_a + 3 * i, m.relA[i]); vstore(rel_b + 3 * i, m.relB[i]); }
    return m.n;
}

void oc_barycentric2(const double* a, const double* b, double tol, double* r) {
    Bary o = bary2(vload(a), vload(b), tol);
    r[0] = o.u; r[1] = o.v; r[2] = o.w; r[3] = o.valid;
}

void oc_barycentric3(const double* a, const double* b, const double* c, double tol, double* r) {
    Bary o = bary3(vload(a), vload(b), vload(c), tol);
    r[0] = o.u; r[1] = o.v; r[2] = o.w; r[3] = o.valid;
}

int32_t oc_supporting_face(const cc_shape_desc* shapes, const cc_hull_data* hulls, int32_t s, const double* dir,
                           const double* pos, const double* quat, double* out) {
    OShapes sh = {shapes, hulls, NULL};
    OFace f;
    Q4 q = {quat[0], quat[1], quat[2], quat[3]};
    supporting_face(&f, &sh, s, vload(dir), vload(pos), q);
    for (int i = 0; i < f.n; i++) vstore(out + 3 * i, f.v[i]);
    return f.n;
}

typedef struct { oc_hit* out; uint32_t* sub; int n, cap; } MeshHits;
static void mesh_hits_cb(void* ctx, const OHit* h, uint32_t sub) {
    MeshHits* m = (MeshHits*)ctx;
    if (m->n < m->cap) {
        oc_hit* o = &m->out[m->n];
        memset(o, 0, sizeof(*o));
        o->hit = 1;
        o->status = h->status;
        vstore(o->point_a, h->pointA); vstore(o->point_b, h->pointB); vstore(o->axis, h->axis);
        o->penetration = h->penetration;
        o->face_a_n = h->faceA.n; o->face_b_n = h->faceB.n;
        for (int i = 0; i < h->faceA.n; i++) vstore(o->face_a + 3 * i, h->faceA.v[i]);
        for (int i = 0; i < h->faceB.n; i++) vstore(o->face_b + 3 * i, h->faceB.v[i]);
        m->sub[m->n] = sub;
    }
    m->n++;
}
int32_t oc_collide_mesh(const cc_shape_desc* shapes, const cc_hull_data* hulls, const cc_mesh_data* meshes, int32_t sa,
                        const double* pa, const double* qa, int32_t sm, const double* pm, const double* qm,
                        double maxSep, const double* move_dir, const cc_settings* s, oc_hit* out, uint32_t* sub_b,
                        int32_t cap) {
    OShapes sh = {shapes, hulls, meshes};
    Q4 QA = {qa[0], qa[1], qa[2], qa[3]}, QM = {qm[0], qm[1], qm[2], qm[3]};
    MeshHits m = {out, sub_b, 0, cap};
    if (shapes[sa].type == CC_SHAPE_SPHERE)
        ocx_collide_sphere_mesh(&sh, sa, vload(pa), sm, vload(pm), QM, s, maxSep, vload(move_dir), mesh_hits_cb, &m);
    else
        ocx_collide_mesh(&sh, sa, vload(pa), QA, sm, vload(pm), QM, s, maxSep, vload(move_dir), mesh_hits_cb, &m);
    return m.n;
}
