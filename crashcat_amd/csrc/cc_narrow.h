// Convex narrowphase device code: support functions, GJK closest points, EPA
// (triangle normal / distance / barycentrics in FP64), supporting faces,
// Sutherland-Hodgman clipping, manifold generation and 4-point pruning.
// One pair per lane.  Citations are to the reference (crashcat) sources.
#pragma once
#include "cc_math.h"

namespace cc {

enum { SHAPE_SPHERE = 0, SHAPE_BOX = 1, SHAPE_CAPSULE = 2, SHAPE_HULL = 3, SHAPE_MESH = 4, SHAPE_CYLINDER = 5, SHAPE_PLANE = 6,
       SHAPE_COMPOUND = 7 };

// A BVH node of a triangle mesh: bounds, then for an internal node the right child (node
// index within the mesh; the left child is the next node) and the split axis, for a leaf the
// first triangle (index within the mesh) and -(count + 1).
template <typename T>
struct MNode {
    T mn[3], mx[3];
    int a, b;
};

template <typename T>
struct ShapeD {
    int type, hfp, hnp, hff, hnf;  // meshes: first triangle, triangles, first node, nodes
    int mfv;  // most vertices a supporting face of the shape can have
    T radius, hh, cr;
    T he[3], com[3], amin[3], amax[3];
    T brad;  // radius of a sphere about the shape origin holding the shape (core + convex radius)
    // decorators over the leaf (scaled.ts, transformed.ts, offset-center-of-mass.ts): deco holds
    // the CC_DECO_* bits, sc the scale of the leaf, lp / lq the local transform
    int deco;
    T sc[3], lp[3], lq[4];
    int cfc, ncc;  // compounds: first child (a shape index) and number of children
};

enum { DECO_SCALED = 1, DECO_TRANSFORMED = 2 };

// The scale a shape's leaf is collided with (1 without a scaled decorator).
template <typename T>
__device__ inline V3<T> shape_scale(const ShapeD<T>& d) {
    return (d.deco & DECO_SCALED) ? mk3<T>(d.sc[0], d.sc[1], d.sc[2]) : mk3<T>(1, 1, 1);
}

// The pose a decorated shape's leaf is collided at: a transformed decorator composes the body
// pose with its own, p + rotate(q, t), q * tq (collideTransformedVsShape transformed.ts:370-420).
template <typename T>
__device__ inline void leaf_pose(const ShapeD<T>& d, V3<T>& p, Q4<T>& q) {
    if (!(d.deco & DECO_TRANSFORMED)) return;
    Q4<T> lq;
    lq.x = d.lq[0]; lq.y = d.lq[1]; lq.z = d.lq[2]; lq.w = d.lq[3];
    p = p + qrotate(q, mk3<T>(d.lp[0], d.lp[1], d.lp[2]));
    q = qmul(q, lq);
}

template <typename T>
struct ShapeTable {
    const ShapeD<T>* shapes;
    const T* pts;     // hull points (COM relative), xyz
    const T* shrunk;  // hull points with the convex radius removed
    const int* ffv;   // face first vertex
    const int* fnv;   // face vertex count
    const int* fvi;   // face vertex indices (relative to the hull's first point)
    const T* pn;      // face plane normals
    const T* mpos;    // mesh vertices, xyz
    const int4* mtri; // mesh triangles: vertex a, b, c (indices into mpos), active-edge bits
    const MNode<T>* mnode;
};

// ---------------------------------------------------------------- supports
enum { SK_ZERO, SK_SPHERE_R, SK_BOX, SK_CAP_NC, SK_CAP_R, SK_HULL, SK_CYL };

template <typename T>
struct Sup {
    int kind;
    int np;
    T r, hh, cr;
    V3<T> he;
    const T* pts;
    bool transformed, addRadius;
    bool hscaled;  // a scaled hull's points, scaled as they are read
    T addR;
    M3<T> R;
    V3<T> t;
    V3<T> hs;
};

// getShapeSupportFunction with the shape's scale (sphere.ts:195-263, uniform |scale.x|;
// support.ts:336-368, box: |scale| per axis, the convex radius by the smallest |scale| clamped to
// the default; capsule.ts:305-404 and cylinder.ts:323-419, uniform; convex-hull.ts:979-1010,
// scaled points and the scaled shrunk points and radius).  A unit scale leaves every value as
// the unscaled functions compute it.
template <typename T>
__device__ inline void sup_for_shape(Sup<T>& s, const ShapeTable<T>& st, int si, bool include) {
    const ShapeD<T>& d = st.shapes[si];
    const V3<T> sc = shape_scale(d);
    const T ax = fabs(sc.x), ay = fabs(sc.y), az = fabs(sc.z);
    s.transformed = false;
    s.addRadius = false;
    s.hscaled = false;
    s.np = 0;
    s.pts = nullptr;
    s.r = s.hh = s.cr = T(0);
    s.he = mk3<T>(0, 0, 0);
    switch (d.type) {
        case SHAPE_SPHERE:
            if (include) { s.kind = SK_SPHERE_R; s.r = ax * d.radius; }
            else { s.kind = SK_ZERO; s.cr = ax * d.radius; }
            break;
        case SHAPE_BOX: {
            s.kind = SK_BOX;
            const V3<T> he = mk3<T>(ax * d.he[0], ay * d.he[1], az * d.he[2]);
            if (!include) {
                T cr = tmin<T>(d.cr * tmin<T>(ax, tmin<T>(ay, az)), T(0.05));
                s.he = mk3<T>(tmax<T>(T(0), he.x - cr), tmax<T>(T(0), he.y - cr), tmax<T>(T(0), he.z - cr));
                s.cr = cr;
            } else {
                s.he = he;
            }
            break;
        }
        case SHAPE_CAPSULE:
            s.hh = ax * d.hh;
            if (include) { s.kind = SK_CAP_R; s.r = ax * d.radius; }
            else { s.kind = SK_CAP_NC; s.cr = ax * d.radius; }
            break;
        case SHAPE_CYLINDER:  // setCylinderWithConvexSupport / setCylinderNoConvexSupport (cylinder.ts:323-397)
            s.kind = SK_CYL;
            if (include) { s.hh = ax * d.hh; s.r = ax * d.radius; }
            else { const T cr = ax * d.cr; s.hh = ax * d.hh - cr; s.r = ax * d.radius - cr; s.cr = cr; }
            break;
        default:
            s.kind = SK_HULL;
            s.np = d.hnp;
            if (include || d.cr == T(0)) {
                s.pts = st.pts + 3 * d.hfp;
                s.hscaled = sc.x != T(1) || sc.y != T(1) || sc.z != T(1);
                s.hs = sc;
            } else {
                s.pts = st.shrunk + 3 * d.hfp;
                s.cr = d.cr * tmin<T>(ax, tmin<T>(ay, az));
            }
            break;
    }
}

template <typename T>
__device__ inline V3<T> sup_base(const Sup<T>& s, V3<T> d) {
    switch (s.kind) {
        case SK_ZERO: return mk3<T>(0, 0, 0);
        case SK_SPHERE_R: {
            T l2 = len2(d);
            if (l2 > T(0)) { T sc = s.r / sqrt(l2); return d * sc; }
            return mk3<T>(0, 0, 0);
        }
        case SK_BOX:
            return mk3<T>(d.x >= T(0) ? s.he.x : -s.he.x, d.y >= T(0) ? s.he.y : -s.he.y, d.z >= T(0) ? s.he.z : -s.he.z);
        case SK_CAP_NC: return mk3<T>(0, d.y > T(0) ? s.hh : -s.hh, 0);
        case SK_CYL: {  // cylinder.ts:295-311
            const T hl = sqrt(d.x * d.x + d.z * d.z);
            V3<T> o = mk3<T>(0, d.y >= T(0) ? s.hh : -s.hh, 0);
            if (hl > T(0)) { const T sc = s.r / hl; o.x = d.x * sc; o.z = d.z * sc; }
            return o;
        }
        case SK_CAP_R: {
            T l2 = len2(d);
            if (l2 > T(0)) {
                T sc = s.r / sqrt(l2);
                return mk3<T>(d.x * sc, d.y * sc + (d.y > T(0) ? s.hh : -s.hh), d.z * sc);
            }
            return mk3<T>(0, s.hh, 0);
        }
        default: {
            T best = -INFINITY;
            V3<T> out = mk3<T>(0, 0, 0);
            for (int i = 0; i < s.np; i++) {
                V3<T> p = mk3<T>(s.pts[3 * i], s.pts[3 * i + 1], s.pts[3 * i + 2]);
                if (s.hscaled) p = mk3<T>(p.x * s.hs.x, p.y * s.hs.y, p.z * s.hs.z);
                T dt = dot(p, d);
                if (dt > best) { best = dt; out = p; }
            }
            return out;
        }
    }
}

// a triangle's support (triangleGetSupport support.ts:83-118): the vertex of largest projection
template <typename T>
struct TriSup {
    V3<T> a, b, c;
};
template <typename T>
__device__ inline V3<T> sup_get(const TriSup<T>& s, V3<T> d) {
    const T d1 = s.a.x * d.x + s.a.y * d.y + s.a.z * d.z;
    const T d2 = s.b.x * d.x + s.b.y * d.y + s.b.z * d.z;
    const T d3 = s.c.x * d.x + s.c.y * d.y + s.c.z * d.z;
    if (d1 > d2) return d1 > d3 ? s.a : s.c;
    return d2 > d3 ? s.b : s.c;
}

// TransformedSupport (support.ts:236-261) + AddConvexRadiusSupport (support.ts:288-303)
template <typename T>
__device__ inline V3<T> sup_get(const Sup<T>& s, V3<T> d) {
    V3<T> out;
    if (s.transformed) {
        V3<T> p = sup_base(s, mul_t(s.R, d));
        out = mul(s.R, p) + s.t;
    } else {
        out = sup_base(s, d);
    }
    if (s.addRadius) {
        T l2 = len2(d);
        if (l2 > T(0)) { T sc = s.addR / sqrt(l2); out = out + d * sc; }
    }
    return out;
}

// ------------------------------------------------------- barycentrics
template <typename T> struct Bary { T u, v, w; };

template <typename T>
__device__ inline Bary<T> bary2(V3<T> a, V3<T> b, T tol2) {  // closest-points.ts:19-53
    Bary<T> r = {0, 0, 0};
    V3<T> ab = b - a;
    T den = len2(ab);
    if (den < tol2) {
        if (len2(a) < len2(b)) { r.u = 1; r.v = 0; }
        else { r.u = 0; r.v = 1; }
        return r;
    }
    r.v = -dot(a, ab) / den;
    r.u = T(1) - r.v;
    return r;
}

template <typename T>
__device__ inline Bary<T> bary3(V3<T> a, V3<T> b, V3<T> c, T tol2) {  // closest-points.ts:57-140
    Bary<T> r = {0, 0, 0};
    V3<T> ab = b - a, ac = c - a, bc = c - b;
    T d00 = len2(ab), d11 = len2(ac), d22 = len2(bc);
    if (d00 <= d22) {
        T d01 = dot(ab, ac);
        T den = d00 * d11 - d01 * d01;
        if (fabs(den) < T(1.0e-12)) {
            if (d00 > d11) { Bary<T> o = bary2(a, b, tol2); r.u = o.u; r.v = o.v; r.w = 0; }
            else { Bary<T> o = bary2(a, c, tol2); r.u = o.u; r.w = o.v; r.v = 0; }
            return r;
        }
        T a0 = dot(a, ab), a1 = dot(a, ac);
        r.v = (d01 * a1 - d11 * a0) / den;
        r.w = (d01 * a0 - d00 * a1) / den;
        r.u = T(1) - r.v - r.w;
    } else {
        T d12 = dot(ac, bc);
        T den = d11 * d22 - d12 * d12;
        if (fabs(den) < T(1.0e-12)) {
            if (d11 > d22) { Bary<T> o = bary2(a, c, tol2); r.u = o.u; r.w = o.v; r.v = 0; }
            else { Bary<T> o = bary2(b, c, tol2); r.v = o.u; r.w = o.v; r.u = 0; }
            return r;
        }
        T c1 = dot(c, ac), c2 = dot(c, bc);
        r.u = (d22 * c1 - d12 * c2) / den;
        r.v = (d11 * c2 - d12 * c1) / den;
        r.w = T(1) - r.u - r.v;
    }
    return r;
}

// ------------------------------------------------------------------- GJK
template <typename T> struct SPt { V3<T> y, p, q; };
template <typename T> struct Simplex { SPt<T> pt[4]; int size; };
template <typename T> struct CPRes { V3<T> point; int set; };

template <typename T> CC_HD T clamp01(T x) { return x < T(0) ? T(0) : (x > T(1) ? T(1) : x); }

template <typename T>
__device__ inline CPRes<T> cp_line(V3<T> a, V3<T> b, T tol2) {  // gjk.ts:33-55
    Bary<T> bc = bary2(a, b, tol2);
    CPRes<T> r;
    if (bc.v <= T(0)) { r.point = a; r.set = 1; }
    else if (bc.u <= T(0)) { r.point = b; r.set = 2; }
    else { r.point = a * bc.u + b * bc.v; r.set = 3; }
    return r;
}

template <typename T>
__device__ inline CPRes<T> cp_triangle(V3<T> inA, V3<T> inB, V3<T> inC, bool mustIncludeC, T tol2) {  // gjk.ts:57-309
    CPRes<T> r;
    bool swapAC = len2(inC - inB) < len2(inC - inA);
    V3<T> a = swapAC ? inC : inA;
    V3<T> c = swapAC ? inA : inC;
    V3<T> ab = inB - a, ac = c - a;
    V3<T> n = cross(ab, ac);
    T n2 = len2(n);
    if (n2 < T(1.0e-10)) {
        int set = 4;
        V3<T> best = inC;
        T bestD = len2(inC);
        if (!mustIncludeC) {
            T a2 = len2(inA);
            if (a2 < bestD) { set = 1; best = inA; bestD = a2; }
            T b2 = len2(inB);
            if (b2 < bestD) { set = 2; best = inB; bestD = b2; }
        }
        V3<T> ac2 = c - a;
        T acl = len2(ac2);
        if (acl > tol2) {
            T v = clamp01(-dot(a, ac2) / acl);
            V3<T> q = a + ac2 * v;
            T d2 = len2(q);
            if (d2 < bestD) { set = 5; best = q; bestD = d2; }
        }
        V3<T> bc2 = inC - inB;
        T bcl = len2(bc2);
        if (bcl > tol2) {
            T v = clamp01(-dot(inB, bc2) / bcl);
            V3<T> q = inB + bc2 * v;
            T d2 = len2(q);
            if (d2 < bestD) { set = 6; best = q; bestD = d2; }
        }
        if (!mustIncludeC) {
            V3<T> ab2 = inB - inA;
            T abl = len2(ab2);
            if (abl > tol2) {
                T v = clamp01(-dot(inA, ab2) / abl);
                V3<T> q = inA + ab2 * v;
                T d2 = len2(q);
                if (d2 < bestD) { set = 3; best = q; }
            }
        }
        r.set = set;
        r.point = best;
        return r;
    }
    V3<T> ap = -a;
    T d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= T(0) && d2 <= T(0)) { r.set = swapAC ? 4 : 1; r.point = a; return r; }
    V3<T> bp = -inB;
    T d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= T(0) && d4 <= d3) { r.set = 2; r.point = inB; return r; }
    if (d1 * d4 <= d3 * d2 && d1 >= T(0) && d3 <= T(0)) {
        T v = d1 / (d1 - d3);
        r.set = swapAC ? 6 : 3;
        r.point = a + ab * v;
        return r;
    }
    V3<T> cp = -c;
    T d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= T(0) && d5 <= d6) { r.set = swapAC ? 1 : 4; r.point = c; return r; }
    if (d5 * d2 <= d1 * d6 && d2 >= T(0) && d6 <= T(0)) {
        T w = d2 / (d2 - d6);
        r.set = 5;
        r.point = a + ac * w;
        return r;
    }
    T d43 = d4 - d3, d56 = d5 - d6;
    if (d3 * d6 <= d5 * d4 && d43 >= T(0) && d56 >= T(0)) {
        T w = d43 / (d43 + d56);
        r.set = swapAC ? 3 : 6;
        r.point = inB + (c - inB) * w;
        return r;
    }
    r.set = 7;
    V3<T> sum = a + inB + c;
    T sc = dot(sum, n) / (T(3) * n2);
    r.point = n * sc;
    return r;
}

template <typename T>
__device__ inline void origin_outside(bool out[4], V3<T> a, V3<T> b, V3<T> c, V3<T> d, T tol) {  // gjk.ts:314-394
    V3<T> ab = b - a, ac = c - a, ad = d - a, bd = d - b, bc = c - b;
    V3<T> abac = cross(ab, ac), acad = cross(ac, ad), adab = cross(ad, ab), bdbc = cross(bd, bc);
    T sp[4] = {dot(a, abac), dot(a, acad), dot(a, adab), dot(b, bdbc)};
    T sd[4] = {dot(ad, abac), dot(ab, acad), dot(ac, adab), -dot(ab, bdbc)};
    if (sd[0] > T(0) && sd[1] > T(0) && sd[2] > T(0) && sd[3] > T(0)) {
        for (int i = 0; i < 4; i++) out[i] = sp[i] >= -tol;
        return;
    }
    if (sd[0] < T(0) && sd[1] < T(0) && sd[2] < T(0) && sd[3] < T(0)) {
        for (int i = 0; i < 4; i++) out[i] = sp[i] <= tol;
        return;
    }
    out[0] = out[1] = out[2] = out[3] = true;
}

template <typename T>
__device__ inline CPRes<T> cp_tetra(V3<T> a, V3<T> b, V3<T> c, V3<T> d, bool mustIncludeD, T tol) {  // gjk.ts:399-487
    T tol2 = tol * tol;
    CPRes<T> r;
    r.set = 15;
    r.point = mk3<T>(0, 0, 0);
    T best = INFINITY;
    bool o[4];
    origin_outside(o, a, b, c, d, tol);
    if (o[0]) {
        if (mustIncludeD) { r.set = 1; r.point = a; }
        else r = cp_triangle(a, b, c, false, tol2);
        best = len2(r.point);
    }
    if (o[1]) {
        CPRes<T> t = cp_triangle(a, c, d, mustIncludeD, tol2);
        T d2 = len2(t.point);
        if (d2 < best) { best = d2; r.point = t.point; r.set = (t.set & 1) + ((t.set & 6) << 1); }
    }
    if (o[2]) {
        CPRes<T> t = cp_triangle(a, b, d, mustIncludeD, tol2);
        T d2 = len2(t.point);
        if (d2 < best) { best = d2; r.point = t.point; r.set = (t.set & 3) + ((t.set & 4) << 1); }
    }
    if (o[3]) {
        CPRes<T> t = cp_triangle(b, c, d, mustIncludeD, tol2);
        T d2 = len2(t.point);
        if (d2 < best) { r.point = t.point; r.set = t.set << 1; }
    }
    return r;
}

template <typename T> struct GjkOut { V3<T> pointA, pointB, axis; T dist2; Simplex<T> simplex; };

// gjkClosestPoints gjk.ts:1294-1546
template <typename T, typename SB>
__device__ inline void gjk_closest_points(GjkOut<T>& out, const Sup<T>& A, const SB& B, T tolerance, V3<T> direction,
                                          T maxDist2) {
    const T GJK_TOL = T(1e-5);
    T tol2 = tolerance * tolerance;
    Simplex<T>& s = out.simplex;
    s.size = 0;
    V3<T> cpt = direction;
    T cd2 = len2(direction);
    int cset = 0;
    T prevDist2 = sizeof(T) == 8 ? T(1.7976931348623157e308) : T(3.402823466e38);
    int it = 0;
    while (it++ < 100) {
        V3<T> p = sup_get(A, cpt);
        V3<T> q = sup_get(B, -cpt);
        V3<T> w = p - q;
        T dt = dot(cpt, w);
        if (dt < T(0) && dt * dt > cd2 * maxDist2) {
            out.dist2 = INFINITY;
            out.axis = cpt;
            s.size = 0;
            out.pointA = out.pointB = mk3<T>(0, 0, 0);
            return;
        }
        {  // append with constant indices only, so the simplex stays in registers
            SPt<T> np_;
            np_.y = w;
            np_.p = p;
            np_.q = q;
            switch (s.size) {
                case 0: s.pt[0] = np_; break;
                case 1: s.pt[1] = np_; break;
                case 2: s.pt[2] = np_; break;
                default: s.pt[3] = np_; break;
            }
        }
        s.size++;
        CPRes<T> cp;
        switch (s.size) {
            case 1: cp.set = 1; cp.point = s.pt[0].y; break;
            case 2: cp = cp_line(s.pt[0].y, s.pt[1].y, T(1e-10)); break;
            case 3: cp = cp_triangle(s.pt[0].y, s.pt[1].y, s.pt[2].y, true, T(1e-10)); break;
            default: cp = cp_tetra(s.pt[0].y, s.pt[1].y, s.pt[2].y, s.pt[3].y, true, T(1e-5)); break;
        }
        T d2 = len2(cp.point);
        if (!(d2 < prevDist2)) { s.size--; break; }
        cpt = cp.point;
        cd2 = d2;
        cset = cp.set;
        if (cset == 0xf) { cpt = mk3<T>(0, 0, 0); cd2 = T(0); break; }
        {  // keep the points named by cset, in order (slot k <- k-th set bit), by selects
            const int keep = cset & ((1 << s.size) - 1);
            const SPt<T> t0 = s.pt[0], t1 = s.pt[1], t2 = s.pt[2], t3 = s.pt[3];
            auto pick = [&](int c) -> SPt<T> { return (c & 1) ? t0 : ((c & 2) ? t1 : ((c & 4) ? t2 : t3)); };
            const int c1 = keep & (keep - 1), c2 = c1 & (c1 - 1), c3 = c2 & (c2 - 1);
            s.pt[0] = pick(keep);
            s.pt[1] = pick(c1);
            s.pt[2] = pick(c2);
            s.pt[3] = pick(c3);
            s.size = __popc((unsigned)keep);
        }
        if (cd2 <= tol2) { cpt = mk3<T>(0, 0, 0); cd2 = T(0); break; }
        T ymax = T(0);
        if (s.size > 0) ymax = tmax<T>(ymax, len2(s.pt[0].y));
        if (s.size > 1) ymax = tmax<T>(ymax, len2(s.pt[1].y));
        if (s.size > 2) ymax = tmax<T>(ymax, len2(s.pt[2].y));
        if (s.size > 3) ymax = tmax<T>(ymax, len2(s.pt[3].y));
        if (cd2 <= GJK_TOL * ymax) { cpt = mk3<T>(0, 0, 0); cd2 = T(0); break; }
        cpt = -cpt;
        if (prevDist2 - cd2 <= GJK_TOL * prevDist2) break;
        prevDist2 = cd2;
    }
    out.pointA = out.pointB = mk3<T>(0, 0, 0);
    if (s.size == 0) { out.dist2 = INFINITY; out.axis = mk3<T>(0, 0, 0); return; }
    if (s.size == 1) { out.pointA = s.pt[0].p; out.pointB = s.pt[0].q; }
    else if (s.size == 2) {
        Bary<T> b = bary2(s.pt[0].y, s.pt[1].y, T(1e-10));
        out.pointA = s.pt[0].p * b.u + s.pt[1].p * b.v;
        out.pointB = s.pt[0].q * b.u + s.pt[1].q * b.v;
    } else if (s.size == 3) {
        Bary<T> b = bary3(s.pt[0].y, s.pt[1].y, s.pt[2].y, T(1e-10));
        out.pointA = s.pt[0].p * b.u + s.pt[1].p * b.v + s.pt[2].p * b.w;
        out.pointB = s.pt[0].q * b.u + s.pt[1].q * b.v + s.pt[2].q * b.w;
    }
    out.axis = cpt;
    out.dist2 = cd2;
}

enum { PD_NOT_COLLIDING = 0, PD_COLLIDING = 1, PD_INDETERMINATE = 2 };
template <typename T> struct PenDepth { int status; V3<T> axis, pointA, pointB; };

// penetrationDepthStepGJK penetration.ts:35-110
template <typename T, typename SB>
__device__ inline void pd_step_gjk(PenDepth<T>& out, Simplex<T>& outSimplex, const Sup<T>& A, const SB& B, T crA, T crB,
                                   V3<T> dir, T tol) {
    T cr = crA + crB, cr2 = cr * cr;
    GjkOut<T> g;
    gjk_closest_points(g, A, B, tol, dir, cr2);
    if (g.dist2 > cr2) {
        out.pointA = out.pointB = out.axis = mk3<T>(0, 0, 0);
        out.status = PD_NOT_COLLIDING;
        return;
    }
    out.pointA = g.pointA;
    out.pointB = g.pointB;
    out.axis = g.axis;
    outSimplex = g.simplex;
    if (g.dist2 > T(0)) {
        T vl = sqrt(g.dist2);
        out.pointA = out.pointA + out.axis * (crA / vl);
        out.pointB = out.pointB + out.axis * (-(crB / vl));
        out.status = PD_COLLIDING;
        return;
    }
    out.status = PD_INDETERMINATE;
}

// ------------------------------------------------------------------- EPA
// epa-convex-hull-builder.ts: triangle pool (256), min-heap on closestLengthSq,
// horizon search with an explicit stack; normal, centroid, distance and the
// barycentric lambdas of each face are kept in FP64.
#define EPA_MAX_TRIANGLES 256
#define EPA_MAX_EDGE_LENGTH 128
#define EPA_POINT_CAP 160
#define EPA_MAX_POINTS_TO_INCLUDE_ORIGIN_IN_HULL 32
#define EPA_MAX_POINTS 128
#define EPA_LANES 64

struct EEdge { short nbT; unsigned char nbE; unsigned char start; };
struct EStack { short tri; short edge; short iter; short pad; };
enum { TNX, TNY, TNZ, TCX, TCY, TCZ, TCLS, TL0, TL1, TNF };  // the FP64 fields of a triangle
enum { TF_REL0, TF_INTERIOR, TF_REMOVED, TF_INQUEUE };     // its flag bytes

// The polytopes of the EPAs of one wavefront, one per lane, interleaved by lane: element k
// of an array of lane l sits at k * 64 + l.  The lanes' polytopes grow alike (the same first
// triangles and point slots), so a wavefront's accesses share memory lines.
template <typename T>
struct EpaWave {  // byte offsets of the arrays in one wavefront's block
    static constexpr size_t L = EPA_LANES;
    static constexpr size_t oY = 0;
    static constexpr size_t oP = oY + sizeof(V3<T>) * EPA_POINT_CAP * L;
    static constexpr size_t oQ = oP + sizeof(V3<T>) * EPA_POINT_CAP * L;
    static constexpr size_t oTD = oQ + sizeof(V3<T>) * EPA_POINT_CAP * L;
    static constexpr size_t oTE = oTD + sizeof(double) * TNF * EPA_MAX_TRIANGLES * L;
    static constexpr size_t oTF = oTE + sizeof(EEdge) * 3 * EPA_MAX_TRIANGLES * L;
    static constexpr size_t oFL = oTF + 4 * EPA_MAX_TRIANGLES * L;
    static constexpr size_t oQU = oFL + sizeof(short) * 2 * EPA_MAX_TRIANGLES * L;
    static constexpr size_t oHC = oQU + sizeof(short) * 2 * EPA_MAX_TRIANGLES * L;
    static constexpr size_t oST = oHC + sizeof(double) * 2 * EPA_MAX_TRIANGLES * L;
    static constexpr size_t oED = oST + sizeof(EStack) * EPA_MAX_EDGE_LENGTH * L;
    static constexpr size_t oNT = oED + sizeof(EEdge) * EPA_MAX_EDGE_LENGTH * L;
    static constexpr size_t bytes = (oNT + sizeof(short) * EPA_MAX_EDGE_LENGTH * L + 255) / 256 * 256;
};

// One lane's polytope: accessors into the wavefront's block, counters in registers.
// Triangles come from a bump allocator and a stack of freed ones (freed first), which hands
// out the same indices as one free list holding 0..255 in pop order.
template <typename T>
struct EpaH {
    char* b;
    int lane;
    int n, bump, nrec, nq, nedges, nnew;
    typedef EpaWave<T> W;
    __device__ V3<T>& y(int i) const { return ((V3<T>*)(b + W::oY))[i * EPA_LANES + lane]; }
    __device__ V3<T>& p(int i) const { return ((V3<T>*)(b + W::oP))[i * EPA_LANES + lane]; }
    __device__ V3<T>& q(int i) const { return ((V3<T>*)(b + W::oQ))[i * EPA_LANES + lane]; }
    __device__ double& td(int f, int t) const { return ((double*)(b + W::oTD))[(f * EPA_MAX_TRIANGLES + t) * EPA_LANES + lane]; }
    __device__ EEdge& te(int t, int e) const { return ((EEdge*)(b + W::oTE))[(e * EPA_MAX_TRIANGLES + t) * EPA_LANES + lane]; }
    __device__ unsigned char& tf(int t, int f) const { return ((unsigned char*)(b + W::oTF))[((t * EPA_LANES + lane) << 2) + f]; }
    __device__ short& fl(int i) const { return ((short*)(b + W::oFL))[i * EPA_LANES + lane]; }
    __device__ short& qu(int i) const { return ((short*)(b + W::oQU))[i * EPA_LANES + lane]; }
    // the closestLengthSq of the triangle in heap slot i, kept beside the slot: a triangle in the
    // heap is never freed or rewritten, so the copy stays equal to its cls
    __device__ double& hc(int i) const { return ((double*)(b + W::oHC))[i * EPA_LANES + lane]; }
    __device__ EStack& st(int i) const { return ((EStack*)(b + W::oST))[i * EPA_LANES + lane]; }
    __device__ EEdge& ed(int i) const { return ((EEdge*)(b + W::oED))[i * EPA_LANES + lane]; }
    __device__ short& nt(int i) const { return ((short*)(b + W::oNT))[i * EPA_LANES + lane]; }
    __device__ double cls(int t) const { return td(TCLS, t); }
};

template <typename T>
__device__ inline bool tri_facing(const EpaH<T>& h, int t, double px, double py, double pz) {
    return h.td(TNX, t) * (px - h.td(TCX, t)) + h.td(TNY, t) * (py - h.td(TCY, t)) + h.td(TNZ, t) * (pz - h.td(TCZ, t)) > 0.0;
}

template <typename T>
__device__ inline void heap_push(EpaH<T>& h, int ti) {
    if (h.nq >= EPA_MAX_TRIANGLES * 2) return;
    const double c = h.cls(ti);
    int cur = h.nq++;
    h.tf(ti, TF_INQUEUE) = 1;
    while (cur > 0) {  // sift up; the new entry is written once, at its final slot
        int par = (cur - 1) >> 1;
        const double cp = h.hc(par);
        if (cp > c) {
            h.qu(cur) = h.qu(par);
            h.hc(cur) = cp;
            cur = par;
        } else break;
    }
    h.qu(cur) = (short)ti;
    h.hc(cur) = c;
}

template <typename T>
__device__ inline int heap_pop(EpaH<T>& h) {
    if (h.nq == 0) return -1;
    const short top = h.qu(0);
    const double topc = h.hc(0);
    const int count = h.nq - 1;
    // the last entry moves to the root and sifts down (the reference swaps it there first)
    const short mv = h.qu(count);
    const double mc = h.hc(count);
    int cur = 0;
    for (;;) {
        int child = (cur << 1) + 1;
        if (child >= count) break;
        int largest = cur;
        double lc = mc;
        const double c1 = h.hc(child);
        if (lc > c1) { largest = child; lc = c1; }
        ++child;
        if (child < count) {
            const double c2 = h.hc(child);
            if (lc > c2) { largest = child; lc = c2; }
        }
        if (largest == cur) break;
        h.qu(cur) = h.qu(largest);
        h.hc(cur) = lc;
        cur = largest;
    }
    h.qu(cur) = mv;
    h.hc(cur) = mc;
    h.qu(count) = top;
    h.hc(count) = topc;
    h.nq = count;
    return top;
}

template <typename T>
__device__ inline void free_tri(EpaH<T>& h, int ti) {  // a free list of 512 entries, as (256 - bump) + nrec
    if ((EPA_MAX_TRIANGLES - h.bump) + h.nrec < EPA_MAX_TRIANGLES * 2) h.fl(h.nrec++) = (short)ti;
}

template <typename T>
__device__ inline int create_tri(EpaH<T>& h, int i1, int i2, int i3) {  // createTriangle :179-345
    int ti;
    if (h.nrec > 0) ti = h.fl(--h.nrec);
    else if (h.bump < EPA_MAX_TRIANGLES) ti = h.bump++;
    else return -1;
    double cls = INFINITY, l0 = 0.0, l1 = 0.0;
    unsigned char rel0 = 0, interior = 0;
    V3<double> y0 = cvt<double>(h.y(i1)), y1 = cvt<double>(h.y(i2)), y2 = cvt<double>(h.y(i3));
    double cx = (y0.x + y1.x + y2.x) / 3.0, cy = (y0.y + y1.y + y2.y) / 3.0, cz = (y0.z + y1.z + y2.z) / 3.0;
    V3<double> cen = mk3<double>(cx, cy, cz);
    V3<double> y10 = y1 - y0, y20 = y2 - y0, y21 = y2 - y1;
    double y20d = len2(y20), y21d = len2(y21);
    V3<double> n;
    if (y20d < y21d) {
        n = cross(y10, y20);
        double nl = len2(n);
        if (nl > 1e-10) {
            double cdn = dot(cen, n);
            cls = fabs(cdn) * cdn / nl;
            double y10d = len2(y10), y10y20 = dot(y10, y20);
            double det = y10d * y20d - y10y20 * y10y20;
            if (det > 0.0) {
                double y0y10 = dot(y0, y10), y0y20 = dot(y0, y20);
                l0 = (y10y20 * y0y20 - y20d * y0y10) / det;
                l1 = (y10y20 * y0y10 - y10d * y0y20) / det;
                rel0 = 1;
                if (l0 > -1e-3 && l1 > -1e-3 && l0 + l1 < 1.0 + 1e-3) interior = 1;
            }
        }
    } else {
        n = cross(y10, y21);
        double nl = len2(n);
        if (nl > 1e-10) {
            double cdn = dot(cen, n);
            cls = fabs(cdn) * cdn / nl;
            double y10d = len2(y10), y10y21 = dot(y10, y21);
            double det = y10d * y21d - y10y21 * y10y21;
            if (det > 0.0) {
                double y1y10 = dot(y1, y10), y1y21 = dot(y1, y21);
                l0 = (y21d * y1y10 - y10y21 * y1y21) / det;
                l1 = (y10y21 * y1y10 - y10d * y1y21) / det;
                rel0 = 0;
                if (l0 > -1e-3 && l1 > -1e-3 && l0 + l1 < 1.0 + 1e-3) interior = 1;
            }
        }
    }
    h.td(TNX, ti) = n.x; h.td(TNY, ti) = n.y; h.td(TNZ, ti) = n.z;
    h.td(TCX, ti) = cx; h.td(TCY, ti) = cy; h.td(TCZ, ti) = cz;
    h.td(TCLS, ti) = cls; h.td(TL0, ti) = l0; h.td(TL1, ti) = l1;
    h.tf(ti, TF_REL0) = rel0; h.tf(ti, TF_INTERIOR) = interior; h.tf(ti, TF_REMOVED) = 0; h.tf(ti, TF_INQUEUE) = 0;
    EEdge e;
    e.nbT = -1; e.nbE = 0;
    e.start = (unsigned char)i1; h.te(ti, 0) = e;
    e.start = (unsigned char)i2; h.te(ti, 1) = e;
    e.start = (unsigned char)i3; h.te(ti, 2) = e;
    return ti;
}

template <typename T>
__device__ inline void link_tri(EpaH<T>& h, int t1, int e1, int t2, int e2) {
    EEdge& a = h.te(t1, e1);
    a.nbT = (short)t2; a.nbE = (unsigned char)e2;
    EEdge& b = h.te(t2, e2);
    b.nbT = (short)t1; b.nbE = (unsigned char)e1;
}

template <typename T>
__device__ inline void unlink_tri(EpaH<T>& h, int ti) {  // unlinkTriangle :451-466
    for (int i = 0; i < 3; i++) {
        EEdge e = h.te(ti, i);
        if (e.nbT >= 0) {
            h.te(e.nbT, e.nbE).nbT = -1;
            h.te(ti, i).nbT = -1;
        }
    }
    if (!h.tf(ti, TF_INQUEUE)) free_tri(h, ti);
}

template <typename T>
__device__ inline bool find_edge(EpaH<T>& h, int facing, double px, double py, double pz) {  // findEdge :468-539
    h.nedges = 0;
    h.tf(facing, TF_REMOVED) = 1;
    int sp = 0;
    EStack s0;
    s0.tri = (short)facing; s0.edge = 0; s0.iter = -1; s0.pad = 0;
    h.st(0) = s0;
    int nextExpected = -1;
    for (;;) {
        EStack cur = h.st(sp);
        if (++cur.iter >= 3) {
            unlink_tri(h, cur.tri);
            if (--sp < 0) break;
        } else {
            h.st(sp).iter = cur.iter;
            EEdge e = h.te(cur.tri, (cur.edge + cur.iter) % 3);
            int n = e.nbT;
            if (n >= 0 && !h.tf(n, TF_REMOVED)) {
                if (tri_facing(h, n, px, py, pz)) {
                    h.tf(n, TF_REMOVED) = 1;
                    if (sp + 1 >= EPA_MAX_EDGE_LENGTH) return false;
                    sp++;
                    EStack s1;
                    s1.tri = (short)n; s1.edge = e.nbE; s1.iter = 0; s1.pad = 0;
                    h.st(sp) = s1;
                } else {
                    if (e.start != nextExpected && nextExpected != -1) return false;
                    nextExpected = h.te(n, e.nbE).start;
                    if (h.nedges >= EPA_MAX_EDGE_LENGTH) return false;
                    h.ed(h.nedges++) = e;
                }
            }
        }
    }
    if (h.nedges > 0 && h.ed(0).start != nextExpected) return false;
    return h.nedges >= 3;
}

template <typename T>
__device__ inline bool add_point(EpaH<T>& h, int facing, int idx, double closestDist2) {  // addPoint :541-593
    V3<double> p = cvt<double>(h.y(idx));
    h.nnew = 0;
    if (!find_edge(h, facing, p.x, p.y, p.z)) return false;
    int ne = h.nedges;
    for (int i = 0; i < ne; i++) {
        int in = (i + 1) % ne;
        int nt = create_tri(h, h.ed(i).start, h.ed(in).start, idx);
        if (nt < 0) return false;
        h.nt(h.nnew++) = (short)nt;
        if ((h.tf(nt, TF_INTERIOR) && h.cls(nt) < closestDist2) || h.cls(nt) < 0.0) heap_push(h, nt);
    }
    for (int i = 0; i < ne; i++) {
        int in = (i + 1) % ne;
        EEdge e = h.ed(i);
        link_tri(h, h.nt(i), 0, e.nbT, e.nbE);
        link_tri(h, h.nt(i), 1, h.nt(in), 2);
    }
    return true;
}

template <typename T, typename SB>
__device__ inline int epa_add(EpaH<T>& P, const Sup<T>& A, const SB& B, V3<T> d) {
    V3<T> p = sup_get(A, d);
    V3<T> q = sup_get(B, -d);
    int i = P.n++;
    P.y(i) = p - q;
    P.p(i) = p;
    P.q(i) = q;
    return i;
}

// penetrationDepthStepEPA penetration.ts:214-614; the polytope lives in device memory.
template <typename T, typename SB>
__device__ inline bool pd_step_epa(PenDepth<T>& out, const Sup<T>& A, const SB& B, T tol, const Simplex<T>& s,
                                   EpaH<T>& H) {
    H.n = s.size;
    for (int i = 0; i < s.size; i++) { H.y(i) = s.pt[i].y; H.p(i) = s.pt[i].p; H.q(i) = s.pt[i].q; }
    if (H.n == 1) {
        H.n--;
        epa_add(H, A, B, mk3<T>(0, 1, 0));
        epa_add(H, A, B, mk3<T>(-1, -1, -1));
        epa_add(H, A, B, mk3<T>(1, -1, -1));
        epa_add(H, A, B, mk3<T>(0, -1, 1));
    } else if (H.n == 2) {
        V3<T> axis = H.y(1) - H.y(0);
        T al = vlen(axis);
        V3<T> an = mk3<T>(axis.x / al, axis.y / al, axis.z / al);
        V3<T> d1;
        if (fabs(an.x) > fabs(an.y)) { T pl = sqrt(an.x * an.x + an.z * an.z); d1 = mk3<T>(an.z / pl, 0, -an.x / pl); }
        else { T pl = sqrt(an.y * an.y + an.z * an.z); d1 = mk3<T>(0, an.z / pl, -an.y / pl); }
        const T c = T(-0.4999999999999998), sn = T(0.8660254037844387), omc = T(1) - c;  // cos/sin of 120 degrees
        T r00 = c + an.x * an.x * omc, r01 = an.x * an.y * omc - an.z * sn, r02 = an.x * an.z * omc + an.y * sn;
        T r10 = an.y * an.x * omc + an.z * sn, r11 = c + an.y * an.y * omc, r12 = an.y * an.z * omc - an.x * sn;
        T r20 = an.z * an.x * omc - an.y * sn, r21 = an.z * an.y * omc + an.x * sn, r22 = c + an.z * an.z * omc;
        V3<T> d2 = mk3<T>(r00 * d1.x + r01 * d1.y + r02 * d1.z, r10 * d1.x + r11 * d1.y + r12 * d1.z, r20 * d1.x + r21 * d1.y + r22 * d1.z);
        V3<T> d3 = mk3<T>(r00 * d2.x + r01 * d2.y + r02 * d2.z, r10 * d2.x + r11 * d2.y + r12 * d2.z, r20 * d2.x + r21 * d2.y + r22 * d2.z);
        epa_add(H, A, B, d1);
        epa_add(H, A, B, d2);
        epa_add(H, A, B, d3);
    }
    // initialize :347-371
    H.bump = 0;
    H.nrec = 0;
    H.nq = 0;
    {
        int t1 = create_tri(H, 0, 1, 2);
        int t2 = create_tri(H, 0, 2, 1);
        link_tri(H, t1, 0, t2, 2);
        link_tri(H, t1, 1, t2, 1);
        link_tri(H, t1, 2, t2, 0);
        heap_push(H, t1);
        heap_push(H, t2);
    }
    for (int i = 3; i < H.n; i++) {
        // findFacingTriangle :420-445
        int best = -1;
        double bestD = 0.0;
        V3<double> p = cvt<double>(H.y(i));
        for (int k = 0; k < H.nq; k++) {
            const int t = H.qu(k);
            if (H.tf(t, TF_REMOVED)) continue;
            const double nx = H.td(TNX, t), ny = H.td(TNY, t), nz = H.td(TNZ, t);
            double dt = nx * (p.x - H.td(TCX, t)) + ny * (p.y - H.td(TCY, t)) + nz * (p.z - H.td(TCZ, t));
            if (dt > 0.0) {
                double d2 = dt * dt / (nx * nx + ny * ny + nz * nz);
                if (d2 > bestD) { best = t; bestD = d2; }
            }
        }
        if (best >= 0) {
            if (!add_point(H, best, i, 1.7976931348623157e308)) { out.status = PD_NOT_COLLIDING; return false; }
        }
    }
    for (;;) {
        if (H.nq == 0) { out.status = PD_NOT_COLLIDING; return false; }
        int t = H.qu(0);
        if (H.tf(t, TF_REMOVED)) {
            heap_pop(H);
            if (H.nq == 0) { out.status = PD_NOT_COLLIDING; return false; }
            free_tri(H, t);
            continue;
        }
        if (H.cls(t) >= 0.0) break;
        heap_pop(H);
        V3<T> nrm = mk3<T>(T(H.td(TNX, t)), T(H.td(TNY, t)), T(H.td(TNZ, t)));
        if (H.n >= EPA_POINT_CAP) { out.status = PD_NOT_COLLIDING; return false; }
        int ni = epa_add(H, A, B, nrm);
        V3<double> w = cvt<double>(H.y(ni));
        if (!tri_facing(H, t, w.x, w.y, w.z) || !add_point(H, t, ni, 1.7976931348623157e308)) {
            out.status = PD_NOT_COLLIDING;
            return false;
        }
        free_tri(H, t);
        if (H.nq == 0 || H.n >= EPA_MAX_POINTS_TO_INCLUDE_ORIGIN_IN_HULL) { out.status = PD_NOT_COLLIDING; return false; }
    }
    double closestDist2 = 1.7976931348623157e308;
    int last = -1;
    bool flip = false;
    do {
        int t = heap_pop(H);
        if (t < 0) { out.status = PD_NOT_COLLIDING; return false; }
        if (H.tf(t, TF_REMOVED)) { free_tri(H, t); continue; }
        const double tcls = H.cls(t);
        if (tcls >= closestDist2) break;
        if (last >= 0) free_tri(H, last);
        last = t;
        const double nx = H.td(TNX, t), ny = H.td(TNY, t), nz = H.td(TNZ, t);
        V3<T> nrm = mk3<T>(T(nx), T(ny), T(nz));
        if (H.n >= EPA_POINT_CAP) break;
        int ni = epa_add(H, A, B, nrm);
        V3<double> w = cvt<double>(H.y(ni));
        double dt = nx * w.x + ny * w.y + nz * w.z;
        if (dt < 0.0) { out.status = PD_NOT_COLLIDING; return false; }
        double d2 = dt * dt / (nx * nx + ny * ny + nz * nz);
        if (d2 - tcls < tcls * double(tol)) break;
        closestDist2 = d2 < closestDist2 ? d2 : closestDist2;
        if (!tri_facing(H, t, w.x, w.y, w.z)) break;
        if (!add_point(H, t, ni, closestDist2)) break;
        bool defect = false;
        for (int i = 0; i < H.nnew; i++) {
            const int q = H.nt(i);
            if (H.td(TNX, q) * H.td(TCX, q) + H.td(TNY, q) * H.td(TCY, q) + H.td(TNZ, q) * H.td(TCZ, q) < 0.0) { defect = true; break; }
        }
        if (defect) {
            V3<T> p2 = sup_get(A, -nrm);
            V3<T> q2 = sup_get(B, nrm);
            V3<double> w2 = cvt<double>(p2 - q2);
            double dot2 = -(H.td(TNX, t) * w2.x + H.td(TNY, t) * w2.y + H.td(TNZ, t) * w2.z);
            if (dot2 < dt) flip = true;
            break;
        }
    } while (H.nq > 0 && H.n < EPA_MAX_POINTS);
    if (last < 0) { out.status = PD_NOT_COLLIDING; return false; }
    const double Lnx = H.td(TNX, last), Lny = H.td(TNY, last), Lnz = H.td(TNZ, last);
    double nl2 = Lnx * Lnx + Lny * Lny + Lnz * Lnz;
    double cdn = H.td(TCX, last) * Lnx + H.td(TCY, last) * Lny + H.td(TCZ, last) * Lnz;
    double sc = cdn / nl2;
    V3<double> pn = mk3<double>(Lnx * sc, Lny * sc, Lnz * sc);
    if (len2(pn) < 1e-10) { out.status = PD_NOT_COLLIDING; return false; }
    if (flip) pn = -pn;
    int i0 = H.te(last, 0).start, i1 = H.te(last, 1).start, i2 = H.te(last, 2).start;
    T l0 = T(H.td(TL0, last)), l1 = T(H.td(TL1, last));
    if (H.tf(last, TF_REL0)) {
        out.pointA = H.p(i0) + (H.p(i1) - H.p(i0)) * l0 + (H.p(i2) - H.p(i0)) * l1;
        out.pointB = H.q(i0) + (H.q(i1) - H.q(i0)) * l0 + (H.q(i2) - H.q(i0)) * l1;
    } else {
        out.pointA = H.p(i1) + (H.p(i0) - H.p(i1)) * l0 + (H.p(i2) - H.p(i1)) * l1;
        out.pointB = H.q(i1) + (H.q(i0) - H.q(i1)) * l0 + (H.q(i2) - H.q(i1)) * l1;
    }
    out.axis = cvt<T>(pn);
    out.status = PD_COLLIDING;
    return true;
}

// ------------------------------------------------------------------ faces
// A face or clipped polygon of at most N vertices.  A supporting face has at most 16
// vertices and clipping two faces of n1 and n2 vertices gives at most n1 + n2, so the general
// capacity is 32; pairs whose shapes' faces have at most four vertices use capacity 8.
#define CC_MAX_FACE 32
#define CC_SMALL_FACE 8
template <typename T, int N = CC_MAX_FACE> struct Face { int n; V3<T> v[N]; };

template <typename T, int N>
__device__ inline void face_transform(Face<T, N>& f, V3<T> pos, Q4<T> q, V3<T> sc) {  // transformFace face.ts:30-84
    M3<T> R = m3_from_quat(q);
    for (int k = 0; k < 3; k++) { R.m[k] *= sc.x; R.m[3 + k] *= sc.y; R.m[6 + k] *= sc.z; }
    for (int i = 0; i < f.n; i++) f.v[i] = mul(R, f.v[i]) + pos;
}

template <typename T, int N>
__device__ inline void supporting_face(Face<T, N>& f, const ShapeTable<T>& st, int si, V3<T> d, V3<T> pos, Q4<T> q) {
    const ShapeD<T>& s = st.shapes[si];
    const V3<T> sc = shape_scale(s);
    const bool insideOut = sc.x * sc.y * sc.z < T(0);  // isScaleInsideOut face.ts:87-89
    if (s.type == SHAPE_SPHERE) { f.n = 0; return; }  // sphere.ts:143-146
    if (s.type == SHAPE_BOX) {  // box.ts:129-332
        T hx = s.he[0], hy = s.he[1], hz = s.he[2];
        T ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
        f.n = 4;
        if (ax >= ay && ax >= az) {
            if (d.x < T(0)) { f.v[0] = mk3<T>(hx, -hy, -hz); f.v[1] = mk3<T>(hx, hy, -hz); f.v[2] = mk3<T>(hx, hy, hz); f.v[3] = mk3<T>(hx, -hy, hz); }
            else { f.v[0] = mk3<T>(-hx, -hy, -hz); f.v[1] = mk3<T>(-hx, -hy, hz); f.v[2] = mk3<T>(-hx, hy, hz); f.v[3] = mk3<T>(-hx, hy, -hz); }
        } else if (ay >= ax && ay >= az) {
            if (d.y < T(0)) { f.v[0] = mk3<T>(-hx, hy, -hz); f.v[1] = mk3<T>(-hx, hy, hz); f.v[2] = mk3<T>(hx, hy, hz); f.v[3] = mk3<T>(hx, hy, -hz); }
            else { f.v[0] = mk3<T>(-hx, -hy, -hz); f.v[1] = mk3<T>(hx, -hy, -hz); f.v[2] = mk3<T>(hx, -hy, hz); f.v[3] = mk3<T>(-hx, -hy, hz); }
        } else {
            if (d.z < T(0)) { f.v[0] = mk3<T>(-hx, -hy, hz); f.v[1] = mk3<T>(hx, -hy, hz); f.v[2] = mk3<T>(hx, hy, hz); f.v[3] = mk3<T>(-hx, hy, hz); }
            else { f.v[0] = mk3<T>(-hx, -hy, -hz); f.v[1] = mk3<T>(-hx, hy, -hz); f.v[2] = mk3<T>(hx, hy, -hz); f.v[3] = mk3<T>(hx, -hy, -hz); }
        }
        if (insideOut) {  // the same face, listed in reverse
            V3<T> t = f.v[0]; f.v[0] = f.v[3]; f.v[3] = t;
            t = f.v[1]; f.v[1] = f.v[2]; f.v[2] = t;
        }
    } else if (s.type == SHAPE_CAPSULE) {  // capsule.ts:193-253
        T len = sqrt(d.x * d.x + d.z * d.z);
        if (len == T(0)) { f.n = 0; return; }
        T sx = (s.radius / len) * d.x, sz = (s.radius / len) * d.z;
        V3<T> top = mk3<T>(-sx, s.hh, -sz), bot = mk3<T>(-sx, -s.hh, -sz);
        T pt = dot(top, d), pb = dot(bot, d);
        if (fabs(pt - pb) < T(0.02) * vlen(d)) { f.n = 2; f.v[0] = top; f.v[1] = bot; }
        else { f.n = 0; return; }
    } else if (s.type == SHAPE_CYLINDER) {  // cylinder.ts:167-240: a side edge, or the cap octagon
        const T xz = d.x * d.x + d.z * d.z, y2 = d.y * d.y;
        if (xz > y2) {
            const T hl = sqrt(xz);
            const T ff = -s.radius / hl;
            const T vx = d.x * ff, vz = d.z * ff;
            f.n = 2;
            f.v[0] = mk3<T>(vx, s.hh, vz);
            f.v[1] = mk3<T>(vx, -s.hh, vz);
        } else {
            const T sx = d.y < T(0) ? s.radius : -s.radius;
            const T sy = d.y < T(0) ? s.hh : -s.hh;
            const T sz = s.radius;
            T rc = T(1), rs = T(0);
            if (xz > T(0.00765427) * y2) {  // more than 5 degrees from the axis: a vertex toward d
                const T hl = sqrt(xz);
                rc = d.x / hl;
                rs = d.z / hl;
            }
            const T c45 = T(0.707106769);
            const T ox[8] = {T(0), c45, T(1), c45, T(0), -c45, T(-1), -c45};
            const T oz[8] = {T(1), c45, T(0), -c45, T(-1), -c45, T(0), c45};
            f.n = 8;
            for (int i = 0; i < 8; i++) {
                const int t = insideOut ? 7 - i : i;
                const T px = sx * ox[t], py = sy * T(1), pz = sz * oz[t];
                f.v[i] = mk3<T>(px * rc - pz * rs, py, px * rs + pz * rc);
            }
        }
    } else {  // convex-hull.ts:479-548: normals through the inverse scale
        const V3<T> inv = mk3<T>(T(1) / sc.x, T(1) / sc.y, T(1) / sc.z);
        int f0 = s.hff;
        V3<T> n0 = mk3<T>(inv.x * st.pn[3 * f0], inv.y * st.pn[3 * f0 + 1], inv.z * st.pn[3 * f0 + 2]);
        T bestDot = dot(n0, d) / vlen(n0);
        int best = 0;
        for (int i = 1; i < s.hnf; i++) {
            V3<T> n = mk3<T>(inv.x * st.pn[3 * (f0 + i)], inv.y * st.pn[3 * (f0 + i) + 1], inv.z * st.pn[3 * (f0 + i) + 2]);
            T dt = dot(n, d) / vlen(n);
            if (dt < bestDot) { bestDot = dt; best = i; }
        }
        int first = st.ffv[f0 + best], nv = st.fnv[f0 + best];
        int delta = (nv + 16 - 1) / 16;
        f.n = 0;
        if (insideOut) {
            for (int i = nv - 1; i >= 0 && f.n < N; i -= delta) {
                int vi = st.fvi[first + i];
                const T* pp = st.pts + 3 * (s.hfp + vi);
                f.v[f.n++] = mk3<T>(pp[0], pp[1], pp[2]);
            }
        } else {
            for (int i = 0; i < nv && f.n < N; i += delta) {
                int vi = st.fvi[first + i];
                const T* pp = st.pts + 3 * (s.hfp + vi);
                f.v[f.n++] = mk3<T>(pp[0], pp[1], pp[2]);
            }
        }
    }
    face_transform(f, pos, q, sc);
}

// ---------------------------------------------------------------- clipping (clip.ts)
template <typename T, int N>
__device__ inline void clip_poly_vs_plane(Face<T, N>& out, const Face<T, N>& poly, V3<T> origin, V3<T> normal) {  // :13-92
    out.n = 0;
    if (poly.n < 2) return;
    V3<T> e1 = poly.v[poly.n - 1];
    T prevNum = dot(origin - e1, normal);
    bool prevIn = prevNum < T(0);
    for (int j = 0; j < poly.n; j++) {
        V3<T> e2 = poly.v[j];
        T num = dot(origin - e2, normal);
        bool curIn = num < T(0);
        if (curIn != prevIn) {
            V3<T> e12 = e2 - e1;
            T den = dot(e12, normal);
            if (den != T(0)) {
                T t = prevNum / den;
                if (out.n < N) out.v[out.n++] = e1 + e12 * t;
            } else curIn = prevIn;
        }
        if (curIn && out.n < N) out.v[out.n++] = e2;
        prevNum = num;
        prevIn = curIn;
        e1 = e2;
    }
}

template <typename T, int N>
__device__ inline void clip_poly_vs_poly(Face<T, N>& out, const Face<T, N>& toClip, const Face<T, N>& clip, V3<T> normal,
                                         Face<T, N>* tmp) {  // :108-175
    out.n = 0;
    if (toClip.n < 2 || clip.n < 3) return;
    int tmpIdx = 0;
    for (int i = 0; i < clip.n; i++) {
        V3<T> c1 = clip.v[i], c2 = clip.v[(i + 1) % clip.n];
        V3<T> cn = cross(normal, c2 - c1);
        const Face<T, N>& src = (i == 0) ? toClip : tmp[tmpIdx];
        tmpIdx ^= 1;
        Face<T, N>& dst = (i == clip.n - 1) ? out : tmp[tmpIdx];
        clip_poly_vs_plane(dst, src, c1, cn);
        if (dst.n < 3) { out.n = 0; break; }
    }
}

template <typename T, int N>
__device__ inline void clip_poly_vs_edge(Face<T, N>& out, const Face<T, N>& poly, V3<T> ev1, V3<T> ev2, V3<T> clipNormal) {  // :190-363
    out.n = 0;
    if (poly.n < 3) return;
    V3<T> en = cross(clipNormal, ev2 - ev1);
    V3<T> p0 = poly.v[0], p1 = poly.v[1], p2 = poly.v[2];
    V3<T> pn = cross(p2 - p0, p1 - p0);
    T pl2 = len2(pn);
    V3<T> v1 = ev1 + pn * (dot(p0 - ev1, pn) / pl2);
    V3<T> v2 = ev2 + pn * (dot(p0 - ev2, pn) / pl2);
    V3<T> v12 = v2 - v1;
    T v12l = len2(v12);
    V3<T> e1 = poly.v[poly.n - 1];
    T prevNum = dot(ev1 - e1, en);
    bool prevIn = prevNum < T(0);
    for (int j = 0; j < poly.n; j++) {
        V3<T> e2 = poly.v[j];
        T num = dot(ev1 - e2, en);
        bool curIn = num < T(0);
        if (curIn != prevIn) {
            V3<T> e12 = e2 - e1;
            T den = dot(e12, en);
            V3<T> cl = (den != T(0)) ? e1 + e12 * (prevNum / den) : e1;
            T pr = dot(cl - v1, v12);
            V3<T> o = pr < T(0) ? v1 : (pr > v12l ? v2 : cl);
            if (out.n < N) out.v[out.n++] = o;
        }
        prevNum = num;
        prevIn = curIn;
        e1 = e2;
    }
}

// ------------------------------------------------------------- manifold
template <typename T, int N = CC_MAX_FACE>
struct Manifold {
    V3<T> base, normal;
    int n;
    V3<T> relA[N], relB[N];
};

// manifoldBetweenTwoFaces manifold.ts:345-446
template <typename T, int N>
__device__ inline void manifold_between_two_faces(Manifold<T, N>& out, V3<T> p1, V3<T> p2, V3<T> axis, T maxDist,
                                                  const Face<T, N>& f1, const Face<T, N>& f2, Face<T, N>* scratch3) {
    out.n = 0;
    if (f1.n < 2 || f2.n < 3) {
        out.relA[0] = p1 - out.base; out.relB[0] = p2 - out.base; out.n = 1;
        return;
    }
    Face<T, N>& clipped = scratch3[2];
    clipped.n = 0;
    if (f1.n >= 3) clip_poly_vs_poly(clipped, f2, f1, axis, scratch3);
    else if (f1.n == 2) clip_poly_vs_edge(clipped, f2, f1.v[0], f1.v[1], axis);
    V3<T> origin = f1.v[0];
    V3<T> first = f1.v[1] - origin;
    V3<T> pnrm = (f1.n >= 3) ? cross(first, f1.v[2] - origin) : cross(cross(first, axis), first);
    T apn = dot(axis, pnrm);
    if (apn != T(0)) {
        T al = vlen(axis);
        for (int i = 0; i < clipped.n; i++) {
            V3<T> q2 = clipped.v[i];
            T dist = dot(q2 - origin, pnrm) / apn;
            if (dist * al < maxDist) {
                V3<T> q1 = q2 + axis * (-dist);
                out.relA[out.n] = q1 - out.base;
                out.relB[out.n] = q2 - out.base;
                out.n++;
                if (out.n >= N) break;
            }
        }
    }
    if (out.n == 0) { out.relA[0] = p1 - out.base; out.relB[0] = p2 - out.base; out.n = 1; }
}

// pruneContactPoints manifold.ts:192-313
template <typename T, int N>
__device__ inline void prune(Manifold<T, N>& m, V3<T> axis) {
    if (m.n <= 4) return;
    const T MIN_D2 = T(1.0e-6);
    T pd2[N];
    V3<T> proj[N];
    for (int i = 0; i < m.n; i++) {
        V3<T> v1 = m.relA[i], v2 = m.relB[i];
        T dp = dot(v1, axis);
        proj[i] = v1 + axis * (-dp);
        pd2[i] = tmax<T>(MIN_D2, len2(v2 - v1));
    }
    int p1 = 0;
    T best = tmax<T>(MIN_D2, len2(proj[0])) * pd2[0];
    for (int i = 1; i < m.n; i++) {
        T me = tmax<T>(MIN_D2, len2(proj[i])) * pd2[i];
        if (me > best) { best = me; p1 = i; }
    }
    int p2 = -1;
    best = -INFINITY;
    for (int i = 0; i < m.n; i++) {
        if (i == p1) continue;
        T me = tmax<T>(MIN_D2, len2(proj[i] - proj[p1])) * pd2[i];
        if (me > best) { best = me; p2 = i; }
    }
    int p3 = -1, p4 = -1;
    T minD = T(0), maxD = T(0);
    V3<T> perp = cross(proj[p2] - proj[p1], axis);
    for (int i = 0; i < m.n; i++) {
        if (i == p1 || i == p2) continue;
        T d = dot(proj[i] - proj[p1], perp);
        if (d < minD) { minD = d; p3 = i; }
        else if (d > maxD) { maxD = d; p4 = i; }
    }
    int order[4] = {p1, p3, p2, p4};
    V3<T> ra[4], rb[4];
    int n = 0;
    for (int k = 0; k < 4; k++)
        if (order[k] != -1) { ra[n] = m.relA[order[k]]; rb[n] = m.relB[order[k]]; n++; }
    for (int k = 0; k < n; k++) { m.relA[k] = ra[k]; m.relB[k] = rb[k]; }
    m.n = n;
}

// -------------------------------------------------------------- hits
template <typename T>
struct Hit {
    V3<T> pointA, pointB, axis;
    T penetration;
};

// The GJK half of collideConvexVsConvexLocal (convex.ts:359-437); returns the status
// and fills the local-space result (pd) or the simplex needed by EPA.
template <typename T>
struct ConvexCtx {
    Q4<T> rel;
    V3<T> relPos;
    T maxSepUse;
};

template <typename T>
__device__ inline void convex_setup(ConvexCtx<T>& c, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb, T maxSep) {
    Q4<T> ia = qconj(qa);
    c.rel.x = ia.w * qb.x + ia.x * qb.w + ia.y * qb.z - ia.z * qb.y;
    c.rel.y = ia.w * qb.y - ia.x * qb.z + ia.y * qb.w + ia.z * qb.x;
    c.rel.z = ia.w * qb.z + ia.x * qb.y - ia.y * qb.x + ia.z * qb.w;
    c.rel.w = ia.w * qb.w - ia.x * qb.x - ia.y * qb.y - ia.z * qb.z;
    c.relPos = qrot2(ia, pb - pa);
    c.maxSepUse = maxSep;
}

}  // namespace cc

// ---------------------------------------------------------------- planes (src/shapes/plane.ts)
// A plane shape: he = unit normal, radius = constant (points p with n.p + constant = 0; the
// solid side is n.p + constant < 0), hh = half extent of its bounded form.  The transforms
// below follow mathcat's mat4 (column-major) as collideConvexVsPlane uses it.
namespace cc {
template <typename T> struct M4 { T m[16]; };

template <typename T>
__device__ inline M4<T> m4_from_rt(Q4<T> q, V3<T> v) {  // mat4.fromRotationTranslation
    const T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    const T xx = q.x * x2, xy = q.x * y2, xz = q.x * z2, yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    const T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    M4<T> o;
    o.m[0] = T(1) - (yy + zz); o.m[1] = xy + wz; o.m[2] = xz - wy; o.m[3] = T(0);
    o.m[4] = xy - wz; o.m[5] = T(1) - (xx + zz); o.m[6] = yz + wx; o.m[7] = T(0);
    o.m[8] = xz + wy; o.m[9] = yz - wx; o.m[10] = T(1) - (xx + yy); o.m[11] = T(0);
    o.m[12] = v.x; o.m[13] = v.y; o.m[14] = v.z; o.m[15] = T(1);
    return o;
}

template <typename T>
__device__ inline M4<T> m4_invert(const M4<T>& A) {  // mat4.invert (cofactors over the determinant)
    const T* a = A.m;
    const T a00 = a[0], a01 = a[1], a02 = a[2], a03 = a[3], a10 = a[4], a11 = a[5], a12 = a[6], a13 = a[7];
    const T a20 = a[8], a21 = a[9], a22 = a[10], a23 = a[11], a30 = a[12], a31 = a[13], a32 = a[14], a33 = a[15];
    const T b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
    const T b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    const T b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
    const T b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    T det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
    M4<T> o;
    det = T(1) / det;
    o.m[0] = (a11 * b11 - a12 * b10 + a13 * b09) * det;
    o.m[1] = (a02 * b10 - a01 * b11 - a03 * b09) * det;
    o.m[2] = (a31 * b05 - a32 * b04 + a33 * b03) * det;
    o.m[3] = (a22 * b04 - a21 * b05 - a23 * b03) * det;
    o.m[4] = (a12 * b08 - a10 * b11 - a13 * b07) * det;
    o.m[5] = (a00 * b11 - a02 * b08 + a03 * b07) * det;
    o.m[6] = (a32 * b02 - a30 * b05 - a33 * b01) * det;
    o.m[7] = (a20 * b05 - a22 * b02 + a23 * b01) * det;
    o.m[8] = (a10 * b10 - a11 * b08 + a13 * b06) * det;
    o.m[9] = (a01 * b08 - a00 * b10 - a03 * b06) * det;
    o.m[10] = (a30 * b04 - a31 * b02 + a33 * b00) * det;
    o.m[11] = (a21 * b02 - a20 * b04 - a23 * b00) * det;
    o.m[12] = (a11 * b07 - a10 * b09 - a12 * b06) * det;
    o.m[13] = (a00 * b09 - a01 * b07 + a02 * b06) * det;
    o.m[14] = (a31 * b01 - a30 * b03 - a32 * b00) * det;
    o.m[15] = (a20 * b03 - a21 * b01 + a22 * b00) * det;
    return o;
}

template <typename T>
__device__ inline M4<T> m4_mul(const M4<T>& A, const M4<T>& B) {  // mat4.multiply: A * B
    M4<T> o;
    for (int j = 0; j < 4; j++) {
        const T b0 = B.m[4 * j], b1 = B.m[4 * j + 1], b2 = B.m[4 * j + 2], b3 = B.m[4 * j + 3];
        for (int i = 0; i < 4; i++) o.m[4 * j + i] = b0 * A.m[i] + b1 * A.m[4 + i] + b2 * A.m[8 + i] + b3 * A.m[12 + i];
    }
    return o;
}

template <typename T>
__device__ inline V3<T> m4_point(const M4<T>& M, V3<T> p) {  // vec3.transformMat4
    const T* m = M.m;
    T w = m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15];
    if (w == T(0)) w = T(1);
    return mk3<T>((m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12]) / w, (m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13]) / w,
                  (m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]) / w);
}

// transformPlane (plane.ts:196-208) of the unit-scale plane (scalePlane :184-194 renormalises it)
template <typename T>
__device__ inline void plane_in(const ShapeD<T>& P, const M4<T>& M, V3<T>& n, T& c) {
    const V3<T> n0 = mk3<T>(P.he[0], P.he[1], P.he[2]);
    const T l = vlen(n0);
    const V3<T> sn = n0 * (T(1) / l);
    const T sc = P.radius / l;
    const T* m = M.m;
    n = mk3<T>(m[0] * sn.x + m[4] * sn.y + m[8] * sn.z, m[1] * sn.x + m[5] * sn.y + m[9] * sn.z,
               m[2] * sn.x + m[6] * sn.y + m[10] * sn.z);
    c = sc - (m[12] * n.x + m[13] * n.y + m[14] * n.z);
}

template <typename T>
struct PlaneHit { V3<T> pA, pB, axis, nl; T pen; };

// collideConvexVsPlane (plane.ts:474-590): convex shape sc at (pc, qc) against plane shape sp at
// (pp, qp), with the convex shape's DEFAULT support (radius included for boxes and cylinders,
// excluded and added back for the others, support.ts:340-368); false when separated by more than
// maxSep.  Points and the axis (from the convex shape toward the plane) in world space; nl is
// the plane normal in the convex shape's frame.
template <typename T>
__device__ inline bool plane_hit(const ShapeTable<T>& st, int sc, V3<T> pc, Q4<T> qc, int sp, V3<T> pp, Q4<T> qp,
                                 T maxSep, PlaneHit<T>& h) {
    const M4<T> M = m4_mul(m4_invert(m4_from_rt(qc, pc)), m4_from_rt(qp, pp));
    V3<T> n;
    T c;
    plane_in(st.shapes[sp], M, n, c);
    const int ty = st.shapes[sc].type;
    Sup<T> S;
    sup_for_shape(S, st, sc, ty == SHAPE_BOX || ty == SHAPE_CYLINDER);
    const V3<T> s = sup_base(S, -n);
    const T dist = dot(n, s) + c;
    const T cr = S.cr;
    const T pen = -dist + cr;
    if (!(pen > -maxSep)) return false;
    const V3<T> p1 = s - mk3<T>(n.x * cr, n.y * cr, n.z * cr);
    const V3<T> p2 = s - mk3<T>(n.x * dist, n.y * dist, n.z * dist);
    const M4<T> TA = m4_from_rt(qc, pc);
    h.pA = m4_point(TA, p1);
    h.pB = m4_point(TA, p2);
    h.axis = qrotate(qc, -n);
    h.pen = pen;
    h.nl = n;
    return true;
}

// getAdaptivePlaneSupportingFace (plane.ts:631-725): a quad on the plane about the convex
// shape's position, its half size the diagonal of the convex shape's local bounds
template <typename T, int N>
__device__ inline void plane_face(Face<T, N>& f, const ShapeTable<T>& st, int sc, V3<T> pc, int sp, V3<T> pp, Q4<T> qp) {
    V3<T> n;
    T c;
    plane_in(st.shapes[sp], m4_from_rt(qp, pp), n, c);
    const T dd = dot(n, pc) + c;  // plane3.projectPoint
    const V3<T> ctr = pc - mk3<T>(n.x * dd, n.y * dd, n.z * dd);
    V3<T> p1 = cross(n, mk3<T>(0, 1, 0));  // getOrthogonalBasis (plane.ts:210-233)
    const T l2 = len2(p1);
    if (l2 < T(1e-6)) p1 = mk3<T>(1, 0, 0);
    else { const T k = T(1) / sqrt(l2); p1 = mk3<T>(p1.x * k, p1.y * k, p1.z * k); }
    V3<T> p2 = normalize(cross(p1, n));
    p1 = normalize(cross(n, p2));
    const ShapeD<T>& S = st.shapes[sc];
    const T size = vlen(mk3<T>(S.amax[0] - S.amin[0], S.amax[1] - S.amin[1], S.amax[2] - S.amin[2]));
    p1 = mk3<T>(p1.x * size, p1.y * size, p1.z * size);
    p2 = mk3<T>(p2.x * size, p2.y * size, p2.z * size);
    f.n = 4;
    f.v[0] = ctr + (p1 + p2);
    f.v[1] = ctr + (p1 - p2);
    f.v[2] = ctr - (p1 + p2);
    f.v[3] = ctr + (-p1 + p2);
}
}  // namespace cc
