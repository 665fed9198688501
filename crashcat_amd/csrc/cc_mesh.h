// Convex shapes against static triangle meshes: collideConvexVsTriangleMesh
// (src/shapes/triangle-mesh.ts:1026-1379) with the reduction collector of the step
// (src/update.ts:333-563), in four passes: each convex-vs-mesh pair walks the mesh BVH from
// the root (closer child first) and counts the triangles that survive the bounds and
// back-face tests; a scan gives each pair a range, a second walk lists them in walk order;
// every listed triangle is collided by GJK (EPA when the cores overlap) on its own lane, its
// normal corrected at inactive edges; then one thread per pair merges its hits into
// manifolds in walk order and creates the pair's run of contact constraints.  The convex
// shape's support modes, the triangle support and the EPA polytope are the convex path's; the
// polytope workspace is the EPA kernel's, free again by then.  A sphere takes
// collideSphereVsTriangleMesh (triangle-mesh.ts:1418-1720) through the same passes: the walk is
// queried from the sphere's centre in the mesh's frame, and a candidate's hit is the closest
// point of its triangle.
#pragma once
#include "cc_kernels.h"

namespace cc {

constexpr int MESH_MAX_MANIFOLDS = 32;  // the collector's manifold pool (update.ts:358-360)
constexpr int MESH_MAX_POINTS = 64;    // points a manifold collects before it is pruned (update.ts:500-518)
constexpr int MESH_STACK = 64;         // BVH traversal stack (deeper trees overflow)
constexpr int MESH_LANES = 64;

template <typename T>
struct MeshHit {  // a candidate triangle's GJK/EPA result in the convex shape's frame
    int hit;
    T pen;
    V3<T> pA, pB, axis;
};

template <typename T>
struct MeshColl {
    Manifold<T, MESH_MAX_POINTS> m[MESH_MAX_MANIFOLDS];
    T depth[MESH_MAX_MANIFOLDS];
    unsigned subA[MESH_MAX_MANIFOLDS], subB[MESH_MAX_MANIFOLDS];
};

// fixNormal src/collision/active-edges.ts:74-153
template <typename T>
__device__ inline V3<T> fix_normal(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> tn, int act, V3<T> p, V3<T> n, V3<T> md) {
    if (act == 7) return n;
    const T nl = vlen(n), tl = vlen(tn);
    if (dot(md, n) * tl < dot(md, tn) * nl) return n;
    if (act == 0) return tn;
    if (dot(tn, n) > T(0.999848) * nl * tl) return n;
    const Bary<T> b = bary3(v0 - p, v1 - p, v2 - p, T(1e-10));
    int edge;
    if (b.u > T(1) - T(1.0e-4)) edge = 5;
    else if (b.v > T(1) - T(1.0e-4)) edge = 3;
    else if (b.w > T(1) - T(1.0e-4)) edge = 6;
    else if (b.u < T(1.0e-4)) edge = 2;
    else if (b.v < T(1.0e-4)) edge = 4;
    else if (b.w < T(1.0e-4)) edge = 1;
    else return tn;
    return (act & edge) ? n : tn;
}

// addHit of the reduction collector (src/update.ts:385-520)
// (collector of arithmetic type K: T, or FP64 for the FP64 narrowphase of an FP32 context)
template <typename T, typename K>
__device__ inline void mesh_add_hit(const Dev<T>& d, MeshColl<K>& C, int& nm, V3<K> base, V3<K> pA, V3<K> pB, V3<K> axis,
                                    K pen, unsigned subA, unsigned subB, const Face<K, CC_MAX_FACE>& fA,
                                    const Face<K, CC_MAX_FACE>& fB, NpScratch<K, CC_MAX_FACE>& S) {
    const NpTol<K> tol = np_tol<T, K>(d);
    const V3<K> normal = normalize(axis);
    int f = -1;
    for (int i = 0; i < nm; i++)
        if (dot(normal, C.m[i].normal) >= tol.normalCos) { f = i; break; }
    if (f < 0) {
        if (nm >= MESH_MAX_MANIFOLDS) {  // pool full: a deeper hit replaces the shallowest manifold
            int sh = 0;
            for (int i = 1; i < nm; i++)
                if (C.depth[i] < C.depth[sh]) sh = i;
            if (pen < C.depth[sh]) return;
            f = sh;
        } else {
            f = nm++;
        }
        C.m[f].n = 0;
        C.m[f].base = base;
        C.m[f].normal = normal;
        C.depth[f] = pen;
        C.subA[f] = subA;
        C.subB[f] = subB;
    } else {
        C.m[f].normal = C.m[f].normal + normal;
    }
    Manifold<K, MESH_MAX_POINTS>& m = C.m[f];
    Manifold<K, CC_MAX_FACE>& t = S.m;
    t.base = base;
    t.n = 0;
    const K maxDist = tol.spec + tol.manifoldTol;
    if (fA.n >= 2 && fB.n >= 3) manifold_between_two_faces(t, pA, pB, axis, maxDist, fA, fB, S.tmp);
    else { t.relA[0] = pA - base; t.relB[0] = pB - base; t.n = 1; }
    for (int i = 0; i < t.n && m.n < MESH_MAX_POINTS; i++) {
        m.relA[m.n] = t.relA[i];
        m.relB[m.n] = t.relB[i];
        m.n++;
    }
    if (pen > C.depth[f]) C.depth[f] = pen;
    if (m.n > 32) prune(m, m.normal);
}

// One convex-vs-mesh pair, as collideTriangleMeshVsConvex / collideConvexVsTriangleMesh set it up
// (triangle-mesh.ts:1058-1131): the convex shape is A of the mesh function; `rev` when the mesh
// is the narrowphase's body A (hits are then inverted, collide-shape-vs-shape.ts:228-330).
// Triangles are read in the convex shape's frame (t21q, t21p); for a sphere (`sphere`,
// triangle-mesh.ts:1448-1500) in the mesh's frame relative to the sphere's centre sc, with the
// walk's query box sc -+ (radius + max separation).  Hits are stored in the frame (hq, hp).
template <typename T>
struct MeshPair {
    int slot, A, Bi;
    bool rev, sphere;
    NpBody<T> RA, RB;
    int cshape;
    V3<T> pa;
    Q4<T> qa;
    Q4<T> t21q;
    V3<T> t21p, b1mn, b1mx, q2mn, q2mx, qc, mdl;
    V3<T> sc;
    T rad;
    Q4<T> hq;
    V3<T> hp;
    int mfn, mnn, mft;
    unsigned mask;
};

template <typename T>
__device__ inline void mesh_pair(const Dev<T>& d, int slot, MeshPair<T>& P) {
    P.slot = slot;
    const int i0 = d.pairA[slot], i1 = d.pairB[slot];
    const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
    const PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
    const bool sw = pb.a != i0;
    P.A = sw ? i1 : i0;
    P.Bi = sw ? i0 : i1;
    P.RA = sw ? r1 : r0;
    P.RB = sw ? r0 : r1;
    P.rev = (P.RA.tinfo & 15) == SHAPE_MESH;
    const NpBody<T>& RC = P.rev ? P.RB : P.RA;
    const NpBody<T>& RM = P.rev ? P.RA : P.RB;
    const V3<T> move = d.npd[P.A].v - d.npd[P.Bi].v;  // activeEdgeMovementDirection (update.ts:730-735)
    const ShapeD<T>& sc = d.st.shapes[RC.shape];
    const ShapeD<T>& sm = d.st.shapes[RM.shape];
    P.cshape = RC.shape;
    P.pa = RC.pos;
    P.qa = RC.q;
    const V3<T> pm = RM.pos;
    const Q4<T> qm = RM.q;
    P.mfn = sm.hff;
    P.mnn = sm.hnf;
    P.mft = sm.hfp;
    const int ntri = sm.hnp;
    const int nbits = ntri > 1 ? 32 - __clz(ntri - 1) : 0;  // pushIndex sub-shape.ts:60-64
    P.mask = nbits >= 32 ? 0xffffffffu : ((1u << nbits) - 1u);
    P.sphere = sc.type == SHAPE_SPHERE;
    if (P.sphere) {  // triangle-mesh.ts:1452-1480
        const Q4<T> iqm = qconj(qm);
        P.rad = sc.radius;
        P.sc = qrotate(iqm, P.pa - pm);
        const T er = P.rad + d.s.spec;
        P.q2mn = mk3<T>(P.sc.x - er, P.sc.y - er, P.sc.z - er);
        P.q2mx = mk3<T>(P.sc.x + er, P.sc.y + er, P.sc.z + er);
        P.qc = P.sc;
        P.mdl = qrotate(iqm, move);  // triangle-mesh.ts:1607-1611
        P.hq = qm;
        P.hp = pm;
        return;
    }
    P.hq = P.qa;
    P.hp = P.pa;
    const Q4<T> iqa = qconj(P.qa);
    P.t21q = qmul(iqa, qm);
    P.t21p = qrotate(iqa, pm - P.pa);
    const Q4<T> iqm = qconj(qm);
    const Q4<T> qAinB = qmul(iqm, P.qa);
    const V3<T> pAinB = qrotate(iqm, P.pa - pm);
    const T ms = d.s.spec;
    const V3<T> e = mk3<T>(ms, ms, ms);
    P.b1mn = mk3<T>(sc.amin[0], sc.amin[1], sc.amin[2]) - e;
    P.b1mx = mk3<T>(sc.amax[0], sc.amax[1], sc.amax[2]) + e;
    const M3<T> R = m3_from_quat(qAinB);
    V3<T> q2mn, q2mx;
    for (int c = 0; c < 8; c++) {  // box3.transformMat4 of the local bounds
        const V3<T> p = mk3<T>((c & 1) ? sc.amax[0] : sc.amin[0], (c & 2) ? sc.amax[1] : sc.amin[1],
                               (c & 4) ? sc.amax[2] : sc.amin[2]);
        const V3<T> wp = mul(R, p) + pAinB;
        if (c == 0) { q2mn = wp; q2mx = wp; }
        else {
            q2mn = mk3<T>(tmin<T>(q2mn.x, wp.x), tmin<T>(q2mn.y, wp.y), tmin<T>(q2mn.z, wp.z));
            q2mx = mk3<T>(tmax<T>(q2mx.x, wp.x), tmax<T>(q2mx.y, wp.y), tmax<T>(q2mx.z, wp.z));
        }
    }
    P.q2mn = q2mn - e;
    P.q2mx = q2mx + e;
    P.qc = mk3<T>((P.q2mn.x + P.q2mx.x) * T(0.5), (P.q2mn.y + P.q2mx.y) * T(0.5), (P.q2mn.z + P.q2mx.z) * T(0.5));
    P.mdl = qrotate(iqa, move);
}

// a triangle of the pair's mesh in the convex shape's frame
template <typename T>
__device__ inline void mesh_tri(const Dev<T>& d, const MeshPair<T>& P, int ti, V3<T>& a, V3<T>& b, V3<T>& c, int& act) {
    const int4 tr = d.st.mtri[P.mft + ti];
    const T* X = d.st.mpos;
    act = tr.w;
    if (P.sphere) {  // rv0..rv2 (triangle-mesh.ts:1540-1545)
        a = mk3<T>(X[3 * tr.x], X[3 * tr.x + 1], X[3 * tr.x + 2]) - P.sc;
        b = mk3<T>(X[3 * tr.y], X[3 * tr.y + 1], X[3 * tr.y + 2]) - P.sc;
        c = mk3<T>(X[3 * tr.z], X[3 * tr.z + 1], X[3 * tr.z + 2]) - P.sc;
        return;
    }
    a = qrotate(P.t21q, mk3<T>(X[3 * tr.x], X[3 * tr.x + 1], X[3 * tr.x + 2])) + P.t21p;
    b = qrotate(P.t21q, mk3<T>(X[3 * tr.y], X[3 * tr.y + 1], X[3 * tr.y + 2])) + P.t21p;
    c = qrotate(P.t21q, mk3<T>(X[3 * tr.z], X[3 * tr.z + 1], X[3 * tr.z + 2])) + P.t21p;
}

// getClosestPointOnTriangle (src/collision/triangle.ts:62-198, degenerate triangles :199-286):
// the point of triangle abc closest to the origin, its squared distance, and the feature it
// lies on (bits of the vertices a, b, c that span it; 7 the interior).
template <typename T>
__device__ inline int closest_on_triangle(V3<T> a, V3<T> b, V3<T> c, V3<T>& p, T& dist2) {
    const V3<T> ab = b - a, ac = c - a;
    const V3<T> n = cross(ab, ac);
    const T nl = len2(n);
    if (nl < T(1e-10)) {
        int f = 4;
        p = c;
        T best = len2(c);
        const T al = len2(a);
        if (al < best) { f = 1; p = a; best = al; }
        const T bl = len2(b);
        if (bl < best) { f = 2; p = b; best = bl; }
        const T eps2 = T(1.1920929e-7 * 1.1920929e-7);
        const T acl = len2(ac);
        if (acl > eps2) {
            const T v = tmax<T>(T(0), tmin<T>(T(1), -dot(a, ac) / acl));
            const V3<T> q = mk3<T>(a.x + v * ac.x, a.y + v * ac.y, a.z + v * ac.z);
            const T dq = len2(q);
            if (dq < best) { f = 5; p = q; best = dq; }
        }
        const V3<T> bc = c - b;
        const T bcl = len2(bc);
        if (bcl > eps2) {
            const T v = tmax<T>(T(0), tmin<T>(T(1), -dot(b, bc) / bcl));
            const V3<T> q = mk3<T>(b.x + v * bc.x, b.y + v * bc.y, b.z + v * bc.z);
            const T dq = len2(q);
            if (dq < best) { f = 6; p = q; best = dq; }
        }
        const T abl = len2(ab);
        if (abl > eps2) {
            const T v = tmax<T>(T(0), tmin<T>(T(1), -dot(a, ab) / abl));
            const V3<T> q = mk3<T>(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
            const T dq = len2(q);
            if (dq < best) { f = 3; p = q; best = dq; }
        }
        dist2 = best;
        return f;
    }
    const V3<T> ap = -a;
    const T d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= T(0) && d2 <= T(0)) { p = a; dist2 = len2(a); return 1; }
    const V3<T> bp = -b;
    const T d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= T(0) && d4 <= d3) { p = b; dist2 = len2(b); return 2; }
    if (d1 * d4 <= d3 * d2 && d1 >= T(0) && d3 <= T(0)) {
        const T v = d1 / (d1 - d3);
        p = mk3<T>(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
        dist2 = len2(p);
        return 3;
    }
    const V3<T> cp = -c;
    const T d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= T(0) && d5 <= d6) { p = c; dist2 = len2(c); return 4; }
    if (d5 * d2 <= d1 * d6 && d2 >= T(0) && d6 <= T(0)) {
        const T w = d2 / (d2 - d6);
        p = mk3<T>(a.x + w * ac.x, a.y + w * ac.y, a.z + w * ac.z);
        dist2 = len2(p);
        return 5;
    }
    const T d43 = d4 - d3, d56 = d5 - d6;
    if (d3 * d6 <= d5 * d4 && d43 >= T(0) && d56 >= T(0)) {
        const T w = d43 / (d43 + d56);
        const V3<T> bc = c - b;
        p = mk3<T>(b.x + w * bc.x, b.y + w * bc.y, b.z + w * bc.z);
        dist2 = len2(p);
        return 6;
    }
    const T cd = (a.x + b.x + c.x) * n.x + (a.y + b.y + c.y) * n.y + (a.z + b.z + c.z) * n.z;
    const T sc = cd / (T(3) * nl);
    p = mk3<T>(n.x * sc, n.y * sc, n.z * sc);
    dist2 = len2(p);
    return 7;
}

// FEATURE_TO_ACTIVE_EDGES (src/collision/sphere-triangle.ts:218-227), 4 bits per feature: the
// edges, one of which must be active for a hit on that feature to keep its own normal
__device__ inline int feature_edges(int f) { return (0x02461350 >> (4 * f)) & 7; }

// One candidate of a sphere pair (triangle-mesh.ts:1546-1640): the hit in the mesh's frame, or false.
template <typename T>
__device__ inline bool sphere_tri_hit(const Dev<T>& d, const MeshPair<T>& P, V3<T> a, V3<T> b, V3<T> c, int act,
                                      MeshHit<T>& out) {
    const V3<T> nrm = cross(b - a, c - a);
    const bool back = dot(nrm, a) > T(0);
    V3<T> p2;
    T dd;
    const int f = closest_on_triangle(a, b, c, p2, dd);
    const T ms = P.rad + d.s.spec;
    if (dd > ms * ms) return false;
    const T pen = P.rad - sqrt(dd);
    V3<T> axis = dd > T(0) ? normalize(p2) : mk3<T>(0, 1, 0);
    const V3<T> p1 = mk3<T>(axis.x * P.rad, axis.y * P.rad, axis.z * P.rad);
    if (d.s.activeEdges && f != 7 && act != 7 && (act & feature_edges(f)) == 0) {
        const V3<T> na = back ? nrm : -nrm;
        if (dot(P.mdl, axis) * vlen(na) >= dot(P.mdl, na)) axis = na;
    }
    out.pen = pen;
    out.pA = P.sc + p1;
    out.pB = P.sc + p2;
    out.axis = axis;
    return true;
}

// The BVH walk (closer child first) calling f(triangle) for each triangle that passes the
// bounds and back-face tests, in the reference's order; false if the stack overflowed.
template <typename T, typename F>
__device__ inline bool mesh_walk(const Dev<T>& d, const MeshPair<T>& P, F&& f) {
    int stack[MESH_STACK];
    int ns = 0;
    const MNode<T>* nodes = d.st.mnode + P.mfn;
    if (P.mnn > 0) stack[ns++] = 0;
    while (ns > 0) {
        const int o = stack[--ns];
        const MNode<T> nd = nodes[o];
        if (!(nd.mn[0] <= P.q2mx.x && nd.mx[0] >= P.q2mn.x && nd.mn[1] <= P.q2mx.y && nd.mx[1] >= P.q2mn.y &&
              nd.mn[2] <= P.q2mx.z && nd.mx[2] >= P.q2mn.z))
            continue;
        if (nd.b < 0) {
            const int t0 = nd.a, cnt = -(nd.b + 1);
            for (int i = 0; i < cnt; i++) {
                V3<T> a, b, c;
                int act;
                mesh_tri(d, P, t0 + i, a, b, c, act);
                const V3<T> tmn = mk3<T>(tmin<T>(tmin<T>(a.x, b.x), c.x), tmin<T>(tmin<T>(a.y, b.y), c.y),
                                         tmin<T>(tmin<T>(a.z, b.z), c.z));
                const V3<T> tmx = mk3<T>(tmax<T>(tmax<T>(a.x, b.x), c.x), tmax<T>(tmax<T>(a.y, b.y), c.y),
                                         tmax<T>(tmax<T>(a.z, b.z), c.z));
                if (!P.sphere && (tmx.x < P.b1mn.x || tmn.x > P.b1mx.x || tmx.y < P.b1mn.y || tmn.y > P.b1mx.y || tmx.z < P.b1mn.z ||
                    tmn.z > P.b1mx.z))
                    continue;
                const V3<T> nrm = cross(b - a, c - a);
                if (!d.s.backfaces && dot(nrm, a) > T(0)) continue;
                f(t0 + i);
            }
        } else {
            const int l = o + 1, r = nd.a;
            const MNode<T> ln = nodes[l], rn = nodes[r];
            const V3<T> lc = mk3<T>((ln.mn[0] + ln.mx[0]) * T(0.5), (ln.mn[1] + ln.mx[1]) * T(0.5), (ln.mn[2] + ln.mx[2]) * T(0.5));
            const V3<T> rc = mk3<T>((rn.mn[0] + rn.mx[0]) * T(0.5), (rn.mn[1] + rn.mx[1]) * T(0.5), (rn.mn[2] + rn.mx[2]) * T(0.5));
            const T ld = len2(lc - P.qc), rd = len2(rc - P.qc);
            if (ns + 2 > MESH_STACK) return false;
            if (ld <= rd) { stack[ns++] = r; stack[ns++] = l; }
            else { stack[ns++] = l; stack[ns++] = r; }
        }
    }
    return true;
}

// 1. candidates per pair (a count, then the scan gives each pair its range)
template <typename T>
__global__ void __launch_bounds__(256) k_mesh_count(Dev<T> d) {
    const int nmesh = min(d.counters[CNT_MESH], d.pairCap);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < d.pairCap; j += gridDim.x * blockDim.x) {
        int n = 0;
        if (j < nmesh) {
            MeshPair<T> P;
            mesh_pair(d, d.meshList[j], P);
            if (!mesh_walk(d, P, [&](int) { n++; })) { atomicOr(&d.counters[CNT_OVERFLOW], 32); n = 0; }
        }
        d.meshCnt[j] = n;
    }
}

// 2. the candidate triangles of each pair, in walk order
template <typename T>
__global__ void __launch_bounds__(256) k_mesh_emit(Dev<T> d) {
    const int nmesh = min(d.counters[CNT_MESH], d.pairCap);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nmesh; j += gridDim.x * blockDim.x) {
        const int n = d.meshCnt[j];
        if (n == 0) continue;
        const int o = d.meshOff[j];
        if (o + n > d.candCap) { atomicOr(&d.counters[CNT_OVERFLOW], 64); continue; }
        MeshPair<T> P;
        mesh_pair(d, d.meshList[j], P);
        int k = 0;
        mesh_walk(d, P, [&](int ti) {
            if (k < n) { d.candTri[o + k] = ti; d.candPair[o + k] = j; }
            k++;
        });
    }
}

// 3. GJK (EPA when the cores overlap) of every candidate, one lane each (triangle-mesh.ts:1190-1297)
template <typename T>
__global__ void __launch_bounds__(64) k_mesh_tri(Dev<T> d) {
    const int nmesh = min(d.counters[CNT_MESH], d.pairCap);
    const int nc = min(d.meshOff[d.pairCap], d.candCap);
    const int worker = blockIdx.x * blockDim.x + threadIdx.x;
    if (worker >= d.meshWorkers || nmesh == 0) return;
    EpaH<T> H;
    H.b = d.epaWork + (size_t)(worker / EPA_LANES) * EpaWave<T>::bytes;
    H.lane = worker % EPA_LANES;
    for (int ci = worker; ci < nc; ci += d.meshWorkers) {
        const int j = d.candPair[ci], ti = d.candTri[ci];
        MeshPair<T> P;
        mesh_pair(d, d.meshList[j], P);
        V3<T> a, b, c;
        int act;
        mesh_tri(d, P, ti, a, b, c, act);
        MeshHit<T>& out = ((MeshHit<T>*)d.candHit)[ci];
        out.hit = 0;
        if (P.sphere) {
            if (sphere_tri_hit(d, P, a, b, c, act, out)) out.hit = 1;
            continue;
        }
        const V3<T> nrm = cross(b - a, c - a);
        const bool back = dot(nrm, a) > T(0);
        V3<T> axis = -nrm;
        if (len2(axis) < T(1e-10)) axis = mk3<T>(1, 0, 0);
        else axis = normalize(axis);
        T maxSep = d.s.spec;
        Sup<T> SA;
        sup_for_shape(SA, d.st, P.cshape, false);
        TriSup<T> TS;
        TS.a = a; TS.b = b; TS.c = c;
        PenDepth<T> pd;
        Simplex<T> simplex;
        simplex.size = 0;
        pd_step_gjk(pd, simplex, SA, TS, SA.cr + maxSep, T(0), axis, d.s.colTol);
        if (pd.status == PD_NOT_COLLIDING) continue;
        if (pd.status == PD_INDETERMINATE) {
            maxSep = tmin<T>(maxSep, T(1));
            Sup<T> SA2;
            sup_for_shape(SA2, d.st, P.cshape, true);
            SA2.addRadius = true;
            SA2.addR = maxSep;
            if (!pd_step_epa(pd, SA2, TS, d.s.penTol, simplex, H)) continue;
        }
        const T pen = sqrt(len2(pd.pointA - pd.pointB)) - maxSep;
        const T al = sqrt(len2(pd.axis));
        if (al > T(0)) { T corr = maxSep / al; pd.pointA = pd.pointA + pd.axis * (-corr); }
        if (d.s.activeEdges && act != 7) pd.axis = fix_normal(a, b, c, back ? nrm : -nrm, act, pd.pointB, pd.axis, P.mdl);
        out.hit = 1;
        out.pen = pen;
        out.pA = pd.pointA;
        out.pB = pd.pointB;
        out.axis = pd.axis;
    }
}

// 4. the reduction collector over each pair's hits in walk order, then its run of constraints
template <typename T>
__global__ void __launch_bounds__(64) k_mesh_collect(Dev<T> d, T dt) {
    const int nmesh = min(d.counters[CNT_MESH], d.pairCap);
    const int worker = blockIdx.x * blockDim.x + threadIdx.x;
    if (worker >= d.meshWorkers) return;
    const int par = *d.parity;
    MeshColl<T>& C = ((MeshColl<T>*)d.meshWork)[worker];
    NpScratch<T, CC_MAX_FACE> S;
    for (int j = worker; j < nmesh; j += d.meshWorkers) {
        const int n = d.meshCnt[j];
        const int o = d.meshOff[j];
        if (n == 0 || o + n > d.candCap) continue;
        MeshPair<T> P;
        mesh_pair(d, d.meshList[j], P);
        int nm = 0;
        for (int k = 0; k < n; k++) {
            const MeshHit<T> h = ((const MeshHit<T>*)d.candHit)[o + k];
            if (!h.hit) continue;
            const int ti = d.candTri[o + k];
            V3<T> a, b, c;
            int act;
            mesh_tri(d, P, ti, a, b, c, act);
            const V3<T> wA = qrotate(P.hq, h.pA) + P.hp;
            const V3<T> wB = qrotate(P.hq, h.pB) + P.hp;
            const V3<T> wAxis = qrotate(P.hq, h.axis);
            const unsigned sub = (0xffffffffu & ~P.mask) | (unsigned)ti;
            if (P.sphere) {  // a sphere has no supporting face: the collector keeps the single point
                S.fA.n = 0;
                S.fB.n = 0;
            } else {
                supporting_face(S.fA, d.st, P.cshape, -h.axis, P.pa, P.qa);
                S.fB.n = 3;
                S.fB.v[0] = a; S.fB.v[1] = b; S.fB.v[2] = c;
                face_transform(S.fB, P.pa, P.qa, mk3<T>(1, 1, 1));
            }
            if (!P.rev) mesh_add_hit(d, C, nm, P.RA.pos, wA, wB, wAxis, h.pen, 0xffffffffu, sub, S.fA, S.fB, S);
            else mesh_add_hit(d, C, nm, P.RA.pos, wB, wA, -wAxis, h.pen, sub, 0xffffffffu, S.fB, S.fA, S);
        }
        // finalizeAndCreateConstraints (update.ts:522-555): the pair's manifolds become a run of
        // consecutive constraint slots
        int run = 0;
        for (int i = 0; i < nm; i++) {
            Manifold<T, MESH_MAX_POINTS>& m = C.m[i];
            m.normal = normalize(m.normal);
            if (m.n > 4) prune(m, m.normal);
            if (m.n > 0) run++;
        }
        if (run == 0) continue;
        const bool dynA = fl_motion(P.RA.flags) == MOTION_DYNAMIC, dynB = fl_motion(P.RB.flags) == MOTION_DYNAMIC;
        if (!((dynA && P.RA.invMass != T(0)) || (dynB && P.RB.invMass != T(0)))) continue;
        const int cs0 = atomicAdd(&d.counters[CNT_CONS], run);
        if (cs0 + run > d.conCap) { atomicOr(&d.counters[CNT_OVERFLOW], 2); continue; }
        int k = 0;
        for (int i = 0; i < nm; i++) {
            if (C.m[i].n <= 0) continue;
            make_contact(d, P.slot, P.A, P.Bi, P.RA, P.RB, C.m[i], C.subA[i], C.subB[i], cs0 + k, k, run, dt, par);
            k++;
        }
    }
}

}  // namespace cc
