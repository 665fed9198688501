// Convex shapes against static triangle meshes: collideConvexVsTriangleMesh
// (src/shapes/triangle-mesh.ts:1026-1379) with the reduction collector of the step
// (src/update.ts:333-563).  One thread per convex-vs-mesh pair walks the mesh BVH from the
// root (closer child first), collides every triangle that survives the bounds tests by GJK
// (EPA when the cores overlap) against the convex shape, corrects the normal at inactive
// edges, merges the hits into manifolds by normal and creates the pair's run of contact
// constraints.  The convex shape's support mode, the triangle support and the EPA polytope
// are the convex path's; the polytope workspace is the EPA kernel's, free again by then.
#pragma once
#include "cc_kernels.h"

namespace cc {

constexpr int MESH_MAX_MANIFOLDS = 32;  // the collector's manifold pool (update.ts:358-360)
constexpr int MESH_MAX_POINTS = 64;    // points a manifold collects before it is pruned (update.ts:500-518)
constexpr int MESH_STACK = 64;         // BVH traversal stack (deeper trees overflow)
constexpr int MESH_LANES = 64;

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
template <typename T>
__device__ inline void mesh_add_hit(const Dev<T>& d, MeshColl<T>& C, int& nm, V3<T> base, V3<T> pA, V3<T> pB, V3<T> axis,
                                    T pen, unsigned subA, unsigned subB, const Face<T, CC_MAX_FACE>& fA,
                                    const Face<T, CC_MAX_FACE>& fB, NpScratch<T, CC_MAX_FACE>& S) {
    const V3<T> normal = normalize(axis);
    int f = -1;
    for (int i = 0; i < nm; i++)
        if (dot(normal, C.m[i].normal) >= d.s.normalCos) { f = i; break; }
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
    Manifold<T, MESH_MAX_POINTS>& m = C.m[f];
    Manifold<T, CC_MAX_FACE>& t = S.m;
    t.base = base;
    t.n = 0;
    const T maxDist = d.s.spec + d.s.manifoldTol;
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

template <typename T>
__global__ void __launch_bounds__(64) k_np_mesh(Dev<T> d, T dt) {
    const int nmesh = min(d.counters[CNT_MESH], d.pairCap);
    const int worker = blockIdx.x * blockDim.x + threadIdx.x;
    if (worker >= d.meshWorkers) return;
    const int par = *d.parity;
    MeshColl<T>& C = ((MeshColl<T>*)d.meshWork)[worker];
    EpaH<T> H;
    H.b = d.epaWork + (size_t)(worker / EPA_LANES) * EpaWave<T>::bytes;
    H.lane = worker % EPA_LANES;
    NpScratch<T, CC_MAX_FACE> S;
    int stack[MESH_STACK];
    for (int j = worker; j < nmesh; j += d.meshWorkers) {
        const int slot = d.meshList[j];
        const int i0 = d.pairA[slot], i1 = d.pairB[slot];
        const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
        const PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
        const bool sw = pb.a != i0;
        const int A = sw ? i1 : i0, Bi = sw ? i0 : i1;
        const NpBody<T>& RA = sw ? r1 : r0;
        const NpBody<T>& RB = sw ? r0 : r1;
        // collideTriangleMeshVsConvex is the convex function with the shapes swapped and the
        // hits inverted (collide-shape-vs-shape.ts:228-330)
        const bool rev = (RA.tinfo & 15) == SHAPE_MESH;
        const NpBody<T>& RC = rev ? RB : RA;
        const NpBody<T>& RM = rev ? RA : RB;
        const V3<T> move = d.npd[A].v - d.npd[Bi].v;  // activeEdgeMovementDirection (update.ts:730-735)
        const ShapeD<T>& sc = d.st.shapes[RC.shape];
        const ShapeD<T>& sm = d.st.shapes[RM.shape];
        const V3<T> pa = RC.pos, pm = RM.pos;
        const Q4<T> qa = RC.q, qm = RM.q;
        const Q4<T> iqa = qconj(qa);
        const Q4<T> t21q = qmul(iqa, qm);
        const V3<T> t21p = qrotate(iqa, pm - pa);
        const Q4<T> iqm = qconj(qm);
        const Q4<T> qAinB = qmul(iqm, qa);
        const V3<T> pAinB = qrotate(iqm, pa - pm);
        const T maxSep0 = d.s.spec;
        const V3<T> e = mk3<T>(maxSep0, maxSep0, maxSep0);
        const V3<T> b1mn = mk3<T>(sc.amin[0], sc.amin[1], sc.amin[2]) - e;
        const V3<T> b1mx = mk3<T>(sc.amax[0], sc.amax[1], sc.amax[2]) + e;
        V3<T> q2mn, q2mx;
        {
            const M3<T> R = m3_from_quat(qAinB);
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
            q2mn = q2mn - e;
            q2mx = q2mx + e;
        }
        Sup<T> SA;
        sup_for_shape(SA, d.st, RC.shape, false);
        const int ntri = sm.hnp;
        const int nbits = ntri > 1 ? 32 - __clz(ntri - 1) : 0;  // pushIndex sub-shape.ts:60-64
        const unsigned mask = nbits >= 32 ? 0xffffffffu : ((1u << nbits) - 1u);
        const V3<T> qc = mk3<T>((q2mn.x + q2mx.x) * T(0.5), (q2mn.y + q2mx.y) * T(0.5), (q2mn.z + q2mx.z) * T(0.5));
        const V3<T> mdl = qrotate(iqa, move);
        const MNode<T>* nodes = d.st.mnode + sm.hff;
        const int4* tris = d.st.mtri + sm.hfp;
        int nm = 0;
        bool ok = true;
        int ns = 0;
        if (sm.hnf > 0) stack[ns++] = 0;
        while (ns > 0 && ok) {
            const int o = stack[--ns];
            const MNode<T> nd = nodes[o];
            if (!(nd.mn[0] <= q2mx.x && nd.mx[0] >= q2mn.x && nd.mn[1] <= q2mx.y && nd.mx[1] >= q2mn.y && nd.mn[2] <= q2mx.z &&
                  nd.mx[2] >= q2mn.z))
                continue;
            if (nd.b < 0) {
                const int t0 = nd.a, cnt = -(nd.b + 1);
                for (int i = 0; i < cnt && ok; i++) {
                    const int ti = t0 + i;
                    const int4 tr = tris[ti];
                    const T* P = d.st.mpos;
                    const V3<T> a = qrotate(t21q, mk3<T>(P[3 * tr.x], P[3 * tr.x + 1], P[3 * tr.x + 2])) + t21p;
                    const V3<T> b = qrotate(t21q, mk3<T>(P[3 * tr.y], P[3 * tr.y + 1], P[3 * tr.y + 2])) + t21p;
                    const V3<T> c = qrotate(t21q, mk3<T>(P[3 * tr.z], P[3 * tr.z + 1], P[3 * tr.z + 2])) + t21p;
                    const V3<T> tmn = mk3<T>(tmin<T>(tmin<T>(a.x, b.x), c.x), tmin<T>(tmin<T>(a.y, b.y), c.y),
                                             tmin<T>(tmin<T>(a.z, b.z), c.z));
                    const V3<T> tmx = mk3<T>(tmax<T>(tmax<T>(a.x, b.x), c.x), tmax<T>(tmax<T>(a.y, b.y), c.y),
                                             tmax<T>(tmax<T>(a.z, b.z), c.z));
                    if (tmx.x < b1mn.x || tmn.x > b1mx.x || tmx.y < b1mn.y || tmn.y > b1mx.y || tmx.z < b1mn.z ||
                        tmn.z > b1mx.z)
                        continue;
                    const V3<T> nrm = cross(b - a, c - a);
                    const bool back = dot(nrm, a) > T(0);
                    if (!d.s.backfaces && back) continue;
                    V3<T> axis = -nrm;
                    if (len2(axis) < T(1e-10)) axis = mk3<T>(1, 0, 0);
                    else axis = normalize(axis);
                    T maxSep = maxSep0;
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
                        sup_for_shape(SA2, d.st, RC.shape, true);
                        SA2.addRadius = true;
                        SA2.addR = maxSep;
                        if (!pd_step_epa(pd, SA2, TS, d.s.penTol, simplex, H)) continue;
                    }
                    const T pen = sqrt(len2(pd.pointA - pd.pointB)) - maxSep;
                    const T al = sqrt(len2(pd.axis));
                    if (al > T(0)) { T corr = maxSep / al; pd.pointA = pd.pointA + pd.axis * (-corr); }
                    if (d.s.activeEdges && tr.w != 7)
                        pd.axis = fix_normal(a, b, c, back ? nrm : -nrm, tr.w, pd.pointB, pd.axis, mdl);
                    const V3<T> wA = qrotate(qa, pd.pointA) + pa;
                    const V3<T> wB = qrotate(qa, pd.pointB) + pa;
                    const V3<T> wAxis = qrotate(qa, pd.axis);
                    const unsigned sub = (0xffffffffu & ~mask) | (unsigned)ti;
                    supporting_face(S.fA, d.st, RC.shape, -pd.axis, pa, qa);
                    S.fB.n = 3;
                    S.fB.v[0] = a; S.fB.v[1] = b; S.fB.v[2] = c;
                    face_transform(S.fB, pa, qa);
                    if (!rev) mesh_add_hit(d, C, nm, RA.pos, wA, wB, wAxis, pen, 0xffffffffu, sub, S.fA, S.fB, S);
                    else mesh_add_hit(d, C, nm, RA.pos, wB, wA, -wAxis, pen, sub, 0xffffffffu, S.fB, S.fA, S);
                }
            } else {
                const int l = o + 1, r = nd.a;
                const MNode<T> ln = nodes[l], rn = nodes[r];
                const V3<T> lc = mk3<T>((ln.mn[0] + ln.mx[0]) * T(0.5), (ln.mn[1] + ln.mx[1]) * T(0.5), (ln.mn[2] + ln.mx[2]) * T(0.5));
                const V3<T> rc = mk3<T>((rn.mn[0] + rn.mx[0]) * T(0.5), (rn.mn[1] + rn.mx[1]) * T(0.5), (rn.mn[2] + rn.mx[2]) * T(0.5));
                const T ld = len2(lc - qc), rd = len2(rc - qc);
                if (ns + 2 > MESH_STACK) { ok = false; break; }
                if (ld <= rd) { stack[ns++] = r; stack[ns++] = l; }
                else { stack[ns++] = l; stack[ns++] = r; }
            }
        }
        if (!ok) { atomicOr(&d.counters[CNT_OVERFLOW], 32); continue; }  // BVH deeper than the stack
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
        const bool dynA = fl_motion(RA.flags) == MOTION_DYNAMIC, dynB = fl_motion(RB.flags) == MOTION_DYNAMIC;
        if (!((dynA && RA.invMass != T(0)) || (dynB && RB.invMass != T(0)))) continue;
        const int cs0 = atomicAdd(&d.counters[CNT_CONS], run);
        if (cs0 + run > d.conCap) { atomicOr(&d.counters[CNT_OVERFLOW], 2); continue; }
        int k = 0;
        for (int i = 0; i < nm; i++) {
            if (C.m[i].n <= 0) continue;
            make_contact(d, slot, A, Bi, RA, RB, C.m[i], C.subA[i], C.subB[i], cs0 + k, k, run, dt, par);
            k++;
        }
    }
}

}  // namespace cc
