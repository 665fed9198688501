// Compound shapes (src/shapes/compound.ts:494-640): a pair with a compound on either side
// collides every child of body A's compound (or A's own shape) with every child of B's (or B's
// own shape), A's children outer, in index order; a child is a convex leaf whose local transform
// is its position and quaternion in the compound, composed with the body pose as the compound
// does (leaf_pose).  Each leaf pair's hit -- the sphere pair's closed form, or GJK, EPA when the
// cores overlap, and both supporting faces -- goes to the step's reduction collector with the
// child's sub-shape id pushed by pushIndex(child, count) (src/body/sub-shape.ts:43-64); the
// collector's manifolds become the pair's run of constraints, as for meshes.  One thread per
// pair owns a collector and an EPA polytope.
#pragma once
#include "cc_mesh.h"

namespace cc {

__device__ inline unsigned sub_push(unsigned value, int bit, int id, int nbits) {  // push sub-shape.ts:43-57
    const unsigned mask = nbits ? (((1u << nbits) - 1u) << bit) : 0u;
    return (value & ~mask) | ((unsigned)id << bit);
}
__device__ inline int sub_bits(int count) { return count > 1 ? 32 - __clz((unsigned)(count - 1)) : 0; }

// collideShapeVsShape of two leaves at their bodies' poses (composed with the leaves' local
// transforms), the hit fed to the collector: collideSphereVsSphere (sphere.ts:319-410) or
// collideConvexVsConvex (convex.ts:257-641) with both supporting faces
template <typename T, typename C>
__device__ inline void leaf_pair_hit(const Dev<T>& d, const ShapeTable<C>& st, const NpTol<C>& tol, int sa,
                                     const NpBody<T>& RA, int sb, const NpBody<T>& RB, unsigned subA, unsigned subB,
                                     EpaH<C>& H, NpScratch<C, CC_MAX_FACE>& S, MeshColl<C>& Col, int& nm) {
    const ShapeD<C>& A = st.shapes[sa];
    const ShapeD<C>& B = st.shapes[sb];
    V3<C> pa = cvt<C>(RA.pos), pb = cvt<C>(RB.pos);
    Q4<C> qa = cvtq<C>(RA.q), qb = cvtq<C>(RB.q);
    leaf_pose(A, pa, qa);
    leaf_pose(B, pb, qb);
    const V3<C> base = cvt<C>(RA.pos);
    if (A.type == SHAPE_SPHERE && B.type == SHAPE_SPHERE) {
        const C ra = A.radius * fabs(shape_scale(A).x), rb = B.radius * fabs(shape_scale(B).x);
        const V3<C> dd = pb - pa;
        const C d2 = len2(dd), rs = ra + rb;
        if (!(d2 < rs * rs)) return;
        const C dist = sqrt(d2);
        const V3<C> n = dist > C(0) ? dd * (C(1) / dist) : mk3<C>(0, 1, 0);
        S.fA.n = 0;
        S.fB.n = 0;
        mesh_add_hit(d, Col, nm, base, pa + n * ra, pb - n * rb, n, rs - dist, subA, subB, S.fA, S.fB, S);
        return;
    }
    ConvexCtx<C> c;
    convex_setup(c, pa, qa, pb, qb, tol.spec);
    Sup<C> SA, SB;
    sup_for_shape(SA, st, sa, false);
    sup_for_shape(SB, st, sb, false);
    SB.transformed = true;
    SB.R = m3_from_quat(c.rel);
    SB.t = c.relPos;
    V3<C> axis0 = c.relPos;
    if (len2(axis0) <= C(1e-12)) axis0 = mk3<C>(1, 0, 0);
    PenDepth<C> pd;
    Simplex<C> simplex;
    simplex.size = 0;
    pd_step_gjk(pd, simplex, SA, SB, SA.cr + c.maxSepUse, SB.cr, axis0, tol.colTol);
    if (pd.status == PD_NOT_COLLIDING) return;
    if (pd.status == PD_INDETERMINATE) {
        c.maxSepUse = tmin<C>(c.maxSepUse, C(1));
        Sup<C> SA2, SB2;
        sup_for_shape(SA2, st, sa, true);
        sup_for_shape(SB2, st, sb, true);
        SA2.addRadius = true;
        SA2.addR = c.maxSepUse;
        SB2.transformed = true;
        SB2.R = SB.R;
        SB2.t = c.relPos;
        if (!pd_step_epa(pd, SA2, SB2, tol.penTol, simplex, H)) return;
    }
    const C pen = sqrt(len2(pd.pointA - pd.pointB)) - c.maxSepUse;
    const C al = sqrt(len2(pd.axis));
    if (al > C(0)) { const C corr = c.maxSepUse / al; pd.pointA = pd.pointA + pd.axis * (-corr); }
    const V3<C> wA = qrot2(qa, pd.pointA) + pa;
    const V3<C> wB = qrot2(qa, pd.pointB) + pa;
    const V3<C> wAxis = qrot2(qa, pd.axis);
    supporting_face(S.fA, st, sa, -pd.axis, pa, qa);
    const Q4<C> crel = qconj(c.rel);
    const V3<C> dirB = qrot2(crel, pd.axis);
    const V3<C> posB = qrot2(qa, c.relPos) + pa;
    Q4<C> qB;
    qB.x = qa.w * c.rel.x + qa.x * c.rel.w + qa.y * c.rel.z - qa.z * c.rel.y;
    qB.y = qa.w * c.rel.y - qa.x * c.rel.z + qa.y * c.rel.w + qa.z * c.rel.x;
    qB.z = qa.w * c.rel.z + qa.x * c.rel.y - qa.y * c.rel.x + qa.z * c.rel.w;
    qB.w = qa.w * c.rel.w - qa.x * c.rel.x - qa.y * c.rel.y - qa.z * c.rel.z;
    supporting_face(S.fB, st, sb, dirB, posB, qB);
    mesh_add_hit(d, Col, nm, base, wA, wB, wAxis, pen, subA, subB, S.fA, S.fB, S);
}

template <typename T, typename C>
__global__ void __launch_bounds__(64) k_compound(Dev<T> d, T dt) {
    const int n = min(d.counters[CNT_COMPOUND], d.pairCap);
    const int worker = blockIdx.x * blockDim.x + threadIdx.x;
    if (worker >= d.cmpWorkers) return;
    const int par = *d.parity;
    const ShapeTable<C>& st = np_shapes<T, C>(d);
    const NpTol<C> tol = np_tol<T, C>(d);
    MeshColl<C>& Col = ((MeshColl<C>*)d.cmpColl)[worker];
    EpaH<C> H;
    H.b = d.cmpEpa + (size_t)(worker / EPA_LANES) * EpaWave<C>::bytes;
    H.lane = worker % EPA_LANES;
    NpScratch<C, CC_MAX_FACE> S;
    for (int j = worker; j < n; j += d.cmpWorkers) {
        const int slot = d.cmpList[j];
        const int i0 = d.pairA[slot], i1 = d.pairB[slot];
        const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
        const PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
        const bool sw = pb.a != i0;
        const NpBody<T>& RA = sw ? r1 : r0;
        const NpBody<T>& RB = sw ? r0 : r1;
        const ShapeD<C>& KA = st.shapes[RA.shape];
        const ShapeD<C>& KB = st.shapes[RB.shape];
        const bool ca = KA.type == SHAPE_COMPOUND, cb = KB.type == SHAPE_COMPOUND;
        const int na = ca ? KA.ncc : 1, nb = cb ? KB.ncc : 1;
        const int ba = ca ? sub_bits(na) : 0, bb = cb ? sub_bits(nb) : 0;
        int nm = 0;
        for (int ia = 0; ia < na; ia++) {
            const int sa = ca ? KA.cfc + ia : RA.shape;
            const unsigned subA = ca ? sub_push(0xffffffffu, 0, ia, ba) : 0xffffffffu;
            for (int ib = 0; ib < nb; ib++) {
                const int sb = cb ? KB.cfc + ib : RB.shape;
                const unsigned subB = cb ? sub_push(0xffffffffu, 0, ib, bb) : 0xffffffffu;
                leaf_pair_hit(d, st, tol, sa, RA, sb, RB, subA, subB, H, S, Col, nm);
            }
        }
        // finalizeAndCreateConstraints (update.ts:522-555): the pair's manifolds become a run of
        // consecutive constraint slots
        int run = 0;
        for (int i = 0; i < nm; i++) {
            Manifold<C, MESH_MAX_POINTS>& m = Col.m[i];
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
            if (Col.m[i].n <= 0) continue;
            make_contact(d, slot, pb.a, pb.b, RA, RB, Col.m[i], Col.subA[i], Col.subB[i], cs0 + k, k, run, dt, par);
            k++;
        }
    }
}

}  // namespace cc
