// Continuous collision detection for MotionQuality.LINEAR_CAST bodies (src/update.ts:896-1677,
// src/ccd.ts), run by the world's workgroup between velocity integration and the position
// iterations: the integration pass delays the move of every dynamic LINEAR_CAST body whose step
// displacement exceeds its threshold and registers it here in body order; each wavefront then
// casts one registered body at a time against the world's bodies (one lane per candidate, the
// earliest hit by a shuffle reduction), and the workgroup ranks the records by time of impact and
// resolves them in that order on one thread, as the reference does.
#pragma once
#include "cc_compound.h"

namespace cc {

// gjkCastShape gjk.ts:853-1267: A (already in B's space) moves by disp against B at rest, both
// without their convex radius.  Returns false for no hit, else the fraction, A's contact point and
// the separating axis in B's space.
template <typename C>
__device__ inline bool gjk_cast_shape(const Sup<C>& A, const Sup<C>& B, V3<C> disp, C tolerance, C crA, C crB, C maxLambda,
                                      C& lambdaOut, V3<C>& pointA, V3<C>& sepAxis) {
    const C BIG = sizeof(C) == 8 ? C(1.7976931348623157e308) : C(3.402823466e38);
    C tol2 = tolerance * tolerance;
    const C sumR = crA + crB;
    Simplex<C> s;
    s.size = 0;
    C lambda = C(0);
    V3<C> x = mk3<C>(0, 0, 0);
    V3<C> q = -sup_get(B, mk3<C>(0, 0, 0));
    V3<C> p = sup_get(A, mk3<C>(0, 0, 0));
    V3<C> v = q - p;
    C vLen2 = BIG;
    bool allowRestart = false;
    V3<C> prevV = mk3<C>(0, 0, 0);
    int it = 0;
    while (it < 100) {
        it++;
        p = sup_get(A, -v);
        q = sup_get(B, v);
        const V3<C> w = x - (q - p);
        const C vw = v.x * w.x + v.y * w.y + v.z * w.z - sumR * sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        if (vw > C(0)) {
            const C vr = v.x * disp.x + v.y * disp.y + v.z * disp.z;
            if (vr >= C(-1.0e-18)) return false;
            const C old = lambda;
            lambda -= vw / vr;
            if (old == lambda) break;
            if (lambda >= maxLambda) return false;
            x = disp * lambda;
            vLen2 = BIG;
            tol2 = tolerance + sumR;
            tol2 = tol2 * tol2;
            allowRestart = true;
        }
        s.pt[s.size].p = p;
        s.pt[s.size].q = q;
        s.size++;
        for (int i = 0; i < s.size; i++) s.pt[i].y = (x - s.pt[i].q) + s.pt[i].p;
        CPRes<C> cp;  // computeClosestPointToSimplex gjk.ts:512-596 with lastPointPartOfClosest false
        switch (s.size) {
            case 1: cp.set = 1; cp.point = s.pt[0].y; break;
            case 2: cp = cp_line(s.pt[0].y, s.pt[1].y, C(1e-10)); break;
            case 3: cp = cp_triangle(s.pt[0].y, s.pt[1].y, s.pt[2].y, false, C(1e-10)); break;
            default: cp = cp_tetra(s.pt[0].y, s.pt[1].y, s.pt[2].y, s.pt[3].y, false, C(1e-5)); break;
        }
        const C d2 = len2(cp.point);
        if (d2 < vLen2) {
            vLen2 = d2;
            v = cp.point;
        } else {
            if (!allowRestart) break;
            allowRestart = false;
            s.pt[0].p = p;
            s.pt[0].q = q;
            s.size = 1;
            v = x - q;
            vLen2 = BIG;
            continue;
        }
        if (cp.set == 0xf) break;
        int ns = 0;  // updatePointSetPQ
        for (int i = 0; i < s.size; i++)
            if (cp.set & (1 << i)) {
                if (ns != i) s.pt[ns] = s.pt[i];
                ns++;
            }
        s.size = ns;
        if (vLen2 <= tol2) break;
        prevV = v;
    }
    for (int i = 0; i < s.size; i++) s.pt[i].y = (x - s.pt[i].q) + s.pt[i].p;
    const C vl = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    V3<C> nv = mk3<C>(0, 0, 0);
    if (vl > C(0)) nv = v * (C(1) / vl);
    V3<C> pa = mk3<C>(0, 0, 0), pb = mk3<C>(0, 0, 0);
    if (s.size == 1) {
        pb = s.pt[0].q + nv * crB;
        pa = lambda > C(0) ? pb : s.pt[0].p + nv * (-crA);
    } else if (s.size == 2) {
        const Bary<C> b = bary2(s.pt[0].y, s.pt[1].y, C(1e-10));
        pb = ((pb + s.pt[0].q * b.u) + s.pt[1].q * b.v) + nv * crB;
        pa = lambda > C(0) ? pb : ((pa + s.pt[0].p * b.u) + s.pt[1].p * b.v) + nv * (-crA);
    } else if (s.size >= 3) {
        const Bary<C> b = bary3(s.pt[0].y, s.pt[1].y, s.pt[2].y, C(1e-10));
        pb = (((pb + s.pt[0].q * b.u) + s.pt[1].q * b.v) + s.pt[2].q * b.w) + nv * crB;
        pa = lambda > C(0) ? pb : (((pa + s.pt[0].p * b.u) + s.pt[1].p * b.v) + s.pt[2].p * b.w) + nv * (-crA);
    }
    lambdaOut = lambda;
    pointA = pa;
    sepAxis = sumR > C(0) ? -v : -prevV;
    return true;
}

// castConvexVsConvex convex.ts:675-875 through penetrationCastShape penetration.ts:621-688 as the
// CCD visitor calls them (update.ts:1218-1237: unit scales, returnDeepestPoint false, so no EPA;
// the supporting faces are never read).  The hit is accepted below a fraction of 1: a body that
// a sequential visit would cast with a smaller bound yields a fraction-plus-slop that loses the
// workgroup's minimum either way.
template <typename C>
__device__ inline bool cast_convex(const ShapeTable<C>& st, int sa, V3<C> pa, Q4<C> qa, V3<C> disp, int sb, V3<C> pb,
                                   Q4<C> qb, bool backfaces, C& fraction, V3<C>& pointA, V3<C>& penAxis) {
    const Q4<C> ib = qconj(qb);
    const V3<C> posAInB = qrotate(ib, pa - pb);
    const Q4<C> quatAInB = qmul(ib, qa);
    const V3<C> dispInB = qrotate(ib, disp);
    Sup<C> A, B;
    sup_for_shape(A, st, sa, false);
    sup_for_shape(B, st, sb, false);
    A.transformed = true;
    A.R = m3_from_quat(quatAInB);
    A.t = posAInB;
    const C tolerance = C(1e-4);
    C lambda;
    V3<C> pA, sep;
    if (!gjk_cast_shape(A, B, dispInB, tolerance, A.cr, B.cr, C(1), lambda, pA, sep)) return false;
    if (len2(sep) < tolerance * tolerance) sep = dispInB;  // contactNormalInvalid
    if (lambda > C(1)) return false;
    if (!backfaces && dot(sep, dispInB) <= C(0)) return false;
    if (lambda >= C(1)) return false;
    fraction = lambda;
    pointA = qrotate(qb, pA) + pb;
    penAxis = qrotate(qb, sep);
    return true;
}

// castConvexVsPlane plane.ts:727-880 as the CCD visitor calls it (unit scales): the convex shape's
// DEFAULT support point is taken to world space with A's transform and measured against the plane
// in the plane shape's own frame; the contact point goes through the plane body's transform
// (translated by the cast for a hit ahead) and the penetration axis is -normal through that same
// matrix, as a point.  The hit is accepted below a fraction of 1 (see cast_convex).
template <typename C>
__device__ inline bool cast_convex_plane(const ShapeTable<C>& st, int sa, V3<C> pa, Q4<C> qa, V3<C> disp, int sp, V3<C> pb,
                                         Q4<C> qb, bool backfaces, C& fraction, V3<C>& pointA, V3<C>& penAxis) {
    const ShapeD<C>& P = st.shapes[sp];
    const V3<C> sn0 = mk3<C>(P.he[0] / C(1), P.he[1] / C(1), P.he[2] / C(1));
    const C len = vlen(sn0);
    const V3<C> n = sn0 * (C(1) / len);
    const C c = P.radius / len;
    const M4<C> start = m4_from_rt(qa, pa);
    const int ty = st.shapes[sa].type;
    Sup<C> S;
    sup_for_shape(S, st, sa, ty == SHAPE_BOX || ty == SHAPE_CYLINDER);
    const V3<C> sw = m4_point(start, sup_get(S, -qrotate(qconj(qa), n)));
    const C sd = dot(n, sw) + c;
    const C cr = S.cr;
    const C pen = -sd + cr;
    const C dt = dot(disp, n);
    M4<C> hitM = m4_from_rt(qb, pb);
    C f;
    if (pen > C(0)) {
        if (!backfaces && dt > C(0)) return false;
        if (pen <= C(-1)) return false;
        f = C(0);
    } else if (dt < C(0)) {
        f = pen / dt;
        if (f >= C(1)) return false;
        const V3<C> o = disp * f;
        const M4<C> p2w = hitM;  // mat4.translate
        hitM.m[12] = p2w.m[0] * o.x + p2w.m[4] * o.y + p2w.m[8] * o.z + p2w.m[12];
        hitM.m[13] = p2w.m[1] * o.x + p2w.m[5] * o.y + p2w.m[9] * o.z + p2w.m[13];
        hitM.m[14] = p2w.m[2] * o.x + p2w.m[6] * o.y + p2w.m[10] * o.z + p2w.m[14];
        hitM.m[15] = p2w.m[3] * o.x + p2w.m[7] * o.y + p2w.m[11] * o.z + p2w.m[15];
    } else {
        return false;
    }
    fraction = f;
    pointA = m4_point(hitM, sw - n * cr);
    penAxis = m4_point(hitM, -n);
    return true;
}

// calculateConstraintPropertiesWithMassOverride axis-constraint-part.ts:222-264 (unit scales)
template <typename T>
__device__ inline void ccd_props(CcdPart<T>& p, bool dynA, bool dynB, T invMA, T invMB, const M3<T>& IA, const M3<T>& IB,
                                 V3<T> r1, V3<T> r2, V3<T> axis, T bias) {
    T inv = T(0);
    if (dynA) {
        p.r1x = cross(r1, axis);
        p.i1 = mul(IA, p.r1x);
        inv += invMA + dot(p.i1, p.r1x);
    }
    if (dynB) {
        p.r2x = cross(r2, axis);
        p.i2 = mul(IB, p.r2x);
        inv += invMB + dot(p.i2, p.r2x);
    }
    if (inv == T(0)) { p.eff = T(0); p.lambda = T(0); }
    else { p.eff = T(1) / inv; p.bias = bias; }
}
// getTotalLambda :578-625 and applyLambda :640-707
template <typename T>
__device__ inline T ccd_total(const CcdPart<T>& p, const Bodies<T>& B, int a, int b, bool movA, bool movB, V3<T> axis) {
    T jv;
    if (movA && movB) jv = dot(axis, B.v[a] - B.v[b]);
    else if (movA) jv = dot(axis, B.v[a]);
    else if (movB) jv = -dot(axis, B.v[b]);
    else jv = T(0);
    if (movA) jv += dot(p.r1x, B.w[a]);
    if (movB) jv -= dot(p.r2x, B.w[b]);
    const T lam = p.eff * (jv - p.bias);
    return p.lambda + lam;
}
template <typename T>
__device__ inline void ccd_apply(CcdPart<T>& p, const Bodies<T>& B, int a, int b, bool dynA, bool dynB, T invMA, T invMB,
                                 V3<T> axis, T total) {
    const T dl = total - p.lambda;
    p.lambda = total;
    if (dl == T(0)) return;
    if (dynA) {
        const T ls = dl * invMA;
        V3<T> v = B.v[a];
        v.x -= axis.x * ls; v.y -= axis.y * ls; v.z -= axis.z * ls;
        B.v[a] = dof_lin(v, fl_dof(B.flags[a]));
        V3<T> w = B.w[a];
        w.x -= p.i1.x * dl; w.y -= p.i1.y * dl; w.z -= p.i1.z * dl;
        B.w[a] = w;
    }
    if (dynB) {
        const T ls = dl * invMB;
        V3<T> v = B.v[b];
        v.x += axis.x * ls; v.y += axis.y * ls; v.z += axis.z * ls;
        B.v[b] = dof_lin(v, fl_dof(B.flags[b]));
        V3<T> w = B.w[b];
        w.x += p.i2.x * dl; w.y += p.i2.y * dl; w.z += p.i2.z * dl;
        B.w[b] = w;
    }
}
// clampLinearVelocity / clampAngularVelocity motion-properties.ts:624-645
template <typename T>
__device__ inline void ccd_clamp(const Bodies<T>& B, int i) {
    const T ml = B.maxLinV[i], ma = B.maxAngV[i];
    V3<T> v = B.v[i], w = B.w[i];
    if (len2(v) > ml * ml) B.v[i] = normalize(v) * ml;
    if (len2(w) > ma * ma) B.w[i] = normalize(w) * ma;
}

// applyCCD update.ts:1457-1593
template <typename T>
__device__ inline void apply_ccd(const Dev<T>& d, CcdPart<T>* parts, const CcdRec<T>& c, int a, int b) {
    const Bodies<T>& B = d.B;
    const int fa = B.flags[a], fb = B.flags[b];
    if (fl_motion(fa) != MOTION_DYNAMIC) return;
    const bool dynB = fl_motion(fb) == MOTION_DYNAMIC, movB = fl_motion(fb) != MOTION_STATIC;
    const V3<T> atHit = B.com[a] + c.delta * c.fraction;
    const V3<T> r1 = c.point - atHit;
    const V3<T> r2 = c.point - B.com[b];
    const V3<T> v1 = B.v[a] + cross(B.w[a], r1);
    const V3<T> v2 = movB ? B.v[b] + cross(B.w[b], r2) : mk3<T>(0, 0, 0);
    const V3<T> rv = v2 - v1;
    const T nv = dot(rv, c.normal);
    if (nv >= T(0)) return;
    T bias = T(0);
    if (c.restitution > T(0) && nv < -d.s.ccdMinVelRest) bias = c.restitution * nv;
    const T invMA = T(1) * B.invMass[a];
    const T invMB = dynB ? T(1) * B.invMass[b] : T(0);
    const M3<T> IA = inv_inertia(B.invI[a], B.iRot[a], B.q[a], fl_dof(fa));
    M3<T> IB;
    for (int k = 0; k < 9; k++) IB.m[k] = T(0);
    if (dynB) IB = inv_inertia(B.invI[b], B.iRot[b], B.q[b], fl_dof(fb));
    CcdPart<T> pn = parts[0];
    ccd_props(pn, true, dynB, invMA, invMB, IA, IB, r1, r2, c.normal, bias);
    T tot = ccd_total(pn, B, a, b, true, movB, c.normal);
    ccd_apply(pn, B, a, b, true, dynB, invMA, invMB, c.normal, tot);
    parts[0] = pn;
    if (c.friction > T(0)) {
        const V3<T> tv = rv - c.normal * nv;
        const T tl2 = len2(tv);
        if (tl2 > T(1e-12)) {
            const V3<T> dir = tv * (T(1) / sqrt(tl2));
            const T maxF = c.friction * pn.lambda;
            CcdPart<T> pf = parts[1];
            ccd_props(pf, true, dynB, invMA, invMB, IA, IB, r1, r2, dir, T(0));
            tot = ccd_total(pf, B, a, b, true, movB, dir);
            ccd_apply(pf, B, a, b, true, dynB, invMA, invMB, dir, tmax<T>(-maxF, tmin<T>(maxF, tot)));
            parts[1] = pf;
        }
    }
    ccd_clamp(B, a);
    if (dynB) ccd_clamp(B, b);
}

// One CCD body's search by one wavefront (findCCDContacts update.ts:1357-1437 with the visitor
// :1110-1240 and the collector :1003-1107): the lanes take the world's bodies [b0, b1) in turn
// through the visitor's filters, a swept-bounds test (A's bounds over its relative displacement,
// widened by 1e-3 m and a relative term for the bounds' rounding, against the body's bounds; it
// stands in for the reference's tree and ray early-outs) and the cast.  The earliest
// fraction-plus-slop wins, equal ones by the lower body index (the reference decides those by its
// tree's visit order).  The manifold block of :1376-1432 never runs there (the collector's reset
// after each visit, :1238, clears hasHit), so the contact point is the collector's.  The searches
// read only the state before the CCD pass, so the workgroup's wavefronts run different records
// at once.
template <typename T, typename C>
__device__ inline void ccd_find(const Dev<T>& d, CcdRec<T>* recs, const int* ccdIdx, int cb0, int e, int b0, int b1, T dt) {
    const int lane = threadIdx.x & 63;
    const Bodies<T>& B = d.B;
    const CcdRec<T> c = recs[cb0 + e];
    const int a = c.body;
    const int fa = B.flags[a];
    const ShapeTable<C>& stc = np_shapes<T, C>(d);
    const V3<double> amn = cvt<double>(B.amin[a]), amx = cvt<double>(B.amax[a]);
    T best = c.fps;
    int bestJ = 0x7fffffff;
    unsigned bSub = 0xffffffffu, bSubA = 0xffffffffu;
    T bFrac = T(0);
    V3<T> bNormal = mk3<T>(0, 0, 0), bPoint = mk3<T>(0, 0, 0);
    for (int j = b0 + lane; j < b1; j += 64) {
        if (j == a) continue;
        const int fb = B.flags[j];
        if (!d.s.layerPairs[fl_layer(fa) * 8 + fl_layer(fb)]) continue;
        if (!((B.groups[a] & B.mask[j]) != 0 && (B.groups[j] & B.mask[a]) != 0)) continue;
        V3<T> rel = c.delta;
        const int cj = ccdIdx[j];
        if (cj >= 0) {
            const CcdRec<T>& cb = recs[cb0 + cj];
            if (a > cb.body) continue;
            rel = rel - cb.delta;
        }
        if (len2(rel) < c.thrSq) continue;
        {  // swept bounds of A against B's bounds
            const V3<double> r = cvt<double>(rel);
            const V3<double> bmn = cvt<double>(B.amin[j]), bmx = cvt<double>(B.amax[j]);
            const double lo[3] = {amn.x + fmin(r.x, 0.0), amn.y + fmin(r.y, 0.0), amn.z + fmin(r.z, 0.0)};
            const double hi[3] = {amx.x + fmax(r.x, 0.0), amx.y + fmax(r.y, 0.0), amx.z + fmax(r.z, 0.0)};
            const double blo[3] = {bmn.x, bmn.y, bmn.z}, bhi[3] = {bmx.x, bmx.y, bmx.z};
            bool sep = false;
            for (int k = 0; k < 3; k++) {
                const double m = 1e-3 + 1e-5 * (fabs(lo[k]) + fabs(hi[k]) + fabs(blo[k]) + fabs(bhi[k]));
                sep = sep || lo[k] - m > bhi[k] || blo[k] - m > hi[k];
            }
            if (sep) continue;
        }
        // a leaf, or every child of a compound at the body pose composed with the child's transform
        // with its sub-shape id pushed (castShapeVsCompound compound.ts)
        // A's leaves: castCompoundVsShape (compound.ts: each child at the composed pose, the
        // displacement turned by A's rotation, its id pushed) or castTransformedVsShape
        // (transformed.ts:537-633, the same for the one transform); scaled and
        // offset-centre-of-mass pass through
        const ShapeD<C>& DA0 = stc.shapes[B.shape[a]];
        const bool cmpA = DA0.type == SHAPE_COMPOUND;
        const int na = cmpA ? DA0.ncc : 1;
        const int abits = cmpA ? sub_bits(na) : 0;
        for (int ka = 0; ka < na; ka++) {
        const int sa = cmpA ? DA0.cfc + ka : B.shape[a];
        const ShapeD<C>& DA = stc.shapes[sa];
        V3<C> pa = cvt<C>(B.pos[a]), da = cvt<C>(rel);
        Q4<C> qa = cvtq<C>(B.q[a]);
        if (DA.deco & DECO_TRANSFORMED) {
            const Q4<C> q0 = qa;
            pa = pa + qrotate(q0, mk3<C>(DA.lp[0], DA.lp[1], DA.lp[2]));
            Q4<C> lq;
            lq.x = DA.lq[0]; lq.y = DA.lq[1]; lq.z = DA.lq[2]; lq.w = DA.lq[3];
            qa = qmul(q0, lq);
            da = qrotate(q0, da);
        }
        const unsigned subA = cmpA ? sub_push(0xffffffffu, 0, ka, abits) : 0xffffffffu;
        const int sj = B.shape[j];
        const ShapeD<C>& DJ = stc.shapes[sj];
        const bool cmp = DJ.type == SHAPE_COMPOUND;
        const int nk = cmp ? DJ.ncc : 1;
        const int nbits = cmp ? sub_bits(nk) : 0;
        const V3<C> pj = cvt<C>(B.pos[j]);
        const Q4<C> qj = cvtq<C>(B.q[j]);
        for (int k = 0; k < nk; k++) {
            int sk = sj;
            V3<C> pk = pj;
            Q4<C> qk = qj;
            unsigned subk = 0xffffffffu;
            if (!cmp && (DJ.deco & DECO_TRANSFORMED)) {  // castShapeVsTransformed (transformed.ts:634-700)
                pk = pj + qrotate(qj, mk3<C>(DJ.lp[0], DJ.lp[1], DJ.lp[2]));
                Q4<C> lq;
                lq.x = DJ.lq[0]; lq.y = DJ.lq[1]; lq.z = DJ.lq[2]; lq.w = DJ.lq[3];
                qk = qmul(qj, lq);
            }
            if (cmp) {
                sk = DJ.cfc + k;
                const ShapeD<C>& K = stc.shapes[sk];
                pk = pj + qrotate(qj, mk3<C>(K.lp[0], K.lp[1], K.lp[2]));
                Q4<C> lq;
                lq.x = K.lq[0]; lq.y = K.lq[1]; lq.z = K.lq[2]; lq.w = K.lq[3];
                qk = qmul(qj, lq);
                subk = sub_push(0xffffffffu, 0, k, nbits);
            }
            C fraction;
            V3<C> pA, pen;
            const bool hit = stc.shapes[sk].type == SHAPE_PLANE
                                 ? cast_convex_plane(stc, sa, pa, qa, da, sk, pk, qk, d.s.backfaces != 0, fraction, pA, pen)
                                 : cast_convex(stc, sa, pa, qa, da, sk, pk, qk, d.s.backfaces != 0, fraction, pA, pen);
            if (!hit) continue;
            // addHit :1045-1098
            const V3<T> n = cvt<T>(normalize(pen));
            const T den = dot(n, c.delta);
            if (den <= c.maxPen) continue;
            const T fps = T(fraction) + c.maxPen / den;
            if (!(fps < best)) continue;
            best = fps;
            bestJ = j;
            bSub = subk;
            bSubA = subA;
            bFrac = T(fraction);
            bNormal = n;
            const V3<T> mb = ((fb & FL_LINEAR_CAST) && fl_motion(fb) != MOTION_STATIC) ? B.v[j] * dt : mk3<T>(0, 0, 0);
            bPoint = len2(mb) > T(1e-12) ? cvt<T>(pA) + mb * T(fraction) : cvt<T>(pA);
        }
        }
    }
    // the wavefront's earliest (fraction-plus-slop, body index)
    T m = best;
    int mj = bestJ;
    for (int o = 32; o > 0; o >>= 1) {
        const T om = __shfl_xor(m, o, 64);
        const int oj = __shfl_xor(mj, o, 64);
        if (om < m || (om == m && oj < mj)) { m = om; mj = oj; }
    }
    if (bestJ != 0x7fffffff && bestJ == mj) {
        CcdRec<T>& r = recs[cb0 + e];
        r.hit = bestJ;
        r.sub1 = bSubA;
        r.sub2 = bSub;
        r.fraction = bFrac;
        r.fps = best;
        r.normal = bNormal;
        r.point = bPoint;
        r.friction = combine(B.friction[a], B.friction[bestJ], fl_fmode(fa), fl_fmode(B.flags[bestJ]));
        r.restitution = combine(B.restitution[a], B.restitution[bestJ], fl_rmode(fa), fl_rmode(B.flags[bestJ]));
    }
}

// sortCCDBodies :1607-1621 (by fraction-plus-slop, then body index) and resolveCCDContacts
// :1626-1677.  The body's ccdBodyIndex keeps its place before the sort, and the cascade check
// looks it up in the sorted order (:1647-1649), as there.
template <typename T, int BS>
__device__ inline void ccd_resolve(const Dev<T>& d, CcdRec<T>* recs, int* ord, const int* ccdIdx, CcdPart<T>* parts, int cb0,
                                   int n) {
    for (int e = threadIdx.x; e < n; e += BS) {
        const CcdRec<T>& c = recs[cb0 + e];
        int rank = 0;
        for (int k = 0; k < n; k++) {
            const CcdRec<T>& o = recs[cb0 + k];
            rank += (o.fps < c.fps || (o.fps == c.fps && o.body < c.body)) ? 1 : 0;
        }
        ord[cb0 + rank] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const Bodies<T>& B = d.B;
        for (int i = 0; i < n; i++) {
            const CcdRec<T>& c = recs[cb0 + ord[cb0 + i]];
            const int a = c.body;
            if (c.hit == -1) {
                if (fl_motion(B.flags[a]) == MOTION_DYNAMIC) {
                    B.pos[a] = B.pos[a] + c.delta;
                    set_transform(B, d.st, a);
                }
                continue;
            }
            const int b = c.hit;
            const int cbi = ccdIdx[b];
            if (cbi >= 0) {
                CcdRec<T>& cb = recs[cb0 + ord[cb0 + cbi]];
                if (cb.hit != -1 && cb.fraction > c.fraction) {
                    cb.hit = -1;
                    cb.fps = c.fraction;
                }
            }
            apply_ccd(d, parts, c, a, b);
            if (fl_motion(B.flags[a]) == MOTION_DYNAMIC) {
                B.pos[a] = B.pos[a] + c.delta * c.fps;
                set_transform(B, d.st, a);
            }
            if (B.sleeping[b]) {
                reset_sleep_timer(B, d.st, b);
                B.sleeping[b] = 0;
            }
        }
    }
    __syncthreads();
}

}  // namespace cc
