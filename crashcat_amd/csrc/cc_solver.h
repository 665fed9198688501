// Per-world solver kernel: one workgroup owns one world for the rest of the step.
//
// crashcat solves each island's contacts serially in key-sorted order (projected
// Gauss-Seidel, src/constraints/contact-constraints.ts:963-1053).  The workgroup keeps
// that order exactly: it sorts the world's constraints by (sortKey, idA, idB) — which is
// the per-island order, since islands never share a body — links every constraint to the
// previous constraint of each of its dynamic bodies, and assigns each constraint the
// length of the longest such chain ending at it (its level).  Constraints of one level
// share no dynamic body, so a level is solved in parallel and levels run in order;
// every body therefore sees exactly the serial sequence of updates.
#pragma once
#include "cc_ccd.h"

namespace cc {

template <int BS>
__device__ inline int block_reduce_max(int v, int* sh) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    int r = sh[0];
    for (int i = 1; i < BS / 64; i++) r = max(r, sh[i]);
    __syncthreads();
    return r;
}

// exclusive scan of one int per thread; returns the prefix, writes the block total
template <int BS>
__device__ inline int block_excl_scan(int v, int* sh, int& total) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < BS / 64; i++) {
        if (i < wave) off += sh[i];
        tot += sh[i];
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

// exclusive scan of arr[0..n) in place, returns the total
template <int BS>
__device__ inline int block_scan_array(int* arr, int n, int* sh) {
    int carry = 0;
    for (int base = 0; base < n; base += BS) {
        int i = base + threadIdx.x;
        int v = i < n ? arr[i] : 0;
        int tot;
        int ex = block_excl_scan<BS>(v, sh, tot);
        if (i < n) arr[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    return carry;
}

// velocities of body i in its velocity record
template <typename T>
__device__ __forceinline__ void vw_load(const Dev<T>& d, int i, V3<T>& v, V3<T>& w) {
    const VBody<T>& r = d.vb[i];
    v = r.v;
    w = r.w;
}
template <typename T>
__device__ __forceinline__ void vw_store(const Dev<T>& d, int i, V3<T> v, V3<T> w) {
    VBody<T>& r = d.vb[i];
    r.v = v;
    r.w = w;
}

// --------------------------------------------------------------- contact constraints
// One contact constraint of the sweeps (solveVelocityConstraint contact-constraints.ts:963-1053):
// friction of every point, then the normal of every point.  All loads are issued before the
// first update, so a constraint costs two dependent memory round trips (its element record,
// then bodies and points).  In the last sweep (rec set) the impulses go to this step's cache
// record (storeAppliedImpulses :1061-1076) instead of the point records.
// TAPE: also record each point's impulses of this sweep for the adjoint (tl: the sweep's records)
// An element's inputs that do not change during a sweep: its points, the bodies' inverse inertia
// and inverse mass.
template <typename T>
struct VIn {
    VPoint<T> P[4];
    M3<T> IA, IB;
    T iA, iB;
};
template <typename T>
__device__ __forceinline__ void vin_load(const Dev<T>& d, const VElem<T>& E, VIn<T>& in) {
    const int np = (E.meta >> 4) & 7;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j < np) in.P[j] = d.vp[E.pt + j];
    in.IA = d.vb[E.a].I;
    in.IB = d.vb[E.b].I;
    in.iA = d.vb[E.a].invMass;
    in.iB = d.vb[E.b].invMass;
}

template <typename T, bool TAPE = false>
__device__ __forceinline__ void solve_velocity_core(const Dev<T>& d, const VElem<T>& E, VIn<T>& in, CacheRec<T>* rec,
                                                    LamRec<T>* tl) {
    const int meta = E.meta;
    const int dyn = meta & 3;
    const int mov = (meta >> 2) & 3, np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    VPoint<T>* P = in.P;
    VPoint<T>* Lm = P;
    V3<T> vA, wA, vB, wB;
    vw_load(d, E.a, vA, wA);
    vw_load(d, E.b, vB, wB);
    const V3<T> n = E.n, t1 = contact_t1(n), t2 = cross(n, t1);
    const T iA = in.iA, iB = in.iB, fr = E.fr;
    const int npS = dyn ? np : 0;  // a constraint without a dynamic body is not solved
    T pre1[4], pre2[4], clp[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        pre1[j] = pre2[j] = clp[j] = T(0);
        if (j >= npS) continue;
        const VPoint<T>& p = P[j];
        VPoint<T>& l = Lm[j];
        if (p.effT1 != T(0) || p.effT2 != T(0)) {
            T l1 = l.lT1 + p.effT1 * axis_jv(dyn, mov, p.r1, p.r2, vA, wA, vB, wB, t1);
            T l2 = l.lT2 + p.effT2 * axis_jv(dyn, mov, p.r1, p.r2, vA, wA, vB, wB, t2);
            T mf = fr * l.lN;
            T fm2 = l1 * l1 + l2 * l2;
            if (TAPE) { pre1[j] = l1; pre2[j] = l2; }
            if (fm2 > mf * mf) {
                T sc = mf / sqrt(fm2);
                l1 *= sc;
                l2 *= sc;
                if (TAPE) clp[j] = T(1);
            }
            T d1 = l1 - l.lT1;
            l.lT1 = l1;
            if (d1 != T(0)) axis_impulse(dyn, dofA, dofB, iA, iB, in.IA, in.IB, p.r1, p.r2, vA, wA, vB, wB, t1, d1);
            T d2 = l2 - l.lT2;
            l.lT2 = l2;
            if (d2 != T(0)) axis_impulse(dyn, dofA, dofB, iA, iB, in.IA, in.IB, p.r1, p.r2, vA, wA, vB, wB, t2, d2);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (j >= npS) continue;
        const VPoint<T>& p = P[j];
        VPoint<T>& l = Lm[j];
        T tl = l.lN + p.effN * (axis_jv(dyn, mov, p.r1, p.r2, vA, wA, vB, wB, n) - p.biasN);
        tl = tmax<T>(T(0), tl);
        T dn = tl - l.lN;
        l.lN = tl;
        if (dn != T(0)) axis_impulse(dyn, dofA, dofB, iA, iB, in.IA, in.IB, p.r1, p.r2, vA, wA, vB, wB, n, dn);
    }
    if (TAPE) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j >= np) continue;
            LamRec<T> r;
            r.n = Lm[j].lN;
            r.t1 = pre1[j];
            r.t2 = pre2[j];
            r.clamp = clp[j];
            tl[E.pt + j] = r;
        }
    }
    if (rec) {
        CacheRec<T>& r = rec[E.slot];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j >= np) continue;
            r.ln[j] = Lm[j].lN;
            r.lt1[j] = Lm[j].lT1;
            r.lt2[j] = Lm[j].lT2;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j >= np) continue;
            VPoint<T>& q = d.vp[E.pt + j];
            q.lN = Lm[j].lN;
            q.lT1 = Lm[j].lT1;
            q.lT2 = Lm[j].lT2;
        }
    }
    if (dyn & 1) vw_store(d, E.a, vA, wA);
    if (dyn & 2) vw_store(d, E.b, vB, wB);
}
template <typename T, bool TAPE = false>
__device__ __forceinline__ void solve_velocity_elem(const Dev<T>& d, const VElem<T>& E, CacheRec<T>* rec,
                                                    LamRec<T>* tl = nullptr) {
    VIn<T> in;
    vin_load(d, E, in);
    solve_velocity_core<T, TAPE>(d, E, in, rec, tl);
}

// warm start of one constraint's points (warmStartVelocity contact-constraints.ts:906-955)
template <typename T>
__device__ inline void warm_points(int dyn, int dofA, int dofB, T iA, T iB, const M3<T>& IA, const M3<T>& IB, V3<T> n,
                                   V3<T> t1, V3<T> t2, int np, VPoint<T>* P, V3<T>& vA, V3<T>& wA, V3<T>& vB,
                                   V3<T>& wB, T ratio) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (j >= np) continue;
        VPoint<T>& p = P[j];
        VPoint<T>& l = P[j];
        if (p.effT1 != T(0) || p.effT2 != T(0)) {
            axis_warm(dyn, dofA, dofB, iA, iB, IA, IB, p.r1, p.r2, vA, wA, vB, wB, t1, l.lT1, ratio);
            axis_warm(dyn, dofA, dofB, iA, iB, IA, IB, p.r1, p.r2, vA, wA, vB, wB, t2, l.lT2, ratio);
        }
        axis_warm(dyn, dofA, dofB, iA, iB, IA, IB, p.r1, p.r2, vA, wA, vB, wB, n, l.lN, ratio);
    }
}

template <typename T>
__device__ inline void warm_start_elem(const Dev<T>& d, const VElem<T>& E, T ratio) {
    const int meta = E.meta;
    const int dyn = meta & 3;
    const int np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    const VBody<T> IA = d.vb[E.a], IB = d.vb[E.b];
    VPoint<T> P[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j < np) P[j] = d.vp[E.pt + j];
    VPoint<T>* Lm = P;
    V3<T> vA = IA.v, wA = IA.w, vB = IB.v, wB = IB.w;
    const V3<T> t1 = contact_t1(E.n);
    warm_points(dyn, dofA, dofB, IA.invMass, IB.invMass, IA.I, IB.I, E.n, t1, cross(E.n, t1), np, P, vA, wA, vB, wB,
                ratio);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (j >= np) continue;
        VPoint<T>& q = d.vp[E.pt + j];
        q.lN = Lm[j].lN;
        q.lT1 = Lm[j].lT1;
        q.lT2 = Lm[j].lT2;
    }
    if (dyn & 1) vw_store(d, E.a, vA, wA);
    if (dyn & 2) vw_store(d, E.b, vB, wB);
}

// the same warm start on a constraint record, in creation order (before the rows are staged)
template <typename T>
__device__ inline void warm_start_con(const Dev<T>& d, int slot, T ratio) {
    Con<T>& k = d.con[slot];
    const int dyn = k.dynMask;
    const int dofA = fl_dof(d.B.flags[k.a]), dofB = fl_dof(d.B.flags[k.b]);
    const VBody<T> IA = d.vb[k.a], IB = d.vb[k.b];
    const int np = k.np;
    VPoint<T> P[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (j >= np) continue;
        const ConPoint<T>& c = k.p[j];
        P[j].r1 = c.r1; P[j].r2 = c.r2;
        P[j].effN = c.effN; P[j].effT1 = c.effT1; P[j].effT2 = c.effT2;
        P[j].lN = c.lN; P[j].lT1 = c.lT1; P[j].lT2 = c.lT2;
    }
    V3<T> vA = IA.v, wA = IA.w, vB = IB.v, wB = IB.w;
    warm_points(dyn, dofA, dofB, k.invMassA, k.invMassB, IA.I, IB.I, k.normal, k.t1, k.t2, np, P, vA, wA, vB, wB, ratio);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (j >= np) continue;
        k.p[j].lN = P[j].lN; k.p[j].lT1 = P[j].lT1; k.p[j].lT2 = P[j].lT2;
    }
    if (dyn & 1) vw_store(d, k.a, vA, wA);
    if (dyn & 2) vw_store(d, k.b, vB, wB);
}

// one contact constraint of a position sweep (solvePositionConstraint :1098-1216) on the pose
// records PA, PB of its bodies and its body-space points LA, LB
template <typename T>
__device__ inline bool solve_position_elem(const Dev<T>& d, const VElem<T>& E, const PBody<T>& PA, const PBody<T>& PB,
                                           const V3<T>* LA, const V3<T>* LB) {
    const int meta = E.meta;
    const int dyn = meta & 3;
    if (!dyn) return false;
    const int np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    const T iMA = PA.invMass, iMB = PB.invMass;
    V3<T> cA = PA.com, cB = PB.com;
    Q4<T> qA = PA.q, qB = PB.q;
    M3<T> RA = m3_from_quat(qA), RB = m3_from_quat(qB);
    M3<T> IA, IB;
    for (int j = 0; j < 9; j++) { IA.m[j] = T(0); IB.m[j] = T(0); }
    if (dyn & 1) IA = inv_inertia(PA.invI, PA.iRot, qA, dofA);
    if (dyn & 2) IB = inv_inertia(PB.invI, PB.iRot, qB, dofB);
    bool any = false;
    const V3<T> n = E.n;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (j >= np) continue;
        V3<T> pA = cA + mul(RA, LA[j]);
        V3<T> pB = cB + mul(RB, LB[j]);
        T sep = dot(pB - pA, n) + d.s.slop;
        sep = tmax<T>(sep, -d.s.maxPen);
        if (sep >= T(0)) continue;
        V3<T> mid = mk3<T>((pA.x + pB.x) * T(0.5), (pA.y + pB.y) * T(0.5), (pA.z + pB.z) * T(0.5));
        V3<T> rA = mid - cA, rB = mid - cB;
        Row<T> rn;
        axis_props(rn, (dyn & 1) != 0, (dyn & 2) != 0, iMA, iMB, IA, IB, rA, rB, n, T(0));
        // solvePositionConstraintWithMassOverride axis-constraint-part.ts:790-847
        if (sep == T(0) || rn.eff == T(0)) continue;
        T lam = -rn.eff * d.s.beta * sep;
        if (dyn & 1) {
            cA = dof_lin(cA - n * (lam * iMA), dofA);
            add_rot_step(qA, rn.i1 * lam, T(-1));
        }
        if (dyn & 2) {
            cB = dof_lin(cB + n * (lam * iMB), dofB);
            add_rot_step(qB, rn.i2 * lam, T(1));
        }
        any = true;
    }
    if (dyn & 1) { PBody<T>& r = d.pb[E.a]; r.com = cA; r.q = qA; }
    if (dyn & 2) { PBody<T>& r = d.pb[E.b]; r.com = cB; r.q = qB; }
    return any;
}

// ---------------------------------------------------------------- point constraint
template <typename T>
__device__ inline void point_props_q(const Dev<T>& d, PointCon<T>& p, Q4<T> qa, Q4<T> qb) {  // point-constraint-part.ts:77-165
    const Bodies<T>& B = d.B;
    int a = p.a, b = p.b;
    M3<double> RA = m3_from_quat(cvtq<double>(qa)), RB = m3_from_quat(cvtq<double>(qb));
    V3<double> r1 = mul(RA, cvt<double>(p.localA)), r2 = mul(RB, cvt<double>(p.localB));
    p.r1 = cvt<T>(r1);
    p.r2 = cvt<T>(r2);
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double sumInv = 0;
    M3<double> X1, X2;
    for (int j = 0; j < 9; j++) { p.i1X.m[j] = T(0); p.i2X.m[j] = T(0); }
    auto accum = [&](int body, V3<double> r, const M3<double>& R, M3<T>& iX) {
        int f = B.flags[body];
        sumInv += double(B.invMass[body]);
        M3<double> I = inv_inertia_d(cvt<double>(B.invI[body]), cvtq<double>(B.iRot[body]), R, fl_dof(f));
        M3<double> X;  // crossProductMatrix
        X.m[0] = 0; X.m[1] = r.z; X.m[2] = -r.y; X.m[3] = -r.z; X.m[4] = 0; X.m[5] = r.x; X.m[6] = r.y; X.m[7] = -r.x; X.m[8] = 0;
        M3<double> IX, XI;
        for (int c = 0; c < 3; c++)
            for (int rr = 0; rr < 3; rr++) {
                IX.m[c * 3 + rr] = I.m[rr] * X.m[c * 3] + I.m[3 + rr] * X.m[c * 3 + 1] + I.m[6 + rr] * X.m[c * 3 + 2];
                XI.m[c * 3 + rr] = X.m[rr] * I.m[c * 3] + X.m[3 + rr] * I.m[c * 3 + 1] + X.m[6 + rr] * I.m[c * 3 + 2];
            }
        for (int j = 0; j < 9; j++) iX.m[j] = T(IX.m[j]);
        for (int c = 0; c < 3; c++)  // XI * X^T
            for (int rr = 0; rr < 3; rr++)
                K[c * 3 + rr] += XI.m[rr] * X.m[c] + XI.m[3 + rr] * X.m[3 + c] + XI.m[6 + rr] * X.m[6 + c];
    };
    if (fl_motion(B.flags[a]) == MOTION_DYNAMIC) accum(a, r1, RA, p.i1X);
    if (fl_motion(B.flags[b]) == MOTION_DYNAMIC) accum(b, r2, RB, p.i2X);
    K[0] += sumInv; K[4] += sumInv; K[8] += sumInv;
    // mat4.invert3x3 in FP64
    double a00 = K[0], a01 = K[1], a02 = K[2], a10 = K[3], a11 = K[4], a12 = K[5], a20 = K[6], a21 = K[7], a22 = K[8];
    double b01 = a22 * a11 - a12 * a21, b11 = -a22 * a10 + a12 * a20, b21 = a21 * a10 - a11 * a20;
    double det = a00 * b01 + a01 * b11 + a02 * b21;
    if (det == 0) {
        p.active = 0;
        for (int j = 0; j < 9; j++) p.eff.m[j] = T(0);
        p.lambda = mk3<T>(0, 0, 0);
        return;
    }
    det = 1.0 / det;
    double o[9] = {b01 * det, (-a22 * a01 + a02 * a21) * det, (a12 * a01 - a02 * a11) * det,
                   b11 * det, (a22 * a00 - a02 * a20) * det, (-a12 * a00 + a02 * a10) * det,
                   b21 * det, (-a21 * a00 + a01 * a20) * det, (a11 * a00 - a01 * a10) * det};
    for (int j = 0; j < 9; j++) p.eff.m[j] = T(o[j]);
    p.active = 1;
}

// applyVelocityStep point-constraint-part.ts:232-262 and solveVelocityConstraint :186-221 on the
// velocity records
template <typename T>
__device__ inline void point_props(const Dev<T>& d, PointCon<T>& p) {
    point_props_q(d, p, d.B.q[p.a], d.B.q[p.b]);
}

template <typename T>
__device__ __forceinline__ void point_apply_vw(const Dev<T>& d, PointCon<T>& p, V3<T> lam) {
    if (len2(lam) == T(0)) return;
    const Bodies<T>& B = d.B;
    V3<T> v, w;
    if (fl_motion(B.flags[p.a]) == MOTION_DYNAMIC) {
        vw_load(d, p.a, v, w);
        vw_store(d, p.a, v + lam * (-B.invMass[p.a]), w - mul(p.i1X, lam));
    }
    if (fl_motion(B.flags[p.b]) == MOTION_DYNAMIC) {
        vw_load(d, p.b, v, w);
        vw_store(d, p.b, v + lam * B.invMass[p.b], w + mul(p.i2X, lam));
    }
}
template <typename T>
__device__ __forceinline__ void point_solve_velocity_vw(const Dev<T>& d, PointCon<T>& p) {
    if (!p.active) return;
    V3<T> va, wa, vb, wb;
    vw_load(d, p.a, va, wa);
    vw_load(d, p.b, vb, wb);
    V3<T> jv = va - cross(p.r1, wa) - vb;
    jv = jv + cross(p.r2, wb);
    V3<T> lam = mul(p.eff, jv);
    p.lambda = p.lambda + lam;
    point_apply_vw(d, p, lam);
}

// solvePositionConstraint point-constraint.ts:211-230 (part :271-316) on the pose records
template <typename T>
__device__ inline bool point_solve_position(const Dev<T>& d, PointCon<T>& p, PBody<T> PA, PBody<T> PB) {
    point_props_q(d, p, PA.q, PB.q);
    if (!p.active) return false;
    const Bodies<T>& B = d.B;
    V3<T> sep = PB.com + p.r2 - PA.com - p.r1;
    if (len2(sep) < T(1e-12)) return false;
    V3<T> lam = mul(p.eff, sep) * (-d.s.beta);
    if (fl_motion(B.flags[p.a]) == MOTION_DYNAMIC) {
        PBody<T>& r = d.pb[p.a];
        r.com = dof_lin(PA.com - lam * B.invMass[p.a], fl_dof(B.flags[p.a]));
        Q4<T> q = PA.q;
        add_rot_step(q, mul(p.i1X, lam), T(-1));
        r.q = q;
    }
    if (fl_motion(B.flags[p.b]) == MOTION_DYNAMIC) {
        PBody<T>& r = d.pb[p.b];
        r.com = dof_lin(PB.com + lam * B.invMass[p.b], fl_dof(B.flags[p.b]));
        Q4<T> q = PB.q;
        add_rot_step(q, mul(p.i2X, lam), T(1));
        r.q = q;
    }
    return true;
}

// ------------------------------------------------------------------- the kernel


template <typename T, int BS, bool ROWS>
__device__ inline void warm_start_levels(const Dev<T>& d, int so, int D, const int* LS, T ratio) {
    for (int l = 0; l < D; l++) {
        const int lo = LS[l], hi = LS[l + 1];
        for (int x = lo + threadIdx.x; x < hi; x += BS) {
            int kind, s;
            VElem<T> E;
            if (ROWS) {
                E = d.ve[so + x];
                kind = E.meta >> 28;
                s = E.a;
            } else {
                s = d.seq[so + d.sched[so + x]];
                kind = s >= 0x40000000 ? VK_INACTIVE : (s >= 0 ? VK_CONTACT : VK_POINT);
                if (s < 0) s = -1 - s;
            }
            if (kind == VK_INACTIVE) continue;
            if (kind == VK_CONTACT) {
                if (ROWS) warm_start_elem(d, E, ratio);
                else warm_start_con(d, s, ratio);
            } else {
                PointCon<T>& pc = d.pc[s];
                if (pc.enabled && !pc.sleeping && pc.active) {
                    pc.lambda = pc.lambda * ratio;
                    point_apply_vw(d, pc, pc.lambda);
                }
            }
        }
        __syncthreads();
    }
}

// creation-order key of sequence element k: contacts in the order they were created
// (ascending (querying body, other body), then the manifolds of one pair in collector order),
// then point constraints by index; the warm start runs in this order
// (src/update.ts:168-175 warm-starts the constraint arrays front to back)
template <typename T>
struct CreationLess {
    const Dev<T>* d;
    int so;
    __device__ void key(int k, unsigned long long& hi, unsigned long long& lo) const {
        // bodies are below 2^23, a run holds at most 32 constraints (co carries the place in bits 24+)
        int s = d->seq[so + k];
        if (s >= 0) {
            const unsigned co = (unsigned)d->con[s].co;
            hi = ((((unsigned long long)(unsigned)d->con[s].cq << 24) | (co & 0xffffffu)) << 5) | (co >> 24);
        } else {
            hi = (1ull << 63) | (unsigned)(-1 - s);
        }
        lo = 0;
    }
};
// solve-order key of sequence element k: contacts first, ordered by (sortKey, idA, idB)
// (sortContactIndices :1240-1264), then point constraints by index (sortConstraintIds
// constraints.ts:248-262, priority 0)
struct SolveLess {
    const unsigned long long* K;
    __device__ void key(int k, unsigned long long& hi, unsigned long long& lo) const {
        hi = K[2 * k];
        lo = K[2 * k + 1];
    }
};
// the same keys, read from the copy kept beside each list entry (x: the entry's position)
struct SolveLessAt {
    const unsigned long long* BK;
    __device__ void key_at(int x, unsigned long long& hi, unsigned long long& lo) const {
        hi = BK[2 * x];
        lo = BK[2 * x + 1];
    }
};
// The graph-coloured solve order (the optional solver mode): each body's list is ordered by a
// hash of the elements' solve keys (splitmix64's finaliser; unique keys, the same in every run,
// unlike element indices) instead of by the keys.  The levels of that random orientation are
// the rounds of a Jones-Plassmann colouring of the conflict graph (elements conflict when they
// share a dynamic body): an element's level is one more than the highest level among its
// neighbours of higher priority, so elements of a level share no dynamic body, and the longest
// chain is short (about 15-20 levels where crashcat's order needs 50-120).
__device__ inline unsigned long long jp_prio(unsigned long long h, unsigned long long l) {
    unsigned long long x = h * 0x9E3779B97F4A7C15ull ^ l;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}
struct ColorLessAt {
    const unsigned long long* BK;
    __device__ void key_at(int x, unsigned long long& hi, unsigned long long& lo) const {
        hi = jp_prio(BK[2 * x], BK[2 * x + 1]);
        lo = BK[2 * x];
    }
};

__device__ inline bool key_lt(unsigned long long h0, unsigned long long l0, unsigned long long h1, unsigned long long l1) {
    return h0 < h1 || (h0 == h1 && l0 < l1);
}

// key of the list entry at position x holding element k (the solve order reads the copy
// stored beside the entry)
template <typename Less>
__device__ inline void list_key(const Less& less, int x, int k, unsigned long long& h, unsigned long long& l) { less.key(k, h, l); }
__device__ inline void list_key(const SolveLessAt& less, int x, int k, unsigned long long& h, unsigned long long& l) { less.key_at(x, h, l); }
__device__ inline void list_key(const ColorLessAt& less, int x, int k, unsigned long long& h, unsigned long long& l) { less.key_at(x, h, l); }

// The same order as one 64-bit key inside the list of one body.  Creation keys are 64-bit
// already.  A solve key (hi, lo = idA << 32 | idB with idA < idB) of a contact of this body
// orders like (hi, the other body): the other body is below this body when this body is B and
// above it when this body is A.  Point constraints (hi = 2^32) follow every contact.
template <typename Less>
__device__ inline unsigned long long list_key64(const Less& less, int x, int k, int body) {
    unsigned long long h, l;
    less.key(k, h, l);
    return h;
}
__device__ inline unsigned long long list_key64(const ColorLessAt& less, int x, int k, int body) {
    unsigned long long h, l;
    less.key_at(x, h, l);
    return h;
}
__device__ inline unsigned long long list_key64(const SolveLessAt& less, int x, int k, int body) {
    unsigned long long h, l;
    less.key_at(x, h, l);
    if (h >> 32) return (1ull << 63) | (l & 0x7fffffffull);
    const unsigned a = (unsigned)(l >> 32), b = (unsigned)l;
    const unsigned other = a == (unsigned)body ? b : a;
    return (h << 31) | other;
}

// Each body's list of sequence elements in the given order -> predecessor / successor on that
// body.  A list entry is k << 1 | side (side 1: the body is the element's B body).  Keys are
// unique, so the order is the same whichever way it is sorted: lists of up to eight entries
// are sorted in registers by a 19-comparator network, longer ones by insertion sort.
template <typename T, int BS, typename Less>
__device__ inline void link_body_lists(const Dev<T>& d, int b0, int bb, int nb, int so, int* BL, unsigned long long* BK,
                                       const int* BC, Less less) {
    for (int i = threadIdx.x; i < nb; i += BS) {
        const int o = d.boff[bb + i], c = BC[i];
        auto link = [&](int ent, int pv, int nx) {
            const int k = ent >> 1;
            if (ent & 1) { d.prevB[so + k] = pv; d.nextB[so + k] = nx; }
            else { d.prevA[so + k] = pv; d.nextA[so + k] = nx; }
        };
        if (c <= 16) {
            int v[16];
            unsigned long long kk[16];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                v[j] = -1;
                kk[j] = ~0ull;
                if (j < c) {
                    v[j] = BL[o + j];
                    kk[j] = list_key64(less, o + j, v[j] >> 1, b0 + i);
                }
            }
#define CC_CX(p, q)                                      \
    {                                                    \
        const bool sw = kk[q] < kk[p];                   \
        const unsigned long long tk = kk[p];             \
        const int tv = v[p];                             \
        kk[p] = sw ? kk[q] : kk[p]; v[p] = sw ? v[q] : v[p]; \
        kk[q] = sw ? tk : kk[q]; v[q] = sw ? tv : v[q];  \
    }
            if (c <= 8) {  // 19 comparators
                CC_CX(0, 2) CC_CX(1, 3) CC_CX(4, 6) CC_CX(5, 7)
                CC_CX(0, 4) CC_CX(1, 5) CC_CX(2, 6) CC_CX(3, 7)
                CC_CX(0, 1) CC_CX(2, 3) CC_CX(4, 5) CC_CX(6, 7)
                CC_CX(2, 4) CC_CX(3, 5)
                CC_CX(1, 4) CC_CX(3, 6)
                CC_CX(1, 2) CC_CX(3, 4) CC_CX(5, 6)
            } else {  // Batcher's odd-even merge sort, 63 comparators
                CC_CX(0, 1) CC_CX(2, 3) CC_CX(4, 5) CC_CX(6, 7) CC_CX(8, 9) CC_CX(10, 11) CC_CX(12, 13) CC_CX(14, 15) CC_CX(0, 2) CC_CX(1, 3) CC_CX(4, 6) CC_CX(5, 7) CC_CX(8, 10) CC_CX(9, 11) CC_CX(12, 14) CC_CX(13, 15) CC_CX(1, 2) CC_CX(5, 6) CC_CX(9, 10) CC_CX(13, 14) CC_CX(0, 4) CC_CX(1, 5) CC_CX(2, 6) CC_CX(3, 7) CC_CX(8, 12) CC_CX(9, 13) CC_CX(10, 14) CC_CX(11, 15) CC_CX(2, 4) CC_CX(3, 5) CC_CX(10, 12) CC_CX(11, 13) CC_CX(1, 2) CC_CX(3, 4) CC_CX(5, 6) CC_CX(9, 10) CC_CX(11, 12) CC_CX(13, 14) CC_CX(0, 8) CC_CX(1, 9) CC_CX(2, 10) CC_CX(3, 11) CC_CX(4, 12) CC_CX(5, 13) CC_CX(6, 14) CC_CX(7, 15) CC_CX(4, 8) CC_CX(5, 9) CC_CX(6, 10) CC_CX(7, 11) CC_CX(2, 4) CC_CX(3, 5) CC_CX(6, 8) CC_CX(7, 9) CC_CX(10, 12) CC_CX(11, 13) CC_CX(1, 2) CC_CX(3, 4) CC_CX(5, 6) CC_CX(7, 8) CC_CX(9, 10) CC_CX(11, 12) CC_CX(13, 14)
            }
#undef CC_CX
#pragma unroll
            for (int j = 0; j < 16; j++)
                if (j < c) link(v[j], j > 0 ? v[j - 1] >> 1 : -1, j + 1 < c ? v[j + 1] >> 1 : -1);
        } else {
            for (int x = o + 1; x < o + c; x++) {  // insertion sort; the stored keys move with the entries
                const int ent = BL[x];
                const unsigned long long bh = BK[2 * x], bl = BK[2 * x + 1];
                unsigned long long h, l;
                list_key(less, x, ent >> 1, h, l);
                int y = x - 1;
                while (y >= o) {
                    unsigned long long hy, ly;
                    list_key(less, y, BL[y] >> 1, hy, ly);
                    if (!key_lt(h, l, hy, ly)) break;
                    BL[y + 1] = BL[y];
                    BK[2 * (y + 1)] = BK[2 * y];
                    BK[2 * (y + 1) + 1] = BK[2 * y + 1];
                    y--;
                }
                BL[y + 1] = ent;
                BK[2 * (y + 1)] = bh;
                BK[2 * (y + 1) + 1] = bl;
            }
            for (int x = o; x < o + c; x++)
                link(BL[x], x > o ? BL[x - 1] >> 1 : -1, x + 1 < o + c ? BL[x + 1] >> 1 : -1);
        }
    }
    __syncthreads();
}

// Kahn's algorithm on the predecessor DAG: level l holds the elements whose longest predecessor
// chain has length l, so elements of one level share no dynamic body; returns the level count
// When DEG is set (LDS), the in-degrees (at most 2) are packed 2 bits each, 16 per word; a
// decrement never borrows from a neighbouring field because no field is decremented below 0.
template <typename T, int BS>
__device__ inline int build_levels(const Dev<T>& d, int so, int L, int* LS, int& sTail, unsigned* DEG) {
    const int tid = threadIdx.x;
    if (tid == 0) sTail = 0;
    __syncthreads();
    if (DEG) {
        for (int wd = tid; wd < (L + 15) >> 4; wd += BS) {
            unsigned word = 0;
            for (int j = 0; j < 16; j++) {
                const int k = (wd << 4) + j;
                if (k >= L) break;
                const int deg = (d.prevA[so + k] >= 0 ? 1 : 0) + (d.prevB[so + k] >= 0 ? 1 : 0);
                word |= (unsigned)deg << (2 * j);
                if (deg == 0) d.sched[so + atomicAdd(&sTail, 1)] = k;
            }
            DEG[wd] = word;
        }
    } else {
        for (int k = tid; k < L; k += BS) {
            int deg = (d.prevA[so + k] >= 0 ? 1 : 0) + (d.prevB[so + k] >= 0 ? 1 : 0);
            d.level[so + k] = deg;
            if (deg == 0) d.sched[so + atomicAdd(&sTail, 1)] = k;
        }
    }
    __syncthreads();
    auto ready = [&](int e) -> bool {  // one predecessor of e done: was it the last?
        if (DEG) {
            const unsigned sh = 2u * (unsigned)(e & 15);
            const unsigned old = atomicSub(&DEG[e >> 4], 1u << sh);
            return ((old >> sh) & 3u) == 1u;
        }
        return atomicSub(&d.level[so + e], 1) == 1;
    };
    int lo = 0, hi = sTail;
    int D = 0;
    __syncthreads();  // every lane has read the first frontier before any lane extends it
    while (lo < hi) {
        if (tid == 0) LS[D] = lo;
        for (int x = lo + tid; x < hi; x += BS) {
            int e = d.sched[so + x];
            int na = d.nextA[so + e], nbx = d.nextB[so + e];
            if (na >= 0 && ready(na)) d.sched[so + atomicAdd(&sTail, 1)] = na;
            if (nbx >= 0 && ready(nbx)) d.sched[so + atomicAdd(&sTail, 1)] = nbx;
        }
        __syncthreads();
        lo = hi;
        hi = sTail;
        D++;
        __syncthreads();
    }
    if (tid == 0) LS[D] = lo;
    __syncthreads();
    return D;
}

// The velocity residual of the solved constraints: for every contact point, the change a further
// sweep would make to its normal impulse and, after the friction-disc projection, to its friction
// impulse, from the final velocities and this step's impulses (in the cache records).
template <typename T, int BS>
__device__ inline void velocity_residual(const Dev<T>& d, int so, int L, const CacheRec<T>* rec) {
    double sD = 0.0, sL = 0.0, mx = 0.0, np = 0.0;
    for (int x = threadIdx.x; x < L; x += BS) {
        const VElem<T> E = d.ve[so + x];
        if ((E.meta >> 28) != VK_CONTACT || (E.iters & 0xffff) == 0) continue;
        const int meta = E.meta, dyn = meta & 3, mov = (meta >> 2) & 3, n_ = (meta >> 4) & 7;
        if (!dyn) continue;
        V3<T> vA, wA, vB, wB;
        vw_load(d, E.a, vA, wA);
        vw_load(d, E.b, vB, wB);
        const V3<T> n = E.n, t1 = contact_t1(n), t2 = cross(n, t1);
        const CacheRec<T>& r = rec[E.slot];
        for (int j = 0; j < n_; j++) {
            const VPoint<T> p = d.vp[E.pt + j];
            const T ln = r.ln[j], lt1 = r.lt1[j], lt2 = r.lt2[j];
            const T tn = tmax<T>(T(0), ln + p.effN * (axis_jv(dyn, mov, p.r1, p.r2, vA, wA, vB, wB, n) - p.biasN));
            T l1 = lt1, l2 = lt2;
            if (p.effT1 != T(0) || p.effT2 != T(0)) {
                l1 = lt1 + p.effT1 * axis_jv(dyn, mov, p.r1, p.r2, vA, wA, vB, wB, t1);
                l2 = lt2 + p.effT2 * axis_jv(dyn, mov, p.r1, p.r2, vA, wA, vB, wB, t2);
                const T mf = E.fr * ln, fm2 = l1 * l1 + l2 * l2;
                if (fm2 > mf * mf) {
                    const T sc = mf / sqrt(fm2);
                    l1 *= sc;
                    l2 *= sc;
                }
            }
            const double dl = fabs((double)tn - (double)ln) + sqrt(((double)l1 - lt1) * ((double)l1 - lt1) +
                                                                   ((double)l2 - lt2) * ((double)l2 - lt2));
            sD += dl;
            sL += (double)ln;
            mx = fmax(mx, dl);
            np += 1.0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        sD += __shfl_xor(sD, o, 64);
        sL += __shfl_xor(sL, o, 64);
        mx = fmax(mx, __shfl_xor(mx, o, 64));
        np += __shfl_xor(np, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && np > 0.0) {
        atomicAdd(&d.resid[0], sD);
        atomicAdd(&d.resid[1], sL);
        atomicMax(reinterpret_cast<unsigned long long*>(&d.resid[2]), (unsigned long long)__double_as_longlong(mx));
        atomicAdd(&d.resid[3], np);
    }
}

constexpr int LS_LDS = 4096;  // level starts kept in LDS (16 KB)

// One dynamic body's or point constraint's overrides (velocity | position << 8) into its
// island's record: the two maxima and a flag per kind for "uses the default" (override 0).
__device__ inline void isl_add_ovr(int* r, int o) {
    const int v = o & 255, p = (o >> 8) & 255;
    atomicMax(&r[0], v);
    atomicMax(&r[1], p);
    const int f = (v == 0 ? 1 : 0) | (p == 0 ? 2 : 0);
    if (f) atomicOr(&r[2], f);
}
// the island's velocity | position << 16 iteration counts (islands.ts:344-386)
template <typename T>
__device__ inline int isl_iters(const Dev<T>& d, int label) {
    if (!d.islIt || label < 0) return d.s.velIters | (d.s.posIters << 16);
    const int* r = d.islIt + 3 * (size_t)label;
    int v = r[0], p = r[1];
    const int f = r[2];
    if (f & 1) v = max(v, d.s.velIters);
    if (f & 2) p = max(p, d.s.posIters);
    return v | (p << 16);
}

// velocity integration of one body (update.ts:896-976); with `ccd`, a dynamic LINEAR_CAST body
// whose displacement exceeds its threshold is only rotated (setQuaternion) and its record is
// returned for the CCD pass instead of moving it (:932-967)
template <typename T>
__device__ inline bool integrate_body(const Dev<T>& d, int i, T dt, bool ccd, CcdRec<T>& rec) {
    const Bodies<T>& B = d.B;
    int f = B.flags[i];
    int mo = fl_motion(f);
    V3<T> v = B.v[i], wv = B.w[i];
    if (mo == MOTION_DYNAMIC) {
        T ml = B.maxLinV[i], ma = B.maxAngV[i];
        if (len2(v) > ml * ml) { v = normalize(v) * ml; B.v[i] = v; }
        if (len2(wv) > ma * ma) { wv = normalize(wv) * ma; B.w[i] = wv; }
    }
    V3<T> rv = wv * dt;
    T ang = vlen(rv);
    Q4<T> q = B.q[i];
    if (ang > T(1e-6)) {
        V3<T> axis = rv * (T(1) / ang);
        q = qnormalize(qmul(qaxis_angle(axis, ang), q));
        B.q[i] = q;
    }
    const V3<T> disp = v * dt;
    if (ccd && (f & FL_LINEAR_CAST) && mo == MOTION_DYNAMIC) {
        const T inner = d.ccdInner[i];
        const T thr = d.s.ccdThr * inner;
        if (len2(disp) > thr * thr) {
            rec.body = i;
            rec.hit = -1;
            rec.sub1 = 0;
            rec.sub2 = 0;
            rec.delta = disp;
            rec.normal = mk3<T>(0, 0, 0);
            rec.point = mk3<T>(0, 0, 0);
            rec.fraction = T(1);
            rec.fps = T(1);
            rec.thrSq = thr * thr;
            rec.maxPen = tmin<T>(d.s.slop, d.s.ccdMaxPen * inner);
            rec.friction = T(0);
            rec.restitution = T(0);
            set_transform(B, d.st, i);
            return true;
        }
    }
    B.com[i] = B.com[i] + disp;
    pos_from_com(B, d.st, i);
    set_transform(B, d.st, i);
    return false;
}

// TAPE: differentiable stepping records the staged constraints' sweeps for the adjoint;
// CCD: the world may hold LINEAR_CAST bodies (the CCD pass after integration)
template <typename T, int BS, bool TAPE = false, bool CCD = false>
__global__ void __launch_bounds__(BS) k_world(Dev<T> d, T dt) {
    __shared__ int sh[BS / 64 + 1];
    __shared__ int sLS[LS_LDS];
    extern __shared__ int dynLds[];  // d.worldLdsInts ints: per-body counts, then packed in-degrees
    __shared__ int sTail;
    const int4 vt = d.vTab[blockIdx.x];
    const int w = vt.x, grp = vt.y, G = vt.z;
    const int tid = threadIdx.x;
    const Bodies<T>& B = d.B;
    const int b0 = d.wStart[w], b1 = d.wStart[w + 1];
    const int nb = b1 - b0;
    const int bb = G > 1 ? vt.w : b0;  // this group's per-body scratch (list offsets, counts)
    const int c0 = d.wConStart[w], c1 = d.wConStart[w + 1];
    const int u0 = d.pcStart[w], u1 = d.pcStart[w + 1];
    // a split world's group takes the constraints of its islands, packed after the groups before it
    int m = c1 - c0, u = u1 - u0;
    int so = c0 + u0;
    const int vid = d.worldVBase[w] + grp;
    if (G > 1) {
        for (int g2 = d.worldVBase[w]; g2 < vid; g2++) so += d.vCnt[2 * g2] + d.vCnt[2 * g2 + 1];
        m = d.vCnt[2 * vid];
        u = d.vCnt[2 * vid + 1];
    }
    const int L = m + u;
    auto mine = [&](int i) { return G == 1 || island_group(B.label[i], G) == grp; };
    const int par = *d.parity;
    const T prevDt = *d.prevDt;
    const T ratio = (d.s.warmStarting && prevDt > T(0)) ? dt / prevDt : T(0);
    unsigned long long* K = d.skeys + 2 * (size_t)so;

    // ---- 1. sequence: the world's contact constraints, then its point constraints (a split
    // world's group: those of its islands, in the same order)
    if (G > 1) {
        int run = 0;
        for (int base = 0; base < c1 - c0; base += BS) {
            const int k = base + tid;
            int s = -1;
            bool f = false;
            if (k < c1 - c0) {
                s = d.conIdx[c0 + k];
                f = island_group(B.label[con_island_body(d.conHead[s])], G) == grp;
            }
            int tot;
            const int r = block_excl_scan<BS>(f ? 1 : 0, sh, tot);
            if (f) d.seq[so + run + r] = s;
            run += tot;
            __syncthreads();
        }
        run = 0;
        for (int base = 0; base < u1 - u0; base += BS) {
            const int k = base + tid;
            bool f = false;
            if (k < u1 - u0) {
                const PointCon<T>& pc = d.pc[u0 + k];
                f = pc_active(d, pc) && island_group(B.label[B.label[pc.a] >= 0 ? pc.a : pc.b], G) == grp;
            }
            int tot;
            const int r = block_excl_scan<BS>(f ? 1 : 0, sh, tot);
            if (f) d.seq[so + m + run + r] = -1 - (u0 + k);
            run += tot;
            __syncthreads();
        }
    }
    for (int k = tid; k < L; k += BS) {
        int s, ea = -1, eb = -1, npk = 0;
        unsigned long long hi, lo;
        if (k < m) {
            s = G > 1 ? d.seq[so + k] : d.conIdx[c0 + k];
            const ConHead c = d.conHead[s];
            npk = (c.bits >> 4) & 7;
            if (c.bits & 1) ea = c.a;
            if (c.bits & 2) eb = c.b;
            hi = (unsigned long long)c.sortKey;
            lo = ((unsigned long long)(unsigned)c.a << 32) | (unsigned)c.b;
        } else {
            int pi = G > 1 ? -1 - d.seq[so + k] : u0 + (k - m);
            s = -1 - pi;
            const PointCon<T>& pc = d.pc[pi];
            bool act = pc.enabled && (B.label[pc.a] >= 0 || B.label[pc.b] >= 0);
            if (act) {
                if (fl_motion(B.flags[pc.a]) == MOTION_DYNAMIC) ea = pc.a;
                if (fl_motion(B.flags[pc.b]) == MOTION_DYNAMIC) eb = pc.b;
            } else {
                s = 0x40000000;  // inactive: skipped by every phase
            }
            hi = 1ull << 32;
            lo = (unsigned long long)pi;
        }
        K[2 * k] = hi;
        K[2 * k + 1] = lo;
        d.elemNp[so + k] = npk;
        d.seq[so + k] = s;
        d.seqA[so + k] = ea;
        d.seqB[so + k] = eb;
        d.prevA[so + k] = -1;
        d.prevB[so + k] = -1;
        d.nextA[so + k] = -1;
        d.nextB[so + k] = -1;
    }
    // per-body counts in LDS when the world's bodies fit, then the packed in-degrees
    const bool cntLds = nb <= d.worldLdsInts;
    int* BC = cntLds ? dynLds : d.bcnt + bb;
    const int degOff = cntLds ? ((nb + 3) & ~3) : 0;
    unsigned* DEG = degOff + ((L + 15) >> 4) <= d.worldLdsInts ? (unsigned*)(dynLds + degOff) : nullptr;
    for (int i = tid; i < nb; i += BS) BC[i] = 0;
    __syncthreads();
    // ---- 2. per-body lists of sequence elements (counting sort by dynamic body); each entry
    // keeps a copy of its element's solve-order key
    for (int k = tid; k < L; k += BS) {
        int ea = d.seqA[so + k], eb = d.seqB[so + k];
        if (ea >= 0) atomicAdd(&BC[ea - b0], 1);
        if (eb >= 0) atomicAdd(&BC[eb - b0], 1);
    }
    __syncthreads();
    for (int i = tid; i < nb; i += BS) d.boff[bb + i] = BC[i];
    __syncthreads();
    block_scan_array<BS>(d.boff + bb, nb, sh);
    for (int i = tid; i < nb; i += BS) BC[i] = 0;
    __syncthreads();
    int* BL = d.blist + 2 * (size_t)so;
    unsigned long long* BK = d.blkey + 4 * (size_t)so;
    for (int k = tid; k < L; k += BS) {
        int ea = d.seqA[so + k], eb = d.seqB[so + k];
        const unsigned long long kh = K[2 * k], kl = K[2 * k + 1];
        if (ea >= 0) {
            const int x = d.boff[bb + ea - b0] + atomicAdd(&BC[ea - b0], 1);
            BL[x] = k << 1;
            BK[2 * x] = kh;
            BK[2 * x + 1] = kl;
        }
        if (eb >= 0) {
            const int x = d.boff[bb + eb - b0] + atomicAdd(&BC[eb - b0], 1);
            BL[x] = (k << 1) | 1;
            BK[2 * x] = kh;
            BK[2 * x + 1] = kl;
        }
    }
    __syncthreads();
    // ---- 3-5. warm start (contact-constraints.ts:906-955, constraints.ts:264-301).  The FP64
    // path applies it in creation order, as the reference does, so that every body sums its
    // warm-start impulses in the same order; the FP32 path uses the solve order, which
    // changes only the rounding of those sums.
    int* LS = d.lvlStart + so + vid;
    for (int k = tid; k < u1 - u0; k += BS) {
        PointCon<T>& pc = d.pc[u0 + k];
        if (G > 1 && island_group(pc_active(d, pc) ? B.label[B.label[pc.a] >= 0 ? pc.a : pc.b] : -1, G) != grp) continue;
        pc.sleeping = B.sleeping[pc.a] && B.sleeping[pc.b];
        if (pc.enabled && !pc.sleeping) point_props(d, pc);
    }
    // island iteration counts from the overrides of its dynamic bodies and point constraints
    if (d.islIt) {
        for (int i = b0 + tid; i < b1; i += BS)
            if (mine(i) && B.label[i] == i) {
                int* r = d.islIt + 3 * (size_t)i;
                r[0] = 0; r[1] = 0; r[2] = 0;
            }
        __syncthreads();
        for (int i = b0 + tid; i < b1; i += BS) {
            const int l = B.label[i];
            if (l < 0 || !mine(i) || fl_motion(B.flags[i]) != MOTION_DYNAMIC) continue;
            isl_add_ovr(d.islIt + 3 * (size_t)l, B.ovr[i]);
        }
        for (int k = tid; k < u1 - u0; k += BS) {
            const PointCon<T>& pc = d.pc[u0 + k];
            if (!pc_active(d, pc)) continue;
            const int l = B.label[B.label[pc.a] >= 0 ? pc.a : pc.b];
            if (G > 1 && island_group(l, G) != grp) continue;
            isl_add_ovr(d.islIt + 3 * (size_t)l, pc.ovr);
        }
        __syncthreads();
    }
    // the velocity records: the inverse inertia is that of the constraint set-up (same pose)
    for (int i = b0 + tid; i < b1; i += BS) {
        if (!mine(i)) continue;
        VBody<T> r;
        r.v = B.v[i];
        r.w = B.w[i];
        for (int j = 0; j < 9; j++) r.I.m[j] = T(0);
        const int f = B.flags[i];
        if (fl_motion(f) == MOTION_DYNAMIC) r.I = inv_inertia(B.invI[i], B.iRot[i], B.q[i], fl_dof(f));
        r.invMass = B.invMass[i];
        d.vb[i] = r;
    }
    __syncthreads();
    constexpr bool kCreationOrderWarmStart = sizeof(T) == 8;
    int D = 0;
    const bool colored = d.colored && !TAPE;
    if (colored) {
        link_body_lists<T, BS>(d, b0, bb, nb, so, BL, BK, BC, ColorLessAt{BK});
    } else {
        if (kCreationOrderWarmStart && d.s.warmStarting) {
            link_body_lists<T, BS>(d, b0, bb, nb, so, BL, BK, BC, CreationLess<T>{&d, so});
            D = build_levels<T, BS>(d, so, L, LS, sTail, DEG);
            warm_start_levels<T, BS, false>(d, so, D, LS, ratio);
        }
        link_body_lists<T, BS>(d, b0, bb, nb, so, BL, BK, BC, SolveLessAt{BK});
    }
    D = build_levels<T, BS>(d, so, L, LS, sTail, DEG);
    if (tid == 0) {
        if (G > 1) atomicMax(&d.stats[2 * w], D);
        else d.stats[2 * w] = D;
    }
    // level starts in LDS when they fit
    const bool lsLds = D + 1 <= LS_LDS;
    if (lsLds)
        for (int l = tid; l <= D; l += BS) sLS[l] = LS[l];
    // ---- stage the constraints in level order; a constraint's points follow each other
    // (dense over the level order)
    int totPts = 0;
    int maxVel = d.s.velIters, maxPos = d.s.posIters;
    {
        int mv = 0, mp = 0;
        for (int x = tid; x < L; x += BS) {  // slot and point count of each element, in level order
            const int k = d.sched[so + x];
            d.svSlot[so + x] = d.seq[so + k];
            d.svPt[so + x] = d.elemNp[so + k];
        }
        __syncthreads();
        totPts = block_scan_array<BS>(d.svPt + so, L, sh);
        for (int x = tid; x < L; x += BS) {
            const int e = so + x;
            const int s = d.svSlot[e];
            const int pbase = 4 * so + d.svPt[e];
            VElem<T> E;
            E.pt = pbase;
            E.slot = s;
            E.fr = T(0);
            E.n = mk3<T>(0, 0, 0);
            E.iters = 0;
            if (s >= 0x40000000) {
                E.a = E.b = 0;
                E.meta = VK_INACTIVE << 28;
                d.ve[e] = E;
                continue;
            }
            if (s < 0) {
                E.a = -1 - s;
                E.b = 0;
                E.meta = VK_POINT << 28;
                const PointCon<T>& pc = d.pc[E.a];
                E.iters = isl_iters(d, B.label[B.label[pc.a] >= 0 ? pc.a : pc.b]);
                mv = max(mv, E.iters & 0xffff);
                mp = max(mp, E.iters >> 16);
                d.ve[e] = E;
                continue;
            }
            const Con<T>& k = d.con[s];
            if (d.diff) {  // pair-order rank: the constraint's place in the world's sequence (G == 1)
                const int rank = c0 + d.sched[so + x];
                d.rec[par][s].rank = rank;
                if (TAPE) {
                    d.tapeRank[e] = rank;
                    for (int j = 0; j < k.np; j++) d.tapeSrc[(size_t)pbase + j] = d.tap[s].src[j];
                }
            }
            E.iters = isl_iters(d, B.label[k.a] >= 0 ? B.label[k.a] : B.label[k.b]);
            mv = max(mv, E.iters & 0xffff);
            mp = max(mp, E.iters >> 16);
            E.a = k.a;
            E.b = k.b;
            E.meta = k.dynMask | (k.movMask << 2) | (k.np << 4) | (fl_dof(B.flags[k.a]) << 8) | (fl_dof(B.flags[k.b]) << 14);
            E.n = k.normal;
            E.fr = k.friction;
            d.ve[e] = E;
            for (int j = 0; j < k.np; j++) {
                const size_t q = (size_t)pbase + j;
                const ConPoint<T>& c = k.p[j];
                VPoint<T> P;
                P.r1 = c.r1; P.r2 = c.r2;
                P.effN = c.effN; P.effT1 = c.effT1; P.effT2 = c.effT2; P.biasN = c.biasN;
                P.lN = c.lN; P.lT1 = c.lT1; P.lT2 = c.lT2;
                d.vp[q] = P;
                d.svLA[q] = c.localA;
                d.svLB[q] = c.localB;
            }
        }
        __syncthreads();
        if (d.islIt) {  // the sweeps run to the largest island count; each element stops at its own
            maxVel = block_reduce_max<BS>(mv, sh);
            maxPos = block_reduce_max<BS>(mp, sh);
        }
    }
    const int* LSp = lsLds ? sLS : LS;
    if ((!kCreationOrderWarmStart || colored) && d.s.warmStarting) warm_start_levels<T, BS, true>(d, so, D, LSp, ratio);
    if (TAPE) {  // the warm-started impulses, sweep -1 of the record
        __syncthreads();
        for (int x = tid; x < L; x += BS) {
            const VElem<T>& E = d.ve[so + x];
            if ((E.meta >> 28) != VK_CONTACT) continue;
            const int np = (E.meta >> 4) & 7;
            for (int j = 0; j < np; j++) {
                const VPoint<T>& P = d.vp[(size_t)E.pt + j];
                LamRec<T> r;
                r.n = P.lN;
                r.t1 = P.lT1;
                r.t2 = P.lT2;
                r.clamp = T(0);
                d.tapeLam[(size_t)E.pt + j] = r;
            }
        }
        __syncthreads();
    }
    // ---- 6. velocity iterations; the last one also stores the impulses (step 7)
    for (int it = 0; it < maxVel; it++) {
        for (int l = 0; l < D; l++) {
            const int lo = LSp[l], hi = LSp[l + 1];
            for (int x = lo + tid; x < hi; x += BS) {
                const VElem<T> E = d.ve[so + x];
                const int kind = E.meta >> 28;
                const int nv = E.iters & 0xffff;
                if (it >= nv) continue;
                CacheRec<T>* rec = it == nv - 1 ? d.rec[par] : nullptr;
                if (TAPE && kind == VK_CONTACT) solve_velocity_elem<T, true>(d, E, rec, d.tapeLam + (size_t)(it + 1) * d.ptsCap);
                else if (kind == VK_CONTACT) solve_velocity_elem(d, E, rec);
                else if (kind == VK_POINT) point_solve_velocity_vw(d, d.pc[E.a]);
            }
            __syncthreads();
        }
    }
    if (d.resid && !TAPE) velocity_residual<T, BS>(d, so, L, d.rec[par]);
    for (int i = b0 + tid; i < b1; i += BS) {
        if (!mine(i)) continue;
        V3<T> v, w;
        vw_load(d, i, v, w);
        B.v[i] = v;
        B.w[i] = w;
    }
    __syncthreads();
    // ---- 7. store impulses into this step's cache records (storeAppliedImpulses :1061-1076)
    // of the islands that ran no velocity iteration to do it
    for (int x = tid; x < (d.s.velIters > 0 && !d.islIt ? 0 : L); x += BS) {
        const int e = so + x;
        const int s = d.svSlot[e];
        if (s < 0 || s >= 0x40000000) continue;
        CacheRec<T>& r = d.rec[par][s];
        const VElem<T>& E = d.ve[e];
        if ((E.iters & 0xffff) != 0) continue;
        const int np = (E.meta >> 4) & 7;
        for (int j = 0; j < np; j++) {
            const VPoint<T>& P = d.vp[(size_t)E.pt + j];
            r.ln[j] = P.lN;
            r.lt1[j] = P.lT1;
            r.lt2[j] = P.lT2;
        }
    }
    // ---- 8. velocity integration (update.ts:896-976)
    bool ccdWorld = false;
    if constexpr (CCD) ccdWorld = d.ccdStart[w + 1] > d.ccdStart[w];
    if (!ccdWorld) {
        for (int i = b0 + tid; i < b1; i += BS) {
            int mo = fl_motion(B.flags[i]);
            if (mo == MOTION_STATIC || B.sleeping[i] || !mine(i)) continue;
            CcdRec<T> rec;
            integrate_body(d, i, dt, false, rec);
        }
    } else if constexpr (CCD) {
        // ---- 8b. CCD (update.ts:131-140): the records in body order (ccd.clear first, src/ccd.ts:107-123),
        // then findCCDContacts, sortCCDBodies and resolveCCDContacts
        const int cb0 = d.ccdStart[w];
        int nccd = 0;
        for (int base = b0; base < b1; base += BS) {
            const int i = base + tid;
            bool reg = false;
            CcdRec<T> rec;
            if (i < b1) {
                const int mo = fl_motion(B.flags[i]);
                if (mo != MOTION_STATIC && !B.sleeping[i]) reg = integrate_body(d, i, dt, true, rec);
            }
            int tot;
            const int r = block_excl_scan<BS>(reg ? 1 : 0, sh, tot);
            if (reg) {
                d.ccd[cb0 + nccd + r] = rec;
                d.ccdIdx[i] = nccd + r;
            } else if (i < b1) {
                d.ccdIdx[i] = -1;
            }
            nccd += tot;
        }
        if (tid == 0) d.ccdCount[w] = nccd;
        __syncthreads();
        if (nccd > 0) {
            using C = typename std::conditional<sizeof(T) == 4, double, T>::type;
            for (int e = tid >> 6; e < nccd; e += BS / 64) {  // one record per wavefront
                if (d.npWide) ccd_find<T, C>(d, d.ccd, d.ccdIdx, cb0, e, b0, b1, dt);
                else ccd_find<T, T>(d, d.ccd, d.ccdIdx, cb0, e, b0, b1, dt);
            }
            __syncthreads();
            ccd_resolve<T, BS>(d, d.ccd, d.ccdOrd, d.ccdIdx, d.ccdPart + 2 * (size_t)w, cb0, nccd);
        }
    }
    for (int i = b0 + tid; i < b1; i += BS) {
        if (!mine(i)) continue;
        d.isl0[i] = 1;
        d.isl1[i] = 0;
        PBody<T> r;
        r.com = B.com[i];
        r.q = B.q[i];
        r.invI = B.invI[i];
        r.iRot = B.iRot[i];
        r.invMass = B.invMass[i];
        r.label = B.label[i];
        d.pb[i] = r;
    }
    __syncthreads();
    // ---- 9. position iterations with the per-island early out (update.ts:187-223)
    int sweeps = 0;
    for (int it = 0; it < maxPos; it++) {
        int anyActive = 0;
        for (int l = 0; l < D; l++) {
            const int lo = LSp[l], hi = LSp[l + 1];
            for (int x = lo + tid; x < hi; x += BS) {
                const VElem<T> E = d.ve[so + x];
                const int kind = E.meta >> 28;
                if (kind == VK_INACTIVE || it >= (E.iters >> 16)) continue;
                const bool con = kind == VK_CONTACT;
                int ia = E.a, ib = E.b;
                if (!con) { ia = d.pc[E.a].a; ib = d.pc[E.a].b; }
                const PBody<T> PA = d.pb[ia], PB = d.pb[ib];
                V3<T> LA[4], LB[4];
                const int np = con ? (E.meta >> 4) & 7 : 0;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (j < np) { LA[j] = d.svLA[E.pt + j]; LB[j] = d.svLB[E.pt + j]; }
                const int isl = PA.label >= 0 ? PA.label : PB.label;
                if (isl < 0 || !d.isl0[isl]) continue;
                anyActive = 1;
                const bool applied = con ? solve_position_elem(d, E, PA, PB, LA, LB) : point_solve_position(d, d.pc[E.a], PA, PB);
                if (applied) d.isl1[isl] = 1;
            }
            __syncthreads();
        }
        if (__syncthreads_or(anyActive)) sweeps = it + 1;
        for (int i = b0 + tid; i < b1; i += BS)
            if (mine(i)) { d.isl0[i] = d.isl1[i]; d.isl1[i] = 0; }
        __syncthreads();
    }
    for (int i = b0 + tid; i < b1; i += BS) {
        if (fl_motion(B.flags[i]) != MOTION_DYNAMIC || !mine(i)) continue;
        const PBody<T>& r = d.pb[i];
        B.com[i] = r.com;
        B.q[i] = r.q;
    }
    if (tid == 0) {
        if (G > 1) atomicMax(&d.stats[2 * w + 1], sweeps);
        else d.stats[2 * w + 1] = sweeps;
    }
    {  // byte model: per sweep 160 B/manifold + 192 B/point, per position sweep 240 B/manifold + 24 B/point
        int awake = 0;
        for (int i = b0 + tid; i < b1; i += BS)
            awake += (fl_motion(B.flags[i]) == MOTION_DYNAMIC && !B.sleeping[i] && mine(i)) ? 1 : 0;
        for (int o = 32; o > 0; o >>= 1) awake += __shfl_xor(awake, o, 64);
        if ((tid & 63) == 0) atomicAdd(&d.acc[ACC_AWAKE], (unsigned long long)awake);
        if (tid == 0) {
            unsigned long long iv = (unsigned long long)d.s.velIters + 1;
            unsigned long long by = iv * (160ull * m + 192ull * totPts) + (unsigned long long)sweeps * (240ull * m + 24ull * totPts);
            atomicAdd(&d.acc[ACC_BYTES_SOLVE], by);
        }
    }
    // ---- 10. positions from the centre of mass + AABBs (update.ts:322-330)
    for (int i = b0 + tid; i < b1; i += BS) {
        int mo = fl_motion(B.flags[i]);
        if (mo == MOTION_STATIC || B.sleeping[i] || !mine(i)) continue;
        pos_from_com(B, d.st, i);
        set_transform(B, d.st, i);
        d.isl0[i] = 1;  // island sleep flags
    }
    __syncthreads();
    // ---- 11. island sleep (islands.ts:392-423)
    if (d.s.allowSleeping) {
        T maxMove = d.s.sleepThresh * d.s.tSleep;
        for (int i = b0 + tid; i < b1; i += BS) {
            int l = B.label[i];
            if (l < 0 || fl_motion(B.flags[i]) != MOTION_DYNAMIC || !mine(i)) continue;
            bool can = update_sleep_state(B, d.st, i, dt, maxMove, d.s.tSleep);
            if (!can) d.isl0[l] = 0;
        }
        __syncthreads();
        for (int i = b0 + tid; i < b1; i += BS) {
            int l = B.label[i];
            if (l < 0 || fl_motion(B.flags[i]) != MOTION_DYNAMIC || !mine(i)) continue;
            if (d.isl0[l]) {
                B.sleeping[i] = 1;
                B.v[i] = mk3<T>(0, 0, 0);
                B.w[i] = mk3<T>(0, 0, 0);
            }
        }
    }
    // ---- 12. clear the forces of the bodies still awake (update.ts:234-238)
    if (d.hasForces) {
        __syncthreads();
        for (int i = b0 + tid; i < b1; i += BS) {
            if (!mine(i) || fl_motion(B.flags[i]) == MOTION_STATIC || B.sleeping[i]) continue;
            B.force[i] = mk3<T>(0, 0, 0);
            B.torque[i] = mk3<T>(0, 0, 0);
        }
    }
}

template <typename T>
__global__ void k_end_step(Dev<T> d, T dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *d.prevDt = dt;
        *d.parity ^= 1;
    }
}

}  // namespace cc
