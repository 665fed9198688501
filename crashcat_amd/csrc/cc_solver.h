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
#include "cc_kernels.h"

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

// packed velocity record of body i: (vx, vy, vz, wx, wy, wz, -, -)
template <typename T>
__device__ inline void vw_load(const Dev<T>& d, int i, V3<T>& v, V3<T>& w) {
    const T* p = d.vw + 8 * (size_t)i;
    v = mk3<T>(p[0], p[1], p[2]);
    w = mk3<T>(p[3], p[4], p[5]);
}
template <typename T>
__device__ inline void vw_store(const Dev<T>& d, int i, V3<T> v, V3<T> w) {
    T* p = d.vw + 8 * (size_t)i;
    p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = w.x; p[4] = w.y; p[5] = w.z;
}

// --------------------------------------------------------------- contact rows
template <typename T>
__device__ inline void solve_velocity_contact(const Dev<T>& d, int slot) {  // :963-1053 for one constraint
    Con<T>& k = d.con[slot];
    int dyn = k.dynMask;
    if (!dyn) return;
    const Bodies<T>& B = d.B;
    int a = k.a, b = k.b;
    V3<T> vA = B.v[a], wA = B.w[a], vB = B.v[b], wB = B.w[b];
    int dofA = fl_dof(B.flags[a]), dofB = fl_dof(B.flags[b]);
    int mov = k.movMask;
    T iA = k.invMassA, iB = k.invMassB;
    int np = k.np;
    for (int j = 0; j < np; j++) {
        ConPoint<T>& cp = k.p[j];
        if (cp.t1.eff != T(0) || cp.t2.eff != T(0)) {
            T l1 = axis_total_lambda(cp.t1, mov, vA, wA, vB, wB, k.t1);
            T l2 = axis_total_lambda(cp.t2, mov, vA, wA, vB, wB, k.t2);
            T mf = k.friction * cp.n.lambda;
            T fm2 = l1 * l1 + l2 * l2;
            if (fm2 > mf * mf) { T sc = mf / sqrt(fm2); l1 *= sc; l2 *= sc; }
            axis_apply(cp.t1, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, k.t1, l1);
            axis_apply(cp.t2, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, k.t2, l2);
        }
    }
    for (int j = 0; j < np; j++) {
        ConPoint<T>& cp = k.p[j];
        T tl = axis_total_lambda(cp.n, mov, vA, wA, vB, wB, k.normal);
        axis_apply(cp.n, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, k.normal, tmax<T>(T(0), tl));
    }
    if (dyn & 1) { B.v[a] = vA; B.w[a] = wA; }
    if (dyn & 2) { B.v[b] = vB; B.w[b] = wB; }
}

template <typename T>
__device__ inline void warm_start_contact(const Dev<T>& d, int slot, T ratio) {  // :906-955 for one constraint
    Con<T>& k = d.con[slot];
    int dyn = k.dynMask;
    const Bodies<T>& B = d.B;
    int a = k.a, b = k.b;
    V3<T> vA = B.v[a], wA = B.w[a], vB = B.v[b], wB = B.w[b];
    int dofA = fl_dof(B.flags[a]), dofB = fl_dof(B.flags[b]);
    T iA = k.invMassA, iB = k.invMassB;
    for (int j = 0; j < k.np; j++) {
        ConPoint<T>& cp = k.p[j];
        if (cp.t1.eff != T(0) || cp.t2.eff != T(0)) {
            axis_warm(cp.t1, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, k.t1, ratio);
            axis_warm(cp.t2, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, k.t2, ratio);
        }
        axis_warm(cp.n, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, k.normal, ratio);
    }
    if (dyn & 1) { B.v[a] = vA; B.w[a] = wA; }
    if (dyn & 2) { B.v[b] = vB; B.w[b] = wB; }
}

template <typename T>
__device__ inline bool solve_position_contact(const Dev<T>& d, int slot) {  // :1098-1216 for one constraint
    Con<T>& k = d.con[slot];
    int dyn = k.dynMask;
    if (!dyn) return false;
    const Bodies<T>& B = d.B;
    int a = k.a, b = k.b;
    int fa = B.flags[a], fb = B.flags[b];
    V3<T> cA = B.com[a], cB = B.com[b];
    Q4<T> qA = B.q[a], qB = B.q[b];
    M3<T> RA = m3_from_quat(qA), RB = m3_from_quat(qB);
    M3<T> IA, IB;
    for (int j = 0; j < 9; j++) { IA.m[j] = T(0); IB.m[j] = T(0); }
    if (dyn & 1) IA = inv_inertia(B.invI[a], B.iRot[a], qA, fl_dof(fa));
    if (dyn & 2) IB = inv_inertia(B.invI[b], B.iRot[b], qB, fl_dof(fb));
    bool any = false;
    V3<T> n = k.normal;
    for (int j = 0; j < k.np; j++) {
        ConPoint<T>& cp = k.p[j];
        V3<T> pA = cA + mul(RA, cp.localA);
        V3<T> pB = cB + mul(RB, cp.localB);
        T sep = dot(pB - pA, n) + d.s.slop;
        sep = tmax<T>(sep, -d.s.maxPen);
        if (sep >= T(0)) continue;
        V3<T> mid = mk3<T>((pA.x + pB.x) * T(0.5), (pA.y + pB.y) * T(0.5), (pA.z + pB.z) * T(0.5));
        V3<T> rA = mid - cA, rB = mid - cB;
        axis_props(cp.n, (dyn & 1) != 0, (dyn & 2) != 0, k.invMassA, k.invMassB, IA, IB, rA, rB, n, T(0));
        // solvePositionConstraintWithMassOverride axis-constraint-part.ts:790-847
        if (sep == T(0) || cp.n.eff == T(0)) continue;
        T lam = -cp.n.eff * d.s.beta * sep;
        if (dyn & 1) {
            cA = dof_lin(cA - n * (lam * k.invMassA), fl_dof(fa));
            add_rot_step(qA, cp.n.i1 * lam, T(-1));
        }
        if (dyn & 2) {
            cB = dof_lin(cB + n * (lam * k.invMassB), fl_dof(fb));
            add_rot_step(qB, cp.n.i2 * lam, T(1));
        }
        any = true;
    }
    if (dyn & 1) { B.com[a] = cA; B.q[a] = qA; }
    if (dyn & 2) { B.com[b] = cB; B.q[b] = qB; }
    return any;
}

// ------------------------------------------------- the same three solves on the level-ordered rows
template <typename T>
__device__ inline void solve_velocity_rows(const Dev<T>& d, int e) {  // :963-1053 for one constraint
    const int meta = d.svMeta[e];
    const int dyn = meta & 3;
    if (!dyn) return;
    const Bodies<T>& B = d.B;
    const int a = d.svA[e], b = d.svB[e];
    V3<T> vA, wA, vB, wB;
    vw_load(d, a, vA, wA);
    vw_load(d, b, vB, wB);
    const int mov = (meta >> 2) & 3, np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    const T iA = d.svIMA[e], iB = d.svIMB[e], fr = d.svFr[e];
    const V3<T> n = d.svN[e], t1 = d.svT1[e], t2 = d.svT2[e];
    const size_t pb = (size_t)d.svPt[e];
    for (int j = 0; j < np; j++) {
        const size_t k = pb + j;
        Row<T> r1 = row_load(d.rT1, k), r2 = row_load(d.rT2, k);
        if (r1.eff != T(0) || r2.eff != T(0)) {
            T l1 = axis_total_lambda(r1, mov, vA, wA, vB, wB, t1);
            T l2 = axis_total_lambda(r2, mov, vA, wA, vB, wB, t2);
            T mf = fr * d.rN.lambda[k];
            T fm2 = l1 * l1 + l2 * l2;
            if (fm2 > mf * mf) { T sc = mf / sqrt(fm2); l1 *= sc; l2 *= sc; }
            axis_apply(r1, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, t1, l1);
            axis_apply(r2, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, t2, l2);
            d.rT1.lambda[k] = r1.lambda;
            d.rT2.lambda[k] = r2.lambda;
        }
    }
    for (int j = 0; j < np; j++) {
        const size_t k = pb + j;
        Row<T> rn = row_load(d.rN, k);
        T tl = axis_total_lambda(rn, mov, vA, wA, vB, wB, n);
        axis_apply(rn, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, n, tmax<T>(T(0), tl));
        d.rN.lambda[k] = rn.lambda;
    }
    if (dyn & 1) vw_store(d, a, vA, wA);
    if (dyn & 2) vw_store(d, b, vB, wB);
}

template <typename T>
__device__ inline void warm_start_rows(const Dev<T>& d, int e, T ratio) {  // :906-955 for one constraint
    const int meta = d.svMeta[e];
    const int dyn = meta & 3;
    const Bodies<T>& B = d.B;
    const int a = d.svA[e], b = d.svB[e];
    V3<T> vA, wA, vB, wB;
    vw_load(d, a, vA, wA);
    vw_load(d, b, vB, wB);
    const int np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    const T iA = d.svIMA[e], iB = d.svIMB[e];
    const V3<T> n = d.svN[e], t1 = d.svT1[e], t2 = d.svT2[e];
    const size_t pb = (size_t)d.svPt[e];
    for (int j = 0; j < np; j++) {
        const size_t k = pb + j;
        Row<T> r1 = row_load(d.rT1, k), r2 = row_load(d.rT2, k), rn = row_load(d.rN, k);
        if (r1.eff != T(0) || r2.eff != T(0)) {
            axis_warm(r1, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, t1, ratio);
            axis_warm(r2, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, t2, ratio);
            d.rT1.lambda[k] = r1.lambda;
            d.rT2.lambda[k] = r2.lambda;
        }
        axis_warm(rn, dyn, dofA, dofB, iA, iB, vA, wA, vB, wB, n, ratio);
        d.rN.lambda[k] = rn.lambda;
    }
    if (dyn & 1) vw_store(d, a, vA, wA);
    if (dyn & 2) vw_store(d, b, vB, wB);
}

template <typename T>
__device__ inline bool solve_position_rows(const Dev<T>& d, int e) {  // :1098-1216 for one constraint
    const int meta = d.svMeta[e];
    const int dyn = meta & 3;
    if (!dyn) return false;
    const Bodies<T>& B = d.B;
    const int a = d.svA[e], b = d.svB[e];
    const int np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    const T iMA = d.svIMA[e], iMB = d.svIMB[e];
    V3<T> cA = B.com[a], cB = B.com[b];
    Q4<T> qA = B.q[a], qB = B.q[b];
    M3<T> RA = m3_from_quat(qA), RB = m3_from_quat(qB);
    M3<T> IA, IB;
    for (int j = 0; j < 9; j++) { IA.m[j] = T(0); IB.m[j] = T(0); }
    if (dyn & 1) IA = inv_inertia(B.invI[a], B.iRot[a], qA, dofA);
    if (dyn & 2) IB = inv_inertia(B.invI[b], B.iRot[b], qB, dofB);
    bool any = false;
    const V3<T> n = d.svN[e];
    const size_t pb = (size_t)d.svPt[e];
    for (int j = 0; j < np; j++) {
        const size_t k = pb + j;
        V3<T> pA = cA + mul(RA, d.svLA[k]);
        V3<T> pB = cB + mul(RB, d.svLB[k]);
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
    if (dyn & 1) { B.com[a] = cA; B.q[a] = qA; }
    if (dyn & 2) { B.com[b] = cB; B.q[b] = qB; }
    return any;
}

// ---------------------------------------------------------------- point constraint
template <typename T>
__device__ inline void point_props(const Dev<T>& d, PointCon<T>& p) {  // point-constraint-part.ts:77-165
    const Bodies<T>& B = d.B;
    int a = p.a, b = p.b;
    Q4<T> qa = B.q[a], qb = B.q[b];
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

template <typename T>
__device__ inline void point_apply(const Dev<T>& d, PointCon<T>& p, V3<T> lam) {  // applyVelocityStep :232-262
    if (len2(lam) == T(0)) return;
    const Bodies<T>& B = d.B;
    if (fl_motion(B.flags[p.a]) == MOTION_DYNAMIC) {
        B.v[p.a] = B.v[p.a] + lam * (-B.invMass[p.a]);
        B.w[p.a] = B.w[p.a] - mul(p.i1X, lam);
    }
    if (fl_motion(B.flags[p.b]) == MOTION_DYNAMIC) {
        B.v[p.b] = B.v[p.b] + lam * B.invMass[p.b];
        B.w[p.b] = B.w[p.b] + mul(p.i2X, lam);
    }
}

// the same two on the packed velocity records
template <typename T>
__device__ inline void point_apply_vw(const Dev<T>& d, PointCon<T>& p, V3<T> lam) {
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
__device__ inline void point_solve_velocity_vw(const Dev<T>& d, PointCon<T>& p) {
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

template <typename T>
__device__ inline void point_solve_velocity(const Dev<T>& d, PointCon<T>& p) {  // :186-221
    if (!p.active) return;
    const Bodies<T>& B = d.B;
    V3<T> jv = B.v[p.a] - cross(p.r1, B.w[p.a]) - B.v[p.b];
    jv = jv + cross(p.r2, B.w[p.b]);
    V3<T> lam = mul(p.eff, jv);
    p.lambda = p.lambda + lam;
    point_apply(d, p, lam);
}

template <typename T>
__device__ inline bool point_solve_position(const Dev<T>& d, PointCon<T>& p) {  // point-constraint.ts:211-230, part :271-316
    point_props(d, p);
    if (!p.active) return false;
    const Bodies<T>& B = d.B;
    V3<T> sep = B.com[p.b] + p.r2 - B.com[p.a] - p.r1;
    if (len2(sep) < T(1e-12)) return false;
    V3<T> lam = mul(p.eff, sep) * (-d.s.beta);
    if (fl_motion(B.flags[p.a]) == MOTION_DYNAMIC) {
        B.com[p.a] = dof_lin(B.com[p.a] - lam * B.invMass[p.a], fl_dof(B.flags[p.a]));
        Q4<T> q = B.q[p.a];
        add_rot_step(q, mul(p.i1X, lam), T(-1));
        B.q[p.a] = q;
    }
    if (fl_motion(B.flags[p.b]) == MOTION_DYNAMIC) {
        B.com[p.b] = dof_lin(B.com[p.b] + lam * B.invMass[p.b], fl_dof(B.flags[p.b]));
        Q4<T> q = B.q[p.b];
        add_rot_step(q, mul(p.i2X, lam), T(1));
        B.q[p.b] = q;
    }
    return true;
}

// ------------------------------------------------------------------- the kernel
template <typename T>
__device__ inline int island_of(const Dev<T>& d, int a, int b) {
    int la = d.B.label[a];
    return la >= 0 ? la : d.B.label[b];
}

// solve-order key of sequence element k: contacts first, ordered by (sortKey, idA, idB)
// (sortContactIndices :1240-1264), then point constraints by index (sortConstraintIds
// constraints.ts:248-262, priority 0)
__device__ inline bool key_less(const unsigned long long* K, int x, int y) {
    unsigned long long hx = K[2 * x], hy = K[2 * y];
    if (hx != hy) return hx < hy;
    return K[2 * x + 1] < K[2 * y + 1];
}

template <typename T, int BS, bool ROWS>
__device__ inline void warm_start_levels(const Dev<T>& d, int so, int D, const int* LS, T ratio) {
    for (int l = 0; l < D; l++) {
        for (int x = LS[l] + threadIdx.x; x < LS[l + 1]; x += BS) {
            int s = ROWS ? d.svSlot[so + x] : d.seq[so + d.sched[so + x]];
            if (s >= 0x40000000) continue;
            if (s >= 0) {
                if (ROWS) warm_start_rows(d, so + x, ratio);
                else warm_start_contact(d, s, ratio);
            }
            else {
                PointCon<T>& pc = d.pc[-1 - s];
                if (pc.enabled && !pc.sleeping && pc.active) {
                    pc.lambda = pc.lambda * ratio;
                    if (ROWS) point_apply_vw(d, pc, pc.lambda);
                    else point_apply(d, pc, pc.lambda);
                }
            }
        }
        __syncthreads();
    }
}

// creation-order key of sequence element k: contacts in the order they were created
// (ascending (querying body, other body), the order of the pair list), then point constraints
// by index; the warm start runs in this order
// (src/update.ts:168-175 warm-starts the constraint arrays front to back)
template <typename T>
struct CreationLess {
    const Dev<T>* d;
    int so;
    __device__ unsigned long long key(int k) const {
        int s = d->seq[so + k];
        if (s >= 0) return ((unsigned long long)(unsigned)d->con[s].cq << 32) | (unsigned)d->con[s].co;
        return (1ull << 63) | (unsigned)(-1 - s);
    }
    __device__ bool operator()(int x, int y) const { return key(x) < key(y); }
};
struct SolveLess {
    const unsigned long long* K;
    __device__ bool operator()(int x, int y) const { return key_less(K, x, y); }
};

// each body's list of sequence elements in the given order -> predecessor / successor on that body
template <typename T, int BS, typename Less>
__device__ inline void link_body_lists(const Dev<T>& d, int b0, int nb, int so, int* BL, Less less) {
    for (int i = threadIdx.x; i < nb; i += BS) {
        int o = d.boff[b0 + i], c = d.bcnt[b0 + i];
        for (int x = o + 1; x < o + c; x++) {  // insertion sort (lists are short)
            int v = BL[x];
            int y = x - 1;
            while (y >= o && less(v, BL[y])) { BL[y + 1] = BL[y]; y--; }
            BL[y + 1] = v;
        }
        for (int x = o; x < o + c; x++) {
            int k = BL[x];
            int pv = x > o ? BL[x - 1] : -1;
            int nx = x + 1 < o + c ? BL[x + 1] : -1;
            if (d.seqA[so + k] == b0 + i) { d.prevA[so + k] = pv; d.nextA[so + k] = nx; }
            else { d.prevB[so + k] = pv; d.nextB[so + k] = nx; }
        }
    }
    __syncthreads();
}

// Kahn's algorithm on the predecessor DAG: level l holds the elements whose longest predecessor
// chain has length l, so elements of one level share no dynamic body; returns the level count
template <typename T, int BS>
__device__ inline int build_levels(const Dev<T>& d, int so, int L, int* LS, int& sTail) {
    const int tid = threadIdx.x;
    if (tid == 0) sTail = 0;
    __syncthreads();
    for (int k = tid; k < L; k += BS) {
        int deg = (d.prevA[so + k] >= 0 ? 1 : 0) + (d.prevB[so + k] >= 0 ? 1 : 0);
        d.level[so + k] = deg;
        if (deg == 0) d.sched[so + atomicAdd(&sTail, 1)] = k;
    }
    __syncthreads();
    int lo = 0, hi = sTail;
    int D = 0;
    __syncthreads();  // every lane has read the first frontier before any lane extends it
    while (lo < hi) {
        if (tid == 0) LS[D] = lo;
        for (int x = lo + tid; x < hi; x += BS) {
            int e = d.sched[so + x];
            int na = d.nextA[so + e], nbx = d.nextB[so + e];
            if (na >= 0 && atomicSub(&d.level[so + na], 1) == 1) d.sched[so + atomicAdd(&sTail, 1)] = na;
            if (nbx >= 0 && atomicSub(&d.level[so + nbx], 1) == 1) d.sched[so + atomicAdd(&sTail, 1)] = nbx;
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

template <typename T, int BS>
__global__ void __launch_bounds__(BS) k_world(Dev<T> d, T dt) {
    __shared__ int sh[BS / 64 + 1];
    __shared__ int sTail;
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const Bodies<T>& B = d.B;
    const int b0 = d.wStart[w], b1 = d.wStart[w + 1];
    const int nb = b1 - b0;
    const int c0 = d.wConStart[w], c1 = d.wConStart[w + 1];
    const int u0 = d.pcStart[w], u1 = d.pcStart[w + 1];
    const int m = c1 - c0, u = u1 - u0, L = m + u;
    const int so = c0 + u0;
    const int par = *d.parity;
    const T prevDt = *d.prevDt;
    const T ratio = (d.s.warmStarting && prevDt > T(0)) ? dt / prevDt : T(0);
    unsigned long long* K = d.skeys + 2 * (size_t)so;

    // ---- 1. sequence: the world's contact constraints, then its point constraints
    for (int k = tid; k < L; k += BS) {
        int s, ea = -1, eb = -1;
        unsigned long long hi, lo;
        if (k < m) {
            s = d.conIdx[c0 + k];
            const Con<T>& c = d.con[s];
            if (c.dynMask & 1) ea = c.a;
            if (c.dynMask & 2) eb = c.b;
            hi = (unsigned long long)c.sortKey;
            lo = ((unsigned long long)(unsigned)c.a << 32) | (unsigned)c.b;
        } else {
            int pi = u0 + (k - m);
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
        d.seq[so + k] = s;
        d.seqA[so + k] = ea;
        d.seqB[so + k] = eb;
        d.prevA[so + k] = -1;
        d.prevB[so + k] = -1;
        d.nextA[so + k] = -1;
        d.nextB[so + k] = -1;
    }
    for (int i = tid; i < nb; i += BS) d.bcnt[b0 + i] = 0;
    __syncthreads();
    // ---- 2. per-body lists of sequence elements (counting sort by dynamic body)
    for (int k = tid; k < L; k += BS) {
        int ea = d.seqA[so + k], eb = d.seqB[so + k];
        if (ea >= 0) atomicAdd(&d.bcnt[ea], 1);
        if (eb >= 0) atomicAdd(&d.bcnt[eb], 1);
    }
    __syncthreads();
    for (int i = tid; i < nb; i += BS) d.boff[b0 + i] = d.bcnt[b0 + i];
    __syncthreads();
    block_scan_array<BS>(d.boff + b0, nb, sh);
    for (int i = tid; i < nb; i += BS) d.bcnt[b0 + i] = 0;
    __syncthreads();
    int* BL = d.blist + 2 * (size_t)so;
    for (int k = tid; k < L; k += BS) {
        int ea = d.seqA[so + k], eb = d.seqB[so + k];
        if (ea >= 0) BL[d.boff[ea] + atomicAdd(&d.bcnt[ea], 1)] = k;
        if (eb >= 0) BL[d.boff[eb] + atomicAdd(&d.bcnt[eb], 1)] = k;
    }
    __syncthreads();
    // ---- 3-5. warm start (contact-constraints.ts:906-955, constraints.ts:264-301).  The FP64
    // path applies it in creation order, as the reference does, so that every body sums its
    // warm-start impulses in the same order; the FP32 path uses the solve order, which
    // changes only the rounding of those sums.
    int* LS = d.lvlStart + so + w;
    for (int k = tid; k < u; k += BS) {
        PointCon<T>& pc = d.pc[u0 + k];
        pc.sleeping = B.sleeping[pc.a] && B.sleeping[pc.b];
        if (pc.enabled && !pc.sleeping) point_props(d, pc);
    }
    __syncthreads();
    constexpr bool kCreationOrderWarmStart = sizeof(T) == 8;
    int D = 0;
    if (kCreationOrderWarmStart && d.s.warmStarting) {
        link_body_lists<T, BS>(d, b0, nb, so, BL, CreationLess<T>{&d, so});
        D = build_levels<T, BS>(d, so, L, LS, sTail);
        warm_start_levels<T, BS, false>(d, so, D, LS, ratio);
    }
    link_body_lists<T, BS>(d, b0, nb, so, BL, SolveLess{K});
    D = build_levels<T, BS>(d, so, L, LS, sTail);
    if (tid == 0) d.stats[2 * w] = D;
    // ---- stage the contact rows in level order; a constraint's points follow each other
    // (point rows of element e at svPt[e] + j, dense over the level order)
    {
        for (int x = tid; x < L; x += BS) {
            const int s = d.seq[so + d.sched[so + x]];
            d.svPt[so + x] = (s >= 0 && s < 0x40000000) ? d.con[s].np : 0;
        }
        __syncthreads();
        block_scan_array<BS>(d.svPt + so, L, sh);
        for (int x = tid; x < L; x += BS) {
            const int e = so + x;
            const int s = d.seq[so + d.sched[so + x]];
            d.svSlot[e] = s;
            const int pbase = 4 * so + d.svPt[e];
            d.svPt[e] = pbase;
            if (s < 0 || s >= 0x40000000) continue;
            const Con<T>& k = d.con[s];
            d.svA[e] = k.a;
            d.svB[e] = k.b;
            d.svMeta[e] = k.dynMask | (k.movMask << 2) | (k.np << 4) | (fl_dof(B.flags[k.a]) << 8) | (fl_dof(B.flags[k.b]) << 14);
            d.svIMA[e] = k.invMassA;
            d.svIMB[e] = k.invMassB;
            d.svFr[e] = k.friction;
            d.svN[e] = k.normal;
            d.svT1[e] = k.t1;
            d.svT2[e] = k.t2;
            for (int j = 0; j < k.np; j++) {
                const size_t q = (size_t)pbase + j;
                row_store(d.rN, q, k.p[j].n);
                row_store(d.rT1, q, k.p[j].t1);
                row_store(d.rT2, q, k.p[j].t2);
                d.svLA[q] = k.p[j].localA;
                d.svLB[q] = k.p[j].localB;
            }
        }
        __syncthreads();
    }
    for (int i = b0 + tid; i < b1; i += BS) vw_store(d, i, B.v[i], B.w[i]);
    __syncthreads();
    if (!kCreationOrderWarmStart && d.s.warmStarting) warm_start_levels<T, BS, true>(d, so, D, LS, ratio);
    // ---- 6. velocity iterations
    for (int it = 0; it < d.s.velIters; it++) {
        for (int l = 0; l < D; l++) {
            for (int x = LS[l] + tid; x < LS[l + 1]; x += BS) {
                int s = d.svSlot[so + x];
                if (s >= 0x40000000) continue;
                if (s >= 0) solve_velocity_rows(d, so + x);
                else point_solve_velocity_vw(d, d.pc[-1 - s]);
            }
            __syncthreads();
        }
    }
    for (int i = b0 + tid; i < b1; i += BS) {
        V3<T> v, w;
        vw_load(d, i, v, w);
        B.v[i] = v;
        B.w[i] = w;
    }
    __syncthreads();
    // ---- 7. store impulses into this step's cache records (storeAppliedImpulses :1061-1076)
    for (int x = tid; x < L; x += BS) {
        const int e = so + x;
        const int s = d.svSlot[e];
        if (s < 0 || s >= 0x40000000) continue;
        CacheRec<T>& r = d.rec[par][s];
        const int np = (d.svMeta[e] >> 4) & 7;
        for (int j = 0; j < np; j++) {
            const size_t q = (size_t)d.svPt[e] + j;
            r.ln[j] = d.rN.lambda[q];
            r.lt1[j] = d.rT1.lambda[q];
            r.lt2[j] = d.rT2.lambda[q];
        }
    }
    // ---- 8. velocity integration (update.ts:896-976)
    for (int i = b0 + tid; i < b1; i += BS) {
        int f = B.flags[i];
        int mo = fl_motion(f);
        if (mo == MOTION_STATIC || B.sleeping[i]) continue;
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
        B.com[i] = B.com[i] + v * dt;
        pos_from_com(B, d.st, i);
        set_transform(B, d.st, i);
    }
    for (int i = b0 + tid; i < b1; i += BS) { d.isl0[i] = 1; d.isl1[i] = 0; }
    __syncthreads();
    // ---- 9. position iterations with the per-island early out (update.ts:187-223)
    int sweeps = 0;
    for (int it = 0; it < d.s.posIters; it++) {
        int anyActive = 0;
        for (int l = 0; l < D; l++) {
            for (int x = LS[l] + tid; x < LS[l + 1]; x += BS) {
                int s = d.svSlot[so + x];
                if (s >= 0x40000000) continue;
                int isl;
                bool applied;
                if (s >= 0) {
                    isl = island_of(d, d.svA[so + x], d.svB[so + x]);
                    if (isl < 0 || !d.isl0[isl]) continue;
                    anyActive = 1;
                    applied = solve_position_rows(d, so + x);
                } else {
                    PointCon<T>& pc = d.pc[-1 - s];
                    isl = island_of(d, pc.a, pc.b);
                    if (isl < 0 || !d.isl0[isl]) continue;
                    anyActive = 1;
                    applied = point_solve_position(d, pc);
                }
                if (applied) d.isl1[isl] = 1;
            }
            __syncthreads();
        }
        if (__syncthreads_or(anyActive)) sweeps = it + 1;
        for (int i = b0 + tid; i < b1; i += BS) { d.isl0[i] = d.isl1[i]; d.isl1[i] = 0; }
        __syncthreads();
    }
    if (tid == 0) d.stats[2 * w + 1] = sweeps;
    {  // byte model: per sweep 160 B/manifold + 192 B/point, per position sweep 240 B/manifold + 24 B/point
        int pts = 0, awake = 0;
        for (int k = tid; k < m; k += BS) pts += d.con[d.conIdx[c0 + k]].np;
        for (int i = b0 + tid; i < b1; i += BS) awake += (fl_motion(B.flags[i]) == MOTION_DYNAMIC && !B.sleeping[i]) ? 1 : 0;
        for (int o = 32; o > 0; o >>= 1) { pts += __shfl_xor(pts, o, 64); awake += __shfl_xor(awake, o, 64); }
        if ((tid & 63) == 0) {
            unsigned long long iv = (unsigned long long)d.s.velIters + 1;
            unsigned long long by = iv * (160ull * (tid == 0 ? m : 0) + 192ull * pts) +
                                    (unsigned long long)sweeps * (240ull * (tid == 0 ? m : 0) + 24ull * pts);
            atomicAdd(&d.acc[ACC_BYTES_SOLVE], by);
            atomicAdd(&d.acc[ACC_AWAKE], (unsigned long long)awake);
        }
    }
    // ---- 10. positions from the centre of mass + AABBs (update.ts:322-330)
    for (int i = b0 + tid; i < b1; i += BS) {
        int mo = fl_motion(B.flags[i]);
        if (mo == MOTION_STATIC || B.sleeping[i]) continue;
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
            if (l < 0 || fl_motion(B.flags[i]) != MOTION_DYNAMIC) continue;
            bool can = update_sleep_state(B, d.st, i, dt, maxMove, d.s.tSleep);
            if (!can) d.isl0[l] = 0;
        }
        __syncthreads();
        for (int i = b0 + tid; i < b1; i += BS) {
            int l = B.label[i];
            if (l < 0 || fl_motion(B.flags[i]) != MOTION_DYNAMIC) continue;
            if (d.isl0[l]) {
                B.sleeping[i] = 1;
                B.v[i] = mk3<T>(0, 0, 0);
                B.w[i] = mk3<T>(0, 0, 0);
            }
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
