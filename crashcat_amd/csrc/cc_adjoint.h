// Adjoint (reverse mode) of the velocity-level contact solve and the integrator, for
// differentiable stepping (cc_fwd / cc_bwd).
//
// The forward step is crashcat's (src/update.ts:45-239).  Its record for one step holds the
// world's contact constraints as the solver staged them (level order, lever arms, effective
// masses, biases), every contact point's impulses after the warm start and after each velocity
// sweep (with the friction impulses before the disc projection and whether it clamped them), the
// warm-start source of each point and the bodies' inverse masses and inertias.  The adjoint runs
// the same level schedule backwards: constraints of one level share no dynamic body, so a level
// is reversed in parallel and levels in reverse order, sweep by sweep from the last.
//
// What is differentiated, per step (reverse order):
//   integration   com' = com + v dt, rotation' = exp(w dt) rotation      (update.ts:896-976)
//   velocity sweeps: friction disc then normal clamp of every point      (contact-constraints.ts:963-1053)
//                 x = lambda + effMass (J v - bias); lambda' = P(x); v' = v - W J^T (lambda' - lambda)
//                 with P = max(0, .) (active set: lambda' > 0) for the normal, and for the friction
//                 pair the projection onto the disc of radius friction * lambda_n (identity inside;
//                 (r / |l|)(I - u u^T) and d/dr = u when clamped)             (:1001-1051)
//   warm start    v' = v - W J^T (ratio * lambda_cached), lambda_cached = the previous step's final
//                 impulse of the matched point                            (:906-955, 651-686)
//   speculative bias  bias = max(0, -penetration / dt), penetration = (pA - pB) . n with the contact
//                 points moving rigidly with their bodies                 (:288-423)
//   acceleration  v' = (v + a dt) (1 - c dt), w' = w (1 - c_w dt)          (update.ts:245-319)
//   friction      combined friction of each contact (combineMaterial, combine-material.ts:40-57)
//   lever arms    r1 = c - com_A, r2 = c - com_B at the contact point c, which enters J (r x axis),
//                 the effective masses and the impulses' angular parts.  The velocities each row saw
//                 are replayed backwards (every impulse undone in reverse order from the sweeps'
//                 final velocities).  The point moves with its bodies: its normal part half with
//                 each, its tangential part with the body that carries it (a sphere or capsule
//                 against a flat shape, the moving body against a static one, else half each).
// Held constant within a step (not differentiated): contact normals, inverse inertias, the
// restitution bias, velocity clamps, and the position (Baumgarte) sweeps.
#pragma once
#include "cc_solver.h"

namespace cc {

template <typename T>
struct Adj {
    // body adjoints: centre of mass, rotation (as a world-space rotation vector), velocities
    V3<T>* xb;
    V3<T>* qb;
    V3<T>* vb;
    V3<T>* wb;
    T* mub;          // [N] friction coefficient of each body
    T* lamBar;       // [3 * ptsCap] impulse adjoints (n, t1, t2) of the staged points
    T* bBar;         // [ptsCap] normal-bias adjoints
    V3<T>* rBar;     // [2 * ptsCap] lever-arm adjoints (r1, r2) of the staged points
    V3<T>* vf;       // [N] the velocities the sweeps saw, replayed backwards
    V3<T>* wf;
    T* seedIn;       // [3 * 4 * conCap] adjoints of this step's final impulses, at rank * 4 + point
    T* seedOut;      // the same for the step before, through the warm-start sources
    // the step's record
    const VElem<T>* ve;
    const VPoint<T>* vp;
    const V3<T>* la;        // body-space contact points on A and B
    const V3<T>* lb;
    const LamRec<T>* lam;   // [(sweeps + 1) * ptsCap]
    const int* src;         // [ptsCap]
    const int* rank;        // [svCap]
    const int* lvl;         // level starts (the solver's lvlStart layout)
    const int* stats;       // [2 * W]: levels of each world
    const int* wcs;         // [W + 1]: each world's range of constraints
    const VBody<T>* vbody;  // inverse mass and world inverse inertia of each body
    const NpBody<T>* npb;   // the collision-time pose of each body
    int ptsCap;
    int sweeps;
    T dt, ratio;
};

// sum over the bodies of vbar' . d v' / d(dl) for an impulse dl along `axis` (applyLambda
// axis-constraint-part.ts:640-707: v_A -= axis dl / m_A, w_A -= I_A (r1 x axis) dl, B the opposite)
template <typename T>
__device__ inline T imp_adj(int dyn, int dofA, int dofB, const VBody<T>& A, const VBody<T>& B, V3<T> r1, V3<T> r2,
                            V3<T> axis, V3<T> vbA, V3<T> wbA, V3<T> vbB, V3<T> wbB) {
    T s = T(0);
    if (dyn & 1) s += -A.invMass * dot(dof_lin(axis, dofA), vbA) - dot(mul(A.I, cross(r1, axis)), wbA);
    if (dyn & 2) s += B.invMass * dot(dof_lin(axis, dofB), vbB) + dot(mul(B.I, cross(r2, axis)), wbB);
    return s;
}

// vbar += g * d(J v)/dv for one axis (getTotalLambda :578-625; only dynamic bodies carry adjoints)
template <typename T>
__device__ inline void jv_adj(int dyn, V3<T> r1, V3<T> r2, V3<T> axis, T g, V3<T>& vbA, V3<T>& wbA, V3<T>& vbB,
                              V3<T>& wbB) {
    if (dyn & 1) { vbA = vbA + axis * g; wbA = wbA + cross(r1, axis) * g; }
    if (dyn & 2) { vbB = vbB - axis * g; wbB = wbB - cross(r2, axis) * g; }
}

// d combine(a, b) / d a (combineMaterial combine-material.ts:40-57)
template <typename T>
__device__ inline T combine_da(T a, T b, int ma, int mb, T c) {
    const int prio[5] = {0, 1, 3, 4, 2};
    const int mode = prio[ma] >= prio[mb] ? ma : mb;
    switch (mode) {
        case 1: return b;
        case 2: return a <= b ? T(1) : T(0);
        case 3: return a >= b ? T(1) : T(0);
        case 4: return a > T(0) ? c / (T(2) * a) : T(0);
        default: return T(0.5);
    }
}

// the friction impulses of a point after sweep s (record index s + 1; index 0: the warm start),
// from the pre-projection values and the disc radius fr * lambda_n of the sweep before
template <typename T>
__device__ inline void fric_post(const LamRec<T>* lam, size_t pts, size_t p, int idx, T fr, T& t1, T& t2) {
    const LamRec<T> r = lam[(size_t)idx * pts + p];
    t1 = r.t1;
    t2 = r.t2;
    if (idx > 0 && r.clamp != T(0)) {
        const T len = sqrt(r.t1 * r.t1 + r.t2 * r.t2);
        const T sc = len > T(0) ? fr * lam[(size_t)(idx - 1) * pts + p].n / len : T(0);
        t1 *= sc;
        t2 *= sc;
    }
}

// undo an impulse dl along `axis` on the replayed velocities (applyLambda's inverse)
template <typename T>
__device__ inline void undo_imp(int dyn, int dofA, int dofB, const VBody<T>& A, const VBody<T>& B, V3<T> r1, V3<T> r2,
                                V3<T> axis, T dl, V3<T>& vA, V3<T>& wA, V3<T>& vB, V3<T>& wB) {
    if (dl == T(0)) return;
    if (dyn & 1) { vA = vA + dof_lin(axis, dofA) * (A.invMass * dl); wA = wA + mul(A.I, cross(r1, axis)) * dl; }
    if (dyn & 2) { vB = vB - dof_lin(axis, dofB) * (B.invMass * dl); wB = wB - mul(B.I, cross(r2, axis)) * dl; }
}

// lever-arm adjoints of one row: through x = lambda_old + eff (J v - bias), where J v holds
// (r x axis) . w and 1 / eff holds (r x axis) . I (r x axis) (so d eff / d r = -2 eff^2
// axis x I (r x axis), and eff (J v - bias) = dx = x - lambda_old), and through the impulse dl's
// angular parts; wA, wB are the velocities the row saw, wbA, wbB the adjoints after it
template <typename T>
__device__ inline void row_r_adj(int dyn, const VBody<T>& A, const VBody<T>& B, V3<T> r1, V3<T> r2, V3<T> axis, T eff,
                                 T xbar, T dx, T dl, V3<T> wA, V3<T> wB, V3<T> wbA, V3<T> wbB, V3<T>& rb1, V3<T>& rb2) {
    if (dyn & 1) {
        rb1 = rb1 + cross(axis, mul(A.I, wbA)) * (-dl);
        if (xbar != T(0)) rb1 = rb1 + (cross(axis, wA) * eff - cross(axis, mul(A.I, cross(r1, axis))) * (T(2) * eff * dx)) * xbar;
    }
    if (dyn & 2) {
        rb2 = rb2 + cross(axis, mul(B.I, wbB)) * dl;
        if (xbar != T(0)) rb2 = rb2 - (cross(axis, wB) * eff + cross(axis, mul(B.I, cross(r2, axis))) * (T(2) * eff * dx)) * xbar;
    }
}

// Reverse of one contact constraint's velocity sweep `it` (solve_velocity_elem): its normals last
// point first, then its friction pairs last point first, each after undoing its own impulse on
// the replayed velocities.
template <typename T>
__device__ inline void adj_elem(const Dev<T>& d, const Adj<T>& a, const VElem<T>& E, int it) {
    const int meta = E.meta;
    const int dyn = meta & 3, np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    if (!dyn) return;
    const VBody<T> IA = a.vbody[E.a], IB = a.vbody[E.b];
    V3<T> vbA = a.vb[E.a], wbA = a.wb[E.a], vbB = a.vb[E.b], wbB = a.wb[E.b];
    V3<T> vA = a.vf[E.a], wA = a.wf[E.a], vB = a.vf[E.b], wB = a.wf[E.b];
    const V3<T> n = E.n, t1 = contact_t1(n), t2 = cross(n, t1);
    const T fr = E.fr;
    const size_t pts = (size_t)a.ptsCap;
    const LamRec<T>* Lold = a.lam + (size_t)it * pts;
    const LamRec<T>* Lnew = a.lam + (size_t)(it + 1) * pts;
    T frBar = T(0);
    for (int j = np - 1; j >= 0; j--) {  // normals: lambda' = max(0, lambda + eff (J v - bias))
        const size_t p = (size_t)E.pt + j;
        const VPoint<T> P = a.vp[p];
        T* lb = a.lamBar + 3 * p;
        const T lnew = Lnew[p].n, lold = Lold[p].n;
        const T dl = lnew - lold;
        undo_imp(dyn, dofA, dofB, IA, IB, P.r1, P.r2, n, dl, vA, wA, vB, wB);
        const T db = imp_adj(dyn, dofA, dofB, IA, IB, P.r1, P.r2, n, vbA, wbA, vbB, wbB);
        const T tot = lb[0] + db;
        T old = -db;
        const bool active = P.effN != T(0) && lnew > T(0);
        const T xb = active ? tot : T(0);
        old += xb;
        V3<T> rb1 = a.rBar[2 * p], rb2 = a.rBar[2 * p + 1];
        row_r_adj(dyn, IA, IB, P.r1, P.r2, n, P.effN, xb, lnew - lold, dl, wA, wB, wbA, wbB, rb1, rb2);
        a.rBar[2 * p] = rb1;
        a.rBar[2 * p + 1] = rb2;
        jv_adj(dyn, P.r1, P.r2, n, P.effN * xb, vbA, wbA, vbB, wbB);
        a.bBar[p] -= P.effN * xb;
        lb[0] = old;
    }
    for (int j = np - 1; j >= 0; j--) {  // friction pairs: the disc of radius fr * lambda_n
        const size_t p = (size_t)E.pt + j;
        const VPoint<T> P = a.vp[p];
        if (P.effT1 == T(0) && P.effT2 == T(0)) continue;
        T* lb = a.lamBar + 3 * p;
        const LamRec<T> rn = Lnew[p];
        const T lnOld = Lold[p].n;
        T o1, o2, n1, n2;  // friction impulses before and after this sweep
        fric_post(a.lam, pts, p, it, fr, o1, o2);
        fric_post(a.lam, pts, p, it + 1, fr, n1, n2);
        const T d1 = n1 - o1, d2 = n2 - o2;
        const V3<T> wbA0 = wbA, wbB0 = wbB;
        undo_imp(dyn, dofA, dofB, IA, IB, P.r1, P.r2, t2, d2, vA, wA, vB, wB);
        undo_imp(dyn, dofA, dofB, IA, IB, P.r1, P.r2, t1, d1, vA, wA, vB, wB);
        const T db1 = imp_adj(dyn, dofA, dofB, IA, IB, P.r1, P.r2, t1, vbA, wbA, vbB, wbB);
        const T db2 = imp_adj(dyn, dofA, dofB, IA, IB, P.r1, P.r2, t2, vbA, wbA, vbB, wbB);
        const T g1 = lb[1] + db1, g2 = lb[2] + db2;  // adjoints of the projected impulses
        T h1 = g1, h2 = g2, mfb = T(0);
        if (rn.clamp != T(0)) {
            const T len = sqrt(rn.t1 * rn.t1 + rn.t2 * rn.t2);
            if (len > T(0)) {
                const T u1 = rn.t1 / len, u2 = rn.t2 / len;
                const T mf = fr * lnOld;
                const T dt_ = u1 * g1 + u2 * g2;
                h1 = (mf / len) * (g1 - u1 * dt_);
                h2 = (mf / len) * (g2 - u2 * dt_);
                mfb = dt_;
            }
        }
        V3<T> rb1 = a.rBar[2 * p], rb2 = a.rBar[2 * p + 1];
        row_r_adj(dyn, IA, IB, P.r1, P.r2, t1, P.effT1, h1, rn.t1 - o1, d1, wA, wB, wbA0, wbB0, rb1, rb2);
        row_r_adj(dyn, IA, IB, P.r1, P.r2, t2, P.effT2, h2, rn.t2 - o2, d2, wA, wB, wbA0, wbB0, rb1, rb2);
        a.rBar[2 * p] = rb1;
        a.rBar[2 * p + 1] = rb2;
        jv_adj(dyn, P.r1, P.r2, t1, P.effT1 * h1, vbA, wbA, vbB, wbB);
        jv_adj(dyn, P.r1, P.r2, t2, P.effT2 * h2, vbA, wbA, vbB, wbB);
        frBar += lnOld * mfb;
        lb[0] += fr * mfb;
        lb[1] = -db1 + h1;
        lb[2] = -db2 + h2;
    }
    if (dyn & 1) { a.vb[E.a] = vbA; a.wb[E.a] = wbA; a.vf[E.a] = vA; a.wf[E.a] = wA; }
    if (dyn & 2) { a.vb[E.b] = vbB; a.wb[E.b] = wbB; a.vf[E.b] = vB; a.wf[E.b] = wB; }
    if (frBar != T(0)) {  // the combined friction of the pair, to the two bodies' coefficients
        const Bodies<T>& B = d.B;
        const T ma = B.friction[E.a], mb = B.friction[E.b];
        const int fa = B.flags[E.a], fb = B.flags[E.b];
        atomicAdd(&a.mub[E.a], frBar * combine_da(ma, mb, fl_fmode(fa), fl_fmode(fb), fr));
        atomicAdd(&a.mub[E.b], frBar * combine_da(mb, ma, fl_fmode(fb), fl_fmode(fa), fr));
    }
}

// Reverse of the warm start of one constraint (all sweeps already reversed): the warm-start
// impulse's adjoint is the sweep-0 impulse adjoint plus its own effect on the velocities; it
// goes, scaled by the dt ratio, to the previous step's final impulse it was taken from.  The
// impulses' angular parts add to the lever-arm adjoints.
template <typename T>
__device__ inline void adj_warm(const Adj<T>& a, const VElem<T>& E) {
    const int meta = E.meta;
    const int dyn = meta & 3, np = (meta >> 4) & 7, dofA = (meta >> 8) & 63, dofB = (meta >> 14) & 63;
    if (!dyn) return;
    const VBody<T> IA = a.vbody[E.a], IB = a.vbody[E.b];
    const V3<T> vbA = a.vb[E.a], wbA = a.wb[E.a], vbB = a.vb[E.b], wbB = a.wb[E.b];
    const V3<T> n = E.n, t1 = contact_t1(n), t2 = cross(n, t1);
    const V3<T> z = mk3<T>(0, 0, 0);
    for (int j = 0; j < np; j++) {
        const size_t p = (size_t)E.pt + j;
        const VPoint<T> P = a.vp[p];
        const LamRec<T> w0 = a.lam[p];  // the warm-started impulses
        V3<T> rb1 = a.rBar[2 * p], rb2 = a.rBar[2 * p + 1];
        if (P.effT1 != T(0) || P.effT2 != T(0)) {
            row_r_adj(dyn, IA, IB, P.r1, P.r2, t1, T(0), T(0), T(0), w0.t1, z, z, wbA, wbB, rb1, rb2);
            row_r_adj(dyn, IA, IB, P.r1, P.r2, t2, T(0), T(0), T(0), w0.t2, z, z, wbA, wbB, rb1, rb2);
        }
        row_r_adj(dyn, IA, IB, P.r1, P.r2, n, T(0), T(0), T(0), w0.n, z, z, wbA, wbB, rb1, rb2);
        a.rBar[2 * p] = rb1;
        a.rBar[2 * p + 1] = rb2;
        const int s = a.src[p];
        if (s < 0 || a.ratio == T(0)) continue;
        const T* lb = a.lamBar + 3 * p;
        T* out = a.seedOut + 3 * (size_t)s;
        if (P.effN != T(0))
            atomicAdd(&out[0], a.ratio * (lb[0] + imp_adj(dyn, dofA, dofB, IA, IB, P.r1, P.r2, n, vbA, wbA, vbB, wbB)));
        if (P.effT1 != T(0))
            atomicAdd(&out[1], a.ratio * (lb[1] + imp_adj(dyn, dofA, dofB, IA, IB, P.r1, P.r2, t1, vbA, wbA, vbB, wbB)));
        if (P.effT2 != T(0))
            atomicAdd(&out[2], a.ratio * (lb[2] + imp_adj(dyn, dofA, dofB, IA, IB, P.r1, P.r2, t2, vbA, wbA, vbB, wbB)));
    }
}

// the share of a contact point's tangential motion that body A carries: a sphere or capsule
// against a flat shape carries it, a moving body against a static one carries it, else half
__device__ inline float carry_a(int tinfoA, int flagsA, int tinfoB, int flagsB) {
    const int ta = tinfoA & 15, tb = tinfoB & 15;
    const bool roundA = ta == SHAPE_SPHERE || ta == SHAPE_CAPSULE, roundB = tb == SHAPE_SPHERE || tb == SHAPE_CAPSULE;
    if (roundA != roundB) return roundA ? 1.0f : 0.0f;
    const bool sa = fl_motion(flagsA) == MOTION_STATIC, sb = fl_motion(flagsB) == MOTION_STATIC;
    if (sa != sb) return sa ? 0.0f : 1.0f;
    return 0.5f;
}

// The contact geometry's adjoints to the bodies' positions and rotations: the speculative bias
// (-penetration / dt) and the lever arms.
template <typename T>
__device__ inline void adj_geometry(const Adj<T>& a, const VElem<T>& E) {
    const int meta = E.meta;
    const int dyn = meta & 3, np = (meta >> 4) & 7;
    if (!dyn) return;
    const V3<T> n = E.n;
    const NpBody<T> RA = a.npb[E.a], RB = a.npb[E.b];
    const T wa = T(carry_a(RA.tinfo, RA.flags, RB.tinfo, RB.flags)), wb = T(1) - wa;
    V3<T> xA = mk3<T>(0, 0, 0), qA = xA, xB = xA, qB = xA;
    for (int j = 0; j < np; j++) {
        const size_t p = (size_t)E.pt + j;
        const VPoint<T> P = a.vp[p];
        if (P.biasN > T(0)) {  // bias = -((pA - pB) . n) / dt, pA and pB moving with their bodies
            const T pb = -a.bBar[p] / a.dt;
            xA = xA + n * pb;
            qA = qA + cross(qrotate(RA.q, a.la[p]), n) * pb;
            xB = xB - n * pb;
            qB = qB - cross(qrotate(RB.q, a.lb[p]), n) * pb;
        }
        const V3<T> r1b = a.rBar[2 * p], r2b = a.rBar[2 * p + 1];
        const V3<T> cb = r1b + r2b;  // the contact point's adjoint
        const V3<T> gn = n * dot(n, cb), gt = cb - gn;
        const V3<T> pA = gt * wa + gn * T(0.5), pB = gt * wb + gn * T(0.5);
        xA = xA - r1b + pA;
        qA = qA + cross(P.r1, pA);
        xB = xB - r2b + pB;
        qB = qB + cross(P.r2, pB);
    }
    if (dyn & 1) { a.xb[E.a] = a.xb[E.a] + xA; a.qb[E.a] = a.qb[E.a] + qA; }
    if (dyn & 2) { a.xb[E.b] = a.xb[E.b] + xB; a.qb[E.b] = a.qb[E.b] + qB; }
}

// The reverse of one recorded step of one world, on one workgroup (worlds are solved unsplit).
template <typename T, int BS>
__global__ void __launch_bounds__(BS) k_adj_world(Dev<T> d, Adj<T> a) {
    const int4 vt = d.vTab[blockIdx.x];
    const int w = vt.x;
    const int tid = threadIdx.x;
    const Bodies<T>& B = d.B;
    const int b0 = d.wStart[w], b1 = d.wStart[w + 1];
    const int c0 = a.wcs[w], c1 = a.wcs[w + 1];
    const int u0 = d.pcStart[w];
    const int so = c0 + u0;
    const int L = c1 - c0;  // no point constraints on this path
    const int vid = d.worldVBase[w];
    const int* LS = a.lvl + so + vid;
    const int D = a.stats[2 * w];
    const T dt = a.dt;
    // integration: com' = com + v dt, rotation' = exp(w dt) rotation; the sweeps' final velocities
    for (int i = b0 + tid; i < b1; i += BS) {
        const VBody<T> r = a.vbody[i];
        a.vf[i] = r.v;
        a.wf[i] = r.w;
        if (fl_motion(B.flags[i]) != MOTION_DYNAMIC) continue;
        a.vb[i] = a.vb[i] + a.xb[i] * dt;
        a.wb[i] = a.wb[i] + a.qb[i] * dt;
    }
    // the adjoints of this step's final impulses (from the next step's warm start)
    for (int x = tid; x < L; x += BS) {
        const VElem<T> E = a.ve[so + x];
        if ((E.meta >> 28) != VK_CONTACT) continue;
        const int np = (E.meta >> 4) & 7;
        const size_t r4 = 4 * (size_t)a.rank[so + x];
        for (int j = 0; j < np; j++) {
            const size_t p = (size_t)E.pt + j;
            T* in = a.seedIn + 3 * (r4 + j);
            for (int c = 0; c < 3; c++) {
                a.lamBar[3 * p + c] = in[c];
                in[c] = T(0);
            }
            a.bBar[p] = T(0);
            a.rBar[2 * p] = mk3<T>(0, 0, 0);
            a.rBar[2 * p + 1] = mk3<T>(0, 0, 0);
        }
    }
    __syncthreads();
    // velocity sweeps, last first; levels in reverse order
    for (int it = a.sweeps - 1; it >= 0; it--) {
        for (int l = D - 1; l >= 0; l--) {
            const int lo = LS[l], hi = LS[l + 1];
            for (int x = lo + tid; x < hi; x += BS) {
                const VElem<T> E = a.ve[so + x];
                if ((E.meta >> 28) != VK_CONTACT || it >= (E.iters & 0xffff)) continue;
                adj_elem(d, a, E, it);
            }
            __syncthreads();
        }
    }
    // warm start
    for (int x = tid; x < L; x += BS) {
        const VElem<T> E = a.ve[so + x];
        if ((E.meta >> 28) == VK_CONTACT) adj_warm(a, E);
    }
    __syncthreads();
    // contact geometry to positions and rotations, level by level (no two share a dynamic body)
    for (int l = 0; l < D; l++) {
        const int lo = LS[l], hi = LS[l + 1];
        for (int x = lo + tid; x < hi; x += BS) {
            const VElem<T> E = a.ve[so + x];
            if ((E.meta >> 28) == VK_CONTACT) adj_geometry(a, E);
        }
        __syncthreads();
    }
    // acceleration: v' = (v + a dt)(1 - c dt) within the DOFs
    for (int i = b0 + tid; i < b1; i += BS) {
        const int f = B.flags[i];
        if (fl_motion(f) != MOTION_DYNAMIC) continue;
        const T ld = tmax<T>(T(0), T(1) - B.linDamp[i] * dt), ad = tmax<T>(T(0), T(1) - B.angDamp[i] * dt);
        a.vb[i] = dof_lin(a.vb[i], fl_dof(f)) * ld;
        a.wb[i] = dof_lin(a.wb[i], fl_dof(f) >> 3) * ad;
    }
}

}  // namespace cc
