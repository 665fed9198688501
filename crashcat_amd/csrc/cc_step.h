// Device state and per-phase device functions of one step (templated on T).
#pragma once
#include "cc_narrow.h"

namespace cc {

enum { MOTION_STATIC = 0, MOTION_KINEMATIC = 1, MOTION_DYNAMIC = 2 };

// bit layout of Bodies::flags
CC_HD int fl_motion(int f) { return f & 3; }
CC_HD int fl_dof(int f) { return (f >> 2) & 63; }
CC_HD int fl_allow_sleep(int f) { return (f >> 8) & 1; }
CC_HD int fl_fmode(int f) { return (f >> 9) & 7; }
CC_HD int fl_rmode(int f) { return (f >> 12) & 7; }
CC_HD int fl_layer(int f) { return (f >> 16) & 255; }

// ---------------------------------------------------------------- CCD records
constexpr int FL_LINEAR_CAST = 1 << 24;  // Bodies::flags: motionQuality == LINEAR_CAST

// CCDBody (src/ccd.ts:19-60)
template <typename T>
struct CcdRec {
    int body, hit;
    unsigned sub1, sub2;  // subShapeId1 / subShapeId2: the hit's sub-shape ids
    V3<T> delta, normal, point;
    T fraction, fps, thrSq, maxPen, friction, restitution;
};
// the AxisConstraintPart applyCCD keeps from one call to the next (module scope in
// src/update.ts:1453-1454): a zero record at upload, then whatever the last call left
template <typename T>
struct CcdPart {
    V3<T> r1x, r2x, i1, i2;
    T eff, lambda, bias;
};

template <typename T>
struct SettingsD {
    T gravity[3];
    T spec, manifoldTol, colTol, penTol, preserveDist2, minVelRest, slop, beta, maxPen, tSleep, sleepThresh;
    int velIters, posIters, gravityEnabled, warmStarting, allowSleeping, pad;
    unsigned char layerPairs[64];
    T normalCos;           // manifolds merge hits whose normals are within this cosine
    int backfaces;         // triangle meshes: collide with back faces
    int activeEdges;       // triangle meshes: correct normals at inactive edges
    T ccdThr, ccdMaxPen, ccdMinVelRest;  // ccd.linearCastThreshold, .linearCastMaxPenetration, .minVelocityForRestitution
};

template <typename T>
struct Bodies {
    V3<T>* pos;
    Q4<T>* q;
    V3<T>* com;
    V3<T>* v;
    V3<T>* w;
    V3<T>* amin;
    V3<T>* amax;
    T* invMass;
    V3<T>* invI;
    Q4<T>* iRot;
    T* friction;
    T* restitution;
    T* gravityFactor;
    T* linDamp;
    T* angDamp;
    T* maxLinV;
    T* maxAngV;
    int* shape;
    int* flags;
    unsigned* groups;
    unsigned* mask;
    int* world;
    int* sleeping;
    V3<T>* sc;     // 3 sleep-test sphere centres per body
    T* sr;         // 3 radii per body
    T* stimer;
    int* label;    // union-find parent / island label (-1 when not in an island)
    int* wake;     // set when a new constraint wakes the body
    V3<T>* force;  // accumulated force and torque, integrated by the next step and then cleared
    V3<T>* torque;
    int* ovr;      // iteration overrides: velocity | position << 8 (0 = world default)
};

// One axis with its cached terms, as the position pass builds it for a contact point.
template <typename T>
struct Row {
    V3<T> r1x, r2x, i1, i2;  // r1 x axis, r2 x axis, I1^-1 (r1 x axis), I2^-1 (r2 x axis)
    T eff, lambda, bias;
};
// Contact constraint record (src/constraints/contact-constraints.ts:31-95) for one pair slot.
template <typename T>
struct ConPoint {
    V3<T> r1, r2;                 // lever arms from the centres of mass to the contact point
    T effN, effT1, effT2, biasN;  // effective masses of the three axes, normal bias
    T lN, lT1, lT2;               // accumulated impulses
    V3<T> localA, localB;
};
template <typename T>
struct Con {
    int a, b, np, dynMask;  // dynMask bit0: A dynamic, bit1: B dynamic
    unsigned sortKey;
    int movMask;            // bit0: A not static, bit1: B not static
    int cq, co;             // the broadphase pair (querying body, other body; bits 24+ of co: the
                            // constraint's place in the pair's run); constraints are created in
                            // ascending (cq, co, place) order
    V3<T> normal, t1, t2;
    T friction, invMassA, invMassB;
    ConPoint<T> p[4];
};
// What the island, cache and ordering passes read of a constraint: bodies, sort key,
// bits = dynMask | movMask << 2 | np << 4.
struct ConHead {
    int a, b;
    unsigned sortKey;
    int bits;  // bits 0-1 dynamic A/B, 2-3 non-static A/B, 4-6 points, 7 not the first of its pair's
               // run, 8-12 (first only) the run's length - 1, 13 found in the previous step's cache
};
// Persistent contact cache entry (src/contacts.ts:9-58): four in-place slots.
template <typename T>
struct CacheRec {
    int a, b, np, run;     // run: constraints the pair made this step (in its first record; 0 in the others)
    unsigned subA, subB;   // sub-shape ids (0xffffffff when empty)
    V3<T> p1[4], p2[4];
    T ln[4], lt1[4], lt2[4];
    int rank;              // differentiable stepping: the constraint's place in pair order this step
};
// What the sweeps read: one record per body, per constraint and per contact point, each read
// with one-line loads (a random body costs one memory line rather than one per field).
// A body: its velocities, its world-space inverse inertia for this step (fixed while the
// velocities are iterated: the pose changes only afterwards) and its inverse mass.
template <typename T>
struct alignas(16) VBody {
    V3<T> v, w;
    M3<T> I;
    T invMass;
};
// A constraint in level order.  meta: dynMask | movMask << 2 | np << 4 | dofA << 8 |
// dofB << 14 | kind << 28 (0 contact, 1 point constraint with index a, 2 inactive); pt: its
// first point record; slot: its constraint slot.
enum { VK_CONTACT = 0, VK_POINT = 1, VK_INACTIVE = 2 };
// iters: the velocity | position << 16 iterations of the element's island
template <typename T>
struct alignas(16) VElem {
    int a, b, meta, pt;
    int slot, iters;
    V3<T> n;
    T fr;
};
// A contact point: lever arms, effective masses of the three axes, normal bias, impulses.
template <typename T>
struct alignas(16) VPoint {
    V3<T> r1, r2;
    T effN, effT1, effT2, biasN;
    T lN, lT1, lT2;
};
// A body as the position sweeps read it: centre of mass and rotation (updated by the sweeps),
// local inverse inertia and its rotation, inverse mass, island label.
template <typename T>
struct alignas(16) PBody {
    V3<T> com;
    Q4<T> q;
    V3<T> invI;
    Q4<T> iRot;
    T invMass;
    int label;
};
// World-space points of a constraint, kept only for the parity taps.
template <typename T>
struct ConTap {
    V3<T> pa[4], pb[4];
    int src[4];  // differentiable stepping: where each point's warm-start impulses came from
                 // (rank * 4 + point of the previous step's constraint), -1 for none
};

// One contact point after one velocity sweep, recorded for the adjoint: its normal impulse, its
// friction impulses before the disc projection, and 1 when the projection clamped them (the
// recorded sweep -1 holds the warm-started impulses).
template <typename T>
struct alignas(16) LamRec {
    T n, t1, t2, clamp;
};

// Ball joint (src/constraints/point-constraint.ts, constraint-part/point-constraint-part.ts).
template <typename T>
struct PointCon {
    int a, b, enabled, sleeping;
    V3<T> localA, localB, r1, r2;
    M3<T> i1X, i2X;   // I1^-1 [r1]x, I2^-1 [r2]x
    M3<T> eff;        // K^-1 (3x3), computed in FP64
    int active;
    V3<T> lambda;
    int ovr;          // iteration overrides: velocity | position << 8 (0 = world default)
};

#define CC_EMPTY_KEY 0xffffffffffffffffull

CC_HD unsigned sort_key(unsigned a, unsigned b, unsigned sa, unsigned sb) {  // computeContactSortKey :1222-1231
    unsigned h = 0;
    h = h * 31u + a; h = h * 31u + b; h = h * 31u + sa; h = h * 31u + sb;
    return h;
}
CC_HD unsigned long long pair_key(int a, int b) { return ((unsigned long long)(unsigned)a << 32) | (unsigned)b; }
CC_HD unsigned hash64(unsigned long long k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
    return (unsigned)k;
}

template <typename T>
CC_HD T combine(T a, T b, int ma, int mb) {  // combineMaterial combine-material.ts:40-57
    const int prio[5] = {0, 1, 3, 4, 2};
    int mode = prio[ma] >= prio[mb] ? ma : mb;
    switch (mode) {
        case 1: return a * b;
        case 2: return tmin<T>(a, b);
        case 3: return tmax<T>(a, b);
        case 4: return sqrt(a * b);
        default: return (a + b) / T(2);
    }
}

// ---------------------------------------------------------------- body helpers
template <typename T>
__device__ inline void update_com(const Bodies<T>& B, const ShapeTable<T>& st, int i) {  // rigid-body.ts:527-536
    const ShapeD<T>& s = st.shapes[B.shape[i]];
    B.com[i] = qrotate(B.q[i], mk3<T>(s.com[0], s.com[1], s.com[2])) + B.pos[i];
}
template <typename T>
__device__ inline void pos_from_com(const Bodies<T>& B, const ShapeTable<T>& st, int i) {  // :545-553
    const ShapeD<T>& s = st.shapes[B.shape[i]];
    B.pos[i] = B.com[i] - qrotate(B.q[i], mk3<T>(s.com[0], s.com[1], s.com[2]));
}
template <typename T>
__device__ inline void update_aabb(const Bodies<T>& B, const ShapeTable<T>& st, int i) {  // updateAABB :561-674
    const ShapeD<T>& s = st.shapes[B.shape[i]];
    M3<T> R = m3_from_quat(B.q[i]);
    V3<T> p = B.pos[i];
    V3<T> lo = p, hi = p;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        V3<T> l = mk3<T>((c & 1) ? s.amax[0] : s.amin[0], (c & 2) ? s.amax[1] : s.amin[1], (c & 4) ? s.amax[2] : s.amin[2]);
        V3<T> wp = mul(R, l) + p;
        if (c == 0) { lo = wp; hi = wp; }
        else {
            lo = mk3<T>(tmin(lo.x, wp.x), tmin(lo.y, wp.y), tmin(lo.z, wp.z));
            hi = mk3<T>(tmax(hi.x, wp.x), tmax(hi.y, wp.y), tmax(hi.z, wp.z));
        }
    }
    B.amin[i] = lo;
    B.amax[i] = hi;
}
template <typename T>
__device__ inline void set_transform(const Bodies<T>& B, const ShapeTable<T>& st, int i) {
    update_com(B, st, i);
    update_aabb(B, st, i);
}

template <typename T>
__device__ inline void sleep_points(const Bodies<T>& B, const ShapeTable<T>& st, int i, V3<T> out[3]) {  // sleep.ts:22-76
    const ShapeD<T>& s = st.shapes[B.shape[i]];
    V3<T> c = B.com[i];
    Q4<T> q = B.q[i];
    out[0] = c;
    T x = (s.amax[0] - s.amin[0]) * T(0.5), y = (s.amax[1] - s.amin[1]) * T(0.5), z = (s.amax[2] - s.amin[2]) * T(0.5);
    int low = (x <= y && x <= z) ? 0 : (y <= z ? 1 : 2);
    V3<T> ax = mk3<T>(1, 0, 0), ay = mk3<T>(0, 1, 0), az = mk3<T>(0, 0, 1);
    if (low == 0) { out[1] = c + qrotate(q, ay) * y; out[2] = c + qrotate(q, az) * z; }
    else if (low == 1) { out[1] = c + qrotate(q, ax) * x; out[2] = c + qrotate(q, az) * z; }
    else { out[1] = c + qrotate(q, ax) * x; out[2] = c + qrotate(q, ay) * y; }
}
template <typename T>
__device__ inline void reset_sleep_timer(const Bodies<T>& B, const ShapeTable<T>& st, int i) {  // sleep.ts:120-127
    if (fl_motion(B.flags[i]) != MOTION_DYNAMIC) return;
    V3<T> p[3];
    sleep_points(B, st, i, p);
    for (int k = 0; k < 3; k++) { B.sc[3 * i + k] = p[k]; B.sr[3 * i + k] = T(0); }
    B.stimer[i] = T(0);
}
template <typename T>
__device__ inline bool update_sleep_state(const Bodies<T>& B, const ShapeTable<T>& st, int i, T dt, T maxMove, T tBefore) {
    if (!fl_allow_sleep(B.flags[i])) return false;  // sleep.ts:88-117
    V3<T> p[3];
    sleep_points(B, st, i, p);
    for (int k = 0; k < 3; k++) {
        T d = vlen(B.sc[3 * i + k] - p[k]);
        T r = tmax<T>(B.sr[3 * i + k], d);
        B.sr[3 * i + k] = r;
        if (r > maxMove) {
            for (int j = 0; j < 3; j++) { B.sc[3 * i + j] = p[j]; B.sr[3 * i + j] = T(0); }
            B.stimer[i] = T(0);
            return false;
        }
    }
    B.stimer[i] += dt;
    return B.stimer[i] >= tBefore;
}

template <typename T>
__device__ inline void add_rot_step(Q4<T>& q, V3<T> s, T sign) {  // add/subRotationStep rigid-body-step.ts:55-97
    T len = vlen(s);
    if (len > T(1.0e-6)) {
        V3<T> axis = s * (T(1) / len);
        q = qnormalize(qmul(qaxis_angle(axis, sign * len), q));
    }
}

// first friction axis of a contact normal (contact-constraints.ts getPerpendicular use in
// addContactConstraint); the set-up and the sweeps compute it with the same operations
template <typename T>
__device__ inline V3<T> contact_t1(V3<T> n) {
    if (fabs(n.x) > fabs(n.y)) { T tl = sqrt(n.x * n.x + n.z * n.z); return mk3<T>(n.z / tl, 0, -n.x / tl); }
    T tl = sqrt(n.y * n.y + n.z * n.z);
    return mk3<T>(0, n.z / tl, -n.y / tl);
}

// ---------------------------------------------------------- constraint parts
template <typename T>
__device__ inline void axis_props(Row<T>& p, bool dynA, bool dynB, T invMA, T invMB, const M3<T>& IA, const M3<T>& IB,
                                  V3<T> r1, V3<T> r2, V3<T> axis, T bias) {  // axis-constraint-part.ts:87-201
    T inv = T(0);
    p.r1x = p.r2x = p.i1 = p.i2 = mk3<T>(0, 0, 0);
    if (dynA) { p.r1x = cross(r1, axis); p.i1 = mul(IA, p.r1x); inv += invMA + dot(p.i1, p.r1x); }
    if (dynB) { p.r2x = cross(r2, axis); p.i2 = mul(IB, p.r2x); inv += invMB + dot(p.i2, p.r2x); }
    if (inv == T(0)) { p.eff = T(0); p.lambda = T(0); p.bias = T(0); }
    else { p.eff = T(1) / inv; p.bias = bias; }
}

// effective mass of one axis from the lever arms (calculateConstraintProperties
// axis-constraint-part.ts:87-201); 0 when the axis has no mass
template <typename T>
__device__ inline T axis_eff(bool dynA, bool dynB, T invMA, T invMB, const M3<T>& IA, const M3<T>& IB, V3<T> r1,
                             V3<T> r2, V3<T> axis) {
    T inv = T(0);
    if (dynA) { V3<T> r1x = cross(r1, axis); V3<T> i1 = mul(IA, r1x); inv += invMA + dot(i1, r1x); }
    if (dynB) { V3<T> r2x = cross(r2, axis); V3<T> i2 = mul(IB, r2x); inv += invMB + dot(i2, r2x); }
    return inv == T(0) ? T(0) : T(1) / inv;
}

// J v of one axis (getTotalLambda :578-625 before the lambda and bias terms).  r x axis is
// recomputed from the lever arm exactly as the constraint set-up computes it, and is zero for
// a body that is not dynamic.
template <typename T>
__device__ inline T axis_jv(int dyn, int mov, V3<T> r1, V3<T> r2, V3<T> vA, V3<T> wA, V3<T> vB, V3<T> wB, V3<T> axis) {
    T jv;
    bool mA = mov & 1, mB = mov & 2;
    if (mA && mB) jv = dot(axis, vA - vB);
    else if (mA) jv = dot(axis, vA);
    else if (mB) jv = -dot(axis, vB);
    else jv = T(0);
    const V3<T> z = mk3<T>(0, 0, 0);
    if (mA) jv += dot((dyn & 1) ? cross(r1, axis) : z, wA);
    if (mB) jv -= dot((dyn & 2) ? cross(r2, axis) : z, wB);
    return jv;
}

// velocity change of an impulse dl along one axis (applyLambda :640-707)
template <typename T>
__device__ inline void axis_impulse(int dyn, int dofA, int dofB, T invMA, T invMB, const M3<T>& IA, const M3<T>& IB,
                                    V3<T> r1, V3<T> r2, V3<T>& vA, V3<T>& wA, V3<T>& vB, V3<T>& wB, V3<T> axis, T dl) {
    if (dyn & 1) {
        T ls = dl * invMA;
        vA = dof_lin(mk3<T>(vA.x - axis.x * ls, vA.y - axis.y * ls, vA.z - axis.z * ls), dofA);
        V3<T> i1 = mul(IA, cross(r1, axis));
        wA = mk3<T>(wA.x - i1.x * dl, wA.y - i1.y * dl, wA.z - i1.z * dl);
    }
    if (dyn & 2) {
        T ls = dl * invMB;
        vB = dof_lin(mk3<T>(vB.x + axis.x * ls, vB.y + axis.y * ls, vB.z + axis.z * ls), dofB);
        V3<T> i2 = mul(IB, cross(r2, axis));
        wB = mk3<T>(wB.x + i2.x * dl, wB.y + i2.y * dl, wB.z + i2.z * dl);
    }
}

// warm start of one axis (warmStart :500-537)
template <typename T>
__device__ inline void axis_warm(int dyn, int dofA, int dofB, T invMA, T invMB, const M3<T>& IA, const M3<T>& IB,
                                 V3<T> r1, V3<T> r2, V3<T>& vA, V3<T>& wA, V3<T>& vB, V3<T>& wB, V3<T> axis, T& lambda,
                                 T ratio) {
    lambda *= ratio;
    if (lambda == T(0)) return;
    if (dyn & 1) { vA = dof_lin(vA - axis * (lambda * invMA), dofA); wA = wA - mul(IA, cross(r1, axis)) * lambda; }
    if (dyn & 2) { vB = dof_lin(vB + axis * (lambda * invMB), dofB); wB = wB + mul(IB, cross(r2, axis)) * lambda; }
}

}  // namespace cc
