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

template <typename T>
struct SettingsD {
    T gravity[3];
    T spec, manifoldTol, colTol, penTol, preserveDist2, minVelRest, slop, beta, maxPen, tSleep, sleepThresh;
    int velIters, posIters, gravityEnabled, warmStarting, allowSleeping, pad;
    unsigned char layerPairs[64];
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
};

// Contact constraint record (src/constraints/contact-constraints.ts:31-95) for one pair slot.
template <typename T>
struct Row {
    V3<T> r1x, r2x, i1, i2;  // r1 x axis, r2 x axis, I1^-1 (r1 x axis), I2^-1 (r2 x axis)
    T eff, lambda, bias;
};
// a row set in structure-of-arrays form: element k of every array belongs to one row
template <typename T>
struct RowSoA {
    V3<T>* r1x;
    V3<T>* r2x;
    V3<T>* i1;
    V3<T>* i2;
    T* eff;
    T* lambda;
    T* bias;
};
template <typename T>
__device__ inline Row<T> row_load(const RowSoA<T>& R, size_t k) {
    Row<T> r;
    r.r1x = R.r1x[k]; r.r2x = R.r2x[k]; r.i1 = R.i1[k]; r.i2 = R.i2[k];
    r.eff = R.eff[k]; r.lambda = R.lambda[k]; r.bias = R.bias[k];
    return r;
}
template <typename T>
__device__ inline void row_store(const RowSoA<T>& R, size_t k, const Row<T>& r) {
    R.r1x[k] = r.r1x; R.r2x[k] = r.r2x; R.i1[k] = r.i1; R.i2[k] = r.i2;
    R.eff[k] = r.eff; R.lambda[k] = r.lambda; R.bias[k] = r.bias;
}
template <typename T>
struct ConPoint {
    Row<T> n, t1, t2;
    V3<T> localA, localB;
};
template <typename T>
struct Con {
    int a, b, np, dynMask;  // dynMask bit0: A dynamic, bit1: B dynamic
    unsigned sortKey;
    int movMask;            // bit0: A not static, bit1: B not static
    int cq, co;             // the broadphase pair (querying body, other body); constraints are
                            // created in ascending (cq, co) order
    V3<T> normal, t1, t2;
    T friction, invMassA, invMassB;
    ConPoint<T> p[4];
};
// Persistent contact cache entry (src/contacts.ts:9-58): four in-place slots.
template <typename T>
struct CacheRec {
    int a, b, np, pad;
    V3<T> p1[4], p2[4];
    T ln[4], lt1[4], lt2[4];
};
// World-space points of a constraint, kept only for the parity taps.
template <typename T>
struct ConTap {
    V3<T> pa[4], pb[4];
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

template <typename T>
__device__ inline T axis_total_lambda(const Row<T>& p, int mov, V3<T> vA, V3<T> wA, V3<T> vB, V3<T> wB, V3<T> axis) {
    // getTotalLambda :578-625
    T jv;
    bool mA = mov & 1, mB = mov & 2;
    if (mA && mB) jv = dot(axis, vA - vB);
    else if (mA) jv = dot(axis, vA);
    else if (mB) jv = -dot(axis, vB);
    else jv = T(0);
    if (mA) jv += dot(p.r1x, wA);
    if (mB) jv -= dot(p.r2x, wB);
    return p.lambda + p.eff * (jv - p.bias);
}

template <typename T>
__device__ inline void axis_apply(Row<T>& p, int dyn, int dofA, int dofB, T invMA, T invMB, V3<T>& vA, V3<T>& wA,
                                  V3<T>& vB, V3<T>& wB, V3<T> axis, T total) {  // applyLambda :640-707
    T dl = total - p.lambda;
    p.lambda = total;
    if (dl == T(0)) return;
    if (dyn & 1) {
        T ls = dl * invMA;
        vA = dof_lin(mk3<T>(vA.x - axis.x * ls, vA.y - axis.y * ls, vA.z - axis.z * ls), dofA);
        wA = mk3<T>(wA.x - p.i1.x * dl, wA.y - p.i1.y * dl, wA.z - p.i1.z * dl);
    }
    if (dyn & 2) {
        T ls = dl * invMB;
        vB = dof_lin(mk3<T>(vB.x + axis.x * ls, vB.y + axis.y * ls, vB.z + axis.z * ls), dofB);
        wB = mk3<T>(wB.x + p.i2.x * dl, wB.y + p.i2.y * dl, wB.z + p.i2.z * dl);
    }
}

template <typename T>
__device__ inline void axis_warm(Row<T>& p, int dyn, int dofA, int dofB, T invMA, T invMB, V3<T>& vA, V3<T>& wA,
                                 V3<T>& vB, V3<T>& wB, V3<T> axis, T ratio) {  // warmStart :500-537
    p.lambda *= ratio;
    if (p.lambda == T(0)) return;
    if (dyn & 1) { vA = dof_lin(vA - axis * (p.lambda * invMA), dofA); wA = wA - p.i1 * p.lambda; }
    if (dyn & 2) { vB = dof_lin(vB + axis * (p.lambda * invMB), dofB); wB = wB + p.i2 * p.lambda; }
}

}  // namespace cc
