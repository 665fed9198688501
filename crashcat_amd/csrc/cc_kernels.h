// Per-phase kernels of one step.  Every kernel that walks a device-sized list uses a
// grid-stride loop bounded by a count read from device memory, so a step launches a
// fixed sequence of kernels with no host round trip (and can be captured in a graph).
#pragma once
#include "cc_step.h"

namespace cc {

template <typename T>
struct Dev {
    SettingsD<T> s;
    ShapeTable<T> st;
    Bodies<T> B;
    int N;              // bodies (all worlds)
    int W;              // worlds
    const int* wStart;  // [W+1] body ranges
    int pairCap, conCap, epaCap, epaWorkers, hashCap;
    // broadphase
    int* bpCount;       // [N+1]
    int* bpOff;         // [N+1] exclusive scan
    int* pairA;         // [pairCap]
    int* pairB;
    int* counters;      // device counters, see CNT_*
    // contact constraints (slots allocated in creation order by an atomic counter)
    Con<T>* con;        // [conCap]
    ConTap<T>* tap;     // [conCap]
    int* wConCnt;       // [W + 1] per-world constraint counts
    int* wConStart;     // [W + 1] exclusive scan of wConCnt
    int* wFill;         // [W]
    int* conIdx;        // [conCap] constraint slots grouped by world
    CacheRec<T>* rec[2];          // [conCap] per parity
    unsigned long long* hkeys[2]; // [hashCap]
    int* hvals[2];
    // EPA jobs and the per-worker polytope workspace
    int* epaPair;       // [epaCap]
    Simplex<T>* epaSimplex;
    EpaState<T>* epaWork;  // [epaWorkers]
    // point constraints
    PointCon<T>* pc;
    int NPC;
    const int* pcStart; // [W+1]
    // per-world solver workspace (sequence = world's contacts then its point constraints)
    unsigned long long* skeys;  // [2 * (conCap + NPC)] solve-order keys (hi, lo)
    int* seq;           // [conCap + NPC]
    int* seqA;
    int* seqB;
    int* prevA;
    int* prevB;
    int* nextA;
    int* nextB;
    int* level;         // in-degree during the level construction
    int* sched;
    int* lvlStart;      // [conCap + NPC + W + 1]
    int* blist;         // [2 * (conCap + NPC)]
    int* bcnt;          // [N]
    int* boff;          // [N]
    int* isl0;          // [N] island flags (by label)
    int* isl1;          // [N]
    int* stats;         // per world: [levels, posSweeps]
    T* prevDt;          // [1]
    int* parity;        // [1]
    unsigned long long* acc;  // [ACC_N] running sums over steps, see ACC_*
};

// running sums (64-bit) over all steps since upload: counts and the algorithmic bytes of
// the byte model (per-item compulsory HBM traffic of each phase, 4-byte words)
enum { ACC_STEPS = 0, ACC_PAIRS, ACC_CONS, ACC_POINTS, ACC_EPA, ACC_AWAKE,
       ACC_BYTES_BODY, ACC_BYTES_BP, ACC_BYTES_NP, ACC_BYTES_SOLVE, ACC_N = 16 };

enum { CNT_PAIRS = 0, CNT_CONS = 1, CNT_EPA = 2, CNT_OVERFLOW = 3, CNT_EPA_TOTAL = 4, CNT_N = 8 };

// ------------------------------------------------------------ acceleration
// accelerationIntegrationUpdate src/update.ts:245-319 (external forces are zero on this path)
template <typename T>
__global__ void k_accel(Dev<T> d, T dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    const Bodies<T>& B = d.B;
    int f = B.flags[i];
    if (fl_motion(f) != MOTION_DYNAMIC || B.sleeping[i]) return;
    T gs = d.s.gravityEnabled ? B.gravityFactor[i] : T(0);
    V3<T> v = B.v[i];
    v.x += (T(0) * B.invMass[i] + d.s.gravity[0] * gs) * dt;
    v.y += (T(0) * B.invMass[i] + d.s.gravity[1] * gs) * dt;
    v.z += (T(0) * B.invMass[i] + d.s.gravity[2] * gs) * dt;
    T ld = tmax<T>(T(0), T(1) - B.linDamp[i] * dt);
    v = v * ld;
    T s2 = len2(v), ml = B.maxLinV[i];
    if (s2 > ml * ml) v = v * (ml / sqrt(s2));
    int dof = fl_dof(f);
    v = dof_lin(v, dof);
    B.v[i] = v;
    V3<T> w = B.w[i];
    T ad = tmax<T>(T(0), T(1) - B.angDamp[i] * dt);
    w = w * ad;
    T a2 = len2(w), ma = B.maxAngV[i];
    if (a2 > ma * ma) w = w * (ma / sqrt(a2));
    w = dof_lin(w, dof >> 3);
    B.w[i] = w;
}

// ---------------------------------------------------------------- broadphase
// Pair predicate of findCollidingPairs + CollisionBodyPairVisitor (src/broadphase/broadphase.ts:44-138,
// 217-277) over the world's bodies; the AABB test is evaluated in FP64 from the stored bounds so that
// the pair set does not depend on the bulk precision.
template <typename T>
__device__ inline bool queries(const Bodies<T>& B, int i) { return fl_motion(B.flags[i]) != MOTION_STATIC && !B.sleeping[i]; }

template <typename T>
__device__ inline bool pair_ok(const Dev<T>& d, int a, int b) {
    const Bodies<T>& B = d.B;
    if (a == b) return false;
    int fa = B.flags[a], fb = B.flags[b];
    if (!d.s.layerPairs[fl_layer(fa) * 8 + fl_layer(fb)]) return false;
    if (!((B.groups[a] & B.mask[b]) != 0 && (B.groups[b] & B.mask[a]) != 0)) return false;
    long long da = queries(B, a) ? a : 0x7fffffffll, db = queries(B, b) ? b : 0x7fffffffll;
    if (da >= db) return false;
    int ma = fl_motion(fa), mb = fl_motion(fb);
    if (ma != MOTION_DYNAMIC && mb != MOTION_DYNAMIC) return false;
    return true;
}

template <typename T>
__device__ inline bool aabb_overlap(const Dev<T>& d, int a, int b, double m) {
    V3<double> amn = cvt<double>(d.B.amin[a]), amx = cvt<double>(d.B.amax[a]);
    V3<double> bmn = cvt<double>(d.B.amin[b]), bmx = cvt<double>(d.B.amax[b]);
    amn = amn - mk3<double>(m, m, m);
    amx = amx + mk3<double>(m, m, m);
    return !(bmx.x < amn.x || bmn.x > amx.x || bmx.y < amn.y || bmn.y > amx.y || bmx.z < amn.z || bmn.z > amx.z);
}

// One querying body per lane; the world's bodies are streamed through LDS tiles.
template <typename T, bool WRITE>
__global__ void k_bp(Dev<T> d, double margin) {
    __shared__ V3<T> smin[256], smax[256];
    __shared__ int sflag[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < d.N && queries(d.B, i);
    // all lanes of a block belong to consecutive bodies; the block covers the union of their worlds
    int iFirst = blockIdx.x * blockDim.x;
    int iLast = min(d.N - 1, iFirst + (int)blockDim.x - 1);
    int w0 = d.B.world[iFirst], w1 = d.B.world[iLast];
    int jb = d.wStart[w0], je = d.wStart[w1 + 1];
    int myW = i < d.N ? d.B.world[i] : -1;
    int jw0 = myW >= 0 ? d.wStart[myW] : 0, jw1 = myW >= 0 ? d.wStart[myW + 1] : 0;
    V3<double> amn, amx;
    if (active) {
        amn = cvt<double>(d.B.amin[i]) - mk3<double>(margin, margin, margin);
        amx = cvt<double>(d.B.amax[i]) + mk3<double>(margin, margin, margin);
    }
    int cnt = 0;
    int out = 0;
    if (WRITE && active) out = d.bpOff[i];
    for (int t0 = jb; t0 < je; t0 += blockDim.x) {
        int j = t0 + threadIdx.x;
        __syncthreads();
        if (j < je) { smin[threadIdx.x] = d.B.amin[j]; smax[threadIdx.x] = d.B.amax[j]; sflag[threadIdx.x] = 1; }
        __syncthreads();
        if (active) {
            int tn = min((int)blockDim.x, je - t0);
            for (int k = 0; k < tn; k++) {
                int jj = t0 + k;
                if (jj < jw0 || jj >= jw1) continue;
                V3<double> bmn = cvt<double>(smin[k]), bmx = cvt<double>(smax[k]);
                if (bmx.x < amn.x || bmn.x > amx.x || bmx.y < amn.y || bmn.y > amx.y || bmx.z < amn.z || bmn.z > amx.z) continue;
                if (!pair_ok(d, i, jj)) continue;
                if (WRITE) {
                    if (out < d.pairCap) { d.pairA[out] = i; d.pairB[out] = jj; }
                    out++;
                } else {
                    cnt++;
                }
            }
        }
    }
    if (!WRITE && i < d.N) d.bpCount[i] = cnt;
}

// --------------------------------------------------------------- scans
// Exclusive scan of n ints (n <= 4096 * 4096) in three launches.
__global__ void k_scan_blocks(const int* in, int* out, int* blockSums, int n);
__global__ void k_scan_sums(int* blockSums, int nb, int* total);
__global__ void k_scan_add(int* out, const int* blockSums, int n);

// ------------------------------------------------------------- narrowphase
template <typename T>
struct PairBodies {
    int a, b;  // after the motion-type swap (src/update.ts:74-78)
};

template <typename T>
__device__ inline PairBodies<T> order_pair(const Dev<T>& d, int a, int b) {
    int ma = fl_motion(d.B.flags[a]), mb = fl_motion(d.B.flags[b]);
    PairBodies<T> r;
    if (ma > mb || (ma == mb && b < a)) { r.a = b; r.b = a; }
    else { r.a = a; r.b = b; }
    return r;
}

// normal velocity bias, src/constraints/contact-constraints.ts:288-423 (no external forces)
template <typename T>
__device__ inline T normal_bias(const Dev<T>& d, int A, int Bi, V3<T> pA, V3<T> pB, V3<T> n, V3<T> rA, V3<T> rB, T e, T dt) {
    const Bodies<T>& B = d.B;
    int ma = fl_motion(B.flags[A]), mb = fl_motion(B.flags[Bi]);
    bool mA = ma != MOTION_STATIC, mB = mb != MOTION_STATIC;
    V3<T> rel;
    if (mA && mB) rel = (B.v[Bi] + cross(B.w[Bi], rB)) - (B.v[A] + cross(B.w[A], rA));
    else if (mA) rel = -(B.v[A] + cross(B.w[A], rA));
    else if (mB) rel = B.v[Bi] + cross(B.w[Bi], rB);
    else return T(0);
    T nv = dot(rel, n);
    T pen = dot(pA - pB, n);
    T spec = tmax<T>(T(0), -pen / dt);
    if (e > T(0) && nv < -d.s.minVelRest) {
        if (nv < -spec) {
            V3<T> g = mk3<T>(d.s.gravity[0], d.s.gravity[1], d.s.gravity[2]);
            V3<T> acc = mk3<T>(0, 0, 0);
            if (mA && mB) acc = g * (B.gravityFactor[Bi] - B.gravityFactor[A]);
            else if (mA) acc = g * (-B.gravityFactor[A]);
            else if (mB) acc = g * B.gravityFactor[Bi];
            T fdv = tmin<T>(T(0), dot(acc, n) * dt);
            return e * (nv - fdv);
        }
        return spec;
    }
    return spec;
}

// Finish one hit: manifold (update.ts:385-563 for the single hit of a convex pair), then
// addContactConstraint (contact-constraints.ts:437-857) into the pair slot.
template <typename T>
__device__ inline void finish_pair(const Dev<T>& d, int slot, int A, int Bi, V3<T> pointA, V3<T> pointB, V3<T> axis,
                                   const Face<T>& fA, const Face<T>& fB, Face<T>* scratch3, Manifold<T>& m, T dt, int par) {
    const Bodies<T>& B = d.B;
    m.base = B.pos[A];
    m.normal = normalize(axis);
    m.n = 0;
    T maxDist = d.s.spec + d.s.manifoldTol;
    if (fA.n >= 2 && fB.n >= 3) manifold_between_two_faces(m, pointA, pointB, axis, maxDist, fA, fB, scratch3);
    else { m.relA[0] = pointA - m.base; m.relB[0] = pointB - m.base; m.n = 1; }
    m.normal = normalize(m.normal);
    if (m.n > 4) prune(m, m.normal);
    if (m.n <= 0) return;
    // swap so that id A < id B (manifold.swapShapes :86-107)
    int ia = A, ib = Bi;
    V3<T> n = m.normal;
    bool sw = ia > ib;
    if (sw) { int t = ia; ia = ib; ib = t; n = -n; }
    int fa = B.flags[ia], fb = B.flags[ib];
    bool dynA = fl_motion(fa) == MOTION_DYNAMIC, dynB = fl_motion(fb) == MOTION_DYNAMIC;
    if (!((dynA && B.invMass[ia] != T(0)) || (dynB && B.invMass[ib] != T(0)))) return;
    int cs = atomicAdd(&d.counters[CNT_CONS], 1);
    if (cs >= d.conCap) { atomicOr(&d.counters[CNT_OVERFLOW], 2); return; }
    Con<T>& k = d.con[cs];
    ConTap<T>& tp = d.tap[cs];
    CacheRec<T>& rc = d.rec[par][cs];
    // previous record of this contact (findContact src/contacts.ts:481-513)
    unsigned long long key = pair_key(ia, ib);
    int prev = -1;
    {
        const unsigned long long* hk = d.hkeys[par ^ 1];
        unsigned h = hash64(key) & (unsigned)(d.hashCap - 1);
        for (int probe = 0; probe < d.hashCap; probe++) {
            unsigned long long kk = hk[h];
            if (kk == CC_EMPTY_KEY) break;
            if (kk == key) { prev = d.hvals[par ^ 1][h]; break; }
            h = (h + 1) & (unsigned)(d.hashCap - 1);
        }
    }
    if (prev >= 0) rc = d.rec[par ^ 1][prev];
    else {
        rc.np = 0;
        for (int j = 0; j < 4; j++) { rc.p1[j] = rc.p2[j] = mk3<T>(0, 0, 0); rc.ln[j] = rc.lt1[j] = rc.lt2[j] = T(0); }
    }
    rc.a = ia; rc.b = ib;
    rc.np = m.n;
    Q4<T> qa = B.q[ia], qb = B.q[ib];
    M3<T> rotA = m3_from_quat(qa), rotB = m3_from_quat(qb);
    T fr = combine(B.friction[ia], B.friction[ib], fl_fmode(fa), fl_fmode(fb));
    T re = combine(B.restitution[ia], B.restitution[ib], fl_rmode(fa), fl_rmode(fb));
    k.a = ia; k.b = ib;
    k.pairIdx = slot;
    k.dynMask = (dynA ? 1 : 0) | (dynB ? 2 : 0);
    k.movMask = (fl_motion(fa) != MOTION_STATIC ? 1 : 0) | (fl_motion(fb) != MOTION_STATIC ? 2 : 0);
    k.sortKey = sort_key((unsigned)ia, (unsigned)ib, 0xffffffffu, 0xffffffffu);
    k.normal = n;
    if (fabs(n.x) > fabs(n.y)) { T tl = sqrt(n.x * n.x + n.z * n.z); k.t1 = mk3<T>(n.z / tl, 0, -n.x / tl); }
    else { T tl = sqrt(n.y * n.y + n.z * n.z); k.t1 = mk3<T>(0, n.z / tl, -n.y / tl); }
    k.t2 = cross(n, k.t1);
    k.friction = fr;
    k.invMassA = B.invMass[ia];
    k.invMassB = B.invMass[ib];
    M3<T> IA, IB;
    for (int j = 0; j < 9; j++) { IA.m[j] = T(0); IB.m[j] = T(0); }
    if (dynA) IA = inv_inertia(B.invI[ia], B.iRot[ia], qa, fl_dof(fa));
    if (dynB) IB = inv_inertia(B.invI[ib], B.iRot[ib], qb, fl_dof(fb));
    V3<T> comA = B.com[ia], comB = B.com[ib];
    k.np = m.n;
    for (int i = 0; i < m.n; i++) {
        ConPoint<T>& cp = k.p[i];
        V3<T> ra_ = sw ? m.relB[i] : m.relA[i];
        V3<T> rb_ = sw ? m.relA[i] : m.relB[i];
        V3<T> pA = m.base + ra_, pB = m.base + rb_;
        tp.pa[i] = pA;
        tp.pb[i] = pB;
        cp.localA = mul_t(rotA, pA - comA);
        cp.localB = mul_t(rotB, pB - comB);
        bool set = false;
        if (prev >= 0) {  // first match over the in-place slots (slots < i already hold this step's points)
            for (int j = 0; j < rc.np; j++) {
                T d1 = len2(cp.localA - rc.p1[j]), d2 = len2(cp.localB - rc.p2[j]);
                if (d1 < d.s.preserveDist2 && d2 < d.s.preserveDist2) {
                    cp.n.lambda = rc.ln[j]; cp.t1.lambda = rc.lt1[j]; cp.t2.lambda = rc.lt2[j];
                    set = true;
                    break;
                }
            }
        }
        if (!set) { cp.n.lambda = cp.t1.lambda = cp.t2.lambda = T(0); }
        V3<T> mid = mk3<T>((pA.x + pB.x) * T(0.5), (pA.y + pB.y) * T(0.5), (pA.z + pB.z) * T(0.5));
        V3<T> rA = mid - comA, rB = mid - comB;
        T bias = normal_bias(d, ia, ib, pA, pB, n, rA, rB, re, dt);
        axis_props(cp.n, dynA, dynB, k.invMassA, k.invMassB, IA, IB, rA, rB, n, bias);
        if (fr > T(0)) {
            axis_props(cp.t1, dynA, dynB, k.invMassA, k.invMassB, IA, IB, rA, rB, k.t1, T(0));
            axis_props(cp.t2, dynA, dynB, k.invMassA, k.invMassB, IA, IB, rA, rB, k.t2, T(0));
        } else {
            cp.t1.eff = T(0); cp.t1.lambda = T(0); cp.t1.bias = T(0);
            cp.t2.eff = T(0); cp.t2.lambda = T(0); cp.t2.bias = T(0);
            cp.t1.r1x = cp.t1.r2x = cp.t1.i1 = cp.t1.i2 = mk3<T>(0, 0, 0);
            cp.t2.r1x = cp.t2.r2x = cp.t2.i1 = cp.t2.i2 = mk3<T>(0, 0, 0);
        }
        rc.p1[i] = cp.localA;
        rc.p2[i] = cp.localB;
        rc.ln[i] = cp.n.lambda;
        rc.lt1[i] = cp.t1.lambda;
        rc.lt2[i] = cp.t2.lambda;
    }
    if (dynA && B.sleeping[ia]) B.wake[ia] = 1;
    if (dynB && B.sleeping[ib]) B.wake[ib] = 1;
}

template <typename T>
struct NpScratch {
    Face<T> fA, fB, tmp[3];
    Manifold<T> m;
};

// world-space hit + supporting faces after GJK/EPA (convex.ts:489-637)
template <typename T>
__device__ inline void finish_convex(const Dev<T>& d, int slot, int A, int Bi, const ConvexCtx<T>& c, PenDepth<T> pd,
                                     NpScratch<T>& S, T dt, int par) {
    const Bodies<T>& B = d.B;
    V3<T> pa = B.pos[A];
    Q4<T> qa = B.q[A];
    T ms = c.maxSepUse;
    // penetration = |pA - pB| - maxSep; the early-out fraction is +inf for the step collector
    T al = sqrt(len2(pd.axis));
    if (al > T(0)) { T corr = ms / al; pd.pointA = pd.pointA + pd.axis * (-corr); }
    V3<T> wA = qrot2(qa, pd.pointA) + pa;
    V3<T> wB = qrot2(qa, pd.pointB) + pa;
    V3<T> wAxis = qrot2(qa, pd.axis);
    supporting_face(S.fA, d.st, B.shape[A], -pd.axis, pa, qa);
    Q4<T> crel = qconj(c.rel);
    V3<T> dirB = qrot2(crel, pd.axis);
    V3<T> posB = qrot2(qa, c.relPos) + pa;
    Q4<T> qB;
    qB.x = qa.w * c.rel.x + qa.x * c.rel.w + qa.y * c.rel.z - qa.z * c.rel.y;
    qB.y = qa.w * c.rel.y - qa.x * c.rel.z + qa.y * c.rel.w + qa.z * c.rel.x;
    qB.z = qa.w * c.rel.z + qa.x * c.rel.y - qa.y * c.rel.x + qa.z * c.rel.w;
    qB.w = qa.w * c.rel.w - qa.x * c.rel.x - qa.y * c.rel.y - qa.z * c.rel.z;
    supporting_face(S.fB, d.st, B.shape[Bi], dirB, posB, qB);
    finish_pair(d, slot, A, Bi, wA, wB, wAxis, S.fA, S.fB, S.tmp, S.m, dt, par);
}

// One pair per lane: sphere-sphere closed form or the GJK step of collideConvexVsConvexLocal;
// pairs whose cores overlap are queued for the EPA kernel.
template <typename T>
__global__ void __launch_bounds__(256) k_narrowphase(Dev<T> d, T dt) {
    int npairs = min(d.counters[CNT_PAIRS], d.pairCap);
    int par = *d.parity;
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < npairs; slot += gridDim.x * blockDim.x) {
        PairBodies<T> pb = order_pair(d, d.pairA[slot], d.pairB[slot]);
        int A = pb.a, Bi = pb.b;
        const Bodies<T>& B = d.B;
        const ShapeD<T>& sa = d.st.shapes[B.shape[A]];
        const ShapeD<T>& sb = d.st.shapes[B.shape[Bi]];
        NpScratch<T> S;
        if (sa.type == SHAPE_SPHERE && sb.type == SHAPE_SPHERE) {  // collideSphereVsSphere sphere.ts:319-410
            V3<T> pa = B.pos[A], pbp = B.pos[Bi];
            T ra = sa.radius, rb = sb.radius;
            V3<T> dd = pbp - pa;
            T d2 = len2(dd), rs = ra + rb;
            if (d2 >= rs * rs) continue;
            T dist = sqrt(d2);
            V3<T> n = dist > T(0) ? dd * (T(1) / dist) : mk3<T>(0, 1, 0);
            S.fA.n = 0;
            S.fB.n = 0;
            finish_pair(d, slot, A, Bi, pa + n * ra, pbp - n * rb, n, S.fA, S.fB, S.tmp, S.m, dt, par);
            continue;
        }
        ConvexCtx<T> c;
        convex_setup(c, B.pos[A], B.q[A], B.pos[Bi], B.q[Bi], d.s.spec);
        Sup<T> SA, SB;
        sup_for_shape(SA, d.st, B.shape[A], false);
        sup_for_shape(SB, d.st, B.shape[Bi], false);
        SB.transformed = true;
        SB.R = m3_from_quat(c.rel);
        SB.t = c.relPos;
        V3<T> axis0 = c.relPos;
        if (len2(axis0) <= T(1e-12)) axis0 = mk3<T>(1, 0, 0);
        PenDepth<T> pd;
        Simplex<T> simplex;
        simplex.size = 0;
        pd_step_gjk(pd, simplex, SA, SB, SA.cr + c.maxSepUse, SB.cr, axis0, d.s.colTol);
        if (pd.status == PD_NOT_COLLIDING) continue;
        if (pd.status == PD_INDETERMINATE) {
            int j = atomicAdd(&d.counters[CNT_EPA], 1);
            if (j < d.epaCap) { d.epaPair[j] = slot; d.epaSimplex[j] = simplex; }
            else atomicOr(&d.counters[CNT_OVERFLOW], 4);
            continue;
        }
        finish_convex(d, slot, A, Bi, c, pd, S, dt, par);
    }
}

template <typename T>
__global__ void __launch_bounds__(64) k_epa(Dev<T> d, T dt) {
    int nj = min(d.counters[CNT_EPA], d.epaCap);
    int par = *d.parity;
    int worker = blockIdx.x * blockDim.x + threadIdx.x;
    if (worker >= d.epaWorkers) return;
    for (int j = worker; j < nj; j += gridDim.x * blockDim.x) {
        int slot = d.epaPair[j];
        PairBodies<T> pb = order_pair(d, d.pairA[slot], d.pairB[slot]);
        int A = pb.a, Bi = pb.b;
        const Bodies<T>& B = d.B;
        ConvexCtx<T> c;
        convex_setup(c, B.pos[A], B.q[A], B.pos[Bi], B.q[Bi], d.s.spec);
        c.maxSepUse = tmin<T>(c.maxSepUse, T(1));
        Sup<T> SA, SB;
        sup_for_shape(SA, d.st, B.shape[A], true);
        sup_for_shape(SB, d.st, B.shape[Bi], true);
        SA.addRadius = true;
        SA.addR = c.maxSepUse;
        SB.transformed = true;
        SB.R = m3_from_quat(c.rel);
        SB.t = c.relPos;
        PenDepth<T> pd;
        Simplex<T> simplex = d.epaSimplex[j];
        if (!pd_step_epa(pd, SA, SB, d.s.penTol, simplex, d.epaWork[worker])) continue;
        NpScratch<T> S;
        finish_convex(d, slot, A, Bi, c, pd, S, dt, par);
    }
}

// wake bodies that received a new constraint (src/update.ts:84-91, src/body/sleep.ts:188-203)
template <typename T>
__global__ void k_wake(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    if (d.B.wake[i]) {
        d.B.wake[i] = 0;
        d.B.sleeping[i] = 0;
        reset_sleep_timer(d.B, d.st, i);
    }
}

// group constraint slots by world (counting sort); the order inside a world is irrelevant
// because the solver orders constraints by their keys
template <typename T>
__global__ void k_con_world_count(Dev<T> d) {
    int nc = min(d.counters[CNT_CONS], d.conCap);
    unsigned long long pts = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        atomicAdd(&d.wConCnt[d.B.world[d.con[k].a]], 1);
        pts += (unsigned long long)d.con[k].np;
    }
    for (int o = 32; o > 0; o >>= 1) pts += __shfl_xor(pts, o, 64);
    if ((threadIdx.x & 63) == 0 && pts) atomicAdd(&d.acc[ACC_POINTS], pts);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&d.acc[ACC_PAIRS], (unsigned long long)min(d.counters[CNT_PAIRS], d.pairCap));
        atomicAdd(&d.acc[ACC_CONS], (unsigned long long)nc);
        atomicAdd(&d.acc[ACC_EPA], (unsigned long long)min(d.counters[CNT_EPA], d.epaCap));
        atomicAdd(&d.acc[ACC_STEPS], 1ull);
    }
}

template <typename T>
__global__ void k_con_world_scatter(Dev<T> d) {
    int nc = min(d.counters[CNT_CONS], d.conCap);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        int w = d.B.world[d.con[k].a];
        d.conIdx[d.wConStart[w] + atomicAdd(&d.wFill[w], 1)] = k;
    }
}

// ----------------------------------------------------------------- islands
// Lock-free union-find: roots are always hooked under the smaller root, so the final
// label of an island is its lowest body index.  Active bodies = non-static, awake.
template <typename T>
__global__ void k_uf_init(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    bool act = fl_motion(d.B.flags[i]) != MOTION_STATIC && !d.B.sleeping[i];
    d.B.label[i] = act ? i : -1;
}

// Parents only ever point to smaller indices and only roots are hooked (by CAS), so a
// non-root x may be re-pointed at any ancestor: path halving stores the grandparent.
__device__ inline int uf_find(int* p, int x) {
    for (;;) {
        int y = __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y == x) return x;
        int z = __hip_atomic_load(&p[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (z == y) return y;
        __hip_atomic_store(&p[x], z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = z;
    }
}
__device__ inline void uf_union(int* p, int a, int b) {
    for (;;) {
        int ra = uf_find(p, a), rb = uf_find(p, b);
        if (ra == rb) return;
        if (ra < rb) { int t = ra; ra = rb; rb = t; }
        if (atomicCAS(&p[ra], ra, rb) == ra) return;
    }
}

template <typename T>
__global__ void k_uf_link(Dev<T> d) {
    int nc = min(d.counters[CNT_CONS], d.conCap);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc + d.NPC; k += gridDim.x * blockDim.x) {
        int a, b;
        if (k < nc) {
            const Con<T>& c = d.con[k];
            a = c.a; b = c.b;
        } else {
            const PointCon<T>& pc = d.pc[k - nc];
            if (!pc.enabled) continue;
            a = pc.a; b = pc.b;
        }
        if (d.B.label[a] >= 0 && d.B.label[b] >= 0) uf_union(d.B.label, a, b);
    }
}

template <typename T>
__global__ void k_uf_compress(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    if (d.B.label[i] >= 0) d.B.label[i] = uf_find(d.B.label, i);
}

// --------------------------------------------------------- contact cache hash
__global__ void k_hash_clear(unsigned long long* keys, int cap);

template <typename T>
__global__ void k_hash_insert(Dev<T> d) {
    int nc = min(d.counters[CNT_CONS], d.conCap);
    int par = *d.parity;
    unsigned long long* hk = d.hkeys[par];
    int* hv = d.hvals[par];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        int slot = k;
        const Con<T>& c = d.con[slot];
        unsigned long long key = pair_key(c.a, c.b);
        unsigned h = hash64(key) & (unsigned)(d.hashCap - 1);
        for (int probe = 0; probe < d.hashCap; probe++) {
            unsigned long long old = atomicCAS(&hk[h], CC_EMPTY_KEY, key);
            if (old == CC_EMPTY_KEY || old == key) { hv[h] = slot; break; }
            h = (h + 1) & (unsigned)(d.hashCap - 1);
        }
    }
}

template <typename T>
__global__ void k_flip_parity(Dev<T> d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d.parity ^= 1;
}

}  // namespace cc
