// Per-phase kernels of one step.  Every kernel that walks a device-sized list uses a
// grid-stride loop bounded by a count read from device memory, so a step launches a
// fixed sequence of kernels with no host round trip (and can be captured in a graph).
#pragma once
#include <type_traits>

#include "cc_step.h"

namespace cc {

// What the narrowphase reads of a body, in two records of one memory line each: the pose part
// (GJK, EPA and manifolds) and the part only the contact set-up reads.  Written once per step
// before the broadphase, when the poses and velocities are those of the collision phase.
// tinfo: shape type | (most vertices of a supporting face) << 4; brad: bounding radius of the
// shape about the body position.
template <typename T>
struct alignas(64) NpBody {
    V3<T> pos;
    Q4<T> q;
    int shape, flags;
    T brad;
    int tinfo;
    V3<T> com;
    T invMass;
};
template <typename T>
struct alignas(64) NpDyn {
    V3<T> v, w;
    V3<T> invI;
    Q4<T> iRot;
    T friction, restitution;
    int sleeping;
};

// a body in the cell-sorted broadphase arrays: one record, so that staging a neighbour cell
// reads its bodies' lines once
template <typename T>
struct alignas(8) CellRec {
    V3<T> mn, mx;
    int body, fl;
    unsigned gr, mk;
};

template <typename T>
struct Dev {
    SettingsD<T> s;
    ShapeTable<T> st;
    Bodies<T> B;
    int N;              // bodies (all worlds)
    int W;              // worlds
    const int* wStart;  // [W+1] body ranges
    // the solver's workgroups: (world, island group, groups of the world, body scratch base),
    // largest worlds first.  A world of few worlds is split into island groups (islands never
    // share a body, so each group is solved alone, in the reference order inside each island)
    const int4* vTab;   // [V]
    int V;
    const int* worldG;      // [W] island groups of each world (1: not split)
    const int* worldVBase;  // [W] index of the world's first group (groups are numbered world by world)
    int* vCnt;              // [2 * V] contacts, active point constraints of each group
    int pairCap, conCap, epaCap, epaWorkers, hashCap;
    // broadphase: uniform grid per world, occupied cells in an open-addressing hash table
    double cellWidth;   // widest box a body may have to be binned (wider ones are "large")
    double cellSize, cellInv;  // cellWidth + 2 * margin, so overlapping small bodies are in adjacent cells
    int cellTabCap;     // power of two, at least 2 * N
    unsigned long long* cellKey;  // [cellTabCap] cell key per table slot (CC_EMPTY_KEY if free)
    int* cellCount;     // [cellTabCap + 1] bodies per slot
    int* cellStart;     // [cellTabCap + 1] exclusive scan of cellCount
    int* cellFill;      // [cellTabCap]
    int* cellList;      // [N] occupied slots (CNT_CELLS of them)
    int* bodySlot;      // [N] table slot of a small body, -1 for a large one
    CellRec<T>* cellRec;  // [N] small bodies grouped by cell: body, bounds and the words of the
                          // pair predicate (flags with bit 30 = querying, groups, mask)
    int* gLargeFlag;    // [N + 1]
    int* gLargeOff;     // [N + 1] exclusive scan of gLargeFlag (bodies are world-ordered)
    int* gLarge;        // [N] large bodies, grouped by world
    int* gjkList;       // [pairCap] pair slots that passed the filter
    int* gjkCombo;      // [pairCap] their (shape A kind, shape B kind) as kindA * 4 + kindB, cylinders with hulls
    int* gjkSorted;     // [pairCap] the same slots ordered by type pair, for GJK
    int* gjkHist;       // [2 * 16] per type pair: count, then next free position
    NpBody<T>* npb;     // [N] narrowphase records of each body
    NpDyn<T>* npd;      // [N]
    T* bBrad;           // [N] bounding radius of each body's shape (fixed at upload)
    int* bTinfo;        // [N] shape type and face size of each body's shape (fixed at upload)
    int* bpCount;       // [N+1] pairs found by each querying body
    int* bpOff;         // [N+1] exclusive scan of bpCount: pairs are stored in body order
    int* bpSlots;       // [N * BP_SLOTS] first BP_SLOTS partners of each querying body
    int* spillA;        // [spillCap] (body, partner, rank within the body) beyond a body's BP_SLOTS
    int* spillB;
    int* spillK;
    int spillCap;
    int* pairCon;       // [pairCap] first constraint slot made from each pair (a run of consecutive
                        // slots, its length in the first head's bits), -1 if none
    int hasMesh;        // a shape is a triangle mesh: pairs may make several constraints (so may compounds)
    int* meshList;      // [pairCap] pair slots of convex-vs-mesh pairs
    int meshWorkers;    // threads of the mesh kernel (each owns a collector and an EPA polytope)
    char* meshWork;     // [meshWorkers * sizeof(MeshColl<T>)] collectors
    int* meshCnt;       // [pairCap + 1] candidate triangles of each mesh pair
    int* meshOff;       // [pairCap + 1] exclusive scan of meshCnt
    int candCap;
    int* candTri;       // [candCap] candidate triangle (index within its mesh)
    int* candPair;      // [candCap] its mesh pair (index into meshList)
    void* candHit;      // [candCap] MeshHit<T>: the candidate's GJK/EPA result
    int* conFlag;       // [pairCap + 1]
    int* conPos;        // [pairCap + 1] exclusive scan of conFlag
    int* pairA;         // [pairCap]
    int* pairB;
    int* counters;      // device counters, see CNT_*
    int hasPlane;       // a plane shape is uploaded
    int* planeList;     // [pairCap] pair slots of convex-vs-plane pairs (when hasPlane)
    // contact constraints (slots allocated in creation order by an atomic counter)
    Con<T>* con;        // [conCap]
    ConHead* conHead;   // [conCap]
    ConTap<T>* tap;     // [conCap]
    int* wConStart;     // [W + 1] range of each world in conIdx
    int* conIdx;        // [conCap] constraint slots in pair order (so grouped by world)
    CacheRec<T>* rec[2];          // [conCap] per parity
    int* recHit[2];               // [conCap] per parity: 0 when written, 1 once the next step matched it
    int* recCnt;                  // [2] records written per parity
    unsigned long long* hkeys[2]; // [hashCap]
    int* hvals[2];
    // narrowphase hits awaiting their manifold: pair slot, kind (HIT_*), and the closest points
    // and axis (in body A's frame for convex pairs, in world space for sphere pairs), kept in FP64
    // so that an FP64 manifold of an FP32 context starts from GJK's and EPA's unrounded results
    int hitCap;
    int* hitSlot;
    int* hitKind;
    V3<double>* hitPA;
    V3<double>* hitPB;
    V3<double>* hitAxis;
    // EPA jobs and the per-worker polytope workspace
    int* epaPair;       // [epaCap]
    Simplex<T>* epaSimplex;
    char* epaWork;         // [ceil(epaWorkers / 64) * max(EpaWave<T>, EpaWave<double>)::bytes] polytopes,
                           // one block per wavefront
    // FP32 contexts with npWide = 1 run GJK and EPA in FP64 from the FP32 poses, with the shapes
    // and tolerances in FP64 beside the FP32 ones and the EPA jobs' simplices in FP64: their
    // discrete choices (EPA's last face, GJK's closest points between parallel features) then go
    // as the FP64 path's from the same poses; the manifolds of convex pairs (supporting faces,
    // clipping, pruning) are then computed in FP64 too
    int npWide;
    ShapeTable<double> st64;
    double spec64, colTol64, penTol64, manifoldTol64, normalCos64;
    // compound pairs (compound.ts): listed by the filter, each solved by one thread that owns a
    // reduction collector (cmpColl) and an EPA polytope (cmpEpa, one block per 64 threads)
    int hasCompound;
    // CCD (MotionQuality.LINEAR_CAST): per world the range of its records [ccdStart[w],
    // ccdStart[w + 1]) (one per dynamic LINEAR_CAST body), the records in body order and their
    // time-of-impact order, each body's record this step (-1 for none), the two constraint parts
    // of each world and each body's shape inner radius
    int hasCcd;
    int* ccdStart;
    int* ccdCount;  // per world: the records of the last step
    CcdRec<T>* ccd;
    int* ccdOrd;
    int* ccdIdx;
    CcdPart<T>* ccdPart;
    T* ccdInner;
    int* cmpList;       // [pairCap]
    int cmpWorkers;
    char* cmpColl;
    char* cmpEpa;
    Simplex<double>* epaSimplex64;  // [epaCap] when npWide
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
    int* elemNp;        // [conCap + NPC] points of each sequence element (0: not a contact)
    int* lvlStart;      // [conCap + NPC + W + 1]
    int* blist;         // [2 * (conCap + NPC)]
    unsigned long long* blkey;  // [4 * (conCap + NPC)] solve-order key (hi, lo) of each blist entry
    int worldLdsInts;   // ints of dynamic LDS each world's workgroup has
    int* bcnt;          // [N]
    int* boff;          // [N]
    // the world's constraints copied in level order for the sweeps (element so + x; its points
    // at ve[e].pt + j, dense over the level order), so a level is read with coalesced loads
    int svCap;          // conCap + NPC
    int* svSlot;        // constraint slot (or -1 - point constraint, or 0x40000000 inactive)
    VElem<T>* ve;       // [svCap]
    VPoint<T>* vp;      // [4 * svCap]
    int* svPt;          // [svCap] point offsets during staging
    V3<T>* svLA;        // [4 * svCap] contact points in body space
    V3<T>* svLB;
    VBody<T>* vb;       // [N] velocities, inverse inertia and inverse mass for the velocity sweeps
    PBody<T>* pb;       // [N] poses for the position sweeps, one record per body
    int* isl0;          // [N] island flags (by label)
    int* isl1;          // [N]
    int* stats;         // per world: [levels, posSweeps]
    int* islIt;         // [3 * N] per island root when a body or point constraint has iteration
                        // overrides (else null): max velocity and position overrides, default flags
    int hasForces;      // a body was given a force or torque: integrate and clear them
    int colored;        // solve order: 0 crashcat's (levels of the order's DAG), 1 graph-coloured
    double* resid;      // [4] when set, the last step's velocity residual: sum |dlambda|, sum lambda_n,
                        // max |dlambda| (bits of a non-negative double), points
    // differentiable stepping (cc_fwd / cc_bwd): constraints record their pair-order rank and the
    // warm-start source of each point; when tapeLam is set the solver also records its sweeps
    int diff;
    LamRec<T>* tapeLam;  // [(velIters + 1) * ptsCap]: sweep s at (s + 1) * ptsCap, the warm start at 0
    int* tapeSrc;        // [ptsCap] warm-start source of each staged point
    int* tapeRank;       // [svCap] pair-order rank of each staged element
    int ptsCap;          // 4 * svCap
    T* prevDt;          // [1]
    int* parity;        // [1]
    unsigned long long* acc;  // [ACC_N] running sums over steps, see ACC_*
};

// running sums (64-bit) over all steps since upload: counts and the algorithmic bytes of
// the byte model (per-item compulsory HBM traffic of each phase, 4-byte words)
enum { ACC_STEPS = 0, ACC_PAIRS, ACC_CONS, ACC_POINTS, ACC_EPA, ACC_AWAKE,
       ACC_BYTES_BODY, ACC_BYTES_BP, ACC_BYTES_NP, ACC_BYTES_SOLVE, ACC_GJK, ACC_N = 16 };

enum { CNT_PAIRS = 0, CNT_CONS = 1, CNT_EPA = 2, CNT_OVERFLOW = 3, CNT_EPA_TOTAL = 4, CNT_SPILL = 5, CNT_CELLS = 6,
       CNT_HITS = 7, CNT_HITS_L = 8, CNT_GJK = 9, CNT_MESH = 10, CNT_PLANE = 11, CNT_COMPOUND = 12, CNT_N = 13 };
// CNT_OVERFLOW bits: 1 pairs, 2 contact constraints, 4 EPA jobs, 8 pair spill list, 16 hits,
// 32 mesh BVH traversal stack, 64 mesh candidate triangles
constexpr int BP_SLOTS = 32;

// orders LDS accesses between the lanes of one wavefront
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One atomic per wave for "give every active lane a slot": lanes are ranked by lane id.
__device__ inline int wave_alloc(int* counter) {
    unsigned long long act = __ballot(1);
    int lane = (int)__lane_id();
    int leader = __ffsll((long long)act) - 1;
    int rank = __popcll(act & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(act));
    base = __shfl(base, leader, 64);
    return base + rank;
}

// ------------------------------------------------------------ acceleration
// accelerationIntegrationUpdate src/update.ts:245-319: a = F / m + g, then damping, the
// clamp and the translation DOFs; omega += I_w^-1 tau dt (I_w^-1 of the current rotation, FP64)
template <typename T>
__global__ void k_accel(Dev<T> d, T dt) {
    if (d.resid && blockIdx.x == 0 && threadIdx.x < 4) d.resid[threadIdx.x] = 0.0;  // this step's residual
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    const Bodies<T>& B = d.B;
    int f = B.flags[i];
    const int sl = B.sleeping[i];
    {  // poses are final for this step's collision detection
        NpBody<T> r;
        r.pos = B.pos[i];
        r.q = B.q[i];
        r.shape = B.shape[i];
        r.flags = f;
        r.brad = d.bBrad[i];
        r.tinfo = d.bTinfo[i];
        r.com = B.com[i];
        r.invMass = B.invMass[i];
        d.npb[i] = r;
    }
    NpDyn<T> rd;
    rd.v = B.v[i];
    rd.w = B.w[i];
    rd.invI = B.invI[i];
    rd.iRot = B.iRot[i];
    rd.friction = B.friction[i];
    rd.restitution = B.restitution[i];
    rd.sleeping = sl;
    if (fl_motion(f) != MOTION_DYNAMIC || sl) {
        d.npd[i] = rd;
        return;
    }
    T gs = d.s.gravityEnabled ? B.gravityFactor[i] : T(0);
    V3<T> v = B.v[i];
    const V3<T> z3 = mk3<T>(0, 0, 0);
    const V3<T> F = d.hasForces ? B.force[i] : z3, tau = d.hasForces ? B.torque[i] : z3;
    v.x += (F.x * B.invMass[i] + d.s.gravity[0] * gs) * dt;
    v.y += (F.y * B.invMass[i] + d.s.gravity[1] * gs) * dt;
    v.z += (F.z * B.invMass[i] + d.s.gravity[2] * gs) * dt;
    T ld = tmax<T>(T(0), T(1) - B.linDamp[i] * dt);
    v = v * ld;
    T s2 = len2(v), ml = B.maxLinV[i];
    if (s2 > ml * ml) v = v * (ml / sqrt(s2));
    int dof = fl_dof(f);
    v = dof_lin(v, dof);
    B.v[i] = v;
    V3<T> w = B.w[i];
    if (tau.x != T(0) || tau.y != T(0) || tau.z != T(0)) {  // a zero torque adds exactly zero
        const M3<T> I = inv_inertia(B.invI[i], B.iRot[i], B.q[i], dof);
        w = w + mul(I, tau) * dt;
    }
    T ad = tmax<T>(T(0), T(1) - B.angDamp[i] * dt);
    w = w * ad;
    T a2 = len2(w), ma = B.maxAngV[i];
    if (a2 > ma * ma) w = w * (ma / sqrt(a2));
    w = dof_lin(w, dof >> 3);
    B.w[i] = w;
    rd.v = v;
    rd.w = w;
    d.npd[i] = rd;
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

// The pair set does not depend on how candidates are found (the reference walks one dynamic
// AABB tree per broadphase layer; the set is the same for any tree).  Here candidates come
// from a uniform grid per world.  A body whose box is at most cellWidth wide on every axis is
// binned by the cell of its box centre; cells are cellWidth + 2 * margin wide, so two small
// bodies whose boxes overlap after growing one by the margin sit in the same or adjacent
// cells.  Wider bodies (floors, big static shapes) go to a per-world list.  Every candidate
// pair is then decided by the exact FP64 predicate (aabb test + pair_ok).
__device__ inline int grid_coord(double x, double inv) {
    double f = floor(x * inv);
    f = fmin(fmax(f, -8190.0), 8190.0);  // 14-bit cell coordinates; far cells merge, which is safe
    return (int)f;
}
__device__ inline unsigned long long cell_key(int w, int cx, int cy, int cz) {
    return ((unsigned long long)(unsigned)w << 42) | ((unsigned long long)(unsigned)(cx + 8192) << 28) |
           ((unsigned long long)(unsigned)(cy + 8192) << 14) | (unsigned long long)(unsigned)(cz + 8192);
}
template <typename T>
__device__ inline int cell_find(const Dev<T>& d, unsigned long long key) {
    unsigned mask = (unsigned)(d.cellTabCap - 1);
    unsigned h = hash64(key) & mask;
    for (int probe = 0; probe < d.cellTabCap; probe++) {
        unsigned long long k = d.cellKey[h];
        if (k == key) return (int)h;
        if (k == CC_EMPTY_KEY) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

template <typename T>
__global__ void k_cells_clear(Dev<T> d) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) { d.counters[CNT_SPILL] = 0; d.counters[CNT_CELLS] = 0; d.cellCount[d.cellTabCap] = 0; }
    for (int h = t; h < d.cellTabCap; h += gridDim.x * blockDim.x) {
        d.cellKey[h] = CC_EMPTY_KEY;
        d.cellCount[h] = 0;
        d.cellFill[h] = 0;
    }
    for (int i = t; i < d.N; i += gridDim.x * blockDim.x) d.bpCount[i] = 0;
}

template <typename T>
__global__ void k_cells_insert(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    V3<double> mn = cvt<double>(d.B.amin[i]), mx = cvt<double>(d.B.amax[i]);
    double ex = fmax(mx.x - mn.x, fmax(mx.y - mn.y, mx.z - mn.z));
    bool large = !(ex <= d.cellWidth);
    d.gLargeFlag[i] = large ? 1 : 0;
    if (large) { d.bodySlot[i] = -1; return; }
    unsigned long long key = cell_key(d.B.world[i], grid_coord(0.5 * (mn.x + mx.x), d.cellInv),
                                      grid_coord(0.5 * (mn.y + mx.y), d.cellInv), grid_coord(0.5 * (mn.z + mx.z), d.cellInv));
    unsigned mask = (unsigned)(d.cellTabCap - 1);
    unsigned h = hash64(key) & mask;
    for (;;) {  // the table has at least twice as many slots as bodies, so a slot is always found
        unsigned long long prev = atomicCAS(&d.cellKey[h], CC_EMPTY_KEY, key);
        if (prev == CC_EMPTY_KEY || prev == key) break;
        h = (h + 1) & mask;
    }
    d.bodySlot[i] = (int)h;
    atomicAdd(&d.cellCount[h], 1);
}

template <typename T>
__global__ void k_cells_scatter(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    int h = d.bodySlot[i];
    if (h < 0) { d.gLarge[d.gLargeOff[i]] = i; return; }
    int k = atomicAdd(&d.cellFill[h], 1);
    int p = d.cellStart[h] + k;
    CellRec<T> r;
    r.body = i;
    r.mn = d.B.amin[i];
    r.mx = d.B.amax[i];
    r.fl = d.B.flags[i] | (queries(d.B, i) ? (1 << 30) : 0);
    r.gr = d.B.groups[i];
    r.mk = d.B.mask[i];
    d.cellRec[p] = r;
    if (k == 0) d.cellList[atomicAdd(&d.counters[CNT_CELLS], 1)] = h;
}

// record pair (query i, other j): the first BP_SLOTS of a body in its own slots, the rest in
// the spill list (k_bp_compact lays them out in body order)
template <typename T>
__device__ inline void bp_emit(const Dev<T>& d, int i, int j) {
    int k = atomicAdd(&d.bpCount[i], 1);
    if (k < BP_SLOTS) {
        d.bpSlots[(size_t)i * BP_SLOTS + k] = j;
    } else {
        int sp = atomicAdd(&d.counters[CNT_SPILL], 1);
        if (sp < d.spillCap) { d.spillA[sp] = i; d.spillB[sp] = j; d.spillK[sp] = k; }
        else atomicOr(&d.counters[CNT_OVERFLOW], 8);
    }
}

template <typename T>
__device__ inline bool box_hit(V3<double> amn, V3<double> amx, V3<double> bmn, V3<double> bmx) {
    return !(bmx.x < amn.x || bmn.x > amx.x || bmx.y < amn.y || bmn.y > amx.y || bmx.z < amn.z || bmn.z > amx.z);
}

// pair_ok from staged per-body words: fl = flags, gr = groups, mk = mask, q = queries
__device__ inline bool pair_ok_words(const unsigned char* layerPairs, int a, int fa, unsigned ga, unsigned ma, bool qa, int b,
                                     int fb, unsigned gb, unsigned mb, bool qb) {
    if (a == b) return false;
    if (!layerPairs[fl_layer(fa) * 8 + fl_layer(fb)]) return false;
    if (!((ga & mb) != 0 && (gb & ma) != 0)) return false;
    long long da = qa ? a : 0x7fffffffll, db = qb ? b : 0x7fffffffll;
    if (da >= db) return false;
    return fl_motion(fa) == MOTION_DYNAMIC || fl_motion(fb) == MOTION_DYNAMIC;
}

// Small-vs-small pairs, one wavefront per occupied cell.  The cell's querying bodies (up to
// 64 at a time) and the bodies of the 27 cells around it (in chunks) are staged in LDS with
// everything the predicate needs; lanes then sweep the (home, candidate) grid.  A body is the
// home of exactly one cell, so its pair counter lives in LDS for the whole kernel; k_bp_large
// runs first and its pairs are counted in.
constexpr int BP_CHUNK = 128;
constexpr int BP_HOME = 64;
template <typename T>
__global__ void __launch_bounds__(256) k_bp_cells(Dev<T> d, double margin) {
    __shared__ V3<T> smn[4][BP_CHUNK], smx[4][BP_CHUNK];
    __shared__ int sbody[4][BP_CHUNK], sfl[4][BP_CHUNK];
    __shared__ unsigned sgr[4][BP_CHUNK], smk[4][BP_CHUNK];
    __shared__ V3<double> hmn[4][BP_HOME], hmx[4][BP_HOME];
    __shared__ int hbody[4][BP_HOME], hfl[4][BP_HOME], hcnt[4][BP_HOME];
    __shared__ unsigned hgr[4][BP_HOME], hmk[4][BP_HOME];
    __shared__ int nStart[4][28], nPre[4][28];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ncells = d.counters[CNT_CELLS];
    const unsigned char* lp = d.s.layerPairs;
    for (int ci = blockIdx.x * 4 + wv; ci < ncells; ci += gridDim.x * 4) {
        const int h = d.cellList[ci];
        const int h0 = d.cellStart[h], nh = d.cellCount[h];
        const CellRec<T> r0 = d.cellRec[h0];
        const int w = d.B.world[r0.body];
        V3<double> c0 = (cvt<double>(r0.mn) + cvt<double>(r0.mx)) * 0.5;
        int hx = grid_coord(c0.x, d.cellInv), hy = grid_coord(c0.y, d.cellInv), hz = grid_coord(c0.z, d.cellInv);
        int cnt = 0, st = 0;
        if (lane < 27) {
            int nx = hx + lane % 3 - 1, ny = hy + (lane / 3) % 3 - 1, nz = hz + lane / 9 - 1;
            int sl = cell_find(d, cell_key(w, nx, ny, nz));
            if (sl >= 0) { cnt = d.cellCount[sl]; st = d.cellStart[sl]; }
        }
        int pre = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            int v = __shfl_up(pre, o, 64);
            if (lane >= o) pre += v;
        }
        if (lane < 27) { nStart[wv][lane] = st; nPre[wv][lane + 1] = pre; }
        if (lane == 0) nPre[wv][0] = 0;
        const int total = __shfl(pre, 26, 64);
        for (int hb0 = 0; hb0 < nh; hb0 += BP_HOME) {
            // stage this batch's querying home bodies, compacted
            CellRec<T> hr;
            bool q = false;
            if (hb0 + lane < nh) {
                hr = d.cellRec[h0 + hb0 + lane];
                q = (hr.fl >> 30) & 1;
            }
            unsigned long long qm = __ballot(q);
            const int nq = __popcll(qm);
            if (q) {
                int r = __popcll(qm & ((1ull << lane) - 1ull));
                hbody[wv][r] = hr.body;
                hmn[wv][r] = cvt<double>(hr.mn) - mk3<double>(margin, margin, margin);
                hmx[wv][r] = cvt<double>(hr.mx) + mk3<double>(margin, margin, margin);
                hfl[wv][r] = hr.fl & 0x3fffffff;
                hgr[wv][r] = hr.gr;
                hmk[wv][r] = hr.mk;
                hcnt[wv][r] = d.bpCount[hr.body];
            }
            wave_sync();
            if (nq == 0) continue;
            for (int cb = 0; cb < total; cb += BP_CHUNK) {
                const int nc = min(BP_CHUNK, total - cb);
                for (int x = lane; x < nc; x += 64) {
                    int f = cb + x;
                    int lo = 0, hi = 26;  // the neighbour cell holding flat index f
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (nPre[wv][mid] <= f) lo = mid; else hi = mid - 1;
                    }
                    int p = nStart[wv][lo] + (f - nPre[wv][lo]);
                    const CellRec<T> cr = d.cellRec[p];
                    sbody[wv][x] = cr.body;
                    smn[wv][x] = cr.mn;
                    smx[wv][x] = cr.mx;
                    sfl[wv][x] = cr.fl;
                    sgr[wv][x] = cr.gr;
                    smk[wv][x] = cr.mk;
                }
                wave_sync();
                const int work = nq * nc;
                int hh = lane / nc, x = lane % nc;  // (home, candidate) of this lane's first cell of work
                for (int t = lane; t < work; t += 64) {
                    const int j = sbody[wv][x];
                    const int fj = sfl[wv][x];
                    if (box_hit<T>(hmn[wv][hh], hmx[wv][hh], cvt<double>(smn[wv][x]), cvt<double>(smx[wv][x])) &&
                        pair_ok_words(lp, hbody[wv][hh], hfl[wv][hh], hgr[wv][hh], hmk[wv][hh], true, j, fj & 0x3fffffff,
                                      sgr[wv][x], smk[wv][x], (fj >> 30) & 1)) {
                        const int ib = hbody[wv][hh];
                        int k = atomicAdd(&hcnt[wv][hh], 1);
                        if (k < BP_SLOTS) {
                            d.bpSlots[(size_t)ib * BP_SLOTS + k] = j;
                        } else {
                            int sp = atomicAdd(&d.counters[CNT_SPILL], 1);
                            if (sp < d.spillCap) { d.spillA[sp] = ib; d.spillB[sp] = j; d.spillK[sp] = k; }
                            else atomicOr(&d.counters[CNT_OVERFLOW], 8);
                        }
                    }
                    x += 64;
                    while (x >= nc) { x -= nc; hh++; }
                }
                wave_sync();
            }
            if (lane < nq) d.bpCount[hbody[wv][lane]] = hcnt[wv][lane];
            wave_sync();
        }
    }
}

// Pairs with a large body: every querying body tests its world's large bodies; a large
// querying body also tests the world's small bodies directly.
template <typename T>
__global__ void k_bp_large(Dev<T> d, double margin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N || !queries(d.B, i)) return;
    const V3<double> amn = cvt<double>(d.B.amin[i]) - mk3<double>(margin, margin, margin);
    const V3<double> amx = cvt<double>(d.B.amax[i]) + mk3<double>(margin, margin, margin);
    const int w = d.B.world[i];
    const int wb0 = d.wStart[w], wb1 = d.wStart[w + 1];
    for (int x = d.gLargeOff[wb0]; x < d.gLargeOff[wb1]; x++) {
        int j = d.gLarge[x];
        if (box_hit<T>(amn, amx, cvt<double>(d.B.amin[j]), cvt<double>(d.B.amax[j])) && pair_ok(d, i, j)) bp_emit(d, i, j);
    }
    if (d.bodySlot[i] >= 0) return;
    for (int j = wb0; j < wb1; j++)
        if (d.bodySlot[j] >= 0 && box_hit<T>(amn, amx, cvt<double>(d.B.amin[j]), cvt<double>(d.B.amax[j])) && pair_ok(d, i, j))
            bp_emit(d, i, j);
}

template <typename T>
__global__ void k_bp_compact(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    int n = min(d.bpCount[i], BP_SLOTS);
    int o = d.bpOff[i];
    const int* slots = d.bpSlots + (size_t)i * BP_SLOTS;
    for (int k = 0; k < n; k++)
        if (o + k < d.pairCap) { d.pairA[o + k] = i; d.pairB[o + k] = slots[k]; }
}

// spilled pairs go to their rank within their body's range
template <typename T>
__global__ void k_bp_spill(Dev<T> d) {
    int ns = min(d.counters[CNT_SPILL], d.spillCap);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        int i = d.spillA[s];
        int p = d.bpOff[i] + d.spillK[s];
        if (p < d.pairCap) { d.pairA[p] = i; d.pairB[p] = d.spillB[s]; }
    }
}

// Differentiable stepping: each querying body's partners in ascending order, so that the pair
// order (and with it the constraints' pair-order ranks) does not depend on which thread found a
// pair first.  Nothing else reads the pair order: constraint creation and solve orders come from
// their own keys.
template <typename T>
__global__ void k_bp_sort(Dev<T> d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    const int lo = d.bpOff[i], hi = min(d.bpOff[i + 1], d.pairCap);
    for (int p = lo + 1; p < hi; p++) {
        const int v = d.pairB[p];
        int q = p - 1;
        while (q >= lo && d.pairB[q] > v) {
            d.pairB[q + 1] = d.pairB[q];
            q--;
        }
        d.pairB[q + 1] = v;
    }
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
__device__ inline PairBodies<T> order_pair_f(int a, int fa, int b, int fb) {
    int ma = fl_motion(fa), mb = fl_motion(fb);
    PairBodies<T> r;
    if (ma > mb || (ma == mb && b < a)) { r.a = b; r.b = a; }
    else { r.a = a; r.b = b; }
    return r;
}

template <typename T>
__device__ inline PairBodies<T> order_pair(const Dev<T>& d, int a, int b) {
    int ma = fl_motion(d.B.flags[a]), mb = fl_motion(d.B.flags[b]);
    PairBodies<T> r;
    if (ma > mb || (ma == mb && b < a)) { r.a = b; r.b = a; }
    else { r.a = a; r.b = b; }
    return r;
}

// The arithmetic type C of GJK, EPA and the convex manifolds: T, or FP64 in an FP32 context with
// npWide (the shape table, tolerances and EPA simplices in FP64 are then the context's FP64 copies).
template <typename T, typename C>
__device__ inline const ShapeTable<C>& np_shapes(const Dev<T>& d) {
    if constexpr (std::is_same<T, C>::value) return d.st;
    else return d.st64;
}
template <typename C> struct NpTol { C spec, colTol, penTol, manifoldTol, normalCos; };
template <typename T, typename C>
__device__ inline NpTol<C> np_tol(const Dev<T>& d) {
    if constexpr (std::is_same<T, C>::value) return NpTol<C>{d.s.spec, d.s.colTol, d.s.penTol, d.s.manifoldTol, d.s.normalCos};
    else return NpTol<C>{C(d.spec64), C(d.colTol64), C(d.penTol64), C(d.manifoldTol64), C(d.normalCos64)};
}
template <typename T, typename C>
__device__ inline Simplex<C>* np_simplices(const Dev<T>& d) {
    if constexpr (std::is_same<T, C>::value) return d.epaSimplex;
    else return d.epaSimplex64;
}

// The poses the two leaves of a pair are collided at, in type C (a transformed decorator
// composes its body's pose with its local transform, leaf_pose).
template <typename C, typename T>
__device__ inline void leaf_poses(const ShapeTable<C>& st, const NpBody<T>& RA, const NpBody<T>& RB, V3<C>& pa,
                                  Q4<C>& qa, V3<C>& pb, Q4<C>& qb) {
    pa = cvt<C>(RA.pos);
    qa = cvtq<C>(RA.q);
    pb = cvt<C>(RB.pos);
    qb = cvtq<C>(RB.q);
    leaf_pose(st.shapes[RA.shape], pa, qa);
    leaf_pose(st.shapes[RB.shape], pb, qb);
}

// normal velocity bias, src/constraints/contact-constraints.ts:288-423
template <typename T>
__device__ inline T normal_bias(const Dev<T>& d, int A, int Bi, int fa, int fb, const NpDyn<T>& DA, const NpDyn<T>& DB,
                                V3<T> pA, V3<T> pB, V3<T> n, V3<T> rA, V3<T> rB, T e, T dt) {
    const Bodies<T>& B = d.B;
    int ma = fl_motion(fa), mb = fl_motion(fb);
    bool mA = ma != MOTION_STATIC, mB = mb != MOTION_STATIC;
    V3<T> rel;
    if (mA && mB) rel = (DB.v + cross(DB.w, rB)) - (DA.v + cross(DA.w, rA));
    else if (mA) rel = -(DA.v + cross(DA.w, rA));
    else if (mB) rel = DB.v + cross(DB.w, rB);
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
            if (d.hasForces && ma == MOTION_DYNAMIC) acc = acc + B.force[A] * (-B.invMass[A]);
            if (d.hasForces && mb == MOTION_DYNAMIC) acc = acc + B.force[Bi] * B.invMass[Bi];
            T fdv = tmin<T>(T(0), dot(acc, n) * dt);
            return e * (nv - fdv);
        }
        return spec;
    }
    return spec;
}

// Finish one hit: manifold (update.ts:385-563 for the single hit of a convex pair), then
// addContactConstraint (contact-constraints.ts:437-857) into the pair slot.
template <typename T, typename C, int N>
__device__ inline void finish_pair(const Dev<T>& d, int slot, int A, int Bi, const NpBody<T>& RA, const NpBody<T>& RB,
                                   V3<C> pointA, V3<C> pointB, V3<C> axis, const Face<C, N>& fA, const Face<C, N>& fB,
                                   Face<C, N>* scratch3, Manifold<C, N>& m, T dt, int par) {
    const NpTol<C> tol = np_tol<T, C>(d);
    m.base = cvt<C>(RA.pos);
    m.normal = normalize(axis);
    m.n = 0;
    C maxDist = tol.spec + tol.manifoldTol;
    if (fA.n >= 2 && fB.n >= 3) manifold_between_two_faces(m, pointA, pointB, axis, maxDist, fA, fB, scratch3);
    else { m.relA[0] = pointA - m.base; m.relB[0] = pointB - m.base; m.n = 1; }
    m.normal = normalize(m.normal);
    if (m.n > 4) prune(m, m.normal);
    if (m.n <= 0) return;
    make_contact(d, slot, A, Bi, RA, RB, m, 0xffffffffu, 0xffffffffu, -1, 0, 1, dt, par);
}

// addContactConstraint (contact-constraints.ts:437-857) for one finished manifold of the pair in
// `slot`: cs is its constraint slot, or -1 to allocate one; the pair's constraints form a run of
// runCnt consecutive slots and this is number runIdx of it.
template <typename T, typename M>
__device__ inline void make_contact(const Dev<T>& d, int slot, int A, int Bi, const NpBody<T>& RA, const NpBody<T>& RB,
                                    const M& m, unsigned subA, unsigned subB, int cs, int runIdx, int runCnt, T dt,
                                    int par) {
    const Bodies<T>& B = d.B;
    // swap so that id A < id B (manifold.swapShapes :86-107)
    int ia = A, ib = Bi;
    V3<T> n = cvt<T>(m.normal);
    bool sw = ia > ib;
    if (sw) { int t = ia; ia = ib; ib = t; n = -n; unsigned u = subA; subA = subB; subB = u; }
    const NpBody<T>& Ria = sw ? RB : RA;
    const NpBody<T>& Rib = sw ? RA : RB;
    int fa = Ria.flags, fb = Rib.flags;
    bool dynA = fl_motion(fa) == MOTION_DYNAMIC, dynB = fl_motion(fb) == MOTION_DYNAMIC;
    if (!((dynA && Ria.invMass != T(0)) || (dynB && Rib.invMass != T(0)))) return;
    const NpDyn<T> DA = d.npd[ia], DB = d.npd[ib];
    if (cs < 0) cs = wave_alloc(&d.counters[CNT_CONS]);
    if (cs >= d.conCap) { atomicOr(&d.counters[CNT_OVERFLOW], 2); return; }
    if (runIdx == 0) d.pairCon[slot] = cs;
    Con<T>& k = d.con[cs];
    ConTap<T>& tp = d.tap[cs];
    CacheRec<T>& rc = d.rec[par][cs];
    // previous record of this contact (findContact src/contacts.ts:481-513): the table maps the
    // body pair to the first record of its run, the run is searched for the sub-shape ids
    unsigned long long key = pair_key(ia, ib);
    int prev = -1;
    {
        const unsigned long long* hk = d.hkeys[par ^ 1];
        unsigned h = hash64(key) & (unsigned)(d.hashCap - 1);
        for (int probe = 0; probe < d.hashCap; probe++) {
            unsigned long long kk = hk[h];
            if (kk == CC_EMPTY_KEY) break;
            if (kk == key) {
                const int p0 = d.hvals[par ^ 1][h];
                const int run = d.rec[par ^ 1][p0].run;
                for (int j = 0; j < run; j++) {
                    const CacheRec<T>& r = d.rec[par ^ 1][p0 + j];
                    if (r.subA == subA && r.subB == subB) { prev = p0 + j; break; }
                }
                break;
            }
            h = (h + 1) & (unsigned)(d.hashCap - 1);
        }
    }
    d.recHit[par][cs] = 0;
    // differentiable stepping: the source of each in-place slot's impulses (rank * 4 + point of
    // the previous step's constraint; slots past its point count hold older values: no source)
    int srcOf[4] = {-1, -1, -1, -1};
    if (prev >= 0) {
        rc = d.rec[par ^ 1][prev];
        d.recHit[par ^ 1][prev] = 1;  // processedThisFrame (src/contacts.ts:468)
        if (d.diff)
            for (int j = 0; j < 4; j++) srcOf[j] = j < rc.np ? rc.rank * 4 + j : -1;
    } else {
        rc.np = 0;
        for (int j = 0; j < 4; j++) { rc.p1[j] = rc.p2[j] = mk3<T>(0, 0, 0); rc.ln[j] = rc.lt1[j] = rc.lt2[j] = T(0); }
    }
    rc.a = ia; rc.b = ib;
    rc.subA = subA; rc.subB = subB;
    rc.run = runIdx == 0 ? runCnt : 0;
    rc.np = m.n;
    Q4<T> qa = Ria.q, qb = Rib.q;
    M3<T> rotA = m3_from_quat(qa), rotB = m3_from_quat(qb);
    T fr = combine(DA.friction, DB.friction, fl_fmode(fa), fl_fmode(fb));
    T re = combine(DA.restitution, DB.restitution, fl_rmode(fa), fl_rmode(fb));
    k.a = ia; k.b = ib;
    k.cq = d.pairA[slot];  // (query, other, place in the pair's run): the creation order of constraints
    k.co = d.pairB[slot] | (runIdx << 24);
    k.dynMask = (dynA ? 1 : 0) | (dynB ? 2 : 0);
    k.movMask = (fl_motion(fa) != MOTION_STATIC ? 1 : 0) | (fl_motion(fb) != MOTION_STATIC ? 2 : 0);
    k.sortKey = sort_key((unsigned)ia, (unsigned)ib, subA, subB);
    k.normal = n;
    k.t1 = contact_t1(n);
    k.t2 = cross(n, k.t1);
    k.friction = fr;
    k.invMassA = Ria.invMass;
    k.invMassB = Rib.invMass;
    M3<T> IA, IB;
    for (int j = 0; j < 9; j++) { IA.m[j] = T(0); IB.m[j] = T(0); }
    if (dynA) IA = inv_inertia(DA.invI, DA.iRot, qa, fl_dof(fa));
    if (dynB) IB = inv_inertia(DB.invI, DB.iRot, qb, fl_dof(fb));
    V3<T> comA = Ria.com, comB = Rib.com;
    k.np = m.n;
    for (int i = 0; i < m.n; i++) {
        ConPoint<T>& cp = k.p[i];
        V3<T> ra_ = cvt<T>(sw ? m.relB[i] : m.relA[i]);
        V3<T> rb_ = cvt<T>(sw ? m.relA[i] : m.relB[i]);
        V3<T> pA = cvt<T>(m.base) + ra_, pB = cvt<T>(m.base) + rb_;
        tp.pa[i] = pA;
        tp.pb[i] = pB;
        cp.localA = mul_t(rotA, pA - comA);
        cp.localB = mul_t(rotB, pB - comB);
        bool set = false;
        int srcI = -1;
        if (prev >= 0) {  // first match over the in-place slots (slots < i already hold this step's points)
            for (int j = 0; j < rc.np; j++) {
                T d1 = len2(cp.localA - rc.p1[j]), d2 = len2(cp.localB - rc.p2[j]);
                if (d1 < d.s.preserveDist2 && d2 < d.s.preserveDist2) {
                    cp.lN = rc.ln[j]; cp.lT1 = rc.lt1[j]; cp.lT2 = rc.lt2[j];
                    srcI = srcOf[j];
                    set = true;
                    break;
                }
            }
        }
        if (d.diff) {
            tp.src[i] = srcI;
            srcOf[i] = srcI;  // slot i now holds this point and its impulses
        }
        if (!set) { cp.lN = cp.lT1 = cp.lT2 = T(0); }
        V3<T> mid = mk3<T>((pA.x + pB.x) * T(0.5), (pA.y + pB.y) * T(0.5), (pA.z + pB.z) * T(0.5));
        V3<T> rA = mid - comA, rB = mid - comB;
        T bias = normal_bias(d, ia, ib, fa, fb, DA, DB, pA, pB, n, rA, rB, re, dt);
        cp.r1 = rA;
        cp.r2 = rB;
        // an axis without mass drops its impulse and bias (axis-constraint-part.ts:87-201)
        cp.effN = axis_eff(dynA, dynB, k.invMassA, k.invMassB, IA, IB, rA, rB, n);
        cp.biasN = cp.effN == T(0) ? T(0) : bias;
        if (cp.effN == T(0)) cp.lN = T(0);
        if (fr > T(0)) {
            cp.effT1 = axis_eff(dynA, dynB, k.invMassA, k.invMassB, IA, IB, rA, rB, k.t1);
            cp.effT2 = axis_eff(dynA, dynB, k.invMassA, k.invMassB, IA, IB, rA, rB, k.t2);
            if (cp.effT1 == T(0)) cp.lT1 = T(0);
            if (cp.effT2 == T(0)) cp.lT2 = T(0);
        } else {
            cp.effT1 = cp.effT2 = T(0);
            cp.lT1 = cp.lT2 = T(0);
        }
        rc.p1[i] = cp.localA;
        rc.p2[i] = cp.localB;
        rc.ln[i] = cp.lN;
        rc.lt1[i] = cp.lT1;
        rc.lt2[i] = cp.lT2;
    }
    ConHead hd;
    hd.a = ia;
    hd.b = ib;
    hd.sortKey = k.sortKey;
    hd.bits = k.dynMask | (k.movMask << 2) | (k.np << 4) | (runIdx > 0 ? 128 : ((runCnt - 1) << 8)) |
              (prev >= 0 ? 8192 : 0);
    d.conHead[cs] = hd;
    if (dynA && DA.sleeping) B.wake[ia] = 1;
    if (dynB && DB.sleeping) B.wake[ib] = 1;
}

template <typename T, int N>
struct NpScratch {
    Face<T, N> fA, fB, tmp[3];
    Manifold<T, N> m;
};

// world-space hit + supporting faces after GJK/EPA (convex.ts:489-637)
template <typename T, typename C, int N>
__device__ inline void finish_convex(const Dev<T>& d, int slot, int A, int Bi, const NpBody<T>& RA,
                                     const NpBody<T>& RB, const ConvexCtx<C>& c, PenDepth<C> pd, NpScratch<C, N>& S, T dt,
                                     int par) {
    const ShapeTable<C>& stc = np_shapes<T, C>(d);
    V3<C> pa = cvt<C>(RA.pos);
    Q4<C> qa = cvtq<C>(RA.q);
    leaf_pose(stc.shapes[RA.shape], pa, qa);
    C ms = c.maxSepUse;
    // penetration = |pA - pB| - maxSep; the early-out fraction is +inf for the step collector
    C al = sqrt(len2(pd.axis));
    if (al > C(0)) { C corr = ms / al; pd.pointA = pd.pointA + pd.axis * (-corr); }
    V3<C> wA = qrot2(qa, pd.pointA) + pa;
    V3<C> wB = qrot2(qa, pd.pointB) + pa;
    V3<C> wAxis = qrot2(qa, pd.axis);
    supporting_face(S.fA, stc, RA.shape, -pd.axis, pa, qa);
    Q4<C> crel = qconj(c.rel);
    V3<C> dirB = qrot2(crel, pd.axis);
    V3<C> posB = qrot2(qa, c.relPos) + pa;
    Q4<C> qB;
    qB.x = qa.w * c.rel.x + qa.x * c.rel.w + qa.y * c.rel.z - qa.z * c.rel.y;
    qB.y = qa.w * c.rel.y - qa.x * c.rel.z + qa.y * c.rel.w + qa.z * c.rel.x;
    qB.z = qa.w * c.rel.z + qa.x * c.rel.y - qa.y * c.rel.x + qa.z * c.rel.w;
    qB.w = qa.w * c.rel.w - qa.x * c.rel.x - qa.y * c.rel.y - qa.z * c.rel.z;
    supporting_face(S.fB, stc, RB.shape, dirB, posB, qB);
    finish_pair(d, slot, A, Bi, RA, RB, wA, wB, wAxis, S.fA, S.fB, S.tmp, S.m, dt, par);
}

enum { HIT_SPHERES = 0, HIT_GJK = 1, HIT_EPA = 2 };

// Hits whose two shapes have supporting faces of at most four vertices go to the front of the
// hit list (manifolds of capacity 8), the others to its back (capacity 32).
template <typename T, typename C>
__device__ inline void push_hit(const Dev<T>& d, int slot, int kind, V3<C> pa, V3<C> pb, V3<C> axis, bool small) {
    int h;
    if (small) h = wave_alloc(&d.counters[CNT_HITS]);
    else h = d.hitCap - 1 - wave_alloc(&d.counters[CNT_HITS_L]);
    if (h < 0 || h >= d.hitCap) { atomicOr(&d.counters[CNT_OVERFLOW], 16); return; }
    d.hitSlot[h] = slot;
    d.hitKind[h] = kind;
    d.hitPA[h] = cvt<double>(pa);
    d.hitPB[h] = cvt<double>(pb);
    d.hitAxis[h] = cvt<double>(axis);
}

// The two narrowphase records of a pair slot, after the motion-type swap (src/update.ts:74-78).
template <typename T>
__device__ inline void np_pair(const Dev<T>& d, int slot, NpBody<T>& RA, NpBody<T>& RB) {
    const int i0 = d.pairA[slot], i1 = d.pairB[slot];
    const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
    PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
    const bool sw = pb.a != i0;
    RA = sw ? r1 : r0;
    RB = sw ? r0 : r1;
}

// Position of this thread's entry when the flagged threads of the block append one entry each
// to a list whose length is *counter: one global atomic per block.  Every thread of the block
// calls it; sh holds BS / 64 + 1 ints.
template <int BS>
__device__ inline int block_append(int* counter, bool flag, int* sh) {
    const unsigned long long m = __ballot(flag);
    const int lane = (int)__lane_id(), wave = threadIdx.x >> 6;
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) sh[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < BS / 64; w++) { int c = sh[w]; sh[w] = tot; tot += c; }
        sh[BS / 64] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const int r = flag ? sh[BS / 64] + sh[wave] + rank : -1;
    __syncthreads();
    return r;
}

// The pair filter: sphere pairs in closed form (collideSphereVsSphere sphere.ts:319-410); pairs
// whose bounding spheres are further apart than the speculative distance stop here (GJK would
// report them separated at once); the rest are listed for GJK with their shape-type pair.
template <typename T>
__global__ void __launch_bounds__(256) k_np_filter(Dev<T> d, T dt) {
    __shared__ int sh[256 / 64 + 1];
    const int npairs = min(d.counters[CNT_PAIRS], d.pairCap);
    for (int base = blockIdx.x * 256; base < npairs; base += gridDim.x * 256) {  // the same trip count in every thread
        const int slot = base + threadIdx.x;
        bool gjk = false, mesh = false, plane = false, cmp = false;
        int combo = 0;
        if (slot < npairs) {
            d.pairCon[slot] = -1;
            NpBody<T> RA, RB;
            np_pair(d, slot, RA, RB);
            const int ta = RA.tinfo & 15, tb = RB.tinfo & 15;
            if (ta == SHAPE_SPHERE && tb == SHAPE_SPHERE) {  // radii times |scale.x|, leaves at their poses
                V3<T> pa, pbp;
                Q4<T> qa, qb;
                leaf_poses(d.st, RA, RB, pa, qa, pbp, qb);
                const ShapeD<T>& SA = d.st.shapes[RA.shape];
                const ShapeD<T>& SB = d.st.shapes[RB.shape];
                T ra = SA.radius * fabs(shape_scale(SA).x), rb = SB.radius * fabs(shape_scale(SB).x);
                V3<T> dd = pbp - pa;
                T d2 = len2(dd), rs = ra + rb;
                if (d2 < rs * rs) {
                    T dist = sqrt(d2);
                    V3<T> n = dist > T(0) ? dd * (T(1) / dist) : mk3<T>(0, 1, 0);
                    push_hit(d, slot, HIT_SPHERES, pa + n * ra, pbp - n * rb, n, true);
                }
            } else if (ta == SHAPE_COMPOUND || tb == SHAPE_COMPOUND) {  // k_compound
                T reach = RA.brad + RB.brad + d.s.spec;
                cmp = !(len2(RB.pos - RA.pos) > reach * reach);
            } else if (ta == SHAPE_PLANE || tb == SHAPE_PLANE) {
                plane = ta != tb && ta != SHAPE_MESH && tb != SHAPE_MESH;  // k_np_plane; none vs a plane or a mesh
            } else if (ta == SHAPE_MESH || tb == SHAPE_MESH) {
                mesh = ta != tb;  // no mesh-vs-mesh function
            } else {
                T reach = RA.brad + RB.brad + d.s.spec;
                gjk = !(len2(RB.pos - RA.pos) > reach * reach);
                combo = min(ta, 3) * 4 + min(tb, 3);  // groups pairs by support kind, cylinders with hulls
            }
        }
        const int idx = block_append<256>(&d.counters[CNT_GJK], gjk, sh);
        if (gjk) {
            d.gjkList[idx] = slot;
            d.gjkCombo[idx] = combo;
        }
        if (d.hasMesh) {
            __syncthreads();
            const int mi = block_append<256>(&d.counters[CNT_MESH], mesh, sh);
            if (mesh) d.meshList[mi] = slot;
        }
        if (d.hasPlane) {
            __syncthreads();
            const int pi = block_append<256>(&d.counters[CNT_PLANE], plane, sh);
            if (plane) d.planeList[pi] = slot;
        }
        if (d.hasCompound) {
            __syncthreads();
            const int ci = block_append<256>(&d.counters[CNT_COMPOUND], cmp, sh);
            if (cmp) d.cmpList[ci] = slot;
        }
    }
}

// counting sort of the GJK list by type pair: counts, then positions
template <typename T>
__global__ void __launch_bounds__(256) k_gjk_count(Dev<T> d) {
    __shared__ int h[16];
    if (threadIdx.x < 16) h[threadIdx.x] = 0;
    __syncthreads();
    const int n = min(d.counters[CNT_GJK], d.pairCap);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) atomicAdd(&h[d.gjkCombo[k]], 1);
    __syncthreads();
    if (threadIdx.x < 16 && h[threadIdx.x]) atomicAdd(&d.gjkHist[threadIdx.x], h[threadIdx.x]);
}
template <typename T>
__global__ void k_gjk_offsets(Dev<T> d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int acc = 0;
        for (int c = 0; c < 16; c++) { int v = d.gjkHist[c]; d.gjkHist[16 + c] = acc; acc += v; }
    }
}
template <typename T>
__global__ void __launch_bounds__(256) k_gjk_scatter(Dev<T> d) {
    __shared__ int h[16], b[16];
    const int n = min(d.counters[CNT_GJK], d.pairCap);
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {  // the same trip count in every thread
        if (threadIdx.x < 16) h[threadIdx.x] = 0;
        __syncthreads();
        const int k = base + threadIdx.x;
        const int c = k < n ? d.gjkCombo[k] : -1;
        const int r = c >= 0 ? atomicAdd(&h[c], 1) : 0;
        __syncthreads();
        if (threadIdx.x < 16) b[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&d.gjkHist[16 + threadIdx.x], h[threadIdx.x]) : 0;
        __syncthreads();
        if (c >= 0) d.gjkSorted[b[c] + r] = d.gjkList[k];
        __syncthreads();
    }
}

// GJK of each listed pair, in type-pair order so that a wavefront's lanes run the same support
// functions: the GJK step of collideConvexVsConvexLocal (convex.ts:257-641), giving a hit for
// k_np_manifold or, when the cores overlap, a job for EPA.  C: arithmetic type (above).
template <typename T, typename C>
__global__ void __launch_bounds__(256, sizeof(C) == 4 ? 3 : 2) k_np_gjk(Dev<T> d, T dt) {
    const int n = min(d.counters[CNT_GJK], d.pairCap);
    const ShapeTable<C>& stc = np_shapes<T, C>(d);
    const NpTol<C> tol = np_tol<T, C>(d);
    Simplex<C>* simplices = np_simplices<T, C>(d);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int slot = d.gjkSorted[k];
        NpBody<T> RA, RB;
        np_pair(d, slot, RA, RB);
        V3<C> pa, pb;
        Q4<C> qa, qb;
        leaf_poses(stc, RA, RB, pa, qa, pb, qb);
        ConvexCtx<C> c;
        convex_setup(c, pa, qa, pb, qb, tol.spec);
        Sup<C> SA, SB;
        sup_for_shape(SA, stc, RA.shape, false);
        sup_for_shape(SB, stc, RB.shape, false);
        SB.transformed = true;
        SB.R = m3_from_quat(c.rel);
        SB.t = c.relPos;
        V3<C> axis0 = c.relPos;
        if (len2(axis0) <= C(1e-12)) axis0 = mk3<C>(1, 0, 0);
        PenDepth<C> pd;
        Simplex<C> simplex;
        simplex.size = 0;
        pd_step_gjk(pd, simplex, SA, SB, SA.cr + c.maxSepUse, SB.cr, axis0, tol.colTol);
        if (pd.status == PD_NOT_COLLIDING) continue;
        if (pd.status == PD_INDETERMINATE) {
            int j = wave_alloc(&d.counters[CNT_EPA]);
            if (j < d.epaCap) { d.epaPair[j] = slot; simplices[j] = simplex; }
            else atomicOr(&d.counters[CNT_OVERFLOW], 4);
            continue;
        }
        push_hit(d, slot, HIT_GJK, pd.pointA, pd.pointB, pd.axis,
                 (RA.tinfo >> 4) <= 4 && (RB.tinfo >> 4) <= 4);
    }
}

template <typename T, typename C>
__global__ void __launch_bounds__(64, 2) k_epa(Dev<T> d, T dt) {
    int nj = min(d.counters[CNT_EPA], d.epaCap);
    int worker = blockIdx.x * blockDim.x + threadIdx.x;
    if (worker >= d.epaWorkers) return;
    const ShapeTable<C>& stc = np_shapes<T, C>(d);
    const NpTol<C> tol = np_tol<T, C>(d);
    const Simplex<C>* simplices = np_simplices<T, C>(d);
    for (int j = worker; j < nj; j += gridDim.x * blockDim.x) {
        int slot = d.epaPair[j];
        const int i0 = d.pairA[slot], i1 = d.pairB[slot];
        const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
        PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
        const bool sw = pb.a != i0;
        const NpBody<T>& RA = sw ? r1 : r0;
        const NpBody<T>& RB = sw ? r0 : r1;
        V3<C> lpa, lpb;
        Q4<C> lqa, lqb;
        leaf_poses(stc, RA, RB, lpa, lqa, lpb, lqb);
        ConvexCtx<C> c;
        convex_setup(c, lpa, lqa, lpb, lqb, tol.spec);
        c.maxSepUse = tmin<C>(c.maxSepUse, C(1));
        Sup<C> SA, SB;
        sup_for_shape(SA, stc, RA.shape, true);
        sup_for_shape(SB, stc, RB.shape, true);
        SA.addRadius = true;
        SA.addR = c.maxSepUse;
        SB.transformed = true;
        SB.R = m3_from_quat(c.rel);
        SB.t = c.relPos;
        PenDepth<C> pd;
        Simplex<C> simplex = simplices[j];
        EpaH<C> H;
        H.b = d.epaWork + (size_t)(worker / EPA_LANES) * EpaWave<C>::bytes;
        H.lane = worker % EPA_LANES;
        if (!pd_step_epa(pd, SA, SB, tol.penTol, simplex, H)) continue;
        push_hit(d, slot, HIT_EPA, pd.pointA, pd.pointB, pd.axis,
                 (RA.tinfo >> 4) <= 4 && (RB.tinfo >> 4) <= 4);
    }
}

// One hit per lane: world-space contact, supporting faces, manifold, contact cache and the
// contact constraint (convex.ts:489-637, src/update.ts:385-563).
// N: face and manifold capacity; the small variant takes the front of the hit list, the
// general one its back.  C: arithmetic type (np_shapes above).
template <typename T, int N, typename C>
__global__ void __launch_bounds__(256) k_np_manifold(Dev<T> d, T dt) {
    const bool small = N == CC_SMALL_FACE;
    const int h0 = small ? 0 : d.hitCap - min(d.counters[CNT_HITS_L], d.hitCap);
    const int h1 = small ? min(d.counters[CNT_HITS], d.hitCap) : d.hitCap;
    int par = *d.parity;
    for (int h = h0 + blockIdx.x * blockDim.x + threadIdx.x; h < h1; h += gridDim.x * blockDim.x) {
        int slot = d.hitSlot[h];
        const int i0 = d.pairA[slot], i1 = d.pairB[slot];
        const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
        PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
        const bool sw = pb.a != i0;
        const NpBody<T>& RA = sw ? r1 : r0;
        const NpBody<T>& RB = sw ? r0 : r1;
        int A = pb.a, Bi = pb.b;
        int kind = d.hitKind[h];
        NpScratch<C, N> S;
        if (kind == HIT_SPHERES) {
            S.fA.n = 0;
            S.fB.n = 0;
            finish_pair(d, slot, A, Bi, RA, RB, cvt<C>(d.hitPA[h]), cvt<C>(d.hitPB[h]), cvt<C>(d.hitAxis[h]), S.fA, S.fB,
                        S.tmp, S.m, dt, par);
            continue;
        }
        V3<C> lpa, lpb;
        Q4<C> lqa, lqb;
        leaf_poses(np_shapes<T, C>(d), RA, RB, lpa, lqa, lpb, lqb);
        ConvexCtx<C> c;
        convex_setup(c, lpa, lqa, lpb, lqb, np_tol<T, C>(d).spec);
        if (kind == HIT_EPA) c.maxSepUse = tmin<C>(c.maxSepUse, C(1));
        PenDepth<C> pd;
        pd.status = 0;
        pd.pointA = cvt<C>(d.hitPA[h]);
        pd.pointB = cvt<C>(d.hitPB[h]);
        pd.axis = cvt<C>(d.hitAxis[h]);
        finish_convex(d, slot, A, Bi, RA, RB, c, pd, S, dt, par);
    }
}

// Convex-vs-plane pairs (collideConvexVsPlane plane.ts:474-590 and its reverse :297-310): the hit,
// the convex shape's supporting face and the adaptive plane face (:550-575), then the pair's
// manifold and constraint, one thread per listed pair.  Kept out of the filter and the manifold
// kernels so that the convex pairs' kernels keep their registers.  C: arithmetic type (np_shapes).
template <typename T, typename C>
__global__ void __launch_bounds__(256) k_np_plane(Dev<T> d, T dt) {
    const int n = min(d.counters[CNT_PLANE], d.pairCap);
    const int par = *d.parity;
    const ShapeTable<C>& stc = np_shapes<T, C>(d);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int slot = d.planeList[k];
        const int i0 = d.pairA[slot], i1 = d.pairB[slot];
        const NpBody<T> r0 = d.npb[i0], r1 = d.npb[i1];
        const PairBodies<T> pb = order_pair_f<T>(i0, r0.flags, i1, r1.flags);
        const bool sw = pb.a != i0;
        const NpBody<T>& RA = sw ? r1 : r0;
        const NpBody<T>& RB = sw ? r0 : r1;
        const bool rev = (RA.tinfo & 15) == SHAPE_PLANE;  // the plane is the narrowphase's body A
        const NpBody<T>& RC = rev ? RB : RA;
        const NpBody<T>& RP = rev ? RA : RB;
        const V3<C> pc = cvt<C>(RC.pos), pp = cvt<C>(RP.pos);
        const Q4<C> qc = cvtq<C>(RC.q), qp = cvtq<C>(RP.q);
        PlaneHit<C> ph;
        if (!plane_hit(stc, RC.shape, pc, qc, RP.shape, pp, qp, np_tol<T, C>(d).spec, ph)) continue;
        NpScratch<C, CC_MAX_FACE> S;
        Face<C, CC_MAX_FACE>& fC = rev ? S.fB : S.fA;
        Face<C, CC_MAX_FACE>& fP = rev ? S.fA : S.fB;
        supporting_face(fC, stc, RC.shape, ph.nl, pc, qc);
        if (fC.n > 0) plane_face(fP, stc, RC.shape, pc, RP.shape, pp, qp);
        else fP.n = 0;
        if (!rev) finish_pair(d, slot, pb.a, pb.b, RA, RB, ph.pA, ph.pB, ph.axis, S.fA, S.fB, S.tmp, S.m, dt, par);
        else finish_pair(d, slot, pb.a, pb.b, RA, RB, ph.pB, ph.pA, -ph.axis, S.fA, S.fB, S.tmp, S.m, dt, par);
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

// The island group of a body of a split world: by its island label; bodies outside every
// island (static, sleeping) go to group 0, which alone writes their solver records.
__device__ inline int island_group(int label, int G) { return label >= 0 ? label % G : 0; }

// the body whose island places a contact constraint: a dynamic one (every constraint has one)
__device__ inline int con_island_body(const ConHead& c) { return (c.bits & 1) ? c.a : c.b; }

// a point constraint takes part in the step when enabled and one of its bodies is in an island
template <typename T>
__device__ inline bool pc_active(const Dev<T>& d, const PointCon<T>& pc) {
    return pc.enabled && (d.B.label[pc.a] >= 0 || d.B.label[pc.b] >= 0);
}

// group sizes of the split worlds: constraints counted by the island group of their body
template <typename T>
__global__ void k_group_zero(Dev<T> d) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < d.V; v += gridDim.x * blockDim.x) {
        d.vCnt[2 * v] = 0;
        d.vCnt[2 * v + 1] = 0;
    }
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < d.W; w += gridDim.x * blockDim.x)
        if (d.worldG[w] > 1) { d.stats[2 * w] = 0; d.stats[2 * w + 1] = 0; }
}
template <typename T>
__global__ void k_group_count(Dev<T> d) {
    const int nc = min(d.counters[CNT_CONS], d.conCap);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc + d.NPC; k += gridDim.x * blockDim.x) {
        int body, kind;
        if (k < nc) {
            body = con_island_body(d.conHead[d.conIdx[k]]);
            kind = 0;
        } else {
            const PointCon<T>& pc = d.pc[k - nc];
            if (!pc_active(d, pc)) continue;
            body = d.B.label[pc.a] >= 0 ? pc.a : pc.b;
            kind = 1;
        }
        const int w = d.B.world[body], G = d.worldG[w];
        if (G == 1) continue;
        atomicAdd(&d.vCnt[2 * (d.worldVBase[w] + island_group(d.B.label[body], G)) + kind], 1);
    }
}

// Constraints in pair order: pairs are stored in body order and bodies in world order, so
// compacting the pair list's constraint slots groups them by world with no sort.
template <typename T>
__global__ void k_con_flags(Dev<T> d) {
    int np = min(d.counters[CNT_PAIRS], d.pairCap);
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < d.pairCap; p += gridDim.x * blockDim.x)
        d.conFlag[p] = (p < np && d.pairCon[p] >= 0)
                           ? (d.hasMesh || d.hasCompound ? 1 + ((d.conHead[d.pairCon[p]].bits >> 8) & 31) : 1) : 0;
}

template <typename T>
__global__ void k_con_compact(Dev<T> d) {
    int np = min(d.counters[CNT_PAIRS], d.pairCap);
    unsigned long long pts = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        int c = d.pairCon[p];
        if (c >= 0) {
            const int bits = d.conHead[c].bits;
            const int run = d.hasMesh || d.hasCompound ? 1 + ((bits >> 8) & 31) : 1;
            for (int t = 0; t < run; t++) {
                d.conIdx[d.conPos[p] + t] = c + t;
                pts += (unsigned long long)((t ? d.conHead[c + t].bits : bits) >> 4 & 7);
            }
        }
    }
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w <= d.W; w += gridDim.x * blockDim.x)
        d.wConStart[w] = d.conPos[min(d.bpOff[d.wStart[w]], np)];
    for (int o = 32; o > 0; o >>= 1) pts += __shfl_xor(pts, o, 64);
    if ((threadIdx.x & 63) == 0 && pts) atomicAdd(&d.acc[ACC_POINTS], pts);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&d.acc[ACC_PAIRS], (unsigned long long)np);
        atomicAdd(&d.acc[ACC_CONS], (unsigned long long)min(d.counters[CNT_CONS], d.conCap));
        atomicAdd(&d.acc[ACC_EPA], (unsigned long long)min(d.counters[CNT_EPA], d.epaCap));
        atomicAdd(&d.acc[ACC_GJK], (unsigned long long)min(d.counters[CNT_GJK], d.pairCap));
        atomicAdd(&d.acc[ACC_STEPS], 1ull);
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
            const ConHead c = d.conHead[k];
            a = c.a; b = c.b;
        } else {
            const PointCon<T>& pc = d.pc[k - nc];
            if (!pc.enabled) continue;
            a = pc.a; b = pc.b;
        }
        if (d.B.label[a] >= 0 && d.B.label[b] >= 0) uf_union(d.B.label, a, b);
    }
}

// the root of x without writing any parent: when every thread stores its own body's root at
// once, a path-halving store of another thread could land after that store and leave an
// intermediate ancestor as the final label (the island would then carry two labels)
__device__ inline int uf_root(const int* p, int x) {
    for (;;) {
        const int y = __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y == x) return x;
        x = y;
    }
}

template <typename T>
__global__ void k_uf_compress(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    if (d.B.label[i] >= 0) {
        const int r = uf_root(d.B.label, i);
        __hip_atomic_store(&d.B.label[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// --------------------------------------------------------- contact cache hash
__global__ void k_hash_clear(unsigned long long* keys, int cap);

template <typename T>
__global__ void k_hash_insert(Dev<T> d) {
    int nc = min(d.counters[CNT_CONS], d.conCap);
    int par = *d.parity;
    unsigned long long* hk = d.hkeys[par];
    int* hv = d.hvals[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) d.recCnt[par] = nc;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        int slot = k;
        const ConHead c = d.conHead[slot];
        if (c.bits & 128) continue;  // not the first constraint of its pair's run
        unsigned long long key = pair_key(c.a, c.b);
        unsigned h = hash64(key) & (unsigned)(d.hashCap - 1);
        for (int probe = 0; probe < d.hashCap; probe++) {
            unsigned long long old = atomicCAS(&hk[h], CC_EMPTY_KEY, key);
            if (old == CC_EMPTY_KEY || old == key) { hv[h] = slot; break; }
            h = (h + 1) & (unsigned)(d.hashCap - 1);
        }
    }
}

// The hash of parity q rebuilt from its n records (a restored snapshot): the first record of
// each pair's run (run > 0) under the pair's key, as k_hash_insert made it.
template <typename T>
__global__ void k_hash_rebuild(Dev<T> d, int q, int n) {
    unsigned long long* hk = d.hkeys[q];
    int* hv = d.hvals[q];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const CacheRec<T>& r = d.rec[q][k];
        if (r.run <= 0) continue;
        unsigned long long key = pair_key(r.a, r.b);
        unsigned h = hash64(key) & (unsigned)(d.hashCap - 1);
        for (int probe = 0; probe < d.hashCap; probe++) {
            unsigned long long old = atomicCAS(&hk[h], CC_EMPTY_KEY, key);
            if (old == CC_EMPTY_KEY || old == key) { hv[h] = k; break; }
            h = (h + 1) & (unsigned)(d.hashCap - 1);
        }
    }
}

// The previous step's records that the last step did not match (destroyUnprocessedContacts,
// src/contacts.ts:531-555), as (a, b, subA, subB) at an atomic position; q is the parity that
// holds them.  The host sorts the list.
template <typename T>
__global__ void k_removed_contacts(Dev<T> d, int q, int4* out, int cap, int* count) {
    const int n = d.recCnt[q];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (d.recHit[q][i]) continue;
        const CacheRec<T>& r = d.rec[q][i];
        const int k = atomicAdd(count, 1);
        if (k < cap) out[k] = make_int4(r.a, r.b, (int)r.subA, (int)r.subB);
    }
}

template <typename T>
__global__ void k_flip_parity(Dev<T> d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d.parity ^= 1;
}

}  // namespace cc
