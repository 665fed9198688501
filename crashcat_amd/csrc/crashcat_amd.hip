// crashcat_amd — host side of the C ABI (include/crashcat_amd.h) and the kernel
// instantiations for FP32 (batched path) and FP64 (plain path).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <array>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/crashcat_amd.h"
#include "cc_mesh_build.h"
#include "cc_mesh.h"
#include "cc_compound.h"
#include "cc_adjoint.h"

using namespace cc;

// ------------------------------------------------------------------ scans
__global__ void __launch_bounds__(1024) cc::k_scan_blocks(const int* in, int* out, int* blockSums, int n) {
    __shared__ int sh[1024 / 64 + 1];
    int base = blockIdx.x * 4096 + threadIdx.x * 4;
    int v[4];
    int s = 0;
    for (int k = 0; k < 4; k++) {
        int i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_excl_scan<1024>(s, sh, tot);
    for (int k = 0; k < 4; k++) {
        int i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) blockSums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) cc::k_scan_sums(int* blockSums, int nb, int* total) {
    __shared__ int sh[1024 / 64 + 1];
    int base = threadIdx.x * 4;
    int v[4];
    int s = 0;
    for (int k = 0; k < 4; k++) {
        int i = base + k;
        v[k] = i < nb ? blockSums[i] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_excl_scan<1024>(s, sh, tot);
    for (int k = 0; k < 4; k++) {
        int i = base + k;
        if (i < nb) blockSums[i] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) *total = tot;
}

__global__ void cc::k_scan_add(int* out, const int* blockSums, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += blockSums[i / 4096];
}

__global__ void cc::k_hash_clear(unsigned long long* keys, int cap) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) keys[i] = CC_EMPTY_KEY;
}

__global__ void k_copy_i32(int* dst, const int* src) { *dst = *src; }

__global__ void k_fill_i32(int* p, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <typename T>
__global__ void k_hash_clear_cur(Dev<T> d) {
    unsigned long long* k = d.hkeys[*d.parity];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.hashCap; i += gridDim.x * blockDim.x) k[i] = CC_EMPTY_KEY;
}

template <typename T>
__global__ void k_set_pair_count(Dev<T> d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int n = d.bpOff[d.N];
        d.counters[CNT_PAIRS] = n;
        d.counters[CNT_EPA] = 0;
        d.counters[CNT_CONS] = 0;
        d.counters[CNT_HITS] = 0;
        d.counters[CNT_HITS_L] = 0;
        d.counters[CNT_GJK] = 0;
        d.counters[CNT_MESH] = 0;
        d.counters[CNT_PLANE] = 0;
        d.counters[CNT_COMPOUND] = 0;
        for (int c = 0; c < 32; c++) d.gjkHist[c] = 0;
        if (n > d.pairCap) atomicOr(&d.counters[CNT_OVERFLOW], 1);
    }
}

template <typename T>
__global__ void k_count_epa(Dev<T> d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        d.counters[CNT_EPA_TOTAL] += d.counters[CNT_EPA];
        // the two hit classes fill the list from either end
        if ((long long)d.counters[CNT_HITS] + d.counters[CNT_HITS_L] > d.hitCap) atomicOr(&d.counters[CNT_OVERFLOW], 16);
    }
}

// the solver records of a body as the world solve packs them (its island group writes them each
// step; those of bodies outside every island, read by any group, are kept valid from here)
template <typename T>
__device__ inline void init_solver_records(Dev<T>& d, int i) {
    const Bodies<T>& B = d.B;
    VBody<T> r;
    r.v = B.v[i];
    r.w = B.w[i];
    for (int j = 0; j < 9; j++) r.I.m[j] = T(0);
    const int f = B.flags[i];
    if (fl_motion(f) == MOTION_DYNAMIC) r.I = inv_inertia(B.invI[i], B.iRot[i], B.q[i], fl_dof(f));
    r.invMass = B.invMass[i];
    d.vb[i] = r;
    PBody<T> p;
    p.com = B.com[i];
    p.q = B.q[i];
    p.invI = B.invI[i];
    p.iRot = B.iRot[i];
    p.invMass = B.invMass[i];
    p.label = B.label[i];
    d.pb[i] = p;
}

// body initialisation at upload: COM, AABB, sleep test spheres (rigid-body.ts:390-434)
template <typename T>
__global__ void k_init_bodies(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    set_transform(d.B, d.st, i);
    reset_sleep_timer(d.B, d.st, i);
    {
        const ShapeD<T>& s = d.st.shapes[d.B.shape[i]];
        d.bBrad[i] = s.brad;
        d.bTinfo[i] = s.type | (s.mfv << 4);
    }
    d.B.wake[i] = 0;
    d.B.label[i] = -1;
    init_solver_records(d, i);
}

template <typename T>
__global__ void k_refresh_bodies(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    set_transform(d.B, d.st, i);
    // d.B.wake holds "requested awake" here: wake a sleeping dynamic body at once, as
    // wake (src/body/sleep.ts:188-203) does when a caller touches a sleeping body
    if (d.B.wake[i] && d.B.sleeping[i] && fl_motion(d.B.flags[i]) == MOTION_DYNAMIC) {
        d.B.sleeping[i] = 0;
        reset_sleep_timer(d.B, d.st, i);
    }
    d.B.wake[i] = 0;
    if (fl_motion(d.B.flags[i]) == MOTION_STATIC) init_solver_records(d, i);
}

// ---------------------------------------------------------------- context
struct Alloc {
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <typename X>
    X* get(size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&p, n * sizeof(X));
        if (e != hipSuccess) { err = e; return nullptr; }
        ptrs.push_back(p);
        return (X*)p;
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

struct Impl {
    int* counters = nullptr;  // device step counters (CNT_*), shared by both precisions
    virtual void state_changed() {}  // a step, body change or restore: a forward record no longer applies
    virtual ~Impl() {}
};

template <typename T>
struct CtxT : Impl {
    Dev<T> d;
    Alloc mem;
    int* scanSums = nullptr;
    int N = 0, W = 0, NPC = 0, numShapes = 0, maxWorld = 0;
    std::vector<int> hostWStart;
    std::vector<hipEvent_t> ev;  // 7 per step of the last cc_step call
    int solverBS = 256;
    int device = 0;
    bool hasMesh = false;
    bool anySplit = false;  // a world is solved as several island groups
    size_t worldLdsNeed = 0;   // ints of LDS the largest world can use
    size_t worldLdsBytes = 0;  // dynamic LDS given to the world kernel
    int worldLdsBS = 0;        // block size worldLdsBytes was sized for
    size_t tapeLdsBytes = 0;   // the same for the recording solver (block size 256)
    bool tapeLdsSet = false;
    size_t ccdLdsBytes = 0;    // the same for the CCD variant (block size 256)
    bool ccdLdsSet = false;
    int bodyBS = 256;          // block size of the one-thread-per-body kernels
    bool tuned = false;        // the autotuner has chosen the block sizes
    int tuneTries = 0;         // steps the autotuner waited for contacts
    // the step's phases as captured graphs, and the arguments they were captured with
    std::vector<hipGraphExec_t> graphs;
    bool graphsValid = false;
    Dev<T> graphDev;
    T graphDt = T(0);
    int graphBS = 0, graphBB = 0;
    // differentiable stepping: checkpoints of the forward run, one step record per step of a
    // segment, the adjoint state
    struct Diff {
        bool valid = false;        // cc_fwd ran and nothing changed the state since
        int n = 0, K = 0;          // steps, checkpoint interval
        double dt = 0;
        std::vector<T> ratio;      // warm-start ratio of each step
        std::vector<char*> ckpt;   // checkpoints at steps 0, K, 2K, ... and after the last step
        size_t ckptBytes = 0;
        int slots = 0;             // step records allocated
        VElem<T>* ve = nullptr;
        VPoint<T>* vp = nullptr;
        V3<T>* la = nullptr;
        V3<T>* lb = nullptr;
        int* lvl = nullptr;
        LamRec<T>* lam = nullptr;
        int* src = nullptr;
        int* rank = nullptr;
        int* stats = nullptr;
        int* wcs = nullptr;
        VBody<T>* vbody = nullptr;
        NpBody<T>* npb = nullptr;
        size_t lvlInts = 0;
        V3<T>* xb = nullptr;       // adjoints
        V3<T>* qb = nullptr;
        V3<T>* vbar = nullptr;
        V3<T>* wbar = nullptr;
        T* mub = nullptr;
        T* lamBar = nullptr;
        T* bBar = nullptr;
        V3<T>* rBar = nullptr;
        V3<T>* vf = nullptr;
        V3<T>* wf = nullptr;
        T* seed[2] = {nullptr, nullptr};
        Alloc mem;                 // everything above
    } diff;
    void state_changed() override { diff.valid = false; }
    ~CtxT() {
        for (auto& g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        diff.mem.release();
        for (char* p : diff.ckpt) (void)hipFree(p);
        mem.release();
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
};

struct cc_ctx {
    cc_config cfg;
    hipStream_t stream = nullptr;
    std::string err;
    Impl* impl = nullptr;
    int steps = 0;
    double phaseMs[12] = {0};
    int nPhase = 0;
};

static int fail(cc_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define HIPCHK(ctx, x)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(ctx, CC_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- allocation-free scratch
// Device memory of the entry points that allocate nothing (cc_download_removed_contacts_noalloc):
// one array per GPU, held until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host it is only a stand-in symbol); it uses the device address that
// hipGetSymbolAddress returned for the device, resolved once per device in cc_create.
#define CC_SCRATCH_RECORDS CC_NOALLOC_MAX_RECORDS  // removed-contact candidates (16 B each) one call stages
#define CC_SCRATCH_BYTES ((size_t)CC_SCRATCH_RECORDS * 16 + 256)
static __device__ __align__(256) unsigned char scratch[CC_SCRATCH_BYTES];
#define CC_MAX_DEVICES 64
static void* g_scratch[CC_MAX_DEVICES];  // device address of `scratch` on each device (null: not resolved)

// resolve the scratch address of the current device (called with that device current)
static hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= CC_MAX_DEVICES) return hipErrorInvalidDevice;
    if (__atomic_load_n(&g_scratch[device], __ATOMIC_ACQUIRE)) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    __atomic_store_n(&g_scratch[device], p, __ATOMIC_RELEASE);  // the same address if two threads race
    return hipSuccess;
}

extern "C" int cc_create(cc_ctx** out, const cc_config* cfg) {
    if (!out) return CC_E_ARG;
    cc_ctx* c = new cc_ctx();
    cc_config def = {CC_PRECISION_FP32, 0, 16, 8, 0, 0, 0, 0, 0, 0, 0};
    c->cfg = cfg ? *cfg : def;
    if (c->cfg.max_pairs_per_body <= 0) c->cfg.max_pairs_per_body = 16;
    if (c->cfg.max_contacts_per_body <= 0) c->cfg.max_contacts_per_body = 8;
    if (c->cfg.precision != CC_PRECISION_FP32 && c->cfg.precision != CC_PRECISION_FP64) {
        delete c;
        return CC_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete c;
        return CC_E_HIP;
    }
    if (hipSetDevice(c->cfg.device) != hipSuccess || resolve_scratch(c->cfg.device) != hipSuccess ||
        hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return CC_E_HIP;
    }
    *out = c;
    return CC_OK;
}

extern "C" int cc_destroy(cc_ctx* c) {
    if (!c) return CC_OK;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c->impl;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CC_OK;
}

extern "C" const char* cc_last_error(const cc_ctx* c) { return c ? c->err.c_str() : "null context"; }

static int next_pow2(long long x) {
    long long p = 1;
    while (p < x) p <<= 1;
    return (int)p;
}

// Grid cell width.  A body's box is the bounds of its rotated local box (the 8 corners,
// src/body/rigid-body.ts:527-553), so its width never exceeds the local box's diagonal.  The
// width is the largest such diagonal over non-static bodies, capped at four times the median
// so that one very large dynamic body does not coarsen the grid for all (it then joins the
// per-world list of large bodies, as floors do).
static double grid_cell_width(const cc_body_desc* bodies, int nb, const cc_shape_desc* shapes) {
    std::vector<double> w;
    for (int i = 0; i < nb; i++) {
        if (bodies[i].motion_type == CC_MOTION_STATIC) continue;
        const cc_shape_desc& q = shapes[bodies[i].shape];
        double dx = q.aabb_max[0] - q.aabb_min[0], dy = q.aabb_max[1] - q.aabb_min[1], dz = q.aabb_max[2] - q.aabb_min[2];
        w.push_back(std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    if (w.empty()) return 1.0;
    std::sort(w.begin(), w.end());
    double cs = std::min(w.back(), 4.0 * w[w.size() / 2]);
    return cs > 1e-6 ? cs : 1.0;
}
// Cells are the width plus twice the margin (and a relative 1e-6): two small bodies whose boxes
// overlap once one is grown by the margin then have centres less than a cell apart per axis.
template <typename T>
static void set_cell_size(Dev<T>& d, double margin) {
    d.cellSize = (d.cellWidth + 2.0 * std::max(margin, 0.0)) * (1.0 + 1e-6);
    d.cellInv = 1.0 / d.cellSize;
}

// Checks of the hull data a scene's hull shapes index, made before anything is allocated: every
// point, face and face-vertex range inside the flattened arrays, and every face vertex inside its
// hull's points (the kernels index with these values unchecked).
static int check_hulls(cc_ctx* c, const cc_shape_desc* shapes, int32_t ns, const cc_hull_data* hulls) {
    const int P = hulls ? hulls->total_points : 0, F = hulls ? hulls->total_faces : 0;
    const int FV = hulls ? hulls->total_face_vertices : 0;
    if (P < 0 || F < 0 || FV < 0) return fail(c, CC_E_ARG, "negative hull totals");
    if ((P > 0 && (!hulls->points || !hulls->shrunk_points)) ||
        (F > 0 && (!hulls->face_first_vertex || !hulls->face_num_vertices || !hulls->plane_normals)) ||
        (FV > 0 && !hulls->face_vertex_indices))
        return fail(c, CC_E_ARG, "hull data pointer is null");
    for (int i = 0; i < ns; i++) {
        const cc_shape_desc& q = shapes[i];
        if (q.type != CC_SHAPE_CONVEX_HULL) continue;
        if (q.hull_first_point < 0 || q.hull_num_points < 1 || q.hull_num_points > 255 ||
            (long long)q.hull_first_point + q.hull_num_points > P || q.hull_first_face < 0 || q.hull_num_faces < 0 ||
            (long long)q.hull_first_face + q.hull_num_faces > F)
            return fail(c, CC_E_ARG, "hull data out of range");
        for (int f = q.hull_first_face; f < q.hull_first_face + q.hull_num_faces; f++) {
            const int fv = hulls->face_first_vertex[f], nv = hulls->face_num_vertices[f];
            if (fv < 0 || nv < 0 || (long long)fv + nv > FV) return fail(c, CC_E_ARG, "hull face vertices out of range");
            for (int k = fv; k < fv + nv; k++)
                if (hulls->face_vertex_indices[k] < 0 || hulls->face_vertex_indices[k] >= q.hull_num_points)
                    return fail(c, CC_E_ARG, "hull face vertex index out of range");
        }
    }
    return CC_OK;
}

template <typename T>
static int upload_into(cc_ctx* c, CtxT<T>* X, const cc_settings* s, int32_t nw, const int32_t* woff,
                       const cc_body_desc* bodies, int32_t nb, const cc_shape_desc* shapes, int32_t ns,
                       const cc_hull_data* hulls, const cc_point_constraint_desc* pcs, int32_t npc,
                       const cc_mesh_data* meshes);

// A scene is built into a fresh context and installed only when the whole upload succeeds; a
// failed upload leaves the context with no scene (later calls report CC_E_STATE).
template <typename T>
static int upload(cc_ctx* c, const cc_settings* s, int32_t nw, const int32_t* woff, const cc_body_desc* bodies, int32_t nb,
                  const cc_shape_desc* shapes, int32_t ns, const cc_hull_data* hulls, const cc_point_constraint_desc* pcs,
                  int32_t npc, const cc_mesh_data* meshes) {
    delete c->impl;
    c->impl = nullptr;
    const int hc = check_hulls(c, shapes, ns, hulls);
    if (hc != CC_OK) return hc;
    auto* X = new CtxT<T>();
    const int rc = upload_into<T>(c, X, s, nw, woff, bodies, nb, shapes, ns, hulls, pcs, npc, meshes);
    if (rc != CC_OK) {
        (void)hipStreamSynchronize(c->stream);
        delete X;
        return rc;
    }
    c->impl = X;
    return CC_OK;
}

// The shape records and hull arrays of a scene in arithmetic type U (the context's bulk type,
// and FP64 beside it in an FP32 context whose GJK and EPA run in FP64).  Mesh fields are set by
// the caller.
template <typename U>
static int shape_table(cc_ctx* c, Alloc& A, const cc_shape_desc* shapes, int32_t ns, const cc_hull_data* hulls,
                       ShapeTable<U>& out) {
    std::vector<ShapeD<U>> sh(std::max(ns, 1));
    for (int i = 0; i < ns; i++) {
        const cc_shape_desc& q = shapes[i];
        ShapeD<U>& o = sh[i];
        o.type = q.type; o.hfp = q.hull_first_point; o.hnp = q.hull_num_points; o.hff = q.hull_first_face;
        o.hnf = q.hull_num_faces;
        o.mfv = q.type == CC_SHAPE_BOX || q.type == CC_SHAPE_PLANE ? 4
                : (q.type == CC_SHAPE_CAPSULE ? 2 : (q.type == CC_SHAPE_CYLINDER ? 8 : 0));
        if (q.type == CC_SHAPE_TRIANGLE_MESH) {
            o.hfp = q.mesh_first_triangle; o.hnp = q.mesh_num_triangles;
            o.hff = q.mesh_first_node; o.hnf = q.mesh_num_nodes;
            o.mfv = 3;
        }
        o.cfc = q.type == CC_SHAPE_COMPOUND ? q.compound_first_child : 0;
        o.ncc = q.type == CC_SHAPE_COMPOUND ? q.compound_num_children : 0;
        if (q.type == CC_SHAPE_COMPOUND) o.mfv = CC_MAX_FACE;
        o.radius = U(q.radius); o.hh = U(q.half_height); o.cr = U(q.convex_radius);
        o.deco = q.decorated;
        for (int k = 0; k < 3; k++) {
            o.sc[k] = q.decorated & CC_DECO_SCALED ? U(q.scale[k]) : U(1);
            o.lp[k] = q.decorated & CC_DECO_TRANSFORMED ? U(q.local_position[k]) : U(0);
        }
        for (int k = 0; k < 4; k++) o.lq[k] = q.decorated & CC_DECO_TRANSFORMED ? U(q.local_quaternion[k]) : U(k == 3);
        for (int k = 0; k < 3; k++) {
            o.he[k] = U(q.half_extents[k]); o.com[k] = U(q.center_of_mass[k]);
            o.amin[k] = U(q.aabb_min[k]); o.amax[k] = U(q.aabb_max[k]);
        }
        // bounding radius, generous (any support point of either mode lies inside it)
        double br = 0.0;
        double he2 = q.half_extents[0] * q.half_extents[0] + q.half_extents[1] * q.half_extents[1] + q.half_extents[2] * q.half_extents[2];
        double com = std::sqrt(q.center_of_mass[0] * q.center_of_mass[0] + q.center_of_mass[1] * q.center_of_mass[1] +
                               q.center_of_mass[2] * q.center_of_mass[2]);
        if (q.type == CC_SHAPE_SPHERE) br = q.radius;
        else if (q.type == CC_SHAPE_BOX) br = std::sqrt(he2) + q.convex_radius;
        else if (q.type == CC_SHAPE_CAPSULE) br = q.half_height + q.radius;
        else if (q.type == CC_SHAPE_CYLINDER) br = std::sqrt(q.radius * q.radius + q.half_height * q.half_height);
        else if (q.type == CC_SHAPE_PLANE) br = 1e30;
        else if (hulls && hulls->points && hulls->shrunk_points) {
            for (int k = 0; k < q.hull_num_points && q.hull_first_point + k < hulls->total_points; k++) {
                const double* pp = hulls->points + 3 * (q.hull_first_point + k);
                const double* ss = hulls->shrunk_points + 3 * (q.hull_first_point + k);
                br = std::max(br, std::sqrt(pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]));
                br = std::max(br, std::sqrt(ss[0] * ss[0] + ss[1] * ss[1] + ss[2] * ss[2]) + q.convex_radius);
            }
            br += com;
        } else br = 1e30;
        if (q.decorated & CC_DECO_SCALED) {  // a bound of the scaled leaf (a scaled hull's shrunk points are scaled already)
            const double ms = std::max(std::fabs(q.scale[0]), std::max(std::fabs(q.scale[1]), std::fabs(q.scale[2])));
            br *= std::max(ms, 1.0);
        }
        if (q.decorated & CC_DECO_TRANSFORMED)  // about the body position, the leaf moved by its local translation
            br += std::sqrt(q.local_position[0] * q.local_position[0] + q.local_position[1] * q.local_position[1] +
                            q.local_position[2] * q.local_position[2]);
        o.brad = U(br * (1.0 + 1e-5) + 1e-6);
    }
    for (int i = 0; i < ns; i++) {  // a compound holds its children's spheres about their positions
        const cc_shape_desc& q = shapes[i];
        if (q.type != CC_SHAPE_COMPOUND) continue;
        double br = 0.0;
        for (int k = q.compound_first_child; k < q.compound_first_child + q.compound_num_children; k++)
            br = std::max(br, (double)sh[k].brad);  // a child's radius counts its local translation already
        sh[i].brad = U(br * (1.0 + 1e-5) + 1e-6);
    }
    int P = hulls ? hulls->total_points : 0, F = hulls ? hulls->total_faces : 0, FV = hulls ? hulls->total_face_vertices : 0;
    for (int i = 0; i < ns; i++) {  // a hull's supporting face takes at most 16 of a face's vertices
        if (shapes[i].type != CC_SHAPE_CONVEX_HULL) continue;
        int mfv = 0;
        for (int f = shapes[i].hull_first_face; f < shapes[i].hull_first_face + shapes[i].hull_num_faces; f++) {
            int nv = hulls->face_num_vertices[f];
            int delta = (nv + 15) / 16;
            mfv = std::max(mfv, delta > 0 ? (nv + delta - 1) / delta : 0);
        }
        sh[i].mfv = mfv;
    }
    std::vector<U> hp(3 * std::max(P, 1)), hs(3 * std::max(P, 1)), hn(3 * std::max(F, 1));
    for (int i = 0; i < 3 * P; i++) { hp[i] = U(hulls->points[i]); hs[i] = U(hulls->shrunk_points[i]); }
    for (int i = 0; i < 3 * F; i++) hn[i] = U(hulls->plane_normals[i]);
    ShapeD<U>* dShapes = A.get<ShapeD<U>>(sh.size());
    U* dPts = A.get<U>(hp.size());
    U* dShr = A.get<U>(hs.size());
    U* dPn = A.get<U>(hn.size());
    int* dFfv = A.get<int>(std::max(F, 1));
    int* dFnv = A.get<int>(std::max(F, 1));
    int* dFvi = A.get<int>(std::max(FV, 1));
    if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed (shapes)");
    HIPCHK(c, hipMemcpy(dShapes, sh.data(), sh.size() * sizeof(ShapeD<U>), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dPts, hp.data(), hp.size() * sizeof(U), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dShr, hs.data(), hs.size() * sizeof(U), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dPn, hn.data(), hn.size() * sizeof(U), hipMemcpyHostToDevice));
    if (F > 0) {
        HIPCHK(c, hipMemcpy(dFfv, hulls->face_first_vertex, F * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dFnv, hulls->face_num_vertices, F * sizeof(int), hipMemcpyHostToDevice));
    }
    if (FV > 0) HIPCHK(c, hipMemcpy(dFvi, hulls->face_vertex_indices, FV * sizeof(int), hipMemcpyHostToDevice));
    out = ShapeTable<U>{dShapes, dPts, dShr, dFfv, dFnv, dFvi, dPn, nullptr, nullptr, nullptr};
    return CC_OK;
}

template <typename T>
static int upload_into(cc_ctx* c, CtxT<T>* X, const cc_settings* s, int32_t nw, const int32_t* woff,
                       const cc_body_desc* bodies, int32_t nb, const cc_shape_desc* shapes, int32_t ns,
                       const cc_hull_data* hulls, const cc_point_constraint_desc* pcs, int32_t npc,
                       const cc_mesh_data* meshes) {
    Dev<T>& d = X->d;
    memset(&d, 0, sizeof(d));
    Alloc& A = X->mem;
    X->N = nb;
    X->W = nw;
    X->NPC = npc;
    // settings
    for (int k = 0; k < 3; k++) d.s.gravity[k] = T(s->gravity[k]);
    d.s.spec = T(s->speculative_contact_distance);
    d.s.manifoldTol = T(s->manifold_tolerance);
    d.s.colTol = T(s->collision_tolerance);
    d.s.penTol = T(s->penetration_tolerance);
    d.s.preserveDist2 = T(s->contact_point_preserve_lambda_max_dist_sq);
    d.s.minVelRest = T(s->min_velocity_for_restitution);
    d.s.slop = T(s->penetration_slop);
    d.s.beta = T(s->baumgarte_factor);
    d.s.maxPen = T(s->max_penetration_distance);
    d.s.tSleep = T(s->time_before_sleep);
    d.s.sleepThresh = T(s->point_velocity_sleep_threshold);
    d.s.velIters = s->velocity_iterations;
    d.s.posIters = s->position_iterations;
    d.s.gravityEnabled = s->gravity_enabled;
    d.s.warmStarting = s->warm_starting;
    d.s.allowSleeping = s->allow_sleeping;
    memcpy(d.s.layerPairs, s->object_layer_pairs, 64);
    d.s.normalCos = T(s->normal_cos_max_delta_rotation);
    d.s.backfaces = s->collide_with_backfaces;
    d.s.activeEdges = s->collide_only_with_active_edges;
    d.s.ccdThr = T(s->ccd_linear_cast_threshold);
    d.s.ccdMaxPen = T(s->ccd_linear_cast_max_penetration);
    d.s.ccdMinVelRest = T(s->ccd_min_velocity_for_restitution);
    // shapes
    bool anyDeco = false, anyMeshOrPlane = false;
    for (int i = 0; i < ns; i++) {
        const cc_shape_desc& q = shapes[i];
        if (q.type < 0 || q.type > CC_SHAPE_COMPOUND) return fail(c, CC_E_UNSUPPORTED, "unsupported shape type " + std::to_string(q.type));
        if (q.decorated & ~(CC_DECO_SCALED | CC_DECO_TRANSFORMED | CC_DECO_OFFSET_COM))
            return fail(c, CC_E_ARG, "unknown decorator bits");
        if (q.decorated) {
            if (q.type == CC_SHAPE_TRIANGLE_MESH || q.type == CC_SHAPE_PLANE)
                return fail(c, CC_E_UNSUPPORTED, "decorated triangle meshes and planes are not on the GPU path");
            if ((q.decorated & CC_DECO_SCALED) && (q.scale[0] == 0 || q.scale[1] == 0 || q.scale[2] == 0))
                return fail(c, CC_E_ARG, "a scale component is zero");
            anyDeco = true;
        }
        anyMeshOrPlane = anyMeshOrPlane || q.type == CC_SHAPE_TRIANGLE_MESH || q.type == CC_SHAPE_PLANE;
        if (q.type == CC_SHAPE_COMPOUND) {  // children: convex leaves (scaled allowed) with their local transform
            anyDeco = true;
            if (q.compound_num_children < 1 || q.compound_first_child < 0 ||
                (long long)q.compound_first_child + q.compound_num_children > ns)
                return fail(c, CC_E_ARG, "compound children out of range");
            for (int k = q.compound_first_child; k < q.compound_first_child + q.compound_num_children; k++) {
                const int t = shapes[k].type;
                if (t == CC_SHAPE_COMPOUND || t == CC_SHAPE_TRIANGLE_MESH || t == CC_SHAPE_PLANE)
                    return fail(c, CC_E_UNSUPPORTED, "compound children must be convex shapes");
            }
        }
        if (q.type == CC_SHAPE_TRIANGLE_MESH) {
            const int MT = meshes ? meshes->total_triangles : 0, MN = meshes ? meshes->total_nodes : 0;
            if (q.mesh_first_triangle < 0 || q.mesh_num_triangles < 0 || q.mesh_first_node < 0 || q.mesh_num_nodes < 0 ||
                q.mesh_first_triangle + q.mesh_num_triangles > MT || q.mesh_first_node + q.mesh_num_nodes > MN ||
                (q.mesh_num_nodes > 0 && (!meshes->nodes || !meshes->triangles || !meshes->positions)))
                return fail(c, CC_E_ARG, "mesh data out of range");
        }
    }
    if (anyDeco && anyMeshOrPlane)
        return fail(c, CC_E_UNSUPPORTED,
                    "decorated or compound shapes in a scene with a triangle mesh or a plane are not on the GPU path");
    {
        const int rc = shape_table<T>(c, A, shapes, ns, hulls, d.st);
        if (rc != CC_OK) return rc;
    }
    // FP32 contexts run GJK and EPA in FP64 unless asked not to: FP64 copies of the shapes and
    // tolerances (the EPA simplices are allocated with the other narrowphase buffers)
    d.npWide = sizeof(T) == 4 && !c->cfg.narrowphase_fp32 ? 1 : 0;
    d.spec64 = s->speculative_contact_distance;
    d.colTol64 = s->collision_tolerance;
    d.penTol64 = s->penetration_tolerance;
    d.manifoldTol64 = s->manifold_tolerance;
    d.normalCos64 = s->normal_cos_max_delta_rotation;
    if (d.npWide) {
        const int rc = shape_table<double>(c, A, shapes, ns, hulls, d.st64);
        if (rc != CC_OK) return rc;
    }
    // triangle meshes: vertices and triangles as given, BVH nodes decoded from the node buffer
    X->hasMesh = false;
    for (int i = 0; i < ns; i++) X->hasMesh = X->hasMesh || shapes[i].type == CC_SHAPE_TRIANGLE_MESH;
    if (X->hasMesh) {
        const int MV = meshes->total_vertices, MT = meshes->total_triangles, MN = meshes->total_nodes;
        std::vector<T> mp(3 * (size_t)std::max(MV, 1));
        for (size_t k = 0; k < 3 * (size_t)MV; k++) mp[k] = T(meshes->positions[k]);
        std::vector<int4> mt((size_t)std::max(MT, 1));
        for (int t = 0; t < MT; t++) {
            const int32_t* r = meshes->triangles + 4 * (size_t)t;
            for (int k = 0; k < 3; k++)
                if (r[k] < 0 || r[k] >= MV) return fail(c, CC_E_ARG, "mesh triangle vertex out of range");
            mt[t] = make_int4(r[0], r[1], r[2], r[3] & 7);
        }
        std::vector<MNode<T>> mn((size_t)std::max(MN, 1));
        for (int i = 0; i < ns; i++) {  // decode each mesh's nodes (offsets count doubles from its first node)
            const cc_shape_desc& q = shapes[i];
            if (q.type != CC_SHAPE_TRIANGLE_MESH) continue;
            for (int j = 0; j < q.mesh_num_nodes; j++) {
                const double* nd = meshes->nodes + 8 * (size_t)(q.mesh_first_node + j);
                MNode<T>& o = mn[(size_t)(q.mesh_first_node + j)];
                for (int k = 0; k < 3; k++) {
                    T lo = T(nd[k]), hi = T(nd[3 + k]);
                    if ((double)lo > nd[k]) lo = std::nextafter(lo, T(-INFINITY));  // FP32: round outwards
                    if ((double)hi < nd[3 + k]) hi = std::nextafter(hi, T(INFINITY));
                    o.mn[k] = lo; o.mx[k] = hi;
                }
                if (nd[7] < 0) {
                    const int st = (int)nd[6], cnt = -((int)nd[7] + 1);
                    if (st < 0 || cnt < 0 || st + cnt > q.mesh_num_triangles) return fail(c, CC_E_ARG, "mesh BVH leaf out of range");
                    o.a = st; o.b = (int)nd[7];
                } else {
                    const long long off = (long long)nd[6];
                    if (off % 8 != 0 || off / 8 <= j || off / 8 >= q.mesh_num_nodes || j + 1 >= q.mesh_num_nodes)
                        return fail(c, CC_E_ARG, "mesh BVH child offset out of range");
                    o.a = (int)(off / 8); o.b = (int)nd[7];
                }
            }
        }
        T* dMp = A.get<T>(mp.size());
        int4* dMt = A.get<int4>(mt.size());
        MNode<T>* dMn = A.get<MNode<T>>(mn.size());
        if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed (meshes)");
        HIPCHK(c, hipMemcpy(dMp, mp.data(), mp.size() * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dMt, mt.data(), mt.size() * sizeof(int4), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dMn, mn.data(), mn.size() * sizeof(MNode<T>), hipMemcpyHostToDevice));
        d.st.mpos = dMp;
        d.st.mtri = dMt;
        d.st.mnode = dMn;
    }
    // bodies
    X->hostWStart.assign(woff, woff + nw + 1);
    std::vector<int> worldOf(std::max(nb, 1));
    for (int w = 0; w < nw; w++)
        for (int i = woff[w]; i < woff[w + 1]; i++) worldOf[i] = w;
    std::vector<V3<T>> pos(nb), com(nb), v(nb), wv(nb), invI(nb);
    std::vector<Q4<T>> q(nb), ir(nb);
    std::vector<T> im(nb), fr(nb), re(nb), gf(nb), ld(nb), ad(nb), ml(nb), ma(nb);
    std::vector<int> shp(nb), flags(nb), slp(nb), ovr(nb);
    std::vector<unsigned> grp(nb), msk(nb);
    std::vector<V3<T>> frc(nb), trq(nb);
    bool anyOvr = false, anyForce = false;
    for (int i = 0; i < nb; i++) {
        const cc_body_desc& b = bodies[i];
        if (b.shape < 0 || b.shape >= ns) return fail(c, CC_E_ARG, "body shape index out of range");
        if (b.motion_type < 0 || b.motion_type > 2) return fail(c, CC_E_ARG, "invalid motion type");
        if (b.object_layer < 0 || b.object_layer >= CC_MAX_OBJECT_LAYERS) return fail(c, CC_E_ARG, "invalid object layer");
        if (X->hasMesh && shapes[b.shape].type == CC_SHAPE_TRIANGLE_MESH && b.motion_type != CC_MOTION_STATIC)
            return fail(c, CC_E_UNSUPPORTED, "triangle-mesh bodies must be static");
        if (shapes[b.shape].type == CC_SHAPE_PLANE && b.motion_type != CC_MOTION_STATIC)
            return fail(c, CC_E_UNSUPPORTED, "plane bodies must be static (plane.ts:53-57)");
        if (b.sensor) return fail(c, CC_E_UNSUPPORTED, "sensor bodies are not on the GPU path (body " + std::to_string(i) + ")");
        if (b.motion_quality != CC_MOTION_QUALITY_DISCRETE && b.motion_quality != CC_MOTION_QUALITY_LINEAR_CAST)
            return fail(c, CC_E_ARG, "invalid motion quality (body " + std::to_string(i) + ")");
        if (b.num_velocity_steps_override < 0 || b.num_velocity_steps_override > 255 || b.num_position_steps_override < 0 ||
            b.num_position_steps_override > 255)
            return fail(c, CC_E_ARG, "iteration overrides must be in [0, 255]");
        anyOvr = anyOvr || (b.motion_type == CC_MOTION_DYNAMIC &&
                            (b.num_velocity_steps_override != 0 || b.num_position_steps_override != 0));
        ovr[i] = b.num_velocity_steps_override | (b.num_position_steps_override << 8);
        frc[i] = mk3<T>(0, 0, 0);
        trq[i] = mk3<T>(0, 0, 0);
        if (b.motion_type == CC_MOTION_DYNAMIC) {  // addForce / addTorque touch dynamic bodies only
            frc[i] = mk3<T>(T(b.force[0]), T(b.force[1]), T(b.force[2]));
            trq[i] = mk3<T>(T(b.torque[0]), T(b.torque[1]), T(b.torque[2]));
            for (int k = 0; k < 3; k++) anyForce = anyForce || b.force[k] != 0.0 || b.torque[k] != 0.0;
        }
        pos[i] = mk3<T>(T(b.position[0]), T(b.position[1]), T(b.position[2]));
        q[i].x = T(b.quaternion[0]); q[i].y = T(b.quaternion[1]); q[i].z = T(b.quaternion[2]); q[i].w = T(b.quaternion[3]);
        v[i] = mk3<T>(T(b.linear_velocity[0]), T(b.linear_velocity[1]), T(b.linear_velocity[2]));
        wv[i] = mk3<T>(T(b.angular_velocity[0]), T(b.angular_velocity[1]), T(b.angular_velocity[2]));
        invI[i] = mk3<T>(T(b.inv_inertia_diagonal[0]), T(b.inv_inertia_diagonal[1]), T(b.inv_inertia_diagonal[2]));
        ir[i].x = T(b.inertia_rotation[0]); ir[i].y = T(b.inertia_rotation[1]); ir[i].z = T(b.inertia_rotation[2]);
        ir[i].w = T(b.inertia_rotation[3]);
        im[i] = T(b.inv_mass); fr[i] = T(b.friction); re[i] = T(b.restitution); gf[i] = T(b.gravity_factor);
        ld[i] = T(b.linear_damping); ad[i] = T(b.angular_damping); ml[i] = T(b.max_linear_velocity);
        ma[i] = T(b.max_angular_velocity);
        shp[i] = b.shape;
        flags[i] = (b.motion_type & 3) | ((b.allowed_dofs & 63) << 2) | ((b.allow_sleeping ? 1 : 0) << 8) |
                   ((b.friction_combine & 7) << 9) | ((b.restitution_combine & 7) << 12) | ((b.object_layer & 255) << 16) |
                   (b.motion_quality == CC_MOTION_QUALITY_LINEAR_CAST ? FL_LINEAR_CAST : 0);
        slp[i] = b.motion_type == CC_MOTION_STATIC ? 0 : (b.sleeping ? 1 : 0);
        grp[i] = b.collision_groups;
        msk[i] = b.collision_mask;
        com[i] = pos[i];
    }
    Bodies<T>& B = d.B;
#define UP(field, vec, type)                                                                    \
    B.field = A.get<type>(std::max<size_t>(vec.size(), 1));                                     \
    if (!B.field) return fail(c, CC_E_HIP, "hipMalloc failed (bodies)");                        \
    if (!vec.empty()) HIPCHK(c, hipMemcpy(B.field, vec.data(), vec.size() * sizeof(type), hipMemcpyHostToDevice));
    UP(pos, pos, V3<T>) UP(q, q, Q4<T>) UP(com, com, V3<T>) UP(v, v, V3<T>) UP(w, wv, V3<T>) UP(invMass, im, T)
    UP(invI, invI, V3<T>) UP(iRot, ir, Q4<T>) UP(friction, fr, T) UP(restitution, re, T) UP(gravityFactor, gf, T)
    UP(linDamp, ld, T) UP(angDamp, ad, T) UP(maxLinV, ml, T) UP(maxAngV, ma, T) UP(shape, shp, int) UP(flags, flags, int)
    UP(groups, grp, unsigned) UP(mask, msk, unsigned) UP(world, worldOf, int) UP(sleeping, slp, int)
    UP(force, frc, V3<T>) UP(torque, trq, V3<T>) UP(ovr, ovr, int)
#undef UP
    B.amin = A.get<V3<T>>(nb);
    B.amax = A.get<V3<T>>(nb);
    B.sc = A.get<V3<T>>(3 * (size_t)nb);
    B.sr = A.get<T>(3 * (size_t)nb);
    B.stimer = A.get<T>(nb);
    B.label = A.get<int>(nb);
    B.wake = A.get<int>(nb);
    // CCD: one record per dynamic LINEAR_CAST body, in a world of undecorated convex leaves (the
    // shape cast of castConvexVsConvex, convex.ts:675-875)
    std::vector<int> ccdStart(nw + 1, 0);
    {
        std::vector<T> inner(nb, T(0));
        for (int w = 0; w < nw; w++) {
            int n = 0;
            for (int i = woff[w]; i < woff[w + 1]; i++)
                n += (bodies[i].motion_quality == CC_MOTION_QUALITY_LINEAR_CAST && bodies[i].motion_type == CC_MOTION_DYNAMIC) ? 1 : 0;
            ccdStart[w + 1] = ccdStart[w] + n;
            if (n == 0) continue;
            for (int i = woff[w]; i < woff[w + 1]; i++) {
                const cc_shape_desc& q = shapes[bodies[i].shape];
                const bool convex = q.type == CC_SHAPE_SPHERE || q.type == CC_SHAPE_BOX || q.type == CC_SHAPE_CAPSULE ||
                                    q.type == CC_SHAPE_CONVEX_HULL || q.type == CC_SHAPE_CYLINDER;
                // the LINEAR_CAST bodies are cast (undecorated convex leaves); the others are hit
                // (convex leaves with their decorators, planes, compounds of convex leaves)
                const bool cast = bodies[i].motion_quality == CC_MOTION_QUALITY_LINEAR_CAST &&
                                  bodies[i].motion_type == CC_MOTION_DYNAMIC;
                // (castScaledVsShape and castOffsetCenterOfMassVsShape pass the pose through; the one
                // transformed decorator a chain holds turns the displacement by the body rotation)
                const bool ok = cast ? convex || q.type == CC_SHAPE_COMPOUND
                                     : convex || q.type == CC_SHAPE_PLANE || q.type == CC_SHAPE_COMPOUND;
                if (!ok)
                    return fail(c, CC_E_UNSUPPORTED, "MotionQuality.LINEAR_CAST (CCD) is on the GPU path for convex LINEAR_CAST "
                                                     "bodies in worlds of convex "
                                                     "shapes, planes and compounds (world " + std::to_string(w) + ", body " +
                                                     std::to_string(i) + ")");
                // getShapeInnerRadius: sphere.ts:147-149, box.ts:333-335, capsule.ts:254-256,
                // cylinder.ts:261-263, convex-hull.ts:550-558 (the nearest face plane),
                // scaled.ts:127-130, compound.ts:283-291 (the smallest child's)
                auto leafInner = [&](const cc_shape_desc& p) {
                    double r = 0;
                    if (p.type == CC_SHAPE_SPHERE || p.type == CC_SHAPE_CAPSULE) r = p.radius;
                    else if (p.type == CC_SHAPE_BOX) r = std::min(p.half_extents[0], std::min(p.half_extents[1], p.half_extents[2]));
                    else if (p.type == CC_SHAPE_CYLINDER) r = std::min(p.half_height, p.radius);
                    else if (p.type == CC_SHAPE_CONVEX_HULL) {
                        r = 1.7976931348623157e308;
                        for (int f = 0; f < p.hull_num_faces; f++) r = std::min(r, -hulls->plane_constants[p.hull_first_face + f]);
                        r = std::max(0.0, r);
                    }
                    if (p.decorated & CC_DECO_SCALED)
                        r = std::min(std::fabs(p.scale[0]), std::min(std::fabs(p.scale[1]), std::fabs(p.scale[2]))) * r;
                    return r;
                };
                double r = 0;
                if (q.type == CC_SHAPE_COMPOUND) {
                    r = 1.7976931348623157e308;
                    for (int k = 0; k < q.compound_num_children; k++) r = std::min(r, leafInner(shapes[q.compound_first_child + k]));
                } else {
                    r = leafInner(q);
                }
                inner[i] = T(r);
            }
        }
        d.hasCcd = ccdStart[nw] > 0 ? 1 : 0;
        if (d.hasCcd) {
            const int NC = ccdStart[nw];
            d.ccdStart = A.get<int>(nw + 1);
            d.ccdCount = A.get<int>(nw);
            d.ccd = A.get<CcdRec<T>>(NC);
            d.ccdOrd = A.get<int>(NC);
            d.ccdIdx = A.get<int>(nb);
            d.ccdPart = A.get<CcdPart<T>>(2 * (size_t)nw);
            d.ccdInner = A.get<T>(nb);
            if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed (CCD)");
            std::vector<CcdPart<T>> zp(2 * (size_t)nw);
            std::memset((void*)zp.data(), 0, zp.size() * sizeof(CcdPart<T>));
            std::vector<int> none(nb, -1);
            HIPCHK(c, hipMemcpy(d.ccdStart, ccdStart.data(), (nw + 1) * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemset(d.ccdCount, 0, nw * sizeof(int)));
            HIPCHK(c, hipMemcpy(d.ccdPart, zp.data(), zp.size() * sizeof(CcdPart<T>), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(d.ccdIdx, none.data(), nb * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(d.ccdInner, inner.data(), nb * sizeof(T), hipMemcpyHostToDevice));
        }
    }
    d.N = nb;
    d.W = nw;
    int* dW = A.get<int>(nw + 1);
    if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed");
    HIPCHK(c, hipMemcpy(dW, woff, (nw + 1) * sizeof(int), hipMemcpyHostToDevice));
    d.wStart = dW;
    // the world solver's workgroups: worlds largest first (longest solves first); in a batch of
    // fewer than 64 worlds each world is split into island groups of about 64 bodies, at most
    // 256 / worlds of them, so that its independent islands fill more compute units
    size_t splitExtra = 0;
    {
        std::vector<int> G(nw), vbase(nw), xbase(nw);
        int V = 0;
        for (int w = 0; w < nw; w++) {
            const int n = woff[w + 1] - woff[w];
            int g = 1;
            if (nw < 64) g = std::max(1, std::min((n + 63) / 64, std::max(1, 256 / nw)));
            if (c->cfg.island_groups > 0) g = std::min(g, (int)c->cfg.island_groups);
            if (ccdStart[w + 1] > ccdStart[w]) g = 1;  // the CCD pass needs the whole world in one workgroup
            G[w] = g;
            vbase[w] = V;
            V += g;
            xbase[w] = nb + (int)splitExtra;
            if (g > 1) splitExtra += (size_t)(g - 1) * n;
        }
        std::vector<int> order(nw);
        for (int w = 0; w < nw; w++) order[w] = w;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return woff[x + 1] - woff[x] > woff[y + 1] - woff[y]; });
        std::vector<int4> vt;
        vt.reserve(V);
        for (int w : order)
            for (int g = 0; g < G[w]; g++)
                vt.push_back(make_int4(w, g, G[w], g == 0 ? woff[w] : xbase[w] + (g - 1) * (woff[w + 1] - woff[w])));
        int4* dV = A.get<int4>(std::max(V, 1));
        int* dG = A.get<int>(nw);
        int* dB = A.get<int>(nw);
        d.vCnt = A.get<int>(2 * (size_t)std::max(V, 1));
        if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed");
        HIPCHK(c, hipMemcpy(dV, vt.data(), V * sizeof(int4), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dG, G.data(), nw * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dB, vbase.data(), nw * sizeof(int), hipMemcpyHostToDevice));
        d.vTab = dV;
        d.V = V;
        d.worldG = dG;
        d.worldVBase = dB;
        X->anySplit = V > nw;
    }
    // point constraints, grouped by world in index order
    std::vector<std::pair<int, int>> pcw;
    for (int i = 0; i < npc; i++) {
        int a = pcs[i].body_a, b = pcs[i].body_b;
        if (a < 0 || a >= nb || b < 0 || b >= nb || worldOf[a] != worldOf[b] || a == b)
            return fail(c, CC_E_ARG, "point constraint bodies invalid or in different worlds");
        const int ov = pcs[i].num_velocity_steps_override, op = pcs[i].num_position_steps_override;
        if (ov < 0 || ov > 255 || op < 0 || op > 255) return fail(c, CC_E_ARG, "iteration overrides must be in [0, 255]");
        anyOvr = anyOvr || ov != 0 || op != 0;
        pcw.push_back({worldOf[a], i});
    }
    std::stable_sort(pcw.begin(), pcw.end(), [](auto& x, auto& y) { return x.first < y.first; });
    std::vector<PointCon<T>> hpc(std::max(npc, 1));
    std::vector<int> pcStart(nw + 1, 0);
    for (int k = 0; k < npc; k++) {
        const cc_point_constraint_desc& p = pcs[pcw[k].second];
        PointCon<T>& o = hpc[k];
        memset(&o, 0, sizeof(o));
        o.a = p.body_a; o.b = p.body_b; o.enabled = p.enabled;
        o.localA = mk3<T>(T(p.local_point_a[0]), T(p.local_point_a[1]), T(p.local_point_a[2]));
        o.localB = mk3<T>(T(p.local_point_b[0]), T(p.local_point_b[1]), T(p.local_point_b[2]));
        o.ovr = p.num_velocity_steps_override | (p.num_position_steps_override << 8);
        pcStart[pcw[k].first + 1]++;
    }
    for (int w = 0; w < nw; w++) pcStart[w + 1] += pcStart[w];
    d.pc = A.get<PointCon<T>>(hpc.size());
    int* dPcs = A.get<int>(nw + 1);
    if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed (constraints)");
    HIPCHK(c, hipMemcpy(d.pc, hpc.data(), hpc.size() * sizeof(PointCon<T>), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dPcs, pcStart.data(), (nw + 1) * sizeof(int), hipMemcpyHostToDevice));
    d.pcStart = dPcs;
    d.NPC = npc;
    // capacities
    long long pairCap = (long long)c->cfg.max_pairs_per_body * std::max(nb, 1);
    if (pairCap > (1ll << 30)) pairCap = 1ll << 30;
    d.pairCap = (int)pairCap;
    d.conCap = (int)std::min<long long>((long long)c->cfg.max_contacts_per_body * std::max(nb, 1), pairCap);
    d.epaCap = std::max(4096, d.pairCap / 16);  // EPA jobs: pairs whose cores overlap
    d.epaWorkers = std::min(d.epaCap, 131072);
    d.hashCap = next_pow2(2ll * d.conCap);
    int maxWorld = 0;
    for (int w = 0; w < nw; w++) maxWorld = std::max(maxWorld, woff[w + 1] - woff[w]);
    if (nb > (1 << 23)) return fail(c, CC_E_CAPACITY, "more than 8388608 bodies in one context");
    d.cellWidth = grid_cell_width(bodies, nb, shapes);
    set_cell_size(d, (double)s->speculative_contact_distance);
    d.cellTabCap = (int)next_pow2(std::max<long long>(2ll * nb, 64));
    d.cellKey = A.get<unsigned long long>(d.cellTabCap);
    d.cellCount = A.get<int>((size_t)d.cellTabCap + 1);
    d.cellStart = A.get<int>((size_t)d.cellTabCap + 1);
    d.cellFill = A.get<int>(d.cellTabCap);
    d.cellList = A.get<int>(nb);
    d.bodySlot = A.get<int>(nb);
    d.cellRec = A.get<CellRec<T>>(nb);
    d.gLargeFlag = A.get<int>(nb + 1);
    d.gLargeOff = A.get<int>(nb + 1);
    d.gLarge = A.get<int>(nb);
    d.hasMesh = X->hasMesh ? 1 : 0;
    d.hasPlane = 0;
    for (int i = 0; i < ns; i++) d.hasPlane = d.hasPlane || shapes[i].type == CC_SHAPE_PLANE;
    if (d.hasPlane) d.planeList = A.get<int>(d.pairCap);
    d.hasCompound = 0;
    for (int i = 0; i < ns; i++) d.hasCompound = d.hasCompound || shapes[i].type == CC_SHAPE_COMPOUND;
    if (d.hasCompound) {  // one collector and one EPA polytope per compound thread
        d.cmpList = A.get<int>(d.pairCap);
        d.cmpWorkers = (int)std::min<long long>(4096, ((long long)d.pairCap + 63) / 64 * 64);
        const size_t coll = d.npWide ? sizeof(MeshColl<double>) : sizeof(MeshColl<T>);
        const size_t wave = d.npWide ? EpaWave<double>::bytes : EpaWave<T>::bytes;
        d.cmpColl = A.get<char>((size_t)d.cmpWorkers * coll);
        d.cmpEpa = A.get<char>((size_t)((d.cmpWorkers + EPA_LANES - 1) / EPA_LANES) * wave);
    }
    if (X->hasMesh) {  // one collector per mesh thread; the EPA polytopes are the EPA kernel's
        d.meshList = A.get<int>(d.pairCap);
        d.meshWorkers = ((d.epaWorkers + EPA_LANES - 1) / EPA_LANES) * EPA_LANES;
        d.meshWork = A.get<char>((size_t)d.meshWorkers * sizeof(MeshColl<T>));
        d.meshCnt = A.get<int>((size_t)d.pairCap + 1);
        d.meshOff = A.get<int>((size_t)d.pairCap + 1);
        d.candCap = (int)std::min<long long>(64ll * std::max(nb, 1), 1ll << 30);
        d.candTri = A.get<int>(d.candCap);
        d.candPair = A.get<int>(d.candCap);
        d.candHit = A.get<MeshHit<T>>(d.candCap);
    }
    d.npb = A.get<NpBody<T>>(nb);
    d.npd = A.get<NpDyn<T>>(nb);
    d.bBrad = A.get<T>(nb);
    d.bTinfo = A.get<int>(nb);
    d.bpCount = A.get<int>(nb + 1);
    d.bpOff = A.get<int>(nb + 1);
    d.bpSlots = A.get<int>((size_t)nb * BP_SLOTS);
    d.spillCap = std::max(nb, 1024);
    d.spillA = A.get<int>(d.spillCap);
    d.spillB = A.get<int>(d.spillCap);
    d.spillK = A.get<int>(d.spillCap);
    d.pairCon = A.get<int>(d.pairCap);
    d.conFlag = A.get<int>((size_t)d.pairCap + 1);
    d.conPos = A.get<int>((size_t)d.pairCap + 1);
    d.pairA = A.get<int>(d.pairCap);
    d.pairB = A.get<int>(d.pairCap);
    d.counters = A.get<int>(CNT_N);
    X->counters = d.counters;
    d.con = A.get<Con<T>>(d.conCap);
    d.conHead = A.get<ConHead>(d.conCap);
    d.tap = A.get<ConTap<T>>(d.conCap);
    d.wConStart = A.get<int>(nw + 1);
    d.conIdx = A.get<int>(d.conCap);
    for (int k = 0; k < 2; k++) {
        d.rec[k] = A.get<CacheRec<T>>(d.conCap);
        d.recHit[k] = A.get<int>(d.conCap);
        d.hkeys[k] = A.get<unsigned long long>(d.hashCap);
        d.hvals[k] = A.get<int>(d.hashCap);
    }
    d.recCnt = A.get<int>(2);
    d.gjkList = A.get<int>(d.pairCap);
    d.gjkCombo = A.get<int>(d.pairCap);
    d.gjkSorted = A.get<int>(d.pairCap);
    d.gjkHist = A.get<int>(32);
    d.hitCap = d.conCap + d.conCap / 8 + 1024;
    d.hitSlot = A.get<int>(d.hitCap);
    d.hitKind = A.get<int>(d.hitCap);
    d.hitPA = A.get<V3<double>>(d.hitCap);
    d.hitPB = A.get<V3<double>>(d.hitCap);
    d.hitAxis = A.get<V3<double>>(d.hitCap);
    d.epaPair = A.get<int>(d.epaCap);
    d.epaSimplex = A.get<Simplex<T>>(d.epaCap);
    if (d.npWide) d.epaSimplex64 = A.get<Simplex<double>>(d.epaCap);
    d.epaWork = A.get<char>((size_t)((d.epaWorkers + EPA_LANES - 1) / EPA_LANES) *
                            std::max(EpaWave<T>::bytes, EpaWave<double>::bytes));
    size_t L = (size_t)d.conCap + npc;
    d.skeys = A.get<unsigned long long>(2 * L);
    d.seq = A.get<int>(L);
    d.seqA = A.get<int>(L);
    d.seqB = A.get<int>(L);
    d.prevA = A.get<int>(L);
    d.prevB = A.get<int>(L);
    d.nextA = A.get<int>(L);
    d.nextB = A.get<int>(L);
    d.level = A.get<int>(L);
    d.sched = A.get<int>(L);
    d.elemNp = A.get<int>(L);
    d.lvlStart = A.get<int>(2 * (L + d.V + 1));
    d.blist = A.get<int>(2 * L);
    d.blkey = A.get<unsigned long long>(4 * L);
    d.svCap = (int)L;
    d.svSlot = A.get<int>(L);
    d.ve = A.get<VElem<T>>(L);
    d.vp = A.get<VPoint<T>>(4 * L);
    d.svPt = A.get<int>(L);
    d.svLA = A.get<V3<T>>(4 * L);
    d.svLB = A.get<V3<T>>(4 * L);
    d.vb = A.get<VBody<T>>(std::max(nb, 1));
    d.pb = A.get<PBody<T>>(std::max(nb, 1));
    d.bcnt = A.get<int>((size_t)nb + splitExtra);
    d.boff = A.get<int>((size_t)nb + splitExtra);
    d.isl0 = A.get<int>(nb);
    d.isl1 = A.get<int>(nb);
    d.stats = A.get<int>(2 * (size_t)nw);
    d.islIt = anyOvr ? A.get<int>(3 * (size_t)std::max(nb, 1)) : nullptr;
    d.hasForces = anyForce ? 1 : 0;
    d.prevDt = A.get<T>(1);
    d.parity = A.get<int>(1);
    d.acc = A.get<unsigned long long>(ACC_N);
    d.colored = c->cfg.solver_order == CC_ORDER_COLORED ? 1 : 0;
    d.resid = c->cfg.measure_residual ? A.get<double>(4) : nullptr;
    {  // scan workspace for the longest scan (pair flags, grid buckets or bodies)
        long long n = std::max<long long>({(long long)d.pairCap + 1, (long long)d.cellTabCap + 1, (long long)nb + 1});
        X->scanSums = A.get<int>((size_t)(n / 4096 + n / (4096ll * 4096) + 64));
    }
    X->maxWorld = maxWorld;
    X->device = c->cfg.device;
    X->worldLdsBS = 0;
    X->worldLdsNeed = 0;
    for (int w = 0; w < nw; w++) {  // per-body counts + in-degrees packed 16 per int
        size_t bw = (size_t)(woff[w + 1] - woff[w]);
        size_t lw = (size_t)c->cfg.max_contacts_per_body * bw + (size_t)(pcStart[w + 1] - pcStart[w]);
        X->worldLdsNeed = std::max(X->worldLdsNeed, ((bw + 3) & ~(size_t)3) + (lw + 15) / 16);
    }
    if (A.err) return fail(c, CC_E_HIP, "hipMalloc failed (workspace)");
    HIPCHK(c, hipMemset(d.counters, 0, CNT_N * sizeof(int)));
    HIPCHK(c, hipMemset(d.prevDt, 0, sizeof(T)));
    HIPCHK(c, hipMemset(d.acc, 0, ACC_N * sizeof(unsigned long long)));
    if (d.resid) HIPCHK(c, hipMemset(d.resid, 0, 4 * sizeof(double)));
    HIPCHK(c, hipMemset(d.parity, 0, sizeof(int)));
    for (int k = 0; k < 2; k++) HIPCHK(c, hipMemset(d.hkeys[k], 0xff, d.hashCap * sizeof(unsigned long long)));
    HIPCHK(c, hipMemset(d.recCnt, 0, 2 * sizeof(int)));

    HIPCHK(c, hipMemset(d.stats, 0, 2 * nw * sizeof(int)));
    int blocks = (nb + 255) / 256;
    if (nb > 0) hipLaunchKernelGGL(k_init_bodies<T>, dim3(blocks), dim3(256), 0, c->stream, d);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int bs = c->cfg.block_solver;
    const int bbq = c->cfg.block_bodies;
    X->bodyBS = (bbq == 64 || bbq == 128 || bbq == 256 || bbq == 512 || bbq == 1024) ? bbq : 256;
    X->tuned = false;
    X->tuneTries = 0;
    if (bs == 64 || bs == 128 || bs == 256 || bs == 512 || bs == 1024) X->solverBS = bs;
    else {
        X->solverBS = maxWorld > 4096 ? 1024 : (maxWorld > 1024 ? 512 : 256);
        // FP64 records need twice the registers: 1024 lanes would spill
        if (sizeof(T) == 8) X->solverBS = std::min(X->solverBS, 512);
    }
    return CC_OK;
}

extern "C" int cc_upload_scene(cc_ctx* c, const cc_settings* s, int32_t nw, const int32_t* woff, const cc_body_desc* bodies,
                               int32_t nb, const cc_shape_desc* shapes, int32_t ns, const cc_hull_data* hulls,
                               const cc_point_constraint_desc* pcs, int32_t npc, const cc_mesh_data* meshes) {
    if (!c || !s || nw <= 0 || !woff || nb < 0 || (nb > 0 && !bodies) || ns < 0 || (ns > 0 && !shapes))
        return fail(c, CC_E_ARG, "invalid arguments to cc_upload_scene");
    if (woff[0] != 0 || woff[nw] != nb) return fail(c, CC_E_ARG, "world offsets must start at 0 and end at num_bodies");
    for (int w = 0; w < nw; w++)
        if (woff[w + 1] < woff[w]) return fail(c, CC_E_ARG, "world offsets must be non-decreasing");
    if (npc > 0 && !pcs) return fail(c, CC_E_ARG, "point constraints pointer is null");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->steps = 0;
    if (c->cfg.precision == CC_PRECISION_FP64)
        return upload<double>(c, s, nw, woff, bodies, nb, shapes, ns, hulls, pcs, npc, meshes);
    return upload<float>(c, s, nw, woff, bodies, nb, shapes, ns, hulls, pcs, npc, meshes);
}

extern "C" int cc_build_triangle_mesh(const double* positions, int32_t nv, const int32_t* indices, int32_t nt,
                                      double degenerate_tolerance, double active_edge_cos_threshold, int32_t max_leaf,
                                      double* out_positions, int32_t* out_triangles, double* out_nodes,
                                      int32_t* out_counts) {
    if (nv < 0 || nt < 0 || (nv > 0 && !positions) || (nt > 0 && !indices) || max_leaf < 1 || !out_counts ||
        (nv > 0 && !out_positions) || (nt > 0 && (!out_triangles || !out_nodes)))
        return CC_E_ARG;
    ccmesh::Builder B;
    if (ccmesh::build_triangle_mesh(positions, nv, indices, nt, degenerate_tolerance, active_edge_cos_threshold, max_leaf,
                                    B) != 0)
        return CC_E_ARG;
    const size_t V = B.pos.size() / 3, T = B.tri.size(), N = B.nodes.size() / 8;
    if (V > (size_t)nv || T > (size_t)nt || N > 2 * (size_t)nt) return CC_E_ARG;
    if (V) memcpy(out_positions, B.pos.data(), sizeof(double) * 3 * V);
    for (size_t t = 0; t < T; t++) {
        out_triangles[4 * t] = B.tri[t].a;
        out_triangles[4 * t + 1] = B.tri[t].b;
        out_triangles[4 * t + 2] = B.tri[t].c;
        out_triangles[4 * t + 3] = B.tri[t].act;
    }
    if (N) memcpy(out_nodes, B.nodes.data(), sizeof(double) * 8 * N);
    out_counts[0] = (int32_t)V;
    out_counts[1] = (int32_t)T;
    out_counts[2] = (int32_t)N;
    return CC_OK;
}

static int grid_for(long long n, int bs, int maxBlocks) {
    long long g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > maxBlocks) g = maxBlocks;
    return (int)g;
}

// exclusive scan of n ints into out[0..n), total in out[n]; work holds the per-block sums of
// every level (about n / 4000 ints)
static void scan_launch(hipStream_t st, const int* in, int* out, int n, int* work) {
    int nb = std::max(1, (n + 4095) / 4096);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(1024), 0, st, in, out, work, n);
    if (nb <= 4096) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, work, nb, out + n);
    } else {
        scan_launch(st, work, work, nb, work + nb + 1);  // in place; its total lands in work[nb]
        hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(1), 0, st, out + n, work + nb);
    }
    hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1), dim3(256), 0, st, out, work, n);
}

// ---------------------------------------------------------------- the step's phases
// A step is a fixed sequence of launches with no host round trip, in phases that each end at a
// HIP event (the phase timers) and are each replayed as one captured hipGraph: the big kernels
// have a phase of their own, so the timers name the dominant kernel.
enum {
    PH_ACCEL = 0,  // k_accel
    PH_GRID,       // cells: clear, insert, scans, scatter; pairs with large bodies
    PH_BPCELLS,    // k_bp_cells
    PH_PAIRS,      // pair layout: scan, compaction, spill
    PH_FILTER,     // k_np_filter and the GJK list's counting sort
    PH_GJK,        // k_np_gjk
    PH_EPA,        // k_epa
    PH_MANIFOLD,   // k_np_manifold<T, 8>
    PH_MANIFOLD_L, // k_np_manifold<T, 32>, planes, triangle meshes
    PH_ISLANDS,    // wake, constraint compaction, union-find
    PH_SOLVER,     // k_world (levels, warm start, sweeps, integration, sleep)
    PH_CACHE,      // contact-cache hash, end of step
    PH_N
};
#define CC_PHASES PH_N

// the world kernel's dynamic LDS: what the largest world can use (its per-body counts and packed
// in-degrees), within what the device leaves beside the kernel's static LDS
template <typename T, int BS, bool TAPE, bool CCD = false>
static size_t world_lds_bytes(CtxT<T>* X, size_t needInts) {
    hipFuncAttributes fa;
    int maxShared = 0;
    size_t dyn = 0;
    if (hipFuncGetAttributes(&fa, (const void*)k_world<T, BS, TAPE, CCD>) == hipSuccess &&
        hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, X->device) == hipSuccess &&
        (size_t)maxShared > fa.sharedSizeBytes) {
        size_t avail = ((size_t)maxShared - fa.sharedSizeBytes) / 4;
        dyn = std::min(avail, needInts) * 4;
    }
    if (dyn > 0 && hipFuncSetAttribute((const void*)k_world<T, BS, TAPE, CCD>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
        dyn = 0;
    return dyn;
}


// set the solver's kernel attributes for the current block size (outside any graph capture)
template <typename T>
static void prepare_world(CtxT<T>* X, bool tape) {
    if (tape) {
        if (!X->tapeLdsSet) {
            X->tapeLdsBytes = world_lds_bytes<T, 256, true>(X, X->worldLdsNeed);
            X->tapeLdsSet = true;
        }
        X->d.worldLdsInts = (int)(X->tapeLdsBytes / 4);
        return;
    }
    if (X->d.hasCcd) {  // worlds with LINEAR_CAST bodies: the CCD variant at 256 threads
        if (!X->ccdLdsSet) {
            X->ccdLdsBytes = world_lds_bytes<T, 256, false, true>(X, X->worldLdsNeed);
            X->ccdLdsSet = true;
        }
        X->d.worldLdsInts = (int)(X->ccdLdsBytes / 4);
        return;
    }
    if (X->worldLdsBS != X->solverBS) {
        switch (X->solverBS) {
            case 64: X->worldLdsBytes = world_lds_bytes<T, 64, false>(X, X->worldLdsNeed); break;
            case 128: X->worldLdsBytes = world_lds_bytes<T, 128, false>(X, X->worldLdsNeed); break;
            case 512: X->worldLdsBytes = world_lds_bytes<T, 512, false>(X, X->worldLdsNeed); break;
            case 1024: X->worldLdsBytes = world_lds_bytes<T, 1024, false>(X, X->worldLdsNeed); break;
            default: X->worldLdsBytes = world_lds_bytes<T, 256, false>(X, X->worldLdsNeed); break;
        }
        X->worldLdsBS = X->solverBS;
    }
    X->d.worldLdsInts = (int)(X->worldLdsBytes / 4);
}

template <typename T>
static void launch_world(CtxT<T>* X, hipStream_t st, T dt, bool tape) {
    Dev<T>& d = X->d;
    if (X->anySplit) {
        hipLaunchKernelGGL(k_group_zero<T>, dim3(grid_for(std::max(d.V, d.W), 256, 1024)), dim3(256), 0, st, d);
        hipLaunchKernelGGL(k_group_count<T>, dim3(grid_for((long long)d.conCap + d.NPC, 256, 4096)), dim3(256), 0, st, d);
    }
    if (tape) {  // the recording solver of differentiable stepping (worlds are never split there)
        hipLaunchKernelGGL((k_world<T, 256, true>), dim3(d.V), dim3(256), X->tapeLdsBytes, st, d, dt);
        return;
    }
    if (d.hasCcd) {
        hipLaunchKernelGGL((k_world<T, 256, false, true>), dim3(d.V), dim3(256), X->ccdLdsBytes, st, d, dt);
        return;
    }
    const size_t lds = X->worldLdsBytes;
    switch (X->solverBS) {
        case 64: hipLaunchKernelGGL((k_world<T, 64>), dim3(d.V), dim3(64), lds, st, d, dt); break;
        case 128: hipLaunchKernelGGL((k_world<T, 128>), dim3(d.V), dim3(128), lds, st, d, dt); break;
        case 512: hipLaunchKernelGGL((k_world<T, 512>), dim3(d.V), dim3(512), lds, st, d, dt); break;
        case 1024: hipLaunchKernelGGL((k_world<T, 1024>), dim3(d.V), dim3(1024), lds, st, d, dt); break;
        default: hipLaunchKernelGGL((k_world<T, 256>), dim3(d.V), dim3(256), lds, st, d, dt); break;
    }
}

// the launches of one phase; BB is the block size of the one-thread-per-body kernels
template <typename T>
static void enqueue_phase(CtxT<T>* X, hipStream_t st, int p, T dt, double margin, bool tape) {
    Dev<T>& d = X->d;
    const int N = d.N;
    const int BB = X->bodyBS;
    const int nbB = (N + BB - 1) / BB;
    switch (p) {
        case PH_ACCEL:
            hipLaunchKernelGGL(k_accel<T>, dim3(nbB), dim3(BB), 0, st, d, dt);
            break;
        case PH_GRID:
            hipLaunchKernelGGL(k_cells_clear<T>, dim3(grid_for(std::max(d.cellTabCap, N), 256, 4096)), dim3(256), 0, st, d);
            hipLaunchKernelGGL(k_cells_insert<T>, dim3(nbB), dim3(BB), 0, st, d);
            scan_launch(st, d.cellCount, d.cellStart, d.cellTabCap, X->scanSums);
            scan_launch(st, d.gLargeFlag, d.gLargeOff, N, X->scanSums);
            hipLaunchKernelGGL(k_cells_scatter<T>, dim3(nbB), dim3(BB), 0, st, d);
            hipLaunchKernelGGL(k_bp_large<T>, dim3(nbB), dim3(BB), 0, st, d, margin);
            break;
        case PH_BPCELLS:
            hipLaunchKernelGGL(k_bp_cells<T>, dim3(grid_for((N + 3) / 4, 1, 8192)), dim3(256), 0, st, d, margin);
            break;
        case PH_PAIRS:
            scan_launch(st, d.bpCount, d.bpOff, N, X->scanSums);
            hipLaunchKernelGGL(k_set_pair_count<T>, dim3(1), dim3(64), 0, st, d);
            hipLaunchKernelGGL(k_bp_compact<T>, dim3(nbB), dim3(BB), 0, st, d);
            hipLaunchKernelGGL(k_bp_spill<T>, dim3(grid_for(d.spillCap, 256, 1024)), dim3(256), 0, st, d);
            if (d.diff) hipLaunchKernelGGL(k_bp_sort<T>, dim3(nbB), dim3(BB), 0, st, d);
            break;
        case PH_FILTER:
            hipLaunchKernelGGL(k_np_filter<T>, dim3(grid_for(d.pairCap, 256, 8192)), dim3(256), 0, st, d, dt);
            hipLaunchKernelGGL(k_gjk_count<T>, dim3(grid_for(d.pairCap, 256, 2048)), dim3(256), 0, st, d);
            hipLaunchKernelGGL(k_gjk_offsets<T>, dim3(1), dim3(64), 0, st, d);
            hipLaunchKernelGGL(k_gjk_scatter<T>, dim3(grid_for(d.pairCap, 256, 4096)), dim3(256), 0, st, d);
            break;
        case PH_GJK:
            if (d.npWide) hipLaunchKernelGGL((k_np_gjk<T, double>), dim3(grid_for(d.pairCap, 256, 8192)), dim3(256), 0, st, d, dt);
            else hipLaunchKernelGGL((k_np_gjk<T, T>), dim3(grid_for(d.pairCap, 256, 8192)), dim3(256), 0, st, d, dt);
            break;
        case PH_EPA:
            if (d.npWide) hipLaunchKernelGGL((k_epa<T, double>), dim3((d.epaWorkers + 63) / 64), dim3(64), 0, st, d, dt);
            else hipLaunchKernelGGL((k_epa<T, T>), dim3((d.epaWorkers + 63) / 64), dim3(64), 0, st, d, dt);
            hipLaunchKernelGGL(k_count_epa<T>, dim3(1), dim3(64), 0, st, d);
            break;
        case PH_MANIFOLD:
            if (d.npWide)
                hipLaunchKernelGGL((k_np_manifold<T, CC_SMALL_FACE, double>), dim3(grid_for(d.hitCap, 256, 8192)), dim3(256), 0, st, d, dt);
            else
                hipLaunchKernelGGL((k_np_manifold<T, CC_SMALL_FACE, T>), dim3(grid_for(d.hitCap, 256, 8192)), dim3(256), 0, st, d, dt);
            break;
        case PH_MANIFOLD_L:
            if (d.npWide)
                hipLaunchKernelGGL((k_np_manifold<T, CC_MAX_FACE, double>), dim3(grid_for(d.hitCap / 8, 256, 1024)), dim3(256), 0, st, d, dt);
            else
                hipLaunchKernelGGL((k_np_manifold<T, CC_MAX_FACE, T>), dim3(grid_for(d.hitCap / 8, 256, 1024)), dim3(256), 0, st, d, dt);
            if (d.hasCompound && d.npWide)
                hipLaunchKernelGGL((k_compound<T, double>), dim3(d.cmpWorkers / 64), dim3(64), 0, st, d, dt);
            else if (d.hasCompound)
                hipLaunchKernelGGL((k_compound<T, T>), dim3(d.cmpWorkers / 64), dim3(64), 0, st, d, dt);
            if (d.hasPlane && d.npWide)
                hipLaunchKernelGGL((k_np_plane<T, double>), dim3(grid_for(d.pairCap, 256, 2048)), dim3(256), 0, st, d, dt);
            else if (d.hasPlane)
                hipLaunchKernelGGL((k_np_plane<T, T>), dim3(grid_for(d.pairCap, 256, 2048)), dim3(256), 0, st, d, dt);
            if (d.hasMesh) {
                hipLaunchKernelGGL(k_mesh_count<T>, dim3(grid_for(d.pairCap, 256, 4096)), dim3(256), 0, st, d);
                scan_launch(st, d.meshCnt, d.meshOff, d.pairCap, X->scanSums);
                hipLaunchKernelGGL(k_mesh_emit<T>, dim3(grid_for(d.pairCap, 256, 4096)), dim3(256), 0, st, d);
                hipLaunchKernelGGL(k_mesh_tri<T>, dim3(d.meshWorkers / MESH_LANES), dim3(MESH_LANES), 0, st, d);
                hipLaunchKernelGGL(k_mesh_collect<T>, dim3(d.meshWorkers / MESH_LANES), dim3(MESH_LANES), 0, st, d, dt);
            }
            break;
        case PH_ISLANDS:
            hipLaunchKernelGGL(k_wake<T>, dim3(nbB), dim3(BB), 0, st, d);
            hipLaunchKernelGGL(k_con_flags<T>, dim3(grid_for(d.pairCap, 256, 4096)), dim3(256), 0, st, d);
            scan_launch(st, d.conFlag, d.conPos, d.pairCap, X->scanSums);
            hipLaunchKernelGGL(k_con_compact<T>, dim3(grid_for(std::max<long long>(d.pairCap, d.W + 1), 256, 4096)),
                               dim3(256), 0, st, d);
            hipLaunchKernelGGL(k_uf_init<T>, dim3(nbB), dim3(BB), 0, st, d);
            hipLaunchKernelGGL(k_uf_link<T>, dim3(grid_for(d.conCap + d.NPC, 256, 2048)), dim3(256), 0, st, d);
            hipLaunchKernelGGL(k_uf_compress<T>, dim3(nbB), dim3(BB), 0, st, d);
            break;
        case PH_SOLVER:
            launch_world<T>(X, st, dt, tape);
            break;
        case PH_CACHE:
            hipLaunchKernelGGL(k_hash_clear_cur<T>, dim3(grid_for(d.hashCap, 256, 2048)), dim3(256), 0, st, d);
            hipLaunchKernelGGL(k_hash_insert<T>, dim3(grid_for(d.conCap, 256, 2048)), dim3(256), 0, st, d);
            hipLaunchKernelGGL(k_end_step<T>, dim3(1), dim3(64), 0, st, d, dt);
            break;
    }
}

// broadphase of one step on its own (cc_broadphase_aabbs): cells, pairs, layout
template <typename T>
static void launch_broadphase(CtxT<T>* X, hipStream_t st, double margin) {
    for (int p = PH_GRID; p <= PH_PAIRS; p++) enqueue_phase<T>(X, st, p, T(0), margin, false);
}

// The phases as graphs, captured again whenever a kernel argument changes (dt, a device
// pointer or count, a block size).  Returns false when capture is not available (the phases
// are then launched directly).
template <typename T>
static bool phase_graphs(CtxT<T>* X, hipStream_t st, T dt, double margin) {
    const bool same = X->graphsValid && X->graphDt == dt && X->graphBS == X->solverBS && X->graphBB == X->bodyBS &&
                      memcmp(&X->graphDev, &X->d, sizeof(Dev<T>)) == 0;
    if (same) return true;
    for (auto& g : X->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    X->graphs.assign(PH_N, nullptr);
    X->graphsValid = false;
    for (int p = 0; p < PH_N; p++) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        enqueue_phase<T>(X, st, p, dt, margin, false);
        if (hipStreamEndCapture(st, &g) != hipSuccess || !g) return false;
        const hipError_t e = hipGraphInstantiate(&X->graphs[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return false;
    }
    X->graphDev = X->d;
    X->graphDt = dt;
    X->graphBS = X->solverBS;
    X->graphBB = X->bodyBS;
    X->graphsValid = true;
    return true;
}

// enqueue nsteps steps (graphs when allowed), recording the phase events of each
template <typename T>
static int enqueue_steps(cc_ctx* c, CtxT<T>* X, T dt, int nsteps, bool tape, bool graphs) {
    hipStream_t st = c->stream;
    const double margin = (double)X->d.s.spec;
    prepare_world<T>(X, tape);
    const bool useGraphs = graphs && !tape && phase_graphs<T>(X, st, dt, margin);
    size_t needEv = (size_t)(PH_N + 1) * nsteps;
    while (X->ev.size() < needEv) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        X->ev.push_back(e);
    }
    for (int s = 0; s < nsteps; s++) {
        hipEvent_t* E = X->ev.data() + (size_t)(PH_N + 1) * s;
        HIPCHK(c, hipEventRecord(E[0], st));
        for (int p = 0; p < PH_N; p++) {
            if (useGraphs) HIPCHK(c, hipGraphLaunch(X->graphs[p], st));
            else enqueue_phase<T>(X, st, p, dt, margin, tape);
            HIPCHK(c, hipEventRecord(E[p + 1], st));
        }
    }
    HIPCHK(c, hipGetLastError());
    return CC_OK;
}

// The autotuner: the solver's block size and that of the one-thread-per-body kernels, chosen by
// timing trial steps from a checkpoint of the state (restored after each: results do not depend
// on block sizes), once per workload (cached by precision, bodies, worlds, largest world, device).
// It runs at the first cc_step after upload whose previous step made constraints, so that the
// trials see contacts.
struct TuneKey {
    int prec, n, w, maxWorld, device, fixedSolver, fixedBodies;
    bool operator<(const TuneKey& o) const {
        return std::tie(prec, n, w, maxWorld, device, fixedSolver, fixedBodies) <
               std::tie(o.prec, o.n, o.w, o.maxWorld, o.device, o.fixedSolver, o.fixedBodies);
    }
};
static std::map<TuneKey, std::array<int, 2>> g_tuned;  // solver block, body block
static std::mutex g_tunedLock;

template <typename T>
static std::vector<std::pair<void*, size_t>> ckpt_segments(Dev<T>& d);
template <typename T>
static hipError_t ckpt_copy(CtxT<T>* X, hipStream_t st, char* buf, bool save);

template <typename T>
static int autotune(cc_ctx* c, CtxT<T>* X, T dt) {
    Dev<T>& d = X->d;
    const TuneKey key{c->cfg.precision, d.N, d.W, X->maxWorld, X->device, c->cfg.block_solver, c->cfg.block_bodies};
    {
        std::lock_guard<std::mutex> lk(g_tunedLock);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            if (c->cfg.block_solver == 0) X->solverBS = it->second[0];
            if (c->cfg.block_bodies == 0) X->bodyBS = it->second[1];
            X->tuned = true;
            return CC_OK;
        }
    }
    hipStream_t st = c->stream;
    int cnt[CNT_N];
    HIPCHK(c, hipMemcpy(cnt, d.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[CNT_CONS] <= 0) {  // not yet: the trials would see no contacts
        if (++X->tuneTries >= 8) X->tuned = true;  // a scene without contacts keeps the rule-based sizes
        return CC_OK;
    }
    size_t bytes = 0;
    for (auto& sg : ckpt_segments(d)) bytes += sg.second;
    char* buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) {  // no room for the trials: keep the rule-based sizes
        (void)hipGetLastError();
        X->tuned = true;
        return CC_OK;
    }
    HIPCHK(c, ckpt_copy(X, st, buf, true));
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    // one timed step from the checkpoint per candidate, after an untimed one (a kernel's first
    // launch loads its code)
    auto trial = [&](int bs, int bb) -> float {
        X->solverBS = bs;
        X->bodyBS = bb;
        prepare_world<T>(X, false);
        float ms = 0;
        for (int r = 0; r < 2; r++) {
            (void)ckpt_copy(X, st, buf, false);
            (void)hipEventRecord(e0, st);
            for (int p = 0; p < PH_N; p++) enqueue_phase<T>(X, st, p, dt, (double)d.s.spec, false);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        return ms;
    };
    std::vector<int> sCand;
    if (c->cfg.block_solver != 0) sCand = {X->solverBS};
    else if (sizeof(T) == 8) sCand = X->maxWorld > 1024 ? std::vector<int>{256, 512} : std::vector<int>{128, 256, 512};
    else sCand = X->maxWorld > 1024 ? std::vector<int>{256, 512, 1024} : std::vector<int>{128, 256, 512};
    std::vector<int> bCand = c->cfg.block_bodies != 0 ? std::vector<int>{X->bodyBS} : std::vector<int>{128, 256, 512};
    const int bb0 = c->cfg.block_bodies != 0 ? X->bodyBS : 256;
    int bestS = sCand[0], bestB = bb0;
    float best = 1e30f;
    for (int bs : sCand) {
        const float t = trial(bs, bb0);
        if (t < best) { best = t; bestS = bs; }
    }
    for (int bb : bCand) {
        if (bb == bb0) continue;
        const float t = trial(bestS, bb);
        if (t < best) { best = t; bestB = bb; }
    }
    X->solverBS = bestS;
    X->bodyBS = bestB;
    prepare_world<T>(X, false);
    HIPCHK(c, ckpt_copy(X, st, buf, false));
    HIPCHK(c, hipStreamSynchronize(st));
    (void)hipFree(buf);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    X->tuned = true;
    std::lock_guard<std::mutex> lk(g_tunedLock);
    g_tuned[key] = {bestS, bestB};
    return CC_OK;
}

template <typename T>
static int step_impl(cc_ctx* c, double dtd, int nsteps, bool tape = false) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    const T dt = T(dtd);
    if (d.N == 0) {  // nothing to step; keep the step count for island reporting
        c->steps += nsteps;
        c->nPhase = 0;
        return CC_OK;
    }
    if (tape && d.hasCcd)
        return fail(c, CC_E_UNSUPPORTED, "differentiable stepping of worlds with LINEAR_CAST (CCD) bodies is not supported");
    if (!tape && !d.diff && !X->tuned && nsteps > 0 && (c->cfg.block_solver == 0 || c->cfg.block_bodies == 0)) {
        const int rc = autotune<T>(c, X, dt);
        if (rc != CC_OK) return rc;
    }
    const int rc = enqueue_steps<T>(c, X, dt, nsteps, tape, c->cfg.disable_graphs == 0);
    if (rc != CC_OK) return rc;
    c->steps += nsteps;
    c->nPhase = nsteps;  // number of steps recorded
    return CC_OK;
}

extern "C" int cc_step(cc_ctx* c, double dt, int32_t nsteps) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (nsteps < 0 || !(dt >= 0)) return fail(c, CC_E_ARG, "invalid dt or nsteps");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    if (c->cfg.precision == CC_PRECISION_FP64) return step_impl<double>(c, dt, nsteps);
    return step_impl<float>(c, dt, nsteps);
}

template <typename T>
static int counters_impl(cc_ctx* c, cc_step_counters* o);

extern "C" int cc_sync(cc_ctx* c) {
    if (!c) return CC_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->impl && c->impl->counters) {
        // a capacity overflow drops pairs, constraints or EPA jobs for the rest of that step,
        // so the state is no longer the reference's: report it (it stays set until re-upload)
        int ovf = 0;
        HIPCHK(c, hipMemcpy(&ovf, c->impl->counters + CNT_OVERFLOW, sizeof(int), hipMemcpyDeviceToHost));
        if (ovf) {
            std::string m = "capacity exceeded:";
            if (ovf & 1) m += " pairs (raise max_pairs_per_body)";
            if (ovf & 2) m += " contact constraints (raise max_contacts_per_body)";
            if (ovf & 4) m += " EPA jobs";
            if (ovf & 8) m += " pair spill list";
            if (ovf & 16) m += " narrowphase hits";
            if (ovf & 32) m += " mesh BVH traversal stack (a BVH deeper than 63 levels)";
            if (ovf & 64) m += " mesh candidate triangles (64 per body)";
            return fail(c, CC_E_CAPACITY, m);
        }
    }
    return CC_OK;
}

template <typename T>
static int download_impl(cc_ctx* c, cc_body_state* out, int n) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    if (n != d.N) return fail(c, CC_E_ARG, "num_bodies mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<V3<T>> pos(n), com(n), v(n), w(n), mn(n), mx(n);
    std::vector<Q4<T>> q(n);
    std::vector<int> sl(n), lab(n);
    if (n > 0) {
        HIPCHK(c, hipMemcpy(pos.data(), d.B.pos, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(com.data(), d.B.com, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(v.data(), d.B.v, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(w.data(), d.B.w, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(mn.data(), d.B.amin, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(mx.data(), d.B.amax, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(q.data(), d.B.q, n * sizeof(Q4<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(sl.data(), d.B.sleeping, n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(lab.data(), d.B.label, n * sizeof(int), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < n; i++) {
        cc_body_state& o = out[i];
        o.position[0] = pos[i].x; o.position[1] = pos[i].y; o.position[2] = pos[i].z;
        o.quaternion[0] = q[i].x; o.quaternion[1] = q[i].y; o.quaternion[2] = q[i].z; o.quaternion[3] = q[i].w;
        o.center_of_mass_position[0] = com[i].x; o.center_of_mass_position[1] = com[i].y; o.center_of_mass_position[2] = com[i].z;
        o.linear_velocity[0] = v[i].x; o.linear_velocity[1] = v[i].y; o.linear_velocity[2] = v[i].z;
        o.angular_velocity[0] = w[i].x; o.angular_velocity[1] = w[i].y; o.angular_velocity[2] = w[i].z;
        o.aabb_min[0] = mn[i].x; o.aabb_min[1] = mn[i].y; o.aabb_min[2] = mn[i].z;
        o.aabb_max[0] = mx[i].x; o.aabb_max[1] = mx[i].y; o.aabb_max[2] = mx[i].z;
        o.sleeping = sl[i];
        o.island = c->steps > 0 ? lab[i] : -1;
    }
    return CC_OK;
}

extern "C" int cc_download_bodies(cc_ctx* c, cc_body_state* out, int32_t n) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (n > 0 && !out) return fail(c, CC_E_ARG, "null output");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return download_impl<double>(c, out, n);
    return download_impl<float>(c, out, n);
}

template <typename T>
static int set_bodies_impl(cc_ctx* c, const cc_body_state* in, int n) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    if (n != d.N) return fail(c, CC_E_ARG, "num_bodies mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<V3<T>> pos(n), v(n), w(n);
    std::vector<Q4<T>> q(n);
    for (int i = 0; i < n; i++) {
        pos[i] = mk3<T>(T(in[i].position[0]), T(in[i].position[1]), T(in[i].position[2]));
        v[i] = mk3<T>(T(in[i].linear_velocity[0]), T(in[i].linear_velocity[1]), T(in[i].linear_velocity[2]));
        w[i] = mk3<T>(T(in[i].angular_velocity[0]), T(in[i].angular_velocity[1]), T(in[i].angular_velocity[2]));
        q[i].x = T(in[i].quaternion[0]); q[i].y = T(in[i].quaternion[1]); q[i].z = T(in[i].quaternion[2]);
        q[i].w = T(in[i].quaternion[3]);
    }
    if (n > 0) {
        HIPCHK(c, hipMemcpy(d.B.pos, pos.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d.B.v, v.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d.B.w, w.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d.B.q, q.data(), n * sizeof(Q4<T>), hipMemcpyHostToDevice));
        std::vector<int> awake(n);
        for (int i = 0; i < n; i++) awake[i] = in[i].sleeping ? 0 : 1;
        HIPCHK(c, hipMemcpy(d.B.wake, awake.data(), n * sizeof(int), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_refresh_bodies<T>, dim3((n + 255) / 256), dim3(256), 0, c->stream, d);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return CC_OK;
}

extern "C" int cc_set_bodies(cc_ctx* c, const cc_body_state* in, int32_t n) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (n > 0 && !in) return fail(c, CC_E_ARG, "null input");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    if (c->cfg.precision == CC_PRECISION_FP64) return set_bodies_impl<double>(c, in, n);
    return set_bodies_impl<float>(c, in, n);
}

// wake the sleeping dynamic bodies flagged in B.wake (sleep.ts:188-203), between steps
template <typename T>
__global__ void k_wake_requested(Dev<T> d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    if (d.B.wake[i] && d.B.sleeping[i] && fl_motion(d.B.flags[i]) == MOTION_DYNAMIC) {
        d.B.sleeping[i] = 0;
        reset_sleep_timer(d.B, d.st, i);
    }
    d.B.wake[i] = 0;
}

template <typename T>
static int add_forces_impl(cc_ctx* c, const double* f, const double* t, int n, int wake) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    if (n != d.N) return fail(c, CC_E_ARG, "num_bodies mismatch");
    if (n == 0 || (!f && !t)) return CC_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<V3<T>> F(n), Tq(n);
    std::vector<int> fl(n), wk(n, 0);
    HIPCHK(c, hipMemcpy(F.data(), d.B.force, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(Tq.data(), d.B.torque, n * sizeof(V3<T>), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(fl.data(), d.B.flags, n * sizeof(int), hipMemcpyDeviceToHost));
    bool anyWake = false;
    for (int i = 0; i < n; i++) {
        if (fl_motion(fl[i]) != MOTION_DYNAMIC) continue;  // addForce / addTorque: dynamic bodies only
        bool nz = false;
        if (f) {
            F[i] = F[i] + mk3<T>(T(f[3 * i]), T(f[3 * i + 1]), T(f[3 * i + 2]));
            nz = nz || f[3 * i] != 0.0 || f[3 * i + 1] != 0.0 || f[3 * i + 2] != 0.0;
        }
        if (t) {
            Tq[i] = Tq[i] + mk3<T>(T(t[3 * i]), T(t[3 * i + 1]), T(t[3 * i + 2]));
            nz = nz || t[3 * i] != 0.0 || t[3 * i + 1] != 0.0 || t[3 * i + 2] != 0.0;
        }
        if (nz) d.hasForces = 1;
        if (nz && wake) { wk[i] = 1; anyWake = true; }
    }
    HIPCHK(c, hipMemcpy(d.B.force, F.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.B.torque, Tq.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
    if (anyWake) {
        HIPCHK(c, hipMemcpy(d.B.wake, wk.data(), n * sizeof(int), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_wake_requested<T>, dim3((n + 255) / 256), dim3(256), 0, c->stream, d);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return CC_OK;
}

extern "C" int cc_add_forces(cc_ctx* c, const double* force, const double* torque, int32_t n, int32_t wake) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (n < 0) return fail(c, CC_E_ARG, "negative body count");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    if (c->cfg.precision == CC_PRECISION_FP64) return add_forces_impl<double>(c, force, torque, n, wake);
    return add_forces_impl<float>(c, force, torque, n, wake);
}

// ---------------------------------------------------------------- snapshot / restore
// Blob layout: SnapHeader, then the device arrays of snap_segments in their order, then the
// last step's contact-cache records.  Only a context holding the same upload reads it back.
struct SnapHeader {
    char magic[8];  // "CCSNAP02"
    int32_t version, tsize, n, npc, w, ncache, parity, hasForces;
    double prevDt;
    int64_t bytes;  // the whole blob
};

template <typename T>
static std::vector<std::pair<void*, size_t>> snap_segments(Dev<T>& d) {
    const size_t n = (size_t)d.N;
    const Bodies<T>& B = d.B;
    return {{B.pos, n * sizeof(V3<T>)},     {B.q, n * sizeof(Q4<T>)},         {B.com, n * sizeof(V3<T>)},
            {B.v, n * sizeof(V3<T>)},       {B.w, n * sizeof(V3<T>)},         {B.amin, n * sizeof(V3<T>)},
            {B.amax, n * sizeof(V3<T>)},    {B.sleeping, n * sizeof(int)},    {B.sc, 3 * n * sizeof(V3<T>)},
            {B.sr, 3 * n * sizeof(T)},      {B.stimer, n * sizeof(T)},        {B.label, n * sizeof(int)},
            {B.wake, n * sizeof(int)},      {B.flags, n * sizeof(int)},       {d.vb, n * sizeof(VBody<T>)},
            {d.pb, n * sizeof(PBody<T>)},   {d.pc, (size_t)d.NPC * sizeof(PointCon<T>)},
            {B.force, n * sizeof(V3<T>)},   {B.torque, n * sizeof(V3<T>)},
            // the constraint parts applyCCD carries from one call to the next
            {d.ccdPart, d.hasCcd ? 2 * (size_t)d.W * sizeof(CcdPart<T>) : 0}};
}

template <typename T>
static int64_t snap_bytes(Dev<T>& d, int ncache) {
    int64_t total = sizeof(SnapHeader);
    for (auto& sg : snap_segments(d)) total += (int64_t)sg.second;
    return total + (int64_t)ncache * (int64_t)sizeof(CacheRec<T>);
}

template <typename T>
static int snapshot_impl(cc_ctx* c, void* out, int64_t cap, int64_t* size) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int par = 0, rc[2] = {0, 0};
    HIPCHK(c, hipMemcpy(&par, d.parity, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(rc, d.recCnt, sizeof(rc), hipMemcpyDeviceToHost));
    const int q = par ^ 1;  // the parity the last step wrote, which the next step reads
    const int nr = std::min(rc[q], d.conCap);
    const int64_t total = snap_bytes(d, nr);
    *size = total;
    if (!out || cap < total) return CC_OK;
    SnapHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, "CCSNAP02", 8);
    h.version = 2;
    h.tsize = (int32_t)sizeof(T);
    h.n = d.N; h.npc = d.NPC; h.w = d.W;
    h.ncache = nr;
    h.parity = par;
    h.hasForces = d.hasForces;
    T pdt;
    HIPCHK(c, hipMemcpy(&pdt, d.prevDt, sizeof(T), hipMemcpyDeviceToHost));
    h.prevDt = (double)pdt;
    h.bytes = total;
    char* p = static_cast<char*>(out);
    memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (auto& sg : snap_segments(d)) {
        if (sg.second) HIPCHK(c, hipMemcpy(p, sg.first, sg.second, hipMemcpyDeviceToHost));
        p += sg.second;
    }
    if (nr > 0) HIPCHK(c, hipMemcpy(p, d.rec[q], (size_t)nr * sizeof(CacheRec<T>), hipMemcpyDeviceToHost));
    return CC_OK;
}

template <typename T>
static int restore_impl(cc_ctx* c, const void* in, int64_t size) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    SnapHeader h;
    if (size < (int64_t)sizeof(h)) return fail(c, CC_E_ARG, "snapshot too short");
    memcpy(&h, in, sizeof(h));
    if (memcmp(h.magic, "CCSNAP02", 8) != 0 || h.version != 2) return fail(c, CC_E_ARG, "not a crashcat_amd snapshot");
    if (h.tsize != (int32_t)sizeof(T) || h.n != d.N || h.npc != d.NPC || h.w != d.W)
        return fail(c, CC_E_ARG, "snapshot of another scene or precision");
    if (h.ncache < 0 || h.ncache > d.conCap || (h.parity != 0 && h.parity != 1))
        return fail(c, CC_E_ARG, "corrupt snapshot header");
    const int64_t total = snap_bytes(d, h.ncache);
    if (h.bytes != total || size != total) return fail(c, CC_E_ARG, "snapshot size mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const char* p = static_cast<const char*>(in) + sizeof(h);
    for (auto& sg : snap_segments(d)) {
        if (sg.second) HIPCHK(c, hipMemcpy(sg.first, p, sg.second, hipMemcpyHostToDevice));
        p += sg.second;
    }
    const int q = h.parity ^ 1;
    const int n = h.ncache;
    if (n > 0) {
        HIPCHK(c, hipMemcpy(d.rec[q], p, (size_t)n * sizeof(CacheRec<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemset(d.recHit[q], 0, (size_t)n * sizeof(int)));
    }
    int rc[2] = {0, 0};
    rc[q] = n;
    HIPCHK(c, hipMemcpy(d.recCnt, rc, sizeof(rc), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.parity, &h.parity, sizeof(int), hipMemcpyHostToDevice));
    d.hasForces = d.hasForces || h.hasForces;
    const T pdt = T(h.prevDt);
    HIPCHK(c, hipMemcpy(d.prevDt, &pdt, sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(d.counters, 0, CNT_N * sizeof(int)));
    HIPCHK(c, hipMemset(d.hkeys[q], 0xff, (size_t)d.hashCap * sizeof(unsigned long long)));
    if (n > 0) hipLaunchKernelGGL(k_hash_rebuild<T>, dim3(grid_for(n, 256, 2048)), dim3(256), 0, c->stream, d, q, n);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CC_OK;
}

// ---------------------------------------------------------------- warm state in plain records
static bool cached_less(const cc_cached_contact& x, const cc_cached_contact& y) {
    return std::tie(x.body_a, x.body_b, x.sub_shape_a, x.sub_shape_b) <
           std::tie(y.body_a, y.body_b, y.sub_shape_a, y.sub_shape_b);
}

template <typename T>
static int set_cache_impl(cc_ctx* c, const cc_cached_contact* in, int n, double prevDt) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    if (n < 0 || (n > 0 && !in)) return fail(c, CC_E_ARG, "contact cache: bad count or null records");
    if (n > d.conCap) return fail(c, CC_E_CAPACITY, "contact cache: more records than the cache holds");
    std::vector<cc_cached_contact> v(in, in + n);
    for (const auto& r : v)
        if (r.body_a < 0 || r.body_b >= d.N || r.body_a >= r.body_b || r.num_points < 1 || r.num_points > 4)
            return fail(c, CC_E_ARG, "contact cache: record with bodies out of range or unordered, or 0 or >4 points");
    std::sort(v.begin(), v.end(), cached_less);
    for (size_t k = 1; k < v.size(); k++)
        if (!cached_less(v[k - 1], v[k])) return fail(c, CC_E_ARG, "contact cache: two records with the same key");
    // records of one body pair form a run; its first record holds the run length (the hash
    // points at it and the lookup scans the run for the sub-shape ids)
    std::vector<CacheRec<T>> rec(v.size());
    for (size_t k = 0; k < v.size(); k++) {
        const auto& r = v[k];
        CacheRec<T> o;
        memset(&o, 0, sizeof(o));
        o.a = r.body_a;
        o.b = r.body_b;
        o.np = r.num_points;
        o.subA = r.sub_shape_a;
        o.subB = r.sub_shape_b;
        for (int j = 0; j < 4; j++) {
            o.p1[j] = mk3<T>(T(r.local_a[j][0]), T(r.local_a[j][1]), T(r.local_a[j][2]));
            o.p2[j] = mk3<T>(T(r.local_b[j][0]), T(r.local_b[j][1]), T(r.local_b[j][2]));
            o.ln[j] = T(r.lambda_n[j]);
            o.lt1[j] = T(r.lambda_t1[j]);
            o.lt2[j] = T(r.lambda_t2[j]);
        }
        if (k == 0 || v[k - 1].body_a != r.body_a || v[k - 1].body_b != r.body_b) {
            size_t e = k;
            while (e < v.size() && v[e].body_a == r.body_a && v[e].body_b == r.body_b) e++;
            o.run = (int)(e - k);
        }
        rec[k] = o;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int par = 0;
    HIPCHK(c, hipMemcpy(&par, d.parity, sizeof(int), hipMemcpyDeviceToHost));
    const int q = par ^ 1;  // the parity the next step reads
    if (n > 0) {
        HIPCHK(c, hipMemcpy(d.rec[q], rec.data(), (size_t)n * sizeof(CacheRec<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemset(d.recHit[q], 0, (size_t)n * sizeof(int)));
    }
    int rc[2] = {0, 0};
    rc[q] = n;
    HIPCHK(c, hipMemcpy(d.recCnt, rc, sizeof(rc), hipMemcpyHostToDevice));
    const T pdt = T(prevDt);
    HIPCHK(c, hipMemcpy(d.prevDt, &pdt, sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(d.counters, 0, CNT_N * sizeof(int)));
    HIPCHK(c, hipMemset(d.hkeys[q], 0xff, (size_t)d.hashCap * sizeof(unsigned long long)));
    if (n > 0) hipLaunchKernelGGL(k_hash_rebuild<T>, dim3(grid_for(n, 256, 2048)), dim3(256), 0, c->stream, d, q, n);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CC_OK;
}

template <typename T>
static int get_cache_impl(cc_ctx* c, cc_cached_contact* out, int cap, int* count, double* prevDt) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int par = 0, rc[2] = {0, 0};
    HIPCHK(c, hipMemcpy(&par, d.parity, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(rc, d.recCnt, sizeof(rc), hipMemcpyDeviceToHost));
    const int q = par ^ 1;
    const int n = std::min(rc[q], d.conCap);
    *count = n;
    if (prevDt) {
        T pdt;
        HIPCHK(c, hipMemcpy(&pdt, d.prevDt, sizeof(T), hipMemcpyDeviceToHost));
        *prevDt = (double)pdt;
    }
    if (!out || cap <= 0) return CC_OK;
    std::vector<CacheRec<T>> rec(n);
    if (n > 0) HIPCHK(c, hipMemcpy(rec.data(), d.rec[q], (size_t)n * sizeof(CacheRec<T>), hipMemcpyDeviceToHost));
    std::vector<cc_cached_contact> v(n);
    for (int k = 0; k < n; k++) {
        const auto& r = rec[k];
        cc_cached_contact o;
        memset(&o, 0, sizeof(o));
        o.body_a = r.a;
        o.body_b = r.b;
        o.sub_shape_a = r.subA;
        o.sub_shape_b = r.subB;
        o.num_points = r.np;
        for (int j = 0; j < r.np && j < 4; j++) {
            o.local_a[j][0] = r.p1[j].x; o.local_a[j][1] = r.p1[j].y; o.local_a[j][2] = r.p1[j].z;
            o.local_b[j][0] = r.p2[j].x; o.local_b[j][1] = r.p2[j].y; o.local_b[j][2] = r.p2[j].z;
            o.lambda_n[j] = r.ln[j];
            o.lambda_t1[j] = r.lt1[j];
            o.lambda_t2[j] = r.lt2[j];
        }
        v[k] = o;
    }
    std::sort(v.begin(), v.end(), cached_less);
    for (int k = 0; k < n && k < cap; k++) out[k] = v[k];
    return CC_OK;
}

template <typename T>
static int set_sleep_impl(cc_ctx* c, const cc_body_sleep* in, int n) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    if (!in || n != d.N) return fail(c, CC_E_ARG, "sleep state: one record per body");
    std::vector<T> timer(n), radius(3 * (size_t)n);
    std::vector<V3<T>> center(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        timer[i] = T(in[i].timer);
        for (int k = 0; k < 3; k++) {
            center[3 * (size_t)i + k] = mk3<T>(T(in[i].center[k][0]), T(in[i].center[k][1]), T(in[i].center[k][2]));
            radius[3 * (size_t)i + k] = T(in[i].radius[k]);
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(d.B.stimer, timer.data(), n * sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.B.sc, center.data(), 3 * (size_t)n * sizeof(V3<T>), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.B.sr, radius.data(), 3 * (size_t)n * sizeof(T), hipMemcpyHostToDevice));
    return CC_OK;
}

template <typename T>
static int set_point_impulses_impl(cc_ctx* c, const double* lam, int n) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    if (n != d.NPC || (n > 0 && !lam)) return fail(c, CC_E_ARG, "point impulses: three per point constraint");
    if (n == 0) return CC_OK;
    std::vector<PointCon<T>> pc(n);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(pc.data(), d.pc, n * sizeof(PointCon<T>), hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++) pc[k].lambda = mk3<T>(T(lam[3 * k]), T(lam[3 * k + 1]), T(lam[3 * k + 2]));
    HIPCHK(c, hipMemcpy(d.pc, pc.data(), n * sizeof(PointCon<T>), hipMemcpyHostToDevice));
    return CC_OK;
}

#define CC_DISPATCH(call) (c->cfg.precision == CC_PRECISION_FP64 ? call<double> : call<float>)

extern "C" int cc_set_contact_cache(cc_ctx* c, const cc_cached_contact* in, int32_t n, double prevDt) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    return CC_DISPATCH(set_cache_impl)(c, in, n, prevDt);
}
extern "C" int cc_download_contact_cache(cc_ctx* c, cc_cached_contact* out, int32_t cap, int32_t* count,
                                         double* prevDt) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    return CC_DISPATCH(get_cache_impl)(c, out, cap, count, prevDt);
}
extern "C" int cc_set_sleep_state(cc_ctx* c, const cc_body_sleep* in, int32_t n) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    return CC_DISPATCH(set_sleep_impl)(c, in, n);
}
extern "C" int cc_set_point_impulses(cc_ctx* c, const double* lam, int32_t n) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    return CC_DISPATCH(set_point_impulses_impl)(c, lam, n);
}
#undef CC_DISPATCH

extern "C" int cc_snapshot(cc_ctx* c, void* out, int64_t cap, int64_t* size) {
    if (!c || !c->impl || !size) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return snapshot_impl<double>(c, out, cap, size);
    return snapshot_impl<float>(c, out, cap, size);
}

extern "C" int cc_restore(cc_ctx* c, const void* in, int64_t size) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (!in) return fail(c, CC_E_ARG, "null snapshot");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->impl->state_changed();
    if (c->cfg.precision == CC_PRECISION_FP64) return restore_impl<double>(c, in, size);
    return restore_impl<float>(c, in, size);
}

template <typename T>
static int pairs_impl(cc_ctx* c, int32_t* out, int cap, int* count) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int cnt[CNT_N];
    HIPCHK(c, hipMemcpy(cnt, d.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    int n = std::min(cnt[CNT_PAIRS], d.pairCap);
    *count = n;
    if (!out || cap <= 0) return CC_OK;
    std::vector<int> a(n), b(n);
    if (n > 0) {
        HIPCHK(c, hipMemcpy(a.data(), d.pairA, n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(b.data(), d.pairB, n * sizeof(int), hipMemcpyDeviceToHost));
    }
    std::vector<std::pair<int, int>> p(n);
    for (int i = 0; i < n; i++) p[i] = {std::min(a[i], b[i]), std::max(a[i], b[i])};
    std::sort(p.begin(), p.end());
    for (int i = 0; i < n && i < cap; i++) { out[2 * i] = p[i].first; out[2 * i + 1] = p[i].second; }
    return CC_OK;
}

extern "C" int cc_download_pairs(cc_ctx* c, int32_t* out, int32_t cap, int32_t* count) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return pairs_impl<double>(c, out, cap, count);
    return pairs_impl<float>(c, out, cap, count);
}

template <typename T>
static int contacts_impl(cc_ctx* c, cc_contact_out* out, int cap, int* count) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int cnt[CNT_N];
    HIPCHK(c, hipMemcpy(cnt, d.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    int n = std::min(cnt[CNT_CONS], d.conCap);
    *count = n;
    if (!out || cap <= 0) return CC_OK;
    std::vector<int> idx(n);
    for (int k = 0; k < n; k++) idx[k] = k;
    std::vector<cc_contact_out> res;
    res.reserve(n);
    std::vector<Con<T>> hcs(n);
    std::vector<ConTap<T>> hts(n);
    std::vector<CacheRec<T>> hrs(n);
    std::vector<ConHead> hhs(n);
    if (n > 0) {
        int par = 0;
        HIPCHK(c, hipMemcpy(&par, d.parity, sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(hcs.data(), d.con, n * sizeof(Con<T>), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(hhs.data(), d.conHead, n * sizeof(ConHead), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(hts.data(), d.tap, n * sizeof(ConTap<T>), hipMemcpyDeviceToHost));
        // the last step wrote its impulses to the records of the parity it ran with
        HIPCHK(c, hipMemcpy(hrs.data(), d.rec[par ^ 1], n * sizeof(CacheRec<T>), hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < n; k++) {
        const Con<T>& h = hcs[idx[k]];
        const ConTap<T>& t = hts[idx[k]];
        const CacheRec<T>& r = hrs[idx[k]];
        cc_contact_out o;
        memset(&o, 0, sizeof(o));
        o.body_a = h.a; o.body_b = h.b; o.num_points = h.np;
        o.persisted = (hhs[idx[k]].bits >> 13) & 1;
        o.sub_shape_a = r.subA; o.sub_shape_b = r.subB;
        o.normal[0] = h.normal.x; o.normal[1] = h.normal.y; o.normal[2] = h.normal.z;
        for (int j = 0; j < h.np && j < 4; j++) {
            o.position_a[j][0] = t.pa[j].x; o.position_a[j][1] = t.pa[j].y; o.position_a[j][2] = t.pa[j].z;
            o.position_b[j][0] = t.pb[j].x; o.position_b[j][1] = t.pb[j].y; o.position_b[j][2] = t.pb[j].z;
            o.lambda_n[j] = r.ln[j]; o.lambda_t1[j] = r.lt1[j]; o.lambda_t2[j] = r.lt2[j];
        }
        res.push_back(o);
    }
    std::sort(res.begin(), res.end(), [](const cc_contact_out& x, const cc_contact_out& y) {
        if (x.body_a != y.body_a) return x.body_a < y.body_a;
        if (x.body_b != y.body_b) return x.body_b < y.body_b;
        if (x.sub_shape_a != y.sub_shape_a) return x.sub_shape_a < y.sub_shape_a;
        return x.sub_shape_b < y.sub_shape_b;
    });
    for (int k = 0; k < n && k < cap; k++) out[k] = res[k];
    return CC_OK;
}

template <typename T>
static int removed_impl(cc_ctx* c, cc_contact_removed* out, int cap, int* count) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int q = 0, rc[2] = {0, 0};
    HIPCHK(c, hipMemcpy(&q, d.parity, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(rc, d.recCnt, sizeof(rc), hipMemcpyDeviceToHost));
    // the step before the last wrote parity q (the last step wrote q ^ 1 and flipped to q)
    const int n = std::min(rc[q], d.conCap);
    *count = 0;
    if (n <= 0) return CC_OK;
    int4* buf = nullptr;
    int* cnt = nullptr;
    HIPCHK(c, hipMalloc(&buf, (size_t)n * sizeof(int4)));
    if (hipMalloc(&cnt, sizeof(int)) != hipSuccess) { (void)hipFree(buf); return fail(c, CC_E_HIP, "hipMalloc failed"); }
    int m = 0;
    std::vector<int4> h;
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof(int), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_removed_contacts<T>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->stream, d, q, buf, n,
                           cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&m, cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && m > 0) {
        h.resize(m);
        e = hipMemcpy(h.data(), buf, (size_t)m * sizeof(int4), hipMemcpyDeviceToHost);
    }
    (void)hipFree(buf);
    (void)hipFree(cnt);
    if (e != hipSuccess) return fail(c, CC_E_HIP, hipGetErrorString(e));
    std::sort(h.begin(), h.end(), [](const int4& x, const int4& y) {
        if (x.x != y.x) return x.x < y.x;
        if (x.y != y.y) return x.y < y.y;
        if ((unsigned)x.z != (unsigned)y.z) return (unsigned)x.z < (unsigned)y.z;
        return (unsigned)x.w < (unsigned)y.w;
    });
    *count = m;
    for (int i = 0; i < m && i < cap && out; i++) {
        out[i].body_a = h[i].x; out[i].body_b = h[i].y;
        out[i].sub_shape_a = (uint32_t)h[i].z; out[i].sub_shape_b = (uint32_t)h[i].w;
    }
    return CC_OK;
}

extern "C" int cc_download_removed_contacts(cc_ctx* c, cc_contact_removed* out, int32_t cap, int32_t* count) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    if (cap < 0) return fail(c, CC_E_ARG, "negative capacity");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return removed_impl<double>(c, out, cap, count);
    return removed_impl<float>(c, out, cap, count);
}

template <typename T>
static int ccd_bodies_impl(cc_ctx* c, cc_ccd_body* out, int cap, int* count) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    *count = 0;
    if (!d.hasCcd) return CC_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int nw = d.W;
    std::vector<int> start(nw + 1), cnt(nw);
    HIPCHK(c, hipMemcpy(start.data(), d.ccdStart, (nw + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(cnt.data(), d.ccdCount, nw * sizeof(int), hipMemcpyDeviceToHost));
    const int NC = start[nw];
    std::vector<CcdRec<T>> rec(NC);
    std::vector<int> ord(NC);
    HIPCHK(c, hipMemcpy(rec.data(), d.ccd, NC * sizeof(CcdRec<T>), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ord.data(), d.ccdOrd, NC * sizeof(int), hipMemcpyDeviceToHost));
    int k = 0;
    for (int w = 0; w < nw; w++)
        for (int i = 0; i < cnt[w]; i++, k++) {
            if (k >= cap || !out) continue;
            const CcdRec<T>& r = rec[start[w] + ord[start[w] + i]];
            cc_ccd_body& o = out[k];
            o.body_index = r.body;
            o.hit_body_index = r.hit;
            const V3<T>* v[3] = {&r.delta, &r.normal, &r.point};
            double* dst[3] = {o.delta_position, o.contact_normal, o.contact_point};
            for (int a = 0; a < 3; a++) { dst[a][0] = v[a]->x; dst[a][1] = v[a]->y; dst[a][2] = v[a]->z; }
            o.fraction = r.fraction;
            o.fraction_plus_slop = r.fps;
            o.linear_cast_threshold_sq = r.thrSq;
            o.max_penetration = r.maxPen;
            o.combined_friction = r.friction;
            o.combined_restitution = r.restitution;
            o.sub_shape_id_1 = r.hit >= 0 ? r.sub1 : 0u;
            o.sub_shape_id_2 = r.hit >= 0 ? r.sub2 : 0u;
        }
    *count = k;
    return CC_OK;
}

extern "C" int cc_download_ccd_bodies(cc_ctx* c, cc_ccd_body* out, int32_t cap, int32_t* count) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    if (cap < 0) return fail(c, CC_E_ARG, "negative capacity");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return ccd_bodies_impl<double>(c, out, cap, count);
    return ccd_bodies_impl<float>(c, out, cap, count);
}

// The removed contacts staged in the device's scratch array instead of per-call allocations; the
// host sorts them in the caller's buffer.
template <typename T>
static int removed_noalloc_impl(cc_ctx* c, cc_contact_removed* out, int cap, int* count) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    char* base = static_cast<char*>(g_scratch[c->cfg.device]);
    if (!base) return fail(c, CC_E_STATE, "scratch memory of this device is not resolved");
    int4* buf = reinterpret_cast<int4*>(base + 256);
    int* cnt = reinterpret_cast<int*>(base);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int q = 0, rc[2] = {0, 0};
    HIPCHK(c, hipMemcpy(&q, d.parity, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(rc, d.recCnt, sizeof(rc), hipMemcpyDeviceToHost));
    const int n = std::min(rc[q], d.conCap);
    *count = 0;
    if (n <= 0) return CC_OK;
    if (n > CC_SCRATCH_RECORDS)
        return fail(c, CC_E_CAPACITY, "the previous step kept " + std::to_string(n) + " contacts; the allocation-free " +
                                          "download stages at most " + std::to_string(CC_SCRATCH_RECORDS));
    // the counter starts at zero on every call (a fill kernel, so that the call can be captured)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, cnt, 0, 1);
    hipLaunchKernelGGL(k_removed_contacts<T>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->stream, d, q, buf, n, cnt);
    HIPCHK(c, hipGetLastError());
    int m = 0;
    HIPCHK(c, hipMemcpyAsync(&m, cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = m;
    if (!out || cap <= 0 || m == 0) return CC_OK;
    if (cap < m)
        return fail(c, CC_E_ARG, "capacity " + std::to_string(cap) + " is below the " + std::to_string(m) +
                                     " removed contacts (query the count with capacity 0)");
    static_assert(sizeof(cc_contact_removed) == sizeof(int4), "cc_contact_removed is four 32-bit fields");
    HIPCHK(c, hipMemcpy(out, buf, (size_t)m * sizeof(int4), hipMemcpyDeviceToHost));
    std::sort(out, out + m, [](const cc_contact_removed& x, const cc_contact_removed& y) {
        if (x.body_a != y.body_a) return x.body_a < y.body_a;
        if (x.body_b != y.body_b) return x.body_b < y.body_b;
        if (x.sub_shape_a != y.sub_shape_a) return x.sub_shape_a < y.sub_shape_a;
        return x.sub_shape_b < y.sub_shape_b;
    });
    return CC_OK;
}

extern "C" int cc_download_removed_contacts_noalloc(cc_ctx* c, cc_contact_removed* out, int32_t cap, int32_t* count) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    if (cap < 0) return fail(c, CC_E_ARG, "negative capacity");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return removed_noalloc_impl<double>(c, out, cap, count);
    return removed_noalloc_impl<float>(c, out, cap, count);
}

extern "C" int cc_download_contacts(cc_ctx* c, cc_contact_out* out, int32_t cap, int32_t* count) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return contacts_impl<double>(c, out, cap, count);
    return contacts_impl<float>(c, out, cap, count);
}

template <typename T>
static int counters_impl(cc_ctx* c, cc_step_counters* o) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int cnt[CNT_N];
    HIPCHK(c, hipMemcpy(cnt, d.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    memset(o, 0, sizeof(*o));
    o->num_pairs = std::min(cnt[CNT_PAIRS], d.pairCap);
    o->num_contacts = std::min(cnt[CNT_CONS], d.conCap);
    o->num_epa = cnt[CNT_EPA];
    o->overflow = cnt[CNT_OVERFLOW];
    int n = d.N;
    std::vector<int> fl(n), sl(n), lab(n);
    if (n > 0) {
        HIPCHK(c, hipMemcpy(fl.data(), d.B.flags, n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(sl.data(), d.B.sleeping, n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(lab.data(), d.B.label, n * sizeof(int), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < n; i++) {
        if (fl_motion(fl[i]) == MOTION_DYNAMIC) {
            o->num_bodies++;
            if (!sl[i]) o->num_awake++;
        }
        if (lab[i] == i) o->num_islands++;
    }
    std::vector<int> st(2 * d.W);
    HIPCHK(c, hipMemcpy(st.data(), d.stats, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int w = 0; w < d.W; w++) {
        o->max_levels = std::max<int64_t>(o->max_levels, st[2 * w]);
        o->position_sweeps = std::max<int64_t>(o->position_sweeps, st[2 * w + 1]);
    }
    int nc = (int)o->num_contacts;
    if (nc > 0) {
        std::vector<Con<T>> hc(nc);
        HIPCHK(c, hipMemcpy(hc.data(), d.con, nc * sizeof(Con<T>), hipMemcpyDeviceToHost));
        for (int k = 0; k < nc; k++) o->num_points += hc[k].np;
    }
    return CC_OK;
}

extern "C" int cc_get_counters(cc_ctx* c, cc_step_counters* o) {
    if (!c || !c->impl || !o) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return counters_impl<double>(c, o);
    return counters_impl<float>(c, o);
}

template <typename T>
static int phase_impl(cc_ctx* c, double* out, int cap, int* count) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double sum[CC_PHASES] = {0};
    for (int s = 0; s < c->nPhase; s++) {
        hipEvent_t* E = X->ev.data() + (size_t)(CC_PHASES + 1) * s;
        for (int p = 0; p < CC_PHASES; p++) {
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, E[p], E[p + 1]));
            sum[p] += ms;
        }
    }
    *count = c->nPhase > 0 ? CC_PHASES : 0;
    for (int p = 0; p < *count && p < cap; p++) out[p] = sum[p];
    return CC_OK;
}

extern "C" int cc_get_phase_times(cc_ctx* c, double* out, int32_t cap, int32_t* count) {
    if (!c || !c->impl || !count) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return phase_impl<double>(c, out, cap, count);
    return phase_impl<float>(c, out, cap, count);
}

template <typename T>
static int totals_impl(cc_ctx* c, cc_totals* o) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long a[ACC_N];
    HIPCHK(c, hipMemcpy(a, X->d.acc, sizeof(a), hipMemcpyDeviceToHost));
    memset(o, 0, sizeof(*o));
    o->steps = (int64_t)a[ACC_STEPS];
    o->pairs = (int64_t)a[ACC_PAIRS];
    o->contacts = (int64_t)a[ACC_CONS];
    o->points = (int64_t)a[ACC_POINTS];
    o->epa = (int64_t)a[ACC_EPA];
    o->awake_body_steps = (int64_t)a[ACC_AWAKE];
    o->solver_bytes = (int64_t)a[ACC_BYTES_SOLVE];
    o->gjk = (int64_t)a[ACC_GJK];
    o->body_steps = (int64_t)a[ACC_STEPS] * X->d.N;
    return CC_OK;
}

template <typename T>
static int residual_impl(cc_ctx* c, double* out) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    if (!X->d.resid) return fail(c, CC_E_STATE, "the context does not measure residuals (cc_config.measure_residual)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, X->d.resid, 4 * sizeof(double), hipMemcpyDeviceToHost));
    return CC_OK;
}

extern "C" int cc_get_solver_residual(cc_ctx* c, double out[4]) {
    if (!c || !c->impl || !out) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return residual_impl<double>(c, out);
    return residual_impl<float>(c, out);
}

extern "C" int cc_get_totals(cc_ctx* c, cc_totals* o) {
    if (!c || !c->impl || !o) return fail(c, CC_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return totals_impl<double>(c, o);
    return totals_impl<float>(c, o);
}

// Broadphase-only tap on caller-supplied AABBs (one world).
template <typename T>
static int bp_impl(cc_ctx* c, const float* mn, const float* mx, const int32_t* mo, const int32_t* sl, const int32_t* la,
                   int n, double margin, int32_t* out, int cap, int32_t* count) {
    if (n == 0) {
        *count = 0;
        return CC_OK;
    }
    // a one-world scene of unit spheres whose bounds are overwritten with the given AABBs
    cc_settings s;
    memset(&s, 0, sizeof(s));
    for (int a = 0; a < CC_MAX_OBJECT_LAYERS; a++)
        for (int b = 0; b < CC_MAX_OBJECT_LAYERS; b++) s.object_layer_pairs[a * CC_MAX_OBJECT_LAYERS + b] = 1;
    s.object_layer_pairs[1 * CC_MAX_OBJECT_LAYERS + 1] = 0;
    s.speculative_contact_distance = margin;
    cc_shape_desc sh;
    memset(&sh, 0, sizeof(sh));
    sh.type = CC_SHAPE_SPHERE;
    sh.radius = 1;
    std::vector<cc_body_desc> bd(std::max(n, 1));
    for (int i = 0; i < n; i++) {
        memset(&bd[i], 0, sizeof(cc_body_desc));
        bd[i].motion_type = mo[i];
        bd[i].sleeping = sl[i];
        bd[i].object_layer = la[i];
        bd[i].collision_groups = bd[i].collision_mask = 0xffffffffu;
        bd[i].quaternion[3] = 1;
        bd[i].allowed_dofs = 63;
    }
    int woff[2] = {0, n};
    int r = upload<T>(c, &s, 1, woff, bd.data(), n, &sh, 1, nullptr, nullptr, 0, nullptr);
    if (r) return r;
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    std::vector<V3<T>> a(n), b(n);
    for (int i = 0; i < n; i++) {
        a[i] = mk3<T>(T(mn[3 * i]), T(mn[3 * i + 1]), T(mn[3 * i + 2]));
        b[i] = mk3<T>(T(mx[3 * i]), T(mx[3 * i + 1]), T(mx[3 * i + 2]));
    }
    if (n > 0) {
        HIPCHK(c, hipMemcpy(d.B.amin, a.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d.B.amax, b.data(), n * sizeof(V3<T>), hipMemcpyHostToDevice));
    }
    // cell width from the given boxes: the widest, capped at four times the median width
    std::vector<double> wd(n);
    for (int i = 0; i < n; i++)
        wd[i] = std::max({(double)mx[3 * i] - mn[3 * i], (double)mx[3 * i + 1] - mn[3 * i + 1], (double)mx[3 * i + 2] - mn[3 * i + 2]});
    std::sort(wd.begin(), wd.end());
    double cw = std::min(wd.back(), 4.0 * wd[wd.size() / 2]);
    d.cellWidth = cw > 1e-6 ? cw : 1.0;
    set_cell_size(d, margin);
    launch_broadphase<T>(X, c->stream, margin);
    HIPCHK(c, hipGetLastError());
    return pairs_impl<T>(c, out, cap, count);
}

extern "C" int cc_broadphase_aabbs(cc_ctx* c, const float* mn, const float* mx, const int32_t* mo, const int32_t* sl,
                                   const int32_t* la, int32_t n, double margin, int32_t* out, int32_t cap, int32_t* count) {
    if (!c || !count || n < 0 || (n > 0 && (!mn || !mx || !mo || !sl || !la))) return fail(c, CC_E_ARG, "invalid arguments");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return bp_impl<double>(c, mn, mx, mo, sl, la, n, margin, out, cap, count);
    return bp_impl<float>(c, mn, mx, mo, sl, la, n, margin, out, cap, count);
}

// ---------------------------------------------------------------- differentiable stepping
// cc_fwd steps like cc_step and keeps a checkpoint of the state every K steps; cc_bwd walks the
// segments backwards: it restores a segment's checkpoint, steps through the segment again with
// the solver recording (bit for bit the same steps), and runs the adjoint of each recorded step
// from the last to the first (cc_adjoint.h).
template <typename T>
static std::vector<std::pair<void*, size_t>> ckpt_segments(Dev<T>& d) {
    auto v = snap_segments(d);
    for (int k = 0; k < 2; k++) {
        v.push_back({d.rec[k], (size_t)d.conCap * sizeof(CacheRec<T>)});
        v.push_back({d.recHit[k], (size_t)d.conCap * sizeof(int)});
        v.push_back({d.hkeys[k], (size_t)d.hashCap * sizeof(unsigned long long)});
        v.push_back({d.hvals[k], (size_t)d.hashCap * sizeof(int)});
    }
    v.push_back({d.recCnt, 2 * sizeof(int)});
    v.push_back({d.parity, sizeof(int)});
    v.push_back({d.prevDt, sizeof(T)});
    v.push_back({d.counters, CNT_N * sizeof(int)});
    return v;
}

template <typename T>
static hipError_t ckpt_copy(CtxT<T>* X, hipStream_t st, char* buf, bool save) {
    size_t off = 0;
    for (auto& sg : ckpt_segments(X->d)) {
        if (sg.second) {
            const hipError_t e = save ? hipMemcpyAsync(buf + off, sg.first, sg.second, hipMemcpyDeviceToDevice, st)
                                      : hipMemcpyAsync(sg.first, buf + off, sg.second, hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return e;
        }
        off += sg.second;
    }
    return hipSuccess;
}

template <typename T>
static int fwd_impl(cc_ctx* c, double dt, int n, int K) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    auto& D = X->diff;
    D.valid = false;
    if (X->anySplit)
        return fail(c, CC_E_UNSUPPORTED, "differentiable stepping solves each world on one workgroup: use "
                                         "island_groups = 1 or a batch of at least 64 worlds");
    if (d.NPC > 0) return fail(c, CC_E_UNSUPPORTED, "differentiable stepping has no point constraints");
    if (d.hasCcd) return fail(c, CC_E_UNSUPPORTED, "differentiable stepping has no LINEAR_CAST (CCD) bodies");
    if (d.colored) return fail(c, CC_E_UNSUPPORTED, "differentiable stepping solves in crashcat's order (CC_ORDER_REFERENCE)");
    if (d.islIt) return fail(c, CC_E_UNSUPPORTED, "differentiable stepping has no iteration overrides");
    if (d.s.allowSleeping)
        return fail(c, CC_E_UNSUPPORTED, "differentiable stepping needs sleeping off (allow_sleeping = 0): a body "
                                         "put to sleep has its velocities zeroed");
    X->solverBS = 256;  // the recording solver's block size (results do not depend on it)
    d.diff = 1;
    hipStream_t st = c->stream;
    const auto segs = ckpt_segments(d);
    size_t bytes = 0;
    for (auto& sg : segs) bytes += sg.second;
    const int nck = (n + K - 1) / K + 1;
    if (D.ckptBytes != bytes) {
        for (char* p : D.ckpt) (void)hipFree(p);
        D.ckpt.clear();
        D.ckptBytes = bytes;
    }
    while ((int)D.ckpt.size() < nck) {
        char* p = nullptr;
        HIPCHK(c, hipMalloc(&p, bytes));
        D.ckpt.push_back(p);
    }
    T pdt;
    HIPCHK(c, hipMemcpy(&pdt, d.prevDt, sizeof(T), hipMemcpyDeviceToHost));
    const T dtT = T(dt);
    D.ratio.assign(n, T(0));
    for (int s = 0; s < n; s++) {
        const T prev = s == 0 ? pdt : dtT;
        D.ratio[s] = (d.s.warmStarting && prev > T(0)) ? dtT / prev : T(0);
    }
    for (int s = 0; s < n; s++) {
        if (s % K == 0) HIPCHK(c, ckpt_copy(X, st, D.ckpt[s / K], true));
        const int rc = step_impl<T>(c, dt, 1);
        if (rc != CC_OK) return rc;
    }
    HIPCHK(c, ckpt_copy(X, st, D.ckpt[nck - 1], true));
    HIPCHK(c, hipStreamSynchronize(st));
    D.n = n;
    D.K = K;
    D.dt = dt;
    D.valid = true;
    return CC_OK;
}

template <typename T>
static int bwd_alloc(cc_ctx* c, CtxT<T>* X) {
    Dev<T>& d = X->d;
    auto& D = X->diff;
    if (D.slots >= D.K && D.xb) return CC_OK;
    D.mem.release();
    const size_t S = (size_t)D.K, N = (size_t)std::max(d.N, 1), sv = (size_t)d.svCap, pts = 4 * sv;
    const size_t W = (size_t)d.W;
    D.lvlInts = 2 * (sv + (size_t)d.V + 1);
    Alloc& A = D.mem;
    D.ve = A.get<VElem<T>>(S * sv);
    D.vp = A.get<VPoint<T>>(S * pts);
    D.la = A.get<V3<T>>(S * pts);
    D.lb = A.get<V3<T>>(S * pts);
    D.lvl = A.get<int>(S * D.lvlInts);
    D.lam = A.get<LamRec<T>>(S * pts * (size_t)(d.s.velIters + 1));
    D.src = A.get<int>(S * pts);
    D.rank = A.get<int>(S * sv);
    D.stats = A.get<int>(S * 2 * W);
    D.wcs = A.get<int>(S * (W + 1));
    D.vbody = A.get<VBody<T>>(S * N);
    D.npb = A.get<NpBody<T>>(S * N);
    D.xb = A.get<V3<T>>(N);
    D.qb = A.get<V3<T>>(N);
    D.vbar = A.get<V3<T>>(N);
    D.wbar = A.get<V3<T>>(N);
    D.mub = A.get<T>(N);
    D.lamBar = A.get<T>(3 * pts);
    D.bBar = A.get<T>(pts);
    D.rBar = A.get<V3<T>>(2 * pts);
    D.vf = A.get<V3<T>>(N);
    D.wf = A.get<V3<T>>(N);
    for (int k = 0; k < 2; k++) D.seed[k] = A.get<T>(3 * 4 * (size_t)d.conCap);
    if (A.err) {
        D.mem.release();
        D.xb = nullptr;
        return fail(c, CC_E_HIP, "hipMalloc failed (step records of the adjoint)");
    }
    D.slots = D.K;
    d.ptsCap = (int)pts;
    return CC_OK;
}

template <typename T>
static int bwd_impl(cc_ctx* c, const double* gcom, const double* gvel, double* gv0, double* gw0, double* gcom0,
                    double* gfr) {
    auto* X = static_cast<CtxT<T>*>(c->impl);
    Dev<T>& d = X->d;
    auto& D = X->diff;
    if (!D.valid) return fail(c, CC_E_STATE, "cc_bwd follows a cc_fwd with no step, body change or restore after it");
    const int rc = bwd_alloc(c, X);
    if (rc != CC_OK) return rc;
    hipStream_t st = c->stream;
    const int N = d.N, W = d.W, n = D.n, K = D.K;
    const size_t sv = (size_t)d.svCap, pts = 4 * sv;
    {  // the adjoints of the final state
        std::vector<V3<T>> x(N), v(N), z(N);
        for (int i = 0; i < N; i++) {
            x[i] = mk3<T>(T(gcom[3 * i]), T(gcom[3 * i + 1]), T(gcom[3 * i + 2]));
            v[i] = gvel ? mk3<T>(T(gvel[3 * i]), T(gvel[3 * i + 1]), T(gvel[3 * i + 2])) : mk3<T>(0, 0, 0);
            z[i] = mk3<T>(0, 0, 0);
        }
        HIPCHK(c, hipMemcpy(D.xb, x.data(), N * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(D.vbar, v.data(), N * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(D.qb, z.data(), N * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(D.wbar, z.data(), N * sizeof(V3<T>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemsetAsync(D.mub, 0, N * sizeof(T), st));
        for (int k = 0; k < 2; k++) HIPCHK(c, hipMemsetAsync(D.seed[k], 0, 3 * 4 * (size_t)d.conCap * sizeof(T), st));
    }
    const Dev<T> keep = d;  // the solver's own staging arrays, restored after each recorded step
    const int nseg = (n + K - 1) / K;
    int rcode = CC_OK;
    for (int seg = nseg - 1; seg >= 0 && rcode == CC_OK; seg--) {
        HIPCHK(c, ckpt_copy(X, st, D.ckpt[seg], false));
        const int s0 = seg * K, s1 = std::min(n, s0 + K);
        for (int s = s0; s < s1; s++) {  // the segment again, recording each step
            const size_t k = (size_t)(s - s0);
            d.ve = D.ve + k * sv;
            d.vp = D.vp + k * pts;
            d.svLA = D.la + k * pts;
            d.svLB = D.lb + k * pts;
            d.lvlStart = D.lvl + k * D.lvlInts;
            d.tapeLam = D.lam + k * pts * (size_t)(d.s.velIters + 1);
            d.tapeSrc = D.src + k * pts;
            d.tapeRank = D.rank + k * sv;
            d.ptsCap = (int)pts;
            rcode = step_impl<T>(c, D.dt, 1, true);
            d.ve = keep.ve; d.vp = keep.vp; d.svLA = keep.svLA; d.svLB = keep.svLB; d.lvlStart = keep.lvlStart;
            d.tapeLam = nullptr; d.tapeSrc = nullptr; d.tapeRank = nullptr;
            if (rcode != CC_OK) break;
            HIPCHK(c, hipMemcpyAsync(D.vbody + k * N, d.vb, N * sizeof(VBody<T>), hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(D.npb + k * N, d.npb, N * sizeof(NpBody<T>), hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(D.stats + k * 2 * W, d.stats, 2 * W * sizeof(int), hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(D.wcs + k * (W + 1), d.wConStart, (W + 1) * sizeof(int), hipMemcpyDeviceToDevice, st));
        }
        if (rcode != CC_OK) break;
        for (int s = s1 - 1; s >= s0; s--) {  // the adjoint of each recorded step, last first
            const size_t k = (size_t)(s - s0);
            Adj<T> a;
            a.xb = D.xb; a.qb = D.qb; a.vb = D.vbar; a.wb = D.wbar; a.mub = D.mub;
            a.lamBar = D.lamBar; a.bBar = D.bBar; a.rBar = D.rBar; a.vf = D.vf; a.wf = D.wf;
            a.seedIn = D.seed[s & 1];
            a.seedOut = D.seed[(s + 1) & 1];
            a.ve = D.ve + k * sv; a.vp = D.vp + k * pts; a.la = D.la + k * pts; a.lb = D.lb + k * pts;
            a.lam = D.lam + k * pts * (size_t)(d.s.velIters + 1);
            a.src = D.src + k * pts; a.rank = D.rank + k * sv; a.lvl = D.lvl + k * D.lvlInts;
            a.stats = D.stats + k * 2 * W; a.wcs = D.wcs + k * (W + 1);
            a.vbody = D.vbody + k * N; a.npb = D.npb + k * N;
            a.ptsCap = (int)pts;
            a.sweeps = d.s.velIters;
            a.dt = T(D.dt);
            a.ratio = D.ratio[s];
            hipLaunchKernelGGL((k_adj_world<T, 256>), dim3(d.V), dim3(256), 0, st, d, a);
        }
        HIPCHK(c, hipGetLastError());
    }
    // leave the context in the state cc_fwd left it in
    HIPCHK(c, ckpt_copy(X, st, D.ckpt[(n + K - 1) / K], false));
    HIPCHK(c, hipStreamSynchronize(st));
    if (rcode != CC_OK) return rcode;
    std::vector<V3<T>> h(N);
    auto out3 = [&](const V3<T>* src, double* dst) -> int {
        if (!dst) return CC_OK;
        HIPCHK(c, hipMemcpy(h.data(), src, N * sizeof(V3<T>), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) {
            dst[3 * i] = (double)h[i].x;
            dst[3 * i + 1] = (double)h[i].y;
            dst[3 * i + 2] = (double)h[i].z;
        }
        return CC_OK;
    };
    int r = out3(D.vbar, gv0);
    if (r == CC_OK) r = out3(D.wbar, gw0);
    if (r == CC_OK) r = out3(D.xb, gcom0);
    if (r != CC_OK) return r;
    if (gfr) {
        std::vector<T> m(N);
        HIPCHK(c, hipMemcpy(m.data(), D.mub, N * sizeof(T), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) gfr[i] = (double)m[i];
    }
    return CC_OK;
}

extern "C" int cc_fwd(cc_ctx* c, double dt, int32_t nsteps, int32_t checkpoint_every) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (nsteps < 1 || checkpoint_every < 1 || !(dt > 0)) return fail(c, CC_E_ARG, "invalid dt, nsteps or checkpoint interval");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return fwd_impl<double>(c, dt, nsteps, checkpoint_every);
    return fwd_impl<float>(c, dt, nsteps, checkpoint_every);
}

extern "C" int cc_bwd(cc_ctx* c, const double* dl_dcom, const double* dl_dvel, double* dl_dv0, double* dl_dw0,
                      double* dl_dcom0, double* dl_dfriction) {
    if (!c || !c->impl) return fail(c, CC_E_STATE, "no scene uploaded");
    if (!dl_dcom) return fail(c, CC_E_ARG, "dl_dcom is required");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cfg.precision == CC_PRECISION_FP64) return bwd_impl<double>(c, dl_dcom, dl_dvel, dl_dv0, dl_dw0, dl_dcom0, dl_dfriction);
    return bwd_impl<float>(c, dl_dcom, dl_dvel, dl_dv0, dl_dw0, dl_dcom0, dl_dfriction);
}
