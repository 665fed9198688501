// Host-side triangle-mesh builder: the data a TriangleMeshShape holds after buildTriangleMesh
// (src/shapes/utils/triangle-mesh-builder.ts:62-220) — welded vertices, non-degenerate unique
// triangles with active-edge bits — and its BVH from triangleMeshBvh.build with the CENTER
// split strategy (src/shapes/utils/triangle-mesh-bvh.ts:368-576, 924-966).  Runs once per mesh
// at scene creation, on the host; the step only reads its output.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace ccmesh {

struct Key3 {
    long long x, y, z;
    bool operator==(const Key3& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct Key3Hash {
    size_t operator()(const Key3& k) const {
        unsigned long long h = (unsigned long long)k.x * 0x9e3779b97f4a7c15ull;
        h ^= (unsigned long long)k.y + 0x7f4a7c159e3779b9ull + (h << 6) + (h >> 2);
        h ^= (unsigned long long)k.z + 0x94d049bb133111ebull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
struct Key4Hash {
    size_t operator()(const std::pair<Key3, long long>& k) const {
        return Key3Hash()(k.first) ^ ((size_t)k.second * 0xbf58476d1ce4e5b9ull);
    }
};

// Math.round: the nearest integer, halves towards +infinity
inline long long js_round(double x) {
    double f = std::floor(x);
    return (long long)(x - f >= 0.5 ? f + 1.0 : f);
}

struct Tri {
    int a, b, c;
    double n[3];
    int act;
    int mat;
};

struct Builder {
    std::vector<double> pos;        // welded vertices
    std::vector<Tri> tri;
    std::vector<double> bd;         // per triangle: cx, hx, cy, hy, cz, hz
    std::vector<double> nodes;      // 8 doubles per node, pre-order
    int maxLeaf = 8;

    // Hoare partition of [s, e) on the triangle centres (partitionTriangles :924-966)
    int partition(int s, int e, int axis, double split) {
        int l = s, r = e - 1;
        const int co = axis * 2;
        while (l <= r) {
            if (bd[6 * (size_t)l + co] < split) {
                l++;
            } else if (bd[6 * (size_t)r + co] >= split) {
                r--;
            } else {
                std::swap(tri[l], tri[r]);
                for (int k = 0; k < 6; k++) std::swap(bd[6 * (size_t)l + k], bd[6 * (size_t)r + k]);
                l++;
                r--;
            }
        }
        return l;
    }

    // buildRecursive :428-576 fused with populateBuffer :398-426 (pre-order; the left child
    // follows its parent, the right child follows the left subtree)
    void build(int s, int e) {
        const size_t at = nodes.size();
        nodes.resize(at + 8);
        const int count = e - s;
        double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
        double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (count > 0) {
            for (int i = s; i < e; i++) {
                const double* b = &bd[6 * (size_t)i];
                for (int k = 0; k < 3; k++) {
                    const double c = b[2 * k], h = b[2 * k + 1];
                    const double lo = c - h, hi = c + h;
                    if (i == s) {
                        mn[k] = lo; mx[k] = hi; cmn[k] = c; cmx[k] = c;
                    } else {
                        if (lo < mn[k]) mn[k] = lo;
                        if (hi > mx[k]) mx[k] = hi;
                        if (c < cmn[k]) cmn[k] = c;
                        if (c > cmx[k]) cmx[k] = c;
                    }
                }
            }
        }
        for (int k = 0; k < 3; k++) { nodes[at + k] = mn[k]; nodes[at + 3 + k] = mx[k]; }
        auto leaf = [&]() {
            nodes[at + 6] = (double)s;
            nodes[at + 7] = -(double)(count + 1);
        };
        if (count <= maxLeaf) { leaf(); return; }
        // CENTER split (getOptimalSplit :578-604): the longest centre extent, at its middle
        const double ex = cmx[0] - cmn[0], ey = cmx[1] - cmn[1], ez = cmx[2] - cmn[2];
        const double MIN_SIZE = 1e-5;
        if (ex < MIN_SIZE && ey < MIN_SIZE && ez < MIN_SIZE) { leaf(); return; }
        const int axis = ex >= ey && ex >= ez ? 0 : (ey >= ez ? 1 : 2);
        const double split = (cmn[axis] + cmx[axis]) / 2;
        const int mid = partition(s, e, axis, split);
        if (mid == s || mid == e) { leaf(); return; }
        build(s, mid);
        nodes[at + 6] = (double)nodes.size();
        nodes[at + 7] = (double)axis;
        build(mid, e);
    }
};

// returns 0 on success, -1 when an index is out of range
inline int build_triangle_mesh(const double* positions, int nv, const int32_t* indices, int nt, double tol,
                               double cosThr, int maxLeaf, Builder& B) {
    // weld on a 1e-8 grid (deduplicateVertices :222-254), first occurrence keeps its slot
    std::vector<int> remap((size_t)nv);
    {
        std::unordered_map<Key3, int, Key3Hash> seen;
        seen.reserve((size_t)nv * 2 + 1);
        for (int i = 0; i < nv; i++) {
            Key3 k{js_round(positions[3 * (size_t)i] * 1e8), js_round(positions[3 * (size_t)i + 1] * 1e8),
                   js_round(positions[3 * (size_t)i + 2] * 1e8)};
            auto it = seen.find(k);
            if (it == seen.end()) {
                const int id = (int)(B.pos.size() / 3);
                seen.emplace(k, id);
                B.pos.push_back(positions[3 * (size_t)i]);
                B.pos.push_back(positions[3 * (size_t)i + 1]);
                B.pos.push_back(positions[3 * (size_t)i + 2]);
                remap[(size_t)i] = id;
            } else {
                remap[(size_t)i] = it->second;
            }
        }
    }
    // non-degenerate triangles with their unit normals (:75-148)
    std::vector<Tri> valid;
    valid.reserve((size_t)nt);
    for (int i = 0; i < nt; i++) {
        int in[3];
        for (int k = 0; k < 3; k++) {
            const int v = indices[3 * (size_t)i + k];
            if (v < 0 || v >= nv) return -1;
            in[k] = remap[(size_t)v];
        }
        const double* a = &B.pos[3 * (size_t)in[0]];
        const double* b = &B.pos[3 * (size_t)in[1]];
        const double* c = &B.pos[3 * (size_t)in[2]];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double l2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
        if (std::sqrt(l2) < tol) continue;
        Tri t;
        t.a = in[0]; t.b = in[1]; t.c = in[2];
        const double il = l2 > 0 ? 1.0 / std::sqrt(l2) : 0.0;
        for (int k = 0; k < 3; k++) t.n[k] = l2 > 0 ? cr[k] * il : cr[k];
        t.act = 7;
        t.mat = -1;
        valid.push_back(t);
    }
    // duplicates by rotation-canonical index triple and material (:150-214)
    {
        struct PairEq {
            bool operator()(const std::pair<Key3, long long>& x, const std::pair<Key3, long long>& y) const {
                return x.first == y.first && x.second == y.second;
            }
        };
        std::unordered_set<std::pair<Key3, long long>, Key4Hash, PairEq> seen2;
        seen2.reserve(valid.size() * 2 + 1);
        for (const Tri& t : valid) {
            int x, y, z;
            if (t.a < t.b) {
                if (t.a < t.c) { x = t.a; y = t.b; z = t.c; }
                else { x = t.c; y = t.a; z = t.b; }
            } else {
                if (t.b < t.c) { x = t.b; y = t.c; z = t.a; }
                else { x = t.c; y = t.a; z = t.b; }
            }
            if (seen2.insert({Key3{x, y, z}, (long long)t.mat}).second) B.tri.push_back(t);
        }
    }
    // active edges by dihedral angle (computeActiveEdges :276-338, processEdge :340-370)
    if (cosThr >= 0.0) {
        struct Refs { int count; int tri0; int tri1; };
        std::unordered_map<unsigned long long, Refs> edges;
        edges.reserve(B.tri.size() * 3 + 1);
        auto process = [&](int ia, int ib, int ti, int bit) {
            const unsigned long long lo = (unsigned long long)(ia < ib ? ia : ib), hi = (unsigned long long)(ia < ib ? ib : ia);
            const unsigned long long key = lo * 67108864ull + hi;
            auto it = edges.find(key);
            if (it == edges.end()) edges.emplace(key, Refs{1, (ti << 3) | bit, 0});
            else if (it->second.count == 1) { it->second.tri1 = (ti << 3) | bit; it->second.count = 2; }
            else { B.tri[(size_t)ti].act |= bit; it->second.count++; }
        };
        for (size_t ti = 0; ti < B.tri.size(); ti++) {
            const Tri& t = B.tri[ti];
            process(t.a, t.b, (int)ti, 1);
            process(t.b, t.c, (int)ti, 2);
            process(t.c, t.a, (int)ti, 4);
        }
        for (const auto& kv : edges) {
            const Refs& r = kv.second;
            if (r.count != 2) continue;
            Tri& t0 = B.tri[(size_t)(r.tri0 >> 3)];
            Tri& t1 = B.tri[(size_t)(r.tri1 >> 3)];
            const double dot = t0.n[0] * t1.n[0] + t0.n[1] * t1.n[1] + t0.n[2] * t1.n[2];
            if (dot >= cosThr) {
                t0.act &= ~(r.tri0 & 7);
                t1.act &= ~(r.tri1 & 7);
            }
        }
    }
    // BVH (build :368-388, precomputeTriangleBuildData :272-342)
    B.maxLeaf = maxLeaf;
    const int n = (int)B.tri.size();
    if (n == 0) return 0;
    B.bd.resize(6 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const Tri& t = B.tri[(size_t)i];
        const double* a = &B.pos[3 * (size_t)t.a];
        const double* b = &B.pos[3 * (size_t)t.b];
        const double* c = &B.pos[3 * (size_t)t.c];
        for (int k = 0; k < 3; k++) {
            double lo = a[k], hi = a[k];
            if (b[k] < lo) lo = b[k];
            if (c[k] < lo) lo = c[k];
            if (b[k] > hi) hi = b[k];
            if (c[k] > hi) hi = c[k];
            B.bd[6 * (size_t)i + 2 * k] = (lo + hi) * 0.5;
            B.bd[6 * (size_t)i + 2 * k + 1] = (hi - lo) * 0.5;
        }
    }
    B.nodes.reserve(16 * (size_t)n);
    B.build(0, n);
    return 0;
}

}  // namespace ccmesh
