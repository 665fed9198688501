/* refcpu internals: FP64 vector math in gl-matrix conventions (column-major
 * matrices, quaternions x,y,z,w), as used by crashcat through its math
 * dependency mathcat@0.0.10 (package.json:50).  Oracle / test infrastructure. */
#ifndef REFCPU_INTERNAL_H
#define REFCPU_INTERNAL_H
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "refcpu.h"

typedef struct { double x, y, z; } V3;
typedef struct { double x, y, z, w; } Q4;
typedef struct { double m[9]; } M3; /* column-major 3x3: m[col*3+row] */

static inline V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vscale(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double vlen2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double vlen(V3 a) { return sqrt(vlen2(a)); }
static inline V3 vnormalize(V3 a) { /* vec3.normalize */
    double l = vlen2(a);
    if (l > 0) { l = 1.0 / sqrt(l); return vscale(a, l); }
    return a;
}
static inline V3 vload(const double* p) { return v3(p[0], p[1], p[2]); }
static inline void vstore(double* p, V3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* mat4.fromQuat (3x3 part) */
static inline M3 m3_from_quat(Q4 q) {
    double x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    double xx = q.x * x2, yx = q.y * x2, yy = q.y * y2, zx = q.z * x2, zy = q.z * y2, zz = q.z * z2;
    double wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    M3 r;
    r.m[0] = 1 - yy - zz; r.m[1] = yx + wz; r.m[2] = zx - wy;
    r.m[3] = yx - wz; r.m[4] = 1 - xx - zz; r.m[5] = zy + wx;
    r.m[6] = zx + wy; r.m[7] = zy - wx; r.m[8] = 1 - xx - yy;
    return r;
}
static inline V3 m3_mul(const M3* m, V3 v) {
    return v3(m->m[0] * v.x + m->m[3] * v.y + m->m[6] * v.z, m->m[1] * v.x + m->m[4] * v.y + m->m[7] * v.z,
              m->m[2] * v.x + m->m[5] * v.y + m->m[8] * v.z);
}
static inline V3 m3_mul_t(const M3* m, V3 v) { /* R^T v */
    return v3(m->m[0] * v.x + m->m[1] * v.y + m->m[2] * v.z, m->m[3] * v.x + m->m[4] * v.y + m->m[5] * v.z,
              m->m[6] * v.x + m->m[7] * v.y + m->m[8] * v.z);
}
/* vec3.transformQuat (gl-matrix form) */
static inline V3 qrotate(Q4 q, V3 a) {
    double uvx = q.y * a.z - q.z * a.y, uvy = q.z * a.x - q.x * a.z, uvz = q.x * a.y - q.y * a.x;
    double uuvx = q.y * uvz - q.z * uvy, uuvy = q.z * uvx - q.x * uvz, uuvz = q.x * uvy - q.y * uvx;
    double w2 = q.w * 2;
    return v3(a.x + uvx * w2 + uuvx * 2, a.y + uvy * w2 + uuvy * 2, a.z + uvz * w2 + uuvz * 2);
}
/* quat.multiply(a, b) */
static inline Q4 qmul(Q4 a, Q4 b) {
    Q4 r;
    r.x = a.x * b.w + a.w * b.x + a.y * b.z - a.z * b.y;
    r.y = a.y * b.w + a.w * b.y + a.z * b.x - a.x * b.z;
    r.z = a.z * b.w + a.w * b.z + a.x * b.y - a.y * b.x;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
static inline Q4 qnormalize(Q4 a) {
    double l = a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w;
    if (l > 0) { l = 1.0 / sqrt(l); a.x *= l; a.y *= l; a.z *= l; a.w *= l; }
    return a;
}
/* sine/cosine: the fdlibm method (Sun Microsystems, 1993; k_sin.c, k_cos.c, and the
 * two-part Cody-Waite reduction of e_rem_pio2.c), which JavaScript's Math.sin/Math.cos
 * descend from; the reference calls them through quat.setAxisAngle (mathcat).  The
 * polynomial coefficients are fdlibm's published minimax constants. */
static inline double o_ksin(double x, double y, int iy) {
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = x * x, v = z * x, r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}
static inline double o_kcos(double x, double y) {
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    uint64_t u;
    memcpy(&u, &x, 8);
    uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    double z = x * x, r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3FD33333u) return 1.0 - (0.5 * z - (z * r - x * y));
    double qx;
    if (ix > 0x3fe90000u) qx = 0.28125;
    else { uint64_t q = (uint64_t)(ix - 0x00200000u) << 32; memcpy(&qx, &q, 8); }
    double hz = 0.5 * z - qx, a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}
static inline void o_sincos(double x, double* s, double* c) {
    const double pio4 = 7.85398163397448278999e-01;
    if (fabs(x) <= pio4) { *s = o_ksin(x, 0.0, 0); *c = o_kcos(x, 0.0); return; }
    const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
                 pio2_1t = 6.07710050650619224932e-11;
    double fn = (double)(long long)(x * invpio2 + (x < 0 ? -0.5 : 0.5));
    long long n = (long long)fn;
    double r = x - fn * pio2_1, w = fn * pio2_1t, y0 = r - w, y1 = (r - y0) - w;
    double ks = o_ksin(y0, y1, 1), kc = o_kcos(y0, y1);
    switch (n & 3) {
        case 0: *s = ks; *c = kc; break;
        case 1: *s = kc; *c = -ks; break;
        case 2: *s = -ks; *c = -kc; break;
        default: *s = -kc; *c = ks; break;
    }
}
static inline Q4 qaxis_angle(V3 axis, double rad) {
    double h = rad * 0.5, s, c;
    o_sincos(h, &s, &c);
    Q4 r = {s * axis.x, s * axis.y, s * axis.z, c};
    return r;
}

/* ---- collision types shared by the two halves --------------------------- */
typedef struct {
    const cc_shape_desc* shapes;
    const cc_hull_data* hulls;
    const cc_mesh_data* meshes;
} OShapes;

#define OC_MAX_FACE 32
typedef struct { int n; V3 v[OC_MAX_FACE]; } OFace;

typedef struct {
    int hit;
    int status;
    V3 pointA, pointB, axis;
    double penetration;
    OFace faceA, faceB;
} OHit;

#define OC_MAX_MANIFOLD_POINTS 64
typedef struct {
    V3 base;
    V3 normal;
    double depth;
    int n;
    V3 relA[OC_MAX_MANIFOLD_POINTS], relB[OC_MAX_MANIFOLD_POINTS];
} OManifold;

int ocx_cast_convex_plane(const OShapes* sh, int sa, V3 pa, Q4 qa, V3 disp, int sp, V3 pb, Q4 qb, int backfaces,
                          double earlyOut, double* fraction, V3* pointA, V3* penAxis);
int ocx_cast_convex(const OShapes* sh, int sa, V3 pa, Q4 qa, V3 disp, int sb, V3 pb, Q4 qb, int backfaces,
                    double earlyOut, double* fraction, V3* pointA, V3* penAxis);
void ocx_collide(const OShapes* sh, int sa, V3 pa, Q4 qa, int sb, V3 pb, Q4 qb, double maxSep,
                 double collisionTol, double penetrationTol, OHit* out);
void ocx_manifold_between_two_faces(OManifold* out, V3 p1, V3 p2, V3 axis, double maxDist, const OFace* f1,
                                    const OFace* f2);
void ocx_prune(OManifold* m, V3 axis);
/* one hit of a convex shape (A) against a triangle of a mesh (B), in traversal order; the
 * triangle's sub-shape id rides along */
typedef void (*ocx_hit_fn)(void* ctx, const OHit* h, uint32_t sub_b);
void ocx_collide_mesh(const OShapes* sh, int sa, V3 pa, Q4 qa, int sm, V3 pm, Q4 qm, const cc_settings* s,
                      double maxSep, V3 moveDir, ocx_hit_fn fn, void* ctx);
void ocx_collide_sphere_mesh(const OShapes* sh, int sa, V3 pa, int sm, V3 pm, Q4 qm, const cc_settings* s,
                             double maxSep, V3 moveDir, ocx_hit_fn fn, void* ctx);
int ocx_mesh_calls(void);
extern int64_t ocx_epa_calls;

#endif
