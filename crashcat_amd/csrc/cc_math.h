// Device/host vector math for the gfx950 step kernels, templated on the bulk
// precision T (float for the batched path, double for the plain path).
// Conventions follow crashcat's math (gl-matrix style, via mathcat@0.0.10):
// column-major matrices, quaternions stored x, y, z, w.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CC_HD __host__ __device__ __forceinline__

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
struct Q4 {
    T x, y, z, w;
};
template <typename T>
struct M3 {
    T m[9];  // column-major: m[col * 3 + row]
};

template <typename T> CC_HD V3<T> mk3(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <typename T> CC_HD V3<T> operator+(V3<T> a, V3<T> b) { return mk3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> CC_HD V3<T> operator-(V3<T> a, V3<T> b) { return mk3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> CC_HD V3<T> operator-(V3<T> a) { return mk3<T>(-a.x, -a.y, -a.z); }
template <typename T> CC_HD V3<T> operator*(V3<T> a, T s) { return mk3<T>(a.x * s, a.y * s, a.z * s); }
template <typename T> CC_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> CC_HD V3<T> cross(V3<T> a, V3<T> b) {
    return mk3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> CC_HD T len2(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <typename T> CC_HD T vlen(V3<T> a) { return sqrt(len2(a)); }
template <typename T> CC_HD V3<T> normalize(V3<T> a) {
    T l = len2(a);
    if (l > T(0)) { l = T(1) / sqrt(l); return a * l; }
    return a;
}
template <typename T> CC_HD T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> CC_HD T tmax(T a, T b) { return a > b ? a : b; }
template <typename U, typename T> CC_HD V3<U> cvt(V3<T> a) { return mk3<U>(U(a.x), U(a.y), U(a.z)); }
template <typename U, typename T> CC_HD Q4<U> cvtq(Q4<T> a) { Q4<U> r; r.x = U(a.x); r.y = U(a.y); r.z = U(a.z); r.w = U(a.w); return r; }

// mat4.fromQuat (3x3 part)
template <typename T> CC_HD M3<T> m3_from_quat(Q4<T> q) {
    T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    T xx = q.x * x2, yx = q.y * x2, yy = q.y * y2, zx = q.z * x2, zy = q.z * y2, zz = q.z * z2;
    T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    M3<T> r;
    r.m[0] = T(1) - yy - zz; r.m[1] = yx + wz; r.m[2] = zx - wy;
    r.m[3] = yx - wz; r.m[4] = T(1) - xx - zz; r.m[5] = zy + wx;
    r.m[6] = zx + wy; r.m[7] = zy - wx; r.m[8] = T(1) - xx - yy;
    return r;
}
template <typename T> CC_HD V3<T> mul(const M3<T>& m, V3<T> v) {
    return mk3<T>(m.m[0] * v.x + m.m[3] * v.y + m.m[6] * v.z, m.m[1] * v.x + m.m[4] * v.y + m.m[7] * v.z,
                  m.m[2] * v.x + m.m[5] * v.y + m.m[8] * v.z);
}
template <typename T> CC_HD V3<T> mul_t(const M3<T>& m, V3<T> v) {
    return mk3<T>(m.m[0] * v.x + m.m[1] * v.y + m.m[2] * v.z, m.m[3] * v.x + m.m[4] * v.y + m.m[5] * v.z,
                  m.m[6] * v.x + m.m[7] * v.y + m.m[8] * v.z);
}
// vec3.transformQuat
template <typename T> CC_HD V3<T> qrotate(Q4<T> q, V3<T> a) {
    T uvx = q.y * a.z - q.z * a.y, uvy = q.z * a.x - q.x * a.z, uvz = q.x * a.y - q.y * a.x;
    T uuvx = q.y * uvz - q.z * uvy, uuvy = q.z * uvx - q.x * uvz, uuvz = q.x * uvy - q.y * uvx;
    T w2 = q.w * T(2);
    return mk3<T>(a.x + uvx * w2 + uuvx * T(2), a.y + uvy * w2 + uuvy * T(2), a.z + uvz * w2 + uuvz * T(2));
}
// v + 2 * cross(q.xyz, cross(q.xyz, v) + q.w * v)  (the inline form of src/shapes/convex.ts)
template <typename T> CC_HD V3<T> qrot2(Q4<T> q, V3<T> v) {
    T uvx = q.y * v.z - q.z * v.y + q.w * v.x;
    T uvy = q.z * v.x - q.x * v.z + q.w * v.y;
    T uvz = q.x * v.y - q.y * v.x + q.w * v.z;
    T uuvx = q.y * uvz - q.z * uvy, uuvy = q.z * uvx - q.x * uvz, uuvz = q.x * uvy - q.y * uvx;
    return mk3<T>(v.x + T(2) * uuvx, v.y + T(2) * uuvy, v.z + T(2) * uuvz);
}
template <typename T> CC_HD Q4<T> qmul(Q4<T> a, Q4<T> b) {
    Q4<T> r;
    r.x = a.x * b.w + a.w * b.x + a.y * b.z - a.z * b.y;
    r.y = a.y * b.w + a.w * b.y + a.z * b.x - a.x * b.z;
    r.z = a.z * b.w + a.w * b.z + a.x * b.y - a.y * b.x;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
template <typename T> CC_HD Q4<T> qnormalize(Q4<T> a) {
    T l = a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w;
    if (l > T(0)) { l = T(1) / sqrt(l); a.x *= l; a.y *= l; a.z *= l; a.w *= l; }
    return a;
}
// FP64 sine and cosine by the fdlibm method (Sun Microsystems, 1993): minimax polynomial
// kernels on [-pi/4, pi/4] and a two-part Cody-Waite reduction by pi/2 outside it.  Only
// IEEE add/mul/div are used, so with contraction off the results are the same bits on the
// GPU and on any host; JavaScript engines descend from the same kernels.
CC_HD double fd_ksin(double x, double y, int iy) {
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = x * x, v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}
CC_HD double fd_kcos(double x, double y) {
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    union { double d; unsigned long long u; } b;
    b.d = x;
    unsigned ix = (unsigned)(b.u >> 32) & 0x7fffffffu;
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3FD33333u) return 1.0 - (0.5 * z - (z * r - x * y));  // |x| < 0.3
    double qx;
    if (ix > 0x3fe90000u) qx = 0.28125;                                // |x| > 0.78125
    else { b.u = (unsigned long long)(ix - 0x00200000u) << 32; qx = b.d; }  // x/4, low word cleared
    double hz = 0.5 * z - qx, a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}
CC_HD void fd_sincos(double x, double& s, double& c) {
    const double pio4 = 7.85398163397448278999e-01;
    if (fabs(x) <= pio4) { s = fd_ksin(x, 0.0, 0); c = fd_kcos(x, 0.0); return; }
    const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
                 pio2_1t = 6.07710050650619224932e-11;
    double fn = (double)(long long)(x * invpio2 + (x < 0 ? -0.5 : 0.5));
    long long n = (long long)fn;
    double r = x - fn * pio2_1, w = fn * pio2_1t;
    double y0 = r - w, y1 = (r - y0) - w;
    double ks = fd_ksin(y0, y1, 1), kc = fd_kcos(y0, y1);
    switch (n & 3) {
        case 0: s = ks; c = kc; break;
        case 1: s = kc; c = -ks; break;
        case 2: s = -ks; c = -kc; break;
        default: s = -kc; c = ks; break;
    }
}
CC_HD void sin_cos(double x, double& s, double& c) { fd_sincos(x, s, c); }
CC_HD void sin_cos(float x, float& s, float& c) { s = sinf(x); c = cosf(x); }

template <typename T> CC_HD Q4<T> qaxis_angle(V3<T> axis, T rad) {
    T h = rad * T(0.5);
    T s, c;
    sin_cos(h, s, c);
    Q4<T> r; r.x = s * axis.x; r.y = s * axis.y; r.z = s * axis.z; r.w = c;
    return r;
}
template <typename T> CC_HD Q4<T> qconj(Q4<T> a) { Q4<T> r; r.x = -a.x; r.y = -a.y; r.z = -a.z; r.w = a.w; return r; }

// getInverseInertiaForRotation (src/body/motion-properties.ts:494-621), computed in FP64.
CC_HD M3<double> inv_inertia_d(V3<double> invDiag, Q4<double> iq, const M3<double>& bR, int dof) {
    M3<double> I = m3_from_quat(iq);
    double r[9];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int row = 0; row < 3; row++)
            r[c * 3 + row] = bR.m[row] * I.m[c * 3 + 0] + bR.m[3 + row] * I.m[c * 3 + 1] + bR.m[6 + row] * I.m[c * 3 + 2];
    double d[3] = {invDiag.x, invDiag.y, invDiag.z};
    M3<double> o;
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            o.m[j * 3 + k] = d[0] * r[0 * 3 + j] * r[0 * 3 + k] + d[1] * r[1 * 3 + j] * r[1 * 3 + k] + d[2] * r[2 * 3 + j] * r[2 * 3 + k];
    int rot = (dof >> 3) & 7;
    if (rot != 7) {
        double m[3] = {(rot & 1) ? 1.0 : 0.0, (rot & 2) ? 1.0 : 0.0, (rot & 4) ? 1.0 : 0.0};
        for (int c = 0; c < 3; c++)
            for (int row = 0; row < 3; row++) o.m[c * 3 + row] *= m[row] * m[c];
    }
    return o;
}

template <typename T> CC_HD M3<T> inv_inertia(V3<T> invDiag, Q4<T> iq, Q4<T> bodyQ, int dof) {
    M3<double> R = m3_from_quat(cvtq<double>(bodyQ));
    M3<double> o = inv_inertia_d(cvt<double>(invDiag), cvtq<double>(iq), R, dof);
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.m[i] = T(o.m[i]);
    return r;
}

template <typename T> CC_HD V3<T> dof_lin(V3<T> v, int dof) {
    if (!(dof & 1)) v.x = T(0);
    if (!(dof & 2)) v.y = T(0);
    if (!(dof & 4)) v.z = T(0);
    return v;
}
