// FP64 device math of the XPBD substep.
//
// Every expression keeps the reference's operation order and association
// (compiled with -ffp-contract=off, no fast-math), so each projection is
// bit-identical to the reference's CPU projection given identical inputs:
// +, -, *, / and sqrt are correctly rounded on both sides. The
// transcendentals atan2, sin, cos and acos are the portable double-double
// functions of portable_math.hpp (correctly rounded in practice, identical
// bits on host and device); remainder is exact.
//
// Vector helpers follow src/softwing/include/softwing/vec3.h (dot and cross
// left to right, length = sqrt(dot)); quaternion and 3x3 helpers follow
// src/softwing/suspension.cpp:553-684.
#pragma once

#include <hip/hip_runtime.h>

#include "portable_math.hpp"

namespace swx {
namespace dev {

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 scale(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d3 divide(d3 a, double s) { return d3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ d3 neg(d3 a) { return d3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double len(d3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ bool fin(double v) { return isfinite(v); }
__device__ __forceinline__ bool fin3(d3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// std::min / std::max / std::clamp semantics (first argument kept on ties
// and when the comparison is false).
__device__ __forceinline__ double stdmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double stdmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double stdclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// Symmetric 3x3 (membrane.h:65-75): xx, yy, zz, xy, xz, yz.
struct s3 {
    double xx, yy, zz, xy, xz, yz;
};
__device__ __forceinline__ d3 mul(const s3& m, d3 v) {
    return d3{m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
              m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
__device__ __forceinline__ bool finS(const s3& m) {
    return isfinite(m.xx) && isfinite(m.yy) && isfinite(m.zz) && isfinite(m.xy) && isfinite(m.xz) && isfinite(m.yz);
}

constexpr double kEps = 2.220446049250313080847263336181640625e-16;

// checkedInverse(SymmetricMatrix3) accept test and cofactors
// (membrane.cpp:208-235). Returns false when the reference throws.
__device__ __forceinline__ bool checkedInverse(const s3& m, s3& inv) {
    if (!finS(m)) return false;
    const double scaleValue =
        stdmax(stdmax(stdmax(stdmax(stdmax(fabs(m.xx), fabs(m.yy)), fabs(m.zz)), fabs(m.xy)), fabs(m.xz)),
               fabs(m.yz));
    const double det = m.xx * (m.yy * m.zz - m.yz * m.yz) - m.xy * (m.xy * m.zz - m.yz * m.xz) +
                       m.xz * (m.xy * m.yz - m.yy * m.xz);
    const double tol = 128.0 * kEps * scaleValue * scaleValue * scaleValue;
    if (!(scaleValue > 0.0) || !isfinite(det) || fabs(det) <= tol) return false;
    const double r = 1.0 / det;
    inv.xx = (m.yy * m.zz - m.yz * m.yz) * r;
    inv.yy = (m.xx * m.zz - m.xz * m.xz) * r;
    inv.zz = (m.xx * m.yy - m.xy * m.xy) * r;
    inv.xy = (m.xz * m.yz - m.xy * m.zz) * r;
    inv.xz = (m.xy * m.yz - m.xz * m.yy) * r;
    inv.yz = (m.xy * m.xz - m.xx * m.yz) * r;
    return true;
}

// isPositiveDefinite (membrane.cpp:186-206).
__device__ __forceinline__ bool positiveDefinite(const s3& m) {
    if (!finS(m) || !(m.xx > 0.0) || !(m.yy > 0.0) || !(m.zz > 0.0)) return false;
    const double sx = sqrt(m.xx), sy = sqrt(m.yy), sz = sqrt(m.zz);
    const double l10 = m.xy / sx / sy;
    const double l20 = m.xz / sx / sz;
    const double p1 = 1.0 - l10 * l10;
    if (!isfinite(p1) || !(p1 > 0.0)) return false;
    const double l11 = sqrt(p1);
    const double l21 = (m.yz / sy / sz - l20 * l10) / l11;
    const double p2 = 1.0 - l20 * l20 - l21 * l21;
    return isfinite(p2) && p2 > 0.0;
}

// Equilibrated Cholesky fallback with residual certificate
// (membrane.cpp:31-127). Returns false when the reference throws.
__device__ bool scaledCholeskySolve(const s3& m, d3 rhs, d3& out) {
    const double sx = sqrt(m.xx), sy = sqrt(m.yy), sz = sqrt(m.zz);
    if (!(sx > 0.0) || !(sy > 0.0) || !(sz > 0.0) || !isfinite(sx) || !isfinite(sy) || !isfinite(sz)) return false;
    const double l10 = m.xy / sx / sy;
    const double l20 = m.xz / sx / sz;
    const double p1 = 1.0 - l10 * l10;
    const double minPivot = 4096.0 * kEps;
    if (!isfinite(p1) || !(p1 > minPivot)) return false;
    const double l11 = sqrt(p1);
    const double l21 = (m.yz / sy / sz - l20 * l10) / l11;
    const double p2 = 1.0 - l20 * l20 - l21 * l21;
    if (!isfinite(l21) || !isfinite(p2) || !(p2 > minPivot)) return false;
    const double l22 = sqrt(p2);
    const double b0 = rhs.x / sx, b1 = rhs.y / sy, b2 = rhs.z / sz;
    const double f0 = b0;
    const double f1 = (b1 - l10 * f0) / l11;
    const double f2 = (b2 - l20 * f0 - l21 * f1) / l22;
    const double rz = f2 / l22;
    const double ry = (f1 - l21 * rz) / l11;
    const double rx = f0 - l10 * ry - l20 * rz;
    const d3 result{rx / sx, ry / sy, rz / sz};
    if (!fin3(result)) return false;
    const d3 residual = sub(mul(m, result), rhs);
    const double rowX = fabs(m.xx * result.x) + fabs(m.xy * result.y) + fabs(m.xz * result.z) + fabs(rhs.x);
    const double rowY = fabs(m.xy * result.x) + fabs(m.yy * result.y) + fabs(m.yz * result.z) + fabs(rhs.y);
    const double rowZ = fabs(m.xz * result.x) + fabs(m.yz * result.y) + fabs(m.zz * result.z) + fabs(rhs.z);
    const double maxRel = 4096.0 * kEps;
    if (!fin3(residual) || !isfinite(rowX) || !isfinite(rowY) || !isfinite(rowZ) ||
        (rowX == 0.0 ? residual.x != 0.0 : fabs(residual.x) > maxRel * rowX) ||
        (rowY == 0.0 ? residual.y != 0.0 : fabs(residual.y) > maxRel * rowY) ||
        (rowZ == 0.0 ? residual.z != 0.0 : fabs(residual.z) > maxRel * rowZ))
        return false;
    out = result;
    return true;
}

// checkedSolve (membrane.cpp:237-257).
__device__ __forceinline__ bool checkedSolve(const s3& m, d3 rhs, d3& out) {
    if (!fin3(rhs)) return false;
    s3 inv;
    if (checkedInverse(m, inv)) {
        out = mul(inv, rhs);
        return true;
    }
    if (!finS(m) || !positiveDefinite(m)) return false;
    return scaledCholeskySolve(m, rhs, out);
}

// ---- Rigid payload helpers (suspension.cpp:553-684, 1018-1102) ---------

struct quat {
    double w, x, y, z;
};
struct m3 {
    double a[9];  // row-major
};

__device__ __forceinline__ double qnorm2(const quat& q) { return q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z; }
__device__ __forceinline__ bool finQ(const quat& q) {
    return isfinite(q.w) && isfinite(q.x) && isfinite(q.y) && isfinite(q.z);
}

// normalizedCanonical; false where the reference fails.
__device__ __forceinline__ bool canonical(const quat& q, quat& out) {
    const double n2 = qnorm2(q);
    if (!finQ(q) || !(n2 > 1.0e-28)) return false;
    const double inv = 1.0 / sqrt(n2);
    quat r{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
    const bool flip = r.w < 0.0 || (r.w == 0.0 && (r.x < 0.0 || (r.x == 0.0 && (r.y < 0.0 || (r.y == 0.0 && r.z < 0.0)))));
    if (flip) {
        r.w = -r.w;
        r.x = -r.x;
        r.y = -r.y;
        r.z = -r.z;
    }
    out = r;
    return true;
}

__device__ __forceinline__ quat qmul(const quat& a, const quat& b) {
    return quat{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}

__device__ __forceinline__ bool rotationMatrix(const quat& orientation, m3& r) {
    quat q;
    if (!canonical(orientation, q)) return false;
    const double xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const double xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const double wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    r.a[0] = 1.0 - 2.0 * (yy + zz);
    r.a[1] = 2.0 * (xy - wz);
    r.a[2] = 2.0 * (xz + wy);
    r.a[3] = 2.0 * (xy + wz);
    r.a[4] = 1.0 - 2.0 * (xx + zz);
    r.a[5] = 2.0 * (yz - wx);
    r.a[6] = 2.0 * (xz - wy);
    r.a[7] = 2.0 * (yz + wx);
    r.a[8] = 1.0 - 2.0 * (xx + yy);
    return true;
}

__device__ __forceinline__ d3 mulM(const m3& m, d3 v) {
    return d3{m.a[0] * v.x + m.a[1] * v.y + m.a[2] * v.z, m.a[3] * v.x + m.a[4] * v.y + m.a[5] * v.z,
              m.a[6] * v.x + m.a[7] * v.y + m.a[8] * v.z};
}

__device__ __forceinline__ m3 mulMM(const m3& a, const m3& b) {
    m3 r;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a.a[row * 3 + k] * b.a[k * 3 + col];
            r.a[row * 3 + col] = s;
        }
    return r;
}

__device__ __forceinline__ m3 transposeM(const m3& a) {
    m3 r;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) r.a[row * 3 + col] = a.a[col * 3 + row];
    return r;
}

__device__ __forceinline__ double detM(const m3& m) {
    return m.a[0] * (m.a[4] * m.a[8] - m.a[5] * m.a[7]) - m.a[1] * (m.a[3] * m.a[8] - m.a[5] * m.a[6]) +
           m.a[2] * (m.a[3] * m.a[7] - m.a[4] * m.a[6]);
}

// inverse(Mat3) with max-magnitude normalisation (suspension.cpp:567-600).
__device__ __forceinline__ bool inverseM(const m3& m, m3& r) {
    // |first largest-magnitude entry| (a running maximum: no dynamic index,
    // so the matrix stays in registers).
    double s = fabs(m.a[0]);
#pragma unroll
    for (int k = 1; k < 9; ++k) s = (s < fabs(m.a[k])) ? fabs(m.a[k]) : s;
    if (!(s > 0.0) || !isfinite(s)) return false;
    m3 n;
    for (int k = 0; k < 9; ++k) n.a[k] = m.a[k] / s;
    const double det = detM(n);
    if (!isfinite(det) || fabs(det) <= 1.0e-14) return false;
    r.a[0] = n.a[4] * n.a[8] - n.a[5] * n.a[7];
    r.a[1] = n.a[2] * n.a[7] - n.a[1] * n.a[8];
    r.a[2] = n.a[1] * n.a[5] - n.a[2] * n.a[4];
    r.a[3] = n.a[5] * n.a[6] - n.a[3] * n.a[8];
    r.a[4] = n.a[0] * n.a[8] - n.a[2] * n.a[6];
    r.a[5] = n.a[2] * n.a[3] - n.a[0] * n.a[5];
    r.a[6] = n.a[3] * n.a[7] - n.a[4] * n.a[6];
    r.a[7] = n.a[1] * n.a[6] - n.a[0] * n.a[7];
    r.a[8] = n.a[0] * n.a[4] - n.a[1] * n.a[3];
    const double denom = det * s;
    for (int k = 0; k < 9; ++k) r.a[k] /= denom;
    return true;
}

// payloadWorldInertia = R * I_body * R^T for a diagonal body inertia.
__device__ __forceinline__ bool worldInertia(const quat& q, const double inertia[3], m3& out) {
    m3 r;
    if (!rotationMatrix(q, r)) return false;
    m3 body{{inertia[0], 0.0, 0.0, 0.0, inertia[1], 0.0, 0.0, 0.0, inertia[2]}};
    out = mulMM(mulMM(r, body), transposeM(r));
    return true;
}

// rotationIncrement (suspension.cpp:643-658).
__device__ __forceinline__ bool rotationIncrement(d3 v, quat& out) {
    if (!fin3(v)) return false;
    const double angle = len(v);
    if (angle < 1.0e-12) return canonical(quat{1.0, 0.5 * v.x, 0.5 * v.y, 0.5 * v.z}, out);
    const double half = 0.5 * angle;
    double sine, cosine;
    pm::sincos(half, &sine, &cosine);
    const double s = sine / angle;
    out = quat{cosine, v.x * s, v.y * s, v.z * s};
    return true;
}

}  // namespace dev
}  // namespace swx
