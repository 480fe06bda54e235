// ORACLE (test infrastructure only; never linked into the product library).
//
// Plain serial C++ restatement of the reference's coupled XPBD substep, in
// the reference's own data layout (array of node records) and sweep orders
// (coloured for workerThreads >= 1, definition order for workerThreads = 0).
// Each function cites the reference lines it
// restates. Compiled with g++ -O2 -ffp-contract=off, like the reference.
//
// Pinned by: bit-identical states against the fixtures the reference itself
// produced (tests/golden, written by oracle/ref_driver against
// /root/reference sources), plus closed-form checks in tests/.
// Build-defined pieces with no reference counterpart (the contact-sphere
// family) are pinned by closed forms only.
#include "xpbd_oracle.hpp"

#include "oracle_colouring.hpp"

#include "../simwing_amd/csrc/portable_math.hpp"

#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <queue>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

// Scalar of every computed quantity: double (the restatement itself), long
// double (the extended build, ORACLE_REAL), or a forward-mode tangent pair
// (the tangent build, ORACLE_TANGENT) that carries one directional
// derivative through the identical operation sequence; the extended build's
// central differences and the tangent build's derivatives are the checks of
// the adjoint.
namespace oracle {
struct Tangent {
    double v = 0.0, d = 0.0;
    Tangent() = default;
    Tangent(double value) : v(value) {}  // constants carry no derivative
    Tangent(double value, double derivative) : v(value), d(derivative) {}
    explicit operator double() const { return v; }
    Tangent& operator+=(const Tangent& o) { d += o.d; v += o.v; return *this; }
    Tangent& operator-=(const Tangent& o) { d -= o.d; v -= o.v; return *this; }
    Tangent& operator*=(const Tangent& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    Tangent& operator/=(const Tangent& o) {
        const double q = v / o.v;
        d = (d - q * o.d) / o.v;
        v = q;
        return *this;
    }
};
inline Tangent operator+(Tangent a, const Tangent& b) { return a += b; }
inline Tangent operator-(Tangent a, const Tangent& b) { return a -= b; }
inline Tangent operator*(Tangent a, const Tangent& b) { return a *= b; }
inline Tangent operator/(Tangent a, const Tangent& b) { return a /= b; }
inline Tangent operator+(Tangent a, double b) { return a += Tangent(b); }
inline Tangent operator-(Tangent a, double b) { return a -= Tangent(b); }
inline Tangent operator*(Tangent a, double b) { return {a.v * b, a.d * b}; }
inline Tangent operator/(Tangent a, double b) { return {a.v / b, a.d / b}; }
inline Tangent operator+(double a, const Tangent& b) { return Tangent(a) += b; }
inline Tangent operator-(double a, const Tangent& b) { return Tangent(a) -= b; }
inline Tangent operator*(double a, const Tangent& b) { return {a * b.v, a * b.d}; }
inline Tangent operator/(double a, const Tangent& b) { return Tangent(a) /= b; }
inline Tangent operator-(const Tangent& a) { return {-a.v, -a.d}; }
inline bool operator<(const Tangent& a, const Tangent& b) { return a.v < b.v; }
inline bool operator>(const Tangent& a, const Tangent& b) { return a.v > b.v; }
inline bool operator<=(const Tangent& a, const Tangent& b) { return a.v <= b.v; }
inline bool operator>=(const Tangent& a, const Tangent& b) { return a.v >= b.v; }
inline bool operator==(const Tangent& a, const Tangent& b) { return a.v == b.v; }
inline bool operator!=(const Tangent& a, const Tangent& b) { return a.v != b.v; }
inline bool operator<(const Tangent& a, double b) { return a.v < b; }
inline bool operator>(const Tangent& a, double b) { return a.v > b; }
inline bool operator<=(const Tangent& a, double b) { return a.v <= b; }
inline bool operator>=(const Tangent& a, double b) { return a.v >= b; }
inline bool operator==(const Tangent& a, double b) { return a.v == b; }
inline bool operator!=(const Tangent& a, double b) { return a.v != b; }
inline bool operator<(double a, const Tangent& b) { return a < b.v; }
inline bool operator>(double a, const Tangent& b) { return a > b.v; }
inline Tangent sqrt(const Tangent& a) {
    const double r = std::sqrt(a.v);
    return {r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
inline Tangent abs(const Tangent& a) { return a.v < 0.0 ? -a : a; }
inline bool isfinite(const Tangent& a) { return std::isfinite(a.v); }
inline Tangent remainder(const Tangent& a, double b) { return {std::remainder(a.v, b), a.d}; }
inline Tangent atan2(const Tangent& y, const Tangent& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return {std::atan2(y.v, x.v), r2 > 0.0 ? (x.v * y.d - y.v * x.d) / r2 : 0.0};
}
inline Tangent sin(const Tangent& a) { return {std::sin(a.v), std::cos(a.v) * a.d}; }
inline Tangent cos(const Tangent& a) { return {std::cos(a.v), -std::sin(a.v) * a.d}; }
inline Tangent acos(const Tangent& a) {
    const double r = 1.0 - a.v * a.v;
    return {std::acos(a.v), r > 0.0 ? -a.d / std::sqrt(r) : 0.0};
}
inline Tangent exp(const Tangent& a) {
    const double e = std::exp(a.v);
    return {e, e * a.d};
}
inline Tangent nextafter(const Tangent& a, double to) { return {std::nextafter(a.v, to), a.d}; }
inline double value(const Tangent& a) { return a.v; }
inline double tangentOf(const Tangent& a) { return a.d; }
inline double value(double a) { return a; }
inline double value(long double a) { return static_cast<double>(a); }
inline double tangentOf(double) { return 0.0; }
inline double tangentOf(long double) { return 0.0; }
// The portable transcendentals (the device kernels' functions): their own
// values for double, the same values with the analytic derivative for the
// tangent pair; the extended build has none and keeps the C library's.
inline double portableAtan2(double y, double x) { return swx::pm::atan2(y, x); }
inline double portableSin(double x) { return swx::pm::sin(x); }
inline double portableCos(double x) { return swx::pm::cos(x); }
inline double portableAcos(double x) { return swx::pm::acos(x); }
inline long double portableAtan2(long double y, long double x) { return std::atan2(y, x); }
inline long double portableSin(long double x) { return std::sin(x); }
inline long double portableCos(long double x) { return std::cos(x); }
inline long double portableAcos(long double x) { return std::acos(x); }
inline Tangent portableAtan2(const Tangent& y, const Tangent& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return {swx::pm::atan2(y.v, x.v), r2 > 0.0 ? (x.v * y.d - y.v * x.d) / r2 : 0.0};
}
inline Tangent portableSin(const Tangent& a) { return {swx::pm::sin(a.v), swx::pm::cos(a.v) * a.d}; }
inline Tangent portableCos(const Tangent& a) { return {swx::pm::cos(a.v), -swx::pm::sin(a.v) * a.d}; }
inline Tangent portableAcos(const Tangent& a) {
    const double r = 1.0 - a.v * a.v;
    return {swx::pm::acos(a.v), r > 0.0 ? -a.d / std::sqrt(r) : 0.0};
}
inline void seedTangent(Tangent& a, double d) { a.d = d; }
inline void seedTangent(double&, double) {}
inline void seedTangent(long double&, double) {}
}  // namespace oracle

#if defined(ORACLE_TANGENT)
using Real = oracle::Tangent;
#elif defined(ORACLE_REAL)
using Real = ORACLE_REAL;
#else
using Real = double;
#endif
constexpr bool kTangentBuild = std::is_same_v<Real, oracle::Tangent>;

using oracle::seedTangent;
using oracle::tangentOf;
using oracle::value;
using std::abs;
using std::acos;
using std::atan2;
using std::cos;
using std::exp;
using std::isfinite;
using std::nextafter;
using std::remainder;
using std::sin;
using std::sqrt;

namespace {

thread_local std::string lastError;

// Transcendentals: the C library's, as the reference calls them (default),
// or the portable Real-Real functions that the device kernels use
// (oracle_set_portable_math), which return the same bits on every platform.
// The tangent build follows the flag too (portable values, analytic
// derivatives); the extended build has only the C library's functions.
thread_local bool portableMath = false;
Real mAtan2(Real y, Real x) { return portableMath ? oracle::portableAtan2(y, x) : atan2(y, x); }
Real mSin(Real x) { return portableMath ? oracle::portableSin(x) : sin(x); }
Real mCos(Real x) { return portableMath ? oracle::portableCos(x) : cos(x); }
Real mAcos(Real x) { return portableMath ? oracle::portableAcos(x) : acos(x); }

struct V {
    Real x = 0.0, y = 0.0, z = 0.0;
};
V operator+(V a, const V& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, const V& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator*(V a, Real s) { return {a.x * s, a.y * s, a.z * s}; }
V operator*(Real s, V a) { return {a.x * s, a.y * s, a.z * s}; }
V operator/(V a, Real s) { return {a.x / s, a.y / s, a.z / s}; }
V operator-(V a) { return {-a.x, -a.y, -a.z}; }
V& operator+=(V& a, const V& b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
V& operator-=(V& a, const V& b) {
    a.x -= b.x;
    a.y -= b.y;
    a.z -= b.z;
    return a;
}
Real dot(const V& a, const V& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V cross(const V& a, const V& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Real length(const V& a) { return sqrt(dot(a, a)); }
bool finite(const V& a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

struct Sym {
    Real xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
};
V operator*(const Sym& m, const V& v) {
    return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
            m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
Sym operator*(Real s, const Sym& m) { return {s * m.xx, s * m.yy, s * m.zz, s * m.xy, s * m.xz, s * m.yz}; }
Sym operator+(const Sym& a, const Sym& b) {
    return {a.xx + b.xx, a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}
bool finite(const Sym& m) {
    return isfinite(m.xx) && isfinite(m.yy) && isfinite(m.zz) && isfinite(m.xy) &&
           isfinite(m.xz) && isfinite(m.yz);
}

// membrane.cpp:208-235
bool symInverse(const Sym& m, Sym& out) {
    if (!finite(m)) return false;
    const Real scale = std::max({abs(m.xx), abs(m.yy), abs(m.zz), abs(m.xy),
                                   abs(m.xz), abs(m.yz)});
    const Real det = m.xx * (m.yy * m.zz - m.yz * m.yz) - m.xy * (m.xy * m.zz - m.yz * m.xz) +
                       m.xz * (m.xy * m.yz - m.yy * m.xz);
    const Real tol = 128.0 * std::numeric_limits<double>::epsilon() * scale * scale * scale;
    if (!(scale > 0.0) || !isfinite(det) || abs(det) <= tol) return false;
    const Real inv = 1.0 / det;
    out = {(m.yy * m.zz - m.yz * m.yz) * inv, (m.xx * m.zz - m.xz * m.xz) * inv,
           (m.xx * m.yy - m.xy * m.xy) * inv, (m.xz * m.yz - m.xy * m.zz) * inv,
           (m.xy * m.yz - m.xz * m.yy) * inv, (m.xy * m.xz - m.xx * m.yz) * inv};
    return true;
}

// membrane.cpp:186-206
bool symPositiveDefinite(const Sym& m) {
    if (!finite(m) || !(m.xx > 0.0) || !(m.yy > 0.0) || !(m.zz > 0.0)) return false;
    const Real sx = sqrt(m.xx), sy = sqrt(m.yy), sz = sqrt(m.zz);
    const Real l10 = m.xy / sx / sy, l20 = m.xz / sx / sz;
    const Real p1 = 1.0 - l10 * l10;
    if (!isfinite(p1) || !(p1 > 0.0)) return false;
    const Real l11 = sqrt(p1);
    const Real l21 = (m.yz / sy / sz - l20 * l10) / l11;
    const Real p2 = 1.0 - l20 * l20 - l21 * l21;
    return isfinite(p2) && p2 > 0.0;
}

// membrane.cpp:31-127
V scaledSolve(const Sym& m, const V& b) {
    const auto fail = [] { throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned"); };
    const Real sx = sqrt(m.xx), sy = sqrt(m.yy), sz = sqrt(m.zz);
    if (!(sx > 0.0) || !(sy > 0.0) || !(sz > 0.0) || !isfinite(sx) || !isfinite(sy) || !isfinite(sz))
        fail();
    const Real l10 = m.xy / sx / sy, l20 = m.xz / sx / sz;
    const Real p1 = 1.0 - l10 * l10;
    const Real minPivot = 4096.0 * std::numeric_limits<double>::epsilon();
    if (!isfinite(p1) || !(p1 > minPivot)) fail();
    const Real l11 = sqrt(p1);
    const Real l21 = (m.yz / sy / sz - l20 * l10) / l11;
    const Real p2 = 1.0 - l20 * l20 - l21 * l21;
    if (!isfinite(l21) || !isfinite(p2) || !(p2 > minPivot)) fail();
    const Real l22 = sqrt(p2);
    const V s{b.x / sx, b.y / sy, b.z / sz};
    const Real f0 = s.x;
    const Real f1 = (s.y - l10 * f0) / l11;
    const Real f2 = (s.z - l20 * f0 - l21 * f1) / l22;
    const Real rz = f2 / l22;
    const Real ry = (f1 - l21 * rz) / l11;
    const Real rx = f0 - l10 * ry - l20 * rz;
    const V r{rx / sx, ry / sy, rz / sz};
    if (!finite(r)) throw std::invalid_argument("SymmetricMatrix3 solve result is non-finite");
    const V res = m * r - b;
    const Real ax = abs(m.xx * r.x) + abs(m.xy * r.y) + abs(m.xz * r.z) + abs(b.x);
    const Real ay = abs(m.xy * r.x) + abs(m.yy * r.y) + abs(m.yz * r.z) + abs(b.y);
    const Real az = abs(m.xz * r.x) + abs(m.yz * r.y) + abs(m.zz * r.z) + abs(b.z);
    const Real rel = 4096.0 * std::numeric_limits<double>::epsilon();
    if (!finite(res) || !isfinite(ax) || !isfinite(ay) || !isfinite(az) ||
        (ax == 0.0 ? res.x != 0.0 : abs(res.x) > rel * ax) ||
        (ay == 0.0 ? res.y != 0.0 : abs(res.y) > rel * ay) ||
        (az == 0.0 ? res.z != 0.0 : abs(res.z) > rel * az))
        fail();
    return r;
}

// membrane.cpp:237-257
V checkedSolve(const Sym& m, const V& b) {
    if (!finite(b)) throw std::invalid_argument("Symmetric solve right-hand side must be finite");
    Sym inv;
    if (symInverse(m, inv)) return inv * b;
    if (!finite(m) || !symPositiveDefinite(m)) throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned");
    return scaledSolve(m, b);
}

struct Q {
    Real w = 1, x = 0, y = 0, z = 0;
};
struct M3 {
    Real a[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    Real& operator()(int r, int c) { return a[r * 3 + c]; }
    Real operator()(int r, int c) const { return a[r * 3 + c]; }
};

struct SuspensionFail : std::runtime_error {
    int phase;
    std::string entity;
    SuspensionFail(int p, std::string e, const std::string& m) : std::runtime_error(m), phase(p), entity(std::move(e)) {}
};
[[noreturn]] void fail(int phase, const std::string& entity, const std::string& message) {
    throw SuspensionFail(phase, entity, message);
}
bool finite(const Q& q) { return isfinite(q.w) && isfinite(q.x) && isfinite(q.y) && isfinite(q.z); }

// suspension.cpp:553-684
M3 transpose(const M3& m) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = m(j, i);
    return r;
}
Real det3(const M3& m) {
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) - m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
           m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}
M3 inverse3(const M3& m) {
    Real scale = abs(m.a[0]);
    for (int k = 1; k < 9; ++k)
        if (scale < abs(m.a[k])) scale = abs(m.a[k]);
    if (!(scale > 0.0) || !isfinite(scale)) fail(1, "inertia", "singular matrix");
    M3 n = m;
    for (Real& v : n.a) v /= scale;
    const Real det = det3(n);
    if (!isfinite(det) || abs(det) <= 1.0e-14) fail(1, "inertia", "singular matrix");
    M3 r;
    r.a[0] = n(1, 1) * n(2, 2) - n(1, 2) * n(2, 1);
    r.a[1] = n(0, 2) * n(2, 1) - n(0, 1) * n(2, 2);
    r.a[2] = n(0, 1) * n(1, 2) - n(0, 2) * n(1, 1);
    r.a[3] = n(1, 2) * n(2, 0) - n(1, 0) * n(2, 2);
    r.a[4] = n(0, 0) * n(2, 2) - n(0, 2) * n(2, 0);
    r.a[5] = n(0, 2) * n(1, 0) - n(0, 0) * n(1, 2);
    r.a[6] = n(1, 0) * n(2, 1) - n(1, 1) * n(2, 0);
    r.a[7] = n(0, 1) * n(2, 0) - n(0, 0) * n(2, 1);
    r.a[8] = n(0, 0) * n(1, 1) - n(0, 1) * n(1, 0);
    for (Real& v : r.a) v /= det * scale;
    return r;
}
V operator*(const M3& m, const V& v) {
    return {m(0, 0) * v.x + m(0, 1) * v.y + m(0, 2) * v.z, m(1, 0) * v.x + m(1, 1) * v.y + m(1, 2) * v.z,
            m(2, 0) * v.x + m(2, 1) * v.y + m(2, 2) * v.z};
}
M3 operator*(const M3& a, const M3& b) {
    M3 r;
    for (Real& v : r.a) v = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) r(i, j) += a(i, k) * b(k, j);
    return r;
}
Q canonical(const Q& q) {
    const Real n2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    if (!finite(q) || !(n2 > 1.0e-28)) fail(1, "orientation", "quaternion is non-finite or near zero");
    const Real inv = 1.0 / sqrt(n2);
    Q r{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
    if (r.w < 0.0 || (r.w == 0.0 && (r.x < 0.0 || (r.x == 0.0 && (r.y < 0.0 || (r.y == 0.0 && r.z < 0.0)))))) {
        r = {-r.w, -r.x, -r.y, -r.z};
    }
    return r;
}
Q operator*(const Q& a, const Q& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
M3 rotation(const Q& orientation) {
    const Q q = canonical(orientation);
    const Real xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z, xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const Real wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3 r;
    const Real v[9] = {1.0 - 2.0 * (yy + zz), 2.0 * (xy - wz),       2.0 * (xz + wy),
                         2.0 * (xy + wz),       1.0 - 2.0 * (xx + zz), 2.0 * (yz - wx),
                         2.0 * (xz - wy),       2.0 * (yz + wx),       1.0 - 2.0 * (xx + yy)};
    std::memcpy(r.a, v, sizeof(v));
    return r;
}
Q increment(const V& v) {
    if (!finite(v)) fail(5, "rotation-increment", "non-finite rotation vector");
    const Real angle = length(v);
    if (angle < 1.0e-12) return canonical({1.0, 0.5 * v.x, 0.5 * v.y, 0.5 * v.z});
    const Real half = 0.5 * angle;
    const Real s = mSin(half) / angle;
    return {mCos(half), v.x * s, v.y * s, v.z * s};
}

struct Node {
    V position, previous, velocity, force;
    Real w = 0.0;
};
struct Constraint {
    std::size_t a, b;
    Real rest, compliance, lambda = 0.0;
    int kind;
};
struct Membrane {
    std::size_t a, b, c;
    Real m00, m01, m10, m11, area;
    Sym compliance;
    Real preW, preT, dampingTime;
    V lambda, resultant;
    Real residual = 0.0;
    Sym stiffness;      // material stiffness matrix (membrane.cpp:259-266)
    Real ratio = 1.0;   // compressionStiffnessRatio
};
struct Hinge {
    std::size_t a, b, c, d;
    Real rest, compliance, lambda = 0.0;
};
struct Segment {
    std::string id;
    long from, to;  // node index, or -(hang + 1)
    Real rest, stiffness, lambda = 0.0;
    std::size_t definition = 0;  // index in the definition's segment arrays
    Real damping = 0.0;          // axial damping, N s/m
};
struct Pair {
    std::size_t node, centre;
    Real radius, compliance, lambda = 0.0;
};
struct Payload {
    V com, v, w, comLocal, prevCom;
    Q q, prevQ;
    Real mass = 0.0;
    M3 inertia;
    std::vector<V> hang;
    std::vector<std::size_t> hangDefinition;  // string-id slot -> definition index
    Real minLength = 1e-10, maxResidual = 2e-4, certified = 0.0;
    int iterations = 12;
};

}  // namespace

struct oracle_body {
    std::vector<Node> nodes;
    std::vector<std::array<std::size_t, 3>> triangles;
    std::vector<Real> pressure;
    std::vector<Constraint> constraints;
    std::vector<Membrane> membranes;
    std::vector<Hinge> hinges;              // colour-major insertion order
    std::vector<std::size_t> hingeOrder;    // insertion slot -> definition index
    std::vector<Pair> pairs;
    bool hasSuspension = false;
    Payload payload;
    std::vector<Segment> segments;          // string-id order
    oracle::Colouring distance, membrane, dihedral;
    std::vector<std::vector<std::size_t>> incidence;
    std::vector<std::size_t> loadPath;
    std::vector<V> external;
    int perturbUlps = 0;
    bool portableMath = false;
};

namespace {

// soft_body.cpp:1598-1640
void projectDistance(oracle_body& b, Constraint& c, Real invDt2) {
    V& pa = b.nodes[c.a].position;
    V& pb = b.nodes[c.b].position;
    const Real wa = b.nodes[c.a].w, wb = b.nodes[c.b].w;
    const V d = pb - pa;
    const Real L = length(d);
    if (L <= 1.0e-12) return;
    const Real value = L - c.rest;
    if (c.kind == SW_CONSTRAINT_CABLE && value <= 0.0 && c.lambda == 0.0) return;
    const Real wsum = wa + wb;
    const Real alpha = c.compliance * invDt2;
    if (wsum + alpha <= 0.0) return;
    const Real dl = (-value - alpha * c.lambda) / (wsum + alpha);
    Real applied = dl;
    if (c.kind == SW_CONSTRAINT_CABLE) {
        const Real next = std::min<Real>(0.0, c.lambda + dl);
        applied = next - c.lambda;
        c.lambda = next;
    } else {
        c.lambda += dl;
    }
    const Real s = applied / L;
    pa -= (wa * s) * d;
    pb += (wb * s) * d;
}

struct MemEval {
    V strain;
    V g[3][3];
};

// soft_body.cpp:204-282 (StrainOnly)
MemEval evaluate(const oracle_body& b, const Membrane& m) {
    const V p0 = b.nodes[m.a].position, p1 = b.nodes[m.b].position, p2 = b.nodes[m.c].position;
    const V e1 = p1 - p0, e2 = p2 - p0;
    const V fw = m.m00 * e1 + m.m10 * e2;
    const V ft = m.m01 * e1 + m.m11 * e2;
    MemEval r;
    r.strain = {0.5 * (dot(fw, fw) - 1.0) + m.preW, 0.5 * (dot(ft, ft) - 1.0) + m.preT, dot(fw, ft)};
    const V byWarp[3] = {fw, V{}, ft};
    const V byWeft[3] = {V{}, ft, fw};
    for (int k = 0; k < 3; ++k) {
        const V g1 = m.m00 * byWarp[k] + m.m01 * byWeft[k];
        const V g2 = m.m10 * byWarp[k] + m.m11 * byWeft[k];
        r.g[k][0] = -(g1 + g2);
        r.g[k][1] = g1;
        r.g[k][2] = g2;
    }
    return r;
}

// effectiveMembraneStiffness (membrane.cpp:273-296) inverted by
// checkedInverse (membrane.cpp:208-235).
Sym stateCompliance(const Membrane& m, const V& greenStrain) {
    const Sym& k = m.stiffness;
    const Real retained = sqrt(m.ratio);
    const Real warp = greenStrain.x < 0.0 ? retained : Real(1.0);
    const Real weft = greenStrain.y < 0.0 ? retained : Real(1.0);
    const Real shear = sqrt(warp * weft);
    const Sym effective{k.xx * warp * warp, k.yy * weft * weft, k.zz * shear * shear,
                        k.xy * warp * weft, k.xz * warp * shear, k.yz * weft * shear};
    Sym inverse;
    if (!symInverse(effective, inverse)) throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned");
    return inverse;
}

// soft_body.cpp:1822-1929
void projectMembrane(oracle_body& b, Membrane& m, Real dt) {
    const MemEval e = evaluate(b, m);
    const std::size_t idx[3] = {m.a, m.b, m.c};
    const Real s = sqrt(m.area);
    V J[3][3];
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) J[k][c] = e.g[k][c] * s;
    Sym M;
    for (int c = 0; c < 3; ++c) {
        const Real w = b.nodes[idx[c]].w;
        M.xx += w * dot(J[0][c], J[0][c]);
        M.yy += w * dot(J[1][c], J[1][c]);
        M.zz += w * dot(J[2][c], J[2][c]);
        M.xy += w * dot(J[0][c], J[1][c]);
        M.xz += w * dot(J[0][c], J[2][c]);
        M.yz += w * dot(J[1][c], J[2][c]);
    }
    V dm;
    for (int c = 0; c < 3; ++c) {
        const V disp = b.nodes[idx[c]].position - b.nodes[idx[c]].previous;
        dm.x += dot(J[0][c], disp);
        dm.y += dot(J[1][c], disp);
        dm.z += dot(J[2][c], disp);
    }
    const Real invDt2 = 1.0 / (dt * dt);
    // A compression ratio below 1 softens the compressed material directions
    // for this projection (soft_body.cpp:1869-1873).
    const Sym at = invDt2 * (m.ratio == 1.0 ? m.compliance : stateCompliance(m, e.strain));
    const Real gamma = m.dampingTime / dt;
    const Sym S = (1.0 + gamma) * M + at;
    const V rhs = -(s * e.strain) - at * m.lambda - gamma * dm;
    const V dl = checkedSolve(S, rhs);
    V corr[3];
    for (int c = 0; c < 3; ++c)
        corr[c] = b.nodes[idx[c]].w * (J[0][c] * dl.x + J[1][c] * dl.y + J[2][c] * dl.z);
    m.lambda += dl;
    for (int c = 0; c < 3; ++c) b.nodes[idx[c]].position += corr[c];
}

Real perturbed(Real v, int ulps) {
    for (int k = 0; k < ulps; ++k) v = nextafter(v, 1e300);
    for (int k = 0; k < -ulps; ++k) v = nextafter(v, -1e300);
    return v;
}

// soft_body.cpp:1648-1711 and 1754-1792
void projectHinge(oracle_body& b, Hinge& h, Real invDt2) {
    const std::size_t idx[4] = {h.a, h.b, h.c, h.d};
    const V a = b.nodes[h.a].position, bb = b.nodes[h.b].position, c = b.nodes[h.c].position,
            d = b.nodes[h.d].position;
    const V edge = bb - a, toC = c - a, toD = d - a;
    const V area1 = cross(edge, toC), area2 = cross(toD, edge);
    const Real el = length(edge), l1 = length(area1), l2 = length(area2);
    if (!(el > 1.0e-12) || !(l1 > 1.0e-12 * el) || !(l2 > 1.0e-12 * el)) return;
    const V eu = edge / el, n1 = area1 / l1, n2 = area2 / l2;
    const Real cosine = std::clamp<Real>(dot(n1, n2), -1.0, 1.0);
    const Real sine = dot(eu, cross(n1, n2));
    const V n1a = cosine * cross(n2, eu) - sine * n2;
    const V n2a = cosine * cross(eu, n1) - sine * n1;
    const V eua = cosine * cross(n1, n2);
    const V a1a = (n1a - dot(n1a, n1) * n1) / l1;
    const V a2a = (n2a - dot(n2a, n2) * n2) / l2;
    V ea = (eua - dot(eua, eu) * eu) / el;
    ea += cross(toC, a1a);
    const V ca = cross(a1a, edge);
    ea += cross(a2a, toD);
    const V da = cross(edge, a2a);
    Real angle = mAtan2(sine, cosine);
    if (b.perturbUlps) angle = perturbed(angle, b.perturbUlps);
    const V g[4] = {-1.0 * (ea + ca + da), ea, ca, da};
    if (!isfinite(angle)) return;
    for (const V& v : g)
        if (!finite(v)) return;
    Real W = 0.0;
    for (int k = 0; k < 4; ++k) W += b.nodes[idx[k]].w * dot(g[k], g[k]);
    const Real alpha = h.compliance * invDt2;
    const Real den = W + alpha;
    if (!(den > 0.0) || !isfinite(den)) return;
    const Real value = remainder(angle - h.rest, 6.283185307179586476925286766559);
    const Real dl = (-value - alpha * h.lambda) / den;
    if (!isfinite(dl)) return;
    for (int k = 0; k < 4; ++k) b.nodes[idx[k]].position += b.nodes[idx[k]].w * dl * g[k];
    h.lambda += dl;
}

M3 worldInertia(const Payload& p, const Q& q) {
    const M3 r = rotation(q);
    return r * p.inertia * transpose(r);
}

V hangPosition(const Payload& p, std::size_t h) { return p.com + rotation(p.q) * (p.hang[h] - p.comLocal); }

// suspension.cpp:1075-1102 within beginSubstep (suspension.cpp:2072-2182)
void beginPayload(Payload& p, Real dt, const V& gravity) {
    p.prevCom = p.com;
    p.prevQ = p.q;
    const V force = V{} + p.mass * gravity;
    const V moment{};
    if (!(dt > 0.0) || !isfinite(dt) || !finite(force) || !finite(moment))
        fail(5, "payload-step", "invalid payload integration input");
    p.q = canonical(p.q);
    const M3 inertia = worldInertia(p, p.q);
    V angular = inertia * p.w;
    p.v += force * (dt / p.mass);
    angular += moment * dt;
    p.w = inverse3(inertia) * angular;
    p.com += p.v * dt;
    p.q = canonical(increment(p.w * dt) * p.q);
    p.w = inverse3(worldInertia(p, p.q)) * angular;
    if (!finite(p.com) || !finite(p.v) || !finite(p.w) || !finite(p.q))
        fail(5, "payload-step", "payload integration produced non-finite state");
}

struct EndState {
    V position, offset;
    Real w = 0.0;
    bool rigid = false, fixed = false;
};

// suspension.cpp:1104-1123
Real effective(const EndState& e, bool rigid, const V& g, const M3& iinv) {
    bool ok = e.w >= 0.0 && isfinite(e.w) && finite(g) && finite(e.offset);
    for (Real v : iinv.a) ok = ok && isfinite(v);
    if (!ok) fail(6, "effective-mass", "non-finite generalized endpoint");
    Real r = e.w;
    if (rigid && e.w > 0.0) {
        const V rot = cross(e.offset, g);
        r += dot(rot, iinv * rot);
    }
    if (!(r >= 0.0) || !isfinite(r)) fail(6, "effective-mass", "invalid generalized effective inverse mass");
    return r;
}

// suspension.cpp:2197-2274 with projectUnilateralLine (1125-1163)
void lineIteration(oracle_body& b, Real dt) {
    Payload& p = b.payload;
    const M3 iinv = inverse3(worldInertia(p, p.q));
    const auto state = [&](long e) {
        EndState s;
        if (e >= 0) {
            s.position = b.nodes[e].position;
            s.w = b.nodes[e].w;
            s.fixed = s.w == 0.0;
        } else {
            s.position = hangPosition(p, static_cast<std::size_t>(-e - 1));
            s.offset = s.position - p.com;
            s.rigid = true;
            s.fixed = false;
            s.w = 1.0 / p.mass;
        }
        return s;
    };
    const auto apply = [&](long e, const EndState& s, const V& g, Real inc) {
        if (s.fixed) return;
        if (e >= 0) {
            b.nodes[e].position += s.w * inc * g;
        } else {
            p.com += s.w * inc * g;
            const V rot = iinv * cross(s.offset, g) * inc;
            p.q = canonical(increment(rot) * p.q);
        }
    };
    for (Segment& seg : b.segments) {
        const EndState from = state(seg.from), to = state(seg.to);
        const V diff = to.position - from.position;
        const Real L = length(diff);
        if (!isfinite(L) || L <= p.minLength) fail(6, seg.id, "line geometry is non-finite or scale-degenerate");
        const V dir = diff / L;
        if (!(L > 0.0) || !isfinite(L) || !(seg.rest > 0.0) || !isfinite(seg.rest) ||
            !(seg.stiffness > 0.0) || !isfinite(seg.stiffness) || !(dt > 0.0) || !isfinite(dt) ||
            !(seg.lambda >= 0.0) || !isfinite(seg.lambda) || abs(length(dir) - 1.0) > 1.0e-10)
            fail(6, "line-projection", "invalid unilateral line projection input");
        const Real eff = effective(from, from.rigid && !from.fixed, dir, iinv) +
                           effective(to, to.rigid && !to.fixed, -dir, iinv);
        const Real at = 1.0 / (seg.stiffness * dt * dt);
        if (!(eff + at > 0.0)) fail(6, "line-projection", "zero generalized effective mass");
        const Real gap = seg.rest - L;
        const Real candidate = seg.lambda + (-gap - at * seg.lambda) / (eff + at);
        const Real next = std::max<Real>(0.0, candidate);
        const Real inc = next - seg.lambda;
        seg.lambda = next;
        apply(seg.from, from, dir, inc);
        apply(seg.to, to, -dir, inc);
    }
}

// Build-defined contact spheres: node held off the (unmoved) centre node.
void projectPair(oracle_body& b, Pair& pr, Real invDt2) {
    const V c = b.nodes[pr.centre].position;
    V& x = b.nodes[pr.node].position;
    const V d = x - c;
    const Real L = length(d);
    if (L <= 1.0e-12) return;
    const Real gap = L - pr.radius;
    if (gap >= 0.0 && pr.lambda == 0.0) return;
    const Real w = b.nodes[pr.node].w;
    const Real alpha = pr.compliance * invDt2;
    if (w + alpha <= 0.0) return;
    const Real dl = (-gap - alpha * pr.lambda) / (w + alpha);
    const Real next = std::max<Real>(0.0, pr.lambda + dl);
    const Real applied = next - pr.lambda;
    pr.lambda = next;
    const Real s = applied / L;
    x += (w * s) * d;
}

// soft_body.cpp:528-629
std::vector<std::size_t> loadPathOrder(const oracle_body& b) {
    const std::size_t n = b.nodes.size();
    std::vector<std::vector<std::size_t>> adj(n);
    std::vector<char> structural(n, 0);
    std::vector<std::size_t> order;
    for (std::size_t i = 0; i < n; ++i) structural[i] = b.nodes[i].w == 0.0;
    for (const auto& t : b.triangles)
        for (auto v : t) structural[v] = 1;
    for (std::size_t k = 0; k < b.constraints.size(); ++k) {
        const Constraint& c = b.constraints[k];
        if (c.kind == SW_CONSTRAINT_DISTANCE || c.kind == SW_CONSTRAINT_SEAM_STITCH) {
            structural[c.a] = structural[c.b] = 1;
            continue;
        }
        order.push_back(k);
        adj[c.a].push_back(c.b);
        adj[c.b].push_back(c.a);
    }
    for (const Hinge& h : b.hinges) structural[h.a] = structural[h.b] = structural[h.c] = structural[h.d] = 1;
    const std::size_t unknown = std::numeric_limits<std::size_t>::max();
    std::vector<std::size_t> depth(n, unknown);
    std::queue<std::size_t> pending;
    for (std::size_t i = 0; i < n; ++i)
        if (structural[i] && !adj[i].empty()) {
            depth[i] = 0;
            pending.push(i);
        }
    const auto visit = [&] {
        while (!pending.empty()) {
            const std::size_t i = pending.front();
            pending.pop();
            for (auto j : adj[i])
                if (depth[j] == unknown) {
                    depth[j] = depth[i] + 1;
                    pending.push(j);
                }
        }
    };
    visit();
    for (;;) {
        std::size_t root = n;
        Real smallest = std::numeric_limits<double>::infinity();
        for (std::size_t i = 0; i < n; ++i)
            if (depth[i] == unknown && !adj[i].empty() && b.nodes[i].w < smallest) {
                root = i;
                smallest = b.nodes[i].w;
            }
        if (root == n) break;
        depth[root] = 0;
        pending.push(root);
        visit();
    }
    std::stable_sort(order.begin(), order.end(), [&](std::size_t l, std::size_t r) {
        return std::max(depth[b.constraints[l].a], depth[b.constraints[l].b]) <
               std::max(depth[b.constraints[r].a], depth[b.constraints[r].b]);
    });
    return order;
}

// soft_body.cpp:118-150
struct Momentum {
    Real mass = 0.0;
    V centre, linear, angular;
};
Momentum measureMomentum(const std::vector<Node>& nodes) {
    Momentum r;
    for (const Node& n : nodes) {
        if (n.w <= 0.0) continue;
        const Real m = 1.0 / n.w;
        r.mass += m;
        r.centre = r.centre + m * n.position;
        r.linear = r.linear + m * n.velocity;
    }
    if (!(r.mass > 0.0)) return r;
    r.centre = r.centre / r.mass;
    for (const Node& n : nodes) {
        if (n.w <= 0.0) continue;
        const Real m = 1.0 / n.w;
        r.angular = r.angular + m * cross(n.position - r.centre, n.velocity);
    }
    return r;
}

// soft_body.cpp:152-190
void restoreMomentum(std::vector<Node>& nodes, const Momentum& target) {
    if (!(target.mass > 0.0)) return;
    Momentum current = measureMomentum(nodes);
    const V translation = (target.linear - current.linear) / target.mass;
    for (Node& n : nodes)
        if (n.w > 0.0) n.velocity = n.velocity + translation;
    current = measureMomentum(nodes);
    Sym inertia;
    for (const Node& n : nodes) {
        if (n.w <= 0.0) continue;
        const Real m = 1.0 / n.w;
        const V r = n.position - current.centre;
        inertia.xx += m * (r.y * r.y + r.z * r.z);
        inertia.yy += m * (r.x * r.x + r.z * r.z);
        inertia.zz += m * (r.x * r.x + r.y * r.y);
        inertia.xy -= m * r.x * r.y;
        inertia.xz -= m * r.x * r.z;
        inertia.yz -= m * r.y * r.z;
    }
    if (!symPositiveDefinite(inertia)) return;
    const V omega = checkedSolve(inertia, target.angular - current.angular);
    for (Node& n : nodes)
        if (n.w > 0.0) n.velocity = n.velocity + cross(omega, n.position - current.centre);
}

// suspension.cpp:1165-1185
Real boundedAxialDampingImpulse(Real relative, Real damping, Real dt, Real effective) {
    if (!isfinite(relative) || !(damping >= 0.0) || !isfinite(damping) || !(dt > 0.0) || !isfinite(dt) ||
        !(effective >= 0.0) || !isfinite(effective))
        fail(6, "line-damping", "invalid axial damping input");
    if (damping == 0.0 || effective == 0.0 || relative == 0.0) return 0.0;
    const Real requested = damping * abs(relative) * dt;
    const Real stopping = abs(relative) / effective;
    const Real bounded = std::min(requested, stopping);  // >= 0; copysign onto a nonzero `relative`
    return relative < 0.0 ? -bounded : bounded;
}

// suspension.cpp:2413-2479 (free payload): axial damping impulses along every
// taut, damped segment in runtime (string-id) order.
void lineDamping(oracle_body& b, Real dt) {
    bool any = false;
    for (const Segment& s : b.segments) any = any || (s.lambda > 0.0 && s.damping > 0.0);
    if (!any) return;
    Payload& p = b.payload;
    const M3 inverseInertia = inverse3(worldInertia(p, p.q));
    const auto offsetOf = [&](long e) { return rotation(p.q) * (p.hang[-e - 1] - p.comLocal); };
    const auto pointVelocity = [&](long e) {
        if (e >= 0) return b.nodes[e].velocity;
        return p.v + cross(p.w, offsetOf(e));
    };
    const auto apply = [&](long e, const V& point, const V& impulse) {
        if (e >= 0) {
            Node& n = b.nodes[e];
            if (n.w > 0.0) n.velocity = n.velocity + impulse * n.w;
        } else {
            p.v = p.v + impulse / p.mass;
            p.w = p.w + inverseInertia * cross(point - p.com, impulse);
        }
    };
    for (const Segment& s : b.segments) {
        if (!(s.lambda > 0.0) || !(s.damping > 0.0)) continue;
        const V from = s.from >= 0 ? b.nodes[s.from].position : hangPosition(p, -s.from - 1);
        const V to = s.to >= 0 ? b.nodes[s.to].position : hangPosition(p, -s.to - 1);
        const V d = to - from;
        const Real magnitude = length(d);
        const V direction = magnitude > 0.0 ? d / magnitude : V{};
        const Real relative = dot(pointVelocity(s.to) - pointVelocity(s.from), direction);
        Real effective = 0.0;
        if (s.from >= 0) {
            effective += b.nodes[s.from].w;
        } else {
            const V rxn = cross(from - p.com, direction);
            effective += 1.0 / p.mass + dot(rxn, inverseInertia * rxn);
        }
        if (s.to >= 0) {
            effective += b.nodes[s.to].w;
        } else {
            const V rxn = cross(to - p.com, direction);
            effective += 1.0 / p.mass + dot(rxn, inverseInertia * rxn);
        }
        if (!(effective > 0.0) || relative == 0.0) continue;
        const Real impulse = boundedAxialDampingImpulse(relative, s.damping, dt, effective);
        const V fromImpulse = direction * impulse;
        apply(s.from, from, fromImpulse);
        apply(s.to, to, -fromImpulse);
    }
}

// soft_body.cpp:1175-1460 (coupled; the coloured sweep for workerThreads >= 1,
// the serial definition-order sweep for workerThreads = 0) for one substep.
void substep(oracle_body& b, Real dt, const sw_step_settings& st) {
    // Pressure gather in ascending triangle order (soft_body.cpp:1492-1534).
    {
        std::vector<V> total(b.nodes.size());
        for (std::size_t i = 0; i < b.nodes.size(); ++i) {
            V f = b.nodes[i].force;
            for (std::size_t t : b.incidence[i]) {
                const auto& tri = b.triangles[t];
                const V& pa = b.nodes[tri[0]].position;
                const V& pb = b.nodes[tri[1]].position;
                const V& pc = b.nodes[tri[2]].position;
                const V areaVector = 0.5 * cross(pb - pa, pc - pa);
                f += b.pressure[t] * areaVector / 3.0;
            }
            total[i] = f;
        }
        for (std::size_t i = 0; i < b.nodes.size(); ++i) b.nodes[i].force = total[i];
    }
    const V gravity{st.gravity[0], st.gravity[1], st.gravity[2]};
    if (b.hasSuspension) beginPayload(b.payload, dt, gravity);
    const Real damping = exp(-st.velocity_damping * dt);
    const V ref{st.damping_reference_velocity[0], st.damping_reference_velocity[1], st.damping_reference_velocity[2]};
    for (Node& n : b.nodes) {
        n.previous = n.position;
        if (n.w == 0.0) {
            n.velocity = {};
            continue;
        }
        const V acc = gravity + n.force * n.w;
        n.velocity = ref + (n.velocity + acc * dt - ref) * damping;
        n.position += n.velocity * dt;
    }
    // soft_body.cpp:1193-1205: an isolated free membrane keeps its predicted
    // momentum (contact spheres stand in for the reference's contact pairs).
    bool preserveMomentum = !b.hasSuspension && (!b.membranes.empty() || !b.hinges.empty()) && b.pairs.empty();
    for (const Node& n : b.nodes) preserveMomentum = preserveMomentum && n.w > 0.0;
    const Momentum predicted = preserveMomentum ? measureMomentum(b.nodes) : Momentum{};
    for (auto& c : b.constraints) c.lambda = 0.0;
    for (auto& m : b.membranes) {
        m.lambda = {};
        m.resultant = {};
        m.residual = 0.0;
    }
    for (auto& h : b.hinges) h.lambda = 0.0;
    for (auto& s : b.segments) s.lambda = 0.0;
    for (auto& p : b.pairs) p.lambda = 0.0;

    const Real invDt2 = 1.0 / (dt * dt);
    const auto cablePair = [&] {
        for (auto it = b.loadPath.rbegin(); it != b.loadPath.rend(); ++it) projectDistance(b, b.constraints[*it], invDt2);
        for (auto k : b.loadPath) projectDistance(b, b.constraints[k], invDt2);
    };
    const int pairs = st.cable_sweep_pairs;
    const int closing = pairs > 0 ? pairs - pairs / 4 : 0;
    const int interleaved = pairs - closing;
    const int iters = st.constraint_iterations;
    for (int it = 0; it < iters; ++it) {
        const int share = interleaved / std::max(1, iters) + (it < interleaved % std::max(1, iters) ? 1 : 0);
        for (int k = 0; k < share; ++k) cablePair();
        if (st.worker_threads == 0) {
            // workerThreads = 0: the serial sweep in definition order
            // (soft_body.cpp:1326-1335, 1347-1351).
            for (auto& c : b.constraints) projectDistance(b, c, invDt2);
            for (auto& m : b.membranes) projectMembrane(b, m, dt);
        } else {
            for (auto k : b.distance.order) projectDistance(b, b.constraints[k], invDt2);
            for (auto k : b.membrane.order) projectMembrane(b, b.membranes[k], dt);
        }
        for (Hinge& h : b.hinges) projectHinge(b, h, invDt2);
        if (b.hasSuspension) lineIteration(b, dt);
        for (Pair& p : b.pairs) projectPair(b, p, invDt2);
    }
    if (iters > 0) {
        for (int k = 0; k < closing; ++k) cablePair();
    } else {
        for (int k = 0; k < pairs; ++k) cablePair();
    }
    if (iters > 0)
        for (auto& c : b.constraints)
            if (c.kind == SW_CONSTRAINT_SEAM_STITCH) projectDistance(b, c, invDt2);
    // soft_body.cpp:2366-2404
    for (Membrane& m : b.membranes) {
        const MemEval e = evaluate(b, m);
        const std::size_t idx[3] = {m.a, m.b, m.c};
        const Real s = sqrt(m.area);
        const Sym at = invDt2 * m.compliance;
        const Real gamma = m.dampingTime / dt;
        V dm;
        for (int c = 0; c < 3; ++c) {
            const V disp = b.nodes[idx[c]].position - b.nodes[idx[c]].previous;
            dm.x += s * dot(e.g[0][c], disp);
            dm.y += s * dot(e.g[1][c], disp);
            dm.z += s * dot(e.g[2][c], disp);
        }
        const V residual = s * e.strain + at * m.lambda + gamma * dm;
        m.residual = length(residual) / s;
        m.resultant = -m.lambda * (invDt2 / s);
    }
    const Real invDt = 1.0 / dt;
    for (Node& n : b.nodes) n.velocity = n.w == 0.0 ? V{} : (n.position - n.previous) * invDt;
    if (preserveMomentum) restoreMomentum(b.nodes, predicted);  // soft_body.cpp:1446-1448
    if (b.hasSuspension) {
        Payload& p = b.payload;
        // suspension.cpp:2386-2411
        p.v = (p.com - p.prevCom) / dt;
        Q delta = p.q * Q{p.prevQ.w, -p.prevQ.x, -p.prevQ.y, -p.prevQ.z};
        delta = canonical(delta);
        const Real half = mAcos(std::clamp<Real>(delta.w, -1.0, 1.0));
        const Real sine = mSin(half);
        if (abs(sine) < 1.0e-12) p.w = V{delta.x, delta.y, delta.z} * (2.0 / dt);
        else p.w = V{delta.x / sine, delta.y / sine, delta.z / sine} * (2.0 * half / dt);
        lineDamping(b, dt);
        // suspension.cpp:2324-2384
        const Real n2 = p.q.w * p.q.w + p.q.x * p.q.x + p.q.y * p.q.y + p.q.z * p.q.z;
        if (!finite(p.com) || !finite(p.q) || !finite(p.v) || !finite(p.w) || abs(n2 - 1.0) > 1.0e-12)
            fail(8, "payload", "payload trial state is non-finite or unnormalized");
        Real worst = 0.0;
        for (const Segment& s : b.segments) {
            const V from = s.from >= 0 ? b.nodes[s.from].position : hangPosition(p, -s.from - 1);
            const V to = s.to >= 0 ? b.nodes[s.to].position : hangPosition(p, -s.to - 1);
            const Real L = length(to - from);
            if (!isfinite(L)) fail(8, s.id, "line trial length is non-finite");
            const Real stretch = std::max<Real>(0.0, L - s.rest);
            const Real compliant = s.lambda / (s.stiffness * dt * dt);
            worst = std::max<Real>(worst, abs(stretch - compliant));
        }
        p.certified = worst;
        if (worst > p.maxResidual) fail(8, "line-residual", "maximum line residual exceeds configured bound");
    }
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const SuspensionFail& e) {
        lastError = std::string(e.what()) + " [" + e.entity + "]";
        return SW_SUSPENSION_ERROR;
    } catch (const std::out_of_range& e) {
        lastError = e.what();
        return SW_OUT_OF_RANGE;
    } catch (const std::invalid_argument& e) {
        lastError = e.what();
        return SW_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        lastError = e.what();
        return SW_LOGIC_ERROR;
    }
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return lastError.c_str(); }

int oracle_create(const sw_definition* d, oracle_body** out) {
    return guard([&] {
        auto b = std::make_unique<oracle_body>();
        b->nodes.resize(d->node_count);
        for (std::size_t i = 0; i < d->node_count; ++i) {
            Node& n = b->nodes[i];
            n.position = {d->node_position[3 * i], d->node_position[3 * i + 1], d->node_position[3 * i + 2]};
            n.previous = n.position;
            n.w = d->node_fixed[i] ? 0.0 : 1.0 / d->node_mass[i];
        }
        b->triangles.resize(d->triangle_count);
        b->incidence.assign(d->node_count, {});
        for (std::size_t t = 0; t < d->triangle_count; ++t) {
            for (int k = 0; k < 3; ++k) b->triangles[t][k] = d->triangle_nodes[3 * t + k];
            for (int k = 0; k < 3; ++k) b->incidence[b->triangles[t][k]].push_back(t);
        }
        b->pressure.assign(d->triangle_count, 0.0);
        std::vector<std::size_t> pairs;
        for (std::size_t c = 0; c < d->constraint_count; ++c) {
            b->constraints.push_back({d->constraint_nodes[2 * c], d->constraint_nodes[2 * c + 1], d->constraint_rest[c],
                                      d->constraint_compliance[c], 0.0, d->constraint_kind[c]});
            pairs.push_back(d->constraint_nodes[2 * c]);
            pairs.push_back(d->constraint_nodes[2 * c + 1]);
        }
        b->distance = oracle::distanceColouring(pairs, d->node_count);
        std::vector<std::size_t> corners;
        for (std::size_t m = 0; m < d->membrane_count; ++m) {
            const std::size_t t = d->membrane_triangle[m];
            const double* chart = &d->membrane_chart[6 * m];
            // soft_body.cpp:663-690: reference matrix, area, inverse, compliance
            const Real r00 = chart[2] - chart[0], r01 = chart[4] - chart[0];
            const Real r10 = chart[3] - chart[1], r11 = chart[5] - chart[1];
            const Real det = r00 * r11 - r01 * r10;
            const Real scale = std::max({abs(r00), abs(r01), abs(r10), abs(r11)});
            if (!(scale > 0.0) || !isfinite(det) ||
                abs(det) <= 64.0 * std::numeric_limits<double>::epsilon() * scale * scale)
                throw std::invalid_argument("Matrix2 is singular or ill-conditioned");
            const Real inv = 1.0 / det;
            const double* mat = &d->membrane_material[8 * m];
            Sym C;
            if (!symInverse(Sym{mat[0], mat[1], mat[3], mat[2], 0.0, 0.0}, C))
                throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned");
            Membrane mem{d->triangle_nodes[3 * t],
                         d->triangle_nodes[3 * t + 1],
                         d->triangle_nodes[3 * t + 2],
                         r11 * inv,
                         -r01 * inv,
                         -r10 * inv,
                         r00 * inv,
                         0.5 * det,
                         C,
                         mat[4],
                         mat[5],
                         mat[6],
                         {},
                         {},
                         0.0};
            mem.stiffness = Sym{mat[0], mat[1], mat[3], mat[2], 0.0, 0.0};
            mem.ratio = mat[7];
            b->membranes.push_back(mem);
            corners.insert(corners.end(), {mem.a, mem.b, mem.c});
        }
        b->membrane = oracle::membraneColouring(corners, d->node_count);
        std::vector<std::size_t> hinges(d->dihedral_nodes, d->dihedral_nodes + 4 * d->dihedral_count);
        b->dihedral = oracle::dihedralColouring(hinges, d->node_count);
        for (std::size_t k : b->dihedral.order) {
            b->hinges.push_back({hinges[4 * k], hinges[4 * k + 1], hinges[4 * k + 2], hinges[4 * k + 3],
                                 d->dihedral_rest_angle[k], d->dihedral_compliance[k], 0.0});
            b->hingeOrder.push_back(k);
        }
        for (std::size_t s = 0; s < d->sphere_count; ++s)
            for (std::uint64_t j = d->sphere_offsets[s]; j < d->sphere_offsets[s + 1]; ++j)
                b->pairs.push_back({d->sphere_nodes[j], d->sphere_centre[s], d->sphere_radius[s], d->sphere_compliance[s]});
        if (d->suspension) {
            const sw_suspension_definition& s = *d->suspension;
            b->hasSuspension = true;
            Payload& p = b->payload;
            p.mass = s.pilot_mass;
            p.com = {s.pilot_com_world[0], s.pilot_com_world[1], s.pilot_com_world[2]};
            p.v = {s.pilot_velocity[0], s.pilot_velocity[1], s.pilot_velocity[2]};
            p.w = {s.pilot_angular_velocity[0], s.pilot_angular_velocity[1], s.pilot_angular_velocity[2]};
            p.comLocal = {s.pilot_com_local[0], s.pilot_com_local[1], s.pilot_com_local[2]};
            p.q = canonical({s.pilot_orientation[0], s.pilot_orientation[1], s.pilot_orientation[2], s.pilot_orientation[3]});
            p.prevCom = p.com;
            p.prevQ = p.q;
            Real in[9] = {s.pilot_inertia[0], 0, 0, 0, s.pilot_inertia[1], 0, 0, 0, s.pilot_inertia[2]};
            std::memcpy(p.inertia.a, in, sizeof(in));
            p.minLength = s.minimum_line_length;
            p.maxResidual = s.maximum_line_residual;
            p.iterations = s.solver_iterations;
            // String-id normalisation (suspension.cpp:996-1005) and endpoint
            // resolution (structure.cpp:460-552, suspension.cpp:1211-1237).
            std::map<std::string, long> endpoint;
            std::vector<std::pair<std::string, std::size_t>> hangs;
            for (std::size_t k = 0; k < s.harness_count; ++k)
                hangs.emplace_back("harness:" + std::to_string(s.harness_ids[k]), k);
            std::sort(hangs.begin(), hangs.end());
            for (std::size_t r = 0; r < hangs.size(); ++r) {
                endpoint[hangs[r].first] = -static_cast<long>(r) - 1;
                const double* l = &s.harness_local[3 * hangs[r].second];
                p.hang.push_back({l[0], l[1], l[2]});
                p.hangDefinition.push_back(hangs[r].second);
            }
            for (std::size_t k = 0; k < s.attachment_count; ++k)
                endpoint["attachment:" + std::to_string(s.attachment_ids[k])] = static_cast<long>(s.attachment_nodes[k]);
            for (std::size_t k = 0; k < s.junction_count; ++k)
                endpoint["junction:" + std::to_string(s.junction_ids[k])] = static_cast<long>(s.junction_nodes[k]);
            const auto name = [](std::uint8_t kind, std::uint64_t id) {
                return std::string(kind == SW_ENDPOINT_ATTACHMENT ? "attachment:"
                                   : kind == SW_ENDPOINT_JUNCTION ? "junction:"
                                                                  : "harness:") +
                       std::to_string(id);
            };
            for (std::size_t k = 0; k < s.segment_count; ++k)
                b->segments.push_back({"segment:" + std::to_string(s.segment_ids[k]),
                                       endpoint.at(name(s.segment_from_kind[k], s.segment_from_id[k])),
                                       endpoint.at(name(s.segment_to_kind[k], s.segment_to_id[k])), s.segment_rest[k],
                                       s.segment_stiffness[k], 0.0, k, s.segment_damping[k]});
            std::sort(b->segments.begin(), b->segments.end(),
                      [](const Segment& l, const Segment& r) { return l.id < r.id; });
        }
        b->loadPath = loadPathOrder(*b);
        b->external.assign(d->node_count, V{});
        *out = b.release();
    });
}

void oracle_destroy(oracle_body* body) { delete body; }

int oracle_set_face_pressure(oracle_body* b, const double* pressure, size_t count) {
    return guard([&] {
        if (count != b->pressure.size()) throw std::out_of_range("Triangle index is out of range");
        std::copy(pressure, pressure + count, b->pressure.begin());
    });
}

int oracle_set_external_forces(oracle_body* b, const double* forces, size_t count) {
    return guard([&] {
        if (count != b->nodes.size()) throw std::invalid_argument("force count must equal the node count");
        for (std::size_t i = 0; i < count; ++i) b->external[i] = {forces[3 * i], forces[3 * i + 1], forces[3 * i + 2]};
    });
}

void oracle_set_angle_perturbation(oracle_body* b, int ulps) { b->perturbUlps = ulps; }

int oracle_set_tangent(oracle_body* b, const double* dLogCompliance, const double* dRest, const double* dPressure) {
    return guard([&] {
        if constexpr (!kTangentBuild) {
            throw std::logic_error("tangent seeds need the tangent build of the oracle");
        } else {
            for (std::size_t m = 0; m < b->membranes.size(); ++m) {
                const double t = dLogCompliance ? dLogCompliance[m] : 0.0;
                Sym& c = b->membranes[m].compliance;
                for (Real* e : {&c.xx, &c.yy, &c.zz, &c.xy, &c.xz, &c.yz}) seedTangent(*e, value(*e) * t);  // d(C e^theta)
            }
            for (Segment& seg : b->segments) seedTangent(seg.rest, dRest ? dRest[seg.definition] : 0.0);
            for (std::size_t t = 0; t < b->pressure.size(); ++t)
                seedTangent(b->pressure[t], dPressure ? dPressure[t] : 0.0);
        }
    });
}

int oracle_tangent_state(const oracle_body* b, double* dPosition) {
    return guard([&] {
        if constexpr (!kTangentBuild) {
            throw std::logic_error("tangents need the tangent build of the oracle");
        } else {
            for (std::size_t i = 0; i < b->nodes.size(); ++i) {
                dPosition[3 * i] = tangentOf(b->nodes[i].position.x);
                dPosition[3 * i + 1] = tangentOf(b->nodes[i].position.y);
                dPosition[3 * i + 2] = tangentOf(b->nodes[i].position.z);
            }
        }
    });
}

void oracle_set_portable_math(oracle_body* b, int portable) { b->portableMath = portable != 0; }

void oracle_math(int portable, int function, const double* a, const double* b, double* out, size_t count) {
    portableMath = portable != 0;
    for (size_t i = 0; i < count; ++i) {
        switch (function) {
        case 2: out[i] = value(mAtan2(a[i], b[i])); break;
        case 4: out[i] = value(mSin(a[i])); break;
        case 5: out[i] = value(mCos(a[i])); break;
        case 6: out[i] = value(mAcos(a[i])); break;
        default: out[i] = std::nan(""); break;
        }
    }
    portableMath = false;
}

void oracle_quad_math(int function, const double* a, const double* b, double* out, size_t count) {
    for (size_t i = 0; i < count; ++i) {
        const __float128 x = a[i], y = b[i];
        __float128 r;
        switch (function) {
        case 2: r = atan2q(x, y); break;
        case 4: r = sinq(x); break;
        case 5: r = cosq(x); break;
        case 6: r = acosq(x); break;
        default: r = nanq(""); break;
        }
        out[i] = static_cast<double>(r);
    }
}

int oracle_step(oracle_body* b, const sw_step_settings* st) {
    portableMath = b->portableMath;
    return guard([&] {
        if (!(st->time_step > 0.0) || st->substeps <= 0 || st->constraint_iterations < 0)
            throw std::invalid_argument("Invalid SoftBody step settings");
        if (b->hasSuspension && st->constraint_iterations != b->payload.iterations)
            throw std::invalid_argument("Coupled structural iterations must equal lineIterations");
        // Structure::step: the external loads, then stepCoupled which
        // replays them for every substep (soft_body.cpp:1021-1070); a
        // failing substep restores the pre-step state.
        const oracle_body saved = *b;
        try {
            const Real dt = static_cast<Real>(st->time_step) / st->substeps;
            for (int k = 0; k < st->substeps; ++k) {
                for (std::size_t i = 0; i < b->nodes.size(); ++i) b->nodes[i].force = b->external[i];
                substep(*b, dt, *st);
            }
            for (Node& n : b->nodes) n.force = {};
        } catch (...) {
            *b = saved;
            throw;
        }
    });
}

int oracle_state(const oracle_body* b, double* position, double* previous, double* velocity, double* constraintLambda,
                 double* membraneLambda, double* membraneResultant, double* membraneResidual, double* dihedralLambda,
                 double* segmentLambda, double* payload, double* suspensionResidual) {
    return guard([&] {
        for (std::size_t i = 0; i < b->nodes.size(); ++i) {
            const Node& n = b->nodes[i];
            if (position) {
                position[3 * i] = value(n.position.x);
                position[3 * i + 1] = value(n.position.y);
                position[3 * i + 2] = value(n.position.z);
            }
            if (previous) {
                previous[3 * i] = value(n.previous.x);
                previous[3 * i + 1] = value(n.previous.y);
                previous[3 * i + 2] = value(n.previous.z);
            }
            if (velocity) {
                velocity[3 * i] = value(n.velocity.x);
                velocity[3 * i + 1] = value(n.velocity.y);
                velocity[3 * i + 2] = value(n.velocity.z);
            }
        }
        if (constraintLambda)
            for (std::size_t k = 0; k < b->constraints.size(); ++k) constraintLambda[k] = value(b->constraints[k].lambda);
        for (std::size_t k = 0; k < b->membranes.size(); ++k) {
            const Membrane& m = b->membranes[k];
            if (membraneLambda) {
                membraneLambda[3 * k] = value(m.lambda.x);
                membraneLambda[3 * k + 1] = value(m.lambda.y);
                membraneLambda[3 * k + 2] = value(m.lambda.z);
            }
            if (membraneResultant) {
                membraneResultant[3 * k] = value(m.resultant.x);
                membraneResultant[3 * k + 1] = value(m.resultant.y);
                membraneResultant[3 * k + 2] = value(m.resultant.z);
            }
            if (membraneResidual) membraneResidual[k] = value(m.residual);
        }
        if (dihedralLambda)
            for (std::size_t s = 0; s < b->hinges.size(); ++s) dihedralLambda[b->hingeOrder[s]] = value(b->hinges[s].lambda);
        if (segmentLambda)
            for (std::size_t k = 0; k < b->segments.size(); ++k) segmentLambda[k] = value(b->segments[k].lambda);
        if (payload && b->hasSuspension) {
            const Payload& p = b->payload;
            const Real v[13] = {p.com.x, p.com.y, p.com.z, p.q.w, p.q.x, p.q.y, p.q.z,
                                  p.v.x,   p.v.y,   p.v.z,   p.w.x, p.w.y, p.w.z};
            for (int k = 0; k < 13; ++k) payload[k] = value(v[k]);
        }
        if (suspensionResidual) *suspensionResidual = value(b->payload.certified);
    });
}

// Diagnostic-frame fields of the current state, restating
// simwing::viewer::buildStructureFrame (structure_frame.cpp:526-612) over
// Structure::suspensionState (structure.cpp:718-776): vertices are the nodes
// then the harness points in definition order, at payloadPointPosition /
// payloadPointVelocity (suspension.cpp:1018-1032); lines are the constraints
// (length |second - first|, violation per constraintViolation,
// structure_frame.cpp:60-73) then the segments in definition order (length
// between the frame vertices of its ends, violation max(0, L - commanded
// rest), the commanded rest being the definition's rest length). The pending
// force of a node is its current external load, of a harness point zero.
int oracle_frame_fields(const oracle_body* b, double* vertexPosition, double* vertexVelocity, double* pendingForce,
                        double* lineLength, double* lineViolation) {
    return guard([&] {
        const std::size_t n = b->nodes.size();
        const Payload& p = b->payload;
        const std::size_t h = b->hasSuspension ? p.hang.size() : 0;
        const auto put = [](double* out, std::size_t i, const V& v) {
            if (!out) return;
            out[3 * i] = value(v.x);
            out[3 * i + 1] = value(v.y);
            out[3 * i + 2] = value(v.z);
        };
        std::vector<V> vertex(n + h);
        for (std::size_t i = 0; i < n; ++i) {
            vertex[i] = b->nodes[i].position;
            put(vertexVelocity, i, b->nodes[i].velocity);
            put(pendingForce, i, b->external[i]);
        }
        std::vector<std::size_t> slotVertex(h);
        for (std::size_t r = 0; r < h; ++r) {
            const std::size_t i = n + p.hangDefinition[r];
            slotVertex[r] = i;
            const V offset = rotation(p.q) * (p.hang[r] - p.comLocal);
            vertex[i] = p.com + offset;
            put(vertexVelocity, i, p.v + cross(p.w, offset));
            put(pendingForce, i, V{});
        }
        for (std::size_t i = 0; i < n + h; ++i) put(vertexPosition, i, vertex[i]);
        const std::size_t c = b->constraints.size();
        for (std::size_t k = 0; k < c; ++k) {
            const Constraint& q = b->constraints[k];
            const Real l = length(b->nodes[q.b].position - b->nodes[q.a].position);
            const Real extension = l - q.rest;
            if (lineLength) lineLength[k] = value(l);
            if (lineViolation)
                lineViolation[k] = value(q.kind == SW_CONSTRAINT_CABLE ? std::max<Real>(0.0, extension) : abs(extension));
        }
        for (const Segment& g : b->segments) {
            const auto end = [&](long e) { return e >= 0 ? vertex[e] : vertex[slotVertex[-e - 1]]; };
            const Real l = length(end(g.to) - end(g.from));
            if (lineLength) lineLength[c + g.definition] = value(l);
            if (lineViolation) lineViolation[c + g.definition] = value(std::max<Real>(0.0, l - g.rest));
        }
    });
}

int oracle_colouring(const oracle_body* b, int family, uint64_t* order, size_t* orderCount, uint64_t* offsets,
                     size_t* offsetCount) {
    return guard([&] {
        const oracle::Colouring* c = family == 0 ? &b->distance : family == 1 ? &b->membrane : &b->dihedral;
        if (orderCount) *orderCount = c->order.size();
        if (offsetCount) *offsetCount = c->offsets.size();
        if (order) std::copy(c->order.begin(), c->order.end(), order);
        if (offsets) std::copy(c->offsets.begin(), c->offsets.end(), offsets);
    });
}

}  // extern "C"
