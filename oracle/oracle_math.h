// TEST INFRASTRUCTURE ONLY — the CPU oracle for parity checks of the HIP path.
// Nothing in the product (dtrt_amd/, include/) may include or link this code;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Double-precision restatement of the reference's math layer (uci-rendering/dtrt).
// Every function cites the reference file:line whose behaviour it restates.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <limits>

namespace oref {

// fwd.h:10-18 (DOUBLE_PRECISION), config.h:5-7. The reference stores the two
// epsilons as float literals promoted to double; keep the same values.
constexpr double EPS = 1e-7;
constexpr double SHADOW_EPS = (double)1e-3f;
constexpr double ANGLE_EPS = (double)1e-4f;
constexpr double PI = 3.14159265358979323846;
constexpr double INV_PI = 0.31830988618379067154;
constexpr double INV_TWOPI = 0.15915494309189534561;
constexpr double INV_FOURPI = 0.07957747154594766788;
constexpr double INF = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------- vectors
struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) { return a / norm(a); }

struct V2 { double x = 0, y = 0; };

// Spectrum = 3 doubles (fwd.h:89).
struct Spec {
    double c[3] = {0, 0, 0};
    Spec() = default;
    Spec(double a, double b, double d) { c[0] = a; c[1] = b; c[2] = d; }
    static Spec ones() { return Spec(1, 1, 1); }
    double operator[](int i) const { return c[i]; }
    double &operator[](int i) { return c[i]; }
    // Eigen isZero(prec): every |coefficient| <= prec.
    bool isZero(double prec = 1e-12) const {
        return std::abs(c[0]) <= prec && std::abs(c[1]) <= prec && std::abs(c[2]) <= prec;
    }
    double minCoeff() const { return std::min(c[0], std::min(c[1], c[2])); }
};
inline Spec operator+(Spec a, Spec b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline Spec operator-(Spec a, Spec b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline Spec operator*(Spec a, Spec b) { return {a[0] * b[0], a[1] * b[1], a[2] * b[2]}; }
inline Spec operator*(Spec a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline Spec operator*(double s, Spec a) { return a * s; }
inline Spec operator/(Spec a, double s) { return {a[0] / s, a[1] / s, a[2] / s}; }
inline Spec &operator+=(Spec &a, Spec b) { a = a + b; return a; }
inline Spec &operator-=(Spec &a, Spec b) { a = a - b; return a; }
inline Spec &operator*=(Spec &a, Spec b) { a = a * b; return a; }
inline Spec &operator*=(Spec &a, double s) { a = a * s; return a; }

// ------------------------------------------------------------ dual numbers
// SimpleAD::Scalar semantics (simpleAD.h:20-240): value + N tangents, chain rule
// per operation, comparisons on the value only.
template <int N> struct D {
    double v = 0;
    double g[N];
    D() { for (int i = 0; i < N; ++i) g[i] = 0; }
    D(double x) : v(x) { for (int i = 0; i < N; ++i) g[i] = 0; }
    bool gradZero(double prec) const {  // Eigen der.isZero(prec)
        for (int i = 0; i < N; ++i) if (std::abs(g[i]) > prec) return false;
        return true;
    }
    double gradAbsMax() const {
        double m = 0;
        for (int i = 0; i < N; ++i) m = std::max(m, std::abs(g[i]));
        return m;
    }
};
template <int N> inline D<N> operator+(const D<N> &a, const D<N> &b) { D<N> r(a.v + b.v); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i]; return r; }
template <int N> inline D<N> operator-(const D<N> &a, const D<N> &b) { D<N> r(a.v - b.v); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i]; return r; }
template <int N> inline D<N> operator-(const D<N> &a) { D<N> r(-a.v); for (int i = 0; i < N; ++i) r.g[i] = -a.g[i]; return r; }
template <int N> inline D<N> operator*(const D<N> &a, const D<N> &b) { D<N> r(a.v * b.v); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i]; return r; }
template <int N> inline D<N> operator/(const D<N> &a, const D<N> &b) { D<N> r(a.v / b.v); double b2 = b.v * b.v; for (int i = 0; i < N; ++i) r.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / b2; return r; }
template <int N> inline D<N> operator+(const D<N> &a, double s) { D<N> r = a; r.v += s; return r; }
template <int N> inline D<N> operator+(double s, const D<N> &a) { return a + s; }
template <int N> inline D<N> operator-(const D<N> &a, double s) { D<N> r = a; r.v -= s; return r; }
template <int N> inline D<N> operator-(double s, const D<N> &a) { D<N> r = -a; r.v += s; return r; }
template <int N> inline D<N> operator*(const D<N> &a, double s) { D<N> r(a.v * s); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * s; return r; }
template <int N> inline D<N> operator*(double s, const D<N> &a) { return a * s; }
template <int N> inline D<N> operator/(const D<N> &a, double s) { D<N> r(a.v / s); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] / s; return r; }
template <int N> inline D<N> operator/(double s, const D<N> &a) { D<N> r(s / a.v); double a2 = a.v * a.v; for (int i = 0; i < N; ++i) r.g[i] = -s * a.g[i] / a2; return r; }
template <int N> inline D<N> &operator+=(D<N> &a, const D<N> &b) { a = a + b; return a; }
template <int N> inline D<N> &operator-=(D<N> &a, const D<N> &b) { a = a - b; return a; }
template <int N> inline D<N> &operator*=(D<N> &a, const D<N> &b) { a = a * b; return a; }
template <int N> inline D<N> &operator/=(D<N> &a, const D<N> &b) { a = a / b; return a; }
// simpleAD.h:202-230
template <int N> inline D<N> dsqrt(const D<N> &a) { double t = std::sqrt(a.v); D<N> r(t); for (int i = 0; i < N; ++i) r.g[i] = 0.5 * a.g[i] / t; return r; }
template <int N> inline D<N> dexp(const D<N> &a) { double e = std::exp(a.v); D<N> r(e); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * e; return r; }
template <int N> inline D<N> dlog(const D<N> &a) { D<N> r(std::log(a.v)); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] / a.v; return r; }
template <int N> inline D<N> dpow(const D<N> &a, double p) { double t = std::pow(a.v, p - 1.0); D<N> r(a.v * t); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * p * t; return r; }
template <int N> inline D<N> dpow(const D<N> &a, const D<N> &p) { return dexp(p * dlog(a)); }
template <int N> inline D<N> dabs(const D<N> &a) { D<N> r(std::abs(a.v)); double s = a.v > 0.0 ? 1.0 : -1.0; for (int i = 0; i < N; ++i) r.g[i] = s * a.g[i]; return r; }
template <int N> inline D<N> dsquare(const D<N> &a) { D<N> r(a.v * a.v); for (int i = 0; i < N; ++i) r.g[i] = 2.0 * a.v * a.g[i]; return r; }
template <int N> inline D<N> dasin(const D<N> &a) { D<N> r(std::asin(a.v)); double q = std::sqrt(1.0 - a.v * a.v); for (int i = 0; i < N; ++i) r.g[i] = a.g[i] / q; return r; }

// Matrix<3,1> AD (simpleAD.h:740-1040) as three dual components.
template <int N> struct DV {
    D<N> x, y, z;
    DV() = default;
    DV(const V3 &v) : x(v.x), y(v.y), z(v.z) {}
    DV(const D<N> &a, const D<N> &b, const D<N> &c) : x(a), y(b), z(c) {}
    V3 val() const { return V3(x.v, y.v, z.v); }
    V3 grad(int i) const { return V3(x.g[i], y.g[i], z.g[i]); }
    void setGrad(int i, const V3 &g) { x.g[i] = g.x; y.g[i] = g.y; z.g[i] = g.z; }
    bool gradZero(double prec) const { return x.gradZero(prec) && y.gradZero(prec) && z.gradZero(prec); }
    D<N> &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const D<N> &operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <int N> inline DV<N> operator+(const DV<N> &a, const DV<N> &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <int N> inline DV<N> operator-(const DV<N> &a, const DV<N> &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <int N> inline DV<N> operator-(const DV<N> &a) { return {-a.x, -a.y, -a.z}; }
template <int N> inline DV<N> operator*(const DV<N> &a, const D<N> &s) { return {a.x * s, a.y * s, a.z * s}; }
template <int N> inline DV<N> operator*(const D<N> &s, const DV<N> &a) { return a * s; }
template <int N> inline DV<N> operator*(const DV<N> &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
template <int N> inline DV<N> operator*(double s, const DV<N> &a) { return a * s; }
template <int N> inline DV<N> operator/(const DV<N> &a, const D<N> &s) { return {a.x / s, a.y / s, a.z / s}; }
template <int N> inline DV<N> operator/(const DV<N> &a, double s) { return {a.x / s, a.y / s, a.z / s}; }
template <int N> inline D<N> ddot(const DV<N> &a, const DV<N> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int N> inline DV<N> dcross(const DV<N> &a, const DV<N> &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
template <int N> inline D<N> dnorm(const DV<N> &a) { return dsqrt(ddot(a, a)); }
template <int N> inline DV<N> dnormalized(const DV<N> &a) { return a / dnorm(a); }

// Array1D<3> AD = SpectrumAD (simpleAD.h:400-740), elementwise.
template <int N> struct DS {
    D<N> c[3];
    DS() = default;
    DS(const Spec &s) { for (int k = 0; k < 3; ++k) c[k] = D<N>(s[k]); }
    static DS fill(const D<N> &a) { DS r; r.c[0] = a; r.c[1] = a; r.c[2] = a; return r; }
    Spec val() const { return Spec(c[0].v, c[1].v, c[2].v); }
    bool isZero(double prec) const { return val().isZero(prec); }
};
template <int N> inline DS<N> operator+(const DS<N> &a, const DS<N> &b) { DS<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] + b.c[k]; return r; }
template <int N> inline DS<N> operator*(const DS<N> &a, const DS<N> &b) { DS<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * b.c[k]; return r; }
template <int N> inline DS<N> operator*(const DS<N> &a, const D<N> &s) { DS<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s; return r; }
template <int N> inline DS<N> operator*(const DS<N> &a, double s) { DS<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s; return r; }
template <int N> inline DS<N> operator/(const DS<N> &a, double s) { DS<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] / s; return r; }
template <int N> inline DS<N> &operator+=(DS<N> &a, const DS<N> &b) { a = a + b; return a; }
template <int N> inline DS<N> &operator*=(DS<N> &a, const DS<N> &b) { a = a * b; return a; }
template <int N> inline DS<N> &operator*=(DS<N> &a, const D<N> &s) { a = a * s; return a; }

// ------------------------------------------------------------------ PCG32
// sampler.cpp:3-41 — PCG32 XSH-RR, stream (idx+1), state offset 0x853c49e6748fea9b+seed,
// doubles built from the 32-bit output placed in the top of the mantissa.
struct Pcg {
    uint64_t state = 0, inc = 1;
    Pcg(uint64_t seed, int idx) {
        state = 0U;
        inc = ((((uint64_t)(int64_t)idx) + 1) << 1u) | 1u;
        next_u32();
        state += (0x853c49e6748fea9bULL + seed);
        next_u32();
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + (inc | 1);
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    double next1D() {
        uint64_t u = ((uint64_t)next_u32() << 20) | 0x3ff0000000000000ULL;
        double d;
        std::memcpy(&d, &u, sizeof(d));
        return d - 1.0;
    }
};

// ---------------------------------------------------------------- frames
// utils.cpp:54-67 — orthonormal basis after a fixed rotation of the normal
// (angle 0.1f about normalize(0.1f,0.2f,0.3f)); rotation matrix by Rodrigues.
struct Rot { double m[3][3]; };
inline const Rot &frameRot() {
    static Rot R = [] {
        Rot r;
        double ax = 0.1f, ay = 0.2f, az = 0.3f;
        double n = std::sqrt(ax * ax + ay * ay + az * az);
        ax /= n; ay /= n; az /= n;
        double th = (double)0.1f, c = std::cos(th), s = std::sin(th), t = 1 - c;
        r.m[0][0] = c + ax * ax * t;      r.m[0][1] = ax * ay * t - az * s; r.m[0][2] = ax * az * t + ay * s;
        r.m[1][0] = ay * ax * t + az * s; r.m[1][1] = c + ay * ay * t;      r.m[1][2] = ay * az * t - ax * s;
        r.m[2][0] = az * ax * t - ay * s; r.m[2][1] = az * ay * t + ax * s; r.m[2][2] = c + az * az * t;
        return r;
    }();
    return R;
}
inline V3 rotApply(const Rot &R, const V3 &v, bool transpose) {
    V3 o;
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) s += (transpose ? R.m[j][i] : R.m[i][j]) * v[j];
        o[i] = s;
    }
    return o;
}
template <int N> inline DV<N> rotApply(const Rot &R, const DV<N> &v, bool transpose) {
    DV<N> o;
    for (int i = 0; i < 3; ++i) {
        D<N> s(0.0);
        for (int j = 0; j < 3; ++j) s += (transpose ? R.m[j][i] : R.m[i][j]) * v[j];
        o[i] = s;
    }
    return o;
}
inline void coordinateSystem(const V3 &n0, V3 &s, V3 &t) {
    const Rot &R = frameRot();
    V3 n = rotApply(R, n0, false);
    if (std::abs(n.x) > std::abs(n.y)) {
        double inv = 1.0 / std::sqrt(n.x * n.x + n.z * n.z);
        t = V3(n.z * inv, 0.0, -n.x * inv);
    } else {
        double inv = 1.0 / std::sqrt(n.y * n.y + n.z * n.z);
        t = V3(0.0, n.z * inv, -n.y * inv);
    }
    s = cross(t, n);
    s = rotApply(R, s, true);
    t = rotApply(R, t, true);
}
// utils.cpp:69-82
template <int N> inline void coordinateSystemAD(const DV<N> &n0, DV<N> &s, DV<N> &t) {
    const Rot &R = frameRot();
    DV<N> n = rotApply(R, n0, false);
    if (std::abs(n.x.v) > std::abs(n.y.v)) {
        D<N> inv = 1.0 / dsqrt(n.x * n.x + n.z * n.z);
        t = DV<N>(n.z * inv, D<N>(0.0), -n.x * inv);
    } else {
        D<N> inv = 1.0 / dsqrt(n.y * n.y + n.z * n.z);
        t = DV<N>(D<N>(0.0), n.z * inv, -n.y * inv);
    }
    s = dcross(t, n);
    s = rotApply(R, s, true);
    t = rotApply(R, t, true);
}

// frame.h:8-79
struct Frame {
    V3 s, t, n;
    Frame() = default;
    Frame(const V3 &a, const V3 &b, const V3 &c) : s(a), t(b), n(c) {}
    explicit Frame(const V3 &nn) : n(nn) { coordinateSystem(nn, s, t); }
    V3 toLocal(const V3 &v) const { return V3(dot(v, s), dot(v, t), dot(v, n)); }
    V3 toWorld(const V3 &v) const { return s * v.x + t * v.y + n * v.z; }
    static double cosTheta(const V3 &v) { return v.z; }
    static double cosTheta2(const V3 &v) { return v.z * v.z; }
    static double tanTheta(const V3 &v) {
        double tmp = 1.0 - v.z * v.z;
        return tmp < EPS ? 0.0 : std::sqrt(tmp) / v.z;
    }
};
// frameAD.h:9-82
template <int N> struct FrameAD {
    DV<N> s, t, n;
    FrameAD() = default;
    explicit FrameAD(const DV<N> &nn) : n(nn) { coordinateSystemAD(nn, s, t); }
    DV<N> toLocal(const DV<N> &v) const { return DV<N>(ddot(v, s), ddot(v, t), ddot(v, n)); }
    DV<N> toWorld(const DV<N> &v) const { return s * v.x + t * v.y + n * v.z; }
    Frame toFrame() const { return Frame(s.val(), t.val(), n.val()); }
    static D<N> tanTheta(const DV<N> &v) {
        D<N> tmp = 1.0 - v.z * v.z;
        return tmp.v < EPS ? D<N>(0.0) : dsqrt(tmp) / v.z;
    }
};

// ----------------------------------------------------------------- warps
// utils.cpp:18-34 (note the reference's sin/cos naming swap; kept as written)
inline V2 squareToUniformDiskConcentric(double sx, double sy) {
    double r1 = 2.0 * sx - 1.0, r2 = 2.0 * sy - 1.0, phi, r;
    if (r1 == 0 && r2 == 0) {
        r = phi = 0;
    } else if (r1 * r1 > r2 * r2) {
        r = r1;
        phi = (PI / 4.0) * (r2 / r1);
    } else {
        r = r2;
        phi = (PI / 2.0) - (r1 / r2) * (PI / 4.0);
    }
    double cosPhi = std::sin(phi), sinPhi = std::cos(phi);
    return V2{r * cosPhi, r * sinPhi};
}
// utils.cpp:36-52
inline V3 squareToCosineHemisphere(double sx, double sy) {
    V2 p = squareToUniformDiskConcentric(sx, sy);
    double z = std::sqrt(std::max(0.0, 1.0 - p.x * p.x - p.y * p.y));
    return V3(p.x, p.y, z);
}
inline V3 squareToUniformSphere(double sx, double sy) {
    double z = 1.0 - 2.0 * sy;
    double r = std::sqrt(1.0 - z * z);
    return V3(r * std::cos(2.0 * PI * sx), r * std::sin(2.0 * PI * sx), z);
}
inline double squareToCosineHemispherePdf(const V3 &d) { return INV_PI * d.z; }

inline double signum(double v) { return std::copysign(1.0, v); }  // math_func.h:30-36
// utils.cpp:84-88
inline double luminance(const Spec &v) { return 0.212671f * v[0] + 0.715160f * v[1] + 0.072169f * v[2]; }

// utils.cpp:119-148
inline double fresnelDielectricExt(double cosThetaI_, double &cosThetaT_, double eta) {
    if (std::abs(eta - 1.0) < EPS) { cosThetaT_ = -cosThetaI_; return 0.0; }
    double scale = (cosThetaI_ > 0.0) ? 1.0 / eta : eta;
    double cosThetaTSqr = 1.0 - (1.0 - cosThetaI_ * cosThetaI_) * (scale * scale);
    if (cosThetaTSqr < EPS) { cosThetaT_ = 0.0; return 1.0; }
    double cosThetaI = std::abs(cosThetaI_), cosThetaT = std::sqrt(cosThetaTSqr);
    double Rs = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
    double Rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    cosThetaT_ = (cosThetaI_ > 0.0) ? -cosThetaT : cosThetaT;
    return 0.5 * (Rs * Rs + Rp * Rp);
}
inline double fresnelDielectricExt(double cosThetaI, double eta) { double c; return fresnelDielectricExt(cosThetaI, c, eta); }
// utils.cpp:150-184
template <int N> inline D<N> fresnelDielectricExtAD(const D<N> &cosThetaI_, const D<N> &eta) {
    if (std::abs(eta.v - 1.0) < EPS) return D<N>(0.0);
    D<N> scale = (cosThetaI_.v > 0.0) ? 1.0 / eta : eta;
    D<N> cosThetaTSqr = 1.0 - (1.0 - dsquare(cosThetaI_)) * dsquare(scale);
    if (cosThetaTSqr.v < EPS) return D<N>(1.0);
    D<N> cosThetaI = dabs(cosThetaI_), cosThetaT = dsqrt(cosThetaTSqr);
    D<N> Rs = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
    D<N> Rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    return 0.5 * (Rs * Rs + Rp * Rp);
}
// utils.h:312, utils.cpp:186-189
inline V3 reflect(const V3 &wi, const V3 &n) { return 2.0 * dot(wi, n) * n - wi; }
inline V3 refract(const V3 &wi, const V3 &n, double eta, double cosThetaT) {
    if (cosThetaT < 0) eta = 1.0 / eta;
    return n * (dot(wi, n) * eta + cosThetaT) - wi * eta;
}

// math_func.cpp:6-23 — Abramowitz & Stegun 7.1.26
inline double erf_as(double x) {
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
                 a5 = 1.061405429, p = 0.3275911;
    double sign = signum(x);
    x = std::abs(x);
    double t = 1.0 / (1.0 + p * x);
    double y = 1.0 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * std::exp(-x * x);
    return sign * y;
}
// math_func.cpp:26-53 — M. Giles, "Approximating the erfinv function" (single-precision fit)
inline double erfinv_giles(double x) {
    double w = -std::log((1.0 - x) * (1.0 + x)), p;
    if (w < 5.0) {
        w = w - 2.5;
        p = 2.81022636e-08;
        p = 3.43273939e-07 + p * w;
        p = -3.5233877e-06 + p * w;
        p = -4.39150654e-06 + p * w;
        p = 0.00021858087 + p * w;
        p = -0.00125372503 + p * w;
        p = -0.00417768164 + p * w;
        p = 0.246640727 + p * w;
        p = 1.50140941 + p * w;
    } else {
        w = std::sqrt(w) - 3.0;
        p = -0.000200214257;
        p = 0.000100950558 + p * w;
        p = 0.00134934322 + p * w;
        p = -0.00367342844 + p * w;
        p = 0.00573950773 + p * w;
        p = -0.0076224613 + p * w;
        p = 0.00943887047 + p * w;
        p = 1.00167406 + p * w;
        p = 2.83297682 + p * w;
    }
    return p * x;
}

// utils.cpp:218-235 — Moller-Trumbore without bounds checks; divisor clamped to +-Epsilon
inline V3 rayIntersectTriangle(const V3 &v0, const V3 &v1, const V3 &v2, const V3 &org, const V3 &dir) {
    V3 e1 = v1 - v0, e2 = v2 - v0;
    V3 pvec = cross(dir, e2);
    double divisor = dot(pvec, e1);
    if (std::abs(divisor) < EPS) divisor = (divisor > 0) ? EPS : -EPS;
    V3 s = org - v0;
    double u = dot(s, pvec) / divisor;
    V3 qvec = cross(s, e1);
    double v = dot(dir, qvec) / divisor;
    double t = dot(e2, qvec) / divisor;
    return V3(u, v, t);
}
// utils.cpp:237-254
template <int N>
inline void rayIntersectTriangleAD(const DV<N> &v0, const DV<N> &v1, const DV<N> &v2, const DV<N> &org,
                                   const DV<N> &dir, D<N> &u, D<N> &v, D<N> &t) {
    DV<N> e1 = v1 - v0, e2 = v2 - v0;
    DV<N> pvec = dcross(dir, e2);
    D<N> divisor = ddot(pvec, e1);
    if (std::abs(divisor.v) < EPS) divisor = D<N>(divisor.v > 0 ? EPS : -EPS);
    DV<N> s = org - v0;
    u = ddot(s, pvec) / divisor;
    DV<N> qvec = dcross(s, e1);
    v = ddot(dir, qvec) / divisor;
    t = ddot(e2, qvec) / divisor;
}

// utils.cpp:256-266
inline double computeIntersectionInTri(const V3 &a, const V3 &b0, const V3 &c0, const V3 &b1, const V3 &c1, double t0) {
    double coeff_b = norm(b0 - a) / norm(b1 - a);
    double coeff_c = norm(c0 - a) / norm(c1 - a);
    double bx = (1 - t0) * coeff_b, by = t0 * coeff_c;
    double s = bx + by;
    return by / s;
}

}  // namespace oref
