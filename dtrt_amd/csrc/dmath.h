// Device math layer for the gfx950 path tracer: dual numbers with N tangent lanes, a
// Philox4x32-10 counter-based RNG, frames, warps, Fresnel and the Moller-Trumbore
// refinement. Semantics follow the reference's SimpleAD (include/simpleAD.h:20-1040)
// and utils (src/utils.cpp).
//
// The estimator's scalar type `real` is fixed per compilation: DT_REAL = double
// (namespace dt::p64, the reference's own precision, include/fwd.h:10-13) or float
// (dt::p32). Scene data stays fp32 in memory (the reference reads fp32 vertices,
// grids and parameters too) and is widened on load; BVH traversal is fp32 in both,
// as Embree's is (src/scene.cpp:17-43), and every hit is refined in `real`.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define DI __device__ __forceinline__
#define HDI __host__ __device__ __forceinline__

#ifndef DT_REAL
#define DT_REAL float
#endif
#ifndef DT_PNS
#define DT_PNS p32
#endif

namespace dt {
namespace DT_PNS {

using real = DT_REAL;
constexpr bool kF64 = sizeof(real) == 8;
// `1.5_r`: a literal of the estimator's type
constexpr real operator""_r(long double v) { return (real)v; }
constexpr real operator""_r(unsigned long long v) { return (real)v; }
// sin and cos of pi * a: the angles the warps draw (2 pi u) and the edge slerp's
// angles, without the general argument reduction (whose double-precision path needs
// many registers); mathematically the same as sincos(pi * a)
DI void rsincospi(float a, float *s, float *c) { sincospif(a, s, c); }
DI void rsincospi(double a, double *s, double *c) { sincospi(a, s, c); }
DI float rsinpi(float a) { return sinpif(a); }
DI double rsinpi(double a) { return sinpi(a); }

// fwd.h:14, config.h:6-7.
constexpr real EPS = 1e-7_r;
constexpr real SHADOW_EPS = 1e-3_r;
constexpr real ANGLE_EPS = 1e-4_r;
constexpr real PI = 3.14159265358979323846_r;
constexpr real INV_PI = 0.31830988618379067154_r;
constexpr real INV_TWOPI = 0.15915494309189534561_r;
constexpr real INV_FOURPI = 0.07957747154594766788_r;
constexpr real FINF = (real)__builtin_huge_val();

// ------------------------------------------------------------------- float3
struct f3 {
    real x, y, z;
    HDI real operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
HDI f3 mk(real a, real b, real c) { return f3{a, b, c}; }
HDI f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HDI f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HDI f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
HDI f3 operator*(f3 a, real s) { return {a.x * s, a.y * s, a.z * s}; }
HDI f3 operator*(real s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
HDI f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
HDI f3 operator/(f3 a, real s) { real r = 1.0_r / s; return {a.x * r, a.y * r, a.z * r}; }
HDI real dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HDI f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HDI real length(f3 a) { return sqrt(dot(a, a)); }
HDI f3 normalize(f3 a) { return a * (1.0_r / sqrt(dot(a, a))); }
HDI bool allZero(f3 a, real prec) { return fabs(a.x) <= prec && fabs(a.y) <= prec && fabs(a.z) <= prec; }
HDI real signum(real v) { return copysign(1.0_r, v); }

// -------------------------------------------------------------- dual numbers
template <int N> struct Df {
    real v;
    real d[N];
    HDI Df() : v(0._r) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0._r;
    }
    HDI Df(real x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0._r;
    }
    HDI bool gradZero(real prec) const {
        bool z = true;
#pragma unroll
        for (int i = 0; i < N; ++i) z = z && fabs(d[i]) <= prec;
        return z;
    }
    HDI real gradAbsMax() const {
        real m = 0._r;
#pragma unroll
        for (int i = 0; i < N; ++i) m = fmax(m, fabs(d[i]));
        return m;
    }
};
#define DF_LOOP _Pragma("unroll") for (int i = 0; i < N; ++i)
template <int N> HDI Df<N> operator+(const Df<N> &a, const Df<N> &b) { Df<N> r(a.v + b.v); DF_LOOP r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> HDI Df<N> operator-(const Df<N> &a, const Df<N> &b) { Df<N> r(a.v - b.v); DF_LOOP r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> HDI Df<N> operator-(const Df<N> &a) { Df<N> r(-a.v); DF_LOOP r.d[i] = -a.d[i]; return r; }
template <int N> HDI Df<N> operator*(const Df<N> &a, const Df<N> &b) { Df<N> r(a.v * b.v); DF_LOOP r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> HDI Df<N> operator/(const Df<N> &a, const Df<N> &b) {
    real ib = 1.0_r / b.v, q = a.v * ib;
    Df<N> r(q);
    DF_LOOP r.d[i] = (a.d[i] - q * b.d[i]) * ib;
    return r;
}
template <int N> HDI Df<N> operator+(const Df<N> &a, real s) { Df<N> r = a; r.v += s; return r; }
template <int N> HDI Df<N> operator+(real s, const Df<N> &a) { Df<N> r = a; r.v += s; return r; }
template <int N> HDI Df<N> operator-(const Df<N> &a, real s) { Df<N> r = a; r.v -= s; return r; }
template <int N> HDI Df<N> operator-(real s, const Df<N> &a) { Df<N> r(s - a.v); DF_LOOP r.d[i] = -a.d[i]; return r; }
template <int N> HDI Df<N> operator*(const Df<N> &a, real s) { Df<N> r(a.v * s); DF_LOOP r.d[i] = a.d[i] * s; return r; }
template <int N> HDI Df<N> operator*(real s, const Df<N> &a) { return a * s; }
template <int N> HDI Df<N> operator/(const Df<N> &a, real s) { return a * (1.0_r / s); }
template <int N> HDI Df<N> operator/(real s, const Df<N> &a) {
    real ia = 1.0_r / a.v, q = s * ia;
    Df<N> r(q);
    DF_LOOP r.d[i] = -q * ia * a.d[i];
    return r;
}
template <int N> HDI Df<N> &operator+=(Df<N> &a, const Df<N> &b) { a = a + b; return a; }
template <int N> HDI Df<N> &operator*=(Df<N> &a, const Df<N> &b) { a = a * b; return a; }
// simpleAD.h:202-230
template <int N> HDI Df<N> dsqrt(const Df<N> &a) { real t = sqrt(a.v), h = 0.5_r / t; Df<N> r(t); DF_LOOP r.d[i] = a.d[i] * h; return r; }
template <int N> HDI Df<N> dexp(const Df<N> &a) { real e = exp(a.v); Df<N> r(e); DF_LOOP r.d[i] = a.d[i] * e; return r; }
template <int N> HDI Df<N> dlog(const Df<N> &a) { real ia = 1.0_r / a.v; Df<N> r(log(a.v)); DF_LOOP r.d[i] = a.d[i] * ia; return r; }
template <int N> HDI Df<N> dpow(const Df<N> &a, real p) { real t = pow(a.v, p - 1.0_r); Df<N> r(a.v * t); DF_LOOP r.d[i] = a.d[i] * p * t; return r; }
template <int N> HDI Df<N> dpow(const Df<N> &a, const Df<N> &p) { return dexp(p * dlog(a)); }
// a^1.5 = a * sqrt(a), with tangents 1.5 sqrt(a) a' (dpow(a, 1.5) without pow)
template <int N> HDI Df<N> dpow15(const Df<N> &a) { real t = sqrt(a.v), h = 1.5_r * t; Df<N> r(a.v * t); DF_LOOP r.d[i] = a.d[i] * h; return r; }
template <int N> HDI Df<N> dabs(const Df<N> &a) { return a.v > 0._r ? a : -a; }

template <int N> struct Dv {
    Df<N> x, y, z;
    HDI Dv() {}
    HDI Dv(const f3 &a) : x(a.x), y(a.y), z(a.z) {}
    HDI Dv(const Df<N> &a, const Df<N> &b, const Df<N> &c) : x(a), y(b), z(c) {}
    HDI f3 val() const { return f3{x.v, y.v, z.v}; }
    HDI f3 grad(int i) const { return f3{x.d[i], y.d[i], z.d[i]}; }
    HDI void setGrad(int i, const f3 &g) { x.d[i] = g.x; y.d[i] = g.y; z.d[i] = g.z; }
    HDI bool gradZero(real prec) const { return x.gradZero(prec) && y.gradZero(prec) && z.gradZero(prec); }
};
template <int N> HDI Dv<N> operator+(const Dv<N> &a, const Dv<N> &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <int N> HDI Dv<N> operator-(const Dv<N> &a, const Dv<N> &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <int N> HDI Dv<N> operator-(const Dv<N> &a) { return {-a.x, -a.y, -a.z}; }
template <int N> HDI Dv<N> operator*(const Dv<N> &a, const Df<N> &s) { return {a.x * s, a.y * s, a.z * s}; }
template <int N> HDI Dv<N> operator*(const Dv<N> &a, real s) { return {a.x * s, a.y * s, a.z * s}; }
template <int N> HDI Dv<N> operator/(const Dv<N> &a, const Df<N> &s) { Df<N> r = 1.0_r / s; return a * r; }
template <int N> HDI Df<N> ddot(const Dv<N> &a, const Dv<N> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int N> HDI Df<N> ddot(const Dv<N> &a, const f3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int N> HDI Dv<N> dcross(const Dv<N> &a, const Dv<N> &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
template <int N> HDI Df<N> dnorm(const Dv<N> &a) { return dsqrt(ddot(a, a)); }
template <int N> HDI Dv<N> dnormalize(const Dv<N> &a) { return a / dnorm(a); }

// Spectrum with tangents (Array1D<3>, simpleAD.h:400-740)
template <int N> struct Ds {
    Df<N> c[3];
    HDI Ds() {}
    HDI Ds(const f3 &s) { c[0] = Df<N>(s.x); c[1] = Df<N>(s.y); c[2] = Df<N>(s.z); }
    HDI static Ds fill(const Df<N> &a) { Ds r; r.c[0] = a; r.c[1] = a; r.c[2] = a; return r; }
    HDI f3 val() const { return f3{c[0].v, c[1].v, c[2].v}; }
    HDI bool isZero(real prec) const { return allZero(val(), prec); }
};
template <int N> HDI Ds<N> operator+(const Ds<N> &a, const Ds<N> &b) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] + b.c[k]; return r; }
template <int N> HDI Ds<N> operator*(const Ds<N> &a, const Ds<N> &b) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * b.c[k]; return r; }
template <int N> HDI Ds<N> operator*(const Ds<N> &a, const Df<N> &s) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s; return r; }
template <int N> HDI Ds<N> operator*(const Ds<N> &a, real s) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s; return r; }
template <int N> HDI Ds<N> &operator+=(Ds<N> &a, const Ds<N> &b) { a = a + b; return a; }

// ------------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): counter-based, so any draw of any path can be regenerated from
// (key, path id, stream, index) — the property adjoint path replay relies on.
struct u4 { uint32_t a, b, c, d; };
HDI uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
HDI u4 philox4x32_10(u4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = mulhi32(M0, ctr.a), lo0 = M0 * ctr.a;
        uint32_t hi1 = mulhi32(M1, ctr.c), lo1 = M1 * ctr.c;
        ctr = u4{hi1 ^ ctr.b ^ k0, lo1, hi0 ^ ctr.d ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// One random stream: key from the render seed; counter = (stream lo, stream hi, block, 0).
// The 4 words of the current block are separate registers picked by a select chain
// (an indexed array would be placed in scratch memory).
struct Rng {
    uint32_t k0, k1, s0, s1, blk;
    uint32_t b0, b1, b2, b3;
    int pos;
    uint32_t nlk;   // density-grid lookups made along this stream's path (work counter)
    HDI void init(uint64_t seed, uint64_t stream) {
        nlk = 0;
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32) ^ 0x5bd1e995u;
        s0 = (uint32_t)stream;
        s1 = (uint32_t)(stream >> 32);
        blk = 0;
        pos = 4;
    }
    HDI real next() {
        if (pos == 4) {
            u4 o = philox4x32_10(u4{s0, s1, blk++, 0x9e3779b9u}, k0, k1);
            b0 = o.a; b1 = o.b; b2 = o.c; b3 = o.d;
            pos = 0;
        }
        uint32_t x = pos == 0 ? b0 : pos == 1 ? b1 : pos == 2 ? b2 : b3;
        pos++;
        if constexpr (kF64) return (real)x * (1.0_r / 4294967296.0_r);   // [0, 1) with 32 bits
        else return (real)(x >> 8) * (1.0_r / 16777216.0_r);             // [0, 1) with 24 bits
    }
};

// ------------------------------------------------------------------- frames
// utils.cpp:54-67 — orthonormal basis built after a fixed rotation of the
// normal (0.1 rad about normalize(0.1,0.2,0.3)); R computed by Rodrigues on the host.
struct FrameRot { real m[9]; };
// Each kernel object defines its own copy under its own name.
#ifndef DT_FRAME_ROT
#define DT_FRAME_ROT c_frame_rot
#endif
extern __constant__ FrameRot DT_FRAME_ROT;

DI f3 rotMul(const real *m, f3 v, bool tr) {
    return tr ? f3{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z, m[2] * v.x + m[5] * v.y + m[8] * v.z}
              : f3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
template <int N> DI Dv<N> rotMul(const real *m, const Dv<N> &v, bool tr) {
    Dv<N> o;
    if (tr) {
        o.x = v.x * m[0] + v.y * m[3] + v.z * m[6];
        o.y = v.x * m[1] + v.y * m[4] + v.z * m[7];
        o.z = v.x * m[2] + v.y * m[5] + v.z * m[8];
    } else {
        o.x = v.x * m[0] + v.y * m[1] + v.z * m[2];
        o.y = v.x * m[3] + v.y * m[4] + v.z * m[5];
        o.z = v.x * m[6] + v.y * m[7] + v.z * m[8];
    }
    return o;
}
struct Frame {
    f3 s, t, n;
    DI Frame() {}
    DI explicit Frame(f3 nn) : n(nn) {
        const real *R = DT_FRAME_ROT.m;
        f3 q = rotMul(R, nn, false);
        if (fabs(q.x) > fabs(q.y)) {
            real il = 1.0_r / sqrt(q.x * q.x + q.z * q.z);
            t = f3{q.z * il, 0._r, -q.x * il};
        } else {
            real il = 1.0_r / sqrt(q.y * q.y + q.z * q.z);
            t = f3{0._r, q.z * il, -q.y * il};
        }
        s = rotMul(R, cross(t, q), true);
        t = rotMul(R, t, true);
    }
    DI f3 toLocal(f3 v) const { return f3{dot(v, s), dot(v, t), dot(v, n)}; }
    DI f3 toWorld(f3 v) const { return s * v.x + t * v.y + n * v.z; }
};
template <int N> struct FrameAD {   // frameAD.h:9-82 via coordinateSystemAD (utils.cpp:69-82)
    Dv<N> s, t, n;
    DI FrameAD() {}
    DI explicit FrameAD(const Dv<N> &nn) : n(nn) {
        const real *R = DT_FRAME_ROT.m;
        Dv<N> q = rotMul(R, nn, false);
        if (fabs(q.x.v) > fabs(q.y.v)) {
            Df<N> il = 1.0_r / dsqrt(q.x * q.x + q.z * q.z);
            t = Dv<N>(q.z * il, Df<N>(0._r), -q.x * il);
        } else {
            Df<N> il = 1.0_r / dsqrt(q.y * q.y + q.z * q.z);
            t = Dv<N>(Df<N>(0._r), q.z * il, -q.y * il);
        }
        s = rotMul(R, dcross(t, q), true);
        t = rotMul(R, t, true);
    }
    DI Dv<N> toLocal(const Dv<N> &v) const { return Dv<N>(ddot(v, s), ddot(v, t), ddot(v, n)); }
    DI Dv<N> toWorld(const Dv<N> &v) const { return s * v.x + t * v.y + n * v.z; }
    DI Frame val() const { Frame f; f.s = s.val(); f.t = t.val(); f.n = n.val(); return f; }
};

// -------------------------------------------------------------------- warps
// utils.cpp:18-52 (the reference's sin/cos naming swap in the concentric map is kept)
DI f3 squareToCosineHemisphere(real sx, real sy) {
    real r1 = 2.0_r * sx - 1.0_r, r2 = 2.0_r * sy - 1.0_r, phi, r;   // phi in units of pi
    if (r1 == 0._r && r2 == 0._r) { r = phi = 0._r; }
    else if (r1 * r1 > r2 * r2) { r = r1; phi = 0.25_r * (r2 / r1); }
    else { r = r2; phi = 0.5_r - (r1 / r2) * 0.25_r; }
    real sp, cp;
    rsincospi(phi, &sp, &cp);
    real px = r * sp, py = r * cp;
    return f3{px, py, sqrt(fmax(0._r, 1.0_r - px * px - py * py))};
}
DI f3 squareToUniformSphere(real sx, real sy) {
    real z = 1.0_r - 2.0_r * sy, r = sqrt(fmax(0._r, 1.0_r - z * z)), sp, cp;
    rsincospi(2.0_r * sx, &sp, &cp);
    return f3{r * cp, r * sp, z};
}
HDI real luminance(f3 v) { return 0.212671_r * v.x + 0.715160_r * v.y + 0.072169_r * v.z; }

// utils.cpp:119-148
DI real fresnelDielectricExt(real ci_, real &ct_, real eta) {
    if (fabs(eta - 1.0_r) < EPS) { ct_ = -ci_; return 0._r; }
    real scale = ci_ > 0._r ? 1.0_r / eta : eta;
    real ct2 = 1.0_r - (1.0_r - ci_ * ci_) * (scale * scale);
    if (ct2 < EPS) { ct_ = 0._r; return 1.0_r; }
    real ci = fabs(ci_), ct = sqrt(ct2);
    real Rs = (ci - eta * ct) / (ci + eta * ct), Rp = (eta * ci - ct) / (eta * ci + ct);
    ct_ = ci_ > 0._r ? -ct : ct;
    return 0.5_r * (Rs * Rs + Rp * Rp);
}
DI real fresnelDielectricExt(real ci, real eta) { real c; return fresnelDielectricExt(ci, c, eta); }
template <int N> DI Df<N> fresnelDielectricExtAD(const Df<N> &ci_, const Df<N> &eta) {   // utils.cpp:150-184
    if (fabs(eta.v - 1.0_r) < EPS) return Df<N>(0._r);
    Df<N> scale = ci_.v > 0._r ? 1.0_r / eta : eta;
    Df<N> ct2 = 1.0_r - (1.0_r - ci_ * ci_) * (scale * scale);
    if (ct2.v < EPS) return Df<N>(1.0_r);
    Df<N> ci = dabs(ci_), ct = dsqrt(ct2);
    Df<N> Rs = (ci - eta * ct) / (ci + eta * ct), Rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5_r * (Rs * Rs + Rp * Rp);
}
DI f3 reflect(f3 wi, f3 n) { return 2.0_r * dot(wi, n) * n - wi; }
DI f3 refract(f3 wi, f3 n, real eta, real ct) { if (ct < 0._r) eta = 1.0_r / eta; return n * (dot(wi, n) * eta + ct) - wi * eta; }

// math_func.cpp:6-23 — Abramowitz & Stegun 7.1.26 (published coefficients)
DI real erf_as(real x) {
    real sign = signum(x);
    x = fabs(x);
    real t = 1.0_r / (1.0_r + 0.3275911_r * x);
    real y = 1.0_r - (((((1.061405429_r * t - 1.453152027_r) * t) + 1.421413741_r) * t - 0.284496736_r) * t + 0.254829592_r) * t * exp(-x * x);
    return sign * y;
}
// math_func.cpp:26-53 — M. Giles, "Approximating the erfinv function", GPU Computing Gems (2011)
DI real erfinv_giles(real x) {
    real w = -log((1.0_r - x) * (1.0_r + x)), p;
    if (w < 5.0_r) {
        w = w - 2.5_r;
        p = 2.81022636e-08_r;
        p = 3.43273939e-07_r + p * w;
        p = -3.5233877e-06_r + p * w;
        p = -4.39150654e-06_r + p * w;
        p = 0.00021858087_r + p * w;
        p = -0.00125372503_r + p * w;
        p = -0.00417768164_r + p * w;
        p = 0.246640727_r + p * w;
        p = 1.50140941_r + p * w;
    } else {
        w = sqrt(w) - 3.0_r;
        p = -0.000200214257_r;
        p = 0.000100950558_r + p * w;
        p = 0.00134934322_r + p * w;
        p = -0.00367342844_r + p * w;
        p = 0.00573950773_r + p * w;
        p = -0.0076224613_r + p * w;
        p = 0.00943887047_r + p * w;
        p = 1.00167406_r + p * w;
        p = 2.83297682_r + p * w;
    }
    return p * x;
}

// utils.cpp:218-254 — refinement of the chosen triangle (no bounds test; divisor
// clamped away from zero), primal and with tangents.
DI f3 rayTriangleUVT(f3 v0, f3 v1, f3 v2, f3 o, f3 d) {
    f3 e1 = v1 - v0, e2 = v2 - v0, pv = cross(d, e2);
    real div = dot(pv, e1);
    if (fabs(div) < EPS) div = div > 0._r ? EPS : -EPS;
    real idv = 1.0_r / div;
    f3 s = o - v0, q = cross(s, e1);
    return f3{dot(s, pv) * idv, dot(d, q) * idv, dot(e2, q) * idv};
}
template <int N>
DI void rayTriangleUVT_AD(const Dv<N> &v0, const Dv<N> &v1, const Dv<N> &v2, const Dv<N> &o, const Dv<N> &d, Df<N> &u,
                          Df<N> &v, Df<N> &t) {
    Dv<N> e1 = v1 - v0, e2 = v2 - v0, pv = dcross(d, e2);
    Df<N> div = ddot(pv, e1);
    if (fabs(div.v) < EPS) div = Df<N>(div.v > 0._r ? EPS : -EPS);
    Df<N> idv = 1.0_r / div;
    Dv<N> s = o - v0, q = dcross(s, e1);
    u = ddot(s, pv) * idv;
    v = ddot(d, q) * idv;
    t = ddot(e2, q) * idv;
}

}  // namespace DT_PNS
}  // namespace dt
