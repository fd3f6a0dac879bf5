// Device math layer for the gfx950 path tracer: fp32 dual numbers with N tangent
// lanes, a Philox4x32-10 counter-based RNG, frames, warps, Fresnel and the
// Moller-Trumbore refinement. Semantics follow the reference's SimpleAD
// (include/simpleAD.h:20-1040) and utils (src/utils.cpp); the arithmetic is fp32.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define DI __device__ __forceinline__
#define HDI __host__ __device__ __forceinline__

namespace dt {

// fwd.h:14, config.h:6-7. The fp32 build keeps the reference's values; tests that
// are meaningless at fp32 resolution are noted where they are made.
constexpr float EPS = 1e-7f;
constexpr float SHADOW_EPS = 1e-3f;
constexpr float ANGLE_EPS = 1e-4f;
constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float INV_TWOPI = 0.15915494309189534561f;
constexpr float INV_FOURPI = 0.07957747154594766788f;
constexpr float FINF = __builtin_huge_valf();

// ------------------------------------------------------------------- float3
struct f3 {
    float x, y, z;
    HDI float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
HDI f3 mk(float a, float b, float c) { return f3{a, b, c}; }
HDI f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HDI f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HDI f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
HDI f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
HDI f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
HDI f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
HDI f3 operator/(f3 a, float s) { float r = 1.0f / s; return {a.x * r, a.y * r, a.z * r}; }
HDI float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HDI f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HDI float length(f3 a) { return sqrtf(dot(a, a)); }
HDI f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
HDI bool allZero(f3 a, float prec) { return fabsf(a.x) <= prec && fabsf(a.y) <= prec && fabsf(a.z) <= prec; }
HDI float signum(float v) { return copysignf(1.0f, v); }

// -------------------------------------------------------------- dual numbers
template <int N> struct Df {
    float v;
    float d[N];
    HDI Df() : v(0.f) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.f;
    }
    HDI Df(float x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.f;
    }
    HDI bool gradZero(float prec) const {
        bool z = true;
#pragma unroll
        for (int i = 0; i < N; ++i) z = z && fabsf(d[i]) <= prec;
        return z;
    }
    HDI float gradAbsMax() const {
        float m = 0.f;
#pragma unroll
        for (int i = 0; i < N; ++i) m = fmaxf(m, fabsf(d[i]));
        return m;
    }
};
#define DF_LOOP _Pragma("unroll") for (int i = 0; i < N; ++i)
template <int N> HDI Df<N> operator+(const Df<N> &a, const Df<N> &b) { Df<N> r(a.v + b.v); DF_LOOP r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> HDI Df<N> operator-(const Df<N> &a, const Df<N> &b) { Df<N> r(a.v - b.v); DF_LOOP r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> HDI Df<N> operator-(const Df<N> &a) { Df<N> r(-a.v); DF_LOOP r.d[i] = -a.d[i]; return r; }
template <int N> HDI Df<N> operator*(const Df<N> &a, const Df<N> &b) { Df<N> r(a.v * b.v); DF_LOOP r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> HDI Df<N> operator/(const Df<N> &a, const Df<N> &b) {
    float ib = 1.0f / b.v, q = a.v * ib;
    Df<N> r(q);
    DF_LOOP r.d[i] = (a.d[i] - q * b.d[i]) * ib;
    return r;
}
template <int N> HDI Df<N> operator+(const Df<N> &a, float s) { Df<N> r = a; r.v += s; return r; }
template <int N> HDI Df<N> operator+(float s, const Df<N> &a) { Df<N> r = a; r.v += s; return r; }
template <int N> HDI Df<N> operator-(const Df<N> &a, float s) { Df<N> r = a; r.v -= s; return r; }
template <int N> HDI Df<N> operator-(float s, const Df<N> &a) { Df<N> r(s - a.v); DF_LOOP r.d[i] = -a.d[i]; return r; }
template <int N> HDI Df<N> operator*(const Df<N> &a, float s) { Df<N> r(a.v * s); DF_LOOP r.d[i] = a.d[i] * s; return r; }
template <int N> HDI Df<N> operator*(float s, const Df<N> &a) { return a * s; }
template <int N> HDI Df<N> operator/(const Df<N> &a, float s) { return a * (1.0f / s); }
template <int N> HDI Df<N> operator/(float s, const Df<N> &a) {
    float ia = 1.0f / a.v, q = s * ia;
    Df<N> r(q);
    DF_LOOP r.d[i] = -q * ia * a.d[i];
    return r;
}
template <int N> HDI Df<N> &operator+=(Df<N> &a, const Df<N> &b) { a = a + b; return a; }
template <int N> HDI Df<N> &operator*=(Df<N> &a, const Df<N> &b) { a = a * b; return a; }
// simpleAD.h:202-230
template <int N> HDI Df<N> dsqrt(const Df<N> &a) { float t = sqrtf(a.v), h = 0.5f / t; Df<N> r(t); DF_LOOP r.d[i] = a.d[i] * h; return r; }
template <int N> HDI Df<N> dexp(const Df<N> &a) { float e = expf(a.v); Df<N> r(e); DF_LOOP r.d[i] = a.d[i] * e; return r; }
template <int N> HDI Df<N> dlog(const Df<N> &a) { float ia = 1.0f / a.v; Df<N> r(logf(a.v)); DF_LOOP r.d[i] = a.d[i] * ia; return r; }
template <int N> HDI Df<N> dpow(const Df<N> &a, float p) { float t = powf(a.v, p - 1.0f); Df<N> r(a.v * t); DF_LOOP r.d[i] = a.d[i] * p * t; return r; }
template <int N> HDI Df<N> dpow(const Df<N> &a, const Df<N> &p) { return dexp(p * dlog(a)); }
template <int N> HDI Df<N> dabs(const Df<N> &a) { return a.v > 0.f ? a : -a; }

template <int N> struct Dv {
    Df<N> x, y, z;
    HDI Dv() {}
    HDI Dv(const f3 &a) : x(a.x), y(a.y), z(a.z) {}
    HDI Dv(const Df<N> &a, const Df<N> &b, const Df<N> &c) : x(a), y(b), z(c) {}
    HDI f3 val() const { return f3{x.v, y.v, z.v}; }
    HDI f3 grad(int i) const { return f3{x.d[i], y.d[i], z.d[i]}; }
    HDI void setGrad(int i, const f3 &g) { x.d[i] = g.x; y.d[i] = g.y; z.d[i] = g.z; }
    HDI bool gradZero(float prec) const { return x.gradZero(prec) && y.gradZero(prec) && z.gradZero(prec); }
};
template <int N> HDI Dv<N> operator+(const Dv<N> &a, const Dv<N> &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <int N> HDI Dv<N> operator-(const Dv<N> &a, const Dv<N> &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <int N> HDI Dv<N> operator-(const Dv<N> &a) { return {-a.x, -a.y, -a.z}; }
template <int N> HDI Dv<N> operator*(const Dv<N> &a, const Df<N> &s) { return {a.x * s, a.y * s, a.z * s}; }
template <int N> HDI Dv<N> operator*(const Dv<N> &a, float s) { return {a.x * s, a.y * s, a.z * s}; }
template <int N> HDI Dv<N> operator/(const Dv<N> &a, const Df<N> &s) { Df<N> r = 1.0f / s; return a * r; }
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
    HDI bool isZero(float prec) const { return allZero(val(), prec); }
};
template <int N> HDI Ds<N> operator+(const Ds<N> &a, const Ds<N> &b) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] + b.c[k]; return r; }
template <int N> HDI Ds<N> operator*(const Ds<N> &a, const Ds<N> &b) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * b.c[k]; return r; }
template <int N> HDI Ds<N> operator*(const Ds<N> &a, const Df<N> &s) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s; return r; }
template <int N> HDI Ds<N> operator*(const Ds<N> &a, float s) { Ds<N> r; for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s; return r; }
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
    HDI float next() {
        if (pos == 4) {
            u4 o = philox4x32_10(u4{s0, s1, blk++, 0x9e3779b9u}, k0, k1);
            b0 = o.a; b1 = o.b; b2 = o.c; b3 = o.d;
            pos = 0;
        }
        uint32_t x = pos == 0 ? b0 : pos == 1 ? b1 : pos == 2 ? b2 : b3;
        pos++;
        return (float)(x >> 8) * (1.0f / 16777216.0f);   // [0, 1) with 24 bits
    }
};

// ------------------------------------------------------------------- frames
// utils.cpp:54-67 — orthonormal basis built after a fixed rotation of the
// normal (0.1 rad about normalize(0.1,0.2,0.3)); R computed by Rodrigues on the host.
struct FrameRot { float m[9]; };
// Each kernel object defines its own copy under its own name.
#ifndef DT_FRAME_ROT
#define DT_FRAME_ROT c_frame_rot
#endif
extern __constant__ FrameRot DT_FRAME_ROT;

DI f3 rotMul(const float *m, f3 v, bool tr) {
    return tr ? f3{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z, m[2] * v.x + m[5] * v.y + m[8] * v.z}
              : f3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
template <int N> DI Dv<N> rotMul(const float *m, const Dv<N> &v, bool tr) {
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
        const float *R = DT_FRAME_ROT.m;
        f3 q = rotMul(R, nn, false);
        if (fabsf(q.x) > fabsf(q.y)) {
            float il = 1.0f / sqrtf(q.x * q.x + q.z * q.z);
            t = f3{q.z * il, 0.f, -q.x * il};
        } else {
            float il = 1.0f / sqrtf(q.y * q.y + q.z * q.z);
            t = f3{0.f, q.z * il, -q.y * il};
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
        const float *R = DT_FRAME_ROT.m;
        Dv<N> q = rotMul(R, nn, false);
        if (fabsf(q.x.v) > fabsf(q.y.v)) {
            Df<N> il = 1.0f / dsqrt(q.x * q.x + q.z * q.z);
            t = Dv<N>(q.z * il, Df<N>(0.f), -q.x * il);
        } else {
            Df<N> il = 1.0f / dsqrt(q.y * q.y + q.z * q.z);
            t = Dv<N>(Df<N>(0.f), q.z * il, -q.y * il);
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
DI f3 squareToCosineHemisphere(float sx, float sy) {
    float r1 = 2.0f * sx - 1.0f, r2 = 2.0f * sy - 1.0f, phi, r;
    if (r1 == 0.f && r2 == 0.f) { r = phi = 0.f; }
    else if (r1 * r1 > r2 * r2) { r = r1; phi = (PI / 4.0f) * (r2 / r1); }
    else { r = r2; phi = (PI / 2.0f) - (r1 / r2) * (PI / 4.0f); }
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float px = r * sp, py = r * cp;
    return f3{px, py, sqrtf(fmaxf(0.f, 1.0f - px * px - py * py))};
}
DI f3 squareToUniformSphere(float sx, float sy) {
    float z = 1.0f - 2.0f * sy, r = sqrtf(fmaxf(0.f, 1.0f - z * z)), sp, cp;
    sincosf(2.0f * PI * sx, &sp, &cp);
    return f3{r * cp, r * sp, z};
}
HDI float luminance(f3 v) { return 0.212671f * v.x + 0.715160f * v.y + 0.072169f * v.z; }

// utils.cpp:119-148
DI float fresnelDielectricExt(float ci_, float &ct_, float eta) {
    if (fabsf(eta - 1.0f) < EPS) { ct_ = -ci_; return 0.f; }
    float scale = ci_ > 0.f ? 1.0f / eta : eta;
    float ct2 = 1.0f - (1.0f - ci_ * ci_) * (scale * scale);
    if (ct2 < EPS) { ct_ = 0.f; return 1.0f; }
    float ci = fabsf(ci_), ct = sqrtf(ct2);
    float Rs = (ci - eta * ct) / (ci + eta * ct), Rp = (eta * ci - ct) / (eta * ci + ct);
    ct_ = ci_ > 0.f ? -ct : ct;
    return 0.5f * (Rs * Rs + Rp * Rp);
}
DI float fresnelDielectricExt(float ci, float eta) { float c; return fresnelDielectricExt(ci, c, eta); }
template <int N> DI Df<N> fresnelDielectricExtAD(const Df<N> &ci_, const Df<N> &eta) {   // utils.cpp:150-184
    if (fabsf(eta.v - 1.0f) < EPS) return Df<N>(0.f);
    Df<N> scale = ci_.v > 0.f ? 1.0f / eta : eta;
    Df<N> ct2 = 1.0f - (1.0f - ci_ * ci_) * (scale * scale);
    if (ct2.v < EPS) return Df<N>(1.0f);
    Df<N> ci = dabs(ci_), ct = dsqrt(ct2);
    Df<N> Rs = (ci - eta * ct) / (ci + eta * ct), Rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5f * (Rs * Rs + Rp * Rp);
}
DI f3 reflect(f3 wi, f3 n) { return 2.0f * dot(wi, n) * n - wi; }
DI f3 refract(f3 wi, f3 n, float eta, float ct) { if (ct < 0.f) eta = 1.0f / eta; return n * (dot(wi, n) * eta + ct) - wi * eta; }

// math_func.cpp:6-23 — Abramowitz & Stegun 7.1.26 (published coefficients)
DI float erf_as(float x) {
    float sign = signum(x);
    x = fabsf(x);
    float t = 1.0f / (1.0f + 0.3275911f * x);
    float y = 1.0f - (((((1.061405429f * t - 1.453152027f) * t) + 1.421413741f) * t - 0.284496736f) * t + 0.254829592f) * t * expf(-x * x);
    return sign * y;
}
// math_func.cpp:26-53 — M. Giles, "Approximating the erfinv function", GPU Computing Gems (2011)
DI float erfinv_giles(float x) {
    float w = -logf((1.0f - x) * (1.0f + x)), p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = 3.43273939e-07f + p * w;
        p = -3.5233877e-06f + p * w;
        p = -4.39150654e-06f + p * w;
        p = 0.00021858087f + p * w;
        p = -0.00125372503f + p * w;
        p = -0.00417768164f + p * w;
        p = 0.246640727f + p * w;
        p = 1.50140941f + p * w;
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = 0.000100950558f + p * w;
        p = 0.00134934322f + p * w;
        p = -0.00367342844f + p * w;
        p = 0.00573950773f + p * w;
        p = -0.0076224613f + p * w;
        p = 0.00943887047f + p * w;
        p = 1.00167406f + p * w;
        p = 2.83297682f + p * w;
    }
    return p * x;
}

// utils.cpp:218-254 — refinement of the chosen triangle (no bounds test; divisor
// clamped away from zero), primal and with tangents.
DI f3 rayTriangleUVT(f3 v0, f3 v1, f3 v2, f3 o, f3 d) {
    f3 e1 = v1 - v0, e2 = v2 - v0, pv = cross(d, e2);
    float div = dot(pv, e1);
    if (fabsf(div) < EPS) div = div > 0.f ? EPS : -EPS;
    float idv = 1.0f / div;
    f3 s = o - v0, q = cross(s, e1);
    return f3{dot(s, pv) * idv, dot(d, q) * idv, dot(e2, q) * idv};
}
template <int N>
DI void rayTriangleUVT_AD(const Dv<N> &v0, const Dv<N> &v1, const Dv<N> &v2, const Dv<N> &o, const Dv<N> &d, Df<N> &u,
                          Df<N> &v, Df<N> &t) {
    Dv<N> e1 = v1 - v0, e2 = v2 - v0, pv = dcross(d, e2);
    Df<N> div = ddot(pv, e1);
    if (fabsf(div.v) < EPS) div = Df<N>(div.v > 0.f ? EPS : -EPS);
    Df<N> idv = 1.0f / div;
    Dv<N> s = o - v0, q = dcross(s, e1);
    u = ddot(s, pv) * idv;
    v = ddot(d, q) * idv;
    t = ddot(e2, q) * idv;
}

}  // namespace dt
