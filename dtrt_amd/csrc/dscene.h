// Device scene services for the gfx950 path tracer: BVH closest-hit traversal
// (the role Embree plays in src/scene.cpp), hit refinement with tangents
// (src/shape.cpp:166-252), heterogeneous/homogeneous media (src/medium/*),
// BSDFs (src/bsdf/*), phase functions (include/phase.h), emitters
// (src/emitter/*), transmittance, emitter lookup through null interfaces and
// next-event estimation (src/scene.cpp:177-524). Everything is fp32.
#pragma once
#include "dmath.h"
#include "scene_types.h"

namespace dt {

// ------------------------------------------------------------ small helpers
DI f3 ld3(const float *p) { return f3{p[0], p[1], p[2]}; }
DI f3 ldg3(const float *p) { return f3{__ldg(p), __ldg(p + 1), __ldg(p + 2)}; }
template <int N> DI Dv<N> ldDv(const float *val, const float *tan) {
    Dv<N> r(ldg3(val));
#pragma unroll
    for (int j = 0; j < N; ++j) r.setGrad(j, ldg3(tan + 3 * j));
    return r;
}
DI f3 xform34(const float *m, f3 p) {
    return f3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
DI f3 xform33(const float *m, f3 p) {
    return f3{m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z, m[8] * p.x + m[9] * p.y + m[10] * p.z};
}

// Per-thread traversal stack in LDS, laid out [depth][lane] so that the lanes of a
// wave touch consecutive banks. Every kernel that traces rays runs blocks of
// exactly TRACE_BLOCK threads.
constexpr int TRACE_BLOCK = 64;
__shared__ int s_bvh_stack[DTRT_BVH_STACK * TRACE_BLOCK];

// ---------------------------------------------------------------- BVH
DI bool boxHit(const float4 &lo, const float4 &hi, f3 o, f3 inv, float tnear, float tfar, float &tentry) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tnear));
    float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tfar));
    tentry = t0;
    return t0 <= t1;
}

// Closest hit in [tnear, tfar]: global face id (or -1), its t, and the box and
// triangle tests done. Inlined at every call site: as an out-of-line call, the path
// kernels save their live path state to scratch around each traversal (measured:
// ~2.5% slower per render, with several GB of spill traffic per launch).
struct TraceHit {
    int face;
    float t;
    uint32_t nbox, ntri;
};

DI TraceHit traceClosest(const float4 *nodes, const float4 *tris, int nnodes, f3 o, f3 d, float tnear,
                                      float tfar) {
    TraceHit res{-1, tfar, 0u, 0u};
    if (nnodes == 0) return res;
    int *stk = s_bvh_stack + threadIdx.x;
    const float BEPS = 1e-6f;
    f3 inv = f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float best = tfar;
    int face = -1;
    int sp = 0;
    int node = 0;
    uint32_t nbox = 0, ntri = 0;
    while (true) {
        if (node >= 0) {
            const float4 *n = nodes + 4 * (size_t)node;
            float4 a = n[0], b = n[1], c = n[2], e = n[3];
            float tl, tr;
            nbox += 2;
            bool hl = boxHit(a, b, o, inv, tnear, best, tl);
            bool hr = boxHit(c, e, o, inv, tnear, best, tr);
            int cl = __float_as_int(a.w), cr = __float_as_int(b.w);
            if (hl && hr) {
                if (tr < tl) { int t = cl; cl = cr; cr = t; }
                stk[TRACE_BLOCK * sp++] = cr;
                node = cl;
                continue;
            } else if (hl) {
                node = cl;
                continue;
            } else if (hr) {
                node = cr;
                continue;
            }
        } else {
            int enc = ~node;
            int first = enc >> 4, cnt = enc & 15;
            ntri += cnt;
            for (int k = first; k < first + cnt; ++k) {
                const float4 *tp = tris + 3 * (size_t)k;
                float4 v0 = tp[0], e1 = tp[1], e2 = tp[2];
                f3 E1{e1.x, e1.y, e1.z}, E2{e2.x, e2.y, e2.z};
                f3 pv = cross(d, E2);
                float det = dot(E1, pv);
                if (det == 0.0f) continue;
                float idet = 1.0f / det;
                f3 s = o - f3{v0.x, v0.y, v0.z};
                float u = dot(s, pv) * idet;
                if (u < -BEPS || u > 1.0f + BEPS) continue;
                f3 q = cross(s, E1);
                float v = dot(d, q) * idet;
                if (v < -BEPS || u + v > 1.0f + BEPS) continue;
                float t = dot(E2, q) * idet;
                if (t >= tnear && t <= best) { best = t; face = __float_as_int(v0.w); }
            }
        }
        if (sp == 0) break;
        node = stk[TRACE_BLOCK * --sp];
    }
    res.face = face;
    res.t = best;
    res.nbox = nbox;
    res.ntri = ntri;
    return res;
}

// ------------------------------------------------------------ hit records
struct Hit {
    int face = -1;
    float t = FINF;
    f3 p, gn, sn, wi;
};
template <int N> struct HitAD {
    int face = -1;
    Df<N> t;
    Dv<N> p, gn, sn, wi;
    DI Hit val() const { Hit h; h.face = face; h.t = t.v; h.p = p.val(); h.gn = gn.val(); h.sn = sn.val(); h.wi = wi.val(); return h; }
};
DI const DShape &shapeOf(const DScene &S, int face) { return S.shapes[__ldg(S.face_shape + face)]; }

// shape.cpp:166-208
DI void refineHit(const DScene &S, int face, f3 o, f3 d, Hit &h) {
    const int *fi = S.F + 3 * (size_t)face;
    int i0 = __ldg(fi), i1 = __ldg(fi + 1), i2 = __ldg(fi + 2);
    f3 uvt = rayTriangleUVT(ldg3(S.V + 3 * (size_t)i0), ldg3(S.V + 3 * (size_t)i1), ldg3(S.V + 3 * (size_t)i2), o, d);
    float u = uvt.x, v = uvt.y, w = 1.0f - (u + v);
    f3 gn = ldg3(S.FN + 3 * (size_t)face), sn = gn;
    if (shapeOf(S, face).has_normals) {
        f3 nn = ldg3(S.Nrm + 3 * (size_t)i0) * w + ldg3(S.Nrm + 3 * (size_t)i1) * u + ldg3(S.Nrm + 3 * (size_t)i2) * v;
        sn = normalize(nn);
        if (dot(gn, sn) < 0.f) gn = -gn;
    }
    h.face = face;
    h.t = uvt.z;
    h.p = o + d * uvt.z;
    h.gn = gn;
    h.sn = sn;
    h.wi = Frame(sn).toLocal(-d);
}
// shape.cpp:210-252
template <int N> DI void refineHitAD(const DScene &S, int face, const Dv<N> &o, const Dv<N> &d, HitAD<N> &h) {
    const int nd = N;
    const int *fi = S.F + 3 * (size_t)face;
    int i0 = __ldg(fi), i1 = __ldg(fi + 1), i2 = __ldg(fi + 2);
    Df<N> u, v, t;
    rayTriangleUVT_AD<N>(ldDv<N>(S.V + 3 * (size_t)i0, S.Vd + (size_t)i0 * nd * 3), ldDv<N>(S.V + 3 * (size_t)i1, S.Vd + (size_t)i1 * nd * 3),
                         ldDv<N>(S.V + 3 * (size_t)i2, S.Vd + (size_t)i2 * nd * 3), o, d, u, v, t);
    Df<N> w = 1.0f - (u + v);
    Dv<N> gn = ldDv<N>(S.FN + 3 * (size_t)face, S.FNd + (size_t)face * nd * 3), sn = gn;
    if (shapeOf(S, face).has_normals) {
        Dv<N> nn = ldDv<N>(S.Nrm + 3 * (size_t)i0, S.Nd + (size_t)i0 * nd * 3) * w + ldDv<N>(S.Nrm + 3 * (size_t)i1, S.Nd + (size_t)i1 * nd * 3) * u +
                   ldDv<N>(S.Nrm + 3 * (size_t)i2, S.Nd + (size_t)i2 * nd * 3) * v;
        sn = dnormalize(nn);
        if (dot(gn.val(), sn.val()) < 0.f) gn = -gn;
    }
    h.face = face;
    h.t = t;
    h.p = o + d * t;
    h.gn = gn;
    h.sn = sn;
    h.wi = FrameAD<N>(sn).toLocal(-d);
}

DI int targetMedium(const DShape &sh, f3 gn, f3 d) { return dot(d, gn) > 0.f ? sh.med_ext : sh.med_int; }
DI bool isTransition(const DShape &sh) { return sh.med_ext >= 0 || sh.med_int >= 0; }
DI const DBsdf &bsdfOf(const DScene &S, const DShape &sh) { return S.bsdfs[sh.bsdf_id]; }

// -------------------------------------------------------------- media
// gridvolume.h:144-201 — node-centred trilinear lookup, zero outside [0, res-2] cells
DI float trilinear(const float *data, const int *res, f3 g, int ch, int nch) {
    float fx0 = floorf(g.x), fy0 = floorf(g.y), fz0 = floorf(g.z);
    // bounds are tested on the floats first so NaN / huge coordinates never reach the int conversion
    if (!(fx0 >= 0.f && fy0 >= 0.f && fz0 >= 0.f && fx0 < (float)(res[0] - 1) && fy0 < (float)(res[1] - 1) && fz0 < (float)(res[2] - 1)))
        return 0.f;
    int x1 = (int)fx0, y1 = (int)fy0, z1 = (int)fz0;
    float fx = g.x - fx0, fy = g.y - fy0, fz = g.z - fz0;
    size_t sy = (size_t)res[0] * nch, sz = sy * res[1];
    const float *b = data + (size_t)z1 * sz + (size_t)y1 * sy + (size_t)x1 * nch + ch;
    float d000 = __ldg(b), d001 = __ldg(b + nch), d010 = __ldg(b + sy), d011 = __ldg(b + sy + nch);
    float d100 = __ldg(b + sz), d101 = __ldg(b + sz + nch), d110 = __ldg(b + sz + sy), d111 = __ldg(b + sz + sy + nch);
    float _fx = 1.f - fx, _fy = 1.f - fy, _fz = 1.f - fz;
    return ((d000 * _fx + d001 * fx) * _fy + (d010 * _fx + d011 * fx) * fy) * _fz +
           ((d100 * _fx + d101 * fx) * _fy + (d110 * _fx + d111 * fx) * fy) * fz;
}
template <int N> DI Df<N> trilinearAD(const float *data, const int *res, const Dv<N> &g, int ch, int nch) {
    float fx0 = floorf(g.x.v), fy0 = floorf(g.y.v), fz0 = floorf(g.z.v);
    if (!(fx0 >= 0.f && fy0 >= 0.f && fz0 >= 0.f && fx0 < (float)(res[0] - 1) && fy0 < (float)(res[1] - 1) && fz0 < (float)(res[2] - 1)))
        return Df<N>(0.f);
    int x1 = (int)fx0, y1 = (int)fy0, z1 = (int)fz0;
    Df<N> fx = g.x - fx0, fy = g.y - fy0, fz = g.z - fz0;
    size_t sy = (size_t)res[0] * nch, sz = sy * res[1];
    const float *b = data + (size_t)z1 * sz + (size_t)y1 * sy + (size_t)x1 * nch + ch;
    float d000 = __ldg(b), d001 = __ldg(b + nch), d010 = __ldg(b + sy), d011 = __ldg(b + sy + nch);
    float d100 = __ldg(b + sz), d101 = __ldg(b + sz + nch), d110 = __ldg(b + sz + sy), d111 = __ldg(b + sz + sy + nch);
    // value and gradient of the trilinear interpolant, chained through g's tangents
    float vx0 = d000 + (d001 - d000) * fx.v, vx1 = d010 + (d011 - d010) * fx.v;
    float vx2 = d100 + (d101 - d100) * fx.v, vx3 = d110 + (d111 - d110) * fx.v;
    float vy0 = vx0 + (vx1 - vx0) * fy.v, vy1 = vx2 + (vx3 - vx2) * fy.v;
    float val = vy0 + (vy1 - vy0) * fz.v;
    float _fy = 1.f - fy.v, _fz = 1.f - fz.v;
    float gx = ((d001 - d000) * _fy + (d011 - d010) * fy.v) * _fz + ((d101 - d100) * _fy + (d111 - d110) * fy.v) * fz.v;
    float gy = (vx1 - vx0) * _fz + (vx3 - vx2) * fz.v;
    float gz = vy1 - vy0;
    Df<N> r(val);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = gx * g.x.d[j] + gy * g.y.d[j] + gz * g.z.d[j];
    return r;
}
DI f3 toGrid(const DMedium &m, f3 pw) {
    f3 pv = xform34(m.w2v, pw);
    return f3{(pv.x - m.bmin[0]) * m.gscale[0], (pv.y - m.bmin[1]) * m.gscale[1], (pv.z - m.bmin[2]) * m.gscale[2]};
}
DI float lookupDensity(const DMedium &m, f3 p) { return trilinear(m.density, m.res, toGrid(m, p), 0, 1); }
DI f3 lookupAlbedo(const DMedium &m, f3 p) {
    f3 pv = xform34(m.w2v, p);
    f3 g{(pv.x - m.amin[0]) * m.agscale[0], (pv.y - m.amin[1]) * m.agscale[1], (pv.z - m.amin[2]) * m.agscale[2]};
    return f3{trilinear(m.albedo_grid, m.ares, g, 0, 3), trilinear(m.albedo_grid, m.ares, g, 1, 3), trilinear(m.albedo_grid, m.ares, g, 2, 3)};
}
// gridvolume.h:301-326: tangent of the grid coordinate under rigid motion of the
// volume about its centre (the albedo grid carries no velocities in the reference).
template <int N> DI Dv<N> toGridAD(const DMedium &m, const Dv<N> &p, bool albedo) {
    f3 pv = xform34(m.w2v, p.val());
    f3 q = p.val() - ld3(m.center);
    const float *mn = albedo ? m.amin : m.bmin;
    const float *gs = albedo ? m.agscale : m.gscale;
    Dv<N> g(f3{(pv.x - mn[0]) * gs[0], (pv.y - mn[1]) * gs[1], (pv.z - mn[2]) * gs[2]});
#pragma unroll
    for (int j = 0; j < N; ++j) {
        f3 a = p.grad(j);
        if (!albedo) a = a - cross(ld3(m.rot_d[j]), q) - ld3(m.trans_d[j]);
        f3 b = xform33(m.w2v, a);
        g.setGrad(j, f3{b.x * gs[0], b.y * gs[1], b.z * gs[2]});
    }
    return g;
}
// Cell of a node-centred grid coordinate, with the bounds test of trilinear().
struct GridCell {
    int x, y, z;
    float fx, fy, fz;
};
DI bool gridCell(const int *res, f3 g, GridCell &c) {
    float fx0 = floorf(g.x), fy0 = floorf(g.y), fz0 = floorf(g.z);
    if (!(fx0 >= 0.f && fy0 >= 0.f && fz0 >= 0.f && fx0 < (float)(res[0] - 1) && fy0 < (float)(res[1] - 1) && fz0 < (float)(res[2] - 1)))
        return false;
    c.x = (int)fx0; c.y = (int)fy0; c.z = (int)fz0;
    c.fx = g.x - fx0; c.fy = g.y - fy0; c.fz = g.z - fz0;
    return true;
}

template <int N> DI Df<N> lookupDensityAD(const DMedium &m, const Dv<N> &p) {
    Df<N> r = trilinearAD<N>(m.density, m.res, toGridAD<N>(m, p, false), 0, 1);
    if (m.vox_any) {   // d(density at p) / d(node v) = trilinear weight of v at p
        GridCell c;
        if (gridCell(m.res, toGrid(m, p.val()), c)) {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const int v = m.vox_lane[j];
                if (v < 0) continue;
                const int dx = v % m.res[0] - c.x, dy = (v / m.res[0]) % m.res[1] - c.y, dz = v / (m.res[0] * m.res[1]) - c.z;
                if ((unsigned)dx > 1u || (unsigned)dy > 1u || (unsigned)dz > 1u) continue;
                r.d[j] += (dx ? c.fx : 1.f - c.fx) * (dy ? c.fy : 1.f - c.fy) * (dz ? c.fz : 1.f - c.fz);
            }
        }
    }
    return r;
}

// Reverse-mode hook of the medium estimators (path-replay adjoint). When set, every
// density-dependent weight of the estimator adds w * d(log weight)/d(node density)
// into grad[node] for the grid of medium *pm; the forward kernels pass none.
struct ScatterCtx {
    float *grad;
    const DMedium *pm;
    float w;
};
// grad[v] += w * (trilinear weight of node v at p) over the 8 nodes of p's cell
DI void scatterDensity(const DMedium &m, f3 p, float w, float *grad) {
    GridCell c;
    if (w == 0.f || !gridCell(m.res, toGrid(m, p), c)) return;
    const size_t sy = (size_t)m.res[0], sz = sy * m.res[1];
    float *b = grad + (size_t)c.z * sz + (size_t)c.y * sy + c.x;
    const float _fx = 1.f - c.fx, _fy = 1.f - c.fy, _fz = 1.f - c.fz;
    unsafeAtomicAdd(b, w * _fx * _fy * _fz);
    unsafeAtomicAdd(b + 1, w * c.fx * _fy * _fz);
    unsafeAtomicAdd(b + sy, w * _fx * c.fy * _fz);
    unsafeAtomicAdd(b + sy + 1, w * c.fx * c.fy * _fz);
    unsafeAtomicAdd(b + sz, w * _fx * _fy * c.fz);
    unsafeAtomicAdd(b + sz + 1, w * c.fx * _fy * c.fz);
    unsafeAtomicAdd(b + sz + sy, w * _fx * c.fy * c.fz);
    unsafeAtomicAdd(b + sz + sy + 1, w * c.fx * c.fy * c.fz);
}
template <int N> DI Ds<N> lookupAlbedoAD(const DMedium &m, const Dv<N> &p) {
    Dv<N> g = toGridAD<N>(m, p, true);
    Ds<N> r;
    for (int k = 0; k < 3; ++k) r.c[k] = trilinearAD<N>(m.albedo_grid, m.ares, g, k, 3);
    return r;
}
template <int N> DI Df<N> scaleAD(const DMedium &m) { Df<N> s(m.scale); for (int j = 0; j < N; ++j) s.d[j] = m.scale_d[j]; return s; }
template <int N> DI Df<N> sigmaTAD(const DMedium &m) { Df<N> s(m.sigma_t); for (int j = 0; j < N; ++j) s.d[j] = m.sigma_t_d[j]; return s; }
template <int N> DI Ds<N> albedoAD(const DMedium &m) {
    Ds<N> a(ld3(m.albedo));
    for (int j = 0; j < N; ++j) for (int k = 0; k < 3; ++k) a.c[k].d[j] = m.albedo_d[j][k];
    return a;
}

// aabb.h:186-220
DI bool aabbIntersect(const float *lo, const float *hi, f3 o, f3 d, float &nearT, float &farT) {
    nearT = -FINF;
    farT = FINF;
    for (int i = 0; i < 3; ++i) {
        float oi = o[i], di = d[i];
        if (di == 0.f) {
            if (oi < lo[i] || oi > hi[i]) return false;
        } else {
            float t1 = (lo[i] - oi) / di, t2 = (hi[i] - oi) / di;
            if (t1 > t2) { float t = t1; t1 = t2; t2 = t; }
            nearT = fmaxf(t1, nearT);
            farT = fminf(t2, farT);
            if (!(nearT <= farT)) return false;
        }
    }
    return true;
}

// heterogeneous.cpp:121-141 — derivative of the optical depth along a segment
// With `sc` (adjoint), each of the 10 points also scatters -w * len/10 * scale.
template <int N>
DI void intSigT(const DMedium &m, const Dv<N> &o, const Dv<N> &d, float tmin, const Df<N> &tmax, Rng &r, float *out,
                const ScatterCtx *scx = nullptr) {
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = 0.f;
    const float len = tmax.v - tmin;
    Df<N> sc = scaleAD<N>(m);
    for (int i = 0; i < 10; ++i) {
        float t = tmin + len * ((i + r.next()) * 0.1f);
        Dv<N> x = o + d * Df<N>(t);
        Df<N> s = sc * lookupDensityAD<N>(m, x);
        if (scx && &m == scx->pm) scatterDensity(m, x.val(), -scx->w * len * 0.1f * m.scale, scx->grad);
        r.nlk++;
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] += s.d[j];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] *= len * 0.1f;
    if (!tmax.gradZero(EPS)) {
        // fp32: the reference steps back 1e-7 from tmax; the value is sampled at
        // the same point with a step representable at fp32 resolution.
        float back = fmaxf(EPS, tmax.v * 1e-6f);
        r.nlk++;
        float l = lookupDensity(m, o.val() + d.val() * (tmax.v - back)) * m.scale;
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] += tmax.d[j] * l;
    }
}

// heterogeneous.cpp:9-37, homogeneous.cpp:6-24
DI bool sampleDistance(const DMedium &m, f3 o, f3 d, float tmax, float r0, float r1, Rng &r, f3 &ps, f3 &thr) {
    if (m.type == MED_HOMOGENEOUS) {
        if (m.sampling_weight > EPS) {
            float t = -logf(r0) / m.sigma_t;
            if (t < tmax) {
                ps = o + d * t;
                if (r1 < m.sampling_weight) thr = thr * (ld3(m.albedo) / m.sampling_weight);
                else thr = f3{0.f, 0.f, 0.f};
                return true;
            }
            return false;
        }
        thr = thr * expf(-tmax * m.sigma_t);
        return false;
    }
    float tN, tF;
    if (!aabbIntersect(m.aabb_min, m.aabb_max, o, d, tN, tF)) return false;
    tN = fmaxf(tN, 0.f);
    tF = fminf(tF, tmax);
    float t = tN;
    for (int dep = 0;; ++dep) {
        t -= logf(r.next()) * m.maxInvDensity;
        if (t > tF) break;
        if (dep > 10000) break;
        f3 p = o + d * t;
        r.nlk++;
        if (r.next() * m.maxDensity < lookupDensity(m, p)) {
            ps = p;
            thr = thr * (m.albedo_grid ? lookupAlbedo(m, p) : ld3(m.albedo));
            return true;
        }
    }
    return false;
}

// heterogeneous.cpp:40-75, homogeneous.cpp:27-49
// With `sc` (adjoint): on a real collision at p the weight's log-derivative is
// sigma_t'(p)/sigma_t(p) - int sigma_t', on a pass-through -int sigma_t'.
template <int N>
DI bool sampleDistanceAD(const DMedium &m, const Dv<N> &o, const Dv<N> &d, const Df<N> &tmax, float r0, float r1, Rng &r,
                         Dv<N> &ps, Ds<N> &thr, const ScatterCtx *scx = nullptr) {
    if (m.type == MED_HOMOGENEOUS) {
        Df<N> st = sigmaTAD<N>(m);
        if (m.sampling_weight > EPS) {
            float t = -logf(r0) / m.sigma_t;
            if (t < tmax.v) {
                ps = o + d * Df<N>(t);
                if (r1 < m.sampling_weight) {
                    Df<N> tmp = st * dexp(Df<N>(-t) * st);
                    thr = thr * (albedoAD<N>(m) * (tmp / (tmp.v * m.sampling_weight)));
                } else {
                    thr = Ds<N>();
                }
                return true;
            }
            Df<N> tmp = dexp(-tmax * st);
            thr = thr * (tmp / tmp.v);
            return false;
        }
        thr = thr * dexp(-tmax * st);
        return false;
    }
    float tN, tF;
    if (!aabbIntersect(m.aabb_min, m.aabb_max, o.val(), d.val(), tN, tF)) return false;
    tN = fmaxf(tN, 0.f);
    tF = fminf(tF, tmax.v);
    float t = tN;
    float is[N];
    for (int dep = 0;; ++dep) {
        t -= logf(r.next()) * m.maxInvDensity;
        if (t > tF) break;
        if (dep > 10000) break;
        Dv<N> p = o + d * Df<N>(t);
        Df<N> dens = lookupDensityAD<N>(m, p);
        r.nlk++;
        if (r.next() * m.maxDensity < dens.v) {
            ps = p;
            Ds<N> alb = m.albedo_grid ? lookupAlbedoAD<N>(m, p) : albedoAD<N>(m);
            Df<N> sigT = scaleAD<N>(m) * dens;
            if (scx && &m == scx->pm) scatterDensity(m, p.val(), scx->w / dens.v, scx->grad);
            intSigT<N>(m, o, d, 0.f, Df<N>(t), r, is, scx);
            // sigT * (1, -intSigT) / sigT.v
            Df<N> w(1.f);
            float iv = 1.0f / sigT.v;
#pragma unroll
            for (int j = 0; j < N; ++j) w.d[j] = sigT.d[j] * iv - is[j];
            thr = thr * (alb * w);
            return true;
        }
    }
    intSigT<N>(m, o, d, 0.f, tmax, r, is, scx);
    Df<N> w(1.f);
#pragma unroll
    for (int j = 0; j < N; ++j) w.d[j] = -is[j];
    thr = thr * w;
    return false;
}

// heterogeneous.cpp:78-90 (residual ratio tracking), homogeneous.cpp:52-54
DI float mediumTransmittance(const DMedium &m, f3 o, f3 d, float tmin, float tmax, Rng &r) {
    if (m.type == MED_HOMOGENEOUS) return expf((tmin - tmax) * m.sigma_t);
    float ret = 1.0f, t = tmin, inv = 1.0f / m.maxDensity;
    while (ret > EPS) {
        t -= logf(r.next()) * m.maxInvDensity;
        if (t > tmax) break;
        r.nlk++;
        ret *= 1.0f - lookupDensity(m, o + d * t) * inv;
    }
    return ret;
}
// With `sc` (adjoint): the estimate's log-derivative is -int sigma_t'.
template <int N>
DI Df<N> mediumTransmittanceAD(const DMedium &m, const Dv<N> &o, const Dv<N> &d, float tmin, const Df<N> &tmax, Rng &r,
                               const ScatterCtx *scx = nullptr) {
    if (m.type == MED_HOMOGENEOUS) return dexp((Df<N>(tmin) - tmax) * sigmaTAD<N>(m));
    Df<N> ret(mediumTransmittance(m, o.val(), d.val(), tmin, tmax.v, r));
    float is[N];
    intSigT<N>(m, o, d, tmin, tmax, r, is, scx);
#pragma unroll
    for (int j = 0; j < N; ++j) ret.d[j] = -ret.v * is[j];
    return ret;
}
DI f3 sigS(const DMedium &m, f3 x) {
    if (m.type == MED_HOMOGENEOUS) return ld3(m.albedo) * m.sigma_t;
    return ld3(m.albedo) * (m.scale * lookupDensity(m, x));
}

// -------------------------------------------------------------- phase
// phase.h:28-58 (sampled direction built in Frame(wi), as the reference does)
DI float phaseEval(const DPhase &ph, f3 wi, f3 wo) {
    if (ph.type == PHASE_ISO) return INV_FOURPI;
    float g = ph.g, tmp = 1.0f + g * g + 2.0f * g * dot(wi, wo);
    return INV_FOURPI * (1.0f - g * g) / (tmp * sqrtf(tmp));
}
template <int N> DI Df<N> phaseEvalAD(const DPhase &ph, const Dv<N> &wi, const f3 &wo) {
    if (ph.type == PHASE_ISO) return Df<N>(INV_FOURPI);
    Df<N> g(ph.g);
    for (int j = 0; j < N; ++j) g.d[j] = ph.g_d[j];
    Df<N> tmp = 1.0f + g * g + 2.0f * g * ddot(wi, wo);
    return INV_FOURPI * (1.0f - g * g) / dpow(tmp, 1.5f);
}
DI f3 phaseSample(const DPhase &ph, f3 wi, float u1, float u2) {
    if (ph.type == PHASE_ISO) return squareToUniformSphere(u1, u2);
    float g = ph.g, ct;
    if (fabsf(g) < EPS) ct = 1.0f - 2.0f * u1;
    else {
        float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    float st = sqrtf(fmaxf(0.f, 1.0f - ct * ct)), sp, cp;
    sincosf(2.0f * PI * u2, &sp, &cp);
    return Frame(wi).toWorld(f3{st * cp, st * sp, ct});
}
template <int N> DI Df<N> phaseSampleWeightAD(const DPhase &ph, const Dv<N> &wi, f3 wo) {
    if (ph.type == PHASE_ISO) return Df<N>(1.f);
    Df<N> e = phaseEvalAD<N>(ph, wi, wo);
    return e.v > EPS ? e / e.v : Df<N>(0.f);
}

// -------------------------------------------------------------- BSDFs
// microfacet.h:10-248 (Beckmann, Smith G1 rational fit, visible-normal sampling)
DI float mfEval(float alpha, f3 m) {
    if (m.z < EPS) return 0.f;
    float c2 = m.z * m.z, be = (m.x * m.x + m.y * m.y) / (alpha * alpha * c2);
    float r = expf(-be) / (PI * alpha * alpha * c2 * c2);
    return r * m.z < 1e-20f ? 0.f : r;
}
DI float tanThetaF(f3 v) { float tmp = 1.0f - v.z * v.z; return tmp < EPS ? 0.f : sqrtf(tmp) / v.z; }
DI float smithG1(float alpha, f3 v, f3 m) {
    if (dot(v, m) * v.z < EPS) return 0.f;
    float tt = fabsf(tanThetaF(v));
    if (tt == 0.f) return 1.f;
    float a = 1.0f / (alpha * tt);
    if (a >= 1.6f) return 1.f;
    float a2 = a * a;
    return (3.535f * a + 2.181f * a2) / (1.0f + 2.276f * a + 2.577f * a2);
}
DI float mfPdf(float alpha, f3 wi, f3 m) {
    if (fabsf(wi.z) < EPS) return 0.f;
    return smithG1(alpha, wi, m) * fabsf(dot(wi, m)) * mfEval(alpha, m) / fabsf(wi.z);
}
DI f3 mfSample(float alpha, f3 wi_, float s0, float s1) {
    f3 wi = normalize(f3{alpha * wi_.x, alpha * wi_.y, wi_.z});
    float theta = 0.f, phi = 0.f;
    if (wi.z < 0.99999f) { theta = acosf(wi.z); phi = atan2f(wi.y, wi.x); }
    float sP, cP;
    sincosf(phi, &sP, &cP);
    float sx, sy;
    const float SQRT_PI_INV = 0.56418958354775628695f;
    if (theta < 1e-4f) {
        float r = sqrtf(-logf(1.0f - s0)), sp, cp;
        sincosf(2.0f * PI * s1, &sp, &cp);
        sx = r * cp;
        sy = r * sp;
    } else {
        float tanI = tanf(theta), cotI = 1.0f / tanI;
        float a = -1.0f, c = erf_as(cotI), smp = fmaxf(s0, 1e-6f);
        float fit = 1.0f + theta * (-0.876f + theta * (0.4265f - 0.0594f * theta));
        float b = c - (1.0f + c) * powf(1.0f - smp, fit);
        float nrm = 1.0f / (1.0f + c + SQRT_PI_INV * tanI * expf(-cotI * cotI));
        for (int it = 1; it < 10; ++it) {
            if (!(b >= a && b <= c)) b = 0.5f * (a + c);
            float ie = erfinv_giles(b);
            float value = nrm * (1.0f + b + SQRT_PI_INV * tanI * expf(-ie * ie)) - smp;
            float deriv = nrm * (1.0f - ie * tanI);
            if (fabsf(value) < 1e-5f) break;
            if (value > 0.f) c = b; else a = b;
            b -= value / deriv;
        }
        sx = erfinv_giles(b);
        sy = erfinv_giles(2.0f * fmaxf(s1, 1e-6f) - 1.0f);
    }
    float rx = (cP * sx - sP * sy) * alpha, ry = (sP * sx + cP * sy) * alpha;
    float n = 1.0f / sqrtf(rx * rx + ry * ry + 1.0f);
    return f3{-rx * n, -ry * n, n};
}
template <int N> DI Df<N> mfEvalAD(const Df<N> &alpha, const Dv<N> &m) {
    if (m.z.v < EPS) return Df<N>(0.f);
    Df<N> c2 = m.z * m.z;
    Df<N> r = dexp(-((m.x * m.x + m.y * m.y) / (alpha * alpha * c2))) / (PI * alpha * alpha * c2 * c2);
    return r.v * m.z.v < 1e-20f ? Df<N>(0.f) : r;
}
template <int N> DI Df<N> smithG1AD(const Df<N> &alpha, const Dv<N> &v, const Dv<N> &m) {
    if (dot(v.val(), m.val()) * v.z.v < EPS) return Df<N>(0.f);
    Df<N> tmp = 1.0f - v.z * v.z;
    if (tmp.v < EPS) return Df<N>(1.f);
    Df<N> tt = dabs(dsqrt(tmp) / v.z);
    if (tt.v == 0.f) return Df<N>(1.f);
    Df<N> a = 1.0f / (alpha * tt);
    if (a.v >= 1.6f) return Df<N>(1.f);
    Df<N> a2 = a * a;
    return (3.535f * a + 2.181f * a2) / (1.0f + 2.276f * a + 2.577f * a2);
}
DI f3 phongReflect(f3 wi) { return f3{-wi.x, -wi.y, wi.z}; }
template <int N> DI Ds<N> specAD(const float *v, const float (*d)[3]) {
    Ds<N> r(ld3(v));
    for (int j = 0; j < N; ++j) for (int k = 0; k < 3; ++k) r.c[k].d[j] = d[j][k];
    return r;
}
template <int N> DI Df<N> scalarAD(float v, const float *d) { Df<N> r(v); for (int j = 0; j < N; ++j) r.d[j] = d[j]; return r; }

// diffuse.cpp:7-38, null.cpp:5-29, phong.cpp:19-95, roughdielectric.cpp:9-206
// importance: adjoint transport (light tracing); drops the radiance scaling of
// refraction, roughdielectric.cpp:50,156
DI f3 bsdfEval(const DBsdf &b, f3 wi, f3 wo, bool importance = false) {
    switch (b.type) {
        case BSDF_DIFFUSE:
            if (wi.z < EPS || wo.z < EPS) return f3{0, 0, 0};
            return ld3(b.refl) * (INV_PI * wo.z);
        case BSDF_NULL: return f3{1, 1, 1};
        case BSDF_PHONG: {
            if (wi.z < EPS || wo.z < EPS) return f3{0, 0, 0};
            f3 r{0, 0, 0};
            float al = dot(wo, phongReflect(wi));
            if (al > 0.f) r = ld3(b.ks) * ((b.ex + 2.0f) * powf(al, b.ex) * INV_TWOPI);
            return (r + ld3(b.kd) * INV_PI) * wo.z;
        }
        default: {
            if (fabsf(wi.z) < EPS || fabsf(wo.z) < EPS) return f3{0, 0, 0};
            bool refl = wi.z * wo.z > 0.f;
            float invEta = 1.0f / b.eta;
            f3 H = refl ? normalize(wo + wi) : normalize(wi + wo * (wi.z > 0.f ? b.eta : invEta));
            H = H * signum(H.z);
            float Dv_ = mfEval(b.alpha, H);
            if (fabsf(Dv_) < EPS) return f3{0, 0, 0};
            float F = fresnelDielectricExt(dot(wi, H), b.eta);
            float G = smithG1(b.alpha, wi, H) * smithG1(b.alpha, wo, H);
            if (refl) { float v = F * Dv_ * G / (4.0f * fabsf(wi.z)); return f3{v, v, v}; }
            float eta = wi.z > 0.f ? b.eta : invEta;
            float sd = dot(wi, H) + eta * dot(wo, H);
            float value = ((1.0f - F) * Dv_ * G * eta * eta * dot(wi, H) * dot(wo, H)) / (wi.z * sd * sd);
            float factor = importance ? 1.0f : (wi.z > 0.f ? invEta : b.eta);
            float v = fabsf(value * factor * factor);
            return f3{v, v, v};
        }
    }
}
DI float bsdfPdf(const DBsdf &b, f3 wi, f3 wo) {
    switch (b.type) {
        case BSDF_DIFFUSE: return (wi.z < EPS || wo.z < EPS) ? 0.f : INV_PI * wo.z;
        case BSDF_NULL: return 1.f;
        case BSDF_PHONG: {
            if (wi.z < EPS || wo.z < EPS) return 0.f;
            float dp = INV_PI * wo.z, spp = 0.f, al = dot(wo, phongReflect(wi));
            if (al > 0.f) spp = powf(al, b.ex) * (b.ex + 1.0f) / (2.0f * PI);
            return b.specW * spp + (1.0f - b.specW) * dp;
        }
        default: {
            bool refl = wi.z * wo.z > 0.f;
            float invEta = 1.0f / b.eta, dwh;
            f3 H;
            if (refl) { H = normalize(wi + wo); dwh = 1.0f / (4.0f * dot(wo, H)); }
            else {
                float eta = wi.z > 0.f ? b.eta : invEta;
                H = normalize(wi + wo * eta);
                float sd = dot(wi, H) + eta * dot(wo, H);
                dwh = (eta * eta * dot(wo, H)) / (sd * sd);
            }
            H = H * signum(H.z);
            float prob = mfPdf(b.alpha, wi * signum(wi.z), H);
            float F = fresnelDielectricExt(dot(wi, H), b.eta);
            prob *= refl ? F : (1.0f - F);
            return fabsf(prob * dwh);
        }
    }
}
DI f3 bsdfSample(const DBsdf &b, f3 wi, float r0, float r1, float r2, f3 &wo, float &pdf, bool importance = false) {
    switch (b.type) {
        case BSDF_DIFFUSE:
            if (wi.z < EPS) return f3{0, 0, 0};
            wo = squareToCosineHemisphere(r0, r1);
            pdf = INV_PI * wo.z;
            return ld3(b.refl);
        case BSDF_NULL:
            wo = -wi;
            pdf = 1.f;
            return f3{1, 1, 1};
        case BSDF_PHONG: {
            if (r2 < b.specW) {
                float sa = sqrtf(fmaxf(0.f, 1.0f - powf(r1, 2.0f / (b.ex + 1.0f)))), ca = powf(r1, 1.0f / (b.ex + 1.0f)), sp, cp;
                sincosf(2.0f * PI * r0, &sp, &cp);
                wo = Frame(phongReflect(wi)).toWorld(f3{sa * cp, sa * sp, ca});
                if (wo.z < EPS) return f3{0, 0, 0};
            } else {
                wo = squareToCosineHemisphere(r0, r1);
            }
            pdf = bsdfPdf(b, wi, wo);
            if (pdf < EPS) return f3{0, 0, 0};
            return bsdfEval(b, wi, wo) / pdf;
        }
        default: {
            float invEta = 1.0f / b.eta;
            f3 wis = wi * signum(wi.z);
            f3 m = mfSample(b.alpha, wis, r0, r1);
            float mpdf = mfPdf(b.alpha, wis, m);
            if (mpdf < EPS) return f3{0, 0, 0};
            pdf = mpdf;
            float ct;
            float F = fresnelDielectricExt(dot(wi, m), ct, b.eta);
            float weight = 1.f, dwh;
            bool srefl = !(r2 > F);
            pdf *= srefl ? F : (1.0f - F);
            if (srefl) {
                wo = reflect(wi, m);
                if (wi.z * wo.z < EPS) return f3{0, 0, 0};
                dwh = 1.0f / (4.0f * dot(wo, m));
            } else {
                if (fabsf(ct) < EPS) return f3{0, 0, 0};
                wo = refract(wi, m, b.eta, ct);
                float eta = ct < 0.f ? b.eta : invEta;
                if (wi.z * wo.z >= 0.f) return f3{0, 0, 0};
                float factor = importance ? 1.0f : (ct < 0.f ? invEta : b.eta);
                weight *= factor * factor;
                float sd = dot(wi, m) + eta * dot(wo, m);
                dwh = (eta * eta * dot(wo, m)) / (sd * sd);
            }
            weight *= smithG1(b.alpha, wo, m);
            pdf *= fabsf(dwh);
            return f3{weight, weight, weight};
        }
    }
}
template <int N> DI Ds<N> bsdfEvalAD(const DBsdf &b, const Dv<N> &wi, const Dv<N> &wo) {
    switch (b.type) {
        case BSDF_DIFFUSE:
            if (wi.z.v < EPS || wo.z.v < EPS) return Ds<N>();
            return specAD<N>(b.refl, b.refl_d) * (wo.z * INV_PI);
        case BSDF_NULL: return Ds<N>(f3{1, 1, 1});
        case BSDF_PHONG: {
            if (wi.z.v < EPS || wo.z.v < EPS) return Ds<N>();
            Ds<N> r;
            Df<N> al = ddot(wo, Dv<N>(-wi.x, -wi.y, wi.z));
            Df<N> e = scalarAD<N>(b.ex, b.ex_d);
            if (al.v > 0.f) r = specAD<N>(b.ks, b.ks_d) * ((e + 2.0f) * dpow(al, e) * INV_TWOPI);
            r += specAD<N>(b.kd, b.kd_d) * Df<N>(INV_PI);
            return r * wo.z;
        }
        default: {
            if (fabsf(wi.z.v) < EPS || fabsf(wo.z.v) < EPS) return Ds<N>();
            Df<N> alpha = scalarAD<N>(b.alpha, b.alpha_d), eta = scalarAD<N>(b.eta, b.eta_d), invEta = 1.0f / eta;
            bool refl = wi.z.v * wo.z.v > 0.f;
            Dv<N> H = refl ? dnormalize(wo + wi) : dnormalize(wi + wo * (wi.z.v > 0.f ? eta : invEta));
            H = H * signum(H.z.v);
            Df<N> Dm = mfEvalAD<N>(alpha, H);
            if (fabsf(Dm.v) < EPS) return Ds<N>();
            Df<N> F = fresnelDielectricExtAD<N>(ddot(wi, H), eta);
            Df<N> G = smithG1AD<N>(alpha, wi, H) * smithG1AD<N>(alpha, wo, H);
            if (refl) return Ds<N>::fill(F * Dm * G / (4.0f * dabs(wi.z)));
            const Df<N> &e = wi.z.v > 0.f ? eta : invEta;
            Df<N> sd = ddot(wi, H) + e * ddot(wo, H);
            Df<N> value = ((1.0f - F) * Dm * G * e * e * ddot(wi, H) * ddot(wo, H)) / (wi.z * sd * sd);
            Df<N> factor = wi.z.v > 0.f ? invEta : eta;
            return Ds<N>::fill(dabs(value * factor * factor));
        }
    }
}

// ------------------------------------------------------------- emitters
// area.cpp:5-15, area2.cpp:5-27
DI f3 emitterEval(const DEmitter &e, f3 n_geo, f3 n_sh, f3 d) {
    if (e.type == EMIT_AREA) return (e.two_sided || dot(n_geo, d) > 0.f) ? ld3(e.I) : f3{0, 0, 0};
    float tmp = dot(n_sh, d);
    return tmp > EPS ? ld3(e.I) * (e.kappa * INV_TWOPI * expf(e.kappa * (tmp - 1.0f))) : f3{0, 0, 0};
}
DI f3 emitterEvalNorm(const DEmitter &e, f3 n, f3 d) { return emitterEval(e, n, n, d); }
template <int N> DI Ds<N> emitterEvalAD(const DEmitter &e, const HitAD<N> &h, const Dv<N> &d) {
    if (e.type == EMIT_AREA) return (e.two_sided || dot(h.gn.val(), d.val()) > 0.f) ? Ds<N>(ld3(e.I)) : Ds<N>();
    Df<N> tmp = ddot(h.sn, d);
    if (tmp.v > EPS) {
        Df<N> k = scalarAD<N>(e.kappa, e.kappa_d);
        return specAD<N>(e.I, e.I_d) * (dexp(k * (tmp - 1.0f)) * k * INV_TWOPI);
    }
    return Ds<N>();
}

// area.cpp:17-19, area2.cpp:29-35 — evaluation against an explicit (AD) normal
template <int N> DI Ds<N> emitterEvalNormAD(const DEmitter &e, const Dv<N> &n, const Dv<N> &d) {
    if (e.type == EMIT_AREA) return (e.two_sided || dot(n.val(), d.val()) > 0.f) ? Ds<N>(ld3(e.I)) : Ds<N>();
    Df<N> tmp = ddot(n, d);
    if (tmp.v > EPS) {
        Df<N> k = scalarAD<N>(e.kappa, e.kappa_d);
        return specAD<N>(e.I, e.I_d) * (dexp(k * (tmp - 1.0f)) * k * INV_TWOPI);
    }
    return Ds<N>();
}

// ---------------------------------------------------- distributions
// pmf.h:103-167 — lower_bound on an unnormalised-then-normalised CDF, sample reuse
DI int sampleReuse(const double *cdf, int n, float &u, float *pdf) {
    int lo = 0, hi = n + 1;
    double uu = u;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (cdf[mid] < uu) lo = mid + 1; else hi = mid; }
    int idx = min(n - 1, max(0, lo - 1));
    while (idx + 1 < n && cdf[idx + 1] - cdf[idx] == 0.0) ++idx;
    double a = cdf[idx], b = cdf[idx + 1];
    if (pdf) *pdf = (float)(b - a);
    u = (float)((uu - a) / (b - a));
    u = fminf(fmaxf(u, 0.f), 1.f);
    return idx;
}
DI int upperBound(const double *cdf, int n, float u) {
    int lo = 0, hi = n;
    double uu = u;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (cdf[mid] <= uu) lo = mid + 1; else hi = mid; }
    return min(max(lo, 0), n - 1);
}

}  // namespace dt
