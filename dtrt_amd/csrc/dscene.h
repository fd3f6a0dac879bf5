// Device scene services for the gfx950 path tracer: BVH closest-hit traversal
// (the role Embree plays in src/scene.cpp), hit refinement with tangents
// (src/shape.cpp:166-252), heterogeneous/homogeneous media (src/medium/*),
// BSDFs (src/bsdf/*), phase functions (include/phase.h), emitters
// (src/emitter/*), transmittance, emitter lookup through null interfaces and
// next-event estimation (src/scene.cpp:177-524). Scene data is read as fp32 and
// computed on in `real`; traversal is fp32.
#pragma once
#include "dmath.h"
#include "scene_types.h"

#include <climits>

namespace dt {
namespace DT_PNS {

// ------------------------------------------------------------ small helpers
// (loads of fp32 scene data, widened to `real`)
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

// Per-thread traversal stack in LDS, laid out [depth][thread of the block] so that the
// lanes of a wave touch consecutive banks. Its size follows the work-group size: every
// kernel that traces rays is launched with traceLds(block, entries) bytes of dynamic
// LDS, entries = DTRT_BVH_STACK or, in the kernels that may keep the tree in LDS after
// the stacks, the scene's stack_depth.
constexpr int TRACE_BLOCK = 64;   // the default work-group size (one wave)
constexpr size_t traceLds(unsigned block, int entries = DTRT_BVH_STACK) { return (size_t)entries * block * sizeof(int); }
extern __shared__ int s_bvh_stack[];

// ---------------------------------------------------------------- BVH (fp32)
struct fv3 {
    float x, y, z;
};
DI fv3 fsub(fv3 a, fv3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
DI float fdot(fv3 a, fv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DI fv3 fcross(fv3 a, fv3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

DI bool boxHit(const float4 &lo, const float4 &hi, fv3 o, fv3 inv, float tnear, float tfar, float &tentry) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tnear));
    float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tfar));
    tentry = t0;
    return t0 <= t1;
}

// Closest hit in [tnear, tfar] of the fp32-rounded ray: global face id (or -1), its
// t, and the box and triangle tests done. Inlined at every call site: as an
// out-of-line call, the path kernels save their live path state to scratch around
// each traversal (measured: ~2.5% slower per render, with several GB of spill
// traffic per launch). Nested loops (inner nodes, then the leaf reached): a wave runs
// its lanes' inner-node steps together and then tests the leaves they reached
// together, instead of running both branches in every step that some lane spends at
// a leaf; nodes are visited in the same order, so hits are the same.
struct TraceHit {
    int face;
    float t;
    uint32_t nbox, ntri;
};

// Part of the tree can live in LDS after the traversal stack: its first lds_nodes nodes
// (the builder numbers the top of the tree breadth-first, so these are its top levels)
// and, when all nodes fit, also the leaf-ordered triangle records, at LDS float4
// 4 * nnodes. treeLdsBytes gives the bytes; the launch chooses lds_nodes.
// Only kernels built with treePrefixInLds<N>() read a part of the nodes from LDS: their
// traversal tests each node's index against lds_nodes, which the others save by
// reading LDS only when the whole tree is there (lds_tris), a launch-uniform test.
// Measured back to back on one GPU: the N = 3 benchmark scene (whole tree in LDS) renders in 0.652 s
// with the uniform test and 0.670 s with the per-node one; the 100K-face dielectric
// scene at N = 8 in 7.02 s without and 6.84 s with a prefix of its tree in LDS.
template <int N> constexpr bool treePrefixInLds() { return N >= 8; }
HDI size_t treeLdsBytes(const DScene &S, int lds_nodes, bool lds_tris) {
    return (size_t)lds_nodes * 64 + (lds_tris ? (size_t)S.ntris * 48 : 0);
}
// The tree's LDS copy starts after stack_depth stack entries per thread.
DI const float4 *ldsTree(int stack_depth) { return reinterpret_cast<const float4 *>(s_bvh_stack + stack_depth * blockDim.x); }
// Copies that part into LDS; called by every thread of the block, at the top of a
// kernel launched with treeLdsBytes more bytes of dynamic LDS.
DI void treeToLds(const DScene &S, int lds_nodes, bool lds_tris) {
    float4 *dst = reinterpret_cast<float4 *>(s_bvh_stack + S.stack_depth * blockDim.x);
    const int nn = 4 * lds_nodes, nt = lds_tris ? 3 * S.ntris : 0;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) dst[i] = S.nodes[i];
    for (int i = threadIdx.x; i < nt; i += blockDim.x) dst[nn + i] = S.tris[i];
    __syncthreads();
}

template <bool PREFIX = false>
DI TraceHit traceClosest(const float4 *nodes, const float4 *tris, int nnodes, f3 o_, f3 d_, real tnear_, real tfar_,
                         int lds_nodes = 0, bool lds_tris = false, int stack_depth = DTRT_BVH_STACK) {
    const fv3 o{(float)o_.x, (float)o_.y, (float)o_.z}, d{(float)d_.x, (float)d_.y, (float)d_.z};
    const float tnear = (float)tnear_, tfar = (float)tfar_;
    TraceHit res{-1, tfar, 0u, 0u};
    if (nnodes == 0) return res;
    int *stk = s_bvh_stack + threadIdx.x;
    const unsigned stride = blockDim.x;
    const float BEPS = 1e-6f;
    const fv3 inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float best = tfar;
    int face = -1;
    int sp = 0;
    int node = 0;
    uint32_t nbox = 0, ntri = 0;
    constexpr int DONE = INT_MIN;   // no leaf encodes to it: leaves hold >= 1 triangle of < 2^27
    while (node != DONE) {
        // inner nodes until this lane reaches a leaf or runs out of nodes; the wave then
        // tests the leaves it reached together
        while (node >= 0) {
            float4 a, b, c, e;
            if (PREFIX ? node < lds_nodes : lds_tris) {
                const float4 *n = ldsTree(stack_depth) + 4 * node;
                a = n[0]; b = n[1]; c = n[2]; e = n[3];
            } else {
                const float4 *n = nodes + 4 * (size_t)node;
                a = n[0]; b = n[1]; c = n[2]; e = n[3];
            }
            float tl, tr;
            nbox += 2;
            bool hl = boxHit(a, b, o, inv, tnear, best, tl);
            bool hr = boxHit(c, e, o, inv, tnear, best, tr);
            int cl = __float_as_int(a.w), cr = __float_as_int(b.w);
            if (hl && hr) {
                if (tr < tl) { int t = cl; cl = cr; cr = t; }
                stk[stride * sp++] = cr;
                node = cl;
            } else if (hl) {
                node = cl;
            } else if (hr) {
                node = cr;
            } else {
                node = sp == 0 ? DONE : stk[stride * --sp];
            }
        }
        if (node == DONE) break;
        int enc = ~node;
        int first = enc >> 4, cnt = enc & 15;
        ntri += cnt;
        for (int k = first; k < first + cnt; ++k) {
            float4 v0, e1, e2;
            if (lds_tris) {
                const float4 *tp = ldsTree(stack_depth) + 4 * nnodes + 3 * k;
                v0 = tp[0]; e1 = tp[1]; e2 = tp[2];
            } else {
                const float4 *tp = tris + 3 * (size_t)k;
                v0 = tp[0]; e1 = tp[1]; e2 = tp[2];
            }
            fv3 E1{e1.x, e1.y, e1.z}, E2{e2.x, e2.y, e2.z};
            fv3 pv = fcross(d, E2);
            float det = fdot(E1, pv);
            if (det == 0.0f) continue;
            float idet = 1.0f / det;
            fv3 s = fsub(o, fv3{v0.x, v0.y, v0.z});
            float u = fdot(s, pv) * idet;
            if (u < -BEPS || u > 1.0f + BEPS) continue;
            fv3 q = fcross(s, E1);
            float v = fdot(d, q) * idet;
            if (v < -BEPS || u + v > 1.0f + BEPS) continue;
            float t = fdot(E2, q) * idet;
            if (t >= tnear && t <= best) { best = t; face = __float_as_int(v0.w); }
        }
        node = sp == 0 ? DONE : stk[stride * --sp];
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
    real t = FINF;
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
    real u = uvt.x, v = uvt.y, w = 1.0_r - (u + v);
    f3 gn = ldg3(S.FN + 3 * (size_t)face), sn = gn;
    if (shapeOf(S, face).has_normals) {
        f3 nn = ldg3(S.Nrm + 3 * (size_t)i0) * w + ldg3(S.Nrm + 3 * (size_t)i1) * u + ldg3(S.Nrm + 3 * (size_t)i2) * v;
        sn = normalize(nn);
        if (dot(gn, sn) < 0._r) gn = -gn;
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
    Df<N> w = 1.0_r - (u + v);
    Dv<N> gn = ldDv<N>(S.FN + 3 * (size_t)face, S.FNd + (size_t)face * nd * 3), sn = gn;
    if (shapeOf(S, face).has_normals) {
        Dv<N> nn = ldDv<N>(S.Nrm + 3 * (size_t)i0, S.Nd + (size_t)i0 * nd * 3) * w + ldDv<N>(S.Nrm + 3 * (size_t)i1, S.Nd + (size_t)i1 * nd * 3) * u +
                   ldDv<N>(S.Nrm + 3 * (size_t)i2, S.Nd + (size_t)i2 * nd * 3) * v;
        sn = dnormalize(nn);
        if (dot(gn.val(), sn.val()) < 0._r) gn = -gn;
    }
    h.face = face;
    h.t = t;
    h.p = o + d * t;
    h.gn = gn;
    h.sn = sn;
    h.wi = FrameAD<N>(sn).toLocal(-d);
}

// The distance of refineHitAD alone, with the geometric normal's value oriented as
// there (towards the shading normal's side): what a walk through a null interface needs.
template <int N> DI Df<N> hitDistanceAD(const DScene &S, int face, const Dv<N> &o, const Dv<N> &d, f3 &gn) {
    const int nd = N;
    const int *fi = S.F + 3 * (size_t)face;
    int i0 = __ldg(fi), i1 = __ldg(fi + 1), i2 = __ldg(fi + 2);
    Df<N> u, v, t;
    rayTriangleUVT_AD<N>(ldDv<N>(S.V + 3 * (size_t)i0, S.Vd + (size_t)i0 * nd * 3), ldDv<N>(S.V + 3 * (size_t)i1, S.Vd + (size_t)i1 * nd * 3),
                         ldDv<N>(S.V + 3 * (size_t)i2, S.Vd + (size_t)i2 * nd * 3), o, d, u, v, t);
    gn = ldg3(S.FN + 3 * (size_t)face);
    if (shapeOf(S, face).has_normals) {
        const real w = 1.0_r - (u.v + v.v);
        const f3 nn = ldg3(S.Nrm + 3 * (size_t)i0) * w + ldg3(S.Nrm + 3 * (size_t)i1) * u.v + ldg3(S.Nrm + 3 * (size_t)i2) * v.v;
        if (dot(gn, nn) < 0._r) gn = -gn;
    }
    return t;
}

DI int targetMedium(const DShape &sh, f3 gn, f3 d) { return dot(d, gn) > 0._r ? sh.med_ext : sh.med_int; }
DI bool isTransition(const DShape &sh) { return sh.med_ext >= 0 || sh.med_int >= 0; }
DI const DBsdf &bsdfOf(const DScene &S, const DShape &sh) { return S.bsdfs[sh.bsdf_id]; }

// -------------------------------------------------------------- media
// gridvolume.h:144-201 — node-centred trilinear lookup, zero outside [0, res-2] cells
DI real trilinear(const float *data, const int *res, f3 g, int ch, int nch) {
    real fx0 = floor(g.x), fy0 = floor(g.y), fz0 = floor(g.z);
    // bounds are tested on the floats first so NaN / huge coordinates never reach the int conversion
    if (!(fx0 >= 0._r && fy0 >= 0._r && fz0 >= 0._r && fx0 < (real)(res[0] - 1) && fy0 < (real)(res[1] - 1) && fz0 < (real)(res[2] - 1)))
        return 0._r;
    int x1 = (int)fx0, y1 = (int)fy0, z1 = (int)fz0;
    real fx = g.x - fx0, fy = g.y - fy0, fz = g.z - fz0;
    size_t sy = (size_t)res[0] * nch, sz = sy * res[1];
    const float *b = data + (size_t)z1 * sz + (size_t)y1 * sy + (size_t)x1 * nch + ch;
    real d000 = __ldg(b), d001 = __ldg(b + nch), d010 = __ldg(b + sy), d011 = __ldg(b + sy + nch);
    real d100 = __ldg(b + sz), d101 = __ldg(b + sz + nch), d110 = __ldg(b + sz + sy), d111 = __ldg(b + sz + sy + nch);
    real _fx = 1._r - fx, _fy = 1._r - fy, _fz = 1._r - fz;
    return ((d000 * _fx + d001 * fx) * _fy + (d010 * _fx + d011 * fx) * fy) * _fz +
           ((d100 * _fx + d101 * fx) * _fy + (d110 * _fx + d111 * fx) * fy) * fz;
}
template <int N> DI Df<N> trilinearAD(const float *data, const int *res, const Dv<N> &g, int ch, int nch) {
    real fx0 = floor(g.x.v), fy0 = floor(g.y.v), fz0 = floor(g.z.v);
    if (!(fx0 >= 0._r && fy0 >= 0._r && fz0 >= 0._r && fx0 < (real)(res[0] - 1) && fy0 < (real)(res[1] - 1) && fz0 < (real)(res[2] - 1)))
        return Df<N>(0._r);
    int x1 = (int)fx0, y1 = (int)fy0, z1 = (int)fz0;
    Df<N> fx = g.x - fx0, fy = g.y - fy0, fz = g.z - fz0;
    size_t sy = (size_t)res[0] * nch, sz = sy * res[1];
    const float *b = data + (size_t)z1 * sz + (size_t)y1 * sy + (size_t)x1 * nch + ch;
    real d000 = __ldg(b), d001 = __ldg(b + nch), d010 = __ldg(b + sy), d011 = __ldg(b + sy + nch);
    real d100 = __ldg(b + sz), d101 = __ldg(b + sz + nch), d110 = __ldg(b + sz + sy), d111 = __ldg(b + sz + sy + nch);
    // value and gradient of the trilinear interpolant, chained through g's tangents
    real vx0 = d000 + (d001 - d000) * fx.v, vx1 = d010 + (d011 - d010) * fx.v;
    real vx2 = d100 + (d101 - d100) * fx.v, vx3 = d110 + (d111 - d110) * fx.v;
    real vy0 = vx0 + (vx1 - vx0) * fy.v, vy1 = vx2 + (vx3 - vx2) * fy.v;
    real val = vy0 + (vy1 - vy0) * fz.v;
    real _fy = 1._r - fy.v, _fz = 1._r - fz.v;
    real gx = ((d001 - d000) * _fy + (d011 - d010) * fy.v) * _fz + ((d101 - d100) * _fy + (d111 - d110) * fy.v) * fz.v;
    real gy = (vx1 - vx0) * _fz + (vx3 - vx2) * fz.v;
    real gz = vy1 - vy0;
    Df<N> r(val);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = gx * g.x.d[j] + gy * g.y.d[j] + gz * g.z.d[j];
    return r;
}
DI f3 toGrid(const DMedium &m, f3 pw) {
    f3 pv = xform34(m.w2v, pw);
    return f3{(pv.x - m.bmin[0]) * m.gscale[0], (pv.y - m.bmin[1]) * m.gscale[1], (pv.z - m.bmin[2]) * m.gscale[2]};
}
DI real lookupDensity(const DMedium &m, f3 p) { return trilinear(m.density, m.res, toGrid(m, p), 0, 1); }
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
    real fx, fy, fz;
};
DI bool gridCell(const int *res, f3 g, GridCell &c) {
    real fx0 = floor(g.x), fy0 = floor(g.y), fz0 = floor(g.z);
    if (!(fx0 >= 0._r && fy0 >= 0._r && fz0 >= 0._r && fx0 < (real)(res[0] - 1) && fy0 < (real)(res[1] - 1) && fz0 < (real)(res[2] - 1)))
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
                r.d[j] += (dx ? c.fx : 1._r - c.fx) * (dy ? c.fy : 1._r - c.fy) * (dz ? c.fz : 1._r - c.fz);
            }
        }
    }
    return r;
}

// Reverse-mode hook of the medium estimators (path-replay adjoint). When set, every
// density-dependent weight of the estimator adds w * d(log weight)/d(node density)
// into grad[node] for the grid of medium *pm; the forward kernels pass none.
struct ScatterCtx {
    float *grad;   // fp32 gradient grid
    const DMedium *pm;
    real w;
};
// grad[v] += w * (trilinear weight of node v at p) over the 8 nodes of p's cell
DI void scatterDensity(const DMedium &m, f3 p, real w, float *grad) {
    GridCell c;
    if (w == 0._r || !gridCell(m.res, toGrid(m, p), c)) return;
    const size_t sy = (size_t)m.res[0], sz = sy * m.res[1];
    float *b = grad + (size_t)c.z * sz + (size_t)c.y * sy + c.x;
    const real _fx = 1._r - c.fx, _fy = 1._r - c.fy, _fz = 1._r - c.fz;
    unsafeAtomicAdd(b, (float)(w * _fx * _fy * _fz));
    unsafeAtomicAdd(b + 1, (float)(w * c.fx * _fy * _fz));
    unsafeAtomicAdd(b + sy, (float)(w * _fx * c.fy * _fz));
    unsafeAtomicAdd(b + sy + 1, (float)(w * c.fx * c.fy * _fz));
    unsafeAtomicAdd(b + sz, (float)(w * _fx * _fy * c.fz));
    unsafeAtomicAdd(b + sz + 1, (float)(w * c.fx * _fy * c.fz));
    unsafeAtomicAdd(b + sz + sy, (float)(w * _fx * c.fy * c.fz));
    unsafeAtomicAdd(b + sz + sy + 1, (float)(w * c.fx * c.fy * c.fz));
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
DI bool aabbIntersect(const float *lo, const float *hi, f3 o, f3 d, real &nearT, real &farT) {
    nearT = -FINF;
    farT = FINF;
    for (int i = 0; i < 3; ++i) {
        real oi = o[i], di = d[i];
        if (di == 0._r) {
            if (oi < lo[i] || oi > hi[i]) return false;
        } else {
            real t1 = (lo[i] - oi) / di, t2 = (hi[i] - oi) / di;
            if (t1 > t2) { real t = t1; t1 = t2; t2 = t; }
            nearT = fmax(t1, nearT);
            farT = fmin(t2, farT);
            if (!(nearT <= farT)) return false;
        }
    }
    return true;
}

// heterogeneous.cpp:121-141 — derivative of the optical depth along a segment
// With `sc` (adjoint), each of the 10 points also scatters -w * len/10 * scale.
template <int N>
DI void intSigT(const DMedium &m, const Dv<N> &o, const Dv<N> &d, real tmin, const Df<N> &tmax, Rng &r, real *out,
                const ScatterCtx *scx = nullptr) {
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = 0._r;
    const real len = tmax.v - tmin;
    Df<N> sc = scaleAD<N>(m);
    for (int i = 0; i < 10; ++i) {
        real t = tmin + len * ((i + r.next()) * 0.1_r);
        Dv<N> x = o + d * Df<N>(t);
        Df<N> s = sc * lookupDensityAD<N>(m, x);
        if (scx && &m == scx->pm) scatterDensity(m, x.val(), -scx->w * len * 0.1_r * m.scale, scx->grad);
        r.nlk++;
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] += s.d[j];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] *= len * 0.1_r;
    if (!tmax.gradZero(EPS)) {
        // The reference steps back Epsilon = 1e-7 from tmax; in fp32 the step is
        // widened to one that is representable at tmax's magnitude.
        const real back = kF64 ? EPS : fmax(EPS, tmax.v * 1e-6_r);
        r.nlk++;
        real l = lookupDensity(m, o.val() + d.val() * (tmax.v - back)) * m.scale;
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] += tmax.d[j] * l;
    }
}

// heterogeneous.cpp:9-37, homogeneous.cpp:6-24
DI bool sampleDistance(const DMedium &m, f3 o, f3 d, real tmax, real r0, real r1, Rng &r, f3 &ps, f3 &thr) {
    if (m.type == MED_HOMOGENEOUS) {
        if (m.sampling_weight > EPS) {
            real t = -log(r0) / m.sigma_t;
            if (t < tmax) {
                ps = o + d * t;
                if (r1 < m.sampling_weight) thr = thr * (ld3(m.albedo) / m.sampling_weight);
                else thr = f3{0._r, 0._r, 0._r};
                return true;
            }
            return false;
        }
        thr = thr * exp(-tmax * m.sigma_t);
        return false;
    }
    real tN, tF;
    if (!aabbIntersect(m.aabb_min, m.aabb_max, o, d, tN, tF)) return false;
    tN = fmax(tN, 0._r);
    tF = fmin(tF, tmax);
    real t = tN;
    for (int dep = 0;; ++dep) {
        t -= log(r.next()) * m.maxInvDensity;
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
DI bool sampleDistanceAD(const DMedium &m, const Dv<N> &o, const Dv<N> &d, const Df<N> &tmax, real r0, real r1, Rng &r,
                         Dv<N> &ps, Ds<N> &thr, const ScatterCtx *scx = nullptr) {
    if (m.type == MED_HOMOGENEOUS) {
        Df<N> st = sigmaTAD<N>(m);
        if (m.sampling_weight > EPS) {
            real t = -log(r0) / m.sigma_t;
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
    real tN, tF;
    if (!aabbIntersect(m.aabb_min, m.aabb_max, o.val(), d.val(), tN, tF)) return false;
    tN = fmax(tN, 0._r);
    tF = fmin(tF, tmax.v);
    real t = tN;
    real is[N];
    for (int dep = 0;; ++dep) {
        t -= log(r.next()) * m.maxInvDensity;
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
            intSigT<N>(m, o, d, 0._r, Df<N>(t), r, is, scx);
            // sigT * (1, -intSigT) / sigT.v
            Df<N> w(1._r);
            real iv = 1.0_r / sigT.v;
#pragma unroll
            for (int j = 0; j < N; ++j) w.d[j] = sigT.d[j] * iv - is[j];
            thr = thr * (alb * w);
            return true;
        }
    }
    intSigT<N>(m, o, d, 0._r, tmax, r, is, scx);
    Df<N> w(1._r);
#pragma unroll
    for (int j = 0; j < N; ++j) w.d[j] = -is[j];
    thr = thr * w;
    return false;
}

// heterogeneous.cpp:78-90 (residual ratio tracking), homogeneous.cpp:52-54
DI real mediumTransmittance(const DMedium &m, f3 o, f3 d, real tmin, real tmax, Rng &r) {
    if (m.type == MED_HOMOGENEOUS) return exp((tmin - tmax) * m.sigma_t);
    real ret = 1.0_r, t = tmin, inv = 1.0_r / m.maxDensity;
    while (ret > EPS) {
        t -= log(r.next()) * m.maxInvDensity;
        if (t > tmax) break;
        r.nlk++;
        ret *= 1.0_r - lookupDensity(m, o + d * t) * inv;
    }
    return ret;
}
// With `sc` (adjoint): the estimate's log-derivative is -int sigma_t'.
template <int N>
DI Df<N> mediumTransmittanceAD(const DMedium &m, const Dv<N> &o, const Dv<N> &d, real tmin, const Df<N> &tmax, Rng &r,
                               const ScatterCtx *scx = nullptr) {
    if (m.type == MED_HOMOGENEOUS) return dexp((Df<N>(tmin) - tmax) * sigmaTAD<N>(m));
    Df<N> ret(mediumTransmittance(m, o.val(), d.val(), tmin, tmax.v, r));
    real is[N];
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
DI real phaseEval(const DPhase &ph, f3 wi, f3 wo) {
    if (ph.type == PHASE_ISO) return INV_FOURPI;
    real g = ph.g, tmp = 1.0_r + g * g + 2.0_r * g * dot(wi, wo);
    return INV_FOURPI * (1.0_r - g * g) / (tmp * sqrt(tmp));
}
template <int N> DI Df<N> phaseEvalAD(const DPhase &ph, const Dv<N> &wi, const f3 &wo) {
    if (ph.type == PHASE_ISO) return Df<N>(INV_FOURPI);
    Df<N> g(ph.g);
    for (int j = 0; j < N; ++j) g.d[j] = ph.g_d[j];
    Df<N> tmp = 1.0_r + g * g + 2.0_r * g * ddot(wi, wo);
    return INV_FOURPI * (1.0_r - g * g) / dpow15(tmp);
}
DI f3 phaseSample(const DPhase &ph, f3 wi, real u1, real u2) {
    if (ph.type == PHASE_ISO) return squareToUniformSphere(u1, u2);
    real g = ph.g, ct;
    if (fabs(g) < EPS) ct = 1.0_r - 2.0_r * u1;
    else {
        real sq = (1.0_r - g * g) / (1.0_r - g + 2.0_r * g * u1);
        ct = (1.0_r + g * g - sq * sq) / (2.0_r * g);
    }
    real st = sqrt(fmax(0._r, 1.0_r - ct * ct)), sp, cp;
    rsincospi(2.0_r * u2, &sp, &cp);
    return Frame(wi).toWorld(f3{st * cp, st * sp, ct});
}
template <int N> DI Df<N> phaseSampleWeightAD(const DPhase &ph, const Dv<N> &wi, f3 wo) {
    if (ph.type == PHASE_ISO) return Df<N>(1._r);
    Df<N> e = phaseEvalAD<N>(ph, wi, wo);
    return e.v > EPS ? e / e.v : Df<N>(0._r);
}

// -------------------------------------------------------------- BSDFs
// microfacet.h:10-248 (Beckmann, Smith G1 rational fit, visible-normal sampling)
DI real mfEval(real alpha, f3 m) {
    if (m.z < EPS) return 0._r;
    real c2 = m.z * m.z, be = (m.x * m.x + m.y * m.y) / (alpha * alpha * c2);
    real r = exp(-be) / (PI * alpha * alpha * c2 * c2);
    return r * m.z < 1e-20_r ? 0._r : r;
}
DI real tanThetaF(f3 v) { real tmp = 1.0_r - v.z * v.z; return tmp < EPS ? 0._r : sqrt(tmp) / v.z; }
DI real smithG1(real alpha, f3 v, f3 m) {
    if (dot(v, m) * v.z < EPS) return 0._r;
    real tt = fabs(tanThetaF(v));
    if (tt == 0._r) return 1._r;
    real a = 1.0_r / (alpha * tt);
    if (a >= 1.6_r) return 1._r;
    real a2 = a * a;
    return (3.535_r * a + 2.181_r * a2) / (1.0_r + 2.276_r * a + 2.577_r * a2);
}
DI real mfPdf(real alpha, f3 wi, f3 m) {
    if (fabs(wi.z) < EPS) return 0._r;
    return smithG1(alpha, wi, m) * fabs(dot(wi, m)) * mfEval(alpha, m) / fabs(wi.z);
}
DI f3 mfSample(real alpha, f3 wi_, real s0, real s1) {
    f3 wi = normalize(f3{alpha * wi_.x, alpha * wi_.y, wi_.z});
    // theta, and phi as its sine and cosine (wi.y / r, wi.x / r: sincos of atan2(wi.y, wi.x))
    real theta = 0._r, sP = 0._r, cP = 1._r;
    if (wi.z < 0.99999_r) {
        theta = acos(wi.z);
        const real rxy = sqrt(wi.x * wi.x + wi.y * wi.y);
        if (rxy > 0._r) { sP = wi.y / rxy; cP = wi.x / rxy; }
    }
    real sx, sy;
    const real SQRT_PI_INV = 0.56418958354775628695_r;
    if (theta < 1e-4_r) {
        real r = sqrt(-log(1.0_r - s0)), sp, cp;
        rsincospi(2.0_r * s1, &sp, &cp);
        sx = r * cp;
        sy = r * sp;
    } else {
        real tanI = tan(theta), cotI = 1.0_r / tanI;
        real a = -1.0_r, c = erf_as(cotI), smp = fmax(s0, 1e-6_r);
        real fit = 1.0_r + theta * (-0.876_r + theta * (0.4265_r - 0.0594_r * theta));
        real b = c - (1.0_r + c) * pow(1.0_r - smp, fit);
        real nrm = 1.0_r / (1.0_r + c + SQRT_PI_INV * tanI * exp(-cotI * cotI));
        for (int it = 1; it < 10; ++it) {
            if (!(b >= a && b <= c)) b = 0.5_r * (a + c);
            real ie = erfinv_giles(b);
            real value = nrm * (1.0_r + b + SQRT_PI_INV * tanI * exp(-ie * ie)) - smp;
            real deriv = nrm * (1.0_r - ie * tanI);
            if (fabs(value) < 1e-5_r) break;
            if (value > 0._r) c = b; else a = b;
            b -= value / deriv;
        }
        sx = erfinv_giles(b);
        sy = erfinv_giles(2.0_r * fmax(s1, 1e-6_r) - 1.0_r);
    }
    real rx = (cP * sx - sP * sy) * alpha, ry = (sP * sx + cP * sy) * alpha;
    real n = 1.0_r / sqrt(rx * rx + ry * ry + 1.0_r);
    return f3{-rx * n, -ry * n, n};
}
template <int N> DI Df<N> mfEvalAD(const Df<N> &alpha, const Dv<N> &m) {
    if (m.z.v < EPS) return Df<N>(0._r);
    Df<N> c2 = m.z * m.z;
    Df<N> r = dexp(-((m.x * m.x + m.y * m.y) / (alpha * alpha * c2))) / (PI * alpha * alpha * c2 * c2);
    return r.v * m.z.v < 1e-20_r ? Df<N>(0._r) : r;
}
template <int N> DI Df<N> smithG1AD(const Df<N> &alpha, const Dv<N> &v, const Dv<N> &m) {
    if (dot(v.val(), m.val()) * v.z.v < EPS) return Df<N>(0._r);
    Df<N> tmp = 1.0_r - v.z * v.z;
    if (tmp.v < EPS) return Df<N>(1._r);
    Df<N> tt = dabs(dsqrt(tmp) / v.z);
    if (tt.v == 0._r) return Df<N>(1._r);
    Df<N> a = 1.0_r / (alpha * tt);
    if (a.v >= 1.6_r) return Df<N>(1._r);
    Df<N> a2 = a * a;
    return (3.535_r * a + 2.181_r * a2) / (1.0_r + 2.276_r * a + 2.577_r * a2);
}
DI f3 phongReflect(f3 wi) { return f3{-wi.x, -wi.y, wi.z}; }
template <int N> DI Ds<N> specAD(const float *v, const float (*d)[3]) {
    Ds<N> r(ld3(v));
    for (int j = 0; j < N; ++j) for (int k = 0; k < 3; ++k) r.c[k].d[j] = d[j][k];
    return r;
}
template <int N> DI Df<N> scalarAD(real v, const float *d) { Df<N> r(v); for (int j = 0; j < N; ++j) r.d[j] = d[j]; return r; }

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
            real al = dot(wo, phongReflect(wi));
            if (al > 0._r) r = ld3(b.ks) * ((b.ex + 2.0_r) * pow(al, b.ex) * INV_TWOPI);
            return (r + ld3(b.kd) * INV_PI) * wo.z;
        }
        default: {
            if (fabs(wi.z) < EPS || fabs(wo.z) < EPS) return f3{0, 0, 0};
            bool refl = wi.z * wo.z > 0._r;
            real invEta = 1.0_r / b.eta;
            f3 H = refl ? normalize(wo + wi) : normalize(wi + wo * (wi.z > 0._r ? b.eta : invEta));
            H = H * signum(H.z);
            real Dv_ = mfEval(b.alpha, H);
            if (fabs(Dv_) < EPS) return f3{0, 0, 0};
            real F = fresnelDielectricExt(dot(wi, H), b.eta);
            real G = smithG1(b.alpha, wi, H) * smithG1(b.alpha, wo, H);
            if (refl) { real v = F * Dv_ * G / (4.0_r * fabs(wi.z)); return f3{v, v, v}; }
            real eta = wi.z > 0._r ? b.eta : invEta;
            real sd = dot(wi, H) + eta * dot(wo, H);
            real value = ((1.0_r - F) * Dv_ * G * eta * eta * dot(wi, H) * dot(wo, H)) / (wi.z * sd * sd);
            real factor = importance ? 1.0_r : (wi.z > 0._r ? invEta : b.eta);
            real v = fabs(value * factor * factor);
            return f3{v, v, v};
        }
    }
}
DI real bsdfPdf(const DBsdf &b, f3 wi, f3 wo) {
    switch (b.type) {
        case BSDF_DIFFUSE: return (wi.z < EPS || wo.z < EPS) ? 0._r : INV_PI * wo.z;
        case BSDF_NULL: return 1._r;
        case BSDF_PHONG: {
            if (wi.z < EPS || wo.z < EPS) return 0._r;
            real dp = INV_PI * wo.z, spp = 0._r, al = dot(wo, phongReflect(wi));
            if (al > 0._r) spp = pow(al, b.ex) * (b.ex + 1.0_r) / (2.0_r * PI);
            return b.specW * spp + (1.0_r - b.specW) * dp;
        }
        default: {
            bool refl = wi.z * wo.z > 0._r;
            real invEta = 1.0_r / b.eta, dwh;
            f3 H;
            if (refl) { H = normalize(wi + wo); dwh = 1.0_r / (4.0_r * dot(wo, H)); }
            else {
                real eta = wi.z > 0._r ? b.eta : invEta;
                H = normalize(wi + wo * eta);
                real sd = dot(wi, H) + eta * dot(wo, H);
                dwh = (eta * eta * dot(wo, H)) / (sd * sd);
            }
            H = H * signum(H.z);
            real prob = mfPdf(b.alpha, wi * signum(wi.z), H);
            real F = fresnelDielectricExt(dot(wi, H), b.eta);
            prob *= refl ? F : (1.0_r - F);
            return fabs(prob * dwh);
        }
    }
}
DI f3 bsdfSample(const DBsdf &b, f3 wi, real r0, real r1, real r2, f3 &wo, real &pdf, bool importance = false) {
    switch (b.type) {
        case BSDF_DIFFUSE:
            if (wi.z < EPS) return f3{0, 0, 0};
            wo = squareToCosineHemisphere(r0, r1);
            pdf = INV_PI * wo.z;
            return ld3(b.refl);
        case BSDF_NULL:
            wo = -wi;
            pdf = 1._r;
            return f3{1, 1, 1};
        case BSDF_PHONG: {
            if (r2 < b.specW) {
                real sa = sqrt(fmax(0._r, 1.0_r - pow(r1, 2.0_r / (b.ex + 1.0_r)))), ca = pow(r1, 1.0_r / (b.ex + 1.0_r)), sp, cp;
                rsincospi(2.0_r * r0, &sp, &cp);
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
            real invEta = 1.0_r / b.eta;
            f3 wis = wi * signum(wi.z);
            f3 m = mfSample(b.alpha, wis, r0, r1);
            real mpdf = mfPdf(b.alpha, wis, m);
            if (mpdf < EPS) return f3{0, 0, 0};
            pdf = mpdf;
            real ct;
            real F = fresnelDielectricExt(dot(wi, m), ct, b.eta);
            real weight = 1._r, dwh;
            bool srefl = !(r2 > F);
            pdf *= srefl ? F : (1.0_r - F);
            if (srefl) {
                wo = reflect(wi, m);
                if (wi.z * wo.z < EPS) return f3{0, 0, 0};
                dwh = 1.0_r / (4.0_r * dot(wo, m));
            } else {
                if (fabs(ct) < EPS) return f3{0, 0, 0};
                wo = refract(wi, m, b.eta, ct);
                real eta = ct < 0._r ? b.eta : invEta;
                if (wi.z * wo.z >= 0._r) return f3{0, 0, 0};
                real factor = importance ? 1.0_r : (ct < 0._r ? invEta : b.eta);
                weight *= factor * factor;
                real sd = dot(wi, m) + eta * dot(wo, m);
                dwh = (eta * eta * dot(wo, m)) / (sd * sd);
            }
            weight *= smithG1(b.alpha, wo, m);
            pdf *= fabs(dwh);
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
            if (al.v > 0._r) r = specAD<N>(b.ks, b.ks_d) * ((e + 2.0_r) * dpow(al, e) * INV_TWOPI);
            r += specAD<N>(b.kd, b.kd_d) * Df<N>(INV_PI);
            return r * wo.z;
        }
        default: {
            if (fabs(wi.z.v) < EPS || fabs(wo.z.v) < EPS) return Ds<N>();
            Df<N> alpha = scalarAD<N>(b.alpha, b.alpha_d), eta = scalarAD<N>(b.eta, b.eta_d), invEta = 1.0_r / eta;
            bool refl = wi.z.v * wo.z.v > 0._r;
            Dv<N> H = refl ? dnormalize(wo + wi) : dnormalize(wi + wo * (wi.z.v > 0._r ? eta : invEta));
            H = H * signum(H.z.v);
            Df<N> Dm = mfEvalAD<N>(alpha, H);
            if (fabs(Dm.v) < EPS) return Ds<N>();
            Df<N> F = fresnelDielectricExtAD<N>(ddot(wi, H), eta);
            Df<N> G = smithG1AD<N>(alpha, wi, H) * smithG1AD<N>(alpha, wo, H);
            if (refl) return Ds<N>::fill(F * Dm * G / (4.0_r * dabs(wi.z)));
            const Df<N> &e = wi.z.v > 0._r ? eta : invEta;
            Df<N> sd = ddot(wi, H) + e * ddot(wo, H);
            Df<N> value = ((1.0_r - F) * Dm * G * e * e * ddot(wi, H) * ddot(wo, H)) / (wi.z * sd * sd);
            Df<N> factor = wi.z.v > 0._r ? invEta : eta;
            return Ds<N>::fill(dabs(value * factor * factor));
        }
    }
}

// ------------------------------------------------------------- emitters
// area.cpp:5-15, area2.cpp:5-27
DI f3 emitterEval(const DEmitter &e, f3 n_geo, f3 n_sh, f3 d) {
    if (e.type == EMIT_AREA) return (e.two_sided || dot(n_geo, d) > 0._r) ? ld3(e.I) : f3{0, 0, 0};
    real tmp = dot(n_sh, d);
    return tmp > EPS ? ld3(e.I) * (e.kappa * INV_TWOPI * exp(e.kappa * (tmp - 1.0_r))) : f3{0, 0, 0};
}
DI f3 emitterEvalNorm(const DEmitter &e, f3 n, f3 d) { return emitterEval(e, n, n, d); }
template <int N> DI Ds<N> emitterEvalAD(const DEmitter &e, const HitAD<N> &h, const Dv<N> &d) {
    if (e.type == EMIT_AREA) return (e.two_sided || dot(h.gn.val(), d.val()) > 0._r) ? Ds<N>(ld3(e.I)) : Ds<N>();
    Df<N> tmp = ddot(h.sn, d);
    if (tmp.v > EPS) {
        Df<N> k = scalarAD<N>(e.kappa, e.kappa_d);
        return specAD<N>(e.I, e.I_d) * (dexp(k * (tmp - 1.0_r)) * k * INV_TWOPI);
    }
    return Ds<N>();
}

// area.cpp:17-19, area2.cpp:29-35 — evaluation against an explicit (AD) normal
template <int N> DI Ds<N> emitterEvalNormAD(const DEmitter &e, const Dv<N> &n, const Dv<N> &d) {
    if (e.type == EMIT_AREA) return (e.two_sided || dot(n.val(), d.val()) > 0._r) ? Ds<N>(ld3(e.I)) : Ds<N>();
    Df<N> tmp = ddot(n, d);
    if (tmp.v > EPS) {
        Df<N> k = scalarAD<N>(e.kappa, e.kappa_d);
        return specAD<N>(e.I, e.I_d) * (dexp(k * (tmp - 1.0_r)) * k * INV_TWOPI);
    }
    return Ds<N>();
}

// ---------------------------------------------------- distributions
// pmf.h:103-167 — lower_bound on an unnormalised-then-normalised CDF, sample reuse
DI int sampleReuse(const double *cdf, int n, real &u, real *pdf) {
    int lo = 0, hi = n + 1;
    double uu = u;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (cdf[mid] < uu) lo = mid + 1; else hi = mid; }
    int idx = min(n - 1, max(0, lo - 1));
    while (idx + 1 < n && cdf[idx + 1] - cdf[idx] == 0.0) ++idx;
    double a = cdf[idx], b = cdf[idx + 1];
    if (pdf) *pdf = (real)(b - a);
    u = (real)((uu - a) / (b - a));
    u = fmin(fmax(u, 0._r), 1._r);
    return idx;
}
DI int upperBound(const double *cdf, int n, real u) {
    int lo = 0, hi = n;
    double uu = u;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (cdf[mid] <= uu) lo = mid + 1; else hi = mid; }
    return min(max(lo, 0), n - 1);
}

}  // namespace DT_PNS
}  // namespace dt
