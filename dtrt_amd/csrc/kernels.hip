// gfx950 kernels of the forward-mode AD volumetric path tracer.
//
//  k_ad_paths<N>    one thread per AD path sample (pixel x spp). Runs the dual-number
//                   estimator LiAD (volpathAD.cpp:145-620) iteratively, accumulates the
//                   interior terms, and pushes every primal sub-path the boundary terms
//                   need (2 per accepted edge sample, 1 per Lins sample) into an HBM queue.
//  k_ad_refill<N>   the same for N >= 8 (one wave per SIMD), with per-bounce lane refill:
//                   a lane whose path has ended starts the wave's next sample.
//  k_children<N>    one thread per queued sub-path: primal Li / Lins estimate, scattered
//                   with its per-lane weights into the derivative accumulators.
//  k_primary<N>     primary-edge samples (integratorAD.cpp:100-165): 2 camera sub-paths each.
//  k_finalize<N>    divide by spp, drop negative values / non-finite tangents like
//                   integratorAD.cpp:59-87, add the primary-edge image, write the output.
#include "tracer.h"
#include "launch.h"

#include <algorithm>
#include <cfloat>
#include <cstdio>

// The file is compiled once per dual width (-DDT_N=1,2,3,4,8,16: that width's AD,
// child, primary-edge, adjoint and finalize kernels and their launchers) and once
// without DT_N (the width-independent kernels and the dispatchers over S.nder), so
// the seven code objects build in parallel. Each object holds its own copy of the
// frame rotation constant, named by DT_FRAME_ROT; setFrameRotation fills all seven.
namespace dt {
namespace DT_PNS {

using ChildHdr = ChildHdrT<real>;
// Edge-sample acceptance tests of volpathAD.cpp:545-584: the slerp weights must sum to
// one within Epsilon and the direction's tangents must be orthogonal to it within
// Epsilon. fp32 rounding needs wider bounds (1e-5, and 1e-4 relative to the tangent size).
constexpr real SLERP_SUM_TOL = kF64 ? 1e-7_r : 1e-5_r;
DI real orthTol(real gmax) { return kF64 ? EPS : 1e-4_r * (1.0_r + gmax); }
// Angle between the edge's two vertices as seen from x: acos of the normalised dot
// product as the reference (volpathAD.cpp:547); in fp32 the atan2 form, which stays
// accurate for small angles.
DI real edgeAngle(f3 e0, f3 e1, f3 c01, real l0, real l1) {
    if constexpr (kF64) return acos(dot(e0, e1) / (l0 * l1));
    else return atan2(length(c01) / (l0 * l1), dot(e0, e1) / (l0 * l1));
}
__constant__ FrameRot DT_FRAME_ROT;

constexpr int BLOCK = TRACE_BLOCK;
constexpr int MAX_BLOCK = 256;   // largest work-group size of the AD-path and child kernels
// Work-group size of the primal path and particle kernels: four waves, so that at their
// occupancy a small scene's tree fits in LDS beside the traversal stacks (treeLdsFor).
constexpr int PRIMAL_BLOCK = 256;
HDI unsigned blockOr64(int b) { return b == 128 || b == 256 ? (unsigned)b : 64u; }
// Minimum resident waves per SIMD requested from the register allocator (HIP's second
// __launch_bounds__ argument): bounds the VGPRs of the two path kernels. Per
// precision: a double takes two registers.
#ifndef CHILD_WAVES_PER_SIMD
#define CHILD_WAVES_PER_SIMD (kF64 ? CHILD_WAVES_F64 : 4)
#endif
#ifndef CHILD_WAVES_F64
#define CHILD_WAVES_F64 2   // heterogeneous-medium scene, 512^2, 128 spp, nder 3, child kernel per render: 2 waves 413 ms, 4 waves 464 ms (1 wave: 1.5x 2)
#endif
#ifndef AD_WAVES_PER_SIMD
#define AD_WAVES_PER_SIMD (kF64 ? AD_WAVES_F64 : 2)
#endif
#ifndef AD_WAVES_F64
#define AD_WAVES_F64 1      // 1 wave: 394 ms, 2 waves: 514 ms per render's AD kernel (same scene and size)
#endif
// Dual widths from which the VolPathTracerAD samples run in k_ad_refill (per-bounce
// lane refill) at one wave per SIMD, with the whole 512-entry register file, rather
// than spilling kilobytes per lane to scratch.
#ifndef AD_REFILL_MIN_N
#define AD_REFILL_MIN_N (kF64 ? AD_REFILL_MIN_N_F64 : 8)
#endif
#ifndef AD_REFILL_MIN_N_F64
#define AD_REFILL_MIN_N_F64 8
#endif
template <int N> constexpr bool adRefill() { return N >= AD_REFILL_MIN_N; }
template <int N> constexpr int adWaves() { return adRefill<N>() ? 1 : AD_WAVES_PER_SIMD; }
// Grid of k_ad_refill: waves per CU sharing a launch's samples (measured on the
// 100K-face dielectric scene at nder 8: in float 64 beats 32 and 128; in double the
// AD kernel takes 8.16, 7.76, 7.62, 7.48, 7.84, 8.78 s at 4, 8, 16, 32, 64, 128).
#ifndef AD_REFILL_WAVES_PER_CU
#define AD_REFILL_WAVES_PER_CU (kF64 ? 32 : 64)
#endif

DI uint32_t xcdRemap(uint32_t b, uint32_t nb) {
    // Blocks are dealt round-robin over the 8 XCDs; give each XCD a contiguous run
    // of logical tiles so neighbouring pixels share that XCD's L2.
    const uint32_t X = 8;
    if (nb % X) return b;
    return (b % X) * (nb / X) + b / X;
}

// View of global AD sample `sid` (views are ordered by their first sample, s0 = 0).
DI int viewOf(const DScene &S, uint64_t sid) {
    int lo = 0, hi = S.nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.views[mid].s0 <= sid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

DI bool writeRecord(const RenderParams &P, char *q, unsigned *qcount, const ChildHdr &h, const real *w, int nw) {
    unsigned idx = atomicAdd(qcount, 1u);
    if (idx >= P.qcap) return false;
    char *rec = q + (size_t)idx * P.rec_stride;
    *reinterpret_cast<ChildHdr *>(rec) = h;
    real *wd = reinterpret_cast<real *>(rec + sizeof(ChildHdr));
    for (int i = 0; i < nw; ++i) wd[i] = w[i];
    return true;
}

// Child pushes of one AD path sample. Children are numbered 1, 2, ... in push order
// and child k draws from RNG stream key + k. When a push finds the queue full the
// sample stops pushing, holds back its own contribution and is listed with the
// number of children already delivered; the replay pass re-runs the sample (same
// streams, hence the same path) and skips those, so every child and every sample
// contribution is counted exactly once.
struct Pusher {
    char *q;
    unsigned *qcount;
    int gpix, view;     // global pixel and view of the sample (written into every record)
    uint64_t key;
    uint32_t child, skip, failed_at;   // failed_at: children delivered before the first failed push, or NO_FAIL
    static constexpr uint32_t NO_FAIL = 0xFFFFFFFFu;
};

DI void push(const RenderParams &P, Pusher &u, ChildHdr &h, const real *w, int nw) {
    const uint32_t k = ++u.child;
    if (k <= u.skip || u.failed_at != Pusher::NO_FAIL) return;
    const uint64_t ck = u.key + k;
    h.key_lo = (uint32_t)ck;
    h.key_hi = (uint32_t)(ck >> 32);
    h.pixel = u.gpix;
    h.view = u.view;
    if (!writeRecord(P, u.q, u.qcount, h, w, nw)) u.failed_at = k - 1;
}

// Wave-reduce the per-thread work counters and add them to work[base..base+3]
// (box tests, triangle tests, density lookups, traced segments).
DI void flushWork(unsigned long long *work, int base, uint32_t nbox, uint32_t ntri, uint32_t nlk, uint32_t nseg) {
    uint32_t v[4] = {nbox, ntri, nlk, nseg};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k)
            if (v[k]) atomicAdd(work + base + k, (unsigned long long)v[k]);
}

template <int N> DI void cameraRayAD(const DCamera &c, real x, real y, Dv<N> &o, Dv<N> &d) {   // camera.cpp:75-83
    x /= c.width;
    y /= c.height;
    f3 n3{(x - 0.5_r) * 2.0_r, (y - 0.5_r) * (-2.0_r) / c.aspect, 1.0_r};
    f3 dc = normalize(f3{c.ndc2c[0] * n3.x + c.ndc2c[1] * n3.y + c.ndc2c[2] * n3.z,
                         c.ndc2c[3] * n3.x + c.ndc2c[4] * n3.y + c.ndc2c[5] * n3.z,
                         c.ndc2c[6] * n3.x + c.ndc2c[7] * n3.y + c.ndc2c[8] * n3.z});
    o = Dv<N>(ld3(c.cpos));
    d = Dv<N>(ld3(c.cs) * dc.x + ld3(c.ct) * dc.y + ld3(c.cn) * dc.z);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        o.setGrad(j, ld3(c.cpos_d[j]));
        d.setGrad(j, ld3(c.cs_d[j]) * dc.x + ld3(c.ct_d[j]) * dc.y + ld3(c.cn_d[j]) * dc.z);
    }
}

// camera.cpp:40-50 (primal, jittered pixel): origin and direction through the
// camera-to-world matrix, as VolPathTracer's camera rays are built.
DI void cameraRay(const DCamera &c, real x, real y, f3 &o, f3 &d) {
    x /= c.width;
    y /= c.height;
    f3 n3{(x - 0.5_r) * 2.0_r, (y - 0.5_r) * (-2.0_r) / c.aspect, 1.0_r};
    f3 dc = normalize(f3{c.ndc2c[0] * n3.x + c.ndc2c[1] * n3.y + c.ndc2c[2] * n3.z,
                         c.ndc2c[3] * n3.x + c.ndc2c[4] * n3.y + c.ndc2c[5] * n3.z,
                         c.ndc2c[6] * n3.x + c.ndc2c[7] * n3.y + c.ndc2c[8] * n3.z});
    o = f3{c.c2w[3], c.c2w[7], c.c2w[11]};
    d = f3{c.c2w[0] * dc.x + c.c2w[1] * dc.y + c.c2w[2] * dc.z, c.c2w[4] * dc.x + c.c2w[5] * dc.y + c.c2w[6] * dc.z,
           c.c2w[8] * dc.x + c.c2w[9] * dc.y + c.c2w[10] * dc.z};
}

// shape.cpp:302-335
DI int isSilhouette(const DScene &S, int4 e, f3 p, bool hasNormals) {
    if (e.z < 0 || e.w < 0) return 2;
    f3 v0 = ldg3(S.V + 3 * (size_t)e.x), v1 = ldg3(S.V + 3 * (size_t)e.y);
    bool ff[2];
    int fs[2] = {e.z, e.w};
    for (int s = 0; s < 2; ++s) {
        const int *fi = S.F + 3 * (size_t)fs[s];
        int k = fi[0];
        if (k == e.x || k == e.y) k = fi[1];
        if (k == e.x || k == e.y) k = fi[2];
        f3 v = ldg3(S.V + 3 * (size_t)k);
        f3 n = s == 0 ? cross(v0 - v, v1 - v) : cross(v1 - v, v0 - v);
        ff[s] = dot(n, p - v) > 0._r;   // sign test; normalisation does not change it
    }
    if (ff[0] != ff[1]) return 2;
    return hasNormals ? 0 : 1;
}

// One boundary (edge) estimate at a path vertex: volpathAD.cpp:207-305 (volume) and
// :513-616 (interface). Accepted samples push two primal sub-paths with +/- weights.
template <int N>
DI void edgeTerm(Tracer<N> &T, const RenderParams &P, Rng &r, bool volume, const Dv<N> &x, const Hit *its0,
                 const DBsdf *b, const DPhase *ph, f3 rd_in, f3 thrv, f3 betav, int med, int depth, Pusher &u,
                 unsigned *stats) {
    const DScene &S = T.S;
    if (!S.edges_valid) return;
    const int ne = P.sppse;
    for (int i = 0; i < ne; ++i) {
        real rnd = (i + r.next()) / ne, pdfs, pdfe;
        int sid = sampleReuse(S.shape_edge_cdf, S.nshapes, rnd, &pdfs);
        const DShape &es = S.shapes[sid];
        if (es.ne <= 0) continue;
        int k = sampleReuse(S.edge_cdf + es.ebase + sid, es.ne, rnd, &pdfe);
        real t = rnd, edgePdf = pdfs * pdfe;
        int4 e = S.edges[es.ebase + k];
        int sil = isSilhouette(S, e, x.val(), es.has_normals);
        if (!(sil > 0 && S.bsdfs[es.bsdf_id].type != BSDF_NULL)) continue;
        Dv<N> e0 = ldDv<N>(S.V + 3 * (size_t)e.x, S.Vd + (size_t)e.x * N * 3) - x;
        Dv<N> e1 = ldDv<N>(S.V + 3 * (size_t)e.y, S.Vd + (size_t)e.y * N * 3) - x;
        f3 c01 = cross(e0.val(), e1.val());
        f3 nrm = normalize(c01);
        real l0 = length(e0.val()), l1 = length(e1.val());
        const real theta = edgeAngle(e0.val(), e1.val(), c01, l0, l1);
        if (theta < EPS || theta > PI - EPS) continue;
        const real thp = theta * INV_PI, sinT = rsinpi(thp);   // sines through sinpi of theta / pi
        real w0 = rsinpi(thp * (1.0_r - t)) / sinT / l0, w1 = rsinpi(thp * t) / sinT / l1;
        real ws = w0 + w1;
        w0 /= ws;
        w1 /= ws;
        if (w0 < -EPS || w1 < -EPS || fabs(w0 + w1 - 1.0_r) > SLERP_SUM_TOL) { atomicAdd(stats + STAT_BAD_EDGE_WEIGHT, 1u); continue; }
        Dv<N> w = e0 * w0 + e1 * w1;
        Df<N> dist = dnorm(w);
        w = w / dist;
        edgePdf /= theta;
        f3 wv = w.val();
        real orth = 0._r, gmax = 0._r;
#pragma unroll
        for (int j = 0; j < N; ++j) { orth = fmax(orth, fabs(dot(w.grad(j), wv))); gmax = fmax(gmax, length(w.grad(j))); }
        if (orth > orthTol(gmax)) { atomicAdd(stats + STAT_NONORTHO, 1u); continue; }
        if (w.gradZero(EPS)) continue;
        f3 fv;
        int med1 = med;
        if (volume) {
            fv = thrv * phaseEval(*ph, -rd_in, wv);
        } else {
            Frame shf(its0->sn);
            fv = thrv * bsdfEval(*b, its0->wi, shf.toLocal(wv));
            const DShape &sh0 = shapeOf(S, its0->face);
            if (isTransition(sh0)) med1 = targetMedium(sh0, its0->gn, wv);
        }
        if (allZero(fv, EPS)) continue;
        real trans = T.evalTransmittance(x.val(), wv, !volume, med1, dist.v, r, depth - 1);
        if (!(trans > EPS)) continue;
        real W[N * 3];
        f3 base = betav * fv * (1.0_r / (edgePdf * ne));
#pragma unroll
        for (int j = 0; j < N; ++j) {
            real nd = dot(nrm, w.grad(j));
            W[3 * j] = base.x * nd; W[3 * j + 1] = base.y * nd; W[3 * j + 2] = base.z * nd;
        }
        ChildHdr h;
        f3 org = x.val();
        h.org[0] = org.x; h.org[1] = org.y; h.org[2] = org.z;
        h.meta = CHILD_LI | ((depth - 1) << 4) | ((med1 + 1) << 12);
        h.face = -1;
        for (int s = 0; s < 2; ++s) {
            f3 w2 = normalize(s == 0 ? wv - nrm * ANGLE_EPS : wv + nrm * ANGLE_EPS);
            h.dir[0] = w2.x; h.dir[1] = w2.y; h.dir[2] = w2.z;
            if (s == 1)
#pragma unroll
                for (int k2 = 0; k2 < N * 3; ++k2) W[k2] = -W[k2];
            push(P, u, h, W, N * 3);
        }
    }
}

// Surface-only forward-mode integrators: PathTracerAD (pathAD.cpp:64-214, DIRECT false)
// and DirectIntegratorAD (directAD.cpp:5-98, DIRECT true), for one AD sample of pixel
// `pixel`. Media are ignored. pathAD's edge terms push their two primal sub-paths as
// CHILD_PATH records (PathTracerAD::Li = path.cpp:11-61); directAD's need only the
// emission seen along two rays and are evaluated here.
template <int N, bool DIRECT>
DI void adSurfaceSample(Tracer<N> &T, const RenderParams &P, const DView &vw, Rng &r, int pixel, Ds<N> &L, Pusher &u,
                        unsigned *stats) {
    const DScene &S = T.S;
    const DCamera &cam = vw.cam;
    const int ix = cam.cropped ? cam.ox + pixel % cam.cw : pixel % cam.width;
    const int iy = cam.cropped ? cam.oy + pixel / cam.cw : pixel / cam.width;
    real jx = 0.5_r, jy = 0.5_r;
    if (vw.sppe >= 0) { jx = r.next(); jy = r.next(); }
    Dv<N> ro, rd;
    cameraRayAD<N>(cam, ix + jx, iy + jy, ro, rd);
    HitAD<N> its;
    // pathAD.cpp:222-227 traces the camera ray from 0, directAD.cpp:11 from ShadowEpsilon
    if (!T.rayIntersectAD(ro, rd, DIRECT, its)) return;
    const int mode = DIRECT ? 0x3F : P.mode;
    int depth = DIRECT ? 1 : P.max_bounces;
    bool incEmission = true;
    Ds<N> beta(f3{1, 1, 1});
    while (true) {
        const DShape &sh = shapeOf(S, its.face);
        if (sh.light_id >= 0 && (mode & 1) && incEmission)
            L += beta * emitterEvalAD<N>(S.emitters[sh.light_id], its, -rd);
        if (depth == 0) break;
        const DBsdf &b = bsdfOf(S, sh);
        const bool isnull = b.type == BSDF_NULL;
        const Hit its0 = its.val();
        FrameAD<N> shf(its.sn);
        if (mode & 2) {   // light sampling
            real r4[4] = {r.next(), r.next(), r.next(), r.next()};
            Dv<N> woL;
            real pdf_nee;
            Ds<N> value = T.neeDirectAD(its, shf, r4, woL, pdf_nee);
            if (!value.isZero(EPS)) {
                Ds<N> bv = bsdfEvalAD<N>(b, its.wi, woL);
                const real bp = bsdfPdf(b, its0.wi, woL.val());
                const real mis = DIRECT ? pdf_nee / (pdf_nee + bp) : misPower(pdf_nee, bp);
                L += beta * (value * bv * mis);
            }
        }
        // BSDF sampling (bsdf.cpp:11-21; null.cpp:20-25)
        real r3[3] = {r.next(), r.next(), r.next()};
        bool cont = false;
        Dv<N> nd;
        Ds<N> bw;
        real bpdf = 1._r;
        if (isnull) {
            nd = rd;
            bw = Ds<N>(f3{1, 1, 1});
            cont = true;
        } else {
            f3 wo0;
            f3 w0 = bsdfSample(b, its0.wi, r3[0], r3[1], r3[2], wo0, bpdf);
            if (!allZero(w0, EPS) && bpdf > EPS) {
                const f3 ww = shf.val().toWorld(wo0);
                bw = bsdfEvalAD<N>(b, its.wi, shf.toLocal(Dv<N>(ww))) * (1.0_r / bpdf);
                if (!bw.isZero(EPS)) { nd = Dv<N>(ww); cont = true; }
            }
        }
        HitAD<N> its1;
        if (cont) {
            cont = T.rayIntersectAD(its.p, nd, true, its1);
            if (cont) {
                const DShape &sh1 = shapeOf(S, its1.face);
                if (sh1.light_id >= 0 && (mode & 2)) {
                    Ds<N> lc = emitterEvalAD<N>(S.emitters[sh1.light_id], its1, -nd);
                    if (!allZero(lc.val(), 0._r)) {
                        const f3 dd = its1.p.val() - its0.p;
                        const real G = its1.wi.z.v / dot(dd, dd);
                        const real pn = S.light_pmf[sh1.light_id] / (S.light_areas[sh1.light_id] * G);
                        const real mis = DIRECT ? bpdf / (pn + bpdf) : misPower(bpdf, pn);
                        L += beta * (bw * lc * mis);
                    }
                }
            }
        }
        // edge sampling at this vertex (pathAD.cpp:112-210; directAD.cpp:55-95)
        if (P.sppse > 0 && (mode & 4) && S.edges_valid) {
            const int ne = P.sppse;
            const Dv<N> &x1 = its.p;
            real der[N * 3];
#pragma unroll
            for (int k = 0; k < N * 3; ++k) der[k] = 0._r;
            for (int i = 0; i < ne; ++i) {
                real rnd = r.next(), pdfs, pdfe;
                const int sid = sampleReuse(S.shape_edge_cdf, S.nshapes, rnd, &pdfs);
                const DShape &es = S.shapes[sid];
                if (es.ne <= 0) continue;
                const int k = sampleReuse(S.edge_cdf + es.ebase + sid, es.ne, rnd, &pdfe);
                const real t = rnd;
                real edgePdf = pdfs * pdfe;
                const int4 e = S.edges[es.ebase + k];
                if (isSilhouette(S, e, x1.val(), es.has_normals) <= 0) continue;
                const Dv<N> v0 = ldDv<N>(S.V + 3 * (size_t)e.x, S.Vd + (size_t)e.x * N * 3);
                const Dv<N> v1 = ldDv<N>(S.V + 3 * (size_t)e.y, S.Vd + (size_t)e.y * N * 3);
                Dv<N> w;
                f3 nrm;
                real dist, geo = 1._r;
                if constexpr (DIRECT) {   // linear in t, pdf per unit length (scene.h:103-108)
                    edgePdf /= length(v1.val() - v0.val());
                    Dv<N> y1 = v0 * (1.0_r - t) + v1 * t;
                    w = y1 - x1;
                    nrm = normalize(cross(v0.val() - x1.val(), v1.val() - x1.val()));
                    Df<N> dAD = dnorm(w);
                    w = w / dAD;
                    dist = dAD.v;
                    const real ct = -dot(w.val(), normalize(v1.val() - v0.val()));
                    geo = sqrt(fmax(0._r, 1.0_r - ct * ct)) / dist;
                } else {   // angle-uniform point on the edge (pathAD.cpp:136-164)
                    const Dv<N> e0 = v0 - x1, e1 = v1 - x1;
                    const f3 c01 = cross(e0.val(), e1.val());
                    nrm = normalize(c01);
                    const real l0 = length(e0.val()), l1 = length(e1.val());
                    const real theta = edgeAngle(e0.val(), e1.val(), c01, l0, l1);
                    if (theta < EPS || theta > PI - EPS) continue;
                    const real thp = theta * INV_PI, sinT = rsinpi(thp);
                    real w0 = rsinpi(thp * (1.0_r - t)) / sinT / l0, w1 = rsinpi(thp * t) / sinT / l1;
                    const real ws = w0 + w1;
                    w0 /= ws;
                    w1 /= ws;
                    if (w0 < -EPS || w1 < -EPS || fabs(w0 + w1 - 1.0_r) > SLERP_SUM_TOL) { atomicAdd(stats + STAT_BAD_EDGE_WEIGHT, 1u); continue; }
                    w = e0 * w0 + e1 * w1;
                    Df<N> dAD = dnorm(w);
                    w = w / dAD;
                    dist = dAD.v;
                    edgePdf /= theta;
                    real orth = 0._r, gmax = 0._r;
#pragma unroll
                    for (int j = 0; j < N; ++j) { orth = fmax(orth, fabs(dot(w.grad(j), w.val()))); gmax = fmax(gmax, length(w.grad(j))); }
                    if (orth > orthTol(gmax)) { atomicAdd(stats + STAT_NONORTHO, 1u); continue; }
                    if (w.gradZero(EPS)) continue;
                }
                const f3 wv = w.val();
                const f3 fv = bsdfEval(b, its0.wi, shf.val().toLocal(wv));
                if (allZero(fv, EPS)) continue;
                real th;   // scene.cpp:150-176: nothing between x1 and the edge point
                if (T.trace(x1.val(), wv, SHADOW_EPS, (1.0_r - SHADOW_EPS) * dist, th) >= 0) continue;
                const f3 base = beta.val() * fv * (geo / (edgePdf * ne));
                if constexpr (DIRECT) {
                    f3 delta{0, 0, 0};
                    for (int sgn = 0; sgn < 2; ++sgn) {
                        const f3 w2 = normalize(sgn == 0 ? wv - nrm * ANGLE_EPS : wv + nrm * ANGLE_EPS);
                        Hit h1;
                        if (T.rayIntersect(x1.val(), w2, true, h1)) {
                            const DShape &hs = shapeOf(S, h1.face);
                            if (hs.light_id >= 0) {
                                const f3 le = emitterEval(S.emitters[hs.light_id], h1.gn, h1.sn, -wv);
                                delta = sgn == 0 ? delta + le : delta - le;
                            }
                        }
                    }
                    if (allZero(delta, EPS)) continue;
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        const real nd_ = dot(nrm, w.grad(j));
                        der[3 * j] += nd_ * base.x * delta.x;
                        der[3 * j + 1] += nd_ * base.y * delta.y;
                        der[3 * j + 2] += nd_ * base.z * delta.z;
                    }
                } else {
                    real W[N * 3];
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        const real nd_ = dot(nrm, w.grad(j));
                        W[3 * j] = base.x * nd_; W[3 * j + 1] = base.y * nd_; W[3 * j + 2] = base.z * nd_;
                    }
                    ChildHdr h;
                    const f3 org = x1.val();
                    h.org[0] = org.x; h.org[1] = org.y; h.org[2] = org.z;
                    h.meta = CHILD_PATH | ((depth - 1) << 4);
                    h.face = -1;
                    for (int sgn = 0; sgn < 2; ++sgn) {
                        const f3 w2 = normalize(sgn == 0 ? wv - nrm * ANGLE_EPS : wv + nrm * ANGLE_EPS);
                        h.dir[0] = w2.x; h.dir[1] = w2.y; h.dir[2] = w2.z;
                        if (sgn == 1)
#pragma unroll
                            for (int k2 = 0; k2 < N * 3; ++k2) W[k2] = -W[k2];
                        push(P, u, h, W, N * 3);
                    }
                }
            }
            if constexpr (DIRECT) {
#pragma unroll
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c) L.c[c].d[j] += der[3 * j + c];
            }
        }
        if (DIRECT || !cont) break;
        beta = beta * bw;
        rd = nd;
        its = its1;
        depth -= 1;
        incEmission = false;
    }
}

// Passes of the path-replay adjoint (k_adjoint); NONE in the forward kernel.
enum class AdjPass { NONE, TOTAL, SCATTER };
struct AdjCtx {
    float *grad;        // d loss / d node density of the grid of *pm (accumulated)
    const DMedium *pm;
    f3 adj;              // d loss / d pixel value of this sample (already / spp)
    f3 Ltot;             // the sample's radiance, from the TOTAL pass
};

// One AD path sample of pixel `pixel` (camera ray through the jittered pixel
// position, then the dual-number estimator of volpathAD.cpp:145-620), accumulated
// into L. A == SCATTER replays the path of a TOTAL pass with the same stream and adds
// the reverse-mode density derivative of every contribution into ac->grad.
// Camera ray (with tangents) through the jittered position of pixel `pixel`.
template <int N> DI void adCameraRay(const DView &vw, Rng &r, int pixel, Dv<N> &ro, Dv<N> &rd) {
    const DCamera &cam = vw.cam;
    int ix = cam.cropped ? cam.ox + pixel % cam.cw : pixel % cam.width;
    int iy = cam.cropped ? cam.oy + pixel / cam.cw : pixel / cam.width;
    real jx = 0.5_r, jy = 0.5_r;
    if (vw.sppe >= 0) { jx = r.next(); jy = r.next(); }
    cameraRayAD<N>(cam, ix + jx, iy + jy, ro, rd);
}

template <int N, AdjPass A>
DI void adPathSample(Tracer<N> &T, const RenderParams &P, const DView &vw, Rng &r, int pixel, Ds<N> &L, Pusher &u,
                     unsigned *stats, const AdjCtx *ac) {
    const DScene &S = T.S;
    const DCamera &cam = vw.cam;
    Dv<N> ro, rd;
    adCameraRay<N>(vw, r, pixel, ro, rd);
    real tcam;
    int hface = T.trace(ro.val(), rd.val(), 0._r, FINF, tcam);
    if (hface >= 0) {
        int med = cam.med_id, depth = P.max_bounces, mode = P.mode;
        Ds<N> beta(f3{1, 1, 1});
        while (mode != 0) {
#define AD_BOUNCE_END break
#include "ad_bounce.inc"
#undef AD_BOUNCE_END
        }
    }
}

// Normal pass: samples s0 .. s0+nsamp-1. Replay pass (replay != nullptr): entries
// {sid lo, sid hi, children delivered, 0} of samples whose pushes overflowed.
// Samples that overflow (again) are appended to ovf.
template <int N, int I>   // I: AdKind
__global__ __launch_bounds__(MAX_BLOCK, adWaves<N>()) void k_ad_paths(DScene S, RenderParams P, uint64_t s0, uint64_t nsamp,
                                                    const uint4 *replay, real *acc, char *q, unsigned *qcount,
                                                    unsigned *stats, uint4 *ovf, unsigned *ovf_count,
                                                    unsigned long long *work) {
    const uint32_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const uint64_t li = (uint64_t)blk * blockDim.x + threadIdx.x;
    Tracer<N> T(S);
    T.useLdsTree(P);
    Ds<N> L;
    const bool active = li < nsamp;
    uint64_t sid = s0 + li;
    uint32_t skip = 0;
    if (replay) {
        uint4 e = active ? replay[li] : make_uint4(0, 0, 0, 0);
        sid = (uint64_t)e.x | ((uint64_t)e.y << 32);
        skip = e.z;
    }
    const int v = viewOf(S, sid);
    const DView &vw = S.views[v];
    const uint64_t ls = sid - vw.s0;   // sample index within the view: keys its random stream
    const int pixel = (int)(ls / (uint64_t)vw.spp), gpix = vw.pix0 + pixel;
    Pusher u{q, qcount, gpix, v, (ls + 1) << 20, 0, skip, Pusher::NO_FAIL};
    Rng r;
    r.init(vw.seed, ls);
    if (active) {
        if constexpr (I == AD_VOLPATH) adPathSample<N, AdjPass::NONE>(T, P, vw, r, pixel, L, u, stats, nullptr);
        else adSurfaceSample<N, I == AD_DIRECT>(T, P, vw, r, pixel, L, u, stats);
    }
    const bool done = u.failed_at == Pusher::NO_FAIL;
    if (active && !done) ovf[atomicAdd(ovf_count, 1u)] = make_uint4((uint32_t)sid, (uint32_t)(sid >> 32), u.failed_at, 0u);
    // accumulate: one atomic per wave when the wave covers one pixel, else per lane
    const int nd = P.npix;
    const bool put = active && done;
    const int gp0 = __shfl(gpix, 0);
    if (!replay && __all(!active || gpix == gp0)) {   // the wave covers one pixel
        real sum[3 * (N + 1)];
        for (int c = 0; c < 3; ++c) {
            sum[c] = put ? L.c[c].v : 0._r;
#pragma unroll
            for (int j = 0; j < N; ++j) sum[3 + 3 * j + c] = put ? L.c[c].d[j] : 0._r;
        }
#pragma unroll
        for (int k = 0; k < 3 * (N + 1); ++k)
            for (int off = 32; off > 0; off >>= 1) sum[k] += __shfl_xor(sum[k], off);
        if ((threadIdx.x & 63) == 0 && active)
            for (int k = 0; k < 3 * (N + 1); ++k)
                if (sum[k] != 0._r) unsafeAtomicAdd(acc + ((size_t)(k / 3) * nd + gpix) * 3 + (k % 3), sum[k]);
    } else if (put) {
        for (int c = 0; c < 3; ++c) {
            if (L.c[c].v != 0._r) unsafeAtomicAdd(acc + (size_t)gpix * 3 + c, L.c[c].v);
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (L.c[c].d[j] != 0._r) unsafeAtomicAdd(acc + ((size_t)(j + 1) * nd + gpix) * 3 + c, L.c[c].d[j]);
        }
    }
    flushWork(work, 0, T.nbox, T.ntri, r.nlk, T.nseg);
}

// State of one AD path between bounces (the locals of adPathSample's loop).
template <int N> struct AdState {
    Dv<N> ro, rd;
    int hface;   // face the ray (ro, rd) hits; the bounce recomputes the hit record from it
    Ds<N> beta;
    int med, depth, mode;
    // The walk along (ro, rd) still to run at the start of the next step: 0 none, 1 first
    // hit only from a surface, 2 emitter walk, 3 the camera ray's first hit (from 0);
    // walk_add: its emitter term counts, with MIS partner pdf walk_cpdf.
    int walk;
    bool walk_surf, walk_add;
    real walk_cpdf;
};

// adPathSample up to its bounce loop: camera ray and first hit. False when the path
// has nothing to do.
template <int N>
DI bool adBegin(const RenderParams &P, const DView &vw, Rng &r, int pixel, AdState<N> &st) {
    adCameraRay<N>(vw, r, pixel, st.ro, st.rd);
    st.walk = 3;   // the first hit is found by the first step's walk, with the other lanes' walks
    st.walk_surf = false;
    st.walk_add = false;
    st.walk_cpdf = 0._r;
    st.hface = -1;
    st.med = vw.cam.med_id;
    st.depth = P.max_bounces;
    st.mode = P.mode;
    st.beta = Ds<N>(f3{1, 1, 1});
    return st.mode != 0;
}

// One iteration of adPathSample's bounce loop (forward pass). False when the path ends.
template <int N>
DI bool adBounce(Tracer<N> &T, const RenderParams &P, Rng &r, int pixel, AdState<N> &st, Ds<N> &L, Pusher &u,
                 unsigned *stats) {
    constexpr AdjPass A = AdjPass::NONE;
    const AdjCtx *ac = nullptr;
    const DScene &S = T.S;
    Dv<N> &ro = st.ro, &rd = st.rd;
    int &hface = st.hface;
    Ds<N> &beta = st.beta;
    int &med = st.med, &depth = st.depth, &mode = st.mode;
    if (st.walk) {   // the walk the last step handed over (or the camera ray's first hit)
        real pdf_nee = 0._r;
        int nface = -1;
        Ds<N> att = T.lookForEmitterAD(ro, rd, st.walk_surf, r, med, depth, nface, pdf_nee, nullptr, nullptr, nullptr,
                                       st.walk != 2);
        hface = nface;
        if (st.walk_add && !att.isZero(EPS)) L += beta * (att * misPower(st.walk_cpdf, pdf_nee));
        const bool camera = st.walk == 3;
        st.walk = 0;
        if (camera && hface < 0) return false;
    }
#define AD_BOUNCE_END return false
#define AD_WALK_DEFERRED(kind, surf, add, cp) \
    { st.walk = (kind); st.walk_surf = (surf); st.walk_add = (add); st.walk_cpdf = (cp); }
#include "ad_bounce.inc"
#undef AD_WALK_DEFERRED
#undef AD_BOUNCE_END
    return mode != 0;
}

// VolPathTracerAD samples with per-bounce lane refill (same estimator and random
// streams as k_ad_paths<N, AD_VOLPATH>), for the dual widths that run one wave per
// SIMD: there the hardware has few other waves to switch to, and the lanes of a wave
// that wait for its longest path dominate. Each wave owns the contiguous range of
// `per_wave` sample indices of its block; a lane whose path has ended takes the
// wave's next sample before the next bounce, until the range is used up. Finished
// samples are summed over the wave per pixel (one atomic per pixel and channel), or
// listed in ovf when a push overflowed the queue. Indices are into replay[] in a
// replay pass.
template <int N>
__global__ __launch_bounds__(BLOCK, adWaves<N>()) void k_ad_refill(DScene S, RenderParams P, uint64_t s0, uint64_t nsamp,
                                                     uint64_t per_wave, const uint4 *replay, real *acc, char *q,
                                                     unsigned *qcount, unsigned *stats, uint4 *ovf, unsigned *ovf_count,
                                                     unsigned long long *work) {
    const uint32_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const uint64_t begin = (uint64_t)blk * per_wave;
    const uint64_t end = begin < nsamp ? min(nsamp, begin + per_wave) : begin;
    Tracer<N> T(S);
    T.useLdsTree(P);
    AdState<N> st;
    Ds<N> L;
    Rng r;
    Pusher u{q, qcount, 0, 0, 0, 0, 0, Pusher::NO_FAIL};
    uint64_t sid = 0, next = begin;
    int pixel = 0;
    uint32_t nlk = 0;
    bool alive = false, have = false, exhausted = false;
    const int nd = P.npix;
    while (true) {
        const bool fin = have && !alive;   // the lane's path has ended: add its sample
        const bool put = fin && u.failed_at == Pusher::NO_FAIL;
        if (fin && !put) ovf[atomicAdd(ovf_count, 1u)] = make_uint4((uint32_t)sid, (uint32_t)(sid >> 32), u.failed_at, 0u);
        uint64_t pm = __ballot(put);
        while (pm) {   // wave-uniform: one pixel per round
            const int lead = __ffsll((unsigned long long)pm) - 1;
            const int gp0 = __shfl(u.gpix, lead);
            const bool mine = put && u.gpix == gp0;
            pm &= ~__ballot(mine);
#pragma unroll
            for (int k = 0; k < 3 * (N + 1); ++k) {
                const int c = k % 3, j = k / 3;
                real x = mine ? (j == 0 ? L.c[c].v : L.c[c].d[j - 1]) : 0._r;
                for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
                if ((int)(threadIdx.x & 63) == lead && x != 0._r) unsafeAtomicAdd(acc + ((size_t)j * nd + gp0) * 3 + c, x);
            }
        }
        if (fin) {
            nlk += r.nlk;
            have = false;
        }
        const bool want = !alive && !exhausted;
        const uint64_t m = __ballot(want);
        if (m == 0 && !__any(alive)) break;   // the wave's range is used up
        if (want) {
            const uint64_t li = next + __popcll(m & __lanemask_lt());
            if (li < end) {
                uint32_t skip = 0;
                sid = s0 + li;
                if (replay) {
                    const uint4 e = replay[li];
                    sid = (uint64_t)e.x | ((uint64_t)e.y << 32);
                    skip = e.z;
                }
                const int v = viewOf(S, sid);
                const DView &vw = S.views[v];
                const uint64_t ls = sid - vw.s0;   // sample index within the view: keys its random stream
                pixel = (int)(ls / (uint64_t)vw.spp);
                u = Pusher{q, qcount, vw.pix0 + pixel, v, (ls + 1) << 20, 0, skip, Pusher::NO_FAIL};
                r.init(vw.seed, ls);
                L = Ds<N>();
                have = true;
                alive = adBegin<N>(P, vw, r, pixel, st);
            } else {
                exhausted = true;
            }
        }
        next += (uint64_t)__popcll(m);
        if (alive) alive = adBounce<N>(T, P, r, pixel, st, L, u, stats);
    }
    flushWork(work, 0, T.nbox, T.ntri, nlk, T.nseg);
}

// The primal integrators: one thread per sample s0 .. s0+n-1 of view 0 (sample k of
// pixel p is p*spp + k), a jittered camera ray and the estimator of KIND:
//   PRIMAL_VOLPATH  VolPathTracer (volpath.cpp:106-141): Tracer::liStep
//                   (volpath.cpp:10-103 is volpathAD.cpp:50-142 with the first ray
//                   leaving the camera);
//   PRIMAL_SIMPLE   VolPathTracerSimple (volpath_simple.cpp:89-124): liSimpleStep;
//   PRIMAL_PATH     PathTracer (path.cpp:64-96) and DirectIntegrator (direct.cpp:
//                   51-85, max depth 1): pathStep, camera ray from tnear ShadowEpsilon.
// Radiance sums go to acc[pixel][3].
template <int KIND>
__global__ __launch_bounds__(PRIMAL_BLOCK, CHILD_WAVES_PER_SIMD) void k_primal_paths(DScene S, RenderParams P, uint64_t s0,
                                                                                     uint64_t nsamp, real *acc,
                                                                                     unsigned long long *work) {
    const uint32_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const uint64_t li = (uint64_t)blk * PRIMAL_BLOCK + threadIdx.x;
    const bool active = li < nsamp;
    const DView &vw = S.views[0];
    const DCamera &cam = vw.cam;
    const uint64_t ls = s0 + li;
    const int pixel = (int)(ls / (uint64_t)vw.spp);
    Tracer<1> T(S);
    T.useLdsTree(P);
    Rng r;
    r.init(vw.seed, ls);
    f3 L{0, 0, 0};
    if (active) {
        const int ix = cam.cropped ? cam.ox + pixel % cam.cw : pixel % cam.width;
        const int iy = cam.cropped ? cam.oy + pixel / cam.cw : pixel / cam.width;
        const real jx = r.next(), jy = r.next();
        f3 ro, rd;
        cameraRay(cam, ix + jx, iy + jy, ro, rd);
        Hit h;
        T.rayIntersect(ro, rd, KIND == PRIMAL_PATH, h);
        LiPath p;
        T.liStartAt(p, ro, rd, P.max_bounces, KIND == PRIMAL_PATH ? -1 : cam.med_id, true, h);
        if constexpr (KIND == PRIMAL_SIMPLE) {
            while (T.liSimpleStep(p, r)) {}
        } else if constexpr (KIND == PRIMAL_PATH) {
            while (T.pathStep(p, r)) {}
        } else {
            while (T.liStep(p, r)) {}
        }
        L = p.ret;
    }
    const int gp0 = __shfl(pixel, 0);
    if (__all(!active || pixel == gp0)) {   // the wave covers one pixel
        real v[3] = {L.x, L.y, L.z};
        for (int c = 0; c < 3; ++c)
            for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_xor(v[c], off);
        if ((threadIdx.x & 63) == 0 && active)
            for (int c = 0; c < 3; ++c)
                if (v[c] != 0._r) unsafeAtomicAdd(acc + (size_t)pixel * 3 + c, v[c]);
    } else if (active) {
        if (L.x != 0._r) unsafeAtomicAdd(acc + (size_t)pixel * 3, L.x);
        if (L.y != 0._r) unsafeAtomicAdd(acc + (size_t)pixel * 3 + 1, L.y);
        if (L.z != 0._r) unsafeAtomicAdd(acc + (size_t)pixel * 3 + 2, L.z);
    }
    flushWork(work, 0, T.nbox, T.ntri, r.nlk, T.nseg);
}

// ------------------------------------------------------------ ParticleTracer
// camera.cpp:129-165 — importance of world point p for the pinhole camera: pixel
// position (px, py) within the output window, direction to the camera; 0 when p is
// behind the near plane or outside the window.
DI real cameraSampleDirect(const DCamera &c, f3 p, real &px, real &py, f3 &dir) {
    const f3 q{c.w2c[0] * p.x + c.w2c[1] * p.y + c.w2c[2] * p.z + c.w2c[3],
               c.w2c[4] * p.x + c.w2c[5] * p.y + c.w2c[6] * p.z + c.w2c[7],
               c.w2c[8] * p.x + c.w2c[9] * p.y + c.w2c[10] * p.z + c.w2c[11]};
    if (q.z < c.clip_near) return 0._r;
    const real f = c.c2ndc[0];
    real inv_area = 0.25_r * f * f * c.aspect;
    int xmin = 0, xmax = c.width, ymin = 0, ymax = c.height;
    if (c.cropped) {
        inv_area *= (real)c.width / (real)c.cw * (real)c.height / (real)c.ch;
        xmin = c.ox; xmax = c.ox + c.cw;
        ymin = c.oy; ymax = c.oy + c.ch;
    }
    const f3 n{c.c2ndc[0] * q.x + c.c2ndc[1] * q.y + c.c2ndc[2] * q.z, c.c2ndc[3] * q.x + c.c2ndc[4] * q.y + c.c2ndc[5] * q.z,
               c.c2ndc[6] * q.x + c.c2ndc[7] * q.y + c.c2ndc[8] * q.z};
    const real sx = (n.x / n.z * 0.5_r + 0.5_r) * c.width, sy = (-(n.y / n.z) * 0.5_r * c.aspect + 0.5_r) * c.height;
    if (!(sx >= xmin && sx <= xmax && sy >= ymin && sy <= ymax)) return 0._r;
    px = sx - xmin;
    py = sy - ymin;
    const real dist = length(q), inv = 1.0_r / dist, invCos = dist / q.z;
    dir = (ld3(c.cpos) - p) * inv;
    return inv * inv * invCos * invCos * invCos * (inv_area * (real)((xmax - xmin) * (ymax - ymin)));
}

// Adds v to pixel (px, py) of a w x h window. A point exactly on the window's closing
// edge (px == w or py == h, which camera.cpp:151-152 accepts) has no pixel and is dropped.
DI void splat(real *acc, int w, int h, real px, real py, f3 v) {
    const int ix = (int)px, iy = (int)py;
    if (ix < 0 || iy < 0 || ix >= w || iy >= h) return;
    real *a = acc + ((size_t)iy * w + ix) * 3;
    if (v.x != 0._r) unsafeAtomicAdd(a, v.x);
    if (v.y != 0._r) unsafeAtomicAdd(a + 1, v.y);
    if (v.z != 0._r) unsafeAtomicAdd(a + 2, v.z);
}

// ptracer.cpp:57-167 — one particle: a point on a light (scene.cpp:795-814), its
// connection to the camera, then a random walk that connects every medium and
// non-null surface vertex to the camera. The emitter's geometric frame is the sampled
// face's normal (the reference leaves that frame unset there).
DI void traceParticle(Tracer<1> &T, const DCamera &cam, int max_bounces, Rng &r, real *acc) {
    const DScene &S = T.S;
    if (S.nlights == 0) return;
    const int W = cam.cropped ? cam.cw : cam.width, H = cam.cropped ? cam.ch : cam.height;
    const f3 cpos = ld3(cam.cpos);
    real r4[4] = {r.next(), r.next(), r.next(), r.next()};
    f3 lp, ln;
    const int l = T.sampleLight(r4, lp, ln);
    const DEmitter &E = S.emitters[l];
    const DShape &lsh = S.shapes[E.shape_id];
    f3 power = ld3(E.I) * (S.light_areas[l] / S.light_pmf[l]);
    {   // ptracer.cpp:118-126
        real px, py;
        f3 dir;
        real v = cameraSampleDirect(cam, lp, px, py, dir);
        if (v != 0._r) {
            const int m = dot(dir, ln) > 0._r ? lsh.med_ext : lsh.med_int;
            v *= T.evalTransmittance(lp, dir, true, m, length(lp - cpos), r, max_bounces);
            if (v != 0._r) splat(acc, W, H, px, py, power * (v * fmax(0._r, dot(ln, dir))));
        }
    }
    const real u0 = r.next(), u1 = r.next();
    f3 d = Frame(ln).toWorld(squareToCosineHemisphere(u0, u1));
    power = power * PI;
    f3 o = lp;
    int med = dot(d, ln) > 0._r ? lsh.med_ext : lsh.med_int;
    f3 thr{1, 1, 1};
    bool onSurf = true;
    Hit its;
    for (int depth = 0; !allZero(thr, 0._r) && depth < max_bounces; ++depth) {
        T.rayIntersect(o, d, onSurf, its);
        const int max_inter = max_bounces - depth - 1;
        bool inside = false;
        if (med >= 0) {
            const real a = r.next(), b = r.next();
            f3 ps;
            inside = sampleDistance(S.media[med], o, d, its.t, a, b, r, ps, thr);
            if (inside) o = ps;
        }
        if (inside) {
            const DPhase &ph = S.phases[S.media[med].phase_id];
            real px, py;
            f3 wo;
            real v = cameraSampleDirect(cam, o, px, py, wo);   // ptracer.cpp:152-165
            if (v != 0._r) {
                v *= T.evalTransmittance(o, wo, false, med, length(o - cpos), r, max_inter);
                if (v != 0._r) splat(acc, W, H, px, py, thr * power * (v * phaseEval(ph, -d, wo)));
            }
            const real a = r.next(), b = r.next();
            d = phaseSample(ph, -d, a, b);   // weight eval/pdf = 1
            onSurf = false;
        } else {
            if (its.face < 0) break;
            const DShape &sh = shapeOf(S, its.face);
            const DBsdf &b = bsdfOf(S, sh);
            const Frame shf(its.sn);
            if (b.type != BSDF_NULL) {   // ptracer.cpp:128-150
                real px, py;
                f3 wo;
                real v = cameraSampleDirect(cam, its.p, px, py, wo);
                if (v != 0._r) {
                    v *= T.evalTransmittance(its.p, wo, true, targetMedium(sh, its.gn, wo), length(its.p - cpos), r, max_inter);
                    if (v != 0._r) {
                        const f3 wi = shf.toWorld(its.wi), wol = shf.toLocal(wo);
                        const real wiG = dot(wi, its.gn), woG = dot(wo, its.gn);
                        if (wiG * its.wi.z > 0._r && woG * wol.z > 0._r) {
                            const real corr = fabs((its.wi.z * woG) / (wol.z * wiG));
                            splat(acc, W, H, px, py, bsdfEval(b, its.wi, wol, true) * thr * power * (v * corr));
                        }
                    }
                }
            }
            const real q0 = r.next(), q1 = r.next(), q2 = r.next();
            f3 wol;
            real bpdf = 0._r;
            const f3 bw = bsdfSample(b, its.wi, q0, q1, q2, wol, bpdf, true);
            if (allZero(bw, 0._r)) break;
            const f3 wo = shf.toWorld(wol);
            const real wiG = dot(-d, its.gn), woG = dot(wo, its.gn);
            if (wiG * its.wi.z <= 0._r || woG * wol.z <= 0._r) break;   // shading-normal light leaks [Veach p. 158]
            thr = thr * bw;
            if (isTransition(sh)) med = woG > 0._r ? sh.med_ext : sh.med_int;
            o = its.p;
            d = wo;
            onSurf = true;
        }
    }
}

#ifndef DT_N
// ParticleTracer::render (ptracer.cpp:10-55): particles s0 .. s0+n-1, each its own
// stream keyed (seed, particle index), splatted into acc[pixel][3] of view 0's window.
__global__ __launch_bounds__(PRIMAL_BLOCK, CHILD_WAVES_PER_SIMD) void k_particles(DScene S, RenderParams P, uint64_t s0,
                                                                                  uint64_t n, real *acc,
                                                                                  unsigned long long *work) {
    const uint32_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const uint64_t li = (uint64_t)blk * PRIMAL_BLOCK + threadIdx.x;
    Tracer<1> T(S);
    T.useLdsTree(P);
    Rng r;
    r.init(S.views[0].seed, s0 + li);
    if (li < n) traceParticle(T, S.views[0].cam, P.max_bounces, r, acc);
    flushWork(work, 0, T.nbox, T.ntri, r.nlk, T.nseg);
}

// out[i] = acc[i] * scale over n floats (the pixel mean of VolPathTracer)
__global__ void k_scale(const real *acc, float *out, uint64_t n, real scale) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = (float)(acc[i] * scale);
}
#endif  // DT_N

// Path-replay adjoint of the density derivative of the interior terms: for each AD
// sample a TOTAL pass computes its radiance, then a SCATTER pass replays the same
// stream and adds d loss / d node density (loss gradient dimg[pixel] / spp) into
// grad. The host restricts the terms to emission, surface and volume interior
// (no edge or Lins sub-paths), so nothing is pushed.
template <int N>
__global__ __launch_bounds__(BLOCK, adWaves<N>()) void k_adjoint(DScene S, RenderParams P, uint64_t s0, uint64_t nsamp,
                                                                   const float *dimg, float *grad, int pmed,
                                                                   unsigned *stats, unsigned long long *work) {
    const uint32_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const uint64_t li = (uint64_t)blk * BLOCK + threadIdx.x;
    Tracer<N> T(S);
    T.useLdsTree(P);
    const uint64_t sid = s0 + li;
    const int v = viewOf(S, sid);
    const DView &vw = S.views[v];
    const uint64_t ls = sid - vw.s0;   // sample index within the view: keys its random stream
    const int pixel = (int)(ls / (uint64_t)vw.spp);
    const size_t gp = (size_t)vw.pix0 + pixel;   // pixel of the concatenated loss-gradient images
    Pusher u{nullptr, nullptr, (int)gp, v, 0, 0, 0, Pusher::NO_FAIL};
    uint32_t nlk = 0;
    if (li < nsamp) {
        const real inv = 1.0_r / vw.spp;
        const f3 adj{dimg[3 * gp] * inv, dimg[3 * gp + 1] * inv, dimg[3 * gp + 2] * inv};
        if (adj.x != 0._r || adj.y != 0._r || adj.z != 0._r) {
            Rng r;
            r.init(vw.seed, ls);
            Ds<N> L;
            adPathSample<N, AdjPass::TOTAL>(T, P, vw, r, pixel, L, u, stats, nullptr);
            nlk += r.nlk;
            const f3 Lt = L.val();
            if (Lt.x != 0._r || Lt.y != 0._r || Lt.z != 0._r) {
                AdjCtx ac{grad, S.media + pmed, adj, Lt};
                Rng r2;
                r2.init(vw.seed, ls);
                Ds<N> L2;
                adPathSample<N, AdjPass::SCATTER>(T, P, vw, r2, pixel, L2, u, stats, &ac);
                nlk += r2.nlk;
            }
        }
    }
    flushWork(work, 0, T.nbox, T.ntri, nlk, T.nseg);
}

#ifndef DT_N
// Graph-safe clears and counters (captured graphs use kernels, not memset nodes).
__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_add_u32(unsigned *dst, const unsigned *src) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst += *src;
}

// Triangle records in leaf order from the current vertex positions: for leaf-order
// slot i of face f (kept in v0.w), v0 | f, e1 = v1 - v0, e2 = v2 - v0 (the BVH build's
// records, rebuilt after a shape's vertices change).
__global__ void k_tri_records(float4 *tris, const int *F, const float *V, int ntris) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntris) return;
    float4 *t = tris + 3 * (size_t)i;
    const float4 old = t[0];
    const int f = __float_as_int(old.w);
    const int *fi = F + 3 * (size_t)f;
    const float *a = V + 3 * (size_t)fi[0], *b = V + 3 * (size_t)fi[1], *c = V + 3 * (size_t)fi[2];
    t[0] = make_float4(a[0], a[1], a[2], old.w);
    t[1] = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.f);
    t[2] = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.f);
}

// Refit of one tree level: each listed node's two child boxes from their contents, a
// leaf child from its triangles' vertices, an inner child from its own two (already
// refitted) child boxes. Links and the tree's shape stay; levels run deepest first.
__global__ void k_refit(float4 *nodes, const float4 *tris, const int *F, const float *V, const int *list, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 *nd = nodes + 4 * (size_t)list[i];
    for (int k = 0; k < 2; ++k) {
        const int link = __float_as_int(k == 0 ? nd[0].w : nd[1].w);
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        if (link >= 0) {
            const float4 *c = nodes + 4 * (size_t)link;
            lo[0] = fminf(c[0].x, c[2].x); lo[1] = fminf(c[0].y, c[2].y); lo[2] = fminf(c[0].z, c[2].z);
            hi[0] = fmaxf(c[1].x, c[3].x); hi[1] = fmaxf(c[1].y, c[3].y); hi[2] = fmaxf(c[1].z, c[3].z);
        } else {
            const int enc = ~link, first = enc >> 4, cnt = enc & 15;
            if (cnt == 0) continue;   // the empty slot of a single-leaf root keeps its box
            for (int t = first; t < first + cnt; ++t) {
                const int *fi = F + 3 * (size_t)__float_as_int(tris[3 * (size_t)t].w);
                for (int q = 0; q < 3; ++q) {
                    const float *p = V + 3 * (size_t)fi[q];
                    for (int a = 0; a < 3; ++a) { lo[a] = fminf(lo[a], p[a]); hi[a] = fmaxf(hi[a], p[a]); }
                }
            }
        }
        float4 &L = nd[2 * k], &H = nd[2 * k + 1];
        L = make_float4(lo[0], lo[1], lo[2], L.w);
        H = make_float4(hi[0], hi[1], hi[2], H.w);
    }
}

// One view of an L2 image loss sum (I - T)^2 over the value channel: dimg = 2 (I - T),
// the loss added into *loss (one double atomic per wave).
__global__ void k_l2_grad(const float *img, const float *target, float *dimg, uint64_t n, double *loss) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const float e = img[i] - target[i];
        dimg[i] = 2.f * e;
        e2 = (double)e * e;
    }
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off);
    if ((threadIdx.x & 63) == 0 && e2 != 0.0) unsafeAtomicAdd(loss, e2);
}

// grad[k] += sum_i dimg[i] * tangent_k[i] for the view image img = [(nder + 1)][n]
// (value, then one tangent image per lane): the loss gradient of each forward-mode
// parameter, grid-stride over the n entries, one double atomic per wave and lane.
__global__ void k_contract(const float *img, const float *dimg, uint64_t n, int nder, double *grad) {
    double acc[DTRT_NDER_MAX];
#pragma unroll
    for (int k = 0; k < DTRT_NDER_MAX; ++k) acc[k] = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double g = dimg[i];
#pragma unroll
        for (int k = 0; k < DTRT_NDER_MAX; ++k)
            if (k < nder) acc[k] += g * (double)img[(uint64_t)(k + 1) * n + i];
    }
#pragma unroll
    for (int k = 0; k < DTRT_NDER_MAX; ++k) {
        if (k >= nder) break;
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0 && v != 0.0) unsafeAtomicAdd(grad + k, v);
    }
}

// dst = max(src, 0) (NaN -> 0), and the maximum into *maxbits (non-negative floats
// order like their bit patterns). dst may alias src.
__global__ void k_load_density(float *dst, const float *src, uint64_t n, unsigned *maxbits) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float v = 0.f;
    if (i < n) {
        v = fmaxf(src[i], 0.f);
        dst[i] = v;
    }
    unsigned b = __float_as_uint(v);
    for (int off = 32; off > 0; off >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, b);
}
#endif  // DT_N

// Adds the finished sub-paths of this wave: lanes with fin set add W * Lv to their
// record's pixel. Lanes that share the first finishing lane's pixel and image are
// summed over the wave first (one atomic per weight); the others add their own.
template <int N>
DI void scatterFinished(const RenderParams &P, const char *q, real *acc, real *acc_pe, bool fin, unsigned rec, f3 Lv) {
    const uint64_t m = __ballot(fin);
    if (!m) return;
    const int nd = P.npix;
    int pix = -1;
    bool prim = false;
    real v[N * 3];
#pragma unroll
    for (int k = 0; k < N * 3; ++k) v[k] = 0._r;
    if (fin) {
        const char *r = q + (size_t)rec * P.rec_stride;
        const ChildHdr *h = reinterpret_cast<const ChildHdr *>(r);
        const real *W = reinterpret_cast<const real *>(r + sizeof(ChildHdr));
        pix = h->pixel;
        prim = (h->meta & 8) != 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            v[3 * j] = W[3 * j] * Lv.x;
            v[3 * j + 1] = W[3 * j + 1] * Lv.y;
            v[3 * j + 2] = W[3 * j + 2] * Lv.z;
        }
    }
    const int lead = __ffsll((unsigned long long)m) - 1;
    const int pix0 = __shfl(pix, lead);
    const bool prim0 = __shfl((int)prim, lead) != 0;
    const bool grp = fin && pix == pix0 && prim == prim0;
    if (fin && !grp) {
        real *dst = prim ? acc_pe : acc + (size_t)nd * 3;
#pragma unroll
        for (int k = 0; k < N * 3; ++k)
            if (v[k] != 0._r) unsafeAtomicAdd(dst + ((size_t)(k / 3) * nd + pix) * 3 + (k % 3), v[k]);
    }
#pragma unroll
    for (int k = 0; k < N * 3; ++k) {
        real sv = grp ? v[k] : 0._r;
        for (int off = 32; off > 0; off >>= 1) sv += __shfl_xor(sv, off);
        v[k] = sv;
    }
    if ((int)(threadIdx.x & 63) == lead) {
        real *dst = prim0 ? acc_pe : acc + (size_t)nd * 3;
#pragma unroll
        for (int k = 0; k < N * 3; ++k)
            if (v[k] != 0._r) unsafeAtomicAdd(dst + ((size_t)(k / 3) * nd + pix0) * 3 + (k % 3), v[k]);
    }
}

// Persistent: each lane holds one primal sub-path and advances it one bounce per
// iteration; a lane whose path has ended takes the next queued record (one atomic
// per wave on qnext), so the lanes of a wave stay busy until the queue is drained.
template <int N>
__global__ __launch_bounds__(MAX_BLOCK, CHILD_WAVES_PER_SIMD) void k_children(DScene S, RenderParams P, const char *q, const unsigned *qcount,
                                                    unsigned *qnext, real *acc, real *acc_pe, unsigned long long *work) {
    Tracer<N> T(S);
    T.useLdsTree(P);
    const unsigned n = min(*qcount, P.qcap);
    const int lane = threadIdx.x & 63;
    uint32_t nlk = 0;
    bool alive = false, exhausted = false;
    unsigned rec = 0;
    LiPath path;
    Rng r;
    f3 pre{0, 0, 0}, cw{0, 0, 0};
    while (true) {
        bool fin = false;
        f3 Lv{0, 0, 0};
        if (!exhausted) {
            const uint64_t need = __ballot(!alive);
            if (need) {
                const int lead = __ffsll((unsigned long long)need) - 1;
                unsigned base = 0;
                if (lane == lead) base = atomicAdd(qnext, (unsigned)__popcll(need));
                base = __shfl(base, lead);
                if (base + __popcll(need) >= n) exhausted = true;
                const unsigned my = base + (unsigned)__popcll(need & ((1ull << lane) - 1ull));
                if (!alive && my < n) {
                    rec = my;
                    const ChildHdr h = *reinterpret_cast<const ChildHdr *>(q + (size_t)rec * P.rec_stride);
                    r.init(S.views[h.view].seed ^ 0x6a09e667f3bcc909ull, ((uint64_t)h.key_hi << 32) | h.key_lo);
                    const int kind = h.meta & 3, depth = (h.meta >> 4) & 255, med = ((h.meta >> 12) & 255) - 1;
                    const f3 org{h.org[0], h.org[1], h.org[2]}, dir{h.dir[0], h.dir[1], h.dir[2]};
                    bool has = true;
                    if (kind == CHILD_LI) {
                        pre = f3{0, 0, 0};
                        cw = f3{1, 1, 1};
                        T.liStart(path, org, dir, depth, med, true);
                    } else {
                        Hit its0;
                        its0.face = h.face;
                        its0.p = org;
                        its0.gn = f3{h.gn[0], h.gn[1], h.gn[2]};
                        its0.sn = f3{h.sn[0], h.sn[1], h.sn[2]};
                        its0.wi = Frame(its0.sn).toLocal(-dir);
                        its0.t = 0._r;
                        pre = T.linsStart(r, its0, dir, med, depth, kind == CHILD_LINS_NULL, path, cw, has);
                    }
                    alive = has;
                    if (!has) { fin = true; Lv = pre; nlk += r.nlk; }
                }
            }
        }
        if (alive && !T.liStep(path, r)) {
            alive = false;
            fin = true;
            Lv = pre + cw * path.ret;
            nlk += r.nlk;
        }
        scatterFinished<N>(P, q, acc, acc_pe, fin, rec, Lv);
        if (exhausted && !__any(alive)) break;
    }
    flushWork(work, 4, T.nbox, T.ntri, nlk, T.nseg);
}

// Persistent drain of CHILD_PATH records (the surface-only integrators' sub-paths,
// path.cpp:11-61 from the record's origin and direction), as k_children.
template <int N>
__global__ __launch_bounds__(MAX_BLOCK, CHILD_WAVES_PER_SIMD) void k_children_path(DScene S, RenderParams P, const char *q,
                                                                            const unsigned *qcount, unsigned *qnext,
                                                                            real *acc, real *acc_pe,
                                                                            unsigned long long *work) {
    Tracer<N> T(S);
    T.useLdsTree(P);
    const unsigned n = min(*qcount, P.qcap);
    const int lane = threadIdx.x & 63;
    uint32_t nlk = 0;
    bool alive = false, exhausted = false;
    unsigned rec = 0;
    LiPath path;
    Rng r;
    while (true) {
        bool fin = false;
        f3 Lv{0, 0, 0};
        if (!exhausted) {
            const uint64_t need = __ballot(!alive);
            if (need) {
                const int lead = __ffsll((unsigned long long)need) - 1;
                unsigned base = 0;
                if (lane == lead) base = atomicAdd(qnext, (unsigned)__popcll(need));
                base = __shfl(base, lead);
                if (base + __popcll(need) >= n) exhausted = true;
                const unsigned my = base + (unsigned)__popcll(need & ((1ull << lane) - 1ull));
                if (!alive && my < n) {
                    rec = my;
                    const ChildHdr h = *reinterpret_cast<const ChildHdr *>(q + (size_t)rec * P.rec_stride);
                    r.init(S.views[h.view].seed ^ 0x6a09e667f3bcc909ull, ((uint64_t)h.key_hi << 32) | h.key_lo);
                    const f3 org{h.org[0], h.org[1], h.org[2]}, dir{h.dir[0], h.dir[1], h.dir[2]};
                    Hit h0;
                    T.rayIntersect(org, dir, true, h0);
                    T.liStartAt(path, org, dir, (h.meta >> 4) & 255, -1, true, h0);
                    alive = true;
                }
            }
        }
        if (alive && !T.pathStep(path, r)) {
            alive = false;
            fin = true;
            Lv = path.ret;
            nlk += r.nlk;
        }
        scatterFinished<N>(P, q, acc, acc_pe, fin, rec, Lv);
        if (exhausted && !__any(alive)) break;
    }
    flushWork(work, 4, T.nbox, T.ntri, nlk, T.nseg);
}

// edge_manager.cpp:112-134 + integratorAD.cpp:126-144 for one primary-edge sample
template <int N>
DI void primaryEdgeSample(Tracer<N> &T, const RenderParams &P, const DView &vw, int view, Rng &r, uint64_t sid,
                          uint64_t total, char *q, unsigned *qcount, unsigned *stats) {
    const DCamera &c = vw.cam;
    real u = r.next();
    int idx = sampleReuse(vw.pe_cdf, vw.npe, u, nullptr);   // edge_manager.cpp:112-134
    const int stride = 16 + 4 * N;
    const float *pe = vw.pe + (size_t)idx * stride;
    double px_ = pe[0] + (double)u * (pe[2] - pe[0]), py_ = pe[1] + (double)u * (pe[3] - pe[1]);
    real pd[N][2];
    bool nz = false;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const float *g = pe + 16 + 4 * j;
        pd[j][0] = g[0] + u * (g[2] - g[0]);
        pd[j][1] = g[1] + u * (g[3] - g[1]);
        nz = nz || fabs(pd[j][0]) > EPS || fabs(pd[j][1]) > EPS;
    }
    if (!nz) return;
    int ipx = (int)floor(px_), ipy = (int)floor(py_);
    double nx = pe[3] - pe[1], ny = pe[0] - pe[2], nl = sqrt(nx * nx + ny * ny);
    nx /= nl;
    ny /= nl;
    // maxD: distance to the edge point in camera space (edge_manager.cpp:119-133)
    double fov = c.fov, asp = c.aspect;
    double tx = (px_ / c.width - 0.5) * 2.0 / fov, ty = (py_ / c.height - 0.5) * (-2.0) / (asp * fov), tz = 1.0;
    double v0p[3] = {pe[4], pe[5], pe[6]}, v1p[3] = {pe[7], pe[8], pe[9]}, v0c[3] = {pe[10], pe[11], pe[12]}, v1c[3] = {pe[13], pe[14], pe[15]};
    double d0 = sqrt((tx - v0p[0]) * (tx - v0p[0]) + (ty - v0p[1]) * (ty - v0p[1]) + (tz - v0p[2]) * (tz - v0p[2]));
    double d01 = sqrt((v1p[0] - v0p[0]) * (v1p[0] - v0p[0]) + (v1p[1] - v0p[1]) * (v1p[1] - v0p[1]) + (v1p[2] - v0p[2]) * (v1p[2] - v0p[2]));
    double t0 = d0 / d01;
    double cb = sqrt(v0p[0] * v0p[0] + v0p[1] * v0p[1] + v0p[2] * v0p[2]) / sqrt(v0c[0] * v0c[0] + v0c[1] * v0c[1] + v0c[2] * v0c[2]);
    double cc = sqrt(v1p[0] * v1p[0] + v1p[1] * v1p[1] + v1p[2] * v1p[2]) / sqrt(v1c[0] * v1c[0] + v1c[1] * v1c[1] + v1c[2] * v1c[2]);
    double bx = (1 - t0) * cb, by = t0 * cc, t1 = by / (bx + by);   // utils.cpp:256-266
    double mx = v0c[0] + t1 * (v1c[0] - v0c[0]), my = v0c[1] + t1 * (v1c[1] - v0c[1]), mz = v0c[2] + t1 * (v1c[2] - v0c[2]);
    real maxD = (real)sqrt(mx * mx + my * my + mz * mz);
    // camera rays in double: the two sides of the edge are PRIMARY_DELTA pixels apart
    auto camdir = [&](double x, double y) {
        x /= c.width;
        y /= c.height;
        double n0 = (x - 0.5) * 2.0, n1 = (y - 0.5) * (-2.0) / asp;
        double dx = c.ndc2c[0] * n0 + c.ndc2c[1] * n1 + c.ndc2c[2], dy = c.ndc2c[3] * n0 + c.ndc2c[4] * n1 + c.ndc2c[5],
               dz = c.ndc2c[6] * n0 + c.ndc2c[7] * n1 + c.ndc2c[8];
        double l = sqrt(dx * dx + dy * dy + dz * dz);
        dx /= l; dy /= l; dz /= l;
        return f3{(real)(c.c2w[0] * dx + c.c2w[1] * dy + c.c2w[2] * dz), (real)(c.c2w[4] * dx + c.c2w[5] * dy + c.c2w[6] * dz),
                  (real)(c.c2w[8] * dx + c.c2w[9] * dy + c.c2w[10] * dz)};
    };
    f3 org{c.c2w[3], c.c2w[7], c.c2w[11]};
    real trans = T.evalTransmittance(org, camdir(px_, py_), false, c.med_id, maxD, r, P.max_bounces - 1);
    if (!(trans > EPS)) return;
    bool inside = c.cropped ? (ipx >= c.ox && ipx < c.ox + c.cw && ipy >= c.oy && ipy < c.oy + c.ch)
                            : (ipx >= 0 && ipx < c.width && ipy >= 0 && ipy < c.height);
    if (!inside) return;
    int pix = vw.pix0 + (c.cropped ? (ipx - c.ox) + c.cw * (ipy - c.oy) : ipx + c.width * ipy);
    real W[N * 3];
    real scale = (real)(vw.pe_sum / (double)total);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        real w = (real)(nx * pd[j][0] + ny * pd[j][1]) * scale;
        W[3 * j] = w; W[3 * j + 1] = w; W[3 * j + 2] = w;
    }
    ChildHdr h;
    h.org[0] = org.x; h.org[1] = org.y; h.org[2] = org.z;
    h.pixel = pix;
    h.view = view;
    // the integrator's pixelColor: volpathAD.cpp:623-629, pathAD.cpp:216-219, directAD.cpp:101-125
    if (P.ad_kind == AD_VOLPATH) h.meta = CHILD_LI | 8 | (P.max_bounces << 4) | ((c.med_id + 1) << 12);
    else h.meta = CHILD_PATH | 8 | ((P.ad_kind == AD_DIRECT ? 1 : P.max_bounces) << 4);
    h.face = -1;
    for (int s = 0; s < 2; ++s) {
        double sg = s == 0 ? -1.0 : 1.0;
        f3 d = camdir(px_ + sg * nx * PRIMARY_DELTA, py_ + sg * ny * PRIMARY_DELTA);
        h.dir[0] = d.x; h.dir[1] = d.y; h.dir[2] = d.z;
        uint64_t ck = ((1ull << 62) | sid) * 4 + s;
        h.key_lo = (uint32_t)ck; h.key_hi = (uint32_t)(ck >> 32);
        if (s == 1)
#pragma unroll
            for (int k = 0; k < N * 3; ++k) W[k] = -W[k];
        // the queue holds two records per primary-edge sample of the launch: cannot fail
        if (!writeRecord(P, q, qcount, h, W, N * 3)) atomicAdd(stats + STAT_QUEUE_OVERFLOW, 1u);
    }
}

template <int N>
__global__ __launch_bounds__(BLOCK) void k_primary(DScene S, RenderParams P, int view, uint64_t s0, uint64_t nsamp,
                                                   uint64_t total, char *q, unsigned *qcount, unsigned *stats,
                                                   unsigned long long *work) {
    Tracer<N> T(S);
    T.useLdsTree(P);
    const uint64_t li = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t sid = s0 + li;
    Rng r;
    const DView &vw = S.views[view];
    r.init(vw.seed, (1ull << 62) | sid);
    if (li < nsamp) primaryEdgeSample<N>(T, P, vw, view, r, sid, total, q, qcount, stats);
    flushWork(work, 8, T.nbox, T.ntri, r.nlk, T.nseg);
}

// One view: pixels pix0 .. pix0+npix of the accumulators into its image.
template <int N>
__global__ void k_finalize(RenderParams P, int npix, int pix0, int spp, int has_pe, const real *acc, const real *acc_pe,
                           float *out, unsigned *stats) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const size_t nd = P.npix, g = (size_t)pix0 + p;   // accumulator stride and pixel; output stride is npix
    if (spp > 0) {
        real inv = 1.0_r / spp;
        real v0 = acc[g * 3] * inv, v1 = acc[g * 3 + 1] * inv, v2 = acc[g * 3 + 2] * inv;
        if (fmin(v0, fmin(v1, v2)) < 0._r) atomicAdd(stats + STAT_NEG_PIXEL, 1u);
        else { out[p * 3] = (float)v0; out[p * 3 + 1] = (float)v1; out[p * 3 + 2] = (float)v2; }
        for (int j = 0; j < N; ++j) {
            const real *a = acc + ((j + 1) * nd + g) * 3;
            real g0 = a[0] * inv, g1 = a[1] * inv, g2 = a[2] * inv;
            if (!isfinite(g0) || !isfinite(g1) || !isfinite(g2)) { atomicAdd(stats + STAT_NONFINITE, 1u); continue; }
            float *o = out + ((size_t)(j + 1) * npix + p) * 3;
            o[0] = (float)g0; o[1] = (float)g1; o[2] = (float)g2;
        }
    }
    if (has_pe)
        for (int j = 0; j < N; ++j) {
            const real *a = acc_pe + (j * nd + g) * 3;
            float *o = out + ((size_t)(j + 1) * npix + p) * 3;
            o[0] = (float)(o[0] + a[0]); o[1] = (float)(o[1] + a[1]); o[2] = (float)(o[2] + a[2]);
        }
}

// ---------------------------------------------------------------- test hooks
#ifndef DT_N
__global__ void k_trace(DScene S, const float *rays, int n, int *out_face, float *out_t) {
    Tracer<1> T(S);
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float *r = rays + (size_t)i * 8;
    f3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
    real t;
    int f = T.trace(o, d, r[6], r[7], t);
    out_face[i] = f;
    if (f >= 0) { Hit h; refineHit(S, f, o, d, h); out_t[i] = (float)h.t; }
    else out_t[i] = (float)FINF;
}
__global__ void k_li(DScene S, const float *rays, int n, int depth, int med, uint64_t seed, float *out) {
    Tracer<1> T(S);
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float *rr = rays + (size_t)i * 6;
    Rng r;
    r.init(seed, (uint64_t)i);
    f3 L = T.Li(r, f3{rr[0], rr[1], rr[2]}, f3{rr[3], rr[4], rr[5]}, depth, med, true);
    out[3 * i] = (float)L.x; out[3 * i + 1] = (float)L.y; out[3 * i + 2] = (float)L.z;
}
__global__ void k_medium(DScene S, int med, int kind, const float *seg, int n, uint64_t seed, float *out) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float *s = seg + (size_t)i * 8;
    Rng r;
    r.init(seed, (uint64_t)i);
    f3 o{s[0], s[1], s[2]}, d{s[3], s[4], s[5]};
    const DMedium &m = S.media[med];
    if (kind == 0) out[i] = (float)mediumTransmittance(m, o, d, s[6], s[7], r);
    else {
        real a = r.next(), b = r.next();
        f3 ps, thr{1, 1, 1};
        out[i] = sampleDistance(m, o, d, s[7], a, b, r, ps, thr) ? 1.f : 0.f;
    }
}
__global__ void k_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t *out) {
    u4 o = philox4x32_10(u4{c0, c1, c2, c3}, k0, k1);
    out[0] = o.a; out[1] = o.b; out[2] = o.c; out[3] = o.d;
}
#endif  // DT_N
template <int N> __global__ void k_grid(DScene S, int med, const float *pts, int n, float *out) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float *p = pts + (size_t)i * (3 + 3 * N);
    Dv<N> x(f3{p[0], p[1], p[2]});
    for (int j = 0; j < N; ++j) x.setGrad(j, f3{p[3 + 3 * j], p[4 + 3 * j], p[5 + 3 * j]});
    Df<N> v = lookupDensityAD<N>(S.media[med], x);
    out[(size_t)i * (1 + N)] = (float)v.v;
    for (int j = 0; j < N; ++j) out[(size_t)i * (1 + N) + 1 + j] = (float)v.d[j];
}

// ------------------------------------------------------------------ launchers
#define DT_FOR_N(X) X(1) X(2) X(3) X(4) X(8) X(16)

// Launchers of one dual width: defined and instantiated in that width's object,
// called through the dispatchers of the width-independent one.
template <int N> struct WidthOps {
    static hipError_t setFrameRotation(const double *m9);
    static void launchAd(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const uint4 *replay, real *acc,
                         char *q, unsigned *qc, unsigned *st, uint4 *ovf, unsigned *ovf_count, unsigned long long *wk,
                         hipStream_t sm);
    static void launchChildren(const DScene &S, const RenderParams &P, const char *q, unsigned *qc, real *acc,
                               real *acc_pe, unsigned long long *wk, unsigned grid, hipStream_t sm);
    static void launchPrimary(const DScene &S, const RenderParams &P, int view, uint64_t s0, uint64_t n, uint64_t total,
                              char *q, unsigned *qc, unsigned *st, unsigned long long *wk, hipStream_t sm);
    static void launchFinalize(const RenderParams &P, int npix, int pix0, int spp, int has_pe, const real *acc,
                               const real *acc_pe, float *out, unsigned *st, hipStream_t sm);
    static void launchGrid(const DScene &S, int med, const float *pts, int n, float *out, hipStream_t sm);
    static void launchAdjoint(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const float *dimg,
                              float *grad, int pmed, unsigned *st, unsigned long long *wk, hipStream_t sm);
    static hipError_t kernelInfo(KernelInfo *out, int cap, int *count);
};

// Appends the resources of kernel `fn`, launched with `block` threads and `lds` bytes of
// dynamic LDS, to out[*count] (nothing beyond cap).
template <class F>
static hipError_t addKernelInfo(F fn, const char *name, int block, size_t lds, KernelInfo *out, int cap, int *count) {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, reinterpret_cast<const void *>(fn));
    if (e != hipSuccess) return e;
    int blocks = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, block, lds);
    if (e != hipSuccess) return e;
    if (*count < cap) {
        KernelInfo &k = out[*count];
        std::snprintf(k.name, sizeof(k.name), "%s", name);
        k.private_bytes = (uint32_t)a.localSizeBytes;
        k.vgprs = (uint32_t)a.numRegs;
        k.block = (uint32_t)block;
        k.waves_per_cu = (uint32_t)(blocks * ((block + 63) / 64));
    }
    ++*count;
    return hipSuccess;
}

// The part of the scene's tree a launch of `kern` with blocks of `blk` keeps in LDS
// (Pl.lds_nodes, Pl.lds_tris), chosen so that a CU holds as many blocks as with the
// traversal stacks alone: all nodes and triangle records when they fit, else, in a
// kernel built to read a prefix (`prefix`: treePrefixInLds), the first nodes (the top
// levels), at least 64 of them. Returns the extra LDS bytes.
template <class K> static size_t treeLdsFor(K kern, unsigned blk, const DScene &S, RenderParams &Pl, bool prefix) {
    Pl.lds_nodes = 0;
    Pl.lds_tris = 0;
    if (S.nnodes <= 0 || Pl.tree_in_hbm) return 0;
    int dev = 0, per_cu = 0, per_block = 0, base = 0, with = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&base, kern, (int)blk, traceLds(blk, S.stack_depth)) != hipSuccess || base <= 0) {
        (void)hipGetLastError();
        return 0;
    }
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(kern)) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    const size_t used = traceLds(blk, S.stack_depth) + fa.sharedSizeBytes;
    const size_t room = std::min<size_t>((size_t)per_cu / (size_t)base, (size_t)per_block);
    if (room <= used) return 0;
    const size_t spare = room - used;
    if (treeLdsBytes(S, S.nnodes, true) <= spare) {
        Pl.lds_nodes = S.nnodes;
        Pl.lds_tris = 1;
    } else if (prefix) {
        Pl.lds_nodes = (int)std::min<size_t>((size_t)S.nnodes, spare / 64);
        if (Pl.lds_nodes < 64) Pl.lds_nodes = 0;
    }
    const size_t tb = treeLdsBytes(S, Pl.lds_nodes, Pl.lds_tris != 0);
    if (tb == 0 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&with, kern, (int)blk, traceLds(blk, S.stack_depth) + tb) != hipSuccess ||
        with < base) {
        (void)hipGetLastError();
        Pl.lds_nodes = 0;
        Pl.lds_tris = 0;
        return 0;
    }
    return tb;
}

template <class Sym> static hipError_t setRotation(const Sym &sym, const double *m9) {
    FrameRot r;
    for (int i = 0; i < 9; ++i) r.m[i] = (real)m9[i];
    return hipMemcpyToSymbol(sym, &r, sizeof(FrameRot));
}

#ifdef DT_N
template <int N> hipError_t WidthOps<N>::setFrameRotation(const double *m9) {
    return setRotation(HIP_SYMBOL(DT_FRAME_ROT), m9);
}
template <int N> void WidthOps<N>::launchAd(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const uint4 *replay,
                                      real *acc, char *q, unsigned *qc, unsigned *st, uint4 *ovf, unsigned *ovf_count,
                                      unsigned long long *wk, hipStream_t sm) {
    const unsigned blk = blockOr64(P.ad_block);
    unsigned nb = (unsigned)((n + blk - 1) / blk);
    RenderParams Pl = P;
    if (P.ad_kind == AD_PATH) {
        const size_t tl = treeLdsFor(k_ad_paths<N, AD_PATH>, blk, S, Pl, treePrefixInLds<N>());
        hipLaunchKernelGGL((k_ad_paths<N, AD_PATH>), dim3(nb), dim3(blk), traceLds(blk, S.stack_depth) + tl, sm, S, Pl, s0, n, replay, acc,
                           q, qc, st, ovf, ovf_count, wk);
    } else if (P.ad_kind == AD_DIRECT) {
        const size_t tl = treeLdsFor(k_ad_paths<N, AD_DIRECT>, blk, S, Pl, treePrefixInLds<N>());
        hipLaunchKernelGGL((k_ad_paths<N, AD_DIRECT>), dim3(nb), dim3(blk), traceLds(blk, S.stack_depth) + tl, sm, S, Pl, s0, n, replay,
                           acc, q, qc, st, ovf, ovf_count, wk);
    } else if constexpr (adRefill<N>()) {
        // one wave per block, AD_REFILL_WAVES_PER_CU of them per CU share the samples
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        const uint64_t waves = (uint64_t)std::max(ncu, 1) * AD_REFILL_WAVES_PER_CU;
        const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((n + BLOCK - 1) / BLOCK, waves));
        const uint64_t per_wave = (n + grid - 1) / grid;
        const size_t tl = treeLdsFor(k_ad_refill<N>, BLOCK, S, Pl, treePrefixInLds<N>());
        // one wave per block: each owns a contiguous range of samples
        hipLaunchKernelGGL(k_ad_refill<N>, dim3((unsigned)grid), dim3(BLOCK), traceLds(BLOCK, S.stack_depth) + tl, sm, S, Pl, s0, n,
                           per_wave, replay, acc, q, qc, st, ovf, ovf_count, wk);
    } else {
        const size_t tl = treeLdsFor(k_ad_paths<N, AD_VOLPATH>, blk, S, Pl, treePrefixInLds<N>());
        hipLaunchKernelGGL((k_ad_paths<N, AD_VOLPATH>), dim3(nb), dim3(blk), traceLds(blk, S.stack_depth) + tl, sm, S, Pl, s0, n, replay,
                           acc, q, qc, st, ovf, ovf_count, wk);
    }
}
template <int N> void WidthOps<N>::launchChildren(const DScene &S, const RenderParams &P, const char *q, unsigned *qc,
                                            real *acc, real *acc_pe, unsigned long long *wk, unsigned grid, hipStream_t sm) {
    const unsigned blk = blockOr64(P.child_block);   // grid: blocks of `blk` threads
    RenderParams Pl = P;
    if (P.ad_kind == AD_VOLPATH) {
        const size_t tl = treeLdsFor(k_children<N>, blk, S, Pl, treePrefixInLds<N>());
        hipLaunchKernelGGL(k_children<N>, dim3(grid), dim3(blk), traceLds(blk, S.stack_depth) + tl, sm, S, Pl, q, qc, qc + 2, acc, acc_pe,
                           wk);
    } else {
        const size_t tl = treeLdsFor(k_children_path<N>, blk, S, Pl, treePrefixInLds<N>());
        hipLaunchKernelGGL(k_children_path<N>, dim3(grid), dim3(blk), traceLds(blk, S.stack_depth) + tl, sm, S, Pl, q, qc, qc + 2, acc,
                           acc_pe, wk);
    }
}
template <int N> void WidthOps<N>::launchPrimary(const DScene &S, const RenderParams &P, int view, uint64_t s0, uint64_t n,
                                           uint64_t total, char *q, unsigned *qc, unsigned *st, unsigned long long *wk,
                                           hipStream_t sm) {
    unsigned nb = (unsigned)((n + BLOCK - 1) / BLOCK);
    RenderParams Pl = P;
    const size_t tl = treeLdsFor(k_primary<N>, BLOCK, S, Pl, treePrefixInLds<N>());
    hipLaunchKernelGGL(k_primary<N>, dim3(nb), dim3(BLOCK), traceLds(BLOCK, S.stack_depth) + tl, sm, S, Pl, view, s0, n, total, q, qc, st,
                       wk);
}
template <int N> void WidthOps<N>::launchFinalize(const RenderParams &P, int npix, int pix0, int spp, int has_pe, const real *acc,
                                            const real *acc_pe, float *out, unsigned *st, hipStream_t sm) {
    hipLaunchKernelGGL(k_finalize<N>, dim3((npix + 255) / 256), dim3(256), 0, sm, P, npix, pix0, spp, has_pe, acc, acc_pe,
                       out, st);
}
template <int N> void WidthOps<N>::launchGrid(const DScene &S, int med, const float *pts, int n, float *out, hipStream_t sm) {
    hipLaunchKernelGGL(k_grid<N>, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, sm, S, med, pts, n, out);
}
template <int N> void WidthOps<N>::launchAdjoint(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n,
                                                 const float *dimg, float *grad, int pmed, unsigned *st,
                                                 unsigned long long *wk, hipStream_t sm) {
    unsigned nb = (unsigned)((n + BLOCK - 1) / BLOCK);
    RenderParams Pl = P;
    const size_t tl = treeLdsFor(k_adjoint<N>, BLOCK, S, Pl, treePrefixInLds<N>());
    hipLaunchKernelGGL(k_adjoint<N>, dim3(nb), dim3(BLOCK), traceLds(BLOCK, S.stack_depth) + tl, sm, S, Pl, s0, n, dimg, grad, pmed, st,
                       wk);
}
template <int N> hipError_t WidthOps<N>::kernelInfo(KernelInfo *out, int cap, int *count) {
    char nm[40];
    hipError_t e = hipSuccess;
#define DT_INFO(FN, LABEL, BLK, LDS)                                             \
    if (e == hipSuccess) {                                                       \
        std::snprintf(nm, sizeof(nm), LABEL "<%d>", N);                          \
        e = addKernelInfo(FN, nm, BLK, LDS, out, cap, count);                     \
    }
    if constexpr (adRefill<N>()) {
        DT_INFO(k_ad_refill<N>, "k_ad_refill", BLOCK, traceLds(BLOCK))
    } else {
        DT_INFO((k_ad_paths<N, AD_VOLPATH>), "k_ad_paths", MAX_BLOCK, traceLds(MAX_BLOCK))
    }
    DT_INFO((k_ad_paths<N, AD_PATH>), "k_ad_paths/path", MAX_BLOCK, traceLds(MAX_BLOCK))
    DT_INFO((k_ad_paths<N, AD_DIRECT>), "k_ad_paths/direct", MAX_BLOCK, traceLds(MAX_BLOCK))
    DT_INFO(k_children<N>, "k_children", MAX_BLOCK, traceLds(MAX_BLOCK))
    DT_INFO(k_children_path<N>, "k_children_path", MAX_BLOCK, traceLds(MAX_BLOCK))
    DT_INFO(k_primary<N>, "k_primary", BLOCK, traceLds(BLOCK))
    DT_INFO(k_adjoint<N>, "k_adjoint", BLOCK, traceLds(BLOCK))
    DT_INFO(k_grid<N>, "k_grid", BLOCK, 0)
    DT_INFO(k_finalize<N>, "k_finalize", 256, 0)
#undef DT_INFO
    return e;
}
template struct WidthOps<DT_N>;

#else  // the width-independent object of this precision: dispatchers over S.nder and the remaining launchers

hipError_t setFrameRotation(const double *m9) {
    hipError_t e = setRotation(HIP_SYMBOL(DT_FRAME_ROT), m9);
#define DT_SET_ROT(n) \
    if (e == hipSuccess) e = WidthOps<n>::setFrameRotation(m9);
    DT_FOR_N(DT_SET_ROT)
#undef DT_SET_ROT
    return e;
}

#define DT_DISPATCH(FN, ...)                              \
    switch (S.nder) {                                     \
        case 1: WidthOps<1>::FN(__VA_ARGS__); break;      \
        case 2: WidthOps<2>::FN(__VA_ARGS__); break;      \
        case 3: WidthOps<3>::FN(__VA_ARGS__); break;      \
        case 4: WidthOps<4>::FN(__VA_ARGS__); break;      \
        case 8: WidthOps<8>::FN(__VA_ARGS__); break;      \
        case 16: WidthOps<16>::FN(__VA_ARGS__); break;    \
        default: return hipErrorInvalidValue;             \
    }                                                     \
    return hipGetLastError();

hipError_t launch_ad_paths(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const uint4 *replay, void *acc,
                           char *q, unsigned *qc, unsigned *st, uint4 *ovf, unsigned *ovf_count, unsigned long long *wk,
                           hipStream_t sm) {
    real *A = static_cast<real *>(acc);
    DT_DISPATCH(launchAd, S, P, s0, n, replay, A, q, qc, st, ovf, ovf_count, wk, sm)
}
hipError_t launch_children(const DScene &S, const RenderParams &P, const char *q, unsigned *qc, void *acc, void *acc_pe,
                           unsigned long long *wk, unsigned grid, hipStream_t sm) {
    real *A = static_cast<real *>(acc), *APE = static_cast<real *>(acc_pe);
    DT_DISPATCH(launchChildren, S, P, q, qc, A, APE, wk, grid, sm)
}
hipError_t launch_primary(const DScene &S, const RenderParams &P, int view, uint64_t s0, uint64_t n, uint64_t total, char *q,
                          unsigned *qc, unsigned *st, unsigned long long *wk, hipStream_t sm) {
    DT_DISPATCH(launchPrimary, S, P, view, s0, n, total, q, qc, st, wk, sm)
}
hipError_t launch_adjoint(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const float *dimg, float *grad,
                          int pmed, unsigned *st, unsigned long long *wk, hipStream_t sm) {
    DT_DISPATCH(launchAdjoint, S, P, s0, n, dimg, grad, pmed, st, wk, sm)
}
hipError_t launch_fill32(void *p, uint64_t n, uint32_t v, hipStream_t sm) {
    const unsigned nb = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill32, dim3(std::max(nb, 1u)), dim3(256), 0, sm, static_cast<uint32_t *>(p), n, v);
    return hipGetLastError();
}
hipError_t launch_add_u32(unsigned *dst, const unsigned *src, hipStream_t sm) {
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, sm, dst, src);
    return hipGetLastError();
}
hipError_t launch_tri_records(const DScene &S, hipStream_t sm) {
    if (S.ntris == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tri_records, dim3((S.ntris + 255) / 256), dim3(256), 0, sm, const_cast<float4 *>(S.tris), S.F, S.V,
                       S.ntris);
    return hipGetLastError();
}
hipError_t launch_refit(const DScene &S, const int *list, int n, hipStream_t sm) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit, dim3((n + 255) / 256), dim3(256), 0, sm, const_cast<float4 *>(S.nodes), S.tris, S.F, S.V,
                       list, n);
    return hipGetLastError();
}
hipError_t launch_l2_grad(const float *img, const float *target, float *dimg, uint64_t n, double *loss, hipStream_t sm) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_l2_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sm, img, target, dimg, n, loss);
    return hipGetLastError();
}
hipError_t launch_contract(const float *img, const float *dimg, uint64_t n, int nder, double *grad, hipStream_t sm) {
    if (n == 0 || nder <= 0) return hipSuccess;
    const unsigned nb = (unsigned)std::min<uint64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_contract, dim3(nb), dim3(256), 0, sm, img, dimg, n, std::min(nder, DTRT_NDER_MAX), grad);
    return hipGetLastError();
}
hipError_t launch_load_density(float *dst, const float *src, uint64_t n, unsigned *maxbits, hipStream_t sm) {
    hipLaunchKernelGGL(k_load_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sm, dst, src, n, maxbits);
    return hipGetLastError();
}
hipError_t launch_finalize(const DScene &S, const RenderParams &P, int npix, int pix0, int spp, int has_pe, const void *acc,
                           const void *acc_pe, float *out, unsigned *st, hipStream_t sm) {
    const real *A = static_cast<const real *>(acc), *APE = static_cast<const real *>(acc_pe);
    DT_DISPATCH(launchFinalize, P, npix, pix0, spp, has_pe, A, APE, out, st, sm)
}
hipError_t launch_grid(const DScene &S, int med, const float *pts, int n, float *out, hipStream_t sm) {
    DT_DISPATCH(launchGrid, S, med, pts, n, out, sm)
}
hipError_t launch_trace(const DScene &S, const float *rays, int n, int *face, float *t, hipStream_t sm) {
    hipLaunchKernelGGL(k_trace, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), traceLds(BLOCK), sm, S, rays, n, face, t);
    return hipGetLastError();
}
hipError_t launch_primal_paths(const DScene &S, const RenderParams &P, int kind, uint64_t s0, uint64_t n, void *acc,
                              unsigned long long *work, hipStream_t sm) {
    const uint64_t nb = (n + PRIMAL_BLOCK - 1) / PRIMAL_BLOCK;
    if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
    real *A = static_cast<real *>(acc);
    RenderParams Pl = P;
    const size_t lds = traceLds(PRIMAL_BLOCK, S.stack_depth);
    if (kind == PRIMAL_SIMPLE) {
        const size_t tl = treeLdsFor(k_primal_paths<PRIMAL_SIMPLE>, PRIMAL_BLOCK, S, Pl, false);
        hipLaunchKernelGGL(k_primal_paths<PRIMAL_SIMPLE>, dim3((unsigned)nb), dim3(PRIMAL_BLOCK), lds + tl, sm, S, Pl, s0,
                           n, A, work);
    } else if (kind == PRIMAL_PATH) {
        const size_t tl = treeLdsFor(k_primal_paths<PRIMAL_PATH>, PRIMAL_BLOCK, S, Pl, false);
        hipLaunchKernelGGL(k_primal_paths<PRIMAL_PATH>, dim3((unsigned)nb), dim3(PRIMAL_BLOCK), lds + tl, sm, S, Pl, s0, n,
                           A, work);
    } else if (kind == PRIMAL_VOLPATH) {
        const size_t tl = treeLdsFor(k_primal_paths<PRIMAL_VOLPATH>, PRIMAL_BLOCK, S, Pl, false);
        hipLaunchKernelGGL(k_primal_paths<PRIMAL_VOLPATH>, dim3((unsigned)nb), dim3(PRIMAL_BLOCK), lds + tl, sm, S, Pl, s0,
                           n, A, work);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_particles(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, void *acc,
                            unsigned long long *work, hipStream_t sm) {
    const uint64_t nb = (n + PRIMAL_BLOCK - 1) / PRIMAL_BLOCK;
    if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
    RenderParams Pl = P;
    const size_t tl = treeLdsFor(k_particles, PRIMAL_BLOCK, S, Pl, false);
    hipLaunchKernelGGL(k_particles, dim3((unsigned)nb), dim3(PRIMAL_BLOCK), traceLds(PRIMAL_BLOCK, S.stack_depth) + tl, sm, S,
                       Pl, s0, n, static_cast<real *>(acc), work);
    return hipGetLastError();
}
hipError_t launch_scale(const DScene &S, const void *acc, float *out, uint64_t n, double scale, hipStream_t sm) {
    (void)S;
    const uint64_t nb = (n + BLOCK - 1) / BLOCK;
    if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)nb), dim3(BLOCK), 0, sm, static_cast<const real *>(acc), out, n, (real)scale);
    return hipGetLastError();
}
hipError_t launch_li(const DScene &S, const float *rays, int n, int depth, int med, uint64_t seed, float *out, hipStream_t sm) {
    hipLaunchKernelGGL(k_li, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), traceLds(BLOCK), sm, S, rays, n, depth, med, seed,
                       out);
    return hipGetLastError();
}
hipError_t launch_medium(const DScene &S, int med, int kind, const float *seg, int n, uint64_t seed, float *out, hipStream_t sm) {
    hipLaunchKernelGGL(k_medium, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, sm, S, med, kind, seg, n, seed, out);
    return hipGetLastError();
}
hipError_t launch_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t sm) {
    hipLaunchKernelGGL(k_philox, dim3(1), dim3(1), 0, sm, ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
    return hipGetLastError();
}
hipError_t kernel_info(const DScene &S, KernelInfo *out, int cap, int *count) {
    *count = 0;
    hipError_t e;
    switch (S.nder) {
#define DT_INFO_N(n) \
        case n: e = WidthOps<n>::kernelInfo(out, cap, count); break;
        DT_FOR_N(DT_INFO_N)
#undef DT_INFO_N
        default: return hipErrorInvalidValue;
    }
    if (e == hipSuccess) e = addKernelInfo(k_primal_paths<PRIMAL_VOLPATH>, "k_primal_paths/volpath", PRIMAL_BLOCK, traceLds(PRIMAL_BLOCK), out, cap, count);
    if (e == hipSuccess) e = addKernelInfo(k_primal_paths<PRIMAL_SIMPLE>, "k_primal_paths/simple", PRIMAL_BLOCK, traceLds(PRIMAL_BLOCK), out, cap, count);
    if (e == hipSuccess) e = addKernelInfo(k_primal_paths<PRIMAL_PATH>, "k_primal_paths/path", PRIMAL_BLOCK, traceLds(PRIMAL_BLOCK), out, cap, count);
    if (e == hipSuccess) e = addKernelInfo(k_particles, "k_particles", PRIMAL_BLOCK, traceLds(PRIMAL_BLOCK), out, cap, count);
    if (e == hipSuccess) e = addKernelInfo(k_refit, "k_refit", 256, 0, out, cap, count);
    return e;
}
#endif  // DT_N

}  // namespace DT_PNS
}  // namespace dt
