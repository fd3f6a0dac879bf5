// Ray-level scene services and the two estimators of the gfx950 path tracer:
//  * Tracer::Li        primal volumetric path tracer      (volpathAD.cpp:50-142)
//  * Tracer::LinsOnce  one sample of the Lins boundary term (volpathAD.cpp:338-441)
// plus transmittance through null interfaces (scene.cpp:177-237), emitter lookup
// (scene.cpp:617-793) and next-event estimation (scene.cpp:399-524).
#pragma once
#include "dscene.h"
#include "launch.h"

namespace dt {
namespace DT_PNS {

DI real misPower(real a, real b) { a *= a; b *= b; return a / (a + b); }   // volpathAD.cpp:43-47

// State of one primal path between bounces (see Tracer::liStep). The hit record of the
// current ray is not carried: each step recomputes it from the face the ray hits.
struct LiPath {
    f3 ro, rd, thr, ret;
    int face;   // face the ray (ro, rd) hits, -1: none
    int depth, max_depth, med;
    bool incEmission;
    // The walk along (ro, rd) that the last bounce set up, run at the start of the next
    // step: 0 none, 1 first hit only (a new ray, or through a null surface), 2 emitter
    // walk after a volume scattering, 3 after a surface scattering. cpdf is the MIS
    // partner pdf of an emitter the walk ends on.
    int walk, walk_max_inter;
    real cpdf;
};

template <int N> struct Tracer {
    const DScene &S;
    uint32_t nbox = 0, ntri = 0, nseg = 0;   // work counters: box tests, triangle tests, traced segments
    int lds_nodes = 0;                       // nodes [0, lds_nodes) are in LDS (treeToLds)
    bool lds_tris = false;                   // all nodes and the triangle records are
    DI explicit Tracer(const DScene &s) : S(s) {}
    // At the top of a kernel, by every thread: copy what the launch made room for into
    // LDS (a kernel that reads no prefix is given the whole tree or nothing).
    DI void useLdsTree(const RenderParams &P) {
        if constexpr (treePrefixInLds<N>()) {
            if (P.lds_nodes > 0) {
                treeToLds(S, P.lds_nodes, P.lds_tris != 0);
                lds_nodes = P.lds_nodes;
                lds_tris = P.lds_tris != 0;
            }
        } else if (P.lds_tris) {
            treeToLds(S, S.nnodes, true);
            lds_tris = true;
        }
    }

    DI int trace(f3 o, f3 d, real tnear, real tfar, real &t) {
        nseg++;
        TraceHit h = traceClosest<treePrefixInLds<N>()>(S.nodes, S.tris, S.nnodes, o, d, tnear, tfar, lds_nodes, lds_tris,
                                                        S.stack_depth);
        nbox += h.nbox;
        ntri += h.ntri;
        t = h.t;
        return h.face;
    }

    // scene.cpp:538-575
    DI bool rayIntersect(f3 o, f3 d, bool onSurface, Hit &h) {
        real t;
        int f = trace(o, d, onSurface ? SHADOW_EPS : 0._r, FINF, t);
        if (f < 0) { h.face = -1; h.t = FINF; return false; }
        refineHit(S, f, o, d, h);
        return true;
    }
    // scene.cpp:577-615
    DI bool rayIntersectAD(const Dv<N> &o, const Dv<N> &d, bool onSurface, HitAD<N> &h) {
        real t;
        int f = trace(o.val(), d.val(), onSurface ? SHADOW_EPS : 0._r, FINF, t);
        if (f < 0) { h.face = -1; h.t = Df<N>(FINF); return false; }
        refineHitAD<N>(S, f, o, d, h);
        return true;
    }

    // scene.cpp:177-237 — shadow-ray transmittance through null interfaces
    DI real evalTransmittance(f3 o, f3 d, bool onSurface, int med, real remaining, Rng &r, int max_inter) {
        real tmin = onSurface ? SHADOW_EPS : 0._r, T = 1._r, tnear = tmin, tfar = (1.0_r - SHADOW_EPS) * remaining;
        int inter = 0;
        while (remaining > 0._r) {
            real tmax = remaining, th;
            int f = trace(o, d, tnear, tfar, th);
            Hit nh;
            if (f >= 0) {
                const DShape &sh = shapeOf(S, f);
                if (inter == max_inter || bsdfOf(S, sh).type != BSDF_NULL) return 0._r;
                refineHit(S, f, o, d, nh);
                if (nh.t < tmax) tmax = nh.t;
            }
            if (med >= 0) T *= mediumTransmittance(S.media[med], o, d, tmin, tmax, r);
            if (f < 0) break;
            const DShape &sh = shapeOf(S, f);
            if (isTransition(sh)) med = targetMedium(sh, nh.gn, d);
            o = o + d * tmax;
            tnear = tmin = SHADOW_EPS;
            remaining -= tmax;
            tfar = (1.0_r - SHADOW_EPS) * remaining;
            inter++;
        }
        return T;
    }

    // scene.cpp:617-701. One trace site: the first hit and each hit past a null
    // interface come from the same traversal. first_only: stop at the first hit (the
    // continuation through a null surface, scene.cpp:538-575).
    DI f3 lookForEmitter(f3 o, f3 d, bool onSurface, Rng &r, int med, int max_inter, Hit &its, real &pdf_nee,
                         bool first_only = false) {
        real tmin = onSurface ? SHADOW_EPS : 0._r, th;
        const f3 scat = o;
        real T = 1._r;
        int inter = 0;
        its.face = -1;
        its.t = FINF;
        while (true) {
            const int f = trace(o, d, tmin, FINF, th);
            if (f < 0) return f3{0, 0, 0};
            Hit nh;
            refineHit(S, f, o, d, nh);
            if (inter == 0) its = nh;
            if (first_only) return f3{0, 0, 0};
            const DShape &sh = shapeOf(S, f);
            real tmax = nh.t;
            if (med >= 0) T *= mediumTransmittance(S.media[med], o, d, tmin, tmax, r);
            if (sh.light_id >= 0) {
                f3 dv = nh.p - scat;
                real G = nh.wi.z / dot(dv, dv);
                pdf_nee = S.light_pmf[sh.light_id] / (S.light_areas[sh.light_id] * G);
                return emitterEval(S.emitters[sh.light_id], nh.gn, nh.sn, -d) * T;
            }
            if (inter == max_inter || bsdfOf(S, sh).type != BSDF_NULL) return f3{0, 0, 0};
            if (isTransition(sh)) med = targetMedium(sh, nh.gn, d);
            o = o + d * tmax;
            tmin = SHADOW_EPS;
            inter++;
        }
    }

    // scene.cpp:703-793 — with tangents; reports the face of the first hit and, when a
    // light ends the walk, the light's face and the walked AD distance. One trace site:
    // the first hit and each hit past a null interface come from the same traversal.
    // first_only: stop at the first hit (the continuation through a null surface,
    // scene.cpp:577-615).
    DI Ds<N> lookForEmitterAD(Dv<N> o, const Dv<N> &d, bool onSurface, Rng &r, int med, int max_inter, int &first_face,
                              real &pdf_nee, int *far_face, real *far_t, const ScatterCtx *scx = nullptr,
                              bool first_only = false) {
        real tmin = onSurface ? SHADOW_EPS : 0._r, th;
        const f3 scat = o.val(), dv = d.val();
        Df<N> T(1._r), dist(0._r);
        int inter = 0;
        first_face = -1;
        while (true) {
            const int f = trace(o.val(), dv, tmin, FINF, th);
            if (f < 0) return Ds<N>();
            if (inter == 0) first_face = f;
            if (first_only) return Ds<N>();
            // At N >= 8 a null crossing takes the distance alone and the full hit record
            // is refined only at a light (the same t, recomputed): fewer dual operations
            // per crossing where each carries many tangents. Measured: dielectric scene
            // at N = 8 8% faster, the N = 3 benchmark scene 5% slower (kept there as is).
            HitAD<N> nh;
            Df<N> tmax;
            f3 gnv;
            if constexpr (N >= 8) {
                tmax = hitDistanceAD<N>(S, f, o, d, gnv);
            } else {
                refineHitAD<N>(S, f, o, d, nh);
                tmax = nh.t;
                gnv = nh.gn.val();
            }
            const DShape &sh = shapeOf(S, f);
            dist += tmax;
            if (med >= 0) T *= mediumTransmittanceAD<N>(S.media[med], o, d, tmin, tmax, r, scx);
            if (sh.light_id >= 0) {
                if constexpr (N >= 8) refineHitAD<N>(S, f, o, d, nh);
                f3 dd = nh.p.val() - scat;
                real G = nh.wi.z.v / dot(dd, dd);
                pdf_nee = S.light_pmf[sh.light_id] / (S.light_areas[sh.light_id] * G);
                if (far_face) { *far_face = f; *far_t = dist.v; }
                return emitterEvalAD<N>(S.emitters[sh.light_id], nh, -d) * T;
            }
            if (inter == max_inter || bsdfOf(S, sh).type != BSDF_NULL) return Ds<N>();
            if (isTransition(sh)) med = targetMedium(sh, gnv, dv);
            o = o + d * tmax;
            tmin = SHADOW_EPS;
            inter++;
        }
    }

    // scene.cpp:401-410 — light by power, triangle by area, point by (1-sqrt u, sqrt u * v)
    DI int sampleLight(const real *r4, f3 &pos, f3 &nrm) {
        int l = upperBound(S.light_cdf, S.nlights, r4[0]);
        const DEmitter &e = S.emitters[l];
        const DShape &sh = S.shapes[e.shape_id];
        int k = upperBound(S.area_cdf_pool + S.area_cdf_begin[l], sh.nf, r4[1]);
        const int *fi = S.F + 3 * (size_t)(sh.fbase + k);
        f3 v0 = ldg3(S.V + 3 * (size_t)fi[0]), v1 = ldg3(S.V + 3 * (size_t)fi[1]), v2 = ldg3(S.V + 3 * (size_t)fi[2]);
        real a = sqrt(r4[2]), b1 = 1.0_r - a, b2 = a * r4[3];
        f3 e1 = v1 - v0, e2 = v2 - v0;
        nrm = normalize(cross(e1, e2));
        pos = v0 + e1 * b1 + e2 * b2;
        return l;
    }

    // A next-event shadow ray, set up but not yet traced (scene.cpp:399-491), so that
    // the scattering events of a bounce share one transmittance walk.
    struct Shadow {
        f3 o, d, ln;
        real dist, pdf;
        int light, med, max_inter;
        bool on_surface, live;
    };
    // scene.cpp:399-416: the light sample of a point in a medium
    DI void neePointSetup(f3 ps, const real *r4, int med, int max_inter, Shadow &q) {
        q.live = false;
        if (S.nlights == 0) return;
        f3 lp, ln;
        const int l = sampleLight(r4, lp, ln);
        f3 dir = lp - ps;
        const real pdf = S.light_pmf[l] / S.light_areas[l];
        if (dot(dir, ln) < 0._r && pdf != 0._r) {
            const real dist = length(dir);
            q.o = ps; q.d = dir / dist; q.ln = ln; q.dist = dist; q.pdf = pdf;
            q.light = l; q.med = med; q.max_inter = max_inter; q.on_surface = false; q.live = true;
        }
    }
    // scene.cpp:456-480: the light sample of a surface point
    DI void neeSurfaceSetup(const Hit &its, const real *r4, int med, int max_inter, Shadow &q) {
        q.live = false;
        if (S.nlights == 0) return;
        f3 lp, ln;
        const int l = sampleLight(r4, lp, ln);
        f3 dir = lp - its.p;
        const real pdf = S.light_pmf[l] / S.light_areas[l];
        const DShape &lsh = S.shapes[S.emitters[l].shape_id];
        const int lbt = bsdfOf(S, lsh).type;
        const bool transm = lbt == BSDF_NULL || lbt == BSDF_ROUGHDIELECTRIC;
        if ((transm || dot(dir, its.sn) > 0._r) && dot(dir, ln) < 0._r && pdf != 0._r && max_inter >= 0) {
            const DShape &sh = shapeOf(S, its.face);
            if (isTransition(sh)) med = targetMedium(sh, its.gn, dir);
            const real dist = length(dir);
            q.o = its.p; q.d = dir / dist; q.ln = ln; q.dist = dist; q.pdf = pdf;
            q.light = l; q.med = med; q.max_inter = max_inter; q.on_surface = true; q.live = true;
        }
    }
    // scene.cpp:417-425, 481-487: the emitted radiance over the pdf, given the walk's
    // transmittance T (0: blocked); pdf becomes the solid-angle pdf.
    DI f3 neeFinish(const Shadow &q, real T, real &pdf) {
        if (!q.live || T == 0._r) return f3{0, 0, 0};
        pdf = q.pdf * (q.dist * q.dist / dot(q.ln, -q.d));
        return emitterEvalNorm(S.emitters[q.light], q.ln, -q.d) * (T / pdf);
    }

    // scene.cpp:399-426 (point in a medium; wo in world space)
    DI f3 neePoint(f3 ps, const real *r4, Rng &r, int med, int max_inter, f3 &wo, real &pdf) {
        if (S.nlights == 0) return f3{0, 0, 0};
        f3 lp, ln;
        int l = sampleLight(r4, lp, ln);
        f3 dir = lp - ps;
        pdf = S.light_pmf[l] / S.light_areas[l];
        if (dot(dir, ln) < 0._r && pdf != 0._r) {
            real dist = length(dir);
            dir = dir / dist;
            real T = evalTransmittance(ps, dir, false, med, dist, r, max_inter);
            if (T != 0._r) {
                pdf *= dist * dist / dot(ln, -dir);
                wo = dir;
                return emitterEvalNorm(S.emitters[l], ln, -dir) * (T / pdf);
            }
        }
        return f3{0, 0, 0};
    }
    // scene.cpp:456-491 (surface point; wo in the shading frame)
    DI f3 neeSurface(const Hit &its, const Frame &shf, const real *r4, Rng &r, int med, int max_inter, f3 &wo, real &pdf,
                     bool flag) {
        if (S.nlights == 0) return f3{0, 0, 0};
        f3 lp, ln;
        int l = sampleLight(r4, lp, ln);
        f3 dir = lp - its.p;
        pdf = S.light_pmf[l] / S.light_areas[l];
        const DShape &lsh = S.shapes[S.emitters[l].shape_id];
        int lbt = bsdfOf(S, lsh).type;
        bool transm = lbt == BSDF_NULL || lbt == BSDF_ROUGHDIELECTRIC;
        if ((transm || dot(dir, its.sn) > 0._r) && dot(dir, ln) < 0._r && pdf != 0._r) {
            const DShape &sh = shapeOf(S, its.face);
            if (isTransition(sh)) med = targetMedium(sh, its.gn, dir);
            real dist = length(dir);
            dir = dir / dist;
            wo = shf.toLocal(dir);
            if (flag && signum(its.wi.z) != signum(wo.z)) --max_inter;
            if (max_inter >= 0) {
                real T = evalTransmittance(its.p, dir, true, med, dist, r, max_inter);
                if (T != 0._r) {
                    pdf *= dist * dist / dot(ln, -dir);
                    return emitterEvalNorm(S.emitters[l], ln, -dir) * (T / pdf);
                }
            }
        }
        return f3{0, 0, 0};
    }
    // A next-event shadow ray with tangents, set up but not yet traced
    // (scene.cpp:428-454, 493-524): the origin carries its tangents, the direction to
    // the light sample none.
    struct ShadowAD {
        Dv<N> o;
        f3 d;
        real dist;
        int light, med, max_inter;
        bool on_surface, live;
    };
    // scene.cpp:428-440: the light sample of a point in a medium
    DI void neePointADSetup(const Dv<N> &ps, const real *r4, int med, int max_inter, ShadowAD &q) {
        q.live = false;
        if (S.nlights == 0) return;
        f3 lp, ln;
        const int l = sampleLight(r4, lp, ln);
        f3 dir = lp - ps.val();
        if (dot(dir, ln) < 0._r) {
            const real dist = length(dir);
            q.o = ps; q.d = dir / dist; q.dist = dist;
            q.light = l; q.med = med; q.max_inter = max_inter; q.on_surface = false; q.live = true;
        }
    }
    // scene.cpp:493-511: the light sample of a surface point
    DI void neeSurfaceADSetup(const HitAD<N> &its, const real *r4, int med, int max_inter, ShadowAD &q) {
        q.live = false;
        if (S.nlights == 0) return;
        f3 lp, ln;
        const int l = sampleLight(r4, lp, ln);
        f3 dir = lp - its.p.val();
        const DShape &lsh = S.shapes[S.emitters[l].shape_id];
        const int lbt = bsdfOf(S, lsh).type;
        const bool transm = lbt == BSDF_NULL || lbt == BSDF_ROUGHDIELECTRIC;
        if ((transm || dot(dir, its.sn.val()) > 0._r) && dot(dir, ln) < 0._r) {
            const DShape &sh = shapeOf(S, its.face);
            if (isTransition(sh)) med = targetMedium(sh, its.gn.val(), dir);
            const real dist = length(dir);
            q.o = its.p; q.d = dir / dist; q.dist = dist;
            q.light = l; q.med = med; q.max_inter = max_inter; q.on_surface = true; q.live = true;
        }
    }
    // scene.cpp:441-453, 512-523: walk the shadow ray; the sampled light's radiance over
    // the solid-angle pdf if the walk ends on it at the sampled distance, else zero.
    DI Ds<N> neeWalkAD(const ShadowAD &q, Rng &r, real &pdf, const ScatterCtx *scx = nullptr) {
        int first = -1, ff = -1;
        real ft = 0._r;
        Ds<N> ret = lookForEmitterAD(q.o, Dv<N>(q.d), q.on_surface, r, q.med, q.max_inter, first, pdf, &ff, &ft, scx);
        if (!ret.isZero(EPS) && pdf > EPS && ff >= 0 && __ldg(S.face_shape + ff) == S.emitters[q.light].shape_id &&
            fabs(ft - q.dist) < SHADOW_EPS)
            return ret * (1.0_r / pdf);
        return Ds<N>();
    }

    // scene.cpp:346-397 — sampleEmitterDirectAD: no media; the shadow ray must first hit
    // the sampled light within ShadowEpsilon of the sampled distance; the value's
    // tangents come from the emitter only (pdf and geometry term are primal)
    DI Ds<N> neeDirectAD(const HitAD<N> &its, const FrameAD<N> &shf, const real *r4, Dv<N> &wo, real &pdf) {
        if (S.nlights == 0) return Ds<N>();
        f3 lp, ln;
        int l = sampleLight(r4, lp, ln);
        f3 dir = lp - its.p.val();
        pdf = S.light_pmf[l] / S.light_areas[l];
        if (dot(dir, its.sn.val()) > 0._r && dot(dir, ln) < 0._r && pdf != 0._r) {
            real dist = length(dir), th;
            dir = dir / dist;
            int f = trace(its.p.val(), dir, SHADOW_EPS, FINF, th);
            if (f >= 0 && __ldg(S.face_shape + f) == S.emitters[l].shape_id) {
                HitAD<N> nh;
                refineHitAD<N>(S, f, its.p, Dv<N>(dir), nh);
                if (fabs(nh.t.v - dist) < SHADOW_EPS) {
                    pdf *= dist * dist / dot(ln, -dir);
                    wo = shf.toLocal(Dv<N>(dir));
                    return emitterEvalNormAD<N>(S.emitters[l], nh.sn, Dv<N>(-dir)) * (1.0_r / pdf);
                }
            }
        }
        return Ds<N>();
    }

    // volpathAD.cpp:50-142 — primal path tracer used by every edge / boundary sub-path
    // The primal estimator as a resumable state machine: liStart traces the first
    // segment, each liStep runs one iteration of the bounce loop of
    // volpathAD.cpp:50-142 (same random-number order), and returns false once the
    // path has ended; the estimate is then in p.ret.
    DI void liStartAt(LiPath &p, f3 ro, f3 rd, int max_depth, int med, bool incEmission, const Hit &first) {
        p.ro = ro;
        p.rd = rd;
        p.thr = f3{1, 1, 1};
        p.ret = f3{0, 0, 0};
        p.depth = 0;
        p.max_depth = max_depth;
        p.med = med;
        p.incEmission = incEmission;
        p.face = first.face;
        p.walk = 0;
        p.walk_max_inter = 0;
        p.cpdf = 0._r;
    }
    // The hit record of the path's current ray (shape.cpp:166-208 on the recorded face).
    DI Hit hitOf(const LiPath &p) const {
        Hit h;
        if (p.face >= 0) refineHit(S, p.face, p.ro, p.rd, h);
        return h;
    }
    // The first hit (from ShadowEpsilon) is found by the first step's walk.
    DI void liStart(LiPath &p, f3 ro, f3 rd, int max_depth, int med, bool incEmission) {
        Hit none;
        liStartAt(p, ro, rd, max_depth, med, incEmission, none);
        p.walk = 1;
    }

    // The volume and surface branches share one shadow walk and one continuation walk,
    // so the lanes of a wave on either kind of event traverse the BVH together. The
    // continuation walk of a bounce runs at the start of the next step, where it also
    // finds the first hit of a new path's ray: paths just taken from the queue trace
    // together with the others.
    DI bool liStep(LiPath &p, Rng &r) {
        if (p.walk) {
            Hit nx;
            real pdf_nee = 0._r;
            const f3 att = lookForEmitter(p.ro, p.rd, p.walk != 2, r, p.med, p.walk_max_inter, nx, pdf_nee, p.walk == 1);
            p.face = nx.face;
            if (p.walk == 2 ? !allZero(att, EPS) : (p.walk == 3 && !allZero(att, 1e-12_r)))
                p.ret = p.ret + p.thr * att * misPower(p.cpdf, pdf_nee);
            p.walk = 0;
        }
        if (p.depth > p.max_depth) return false;
        const int max_inter = p.max_depth - p.depth - 1;
        const Hit its = hitOf(p);
        real r1a = r.next(), r1b = r.next();
        real r2[4] = {r.next(), r.next(), r.next(), r.next()};
        real r3[3] = {r.next(), r.next(), r.next()};
        bool inside = false;
        if (p.med >= 0) {
            f3 ps;
            inside = sampleDistance(S.media[p.med], p.ro, p.rd, its.t, r1a, r1b, r, ps, p.thr);
            if (inside) p.ro = ps;
        }
        Shadow q;
        q.live = false;
        int bt = BSDF_NULL;
        if (inside) {
            if (p.depth >= p.max_depth) return false;
            if (allZero(p.thr, EPS)) return false;
            neePointSetup(p.ro, r2, p.med, max_inter, q);
        } else {
            if (its.face < 0) return false;
            const DShape &sh = shapeOf(S, its.face);
            if (sh.light_id >= 0 && p.incEmission)
                p.ret = p.ret + p.thr * emitterEval(S.emitters[sh.light_id], its.gn, its.sn, -p.rd);
            if (p.depth >= p.max_depth) return false;
            bt = bsdfOf(S, sh).type;
            if (bt != BSDF_NULL) neeSurfaceSetup(its, r2, p.med, max_inter, q);
        }
        real pdf_nee = 0._r;
        const real Tq = q.live ? evalTransmittance(q.o, q.d, q.on_surface, q.med, q.dist, r, q.max_inter) : 0._r;
        const f3 value = neeFinish(q, Tq, pdf_nee);
        f3 wo;
        bool first_only = false;
        real bpdf = 0._r;
        f3 bw{1, 1, 1};
        if (inside) {
            const DPhase &ph = S.phases[S.media[p.med].phase_id];
            if (!allZero(value, EPS)) {
                real pv = phaseEval(ph, -p.rd, q.d);
                if (pv > EPS) p.ret = p.ret + p.thr * value * (misPower(pdf_nee, pv) * pv);
            }
            wo = phaseSample(ph, -p.rd, r3[0], r3[1]);
            p.rd = wo;
        } else {
            const DShape &sh = shapeOf(S, its.face);
            const DBsdf &b = bsdfOf(S, sh);
            const Frame shf(its.sn);
            if (!allZero(value, EPS)) {
                const f3 wl = shf.toLocal(q.d);
                p.ret = p.ret + p.thr * value * bsdfEval(b, its.wi, wl) * misPower(pdf_nee, bsdfPdf(b, its.wi, wl));
            }
            bw = bsdfSample(b, its.wi, r3[0], r3[1], r3[2], wo, bpdf);
            if (allZero(bw, EPS)) return false;
            wo = shf.toWorld(wo);
            p.ro = its.p;
            p.rd = wo;
            if (isTransition(sh)) p.med = targetMedium(sh, its.gn, wo);
            first_only = bt == BSDF_NULL;
        }
        p.walk = inside ? 2 : (first_only ? 1 : 3);
        p.walk_max_inter = max_inter;
        if (inside) {
            p.cpdf = phaseEval(S.phases[S.media[p.med].phase_id], -wo, wo);
            p.incEmission = false;
        } else if (!first_only) {
            p.thr = p.thr * bw;
            p.cpdf = bpdf;
            p.incEmission = false;
        }
        p.depth++;
        return true;   // the walk runs in the next step, also when this was the last bounce
    }

    // volpath_simple.cpp:10-86 — VolPathTracerSimple's bounce: next-event estimation
    // without MIS; emission found by a traced ray counts only before the first
    // non-null scattering event. Same random-number order as liStep.
    DI bool liSimpleStep(LiPath &p, Rng &r) {
        if (p.depth > p.max_depth) return false;
        const int max_inter = p.max_depth - p.depth - 1;
        const Hit its = hitOf(p);
        real r1a = r.next(), r1b = r.next();
        real r2[4] = {r.next(), r.next(), r.next(), r.next()};
        real r3[3] = {r.next(), r.next(), r.next()};
        bool inside = false;
        if (p.med >= 0) {
            f3 ps;
            inside = sampleDistance(S.media[p.med], p.ro, p.rd, its.t, r1a, r1b, r, ps, p.thr);
            if (inside) p.ro = ps;
        }
        if (inside) {
            if (p.depth >= p.max_depth) return false;
            if (allZero(p.thr, 0._r)) return false;
            const DPhase &ph = S.phases[S.media[p.med].phase_id];
            f3 wo;
            real pdf_nee;
            f3 value = neePoint(p.ro, r2, r, p.med, max_inter, wo, pdf_nee);
            if (!allZero(value, 0._r)) {
                real pv = phaseEval(ph, -p.rd, wo);
                if (pv != 0._r) p.ret = p.ret + p.thr * value * pv;
            }
            p.rd = phaseSample(ph, -p.rd, r3[0], r3[1]);   // weight eval/pdf = 1
            Hit nx;
            rayIntersect(p.ro, p.rd, false, nx);
            p.face = nx.face;
            p.incEmission = false;
        } else {
            if (its.face < 0) return false;
            const DShape &sh = shapeOf(S, its.face);
            if (sh.light_id >= 0 && p.incEmission)
                p.ret = p.ret + p.thr * emitterEval(S.emitters[sh.light_id], its.gn, its.sn, -p.rd);
            if (p.depth >= p.max_depth) return false;
            const DBsdf &b = bsdfOf(S, sh);
            Frame shf(its.sn);
            f3 wo;
            if (b.type != BSDF_NULL) {
                real pdf_nee;
                f3 value = neeSurface(its, shf, r2, r, p.med, max_inter, wo, pdf_nee, false);
                if (!allZero(value, 0._r)) p.ret = p.ret + p.thr * value * bsdfEval(b, its.wi, wo);
                p.incEmission = false;
            }
            real bpdf = 0._r;
            f3 bw = bsdfSample(b, its.wi, r3[0], r3[1], r3[2], wo, bpdf);
            if (allZero(bw, 0._r)) return false;
            p.thr = p.thr * bw;
            wo = shf.toWorld(wo);
            p.ro = its.p;
            p.rd = wo;
            if (isTransition(sh)) p.med = targetMedium(sh, its.gn, wo);
            Hit nx;
            rayIntersect(p.ro, p.rd, true, nx);
            p.face = nx.face;
        }
        p.depth++;
        return p.depth <= p.max_depth;
    }

    // scene.cpp:302-344 — sampleEmitterDirect: no media; the light point is visible
    // when no surface lies within [ShadowEpsilon, (1 - ShadowEpsilon) * dist]
    DI f3 neeDirect(const Hit &its, const Frame &shf, const real *r4, f3 &wo, real &pdf) {
        if (S.nlights == 0) return f3{0, 0, 0};
        f3 lp, ln;
        int l = sampleLight(r4, lp, ln);
        f3 dir = lp - its.p;
        pdf = S.light_pmf[l] / S.light_areas[l];
        if (dot(dir, its.sn) > 0._r && dot(dir, ln) < 0._r && pdf != 0._r) {
            real dist = length(dir), th;
            dir = dir / dist;
            if (trace(its.p, dir, SHADOW_EPS, (1.0_r - SHADOW_EPS) * dist, th) < 0) {
                pdf *= dist * dist / dot(ln, -dir);
                wo = shf.toLocal(dir);
                return emitterEvalNorm(S.emitters[l], ln, -dir) * (1.0_r / pdf);
            }
        }
        return f3{0, 0, 0};
    }

    // path.cpp:11-61 — one bounce of PathTracer::radiance: surfaces only (media are
    // ignored), emission seen directly at the first hit, NEE and BSDF sampling with the
    // power heuristic. p.face is the current ray's hit; the caller traces the camera ray
    // with tnear = ShadowEpsilon, as path.cpp:14 does.
    DI bool pathStep(LiPath &p, Rng &r) {
        if (p.depth > p.max_depth || p.face < 0) return false;
        const Hit its = hitOf(p);
        const DShape &sh = shapeOf(S, its.face);
        if (sh.light_id >= 0 && p.depth == 0)
            p.ret = p.ret + p.thr * emitterEval(S.emitters[sh.light_id], its.gn, its.sn, -p.rd);
        if (p.depth >= p.max_depth) return false;
        const DBsdf &b = bsdfOf(S, sh);
        Frame shf(its.sn);
        real r4[4] = {r.next(), r.next(), r.next(), r.next()};
        f3 wo;
        real pdf_nee;
        f3 value = neeDirect(its, shf, r4, wo, pdf_nee);
        if (!allZero(value, 0._r))
            p.ret = p.ret + p.thr * value * bsdfEval(b, its.wi, wo) * misPower(pdf_nee, bsdfPdf(b, its.wi, wo));
        real r3[3] = {r.next(), r.next(), r.next()};
        real bpdf = 0._r;
        f3 bw = bsdfSample(b, its.wi, r3[0], r3[1], r3[2], wo, bpdf);
        if (allZero(bw, 0._r)) return false;
        p.ro = its.p;
        p.rd = shf.toWorld(wo);
        Hit nx;
        if (!rayIntersect(p.ro, p.rd, true, nx)) { p.face = -1; return false; }
        p.face = nx.face;
        p.thr = p.thr * bw;
        const DShape &hs = shapeOf(S, nx.face);
        if (hs.light_id >= 0) {
            f3 lc = emitterEval(S.emitters[hs.light_id], nx.gn, nx.sn, -p.rd);
            if (!allZero(lc, 0._r)) {
                f3 dd = nx.p - p.ro;
                real G = nx.wi.z / dot(dd, dd);
                real pn = S.light_pmf[hs.light_id] / (S.light_areas[hs.light_id] * G);
                p.ret = p.ret + p.thr * lc * misPower(bpdf, pn);
            }
        }
        p.depth++;
        return true;
    }

    DI f3 Li(Rng &r, f3 ro, f3 rd, int max_depth, int med, bool incEmission) {   // volpathAD.cpp:50-142
        LiPath p;
        liStart(p, ro, rd, max_depth, med, incEmission);
        while (liStep(p, r)) {}
        return p.ret;
    }

    // volpathAD.cpp:338-441 — one sample of Lins at a boundary hit (ALTERNATIVE_LINS_BOUNDARY_SAMPLING).
    // Returns the terms computed at the hit; every branch ends in at most one primal
    // sub-path, which is returned as `cont` (started, weight `cw`) for the caller to
    // run: the sample is the returned value + cw * cont.ret. `has_cont` = false: none.
    DI f3 linsStart(Rng &r, const Hit &its0, f3 rdv, int med, int max_depth, bool null_bsdf, LiPath &cont, f3 &cw,
                    bool &has_cont) {
        has_cont = false;
        const DShape &shp = shapeOf(S, its0.face);
        const DBsdf &b = bsdfOf(S, shp);
        const DPhase &ph = S.phases[S.media[med].phase_id];
        Frame shf(its0.sn);
        f3 n = dot(its0.gn, rdv) < 0._r ? -its0.gn : its0.gn;
        f3 Lins{0, 0, 0}, wo;
        real pdf_nee, pv = 0._r;
        real r4[4] = {r.next(), r.next(), r.next(), r.next()};
        if (null_bsdf) {
            f3 att = neeSurface(its0, shf, r4, r, med, max_depth - 1, wo, pdf_nee, true);
            if (!allZero(att, EPS)) {
                wo = shf.toWorld(wo);
                pv = phaseEval(ph, -rdv, wo);
                if (pv > EPS) Lins = Lins + att * (pv * misPower(pdf_nee, pv));
            }
            real u1 = r.next(), u2 = r.next();
            wo = phaseSample(ph, -rdv, u1, u2);
            pv = 1._r;
            int md1 = max_depth - 1;
            if (dot(wo, n) > EPS) {
                if (shp.light_id >= 0) Lins = Lins + emitterEval(S.emitters[shp.light_id], its0.gn, its0.sn, -wo) * pv;
                --md1;
            }
            if (md1 > 0) {
                int med1 = isTransition(shp) ? targetMedium(shp, its0.gn, wo) : med;
                Hit h1;
                f3 att = lookForEmitter(its0.p, wo, true, r, med1, md1, h1, pdf_nee);
                if (!allZero(att, EPS)) Lins = Lins + att * (pv * misPower(phaseEval(ph, -rdv, wo), pdf_nee));
                liStartAt(cont, its0.p, wo, md1, med1, false, h1);   // h1 = first hit of this ray
                cw = f3{pv, pv, pv};
                has_cont = true;
            }
        } else {
            f3 att = neeSurface(its0, shf, r4, r, med, max_depth - 1, wo, pdf_nee, false);
            if (!allZero(att, EPS)) {
                wo = shf.toWorld(wo);
                if (dot(wo, n) < -EPS) {
                    pv = phaseEval(ph, -rdv, wo);
                    if (pv > EPS) Lins = Lins + att * (pv * misPower(pdf_nee, pv));
                }
            }
            real u1 = r.next(), u2 = r.next();
            wo = phaseSample(ph, -rdv, u1, u2);
            pv = 1._r;
            real wdn = dot(wo, n);
            if (wdn < -EPS) {
                Hit h1;
                f3 att2 = lookForEmitter(its0.p, wo, true, r, med, max_depth - 1, h1, pdf_nee);
                if (!allZero(att2, EPS)) Lins = Lins + att2 * (pv * misPower(phaseEval(ph, -rdv, wo), pdf_nee));
                liStartAt(cont, its0.p, wo, max_depth - 1, med, false, h1);
                cw = f3{pv, pv, pv};
                has_cont = true;
            } else if (wdn > EPS) {
                if (shp.light_id >= 0) Lins = Lins + emitterEval(S.emitters[shp.light_id], its0.gn, its0.sn, -wo) * pv;
                if (max_depth > 1) {
                    Hit itsb = its0;
                    itsb.wi = shf.toLocal(-wo);
                    real r4b[4] = {r.next(), r.next(), r.next(), r.next()};
                    f3 wl;
                    f3 att3 = neeSurface(itsb, shf, r4b, r, med, max_depth - 2, wl, pdf_nee, false);
                    if (!allZero(att3, EPS)) Lins = Lins + bsdfEval(b, itsb.wi, wl) * att3 * (pv * misPower(pdf_nee, bsdfPdf(b, itsb.wi, wl)));
                    real q0 = r.next(), q1 = r.next(), q2 = r.next(), bpdf = 0._r;
                    f3 bv = bsdfSample(b, itsb.wi, q0, q1, q2, wl, bpdf);
                    if (!allZero(bv, EPS)) {
                        f3 thr1 = bv * pv;
                        f3 wow = shf.toWorld(wl);
                        int med1 = isTransition(shp) ? targetMedium(shp, itsb.gn, wow) : med;
                        Hit itsc;
                        f3 att4 = lookForEmitter(itsb.p, wow, true, r, med1, max_depth - 2, itsc, pdf_nee);
                        if (!allZero(att4, 1e-12_r)) Lins = Lins + thr1 * att4 * misPower(bpdf, pdf_nee);
                        liStartAt(cont, itsb.p, wow, max_depth - 2, med1, false, itsc);
                        cw = thr1;
                        has_cont = true;
                    }
                }
            }
        }
        return Lins;
    }

    DI f3 LinsOnce(Rng &r, const Hit &its0, f3 rdv, int med, int max_depth, bool null_bsdf) {
        LiPath c;
        f3 cw{0, 0, 0};
        bool has;
        f3 L = linsStart(r, its0, rdv, med, max_depth, null_bsdf, c, cw, has);
        if (has) {
            while (liStep(c, r)) {}
            L = L + cw * c.ret;
        }
        return L;
    }
};

}  // namespace DT_PNS
}  // namespace dt
