// TEST INFRASTRUCTURE ONLY — CPU oracle (see oracle_math.h header).
//
// Double-precision, OpenMP restatement of dtrt's forward-mode-AD volumetric path
// tracer (VolPathTracerAD) and everything it calls, with the reference's PCG32
// streams and draw order. Recursive like the reference. Exposed through a C API
// whose object constructors mirror src/dtrt.cpp:35-115 (prefix "oracle_").
#pragma once
#include "oracle_scene.h"
#include <atomic>
#include <cstdio>
#include <map>
#include <omp.h>

namespace oref {

// ------------------------------------------------------------ diagnostics
struct Counters {
    std::atomic<long long> neg_pixels{0}, nonfinite_der{0}, dt_cap{0}, bad_edge_weight{0}, nonortho{0},
        med_inconsistent{0};
    void reset() { neg_pixels = 0; nonfinite_der = 0; dt_cap = 0; bad_edge_weight = 0; nonortho = 0; med_inconsistent = 0; }
};
static Counters g_counters;

// ============================================================== geometry
// shape.cpp:109-126 — face normals with tangents; a near-degenerate face keeps
// its unnormalised cross product, as in the reference.
void OShape::computeFaceNormals() {
    fn.assign(nf, V3());
    fnd.assign((size_t)nder * nf, V3());
    for (int k = 0; k < nf; ++k) {
        const auto &ix = f[k];
        V3 e1 = v[ix[1]] - v[ix[0]], e2 = v[ix[2]] - v[ix[0]];
        V3 c = cross(e1, e2);
        double len = norm(c);
        std::vector<V3> dc(nder);
        for (int j = 0; j < nder; ++j) {
            V3 de1 = vd[(size_t)j * nv + ix[1]] - vd[(size_t)j * nv + ix[0]];
            V3 de2 = vd[(size_t)j * nv + ix[2]] - vd[(size_t)j * nv + ix[0]];
            dc[j] = cross(de1, e2) + cross(e1, de2);
        }
        if (len < EPS) {
            fn[k] = c;
            for (int j = 0; j < nder; ++j) fnd[(size_t)j * nf + k] = dc[j];
        } else {
            V3 nrm = c / len;
            fn[k] = nrm;
            for (int j = 0; j < nder; ++j)  // d(c/|c|) = (dc - n (n.dc)) / |c|
                fnd[(size_t)j * nf + k] = (dc[j] - nrm * dot(nrm, dc[j])) / len;
        }
    }
}

// shape.cpp:254-300 — edges keyed by sorted vertex pair, visited in key order.
void OShape::constructEdges() {
    std::map<std::pair<int, int>, std::vector<int>> em;
    for (int k = 0; k < nf; ++k)
        for (int e = 0; e < 3; ++e) {
            int a = f[k][e], b = f[k][(e + 1) % 3];
            em[a < b ? std::make_pair(a, b) : std::make_pair(b, a)].push_back(k);
        }
    edges.clear();
    for (auto const &it : em) {
        double length = norm(v[it.first.first] - v[it.first.second]);
        if (it.second.size() > 2) { ok = false; err = "Every edge can be shared by at most 2 faces"; return; }
        if (it.second.size() == 2) {
            if (it.second[0] == it.second[1]) { ok = false; err = "Duplicated faces"; return; }
            double val = dot(fn[it.second[0]], fn[it.second[1]]);
            if (val < -1.0 + EPS) { ok = false; err = "Inconsistent normal orientation"; return; }
            if (val < 1.0 - EPS) edges.push_back(Edge{it.first.first, it.first.second, it.second[0], it.second[1], length});
        } else {
            edges.push_back(Edge{it.first.first, it.first.second, it.second[0], -1, length});
        }
    }
    edge_distrb.clear();
    for (const Edge &e : edges) edge_distrb.append(e.length);
    edge_distrb.normalize();
}

// shape.cpp:302-335
int OShape::isSilhouette(const Edge &e, const V3 &p) const {
    if (e.f0 == -1 || e.f1 == -1) return 2;
    const V3 &v0 = v[e.v0], &v1 = v[e.v1];
    bool ff0 = false, ff1 = false;
    for (int i = 0; i < 3; ++i) {
        int k = f[e.f0][i];
        if (k != e.v0 && k != e.v1) {
            V3 n0 = normalized(cross(v0 - v[k], v1 - v[k]));
            ff0 = dot(n0, p - v[k]) > 0.0;
            break;
        }
    }
    for (int i = 0; i < 3; ++i) {
        int k = f[e.f1][i];
        if (k != e.v0 && k != e.v1) {
            V3 n1 = normalized(cross(v1 - v[k], v0 - v[k]));
            ff1 = dot(n1, p - v[k]) > 0.0;
            break;
        }
    }
    if ((ff0 && !ff1) || (!ff0 && ff1)) return 2;
    return hasNormals() ? 0 : 1;
}

// ---------------------------------------------------------- BVH (Embree stand-in)
static void bbox_of(const OScene &s, int pi, V3 &lo, V3 &hi) {
    const OShape &sh = *s.shapes[s.prims[pi].first];
    const auto &ix = sh.f[s.prims[pi].second];
    lo = hi = sh.v[ix[0]];
    for (int k = 1; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], sh.v[ix[k]][a]);
            hi[a] = std::max(hi[a], sh.v[ix[k]][a]);
        }
}

void OScene::buildBvh() {
    prims.clear();
    for (int s = 0; s < (int)shapes.size(); ++s)
        for (int k = 0; k < shapes[s]->nf; ++k) prims.push_back({s, k});
    nodes.clear();
    if (prims.empty()) return;
    std::vector<V3> blo(prims.size()), bhi(prims.size()), cen(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) {
        bbox_of(*this, (int)i, blo[i], bhi[i]);
        cen[i] = (blo[i] + bhi[i]) * 0.5;
    }
    std::vector<int> order(prims.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    // binned SAH (16 bins per axis, as Embree's builders price splits) with leaves of
    // <= 4 triangles; a median split on the widest centroid axis where SAH finds no
    // split or the depth reaches 64, which bounds the traversal stack
    struct Task { int node, first, count, depth; };
    nodes.push_back(BvhNode());
    std::vector<Task> stack{{0, 0, (int)prims.size(), 0}};
    auto area = [](const V3 &l, const V3 &h) {
        V3 e(std::max(0.0, h.x - l.x), std::max(0.0, h.y - l.y), std::max(0.0, h.z - l.z));
        return e.x * e.y + e.y * e.z + e.z * e.x;
    };
    while (!stack.empty()) {
        Task tk = stack.back();
        stack.pop_back();
        BvhNode &nd = nodes[tk.node];
        V3 lo(INF, INF, INF), hi(-INF, -INF, -INF), clo = lo, chi = hi;
        for (int i = tk.first; i < tk.first + tk.count; ++i) {
            int p = order[i];
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], blo[p][a]); hi[a] = std::max(hi[a], bhi[p][a]);
                clo[a] = std::min(clo[a], cen[p][a]); chi[a] = std::max(chi[a], cen[p][a]);
            }
        }
        nd.lo = lo; nd.hi = hi;
        int axis = 0;
        V3 ext = chi - clo;
        if (ext.y > ext[axis]) axis = 1;
        if (ext.z > ext[axis]) axis = 2;
        if (tk.count <= 4 || ext[axis] <= 0) { nd.first = tk.first; nd.count = tk.count; continue; }
        const int NB = 16;
        int bestAxis = -1, bestBin = -1;
        double bestCost = INF;
        if (tk.depth < 64)
            for (int a = 0; a < 3; ++a) {
                if (!(ext[a] > 0)) continue;
                int cnt[NB] = {0};
                V3 bl[NB], bh[NB];
                for (int b = 0; b < NB; ++b) { bl[b] = V3(INF, INF, INF); bh[b] = V3(-INF, -INF, -INF); }
                for (int i = tk.first; i < tk.first + tk.count; ++i) {
                    int p = order[i];
                    int b = std::min(NB - 1, (int)((cen[p][a] - clo[a]) / ext[a] * NB));
                    cnt[b]++;
                    for (int c = 0; c < 3; ++c) { bl[b][c] = std::min(bl[b][c], blo[p][c]); bh[b][c] = std::max(bh[b][c], bhi[p][c]); }
                }
                double rA[NB];
                int rC[NB], rc = 0;
                V3 rl(INF, INF, INF), rh(-INF, -INF, -INF);
                for (int b = NB - 1; b > 0; --b) {
                    rc += cnt[b];
                    for (int c = 0; c < 3; ++c) { rl[c] = std::min(rl[c], bl[b][c]); rh[c] = std::max(rh[c], bh[b][c]); }
                    rA[b] = area(rl, rh);
                    rC[b] = rc;
                }
                int lc = 0;
                V3 ll(INF, INF, INF), lh(-INF, -INF, -INF);
                for (int b = 0; b < NB - 1; ++b) {
                    lc += cnt[b];
                    for (int c = 0; c < 3; ++c) { ll[c] = std::min(ll[c], bl[b][c]); lh[c] = std::max(lh[c], bh[b][c]); }
                    if (lc == 0 || rC[b + 1] == 0) continue;
                    double cost = area(ll, lh) * lc + rA[b + 1] * rC[b + 1];
                    if (cost < bestCost) { bestCost = cost; bestAxis = a; bestBin = b; }
                }
            }
        int mid;
        if (bestAxis >= 0) {
            const int a = bestAxis;
            mid = (int)(std::partition(order.begin() + tk.first, order.begin() + tk.first + tk.count, [&](int p) {
                       return std::min(NB - 1, (int)((cen[p][a] - clo[a]) / ext[a] * NB)) <= bestBin;
                   }) - order.begin());
        } else {
            mid = tk.first + tk.count / 2;
            std::nth_element(order.begin() + tk.first, order.begin() + mid, order.begin() + tk.first + tk.count,
                             [&](int a, int b) { return cen[a][axis] < cen[b][axis]; });
        }
        int l = (int)nodes.size();
        nodes.push_back(BvhNode());
        nodes.push_back(BvhNode());
        nodes[tk.node].left = l;
        nodes[tk.node].right = l + 1;
        stack.push_back({l, tk.first, mid - tk.first, tk.depth + 1});
        stack.push_back({l + 1, mid, tk.first + tk.count - mid, tk.depth + 1});
    }
    std::vector<std::pair<int, int>> np(prims.size());
    for (size_t i = 0; i < order.size(); ++i) np[i] = prims[order[i]];
    prims.swap(np);
}

// Triangle test used to choose the closest primitive: double Moller-Trumbore with
// barycentric bounds (Embree's role; its float/watertight details are not restated).
static inline bool triHit(const V3 &v0, const V3 &v1, const V3 &v2, const V3 &o, const V3 &d, double tnear,
                          double tfar, double &t) {
    V3 e1 = v1 - v0, e2 = v2 - v0, pv = cross(d, e2);
    double det = dot(e1, pv);
    if (det == 0.0) return false;
    double inv = 1.0 / det;
    V3 s = o - v0;
    double u = dot(s, pv) * inv;
    if (u < 0.0 || u > 1.0) return false;
    V3 q = cross(s, e1);
    double w = dot(d, q) * inv;
    if (w < 0.0 || u + w > 1.0) return false;
    t = dot(e2, q) * inv;
    return t >= tnear && t <= tfar;
}

static inline bool boxHit(const V3 &lo, const V3 &hi, const V3 &o, const V3 &inv, double tnear, double tfar) {
    double t0 = tnear, t1 = tfar;
    for (int a = 0; a < 3; ++a) {
        double ta = (lo[a] - o[a]) * inv[a], tb = (hi[a] - o[a]) * inv[a];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
        if (t0 > t1) return false;
    }
    return true;
}

bool OScene::intersect(const V3 &o, const V3 &d, double tnear, double tfar, int &shape, int &tri) const {
    shape = -1;
    tri = -1;
    double best = tfar;
    if (brute_force) {
        for (int s = 0; s < (int)shapes.size(); ++s)
            for (int k = 0; k < shapes[s]->nf; ++k) {
                const auto &ix = shapes[s]->f[k];
                double t;
                if (triHit(shapes[s]->v[ix[0]], shapes[s]->v[ix[1]], shapes[s]->v[ix[2]], o, d, tnear, best, t)) {
                    best = t; shape = s; tri = k;
                }
            }
        return shape >= 0;
    }
    if (nodes.empty()) return false;
    V3 inv(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    int stack[160];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BvhNode &nd = nodes[stack[--sp]];
        if (!boxHit(nd.lo, nd.hi, o, inv, tnear, best)) continue;
        if (nd.left < 0) {
            for (int i = nd.first; i < nd.first + nd.count; ++i) {
                const OShape &sh = *shapes[prims[i].first];
                const auto &ix = sh.f[prims[i].second];
                double t;
                if (triHit(sh.v[ix[0]], sh.v[ix[1]], sh.v[ix[2]], o, d, tnear, best, t)) {
                    best = t; shape = prims[i].first; tri = prims[i].second;
                }
            }
        } else {   // the child whose centre lies nearer along the ray is visited first
            const BvhNode &L = nodes[nd.left], &R = nodes[nd.right];
            const double cl = dot((L.lo + L.hi) * 0.5 - o, d), cr = dot((R.lo + R.hi) * 0.5 - o, d);
            if (cl <= cr) { stack[sp++] = nd.right; stack[sp++] = nd.left; }
            else { stack[sp++] = nd.left; stack[sp++] = nd.right; }
        }
    }
    return shape >= 0;
}

// scene.cpp:45-98 — light pmf proportional to area * luminance * pi; per-light
// triangle CDF proportional to area.
void OScene::buildLights() {
    int nl = (int)emitters.size();
    light_pmf.assign(nl, 0);
    light_cdf.assign(nl, 0);
    light_areas.assign(nl, 0);
    area_cdf_begin.assign(nl, 0);
    int total = 0;
    for (int l = 0; l < nl; ++l) { area_cdf_begin[l] = total; total += shapes[emitters[l]->shape_id]->nf; }
    area_cdf_pool.assign(total, 0);
    double tot_imp = 0;
    for (int l = 0; l < nl; ++l) {
        const OShape &sh = *shapes[emitters[l]->shape_id];
        double s = 0;
        for (int k = 0; k < sh.nf; ++k) { s += sh.area(k); area_cdf_pool[area_cdf_begin[l] + k] = s; }
        light_areas[l] = s;
        for (int k = 0; k < sh.nf; ++k) area_cdf_pool[area_cdf_begin[l] + k] /= s;
        light_pmf[l] = s * luminance(emitters[l]->intensity()) * PI;
        tot_imp += light_pmf[l];
    }
    if (nl > 0 && !(tot_imp > 0)) { ok = false; err = "total light importance must be positive"; return; }
    for (int l = 0; l < nl; ++l) light_pmf[l] /= tot_imp;
    if (nl > 0) {
        light_cdf[0] = light_pmf[0];
        for (int l = 1; l < nl; ++l) light_cdf[l] = light_cdf[l - 1] + light_pmf[l];
    }
}

// line_clip.h:37-95 (Cohen-Sutherland on AD screen points)
struct ADPoint2 { V2 v; V2 d[NDER_MAX]; };
static int outCode(const V2 &p, double xmin, double xmax, double ymin, double ymax) {
    int c = 0;
    if (p.x < xmin - EPS) c |= 1; else if (p.x > xmax + EPS) c |= 2;
    if (p.y < ymin - EPS) c |= 4; else if (p.y > ymax + EPS) c |= 8;
    return c;
}
static ADPoint2 lerpAD(const ADPoint2 &a, const ADPoint2 &b, double s, int nd) {
    ADPoint2 r;
    r.v = V2{a.v.x + (b.v.x - a.v.x) * s, a.v.y + (b.v.y - a.v.y) * s};
    for (int j = 0; j < nd; ++j)
        r.d[j] = V2{a.d[j].x + (b.d[j].x - a.d[j].x) * s, a.d[j].y + (b.d[j].y - a.d[j].y) * s};
    return r;
}
static bool clipLine(const ADPoint2 &v0, const ADPoint2 &v1, ADPoint2 &c0, ADPoint2 &c1, double xmin, double xmax,
                     double ymin, double ymax, int nd) {
    int o0 = outCode(v0.v, xmin, xmax, ymin, ymax), o1 = outCode(v1.v, xmin, xmax, ymin, ymax);
    c0 = v0;
    c1 = v1;
    for (int guard = 0; guard < 64; ++guard) {
        if (!(o0 | o1)) return true;
        if (o0 & o1) return false;
        ADPoint2 p;
        int oo = o0 ? o0 : o1;
        if (oo & 8) p = lerpAD(c0, c1, (ymax - c0.v.y) / (c1.v.y - c0.v.y), nd);
        else if (oo & 4) p = lerpAD(c0, c1, (ymin - c0.v.y) / (c1.v.y - c0.v.y), nd);
        else if (oo & 2) p = lerpAD(c0, c1, (xmax - c0.v.x) / (c1.v.x - c0.v.x), nd);
        else p = lerpAD(c0, c1, (xmin - c0.v.x) / (c1.v.x - c0.v.x), nd);
        if (oo == o0) { c0 = p; o0 = outCode(c0.v, xmin, xmax, ymin, ymax); }
        else { c1 = p; o1 = outCode(c1.v, xmin, xmax, ymin, ymax); }
    }
    return false;
}

template <int N>
static void projectEdgeToScreen(OScene &sc, DV<N> v0, DV<N> v1) {  // edge_manager.cpp:31-110
    const OCamera &cam = *sc.cam;
    DV<N> cpos = cam.cposAD<N>(), fs, ft, fnn;
    cam.frameAD(fs, ft, fnn);
    double clip_near = cam.clip_near;
    double z0 = ddot(v0 - cpos, fnn).v, z1 = ddot(v1 - cpos, fnn).v;
    if (z0 < clip_near && z1 < clip_near) return;
    else if (z0 < clip_near) v0 = v0 + (v1 - v0) * ((clip_near - z0) / (z1 - z0));
    else if (z1 < clip_near) v1 = v1 + (v0 - v1) * ((clip_near - z1) / (z0 - z1));
    auto xfm = [&](const V3 &p) {
        double h[4];
        for (int i = 0; i < 4; ++i) h[i] = cam.w2c[i][0] * p.x + cam.w2c[i][1] * p.y + cam.w2c[i][2] * p.z + cam.w2c[i][3];
        return V3(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
    };
    V3 v0_local = xfm(v0.val()), v1_local = xfm(v1.val());
    v0 = (v0 - cpos) / ddot(v0 - cpos, fnn);
    v1 = (v1 - cpos) / ddot(v1 - cpos, fnn);
    v0 = DV<N>(ddot(v0, fs), ddot(v0, ft), ddot(v0, fnn));
    v1 = DV<N>(ddot(v1, fs), ddot(v1, ft), ddot(v1, fnn));
    double fov = cam.c2ndc[0][0], aspect = double(cam.width) / double(cam.height);
    ADPoint2 r0, r1;
    D<N> a = v0.x * fov * 0.5 + 0.5, b = v0.y * fov * (-0.5) * aspect + 0.5;
    D<N> c = v1.x * fov * 0.5 + 0.5, d = v1.y * fov * (-0.5) * aspect + 0.5;
    r0.v = V2{a.v, b.v};
    r1.v = V2{c.v, d.v};
    for (int j = 0; j < N; ++j) { r0.d[j] = V2{a.g[j], b.g[j]}; r1.d[j] = V2{c.g[j], d.g[j]}; }
    double xmin = 0, xmax = 1, ymin = 0, ymax = 1;
    if (cam.cropped()) {
        xmin = (float)cam.ox / cam.width;
        xmax = (float)(cam.ox + cam.cw) / cam.width;
        ymin = (float)cam.oy / cam.height;
        ymax = (float)(cam.oy + cam.ch) / cam.height;
    }
    ADPoint2 s0, s1;
    if (!clipLine(r0, r1, s0, s1, xmin, xmax, ymin, ymax, N)) return;
    V3 v0p = v0.val(), v1p = v1.val();
    V3 tmp0 = v0_local, tmp1 = v1_local;
    double dv0 = std::hypot(r0.v.x - s0.v.x, r0.v.y - s0.v.y);
    if (dv0 > EPS) {
        V3 pc((s0.v.x - 0.5) * 2.0 / fov, (s0.v.y - 0.5) * (-2.0) / (aspect * fov), 1.0);
        double t0 = norm(pc - v0.val()) / norm(v1.val() - v0.val());
        v0p = v0.val() + t0 * (v1.val() - v0.val());
        double t1 = computeIntersectionInTri(V3(), v0.val(), v1.val(), tmp0, tmp1, t0);
        v0_local = tmp0 + t1 * (tmp1 - tmp0);
    }
    double dv1 = std::hypot(r1.v.x - s1.v.x, r1.v.y - s1.v.y);
    if (dv1 > EPS) {
        V3 pc((s1.v.x - 0.5) * 2.0 / fov, (s1.v.y - 0.5) * (-2.0) / (aspect * fov), 1.0);
        double t0 = norm(pc - v0.val()) / norm(v1.val() - v0.val());
        v1p = v0.val() + t0 * (v1.val() - v0.val());
        double t1 = computeIntersectionInTri(V3(), v0.val(), v1.val(), tmp0, tmp1, t0);
        v1_local = tmp0 + t1 * (tmp1 - tmp0);
    }
    PrimaryEdge pe;
    pe.v0s = V2{s0.v.x * cam.width, s0.v.y * cam.height};
    pe.v1s = V2{s1.v.x * cam.width, s1.v.y * cam.height};
    for (int j = 0; j < N; ++j) {
        pe.v0s_d[j] = V2{s0.d[j].x * cam.width, s0.d[j].y * cam.height};
        pe.v1s_d[j] = V2{s1.d[j].x * cam.width, s1.d[j].y * cam.height};
    }
    pe.v0p = v0p; pe.v1p = v1p; pe.v0c = v0_local; pe.v1c = v1_local;
    sc.primary_edges.push_back(pe);
    sc.primary_distrb.append(std::hypot(pe.v0s.x - pe.v1s.x, pe.v0s.y - pe.v1s.y));
}

template <int N> static void buildPrimaryEdges(OScene &sc) {  // edge_manager.cpp:6-29
    sc.primary_edges.clear();
    sc.primary_distrb.clear();
    for (const OShape *sh : sc.shapes) {
        if (!(!sc.bsdfs[sh->bsdf_id]->isNull() || sh->light_id >= 0)) continue;
        for (const Edge &e : sh->edges)
            if (sh->isSilhouette(e, sc.cam->cpos) > 0)
                projectEdgeToScreen<N>(sc, sh->vertexAD<N>(e.v0), sh->vertexAD<N>(e.v1));
    }
    if (sc.primary_distrb.size() > 0) sc.primary_distrb.normalize();
}

// bruteforce.cpp:4-20 (plus the primary edges of its EdgeManager base)
void OScene::initEdges(const double *weights) {
    std::vector<double> w(shapes.size());
    for (size_t i = 0; i < shapes.size(); ++i) {
        bool ignore = bsdfs[shapes[i]->bsdf_id]->isNull() && shapes[i]->light_id < 0;
        w[i] = (ignore ? 0.0 : shapes[i]->edge_distrb.sum) * (weights ? weights[i] : 1.0);
    }
    edges_distrb.clear();
    double s = 0;
    for (double x : w) s += x;
    if (s > EPS) {
        for (double x : w) edges_distrb.append(x);
        edges_distrb.normalize();
    }
    switch (nder) {
        case 1: buildPrimaryEdges<1>(*this); break;
        case 2: buildPrimaryEdges<2>(*this); break;
        case 3: buildPrimaryEdges<3>(*this); break;
        case 4: buildPrimaryEdges<4>(*this); break;
        case 8: buildPrimaryEdges<8>(*this); break;
        case 16: buildPrimaryEdges<16>(*this); break;
        default: ok = false; err = "unsupported nder";
    }
}

// ============================================================ intersections
struct Its {   // intersection.h:8-51
    int shape = -1;
    double t = INF;
    V3 p, wi;
    Frame geo, sh;
};
template <int N> struct ItsAD {   // intersectionAD.h:8-51
    int shape = -1;
    D<N> t;
    DV<N> p, wi;
    FrameAD<N> geo, sh;
    Its toIts() const {
        Its r;
        r.shape = shape; r.t = t.v; r.p = p.val(); r.wi = wi.val();
        r.geo = geo.toFrame(); r.sh = sh.toFrame();
        return r;
    }
};

// shape.cpp:166-208
static void shapeIntersect(const OShape &sh, int k, const V3 &o, const V3 &d, Its &its) {
    const auto &ix = sh.f[k];
    V3 uvt = rayIntersectTriangle(sh.v[ix[0]], sh.v[ix[1]], sh.v[ix[2]], o, d);
    double u = uvt.x, v = uvt.y, w = 1.0 - (u + v), t = uvt.z;
    V3 gn = sh.fn[k], sn = gn;
    if (sh.hasNormals()) {
        V3 nn = w * sh.n[ix[0]] + u * sh.n[ix[1]] + v * sh.n[ix[2]];
        sn = normalized(nn);
        if (dot(gn, sn) < 0.0) gn = -gn;
    }
    its.geo = Frame(gn);
    its.sh = Frame(sn);
    its.p = o + d * t;
    its.t = t;
    its.wi = its.sh.toLocal(-d);
}
// shape.cpp:210-252
template <int N>
static void shapeIntersectAD(const OShape &sh, int k, const DV<N> &o, const DV<N> &d, ItsAD<N> &its) {
    const auto &ix = sh.f[k];
    D<N> u, v, t;
    rayIntersectTriangleAD<N>(sh.vertexAD<N>(ix[0]), sh.vertexAD<N>(ix[1]), sh.vertexAD<N>(ix[2]), o, d, u, v, t);
    D<N> w = 1.0 - (u + v);
    DV<N> gn = sh.faceNormalAD<N>(k), sn = gn;
    if (sh.hasNormals()) {
        DV<N> nn = sh.normalAD<N>(ix[0]) * w + sh.normalAD<N>(ix[1]) * u + sh.normalAD<N>(ix[2]) * v;
        sn = dnormalized(nn);
        if (dot(gn.val(), sn.val()) < 0.0) gn = -gn;
    }
    its.geo = FrameAD<N>(gn);
    its.sh = FrameAD<N>(sn);
    its.p = o + d * t;
    its.t = t;
    its.wi = its.sh.toLocal(-d);
}

// ================================================================= media
// gridvolume.h:144-171, 296-317
static inline V3 worldToVolume(const OMedium &m, const V3 &p) {
    V3 r;
    for (int i = 0; i < 3; ++i) r[i] = m.w2v[i][0] * p.x + m.w2v[i][1] * p.y + m.w2v[i][2] * p.z + m.w2v[i][3];
    return r;
}
static double gridTrilinear(const float *data, const int *res, const V3 &g, int ch, int nch) {
    int x1 = (int)std::floor(g.x), y1 = (int)std::floor(g.y), z1 = (int)std::floor(g.z);
    int x2 = x1 + 1, y2 = y1 + 1, z2 = z1 + 1;
    if (x1 < 0 || x2 >= res[0] || y1 < 0 || y2 >= res[1] || z1 < 0 || z2 >= res[2]) return 0.0;
    double fx = g.x - x1, fy = g.y - y1, fz = g.z - z1, _fx = 1 - fx, _fy = 1 - fy, _fz = 1 - fz;
    auto at = [&](int x, int y, int z) { return (double)data[((size_t)(z * res[1] + y) * res[0] + x) * nch + ch]; };
    return ((at(x1, y1, z1) * _fx + at(x2, y1, z1) * fx) * _fy + (at(x1, y2, z1) * _fx + at(x2, y2, z1) * fx) * fy) * _fz +
           ((at(x1, y1, z2) * _fx + at(x2, y1, z2) * fx) * _fy + (at(x1, y2, z2) * _fx + at(x2, y2, z2) * fx) * fy) * fz;
}
static V3 toGrid(const V3 &pv, const V3 &mn, const V3 &ext, const int *res) {
    return V3((pv.x - mn.x) * ((res[0] - 1) / ext.x), (pv.y - mn.y) * ((res[1] - 1) / ext.y), (pv.z - mn.z) * ((res[2] - 1) / ext.z));
}
static double lookupFloat(const OMedium &m, const V3 &p) {
    V3 pv = worldToVolume(m, p);
    V3 g((pv.x - m.bmin.x) * (m.res[0] - 1) / m.extent.x, (pv.y - m.bmin.y) * (m.res[1] - 1) / m.extent.y,
         (pv.z - m.bmin.z) * (m.res[2] - 1) / m.extent.z);
    return gridTrilinear(m.density.data(), m.res, g, 0, 1);
}
static Spec lookupSpectrum(const OMedium &m, const V3 &p) {   // gridvolume.h:204-232
    V3 pv = worldToVolume(m, p);
    V3 g((pv.x - m.amin.x) * (m.albedo_res[0] - 1) / m.aextent.x, (pv.y - m.amin.y) * (m.albedo_res[1] - 1) / m.aextent.y,
         (pv.z - m.amin.z) * (m.albedo_res[2] - 1) / m.aextent.z);
    return Spec(gridTrilinear(m.albedo_grid.data(), m.albedo_res, g, 0, 3), gridTrilinear(m.albedo_grid.data(), m.albedo_res, g, 1, 3),
                gridTrilinear(m.albedo_grid.data(), m.albedo_res, g, 2, 3));
}
// gridvolume.h:301-312 (rigid motion about the volume centre) and :320-326
// The albedo grid's velocities are never initialised by the reference
// (heterogeneous.h:203-215 only moves m_densityVol), so it ignores rigid motion.
template <int N> static DV<N> worldToGridAD(const OMedium &m, const DV<N> &p, const V3 &mn, const V3 &ext, const int *res,
                                            bool moving) {
    DV<N> pv(worldToVolume(m, p.val()));
    V3 q = p.val() - m.center;
    for (int j = 0; j < N; ++j) {
        V3 a = moving ? p.grad(j) - cross(m.rot_d[j], q) - m.trans_d[j] : p.grad(j);
        V3 b;
        for (int i = 0; i < 3; ++i) b[i] = m.w2v[i][0] * a.x + m.w2v[i][1] * a.y + m.w2v[i][2] * a.z;
        pv.setGrad(j, b);
    }
    DV<N> g;
    g.x = (pv.x - mn.x) * ((double)(res[0] - 1) / ext.x);
    g.y = (pv.y - mn.y) * ((double)(res[1] - 1) / ext.y);
    g.z = (pv.z - mn.z) * ((double)(res[2] - 1) / ext.z);
    return g;
}
template <int N> static D<N> trilinearAD(const float *data, const int *res, const DV<N> &g, int ch, int nch) {
    int x1 = (int)std::floor(g.x.v), y1 = (int)std::floor(g.y.v), z1 = (int)std::floor(g.z.v);
    int x2 = x1 + 1, y2 = y1 + 1, z2 = z1 + 1;
    if (x1 < 0 || x2 >= res[0] || y1 < 0 || y2 >= res[1] || z1 < 0 || z2 >= res[2]) return D<N>(0.0);
    D<N> fx = g.x - (double)x1, fy = g.y - (double)y1, fz = g.z - (double)z1;
    D<N> _fx = 1.0 - fx, _fy = 1.0 - fy, _fz = 1.0 - fz;
    auto at = [&](int x, int y, int z) { return (double)data[((size_t)(z * res[1] + y) * res[0] + x) * nch + ch]; };
    return ((at(x1, y1, z1) * _fx + at(x2, y1, z1) * fx) * _fy + (at(x1, y2, z1) * _fx + at(x2, y2, z1) * fx) * fy) * _fz +
           ((at(x1, y1, z2) * _fx + at(x2, y1, z2) * fx) * _fy + (at(x1, y2, z2) * _fx + at(x2, y2, z2) * fx) * fy) * fz;
}
template <int N> static D<N> lookupFloatAD(const OMedium &m, const DV<N> &p) {  // gridvolume.h:174-201
    return trilinearAD<N>(m.density.data(), m.res, worldToGridAD<N>(m, p, m.bmin, m.extent, m.res, true), 0, 1);
}
template <int N> static DS<N> lookupSpectrumAD(const OMedium &m, const DV<N> &p) {  // gridvolume.h:235-263
    DV<N> g = worldToGridAD<N>(m, p, m.amin, m.aextent, m.albedo_res, false);
    DS<N> r;
    for (int k = 0; k < 3; ++k) r.c[k] = trilinearAD<N>(m.albedo_grid.data(), m.albedo_res, g, k, 3);
    return r;
}
template <int N> static D<N> densityScaleAD(const OMedium &m) {
    D<N> s(m.scale);
    for (int j = 0; j < N; ++j) s.g[j] = m.scale_d[j];
    return s;
}
template <int N> static DS<N> albedoAD(const OMedium &m) {
    DS<N> a(m.albedo);
    for (int j = 0; j < N; ++j) for (int k = 0; k < 3; ++k) a.c[k].g[j] = m.albedo_d[j][k];
    return a;
}
template <int N> static D<N> sigmaTAD(const OMedium &m) {
    D<N> s(m.sigma_t);
    for (int j = 0; j < N; ++j) s.g[j] = m.sigma_t_d[j];
    return s;
}

// aabb.h:186-220 — slab test with the parallel-ray special case
static bool aabbIntersect(const V3 &lo, const V3 &hi, const V3 &o, const V3 &d, double &nearT, double &farT) {
    nearT = -INF;
    farT = INF;
    for (int i = 0; i < 3; ++i) {
        if (d[i] == 0) {
            if (o[i] < lo[i] || o[i] > hi[i]) return false;
        } else {
            double t1 = (lo[i] - o[i]) / d[i], t2 = (hi[i] - o[i]) / d[i];
            if (t1 > t2) std::swap(t1, t2);
            nearT = std::max(t1, nearT);
            farT = std::min(t2, farT);
            if (!(nearT <= farT)) return false;
        }
    }
    return true;
}

// heterogeneous.cpp:121-141
template <int N>
static void intSigT(const OMedium &m, const DV<N> &org, const DV<N> &dir, const D<N> &tmin, const D<N> &tmax, Pcg &s,
                    double *out) {
    for (int j = 0; j < N; ++j) out[j] = 0;
    const double len = tmax.v - tmin.v;
    D<N> sc = densityScaleAD<N>(m);
    for (int i = 0; i < 10; ++i) {
        double rnd = (i + s.next1D()) / 10;
        D<N> tt = tmin + len * rnd;
        D<N> sig = sc * lookupFloatAD<N>(m, org + dir * tt);
        for (int j = 0; j < N; ++j) out[j] += sig.g[j];
    }
    for (int j = 0; j < N; ++j) out[j] *= len / 10.0;
    V3 o = org.val(), d = dir.val();
    if (!tmax.gradZero(EPS)) {
        double l = lookupFloat(m, o + d * (tmax.v - EPS));
        for (int j = 0; j < N; ++j) out[j] += tmax.g[j] * m.scale * l;
    }
    if (!tmin.gradZero(EPS)) {
        double l = lookupFloat(m, o + d * (tmin.v + EPS));
        for (int j = 0; j < N; ++j) out[j] -= tmin.g[j] * m.scale * l;
    }
}

// heterogeneous.cpp:9-37, homogeneous.cpp:6-24
static bool sampleDistance(const OMedium &m, const V3 &o, const V3 &d, double tmax, const double *rnd2, Pcg &s,
                           V3 &p_scatter, Spec &thr) {
    if (m.type == 0) {
        if (m.sampling_weight > EPS) {
            double t = -std::log(rnd2[0]) / m.sigma_t;
            if (t < tmax) {
                p_scatter = o + d * t;
                if (rnd2[1] < m.sampling_weight) thr *= m.albedo / m.sampling_weight;
                else thr = Spec();
                return true;
            }
            return false;
        }
        thr *= std::exp(-tmax * m.sigma_t);
        return false;
    }
    double tNear, tFar;
    if (!aabbIntersect(m.aabb_min, m.aabb_max, o, d, tNear, tFar)) return false;
    tNear = std::max(tNear, 0.0);
    tFar = std::min(tFar, tmax);
    double t = tNear;
    for (int dep = 0;; ++dep) {
        t -= std::log(s.next1D()) * m.maxInvDensity;
        if (t > tFar) break;
        if (dep > 10000) { g_counters.dt_cap++; break; }
        V3 p = o + d * t;
        double dens = lookupFloat(m, p);
        if (s.next1D() * m.maxDensity < dens) {
            p_scatter = p;
            thr *= m.hasAlbedoGrid() ? lookupSpectrum(m, p) : m.albedo;
            return true;
        }
    }
    return false;
}

// heterogeneous.cpp:40-75, homogeneous.cpp:27-49
template <int N>
static bool sampleDistanceAD(const OMedium &m, const DV<N> &org, const DV<N> &dir, const D<N> &tmax, const double *rnd2,
                             Pcg &s, DV<N> &p_scatter, DS<N> &thr) {
    if (m.type == 0) {
        D<N> st = sigmaTAD<N>(m);
        if (m.sampling_weight > EPS) {
            double t = -std::log(rnd2[0]) / m.sigma_t;
            if (t < tmax.v) {
                p_scatter = org + dir * D<N>(t);
                if (rnd2[1] < m.sampling_weight) {
                    D<N> tmp = st * dexp(-t * st);
                    thr *= albedoAD<N>(m) * (tmp / (tmp.v * m.sampling_weight));
                } else {
                    thr = DS<N>();
                }
                return true;
            }
            D<N> tmp = dexp(-tmax * st);
            thr *= tmp / tmp.v;
            return false;
        }
        thr *= dexp(-tmax * st);
        return false;
    }
    V3 o = org.val(), d = dir.val();
    double tNear, tFar;
    if (!aabbIntersect(m.aabb_min, m.aabb_max, o, d, tNear, tFar)) return false;
    tNear = std::max(tNear, 0.0);
    tFar = std::min(tFar, tmax.v);
    double t = tNear;
    for (int dep = 0;; ++dep) {
        t -= std::log(s.next1D()) * m.maxInvDensity;
        if (t > tFar) break;
        if (dep > 10000) { g_counters.dt_cap++; break; }
        DV<N> p = org + dir * D<N>(t);
        D<N> dens = lookupFloatAD<N>(m, p);
        if (s.next1D() * m.maxDensity < dens.v) {
            p_scatter = p;
            DS<N> alb = m.hasAlbedoGrid() ? lookupSpectrumAD<N>(m, p) : albedoAD<N>(m);
            D<N> sigT = densityScaleAD<N>(m) * dens;
            D<N> tmp(1.0);
            double is[N];
            intSigT<N>(m, org, dir, D<N>(0.0), D<N>(t), s, is);
            for (int j = 0; j < N; ++j) tmp.g[j] = -is[j];
            thr *= alb * (sigT * tmp / sigT.v);
            return true;
        }
    }
    D<N> tmp(1.0);
    double is[N];
    intSigT<N>(m, org, dir, D<N>(0.0), tmax, s, is);
    for (int j = 0; j < N; ++j) tmp.g[j] = -is[j];
    thr *= tmp;
    return false;
}

// heterogeneous.cpp:78-90 (residual ratio tracking; m_residualRatioTracking is
// always true), homogeneous.cpp:52-54
static double mediumTransmittance(const OMedium &m, const V3 &o, const V3 &d, double tmin, double tmax, Pcg &s) {
    if (m.type == 0) return std::exp((tmin - tmax) * m.sigma_t);
    double ret = 1.0, t = tmin;
    while (ret > EPS) {
        t -= std::log(s.next1D()) * m.maxInvDensity;
        if (t > tmax) break;
        ret *= 1.0 - lookupFloat(m, o + d * t) / m.maxDensity;
    }
    return ret;
}
// heterogeneous.cpp:113-118, homogeneous.cpp:57-59
template <int N>
static D<N> mediumTransmittanceAD(const OMedium &m, const DV<N> &org, const DV<N> &dir, const D<N> &tmin, const D<N> &tmax,
                                  Pcg &s) {
    if (m.type == 0) return dexp((tmin - tmax) * sigmaTAD<N>(m));
    D<N> ret(mediumTransmittance(m, org.val(), dir.val(), tmin.v, tmax.v, s));
    double is[N];
    intSigT<N>(m, org, dir, tmin, tmax, s, is);
    for (int j = 0; j < N; ++j) ret.g[j] = -ret.v * is[j];
    return ret;
}
static Spec sigS(const OMedium &m, const V3 &x) {   // heterogeneous.h:73-75, homogeneous.h:48
    if (m.type == 0) return m.albedo * m.sigma_t;
    return m.albedo * (m.scale * lookupFloat(m, x));
}

// ================================================================ phase
// phase.h:28-58 — note: the sampled direction is built in Frame(wi), where the
// evaluation treats wi.wo = -1 as forward; restated as written.
static double phaseEval(const OPhase &ph, const V3 &wi, const V3 &wo) {
    if (ph.type == 1) return INV_FOURPI;
    double g = ph.g, temp = 1.0 + g * g + 2.0 * g * dot(wi, wo);
    return INV_FOURPI * (1.0 - g * g) / (temp * std::sqrt(temp));
}
template <int N> static D<N> phaseEvalAD(const OPhase &ph, const DV<N> &wi, const DV<N> &wo) {
    if (ph.type == 1) return D<N>(INV_FOURPI);
    D<N> g(ph.g);
    for (int j = 0; j < N; ++j) g.g[j] = ph.g_d[j];
    D<N> temp = 1.0 + g * g + 2.0 * g * ddot(wi, wo);
    return INV_FOURPI * (1.0 - g * g) / dpow(temp, 1.5);
}
static double phaseSample(const OPhase &ph, const V3 &wi, double u1, double u2, V3 &wo) {
    if (ph.type == 1) { wo = squareToUniformSphere(u1, u2); return 1.0; }
    double g = ph.g, cosTheta;
    if (std::abs(g) < EPS) {
        cosTheta = 1.0 - 2.0 * u1;
    } else {
        double sqrTerm = (1.0 - g * g) / (1.0 - g + 2.0 * g * u1);
        cosTheta = (1.0 + g * g - sqrTerm * sqrTerm) / (2.0 * g);
    }
    double sinTheta = std::sqrt(1.0 - cosTheta * cosTheta);
    double sinPhi = std::sin(2.0 * PI * u2), cosPhi = std::cos(2.0 * PI * u2);
    wo = Frame(wi).toWorld(V3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta));
    return 1.0;
}
template <int N> static D<N> phaseSampleAD(const OPhase &ph, const DV<N> &wi, double u1, double u2, DV<N> &wo) {
    V3 w;
    phaseSample(ph, wi.val(), u1, u2, w);
    wo = DV<N>(w);
    if (ph.type == 1) {  // base PhaseFunction::sampleAD (phase.h:12-17)
        return INV_FOURPI > EPS ? D<N>(1.0) : D<N>(0.0);
    }
    D<N> tmp = phaseEvalAD<N>(ph, wi, wo);
    return tmp.v > EPS ? tmp / tmp.v : D<N>(0.0);
}

// ================================================================= BSDFs
// microfacet.h:10-248
struct Microfacet {
    double alpha;
    double eval(const V3 &m) const {
        if (Frame::cosTheta(m) < EPS) return 0.0;
        double c2 = Frame::cosTheta2(m);
        double be = (m.x * m.x + m.y * m.y) / (alpha * alpha * c2);
        double r = std::exp(-be) / (PI * alpha * alpha * c2 * c2);
        if (r * Frame::cosTheta(m) < 1e-20) r = 0.0;
        return r;
    }
    double smithG1(const V3 &v, const V3 &m) const {
        if (dot(v, m) * Frame::cosTheta(v) < EPS) return 0.0;
        double tanTheta = std::abs(Frame::tanTheta(v));
        if (tanTheta == 0.0) return 1.0;
        double a = 1.0 / (alpha * tanTheta);
        if (a >= 1.6f) return 1.0;
        double a2 = a * a;
        return (3.535f * a + 2.181f * a2) / (1.0 + 2.276f * a + 2.577f * a2);
    }
    double G(const V3 &wi, const V3 &wo, const V3 &m) const { return smithG1(wi, m) * smithG1(wo, m); }
    double pdf(const V3 &wi, const V3 &m) const {
        if (std::abs(Frame::cosTheta(wi)) < EPS) return 0.0;
        return smithG1(wi, m) * std::abs(dot(wi, m)) * eval(m) / std::abs(Frame::cosTheta(wi));
    }
    V2 sampleVisible11(double thetaI, double s0, double s1) const {
        const double SQRT_PI_INV = 1.0 / std::sqrt(PI);
        if (thetaI < 1e-4f) {
            double r = std::sqrt(-std::log(1.0 - s0));
            double ph = 2.0 * PI * s1;
            return V2{r * std::cos(ph), r * std::sin(ph)};
        }
        double tanThetaI = std::tan(thetaI), cotThetaI = 1.0 / tanThetaI;
        double a = -1.0, c = erf_as(cotThetaI);
        double sx = std::max(s0, 1e-6);
        double fit = 1.0 + thetaI * (-0.876f + thetaI * (0.4265f - 0.0594f * thetaI));
        double b = c - (1.0 + c) * std::pow(1.0 - sx, fit);
        double normalization = 1.0 / (1.0 + c + SQRT_PI_INV * tanThetaI * std::exp(-cotThetaI * cotThetaI));
        int it = 0;
        while (++it < 10) {
            if (!(b >= a && b <= c)) b = 0.5 * (a + c);
            double invErf = erfinv_giles(b);
            double value = normalization * (1.0 + b + SQRT_PI_INV * tanThetaI * std::exp(-invErf * invErf)) - sx;
            double derivative = normalization * (1.0 - invErf * tanThetaI);
            if (std::abs(value) < 1e-5f) break;
            if (value > 0) c = b; else a = b;
            b -= value / derivative;
        }
        return V2{erfinv_giles(b), erfinv_giles(2.0 * std::max(s1, (double)1e-6f) - 1.0)};
    }
    V3 sample(const V3 &wi_, double s0, double s1) const {
        V3 wi = normalized(V3(alpha * wi_.x, alpha * wi_.y, wi_.z));
        double theta = 0, phi = 0;
        if (Frame::cosTheta(wi) < 0.99999) { theta = std::acos(wi.z); phi = std::atan2(wi.y, wi.x); }
        double sinPhi = std::sin(phi), cosPhi = std::cos(phi);
        V2 sl = sampleVisible11(theta, s0, s1);
        V2 r{cosPhi * sl.x - sinPhi * sl.y, sinPhi * sl.x + cosPhi * sl.y};
        r.x *= alpha; r.y *= alpha;
        double nrm = 1.0 / std::sqrt(r.x * r.x + r.y * r.y + 1.0);
        return V3(-r.x * nrm, -r.y * nrm, nrm);
    }
};
template <int N> struct MicrofacetAD {
    D<N> alpha;
    D<N> eval(const DV<N> &m) const {
        if (m.z.v < EPS) return D<N>(0.0);
        D<N> c2 = m.z * m.z;
        D<N> be = (m.x * m.x + m.y * m.y) / (alpha * alpha * c2);
        D<N> r = dexp(-be) / (PI * alpha * alpha * c2 * c2);
        if (r.v * m.z.v < 1e-20) r = D<N>(0.0);
        return r;
    }
    D<N> smithG1(const DV<N> &v, const DV<N> &m) const {
        if (dot(v.val(), m.val()) * v.z.v < EPS) return D<N>(0.0);
        D<N> tanTheta = dabs(FrameAD<N>::tanTheta(v));
        if (tanTheta.v == 0.0) return D<N>(1.0);
        D<N> a = 1.0 / (alpha * tanTheta);
        if (a.v >= 1.6f) return D<N>(1.0);
        D<N> a2 = a * a;
        return ((double)3.535f * a + (double)2.181f * a2) / (1.0 + (double)2.276f * a + (double)2.577f * a2);
    }
};

static Microfacet mfOf(const OBsdf &b) { return Microfacet{b.alpha}; }
template <int N> static MicrofacetAD<N> mfADOf(const OBsdf &b) {
    MicrofacetAD<N> m;
    m.alpha = D<N>(b.alpha);
    for (int j = 0; j < N; ++j) m.alpha.g[j] = b.alpha_d[j];
    return m;
}
template <int N> static D<N> etaAD(const OBsdf &b) {
    D<N> e(b.eta);
    for (int j = 0; j < N; ++j) e.g[j] = b.eta_d[j];
    return e;
}
template <int N> static DS<N> specAD(const Spec &v, const std::vector<Spec> &d) {
    DS<N> r(v);
    for (int j = 0; j < N; ++j) for (int k = 0; k < 3; ++k) r.c[k].g[j] = d[j][k];
    return r;
}

static V3 phongReflect(const V3 &wi) { return V3(-wi.x, -wi.y, wi.z); }

// diffuse.cpp:7-38, null.cpp:5-29, phong.cpp:19-95, roughdielectric.cpp:9-206
// importance: adjoint transport (ptracer.cpp), roughdielectric.cpp:50,156
static Spec bsdfEval(const OBsdf &b, const Its &its, const V3 &wo, bool importance = false) {
    const V3 &wi = its.wi;
    switch (b.type) {
        case 0:
            if (wi.z < EPS || wo.z < EPS) return Spec();
            return b.refl * (INV_PI * wo.z);
        case 1: return Spec::ones();
        case 2: {
            if (wi.z < EPS || wo.z < EPS) return Spec();
            Spec r;
            double alpha = dot(wo, phongReflect(wi));
            if (alpha > 0.0) r += b.ks * ((b.ex + 2.0) * std::pow(alpha, b.ex) * INV_TWOPI);
            r += b.kd * INV_PI;
            return r * wo.z;
        }
        default: {
            if (std::abs(wi.z) < EPS || std::abs(wo.z) < EPS) return Spec();
            Microfacet mf = mfOf(b);
            bool refl = wi.z * wo.z > 0;
            double invEta = 1.0 / b.eta;
            V3 H;
            if (refl) H = normalized(wo + wi);
            else { double eta = wi.z > 0 ? b.eta : invEta; H = normalized(wi + wo * eta); }
            H = H * signum(H.z);
            double Dv = mf.eval(H);
            if (std::abs(Dv) < EPS) return Spec();
            double F = fresnelDielectricExt(dot(wi, H), b.eta);
            double G = mf.G(wi, wo, H);
            if (refl) return Spec::ones() * (F * Dv * G / (4.0 * std::abs(wi.z)));
            double eta = wi.z > 0.0 ? b.eta : invEta;
            double sqrtDenom = dot(wi, H) + eta * dot(wo, H);
            double value = ((1.0 - F) * Dv * G * eta * eta * dot(wi, H) * dot(wo, H)) / (wi.z * sqrtDenom * sqrtDenom);
            double factor = importance ? 1.0 : (wi.z > 0 ? invEta : b.eta);
            return Spec::ones() * std::abs(value * factor * factor);
        }
    }
}

static double bsdfPdf(const OBsdf &b, const Its &its, const V3 &wo) {
    const V3 &wi = its.wi;
    switch (b.type) {
        case 0:
            if (wi.z < EPS || wo.z < EPS) return 0.0;
            return squareToCosineHemispherePdf(wo);
        case 1: return 1.0;
        case 2: {
            if (wi.z < EPS || wo.z < EPS) return 0.0;
            double diffP = squareToCosineHemispherePdf(wo), specP = 0.0;
            double alpha = dot(wo, phongReflect(wi));
            if (alpha > 0) specP = std::pow(alpha, b.ex) * (b.ex + 1.0) / (2.0 * PI);
            return b.specW * specP + (1.0 - b.specW) * diffP;
        }
        default: {
            Microfacet mf = mfOf(b);
            double invEta = 1.0 / b.eta;
            bool refl = wi.z * wo.z > 0;
            V3 H;
            double dwh;
            if (refl) {
                H = normalized(wi + wo);
                dwh = 1.0 / (4.0 * dot(wo, H));
            } else {
                double eta = wi.z > 0 ? b.eta : invEta;
                H = normalized(wi + eta * wo);
                double sd = dot(wi, H) + eta * dot(wo, H);
                dwh = (eta * eta * dot(wo, H)) / (sd * sd);
            }
            H = H * signum(H.z);
            double prob = mf.pdf(wi * signum(wi.z), H);
            double F = fresnelDielectricExt(dot(wi, H), b.eta);
            prob *= refl ? F : (1.0 - F);
            return std::abs(prob * dwh);
        }
    }
}

static Spec bsdfSample(const OBsdf &b, const Its &its, const double *rnd, V3 &wo, double &pdf, double &eta,
                       bool importance = false) {
    const V3 &wi = its.wi;
    switch (b.type) {
        case 0:
            if (wi.z < EPS) return Spec();
            wo = squareToCosineHemisphere(rnd[0], rnd[1]);
            eta = 1.0;
            pdf = squareToCosineHemispherePdf(wo);
            return b.refl;
        case 1:
            wo = -wi; pdf = 1.0; eta = 1.0;
            return Spec::ones();
        case 2: {
            if (rnd[2] < b.specW) {
                V3 R = phongReflect(wi);
                double sinA = std::sqrt(1.0 - std::pow(rnd[1], 2.0 / (b.ex + 1.0)));
                double cosA = std::pow(rnd[1], 1.0 / (b.ex + 1.0));
                double phi = (2.0 * PI) * rnd[0];
                wo = Frame(R).toWorld(V3(sinA * std::cos(phi), sinA * std::sin(phi), cosA));
                if (wo.z < EPS) return Spec();
            } else {
                wo = squareToCosineHemisphere(rnd[0], rnd[1]);
            }
            eta = 1.0;
            pdf = bsdfPdf(b, its, wo);
            if (pdf < EPS) return Spec();
            return bsdfEval(b, its, wo) / pdf;
        }
        default: {
            Microfacet mf = mfOf(b);
            double invEta = 1.0 / b.eta;
            V3 wis = wi * signum(wi.z);
            V3 m = mf.sample(wis, rnd[0], rnd[1]);
            double mpdf = mf.pdf(wis, m);
            if (mpdf < EPS) return Spec();
            pdf = mpdf;
            double cosThetaT;
            double F = fresnelDielectricExt(dot(wi, m), cosThetaT, b.eta);
            double weight = 1.0, dwh;
            bool sampleRefl;
            if (rnd[2] > F) { sampleRefl = false; pdf *= 1.0 - F; }
            else { sampleRefl = true; pdf *= F; }
            if (sampleRefl) {
                wo = reflect(wi, m);
                eta = 1.0;
                if (wi.z * wo.z < EPS) return Spec();
                dwh = 1.0 / (4.0 * dot(wo, m));
            } else {
                if (std::abs(cosThetaT) < EPS) return Spec();
                wo = refract(wi, m, b.eta, cosThetaT);
                eta = cosThetaT < 0 ? b.eta : invEta;
                if (wi.z * wo.z >= 0) return Spec();
                double factor = importance ? 1.0 : (cosThetaT < 0 ? invEta : b.eta);
                weight *= factor * factor;
                double sd = dot(wi, m) + eta * dot(wo, m);
                dwh = (eta * eta * dot(wo, m)) / (sd * sd);
            }
            weight *= mf.smithG1(wo, m);
            pdf *= std::abs(dwh);
            return Spec::ones() * weight;
        }
    }
}

template <int N> static DS<N> bsdfEvalAD(const OBsdf &b, const ItsAD<N> &its, const DV<N> &wo) {
    const DV<N> &wi = its.wi;
    switch (b.type) {
        case 0:
            if (wi.z.v < EPS || wo.z.v < EPS) return DS<N>();
            return specAD<N>(b.refl, b.refl_d) * INV_PI * wo.z;
        case 1: return DS<N>(Spec::ones());
        case 2: {
            if (wi.z.v < EPS || wo.z.v < EPS) return DS<N>();
            DS<N> r;
            D<N> alpha = ddot(wo, DV<N>(-wi.x, -wi.y, wi.z));
            D<N> e(b.ex);
            for (int j = 0; j < N; ++j) e.g[j] = b.ex_d[j];
            if (alpha.v > 0.0) r += specAD<N>(b.ks, b.ks_d) * ((e + 2.0) * dpow(alpha, e) * INV_TWOPI);
            r += specAD<N>(b.kd, b.kd_d) * INV_PI;
            return r * wo.z;
        }
        default: {
            if (std::abs(wi.z.v) < EPS || std::abs(wo.z.v) < EPS) return DS<N>();
            MicrofacetAD<N> mf = mfADOf<N>(b);
            D<N> eta = etaAD<N>(b), invEta = 1.0 / eta;
            bool refl = wi.z.v * wo.z.v > 0;
            DV<N> H;
            if (refl) H = dnormalized(wo + wi);
            else H = dnormalized(wi + wo * (wi.z.v > 0 ? eta : invEta));
            H = H * signum(H.z.v);
            D<N> Dv = mf.eval(H);
            if (std::abs(Dv.v) < EPS) return DS<N>();
            D<N> F = fresnelDielectricExtAD<N>(ddot(wi, H), eta);
            D<N> G = mf.smithG1(wi, H) * mf.smithG1(wo, H);
            if (refl) return DS<N>::fill(F * Dv * G / (4.0 * dabs(wi.z)));
            const D<N> &e = wi.z.v > 0.0 ? eta : invEta;
            D<N> sd = ddot(wi, H) + e * ddot(wo, H);
            D<N> value = ((1.0 - F) * Dv * G * e * e * ddot(wi, H) * ddot(wo, H)) / (wi.z * sd * sd);
            D<N> factor = wi.z.v > 0 ? invEta : eta;
            return DS<N>::fill(dabs(value * factor * factor));
        }
    }
}

// bsdf.cpp:11-21 (generic AD sampling) and null.cpp:18-23
template <int N>
static DS<N> bsdfSampleAD(const OBsdf &b, const ItsAD<N> &its, const double *rnd, DV<N> &wo, double &pdf, double &eta) {
    if (b.type == 1) {
        wo = -its.wi; pdf = 1.0; eta = 1.0;
        return DS<N>(Spec::ones());
    }
    DS<N> ret;
    Its its0 = its.toIts();
    V3 wo0;
    if (!bsdfSample(b, its0, rnd, wo0, pdf, eta).isZero(EPS))
        if (pdf > EPS) {
            wo = its.sh.toLocal(DV<N>(its0.sh.toWorld(wo0)));
            DS<N> e = bsdfEvalAD<N>(b, its, wo);
            ret = e / pdf;
        }
    return ret;
}

// ============================================================== emitters
static Spec emitterEvalIts(const OEmitter &e, const Its &its, const V3 &d) {   // area.cpp:5-7, area2.cpp:5-11
    if (e.type == 0) return (e.two_sided || dot(its.geo.n, d) > 0) ? e.I : Spec();
    double tmp = dot(its.sh.n, d);
    if (tmp > EPS) return e.I * (e.kappa * INV_TWOPI * std::exp(e.kappa * (tmp - 1.0)));
    return Spec();
}
static Spec emitterEvalNorm(const OEmitter &e, const V3 &nrm, const V3 &d) {   // area.cpp:13-15, area2.cpp:21-27
    if (e.type == 0) return (e.two_sided || dot(nrm, d) > 0) ? e.I : Spec();
    double tmp = dot(nrm, d);
    if (tmp > EPS) return e.I * (e.kappa * INV_TWOPI * std::exp(e.kappa * (tmp - 1.0)));
    return Spec();
}
template <int N> static DS<N> emitterEvalNormAD(const OEmitter &e, const DV<N> &n, const DV<N> &d) {   // area.cpp:17-19, area2.cpp:29-35
    if (e.type == 0) return (e.two_sided || dot(n.val(), d.val()) > 0) ? DS<N>(e.I) : DS<N>();
    D<N> tmp = ddot(n, d);
    if (tmp.v > EPS) {
        D<N> k(e.kappa);
        for (int j = 0; j < N; ++j) k.g[j] = e.kappa_d[j];
        return specAD<N>(e.I, e.I_d) * (dexp(k * (tmp - 1.0)) * k * INV_TWOPI);
    }
    return DS<N>();
}
template <int N> static DS<N> emitterEvalAD(const OEmitter &e, const ItsAD<N> &its, const DV<N> &d) {   // area.cpp:9-11, area2.cpp:13-19
    if (e.type == 0) return (e.two_sided || dot(its.geo.n.val(), d.val()) > 0) ? DS<N>(e.I) : DS<N>();
    D<N> tmp = ddot(its.sh.n, d);
    if (tmp.v > EPS) {
        D<N> k(e.kappa);
        for (int j = 0; j < N; ++j) k.g[j] = e.kappa_d[j];
        return specAD<N>(e.I, e.I_d) * (dexp(k * (tmp - 1.0)) * k * INV_TWOPI);
    }
    return DS<N>();
}

// ======================================================== scene services
template <int N> struct Tracer {
    const OScene &sc;
    explicit Tracer(const OScene &s) : sc(s) {}

    const OShape &shape(int s) const { return *sc.shapes[s]; }
    const OBsdf &bsdfOf(int s) const { return *sc.bsdfs[sc.shapes[s]->bsdf_id]; }
    const OMedium *med(int id) const { return id >= 0 ? sc.media[id] : nullptr; }
    int targetMedium(int s, const V3 &geo_n, const V3 &d) const {
        return dot(d, geo_n) > 0 ? sc.shapes[s]->med_ext : sc.shapes[s]->med_int;
    }
    bool isTransition(int s) const { return sc.shapes[s]->isMediumTransition(); }

    // scene.cpp:538-575
    bool rayIntersect(const V3 &o, const V3 &d, bool onSurface, Its &its) const {
        int s, k;
        if (!sc.intersect(o, d, onSurface ? SHADOW_EPS : 0.0, INF, s, k)) { its.t = INF; its.shape = -1; return false; }
        its.shape = s;
        shapeIntersect(shape(s), k, o, d, its);
        return true;
    }
    // scene.cpp:577-615
    bool rayIntersectAD(const DV<N> &o, const DV<N> &d, bool onSurface, ItsAD<N> &its) const {
        int s, k;
        if (!sc.intersect(o.val(), d.val(), onSurface ? SHADOW_EPS : 0.0, INF, s, k)) { its.t = D<N>(INF); its.shape = -1; return false; }
        its.shape = s;
        shapeIntersectAD<N>(shape(s), k, o, d, its);
        return true;
    }

    // scene.cpp:177-237
    double evalTransmittance(V3 o, const V3 &d, bool onSurface, int medium, double remaining, Pcg &s, int max_inter) const {
        double tmin = onSurface ? SHADOW_EPS : 0.0, T = 1.0, tnear = tmin, tfar = (1.0 - SHADOW_EPS) * remaining;
        int inter = 0;
        while (remaining > 0) {
            double tmax = remaining;
            int hs, hk;
            bool hit = sc.intersect(o, d, tnear, tfar, hs, hk);
            Its nits;
            if (hit) {
                if (inter == max_inter || !bsdfOf(hs).isNull()) return 0.0;
                shapeIntersect(shape(hs), hk, o, d, nits);
                if (nits.t < tmax) tmax = nits.t;
            }
            if (medium >= 0) T *= mediumTransmittance(*sc.media[medium], o, d, tmin, tmax, s);
            if (!hit) break;
            if (isTransition(hs)) medium = targetMedium(hs, nits.geo.n, d);
            o = o + d * tmax;
            tnear = tmin = SHADOW_EPS;
            remaining -= tmax;
            tfar = (1.0 - SHADOW_EPS) * remaining;
            inter++;
        }
        return T;
    }

    void sampleLight(const double *rnd4, int &light, int &tri, V3 &pos, V3 &nrm) const {   // scene.cpp:401-410
        int nl = (int)sc.emitters.size();
        light = (int)(std::upper_bound(sc.light_cdf.begin(), sc.light_cdf.end(), rnd4[0]) - sc.light_cdf.begin());
        light = std::min(std::max(light, 0), nl - 1);
        const OShape &sh = shape(sc.emitters[light]->shape_id);
        auto b = sc.area_cdf_pool.begin() + sc.area_cdf_begin[light];
        tri = (int)(std::upper_bound(b, b + sh.nf, rnd4[1]) - b);
        tri = std::min(std::max(tri, 0), sh.nf - 1);
        sh.samplePosition(tri, rnd4[2], rnd4[3], pos, nrm);
    }

    // scene.cpp:617-701
    Spec rayIntersectAndLookForEmitter(V3 o, const V3 &d, bool onSurface, Pcg &s, int medium, int max_inter, Its &its,
                                       double &pdf_nee) const {
        double tmin = onSurface ? SHADOW_EPS : 0.0;
        int inter = 0;
        V3 scat = o;
        int hs, hk;
        bool hit = sc.intersect(o, d, tmin, INF, hs, hk);
        its.shape = -1;
        if (!hit) { its.t = INF; return Spec(); }
        its.shape = hs;
        shapeIntersect(shape(hs), hk, o, d, its);
        double T = 1.0;
        while (true) {
            if (!hit) return Spec();
            Its nits;
            shapeIntersect(shape(hs), hk, o, d, nits);
            double tmax = nits.t;
            if (medium >= 0) T *= mediumTransmittance(*sc.media[medium], o, d, tmin, tmax, s);
            if (shape(hs).isEmitter()) {
                V3 dv = nits.p - scat;
                double G = nits.wi.z / dot(dv, dv);
                int l = shape(hs).light_id;
                pdf_nee = sc.light_pmf[l] / (sc.light_areas[l] * G);
                return emitterEvalIts(*sc.emitters[l], nits, -d) * T;
            }
            if (inter == max_inter || !bsdfOf(hs).isNull()) return Spec();
            if (isTransition(hs)) medium = targetMedium(hs, nits.geo.n, d);
            o = o + d * tmax;
            tmin = SHADOW_EPS;
            inter++;
            hit = sc.intersect(o, d, SHADOW_EPS, INF, hs, hk);
        }
    }

    // scene.cpp:703-793
    DS<N> rayIntersectAndLookForEmitterAD(DV<N> o, const DV<N> &d, bool onSurface, Pcg &s, int medium, int max_inter,
                                          ItsAD<N> &its, double &pdf_nee, ItsAD<N> *itsFar) const {
        double tmin = onSurface ? SHADOW_EPS : 0.0;
        int inter = 0;
        V3 scat = o.val(), dv = d.val();
        int hs, hk;
        bool hit = sc.intersect(o.val(), dv, tmin, INF, hs, hk);
        its.shape = -1;
        if (!hit) { its.t = D<N>(INF); return DS<N>(); }
        its.shape = hs;
        shapeIntersectAD<N>(shape(hs), hk, o, d, its);
        D<N> T(1.0), dist(0.0);
        while (true) {
            if (!hit) return DS<N>();
            ItsAD<N> nits;
            shapeIntersectAD<N>(shape(hs), hk, o, d, nits);
            nits.shape = hs;
            D<N> tmax = nits.t;
            dist += nits.t;
            if (medium >= 0) T *= mediumTransmittanceAD<N>(*sc.media[medium], o, d, D<N>(tmin), tmax, s);
            if (shape(hs).isEmitter()) {
                V3 dd = nits.p.val() - scat;
                double G = nits.wi.z.v / dot(dd, dd);
                int l = shape(hs).light_id;
                pdf_nee = sc.light_pmf[l] / (sc.light_areas[l] * G);
                if (itsFar) { *itsFar = nits; itsFar->t = dist; }
                return emitterEvalAD<N>(*sc.emitters[l], nits, -d) * T;
            }
            if (inter == max_inter || !bsdfOf(hs).isNull()) return DS<N>();
            if (isTransition(hs)) medium = targetMedium(hs, nits.geo.n.val(), dv);
            o = o + d * tmax;
            tmin = SHADOW_EPS;
            inter++;
            hit = sc.intersect(o.val(), dv, SHADOW_EPS, INF, hs, hk);
        }
    }

    // scene.cpp:399-426 (point version; wo in world space)
    Spec sampleAttenuatedEmitterDirectPt(const V3 &ps, const double *rnd4, Pcg &s, int medium, int max_inter, V3 &wo,
                                         double &pdf) const {
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);
        V3 dir = lp - ps;
        pdf = sc.light_pmf[l] / sc.light_areas[l];
        if (dot(dir, ln) < 0 && pdf != 0) {
            double dist = norm(dir);
            dir = dir / dist;
            double T = evalTransmittance(ps, dir, false, medium, dist, s, max_inter);
            if (T != 0) {
                pdf *= dist * dist / dot(ln, -dir);
                wo = dir;
                return emitterEvalNorm(*sc.emitters[l], ln, -dir) * T / pdf;
            }
        }
        return Spec();
    }
    // scene.cpp:456-491 (surface version; wo in local space)
    Spec sampleAttenuatedEmitterDirectIts(const Its &its, const double *rnd4, Pcg &s, int medium, int max_inter, V3 &wo,
                                          double &pdf, bool flag) const {
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);
        V3 dir = lp - its.p;
        pdf = sc.light_pmf[l] / sc.light_areas[l];
        bool transm = bsdfOf(sc.emitters[l]->shape_id).isTransmissive();
        if ((transm || dot(dir, its.sh.n) > 0) && dot(dir, ln) < 0 && pdf != 0) {
            if (isTransition(its.shape)) medium = targetMedium(its.shape, its.geo.n, dir);
            double dist = norm(dir);
            dir = dir / dist;
            wo = its.sh.toLocal(dir);
            if (flag && signum(its.wi.z) != signum(wo.z)) --max_inter;
            if (max_inter >= 0) {
                double T = evalTransmittance(its.p, dir, true, medium, dist, s, max_inter);
                if (T != 0) {
                    pdf *= dist * dist / dot(ln, -dir);
                    return emitterEvalNorm(*sc.emitters[l], ln, -dir) * T / pdf;
                }
            }
        }
        return Spec();
    }
    // scene.cpp:302-344 — sampleEmitterDirect: no media; visible when no surface lies
    // within [ShadowEpsilon, (1 - ShadowEpsilon) * dist] along the shadow ray
    Spec sampleEmitterDirect(const Its &its, const double *rnd4, V3 &wo, double &pdf) const {
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);
        V3 dir = lp - its.p;
        pdf = sc.light_pmf[l] / sc.light_areas[l];
        if (dot(dir, its.sh.n) > 0 && dot(dir, ln) < 0 && pdf != 0) {
            double dist = norm(dir);
            dir = dir / dist;
            int hs, hk;
            if (!sc.intersect(its.p, dir, SHADOW_EPS, (1.0 - SHADOW_EPS) * dist, hs, hk)) {
                pdf *= dist * dist / dot(ln, -dir);
                wo = its.sh.toLocal(dir);
                return emitterEvalNorm(*sc.emitters[l], ln, -dir) / pdf;
            }
        }
        return Spec();
    }
    // scene.cpp:428-454
    DS<N> sampleAttenuatedEmitterDirectPtAD(const DV<N> &ps, const double *rnd4, Pcg &s, int medium, int max_inter,
                                            DV<N> &wo, double &pdf) const {
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);
        V3 dir = lp - ps.val();
        if (dot(dir, ln) < 0) {
            double dist = norm(dir);
            dir = dir / dist;
            ItsAD<N> inear, ifar;
            DS<N> ret = rayIntersectAndLookForEmitterAD(ps, DV<N>(dir), false, s, medium, max_inter, inear, pdf, &ifar);
            if (!ret.isZero(EPS) && pdf > EPS && ifar.shape == sc.emitters[l]->shape_id && std::abs(ifar.t.v - dist) < SHADOW_EPS) {
                wo = DV<N>(dir);
                return ret / pdf;
            }
        }
        return DS<N>();
    }
    // scene.cpp:493-524
    DS<N> sampleAttenuatedEmitterDirectItsAD(const ItsAD<N> &its, const double *rnd4, Pcg &s, int medium, int max_inter,
                                             DV<N> &wo, double &pdf) const {
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);
        V3 dir = lp - its.p.val();
        bool transm = bsdfOf(sc.emitters[l]->shape_id).isTransmissive();
        if ((transm || dot(dir, its.sh.n.val()) > 0) && dot(dir, ln) < 0) {
            if (isTransition(its.shape)) medium = targetMedium(its.shape, its.geo.n.val(), dir);
            double dist = norm(dir);
            dir = dir / dist;
            ItsAD<N> inear, ifar;
            DS<N> ret = rayIntersectAndLookForEmitterAD(its.p, DV<N>(dir), true, s, medium, max_inter, inear, pdf, &ifar);
            if (!ret.isZero(EPS) && pdf > EPS && ifar.shape == sc.emitters[l]->shape_id && std::abs(ifar.t.v - dist) < SHADOW_EPS) {
                wo = its.sh.toLocal(DV<N>(dir));
                return ret / pdf;
            }
        }
        return DS<N>();
    }

    // scene.h:92-94 -> bruteforce.cpp:22-26 -> shape.cpp:337-342: shape by weight,
    // edge by length, sample reused. This is the 5-argument overload LiAD calls, so
    // the pdf is P(shape) * P(edge | shape) with no division by the edge length.
    const Edge *sampleEdge(double &rnd, int &shape_id, double &pdf) const {
        shape_id = (int)sc.edges_distrb.sampleReuse(rnd, pdf);
        const OShape &sh = shape(shape_id);
        double p1;
        int e = (int)sh.edge_distrb.sampleReuse(rnd, p1);
        pdf *= p1;
        return &sh.edges[e];
    }

    static double mis(double a, double b) { a *= a; b *= b; return a / (a + b); }   // volpathAD.cpp:43-47

    // ===================================================== integrator
    // volpathAD.cpp:50-142 — primal volumetric path tracer. With `volpath` set it is
    // VolPathTracer::radiance (volpath.cpp:10-103) instead: the same loop, but the first
    // ray leaves the camera (tnear 0) and paths end on exactly-zero throughput, phase
    // value or BSDF weight rather than below Epsilon.
    Spec Li(Pcg &s, V3 ro, V3 rd, int max_depth, int med, bool incEmission, bool volpath = false) const {
        Its its;
        Spec ret, thr = Spec::ones();
        const double zeps = volpath ? 1e-12 : EPS;   // Eigen isZero() vs isZero(Epsilon)
        rayIntersect(ro, rd, !volpath, its);
        for (int depth = 0; depth <= max_depth; ++depth) {
            int max_inter = max_depth - depth - 1;
            double rnd1[2] = {s.next1D(), s.next1D()};
            double rnd2[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
            double rnd3[3] = {s.next1D(), s.next1D(), s.next1D()};
            bool inside = false;
            if (med >= 0) {
                V3 ps;
                inside = sampleDistance(*sc.media[med], ro, rd, its.t, rnd1, s, ps, thr);
                if (inside) ro = ps;
            }
            if (inside) {
                if (depth >= max_depth) break;
                if (thr.isZero(zeps)) break;
                const OPhase &ph = *sc.phases[sc.media[med]->phase_id];
                V3 wo;
                double pdf_nee;
                Spec value = sampleAttenuatedEmitterDirectPt(ro, rnd2, s, med, max_inter, wo, pdf_nee);
                if (!value.isZero(zeps)) {
                    double pv = phaseEval(ph, -rd, wo);
                    if (volpath ? pv != 0.0 : pv > EPS) ret += thr * value * (mis(pdf_nee, phaseEval(ph, -rd, wo)) * pv);
                }
                double pv = phaseSample(ph, -rd, rnd3[0], rnd3[1], wo);
                if (volpath ? pv == 0.0 : pv < EPS) break;
                thr *= pv;
                rd = wo;
                Spec att = rayIntersectAndLookForEmitter(ro, rd, false, s, med, max_inter, its, pdf_nee);
                if (!att.isZero(zeps)) {
                    double ppdf = phaseEval(ph, -rd, wo);
                    ret += thr * att * mis(ppdf, pdf_nee);
                }
                incEmission = false;
            } else {
                if (its.shape < 0) break;
                const OShape &shp = shape(its.shape);
                if (shp.isEmitter() && incEmission) ret += thr * emitterEvalIts(*sc.emitters[shp.light_id], its, -rd);
                if (depth >= max_depth) break;
                const OBsdf &b = bsdfOf(its.shape);
                V3 wo;
                double pdf_nee;
                if (!b.isNull()) {
                    Spec value = sampleAttenuatedEmitterDirectIts(its, rnd2, s, med, max_inter, wo, pdf_nee, false);
                    if (!value.isZero(zeps)) {
                        Spec bv = bsdfEval(b, its, wo);
                        double bp = bsdfPdf(b, its, wo);
                        ret += thr * value * bv * mis(pdf_nee, bp);
                    }
                }
                double bpdf, beta;
                Spec bw = bsdfSample(b, its, rnd3, wo, bpdf, beta);
                if (bw.isZero(zeps)) break;
                wo = its.sh.toWorld(wo);
                ro = its.p;
                rd = wo;
                if (isTransition(its.shape)) med = targetMedium(its.shape, its.geo.n, wo);
                if (b.isNull()) {
                    rayIntersect(ro, rd, true, its);
                    continue;
                }
                Spec att = rayIntersectAndLookForEmitter(ro, rd, true, s, med, max_inter, its, pdf_nee);
                thr *= bw;
                if (!att.isZero()) ret += thr * att * mis(bpdf, pdf_nee);
                incEmission = false;
            }
        }
        return ret;
    }

    // volpath_simple.cpp:10-86 — VolPathTracerSimple::radiance: next-event estimation
    // without MIS; traced rays add emission only before the first non-null scattering.
    Spec LiSimple(Pcg &s, V3 ro, V3 rd, int max_depth, int med) const {
        Its its;
        Spec ret, thr = Spec::ones();
        bool incEmission = true;
        rayIntersect(ro, rd, false, its);
        for (int depth = 0; depth <= max_depth; ++depth) {
            int max_inter = max_depth - depth - 1;
            double rnd1[2] = {s.next1D(), s.next1D()};
            double rnd2[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
            double rnd3[3] = {s.next1D(), s.next1D(), s.next1D()};
            bool inside = false;
            if (med >= 0) {
                V3 ps;
                inside = sampleDistance(*sc.media[med], ro, rd, its.t, rnd1, s, ps, thr);
                if (inside) ro = ps;
            }
            if (inside) {
                if (depth >= max_depth) break;
                if (thr.isZero()) break;
                const OPhase &ph = *sc.phases[sc.media[med]->phase_id];
                V3 wo;
                double pdf_nee;
                Spec value = sampleAttenuatedEmitterDirectPt(ro, rnd2, s, med, max_inter, wo, pdf_nee);
                if (!value.isZero()) {
                    double pv = phaseEval(ph, -rd, wo);
                    if (pv != 0.0) ret += thr * value * pv;
                }
                double pv = phaseSample(ph, -rd, rnd3[0], rnd3[1], wo);
                if (pv == 0.0) break;
                thr *= pv;
                rd = wo;
                rayIntersect(ro, rd, false, its);
                incEmission = false;
            } else {
                if (its.shape < 0) break;
                const OShape &shp = shape(its.shape);
                if (shp.isEmitter() && incEmission) ret += thr * emitterEvalIts(*sc.emitters[shp.light_id], its, -rd);
                if (depth >= max_depth) break;
                const OBsdf &b = bsdfOf(its.shape);
                V3 wo;
                if (!b.isNull()) {
                    double pdf_nee;
                    Spec value = sampleAttenuatedEmitterDirectIts(its, rnd2, s, med, max_inter, wo, pdf_nee, false);
                    if (!value.isZero()) ret += thr * value * bsdfEval(b, its, wo);
                    incEmission = false;
                }
                double bpdf, beta;
                Spec bw = bsdfSample(b, its, rnd3, wo, bpdf, beta);
                if (bw.isZero()) break;
                thr *= bw;
                wo = its.sh.toWorld(wo);
                ro = its.p;
                rd = wo;
                if (isTransition(its.shape)) med = targetMedium(its.shape, its.geo.n, wo);
                rayIntersect(ro, rd, true, its);
            }
        }
        return ret;
    }

    // path.cpp:11-61 — PathTracer::radiance: surfaces only (media are ignored), emission
    // seen directly at the first hit, NEE + BSDF sampling with the power heuristic.
    // DirectIntegrator::radiance (direct.cpp:7-48) is this with max_depth 1.
    Spec LiPath(Pcg &s, V3 ro, V3 rd, int max_depth) const {
        Its its;
        Spec ret;
        if (!rayIntersect(ro, rd, true, its)) return ret;
        Spec thr = Spec::ones();
        for (int depth = 0; depth <= max_depth && its.shape >= 0; ++depth) {
            const OShape &shp = shape(its.shape);
            if (shp.isEmitter() && depth == 0) ret += thr * emitterEvalIts(*sc.emitters[shp.light_id], its, -rd);
            if (depth >= max_depth) break;
            const OBsdf &b = bsdfOf(its.shape);
            double rnd4[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
            V3 wo;
            double pdf_nee;
            Spec value = sampleEmitterDirect(its, rnd4, wo, pdf_nee);
            if (!value.isZero()) ret += thr * value * bsdfEval(b, its, wo) * mis(pdf_nee, bsdfPdf(b, its, wo));
            double rnd3[3] = {s.next1D(), s.next1D(), s.next1D()};
            double bpdf, beta;
            Spec bw = bsdfSample(b, its, rnd3, wo, bpdf, beta);
            if (bw.isZero()) break;
            wo = its.sh.toWorld(wo);
            ro = its.p;
            rd = wo;
            if (!rayIntersect(ro, rd, true, its)) break;
            thr *= bw;
            const OShape &hs = shape(its.shape);
            if (hs.isEmitter()) {
                Spec lc = emitterEvalIts(*sc.emitters[hs.light_id], its, -rd);
                if (!lc.isZero()) {
                    V3 dd = its.p - ro;
                    double G = its.wi.z / dot(dd, dd);
                    double pn = sc.light_pmf[hs.light_id] / sc.light_areas[hs.light_id] / G;
                    ret += thr * lc * mis(bpdf, pn);
                }
            }
        }
        return ret;
    }

    // ptracer.cpp:57-167 — ParticleTracer::traceParticle into img[pixel] (window size
    // w x h). The emitter's geometric frame is the sampled face normal (the reference
    // leaves it unset); a splat on the window's closing edge has no pixel and is dropped.
    void traceParticle(Pcg &s, int max_bounces, std::vector<Spec> &img) const {
        const OCamera &cam = *sc.cam;
        const int W = cam.cropped() ? cam.cw : cam.width, H = cam.cropped() ? cam.ch : cam.height;
        auto splat = [&](double px, double py, const Spec &v) {
            int ix = (int)px, iy = (int)py;
            if (ix >= 0 && iy >= 0 && ix < W && iy < H) img[(size_t)iy * W + ix] += v;
        };
        if (sc.emitters.empty()) return;
        double rnd4[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);   // scene.cpp:795-814
        const OEmitter &E = *sc.emitters[l];
        const OShape &lsh = shape(E.shape_id);
        Spec power = E.I * (sc.light_areas[l] / sc.light_pmf[l]);
        {
            double px, py;
            V3 dir;
            double v = cam.sampleDirect(lp, px, py, dir);
            if (v != 0.0) {
                int m = dot(dir, ln) > 0 ? lsh.med_ext : lsh.med_int;
                v *= evalTransmittance(lp, dir, true, m, norm(lp - cam.cpos), s, max_bounces);
                if (v != 0.0) splat(px, py, power * (v * std::max(0.0, dot(ln, dir))));
            }
        }
        double u0 = s.next1D(), u1 = s.next1D();
        V3 d = Frame(ln).toWorld(squareToCosineHemisphere(u0, u1));
        power = power * PI;
        V3 o = lp;
        int medium = dot(d, ln) > 0 ? lsh.med_ext : lsh.med_int;
        Spec thr = Spec::ones();
        bool onSurf = true;
        Its its;
        for (int depth = 0; !thr.isZero() && depth < max_bounces; ++depth) {
            rayIntersect(o, d, onSurf, its);
            const int max_inter = max_bounces - depth - 1;
            bool inside = false;
            if (medium >= 0) {
                double rnd1[2] = {s.next1D(), s.next1D()};
                V3 ps;
                inside = sampleDistance(*sc.media[medium], o, d, its.t, rnd1, s, ps, thr);
                if (inside) o = ps;
            }
            if (inside) {
                const OPhase &ph = *sc.phases[sc.media[medium]->phase_id];
                double px, py;
                V3 wo;
                double v = cam.sampleDirect(o, px, py, wo);
                if (v != 0.0) {
                    v *= evalTransmittance(o, wo, false, medium, norm(o - cam.cpos), s, max_inter);
                    if (v != 0.0) splat(px, py, thr * power * (v * phaseEval(ph, -d, wo)));
                }
                double a = s.next1D(), b = s.next1D();
                V3 wn;
                double pv = phaseSample(ph, -d, a, b, wn);
                if (pv == 0.0) break;
                thr *= pv;
                d = wn;
                onSurf = false;
            } else {
                if (its.shape < 0) break;
                const OBsdf &b = bsdfOf(its.shape);
                const OShape &sh = shape(its.shape);
                if (!b.isNull()) {
                    double px, py;
                    V3 wo;
                    double v = cam.sampleDirect(its.p, px, py, wo);
                    if (v != 0.0) {
                        int m = dot(wo, its.geo.n) > 0 ? sh.med_ext : sh.med_int;
                        v *= evalTransmittance(its.p, wo, true, m, norm(its.p - cam.cpos), s, max_inter);
                        if (v != 0.0) {
                            V3 wi = its.sh.toWorld(its.wi), wol = its.sh.toLocal(wo);
                            double wiG = dot(wi, its.geo.n), woG = dot(wo, its.geo.n);
                            if (wiG * its.wi.z > 0 && woG * wol.z > 0) {
                                double corr = std::abs((its.wi.z * woG) / (wol.z * wiG));
                                splat(px, py, bsdfEval(b, its, wol, true) * thr * power * (v * corr));
                            }
                        }
                    }
                }
                double rnd3[3] = {s.next1D(), s.next1D(), s.next1D()};
                V3 wol;
                double bpdf, beta;
                Spec bw = bsdfSample(b, its, rnd3, wol, bpdf, beta, true);
                if (bw.isZero()) break;
                V3 wo = its.sh.toWorld(wol);
                double wiG = dot(-d, its.geo.n), woG = dot(wo, its.geo.n);
                if (wiG * its.wi.z <= 0 || woG * wol.z <= 0) break;
                thr *= bw;
                if (isTransition(its.shape)) medium = woG > 0 ? sh.med_ext : sh.med_int;
                o = its.p;
                d = wo;
                onSurf = true;
            }
        }
    }
    // ptracer.cpp:10-55 — 100000 * spp particles, one stream each, image / count
    void renderParticles(uint64_t seed, int spp, int max_bounces, float *out) const {
        const OCamera &cam = *sc.cam;
        const int npix = cam.cropped() ? cam.cw * cam.ch : cam.width * cam.height;
        const int64_t total = (int64_t)100000 * spp;
        std::vector<Spec> img(npix);
#pragma omp parallel
        {
            std::vector<Spec> loc(npix);
#pragma omp for schedule(dynamic, 1024)
            for (int64_t i = 0; i < total; ++i) {
                Pcg s(seed, (uint64_t)i);
                traceParticle(s, max_bounces, loc);
            }
#pragma omp critical
            for (int p = 0; p < npix; ++p) img[p] += loc[p];
        }
        for (int p = 0; p < npix; ++p)
            for (int c = 0; c < 3; ++c) out[(size_t)p * 3 + c] = (float)(img[p][c] / (double)total);
    }

    // volpathAD.cpp:145-620 — forward-mode differential estimator
    DS<N> LiAD(const ItsAD<N> &its, Pcg &s, const DV<N> &ray_o, const DV<N> &ray_d, int med, int max_depth, int mode,
               int nEdge, double ddistCoeff) const {
        if (mode == 0) return DS<N>();
        DV<N> ro = ray_o, rd = ray_d;
        double rnd1[2] = {s.next1D(), s.next1D()};
        double rnd2[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
        double rnd3[3] = {s.next1D(), s.next1D(), s.next1D()};
        DS<N> ret, thr(Spec::ones());
        ItsAD<N> its1;
        bool volume = false;
        if (med >= 0) {
            DV<N> ps;
            volume = sampleDistanceAD<N>(*sc.media[med], ro, rd, its.t, rnd1, s, ps, thr);
            if (volume) ro = ps;
        }
        if (volume) {
            if (max_depth == 0 || thr.isZero(EPS)) return ret;
            const OPhase &ph = *sc.phases[sc.media[med]->phase_id];
            {
                double pdf_nee;
                DV<N> wo;
                if (mode & 8) {
                    DS<N> value = sampleAttenuatedEmitterDirectPtAD(ro, rnd2, s, med, max_depth - 1, wo, pdf_nee);
                    if (!value.isZero(EPS)) {
                        D<N> pv = phaseEvalAD<N>(ph, -rd, wo);
                        if (pv.v > EPS) {
                            double ppdf = phaseEval(ph, -rd.val(), wo.val());
                            ret += thr * value * (mis(pdf_nee, ppdf) * pv);
                        }
                    }
                }
                D<N> pv = phaseSampleAD<N>(ph, -rd, rnd3[0], rnd3[1], wo);
                if (pv.v > EPS) {
                    DS<N> thr1 = thr * pv;
                    rd = wo;
                    DS<N> att = rayIntersectAndLookForEmitterAD(ro, rd, false, s, med, max_depth - 1, its1, pdf_nee, nullptr);
                    if (!att.isZero(EPS) && (mode & 8)) {
                        double ppdf = phaseEval(ph, -rd.val(), wo.val());
                        ret += thr1 * att * mis(ppdf, pdf_nee);
                    }
                    ret += thr1 * LiAD(its1, s, ro, rd, med, max_depth - 1, mode & ~1, nEdge, ddistCoeff);
                }
            }
            if (nEdge > 0 && (mode & 32)) {   // volume edge term (volpathAD.cpp:207-305)
                double der[N][3] = {};
                const DV<N> &x1 = ro;
                for (int i = 0; i < nEdge; ++i) {
                    double rnd = (i + s.next1D()) / nEdge;
                    int sid;
                    double edgePdf;
                    const Edge *e = sampleEdge(rnd, sid, edgePdf);
                    double t = rnd;
                    const OShape &shp = shape(sid);
                    int sil = shp.isSilhouette(*e, x1.val());
                    bool valid = sil > 0 && !bsdfOf(sid).isNull();
                    if (!valid) continue;
                    DV<N> e0 = shp.vertexAD<N>(e->v0) - x1, e1 = shp.vertexAD<N>(e->v1) - x1;
                    V3 nrm = normalized(cross(e0.val(), e1.val()));
                    DV<N> w1;
                    double dist;
                    {
                        V3 ne0 = normalized(e0.val()), ne1 = normalized(e1.val());
                        double theta = std::acos(dot(ne0, ne1)), sinTheta = std::sin(theta);
                        if (theta < EPS || theta > PI - EPS) continue;
                        double w0 = std::sin(theta * (1.0 - t)) / sinTheta, ww1 = std::sin(theta * t) / sinTheta;
                        w0 /= norm(e0.val());
                        ww1 /= norm(e1.val());
                        double tmp = w0 + ww1;
                        w0 /= tmp;
                        ww1 /= tmp;
                        if (w0 < -EPS || ww1 < -EPS || std::abs(w0 + ww1 - 1.0) > EPS) { g_counters.bad_edge_weight++; continue; }
                        w1 = e0 * w0 + e1 * ww1;
                        D<N> dAD = dnorm(w1);
                        w1 = w1 / dAD;
                        dist = dAD.v;
                        edgePdf /= theta;
                    }
                    double tmpo = 0;
                    for (int j = 0; j < N; ++j) tmpo = std::max(tmpo, std::abs(dot(w1.grad(j), w1.val())));
                    if (tmpo > EPS) { g_counters.nonortho++; continue; }
                    if (!w1.gradZero(EPS)) {
                        double pe = phaseEval(ph, -ray_d.val(), w1.val());
                        Spec phaseVal = thr.val() * pe;
                        if (!phaseVal.isZero(EPS)) {
                            double trans = evalTransmittance(x1.val(), w1.val(), false, med, dist, s, max_depth - 1);
                            if (trans > EPS) {
                                V3 w2 = normalized(w1.val() - ANGLE_EPS * nrm);
                                Spec delta = Li(s, x1.val(), w2, max_depth - 1, med, true);
                                w2 = normalized(w1.val() + ANGLE_EPS * nrm);
                                delta -= Li(s, x1.val(), w2, max_depth - 1, med, true);
                                if (!delta.isZero(EPS))
                                    for (int j = 0; j < N; ++j) {
                                        double nd = dot(nrm, w1.grad(j));
                                        for (int c = 0; c < 3; ++c) der[j][c] += nd * phaseVal[c] * delta[c] / edgePdf;
                                    }
                            }
                        }
                    }
                }
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c) ret.c[c].g[j] += der[j][c] / nEdge;
            }
        } else {
            if (its.shape < 0 || thr.isZero(EPS)) return ret;
            const OShape &shp = shape(its.shape);
            if (shp.isMediumTransition() && targetMedium(its.shape, its.geo.n.val(), -rd.val()) != med) {
                g_counters.med_inconsistent++;
                return ret;
            }
            if (shp.isEmitter() && (mode & 1)) ret += thr * emitterEvalAD<N>(*sc.emitters[shp.light_id], its, -rd);
            if (max_depth == 0) return ret;
            const OBsdf &b = bsdfOf(its.shape);

            // Lins boundary term (volpathAD.cpp:321-467)
            if (med >= 0 && !its.t.gradZero(EPS) && (mode & 16)) {
                const double derVal = its.t.gradAbsMax();
                Its its0 = its.toIts();
                V3 n = its0.geo.n;
                if (dot(n, rd.val()) < 0.0) n = -n;
                Spec sgs = sigS(*sc.media[med], its0.p - SHADOW_EPS * n);
                const OPhase &ph = *sc.phases[sc.media[med]->phase_id];
                int nsamples = std::min(std::max((int)(ddistCoeff * derVal), 1), 100);
                Spec Lins;
                V3 rdv = rd.val();
                for (int i = 0; i < nsamples; ++i) {
                    V3 wo;
                    double pdf_nee, pdf_phase, pdf_bsdf, pv = 0;
                    Spec bv, att;
                    Its its1p;
                    if (b.isNull()) {
                        double r4[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
                        att = sampleAttenuatedEmitterDirectIts(its0, r4, s, med, max_depth - 1, wo, pdf_nee, true);
                        if (!att.isZero(EPS)) {
                            wo = its0.sh.toWorld(wo);
                            pv = phaseEval(ph, -rdv, wo);
                            if (pv > EPS) {
                                pdf_phase = phaseEval(ph, -rdv, wo);
                                Lins += att * (pv * mis(pdf_nee, pdf_phase));
                            }
                        }
                        double u1 = s.next1D(), u2 = s.next1D();
                        if ((pv = phaseSample(ph, -rdv, u1, u2, wo)) > EPS) {
                            int md1 = max_depth - 1;
                            if (dot(wo, n) > EPS) {
                                if (shp.isEmitter()) Lins += emitterEvalIts(*sc.emitters[shp.light_id], its0, -wo) * pv;
                                --md1;
                            }
                            if (md1 > 0) {
                                int med1 = shp.isMediumTransition() ? targetMedium(its.shape, its0.geo.n, wo) : med;
                                att = rayIntersectAndLookForEmitter(its0.p, wo, true, s, med1, md1, its1p, pdf_nee);
                                if (!att.isZero(EPS)) {
                                    pdf_phase = phaseEval(ph, -rdv, wo);
                                    Lins += att * (pv * mis(pdf_phase, pdf_nee));
                                }
                                Lins += Li(s, its0.p, wo, md1, med1, false) * pv;
                            }
                        }
                    } else {
                        double r4[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
                        att = sampleAttenuatedEmitterDirectIts(its0, r4, s, med, max_depth - 1, wo, pdf_nee, false);
                        if (!att.isZero(EPS)) {
                            wo = its0.sh.toWorld(wo);
                            if (dot(wo, n) < -EPS) {
                                pv = phaseEval(ph, -rdv, wo);
                                if (pv > EPS) {
                                    pdf_phase = phaseEval(ph, -rdv, wo);
                                    Lins += att * (pv * mis(pdf_nee, pdf_phase));
                                }
                            }
                        }
                        double u1 = s.next1D(), u2 = s.next1D();
                        if ((pv = phaseSample(ph, -rdv, u1, u2, wo)) > EPS) {
                            double woDotN = dot(wo, n);
                            if (woDotN < -EPS) {
                                att = rayIntersectAndLookForEmitter(its0.p, wo, true, s, med, max_depth - 1, its1p, pdf_nee);
                                if (!att.isZero(EPS)) {
                                    pdf_phase = phaseEval(ph, -rdv, wo);
                                    Lins += att * (pv * mis(pdf_phase, pdf_nee));
                                }
                                Lins += Li(s, its0.p, wo, max_depth - 1, med, false) * pv;
                            } else if (woDotN > EPS) {
                                if (shp.isEmitter()) Lins += emitterEvalIts(*sc.emitters[shp.light_id], its0, -wo) * pv;
                                if (max_depth > 1) {
                                    Its itsb = its0;
                                    itsb.wi = itsb.sh.toLocal(-wo);
                                    double r4b[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
                                    att = sampleAttenuatedEmitterDirectIts(itsb, r4b, s, med, max_depth - 2, wo, pdf_nee, false);
                                    if (!att.isZero(EPS)) {
                                        bv = bsdfEval(b, itsb, wo);
                                        pdf_bsdf = bsdfPdf(b, itsb, wo);
                                        Lins += bv * att * (pv * mis(pdf_nee, pdf_bsdf));
                                    }
                                    double r3[3] = {s.next1D(), s.next1D(), s.next1D()};
                                    double eta;
                                    bv = bsdfSample(b, itsb, r3, wo, pdf_bsdf, eta);
                                    if (!bv.isZero(EPS)) {
                                        Spec thr1 = bv * pv;
                                        wo = itsb.sh.toWorld(wo);
                                        int med1 = shp.isMediumTransition() ? targetMedium(its.shape, itsb.geo.n, wo) : med;
                                        Its itsc;
                                        att = rayIntersectAndLookForEmitter(itsb.p, wo, true, s, med1, max_depth - 2, itsc, pdf_nee);
                                        if (!att.isZero()) Lins += thr1 * att * mis(pdf_bsdf, pdf_nee);
                                        Lins += thr1 * Li(s, itsb.p, wo, max_depth - 2, med1, false);
                                    }
                                }
                            }
                        }
                    }
                }
                Lins = Lins / (double)nsamples;
                Spec tv = thr.val();
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c) ret.c[c].g[j] += its.t.g[j] * sgs[c] * Lins[c] * tv[c];
            }

            // area sampling (volpathAD.cpp:469-511)
            {
                DV<N> wo;
                double pdf_nee;
                if (!b.isNull() && (mode & 2)) {
                    DS<N> value = sampleAttenuatedEmitterDirectItsAD(its, rnd2, s, med, max_depth - 1, wo, pdf_nee);
                    if (!value.isZero(EPS)) {
                        DS<N> bv = bsdfEvalAD<N>(b, its, wo);
                        double bp = bsdfPdf(b, its.toIts(), wo.val());
                        ret += thr * value * bv * mis(pdf_nee, bp);
                    }
                }
                double bpdf, beta;
                DS<N> bw = bsdfSampleAD<N>(b, its, rnd3, wo, bpdf, beta);
                if (!bw.isZero(EPS)) {
                    wo = its.sh.toWorld(wo);
                    DV<N> no = its.p, nd = wo;
                    int med1 = shp.isMediumTransition() ? targetMedium(its.shape, its.geo.n.val(), wo.val()) : med;
                    DS<N> thr1 = thr;
                    int mode1;
                    if (b.isNull()) {
                        rayIntersectAD(no, nd, true, its1);
                        mode1 = mode;
                    } else {
                        DS<N> att = rayIntersectAndLookForEmitterAD(no, nd, true, s, med1, max_depth - 1, its1, pdf_nee, nullptr);
                        thr1 *= bw;
                        if (!att.isZero(EPS) && (mode & 2)) ret += thr1 * att * mis(bpdf, pdf_nee);
                        mode1 = mode & ~1;
                    }
                    ret += thr1 * LiAD(its1, s, no, nd, med1, max_depth - 1, mode1, nEdge, ddistCoeff);
                }
            }

            // interface edge term (volpathAD.cpp:513-616)
            if (!b.isNull() && nEdge > 0 && (mode & 4)) {
                double der[N][3] = {};
                const DV<N> &x0 = its.p;
                Its its0 = its.toIts();
                for (int i = 0; i < nEdge; ++i) {
                    double rnd = (i + s.next1D()) / nEdge;
                    int sid;
                    double edgePdf;
                    const Edge *e = sampleEdge(rnd, sid, edgePdf);
                    double t = rnd;
                    const OShape &esh = shape(sid);
                    int sil = esh.isSilhouette(*e, x0.val());
                    bool valid = sil > 0 && !bsdfOf(sid).isNull();
                    if (!valid) continue;
                    DV<N> e0 = esh.vertexAD<N>(e->v0) - x0, e1 = esh.vertexAD<N>(e->v1) - x0;
                    V3 nrm = normalized(cross(e0.val(), e1.val()));
                    DV<N> w1;
                    double dist;
                    {
                        V3 ne0 = normalized(e0.val()), ne1 = normalized(e1.val());
                        double theta = std::acos(dot(ne0, ne1)), sinTheta = std::sin(theta);
                        if (theta < EPS || theta > PI - EPS) continue;
                        double w0 = std::sin(theta * (1.0 - t)) / sinTheta, ww1 = std::sin(theta * t) / sinTheta;
                        w0 /= norm(e0.val());
                        ww1 /= norm(e1.val());
                        double tmp = w0 + ww1;
                        w0 /= tmp;
                        ww1 /= tmp;
                        if (w0 < -EPS || ww1 < -EPS || std::abs(w0 + ww1 - 1.0) > EPS) { g_counters.bad_edge_weight++; continue; }
                        w1 = e0 * w0 + e1 * ww1;
                        D<N> dAD = dnorm(w1);
                        w1 = w1 / dAD;
                        dist = dAD.v;
                        edgePdf /= theta;
                    }
                    double tmpo = 0;
                    for (int j = 0; j < N; ++j) tmpo = std::max(tmpo, std::abs(dot(w1.grad(j), w1.val())));
                    if (tmpo > EPS) { g_counters.nonortho++; continue; }
                    if (!w1.gradZero(EPS)) {
                        Spec bsdfVal = thr.val() * bsdfEval(b, its0, its0.sh.toLocal(w1.val()));
                        if (!bsdfVal.isZero(EPS)) {
                            int med1 = shp.isMediumTransition() ? targetMedium(its.shape, its0.geo.n, w1.val()) : med;
                            double trans = evalTransmittance(x0.val(), w1.val(), true, med1, dist, s, max_depth - 1);
                            if (trans > EPS) {
                                V3 w2 = normalized(w1.val() - ANGLE_EPS * nrm);
                                Spec delta = Li(s, x0.val(), w2, max_depth - 1, med1, true);
                                w2 = normalized(w1.val() + ANGLE_EPS * nrm);
                                delta -= Li(s, x0.val(), w2, max_depth - 1, med1, true);
                                if (!delta.isZero(EPS))
                                    for (int j = 0; j < N; ++j) {
                                        double ndot = dot(nrm, w1.grad(j));
                                        for (int c = 0; c < 3; ++c) der[j][c] += ndot * bsdfVal[c] * delta[c] / edgePdf;
                                    }
                            }
                        }
                    }
                }
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c) ret.c[c].g[j] += der[j][c] / nEdge;
            }
        }
        return ret;
    }

    // scene.cpp:346-397 — sampleEmitterDirectAD (no media)
    DS<N> sampleEmitterDirectAD(const ItsAD<N> &its, const double *rnd4, DV<N> &wo, double &pdf) const {
        int l, k;
        V3 lp, ln;
        sampleLight(rnd4, l, k, lp, ln);
        V3 dir = lp - its.p.val();
        pdf = sc.light_pmf[l] / sc.light_areas[l];
        if (dot(dir, its.sh.n.val()) > 0 && dot(dir, ln) < 0 && pdf != 0) {
            double dist = norm(dir);
            dir = dir / dist;
            ItsAD<N> nits;
            if (rayIntersectAD(its.p, DV<N>(dir), true, nits) && nits.shape == sc.emitters[l]->shape_id &&
                std::abs(nits.t.v - dist) < SHADOW_EPS) {
                pdf *= dist * dist / dot(ln, -dir);
                wo = its.sh.toLocal(DV<N>(dir));
                return emitterEvalNormAD<N>(*sc.emitters[l], nits.sh.n, DV<N>(-dir)) / pdf;
            }
        }
        return DS<N>();
    }

    // pathAD.cpp:64-214 (DIRECT: directAD.cpp:5-98) — surface-only forward-mode
    // estimators; media are ignored. `its` is the hit of the incoming ray direction rd.
    DS<N> LiSurfaceAD(const ItsAD<N> &its, Pcg &s, const DV<N> &rd, int max_depth, int mode, bool incEmission, int nEdge,
                      bool direct) const {
        DS<N> ret;
        if (its.shape < 0) return ret;
        const OShape &shp = shape(its.shape);
        if (shp.isEmitter() && (mode & 1) && incEmission) ret += emitterEvalAD<N>(*sc.emitters[shp.light_id], its, -rd);
        if (max_depth == 0) return ret;
        const OBsdf &b = bsdfOf(its.shape);
        Its its0 = its.toIts();
        if (mode & 2) {
            double rnd4[4] = {s.next1D(), s.next1D(), s.next1D(), s.next1D()};
            DV<N> wo;
            double pdf_nee;
            DS<N> value = sampleEmitterDirectAD(its, rnd4, wo, pdf_nee);
            if (!value.isZero(EPS)) {
                DS<N> bv = bsdfEvalAD<N>(b, its, wo);
                double bp = bsdfPdf(b, its0, wo.val());
                ret += value * bv * (direct ? pdf_nee / (pdf_nee + bp) : mis(pdf_nee, bp));
            }
        }
        {
            double rnd3[3] = {s.next1D(), s.next1D(), s.next1D()};
            DV<N> wo;
            double bpdf, beta;
            DS<N> value = bsdfSampleAD<N>(b, its, rnd3, wo, bpdf, beta);
            if (!value.isZero(EPS)) {
                DV<N> nd = b.isNull() ? rd : its.sh.toWorld(wo);
                ItsAD<N> its1;
                if (rayIntersectAD(its.p, nd, true, its1)) {
                    const OShape &s1 = shape(its1.shape);
                    if (s1.isEmitter() && (mode & 2)) {
                        DS<N> lc = emitterEvalAD<N>(*sc.emitters[s1.light_id], its1, -nd);
                        if (!lc.val().isZero()) {
                            V3 dd = its1.p.val() - its.p.val();
                            double G = its1.wi.val().z / dot(dd, dd);
                            double pn = sc.light_pmf[s1.light_id] / sc.light_areas[s1.light_id] / G;
                            ret += value * lc * (direct ? bpdf / (pn + bpdf) : mis(bpdf, pn));
                        }
                    }
                    if (!direct) ret += value * LiSurfaceAD(its1, s, nd, max_depth - 1, mode, false, nEdge, false);
                }
            }
        }
        if (nEdge > 0 && (mode & 4) && sc.edges_distrb.sum > 0) {
            double der[N][3] = {};
            const DV<N> &x1 = its.p;
            for (int i = 0; i < nEdge; ++i) {
                double rnd = s.next1D();
                int sid;
                double edgePdf;
                const Edge *e = sampleEdge(rnd, sid, edgePdf);
                double t = rnd;
                const OShape &esh = shape(sid);
                if (esh.isSilhouette(*e, x1.val()) <= 0) continue;
                DV<N> v0 = esh.vertexAD<N>(e->v0), v1 = esh.vertexAD<N>(e->v1);
                DV<N> w1;
                V3 nrm;
                double dist, geo = 1.0;
                if (direct) {   // scene.h:103-108: pdf per unit length, linear in t
                    edgePdf /= e->length;
                    DV<N> y1 = v0 * (1.0 - t) + v1 * t;
                    w1 = y1 - x1;
                    nrm = normalized(cross(v0.val() - x1.val(), v1.val() - x1.val()));
                    D<N> dAD = dnorm(w1);
                    w1 = w1 / dAD;
                    dist = dAD.v;
                    double ct = -dot(w1.val(), normalized(v1.val() - v0.val()));
                    geo = std::sqrt(1.0 - ct * ct) / dist;
                } else {
                    DV<N> e0 = v0 - x1, e1 = v1 - x1;
                    nrm = normalized(cross(e0.val(), e1.val()));
                    V3 ne0 = normalized(e0.val()), ne1 = normalized(e1.val());
                    double theta = std::acos(dot(ne0, ne1)), sinTheta = std::sin(theta);
                    if (theta < EPS || theta > PI - EPS) continue;
                    double w0 = std::sin(theta * (1.0 - t)) / sinTheta, ww1 = std::sin(theta * t) / sinTheta;
                    w0 /= norm(e0.val());
                    ww1 /= norm(e1.val());
                    double tmp = w0 + ww1;
                    w0 /= tmp;
                    ww1 /= tmp;
                    if (w0 < -EPS || ww1 < -EPS || std::abs(w0 + ww1 - 1.0) > EPS) { g_counters.bad_edge_weight++; continue; }
                    w1 = e0 * w0 + e1 * ww1;
                    D<N> dAD = dnorm(w1);
                    w1 = w1 / dAD;
                    dist = dAD.v;
                    edgePdf /= theta;
                    double tmpo = 0;
                    for (int j = 0; j < N; ++j) tmpo = std::max(tmpo, std::abs(dot(w1.grad(j), w1.val())));
                    if (tmpo > EPS) { g_counters.nonortho++; continue; }
                    if (w1.gradZero(EPS)) continue;
                }
                Spec bsdfVal = bsdfEval(b, its0, its0.sh.toLocal(w1.val()));
                if (bsdfVal.isZero(EPS)) continue;
                int hs, hk;   // scene.cpp:150-176
                if (sc.intersect(x1.val(), w1.val(), SHADOW_EPS, (1.0 - SHADOW_EPS) * dist, hs, hk)) continue;
                Spec delta;
                for (int sgn = 0; sgn < 2; ++sgn) {
                    V3 w2 = normalized(sgn == 0 ? w1.val() - ANGLE_EPS * nrm : w1.val() + ANGLE_EPS * nrm);
                    Spec v;
                    if (direct) {
                        Its i1;
                        if (rayIntersect(x1.val(), w2, true, i1) && shape(i1.shape).isEmitter())
                            v = emitterEvalIts(*sc.emitters[shape(i1.shape).light_id], i1, -w1.val());
                    } else {
                        v = LiPath(s, x1.val(), w2, max_depth - 1);
                    }
                    if (sgn == 0) delta += v; else delta -= v;
                }
                if (delta.isZero(EPS)) continue;
                for (int j = 0; j < N; ++j) {
                    double ndot = dot(nrm, w1.grad(j));
                    for (int c = 0; c < 3; ++c) der[j][c] += ndot * bsdfVal[c] * delta[c] * geo / edgePdf;
                }
            }
            for (int j = 0; j < N; ++j)
                for (int c = 0; c < 3; ++c) ret.c[c].g[j] += der[j][c] / nEdge;
        }
        return ret;
    }

    int adKind = 0;   // 0 VolPathTracerAD, 1 PathTracerAD, 2 DirectAD

    // volpathAD.cpp:623-629; pathAD.cpp:216-219 / directAD (the path tracer's radiance,
    // depth 1 for directAD)
    Spec pixelColor(Pcg &s, int max_bounces, double x, double y) const {
        V3 o, d;
        sc.cam->samplePrimaryRay(x, y, o, d);
        if (adKind) return LiPath(s, o, d, adKind == 2 ? 1 : max_bounces);
        return Li(s, o, d, max_bounces, sc.cam->med_id, true);
    }
    // volpathAD.cpp:632-643
    DS<N> pixelColorAD(Pcg &s, int max_bounces, int mode, int nEdge, double ddist, double x, double y) const {
        if (adKind) {   // pathAD.cpp:222-227 (camera ray from 0); directAD.cpp:11 (from ShadowEpsilon)
            DV<N> o, d;
            sc.cam->samplePrimaryRayAD<N>(x, y, o, d);
            ItsAD<N> its;
            if (!rayIntersectAD(o, d, adKind == 2, its)) return DS<N>();
            return LiSurfaceAD(its, s, d, adKind == 2 ? 1 : max_bounces, adKind == 2 ? 0x3F : (mode >= 0 ? mode : 0x3F),
                               true, nEdge, adKind == 2);
        }
        DV<N> o, d;
        sc.cam->samplePrimaryRayAD<N>(x, y, o, d);
        ItsAD<N> its;
        if (rayIntersectAD(o, d, false, its))
            return LiAD(its, s, o, d, sc.cam->med_id, max_bounces, mode >= 0 ? mode : 0x3F, nEdge, ddist);
        return DS<N>();
    }

    // edge_manager.cpp:112-134
    double samplePrimaryEdge(double rnd, int &px, int &py, V2 &p, V2 *pd, V2 &nrm) const {
        const OCamera &cam = *sc.cam;
        int idx = (int)sc.primary_distrb.sampleReuse(rnd);
        const PrimaryEdge &pe = sc.primary_edges[idx];
        V2 pv{pe.v0s.x + rnd * (pe.v1s.x - pe.v0s.x), pe.v0s.y + rnd * (pe.v1s.y - pe.v0s.y)};
        for (int j = 0; j < N; ++j)
            pd[j] = V2{pe.v0s_d[j].x + rnd * (pe.v1s_d[j].x - pe.v0s_d[j].x), pe.v0s_d[j].y + rnd * (pe.v1s_d[j].y - pe.v0s_d[j].y)};
        double fov = cam.c2ndc[0][0], aspect = double(cam.width) / double(cam.height);
        V3 tmp((pv.x / cam.width - 0.5) * 2.0 / fov, (pv.y / cam.height - 0.5) * (-2.0) / (aspect * fov), 1.0);
        px = (int)std::floor(pv.x);
        py = (int)std::floor(pv.y);
        nrm = V2{pe.v1s.y - pe.v0s.y, pe.v0s.x - pe.v1s.x};
        double l = std::hypot(nrm.x, nrm.y);
        nrm.x /= l; nrm.y /= l;
        p = V2{pv.x - px, pv.y - py};
        double t0 = norm(tmp - pe.v0p) / norm(pe.v1p - pe.v0p);
        double t1 = computeIntersectionInTri(V3(), pe.v0p, pe.v1p, pe.v0c, pe.v1c, t0);
        return norm(pe.v0c + t1 * (pe.v1c - pe.v0c));
    }

    // volpath.cpp:106-141 — VolPathTracer::render (simple: volpath_simple.cpp:89-124,
    // VolPathTracerSimple::render): one stream per pixel, a jittered camera ray per
    // sample, the mean written to float32[H][W][3] (crop size if cropped).
    // kind: 1 VolPathTracer, 2 VolPathTracerSimple, 3 PathTracer (path.cpp:64-96),
    // 4 DirectIntegrator (direct.cpp:51-85, = PathTracer at depth 1)
    void renderPrimal(uint64_t seed, int spp, int max_bounces, float *out, int kind = 1) const {
        const OCamera &cam = *sc.cam;
        const bool cropped = cam.cropped();
        const int npix = cropped ? cam.cw * cam.ch : cam.width * cam.height;
#pragma omp parallel for schedule(dynamic, 4)
        for (int ip = 0; ip < npix; ++ip) {
            int ix = cropped ? cam.ox + ip % cam.cw : ip % cam.width;
            int iy = cropped ? cam.oy + ip / cam.cw : ip / cam.width;
            Pcg s(seed, ip);
            Spec pix;
            for (int k = 0; k < spp; ++k) {
                double rx = s.next1D(), ry = s.next1D();
                V3 o, d;
                cam.samplePrimaryRay(ix + rx, iy + ry, o, d);
                if (kind == 2) pix += LiSimple(s, o, d, max_bounces, cam.med_id);
                else if (kind >= 3) pix += LiPath(s, o, d, kind == 4 ? 1 : max_bounces);
                else pix += Li(s, o, d, max_bounces, cam.med_id, true, true);
            }
            for (int c = 0; c < 3; ++c) out[(size_t)ip * 3 + c] = (float)(pix[c] / spp);
        }
    }

    // integratorAD.cpp:19-166
    void render(uint64_t seed, int spp, int max_bounces, int sppe, int sppse, int mode, double ddist, float *out) const {
        const OCamera &cam = *sc.cam;
        const bool cropped = cam.cropped();
        const int npix = cropped ? cam.cw * cam.ch : cam.width * cam.height;
        if (spp > 0) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int ip = 0; ip < npix; ++ip) {
                int ix = cropped ? cam.ox + ip % cam.cw : ip % cam.width;
                int iy = cropped ? cam.oy + ip / cam.cw : ip / cam.width;
                Pcg s(seed, ip);
                DS<N> pix;
                for (int k = 0; k < spp; ++k) {
                    double rx = 0.5, ry = 0.5;
                    if (sppe >= 0) { rx = s.next1D(); ry = s.next1D(); }
                    pix += pixelColorAD(s, max_bounces, mode, sppse, ddist, ix + rx, iy + ry);
                }
                pix = pix / (double)spp;
                Spec v = pix.val();
                if (v.minCoeff() < 0.0) g_counters.neg_pixels++;
                else for (int c = 0; c < 3; ++c) out[(size_t)ip * 3 + c] = (float)v[c];
                for (int j = 0; j < N; ++j) {
                    double g0 = pix.c[0].g[j], g1 = pix.c[1].g[j], g2 = pix.c[2].g[j];
                    if (!std::isfinite(g0) || !std::isfinite(g1) || !std::isfinite(g2)) { g_counters.nonfinite_der++; continue; }
                    size_t off = ((size_t)(j + 1) * npix + ip) * 3;
                    out[off] = (float)g0; out[off + 1] = (float)g1; out[off + 2] = (float)g2;
                }
            }
        }
        if (sppe > 0 && !sc.primary_edges.empty()) {
            const double imageDelta = (double)1e-4f;
            const int per_block = 128;
            const int nblock = (int)std::ceil(sc.primary_distrb.sum * sppe / per_block);
            const int nsamples = nblock * per_block;
            const int med_cam = cam.med_id;
            const double pdfsum = sc.primary_distrb.sum;
            std::vector<double> acc((size_t)npix * N * 3, 0.0);
#pragma omp parallel
            {
                std::vector<double> loc((size_t)npix * N * 3, 0.0);
#pragma omp for schedule(dynamic, 1)
                for (int blk = 0; blk < nblock; ++blk) {
                    // One stream per 128-sample block (the reference keys streams by
                    // OpenMP worker id; keying by block makes the oracle deterministic).
                    Pcg s(seed, blk);
                    for (int k = 0; k < per_block; ++k) {
                        int px, py;
                        V2 p, pd[NDER_MAX], nrm;
                        double maxD = samplePrimaryEdge(s.next1D(), px, py, p, pd, nrm);
                        bool nz = false;
                        for (int j = 0; j < N; ++j) if (std::abs(pd[j].x) > EPS || std::abs(pd[j].y) > EPS) nz = true;
                        if (!nz) continue;
                        V3 ro, rd;
                        cam.samplePrimaryRay(px + p.x, py + p.y, ro, rd);
                        double trans = evalTransmittance(ro, rd, false, med_cam, maxD, s, max_bounces - 1);
                        if (trans > EPS) {
                            Spec dv = pixelColor(s, max_bounces, px + p.x - nrm.x * imageDelta, py + p.y - nrm.y * imageDelta);
                            dv -= pixelColor(s, max_bounces, px + p.x + nrm.x * imageDelta, py + p.y + nrm.y * imageDelta);
                            int ipx = cropped ? (px - cam.ox) + cam.cw * (py - cam.oy) : px + cam.width * py;
                            bool inside = cropped ? (px >= cam.ox && px < cam.ox + cam.cw && py >= cam.oy && py < cam.oy + cam.ch)
                                                  : (px >= 0 && px < cam.width && py >= 0 && py < cam.height);
                            if (!inside) continue;
                            for (int j = 0; j < N; ++j) {
                                double w = (nrm.x * pd[j].x + nrm.y * pd[j].y) * pdfsum;
                                for (int c = 0; c < 3; ++c) loc[((size_t)j * npix + ipx) * 3 + c] += w * dv[c];
                            }
                        }
                    }
                }
#pragma omp critical
                for (size_t i = 0; i < acc.size(); ++i) acc[i] += loc[i];
            }
            for (int ip = 0; ip < npix; ++ip)
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c)
                        out[((size_t)(j + 1) * npix + ip) * 3 + c] += (float)(acc[((size_t)j * npix + ip) * 3 + c] / nsamples);
        }
    }
};

}  // namespace oref
