// Host side of libdtrt_mi355x: the C-ABI declared in include/dtrt.h.
//
// Objects copy their inputs (shape.cpp:14-32, camera.cpp:13-27). Scene creation
// (scene.cpp:9-136) flattens the shapes, builds the BVH that replaces Embree, the
// light, edge and primary-edge distributions, and uploads everything to HBM once.
// A render is a sequence of kernel launches on the scene's stream.
#include "launch.h"
#include "../../include/dtrt.h"
#include "../../include/dtrt_testing.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using namespace dt;

namespace {

thread_local std::string g_err;
int g_nder = -1;

int nderNow() {
    if (g_nder < 0) {
        const char *e = std::getenv("DTRT_NDER");
        int n = e ? std::atoi(e) : 4;
        g_nder = (n == 1 || n == 2 || n == 3 || n == 4 || n == 8 || n == 16) ? n : 4;
    }
    return g_nder;
}

// Estimator precision of scenes built from now on: 64 (double, the reference's
// precision, include/fwd.h:10-13) unless DTRT_PRECISION=32 or dtrt_set_precision(32).
int g_prec = -1;
int precisionNow() {
    if (g_prec < 0) {
        const char *e = std::getenv("DTRT_PRECISION");
        g_prec = (e && std::atoi(e) == 32) ? 32 : 64;
    }
    return g_prec;
}

bool fail(const std::string &m) { g_err = m; return false; }

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);                      \
            return false;                                                                \
        }                                                                                \
    } while (0)

struct Obj {
    int tag;
    explicit Obj(int t) : tag(t) {}
    virtual ~Obj() = default;
};
enum { T_CAMERA = 1, T_SHAPE, T_BSDF, T_PHASE, T_EMITTER, T_MEDIUM, T_SCENE };

struct D3 { double x, y, z; };
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross3(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm3(D3 a) { return std::sqrt(dot3(a, a)); }

bool invert4(const double m[4][4], double o[4][4]) {
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i][j] : (j - 4 == i);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r) if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0) return false;
        for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) { double f = a[r][c]; for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j]; }
    }
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) o[i][j] = a[i][j + 4];
    return true;
}

// ------------------------------------------------------------------ objects
uint64_t g_camera_uid = 0;   // identifies a camera configuration (primary edges are cached by it)

struct Camera : Obj {
    Camera() : Obj(T_CAMERA), uid(++g_camera_uid) {}
    uint64_t uid;
    int w = 0, h = 0, ox = 0, oy = 0, cw = -1, ch = -1, med = -1, nder = 0;
    float clip_near = 0;
    double c2w[4][4], w2c[4][4], c2ndc[3][3], ndc2c[3][3];
    D3 cpos, cs, ct, cn, cpos_d[DTRT_NDER_MAX], cs_d[DTRT_NDER_MAX], ct_d[DTRT_NDER_MAX], cn_d[DTRT_NDER_MAX];
    bool cropped() const { return cw > 0 && ch > 0; }
};

struct EdgeH { int v0, v1, f0, f1; double len; };

struct Shape : Obj {
    Shape() : Obj(T_SHAPE) {}
    int nv = 0, nf = 0, nder = 0, light_id = -1, bsdf_id = -1, med_int = -1, med_ext = -1;
    std::vector<D3> v, vd, n, nd, fn, fnd;   // tangents entity-major: vd[i*nder + j]
    std::vector<int> f;
    std::vector<EdgeH> edges;
    std::vector<double> edge_cdf;            // normalised, ne+1 entries
    double edge_sum = 0;
    bool hasNormals() const { return !n.empty(); }
};

struct Bsdf : Obj { Bsdf() : Obj(T_BSDF) { std::memset(&d, 0, sizeof(d)); } DBsdf d; };
struct Phase : Obj { Phase() : Obj(T_PHASE) { std::memset(&d, 0, sizeof(d)); } DPhase d; };
struct Emitter : Obj { Emitter() : Obj(T_EMITTER) { std::memset(&d, 0, sizeof(d)); } DEmitter d; };
struct Medium : Obj {
    Medium() : Obj(T_MEDIUM) { std::memset(&d, 0, sizeof(d)); }
    DMedium d;
    std::vector<float> density, albedo;
};

// shape.cpp:109-126 — face normals and their tangents (double)
// shape.cpp:37-51 — vertex tangents [nder][nv][3]; with normals, the normal tangents
// follow as one [nv][3] block that every lane reads (shape.cpp:45, no lane offset)
void setShapeVelocities(Shape &s, const float *vel) {
    const int nd = s.nder, nv = s.nv;
    for (int j = 0; j < nd; ++j)
        for (int i = 0; i < nv; ++i) {
            const float *p = vel + ((size_t)j * nv + i) * 3;
            s.vd[(size_t)i * nd + j] = D3{p[0], p[1], p[2]};
            if (s.hasNormals()) {
                const float *q = vel + (size_t)nv * 3 * nd + (size_t)i * 3;
                s.nd[(size_t)i * nd + j] = D3{q[0], q[1], q[2]};
            }
        }
}

void computeFaceNormals(Shape &s) {
    s.fn.assign(s.nf, D3{0, 0, 0});
    s.fnd.assign((size_t)s.nf * s.nder, D3{0, 0, 0});
    for (int k = 0; k < s.nf; ++k) {
        int a = s.f[3 * k], b = s.f[3 * k + 1], c = s.f[3 * k + 2];
        D3 e1 = s.v[b] - s.v[a], e2 = s.v[c] - s.v[a], cr = cross3(e1, e2);
        double len = norm3(cr);
        for (int j = 0; j < s.nder; ++j) {
            D3 de1 = s.vd[(size_t)b * s.nder + j] - s.vd[(size_t)a * s.nder + j];
            D3 de2 = s.vd[(size_t)c * s.nder + j] - s.vd[(size_t)a * s.nder + j];
            D3 dc = cross3(de1, e2) + cross3(e1, de2);
            if (len < 1e-7) s.fnd[(size_t)k * s.nder + j] = dc;
            else {
                D3 nn = cr * (1.0 / len);
                s.fnd[(size_t)k * s.nder + j] = (dc - nn * dot3(nn, dc)) * (1.0 / len);
            }
        }
        s.fn[k] = len < 1e-7 ? cr : cr * (1.0 / len);
    }
}

// shape.cpp:254-300 — edges ordered by (min vertex, max vertex); crease or boundary only
bool constructEdges(Shape &s) {
    std::vector<std::array<int, 3>> ev;
    ev.reserve((size_t)s.nf * 3);
    for (int k = 0; k < s.nf; ++k)
        for (int e = 0; e < 3; ++e) {
            int a = s.f[3 * k + e], b = s.f[3 * k + (e + 1) % 3];
            ev.push_back({std::min(a, b), std::max(a, b), k});
        }
    std::sort(ev.begin(), ev.end());
    s.edges.clear();
    for (size_t i = 0; i < ev.size();) {
        size_t j = i;
        while (j < ev.size() && ev[j][0] == ev[i][0] && ev[j][1] == ev[i][1]) ++j;
        size_t cnt = j - i;
        int a = ev[i][0], b = ev[i][1];
        double len = norm3(s.v[a] - s.v[b]);
        if (cnt > 2) return fail("shape: every edge can be shared by at most 2 faces");
        if (cnt == 2) {
            int f0 = ev[i][2], f1 = ev[i + 1][2];
            if (f0 == f1) return fail("shape: duplicated faces");
            double val = dot3(s.fn[f0], s.fn[f1]);
            if (val < -1.0 + 1e-7) return fail("shape: inconsistent normal orientation");
            if (val < 1.0 - 1e-7) s.edges.push_back({a, b, f0, f1, len});
        } else {
            s.edges.push_back({a, b, ev[i][2], -1, len});
        }
        i = j;
    }
    s.edge_cdf.assign(s.edges.size() + 1, 0.0);
    for (size_t e = 0; e < s.edges.size(); ++e) s.edge_cdf[e + 1] = s.edge_cdf[e] + s.edges[e].len;
    s.edge_sum = s.edge_cdf.back();
    if (s.edge_sum > 0) {
        for (double &c : s.edge_cdf) c /= s.edge_sum;
        s.edge_cdf.back() = 1.0;
    }
    return true;
}

// ---------------------------------------------------------------- BVH build
// Binned-SAH binary BVH; GPU nodes hold both children's boxes (64 B) and leaves
// are encoded in the child slot as ~((first << 4) | count), count <= 15.
struct BuildNode { float lo[3], hi[3]; int left = -1, right = -1, first = 0, count = 0; };

struct BvhOut {
    std::vector<float4> nodes, tris;
    int depth = 0;
};

static float boxArea(const float *l, const float *h) {
    float dx = std::max(0.f, h[0] - l[0]), dy = std::max(0.f, h[1] - l[1]), dz = std::max(0.f, h[2] - l[2]);
    return dx * dy + dy * dz + dz * dx;
}

bool buildBvh(const std::vector<std::array<float, 9>> &tri, const std::vector<int> &shape_of, BvhOut &out) {
    const int n = (int)tri.size();
    std::vector<float> cen((size_t)n * 3), lo((size_t)n * 3), hi((size_t)n * 3);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float x0 = tri[i][a], x1 = tri[i][3 + a], x2 = tri[i][6 + a];
            lo[3 * i + a] = std::min(x0, std::min(x1, x2));
            hi[3 * i + a] = std::max(x0, std::max(x1, x2));
            cen[3 * i + a] = 0.5f * (lo[3 * i + a] + hi[3 * i + a]);
        }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    // Enclosing shapes. SAH prices a split by the chance that a ray from outside the
    // node enters each child; paths start inside the scene, and a shape that encloses
    // the others (an environment sphere, a medium's boundary box around its contents)
    // is then hit only after everything it holds has been tested. Its triangles are
    // split off as the node's second child when its box holds the boxes of all other
    // shapes in the node and has at least 4 times their union's area; traversal visits
    // the first child first when both entries tie (origins inside both boxes).
    const int nshape = shape_of.empty() ? 0 : 1 + *std::max_element(shape_of.begin(), shape_of.end());
    std::vector<float> sbox((size_t)nshape * 6);
    std::vector<int> present;
    auto enclosingShape = [&](int first, int count) -> int {
        const int s0 = shape_of[idx[first]];
        int i = first;
        while (i < first + count && shape_of[idx[i]] == s0) ++i;
        if (i == first + count) return -1;   // one shape only
        present.clear();
        for (int k = first; k < first + count; ++k) {
            const int p = idx[k], sh = shape_of[p];
            float *b = &sbox[(size_t)sh * 6];
            if (std::find(present.begin(), present.end(), sh) == present.end()) {
                present.push_back(sh);
                for (int a = 0; a < 3; ++a) { b[a] = INFINITY; b[3 + a] = -INFINITY; }
            }
            for (int a = 0; a < 3; ++a) { b[a] = std::min(b[a], lo[3 * p + a]); b[3 + a] = std::max(b[3 + a], hi[3 * p + a]); }
            if (present.size() > 64) return -1;   // many small shapes: leave it to SAH
        }
        int best = -1;
        float bestArea = 0.f;
        for (int sh : present) {
            float ul[3] = {INFINITY, INFINITY, INFINITY}, uh[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int o : present)
                if (o != sh)
                    for (int a = 0; a < 3; ++a) {
                        ul[a] = std::min(ul[a], sbox[(size_t)o * 6 + a]);
                        uh[a] = std::max(uh[a], sbox[(size_t)o * 6 + 3 + a]);
                    }
            const float *b = &sbox[(size_t)sh * 6];
            bool holds = true;
            for (int a = 0; a < 3; ++a) holds = holds && b[a] <= ul[a] && b[3 + a] >= uh[a];
            const float sa = boxArea(b, b + 3);
            if (holds && sa >= 4.f * boxArea(ul, uh) && sa > bestArea) { best = sh; bestArea = sa; }
        }
        return best;
    };
    std::vector<BuildNode> bn;
    bn.reserve(2 * (size_t)n + 1);
    struct Task { int node, first, count, depth, shell; float cx, cy, cz; };
    bn.push_back(BuildNode());
    std::vector<Task> st{{0, 0, n, 1, 0, 0.f, 0.f, 0.f}};
    const int NB = 16, MAXLEAF = 15;
    int maxDepth = 0;
    auto area = [](const float *l, const float *h) {
        float dx = std::max(0.f, h[0] - l[0]), dy = std::max(0.f, h[1] - l[1]), dz = std::max(0.f, h[2] - l[2]);
        return dx * dy + dy * dz + dz * dx;
    };
    while (!st.empty()) {
        Task t = st.back();
        st.pop_back();
        maxDepth = std::max(maxDepth, t.depth);
        float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = t.first; i < t.first + t.count; ++i) {
            int p = idx[i];
            for (int a = 0; a < 3; ++a) {
                nlo[a] = std::min(nlo[a], lo[3 * p + a]); nhi[a] = std::max(nhi[a], hi[3 * p + a]);
                clo[a] = std::min(clo[a], cen[3 * p + a]); chi[a] = std::max(chi[a], cen[3 * p + a]);
            }
        }
        BuildNode &nd = bn[t.node];
        std::memcpy(nd.lo, nlo, sizeof(nlo));
        std::memcpy(nd.hi, nhi, sizeof(nhi));
        bool mustSplit = t.count > MAXLEAF;
        // nodes of at most 2 triangles are leaves (measured: 2 beats 1, 4 and 8 on the
        // medium, dielectric and surface scenes)
        if (t.count <= 2 && !mustSplit) { nd.first = t.first; nd.count = t.count; continue; }
        // binned SAH over the three axes
        int bestAxis = -1, bestSplit = -1;
        float bestCost = INFINITY;
        for (int a = 0; a < 3; ++a) {
            float ext = chi[a] - clo[a];
            if (!(ext > 0)) continue;
            int cnt[NB] = {0};
            float blo[NB][3], bhi[NB][3];
            for (int b = 0; b < NB; ++b) for (int c = 0; c < 3; ++c) { blo[b][c] = INFINITY; bhi[b][c] = -INFINITY; }
            for (int i = t.first; i < t.first + t.count; ++i) {
                int p = idx[i];
                int b = std::min(NB - 1, (int)((cen[3 * p + a] - clo[a]) / ext * NB));
                cnt[b]++;
                for (int c = 0; c < 3; ++c) { blo[b][c] = std::min(blo[b][c], lo[3 * p + c]); bhi[b][c] = std::max(bhi[b][c], hi[3 * p + c]); }
            }
            float rA[NB], rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int rC[NB], rc = 0;
            for (int b = NB - 1; b > 0; --b) {
                rc += cnt[b];
                for (int c = 0; c < 3; ++c) { rl[c] = std::min(rl[c], blo[b][c]); rh[c] = std::max(rh[c], bhi[b][c]); }
                rA[b] = rc ? area(rl, rh) : 0.f;
                rC[b] = rc;
            }
            float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int lc = 0;
            for (int b = 0; b < NB - 1; ++b) {
                lc += cnt[b];
                for (int c = 0; c < 3; ++c) { ll[c] = std::min(ll[c], blo[b][c]); lh[c] = std::max(lh[c], bhi[b][c]); }
                if (lc == 0 || rC[b + 1] == 0) continue;
                float cost = area(ll, lh) * lc + rA[b + 1] * rC[b + 1];
                if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = b; }
            }
        }
        float leafCost = area(nlo, nhi) * t.count;
        // levels a balanced split of this node still needs before its leaves fit
        int need = 0;
        for (int c = t.count; c > MAXLEAF; c = (c + 1) / 2) ++need;
        const bool tight = t.depth + need >= DTRT_BVH_STACK - 2;
        if (tight && t.depth + need > DTRT_BVH_STACK - 2) return fail("bvh: too many triangles for the traversal stack");
        int mid;
        const int encl = tight || shape_of.empty() || t.shell ? -1 : enclosingShape(t.first, t.count);
        // Below an enclosing shape's split (of at least 64 triangles), its triangles are
        // split by direction from the centre of what it encloses (signed dominant axis of
        // centroid - centre, six classes) until one class is left: the boxes of such
        // groups keep clear of the interior, where the paths start, unlike SAH's planes
        // through the centre.
        auto dirClass = [&](int p) {
            const float d[3] = {cen[3 * p] - t.cx, cen[3 * p + 1] - t.cy, cen[3 * p + 2] - t.cz};
            int a = 0;
            for (int k = 1; k < 3; ++k)
                if (std::fabs(d[k]) > std::fabs(d[a])) a = k;
            return 2 * a + (d[a] < 0.f ? 1 : 0);
        };
        unsigned classes = 0;
        if (t.shell && !tight)
            for (int k = t.first; k < t.first + t.count; ++k) classes |= 1u << dirClass(idx[k]);
        int childShell = 0;   // 1: children keep the direction split, 2: second child starts it
        float ccx = t.cx, ccy = t.cy, ccz = t.cz;
        if (__builtin_popcount(classes) > 1) {   // half of the present classes to each side
            int cls[6], nc = 0;
            for (int c = 0; c < 6; ++c)
                if (classes >> c & 1) cls[nc++] = c;
            unsigned left = 0;
            for (int i = 0; i < nc / 2; ++i) left |= 1u << cls[i];
            mid = (int)(std::partition(idx.begin() + t.first, idx.begin() + t.first + t.count,
                                       [&](int p) { return (left >> dirClass(p) & 1) != 0; }) - idx.begin());
            childShell = 1;
        } else if (encl >= 0) {   // the contents first, the enclosing shape second
            mid = (int)(std::partition(idx.begin() + t.first, idx.begin() + t.first + t.count,
                                       [&](int p) { return shape_of[p] != encl; }) - idx.begin());
            float ul[3] = {INFINITY, INFINITY, INFINITY}, uh[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int k = t.first; k < mid; ++k)
                for (int a = 0; a < 3; ++a) {
                    ul[a] = std::min(ul[a], lo[3 * idx[k] + a]);
                    uh[a] = std::max(uh[a], hi[3 * idx[k] + a]);
                }
            ccx = 0.5f * (ul[0] + uh[0]);
            ccy = 0.5f * (ul[1] + uh[1]);
            ccz = 0.5f * (ul[2] + uh[2]);
            childShell = t.first + t.count - mid >= 64 ? 2 : 0;   // small shells: SAH is as good
        } else if (tight) {
            // depth budget exhausted: a node that fits a leaf becomes one (a split here
            // could leave a child too large for the last level), a larger one is split at
            // the median along the widest centroid axis, which keeps every leaf below at
            // depth <= DTRT_BVH_STACK - 2
            if (!mustSplit) { nd.first = t.first; nd.count = t.count; continue; }
            int ax = 0;
            for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
            mid = t.first + (t.count + 1) / 2;
            std::nth_element(idx.begin() + t.first, idx.begin() + mid, idx.begin() + t.first + t.count,
                             [&](int p, int q) { return cen[3 * p + ax] < cen[3 * q + ax]; });
        } else if (bestAxis < 0 || (!mustSplit && bestCost >= leafCost)) {
            if (!mustSplit) { nd.first = t.first; nd.count = t.count; continue; }
            mid = t.first + (t.count + 1) / 2;   // identical centroids: split by index
        } else {
            float ext = chi[bestAxis] - clo[bestAxis];
            auto it = std::partition(idx.begin() + t.first, idx.begin() + t.first + t.count, [&](int p) {
                return std::min(NB - 1, (int)((cen[3 * p + bestAxis] - clo[bestAxis]) / ext * NB)) <= bestSplit;
            });
            mid = (int)(it - idx.begin());
            if (mid == t.first || mid == t.first + t.count) mid = t.first + (t.count + 1) / 2;
        }
        int l = (int)bn.size();
        bn.push_back(BuildNode());
        bn.push_back(BuildNode());
        bn[t.node].left = l;
        bn[t.node].right = l + 1;
        st.push_back({l, t.first, mid - t.first, t.depth + 1, childShell == 1, ccx, ccy, ccz});
        st.push_back({l + 1, mid, t.first + t.count - mid, t.depth + 1, childShell != 0, ccx, ccy, ccz});
    }
    // triangles in leaf order
    out.tris.resize((size_t)n * 3);
    for (int i = 0; i < n; ++i) {
        const auto &T = tri[idx[i]];
        float4 v0{T[0], T[1], T[2], 0.f};
        int fid = idx[i];
        std::memcpy(&v0.w, &fid, 4);
        out.tris[3 * (size_t)i] = v0;
        out.tris[3 * (size_t)i + 1] = float4{T[3] - T[0], T[4] - T[1], T[5] - T[2], 0.f};
        out.tris[3 * (size_t)i + 2] = float4{T[6] - T[0], T[7] - T[1], T[8] - T[2], 0.f};
    }
    // GPU nodes: inner build nodes only; a leaf root gets a wrapper node
    std::vector<int> gpuIndex(bn.size(), -1);
    std::vector<int> order;
    if (bn[0].left < 0) {
        out.nodes.resize(4);
        const BuildNode &r = bn[0];
        int leaf = ~((r.first << 4) | r.count), empty = ~0;
        int il, ir;
        std::memcpy(&il, &leaf, 4);
        std::memcpy(&ir, &empty, 4);
        float fl, fr;
        std::memcpy(&fl, &leaf, 4);
        std::memcpy(&fr, &empty, 4);
        out.nodes[0] = float4{r.lo[0], r.lo[1], r.lo[2], fl};
        out.nodes[1] = float4{r.hi[0], r.hi[1], r.hi[2], fr};
        out.nodes[2] = float4{1.f, 1.f, 1.f, 0.f};
        out.nodes[3] = float4{-1.f, -1.f, -1.f, 0.f};
        out.depth = 1;
        return true;
    }
    // Inner nodes are numbered breadth-first for the first BVH_TOP_NODES (the top levels,
    // which a launch may keep in LDS as a prefix of the node array), then depth-first
    // (preorder) below each remaining subtree root in that order, which keeps a subtree
    // together in memory.
    constexpr size_t BVH_TOP_NODES = 1024;
    std::vector<int> level{0};
    size_t head = 0;
    while (head < level.size() && order.size() < BVH_TOP_NODES) {
        const int b = level[head++];
        gpuIndex[b] = (int)order.size();
        order.push_back(b);
        for (int c : {bn[b].left, bn[b].right})
            if (bn[c].left >= 0) level.push_back(c);
    }
    std::vector<int> stack;
    for (size_t k = head; k < level.size(); ++k) {
        stack.push_back(level[k]);
        while (!stack.empty()) {
            const int b = stack.back();
            stack.pop_back();
            gpuIndex[b] = (int)order.size();
            order.push_back(b);
            for (int c : {bn[b].right, bn[b].left})
                if (bn[c].left >= 0) stack.push_back(c);
        }
    }
    out.nodes.resize(order.size() * 4);
    auto enc = [&](int c) {
        const BuildNode &x = bn[c];
        int v = x.left >= 0 ? gpuIndex[c] : ~((x.first << 4) | x.count);
        float f;
        std::memcpy(&f, &v, 4);
        return f;
    };
    for (size_t g = 0; g < order.size(); ++g) {
        const BuildNode &p = bn[order[g]], &L = bn[p.left], &R = bn[p.right];
        out.nodes[4 * g] = float4{L.lo[0], L.lo[1], L.lo[2], enc(p.left)};
        out.nodes[4 * g + 1] = float4{L.hi[0], L.hi[1], L.hi[2], enc(p.right)};
        out.nodes[4 * g + 2] = float4{R.lo[0], R.lo[1], R.lo[2], 0.f};
        out.nodes[4 * g + 3] = float4{R.hi[0], R.hi[1], R.hi[2], 0.f};
    }
    out.depth = maxDepth;
    return true;
}

// ------------------------------------------------------------ primary edges
// edge_manager.cpp:31-110 with line_clip.h:37-95, in double on the host.
struct P2 { double x, y; double dx[DTRT_NDER_MAX], dy[DTRT_NDER_MAX]; };
struct PEdge { P2 s0, s1; D3 p0, p1, c0, c1; };

struct ADV { D3 v; D3 g[DTRT_NDER_MAX]; };

int outcode(double x, double y, double xmin, double xmax, double ymin, double ymax) {
    int c = 0;
    if (x < xmin - 1e-7) c |= 1; else if (x > xmax + 1e-7) c |= 2;
    if (y < ymin - 1e-7) c |= 4; else if (y > ymax + 1e-7) c |= 8;
    return c;
}
P2 lerp2(const P2 &a, const P2 &b, double s, int nd) {
    P2 r;
    r.x = a.x + (b.x - a.x) * s;
    r.y = a.y + (b.y - a.y) * s;
    for (int j = 0; j < nd; ++j) { r.dx[j] = a.dx[j] + (b.dx[j] - a.dx[j]) * s; r.dy[j] = a.dy[j] + (b.dy[j] - a.dy[j]) * s; }
    return r;
}
bool clipSeg(P2 a, P2 b, P2 &c0, P2 &c1, double xmin, double xmax, double ymin, double ymax, int nd) {
    int o0 = outcode(a.x, a.y, xmin, xmax, ymin, ymax), o1 = outcode(b.x, b.y, xmin, xmax, ymin, ymax);
    c0 = a;
    c1 = b;
    for (int it = 0; it < 64; ++it) {
        if (!(o0 | o1)) return true;
        if (o0 & o1) return false;
        int oo = o0 ? o0 : o1;
        P2 p;
        if (oo & 8) p = lerp2(c0, c1, (ymax - c0.y) / (c1.y - c0.y), nd);
        else if (oo & 4) p = lerp2(c0, c1, (ymin - c0.y) / (c1.y - c0.y), nd);
        else if (oo & 2) p = lerp2(c0, c1, (xmax - c0.x) / (c1.x - c0.x), nd);
        else p = lerp2(c0, c1, (xmin - c0.x) / (c1.x - c0.x), nd);
        if (oo == o0) { c0 = p; o0 = outcode(c0.x, c0.y, xmin, xmax, ymin, ymax); }
        else { c1 = p; o1 = outcode(c1.x, c1.y, xmin, xmax, ymin, ymax); }
    }
    return false;
}
double inTri(D3 b0, D3 c0, D3 b1, D3 c1, double t0) {   // utils.cpp:256-266 with a = origin
    double cb = norm3(b0) / norm3(b1), cc = norm3(c0) / norm3(c1);
    double bx = (1 - t0) * cb, by = t0 * cc;
    return by / (bx + by);
}

void projectEdge(const Camera &cam, ADV a, ADV b, int nd, std::vector<PEdge> &out, std::vector<double> &w) {
    const D3 n = cam.cn;
    double z0 = dot3(a.v - cam.cpos, n), z1 = dot3(b.v - cam.cpos, n), cn = cam.clip_near;
    auto addScaled = [&](ADV &x, const ADV &y, const ADV &z, double s) {   // x += s * (y - z)
        x.v = x.v + (y.v - z.v) * s;
        for (int j = 0; j < nd; ++j) x.g[j] = x.g[j] + (y.g[j] - z.g[j]) * s;
    };
    if (z0 < cn && z1 < cn) return;
    if (z0 < cn) { ADV a0 = a; addScaled(a, b, a0, (cn - z0) / (z1 - z0)); }
    else if (z1 < cn) { ADV b0 = b; addScaled(b, a, b0, (cn - z1) / (z0 - z1)); }
    auto xfm = [&](D3 p) {
        double hh[4];
        for (int i = 0; i < 4; ++i) hh[i] = cam.w2c[i][0] * p.x + cam.w2c[i][1] * p.y + cam.w2c[i][2] * p.z + cam.w2c[i][3];
        return D3{hh[0] / hh[3], hh[1] / hh[3], hh[2] / hh[3]};
    };
    D3 l0 = xfm(a.v), l1 = xfm(b.v);
    // perspective division and camera-frame coordinates, with tangents from the
    // vertex motion and the camera pose tangents
    auto toPlane = [&](const ADV &p, D3 &pl, D3 *plg) {
        D3 q = p.v - cam.cpos;
        double z = dot3(q, n);
        D3 r = q * (1.0 / z);
        D3 loc{dot3(r, cam.cs), dot3(r, cam.ct), dot3(r, cam.cn)};
        pl = loc;
        for (int j = 0; j < nd; ++j) {
            D3 dq = p.g[j] - cam.cpos_d[j];
            double dz = dot3(dq, n) + dot3(q, cam.cn_d[j]);
            D3 dr = (dq - r * dz) * (1.0 / z);
            plg[j] = D3{dot3(dr, cam.cs) + dot3(r, cam.cs_d[j]), dot3(dr, cam.ct) + dot3(r, cam.ct_d[j]),
                        dot3(dr, cam.cn) + dot3(r, cam.cn_d[j])};
        }
    };
    D3 p0, p1, g0[DTRT_NDER_MAX], g1[DTRT_NDER_MAX];
    toPlane(a, p0, g0);
    toPlane(b, p1, g1);
    double fov = cam.c2ndc[0][0], asp = double(cam.w) / double(cam.h);
    P2 r0, r1;
    r0.x = p0.x * fov * 0.5 + 0.5; r0.y = p0.y * fov * (-0.5) * asp + 0.5;
    r1.x = p1.x * fov * 0.5 + 0.5; r1.y = p1.y * fov * (-0.5) * asp + 0.5;
    for (int j = 0; j < nd; ++j) {
        r0.dx[j] = g0[j].x * fov * 0.5; r0.dy[j] = g0[j].y * fov * (-0.5) * asp;
        r1.dx[j] = g1[j].x * fov * 0.5; r1.dy[j] = g1[j].y * fov * (-0.5) * asp;
    }
    double xmin = 0, xmax = 1, ymin = 0, ymax = 1;
    if (cam.cropped()) {
        xmin = (float)cam.ox / cam.w; xmax = (float)(cam.ox + cam.cw) / cam.w;
        ymin = (float)cam.oy / cam.h; ymax = (float)(cam.oy + cam.ch) / cam.h;
    }
    P2 s0, s1;
    if (!clipSeg(r0, r1, s0, s1, xmin, xmax, ymin, ymax, nd)) return;
    D3 pp0 = p0, pp1 = p1, c0 = l0, c1 = l1;
    if (std::hypot(r0.x - s0.x, r0.y - s0.y) > 1e-7) {
        D3 pc{(s0.x - 0.5) * 2.0 / fov, (s0.y - 0.5) * (-2.0) / (asp * fov), 1.0};
        double t0 = norm3(pc - p0) / norm3(p1 - p0);
        pp0 = p0 + (p1 - p0) * t0;
        c0 = l0 + (l1 - l0) * inTri(p0, p1, l0, l1, t0);
    }
    if (std::hypot(r1.x - s1.x, r1.y - s1.y) > 1e-7) {
        D3 pc{(s1.x - 0.5) * 2.0 / fov, (s1.y - 0.5) * (-2.0) / (asp * fov), 1.0};
        double t0 = norm3(pc - p0) / norm3(p1 - p0);
        pp1 = p0 + (p1 - p0) * t0;
        c1 = l0 + (l1 - l0) * inTri(p0, p1, l0, l1, t0);
    }
    PEdge e;
    e.s0 = s0; e.s1 = s1;
    e.s0.x *= cam.w; e.s0.y *= cam.h; e.s1.x *= cam.w; e.s1.y *= cam.h;
    for (int j = 0; j < nd; ++j) { e.s0.dx[j] *= cam.w; e.s0.dy[j] *= cam.h; e.s1.dx[j] *= cam.w; e.s1.dy[j] *= cam.h; }
    e.p0 = pp0; e.p1 = pp1; e.c0 = c0; e.c1 = c1;
    out.push_back(e);
    w.push_back(std::hypot(e.s0.x - e.s1.x, e.s0.y - e.s1.y));
}

// shape.cpp:302-335 on the host (primary-edge selection from the camera position)
int silhouetteHost(const Shape &s, const EdgeH &e, D3 p) {
    if (e.f0 < 0 || e.f1 < 0) return 2;
    bool ff[2];
    int fs[2] = {e.f0, e.f1};
    for (int k = 0; k < 2; ++k) {
        const int *fi = &s.f[3 * fs[k]];
        int o = fi[0];
        if (o == e.v0 || o == e.v1) o = fi[1];
        if (o == e.v0 || o == e.v1) o = fi[2];
        D3 nn = k == 0 ? cross3(s.v[e.v0] - s.v[o], s.v[e.v1] - s.v[o]) : cross3(s.v[e.v1] - s.v[o], s.v[e.v0] - s.v[o]);
        ff[k] = dot3(nn, p - s.v[o]) > 0.0;
    }
    if (ff[0] != ff[1]) return 2;
    return s.hasNormals() ? 0 : 1;
}

// ------------------------------------------------------------------- scene
// camera.cpp:8-83 data for the device (ray generation in float, tangents per lane)
void fillDCamera(const Camera &cam, int nder, DCamera &c) {
    std::memset(&c, 0, sizeof(c));
    c.width = cam.w; c.height = cam.h; c.ox = cam.ox; c.oy = cam.oy; c.cw = cam.cw; c.ch = cam.ch;
    c.cropped = cam.cropped(); c.med_id = cam.med;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) c.c2w[4 * i + j] = (float)cam.c2w[i][j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c.ndc2c[3 * i + j] = (float)cam.ndc2c[i][j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) c.w2c[4 * i + j] = (float)cam.w2c[i][j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c.c2ndc[3 * i + j] = (float)cam.c2ndc[i][j];
    c.fov = (float)cam.c2ndc[0][0];
    c.aspect = (float)cam.w / (float)cam.h;
    c.clip_near = cam.clip_near;
    auto put = [](float *dst, D3 v) { dst[0] = (float)v.x; dst[1] = (float)v.y; dst[2] = (float)v.z; };
    put(c.cpos, cam.cpos); put(c.cs, cam.cs); put(c.ct, cam.ct); put(c.cn, cam.cn);
    for (int j = 0; j < nder; ++j) { put(c.cpos_d[j], cam.cpos_d[j]); put(c.cs_d[j], cam.cs_d[j]); put(c.ct_d[j], cam.ct_d[j]); put(c.cn_d[j], cam.cn_d[j]); }
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Bump allocation inside a caller-provided device area (the scratch-only render):
// 256-byte aligned pieces, null when the area is used up.
struct Bump {
    char *p = nullptr;
    uint64_t left = 0;
    void *take(uint64_t n) {
        n = (n + 255) / 256 * 256;
        if (n > left) return nullptr;
        void *r = p;
        p += n;
        left -= n;
        return r;
    }
};
// The scratch-only render's child queue must hold at least this much; below it the
// render is refused (the exact overflow replay would still finish, in many rounds).
constexpr uint64_t NOALLOC_MIN_QUEUE = uint64_t(16) << 20;

struct Scene : Obj {
    Scene() : Obj(T_SCENE) {}
    ~Scene() override {
        if (stream) (void)hipStreamSynchronize(stream);
        graph.reset();
        stepg.reset();
        if (stepbuf.p) (void)hipFree(stepbuf.p);
        for (void *p : pe_bufs) (void)hipFree(p);
        for (auto *b : {&edgebuf, &ecdfbuf, &lightbuf, &refitbuf, &shapecdfbuf}) if (b->p) (void)hipFree(b->p);
        for (auto &b : bufs) if (b.p) (void)hipFree(b.p);
        for (auto *b : {&acc, &acc_pe, &queue, &qcount, &stats, &work, &outbuf, &ovf[0], &ovf[1], &viewbuf, &tunebuf})
            if (b->p) (void)hipFree(b->p);
        for (auto &e : ev) if (e) (void)hipEventDestroy(e);
        if (caller_ev) (void)hipEventDestroy(caller_ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int nder = 0;
    Camera cam;
    std::vector<Shape *> shapes;
    std::vector<Bsdf *> bsdfs;
    std::vector<Emitter *> emitters;
    std::vector<Phase *> phases;
    std::vector<Medium *> media;
    DScene S;
    std::vector<DevBuf> bufs;
    DevBuf acc, acc_pe, queue, qcount, stats, work, outbuf, ovf[2], viewbuf;
    std::vector<DMedium> hmed;   // host copy of the device media records
    struct PECache { int npe = 0; double pe_sum = 0; const double *cdf = nullptr; const float *pe = nullptr; };
    std::vector<void *> pe_bufs;   // device buffers of the primary-edge cache (freed when geometry changes)
    std::vector<uint64_t> last_chunks;   // AD chunk sizes of the last eager renderViews
    uint64_t last_replayed = 0;
    // A captured launch sequence of renderViews (fixed chunk schedule), replayed with new seeds.
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        std::vector<Camera> cams;
        std::vector<dtrt_render_options> opts;
        std::vector<float *> images;
        std::vector<DView> hv;
        std::vector<void *> bufs;   // device buffers the graph was captured with
        void reset() {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
            x = nullptr;
            g = nullptr;
        }
    } graph;
    std::map<uint64_t, PECache> pecache;   // primary edges per camera configuration (Camera::uid)
    hipStream_t stream = nullptr;
    // The caller's stream (default: the legacy null stream). The scene's own stream is
    // non-blocking, so every call that reads or writes caller memory first makes it wait
    // for the work the caller has queued there (a zeroed output, a loss gradient).
    hipStream_t caller = nullptr;
    hipEvent_t caller_ev = nullptr;
    bool afterCaller() {
        if (!caller_ev) HIPCHK(hipEventCreateWithFlags(&caller_ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(caller_ev, caller));
        HIPCHK(hipStreamWaitEvent(stream, caller_ev, 0));
        return true;
    }
    std::vector<hipEvent_t> ev;
    // last render
    int64_t last_stats[10] = {0};
    double kernel_ms[4] = {0};
    int64_t kernel_launches[4] = {0};
    uint64_t work_counts[12] = {0};
    int edge_ebase_total = 0;
    size_t queue_budget = size_t(6) << 30;   // bytes of queued child records
    bool tree_in_hbm = false;                 // test setting: traversal reads the BVH from HBM only
    // launch settings chosen by autotune(): persistent child blocks per CU, the share of
    // the queue an AD chunk aims to fill, and the largest AD chunk
    int tune_child_blocks = 32;
    double tune_fill = 0.75;
    uint64_t tune_max_chunk = uint64_t(1) << 22;
    // work-group sizes of the AD-path and child kernels. The child kernel's 256 lets a
    // small scene's tree sit in LDS beside the traversal stacks at the kernel's full
    // occupancy (treeLdsFor in kernels.hip): benchmark scene 390 -> 318 ms in double,
    // 262 -> 220 ms in float, against 64.
    int tune_ad_block = 64, tune_child_block = 256;
    DevBuf tunebuf;

    template <class T> const T *upload(const std::vector<T> &v) {
        if (v.empty()) return nullptr;
        DevBuf b;
        b.bytes = v.size() * sizeof(T);
        if (hipMalloc(&b.p, b.bytes) != hipSuccess) return nullptr;
        if (hipMemcpy(b.p, v.data(), b.bytes, hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(b.p); return nullptr; }
        bufs.push_back(b);
        return static_cast<const T *>(b.p);
    }
    bool ensure(DevBuf &b, size_t bytes) {
        if (b.bytes >= bytes) return true;
        if (b.p) { (void)hipStreamSynchronize(stream); (void)hipFree(b.p); b.p = nullptr; b.bytes = 0; }
        HIPCHK(hipMalloc(&b.p, bytes));
        b.bytes = bytes;
        return true;
    }
    bool build();
    bool initEdges(const float *weights);
    bool render(const dtrt_render_options &o, float *out, bool out_is_device, int ad_kind = AD_VOLPATH) {
        const Camera *c = &cam;
        float *outs[1] = {out};
        return renderViews(1, &c, &o, outs, out_is_device, ad_kind);
    }
    bool renderViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, float *const *outs, bool on_device,
                     int ad_kind = AD_VOLPATH, Bump *pool = nullptr);
    bool primaryEdges(const Camera &c, PECache &out);
    void projectPrimaryEdges(const Camera &cam, std::vector<double> &cdf, std::vector<float> &packed, int &npe,
                             double &pe_sum) const;
    // dtrt_render_device_noalloc: one view into a device image, every device buffer of
    // the render taken from the library's scratch area; no device allocation.
    bool renderNoAlloc(const dtrt_render_options &o, float *out_device, int ad_kind) {
        void *base = nullptr;
        uint64_t bytes = 0;
        HIPCHK(scratchArea(&base, &bytes));
        Bump pool{static_cast<char *>(base), bytes};
        const Camera *c = &cam;
        float *outs[1] = {out_device};
        return renderViews(1, &c, &o, outs, true, ad_kind, &pool);
    }
    bool captureViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, float *const *images);
    int replayViews(const uint64_t *seeds);
    std::vector<void *> graphBuffers() const {
        return {viewbuf.p, acc.p, acc_pe.p, queue.p, qcount.p, stats.p, work.p, ovf[0].p, ovf[1].p};
    }
    bool setupViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, std::vector<DView> &hv,
                    std::vector<uint64_t> &pe_total, Bump *pool = nullptr);
    bool renderAdjoint(const dtrt_render_options &o, int med, const float *dimg, float *grad) {
        const Camera *c = &cam;
        return renderAdjointViews(1, &c, &o, med, dimg, grad);
    }
    bool renderAdjointViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, int med, const float *dimg,
                            float *grad, Bump *pool = nullptr, bool launch = true, RenderParams *Pout = nullptr,
                            uint64_t *total_out = nullptr);
    // One L2 inverse-rendering step (dtrt_scene_capture_step): forward views into
    // `images`, d loss / d I = 2 (I - T), its contraction with the tangent images into
    // pgrad[nder], the density adjoint into dgrid. Eager; captureStep records the same
    // launches as a graph, replayStep relaunches it with new seeds.
    struct StepArgs {
        std::vector<Camera> cams;
        std::vector<dtrt_render_options> fopts, aopts;
        std::vector<float *> images;
        std::vector<const float *> targets;
        int med = -1;
        float *dgrid = nullptr;
        double *pgrad = nullptr, *loss = nullptr;
    };
    bool runStep(StepArgs &a);
    bool stepLoss(StepArgs &a, const std::vector<DView> &hv, hipStream_t sm);
    bool captureStep(StepArgs &a);
    int replayStep(const uint64_t *seeds);
    struct StepGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        StepArgs args;
        std::vector<DView> hvf, hva;
        std::vector<void *> bufs;
        void reset() {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
            x = nullptr;
            g = nullptr;
        }
    } stepg;
    DevBuf stepbuf;   // the step's d loss / d I images and its adjoint view table
    // dtrt_scene_update_shape: new vertex positions (and optionally normals and tangents)
    // of one shape; face normals, edges and light CDFs are recomputed on the host, the
    // triangle records and BVH boxes are rebuilt on the GPU over the existing tree.
    bool updateShape(int si, const float *v, const float *n, const float *vel);
    bool uploadEdges();    // edges, edge CDFs and per-shape edge ranges of all shapes
    bool uploadLights();   // light CDFs, pmf and areas (scene.cpp:45-98)
    void dropPrimaryEdges();
    DevBuf edgebuf, ecdfbuf, lightbuf, refitbuf, shapecdfbuf;
    std::vector<int> refit_level_end;   // node ranges of refitbuf per depth, deepest first
    std::vector<float> last_edge_weights;   // the weights of the last initEdges call (empty: none)
    bool renderPrimal(const dtrt_render_options &o, float *out, bool out_is_device, int integrator);
    bool autotune(const dtrt_render_options &o, int ad_kind, double *best_ms, int64_t *settings3);
    bool setDensity(int med, const float *density, bool on_device);
    bool setVoxelLanes(int med, const int *vox, int n);
    bool uploadMedia() {
        if (hmed.empty()) return true;
        HIPCHK(hipMemcpyAsync(const_cast<DMedium *>(S.media), hmed.data(), hmed.size() * sizeof(DMedium), hipMemcpyHostToDevice,
                              stream));
        HIPCHK(hipStreamSynchronize(stream));
        return true;
    }
    size_t accBytes() const { return S.f64 ? 8 : 4; }   // accumulator element: the estimator's type
    bool gridMedium(int med) const {
        if (med < 0 || med >= (int)hmed.size() || hmed[med].type != MED_HETEROGENEOUS || !hmed[med].density)
            return fail("medium index does not name a heterogeneous grid medium of this scene");
        return true;
    }
};

bool frameRotationReady() {
    // utils.cpp:55 — rotation by 0.1f rad about normalize(0.1f, 0.2f, 0.3f), Rodrigues' formula
    static std::mutex mu;
    static std::vector<int> done;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail("no HIP device");
    std::lock_guard<std::mutex> lk(mu);
    if (std::find(done.begin(), done.end(), dev) != done.end()) return true;
    double ax = 0.1f, ay = 0.2f, az = 0.3f, l = std::sqrt(ax * ax + ay * ay + az * az);
    ax /= l; ay /= l; az /= l;
    double th = (double)0.1f, c = std::cos(th), s = std::sin(th), t = 1 - c;
    const double m[9] = {c + ax * ax * t, ax * ay * t - az * s, ax * az * t + ay * s,
                         ay * ax * t + az * s, c + ay * ay * t, ay * az * t - ax * s,
                         az * ax * t - ay * s, az * ay * t + ax * s, c + az * az * t};
    HIPCHK(setFrameRotation(m));
    done.push_back(dev);
    return true;
}

// Scratch limits of the kernels a scene launches. A kernel's private segment (register
// spills, private arrays) lives in per-lane scratch that the runtime provisions for every
// wave the device holds at once. The limits are the largest figures that a kernel has
// passed the GPU parity suite with, rounded up: 10,360 bytes per lane and 679 MB over the
// device (a build of the double-precision nder-16 refill kernel: 10,360 B x 64 lanes x 4
// waves per CU x 256 CUs; the current build needs 7,104 B). A width-16 build of the
// single-precision AD-path kernel held to two waves per SIMD, whose spills needed 7.9 KB
// per lane (1.04 GB over the device), rendered empty images: every path ended after two
// segments. Its cause is not known; per-lane size alone is not it (the 10,360-byte
// kernel was correct), the device-wide figure is not ruled out. So a kernel beyond
// either limit is refused, not launched.
constexpr uint64_t kMaxPrivateBytesPerLane = 12u << 10;
constexpr uint64_t kMaxScratchBytesPerDevice = 768ull << 20;

// Device-wide scratch of kernel k at full occupancy on a device of `cus` compute units.
uint64_t deviceScratch(const KernelInfo &k, int cus) {
    return (uint64_t)k.private_bytes * 64 * k.waves_per_cu * (uint64_t)std::max(cus, 1);
}

// The limits in force (lowered only by dtrt_test_set_scratch_limits) and the device,
// width and precision combinations already checked against them.
struct ScratchLimits {
    std::mutex mu;
    uint64_t lane = kMaxPrivateBytesPerLane, device = kMaxScratchBytesPerDevice;
    std::vector<std::array<int, 3>> done;
};
ScratchLimits &scratchLimits() {
    static ScratchLimits L;
    return L;
}

// Checks, once per device, estimator precision and width, that every kernel a scene of
// that kind launches stays within the scratch limits.
bool kernelsWithinScratchLimits(const DScene &S) {
    ScratchLimits &L = scratchLimits();
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail("no HIP device");
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const std::array<int, 3> key{dev, S.nder, S.f64};
    std::lock_guard<std::mutex> lk(L.mu);
    if (std::find(L.done.begin(), L.done.end(), key) != L.done.end()) return true;
    KernelInfo info[32];
    int n = 0;
    HIPCHK(kernel_info(S, info, 32, &n));
    for (int i = 0; i < std::min(n, 32); ++i) {
        const KernelInfo &k = info[i];
        const uint64_t dev_bytes = deviceScratch(k, cus);
        if (k.private_bytes > L.lane || dev_bytes > L.device) {
            char msg[256];
            std::snprintf(msg, sizeof(msg),
                          "%s (%s, nder %d) needs %u bytes of scratch per lane, %.1f MB over the device: above "
                          "the checked limits of %llu bytes and %.1f MB",
                          k.name, S.f64 ? "double" : "float", S.nder, k.private_bytes, dev_bytes / 1e6,
                          (unsigned long long)L.lane, L.device / 1e6);
            return fail(msg);
        }
    }
    L.done.push_back(key);
    return true;
}

bool Scene::build() {
    if (!frameRotationReady()) return false;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&caller_ev, hipEventDisableTiming));
    std::memset(&S, 0, sizeof(S));
    S.nder = nder;
    S.f64 = precisionNow() == 64;
    if (!kernelsWithinScratchLimits(S)) return false;
    S.nshapes = (int)shapes.size();
    S.nlights = (int)emitters.size();
    S.nmedia = (int)media.size();
    // flatten geometry
    std::vector<float> V, Vd, Nrm, Nd, FN, FNd;
    std::vector<int> F, fshape;
    std::vector<DShape> dsh;
    std::vector<int4> edges;
    std::vector<double> ecdf;
    std::vector<std::array<float, 9>> tri;
    int vbase = 0, fbase = 0, ebase = 0;
    for (size_t si = 0; si < shapes.size(); ++si) {
        const Shape &s = *shapes[si];
        DShape d{};
        d.vbase = vbase; d.fbase = fbase; d.nv = s.nv; d.nf = s.nf;
        d.light_id = s.light_id; d.bsdf_id = s.bsdf_id; d.med_int = s.med_int; d.med_ext = s.med_ext;
        d.has_normals = s.hasNormals();
        d.ebase = ebase; d.ne = (int)s.edges.size();
        dsh.push_back(d);
        for (int i = 0; i < s.nv; ++i) {
            V.insert(V.end(), {(float)s.v[i].x, (float)s.v[i].y, (float)s.v[i].z});
            D3 nn = s.hasNormals() ? s.n[i] : D3{0, 0, 0};
            Nrm.insert(Nrm.end(), {(float)nn.x, (float)nn.y, (float)nn.z});
            for (int j = 0; j < nder; ++j) {
                const D3 &g = s.vd[(size_t)i * nder + j];
                Vd.insert(Vd.end(), {(float)g.x, (float)g.y, (float)g.z});
                D3 h = s.hasNormals() ? s.nd[(size_t)i * nder + j] : D3{0, 0, 0};
                Nd.insert(Nd.end(), {(float)h.x, (float)h.y, (float)h.z});
            }
        }
        for (int k = 0; k < s.nf; ++k) {
            int a = s.f[3 * k] + vbase, b = s.f[3 * k + 1] + vbase, c = s.f[3 * k + 2] + vbase;
            F.insert(F.end(), {a, b, c});
            fshape.push_back((int)si);
            FN.insert(FN.end(), {(float)s.fn[k].x, (float)s.fn[k].y, (float)s.fn[k].z});
            for (int j = 0; j < nder; ++j) {
                const D3 &g = s.fnd[(size_t)k * nder + j];
                FNd.insert(FNd.end(), {(float)g.x, (float)g.y, (float)g.z});
            }
            std::array<float, 9> t;
            for (int q = 0; q < 3; ++q) {
                const D3 &p = s.v[s.f[3 * k + q]];
                t[3 * q] = (float)p.x; t[3 * q + 1] = (float)p.y; t[3 * q + 2] = (float)p.z;
            }
            tri.push_back(t);
        }
        for (const EdgeH &e : s.edges)
            edges.push_back(int4{e.v0 + vbase, e.v1 + vbase, e.f0 + fbase, e.f1 >= 0 ? e.f1 + fbase : -1});
        ecdf.insert(ecdf.end(), s.edge_cdf.begin(), s.edge_cdf.end());   // ne+1 entries at ebase + si
        vbase += s.nv;
        fbase += s.nf;
        ebase += (int)s.edges.size();
    }
    S.NV = vbase;
    S.NF = fbase;
    BvhOut bvh;
    if (!buildBvh(tri, fshape, bvh)) return false;
    S.nnodes = (int)(bvh.nodes.size() / 4);
    S.ntris = (int)tri.size();
    // a lane's stack holds at most one entry per inner node on its path (depth - 1)
    S.stack_depth = std::min(DTRT_BVH_STACK, std::max(1, bvh.depth));
    S.nodes = upload(bvh.nodes);
    S.tris = upload(bvh.tris);
    S.V = upload(V); S.Vd = upload(Vd); S.Nrm = upload(Nrm); S.Nd = upload(Nd);
    S.F = upload(F); S.face_shape = upload(fshape); S.FN = upload(FN); S.FNd = upload(FNd);
    S.shapes = upload(dsh);
    S.edges = upload(edges);
    S.edge_cdf = upload(ecdf);
    // materials
    std::vector<DBsdf> db;
    for (auto *b : bsdfs) db.push_back(b->d);
    std::vector<DPhase> dp;
    for (auto *p : phases) dp.push_back(p->d);
    std::vector<DEmitter> de;
    for (auto *e : emitters) de.push_back(e->d);
    S.bsdfs = upload(db);
    S.phases = upload(dp);
    S.emitters = upload(de);
    std::vector<DMedium> dm;
    for (auto *m : media) {
        DMedium d = m->d;
        d.density = m->density.empty() ? nullptr : upload(m->density);
        d.albedo_grid = m->albedo.empty() ? nullptr : upload(m->albedo);
        for (int j = 0; j < DTRT_NDER_MAX; ++j) d.vox_lane[j] = -1;
        d.vox_any = 0;
        dm.push_back(d);
    }
    S.media = upload(dm);
    hmed = dm;
    if (!uploadLights()) return false;
    for (const auto &b : bufs) if (!b.p) return fail("scene: device allocation failed");
    return initEdges(nullptr);
}

// scene.cpp:45-98 — light pmf proportional to area * luminance * pi, per light a
// triangle CDF proportional to area; one device buffer, rewritten when geometry changes.
bool Scene::uploadLights() {
    int nl = (int)emitters.size();
    std::vector<double> lcdf(nl), pool;
    std::vector<float> lpmf(nl), lareas(nl);
    std::vector<int> abeg(nl);
    double tot = 0;
    std::vector<double> pmf(nl);
    for (int l = 0; l < nl; ++l) {
        const Shape &s = *shapes[emitters[l]->d.shape_id];
        abeg[l] = (int)pool.size();
        double sum = 0;
        std::vector<double> c;
        for (int k = 0; k < s.nf; ++k) {
            const int *fi = &s.f[3 * k];
            sum += 0.5 * norm3(cross3(s.v[fi[1]] - s.v[fi[0]], s.v[fi[2]] - s.v[fi[0]]));
            c.push_back(sum);
        }
        for (double &x : c) x /= sum;
        pool.insert(pool.end(), c.begin(), c.end());
        lareas[l] = (float)sum;
        const float *I = emitters[l]->d.I;
        pmf[l] = sum * (0.212671f * I[0] + 0.715160f * I[1] + 0.072169f * I[2]) * M_PI;
        tot += pmf[l];
    }
    if (nl > 0 && !(tot > 0)) return fail("scene: total light importance must be positive");
    double run = 0;
    for (int l = 0; l < nl; ++l) { pmf[l] /= tot; lpmf[l] = (float)pmf[l]; run += pmf[l]; lcdf[l] = run; }
    // one buffer: [light cdf | area cdf pool] doubles, then [pmf | areas] floats, then begins
    const size_t nd = lcdf.size() + pool.size(), bytes = nd * 8 + (size_t)nl * 8 + (size_t)nl * 4 + 64;
    if (!ensure(lightbuf, bytes)) return false;
    char *b = static_cast<char *>(lightbuf.p);
    std::vector<char> host(bytes, 0);
    std::memcpy(host.data(), lcdf.data(), lcdf.size() * 8);
    std::memcpy(host.data() + lcdf.size() * 8, pool.data(), pool.size() * 8);
    std::memcpy(host.data() + nd * 8, lpmf.data(), (size_t)nl * 4);
    std::memcpy(host.data() + nd * 8 + (size_t)nl * 4, lareas.data(), (size_t)nl * 4);
    std::memcpy(host.data() + nd * 8 + (size_t)nl * 8, abeg.data(), (size_t)nl * 4);
    HIPCHK(hipStreamSynchronize(stream));   // no kernel of this scene is reading the old values
    HIPCHK(hipMemcpy(b, host.data(), bytes, hipMemcpyHostToDevice));
    S.light_cdf = reinterpret_cast<const double *>(b);
    S.area_cdf_pool = reinterpret_cast<const double *>(b + lcdf.size() * 8);
    S.light_pmf = reinterpret_cast<const float *>(b + nd * 8);
    S.light_areas = reinterpret_cast<const float *>(b + nd * 8 + (size_t)nl * 4);
    S.area_cdf_begin = reinterpret_cast<const int *>(b + nd * 8 + (size_t)nl * 8);
    return true;
}

// Edges of all shapes (global vertex and face ids), their per-shape CDFs (ne + 1
// entries at ebase + shape index) and each shape's edge range in the shape records.
bool Scene::uploadEdges() {
    std::vector<int4> edges;
    std::vector<double> ecdf;
    std::vector<DShape> dsh(shapes.size());
    HIPCHK(hipMemcpy(dsh.data(), S.shapes, dsh.size() * sizeof(DShape), hipMemcpyDeviceToHost));
    int ebase = 0;
    for (size_t si = 0; si < shapes.size(); ++si) {
        const Shape &s = *shapes[si];
        dsh[si].ebase = ebase;
        dsh[si].ne = (int)s.edges.size();
        for (const EdgeH &e : s.edges)
            edges.push_back(int4{e.v0 + dsh[si].vbase, e.v1 + dsh[si].vbase, e.f0 + dsh[si].fbase,
                                 e.f1 >= 0 ? e.f1 + dsh[si].fbase : -1});
        ecdf.insert(ecdf.end(), s.edge_cdf.begin(), s.edge_cdf.end());
        ebase += (int)s.edges.size();
    }
    if (!ensure(edgebuf, std::max<size_t>(1, edges.size()) * sizeof(int4)) ||
        !ensure(ecdfbuf, std::max<size_t>(1, ecdf.size()) * sizeof(double)))
        return false;
    HIPCHK(hipStreamSynchronize(stream));
    if (!edges.empty()) HIPCHK(hipMemcpy(edgebuf.p, edges.data(), edges.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ecdfbuf.p, ecdf.data(), ecdf.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(const_cast<DShape *>(S.shapes), dsh.data(), dsh.size() * sizeof(DShape), hipMemcpyHostToDevice));
    S.edges = static_cast<const int4 *>(edgebuf.p);
    S.edge_cdf = static_cast<const double *>(ecdfbuf.p);
    return true;
}

void Scene::dropPrimaryEdges() {
    (void)hipStreamSynchronize(stream);
    for (void *p : pe_bufs) (void)hipFree(p);
    pe_bufs.clear();
    pecache.clear();
}

bool Scene::updateShape(int si, const float *v, const float *n, const float *vel) {
    if (si < 0 || si >= (int)shapes.size() || !v) return fail("update_shape: bad shape index or null vertices");
    Shape &sh = *shapes[si];
    if (n && !sh.hasNormals()) return fail("update_shape: the shape was created without normals");
    // host geometry, as the Shape constructor builds it (shape.cpp:10-126, 254-300)
    for (int i = 0; i < sh.nv; ++i) sh.v[i] = D3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    if (n)
        for (int i = 0; i < sh.nv; ++i) {
            D3 q{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
            sh.n[i] = q * (1.0 / norm3(q));
        }
    if (vel) setShapeVelocities(sh, vel);
    computeFaceNormals(sh);
    if (!constructEdges(sh)) return false;
    // device copies of this shape's vertex, normal and face records
    DShape d;
    HIPCHK(hipMemcpy(&d, S.shapes + si, sizeof(DShape), hipMemcpyDeviceToHost));
    std::vector<float> V, Vd, Nrm, Nd, FN, FNd;
    for (int i = 0; i < sh.nv; ++i) {
        V.insert(V.end(), {(float)sh.v[i].x, (float)sh.v[i].y, (float)sh.v[i].z});
        D3 nn = sh.hasNormals() ? sh.n[i] : D3{0, 0, 0};
        Nrm.insert(Nrm.end(), {(float)nn.x, (float)nn.y, (float)nn.z});
        for (int j = 0; j < nder; ++j) {
            const D3 &g = sh.vd[(size_t)i * nder + j];
            Vd.insert(Vd.end(), {(float)g.x, (float)g.y, (float)g.z});
            D3 h = sh.hasNormals() ? sh.nd[(size_t)i * nder + j] : D3{0, 0, 0};
            Nd.insert(Nd.end(), {(float)h.x, (float)h.y, (float)h.z});
        }
    }
    for (int k = 0; k < sh.nf; ++k) {
        FN.insert(FN.end(), {(float)sh.fn[k].x, (float)sh.fn[k].y, (float)sh.fn[k].z});
        for (int j = 0; j < nder; ++j) {
            const D3 &g = sh.fnd[(size_t)k * nder + j];
            FNd.insert(FNd.end(), {(float)g.x, (float)g.y, (float)g.z});
        }
    }
    HIPCHK(hipStreamSynchronize(stream));
    auto put = [&](const float *dst, const std::vector<float> &src, size_t first) -> hipError_t {
        return hipMemcpy(const_cast<float *>(dst) + first, src.data(), src.size() * sizeof(float), hipMemcpyHostToDevice);
    };
    HIPCHK(put(S.V, V, (size_t)d.vbase * 3));
    HIPCHK(put(S.Vd, Vd, (size_t)d.vbase * nder * 3));
    HIPCHK(put(S.Nrm, Nrm, (size_t)d.vbase * 3));
    HIPCHK(put(S.Nd, Nd, (size_t)d.vbase * nder * 3));
    HIPCHK(put(S.FN, FN, (size_t)d.fbase * 3));
    HIPCHK(put(S.FNd, FNd, (size_t)d.fbase * nder * 3));
    if (!uploadEdges() || !initEdges(last_edge_weights.empty() ? nullptr : last_edge_weights.data()) || !uploadLights())
        return false;
    // triangle records in leaf order, then node boxes level by level from the deepest
    if (refit_level_end.empty()) {
        std::vector<int> depth(S.nnodes, -1), order;
        std::vector<float4> nodes((size_t)S.nnodes * 4);
        HIPCHK(hipMemcpy(nodes.data(), S.nodes, nodes.size() * sizeof(float4), hipMemcpyDeviceToHost));
        std::vector<std::vector<int>> levels;
        std::vector<int> q{0};
        depth[0] = 0;
        for (size_t h = 0; h < q.size(); ++h) {
            const int nd_ = q[h];
            if ((int)levels.size() <= depth[nd_]) levels.resize(depth[nd_] + 1);
            levels[depth[nd_]].push_back(nd_);
            for (int k = 0; k < 2; ++k) {
                int link;
                std::memcpy(&link, &nodes[4 * (size_t)nd_ + k].w, 4);
                if (link >= 0) { depth[link] = depth[nd_] + 1; q.push_back(link); }
            }
        }
        for (int l = (int)levels.size() - 1; l >= 0; --l) {
            order.insert(order.end(), levels[l].begin(), levels[l].end());
            refit_level_end.push_back((int)order.size());
        }
        if (!ensure(refitbuf, order.size() * sizeof(int))) return false;
        HIPCHK(hipMemcpy(refitbuf.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPCHK(launch_tri_records(S, stream));
    int begin = 0;
    for (int end : refit_level_end) {
        HIPCHK(launch_refit(S, static_cast<const int *>(refitbuf.p) + begin, end - begin, stream));
        begin = end;
    }
    HIPCHK(hipStreamSynchronize(stream));
    // what was built from the old geometry: projected primary edges and captured graphs
    dropPrimaryEdges();
    graph.reset();
    stepg.reset();
    return true;
}

// bruteforce.cpp:4-20 and edge_manager.cpp:6-16
bool Scene::initEdges(const float *weights) {
    if (weights && weights != last_edge_weights.data()) last_edge_weights.assign(weights, weights + shapes.size());
    std::vector<double> w(shapes.size() + 1, 0.0);
    double s = 0;
    for (size_t i = 0; i < shapes.size(); ++i) {
        bool ignore = bsdfs[shapes[i]->bsdf_id]->d.type == BSDF_NULL && shapes[i]->light_id < 0;
        double x = (ignore ? 0.0 : shapes[i]->edge_sum) * (weights ? (double)weights[i] : 1.0);
        if (x < -1e-7) return fail("initEdges: negative weight");
        s += x;
        w[i + 1] = w[i] + x;
    }
    S.edges_valid = s > 1e-7;
    if (S.edges_valid) { for (double &x : w) x /= s; w.back() = 1.0; }
    if (!ensure(shapecdfbuf, w.size() * sizeof(double))) return false;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(shapecdfbuf.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    S.shape_edge_cdf = static_cast<const double *>(shapecdfbuf.p);
    return true;
}

// edge_manager.cpp:6-134 for one camera: silhouette edges of non-null or light shapes,
// projected and clipped to the camera's window, sampled in proportion to screen length.
bool Scene::primaryEdges(const Camera &cam, PECache &out) {
    auto it = pecache.find(cam.uid);
    if (it != pecache.end()) { out = it->second; return true; }
    PECache c;
    std::vector<double> cdf;
    std::vector<float> packed;
    projectPrimaryEdges(cam, cdf, packed, c.npe, c.pe_sum);
    auto put = [&](const void *src, size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 1)) != hipSuccess) return nullptr;
        pe_bufs.push_back(p);
        if (bytes && hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return p;
    };
    c.cdf = static_cast<const double *>(put(cdf.data(), cdf.size() * sizeof(double)));
    c.pe = packed.empty() ? nullptr : static_cast<const float *>(put(packed.data(), packed.size() * sizeof(float)));
    if (!c.cdf || (!packed.empty() && !c.pe)) return fail("scene: device allocation failed");
    pecache[cam.uid] = c;
    out = c;
    return true;
}

// The primary edges of `cam` on the host: the sampling CDF (npe + 1 entries) and the
// packed records [npe][16 + 4 nder] the kernels read.
void Scene::projectPrimaryEdges(const Camera &cam, std::vector<double> &cdf, std::vector<float> &packed, int &npe,
                                double &pe_sum) const {
    std::vector<PEdge> pe;
    std::vector<double> pw;
    for (const Shape *sh : shapes) {
        if (!(bsdfs[sh->bsdf_id]->d.type != BSDF_NULL || sh->light_id >= 0)) continue;
        for (const EdgeH &e : sh->edges) {
            if (silhouetteHost(*sh, e, cam.cpos) <= 0) continue;
            ADV a, b;
            a.v = sh->v[e.v0];
            b.v = sh->v[e.v1];
            for (int j = 0; j < nder; ++j) { a.g[j] = sh->vd[(size_t)e.v0 * nder + j]; b.g[j] = sh->vd[(size_t)e.v1 * nder + j]; }
            projectEdge(cam, a, b, nder, pe, pw);
        }
    }
    npe = (int)pe.size();
    cdf.assign(pe.size() + 1, 0.0);
    for (size_t i = 0; i < pw.size(); ++i) cdf[i + 1] = cdf[i] + pw[i];
    pe_sum = cdf.back();
    if (pe_sum > 0) { for (double &x : cdf) x /= pe_sum; cdf.back() = 1.0; }
    const int stride = 16 + 4 * nder;
    packed.assign(pe.size() * stride, 0.f);
    for (size_t i = 0; i < pe.size(); ++i) {
        float *p = &packed[i * stride];
        const PEdge &e = pe[i];
        double v[16] = {e.s0.x, e.s0.y, e.s1.x, e.s1.y, e.p0.x, e.p0.y, e.p0.z, e.p1.x, e.p1.y, e.p1.z,
                        e.c0.x, e.c0.y, e.c0.z, e.c1.x, e.c1.y, e.c1.z};
        for (int k = 0; k < 16; ++k) p[k] = (float)v[k];
        for (int j = 0; j < nder; ++j) {
            p[16 + 4 * j] = (float)e.s0.dx[j]; p[17 + 4 * j] = (float)e.s0.dy[j];
            p[18 + 4 * j] = (float)e.s1.dx[j]; p[19 + 4 * j] = (float)e.s1.dy[j];
        }
    }
}

constexpr uint64_t MAX_CHUNK = uint64_t(1) << 22;    // AD samples per launch

// Per-view device records: camera, sample and pixel offsets, seed, primary edges.
bool Scene::setupViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, std::vector<DView> &hv,
                       std::vector<uint64_t> &pe_total, Bump *pool) {
    hv.assign(nv, DView{});
    pe_total.assign(nv, 0);
    uint64_t s0 = 0;
    int64_t pix0 = 0;
    for (int v = 0; v < nv; ++v) {
        const Camera &c = *cams[v];
        const dtrt_render_options &o = opts[v];
        if (c.nder != nder) return fail("render: a view's camera was created with another nder");
        if (c.med >= (int)media.size()) return fail("render: a view's camera medium is out of range");
        if (o.spp < 0) return fail("render: negative spp");
        DView &d = hv[v];
        fillDCamera(c, nder, d.cam);
        d.npix = c.cropped() ? c.cw * c.ch : c.w * c.h;
        d.pix0 = (int)pix0;
        d.s0 = s0;
        d.seed = o.seed;
        d.spp = o.spp;
        d.sppe = o.sppe;
        PECache pe;
        if (o.sppe > 0 && pool && !pecache.count(c.uid)) {   // project into the pool, do not cache
            std::vector<double> cdf;
            std::vector<float> packed;
            projectPrimaryEdges(c, cdf, packed, pe.npe, pe.pe_sum);
            void *a = pool->take(cdf.size() * sizeof(double)), *b = pool->take(packed.size() * sizeof(float));
            if (!a || !b) return fail("render: the primary edges do not fit the scratch area");
            HIPCHK(hipMemcpy(a, cdf.data(), cdf.size() * sizeof(double), hipMemcpyHostToDevice));
            if (!packed.empty()) HIPCHK(hipMemcpy(b, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice));
            pe.cdf = static_cast<const double *>(a);
            pe.pe = packed.empty() ? nullptr : static_cast<const float *>(b);
        } else if (o.sppe > 0) {
            if (!primaryEdges(c, pe)) return false;
        }
        if (o.sppe > 0) {
            d.npe = pe.npe;
            d.pe_sum = pe.pe_sum;
            d.pe_cdf = pe.cdf;
            d.pe = pe.pe;
            if (pe.npe > 0) pe_total[v] = (uint64_t)std::ceil(pe.pe_sum * o.sppe / 128.0) * 128;
        }
        s0 += (uint64_t)d.npix * o.spp;
        pix0 += d.npix;
        if (pix0 > 0x7FFFFFFF / 4) return fail("render: too many pixels in one call");
    }
    void *vt = nullptr;
    if (pool) {
        vt = pool->take((uint64_t)nv * sizeof(DView));
        if (!vt) return fail("render: the view table does not fit the scratch area");
    } else {
        if (!ensure(viewbuf, (size_t)nv * sizeof(DView))) return false;
        vt = viewbuf.p;
    }
    HIPCHK(hipMemcpyAsync(vt, hv.data(), (size_t)nv * sizeof(DView), hipMemcpyHostToDevice, stream));
    S.views = static_cast<const DView *>(vt);
    S.nviews = nv;
    return true;
}

constexpr uint64_t FIRST_CHUNK = uint64_t(1) << 16;  // before the children-per-sample rate is known

// All views in one pass: their AD samples form one global sample space (launched in
// the same chunks), their primary-edge samples follow view by view, and each view's
// pixels occupy their own range of the accumulators. Every view's image equals what
// a render of that view alone gives (same streams), up to float summation order.
bool Scene::renderViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, float *const *outs,
                        bool out_is_device, int ad_kind, Bump *pool) {
    if (nv <= 0) return fail("render: no views");
    if (pool && !out_is_device) return fail("render: the scratch-only render writes device images only");
    if (!afterCaller()) return false;
    const dtrt_render_options &o0 = opts[0];
    for (int v = 1; v < nv; ++v)
        if (opts[v].max_bounces != o0.max_bounces || opts[v].mode != o0.mode || opts[v].sppse != o0.sppse ||
            opts[v].ddist_coeff != o0.ddist_coeff)
            return fail("render: views of one call must share max_bounces, mode, sppse and ddistCoeff");
    RenderParams P{};
    P.ad_kind = ad_kind;
    P.max_bounces = o0.max_bounces; P.sppse = std::max(0, o0.sppse);
    P.mode = o0.mode >= 0 ? o0.mode : 0x3F; P.ddist = o0.ddist_coeff;
    if (P.max_bounces < 0 || P.max_bounces > 200) return fail("render: max_bounces out of range");
    P.rec_stride = childRecordStride(nder, S.f64 != 0);
    P.ad_block = tune_ad_block;
    P.child_block = tune_child_block;
    P.tree_in_hbm = tree_in_hbm ? 1 : 0;
    std::vector<DView> hv;
    std::vector<uint64_t> pe_total;
    if (!setupViews(nv, cams, opts, hv, pe_total, pool)) return false;
    int64_t npix_all = 0;
    uint64_t total = 0, pe_all = 0;
    for (int v = 0; v < nv; ++v) {
        npix_all += hv[v].npix;
        total += (uint64_t)hv[v].npix * hv[v].spp;
        pe_all += pe_total[v];
    }
    P.npix = (int)npix_all;
    const size_t npix = P.npix, img = (size_t)(nder + 1) * npix * 3;
    // One queue of child records for both passes. AD samples are launched in chunks
    // sized from the children-per-sample rate seen so far; a sample whose pushes
    // overflow is replayed exactly (see k_ad_paths). A primary-edge sample pushes at
    // most 2 records, so its chunks never overflow.
    const size_t es = accBytes();
    void *A = nullptr, *APE = nullptr;
    char *Q = nullptr;
    unsigned *qc = nullptr;   // [0] queued records, [1] overflowed samples, [2] consumed records
    unsigned *st = nullptr;   // [0..STAT_COUNT) counters, [32..) discarded (replays)
    unsigned long long *wk = nullptr;
    uint4 *OV[2] = {nullptr, nullptr};
    uint64_t cap = 0, max_chunk = 0;
    if (pool) {   // everything from the scratch area: fixed parts first, the queue gets the rest
        max_chunk = std::min<uint64_t>(std::min<uint64_t>(tune_max_chunk, uint64_t(1) << 18), std::max<uint64_t>(total, 64));
        A = pool->take(img * es);
        APE = pool->take((uint64_t)nder * npix * 3 * es);
        qc = static_cast<unsigned *>(pool->take(64));
        st = static_cast<unsigned *>(pool->take(64 * 4));
        wk = static_cast<unsigned long long *>(pool->take(16 * 8));
        OV[0] = static_cast<uint4 *>(pool->take(max_chunk * 16));
        OV[1] = static_cast<uint4 *>(pool->take(max_chunk * 16));
        const uint64_t qbytes = pool->left / 256 * 256;
        if (!A || !APE || !qc || !st || !wk || !OV[0] || !OV[1] || qbytes < NOALLOC_MIN_QUEUE)
            return fail("render_device_noalloc: the image needs more than the " +
                        std::to_string(DTRT_SCRATCH_BYTES >> 20) + " MiB scratch area holds (accumulators " +
                        std::to_string((img + (size_t)nder * npix * 3) * es >> 20) + " MiB, plus a queue of at least " +
                        std::to_string(NOALLOC_MIN_QUEUE >> 20) + " MiB)");
        Q = static_cast<char *>(pool->take(qbytes));
        cap = std::min<uint64_t>(qbytes / P.rec_stride, 0xFFFFFFF0ull);
    } else {
        cap = std::max<uint64_t>(128, std::min<uint64_t>(queue_budget / P.rec_stride, 0xFFFFFFF0ull));
        max_chunk = std::min<uint64_t>(tune_max_chunk, std::max<uint64_t>(total, 64));
        if (!ensure(acc, img * es) || !ensure(acc_pe, (size_t)nder * npix * 3 * es) || !ensure(qcount, 64) ||
            !ensure(stats, 64 * 4) || !ensure(work, 16 * 8) || !ensure(queue, (size_t)cap * P.rec_stride) ||
            !ensure(ovf[0], (size_t)max_chunk * 16) || !ensure(ovf[1], (size_t)max_chunk * 16))
            return false;
        A = acc.p;
        APE = acc_pe.p;
        Q = static_cast<char *>(queue.p);
        qc = static_cast<unsigned *>(qcount.p);
        st = static_cast<unsigned *>(stats.p);
        wk = static_cast<unsigned long long *>(work.p);
        OV[0] = static_cast<uint4 *>(ovf[0].p);
        OV[1] = static_cast<uint4 *>(ovf[1].p);
    }
    const uint64_t pe_chunk = std::max<uint64_t>(64, std::min<uint64_t>(cap / 2 / 64 * 64, uint64_t(1) << 23));
    P.qcap = (unsigned)cap;
    // clears by fill kernels (the scratch area keeps the last call's contents)
    HIPCHK(launch_fill32(A, img * es / 4, 0u, stream));
    HIPCHK(launch_fill32(APE, (size_t)nder * npix * 3 * es / 4, 0u, stream));
    HIPCHK(launch_fill32(st, 64, 0u, stream));
    HIPCHK(launch_fill32(wk, 32, 0u, stream));
    // per-view output images in device memory (staged through outbuf for host outputs)
    std::vector<float *> dout(nv);
    if (out_is_device) {
        for (int v = 0; v < nv; ++v) dout[v] = outs[v];
    } else {
        if (!ensure(outbuf, img * 4)) return false;
        size_t off = 0;
        for (int v = 0; v < nv; ++v) {
            dout[v] = static_cast<float *>(outbuf.p) + off;
            const size_t n = (size_t)(nder + 1) * hv[v].npix * 3;
            HIPCHK(hipMemcpyAsync(dout[v], outs[v], n * 4, hipMemcpyHostToDevice, stream));
            off += n;
        }
    }
    if (!pool && ev.empty()) {
        ev.resize(2);
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    }
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned childGrid = (unsigned)std::max(1, ncu) * (unsigned)tune_child_blocks;
    std::memset(kernel_ms, 0, sizeof(kernel_ms));
    std::memset(kernel_launches, 0, sizeof(kernel_launches));
    auto timed = [&](int slot, auto &&fn) -> bool {   // per-kernel HIP-event times (not in the scratch-only render)
        if (pool) {
            HIPCHK(fn());
            kernel_launches[slot]++;
            return true;
        }
        HIPCHK(hipEventRecord(ev[0], stream));
        HIPCHK(fn());
        HIPCHK(hipEventRecord(ev[1], stream));
        HIPCHK(hipEventSynchronize(ev[1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        kernel_ms[slot] += ms;
        kernel_launches[slot]++;
        return true;
    };
    uint64_t children = 0, replayed = 0;
    auto counts = [&](unsigned h[2]) -> bool {
        HIPCHK(hipMemcpyAsync(h, qc, 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return true;
    };
    uint64_t chunk = std::min<uint64_t>(FIRST_CHUNK, max_chunk);
    last_chunks.clear();
    for (uint64_t s0 = 0; s0 < total;) {
        const uint64_t n = std::min(chunk, total - s0);
        last_chunks.push_back(n);
        uint64_t delivered = 0;
        unsigned h[2];
        HIPCHK(launch_fill32(qc, 3, 0u, stream));
        if (!timed(0, [&] { return launch_ad_paths(S, P, s0, n, nullptr, A, Q, qc, st, OV[0], qc + 1, wk, stream); })) return false;
        if (!timed(1, [&] { return launch_children(S, P, Q, qc, A, APE, wk, childGrid, stream); })) return false;
        if (!counts(h)) return false;
        delivered += std::min<uint64_t>(h[0], cap);
        int in = 0;
        for (unsigned nov = h[1]; nov > 0; in ^= 1) {   // each round delivers >= 1 record: terminates
            replayed += nov;
            HIPCHK(launch_fill32(qc, 3, 0u, stream));
            if (!timed(0, [&] { return launch_ad_paths(S, P, 0, nov, OV[in], A, Q, qc, st + 32, OV[in ^ 1], qc + 1, wk, stream); }))
                return false;
            if (!timed(1, [&] { return launch_children(S, P, Q, qc, A, APE, wk, childGrid, stream); })) return false;
            if (!counts(h)) return false;
            delivered += std::min<uint64_t>(h[0], cap);
            nov = h[1];
        }
        children += delivered;
        s0 += n;
        // next chunk: fill ~3/4 of the queue at the observed rate, in whole waves
        const double rate = (double)delivered / (double)n;
        uint64_t next = rate > 0 ? (uint64_t)(tune_fill * (double)cap / rate) : max_chunk;
        chunk = std::max<uint64_t>(64, std::min(next, max_chunk) / 64 * 64);
    }
    for (int v = 0; v < nv; ++v)
        for (uint64_t s0 = 0; s0 < pe_total[v]; s0 += pe_chunk) {
            uint64_t n = std::min(pe_chunk, pe_total[v] - s0);
            HIPCHK(launch_fill32(qc, 3, 0u, stream));
            if (!timed(2, [&] { return launch_primary(S, P, v, s0, n, pe_total[v], Q, qc, st, wk, stream); })) return false;
            if (!timed(1, [&] { return launch_children(S, P, Q, qc, A, APE, wk, childGrid, stream); })) return false;
            unsigned h[2];
            if (!counts(h)) return false;
            children += std::min<uint64_t>(h[0], cap);
        }
    for (int v = 0; v < nv; ++v)
        if (!timed(3, [&] {
                return launch_finalize(S, P, hv[v].npix, hv[v].pix0, hv[v].spp, pe_total[v] > 0, A, APE, dout[v], st, stream);
            }))
            return false;
    if (!out_is_device)
        for (int v = 0; v < nv; ++v)
            HIPCHK(hipMemcpyAsync(outs[v], dout[v], (size_t)(nder + 1) * hv[v].npix * 3 * 4, hipMemcpyDeviceToHost, stream));
    unsigned hst[STAT_COUNT];
    unsigned long long hw[12];
    HIPCHK(hipMemcpyAsync(hst, st, sizeof(hst), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(hw, wk, sizeof(hw), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int i = 0; i < 6; ++i) last_stats[i] = hst[i];
    last_stats[6] = (int64_t)replayed;
    last_replayed = replayed;
    last_stats[7] = (int64_t)total;
    last_stats[8] = (int64_t)children;
    last_stats[9] = (int64_t)pe_all;
    for (int i = 0; i < 12; ++i) work_counts[i] = hw[i];
    if (hst[STAT_QUEUE_OVERFLOW]) return fail("render: primary-edge child queue overflow");
    return true;
}

// Render the views eagerly once (which fixes the adaptive chunk schedule), then
// capture the same launch sequence as a hipGraph: clears as fill kernels, per chunk
// AD paths + children, per view primary edges + finalize into `images` (device).
// An overflowing sample cannot be replayed inside a graph, so every chunk's overflow
// count is summed into qcount[3]; replayViews checks it and falls back to the eager
// path, so a replay is always exact.
bool Scene::captureViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, float *const *images) {
    graph.reset();
    if (!renderViews(nv, cams, opts, images, true)) return false;
    if (last_replayed) return fail("capture: the schedule overflowed the child queue; render in smaller batches");
    RenderParams P{};
    const dtrt_render_options &o0 = opts[0];
    P.max_bounces = o0.max_bounces; P.sppse = std::max(0, o0.sppse);
    P.mode = o0.mode >= 0 ? o0.mode : 0x3F; P.ddist = o0.ddist_coeff;
    P.rec_stride = childRecordStride(nder, S.f64 != 0);
    P.ad_block = tune_ad_block;
    P.child_block = tune_child_block;
    P.tree_in_hbm = tree_in_hbm ? 1 : 0;
    std::vector<DView> hv;
    std::vector<uint64_t> pe_total;
    if (!setupViews(nv, cams, opts, hv, pe_total)) return false;
    HIPCHK(hipStreamSynchronize(stream));
    int64_t npix_all = 0;
    for (int v = 0; v < nv; ++v) npix_all += hv[v].npix;
    P.npix = (int)npix_all;
    const size_t npix = P.npix, img = (size_t)(nder + 1) * npix * 3;
    const uint64_t cap = std::max<uint64_t>(128, std::min<uint64_t>(queue_budget / P.rec_stride, 0xFFFFFFF0ull));
    const uint64_t pe_chunk = std::max<uint64_t>(64, std::min<uint64_t>(cap / 2 / 64 * 64, uint64_t(1) << 23));
    P.qcap = (unsigned)cap;
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned childGrid = (unsigned)std::max(1, ncu) * (unsigned)tune_child_blocks;
    unsigned *qc = static_cast<unsigned *>(qcount.p);
    unsigned *st = static_cast<unsigned *>(stats.p);
    unsigned long long *wk = static_cast<unsigned long long *>(work.p);
    void *A = acc.p, *APE = acc_pe.p;
    char *Q = static_cast<char *>(queue.p);
    uint4 *OV = static_cast<uint4 *>(ovf[0].p);
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    auto issue = [&]() -> hipError_t {
        hipError_t e;
#define GCHK(x) if ((e = (x)) != hipSuccess) return e
        GCHK(launch_fill32(A, img * accBytes() / 4, 0u, stream));
        GCHK(launch_fill32(APE, (size_t)nder * npix * 3 * accBytes() / 4, 0u, stream));
        GCHK(launch_fill32(st, 64, 0u, stream));
        GCHK(launch_fill32(wk, 32, 0u, stream));
        for (int v = 0; v < nv; ++v) GCHK(launch_fill32(images[v], (size_t)(nder + 1) * hv[v].npix * 3, 0u, stream));
        uint64_t s0 = 0;
        for (uint64_t n : last_chunks) {
            GCHK(launch_fill32(qc, 3, 0u, stream));
            GCHK(launch_ad_paths(S, P, s0, n, nullptr, A, Q, qc, st, OV, qc + 1, wk, stream));
            GCHK(launch_add_u32(qc + 3, qc + 1, stream));
            GCHK(launch_children(S, P, Q, qc, A, APE, wk, childGrid, stream));
            s0 += n;
        }
        for (int v = 0; v < nv; ++v)
            for (uint64_t p0 = 0; p0 < pe_total[v]; p0 += pe_chunk) {
                GCHK(launch_fill32(qc, 3, 0u, stream));
                GCHK(launch_primary(S, P, v, p0, std::min(pe_chunk, pe_total[v] - p0), pe_total[v], Q, qc, st, wk, stream));
                GCHK(launch_children(S, P, Q, qc, A, APE, wk, childGrid, stream));
            }
        for (int v = 0; v < nv; ++v)
            GCHK(launch_finalize(S, P, hv[v].npix, hv[v].pix0, hv[v].spp, pe_total[v] > 0, A, APE, images[v], st, stream));
#undef GCHK
        return hipSuccess;
    };
    hipError_t e = issue();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(std::string("capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    graph.g = g;
    HIPCHK(hipGraphInstantiate(&graph.x, g, nullptr, nullptr, 0));
    graph.cams.clear();
    for (int v = 0; v < nv; ++v) graph.cams.push_back(*cams[v]);
    graph.opts.assign(opts, opts + nv);
    graph.images.assign(images, images + nv);
    graph.hv = hv;
    graph.bufs = graphBuffers();
    return true;
}

// Replay the captured views with new seeds (one per view). Returns 0 after a graph
// replay, 1 if the fixed schedule overflowed and the views were rendered eagerly
// instead, -1 on error.
int Scene::replayViews(const uint64_t *seeds) {
    if (!graph.x) { g_err = "replay: no captured views"; return -1; }
    if (graphBuffers() != graph.bufs) { g_err = "replay: an eager render reallocated the buffers; capture again"; return -1; }
    if (!afterCaller()) return -1;
    const int nv = (int)graph.hv.size();
    for (int v = 0; v < nv; ++v) {
        graph.hv[v].seed = seeds[v];
        graph.opts[v].seed = seeds[v];
    }
    unsigned *qc = static_cast<unsigned *>(qcount.p);
    if (hipMemcpyAsync(viewbuf.p, graph.hv.data(), (size_t)nv * sizeof(DView), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemsetAsync(qc + 3, 0, 4, stream) != hipSuccess || hipEventRecord(ev[0], stream) != hipSuccess ||
        hipGraphLaunch(graph.x, stream) != hipSuccess || hipEventRecord(ev[1], stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        g_err = "replay: graph launch failed";
        return -1;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    unsigned flag = 0;
    unsigned hst[STAT_COUNT];
    if (hipMemcpy(&flag, qc + 3, 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hst, stats.p, sizeof(hst), hipMemcpyDeviceToHost) != hipSuccess) {
        g_err = "replay: readback failed";
        return -1;
    }
    if (flag) {   // some sample overflowed the queue: render these seeds eagerly (exact)
        std::vector<const Camera *> cp;
        for (auto &c : graph.cams) cp.push_back(&c);
        if (!renderViews(nv, cp.data(), graph.opts.data(), graph.images.data(), true)) return -1;
        return 1;
    }
    std::memset(kernel_ms, 0, sizeof(kernel_ms));
    std::memset(kernel_launches, 0, sizeof(kernel_launches));
    kernel_ms[0] = ms;
    kernel_launches[0] = 1;
    for (int i = 0; i < 6; ++i) last_stats[i] = hst[i];
    last_stats[6] = 0;
    return 0;
}

// d loss / d I of every view into stepbuf (views concatenated like the adjoint's
// loss-gradient input), the loss into a.loss, the lanes' gradient into a.pgrad.
bool Scene::stepLoss(StepArgs &a, const std::vector<DView> &hv, hipStream_t sm) {
    float *dimg = static_cast<float *>(stepbuf.p);
    for (size_t v = 0; v < hv.size(); ++v) {
        const uint64_t n = (uint64_t)hv[v].npix * 3;
        float *dv = dimg + (size_t)hv[v].pix0 * 3;
        HIPCHK(launch_l2_grad(a.images[v], a.targets[v], dv, n, a.loss, sm));
        if (a.pgrad) HIPCHK(launch_contract(a.images[v], dv, n, nder, a.pgrad, sm));
    }
    return true;
}

static bool stepOptionsValid(const Scene::StepArgs &a) {
    for (size_t v = 0; v < a.fopts.size(); ++v)
        if (!a.images[v] || !a.targets[v] || a.fopts[v].spp <= 0 || a.aopts[v].spp <= 0)
            return fail("step: every view needs an image, a target and positive spp");
    return true;
}

bool Scene::runStep(StepArgs &a) {
    const int nv = (int)a.cams.size();
    if (!gridMedium(a.med) || !a.dgrid) return false;
    if (!stepOptionsValid(a)) return false;
    std::vector<const Camera *> cp;
    for (auto &c : a.cams) cp.push_back(&c);
    if (!afterCaller()) return false;
    for (int v = 0; v < nv; ++v) {   // the forward render adds into zeroed images
        const Camera &c = a.cams[v];
        const uint64_t np = c.cropped() ? (uint64_t)c.cw * c.ch : (uint64_t)c.w * c.h;
        HIPCHK(launch_fill32(a.images[v], (uint64_t)(nder + 1) * np * 3, 0u, stream));
    }
    if (!renderViews(nv, cp.data(), a.fopts.data(), a.images.data(), true)) return false;
    std::vector<DView> hv;
    std::vector<uint64_t> pe_total;
    if (!setupViews(nv, cp.data(), a.fopts.data(), hv, pe_total)) return false;
    int64_t npix = 0;
    for (auto &d : hv) npix += d.npix;
    if (!ensure(stepbuf, (size_t)npix * 3 * 4 + (size_t)nv * sizeof(DView) + 512)) return false;
    const DMedium &m = hmed[a.med];
    HIPCHK(launch_fill32(a.dgrid, (uint64_t)m.res[0] * m.res[1] * m.res[2], 0u, stream));
    if (a.pgrad) HIPCHK(launch_fill32(a.pgrad, 2 * (uint64_t)nder, 0u, stream));
    if (a.loss) HIPCHK(launch_fill32(a.loss, 2, 0u, stream));
    if (!stepLoss(a, hv, stream)) return false;
    return renderAdjointViews(nv, cp.data(), a.aopts.data(), a.med, static_cast<const float *>(stepbuf.p), a.dgrid);
}

// The eager step fixes the forward chunk schedule; the graph then holds: clears, the
// forward chunks and primary-edge passes, finalize, the loss gradient and lane
// contraction, the grid clear and the adjoint chunks. The forward views keep the view
// table in viewbuf, the adjoint's lies in stepbuf, so replays rewrite both.
bool Scene::captureStep(StepArgs &a) {
    stepg.reset();
    if (!runStep(a)) return false;
    if (last_replayed) return fail("capture_step: the schedule overflowed the child queue; render in smaller batches");
    const int nv = (int)a.cams.size();
    std::vector<const Camera *> cp;
    for (auto &c : a.cams) cp.push_back(&c);
    RenderParams P{};
    const dtrt_render_options &o0 = a.fopts[0];
    P.ad_kind = AD_VOLPATH;
    P.max_bounces = o0.max_bounces; P.sppse = std::max(0, o0.sppse);
    P.mode = o0.mode >= 0 ? o0.mode : 0x3F; P.ddist = o0.ddist_coeff;
    P.rec_stride = childRecordStride(nder, S.f64 != 0);
    P.ad_block = tune_ad_block;
    P.child_block = tune_child_block;
    P.tree_in_hbm = tree_in_hbm ? 1 : 0;
    std::vector<DView> hvf, hva;
    std::vector<uint64_t> pe_total, pe_unused;
    if (!setupViews(nv, cp.data(), a.fopts.data(), hvf, pe_total)) return false;
    const DScene Sf = S;
    int64_t npix_all = 0;
    for (int v = 0; v < nv; ++v) npix_all += hvf[v].npix;
    P.npix = (int)npix_all;
    char *base = static_cast<char *>(stepbuf.p);
    Bump apool{base + ((size_t)npix_all * 3 * 4 + 255) / 256 * 256, (uint64_t)nv * sizeof(DView) + 256};
    RenderParams PA{};
    uint64_t atotal = 0;
    if (!renderAdjointViews(nv, cp.data(), a.aopts.data(), a.med, static_cast<const float *>(stepbuf.p), a.dgrid, &apool,
                            false, &PA, &atotal))
        return false;
    const DScene Sa = S;
    hva.assign(nv, DView{});
    HIPCHK(hipMemcpyAsync(hva.data(), Sa.views, (size_t)nv * sizeof(DView), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    S = Sf;
    const size_t npix = P.npix, img = (size_t)(nder + 1) * npix * 3;
    const uint64_t cap = std::max<uint64_t>(128, std::min<uint64_t>(queue_budget / P.rec_stride, 0xFFFFFFF0ull));
    const uint64_t pe_chunk = std::max<uint64_t>(64, std::min<uint64_t>(cap / 2 / 64 * 64, uint64_t(1) << 23));
    P.qcap = (unsigned)cap;
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned childGrid = (unsigned)std::max(1, ncu) * (unsigned)tune_child_blocks;
    unsigned *qc = static_cast<unsigned *>(qcount.p);
    unsigned *st = static_cast<unsigned *>(stats.p);
    unsigned long long *wk = static_cast<unsigned long long *>(work.p);
    void *A = acc.p, *APE = acc_pe.p;
    char *Q = static_cast<char *>(queue.p);
    uint4 *OV = static_cast<uint4 *>(ovf[0].p);
    const DMedium &m = hmed[a.med];
    const uint64_t ngrid = (uint64_t)m.res[0] * m.res[1] * m.res[2];
    const uint64_t achunk = uint64_t(1) << 22;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    auto issue = [&]() -> hipError_t {
        hipError_t e;
#define GCHK(x) if ((e = (x)) != hipSuccess) return e
        GCHK(launch_fill32(A, img * accBytes() / 4, 0u, stream));
        GCHK(launch_fill32(APE, (size_t)nder * npix * 3 * accBytes() / 4, 0u, stream));
        GCHK(launch_fill32(st, 64, 0u, stream));
        GCHK(launch_fill32(wk, 32, 0u, stream));
        for (int v = 0; v < nv; ++v) GCHK(launch_fill32(a.images[v], (size_t)(nder + 1) * hvf[v].npix * 3, 0u, stream));
        uint64_t s0 = 0;
        for (uint64_t n : last_chunks) {
            GCHK(launch_fill32(qc, 3, 0u, stream));
            GCHK(launch_ad_paths(Sf, P, s0, n, nullptr, A, Q, qc, st, OV, qc + 1, wk, stream));
            GCHK(launch_add_u32(qc + 3, qc + 1, stream));
            GCHK(launch_children(Sf, P, Q, qc, A, APE, wk, childGrid, stream));
            s0 += n;
        }
        for (int v = 0; v < nv; ++v)
            for (uint64_t p0 = 0; p0 < pe_total[v]; p0 += pe_chunk) {
                GCHK(launch_fill32(qc, 3, 0u, stream));
                GCHK(launch_primary(Sf, P, v, p0, std::min(pe_chunk, pe_total[v] - p0), pe_total[v], Q, qc, st, wk, stream));
                GCHK(launch_children(Sf, P, Q, qc, A, APE, wk, childGrid, stream));
            }
        for (int v = 0; v < nv; ++v)
            GCHK(launch_finalize(Sf, P, hvf[v].npix, hvf[v].pix0, hvf[v].spp, pe_total[v] > 0, A, APE, a.images[v], st, stream));
        if (a.pgrad) GCHK(launch_fill32(a.pgrad, 2 * (uint64_t)nder, 0u, stream));
        if (a.loss) GCHK(launch_fill32(a.loss, 2, 0u, stream));
        float *dimg = static_cast<float *>(stepbuf.p);
        for (int v = 0; v < nv; ++v) {
            const uint64_t n = (uint64_t)hvf[v].npix * 3;
            GCHK(launch_l2_grad(a.images[v], a.targets[v], dimg + (size_t)hvf[v].pix0 * 3, n, a.loss, stream));
            if (a.pgrad) GCHK(launch_contract(a.images[v], dimg + (size_t)hvf[v].pix0 * 3, n, nder, a.pgrad, stream));
        }
        GCHK(launch_fill32(a.dgrid, ngrid, 0u, stream));
        for (uint64_t a0 = 0; a0 < atotal; a0 += achunk)
            GCHK(launch_adjoint(Sa, PA, a0, std::min(achunk, atotal - a0), dimg, a.dgrid, a.med, st + 32, wk, stream));
#undef GCHK
        return hipSuccess;
    };
    hipError_t e = issue();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(std::string("capture_step: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    stepg.g = g;
    HIPCHK(hipGraphInstantiate(&stepg.x, g, nullptr, nullptr, 0));
    stepg.args = a;
    stepg.hvf = hvf;
    stepg.hva = hva;
    stepg.bufs = graphBuffers();
    stepg.bufs.push_back(stepbuf.p);
    return true;
}

// Replay the captured step with new seeds: seeds[2v] forward, seeds[2v + 1] adjoint
// of view v. 0 after a graph launch, 1 if the forward schedule overflowed the queue
// and the step ran eagerly instead (exact either way), -1 on error.
int Scene::replayStep(const uint64_t *seeds) {
    if (!stepg.x) { g_err = "replay_step: no captured step"; return -1; }
    std::vector<void *> now = graphBuffers();
    now.push_back(stepbuf.p);
    if (now != stepg.bufs) { g_err = "replay_step: an eager render reallocated the buffers; capture again"; return -1; }
    if (!afterCaller()) return -1;
    const int nv = (int)stepg.hvf.size();
    for (int v = 0; v < nv; ++v) {
        stepg.hvf[v].seed = stepg.args.fopts[v].seed = seeds[2 * v];
        stepg.hva[v].seed = stepg.args.aopts[v].seed = seeds[2 * v + 1];
    }
    char *base = static_cast<char *>(stepbuf.p);
    int64_t npix_all = 0;
    for (auto &d : stepg.hvf) npix_all += d.npix;
    void *avt = base + ((size_t)npix_all * 3 * 4 + 255) / 256 * 256;
    unsigned *qc = static_cast<unsigned *>(qcount.p);
    if (ev.empty()) {
        ev.resize(2);
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    }
    if (hipMemcpyAsync(viewbuf.p, stepg.hvf.data(), (size_t)nv * sizeof(DView), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(avt, stepg.hva.data(), (size_t)nv * sizeof(DView), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemsetAsync(qc + 3, 0, 4, stream) != hipSuccess || hipEventRecord(ev[0], stream) != hipSuccess ||
        hipGraphLaunch(stepg.x, stream) != hipSuccess || hipEventRecord(ev[1], stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        g_err = "replay_step: graph launch failed";
        return -1;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    unsigned flag = 0;
    if (hipMemcpy(&flag, qc + 3, 4, hipMemcpyDeviceToHost) != hipSuccess) { g_err = "replay_step: readback failed"; return -1; }
    if (flag) {   // some forward sample overflowed the queue: run these seeds eagerly (exact)
        StepArgs a = stepg.args;
        if (!runStep(a)) return -1;
        return 1;
    }
    std::memset(kernel_ms, 0, sizeof(kernel_ms));
    std::memset(kernel_launches, 0, sizeof(kernel_launches));
    kernel_ms[0] = ms;
    kernel_launches[0] = 1;
    return 0;
}

// Replace the density grid of medium `med` (same resolution): negative values are
// clamped to 0 as the VOL reader does, and the majorant is recomputed.
bool Scene::setDensity(int med, const float *density, bool on_device) {
    if (!gridMedium(med) || !density) return fail("set_density: bad arguments");
    if (on_device && !afterCaller()) return false;
    DMedium &d = hmed[med];
    const uint64_t n = (uint64_t)d.res[0] * d.res[1] * d.res[2];
    float *dst = const_cast<float *>(d.density);
    if (!ensure(stats, 64 * 4)) return false;
    unsigned *mx = static_cast<unsigned *>(stats.p) + 48;
    HIPCHK(hipMemsetAsync(mx, 0, 4, stream));
    if (!on_device) HIPCHK(hipMemcpyAsync(dst, density, n * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(launch_load_density(dst, on_device ? density : dst, n, mx, stream));
    unsigned bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, mx, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    float m;
    std::memcpy(&m, &bits, 4);
    if (!(m > 0.f)) return fail("set_density: density grid is all zero");
    d.maxDensity = m;
    d.maxInvDensity = (float)(1.0 / ((double)d.scale * m));
    return uploadMedia();
}

bool Scene::setVoxelLanes(int med, const int *vox, int n) {
    if (!gridMedium(med) || n < 0 || n > nder || (n > 0 && !vox)) return fail("voxel lanes: bad arguments");
    DMedium &d = hmed[med];
    const int64_t nv = (int64_t)d.res[0] * d.res[1] * d.res[2];
    d.vox_any = 0;
    for (int j = 0; j < DTRT_NDER_MAX; ++j) {
        int v = j < n ? vox[j] : -1;
        if (v >= nv) return fail("voxel lanes: node index outside the grid");
        d.vox_lane[j] = v < 0 ? -1 : v;
        d.vox_any |= v >= 0;
    }
    return uploadMedia();
}

// Path-replay adjoint of the interior terms (emission, surface, volume) w.r.t. the
// node densities of medium `med`; dimg = d loss / d image value [npix][3], grad
// accumulates [z][y][x]. Both are device pointers.
// Views adjoint: dimg holds the views' loss gradients back to back ([h][w][3] each, in
// view order); every view's samples replay the streams of its forward render.
bool Scene::renderAdjointViews(int nv, const Camera *const *cams, const dtrt_render_options *opts, int med,
                               const float *dimg, float *grad, Bump *pool, bool launch, RenderParams *Pout,
                               uint64_t *total_out) {
    if (!gridMedium(med) || !dimg || !grad || nv <= 0) return false;
    if (launch && !afterCaller()) return false;
    const dtrt_render_options &o = opts[0];
    RenderParams P{};
    P.max_bounces = o.max_bounces; P.sppse = 0;
    P.mode = (o.mode >= 0 ? o.mode : 0x3F) & (1 | 2 | 8);
    P.tree_in_hbm = tree_in_hbm ? 1 : 0;
    P.ddist = 0.f;
    std::vector<dtrt_render_options> ov(opts, opts + nv);
    for (auto &x : ov) {
        if (x.spp <= 0 || x.max_bounces != o.max_bounces || x.mode != o.mode || x.max_bounces < 0 || x.max_bounces > 200)
            return fail("adjoint: bad options (views must share max_bounces and mode)");
        x.sppe = x.sppe < 0 ? -1 : 0;   // pixel jitter as in the forward render; no primary-edge pass
    }
    std::vector<DView> hv;
    std::vector<uint64_t> pe_total;
    if (!setupViews(nv, cams, ov.data(), hv, pe_total, pool)) return false;
    uint64_t total = 0;
    P.npix = 0;
    for (int v = 0; v < nv; ++v) { P.npix += hv[v].npix; total += (uint64_t)hv[v].npix * hv[v].spp; }
    if (Pout) *Pout = P;
    if (total_out) *total_out = total;
    if (!launch) return true;
    if (!ensure(stats, 64 * 4) || !ensure(work, 16 * 8)) return false;
    HIPCHK(hipMemsetAsync(work.p, 0, 16 * 8, stream));
    if (ev.empty()) {
        ev.resize(2);
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    }
    unsigned *st = static_cast<unsigned *>(stats.p) + 32;   // warnings of the replays are not reported
    unsigned long long *wk = static_cast<unsigned long long *>(work.p);
    const uint64_t chunk = uint64_t(1) << 22;
    std::memset(kernel_ms, 0, sizeof(kernel_ms));
    std::memset(kernel_launches, 0, sizeof(kernel_launches));
    HIPCHK(hipEventRecord(ev[0], stream));
    for (uint64_t s0 = 0; s0 < total; s0 += chunk)
        HIPCHK(launch_adjoint(S, P, s0, std::min(chunk, total - s0), dimg, grad, med, st, wk, stream));
    HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(hipEventSynchronize(ev[1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    kernel_ms[0] = ms;
    kernel_launches[0] = (int64_t)((total + chunk - 1) / chunk);
    unsigned long long hw[12];
    HIPCHK(hipMemcpy(hw, wk, sizeof(hw), hipMemcpyDeviceToHost));
    for (int i = 0; i < 12; ++i) work_counts[i] = hw[i];
    return true;
}

// The primal integrators' render (volpath.cpp:106-141, volpath_simple.cpp:89-124,
// path.cpp:64-96, direct.cpp:51-85, ptracer.cpp:10-55): float32[H][W][3] (crop size if
// cropped), every pixel written. integrator: 1 VolPathTracer, 2 VolPathTracerSimple,
// 3 PathTracer, 4 DirectIntegrator (each pixel the mean of spp jittered samples);
// 5 ParticleTracer (100000 * spp light particles, splatted and divided by their count).
bool Scene::renderPrimal(const dtrt_render_options &o, float *out, bool out_is_device, int integrator) {
    if (o.spp <= 0 || o.max_bounces < 0 || o.max_bounces > 200) return fail("render: bad options");
    if (!afterCaller()) return false;
    RenderParams P{};
    P.max_bounces = integrator == 4 ? 1 : o.max_bounces;
    const int kind = integrator == 2 ? PRIMAL_SIMPLE : integrator >= 3 ? PRIMAL_PATH : PRIMAL_VOLPATH;
    dtrt_render_options ov = o;
    ov.sppe = 0;   // no primary-edge pass
    const Camera *c = &cam;
    std::vector<DView> hv;
    std::vector<uint64_t> pe_total;
    if (!setupViews(1, &c, &ov, hv, pe_total)) return false;
    P.npix = hv[0].npix;
    const size_t n3 = (size_t)P.npix * 3;
    if (!ensure(acc, n3 * accBytes()) || !ensure(work, 16 * 8)) return false;
    float *dout = out;
    if (!out_is_device) {
        if (!ensure(outbuf, n3 * 4)) return false;
        dout = static_cast<float *>(outbuf.p);
    }
    HIPCHK(hipMemsetAsync(acc.p, 0, n3 * accBytes(), stream));
    HIPCHK(hipMemsetAsync(work.p, 0, 16 * 8, stream));
    if (ev.empty()) {
        ev.resize(2);
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    }
    unsigned long long *wk = static_cast<unsigned long long *>(work.p);
    void *A = acc.p;
    // ParticleTracer: spp blocks of 100000 particles (ptracer.cpp:11-12); others: spp per pixel
    const uint64_t total = integrator == 5 ? (uint64_t)100000 * o.spp : (uint64_t)P.npix * o.spp;
    const uint64_t chunk = uint64_t(1) << (integrator == 5 ? 22 : 26);
    std::memset(kernel_ms, 0, sizeof(kernel_ms));
    std::memset(kernel_launches, 0, sizeof(kernel_launches));
    HIPCHK(hipEventRecord(ev[0], stream));
    for (uint64_t s0 = 0; s0 < total; s0 += chunk) {
        if (integrator == 5) HIPCHK(launch_particles(S, P, s0, std::min(chunk, total - s0), A, wk, stream));
        else HIPCHK(launch_primal_paths(S, P, kind, s0, std::min(chunk, total - s0), A, wk, stream));
    }
    HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(launch_scale(S, A, dout, n3, 1.0 / (double)(integrator == 5 ? total : o.spp), stream));
    if (!out_is_device) HIPCHK(hipMemcpyAsync(out, dout, n3 * 4, hipMemcpyDeviceToHost, stream));
    unsigned long long hw[12];
    HIPCHK(hipMemcpyAsync(hw, wk, sizeof(hw), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    kernel_ms[0] = ms;
    kernel_launches[0] = (int64_t)((total + chunk - 1) / chunk);
    for (int i = 0; i < 12; ++i) work_counts[i] = hw[i];
    for (int i = 0; i < 10; ++i) last_stats[i] = 0;
    last_stats[7] = (int64_t)total;
    return true;
}

// Coordinate descent over the launch settings, each candidate timed on one full render
// of `o` (after one warm-up): persistent child blocks per CU, then the queue fill target
// of the AD chunks, then the largest AD chunk. The fastest values stay in the scene.
// settings3 = {child blocks per CU, fill target in 1/100, max AD chunk}.
bool Scene::autotune(const dtrt_render_options &o, int ad_kind, double *best_ms, int64_t *settings3) {
    const size_t npix = cam.cropped() ? (size_t)cam.cw * cam.ch : (size_t)cam.w * cam.h;
    if (!ensure(tunebuf, (size_t)(nder + 1) * npix * 3 * 4)) return false;
    float *out = static_cast<float *>(tunebuf.p);
    auto run = [&](double &ms) -> bool {
        auto t0 = std::chrono::steady_clock::now();
        if (!render(o, out, true, ad_kind)) return false;
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return true;
    };
    double best = 0;
    if (!run(best) || !run(best)) return false;   // warm-up, then the current settings
    const int blocks[] = {16, 24, 32, 48};
    const double fills[] = {0.5, 0.75, 0.9};
    const uint64_t chunks[] = {uint64_t(1) << 21, uint64_t(1) << 22, uint64_t(1) << 23};
    const size_t budgets[] = {size_t(1) << 30, size_t(2) << 30, size_t(4) << 30, size_t(6) << 30};
    const int groups[] = {64, 128, 256};
    auto tryKnob = [&](auto &knob, auto value) -> bool {   // keep `value` if it renders faster
        if (knob == value) return true;
        const auto keep = knob;
        knob = value;
        double ms;
        if (!run(ms)) return false;
        if (ms < best) best = ms; else knob = keep;
        return true;
    };
    for (size_t q : budgets)
        if (!tryKnob(queue_budget, q)) return false;
    for (int g : groups)
        if (!tryKnob(tune_ad_block, g)) return false;
    for (int g : groups)
        if (!tryKnob(tune_child_block, g)) return false;
    for (int b : blocks) {
        if (b == tune_child_blocks) continue;
        const int keep = tune_child_blocks;
        tune_child_blocks = b;
        double ms;
        if (!run(ms)) return false;
        if (ms < best) best = ms; else tune_child_blocks = keep;
    }
    for (double f : fills) {
        if (f == tune_fill) continue;
        const double keep = tune_fill;
        tune_fill = f;
        double ms;
        if (!run(ms)) return false;
        if (ms < best) best = ms; else tune_fill = keep;
    }
    for (uint64_t c : chunks) {
        if (c == tune_max_chunk) continue;
        const uint64_t keep = tune_max_chunk;
        tune_max_chunk = c;
        double ms;
        if (!run(ms)) return false;
        if (ms < best) best = ms; else tune_max_chunk = keep;
    }
    if (best_ms) *best_ms = best;
    if (settings3) {
        settings3[0] = tune_child_blocks;
        settings3[1] = (int64_t)std::lround(tune_fill * 100.0);
        settings3[2] = (int64_t)tune_max_chunk;
    }
    return true;
}

template <class T> T *as(dtrt_object *o, int tag) {
    Obj *b = reinterpret_cast<Obj *>(o);
    if (!b || b->tag != tag) { g_err = "wrong or null object handle"; return nullptr; }
    return static_cast<T *>(b);
}
dtrt_object *handle(Obj *o) { return reinterpret_cast<dtrt_object *>(o); }

void fillSpecD(float (*dst)[3], const float *src, int nd) {
    for (int j = 0; j < nd; ++j) for (int k = 0; k < 3; ++k) dst[j][k] = src ? src[3 * j + k] : 0.f;
}
void fillScalarD(float *dst, const float *src, int nd) { for (int j = 0; j < nd; ++j) dst[j] = src ? src[j] : 0.f; }

bool readVol(const char *fn, std::vector<float> &data, int *res, int &ch, float *mn, float *mx) {
    FILE *f = std::fopen(fn, "rb");
    if (!f) return fail(std::string("cannot open volume file ") + fn);
    char hdr[3];
    uint8_t ver = 0;
    int type = 0;
    bool ok = std::fread(hdr, 1, 3, f) == 3 && std::strncmp(hdr, "VOL", 3) == 0 && std::fread(&ver, 1, 1, f) == 1 &&
              ver == 3 && std::fread(&type, 4, 1, f) == 1 && type == 1 && std::fread(res, 4, 3, f) == 3 &&
              std::fread(&ch, 4, 1, f) == 1 && (ch == 1 || ch == 3) && res[0] > 1 && res[1] > 1 && res[2] > 1 &&
              std::fread(mn, 4, 3, f) == 3 && std::fread(mx, 4, 3, f) == 3;
    if (ok) {
        // the header's voxel count must match the bytes that follow it before anything
        // is allocated: a corrupt header must not ask for an arbitrary amount of memory
        const long here = std::ftell(f);
        ok = here >= 0 && std::fseek(f, 0, SEEK_END) == 0;
        const long end = ok ? std::ftell(f) : -1;
        ok = ok && end >= here && std::fseek(f, here, SEEK_SET) == 0;
        const uint64_t n = (uint64_t)res[0] * (uint64_t)res[1] * (uint64_t)res[2] * (uint64_t)ch;
        ok = ok && n <= ((uint64_t)1 << 36) && n * 4 <= (uint64_t)(end - here);
        if (ok) {
            data.resize((size_t)n);
            ok = std::fread(data.data(), 4, (size_t)n, f) == (size_t)n;
        }
    }
    std::fclose(f);
    if (!ok) return fail(std::string("invalid or truncated VOL v3 file ") + fn);
    for (float &x : data) x = std::max(x, 0.0f);
    return true;
}

// heterogeneous.h:13-106, gridvolume.h:49-141
bool finishHetero(Medium &m, const float *to_world, float scale, const float *albedo3, int phase_id, const float *trans,
                  const float *rot, const float *dscale, const float *dalb, const float *bmin, const float *bmax) {
    DMedium &d = m.d;
    int nd = nderNow();
    d.type = MED_HETEROGENEOUS;
    d.phase_id = phase_id;
    double v2w[4][4], w2v[4][4];
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) v2w[i][j] = to_world ? to_world[4 * i + j] : (i == j);
    if (std::fabs(v2w[3][0]) > 1e-7 || std::fabs(v2w[3][1]) > 1e-7 || std::fabs(v2w[3][2]) > 1e-7 || std::fabs(v2w[3][3] - 1) > 1e-7)
        return fail("heterogeneous: invalid to-world transformation");
    if (!invert4(v2w, w2v)) return fail("heterogeneous: singular to-world transformation");
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) d.w2v[4 * i + j] = (float)w2v[i][j];
    float mx = 0;
    for (float x : m.density) mx = std::max(mx, x);
    if (!(mx > 0)) return fail("heterogeneous: density grid is all zero");
    d.maxDensity = mx;
    d.scale = scale;
    d.maxInvDensity = (float)(1.0 / ((double)scale * mx));
    for (int a = 0; a < 3; ++a) {
        d.bmin[a] = bmin[a];
        d.gscale[a] = (float)((double)(d.res[a] - 1) / ((double)bmax[a] - bmin[a]));
        d.albedo[a] = albedo3 ? albedo3[a] : 0.f;
    }
    fillScalarD(d.scale_d, dscale, nd);
    fillSpecD(d.albedo_d, dalb, nd);
    fillSpecD(d.trans_d, trans, nd);
    fillSpecD(d.rot_d, rot, nd);
    double cx = 0.5 * ((double)bmin[0] + bmax[0]), cy = 0.5 * ((double)bmin[1] + bmax[1]), cz = 0.5 * ((double)bmin[2] + bmax[2]);
    for (int i = 0; i < 3; ++i) d.center[i] = (float)(v2w[i][0] * cx + v2w[i][1] * cy + v2w[i][2] * cz + v2w[i][3]);
    for (int i = 0; i < 3; ++i) { d.aabb_min[i] = INFINITY; d.aabb_max[i] = -INFINITY; }
    for (int c = 0; c < 8; ++c) {
        double p[3] = {c & 1 ? bmax[0] : bmin[0], c & 2 ? bmax[1] : bmin[1], c & 4 ? bmax[2] : bmin[2]};
        for (int i = 0; i < 3; ++i) {
            float w = (float)(v2w[i][0] * p[0] + v2w[i][1] * p[1] + v2w[i][2] * p[2] + v2w[i][3]);
            d.aabb_min[i] = std::min(d.aabb_min[i], w);
            d.aabb_max[i] = std::max(d.aabb_max[i], w);
        }
    }
    return true;
}


template <class T> bool devTmp(std::vector<void *> &keep, T *&p, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) { g_err = "hipMalloc failed"; return false; }
    keep.push_back(q);
    p = static_cast<T *>(q);
    return true;
}
struct TmpFree {
    std::vector<void *> v;
    ~TmpFree() { for (void *p : v) (void)hipFree(p); }
};

// Runs f and turns any C++ exception into an error message and the error return:
// nothing thrown crosses the C-ABI.
template <class R, class F> R guarded(R on_error, F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        g_err = "out of host memory";
    } catch (const std::exception &e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown error";
    }
    return on_error;
}

}  // namespace

extern "C" {

int dtrt_abi_version(void) { return 1; }
int dtrt_set_nder(int n) {
    if (!(n == 1 || n == 2 || n == 3 || n == 4 || n == 8 || n == 16)) { g_err = "nder must be one of 1, 2, 3, 4, 8, 16"; return -1; }
    g_nder = n;
    return 0;
}
int dtrt_nder(void) { return nderNow(); }
int dtrt_set_precision(int bits) {
    if (bits != 32 && bits != 64) { g_err = "precision must be 32 or 64 bits"; return -1; }
    g_prec = bits;
    return 0;
}
int dtrt_precision(void) { return precisionNow(); }
double dtrt_angle_eps(void) { return (double)1e-4f; }
const char *dtrt_last_error(void) { return g_err.c_str(); }
void dtrt_destroy(dtrt_object *o) { delete reinterpret_cast<Obj *>(o); }

dtrt_object *dtrt_camera_create(int w, int h, const float *c2w, const float *c2ndc, float clip_near, int med_id, const float *vel) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!c2w || !c2ndc || w <= 0 || h <= 0) { g_err = "camera: bad arguments"; return nullptr; }
        auto c = std::make_unique<Camera>();
        int nd = nderNow();
        c->w = w; c->h = h; c->clip_near = clip_near; c->med = med_id; c->nder = nd;
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) c->c2w[i][j] = c2w[4 * i + j];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c->c2ndc[i][j] = c2ndc[3 * i + j];
        if (c->c2ndc[0][0] != c->c2ndc[1][1]) { g_err = "camera: cam_to_ndc[0][0] must equal cam_to_ndc[1][1]"; return nullptr; }
        if (!invert4(c->c2w, c->w2c)) { g_err = "camera: singular cam_to_world"; return nullptr; }
        double m4[4][4] = {{0}}, i4[4][4];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m4[i][j] = c->c2ndc[i][j];
        m4[3][3] = 1;
        if (!invert4(m4, i4)) { g_err = "camera: singular cam_to_ndc"; return nullptr; }
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c->ndc2c[i][j] = i4[i][j];
        c->cpos = D3{c->c2w[0][3], c->c2w[1][3], c->c2w[2][3]};
        c->cs = D3{c->c2w[0][0], c->c2w[1][0], c->c2w[2][0]};
        c->ct = D3{c->c2w[0][1], c->c2w[1][1], c->c2w[2][1]};
        c->cn = D3{c->c2w[0][2], c->c2w[1][2], c->c2w[2][2]};
        for (int j = 0; j < nd; ++j) {   // camera.cpp:92-111
            D3 dx = vel ? D3{vel[3 * j], vel[3 * j + 1], vel[3 * j + 2]} : D3{0, 0, 0};
            D3 dw = vel ? D3{vel[3 * nd + 3 * j], vel[3 * nd + 3 * j + 1], vel[3 * nd + 3 * j + 2]} : D3{0, 0, 0};
            c->cpos_d[j] = dx;
            c->cs_d[j] = cross3(dw, c->cs);
            c->ct_d[j] = cross3(dw, c->ct);
            c->cn_d[j] = cross3(dw, c->cn);
        }
        return handle(c.release());
    });
}
int dtrt_camera_set_rect(dtrt_object *cam, int ox, int oy, int w, int h) {
    return guarded<int>(-1, [&]() -> int {
        Camera *c = as<Camera>(cam, T_CAMERA);
        if (!c) return -1;
        c->ox = ox; c->oy = oy; c->cw = w; c->ch = h;
        c->uid = ++g_camera_uid;   // the crop clips the primary edges
        if (c->cropped() && (ox < 0 || oy < 0 || ox + w > c->w || oy + h > c->h)) { g_err = "set_rect: crop outside the image"; return -1; }
        return 0;
    });
}

dtrt_object *dtrt_shape_create(const float *v, const int *idx, const float *uvs, const float *n, int nv, int nf, int light_id,
                               int bsdf_id, int med_int, int med_ext, const float *vel) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        (void)uvs;   // no BSDF on the path reads texture coordinates
        if (!v || !idx || nv <= 0 || nf <= 0) { g_err = "shape: bad arguments"; return nullptr; }
        auto s = std::make_unique<Shape>();
        int nd = nderNow();
        s->nv = nv; s->nf = nf; s->nder = nd; s->light_id = light_id; s->bsdf_id = bsdf_id; s->med_int = med_int; s->med_ext = med_ext;
        s->v.resize(nv);
        for (int i = 0; i < nv; ++i) s->v[i] = D3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        s->f.assign(idx, idx + 3 * (size_t)nf);
        for (int x : s->f) if (x < 0 || x >= nv) { g_err = "shape: vertex index out of range"; return nullptr; }
        if (n) {
            s->n.resize(nv);
            for (int i = 0; i < nv; ++i) {
                D3 q{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
                s->n[i] = q * (1.0 / norm3(q));
            }
        }
        s->vd.assign((size_t)nv * nd, D3{0, 0, 0});
        s->nd.assign(n ? (size_t)nv * nd : 0, D3{0, 0, 0});
        if (vel) setShapeVelocities(*s, vel);
        computeFaceNormals(*s);
        if (!constructEdges(*s)) return nullptr;
        return handle(s.release());
    });
}

dtrt_object *dtrt_bsdf_diffuse_create(const float *rgb, const float *d) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!rgb) { g_err = "BSDF_diffuse: null reflectance"; return nullptr; }
        auto b = std::make_unique<Bsdf>();
        b->d.type = BSDF_DIFFUSE;
        for (int k = 0; k < 3; ++k) b->d.refl[k] = rgb[k];
        fillSpecD(b->d.refl_d, d, nderNow());
        return handle(b.release());
    });
}
dtrt_object *dtrt_bsdf_null_create(void) { auto b = std::make_unique<Bsdf>(); b->d.type = BSDF_NULL; return handle(b.release()); }
dtrt_object *dtrt_bsdf_phong_create(const float *kd, const float *ks, float ex, const float *dkd, const float *dks, const float *dex) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!kd || !ks) { g_err = "BSDF_Phong: null reflectance"; return nullptr; }
        auto b = std::make_unique<Bsdf>();
        b->d.type = BSDF_PHONG;
        for (int k = 0; k < 3; ++k) { b->d.kd[k] = kd[k]; b->d.ks[k] = ks[k]; }
        b->d.ex = ex;
        double v0 = ((double)kd[0] + kd[1] + kd[2]) / 3, v1 = ((double)ks[0] + ks[1] + ks[2]) / 3;
        b->d.specW = (float)(v1 / (v0 + v1));
        int nd = nderNow();
        fillSpecD(b->d.kd_d, dkd, nd);
        fillSpecD(b->d.ks_d, dks, nd);
        fillScalarD(b->d.ex_d, dex, nd);
        return handle(b.release());
    });
}
dtrt_object *dtrt_bsdf_roughdielectric_create(float alpha, float int_ior, float ext_ior, const float *dalpha, const float *deta) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        auto b = std::make_unique<Bsdf>();
        b->d.type = BSDF_ROUGHDIELECTRIC;
        b->d.alpha = alpha < 1e-4f ? 1e-4f : alpha;   // microfacet.h:11-14
        b->d.eta = int_ior / ext_ior;
        int nd = nderNow();
        fillScalarD(b->d.alpha_d, dalpha, nd);
        fillScalarD(b->d.eta_d, deta, nd);
        return handle(b.release());
    });
}
dtrt_object *dtrt_phase_hg_create(float g, const float *dg) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        auto p = std::make_unique<Phase>();
        p->d.type = PHASE_HG;
        p->d.g = g;
        fillScalarD(p->d.g_d, dg, nderNow());
        return handle(p.release());
    });
}
dtrt_object *dtrt_phase_isotropic_create(void) { auto p = std::make_unique<Phase>(); p->d.type = PHASE_ISO; return handle(p.release()); }
dtrt_object *dtrt_area_light_create(int shape_id, const float *I, int two_sided) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!I) { g_err = "AreaLight: null intensity"; return nullptr; }
        auto e = std::make_unique<Emitter>();
        e->d.type = EMIT_AREA; e->d.shape_id = shape_id; e->d.two_sided = two_sided != 0;
        for (int k = 0; k < 3; ++k) e->d.I[k] = I[k];
        return handle(e.release());
    });
}
dtrt_object *dtrt_area_light_ex_create(int shape_id, const float *I, float kappa, const float *dI, const float *dk) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!I) { g_err = "AreaLightEx: null intensity"; return nullptr; }
        auto e = std::make_unique<Emitter>();
        e->d.type = EMIT_AREA_EX; e->d.shape_id = shape_id; e->d.kappa = kappa;
        for (int k = 0; k < 3; ++k) e->d.I[k] = I[k];
        int nd = nderNow();
        fillSpecD(e->d.I_d, dI, nd);
        fillScalarD(e->d.kappa_d, dk, nd);
        return handle(e.release());
    });
}
dtrt_object *dtrt_homogeneous_create(float sigma_t, const float *alb, int phase_id, const float *dsig, const float *dalb) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!alb) { g_err = "Homogeneous: null albedo"; return nullptr; }
        auto m = std::make_unique<Medium>();
        m->d.type = MED_HOMOGENEOUS; m->d.phase_id = phase_id; m->d.sigma_t = sigma_t;
        for (int k = 0; k < 3; ++k) m->d.albedo[k] = alb[k];
        m->d.sampling_weight = std::max(alb[0], std::max(alb[1], alb[2]));
        int nd = nderNow();
        fillScalarD(m->d.sigma_t_d, dsig, nd);
        fillSpecD(m->d.albedo_d, dalb, nd);
        return handle(m.release());
    });
}
dtrt_object *dtrt_heterogeneous_create(const char *file, const float *to_world, float scale, const float *albedo3,
                                       const char *albedo_file, int phase_id, const float *trans, const float *rot,
                                       const float *dscale, const float *dalb) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!file) { g_err = "Heterogeneous: null density file"; return nullptr; }
        auto m = std::make_unique<Medium>();
        int ch = 0;
        float mn[3], mx[3];
        if (!readVol(file, m->density, m->d.res, ch, mn, mx)) return nullptr;
        if (ch != 1) { g_err = "Heterogeneous: density grid must have one channel"; return nullptr; }
        if (albedo_file) {
            int ach = 0;
            float amn[3], amx[3];
            if (!readVol(albedo_file, m->albedo, m->d.ares, ach, amn, amx)) return nullptr;
            if (ach != 3) { g_err = "Heterogeneous: albedo grid must have three channels"; return nullptr; }
            for (int a = 0; a < 3; ++a) { m->d.amin[a] = amn[a]; m->d.agscale[a] = (float)((double)(m->d.ares[a] - 1) / ((double)amx[a] - amn[a])); }
            albedo3 = nullptr;   // the reference leaves the constant albedo zero in this variant
        }
        if (!finishHetero(*m, to_world, scale, albedo3, phase_id, trans, rot, dscale, dalb, mn, mx)) return nullptr;
        return handle(m.release());
    });
}
dtrt_object *dtrt_heterogeneous_create_grid(const float *density, int rx, int ry, int rz, const float *bmin, const float *bmax,
                                            const float *to_world, float scale, const float *albedo3, int phase_id,
                                            const float *trans, const float *rot, const float *dscale, const float *dalb) {
    return guarded<dtrt_object *>(nullptr, [&]() -> dtrt_object * {
        if (!density || rx < 2 || ry < 2 || rz < 2 || !bmin || !bmax) { g_err = "grid: bad arguments"; return nullptr; }
        auto m = std::make_unique<Medium>();
        m->d.res[0] = rx; m->d.res[1] = ry; m->d.res[2] = rz;
        m->density.assign(density, density + (size_t)rx * ry * rz);
        for (float &x : m->density) x = std::max(x, 0.0f);
        if (!finishHetero(*m, to_world, scale, albedo3, phase_id, trans, rot, dscale, dalb, bmin, bmax)) return nullptr;
        return handle(m.release());
    });
}

dtrt_object *dtrt_scene_create(dtrt_object *camera, dtrt_object *const *shapes, int ns, dtrt_object *const *bsdfs, int nb,
                               dtrt_object *const *lights, int nl, dtrt_object *const *phases, int np,
                               dtrt_object *const *media, int nm) {
    try {
        auto sc = std::make_unique<Scene>();
        Camera *c = as<Camera>(camera, T_CAMERA);
        if (!c) return nullptr;
        sc->nder = nderNow();
        if (c->nder != sc->nder) { g_err = "scene: camera was created with a different nder"; return nullptr; }
        sc->cam = *c;
        for (int i = 0; i < ns; ++i) { Shape *s = as<Shape>(shapes[i], T_SHAPE); if (!s) return nullptr; sc->shapes.push_back(s); }
        for (int i = 0; i < nb; ++i) { Bsdf *b = as<Bsdf>(bsdfs[i], T_BSDF); if (!b) return nullptr; sc->bsdfs.push_back(b); }
        for (int i = 0; i < nl; ++i) { Emitter *e = as<Emitter>(lights[i], T_EMITTER); if (!e) return nullptr; sc->emitters.push_back(e); }
        for (int i = 0; i < np; ++i) { Phase *p = as<Phase>(phases[i], T_PHASE); if (!p) return nullptr; sc->phases.push_back(p); }
        for (int i = 0; i < nm; ++i) { Medium *m = as<Medium>(media[i], T_MEDIUM); if (!m) return nullptr; sc->media.push_back(m); }
        for (Shape *s : sc->shapes)
            if (s->nder != sc->nder || s->bsdf_id < 0 || s->bsdf_id >= nb || s->light_id >= nl || s->med_int >= nm ||
                s->med_ext >= nm) { g_err = "scene: a shape references an object that does not exist"; return nullptr; }
        for (Emitter *e : sc->emitters)
            if (e->d.shape_id < 0 || e->d.shape_id >= ns) { g_err = "scene: light shape id out of range"; return nullptr; }
        for (Medium *m : sc->media)
            if (m->d.phase_id < 0 || m->d.phase_id >= np) { g_err = "scene: medium phase id out of range"; return nullptr; }
        if (c->med >= nm) { g_err = "scene: camera medium out of range"; return nullptr; }
        if (!sc->build()) return nullptr;
        return handle(sc.release());
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}
int dtrt_scene_init_edges(dtrt_object *scene, const float *w) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || !w) { if (s) g_err = "initEdges: null weights"; return -1; }
        return s->initEdges(w) ? 0 : -1;
    });
}

int dtrt_render_views(int integrator, dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                      const dtrt_render_options *options, float *const *images, int images_on_device) {
    if (integrator != 0) { g_err = "render_views: only VolPathTracerAD (integrator 0) is provided"; return -2; }
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || nviews <= 0 || !options || !images) { if (s) g_err = "render_views: bad arguments"; return -1; }
    std::vector<const Camera *> cams(nviews);
    for (int v = 0; v < nviews; ++v) {
        if (!images[v]) { g_err = "render_views: null image"; return -1; }
        if (cameras && cameras[v]) {
            cams[v] = as<Camera>(cameras[v], T_CAMERA);
            if (!cams[v]) return -1;
        } else {
            cams[v] = &s->cam;
        }
    }
    try {
        return s->renderViews(nviews, cams.data(), options, images, images_on_device != 0) ? 0 : -1;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

int dtrt_render_scenes(int integrator, int nscenes, dtrt_object *const *scenes, const int *nviews,
                       dtrt_object *const *cameras, const dtrt_render_options *options, float *const *images,
                       int images_on_device) {
    if (integrator != 0) { g_err = "render_scenes: only VolPathTracerAD (integrator 0) is provided"; return -2; }
    if (nscenes <= 0 || !scenes || !nviews || !options || !images) { g_err = "render_scenes: bad arguments"; return -1; }
    std::vector<size_t> first((size_t)nscenes + 1, 0);
    for (int i = 0; i < nscenes; ++i) {
        if (!as<Scene>(scenes[i], T_SCENE)) return -1;
        if (nviews[i] <= 0) { g_err = "render_scenes: every scene needs at least one view"; return -1; }
        for (int j = 0; j < i; ++j)
            if (scenes[j] == scenes[i]) { g_err = "render_scenes: a scene is listed twice"; return -1; }
        first[i + 1] = first[i] + (size_t)nviews[i];
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { g_err = "render_scenes: no HIP device"; return -1; }
    // Each scene renders on its own stream from its own host thread (a render waits on
    // its child counts between launches), so the scenes' kernels run side by side.
    std::vector<int> rc((size_t)nscenes, 0);
    std::vector<std::string> err((size_t)nscenes);
    auto one = [&](int i) {
        if (hipSetDevice(dev) != hipSuccess) { rc[i] = -1; err[i] = "hipSetDevice failed"; return; }
        rc[i] = dtrt_render_views(integrator, scenes[i], nviews[i], cameras ? cameras + first[i] : nullptr,
                                  options + first[i], images + first[i], images_on_device);
        if (rc[i] != 0) err[i] = g_err;
    };
    std::vector<std::thread> pool;
    try {
        for (int i = 1; i < nscenes; ++i) pool.emplace_back(one, i);
    } catch (const std::exception &e) {   // could not start a thread: render the rest here
        for (int i = (int)pool.size() + 1; i < nscenes; ++i) one(i);
    }
    one(0);
    for (std::thread &t : pool) t.join();
    for (int i = 0; i < nscenes; ++i)
        if (rc[i] != 0) {
            g_err = "render_scenes: scene " + std::to_string(i) + ": " + err[i];
            return rc[i];
        }
    return 0;
}

int dtrt_scene_capture_views(dtrt_object *scene, int nviews, dtrt_object *const *cameras, const dtrt_render_options *options,
                             float *const *images_device) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || nviews <= 0 || !options || !images_device) { if (s) g_err = "capture_views: bad arguments"; return -1; }
    std::vector<const Camera *> cams(nviews);
    for (int v = 0; v < nviews; ++v) {
        if (!images_device[v]) { g_err = "capture_views: null image"; return -1; }
        cams[v] = (cameras && cameras[v]) ? as<Camera>(cameras[v], T_CAMERA) : &s->cam;
        if (!cams[v]) return -1;
    }
    try {
        return s->captureViews(nviews, cams.data(), options, images_device) ? 0 : -1;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

static bool stepArgs(Scene *s, int nviews, dtrt_object *const *cameras, const dtrt_render_options *options,
                     float *const *images, const float *const *targets, int medium, const dtrt_render_options *adj_options,
                     float *d_density, double *param_grad, double *loss, Scene::StepArgs &a) {
    if (nviews <= 0 || !options || !adj_options || !images || !targets || !d_density)
        return fail("step: bad arguments");
    for (int v = 0; v < nviews; ++v) {
        const Camera *c = (cameras && cameras[v]) ? as<Camera>(cameras[v], T_CAMERA) : &s->cam;
        if (!c) return false;
        a.cams.push_back(*c);
    }
    a.fopts.assign(options, options + nviews);
    a.aopts.assign(adj_options, adj_options + nviews);
    a.images.assign(images, images + nviews);
    a.targets.assign(targets, targets + nviews);
    a.med = medium;
    a.dgrid = d_density;
    a.pgrad = param_grad;
    a.loss = loss;
    return true;
}

int dtrt_scene_step(dtrt_object *scene, int nviews, dtrt_object *const *cameras, const dtrt_render_options *options,
                    float *const *images, const float *const *targets, int medium,
                    const dtrt_render_options *adj_options, float *d_density, double *param_grad, double *loss) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s) return -1;
    return guarded<int>(-1, [&]() -> int {
        Scene::StepArgs a;
        if (!stepArgs(s, nviews, cameras, options, images, targets, medium, adj_options, d_density, param_grad, loss, a))
            return -1;
        return s->runStep(a) ? 0 : -1;
    });
}

int dtrt_scene_capture_step(dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                            const dtrt_render_options *options, float *const *images, const float *const *targets,
                            int medium, const dtrt_render_options *adj_options, float *d_density, double *param_grad,
                            double *loss) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s) return -1;
    return guarded<int>(-1, [&]() -> int {
        Scene::StepArgs a;
        if (!stepArgs(s, nviews, cameras, options, images, targets, medium, adj_options, d_density, param_grad, loss, a))
            return -1;
        return s->captureStep(a) ? 0 : -1;
    });
}

int dtrt_scene_replay_step(dtrt_object *scene, const uint64_t *seeds) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !seeds) { if (s) g_err = "replay_step: bad arguments"; return -1; }
    return guarded<int>(-1, [&]() -> int { return s->replayStep(seeds); });
}

int dtrt_scene_update_shape(dtrt_object *scene, int shape, const float *vertices, const float *normals,
                            const float *velocities) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s) return -1;
    return guarded<int>(-1, [&]() -> int { return s->updateShape(shape, vertices, normals, velocities) ? 0 : -1; });
}

int dtrt_scene_replay_views(dtrt_object *scene, const uint64_t *seeds) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || !seeds) { if (s) g_err = "replay_views: bad arguments"; return -1; }
        return s->replayViews(seeds);
    });
}

static int renderCommon(int integrator, dtrt_object *scene, const dtrt_render_options *o, float *img, bool dev) {
    if (integrator < 0 || integrator > 7) {
        g_err = "render: integrators are VolPathTracerAD (0), VolPathTracer (1), VolPathTracerSimple (2), "
                "PathTracer (3), DirectIntegrator (4), ParticleTracer (5), PathTracerAD (6) and DirectAD (7)";
        return -2;
    }
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !o || !img) { if (s) g_err = "render: null argument"; return -1; }
    try {
        if (integrator >= 1 && integrator <= 5) return s->renderPrimal(*o, img, dev, integrator) ? 0 : -1;
        const int ad = integrator == 6 ? AD_PATH : integrator == 7 ? AD_DIRECT : AD_VOLPATH;
        return s->render(*o, img, dev, ad) ? 0 : -1;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}
int dtrt_render(int integrator, dtrt_object *scene, const dtrt_render_options *o, float *image) {
    return renderCommon(integrator, scene, o, image, false);
}
int dtrt_render_device(int integrator, dtrt_object *scene, const dtrt_render_options *o, float *image) {
    return renderCommon(integrator, scene, o, image, true);
}
int dtrt_render_device_noalloc(int integrator, dtrt_object *scene, const dtrt_render_options *o, float *image) {
    if (integrator != 0 && integrator != 6 && integrator != 7) {
        g_err = "render_device_noalloc: forward-mode integrators only: VolPathTracerAD (0), PathTracerAD (6), DirectAD (7)";
        return -2;
    }
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !o || !image) { if (s) g_err = "render_device_noalloc: null argument"; return -1; }
    return guarded<int>(-1, [&]() -> int {
        const int ad = integrator == 6 ? AD_PATH : integrator == 7 ? AD_DIRECT : AD_VOLPATH;
        return s->renderNoAlloc(*o, image, ad) ? 0 : -1;
    });
}
int dtrt_render_stats(dtrt_object *scene, int64_t *out) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || !out) return -1;
        for (int i = 0; i < 10; ++i) out[i] = s->last_stats[i];
        return 0;
    });
}

int dtrt_scene_precision(dtrt_object *scene) {
    Scene *s = as<Scene>(scene, T_SCENE);
    return s ? (s->S.f64 ? 64 : 32) : -1;
}
int dtrt_scene_launch_settings(dtrt_object *scene, int64_t *out6) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !out6) return -1;
    out6[0] = s->tune_child_blocks;
    out6[1] = (int64_t)std::lround(s->tune_fill * 100.0);
    out6[2] = (int64_t)s->tune_max_chunk;
    out6[3] = (int64_t)s->queue_budget;
    out6[4] = s->tune_ad_block;
    out6[5] = s->tune_child_block;
    return 0;
}
int dtrt_scene_set_launch_settings(dtrt_object *scene, const int64_t *in6) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !in6) return -1;
    auto group = [](int64_t g) { return g == 64 || g == 128 || g == 256; };
    if (in6[0] < 1 || in6[0] > 1024 || in6[1] < 1 || in6[1] > 100 || in6[2] < 64 || in6[3] < (int64_t(1) << 20) ||
        !group(in6[4]) || !group(in6[5])) {
        g_err = "set_launch_settings: out of range (work-group sizes must be 64, 128 or 256)";
        return -1;
    }
    s->tune_child_blocks = (int)in6[0];
    s->tune_fill = in6[1] / 100.0;
    s->tune_max_chunk = (uint64_t)in6[2];
    s->queue_budget = (size_t)in6[3];
    s->tune_ad_block = (int)in6[4];
    s->tune_child_block = (int)in6[5];
    return 0;
}
int dtrt_scene_set_stream(dtrt_object *scene, void *stream) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s) return -1;
    s->caller = static_cast<hipStream_t>(stream);
    return 0;
}

int dtrt_scene_autotune(int integrator, dtrt_object *scene, const dtrt_render_options *options, double *best_ms,
                        int64_t *settings3) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !options) { if (s) g_err = "autotune: bad arguments"; return -1; }
    if (integrator != 0 && integrator != 6 && integrator != 7) { g_err = "autotune: forward-mode integrators only"; return -2; }
    try {
        return s->autotune(*options, integrator == 6 ? AD_PATH : integrator == 7 ? AD_DIRECT : AD_VOLPATH, best_ms,
                           settings3) ? 0 : -1;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

int dtrt_scene_set_density(dtrt_object *scene, int medium, const float *density, int on_device) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s) return -1;
        return s->setDensity(medium, density, on_device != 0) ? 0 : -1;
    });
}

int dtrt_render_views_adjoint(int integrator, dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                              const dtrt_render_options *options, int medium, const float *d_images_device,
                              float *d_density_device) {
    Scene *s = as<Scene>(scene, T_SCENE);
    if (!s || !options || nviews <= 0) { if (s) g_err = "render_views_adjoint: bad arguments"; return -1; }
    if (integrator != 0) { g_err = "render_views_adjoint: only VolPathTracerAD (0) is implemented"; return -1; }
    std::vector<const Camera *> cams(nviews);
    for (int v = 0; v < nviews; ++v) {
        cams[v] = (cameras && cameras[v]) ? as<Camera>(cameras[v], T_CAMERA) : &s->cam;
        if (!cams[v]) return -1;
    }
    try {
        return s->renderAdjointViews(nviews, cams.data(), options, medium, d_images_device, d_density_device) ? 0 : -1;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

int dtrt_render_adjoint(int integrator, dtrt_object *scene, const dtrt_render_options *options, int medium,
                        const float *d_image_device, float *d_density_device) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || !options) { g_err = "render_adjoint: bad arguments"; return -1; }
        if (integrator != 0) { g_err = "render_adjoint: only VolPathTracerAD (0) is implemented"; return -1; }
        return s->renderAdjoint(*options, medium, d_image_device, d_density_device) ? 0 : -1;
    });
}

// ------------------------------------------------------------- test / bench hooks
int dtrt_test_set_voxel_lanes(dtrt_object *scene, int medium, const int *voxels, int n) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s) return -1;
        return s->setVoxelLanes(medium, voxels, n) ? 0 : -1;
    });
}

int dtrt_test_set_queue_budget(dtrt_object *scene, uint64_t bytes) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s) return -1;
        s->queue_budget = bytes ? (size_t)bytes : (size_t(6) << 30);
        return 0;
    });
}

int dtrt_test_set_tree_in_hbm(dtrt_object *scene, int on) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s) return -1;
        s->tree_in_hbm = on != 0;
        return 0;
    });
}

int dtrt_test_kernel_times(dtrt_object *scene, double *ms4, int64_t *launches4, uint64_t *work12) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s) return -1;
        for (int i = 0; i < 4; ++i) { ms4[i] = s->kernel_ms[i]; launches4[i] = s->kernel_launches[i]; }
        for (int i = 0; i < 12; ++i) work12[i] = s->work_counts[i];
        return 0;
    });
}
int dtrt_test_build_bvh(const float *tri9, const int *shape_of, int ntri, int64_t *out8) {
    return guarded<int>(-1, [&]() -> int {
        if (!tri9 || !out8 || ntri <= 0) {
            fail("build_bvh: empty input");
            return -1;
        }
        std::vector<std::array<float, 9>> tri((size_t)ntri);
        for (int i = 0; i < ntri; ++i)
            for (int k = 0; k < 9; ++k) tri[i][k] = tri9[(size_t)i * 9 + k];
        std::vector<int> sh;
        if (shape_of) sh.assign(shape_of, shape_of + ntri);
        BvhOut b;
        if (!buildBvh(tri, sh, b)) return -1;
        // walk the GPU nodes: depth, pushes, leaf sizes, root child sizes
        const int nn = (int)(b.nodes.size() / 4);
        int64_t maxDepth = 0, leaves = 0, maxLeaf = 0, rootCount[2] = {0, 0};
        struct W { int link, depth, root; };
        std::vector<W> stack{{0, 1, -1}};
        while (!stack.empty()) {
            const W w = stack.back();
            stack.pop_back();
            if (w.link < 0) {
                const int enc = ~w.link, cnt = enc & 15;
                if (cnt == 0) continue;   // the empty slot of a single-leaf root
                ++leaves;
                maxLeaf = std::max<int64_t>(maxLeaf, cnt);
                maxDepth = std::max<int64_t>(maxDepth, w.depth);
                if (w.root >= 0) rootCount[w.root] += cnt;
                continue;
            }
            const float4 *n = b.nodes.data() + 4 * (size_t)w.link;
            int c[2];
            std::memcpy(&c[0], &n[0].w, 4);
            std::memcpy(&c[1], &n[1].w, 4);
            for (int k = 0; k < 2; ++k) stack.push_back({c[k], w.depth + 1, w.link == 0 ? k : w.root});
        }
        out8[0] = nn; out8[1] = maxDepth; out8[2] = rootCount[0]; out8[3] = rootCount[1];
        out8[4] = maxDepth - 1; out8[5] = leaves; out8[6] = maxLeaf; out8[7] = 0;
        return 0;
    });
}
int dtrt_test_scene_info(dtrt_object *scene, int64_t *out8) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s) return -1;
        int64_t ne = 0;
        for (Shape *sh : s->shapes) ne += (int64_t)sh->edges.size();
        Scene::PECache pe;
        if (!s->primaryEdges(s->cam, pe)) return -1;
        out8[0] = s->S.nnodes; out8[1] = s->S.ntris; out8[2] = ne; out8[3] = pe.npe;
        std::memcpy(&out8[4], &pe.pe_sum, 8);
        out8[5] = s->S.edges_valid; out8[6] = s->S.NV; out8[7] = s->S.NF;
        return 0;
    });
}
int dtrt_test_medium_data(dtrt_object *medium, int which, float *out, int64_t cap, double *meta8) {
    Medium *m = as<Medium>(medium, T_MEDIUM);
    if (!m || !meta8) return -1;
    const std::vector<float> &v = which ? m->albedo : m->density;
    const int *res = which ? m->d.ares : m->d.res;
    const double meta[8] = {(double)res[0], (double)res[1], (double)res[2], which ? 3.0 : 1.0, (double)v.size(),
                            (double)m->d.maxDensity, (double)m->d.scale, 0.0};
    std::memcpy(meta8, meta, sizeof(meta));
    if (out) std::memcpy(out, v.data(), (size_t)std::min<int64_t>(cap, (int64_t)v.size()) * sizeof(float));
    return 0;
}
int dtrt_test_kernel_resources(int nder, int bits, char *names, int name_len, uint64_t *vals, int cap) {
    return guarded<int>(-1, [&]() -> int {
        if (!(nder == 1 || nder == 2 || nder == 3 || nder == 4 || nder == 8 || nder == 16) || (bits != 32 && bits != 64) ||
            cap < 0 || (cap > 0 && (!vals || !names || name_len < 1))) { g_err = "kernel_resources: bad arguments"; return -1; }
        DScene S;
        std::memset(&S, 0, sizeof(S));
        S.nder = nder;
        S.f64 = bits == 64;
        int dev = 0, cus = 0, n = 0;
        KernelInfo info[32];
        auto query = [&]() -> bool {
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            HIPCHK(kernel_info(S, info, 32, &n));
            return true;
        };
        if (!query()) return -1;
        for (int i = 0; i < std::min({n, cap, 32}); ++i) {
            std::snprintf(names + (size_t)i * name_len, (size_t)name_len, "%s", info[i].name);
            const uint64_t v[5] = {info[i].private_bytes, info[i].vgprs, info[i].block, info[i].waves_per_cu,
                                   deviceScratch(info[i], cus)};
            std::memcpy(vals + 5 * (size_t)i, v, sizeof(v));
        }
        return n;
    });
}
int dtrt_test_scratch_limits(uint64_t *out6) {
    return guarded<int>(-1, [&]() -> int {
        if (!out6) return -1;
        {
            ScratchLimits &L = scratchLimits();
            std::lock_guard<std::mutex> lk(L.mu);
            out6[0] = L.lane;
            out6[1] = L.device;
        }
        const hipLimit_t lim[4] = {hipLimitStackSize, hipExtLimitScratchMin, hipExtLimitScratchMax,
                                   hipExtLimitScratchCurrent};
        for (int i = 0; i < 4; ++i) {
            size_t v = 0;
            out6[2 + i] = hipDeviceGetLimit(&v, lim[i]) == hipSuccess ? v : 0;
        }
        (void)hipGetLastError();   // an unsupported limit leaves no error behind
        return 0;
    });
}
int dtrt_test_set_scratch_limits(uint64_t lane_bytes, uint64_t device_bytes) {
    ScratchLimits &L = scratchLimits();
    std::lock_guard<std::mutex> lk(L.mu);
    L.lane = lane_bytes ? std::min<uint64_t>(lane_bytes, kMaxPrivateBytesPerLane) : kMaxPrivateBytesPerLane;
    L.device = device_bytes ? std::min<uint64_t>(device_bytes, kMaxScratchBytesPerDevice) : kMaxScratchBytesPerDevice;
    L.done.clear();
    return 0;
}
int dtrt_test_shape_edges(dtrt_object *shape, int *out_v0v1f0f1, double *out_len, int cap) {
    return guarded<int>(-1, [&]() -> int {
        Shape *s = as<Shape>(shape, T_SHAPE);
        if (!s) return -1;
        int n = (int)s->edges.size();
        for (int i = 0; i < std::min(n, cap); ++i) {
            out_v0v1f0f1[4 * i] = s->edges[i].v0; out_v0v1f0f1[4 * i + 1] = s->edges[i].v1;
            out_v0v1f0f1[4 * i + 2] = s->edges[i].f0; out_v0v1f0f1[4 * i + 3] = s->edges[i].f1;
            out_len[i] = s->edges[i].len;
        }
        return n;
    });
}

int dtrt_test_trace(dtrt_object *scene, const float *rays8, int n, int *out_face, float *out_t) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || n < 0) return -1;
        TmpFree tf;
        float *dr, *dt_;
        int *df;
        if (!devTmp(tf.v, dr, (size_t)n * 8) || !devTmp(tf.v, df, n) || !devTmp(tf.v, dt_, n)) return -1;
        if (hipMemcpy(dr, rays8, (size_t)n * 8 * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
        if (launch_trace(s->S, dr, n, df, dt_, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { g_err = "trace failed"; return -1; }
        (void)hipMemcpy(out_face, df, (size_t)n * 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(out_t, dt_, (size_t)n * 4, hipMemcpyDeviceToHost);
        return 0;
    });
}
int dtrt_test_li(dtrt_object *scene, const float *rays6, int n, int depth, int med, uint64_t seed, float *out3) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || n < 0 || med >= (int)s->media.size()) return -1;
        TmpFree tf;
        float *dr, *dout;
        if (!devTmp(tf.v, dr, (size_t)n * 6) || !devTmp(tf.v, dout, (size_t)n * 3)) return -1;
        if (hipMemcpy(dr, rays6, (size_t)n * 6 * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
        if (launch_li(s->S, dr, n, depth, med, seed, dout, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { g_err = "li failed"; return -1; }
        (void)hipMemcpy(out3, dout, (size_t)n * 3 * 4, hipMemcpyDeviceToHost);
        return 0;
    });
}
int dtrt_test_medium(dtrt_object *scene, int med, int kind, const float *seg8, int n, uint64_t seed, float *out) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || med < 0 || med >= (int)s->media.size() || n < 0) { g_err = "medium test: bad arguments"; return -1; }
        TmpFree tf;
        float *dr, *dout;
        if (!devTmp(tf.v, dr, (size_t)n * 8) || !devTmp(tf.v, dout, n)) return -1;
        if (hipMemcpy(dr, seg8, (size_t)n * 8 * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
        if (launch_medium(s->S, med, kind, dr, n, seed, dout, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { g_err = "medium failed"; return -1; }
        (void)hipMemcpy(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost);
        return 0;
    });
}
int dtrt_test_grid(dtrt_object *scene, int med, const float *pts, int n, float *out) {
    return guarded<int>(-1, [&]() -> int {
        Scene *s = as<Scene>(scene, T_SCENE);
        if (!s || med < 0 || med >= (int)s->media.size() || s->media[med]->d.type != MED_HETEROGENEOUS || n < 0) { g_err = "grid test: bad arguments"; return -1; }
        const int nd = s->nder;
        TmpFree tf;
        float *dp, *dout;
        if (!devTmp(tf.v, dp, (size_t)n * (3 + 3 * nd)) || !devTmp(tf.v, dout, (size_t)n * (1 + nd))) return -1;
        if (hipMemcpy(dp, pts, (size_t)n * (3 + 3 * nd) * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
        if (launch_grid(s->S, med, dp, n, dout, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) { g_err = "grid failed"; return -1; }
        (void)hipMemcpy(out, dout, (size_t)n * (1 + nd) * 4, hipMemcpyDeviceToHost);
        return 0;
    });
}
int dtrt_test_philox(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4) {
    return guarded<int>(-1, [&]() -> int {
        TmpFree tf;
        uint32_t *d;
        if (!devTmp(tf.v, d, 4)) return -1;
        if (launch_philox(ctr4, key2, d, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) { g_err = "philox failed"; return -1; }
        (void)hipMemcpy(out4, d, 16, hipMemcpyDeviceToHost);
        return 0;
    });
}

}  // extern "C"
