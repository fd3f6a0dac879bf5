// TEST INFRASTRUCTURE ONLY — CPU oracle (see oracle_math.h header).
// Scene-side restatement: shapes/edges, distributions, BVH (stand-in for Embree),
// camera, media, BSDFs, phase functions, emitters and the scene services.
#pragma once
#include "oracle_math.h"
#include <vector>
#include <array>
#include <string>
#include <memory>

namespace oref {

constexpr int NDER_MAX = 16;

// pmf.h:8-190 — discrete distribution over an unnormalised CDF.
struct Distrib {
    std::vector<double> cdf{0.0};
    double sum = 0, normalization = 0;
    bool normalized = false;
    void clear() { cdf.assign(1, 0.0); sum = normalization = 0; normalized = false; }
    void append(double p) { cdf.push_back(cdf.back() + p); }
    size_t size() const { return cdf.size() - 1; }
    double operator[](size_t e) const { return cdf[e + 1] - cdf[e]; }
    double normalize() {
        sum = cdf.back();
        if (sum > 0) {
            normalization = 1.0 / sum;
            for (size_t i = 1; i < cdf.size(); ++i) cdf[i] *= normalization;
            cdf.back() = 1.0;
            normalized = true;
        } else {
            normalization = 0.0;
        }
        return sum;
    }
    size_t sample(double u) const {
        auto it = std::lower_bound(cdf.begin(), cdf.end(), u);
        ptrdiff_t k = it - cdf.begin() - 1;
        size_t index = std::min(cdf.size() - 2, (size_t)std::max((ptrdiff_t)0, k));
        // zero-probability entries are skipped (pmf.h:111); bounded here so the
        // scan cannot read past the last entry.
        while (index + 1 < cdf.size() - 1 && (*this)[index] == 0) ++index;
        return index;
    }
    size_t sampleReuse(double &u) const {
        size_t index = sample(u);
        u = (u - cdf[index]) / (cdf[index + 1] - cdf[index]);
        return index;
    }
    size_t sampleReuse(double &u, double &pdf) const {
        size_t index = sample(u);
        pdf = (*this)[index];
        u = (u - cdf[index]) / (cdf[index + 1] - cdf[index]);
        return index;
    }
};

// edge_manager.h:14-26
struct Edge {
    int v0 = -1, v1 = -1, f0 = -1, f1 = -1;
    double length = 0;
};

// shape.h / shape.cpp:10-52 — vertices with per-lane tangents.
struct OShape {
    int nv = 0, nf = 0, nder = 0;
    std::vector<V3> v, vd;        // vd[j*nv + i]
    std::vector<V3> n, nd;        // shading normals (optional) and their tangents
    std::vector<std::array<int, 3>> f;
    std::vector<V3> fn, fnd;      // face normals and tangents fnd[j*nf + k]
    int light_id = -1, bsdf_id = -1, med_int = -1, med_ext = -1;
    std::vector<Edge> edges;
    Distrib edge_distrb;
    bool ok = true;
    std::string err;

    bool hasNormals() const { return !n.empty(); }
    bool isMediumTransition() const { return med_ext >= 0 || med_int >= 0; }
    bool isEmitter() const { return light_id >= 0; }
    template <int N> DV<N> vertexAD(int i) const {
        DV<N> r(v[i]);
        for (int j = 0; j < N; ++j) r.setGrad(j, vd[(size_t)j * nv + i]);
        return r;
    }
    template <int N> DV<N> normalAD(int i) const {
        DV<N> r(n[i]);
        for (int j = 0; j < N; ++j) r.setGrad(j, nd[(size_t)j * nv + i]);
        return r;
    }
    template <int N> DV<N> faceNormalAD(int k) const {
        DV<N> r(fn[k]);
        for (int j = 0; j < N; ++j) r.setGrad(j, fnd[(size_t)j * nf + k]);
        return r;
    }
    double area(int k) const {  // shape.cpp:135-141
        const auto &i = f[k];
        return 0.5 * norm(cross(v[i[1]] - v[i[0]], v[i[2]] - v[i[0]]));
    }
    // shape.cpp:151-164
    void samplePosition(int k, double u, double w, V3 &pos, V3 &nrm) const {
        const auto &i = f[k];
        double a = std::sqrt(u), b1 = 1.0 - a, b2 = a * w;
        V3 e1 = v[i[1]] - v[i[0]], e2 = v[i[2]] - v[i[0]];
        nrm = normalized(cross(e1, e2));
        pos = v[i[0]] + e1 * b1 + e2 * b2;
    }
    int isSilhouette(const Edge &e, const V3 &p) const;
    void computeFaceNormals();
    void constructEdges();
};

// camera.h / camera.cpp:8-83
struct OCamera {
    int width = 0, height = 0;
    int ox = 0, oy = 0, cw = -1, ch = -1;   // CropRectangle (camera.h:5-12)
    double c2w[4][4], w2c[4][4], c2ndc[3][3], ndc2c[3][3];
    float clip_near = 0;
    int med_id = -1;
    int nder = 0;
    V3 cpos, cs, ct, cn;
    V3 cpos_d[NDER_MAX], cs_d[NDER_MAX], ct_d[NDER_MAX], cn_d[NDER_MAX];
    bool cropped() const { return cw > 0 && ch > 0; }
    template <int N> DV<N> cposAD() const { DV<N> r(cpos); for (int j = 0; j < N; ++j) r.setGrad(j, cpos_d[j]); return r; }
    template <int N> void frameAD(DV<N> &s, DV<N> &t, DV<N> &nn) const {
        s = DV<N>(cs); t = DV<N>(ct); nn = DV<N>(cn);
        for (int j = 0; j < N; ++j) { s.setGrad(j, cs_d[j]); t.setGrad(j, ct_d[j]); nn.setGrad(j, cn_d[j]); }
    }
    V3 camDir(double x, double y) const {
        x /= width; y /= height;
        double aspect = double(width) / double(height);
        V3 ndc((x - 0.5) * 2.0, (y - 0.5) * (-2.0) / aspect, 1.0);
        V3 d;
        for (int i = 0; i < 3; ++i) d[i] = ndc2c[i][0] * ndc.x + ndc2c[i][1] * ndc.y + ndc2c[i][2] * ndc.z;
        return normalized(d);
    }
    // camera.cpp:129-165 — importance of world point p; pixel position within the window
    double sampleDirect(const V3 &p, double &px, double &py, V3 &dir) const {
        V3 q;
        for (int i = 0; i < 3; ++i) q[i] = w2c[i][0] * p.x + w2c[i][1] * p.y + w2c[i][2] * p.z + w2c[i][3];
        if (q.z < clip_near) return 0.0;
        double f = c2ndc[0][0], aspect = double(width) / double(height);
        double inv_area = 0.25 * f * f * aspect;
        int xmin = 0, xmax = width, ymin = 0, ymax = height;
        if (cropped()) {
            inv_area *= double(width) / cw * double(height) / ch;
            xmin = ox; xmax = ox + cw; ymin = oy; ymax = oy + ch;
        }
        V3 n;
        for (int i = 0; i < 3; ++i) n[i] = c2ndc[i][0] * q.x + c2ndc[i][1] * q.y + c2ndc[i][2] * q.z;
        double sx = (n.x / n.z * 0.5 + 0.5) * width, sy = (-(n.y / n.z) * 0.5 * aspect + 0.5) * height;
        if (!(sx >= xmin && sx <= xmax && sy >= ymin && sy <= ymax)) return 0.0;
        px = sx - xmin;
        py = sy - ymin;
        double dist = norm(q), inv = 1.0 / dist, invCos = dist / q.z;
        dir = (cpos - p) * inv;
        return inv * inv * invCos * invCos * invCos * inv_area * double(xmax - xmin) * double(ymax - ymin);
    }
    void samplePrimaryRay(double x, double y, V3 &org, V3 &dir) const {  // camera.cpp:52-60
        V3 d = camDir(x, y);
        double w = c2w[3][3];
        org = V3(c2w[0][3] / w, c2w[1][3] / w, c2w[2][3] / w);
        dir = V3(c2w[0][0] * d.x + c2w[0][1] * d.y + c2w[0][2] * d.z,
                 c2w[1][0] * d.x + c2w[1][1] * d.y + c2w[1][2] * d.z,
                 c2w[2][0] * d.x + c2w[2][1] * d.y + c2w[2][2] * d.z);
    }
    template <int N> void samplePrimaryRayAD(double x, double y, DV<N> &org, DV<N> &dir) const {  // camera.cpp:75-83
        V3 d = camDir(x, y);
        DV<N> s, t, nn;
        frameAD(s, t, nn);
        org = cposAD<N>();
        dir = s * d.x + t * d.y + nn * d.z;
    }
};

// BSDFs (bsdf/*.cpp). type: 0 diffuse, 1 null, 2 phong, 3 rough dielectric
struct OBsdf {
    int type = 0;
    Spec refl; std::vector<Spec> refl_d;                       // diffuse
    Spec kd, ks; std::vector<Spec> kd_d, ks_d; double ex = 0; std::vector<double> ex_d; double specW = 0;  // phong
    double alpha = 0; std::vector<double> alpha_d; double eta = 1; std::vector<double> eta_d;            // rough dielectric
    bool isNull() const { return type == 1; }
    bool isTransmissive() const { return type == 1 || type == 3; }
};

// phase.h:24-79. type: 0 HG, 1 isotropic
struct OPhase {
    int type = 0;
    double g = 0;
    std::vector<double> g_d;
};

// emitter/area.cpp, area2.cpp. type: 0 AreaLight, 1 AreaLightEx
struct OEmitter {
    int type = 0, shape_id = -1;
    bool two_sided = false;
    Spec I; std::vector<Spec> I_d; double kappa = 0; std::vector<double> kappa_d;
    Spec intensity() const { return I; }
};

// medium/*.h. type: 0 homogeneous, 1 heterogeneous
struct OMedium {
    int type = 0, phase_id = -1, nder = 0;
    // homogeneous (homogeneous.h:10-58)
    double sigma_t = 0; std::vector<double> sigma_t_d; Spec albedo; std::vector<Spec> albedo_d;
    double sampling_weight = 0;
    // heterogeneous (heterogeneous.h:13-106, gridvolume.h:15-351)
    int res[3] = {0, 0, 0};
    std::vector<float> density;
    std::vector<float> albedo_grid;   // 3 channels when present
    int albedo_res[3] = {0, 0, 0};
    V3 bmin, bmax, extent, center;
    V3 amin, amax, aextent;          // albedo grid data box
    double v2w[4][4], w2v[4][4];
    V3 aabb_min, aabb_max;
    double maxDensity = 0, maxInvDensity = 0;
    double scale = 1; std::vector<double> scale_d;
    V3 trans_d[NDER_MAX], rot_d[NDER_MAX];
    bool hasAlbedoGrid() const { return !albedo_grid.empty(); }
};

struct PrimaryEdge {   // edge_manager.h:28-33
    V2 v0s, v1s;
    V2 v0s_d[NDER_MAX], v1s_d[NDER_MAX];
    V3 v0p, v1p, v0c, v1c;
};

struct BvhNode {
    V3 lo, hi;
    int left = -1, right = -1, first = 0, count = 0;
};

struct OScene {
    int nder = 0;
    const OCamera *cam = nullptr;
    std::vector<const OShape *> shapes;
    std::vector<const OBsdf *> bsdfs;
    std::vector<const OEmitter *> emitters;
    std::vector<const OPhase *> phases;
    std::vector<const OMedium *> media;
    // light sampling (scene.cpp:45-98)
    std::vector<double> light_pmf, light_cdf, light_areas, area_cdf_pool;
    std::vector<int> area_cdf_begin;
    // edges (edge_manager.cpp, bruteforce.cpp)
    Distrib edges_distrb;
    std::vector<PrimaryEdge> primary_edges;
    Distrib primary_distrb;
    // triangle BVH in place of Embree
    std::vector<BvhNode> nodes;
    std::vector<std::pair<int, int>> prims;   // (shape, triangle)
    bool brute_force = false;
    bool ok = true;
    std::string err;

    void buildLights();
    void initEdges(const double *weights);
    void buildBvh();
    // closest hit with tnear <= t <= tfar; returns false on miss.
    bool intersect(const V3 &org, const V3 &dir, double tnear, double tfar, int &shape, int &tri) const;
};

}  // namespace oref
