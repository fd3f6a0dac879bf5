// TEST INFRASTRUCTURE ONLY — CPU oracle C API (see oracle_math.h header).
// Constructors mirror the reference's pybind surface (src/dtrt.cpp:35-115) with
// the same argument meaning and tensor layouts; prefix "oracle_".
#include "oracle_core.h"
#include <cstdio>
#include <cstring>
#include <string>

using namespace oref;

namespace {
thread_local std::string g_err;
int g_nder = 4;

struct Handle {
    int tag;
    explicit Handle(int t) : tag(t) {}
    virtual ~Handle() = default;
};
enum { T_CAM = 1, T_SHAPE, T_BSDF, T_PHASE, T_EMIT, T_MED, T_SCENE };
struct HCam : Handle { OCamera c; HCam() : Handle(T_CAM) {} };
struct HShape : Handle { OShape s; HShape() : Handle(T_SHAPE) {} };
struct HBsdf : Handle { OBsdf b; HBsdf() : Handle(T_BSDF) {} };
struct HPhase : Handle { OPhase p; HPhase() : Handle(T_PHASE) {} };
struct HEmit : Handle { OEmitter e; HEmit() : Handle(T_EMIT) {} };
struct HMed : Handle { OMedium m; HMed() : Handle(T_MED) {} };
struct HScene : Handle { OScene s; HScene() : Handle(T_SCENE) {} };

bool inv4(const double m[4][4], double o[4][4]) {
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r) if (std::abs(a[r][c]) > std::abs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return false;
        if (p != c) for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) o[i][j] = a[i][j + 4];
    return true;
}
bool inv3(const double m[3][3], double o[3][3]) {
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (det == 0) return false;
    o[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    return true;
}

// gridvolume.h:88-132 — VOL version 3, float32 data, 1 or 3 channels.
bool loadVol(const char *fn, std::vector<float> &data, int *res, int &ch, V3 &mn, V3 &mx) {
    FILE *f = std::fopen(fn, "rb");
    if (!f) { g_err = std::string("cannot open volume file ") + fn; return false; }
    char hdr[3];
    uint8_t ver;
    int type;
    float pmin[3], pmax[3];
    bool ok = std::fread(hdr, 1, 3, f) == 3 && std::strncmp(hdr, "VOL", 3) == 0 && std::fread(&ver, 1, 1, f) == 1 && ver == 3 &&
              std::fread(&type, 4, 1, f) == 1 && std::fread(res, 4, 3, f) == 3 && std::fread(&ch, 4, 1, f) == 1 && type == 1 &&
              (ch == 1 || ch == 3) && res[0] > 1 && res[1] > 1 && res[2] > 1 && std::fread(pmin, 4, 3, f) == 3 &&
              std::fread(pmax, 4, 3, f) == 3;
    if (ok) {
        size_t n = (size_t)res[0] * res[1] * res[2] * ch;
        data.resize(n);
        ok = std::fread(data.data(), 4, n, f) == n;
    }
    std::fclose(f);
    if (!ok) { g_err = std::string("invalid VOL v3 file ") + fn; return false; }
    mn = V3(pmin[0], pmin[1], pmin[2]);
    mx = V3(pmax[0], pmax[1], pmax[2]);
    return true;
}

void finishHetero(OMedium &m, const float *to_world, float scale, int phase_id, const float *trans, const float *rot,
                  const float *dscale, const float *dalb, const float *albedo3) {
    m.type = 1;
    m.phase_id = phase_id;
    m.nder = g_nder;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) m.v2w[i][j] = to_world ? to_world[i * 4 + j] : (i == j);
    inv4(m.v2w, m.w2v);
    m.maxDensity = 0;
    for (float &x : m.density) { x = std::max(x, 0.0f); if (x > m.maxDensity) m.maxDensity = x; }
    m.extent = m.bmax - m.bmin;
    m.scale = scale;
    m.scale_d.assign(g_nder, 0.0);
    m.albedo = albedo3 ? Spec(albedo3[0], albedo3[1], albedo3[2]) : Spec();
    m.albedo_d.assign(g_nder, Spec());
    for (int j = 0; j < g_nder; ++j) {
        m.trans_d[j] = trans ? V3(trans[3 * j], trans[3 * j + 1], trans[3 * j + 2]) : V3();
        m.rot_d[j] = rot ? V3(rot[3 * j], rot[3 * j + 1], rot[3 * j + 2]) : V3();
        if (dscale) m.scale_d[j] = dscale[j];
        if (dalb) m.albedo_d[j] = Spec(dalb[3 * j], dalb[3 * j + 1], dalb[3 * j + 2]);
    }
    m.maxInvDensity = 1.0 / (m.scale * m.maxDensity);
    auto v2w = [&](const V3 &p) {
        V3 r;
        for (int i = 0; i < 3; ++i) r[i] = m.v2w[i][0] * p.x + m.v2w[i][1] * p.y + m.v2w[i][2] * p.z + m.v2w[i][3];
        return r;
    };
    m.center = v2w((m.bmin + m.bmax) * 0.5);   // gridvolume.h:135-141
    m.aabb_min = V3(INF, INF, INF);
    m.aabb_max = V3(-INF, -INF, -INF);
    for (int c = 0; c < 8; ++c) {
        V3 p(c & 1 ? m.bmax.x : m.bmin.x, c & 2 ? m.bmax.y : m.bmin.y, c & 4 ? m.bmax.z : m.bmin.z);
        V3 w = v2w(p);
        for (int a = 0; a < 3; ++a) { m.aabb_min[a] = std::min(m.aabb_min[a], w[a]); m.aabb_max[a] = std::max(m.aabb_max[a], w[a]); }
    }
}

}  // namespace

extern "C" {

typedef struct oracle_render_options {
    uint64_t seed;
    int32_t spp, max_bounces, sppe, sppse, quiet, mode;
    float ddist_coeff;
} oracle_render_options;

int oracle_abi_version(void) { return 1; }
const char *oracle_last_error(void) { return g_err.c_str(); }
int oracle_set_nder(int n) {
    if (!(n == 1 || n == 2 || n == 3 || n == 4 || n == 8 || n == 16)) { g_err = "nder must be one of 1,2,3,4,8,16"; return -1; }
    g_nder = n;
    return 0;
}
int oracle_nder(void) { return g_nder; }
double oracle_angle_eps(void) { return ANGLE_EPS; }
void oracle_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

void oracle_destroy(void *h) { delete static_cast<Handle *>(h); }

// camera.cpp:8-38
void *oracle_camera_create(int w, int h, const float *c2w, const float *c2ndc, float clip_near, int med_id, const float *vel) {
    if (!c2w || !c2ndc || w <= 0 || h <= 0) { g_err = "camera: bad arguments"; return nullptr; }
    auto *H = new HCam();
    OCamera &c = H->c;
    c.width = w; c.height = h; c.clip_near = clip_near; c.med_id = med_id; c.nder = g_nder;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) c.c2w[i][j] = c2w[i * 4 + j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c.c2ndc[i][j] = c2ndc[i * 3 + j];
    if (c.c2ndc[0][0] != c.c2ndc[1][1]) { g_err = "camera: cam_to_ndc[0][0] must equal [1][1]"; delete H; return nullptr; }
    inv4(c.c2w, c.w2c);
    inv3(c.c2ndc, c.ndc2c);
    c.cpos = V3(c.c2w[0][3], c.c2w[1][3], c.c2w[2][3]);
    c.cs = V3(c.c2w[0][0], c.c2w[1][0], c.c2w[2][0]);
    c.ct = V3(c.c2w[0][1], c.c2w[1][1], c.c2w[2][1]);
    c.cn = V3(c.c2w[0][2], c.c2w[1][2], c.c2w[2][2]);
    for (int j = 0; j < g_nder; ++j) {
        V3 dx = vel ? V3(vel[3 * j], vel[3 * j + 1], vel[3 * j + 2]) : V3();
        V3 dw = vel ? V3(vel[3 * g_nder + 3 * j], vel[3 * g_nder + 3 * j + 1], vel[3 * g_nder + 3 * j + 2]) : V3();
        c.cpos_d[j] = dx;
        c.cs_d[j] = cross(dw, c.cs);
        c.ct_d[j] = cross(dw, c.ct);
        c.cn_d[j] = cross(dw, c.cn);
    }
    return H;
}
int oracle_camera_set_rect(void *cam, int ox, int oy, int w, int h) {
    auto *H = static_cast<HCam *>(cam);
    H->c.ox = ox; H->c.oy = oy; H->c.cw = w; H->c.ch = h;
    return 0;
}

// shape.cpp:10-52 (normal tangents read from offset nv*3*nder for every lane, as written)
void *oracle_shape_create(const float *v, const int *idx, const float *uv, const float *n, int nv, int nf, int light_id,
                          int bsdf_id, int med_int, int med_ext, const float *vel) {
    (void)uv;
    if (!v || !idx || nv <= 0 || nf <= 0) { g_err = "shape: bad arguments"; return nullptr; }
    auto *H = new HShape();
    OShape &s = H->s;
    s.nv = nv; s.nf = nf; s.nder = g_nder;
    s.light_id = light_id; s.bsdf_id = bsdf_id; s.med_int = med_int; s.med_ext = med_ext;
    s.v.resize(nv);
    for (int i = 0; i < nv; ++i) s.v[i] = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    if (n) {
        s.n.resize(nv);
        for (int i = 0; i < nv; ++i) s.n[i] = normalized(V3(n[3 * i], n[3 * i + 1], n[3 * i + 2]));
    }
    s.f.resize(nf);
    for (int k = 0; k < nf; ++k) {
        s.f[k] = {idx[3 * k], idx[3 * k + 1], idx[3 * k + 2]};
        for (int a = 0; a < 3; ++a)
            if (s.f[k][a] < 0 || s.f[k][a] >= nv) { g_err = "shape: index out of range"; delete H; return nullptr; }
    }
    s.vd.assign((size_t)g_nder * nv, V3());
    s.nd.assign(n ? (size_t)g_nder * nv : 0, V3());
    if (vel) {
        for (int j = 0; j < g_nder; ++j)
            for (int i = 0; i < nv; ++i) {
                const float *p = vel + ((size_t)j * nv + i) * 3;
                s.vd[(size_t)j * nv + i] = V3(p[0], p[1], p[2]);
                if (n) {
                    const float *q = vel + (size_t)nv * 3 * g_nder + (size_t)i * 3;
                    s.nd[(size_t)j * nv + i] = V3(q[0], q[1], q[2]);
                }
            }
    }
    s.computeFaceNormals();
    s.constructEdges();
    if (!s.ok) { g_err = "shape: " + s.err; delete H; return nullptr; }
    return H;
}

void *oracle_bsdf_diffuse_create(const float *rgb, const float *d) {
    auto *H = new HBsdf();
    H->b.type = 0;
    H->b.refl = Spec(rgb[0], rgb[1], rgb[2]);
    H->b.refl_d.assign(g_nder, Spec());
    if (d) for (int j = 0; j < g_nder; ++j) H->b.refl_d[j] = Spec(d[3 * j], d[3 * j + 1], d[3 * j + 2]);
    return H;
}
void *oracle_bsdf_null_create(void) { auto *H = new HBsdf(); H->b.type = 1; return H; }
void *oracle_bsdf_phong_create(const float *kd, const float *ks, float ex, const float *dkd, const float *dks, const float *dex) {
    auto *H = new HBsdf();
    OBsdf &b = H->b;
    b.type = 2;
    b.kd = Spec(kd[0], kd[1], kd[2]);
    b.ks = Spec(ks[0], ks[1], ks[2]);
    b.ex = ex;
    double v0 = (b.kd[0] + b.kd[1] + b.kd[2]) / 3.0, v1 = (b.ks[0] + b.ks[1] + b.ks[2]) / 3.0;
    b.specW = v1 / (v0 + v1);
    b.kd_d.assign(g_nder, Spec()); b.ks_d.assign(g_nder, Spec()); b.ex_d.assign(g_nder, 0.0);
    for (int j = 0; j < g_nder; ++j) {
        if (dkd) b.kd_d[j] = Spec(dkd[3 * j], dkd[3 * j + 1], dkd[3 * j + 2]);
        if (dks) b.ks_d[j] = Spec(dks[3 * j], dks[3 * j + 1], dks[3 * j + 2]);
        if (dex) b.ex_d[j] = dex[j];
    }
    return H;
}
// roughdielectric.h:10-38, microfacet.h:11-14 (alpha clamped to >= 1e-4f)
void *oracle_bsdf_roughdielectric_create(float alpha, float int_ior, float ext_ior, const float *dalpha, const float *deta) {
    auto *H = new HBsdf();
    OBsdf &b = H->b;
    b.type = 3;
    b.alpha = alpha < 1e-4f ? (double)1e-4f : (double)alpha;
    b.eta = (double)int_ior / (double)ext_ior;
    b.alpha_d.assign(g_nder, 0.0); b.eta_d.assign(g_nder, 0.0);
    for (int j = 0; j < g_nder; ++j) {
        if (dalpha) b.alpha_d[j] = dalpha[j];
        if (deta) b.eta_d[j] = deta[j];
    }
    return H;
}
void *oracle_phase_hg_create(float g, const float *dg) {
    auto *H = new HPhase();
    H->p.type = 0; H->p.g = g; H->p.g_d.assign(g_nder, 0.0);
    if (dg) for (int j = 0; j < g_nder; ++j) H->p.g_d[j] = dg[j];
    return H;
}
void *oracle_phase_isotropic_create(void) { auto *H = new HPhase(); H->p.type = 1; H->p.g_d.assign(g_nder, 0.0); return H; }
void *oracle_area_light_create(int shape_id, const float *I, int two_sided) {
    auto *H = new HEmit();
    H->e.type = 0; H->e.shape_id = shape_id; H->e.two_sided = two_sided != 0;
    H->e.I = Spec(I[0], I[1], I[2]);
    H->e.I_d.assign(g_nder, Spec()); H->e.kappa_d.assign(g_nder, 0.0);
    return H;
}
void *oracle_area_light_ex_create(int shape_id, const float *I, float kappa, const float *dI, const float *dk) {
    auto *H = new HEmit();
    H->e.type = 1; H->e.shape_id = shape_id; H->e.I = Spec(I[0], I[1], I[2]); H->e.kappa = kappa;
    H->e.I_d.assign(g_nder, Spec()); H->e.kappa_d.assign(g_nder, 0.0);
    for (int j = 0; j < g_nder; ++j) {
        if (dI) H->e.I_d[j] = Spec(dI[3 * j], dI[3 * j + 1], dI[3 * j + 2]);
        if (dk) H->e.kappa_d[j] = dk[j];
    }
    return H;
}
void *oracle_homogeneous_create(float sigma_t, const float *alb, int phase_id, const float *dsig, const float *dalb) {
    auto *H = new HMed();
    OMedium &m = H->m;
    m.type = 0; m.phase_id = phase_id; m.nder = g_nder;
    m.sigma_t = sigma_t;
    m.albedo = Spec(alb[0], alb[1], alb[2]);
    m.sampling_weight = std::max(alb[0], std::max(alb[1], alb[2]));
    m.sigma_t_d.assign(g_nder, 0.0); m.albedo_d.assign(g_nder, Spec());
    for (int j = 0; j < g_nder; ++j) {
        if (dsig) m.sigma_t_d[j] = dsig[j];
        if (dalb) m.albedo_d[j] = Spec(dalb[3 * j], dalb[3 * j + 1], dalb[3 * j + 2]);
    }
    return H;
}
// heterogeneous.h:13-106 — density (and optional albedo) grids from VOL files.
void *oracle_heterogeneous_create(const char *density_file, const float *to_world, float scale, const float *albedo3,
                                  const char *albedo_file, int phase_id, const float *trans, const float *rot,
                                  const float *dscale, const float *dalb) {
    auto *H = new HMed();
    OMedium &m = H->m;
    int ch;
    if (!loadVol(density_file, m.density, m.res, ch, m.bmin, m.bmax) || ch != 1) {
        if (g_err.empty()) g_err = "density grid must have one channel";
        delete H;
        return nullptr;
    }
    if (albedo_file) {
        int ach;
        V3 amn, amx;
        if (!loadVol(albedo_file, m.albedo_grid, m.albedo_res, ach, amn, amx) || ach != 3) { delete H; return nullptr; }
        for (float &x : m.albedo_grid) x = std::max(x, 0.0f);
        m.amin = amn; m.amax = amx; m.aextent = amx - amn;
        albedo3 = nullptr;   // the reference leaves m_albedo zero in this variant
    }
    finishHetero(m, to_world, scale, phase_id, trans, rot, dscale, dalb, albedo3);
    return H;
}
void *oracle_heterogeneous_create_grid(const float *density, int rx, int ry, int rz, const float *bmin, const float *bmax,
                                       const float *to_world, float scale, const float *albedo3, int phase_id,
                                       const float *trans, const float *rot, const float *dscale, const float *dalb) {
    if (!density || rx < 2 || ry < 2 || rz < 2) { g_err = "grid: bad arguments"; return nullptr; }
    auto *H = new HMed();
    OMedium &m = H->m;
    m.res[0] = rx; m.res[1] = ry; m.res[2] = rz;
    m.density.assign(density, density + (size_t)rx * ry * rz);
    m.bmin = V3(bmin[0], bmin[1], bmin[2]);
    m.bmax = V3(bmax[0], bmax[1], bmax[2]);
    finishHetero(m, to_world, scale, phase_id, trans, rot, dscale, dalb, albedo3);
    return H;
}

// scene.cpp:9-136
void *oracle_scene_create(void *cam, void *const *shapes, int ns, void *const *bsdfs, int nb, void *const *lights, int nl,
                          void *const *phases, int np, void *const *media, int nm) {
    auto *H = new HScene();
    OScene &s = H->s;
    s.nder = g_nder;
    s.cam = &static_cast<HCam *>(cam)->c;
    for (int i = 0; i < ns; ++i) s.shapes.push_back(&static_cast<HShape *>(shapes[i])->s);
    for (int i = 0; i < nb; ++i) s.bsdfs.push_back(&static_cast<HBsdf *>(bsdfs[i])->b);
    for (int i = 0; i < nl; ++i) s.emitters.push_back(&static_cast<HEmit *>(lights[i])->e);
    for (int i = 0; i < np; ++i) s.phases.push_back(&static_cast<HPhase *>(phases[i])->p);
    for (int i = 0; i < nm; ++i) s.media.push_back(&static_cast<HMed *>(media[i])->m);
    for (auto *sh : s.shapes)
        if (sh->bsdf_id < 0 || sh->bsdf_id >= nb || sh->light_id >= nl || sh->med_int >= nm || sh->med_ext >= nm) {
            g_err = "scene: shape references an object that does not exist";
            delete H;
            return nullptr;
        }
    for (auto *e : s.emitters)
        if (e->shape_id < 0 || e->shape_id >= ns) { g_err = "scene: light shape id out of range"; delete H; return nullptr; }
    s.buildBvh();
    s.buildLights();
    s.initEdges(nullptr);
    if (!s.ok) { g_err = "scene: " + s.err; delete H; return nullptr; }
    return H;
}
int oracle_scene_init_edges(void *scene, const float *w) {
    OScene &s = static_cast<HScene *>(scene)->s;
    std::vector<double> wd(s.shapes.size());
    for (size_t i = 0; i < wd.size(); ++i) wd[i] = w[i];
    s.initEdges(wd.data());
    return s.ok ? 0 : -1;
}
void oracle_scene_set_brute_force(void *scene, int on) { static_cast<HScene *>(scene)->s.brute_force = on != 0; }

// Scene facts for the host-logic tests: counts and distribution sums.
int oracle_scene_query(void *scene, const char *what, int arg, double *out, int nout) {
    OScene &s = static_cast<HScene *>(scene)->s;
    std::string w(what);
    if (w == "num_edges") { out[0] = (double)s.shapes.at(arg)->edges.size(); out[1] = s.shapes[arg]->edge_distrb.sum; return 2; }
    if (w == "primary_edges") { out[0] = (double)s.primary_edges.size(); out[1] = s.primary_distrb.sum; return 2; }
    if (w == "edge_shape_pmf") {
        int n = (int)std::min((size_t)nout, s.edges_distrb.size());
        for (int i = 0; i < n; ++i) out[i] = s.edges_distrb[i];
        return n;
    }
    if (w == "light_pmf") {
        int n = std::min(nout, (int)s.light_pmf.size());
        for (int i = 0; i < n; ++i) out[i] = s.light_pmf[i];
        return n;
    }
    if (w == "primary_edge") {   // v0s, v1s, grads of both (2 + 2 + 4*nder)
        const PrimaryEdge &pe = s.primary_edges.at(arg);
        int k = 0;
        out[k++] = pe.v0s.x; out[k++] = pe.v0s.y; out[k++] = pe.v1s.x; out[k++] = pe.v1s.y;
        for (int j = 0; j < s.nder; ++j) { out[k++] = pe.v0s_d[j].x; out[k++] = pe.v0s_d[j].y; out[k++] = pe.v1s_d[j].x; out[k++] = pe.v1s_d[j].y; }
        return k;
    }
    g_err = "unknown query";
    return -1;
}

// Closest hit for n rays; rays[i] = {ox,oy,oz,dx,dy,dz,tnear,tfar}.
int oracle_scene_intersect(void *scene, const double *rays, int n, int *out_shape, int *out_tri, double *out_t) {
    OScene &s = static_cast<HScene *>(scene)->s;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const double *r = rays + (size_t)i * 8;
        int sh, tr;
        V3 o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
        bool hit = s.intersect(o, d, r[6], r[7], sh, tr);
        out_shape[i] = hit ? sh : -1;
        out_tri[i] = hit ? tr : -1;
        if (hit) {
            Its its;
            shapeIntersect(*s.shapes[sh], tr, o, d, its);
            out_t[i] = its.t;
        } else {
            out_t[i] = INF;
        }
    }
    return 0;
}

void oracle_counters(long long *out6) {
    out6[0] = g_counters.neg_pixels; out6[1] = g_counters.nonfinite_der; out6[2] = g_counters.dt_cap;
    out6[3] = g_counters.bad_edge_weight; out6[4] = g_counters.nonortho; out6[5] = g_counters.med_inconsistent;
}

// integrator kind 0: VolPathTracerAD (integratorAD.cpp:19-166); kinds 1-4, float32[H][W][3]:
// VolPathTracer (volpath.cpp:106-141), VolPathTracerSimple (volpath_simple.cpp:89-124),
// PathTracer (path.cpp:64-96), DirectIntegrator (direct.cpp:51-85); kind 5: ParticleTracer
// (ptracer.cpp:10-55); kinds 6, 7: PathTracerAD (pathAD.cpp), DirectAD (directAD.cpp)
// through integratorAD.cpp:19-166
int oracle_render(int integrator, void *scene, const oracle_render_options *o, float *out) {
    if (integrator < 0 || integrator > 7) {
        g_err = "integrators: VolPathTracerAD (0), VolPathTracer (1), VolPathTracerSimple (2), PathTracer (3), "
                "DirectIntegrator (4), ParticleTracer (5), PathTracerAD (6), DirectAD (7)";
        return -2;
    }
    OScene &s = static_cast<HScene *>(scene)->s;
    g_counters.reset();
    if (integrator == 5) {
        Tracer<1>(s).renderParticles(o->seed, o->spp, o->max_bounces, out);
        return 0;
    }
    if (integrator >= 1 && integrator <= 4) {
        Tracer<1>(s).renderPrimal(o->seed, o->spp, o->max_bounces, out, integrator);
        return 0;
    }
#define OR_CASE(K) case K: { Tracer<K> tr(s); tr.adKind = integrator == 6 ? 1 : integrator == 7 ? 2 : 0; \
        tr.render(o->seed, o->spp, o->max_bounces, o->sppe, o->sppse, o->mode, o->ddist_coeff, out); } break;
    switch (s.nder) {
        OR_CASE(1) OR_CASE(2) OR_CASE(3) OR_CASE(4) OR_CASE(8) OR_CASE(16)
        default: g_err = "unsupported nder"; return -1;
    }
#undef OR_CASE
    return 0;
}

// Primal radiance of explicit rays (Li with incEmission, as edge sub-paths call it):
// rays[i] = {ox,oy,oz,dx,dy,dz}; one PCG stream per ray keyed (seed, i).
int oracle_trace_li(void *scene, const double *rays, int n, int max_depth, int medium, uint64_t seed, double *out3) {
    OScene &s = static_cast<HScene *>(scene)->s;
    Tracer<1> tr(s);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        const double *r = rays + (size_t)i * 6;
        Pcg p(seed, i);
        Spec L = tr.Li(p, V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), max_depth, medium, true);
        for (int c = 0; c < 3; ++c) out3[(size_t)i * 3 + c] = L[c];
    }
    return 0;
}

// Transmittance estimators of one medium along explicit segments:
// seg[i] = {ox,oy,oz,dx,dy,dz,tmin,tmax}; kind 0 = ratio tracking value,
// kind 1 = delta-tracking scatter indicator (1 if a scatter event is sampled).
int oracle_medium_estimate(void *medium, int kind, const double *seg, int n, uint64_t seed, double *out) {
    const OMedium &m = static_cast<HMed *>(medium)->m;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const double *r = seg + (size_t)i * 8;
        Pcg p(seed, i);
        V3 o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
        if (kind == 0) {
            out[i] = mediumTransmittance(m, o, d, r[6], r[7], p);
        } else {
            double rnd2[2] = {p.next1D(), p.next1D()};
            V3 ps;
            Spec thr = Spec::ones();
            out[i] = sampleDistance(m, o, d, r[7], rnd2, p, ps, thr) ? 1.0 : 0.0;
        }
    }
    return 0;
}

// Density lookups with tangents: pts[i] = {x,y,z, then nder tangent vectors}.
// out[i] = {value, d/dtheta_j for j < nder}.
int oracle_grid_lookup(void *medium, const double *pts, int n, double *out) {
    const OMedium &m = static_cast<HMed *>(medium)->m;
    const int nd = m.nder, stride = 3 + 3 * nd;
    for (int i = 0; i < n; ++i) {
        const double *r = pts + (size_t)i * stride;
        DV<16> p(V3(r[0], r[1], r[2]));
        for (int j = 0; j < nd; ++j) p.setGrad(j, V3(r[3 + 3 * j], r[4 + 3 * j], r[5 + 3 * j]));
        // lanes beyond nder carry zero motion
        OMedium mm = m;
        for (int j = nd; j < 16; ++j) { mm.trans_d[j] = V3(); mm.rot_d[j] = V3(); }
        D<16> v = lookupFloatAD<16>(mm, p);
        out[(size_t)i * (1 + nd)] = v.v;
        for (int j = 0; j < nd; ++j) out[(size_t)i * (1 + nd) + 1 + j] = v.g[j];
    }
    return 0;
}

// Unit functions of the math layer for closed-form pinning tests.
int oracle_unit(const char *name, const double *in, int nin, double *out, int nout) {
    std::string w(name);
    (void)nin;
    (void)nout;
    if (w == "pcg32") {   // in: seed_hi, seed_lo, idx, count -> out: u32 values
        uint64_t seed = ((uint64_t)in[0] << 32) | (uint64_t)in[1];
        Pcg p(seed, (int)in[2]);
        for (int i = 0; i < (int)in[3]; ++i) out[i] = p.next_u32();
        return (int)in[3];
    }
    if (w == "pcg_next1d") {
        uint64_t seed = ((uint64_t)in[0] << 32) | (uint64_t)in[1];
        Pcg p(seed, (int)in[2]);
        for (int i = 0; i < (int)in[3]; ++i) out[i] = p.next1D();
        return (int)in[3];
    }
    if (w == "erf") { out[0] = erf_as(in[0]); return 1; }
    if (w == "erfinv") { out[0] = erfinv_giles(in[0]); return 1; }
    if (w == "fresnel") { double ct; out[0] = fresnelDielectricExt(in[0], ct, in[1]); out[1] = ct; return 2; }
    if (w == "frame") {
        Frame f(normalized(V3(in[0], in[1], in[2])));
        double v[9] = {f.s.x, f.s.y, f.s.z, f.t.x, f.t.y, f.t.z, f.n.x, f.n.y, f.n.z};
        std::memcpy(out, v, sizeof(v));
        return 9;
    }
    if (w == "hg_eval") { OPhase ph; ph.type = 0; ph.g = in[0]; out[0] = phaseEval(ph, V3(in[1], in[2], in[3]), V3(in[4], in[5], in[6])); return 1; }
    if (w == "hg_sample") {
        OPhase ph; ph.type = 0; ph.g = in[0];
        V3 wo;
        phaseSample(ph, V3(in[1], in[2], in[3]), in[4], in[5], wo);
        out[0] = wo.x; out[1] = wo.y; out[2] = wo.z;
        return 3;
    }
    if (w == "cos_hemisphere") { V3 d = squareToCosineHemisphere(in[0], in[1]); out[0] = d.x; out[1] = d.y; out[2] = d.z; return 3; }
    if (w == "uniform_sphere") { V3 d = squareToUniformSphere(in[0], in[1]); out[0] = d.x; out[1] = d.y; out[2] = d.z; return 3; }
    if (w == "beckmann_eval") { Microfacet mf{in[0]}; out[0] = mf.eval(V3(in[1], in[2], in[3])); return 1; }
    if (w == "beckmann_pdf") { Microfacet mf{in[0]}; out[0] = mf.pdf(V3(in[1], in[2], in[3]), V3(in[4], in[5], in[6])); return 1; }
    if (w == "beckmann_sample") {
        Microfacet mf{in[0]};
        V3 m = mf.sample(V3(in[1], in[2], in[3]), in[4], in[5]);
        out[0] = m.x; out[1] = m.y; out[2] = m.z;
        return 3;
    }
    if (w == "smith_g1") { Microfacet mf{in[0]}; out[0] = mf.smithG1(V3(in[1], in[2], in[3]), V3(in[4], in[5], in[6])); return 1; }
    if (w == "bsdf_eval" || w == "bsdf_pdf" || w == "bsdf_sample") {
        // in: type, p0..p5 (diffuse: r,g,b | phong: kd(3), ks(3) + ex in in[7] | dielectric: alpha, eta),
        //     then wi(3), wo(3) or rnd(3) at in[8..13]
        OBsdf b;
        b.type = (int)in[0];
        b.refl = Spec(in[1], in[2], in[3]);
        b.kd = Spec(in[1], in[2], in[3]); b.ks = Spec(in[4], in[5], in[6]); b.ex = in[7];
        double v0 = (b.kd[0] + b.kd[1] + b.kd[2]) / 3, v1 = (b.ks[0] + b.ks[1] + b.ks[2]) / 3;
        b.specW = v1 / (v0 + v1);
        b.alpha = in[1] < 1e-4f ? (double)1e-4f : in[1];
        b.eta = in[2];
        Its its;
        its.wi = V3(in[8], in[9], in[10]);
        V3 wo(in[11], in[12], in[13]);
        if (w == "bsdf_eval") { Spec r = bsdfEval(b, its, wo); out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; return 3; }
        if (w == "bsdf_pdf") { out[0] = bsdfPdf(b, its, wo); return 1; }
        double pdf = 0, eta = 0;
        V3 wo2;
        Spec r = bsdfSample(b, its, &in[11], wo2, pdf, eta);
        out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; out[3] = wo2.x; out[4] = wo2.y; out[5] = wo2.z; out[6] = pdf; out[7] = eta;
        return 8;
    }
    if (w == "ray_triangle") {
        V3 r = rayIntersectTriangle(V3(in[0], in[1], in[2]), V3(in[3], in[4], in[5]), V3(in[6], in[7], in[8]),
                                    V3(in[9], in[10], in[11]), V3(in[12], in[13], in[14]));
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
        return 3;
    }
    if (w == "ray_triangle_ad") {   // one tangent lane: v0,v1,v2,o,d values then their tangents
        DV<1> a(V3(in[0], in[1], in[2])), b(V3(in[3], in[4], in[5])), c(V3(in[6], in[7], in[8])),
            o(V3(in[9], in[10], in[11])), d(V3(in[12], in[13], in[14]));
        a.setGrad(0, V3(in[15], in[16], in[17])); b.setGrad(0, V3(in[18], in[19], in[20]));
        c.setGrad(0, V3(in[21], in[22], in[23])); o.setGrad(0, V3(in[24], in[25], in[26]));
        d.setGrad(0, V3(in[27], in[28], in[29]));
        D<1> u, v, t;
        rayIntersectTriangleAD<1>(a, b, c, o, d, u, v, t);
        out[0] = u.v; out[1] = v.v; out[2] = t.v; out[3] = u.g[0]; out[4] = v.g[0]; out[5] = t.g[0];
        return 6;
    }
    if (w == "distrib_sample") {   // in: n, weights..., u -> index, pdf, reused u
        Distrib d;
        int n = (int)in[0];
        for (int i = 0; i < n; ++i) d.append(in[1 + i]);
        d.normalize();
        double u = in[1 + n], pdf;
        out[0] = (double)d.sampleReuse(u, pdf);
        out[1] = pdf;
        out[2] = u;
        return 3;
    }
    g_err = "unknown unit " + w;
    return -1;
}

}  // extern "C"
