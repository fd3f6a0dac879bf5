// Device scene layout (plain data shared by the host builder and the kernels).
//
// HBM layout: all shapes are flattened into global vertex / face / edge arrays;
// per-lane tangents are stored entity-major ([entity][nder][3]) so the N lanes of
// one vertex are one contiguous gather. The BVH is a binary tree of 64-byte nodes
// holding both children's boxes; triangles are stored in leaf order as three
// float4 (v0 | global face id, e1, e2) so a leaf is one contiguous 48*k-byte run.
#pragma once
#include <cstdint>

#define DTRT_NDER_MAX 16
// Traversal stack entries per thread (LDS, 4 B each): the builder keeps every leaf
// at depth <= DTRT_BVH_STACK - 2, so traversal never pushes more than that.
#define DTRT_BVH_STACK 32

namespace dt {

enum BsdfType { BSDF_DIFFUSE = 0, BSDF_NULL = 1, BSDF_PHONG = 2, BSDF_ROUGHDIELECTRIC = 3 };
enum PhaseType { PHASE_HG = 0, PHASE_ISO = 1 };
enum EmitterType { EMIT_AREA = 0, EMIT_AREA_EX = 1 };
enum MediumType { MED_HOMOGENEOUS = 0, MED_HETEROGENEOUS = 1 };

struct DCamera {
    int width, height, ox, oy, cw, ch, cropped, med_id;
    float c2w[12];          // rows of the 3x4 camera-to-world transform
    float ndc2c[9];
    float fov, aspect, clip_near, pad;
    float cpos[3], cs[3], ct[3], cn[3];
    float cpos_d[DTRT_NDER_MAX][3], cs_d[DTRT_NDER_MAX][3], ct_d[DTRT_NDER_MAX][3], cn_d[DTRT_NDER_MAX][3];
    float w2c[12];          // rows of the 3x4 world-to-camera transform (light tracing)
    float c2ndc[9];
};

struct DShape {
    int vbase, fbase, nv, nf;
    int light_id, bsdf_id, med_int, med_ext;
    int has_normals, ebase, ne, pad;
};

struct DBsdf {
    int type, pad0;
    float refl[3], kd[3], ks[3], ex, specW, alpha, eta, pad1;
    float refl_d[DTRT_NDER_MAX][3], kd_d[DTRT_NDER_MAX][3], ks_d[DTRT_NDER_MAX][3];
    float ex_d[DTRT_NDER_MAX], alpha_d[DTRT_NDER_MAX], eta_d[DTRT_NDER_MAX];
};

struct DPhase {
    int type;
    float g;
    float g_d[DTRT_NDER_MAX];
};

struct DEmitter {
    int type, shape_id, two_sided;
    float I[3], kappa;
    float I_d[DTRT_NDER_MAX][3], kappa_d[DTRT_NDER_MAX];
};

struct DMedium {
    int type, phase_id;
    // homogeneous
    float sigma_t, sampling_weight;
    float albedo[3];
    float sigma_t_d[DTRT_NDER_MAX], albedo_d[DTRT_NDER_MAX][3];
    // heterogeneous: node-centred grid, trilinear, rigid motion about `center`
    int res[3], ares[3];
    const float *density;       // [z][y][x]
    const float *albedo_grid;   // [z][y][x][3] or null
    float w2v[12];
    float bmin[3], gscale[3];   // grid coordinate = (p - bmin) * gscale
    float amin[3], agscale[3];
    float center[3], aabb_min[3], aabb_max[3];
    float maxDensity, maxInvDensity, scale;
    float scale_d[DTRT_NDER_MAX], trans_d[DTRT_NDER_MAX][3], rot_d[DTRT_NDER_MAX][3];
    // optional per-voxel tangent lanes: lane j differentiates w.r.t. the density of
    // grid node vox_lane[j] (flat [z][y][x] index, -1 = none); vox_any = any set
    int vox_any;
    int vox_lane[DTRT_NDER_MAX];
};

struct DScene {
    int nder, nshapes, nlights, nmedia;
    int f64;                    // estimator precision of the scene's kernels: 1 double, 0 float
    int NV, NF, ntris, nnodes;
    int stack_depth;            // traversal stack entries a lane can need: the tree's depth (<= DTRT_BVH_STACK)
    const float4 *nodes;        // 4 float4 per node
    const float4 *tris;         // 3 float4 per triangle, leaf order
    const float *V;             // [NV][3]
    const float *Vd;            // [NV][nder][3]
    const float *Nrm;           // [NV][3] shading normals (shapes with normals only)
    const float *Nd;            // [NV][nder][3]
    const int *F;               // [NF][3] global vertex ids
    const int *face_shape;      // [NF]
    const float *FN;            // [NF][3]
    const float *FNd;           // [NF][nder][3]
    const DShape *shapes;
    const DBsdf *bsdfs;
    const DPhase *phases;
    const DEmitter *emitters;
    const DMedium *media;
    // light sampling (scene.cpp:45-98)
    const double *light_cdf;    // [nlights]
    const float *light_pmf, *light_areas;
    const int *area_cdf_begin;
    const double *area_cdf_pool;
    // secondary edges (bruteforce.cpp, shape.cpp:254-300)
    const int4 *edges;          // v0, v1, f0, f1 (global ids; f1 = -1 for boundary edges)
    const double *edge_cdf;     // per shape: ne+1 entries starting at ebase + shape index
    const double *shape_edge_cdf;   // nshapes+1
    int edges_valid;
    int nviews;
    const struct DView *views;  // the cameras of the render (one for dtrt_render)
};

// One view of a render: camera and output window, its share of the global AD sample
// and pixel index spaces, and the primary edges projected for its camera.
struct DView {
    DCamera cam;
    uint64_t s0;                // first global AD sample (samples of view v: s0 .. s0 + npix*spp)
    uint64_t seed;              // render seed of this view
    int spp, sppe;
    int npix, pix0;             // pixels of the output window; first global (accumulator) pixel
    int npe, pad;
    double pe_sum;
    const double *pe_cdf;       // npe+1 (edge_manager.cpp:6-134)
    const float *pe;            // [npe][16 + 4*nder]
};

// Child primal sub-path records written by the AD-path kernel and the
// primary-edge kernel, consumed by the child kernel.
enum ChildKind { CHILD_LI = 0, CHILD_LINS_NULL = 1, CHILD_LINS_SURF = 2, CHILD_PATH = 3 };   // PATH: path.cpp radiance

// R: the estimator's scalar type (float or double), so a sub-path starts from the
// exact point and direction its parent computed.
template <class R> struct ChildHdrT {
    R org[3];
    R dir[3];                   // CHILD_LI: ray direction; LINS: incoming ray direction
    int pixel;
    int meta;                   // kind | (max_depth << 4) | ((med + 1) << 12) | (face id in LINS records via `face`)
    uint32_t key_lo, key_hi;    // RNG stream of this child
    int face;                   // LINS: global face id of the boundary hit
    R gn[3], sn[3];             // LINS: geometric and shading normals at the hit
    int view;                   // view of the sample that pushed the record (its seed keys the stream)
};
// followed by R w[nder][3] (weights applied to the child's RGB radiance); records are
// 16-byte aligned
inline int childRecordStride(int nder, bool f64) {
    const size_t h = f64 ? sizeof(ChildHdrT<double>) : sizeof(ChildHdrT<float>);
    return (int)((h + 3 * (size_t)nder * (f64 ? 8 : 4) + 15) / 16 * 16);
}

}  // namespace dt
