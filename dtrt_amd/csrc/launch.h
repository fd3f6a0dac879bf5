// Host-visible launch interface of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "scene_types.h"

namespace dt {

// Per-render parameters (RenderOptions, integrator.h:10-25, plus the image layout).
// Settings shared by all views of a render (per-view ones are in DView).
// The forward-mode integrator of a render (IntegratorAD subclasses).
enum AdKind { AD_VOLPATH = 0, AD_PATH = 1, AD_DIRECT = 2 };   // volpathAD.cpp, pathAD.cpp, directAD.cpp

struct RenderParams {
    int ad_kind;             // AdKind
    int max_bounces, sppse, mode;
    float ddist;
    int npix;                // pixels of all views (accumulator stride)
    int rec_stride;          // bytes per child record (ChildHdr + 12*nder, 16-byte aligned)
    unsigned qcap;           // child records the queue can hold
    int ad_block, child_block;   // work-group sizes of the AD-path and child kernels (64, 128, 256; 0 = 64)
    int lds_nodes, lds_tris;     // set per launch: the part of the BVH the kernel copies into LDS (treeToLds)
    int tree_in_hbm;             // nonzero: no part of the BVH goes into LDS (a test setting)
};

// Counters written by the kernels (warnings the reference prints, as counts).
enum Stat {
    STAT_NEG_PIXEL = 0, STAT_NONFINITE, STAT_DT_CAP, STAT_BAD_EDGE_WEIGHT, STAT_NONORTHO, STAT_MED_INCONSISTENT,
    STAT_QUEUE_OVERFLOW, STAT_COUNT
};

// Screen-space half-gap between the two primary-edge sub-paths, in pixels
// (integratorAD.cpp:107). The estimate is not insensitive to it: faces seen nearly
// edge-on near a silhouette are narrower than 1e-3 px, and a wider gap steps over
// them (a 24-px view of a subdivided sphere moves by ~7% between 1e-3 and 1e-4).
// The side rays are built in double and rounded once; 1e-4 px stays several ulp of
// an fp32 direction apart up to ~1000-px-wide images.
constexpr double PRIMARY_DELTA = 1e-4;

// Compiled resources of one kernel (hipFuncGetAttributes) at the largest work-group size
// it is launched with, and the waves of it one CU holds at that size
// (hipOccupancyMaxActiveBlocksPerMultiprocessor).
struct KernelInfo {
    char name[40];
    uint32_t private_bytes;   // private segment per lane: register spills and private arrays
    uint32_t vgprs;
    uint32_t block;
    uint32_t waves_per_cu;
};

// The launchers exist once per estimator precision: in namespace p32 (float) and p64
// (double), each defined by the kernel objects of that precision, and once in
// namespace dt itself, where they dispatch on S.f64 (dispatch.cpp). Accumulators
// (`acc`, `acc_pe`) hold the estimator's type: float for p32, double for p64.
//
// setFrameRotation: the fixed local-frame rotation (row-major 3x3) into every kernel object.
// launch_ad_paths: work[0..3] AD paths, [4..7] child sub-paths, [8..11] primary-edge
// samples: box tests, triangle tests, density lookups, traced segments.
// AD samples s0 .. s0+n-1, or (replay != nullptr) the n listed {sid lo, sid hi,
// children delivered, 0} entries; samples whose child pushes overflow the queue are
// appended to ovf[*ovf_count] (capacity >= n) and contribute nothing in this launch.
// launch_children: drains the qc[0] queued records (qc[2] = 0 on entry: the consumption counter).
// launch_primary: primary-edge samples s0 .. s0+n-1 (of `total`) of view `view`.
// launch_finalize: one view, accumulator pixels pix0 .. pix0+npix into `out` (that view's image).
// launch_adjoint: path-replay adjoint over AD samples s0 .. s0+n-1 (interior terms only):
// adds d loss / d node density of medium pmed's grid into grad, given d loss / d pixel.
// launch_fill32: p[0..n) = v (32-bit words); launch_add_u32: *dst += *src. Kernels, so
// that captured graphs hold no memset nodes.
// launch_load_density: dst = max(src, 0) over n floats; atomicMax of the result's bits
// into *maxbits.
// launch_l2_grad: one view of the L2 loss sum (I - T)^2: dimg = 2 (I - T) over n floats,
// the loss added into *loss. launch_contract: grad[k] += sum dimg * (tangent image k)
// for a view image [(nder + 1)][n] (forward-mode lanes contracted with d loss / d I).
// launch_tri_records: the leaf-order triangle records from S.V; launch_refit: the boxes of
// the n nodes listed (one tree level, children refitted first).
// launch_primal_paths: samples s0 .. s0+n-1 of view 0, radiance sums into acc[pixel][3].
// launch_particles: ParticleTracer particles s0 .. s0+n-1 splatted into acc[pixel][3]
// of view 0's window.
// launch_scale: out[i] = acc[i] * scale, i < n.
// kernel_info: the kernels a render at S.nder launches, then the width-independent ones;
// *count of them written into out[0 .. cap).
#define DT_LAUNCH_API                                                                                                   \
    hipError_t setFrameRotation(const double *m9);                                                                       \
    hipError_t launch_ad_paths(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const uint4 *replay,    \
                               void *acc, char *q, unsigned *qc, unsigned *st, uint4 *ovf, unsigned *ovf_count,          \
                               unsigned long long *work, hipStream_t sm);                                                \
    hipError_t launch_children(const DScene &S, const RenderParams &P, const char *q, unsigned *qc, void *acc,          \
                               void *acc_pe, unsigned long long *work, unsigned grid, hipStream_t sm);                   \
    hipError_t launch_primary(const DScene &S, const RenderParams &P, int view, uint64_t s0, uint64_t n,                 \
                              uint64_t total, char *q, unsigned *qc, unsigned *st, unsigned long long *work,             \
                              hipStream_t sm);                                                                           \
    hipError_t launch_finalize(const DScene &S, const RenderParams &P, int npix, int pix0, int spp, int has_pe,          \
                               const void *acc, const void *acc_pe, float *out, unsigned *st, hipStream_t sm);           \
    hipError_t launch_adjoint(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const float *dimg,        \
                              float *grad, int pmed, unsigned *st, unsigned long long *work, hipStream_t sm);            \
    hipError_t launch_fill32(void *p, uint64_t n, uint32_t v, hipStream_t sm);                                           \
    hipError_t launch_add_u32(unsigned *dst, const unsigned *src, hipStream_t sm);                                       \
    hipError_t launch_load_density(float *dst, const float *src, uint64_t n, unsigned *maxbits, hipStream_t sm);         \
    hipError_t launch_l2_grad(const float *img, const float *target, float *dimg, uint64_t n, double *loss,             \
                              hipStream_t sm);                                                                           \
    hipError_t launch_contract(const float *img, const float *dimg, uint64_t n, int nder, double *grad, hipStream_t sm); \
    hipError_t launch_tri_records(const DScene &S, hipStream_t sm);                                                     \
    hipError_t launch_refit(const DScene &S, const int *list, int n, hipStream_t sm);                                   \
    hipError_t launch_primal_paths(const DScene &S, const RenderParams &P, int kind, uint64_t s0, uint64_t n, void *acc,  \
                                   unsigned long long *work, hipStream_t sm);                                            \
    hipError_t launch_particles(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, void *acc,             \
                                unsigned long long *work, hipStream_t sm);                                               \
    hipError_t launch_scale(const DScene &S, const void *acc, float *out, uint64_t n, double scale, hipStream_t sm);    \
    hipError_t launch_trace(const DScene &S, const float *rays, int n, int *face, float *t, hipStream_t sm);             \
    hipError_t launch_li(const DScene &S, const float *rays, int n, int depth, int med, uint64_t seed, float *out,       \
                         hipStream_t sm);                                                                                \
    hipError_t launch_medium(const DScene &S, int med, int kind, const float *seg, int n, uint64_t seed, float *out,     \
                             hipStream_t sm);                                                                            \
    hipError_t launch_grid(const DScene &S, int med, const float *pts, int n, float *out, hipStream_t sm);              \
    hipError_t launch_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t sm);                \
    hipError_t kernel_info(const DScene &S, KernelInfo *out, int cap, int *count);

enum PrimalKind { PRIMAL_VOLPATH = 1, PRIMAL_SIMPLE = 2, PRIMAL_PATH = 3 };

// The allocation-free render's device scratch (scratch.hip): bytes per GPU, and the
// device address of the current device's copy (looked up once per device).
#define DTRT_SCRATCH_BYTES (uint64_t(256) << 20)
hipError_t scratchArea(void **ptr, uint64_t *bytes);

namespace p32 {
DT_LAUNCH_API
}
namespace p64 {
DT_LAUNCH_API
}
DT_LAUNCH_API

}  // namespace dt
