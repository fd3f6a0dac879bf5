/*
 * dtrt_testing.h — measurement and unit-test hooks of libdtrt_mi355x.
 *
 * These entry points have no counterpart in the reference. They exist so that the
 * parity tests can drive individual device stages (BVH closest hit, primal Li,
 * medium estimators, grid lookups with tangents, the RNG) on explicit inputs, and
 * so that the benchmark can read per-kernel HIP-event times and work counters of
 * the last render.
 */
#ifndef DTRT_MI355X_TESTING_H
#define DTRT_MI355X_TESTING_H

#include <stdint.h>
#include "dtrt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel times of the last render (HIP events on the scene's stream), in ms, summed
 * per kernel: [0] AD paths, [1] child sub-paths, [2] primary edges, [3] finalize;
 * launches4 = launch counts. work12: [0..3] AD-path kernel, [4..7] child kernel,
 * [8..11] primary-edge kernel; each = box tests, triangle tests, grid lookups,
 * traced segments. */
int dtrt_test_kernel_times(dtrt_object *scene, double *ms4, int64_t *launches4, uint64_t *work12);
/* Forward-mode tangent lanes w.r.t. single grid nodes: lane j (< n) of medium
 * `medium` differentiates w.r.t. the density of node voxels[j] (flat [z][y][x]
 * index, -1 = none); n = 0 clears. Other tangent sources still add to the lanes.
 * Used to check dtrt_render_adjoint lane by lane. */
int dtrt_test_set_voxel_lanes(dtrt_object *scene, int medium, const int *voxels, int n);
/* Bytes of child-record queue the scene's renders may use (0 = default 6 GiB).
 * A small budget forces the exact overflow-replay path. */
int dtrt_test_set_queue_budget(dtrt_object *scene, uint64_t bytes);
/* on != 0: the scene's kernels read the BVH from HBM only; 0 (the default): a launch
 * keeps the tree, or its top levels, in LDS where that costs no occupancy. Images are
 * the same either way up to summation order. */
int dtrt_test_set_tree_in_hbm(dtrt_object *scene, int on);
/* [0] BVH nodes, [1] triangles, [2] edges, [3] primary edges, [4] primary-edge
 * length sum (double bits), [5] secondary edges usable, [6] vertices, [7] faces. */
int dtrt_test_scene_info(dtrt_object *scene, int64_t *out8);
/* Host copy of a heterogeneous medium's grid as the constructor stored it: which = 0
 * density [z][y][x], 1 albedo [z][y][x][3]; up to `cap` floats into out (may be NULL).
 * meta8: {rx, ry, rz, channels, floats stored, max density, scale, 0} (doubles).
 * Returns 0, or -1 for a handle that is not a medium. */
int dtrt_test_medium_data(dtrt_object *medium, int which, float *out, int64_t cap, double *meta8);
/* Compiled resources of the kernels a scene of dual width nder (1, 2, 3, 4, 8, 16)
 * and estimator precision `bits` (32, 64) launches: up to cap kernels, each a
 * NUL-terminated name in names[i * name_len ..] and five values in vals[5 i ..]:
 * private (scratch) bytes per lane, VGPRs, work-group size, waves per CU at that size,
 * scratch bytes over the device at that occupancy. Returns the number of kernels
 * (which may exceed cap), or -1. */
int dtrt_test_kernel_resources(int nder, int bits, char *names, int name_len, uint64_t *vals, int cap);
/* out6: [0] scratch bytes per lane and [1] over the device above which a scene's
 * kernels are refused; the runtime's [2] stack-size limit per thread and [3..5] scratch
 * limit minimum, maximum and current value (hipDeviceGetLimit; 0 where unsupported). */
int dtrt_test_scratch_limits(uint64_t *out6);
/* Lowers the scratch limits scenes are checked against (0 restores a default; values
 * above the defaults are clamped to them), so that the refusal can be tested. */
int dtrt_test_set_scratch_limits(uint64_t lane_bytes, uint64_t device_bytes);
/* Edge list of a shape (v0, v1, f0, f1 local ids; f1 = -1 on boundaries) and lengths;
 * returns the number of edges. */
int dtrt_test_shape_edges(dtrt_object *shape, int *out_v0v1f0f1, double *out_len, int cap);
/* Closest hit of n rays {ox,oy,oz,dx,dy,dz,tnear,tfar}: global face id (-1 = miss)
 * and the refined distance. */
int dtrt_test_trace(dtrt_object *scene, const float *rays8, int n, int *out_face, float *out_t);
/* Primal radiance Li of n rays {o, d} (one RNG stream per ray). */
int dtrt_test_li(dtrt_object *scene, const float *rays6, int n, int max_depth, int medium, uint64_t seed, float *out3);
/* Medium estimators along segments {o, d, tmin, tmax}: kind 0 ratio-tracking
 * transmittance, kind 1 delta-tracking scatter indicator. */
int dtrt_test_medium(dtrt_object *scene, int medium, int kind, const float *seg8, int n, uint64_t seed, float *out);
/* Density lookups with tangents: points {x, y, z, nder tangent vectors} ->
 * {value, nder derivatives}. */
int dtrt_test_grid(dtrt_object *scene, int medium, const float *pts, int n, float *out);
/* Host-only BVH build of ntri triangles {v0, v1, v2} (9 floats each) with the shape
 * index of each (may be NULL: one shape). out8: [0] inner nodes, [1] deepest leaf
 * depth (root = 1), [2] triangles under the root's first child, [3] under its second
 * child, [4] most stack entries a traversal can push (binary: one per level below the
 * root), [5] leaves, [6] largest leaf. Needs no GPU. */
int dtrt_test_build_bvh(const float *tri9, const int *shape_of, int ntri, int64_t *out8);
/* One Philox4x32-10 block. */
int dtrt_test_philox(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4);

#ifdef __cplusplus
}
#endif
#endif
