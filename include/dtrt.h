/*
 * dtrt.h — C-ABI of the MI355X (gfx950) differentiable volumetric path tracer.
 *
 * Drop-in for the reference's pybind11 module `dtrt` (src/dtrt.cpp:24-155): the
 * object constructors take the same arguments, with the same meaning and tensor
 * layouts, as the pybind constructors they replace; the render call fills the same
 * caller-owned float32[(nder+1)][H][W][3] image (integratorAD.cpp:66-85).
 *
 * Conventions
 *  - Pointers are plain host pointers to float32 / int32 data; NULL where the
 *    reference accepts a null `float_ptr(0)` (src/dtrt.cpp:27-30, include/ptr.h).
 *  - Constructors copy their input arrays (shape.cpp:14-32, camera.cpp:13-27);
 *    a scene references its camera/shapes/materials: keep them alive until the
 *    scene is destroyed (scene.h:42-46 keeps raw pointers as well).
 *  - Tangent ("velocity") layouts, with nder = dtrt_nder():
 *      camera      [2][nder][3]  translations, then rotation axes (camera.cpp:29-37)
 *      shape       [nder][nv][3] vertex tangents; if normals are given, the normal
 *                  tangents are read from offset nder*nv*3 for every lane, exactly
 *                  as shape.cpp:45 does
 *      spectra     [nder][3]; scalars [nder]
 *      medium      trans/rot [nder][3] (heterogeneous.h:187-201)
 *  - Functions returning int give 0 on success and a negative code on failure;
 *    constructors return NULL on failure. dtrt_last_error() describes the last
 *    failure of the calling thread. Nothing aborts the host process.
 *  - One render per scene at a time; different scenes may render concurrently.
 */
#ifndef DTRT_MI355X_H
#define DTRT_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dtrt_object dtrt_object;   /* every handle below is a dtrt_object */

/* RenderOptions (include/integrator.h:10-25; bound at src/dtrt.cpp:104-115). */
typedef struct dtrt_render_options {
    uint64_t seed;
    int32_t spp;           /* num_samples */
    int32_t max_bounces;
    int32_t sppe;          /* num_samples_primary_edge (<0: pixel centres, no jitter) */
    int32_t sppse;         /* num_samples_secondary_edge */
    int32_t quiet;
    int32_t mode;          /* EMode bitmask (integratorAD.h:12-24); -1 = all terms */
    float ddist_coeff;     /* ddistCoeff */
} dtrt_render_options;

/* Module attributes: `dtrt.nder`, `dtrt.angleEps` (src/dtrt.cpp:152-153). The
 * reference fixes nder at compile time (include/config.h:5); here it is a process
 * setting (default: env DTRT_NDER, else 4) that must be chosen before objects are
 * created. Supported: 1, 2, 3, 4, 8, 16. */
int dtrt_abi_version(void);
int dtrt_set_nder(int nder);
int dtrt_nder(void);
/* Estimator precision of scenes created from now on, in bits: 64 (double, the
 * reference's Float, include/fwd.h:10-13; the default) or 32 (float). Like nder the
 * reference fixes it at compile time; here it is a process setting (env
 * DTRT_PRECISION, else 64) read when a scene is created. Dual numbers, tracking,
 * shading and accumulation use it; BVH traversal is fp32 in both, as Embree's is
 * (src/scene.cpp:17-43), and outputs are float32 as the reference's. */
int dtrt_set_precision(int bits);
int dtrt_precision(void);
/* The precision a scene was created with (32 or 64), -1 for a bad handle. */
int dtrt_scene_precision(dtrt_object *scene);
double dtrt_angle_eps(void);
const char *dtrt_last_error(void);
void dtrt_destroy(dtrt_object *obj);

/* dtrt.Camera(w, h, cam_to_world[16] row-major, cam_to_ndc[9], clip_near, med_id, velocities)
 * — src/dtrt.cpp:35-37, camera.cpp:8-38. */
dtrt_object *dtrt_camera_create(int width, int height, const float *cam_to_world, const float *cam_to_ndc, float clip_near,
                                int med_id, const float *velocities);
/* Camera.set_rect(ox, oy, w, h) — src/dtrt.cpp:37, camera.h:27. */
int dtrt_camera_set_rect(dtrt_object *camera, int offset_x, int offset_y, int crop_width, int crop_height);

/* dtrt.Shape(vertices[nv][3], indices[nf][3], uvs[nv][2], normals[nv][3], nv, nf, light_id, bsdf_id,
 *            med_int_id, med_ext_id, velocities) — src/dtrt.cpp:39-50, shape.cpp:10-52. */
dtrt_object *dtrt_shape_create(const float *vertices, const int *indices, const float *uvs, const float *normals,
                               int num_vertices, int num_triangles, int light_id, int bsdf_id, int med_int_id,
                               int med_ext_id, const float *velocities);

/* dtrt.BSDF_diffuse(Spectrum3f[, d_reflectance]) — src/dtrt.cpp:56-58, bsdf/diffuse.h:9-14. */
dtrt_object *dtrt_bsdf_diffuse_create(const float *reflectance3, const float *d_reflectance);
/* dtrt.BSDF_null() — src/dtrt.cpp:59-60. */
dtrt_object *dtrt_bsdf_null_create(void);
/* dtrt.BSDF_Phong(kd, ks, exponent[, dkd, dks, dexp]) — src/dtrt.cpp:61-63, bsdf/phong.h:9-39. */
dtrt_object *dtrt_bsdf_phong_create(const float *kd3, const float *ks3, float exponent, const float *d_kd,
                                    const float *d_ks, const float *d_exponent);
/* dtrt.BSDF_roughdielectric(alpha, intIOR, extIOR[, d_alpha[, d_eta]]) — src/dtrt.cpp:64-67. */
dtrt_object *dtrt_bsdf_roughdielectric_create(float alpha, float int_ior, float ext_ior, const float *d_alpha,
                                              const float *d_eta);
/* dtrt.HG(g[, d_g]) and dtrt.Isotropic() — src/dtrt.cpp:70-74, include/phase.h:24-79. */
dtrt_object *dtrt_phase_hg_create(float g, const float *d_g);
dtrt_object *dtrt_phase_isotropic_create(void);
/* dtrt.AreaLight(shape_id, intensity_ptr, two_sided) — src/dtrt.cpp:77-78, emitter/area.h:10-15. */
dtrt_object *dtrt_area_light_create(int shape_id, const float *intensity3, int two_sided);
/* dtrt.AreaLightEx(shape_id, intensity, kappa[, d_intensity, d_kappa]) — src/dtrt.cpp:79-81. */
dtrt_object *dtrt_area_light_ex_create(int shape_id, const float *intensity3, float kappa, const float *d_intensity,
                                       const float *d_kappa);
/* dtrt.Homogeneous(sigma_t, albedo, phase_id[, d_sigma_t, d_albedo]) — src/dtrt.cpp:84-86. */
dtrt_object *dtrt_homogeneous_create(float sigma_t, const float *albedo3, int phase_id, const float *d_sigma_t,
                                     const float *d_albedo);
/* dtrt.Heterogeneous(density_file, toWorld[16], scale, albedo | albedo_file, phase_id
 *                    [, trans_vel, rot_vel, d_density[, d_albedo]]) — src/dtrt.cpp:87-93,
 * medium/heterogeneous.h:13-106. Exactly one of albedo3 / albedo_file is used
 * (albedo_file != NULL selects the VOL albedo grid). */
dtrt_object *dtrt_heterogeneous_create(const char *density_file, const float *to_world, float density_scale,
                                       const float *albedo3, const char *albedo_file, int phase_id,
                                       const float *trans_velocities, const float *rot_velocities,
                                       const float *d_density, const float *d_albedo);
/* In-memory variant of the above (no VOL file): density float32[rz][ry][rx] on the
 * volume-space box [bmin, bmax]. No counterpart in the reference, which only reads
 * VOL files (gridvolume.h:88-132); it is for callers that generate the grid. */
dtrt_object *dtrt_heterogeneous_create_grid(const float *density, int res_x, int res_y, int res_z, const float *bmin3,
                                            const float *bmax3, const float *to_world, float density_scale,
                                            const float *albedo3, int phase_id, const float *trans_velocities,
                                            const float *rot_velocities, const float *d_density,
                                            const float *d_albedo);

/* dtrt.Scene(camera, shapes, bsdfs, area_lights, phases, media) — src/dtrt.cpp:95-101,
 * scene.cpp:9-136 (builds the BVH, light and edge distributions, uploads to HBM). */
dtrt_object *dtrt_scene_create(dtrt_object *camera, dtrt_object *const *shapes, int num_shapes, dtrt_object *const *bsdfs,
                               int num_bsdfs, dtrt_object *const *area_lights, int num_lights,
                               dtrt_object *const *phases, int num_phases, dtrt_object *const *media, int num_media);
/* Scene.initEdges(weights[num_shapes]) — src/dtrt.cpp:102, scene.h:29-32. */
int dtrt_scene_init_edges(dtrt_object *scene, const float *weights);

/* Integrator.render(scene, options, float_ptr image) — src/dtrt.cpp:148-150,
 * integratorAD.cpp:19-166. integrator 0 = VolPathTracerAD. `image` is host memory:
 * float32[(nder+1)][H][W][3] (crop size if cropped), zeroed by the caller; values
 * and tangents the reference would drop (negative pixel, non-finite tangent) are
 * left untouched, primary-edge contributions are added. Synchronous.
 * The primal integrators write the image float32[H][W][3], every pixel the mean of spp
 * jittered samples; sppe, sppse, mode and ddistCoeff are ignored:
 *   1 = VolPathTracer       (src/dtrt.cpp:132-134, volpath.cpp:106-141)
 *   2 = VolPathTracerSimple (src/dtrt.cpp:136-138, volpath_simple.cpp:89-124)
 *   3 = PathTracer          (src/dtrt.cpp:123-125, path.cpp:64-96; surfaces only)
 *   4 = DirectIntegrator    (src/dtrt.cpp:119-121, direct.cpp:51-85; surfaces only)
 *   5 = ParticleTracer      (src/dtrt.cpp:127-129, ptracer.cpp:10-167): 100000 * spp
 *       light particles splatted into the image and divided by their count
 * The surface-only forward-mode integrators fill the VolPathTracerAD layout (media
 * ignored; primary-edge pass included):
 *   6 = PathTracerAD        (src/dtrt.cpp:144-146, pathAD.cpp)
 *   7 = DirectAD            (src/dtrt.cpp:140-142, directAD.cpp) */
int dtrt_render(int integrator, dtrt_object *scene, const dtrt_render_options *options, float *image);
/* Same, with `image` in device memory of the current HIP device (no host copies). */
int dtrt_render_device(int integrator, dtrt_object *scene, const dtrt_render_options *options, float *image_device);

/* Autotuner (no counterpart in the reference): renders `options` with the scene's
 * camera several times, by coordinate descent over the launch settings (child queue
 * budget, work-group sizes of the AD-path and child kernels, persistent child blocks
 * per CU, the queue share an AD chunk aims at, the largest AD chunk), and
 * keeps the fastest in the scene for later renders, view batches and graph captures.
 * integrator: 0, 6 or 7. best_ms (may be NULL): the fastest render; settings3 (may be
 * NULL): {blocks per CU, fill target in percent, max AD chunk}. Images and stats of
 * the last call are overwritten. */
int dtrt_scene_autotune(int integrator, dtrt_object *scene, const dtrt_render_options *options, double *best_ms,
                        int64_t *settings3);
/* The scene's launch settings, as the autotuner left them: out6 = {persistent child
 * blocks per CU, AD chunk fill target in percent, largest AD chunk in samples, child
 * queue budget in bytes, AD-path work-group size, child work-group size}. */
int dtrt_scene_launch_settings(dtrt_object *scene, int64_t *out6);
/* Sets them (same order), e.g. to reuse a tuned configuration; work-group sizes 64,
 * 128 or 256. Settings change speed only: images are the same up to summation order. */
int dtrt_scene_set_launch_settings(dtrt_object *scene, const int64_t *in6);

/* Several views of one scene in one pass. View v renders cameras[v] (NULL entry or
 * NULL array: the scene's camera) with options[v] into images[v], laid out as for
 * dtrt_render for that camera's size/crop. Views may differ in camera, resolution,
 * crop, spp, sppe and seed; max_bounces, mode, sppse and ddist_coeff must be equal.
 * Each image equals a dtrt_render of that view alone with the same options (same
 * random streams), up to float summation order. Their AD samples share the
 * launches, so small views fill the GPU together. Cameras must have been created
 * with the scene's nder. images_on_device: 1 = device pointers, 0 = host memory.
 * No counterpart in the reference (one camera per Scene). */
int dtrt_render_views(int integrator, dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                      const dtrt_render_options *options, float *const *images, int images_on_device);

/* Views of several scenes in one call: scene i renders its nviews[i] views (the next
 * nviews[i] entries of cameras, options and images, as for dtrt_render_views) on its
 * own HIP stream, driven by its own host thread, so the scenes' launches overlap on
 * the GPU. Each image equals that scene's dtrt_render_views result. A scene may be
 * listed once. Returns 0, or the first failing scene's status with its error.
 * No counterpart in the reference (one scene per render call). */
int dtrt_render_scenes(int integrator, int nscenes, dtrt_object *const *scenes, const int *nviews,
                       dtrt_object *const *cameras, const dtrt_render_options *options, float *const *images,
                       int images_on_device);

/* Render views once with dtrt_render_views (device images), then capture that launch
 * sequence (its chunk schedule fixed) as a HIP graph owned by the scene, replacing any
 * earlier one. dtrt_scene_replay_views re-renders the same views into the same images
 * with new seeds (one per view) as a single graph launch: no host round trips between
 * kernels. A replay returns 0; if a sample of the fixed schedule overflowed the child
 * queue, the views are rendered eagerly instead (exact) and it returns 1; -1 on error
 * (also when a later render reallocated the scene's buffers: capture again). Images
 * are overwritten (pixels the reference drops read 0). No counterpart in the reference. */
int dtrt_scene_capture_views(dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                             const dtrt_render_options *options, float *const *images_device);
int dtrt_scene_replay_views(dtrt_object *scene, const uint64_t *seeds);

/* Replace the density grid of heterogeneous medium `medium` (index into the scene's
 * media) with `density` float32[rz][ry][rx] of the same resolution, host memory
 * (on_device = 0) or device memory (1). Negative values are clamped to 0, as the
 * VOL reader does (gridvolume.h:88-132), and the majorant is recomputed. Only the
 * scene's copy changes. No counterpart in the reference, whose grids are fixed at
 * construction; it is how an optimiser updates the parameters between renders. */
int dtrt_scene_set_density(dtrt_object *scene, int medium, const float *density, int on_device);

/* Reverse-mode (path-replay) gradient of a loss w.r.t. every node density of
 * medium `medium`: d_density[z][y][x] += sum over pixels of d loss/d image value
 * times d image value/d node density. `d_image_device` is float32[H][W][3] (crop
 * size if cropped), i.e. d loss / d channel 0 of the image dtrt_render writes. The
 * derivative is that of the forward-mode estimator's interior terms (emission,
 * surface, volume: EMode bits 1|2|8 of options->mode), with the reference's
 * delta-tracking and ratio-tracking weights (heterogeneous.cpp:40-141); each AD
 * sample is evaluated once and replayed from its random stream. Device pointers;
 * synchronous. No counterpart in the reference, whose forward mode needs one
 * tangent lane per parameter. */
/* The same over several views (no counterpart in the reference): cameras/options as
 * for dtrt_render_views (views share max_bounces and mode); d_images_device holds
 * every view's d loss / d image value back to back ([h][w][3] per view, in view
 * order). The gradient of all views is accumulated into d_density_device. */
int dtrt_render_views_adjoint(int integrator, dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                              const dtrt_render_options *options, int medium, const float *d_images_device,
                              float *d_density_device);
int dtrt_render_adjoint(int integrator, dtrt_object *scene, const dtrt_render_options *options, int medium,
                        const float *d_image_device, float *d_density_device);

/* No pybind counterpart: the reference builds a new Scene (and Embree BVH, scene.cpp:
 * 17-43) for every optimisation step. New vertex positions for shape `shape` of the
 * scene (host, [nv][3], the constructor's layout), optionally new normals ([nv][3], only
 * for shapes created with normals) and new tangents (the constructor's velocity layout),
 * NULL keeping the old ones. Face normals, edges (crease/boundary classification and
 * lengths), edge CDFs with the last initEdges weights and light CDFs are recomputed on
 * the host; the triangle records and every BVH box are rebuilt on the GPU over the
 * existing tree (a refit: same topology, so traversal stays exact but may slow if the
 * shape moves far). Also updates the Shape object. Primary-edge projections are
 * dropped and captured graphs (views, steps) released: capture again after. */
int dtrt_scene_update_shape(dtrt_object *scene, int shape, const float *vertices, const float *normals,
                            const float *velocities);

/* No pybind counterpart (the batched inverse-rendering step; the reference's examples
 * run it as a Python loop of render, loss and backward calls, e.g.
 * examples/glass_1/optimize.py:51-68). One step of an L2 image loss
 *   L = sum over views, pixels and colour channels of (I - T)^2
 * over nviews views (cameras[v], or the scene's camera where NULL): the forward-mode
 * render of every view into images[v] ([(nder+1)][h][w][3], device), the loss into
 * *loss (double, device, may be NULL), the loss gradient of every tangent lane,
 * sum 2 (I - T) dI/dtheta_k, into param_grad[nder] (double, device, may be NULL), and
 * the path-replay density adjoint of medium `medium` with adj_options[v] (interior
 * EMode terms) into d_density (float, device; overwritten, not accumulated). targets[v]
 * are [h][w][3] device images. dtrt_scene_step runs it eagerly; capture_step runs it
 * once and records the same launch sequence as one HIP graph (the forward chunk
 * schedule is that run's); replay_step relaunches the graph with new seeds
 * (seeds[2v] forward, seeds[2v+1] adjoint) and returns 0, or 1 when a forward sample
 * overflowed the fixed schedule's queue and the step ran eagerly instead (exact
 * either way), -1 on error. An eager render on the scene in between may reallocate
 * the buffers the graph uses; replay then fails and asks for a new capture. */
int dtrt_scene_step(dtrt_object *scene, int nviews, dtrt_object *const *cameras, const dtrt_render_options *options,
                    float *const *images, const float *const *targets, int medium,
                    const dtrt_render_options *adj_options, float *d_density, double *param_grad, double *loss);
int dtrt_scene_capture_step(dtrt_object *scene, int nviews, dtrt_object *const *cameras,
                            const dtrt_render_options *options, float *const *images, const float *const *targets,
                            int medium, const dtrt_render_options *adj_options, float *d_density, double *param_grad,
                            double *loss);
int dtrt_scene_replay_step(dtrt_object *scene, const uint64_t *seeds);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same as dtrt_render_device for one view (integrator 0, 6 or 7), but
 * every device buffer of the render (accumulators, child queue, overflow lists, view
 * table, primary edges not yet cached for the camera) comes from one fixed scratch
 * area of 256 MiB per GPU that the library holds from load until the process ends;
 * the call makes no device allocation. Renders whose accumulators leave less than a
 * 16 MiB child queue in that area are refused with an error (at nder 3 in double,
 * 168 bytes of accumulators per pixel: above about 1.4 megapixels). The area keeps the last call's contents and is cleared by
 * each call; calls that overlap on one GPU (several threads or scenes at once) race
 * on it. Kernel times are not recorded (dtrt_test_kernel_times reads zeros). */
int dtrt_render_device_noalloc(int integrator, dtrt_object *scene, const dtrt_render_options *options,
                               float *image_device);

/* No pybind counterpart (the reference's render reads host memory only). Names the
 * caller's HIP stream (hipStream_t; NULL = the legacy null stream, the default). Every
 * later call on `scene` that reads or writes caller device memory (dtrt_render_device,
 * device images of dtrt_render_views, the adjoints, dtrt_scene_set_density from device
 * memory, graph replays) first waits for the work queued on that stream, so a zeroed
 * output or a loss gradient the caller has just queued is complete before the library
 * reads it. The calls return after their own work has finished. */
int dtrt_scene_set_stream(dtrt_object *scene, void *stream);

/* Counters of the last render on `scene` (the warnings the reference prints):
 * out[0] negative pixels, [1] non-finite tangents, [2] delta-tracking caps,
 * [3] invalid edge weights, [4] non-orthogonal edge tangents, [5] medium
 * inconsistencies, [6] AD samples re-run after their child pushes overflowed the
 * queue (exact replay; the image is unaffected), [7] AD path samples,
 * [8] child sub-paths, [9] primary-edge samples. */
int dtrt_render_stats(dtrt_object *scene, int64_t *out10);

#ifdef __cplusplus
}
#endif
#endif /* DTRT_MI355X_H */
