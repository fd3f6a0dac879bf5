// Launchers of the gfx950 kernels by estimator precision: each forwards to the set of
// the scene's precision (dt::p64, double, or dt::p32, float; S.f64). Host code only.
#include "launch.h"

namespace dt {

#define DT_BY_PREC(S, FN, ...) ((S).f64 ? p64::FN(__VA_ARGS__) : p32::FN(__VA_ARGS__))

hipError_t setFrameRotation(const double *m9) {
    hipError_t e = p32::setFrameRotation(m9);
    return e == hipSuccess ? p64::setFrameRotation(m9) : e;
}
hipError_t launch_ad_paths(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const uint4 *replay, void *acc,
                           char *q, unsigned *qc, unsigned *st, uint4 *ovf, unsigned *ovf_count, unsigned long long *work,
                           hipStream_t sm) {
    return DT_BY_PREC(S, launch_ad_paths, S, P, s0, n, replay, acc, q, qc, st, ovf, ovf_count, work, sm);
}
hipError_t launch_children(const DScene &S, const RenderParams &P, const char *q, unsigned *qc, void *acc, void *acc_pe,
                           unsigned long long *work, unsigned grid, hipStream_t sm) {
    return DT_BY_PREC(S, launch_children, S, P, q, qc, acc, acc_pe, work, grid, sm);
}
hipError_t launch_primary(const DScene &S, const RenderParams &P, int view, uint64_t s0, uint64_t n, uint64_t total, char *q,
                          unsigned *qc, unsigned *st, unsigned long long *work, hipStream_t sm) {
    return DT_BY_PREC(S, launch_primary, S, P, view, s0, n, total, q, qc, st, work, sm);
}
hipError_t launch_finalize(const DScene &S, const RenderParams &P, int npix, int pix0, int spp, int has_pe, const void *acc,
                           const void *acc_pe, float *out, unsigned *st, hipStream_t sm) {
    return DT_BY_PREC(S, launch_finalize, S, P, npix, pix0, spp, has_pe, acc, acc_pe, out, st, sm);
}
hipError_t launch_adjoint(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, const float *dimg, float *grad,
                          int pmed, unsigned *st, unsigned long long *work, hipStream_t sm) {
    return DT_BY_PREC(S, launch_adjoint, S, P, s0, n, dimg, grad, pmed, st, work, sm);
}
hipError_t launch_fill32(void *p, uint64_t n, uint32_t v, hipStream_t sm) { return p32::launch_fill32(p, n, v, sm); }
hipError_t launch_add_u32(unsigned *dst, const unsigned *src, hipStream_t sm) { return p32::launch_add_u32(dst, src, sm); }
hipError_t launch_load_density(float *dst, const float *src, uint64_t n, unsigned *maxbits, hipStream_t sm) {
    return p32::launch_load_density(dst, src, n, maxbits, sm);
}
hipError_t launch_l2_grad(const float *img, const float *target, float *dimg, uint64_t n, double *loss, hipStream_t sm) {
    return p32::launch_l2_grad(img, target, dimg, n, loss, sm);
}
hipError_t launch_contract(const float *img, const float *dimg, uint64_t n, int nder, double *grad, hipStream_t sm) {
    return p32::launch_contract(img, dimg, n, nder, grad, sm);
}
hipError_t launch_tri_records(const DScene &S, hipStream_t sm) { return p32::launch_tri_records(S, sm); }
hipError_t launch_refit(const DScene &S, const int *list, int n, hipStream_t sm) {
    return p32::launch_refit(S, list, n, sm);
}
hipError_t launch_primal_paths(const DScene &S, const RenderParams &P, int kind, uint64_t s0, uint64_t n, void *acc,
                               unsigned long long *work, hipStream_t sm) {
    return DT_BY_PREC(S, launch_primal_paths, S, P, kind, s0, n, acc, work, sm);
}
hipError_t launch_particles(const DScene &S, const RenderParams &P, uint64_t s0, uint64_t n, void *acc,
                            unsigned long long *work, hipStream_t sm) {
    return DT_BY_PREC(S, launch_particles, S, P, s0, n, acc, work, sm);
}
hipError_t launch_scale(const DScene &S, const void *acc, float *out, uint64_t n, double scale, hipStream_t sm) {
    return DT_BY_PREC(S, launch_scale, S, acc, out, n, scale, sm);
}
hipError_t launch_trace(const DScene &S, const float *rays, int n, int *face, float *t, hipStream_t sm) {
    return DT_BY_PREC(S, launch_trace, S, rays, n, face, t, sm);
}
hipError_t launch_li(const DScene &S, const float *rays, int n, int depth, int med, uint64_t seed, float *out,
                     hipStream_t sm) {
    return DT_BY_PREC(S, launch_li, S, rays, n, depth, med, seed, out, sm);
}
hipError_t launch_medium(const DScene &S, int med, int kind, const float *seg, int n, uint64_t seed, float *out,
                         hipStream_t sm) {
    return DT_BY_PREC(S, launch_medium, S, med, kind, seg, n, seed, out, sm);
}
hipError_t launch_grid(const DScene &S, int med, const float *pts, int n, float *out, hipStream_t sm) {
    return DT_BY_PREC(S, launch_grid, S, med, pts, n, out, sm);
}
hipError_t launch_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t sm) {
    return p32::launch_philox(ctr, key, out, sm);
}
hipError_t kernel_info(const DScene &S, KernelInfo *out, int cap, int *count) {
    return DT_BY_PREC(S, kernel_info, S, out, cap, count);
}

}  // namespace dt
