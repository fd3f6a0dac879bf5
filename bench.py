"""Benchmark: forward-mode AD path samples per second on one MI355X.

Workload (one "step"): a full render of the heterogeneous-medium configuration —
a null cube holding a 64^3 seeded fBm density grid (HG g = 0.3, albedo 0.8), a
quad light and an enclosing emissive sphere, 512x512 pixels, 128 spp, depth 10,
3 tangent lanes (medium + boundary translation in x, y, z), 10 secondary edge
samples per vertex and the primary-edge pass at 4096 samples per unit length.
Inputs (scene, BVH, grid) are resident in HBM before timing starts; the image is
written to device memory.

The estimator runs in double precision, the reference's (include/fwd.h:10-13);
`fp32_leg` repeats the workload with the float kernel set.

Prints one JSON line (see the keys below). `roofline` is computed for the kernel
with the largest share of GPU time from its work counters (box tests, triangle
tests, density lookups, traced segments) and its HIP-event time on the scene's
stream; `cpu_baseline` times the CPU oracle (a restatement of the reference
integrator, OpenMP over all host threads) on a centred crop of the same workload,
and `gpu_same_crop` times the GPU on that same crop (`gpu_over_cpu` is their ratio).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = ["ad_paths", "children", "primary_edges", "finalize"]


def algorithmic_bytes(work, nder, spp, npix, scalar_bytes=8):
    """Algorithmic HBM bytes of each kernel from its work counters.

    Per event: 32 B per density lookup (8 fp32 voxels), 36 B per triangle test
    (3 x float3), 32 B per box test (one fp32 AABB + link), and per traced segment the
    carried path state (org, dir, throughput, accumulator: 12 scalars of the estimator's
    type, `scalar_bytes` each) read + written: 24 * scalar_bytes * P bytes, with
    P = 1 + nder lanes on AD paths and P = 1 on primal sub-paths. The AD kernel also
    writes its pixel (fp32): 12 * P bytes per pixel.
    """
    P = 1 + nder
    seg_bytes = 24 * scalar_bytes
    out = {}
    for k, base, p in (("ad_paths", 0, P), ("children", 4, 1), ("primary_edges", 8, 1)):
        box, tri, lk, seg = [int(x) for x in work[base:base + 4]]
        out[k] = 32 * lk + 36 * tri + 32 * box + seg_bytes * p * seg
    out["ad_paths"] += 12 * P * npix
    out["finalize"] = 3 * npix * (scalar_bytes * (2 * P - 1) + 4 * P)
    return out


def cpu_baseline(desc, target_s=15.0):
    """Time the CPU oracle on a centred crop sized for roughly `target_s` seconds."""
    from oracle import pyoracle
    from dtrt_amd import scenes
    o = pyoracle.load()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o.set_threads(threads)

    def run(side):
        d = scenes.cube_medium(nder=desc.nder, res=desc.width, spp=desc.spp,
                           crop=((desc.width - side) // 2, (desc.height - side) // 2, side, side))
        sc = scenes.instantiate(o, d)
        t = time.perf_counter()
        scenes.render(o, d, seed=0, scene=sc)
        return time.perf_counter() - t, side * side * d.spp

    side = 8
    t, n = run(side)
    side = int(min(160, max(8, side * (target_s / max(t, 1e-3)) ** 0.5)))
    t, n = run(side)
    return {"value": n / t, "unit": "AD path samples/s", "cores": threads, "kind": "port", "crop": side,
            "sample": "CPU oracle (fp64 restatement of volpathAD/integratorAD, binned-SAH BVH, OpenMP) on a centred "
                      "%dx%d crop of the same 512x512 scene at 128 spp, depth 10, nder 3, sppse 10, sppe 4096: "
                      "%d samples in %.1f s" % (side, side, n, t)}


def gpu_crop_rate(desc, side, reps=5):
    """The GPU on the CPU baseline's own centred crop (same spp, depth, nder, edge
    sampling), at the headline precision: AD path samples/s over `reps` renders."""
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    d = scenes.cube_medium(nder=desc.nder, res=desc.width, spp=desc.spp,
                       crop=((desc.width - side) // 2, (desc.height - side) // 2, side, side))
    sc = scenes.instantiate(G, d)
    out = torch.zeros(d.out_shape(), dtype=torch.float32, device="cuda")
    opts = G.RenderOptions(0, d.spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)
    G.render_device(0, sc, opts, out.data_ptr())
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(reps):
        opts.seed = 1 + k
        G.render_device(0, sc, opts, out.data_ptr())
    torch.cuda.synchronize()
    return side * side * d.spp * reps / (time.perf_counter() - t)


def precision_leg(desc, bits, steps):
    """The headline workload with the other estimator precision's kernel set."""
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    old = G.precision()
    G.set_precision(bits)
    try:
        sc = scenes.instantiate(G, desc)
    finally:
        G.set_precision(old)
    out = torch.zeros(desc.out_shape(), dtype=torch.float32, device="cuda")
    opts = G.RenderOptions(0, desc.spp, desc.max_bounces, desc.sppe, desc.sppse, True, -1, desc.ddist)
    G.render_device(0, sc, opts, out.data_ptr())
    torch.cuda.synchronize()
    ms = 0
    t = time.perf_counter()
    for k in range(steps):
        out.zero_()
        opts.seed = k
        G.render_device(0, sc, opts, out.data_ptr())
        ms = ms + G.kernel_times(sc)[0]
    torch.cuda.synchronize()
    el = (time.perf_counter() - t) / steps
    return {"dtype": "f%d" % bits, "value": desc.width * desc.height * desc.spp / el, "unit": "AD path samples/s",
            "ms_per_step": el * 1e3, "kernel_ms_per_step": {k: float(ms[i]) / steps for i, k in enumerate(KERNELS)}}


def adjoint_leg(spp=64, res=512, grid_res=128):
    """Time to the full density gradient: one path-replay adjoint pass against forward mode.

    Workload: the heterogeneous-medium scene with a 128^3 grid (2,097,152 densities),
    512x512, 64 spp, depth 10, interior terms, L2 loss against a render of a different
    seeded grid. The forward-mode time is one 16-lane pass (16 node lanes) times the
    ceil(N / 16) passes the full gradient needs. Per sample both evaluate the same
    derivative estimator (the adjoint is its reverse mode), so equal spp is equal variance.
    """
    import numpy as np
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    interior = 1 | 2 | 8

    def desc(nder, seed):
        d = scenes.cube_medium(nder=nder, res=res, spp=spp, grid_res=grid_res, seed=seed)
        for sh in d.shapes:
            if sh.velocities is not None:
                sh.velocities = np.zeros_like(sh.velocities)
        d.media[0].trans_vel = np.zeros_like(d.media[0].trans_vel)
        return d

    def timed_render(sc, d, seed):
        out = torch.zeros(d.out_shape(), dtype=torch.float32, device="cuda")
        opts = G.RenderOptions(seed, spp, d.max_bounces, 0, 0, True, interior)
        torch.cuda.synchronize()
        t = time.perf_counter()
        G.render_device(0, sc, opts, out.data_ptr())
        torch.cuda.synchronize()
        return out, time.perf_counter() - t

    d8 = desc(1, 8)
    tgt, _ = timed_render(scenes.instantiate(G, d8), d8, 99)
    d = desc(1, 7)
    sc = scenes.instantiate(G, d)
    img, t_fwd1 = timed_render(sc, d, 1)
    dimg = (2.0 / img[0].numel()) * (img[0] - tgt[0])
    grad = torch.zeros(d.media[0].density.shape, dtype=torch.float32, device="cuda")
    opts = G.RenderOptions(2, spp, d.max_bounces, 0, 0, True, interior)
    G.render_adjoint(sc, opts, 0, dimg.contiguous().data_ptr(), grad.data_ptr())   # warm-up
    grad.zero_()
    torch.cuda.synchronize()
    t = time.perf_counter()
    G.render_adjoint(sc, opts, 0, dimg.contiguous().data_ptr(), grad.data_ptr())
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t
    nparam = int(np.prod(d.media[0].density.shape))
    d16 = desc(16, 7)
    sc16 = scenes.instantiate(G, d16)
    rng = np.random.default_rng(99)
    G.set_voxel_lanes(sc16, 0, rng.choice(nparam, 16, replace=False))
    timed_render(sc16, d16, 1)   # warm-up
    _, t16 = timed_render(sc16, d16, 1)
    passes = -(-nparam // 16)
    return {"workload": "heterogeneous-medium scene, %d^3 density grid (%d params), %dx%d, %d spp, depth %d, "
                        "interior terms, L2 loss vs a render of grid seed 8" % (grid_res, nparam, res, res, spp, d.max_bounces),
            "adjoint_s": t_adj, "forward_render_s": t_fwd1, "forward_mode_pass_s_16_lanes": t16,
            "forward_mode_passes": passes, "forward_mode_total_s": t16 * passes,
            "adjoint_speedup": t16 * passes / t_adj, "grad_nonzero_nodes": int((grad != 0).sum().item()),
            "equal_variance": "same per-sample derivative estimator in reverse and forward mode, equal spp"}


def views_leg(nviews=32):
    """Batched multi-view render (dtrt_render_views) against the same views one call each.

    The heterogeneous-medium scene (64^3 grid, nder 3, depth 10, sppse 10, sppe 4096)
    seen by 32 orbiting cameras. Resolutions are uneven, 128^2..768^2 in steps of 64, and
    spp is one of 16, 32, 64, all drawn with PCG32 seed 5.
    """
    import numpy as np
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    d = scenes.cube_medium(nder=3, res=512, spp=64)
    sc = scenes.instantiate(G, d)
    views = scenes.orbit_views(nviews)
    cams = [G.Camera(w, h, c2w.astype(np.float32), d.cam_to_ndc.astype(np.float32), d.clip_near, -1, None)
            for c2w, w, h, _ in views]
    opts = [G.RenderOptions(100 + i, spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)
            for i, (_, _, _, spp) in enumerate(views)]
    outs = [torch.zeros((d.nder + 1, h, w, 3), device="cuda") for _, w, h, _ in views]
    samples = sum(w * h * spp for _, w, h, spp in views)
    G.render_views(sc, cams, opts, outs)   # warm-up; projects and caches each camera's primary edges
    torch.cuda.synchronize()
    t = time.perf_counter()
    G.render_views(sc, cams, opts, outs)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t
    t = time.perf_counter()
    for i in range(nviews):
        G.render_views(sc, [cams[i]], [opts[i]], [outs[i]])
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t
    # small views: one captured HIP graph per optimisation step against the eager batch
    small = scenes.orbit_views(8, res_choices=(128,), spp_choices=(16,))
    scams = [G.Camera(w, h, c2w.astype(np.float32), d.cam_to_ndc.astype(np.float32), d.clip_near, -1, None)
             for c2w, w, h, _ in small]
    sopts = [G.RenderOptions(200 + i, spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)
             for i, (_, _, _, spp) in enumerate(small)]
    souts = [torch.zeros((d.nder + 1, h, w, 3), device="cuda") for _, w, h, _ in small]
    G.capture_views(sc, scams, sopts, souts)
    reps = 10
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(reps):
        G.render_views(sc, scams, sopts, souts)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    fallbacks = sum(G.replay_views(sc, [300 + 8 * k + i for i in range(len(small))]) for k in range(reps))
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t) / reps
    # several small scenes (own density grids) in one call, each on its own stream
    nsc = 8
    mds = [scenes.cube_medium(nder=3, res=128, spp=16, seed=20 + i) for i in range(nsc)]
    mscs = [scenes.instantiate(G, m) for m in mds]
    mopts = [[G.RenderOptions(400 + i, m.spp, m.max_bounces, m.sppe, m.sppse, True, -1, m.ddist)] for i, m in enumerate(mds)]
    mouts = [[torch.zeros(m.out_shape(), device="cuda")] for m in mds]
    G.render_scenes(mscs, None, mopts, mouts)   # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(reps):
        for i in range(nsc):
            G.render_views(mscs[i], None, mopts[i], mouts[i])
    torch.cuda.synchronize()
    t_ms_seq = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for k in range(reps):
        G.render_scenes(mscs, None, mopts, mouts)
    torch.cuda.synchronize()
    t_ms_tog = (time.perf_counter() - t) / reps
    return {"workload": "%d orbit views of the heterogeneous-medium scene, res %d^2..%d^2, spp 16/32/64 (PCG32 seed 5), "
                        "nder 3, depth 10, sppse 10, sppe 4096" % (nviews, min(v[1] for v in views), max(v[1] for v in views)),
            "multi_scene": {"scenes": nsc, "workload": "heterogeneous-medium scene with 8 different 64^3 grids (seeds 20..27), "
                                                       "one 128^2 view at 16 spp each",
                            "one_call_per_scene_ms": t_ms_seq * 1e3, "one_call_ms": t_ms_tog * 1e3,
                            "speedup": t_ms_seq / t_ms_tog},
            "samples": samples, "batched_s": t_batch, "one_call_per_view_s": t_seq,
            "batched_samples_per_s": samples / t_batch, "batching_speedup": t_seq / t_batch,
            "small_batch": {"views": len(small), "res": 128, "spp": 16, "eager_ms": t_eager * 1e3,
                            "graph_replay_ms": t_graph * 1e3, "graph_speedup": t_eager / t_graph,
                            "replays_that_fell_back": int(fallbacks)}}


def primal_leg(res=512, spp=128, reps=2):
    """VolPathTracer and VolPathTracerSimple (primal target renders) on the same
    heterogeneous-medium scene and size as the main workload."""
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    d = scenes.cube_medium(nder=3, res=res, spp=spp)
    sc = scenes.instantiate(G, d)
    out = torch.zeros((d.height, d.width, 3), device="cuda")
    rec = {"workload": "primal render of the heterogeneous-medium scene, %d^2, %d spp, depth %d" % (res, spp, d.max_bounces)}
    for kind, name in ((1, "VolPathTracer"), (2, "VolPathTracerSimple")):
        opts = G.RenderOptions(7, spp, d.max_bounces, 0, 0, True)
        G.render_device(kind, sc, opts, out.data_ptr())   # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(reps):
            opts.seed = 8 + k
            G.render_device(kind, sc, opts, out.data_ptr())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        rec[name] = {"s_per_render": dt, "samples_per_s": res * res * spp / dt}
    return rec


def glass_leg(iters=4):
    """One optimisation step of the reference's glass_1 example (examples/glass_1/
    optimize.py: 256^2, spp 1024, sppe 32768, sppse 10, 4 bounces, 4 variables, 10-level
    pyramid loss, Adam): AD render + loss + backward + parameter update, averaged."""
    import numpy as np
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import loaders
    from dtrt_amd.harness import ADLossFunc, SceneManager, SceneTransform
    d, integ = loaders.load_mitsuba(os.path.join(ROOT, "tests", "golden", "scenes", "glass_1", "scene.xml"), nder=4)
    opts = G.RenderOptions(0, 1024, 4, 32768, 10, True)
    sm = SceneManager(d, [[SceneTransform("SHAPE_TRANSLATE", [1.0, 0, 0], 0)],
                          [SceneTransform("SHAPE_TRANSLATE", [0, 1.0, 0], 0)],
                          [SceneTransform("SHAPE_TRANSLATE", [0, 0, 1.0], 0)],
                          [SceneTransform("BSDF_VARY", "roughdielectric", 0, 0.1, 0.0)]])
    sm.set_arguments([-5.0, -3.0, -4.0, -2.5])
    tgt = torch.zeros(d.out_shape(), device="cuda")
    G.render_device(loaders.integrator_for(G, integ).kind, sm.scene(G), opts, tgt.data_ptr())
    target = tgt.view(-1)[: d.height * d.width * 3].view(d.height, d.width, 3).clone()
    sm.reset()
    var = torch.zeros(4, requires_grad=True)
    optim = torch.optim.Adam([var], lr=2e-1)
    oor = torch.zeros(4)
    times, losses = [], []
    for t in range(iters + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        optim.zero_grad()
        opts.seed = t + 1
        img = ADLossFunc.apply(sm, G.VolPathTracerAD(), opts, var, oor, torch.tensor([10000.0] * 4), 10)
        loss = (img - target).pow(2).sum()
        loss.backward()
        optim.step()
        oor = torch.as_tensor(sm.set_arguments(var.detach().numpy()))
        torch.cuda.synchronize()
        if t > 0:   # the first step also builds the scene's device buffers
            times.append(time.perf_counter() - t0)
        losses.append(loss.item())
    return {"workload": "examples/glass_1 optimisation step at the script's settings (256^2, spp 1024, sppe 32768, "
                        "sppse 10, nder 4)", "s_per_step": float(np.mean(times)), "steps_timed": iters,
            "loss_first_steps": losses, "reference_loss_first_steps": [820.2951, 721.77826, 631.98737, 548.87231, 474.86902]}


def configs_leg():
    """The other forward-mode scenes at their full sizes, one timed render each after a
    warm-up: the surface scene (ground + icosphere, no media, 256^2, 64 spp, depth 5,
    nder 1) and the dielectric scene (100,820-face displaced rough-dielectric sphere in a
    128^3 fBm medium, 512^2, 256 spp, depth 12, nder 8)."""
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    rec = {}
    for name, d in (("surface_icosphere_256sq_64spp_nder1", scenes.ico_surface()),
                    ("dielectric_sphere_in_medium_512sq_256spp_nder8", scenes.dielectric_medium())):
        sc = scenes.instantiate(G, d)
        out = torch.zeros(d.out_shape(), device="cuda")
        opts = G.RenderOptions(1, d.spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)
        G.render_device(0, sc, opts, out.data_ptr())
        torch.cuda.synchronize()
        opts.seed = 2
        out.zero_()
        t = time.perf_counter()
        G.render_device(0, sc, opts, out.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        n = d.width * d.height * d.spp
        st = G.render_stats(sc)
        rec[name] = {"width": d.width, "height": d.height, "spp": d.spp, "max_bounces": d.max_bounces, "nder": d.nder,
                     "faces": int(sum(len(s.indices) for s in d.shapes)), "s_per_render": dt,
                     "ad_samples_per_s": n / dt, "children": int(st["children"]), "pe_samples": int(st["pe_samples"])}
        del sc
    return rec


def inverse_leg(spp_divisor=16):
    """One batched inverse-rendering step on 32 orbit views (res 128^2..768^2, spp
    16..256 drawn with PCG32 seed 5, each divided by spp_divisor) of a 1,003,520-face
    displaced sphere in a 256^3 fBm medium, through dtrt_scene_step: the forward render
    of all views with 8 shape-deformation tangent lanes (all terms), the L2 loss against
    target images, its gradient for the 8 shape lanes, and the density adjoint over all
    16.8M grid nodes. Timed eagerly, then as one captured HIP graph replayed with new
    seeds; the forward pass is also timed batched against one call per view."""
    import numpy as np
    import torch
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes
    t0 = time.perf_counter()
    d = scenes.displaced_views()
    views = scenes.orbit_views(32, seed=5, spp_choices=(16, 32, 64, 128, 256), radius=3.2)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    sc = scenes.instantiate(G, d)
    cams = [G.Camera(w, h, c2w.astype(np.float32), d.cam_to_ndc.astype(np.float32), d.clip_near, -1, None)
            for c2w, w, h, _ in views]
    spps = [max(1, spp // spp_divisor) for _, _, _, spp in views]
    fo = [G.RenderOptions(500 + i, spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist) for i, spp in enumerate(spps)]
    ao = [G.RenderOptions(900 + i, spp, d.max_bounces, 0, 0, True, 1 | 2 | 8, 0.0) for i, spp in enumerate(spps)]
    ims = [torch.zeros((d.nder + 1, h, w, 3), device="cuda") for _, w, h, _ in views]
    rng = np.random.default_rng(6)
    tgts = [torch.tensor(rng.random((h, w, 3)) * 0.2, dtype=torch.float32, device="cuda") for _, w, h, _ in views]
    grad = torch.zeros(d.media[0].density.shape, dtype=torch.float32, device="cuda")
    pgrad = torch.zeros(d.nder, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    samples = sum(w * h * spp for (_, w, h, _), spp in zip(views, spps))
    G.render_views(sc, cams[:1], fo[:1], [torch.zeros_like(ims[0])])   # device buffers, primary edges of view 0
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0

    def clock(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t, r

    t_capture, _ = clock(lambda: G.capture_step(sc, cams, fo, ims, tgts, 0, ao, grad, pgrad, loss))
    t_graph, rc = clock(lambda: G.replay_step(sc, [700 + i for i in range(2 * len(views))]))
    t_eager, _ = clock(lambda: G.step(sc, cams, fo, ims, tgts, 0, ao, grad, pgrad, loss))
    t_batch, _ = clock(lambda: G.render_views(sc, cams, fo, ims))

    def singles():
        for i in range(len(views)):
            G.render_views(sc, [cams[i]], [fo[i]], [ims[i]])
    t_single, _ = clock(singles)
    # a shape step: new vertex positions for the 1M-face sphere, BVH refitted on the GPU
    v_new = d.shapes[0].vertices + 1e-3 * d.shapes[0].velocities[0]
    t_update, _ = clock(lambda: G.update_shape(sc, 0, v_new))
    return {"workload": "batched inverse-rendering step: 32 orbit views (res 128^2..768^2, spp 16..256 / %d, PCG32 seed 5) "
                        "of a 1,003,520-face displaced sphere in a 256^3 fBm medium; forward nder 8 (shape lanes), all "
                        "terms; L2 loss vs target images; gradient of the 8 lanes; density adjoint over 16,777,216 nodes"
                        % spp_divisor,
            "spp_divisor": spp_divisor, "samples": samples, "scene_generation_s": t_gen, "setup_s": t_setup,
            "vertex_update_s": t_update,
            "eager_step_s": t_eager, "graph_step_s": t_graph, "graph_fell_back_to_eager": bool(rc == 1),
            "graph_speedup": t_eager / t_graph, "capture_s_including_one_eager_step": t_capture,
            "step_samples_per_s": samples / t_graph, "forward_batched_s": t_batch, "forward_one_call_per_view_s": t_single,
            "batching_speedup": t_single / t_batch, "loss": float(loss.item()),
            "shape_lane_grad": [float(x) for x in pgrad.cpu()], "grad_nonzero_nodes": int((grad != 0).sum().item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--spp", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the float-precision leg of the headline workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-vs-forward gradient timing")
    ap.add_argument("--no-views", action="store_true", help="skip the batched multi-view timing")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch settings at the scene's first use (the default keeps the built-in ones, "
                         "which the profiles under profiles/ were taken with)")
    ap.add_argument("--no-primal", action="store_true", help="skip the primal-integrator timing")
    ap.add_argument("--no-examples", action="store_true", help="skip the glass_1 optimisation-step timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the surface / dielectric scene timings")
    ap.add_argument("--inverse-spp-divisor", type=int, default=16,
                    help="divide the batched inverse step's per-view spp by this (0: skip the leg)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the renderer runs on a single device (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.calibrate_copy:
        n = args.calibrate_copy << 18
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return
    import dtrt_amd.dtrt as G
    from dtrt_amd import scenes

    G.set_precision(64)   # the reference's Float (include/fwd.h:10-13); every leg below runs at it
    desc = scenes.cube_medium(nder=3, res=args.res, spp=args.spp)
    scene = scenes.instantiate(G, desc)
    out = torch.zeros(desc.out_shape(), dtype=torch.float32, device="cuda")
    opts = G.RenderOptions(0, desc.spp, desc.max_bounces, desc.sppe, desc.sppse, True, -1, desc.ddist)

    def step(seed):
        out.zero_()
        opts.seed = seed
        G.render_device(0, scene, opts, out.data_ptr())

    tune = None
    if args.autotune:   # the scene's first use: pick launch settings, then time with them
        tms, tset = G.autotune(scene, G.RenderOptions(999, desc.spp, desc.max_bounces, desc.sppe, desc.sppse, True, -1,
                                                      desc.ddist))
        tune = dict(tset, best_render_ms=tms)
    for w in range(args.warmup):
        step(1000 + w)
    torch.cuda.synchronize()
    ms = np.zeros(4)
    launches = np.zeros(4, np.int64)
    work = np.zeros(12, np.uint64)
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(s)
        m, l, w = G.kernel_times(scene)
        ms += m
        launches += l
        work += w
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = G.render_stats(scene)
    npix = desc.width * desc.height
    samples = npix * desc.spp * args.steps
    value = samples / elapsed
    alg = algorithmic_bytes(work, desc.nder, desc.spp, npix * args.steps, scalar_bytes=8)
    dom = int(np.argmax(ms))
    name = KERNELS[dom]
    achieved = alg[name] / (ms[dom] * 1e-3) / 1e9
    # HBM bytes per launch of the dominant kernel, as measured: the copy-calibrated
    # rocprofv3 FETCH_SIZE and WRITE_SIZE counters of this command (bench.py, this
    # workload and precision) averaged over the kernel's launches of the profiled run's
    # timed render. Null when no profile of this workload, precision and kernel exists.
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", "ad_render_heterogeneous_512x512_128spp_nder3_f64_hbm_traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        entry = t.get("kernels", {}).get("%s<%d>" % (name, desc.nder))
        same = (t["width"], t["height"], t["spp"], t["nder"], t["precision"]) == (desc.width, desc.height, desc.spp,
                                                                                 desc.nder, 64)
        if same and entry and not args.autotune and entry["launches"] * args.steps == int(launches[dom]):
            traffic = entry["hbm_bytes_per_launch"]
            traffic_src = "measured: %s (%d launches, %s)" % (os.path.basename(tfile), entry["launches"], t["command"])
    rec = {
        "metric": "AD path samples/s at 1 GPU; adjoint vs forward grad time at equal variance",
        "value": value,
        "unit": "AD path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded fBm density grid, generated meshes)",
        "config": {"workload": "heterogeneous-medium AD render: null cube + %d^3 fBm grid, HG 0.3, quad light + "
                               "emissive env sphere" % 64, "width": desc.width, "height": desc.height,
                   "spp": desc.spp, "max_bounces": desc.max_bounces, "nder": desc.nder, "sppse": desc.sppse,
                   "sppe": desc.sppe, "primary_edge_samples_per_step": int(st["pe_samples"])},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "avg_launch_ms": ms[dom] / max(1, launches[dom]), "launches_timed": int(launches[dom]),
                     "algorithmic_bytes_per_launch": alg[name] / max(1, launches[dom])},
        "launches_per_step": {k: int(launches[i]) // args.steps for i, k in enumerate(KERNELS)},
        "algorithmic_bytes_per_step": {k: float(alg[k]) / args.steps for k in KERNELS if k in alg},
        "kernel_ms_per_step": {k: ms[i] / args.steps for i, k in enumerate(KERNELS)},
        "work_per_step": {k: [int(x) // args.steps for x in work[b:b + 4]]
                          for k, b in (("ad_paths", 0), ("children", 4), ("primary_edges", 8))},
        "render_stats": {k: int(v) for k, v in st.items()},
        "autotune": tune,
    }
    del scene
    if not args.no_fp32:
        rec["fp32_leg"] = precision_leg(desc, 32, args.steps)
    if not args.no_views:
        rec["views"] = views_leg()
    if not args.no_adjoint:
        rec["adjoint"] = adjoint_leg()
    if not args.no_primal:
        rec["primal"] = primal_leg(args.res, args.spp)
    if not args.no_examples:
        rec["glass_1"] = glass_leg()
    if not args.no_configs:
        rec["other_scenes"] = configs_leg()
    if args.inverse_spp_divisor > 0:
        rec["inverse_step"] = inverse_leg(args.inverse_spp_divisor)
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(desc)
        crop = rec["cpu_baseline"]["crop"]
        rec["gpu_same_crop"] = {"crop": crop, "value": gpu_crop_rate(desc, crop), "unit": "AD path samples/s",
                                "dtype": "f64"}
        rec["gpu_over_cpu"] = rec["gpu_same_crop"]["value"] / rec["cpu_baseline"]["value"]
        rec["gpu_full_frame_over_cpu_crop"] = value / rec["cpu_baseline"]["value"]
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
