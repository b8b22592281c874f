"""Throughput benchmark of the gfx950 path tracer on the 1080p spectral workload.

Workload: a geodesic icosphere at frequency 229 (1,048,820 triangles) displaced
by seeded 3-octave value noise, split into 64 material IDs (GGX roughness
U(0.05,1), sRGB albedo U(0.05,0.95), 20% dielectric IOR 1.5), lit by a
synthetic 2048x1024 RGBA32F sky with a 0.5 degree sun of radiance 1e4;
1920x1080, 64 samples per pixel per frame, path depth 12 (Russian roulette
from depth 4), hero-wavelength spectral mode, PCG32 streams with seed 1.

A step is one VKRT_draw: one full frame of 1920*1080*64 paths traced through
the wavefront pipeline and accumulated. The metric counts bounces: one
extension ray per path per depth iteration (the reference's depth loop,
src/shaders/integrator/path/spectral_hero/integrator.slang:25-27), counted on
the device with an atomic counter.

Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published peak

# Algorithmic bytes per bounce (one fixed formula for the whole pipeline):
#   B_ray   ray write 32 + ray read 32 + hit write 20 + hit read 20
#   B_hit   3 x u32 indices + 3 x 48 B ShaderVertex
#   B_state hero path state read + write
#   B_nee   per shadow ray: 2x32 ray + 2x4 result + 2x(4+4) alias + 48 EmissiveTriangle + 32 EmissiveMesh
#   B_trav  BVH nodes fetched x node bytes + triangle records fetched x record bytes (instrumented traversal)
#   B_accum 72 B per pixel per frame
B_RAY = 104
B_HIT = 156
B_STATE = 192
B_NEE = 170
B_ACCUM_PER_PIXEL = 72
NODE_BYTES = 96   # 8-wide node: origin + exponents 16 B, 8-bit child planes 48 B, 8 child references 32 B
TRI_BYTES = 48    # triangle record: 3 vertices x float3 + global triangle id
# The traversal kernel's own share of B_ray: ray read 32 B + hit write 20 B.
B_EXTEND_IO = 32 + 20

WIDTH, HEIGHT, MIN_DEPTH = 1920, 1080, 4
CPU_ROWS = 512  # centred band of full-width rows for the CPU baseline sample

# The headline workload (spectral) and the incoherent-traversal stress workload.
WORKLOADS = {
    "spectral": dict(spp=64, depth=12, mode="spectral",
                     name="displaced icosphere 1,048,820 triangles, 64 materials (20% dielectric IOR 1.5), "
                          "2048x1024 HDR sky + 0.5 deg sun 1e4; 1920x1080, 64 spp, depth 12, hero spectral, "
                          "PCG32 seed 1"),
    "stress": dict(spp=16, depth=16, mode="rgb",
                   name="16 rigidly transformed instances of the displaced icosphere + 4M random triangles in the "
                        "unit cube (20,781,120 triangles), all diffuse, 2048x1024 sky; 1920x1080, 16 spp, depth 16, "
                        "RGB, PCG32 seed 1"),
}


def traversal_digest() -> str:
    """SHA-256 of the code the closest-hit kernel and the tree it walks are
    built from (the traversal and common device headers, the BVH builder and
    kExtend's definition), recorded with the committed traffic counters so that
    a changed kernel or tree layout does not report them."""
    import hashlib

    h = hashlib.sha256()
    hip = os.path.join(ROOT, "vkrt_amd", "csrc", "hip")
    for name in ("dev_traverse.h", "dev_common.h", "bvh.hip"):
        with open(os.path.join(hip, name), "rb") as f:
            h.update(f.read())
    with open(os.path.join(hip, "render.hip")) as f:
        text = f.read()
    start = text.index("// Closest-hit traversal, persistent with lane refill")
    end = text.index("\n}\n", text.index("void kExtend(", start)) + 3
    h.update(text[start:end].encode())
    return h.hexdigest()


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing leg")
    p.add_argument("--cpu-rows", type=int, default=CPU_ROWS)
    p.add_argument("--workload", choices=sorted(WORKLOADS), default="spectral",
                   help="spectral: the headline 1080p spectral workload; stress: the 20.8M-triangle traversal test")
    p.add_argument("--prb-iters", type=int, default=200,
                   help="Adam iterations of the inverse-rendering variant (0: forward/backward timings only, "
                        "-1: no PRB variant)")
    p.add_argument("--nrc-frames", type=int, default=240,
                   help="orbit frames of the neural radiance cache variant (0 skips it)")
    p.add_argument("--no-stress", dest="stress", action="store_false",
                   help="skip the 20.8M-triangle stress variant of the default run")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run device-to-device copies of MIB mebibytes (reference kernel for HBM counters)")
    return p.parse_args()


def copy_calibration(mib: int):
    """Copies a buffer of `mib` MiB device-to-device 5 times with the HIP
    runtime's copy kernel, so that rocprofv3's FETCH_SIZE / WRITE_SIZE can be
    scaled: each dispatch reads and writes exactly mib MiB."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copied {mib} MiB device-to-device 5 times")


def build_workload(kind):
    import vkrt_amd as V
    from vkrt_amd import scenes

    w = WORKLOADS[kind]
    t0 = time.time()
    scene = scenes.displaced_sphere(229) if kind == "spectral" else scenes.stress_scene()
    scene.environment = scenes.sky_environment(2048, 1024)
    r = V.Renderer(WIDTH, HEIGHT)
    scenes.load(r, scene)
    mode = V.RENDER_MODE_SPECTRAL if w["mode"] == "spectral" else V.RENDER_MODE_RGB
    r.configure(spp=w["spp"], max_depth=w["depth"], min_depth=MIN_DEPTH, mode=mode, spectral=V.SPECTRAL_HERO,
                rng=V.RNG_PCG32, seed=1)
    r.start(WIDTH, HEIGHT, 0)
    log(f"{kind} scene ready in {time.time() - t0:.1f} s: {scene.triangle_count} triangles")
    return r, scene


def stress_variant(steps: int = 8, warmup: int = 2):
    """The 20,781,120-triangle incoherent-traversal configuration at 1080p,
    16 spp, depth 16: BVH build time, Gbounces/s and the traversal kernel's
    fraction of the HBM roof on the same algorithmic-bytes model."""
    r, scene = build_workload("stress")
    for _ in range(warmup):
        r.draw()
    st = r.stats()
    build_ms = st.bvhBuildMs
    r.synchronize()
    t0 = time.perf_counter()
    bounces = ext_rays = ext_ms = launches = tail_rays = tail_ms = 0
    tail_depth = 0
    for _ in range(steps):
        r.draw()
        s = r.stats()
        bounces += s.extensionRays
        ext_rays += s.raysBeforeTail  # the rays kExtend traced (the fused tail traced the rest)
        ext_ms += s.extendMs
        launches += s.extendLaunches
        tail_rays += s.tailRays
        tail_ms += s.tailMs
        tail_depth = s.tailStartDepth
    r.synchronize()
    elapsed = time.perf_counter() - t0
    r.set_traversal_statistics(True)
    r.draw()
    s = r.stats()
    r.set_traversal_statistics(False)
    # per kExtend ray: the depths before the one the fused tail starts at
    nodes = s.nodeVisitsBeforeTail / max(1, s.raysBeforeTail)
    tris = s.triangleTestsBeforeTail / max(1, s.raysBeforeTail)
    per_ray = B_EXTEND_IO + nodes * NODE_BYTES + tris * TRI_BYTES
    achieved = ext_rays * per_ray / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    out = {
        "workload": WORKLOADS["stress"]["name"],
        "triangles": scene.triangle_count,
        "value": round(bounces / elapsed / 1e9, 4),
        "unit": "Gbounces/s",
        "rays_per_s": round(bounces / elapsed),
        "steps": steps,
        "ms_per_step": round(elapsed * 1e3 / steps, 3),
        "bounces_per_step": bounces // steps,
        "bvh": {"build_ms": round(build_ms, 2), "nodes": st.bvhNodeCount, "depth": st.bvhMaxDepth},
        "kExtend": {"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBPS, 4), "unit": "GB/s",
                    "avg_launch_ms": round(ext_ms / max(1, launches), 4), "launches": launches,
                    "bytes_per_ray": round(per_ray, 1), "nodes_per_ray": round(nodes, 3),
                    "tris_per_ray": round(tris, 3), "rays_per_step": ext_rays // steps},
        "tail": {"start_depth": tail_depth, "rays_per_step": tail_rays // steps,
                 "ms_per_step": round(tail_ms / steps, 3)},
    }
    log(f"stress: {scene.triangle_count} triangles, BVH build {build_ms:.1f} ms, {out['value']:.3f} Gbounces/s, "
        f"kExtend {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBPS:.3f} of peak), {nodes:.2f} nodes/ray, "
        f"{tris:.2f} triangles/ray")
    r.close()
    return out


def cpu_baseline(renderer, rows: int, spp: int, depth: int):
    """Times the CPU oracle (a C restatement of the reference integrator) on a
    centred band of `rows` full-width rows of the same frame."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    # every core this process may run on, within the host share the job is given
    # (OMP_NUM_THREADS, 16 per GPU on a one-GPU box)
    visible = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(visible, share) if share > 0 else visible)
    buffers = renderer.export()
    t0 = time.time()
    osc = O.OracleScene(buffers)
    build_s = time.time() - t0
    r0 = HEIGHT // 2 - rows // 2
    accum = np.zeros((HEIGHT, WIDTH, 4), dtype=np.float32)
    t0 = time.time()
    _, bounces = osc.render(WIDTH, HEIGHT, accum=accum, rows=(r0, r0 + rows), threads=threads)
    dt = time.time() - t0
    osc.close()
    log(f"cpu oracle: {bounces} bounces in {dt:.2f} s on {threads} threads (BVH build {build_s:.1f} s, excluded)")
    return {
        "value": bounces / dt / 1e9,
        "unit": "Gbounces/s",
        "cores": threads,
        "kind": "port",
        "sample": f"rows {r0}-{r0 + rows - 1} of the 1920x1080 frame, all {WIDTH} columns, {spp} spp, depth "
                  f"{depth}, same scene, mode and seeds; FP32 C oracle with OpenMP on {threads} threads "
                  f"({visible} cores in the affinity mask, host share {share or 'unset'}); "
                  f"{bounces} bounces in {dt:.2f} s; oracle BVH build ({build_s:.1f} s) excluded",
    }


PRB_LR = {"albedo": 0.01, "roughness": 0.01, "emission": 0.01, "environment": 0.002}


def prb_variant(iters: int, lr=None):
    """Differentiable configuration: the Cornell geometry with 64 materials (one
    per face patch) and a 512x256 environment, 512x512, 32 spp forward and 32 spp
    backward from 3 seeded views. Times VKRT_draw and VKRT_renderBackward per view
    (wall clock around the C calls, device synchronised), then runs `iters` Adam
    iterations against targets rendered with parameters perturbed by +-20 %."""
    import torch

    import vkrt_amd as V
    from vkrt_amd import autograd as A
    from vkrt_amd import scenes

    size, spp = 512, 32
    r = V.Renderer(size, size)
    scene = scenes.differentiable_cornell()
    scenes.load(r, scene)
    r.configure(spp=spp, max_depth=8, min_depth=MIN_DEPTH, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=1)
    views = scenes.prb_views()
    env_shape = scene.environment.shape[:2]
    d_image = np.random.default_rng(0).normal(size=(size, size, 3)).astype(np.float32)
    fwd, bwd, bwd_dev = [], [], []
    for rep in range(3):
        for v in views:
            r.set_camera(*v)
            r.start(size, size, 0)
            r.draw()  # the first frame after a camera change also re-packs the scene
            r.start(size, size, 0)
            r.synchronize()
            t0 = time.perf_counter()
            r.draw()
            r.synchronize()
            t1 = time.perf_counter()
            r.start(size, size, 0)
            r.synchronize()
            t2 = time.perf_counter()
            r.render_backward(d_image, environment_shape=env_shape)
            t3 = time.perf_counter()
            if rep > 0:
                fwd.append((t1 - t0) * 1e3)
                bwd.append((t3 - t2) * 1e3)
                bwd_dev.append(r.stats().backwardMs)
    fwd_ms, bwd_ms = float(np.mean(fwd)), float(np.mean(bwd))
    log(f"PRB: forward {fwd_ms:.2f} ms, backward {bwd_ms:.2f} ms per view (512x512, 32 spp) -> {bwd_ms / fwd_ms:.2f}x")
    out = {"fwd_ms": round(fwd_ms, 3), "bwd_ms": round(bwd_ms, 3), "bwd_device_ms": round(float(np.mean(bwd_dev)), 3),
           "bwd_over_fwd": round(bwd_ms / fwd_ms, 3), "width": size, "height": size, "spp": spp, "views": len(views),
           "parameters": 64 * 7 + env_shape[0] * env_shape[1] * 3,
           "scene": "Cornell geometry, 68 face patches + light, 64 materials, 512x256 sky; RGB, depth 8, PCG32 seed 1"}
    if iters > 0:
        start = A.SceneParameters.from_renderer(r, scene.environment)
        rng = np.random.default_rng(4)
        target = A.SceneParameters(*[t * torch.tensor(rng.uniform(0.8, 1.2, t.shape), dtype=torch.float32)
                                     for t in start.tensors()])
        target.clamp_()
        held_out = scenes.prb_views(seed=1, count=1)[0]  # a fourth view, not optimised on
        with torch.no_grad():
            ref = [A.render(r, target, v) for v in views]
            ref_held = A.render(r, target, held_out)
        p = A.SceneParameters(*[t.clone().requires_grad_(True) for t in start.tensors()])

        def errors(q, mask):
            """Mean absolute error of each parameter kind (seen materials only)."""
            with torch.no_grad():
                return {"albedo": float((q.base_color - target.base_color).abs()[mask].mean()),
                        "roughness": float((q.roughness - target.roughness).abs()[mask].mean()),
                        "emission": float((q.emission - target.emission).abs()[mask].mean()),
                        "environment": float((q.environment - target.environment).abs().mean())}

        def held_loss():
            with torch.no_grad():
                return float(((A.render(r, p, held_out) - ref_held) ** 2).mean())

        held_first = held_loss()
        # step sizes per parameter kind: albedo and roughness live in [0, 1]; emission
        # of non-emitting patches starts at 0 and, like the sky texels, is only weakly
        # observed (Adam moves every parameter by about its step size per iteration)
        lr = dict(PRB_LR, **(lr or {}))
        opt = torch.optim.Adam([{"params": [p.base_color], "lr": lr["albedo"]},
                                {"params": [p.roughness], "lr": lr["roughness"]},
                                {"params": [p.emission], "lr": lr["emission"]},
                                {"params": [p.environment], "lr": lr["environment"]}])
        losses = []
        seen = None
        t0 = time.perf_counter()
        for it in range(iters):
            opt.zero_grad()
            loss = sum(((A.render(r, p, v) - t) ** 2).mean() for v, t in zip(views, ref))
            loss.backward()
            if it == 0:  # materials the views see: those with a non-negligible albedo gradient
                gm = p.base_color.grad.abs().amax(dim=1)
                seen = gm > 1e-3 * float(gm.max())
                err0 = float((p.base_color - target.base_color).detach().abs()[seen].mean())
                all0 = errors(p, seen)
            opt.step()
            p.clamp_()
            losses.append(float(loss.detach()))
        el = time.perf_counter() - t0
        err1 = float((p.base_color - target.base_color).detach().abs()[seen].mean())
        held_last = held_loss()
        log(f"PRB: {iters} Adam iterations in {el:.1f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e}, held-out view "
            f"{held_first:.3e} -> {held_last:.3e}, albedo error of {int(seen.sum())} seen materials "
            f"{err0:.4f} -> {err1:.4f}")
        out["optimisation"] = {"iterations": iters, "seconds": round(el, 2), "learning_rates": lr,
                               "loss_first": losses[0],
                               "loss_last": losses[-1], "held_out_view_loss_first": held_first,
                               "held_out_view_loss_last": held_last, "seen_materials": int(seen.sum()),
                               "albedo_error_first": round(err0, 5), "albedo_error_last": round(err1, 5),
                               "errors_first": all0, "errors_last": errors(p, seen),
                               "parameters_optimised": "albedo, roughness, emission of 64 materials + 512x256 sky"}
        # The same problem with only the albedos free and every other parameter at
        # its target: if the adjoint's albedo gradient is right, the albedo error
        # falls; the all-parameter run's albedo drift then comes from trading
        # albedo against emission, roughness and the sky, which the three views do
        # not tell apart.
        q = A.SceneParameters(start.base_color.clone().requires_grad_(True), target.roughness.clone(),
                              target.emission.clone(), target.environment.clone())
        opt = torch.optim.Adam([q.base_color], lr=lr["albedo"])
        qe0 = float((q.base_color - target.base_color).detach().abs()[seen].mean())
        qloss = []
        t0 = time.perf_counter()
        for it in range(iters):
            opt.zero_grad()
            loss = sum(((A.render(r, q, v) - t) ** 2).mean() for v, t in zip(views, ref))
            loss.backward()
            opt.step()
            q.clamp_()
            qloss.append(float(loss.detach()))
        qe1 = float((q.base_color - target.base_color).detach().abs()[seen].mean())
        log(f"PRB albedo only: {iters} iterations in {time.perf_counter() - t0:.1f} s, loss {qloss[0]:.3e} -> "
            f"{qloss[-1]:.3e}, albedo error {qe0:.4f} -> {qe1:.4f}; all-parameter run errors "
            f"{out['optimisation']['errors_first']} -> {out['optimisation']['errors_last']}")
        out["optimisation_albedo_only"] = {"iterations": iters, "loss_first": qloss[0], "loss_last": qloss[-1],
                                           "albedo_error_first": round(qe0, 5), "albedo_error_last": round(qe1, 5),
                                           "fixed_at_target": "roughness, emission, sky"}
    r.close()
    return out


_NRC_REFERENCES = {}


def nrc_variant(frames: int, steps: int = 1, batch: int = 16384, prefix: int = 1, spectral: bool = False):
    """Neural radiance cache configuration: the Cornell box at 1920x1080, 1 spp per
    frame, depth 8, RGB (or hero spectral, the cache then in CIE XYZ). The cache
    trains online along `frames` frames of a slow orbit (12 degrees about the
    box's vertical axis) that ends at the default view, querying after `prefix`
    bounces; then, at that view, the plain tracer and the cached tracer (still
    training) each accumulate 1-spp frames, and relMSE = mean((x - ref)^2 /
    (ref^2 + 1e-2)) against a 2048-spp plain-tracer image of the view is recorded
    with wall time at 1, 2, 4, ..., 64 frames (readbacks excluded). time_ratio is
    the wall time the cached tracer needs to reach the plain tracer's 64-frame
    relMSE, divided by the plain tracer's time for it (null if it never does).
    The cache takes `steps` Adam steps of `batch` records per frame from the
    65,536 training paths' records."""
    import math

    import vkrt_amd as V
    from vkrt_amd import scenes

    w, h = 1920, 1080
    radius = 1.85

    def pose(deg):
        th = math.radians(deg)
        return ((0.5 + radius * math.sin(th), 0.5 - radius * math.cos(th), 0.5), (0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 40.0)

    settings = []

    mode = V.RENDER_MODE_SPECTRAL if spectral else V.RENDER_MODE_RGB

    def make(spp, cache):
        r = V.Renderer(w, h)
        scenes.load(r, scenes.cornell_box())
        r.configure(spp=spp, max_depth=8, min_depth=MIN_DEPTH, mode=mode, spectral=V.SPECTRAL_HERO,
                    rng=V.RNG_PCG32, seed=3)
        if cache:
            r.nrc_reset(1)
            settings.append(r.set_neural_cache(True, prefixBounces=prefix, trainingSteps=steps, batchSize=batch))
        return r

    def relmse(x, ref):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    ref_spp, ref_draws = 256, 8
    if mode not in _NRC_REFERENCES:  # one plain reference image per render mode
        rr = make(ref_spp, False)
        rr.set_camera(*pose(0.0))
        rr.start(w, h, 0)
        for _ in range(ref_draws):
            rr.draw()
        _NRC_REFERENCES[mode] = rr.accumulation()[..., :3].astype(np.float64)
        rr.close()
    ref = _NRC_REFERENCES[mode]

    r = make(1, True)
    r.synchronize()
    t0 = time.perf_counter()
    for deg in np.linspace(-12.0, 0.0, frames):
        r.set_camera(*pose(float(deg)))
        r.start(w, h, 0)
        r.draw()
    r.synchronize()
    orbit_ms = (time.perf_counter() - t0) * 1e3 / frames
    st = r.neural_cache_stats()
    log(f"NRC ({'hero' if spectral else 'rgb'}, prefix {prefix}, {steps} x {batch}): {frames} orbit frames at {orbit_ms:.2f} ms/frame (inference {st.inferMs:.3f} ms, training "
        f"{st.trainMs:.3f} ms, {st.queries} queries, {st.trainingRecords} records, loss {st.lastLoss:.3f})")

    def converge(rend, n=64):
        rend.set_camera(*pose(0.0))
        rend.start(w, h, 0)
        marks, elapsed = [], 0.0
        for i in range(1, n + 1):
            rend.synchronize()
            t = time.perf_counter()
            rend.draw()
            rend.synchronize()
            elapsed += time.perf_counter() - t
            if i & (i - 1) == 0:
                marks.append((i, elapsed * 1e3, relmse(rend.accumulation()[..., :3].astype(np.float64), ref)))
        return marks

    nrc_marks = converge(r)
    r.close()
    plain = make(1, False)
    plain_marks = converge(plain)
    plain.close()
    target_err, target_ms = plain_marks[-1][2], plain_marks[-1][1]
    ratio = None
    prev = None
    for frames_i, ms, err in nrc_marks:
        if err <= target_err:
            if prev is None:
                ratio = ms / target_ms
            else:  # interpolate the time in log error between the two checkpoints
                (pms, perr) = prev
                f = (math.log(perr) - math.log(target_err)) / max(math.log(perr) - math.log(err), 1e-12)
                ratio = (pms + f * (ms - pms)) / target_ms
            break
        prev = (ms, err)
    for (f, pms, perr), (_, nms, nerr) in zip(plain_marks, nrc_marks):
        log(f"NRC equal time: {f:3d} frames plain {pms:8.1f} ms relMSE {perr:.4e} | cached {nms:8.1f} ms relMSE {nerr:.4e}")
    log(f"NRC ({'hero' if spectral else 'rgb'}, prefix {prefix}, {steps} x {batch}): time to reach the plain tracer's "
        f"64-frame relMSE, cached / plain = {ratio}")
    cs = settings[0]
    return {"width": w, "height": h, "spp_per_frame": 1, "orbit_frames": frames, "orbit_ms_per_frame": round(orbit_ms, 3),
            "reference_spp": ref_spp * ref_draws, "infer_ms": round(st.inferMs, 3), "train_ms": round(st.trainMs, 3),
            "queries_per_frame": st.queries, "training_records_per_frame": st.trainingRecords,
            "checkpoints": [{"frames": f, "plain_ms": round(pms, 2), "plain_relmse": perr, "cached_ms": round(nms, 2),
                             "cached_relmse": nerr}
                            for (f, pms, perr), (_, nms, nerr) in zip(plain_marks, nrc_marks)],
            "time_ratio": None if ratio is None else round(ratio, 3),
            "scene": f"Cornell box (36 triangles), depth 8, {'hero spectral (cache in CIE XYZ)' if spectral else 'RGB'}, "
                     f"PCG32 seed 3; cache after {cs.prefixBounces} bounce(s), {cs.trainingPaths} training paths and "
                     f"{steps} Adam step(s) of {batch} records at learning rate {cs.learningRate:g} per frame"}


def main():
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the path shards by sample range with one all-reduce of the accumulation buffer per frame, "
              "which this build does not run.", file=sys.stderr, flush=True)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    wl = WORKLOADS[args.workload]
    SPP, MAX_DEPTH = wl["spp"], wl["depth"]
    r, scene = build_workload(args.workload)
    log(f"warmup {args.warmup} frame(s) (the first uploads the scene and builds the BVH)")
    for _ in range(args.warmup):
        r.draw()
    st = r.stats()
    bvh = {"build_ms": round(st.bvhBuildMs, 2), "nodes": st.bvhNodeCount, "depth": st.bvhMaxDepth,
           "triangles": st.triangleCount}
    log(f"BVH build {st.bvhBuildMs:.1f} ms, {st.bvhNodeCount} nodes, depth {st.bvhMaxDepth}")

    r.synchronize()
    t0 = time.perf_counter()
    bounces = shadow = paths = 0
    ext_ms = shade_ms = ext0_ms = shade0_ms = 0.0
    ext_launches = 0
    ext_rays = tail_rays = 0
    tail_ms = 0.0
    tail_depth = 0
    for i in range(args.steps):
        r.draw()
        s = r.stats()
        bounces += s.extensionRays
        shadow += s.shadowRays
        paths += s.paths
        ext_ms += s.extendMs
        shade_ms += s.shadeMs
        ext0_ms += s.extendMsDepth0
        shade0_ms += s.shadeMsDepth0
        ext_launches += s.extendLaunches
        ext_rays += s.raysBeforeTail  # the rays kExtend traced (the fused tail traced the rest)
        tail_rays += s.tailRays
        tail_ms += s.tailMs
        tail_depth = s.tailStartDepth
    r.synchronize()
    elapsed = time.perf_counter() - t0
    value = bounces / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    log(f"timed {args.steps} frames: {elapsed:.3f} s, {bounces} bounces, {shadow} shadow rays -> {value:.3f} Gbounces/s")

    acc = r.accumulation()
    expected = SPP * (args.warmup + args.steps)
    if not np.isfinite(acc).all() or int(acc[HEIGHT // 2, WIDTH // 2, 3]) != expected:
        log(f"accumulation check failed: finite={np.isfinite(acc).all()} samples={acc[HEIGHT // 2, WIDTH // 2, 3]}")
        sys.exit(1)

    # One more frame with the instrumented traversal for nodes and triangles per ray.
    r.set_traversal_statistics(True)
    r.draw()
    s = r.stats()
    r.set_traversal_statistics(False)
    # every ray (the whole-pipeline model) and kExtend's rays (the depths before the fused tail)
    b_trav_all = (s.nodeVisits * NODE_BYTES + s.triangleTests * TRI_BYTES) / max(1, s.extensionRays)
    nodes_per_ray = s.nodeVisitsBeforeTail / max(1, s.raysBeforeTail)
    tris_per_ray = s.triangleTestsBeforeTail / max(1, s.raysBeforeTail)
    b_trav = nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES
    log(f"traversal (kExtend rays): {nodes_per_ray:.2f} nodes/ray, {tris_per_ray:.2f} triangles/ray -> {b_trav:.0f} B/ray;"
        f" all rays {b_trav_all:.0f} B/ray")

    # Dominant kernel: the closest-hit traversal (kExtend), timed with HIP events on its stream.
    ext_bytes = ext_rays * (B_EXTEND_IO + b_trav)
    ext_avg_ms = ext_ms / max(1, ext_launches)
    achieved = ext_bytes / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    # Memory-fabric bytes per launch from the committed FETCH_SIZE / WRITE_SIZE
    # passes (copy-calibrated), per ray, times this run's rays per launch.
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "pmc_traffic_1080p_spectral.json")
    if args.workload == "spectral" and os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f)
        if pmc.get("kernel_sources_sha256") == traversal_digest():
            traffic = round(pmc["kExtend"]["hbm_bytes_per_ray"] * ext_rays / max(1, ext_launches))
        else:
            log("the committed traffic counters were measured on another kExtend build: traffic not reported")
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4),
        "traffic": traffic,
        "kernel": "kExtend<false, false, false> (closest-hit traversal, 8-wide quantised BVH)",
        "launches": ext_launches,
        "rays_per_launch": ext_rays // max(1, ext_launches),
        "avg_launch_ms": round(ext_avg_ms, 4),
        "bytes_per_ray": round(B_EXTEND_IO + b_trav, 1),
        "nodes_per_ray": round(nodes_per_ray, 3),
        "tris_per_ray": round(tris_per_ray, 3),
    }
    # Whole-pipeline algorithmic bytes per bounce over wall time.
    bounces_per_px = bounces / (WIDTH * HEIGHT * args.steps)
    b_bounce = (B_RAY + B_HIT + B_STATE + b_trav_all + B_NEE * shadow / max(1, bounces)
                + B_ACCUM_PER_PIXEL / max(bounces_per_px, 1e-9))
    pipeline = {
        "bytes_per_bounce": round(b_bounce, 1),
        "achieved": round(b_bounce * bounces / elapsed / 1e9, 1),
        "frac": round(b_bounce * bounces / elapsed / 1e9 / HBM_PEAK_GBPS, 4),
        "unit": "GB/s",
        "extend_share_of_time": round(ext_ms * 1e-3 / elapsed, 3),
    }
    # Every path traces one primary ray (depth 0); on this view most of them see
    # only the sky. The split shows what the deeper, incoherent bounces cost.
    depth_split = {
        "depth0": {"bounces_per_step": paths // max(1, args.steps),
                   "extend_ms_per_step": round(ext0_ms / max(1, args.steps), 3),
                   "shade_ms_per_step": round(shade0_ms / max(1, args.steps), 3)},
        "depth_ge1": {"bounces_per_step": (bounces - paths) // max(1, args.steps),
                      "extend_ms_per_step": round((ext_ms - ext0_ms) / max(1, args.steps), 3),
                      "shade_ms_per_step": round((shade_ms - shade0_ms) / max(1, args.steps), 3),
                      "tail_ms_per_step": round(tail_ms / max(1, args.steps), 3),
                      "gbounces_per_s_of_kernel_time": round((bounces - paths) / max(1e-9, (ext_ms - ext0_ms + shade_ms
                                                                                          - shade0_ms + tail_ms) * 1e-3)
                                                             / 1e9, 4)},
        "tail": {"start_depth": tail_depth, "rays_per_step": tail_rays // max(1, args.steps),
                 "ms_per_step": round(tail_ms / max(1, args.steps), 3)},
    }
    log(f"depth 0: {depth_split['depth0']}; depth >= 1: {depth_split['depth_ge1']}")

    # Variant: the same workload with environment NEE (an addition of this
    # build; the reference gathers the environment only on BSDF misses).
    variants = {}
    if args.steps > 0 and args.workload == "spectral":
        r.configure(env_importance=True)
        r.draw()  # first frame of the new setting
        r.synchronize()
        t0 = time.perf_counter()
        vb = vs = 0
        vsteps = min(args.steps, 3)
        for _ in range(vsteps):
            r.draw()
            st2 = r.stats()
            vb += st2.extensionRays
            vs += st2.shadowRays
        r.synchronize()
        vel = time.perf_counter() - t0
        r.configure(env_importance=False)
        variants["env_nee"] = {"value": round(vb / vel / 1e9, 4), "unit": "Gbounces/s", "steps": vsteps,
                               "ms_per_step": round(vel * 1e3 / vsteps, 3),
                               "shadow_rays_per_step": vs // vsteps}
        log(f"env NEE variant: {vb / vel / 1e9:.3f} Gbounces/s, {vs // vsteps} shadow rays per frame")

    if args.workload == "spectral" and args.stress and args.steps > 0:
        # before the long differentiable and neural-cache runs, whose
        # allocations and clock history would otherwise precede it
        variants["stress"] = stress_variant()
    if args.workload == "spectral" and args.prb_iters >= 0 and args.steps > 0:
        variants["prb"] = prb_variant(args.prb_iters)
    if args.workload == "spectral" and args.nrc_frames > 0 and args.steps > 0:
        # the configuration as specified: query after two bounces, 2^16 training
        # records per frame; then the same in hero spectral mode; then the
        # settings that measured fastest (one bounce, one step of 16384)
        variants["nrc"] = nrc_variant(args.nrc_frames, steps=1, batch=65536, prefix=2)
        variants["nrc_spectral"] = nrc_variant(args.nrc_frames, steps=1, batch=65536, prefix=2, spectral=True)
        variants["nrc_tuned"] = nrc_variant(args.nrc_frames)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(r, args.cpu_rows if args.workload == "spectral" else 64, SPP, MAX_DEPTH)
    r.close()

    line = {
        "metric": "Gbounces/s at 1080p spectral" if args.workload == "spectral" else "Gbounces/s at 1080p, 20.8M-triangle stress",
        "value": round(value, 4),
        "unit": "Gbounces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural displaced icosphere and HDR sky generated in-process)",
        "config": {
            "workload": wl["name"],
            "width": WIDTH,
            "height": HEIGHT,
            "spp": SPP,
            "max_depth": MAX_DEPTH,
            "triangles": scene.triangle_count,
            "paths_per_step": paths // max(1, args.steps),
            "bounces_per_step": bounces // max(1, args.steps),
            "shadow_rays_per_step": shadow // max(1, args.steps),
        },
        "roofline": roofline,
        "pipeline": pipeline,
        "depth_split": depth_split,
        "bvh": bvh,
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
