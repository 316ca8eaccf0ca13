"""Benchmark: Mrays/s of the path tracer on the 1.51 M-triangle 1080p workload.

Workload (config 1 of the synthetic generators): icosphere subdivided 8x
(1,310,720 triangles, radius 3, value-noise displacement 0.1) + 200,000 random
small triangles in [-5,5]^3, one quad light, 512x256 lognormal env map,
1920x1080, 1 spp, 8 bounces, binned-SAH BVH4 built on the host (not timed).

A step is one frame (1920x1080, 1 spp, up to 8 bounces, accumulated into the
running mean), issued as one dmt_render call: camera rays into the path
queue (k_generate), the frame-completion kernel (k_tail) that runs every path
to its end, the accumulation pass (k_accumulate).  The library keeps four
frames in flight on four HIP streams (the accumulation passes stay ordered
frame after frame, so the image equals the sequential result).  value =
(primary + extension + shadow rays) / wall time of the K frames, inputs
resident in HBM.

roofline: the kernel with the most device time per frame (k_tail on the
default schedule; k_extend, the closest-hit kernel, when the camera bounce
runs through the per-bounce kernels), HBM-bound by design (no dense
contraction).  Algorithmic bytes per launch = node visits x 128 B (BVH4 node)
+ triangle tests x 48 B (TriRec) + sphere tests x 20 B (+ rays x 40 B for
k_extend: 32 B ray record in, 8 B hit out), counted by an instrumented
re-render of the same frame; divided by the HIP-event duration of that
kernel's launches in frames rendered one at a time after the timed region (in
the timed region four frames share the GPU, and a launch's duration there
includes the kernels of the other frames).  traffic: the measured bytes per
launch beyond L2 (rocprofv3 FETCH_SIZE + WRITE_SIZE passes of this workload,
scaled by a 512 MiB copy), read from profiles/<kernel>_c1_1080p_pmc.txt.

cpu_baseline: the CPU oracle (reference semantics, fp32) rendering whole frames
of the same workload for about 10 s, on this host's cores (hw - 4 threads as
the reference's software tracer uses, and all cores).

--workload batch: the batched form.  64 scenes of the batched workload
generator (resolution uniform in {128, 256, 512, 1024} squared, triangles
log-uniform in 10k..2M, half with 64 spheres; 1 spp, depth 6).  The plain
form renders each frame as 64 sequential per-scene dmt_render calls, each to
completion; the batched form renders each frame of all 64 scenes with one
persistent-kernel launch replayed from a captured hipGraph, with parameters
picked by dmt_batch_autotune.  All run the same frames from frame 0; the
images must be identical.  value = plain time / batched time over the timed
frames.  vs_pipelined_plain compares with the same dmt_render calls issued
back to back, the library keeping up to four frames in flight.

--workload adjoint: the differentiable form.  Config 4: 256 spheres with 256
distinct materials over the ground plane, 512x512, 16 spp, depth 4, Russian
roulette off.  The target image is rendered with the generator's materials;
the spheres' parameters are then perturbed by N(0, 0.1) (albedo kept in
[0.05, 1], emission >= 0.02).  A forward step is one dmt_render of the frame
(to completion), a backward step one dmt_adjoint with dLoss/dI = 2 (I - target)
(to completion).  value = backward time / forward time (lower is better).

--workload features: the camera-sample features on the config-1 workload:
thin lens (aperture 0.05, focus 12, six blades, cat's-eye 0.3), motion blur
(the mesh translated from (-0.1, 0, 0) to its rest position over the
shutter), hero-wavelength dispersion (C = 0.0042 um^2).  Such frames render
through the feature variant of k_tail.  value = their Mrays/s over K
frames issued back to back; vs_plain = that rate over the plain frame's rate
(the same scene and camera without the features) measured the same way.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "demont-engine_amd", "python"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 1080p workload, from the
    recorded rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a copy kernel
    of known size; a profiler cannot run inside this process)."""
    try:
        for line in open(os.path.join(ROOT, "profiles", f"{kernel}_c1_1080p_pmc.txt")):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("c1", "c2", "batch", "adjoint", "features"), default="c1")
    ap.add_argument("--steps", type=int, default=None, help="timed frames (c1: 20, batch: 100)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed frames (c1: 3, batch: 2)")
    ap.add_argument("--scenes", type=int, default=64, help="batch workload: number of scenes")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--schedule", type=int, default=None,
                    help="per-bounce iterations before the frame-completion launch (-1: per-bounce kernels only; "
                         "default: the library's)")
    ap.add_argument("--roofline-frames", type=int, default=10,
                    help="frames rendered one at a time to time the closest-hit kernel alone")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import dmt

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    if args.workload == "batch":
        args.steps = 100 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        return bench_batch(args, np, torch, dmt)
    if args.workload == "adjoint":
        args.steps = 10 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        return bench_adjoint(args, np, torch, dmt)
    if args.workload == "features":
        args.steps = 10 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        return bench_features(args, np, torch, dmt)
    args.steps = 20 if args.steps is None else args.steps
    args.warmup = 3 if args.warmup is None else args.warmup
    if args.calibration_copy_mb > 0:
        # one plain device-to-device copy of known size, so that the HBM
        # counters of a profiler pass can be scaled to bytes
        n = args.calibration_copy_mb * (1 << 20) // 4
        src = torch.arange(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibration_copy_mb} MiB done")
        del src, dst
    cfg = 2 if args.workload == "c2" else 1
    t0 = time.time()
    sd = dmt.generate(cfg, seed=args.seed)
    log(f"[bench] scene: {sd.indices.shape[0]} triangles, {sd.frame.width}x{sd.frame.height}, "
        f"{sd.frame.max_bounces} bounces ({time.time() - t0:.1f}s)")
    dev = dmt.Device(0)
    t0 = time.time()
    scene = dmt.Scene(dev, sd)
    log(f"[bench] BVH build + upload {time.time() - t0:.1f}s")
    frame = sd.frame.copy()
    if args.schedule is not None:
        scene.set_schedule(dmt.SCHEDULE_WAVEFRONT if args.schedule < 0 else args.schedule)
    schedule = scene.schedule()

    def run_frames(k0, n):
        f = frame.copy()
        f.frame_index = k0
        f.reset_accum = 1 if k0 == 0 else 0
        scene.render_frames(f, n)

    def timed(fn):
        dev.synchronize()
        torch.cuda.synchronize()
        scene.reset_stats()
        scene.set_profiling(True)
        dev.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        dev.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = scene.stats()
        scene.set_profiling(False)
        return dt, st

    # K frames, one dmt_render call each (the library keeps four in flight)
    def per_frame():
        for k in range(args.steps):
            f = frame.copy()
            f.frame_index = args.warmup + k
            f.reset_accum = 0
            scene.render(f)

    run_frames(0, args.warmup)
    dt, st = timed(per_frame)
    rays = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
    mrays = rays / dt / 1e6
    log(f"[bench] {args.steps} frames in {dt * 1e3:.1f} ms: {mrays:.1f} Mrays/s "
        f"(primary {st['primary_rays']}, extension {st['extension_rays']}, shadow {st['shadow_rays']})")

    # Algorithmic bytes of the closest-hit kernel (k_extend) and of the
    # frame-completion kernel (k_tail), from an instrumented re-render.
    f = frame.copy()
    f.frame_index = args.warmup
    tr = scene.measure_traffic(f)
    ext_bytes = (tr["ext_node_visits"] * 128 + tr["ext_tri_tests"] * 48 + tr["ext_sphere_tests"] * 20
                 + tr["ext_rays"] * 40)
    tail_bytes = tr["tail_node_visits"] * 128 + tr["tail_tri_tests"] * 48 + tr["tail_sphere_tests"] * 20
    # In the timed region four frames share the GPU, so a launch's event
    # duration there includes other frames' kernels running beside it.  A
    # kernel's own launch duration comes from frames rendered one at a time
    # (each waits for the previous one): the same frame as the instrumented
    # re-render, so its algorithmic bytes are exact.
    def isolated():
        for _ in range(args.roofline_frames):
            fi = frame.copy()
            fi.frame_index = args.warmup
            fi.reset_accum = 0
            scene.render(fi)
            dev.synchronize()

    _, st_iso = timed(isolated)
    nf = args.roofline_frames
    kern = {}
    for name, key, nbytes in (("k_extend", "extend", ext_bytes), ("k_tail", "tail", tail_bytes)):
        launches = st_iso[f"{key}_launches"]
        if launches == 0:
            continue
        per_launch = nbytes * nf / launches
        avg_ms = st_iso[f"{key}_ms"] / launches
        shared = st[f"{key}_ms"] / max(st[f"{key}_launches"], 1)
        kern[name] = dict(bytes_per_launch=per_launch, avg_ms=avg_ms, ms_per_frame=st_iso[f"{key}_ms"] / nf,
                          launches=launches, achieved=per_launch / (avg_ms * 1e-3) / 1e9)
        log(f"[bench] {name}: {launches / nf:.0f} launch(es)/frame, {per_launch / 1e6:.1f} MB/launch, "
            f"avg launch {avg_ms:.3f} ms alone -> {kern[name]['achieved']:.0f} GB/s "
            f"({shared:.3f} ms with four frames in flight)")
    log(f"[bench] k_extend: {tr['ext_rays']} rays, {tr['ext_node_visits'] / max(tr['ext_rays'], 1):.1f} nodes/ray, "
        f"{tr['ext_tri_tests'] / max(tr['ext_rays'], 1):.2f} tri tests/ray; k_tail: {tr['tail_node_visits']} node "
        f"visits, {tr['tail_tri_tests']} tri tests, {tr['tail_sphere_tests']} sphere tests")
    dom = max(kern, key=lambda k: kern[k]["ms_per_frame"])
    K = kern[dom]
    bytes_per_launch, avg_launch_ms, achieved = K["bytes_per_launch"], K["avg_ms"], K["achieved"]

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ctypes as orc
        ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, ncpu - 4) if ncpu > 4 else ncpu   # SoftwareTracer.cpp:867-868 policy
        oscene = orc.OracleScene.from_scene_data(sd)

        def cpu_run(nthreads):
            # whole frames (frame_index 0, 1, ...) until at least cpu_seconds of work
            crays, t0 = 0, time.perf_counter()
            k = 0
            while True:
                f = frame.copy()
                f.frame_index = k
                f.reset_accum = 1
                _, cnt = oscene.render(f, 32, nthreads)
                crays += cnt["primary"] + cnt["extension"] + cnt["shadow"]
                k += 1
                el = time.perf_counter() - t0
                if el >= args.cpu_seconds or k >= 200:
                    return crays, el, k

        crays, cdt, nfr = cpu_run(threads)
        cpu = {"value": round(crays / cdt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{nfr} whole {frame.width}x{frame.height} frames ({crays} rays, {cdt:.1f} s) of the fp32 CPU oracle "
                         f"(reference semantics, same scene, median BVH), {threads} threads of {ncpu} "
                         f"(hw - 4, SoftwareTracer.cpp:867-868)"}
        if threads != ncpu:
            crays2, cdt2, nfr2 = cpu_run(ncpu)
            cpu["all_cores"] = {"value": round(crays2 / cdt2 / 1e6, 3), "cores": ncpu,
                                "sample": f"{nfr2} frames, {cdt2:.1f} s"}
        log(f"[bench] cpu baseline {cpu}")

    if cfg == 1:
        metric, workload = "Mrays/s @1080p 1M tris 1spp", (
            "icosphere-8 displaced + 200k random tris (1,510,720 tris), 512x256 lognormal env, 1 quad light, "
            "1920x1080, 1 spp, max depth 8, binned-SAH BVH4")
    else:
        metric, workload = "Mrays/s (many-light mixed scene)", (
            "4,096 spheres + 500k-tri mesh lit by 4,096 emissive triangles through the light tree, all-zero env, "
            "1280x720, 4 spp, max depth 6, binned-SAH BVH4")
    line = {
        "metric": metric,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded generator, config {cfg})",
        "config": {"workload": workload,
                   "triangles": int(sd.indices.shape[0]), "width": frame.width, "height": frame.height,
                   "spp": frame.spp, "max_bounces": frame.max_bounces, "rays_per_frame": rays // args.steps,
                   "schedule": ("per-bounce kernels" if schedule == dmt.SCHEDULE_WAVEFRONT else
                                f"{schedule} per-bounce iteration(s), then one frame-completion launch")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom) if cfg == 1 else None,
                     "kernel": dom, "bytes_per_launch": int(bytes_per_launch),
                     "avg_launch_ms": round(avg_launch_ms, 5),
                     "timing": f"{int(K['launches'])} launches of {args.roofline_frames} frames rendered "
                               "one at a time, HIP events on the launch stream",
                     "other_kernels": {k: {"achieved": round(v["achieved"], 1), "avg_launch_ms": round(v["avg_ms"], 5),
                                           "bytes_per_launch": int(v["bytes_per_launch"])}
                                       for k, v in kern.items() if k != dom}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    scene.close()
    dev.close()


def bench_batch(args, np, torch, dmt):
    t0 = time.time()
    dev = dmt.Device(0)
    scenes, frames = [], []
    tris = pixels = 0
    for i in range(args.scenes):
        sd = dmt.generate_batch_scene(args.seed, i)
        scenes.append(dmt.Scene(dev, sd))
        frames.append(sd.frame.copy())
        tris += int(sd.indices.shape[0])
        pixels += sd.frame.width * sd.frame.height
        del sd
    log(f"[bench] {len(scenes)} scenes, {pixels / 1e6:.2f} M pixels, {tris / 1e6:.2f} M triangles "
        f"(generate + build + upload {time.time() - t0:.1f}s)")

    def sync():
        dev.synchronize()
        torch.cuda.synchronize()

    # plain: each frame is one dmt_render per scene.  Sequential: every
    # render runs to completion before the next starts (host sync after each
    # scene's frame).  Pipelined: the library's own frames in flight (up to
    # four renders overlap on the device).
    def plain(k0, n, each=None):
        for k in range(k0, k0 + n):
            for sc, f in zip(scenes, frames):
                g = f.copy()
                g.frame_index = f.frame_index + k
                g.reset_accum = 1 if k == 0 else 0
                sc.render(g)
                if each:
                    each()

    def timed_plain(each):
        plain(0, args.warmup, each)
        sync()
        for sc in scenes:
            sc.reset_stats()
        sync()
        t0 = time.perf_counter()
        plain(args.warmup, args.steps, each)
        sync()
        dt = time.perf_counter() - t0
        rays = 0
        for sc in scenes:
            st = sc.stats()
            rays += st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
        return dt, rays, [sc.readback(f.width, f.height) for sc, f in zip(scenes, frames)]

    dt_pipe, rays_plain, images = timed_plain(None)
    log(f"[bench] plain, pipelined: {args.steps} frames in {dt_pipe * 1e3:.1f} ms "
        f"({rays_plain / dt_pipe / 1e6:.1f} Mrays/s)")
    dt_plain, rays_seq, images_seq = timed_plain(dev.synchronize)
    same_seq = all(np.array_equal(a, b) for a, b in zip(images, images_seq)) and rays_seq == rays_plain
    log(f"[bench] plain, sequential: {args.steps} frames in {dt_plain * 1e3:.1f} ms "
        f"({rays_plain / dt_plain / 1e6:.1f} Mrays/s), images equal to the pipelined renders {same_seq}")

    batch = dmt.Batch(dev, scenes, frames)
    t0 = time.time()
    tuned, tune_ms = batch.autotune(frames_per_trial=2)
    log(f"[bench] autotune {tuned.as_dict()} {tune_ms:.2f} ms/frame ({time.time() - t0:.1f}s)")
    batch.render(0, args.warmup)
    sync()
    batch.reset_stats()
    sync()
    t0 = time.perf_counter()
    batch.render(args.warmup, args.steps)
    sync()
    dt_batch = time.perf_counter() - t0
    st = batch.stats()
    rays_batch = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
    same = [bool(np.array_equal(sc.readback(f.width, f.height), img)) for sc, f, img in zip(scenes, frames, images)]
    log(f"[bench] batch: {args.steps} frames in {dt_batch * 1e3:.1f} ms ({rays_batch / dt_batch / 1e6:.1f} Mrays/s), "
        f"identical images {sum(same)}/{len(same)}, rays equal {rays_batch == rays_plain}")
    speedup = dt_plain / dt_batch
    line = {
        "metric": "batched vs plain speedup",
        "value": round(speedup, 3),
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_batch * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded batched-workload generator)",
        "config": {"workload": f"{len(scenes)} scenes: resolution uniform in {{128,256,512,1024}}^2, triangles "
                               "log-uniform 10k..2M (icosphere + random tris), odd scenes + 64 spheres; "
                               "1 spp, max depth 6; one persistent launch per frame from a captured hipGraph "
                               "vs 64 sequential per-scene renders (each to completion)",
                   "scenes": len(scenes), "pixels": pixels, "triangles": tris, "seed": args.seed,
                   "params": tuned.as_dict()},
        "plain_ms_per_step": round(dt_plain * 1e3 / args.steps, 4),
        "batch_mrays": round(rays_batch / dt_batch / 1e6, 2),
        "plain_mrays": round(rays_plain / dt_plain / 1e6, 2),
        "vs_pipelined_plain": round(dt_pipe / dt_batch, 3),
        "pipelined_plain_ms_per_step": round(dt_pipe * 1e3 / args.steps, 4),
        "identical_images": all(same) and rays_batch == rays_plain and same_seq,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    batch.close()
    for sc in scenes:
        sc.close()
    dev.close()


def bench_adjoint(args, np, torch, dmt):
    sd = dmt.generate(4, seed=args.seed)
    dev = dmt.Device(0)
    sc = dmt.Scene(dev, sd)
    f = sd.frame.copy()
    f.flags = dmt.DMT_FRAME_NO_RUSSIAN_ROULETTE
    f.reset_accum = 1
    W, H = f.width, f.height
    gt = sc.materials()
    sc.render(f)
    target = sc.readback(W, H)[..., :3].astype(np.float64)
    rng = np.random.default_rng(args.seed + 100)
    init = gt.copy()
    init[1:] = gt[1:] + rng.normal(0.0, 0.1, size=gt[1:].shape).astype(np.float32)
    init[1:, 0:3] = np.clip(init[1:, 0:3], 0.05, 1.0)
    init[1:, 4:7] = np.maximum(init[1:, 4:7], 0.02)
    sc.set_materials(init.astype(np.float32))

    def sync():
        dev.synchronize()
        torch.cuda.synchronize()

    for _ in range(args.warmup):
        sc.render(f)
    sync()
    fwd = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        sc.render(f)
        sync()
        fwd.append(time.perf_counter() - t0)
    img = sc.readback(W, H)[..., :3].astype(np.float64)
    loss = float(((img - target) ** 2).sum())
    dimg = torch.zeros((H, W, 4), device="cuda", dtype=torch.float32)
    dimg[..., :3] = torch.from_numpy((2.0 * (img - target)).astype(np.float32)).cuda()
    grads = torch.empty((sd.prims.shape[0], dmt.MATERIAL_FLOATS), device="cuda", dtype=torch.float32)
    sync()
    for _ in range(args.warmup):
        sc.adjoint(f, dimg.data_ptr(), grads.data_ptr())
    sync()
    bwd = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        sc.adjoint(f, dimg.data_ptr(), grads.data_ptr())
        sync()
        bwd.append(time.perf_counter() - t0)
    g = grads.cpu().numpy()
    fwd_ms = 1e3 * float(np.median(fwd))
    bwd_ms = 1e3 * float(np.median(bwd))
    log(f"[bench] forward {fwd_ms:.3f} ms, backward {bwd_ms:.3f} ms, loss {loss:.4g}, "
        f"|grad| max {np.abs(g).max():.4g}, finite {bool(np.isfinite(g).all())}")
    line = {
        "metric": "bwd/fwd time ratio",
        "value": round(bwd_ms / fwd_ms, 3),
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bwd_ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, config 4)",
        "config": {"workload": "256 spheres with 256 materials + ground plane, quad light, 512x512, 16 spp, "
                               "max depth 4, Russian roulette off; L2 image loss; gradients of 256 x (3 albedo + "
                               "1 roughness + 3 emission)",
                   "width": W, "height": H, "spp": f.spp, "max_bounces": f.max_bounces,
                   "params": int(sd.prims.shape[0] - 1) * 7},
        "forward_ms": round(fwd_ms, 4),
        "backward_ms": round(bwd_ms, 4),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    sc.close()
    dev.close()


def bench_features(args, np, torch, dmt):
    sd = dmt.generate(1, seed=args.seed)
    dev = dmt.Device(0)
    scene = dmt.Scene(dev, sd)
    plain = sd.frame.copy()
    feat = plain.copy()
    feat.dof = (0.05, 12.0, 6.0, 0.2)
    feat.cat_eye = 0.3
    feat.motion_blur = 1
    feat.spectral = 1
    feat.cauchy_c = 0.0042
    feat.mesh_shift_active = 1
    feat.mesh_shift_prev = (-0.1, 0.0, 0.0)
    feat.mesh_shift_curr = (0.0, 0.0, 0.0)

    def sync():
        dev.synchronize()
        torch.cuda.synchronize()

    def run(fr):
        for k in range(args.warmup):
            f = fr.copy()
            f.frame_index, f.reset_accum = k, 1 if k == 0 else 0
            scene.render(f)
        sync()
        scene.reset_stats()
        sync()
        t0 = time.perf_counter()
        for k in range(args.steps):
            f = fr.copy()
            f.frame_index, f.reset_accum = args.warmup + k, 0
            scene.render(f)
        sync()
        dt = time.perf_counter() - t0
        st = scene.stats()
        rays = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
        return rays / dt / 1e6, dt * 1e3 / args.steps, rays / args.steps

    mp, msp, rp = run(plain)
    mf, msf, rf = run(feat)
    log(f"[bench] plain frame {msp:.3f} ms, {mp:.1f} Mrays/s ({rp:.0f} rays); feature frame {msf:.3f} ms, "
        f"{mf:.1f} Mrays/s ({rf:.0f} rays)")
    line = {
        "metric": "Mrays/s with camera-sample features",
        "value": round(mf, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msf, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, config 1)",
        "config": {"workload": "config 1 (1.51 M triangles, 1920x1080, 1 spp, depth 8) with thin lens "
                               "(aperture 0.05, focus 12, 6 blades, cat's eye 0.3), motion blur (mesh translated "
                               "over the shutter) and hero-wavelength dispersion",
                   "width": plain.width, "height": plain.height, "spp": plain.spp,
                   "max_bounces": plain.max_bounces},
        "vs_plain": round(mf / mp, 3),
        "plain_mrays": round(mp, 2),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    scene.close()
    dev.close()


if __name__ == "__main__":
    main()
