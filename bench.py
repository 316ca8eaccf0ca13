"""Benchmark of the path tracer: the three figures of its headline metric in
one run -- Mrays/s on the 1.51 M-triangle 1080p workload, the batched form's
speedup over plain per-scene launches, the backward/forward time ratio of the
differentiable form -- plus Mrays/s on the many-light workload.

`python bench.py` (the default, --workload all) prints ONE JSON line: the
1080p workload's figures at top level (value = its Mrays/s), and the objects
"c2", "batched" and "adjoint" for the other workloads.  --workload c1 / c2 /
batch / adjoint / features runs one of them alone (its own line).

c1 (top level): icosphere subdivided 8x (1,310,720 triangles, radius 3, value-
noise displacement 0.1) + 200,000 random small triangles in [-5,5]^3, one quad
light, 512x256 lognormal env map, 1920x1080, 1 spp, 8 bounces, binned-SAH BVH4
built on the host (not timed).  A step is one frame (1920x1080, 1 spp, up to 8
bounces, accumulated into the running mean), issued as one dmt_render call:
camera rays into the path queue (k_generate), the frame-completion kernel
(k_tail) that runs every path to its end, the accumulation pass
(k_accumulate).  The library keeps four frames in flight on four HIP streams
(the accumulation passes stay ordered frame after frame, so the image equals
the sequential result).  value = (primary + extension + shadow rays) / wall
time of the K frames, inputs resident in HBM.

c2: 4,096 analytic objects -- 3,584 spheres and 512 sphere-traced SDF clusters
(boxes, rounded boxes, tori, hard and smooth CSG unions, a subtraction), a
third each Lambert, metal, glass -- + a 500k-triangle mesh lit by 4,096
emissive triangles through the light tree, 1280x720, 4 spp, depth 6: the
workload that exercises next-event estimation, the light tree, shadow
transmittance and the SDF clusters.  Same step and value definitions.

roofline (c1 and c2): the frame-completion kernel k_tail, which has no dense
contraction (no MFMA), against the 8 TB/s HBM peak.  achieved = HBM-side bytes
per launch, measured (rocprofv3 FETCH_SIZE + WRITE_SIZE passes of the same
workload, scaled by a 512 MiB copy; recorded in profiles/k_tail_<c>.txt) /
the kernel's launch duration, measured here with HIP events on its stream in
frames rendered one at a time after the timed region.  frac = achieved /
peak.  The profile records a hash of the kernel sources it was taken with;
"profile_matches_sources" says whether the current sources hash the same.
"algorithmic" gives the bytes the traversal reads at the node level: 128 B
per BVH4 node visit + 48 B per triangle test + 20 B per sphere or SDF-cluster
test (a cluster's trace steps read its record from the L1/L2), counted
by an instrumented re-render of the same frame, and their rate; most of them
are served by the L2 caches and the 256 MB MALL (l2_hit), which is why that
rate can exceed the HBM peak.  l2_hit, valu_busy and valu_lane_utilisation
come from the same profile; "limiter" names what they show.

batched: 64 scenes of the batched workload generator (resolution uniform in
{128, 256, 512, 1024} squared, triangles log-uniform in 10k..2M, half with 64
spheres; 1 spp, depth 6).  The plain form renders each frame as 64 per-scene
dmt_render calls; the batched form renders each frame of all 64 scenes with
one persistent-kernel launch replayed from a captured hipGraph, with
parameters picked by dmt_batch_autotune.  All run the same frames from frame
0; the images must be identical.  value = plain time / batched time over the
timed frames, the plain renders issued back to back (the library keeps up to
four frames in flight); vs_sequential_plain compares with the plain renders
each run to completion (host sync after every scene's frame).

adjoint: config 4: 256 spheres with 256 distinct materials over the ground
plane, 512x512, 16 spp, depth 4, Russian roulette off.  The target image is
rendered with the generator's materials; the spheres' parameters are then
perturbed by N(0, 0.1) (albedo kept in [0.05, 1], emission >= 0.02).  A
forward step is one dmt_render of the frame (to completion), a backward step
one dmt_adjoint with dLoss/dI = 2 (I - target) (to completion).  value =
backward time / forward time (lower is better).

--workload features: the camera-sample features on the config-1 workload:
thin lens (aperture 0.05, focus 12, six blades, cat's-eye 0.3), motion blur
(the mesh translated from (-0.1, 0, 0) to its rest position over the
shutter), hero-wavelength dispersion (C = 0.0042 um^2).  value = their Mrays/s
over K frames issued back to back; vs_plain = that rate over the plain
frame's rate (the same scene and camera without the features).

cpu_baseline (c1): the CPU oracle (reference semantics, fp32) rendering whole
frames of the same workload for about 10 s, on this host's cores (hw - 4
threads as the reference's software tracer uses, and all cores).
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "demont-engine_amd", "python"))

HBM_PEAK_GBS = 8000.0
PROFILES = {"c1": "k_tail_c1_1080p.txt", "c2": "k_tail_c2_720p.txt"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_sources_sha256():
    """Hash of the device sources (kernel headers and sources, the C ABI's
    launch code): a profile taken with other sources is flagged."""
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "demont-engine_amd", "csrc")
    files = sorted(glob.glob(os.path.join(csrc, "kernels", "*.h")) + glob.glob(os.path.join(csrc, "kernels", "*.hip")))
    files.append(os.path.join(csrc, "capi.hip"))
    for p in files:
        h.update(os.path.relpath(p, ROOT).encode())
        with open(p, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def read_profile(workload):
    """The figures section of profiles/k_tail_<workload>.txt (key: value)."""
    name = PROFILES.get(workload)
    if not name:
        return None
    path = os.path.join(ROOT, "profiles", name)
    out = {"path": os.path.relpath(path, ROOT)}
    try:
        with open(path) as fh:
            figures = False
            for line in fh:
                if line.startswith("## figures"):
                    figures = True
                    continue
                if figures and ":" in line:
                    k, v = line.split(":", 1)
                    out[k.strip()] = v.strip()
    except OSError:
        return None
    return out if "traffic_bytes_per_launch" in out else None


def roofline(workload, kernel, avg_ms, algo_bytes, launches, frames):
    r = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None,
         "kernel": kernel, "avg_launch_ms": round(avg_ms, 5),
         "timing": f"{launches} launches of {frames} frames rendered one at a time, HIP events on the launch stream"}
    prof = read_profile(workload)
    if prof:
        traffic = int(prof["traffic_bytes_per_launch"])
        achieved = traffic / (avg_ms * 1e-3) / 1e9
        busy, util = float(prof["valu_busy"]), float(prof["valu_lane_utilisation"])
        r.update(achieved=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                 profile=prof["path"], profile_matches_sources=prof.get("kernel_sources_sha256") == kernel_sources_sha256(),
                 l2_hit=float(prof["l2_hit"]), valu_busy=busy, valu_lane_utilisation=util,
                 limiter=(f"VALU issue: {busy:.0%} VALU busy at {util:.0%} lane utilisation (divergence), "
                          f"{achieved / HBM_PEAK_GBS:.0%} of HBM bandwidth" if busy >= 0.5 else
                          f"dependent fetch latency: {busy:.0%} VALU busy, {achieved / HBM_PEAK_GBS:.0%} of HBM bandwidth"))
    r["algorithmic"] = {"bytes_per_launch": int(algo_bytes),
                        "per_unit": "128 B per BVH4 node visit + 48 B per triangle test + 20 B per sphere or "
                                    "SDF-cluster test",
                        "GBps": round(algo_bytes / (avg_ms * 1e-3) / 1e9, 1),
                        "served_by": "L2 and MALL for the most part (l2_hit), so this rate is not HBM bandwidth"}
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("all", "c1", "c2", "batch", "adjoint", "features"), default="all")
    ap.add_argument("--steps", type=int, default=None, help="timed frames (c1/c2: 20, batch: 100, adjoint: 10)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed frames (c1/c2: 3, batch/adjoint: 2)")
    ap.add_argument("--scenes", type=int, default=64, help="batch workload: number of scenes")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--schedule", type=int, default=None,
                    help="per-bounce iterations before the frame-completion launch (-1: per-bounce kernels only; "
                         "default: the library's)")
    ap.add_argument("--roofline-frames", type=int, default=10,
                    help="frames rendered one at a time to time the frame-completion kernel alone")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library renders a frame on one GPU and has no multi-GPU path "
              f"(asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import dmt

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
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

    def steps(default):
        return default if args.steps is None else args.steps

    def warm(default):
        return default if args.warmup is None else args.warmup

    if args.workload == "batch":
        line = bench_batch(np, torch, dmt, args.scenes, args.seed, steps(100), warm(2))
    elif args.workload == "adjoint":
        line = bench_adjoint(np, torch, dmt, args.seed, steps(10), warm(2))
    elif args.workload == "features":
        line = bench_features(np, torch, dmt, args.seed, steps(10), warm(2))
    elif args.workload in ("c1", "c2"):
        line = bench_frames(np, torch, dmt, args, 1 if args.workload == "c1" else 2, steps(20), warm(3),
                            cpu=not args.no_cpu_baseline and args.workload == "c1")
    else:
        line = bench_frames(np, torch, dmt, args, 1, steps(20), warm(3), cpu=not args.no_cpu_baseline)
        c2 = bench_frames(np, torch, dmt, args, 2, 20, 3, cpu=False)
        line["c2"] = {k: c2[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "warmup", "config",
                                         "roofline")}
        b = bench_batch(np, torch, dmt, args.scenes, args.seed, 100, 2)
        line["batched"] = {k: b[k] for k in b if k not in ("n_gpus", "higher_is_better", "scaling", "vs_baseline",
                                                          "dtype", "data", "roofline", "cpu_baseline")}
        a = bench_adjoint(np, torch, dmt, args.seed, 10, 2)
        line["adjoint"] = {k: a[k] for k in a if k not in ("n_gpus", "scaling", "vs_baseline", "dtype", "data",
                                                          "roofline", "cpu_baseline")}
        line["metric"] = "Mrays/s @1080p 1M tris 1spp; batched vs plain speedup; bwd/fwd time ratio"
    print(json.dumps(line), flush=True)


def bench_frames(np, torch, dmt, args, cfg, nsteps, nwarm, cpu):
    t0 = time.time()
    sd = dmt.generate(cfg, seed=args.seed)
    log(f"[bench] c{cfg} scene: {sd.indices.shape[0]} triangles, {sd.prims.shape[0]} prims, {sd.lights.shape[0]} lights,"
        f" {sd.frame.width}x{sd.frame.height}, {sd.frame.spp} spp, {sd.frame.max_bounces} bounces "
        f"({time.time() - t0:.1f}s)")
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
        for k in range(nsteps):
            f = frame.copy()
            f.frame_index = nwarm + k
            f.reset_accum = 0
            scene.render(f)

    run_frames(0, nwarm)
    dt, st = timed(per_frame)
    rays = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
    mrays = rays / dt / 1e6
    log(f"[bench] {nsteps} frames in {dt * 1e3:.1f} ms: {mrays:.1f} Mrays/s "
        f"(primary {st['primary_rays']}, extension {st['extension_rays']}, shadow {st['shadow_rays']})")

    # Node-level bytes of the closest-hit kernel (k_extend) and of the
    # frame-completion kernel (k_tail), from an instrumented re-render.
    f = frame.copy()
    f.frame_index = nwarm
    tr = scene.measure_traffic(f)
    ext_bytes = (tr["ext_node_visits"] * 128 + tr["ext_tri_tests"] * 48 + tr["ext_sphere_tests"] * 20
                 + tr["ext_rays"] * 40)
    tail_bytes = tr["tail_node_visits"] * 128 + tr["tail_tri_tests"] * 48 + tr["tail_sphere_tests"] * 20

    # In the timed region four frames share the GPU, so a launch's event
    # duration there includes other frames' kernels running beside it.  A
    # kernel's own launch duration comes from frames rendered one at a time
    # (each waits for the previous one): the same frame as the instrumented
    # re-render, so its node-level bytes are exact.
    def isolated():
        for _ in range(args.roofline_frames):
            fi = frame.copy()
            fi.frame_index = nwarm
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
                          launches=launches)
        log(f"[bench] {name}: {launches / nf:.0f} launch(es)/frame, {per_launch / 1e6:.1f} MB node-level/launch, "
            f"avg launch {avg_ms:.3f} ms alone ({shared:.3f} ms with four frames in flight)")
    log(f"[bench] k_tail: {tr['tail_node_visits']} node visits, {tr['tail_tri_tests']} tri tests, "
        f"{tr['tail_sphere_tests']} sphere and SDF-cluster tests per frame")
    dom = max(kern, key=lambda k: kern[k]["ms_per_frame"])
    K = kern[dom]
    roof = roofline(f"c{cfg}", dom, K["avg_ms"], K["bytes_per_launch"], int(K["launches"]), nf)
    log(f"[bench] roofline {roof}")

    cpu_line = None
    if cpu:
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
        cpu_line = {"value": round(crays / cdt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                    "sample": f"{nfr} whole {frame.width}x{frame.height} frames ({crays} rays, {cdt:.1f} s) of the fp32 "
                              f"CPU oracle (reference semantics, same scene, median BVH), {threads} threads of {ncpu} "
                              f"(hw - 4, SoftwareTracer.cpp:867-868)"}
        if threads != ncpu:
            crays2, cdt2, nfr2 = cpu_run(ncpu)
            cpu_line["all_cores"] = {"value": round(crays2 / cdt2 / 1e6, 3), "cores": ncpu,
                                     "sample": f"{nfr2} frames, {cdt2:.1f} s"}
        log(f"[bench] cpu baseline {cpu_line}")

    if cfg == 1:
        metric, workload = "Mrays/s @1080p 1M tris 1spp", (
            "icosphere-8 displaced + 200k random tris (1,510,720 tris), 512x256 lognormal env, 1 quad light, "
            "1920x1080, 1 spp, max depth 8, binned-SAH BVH4")
    else:
        metric, workload = "Mrays/s (many-light mixed scene)", (
            "3,584 spheres + 512 SDF clusters (boxes, rounded boxes, tori, CSG unions; Lambert, metal, glass) + "
            "500k-tri mesh lit by 4,096 emissive triangles through the light tree, all-zero env, 1280x720, 4 spp, "
            "max depth 6, binned-SAH BVH4")
    line = {
        "metric": metric,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": nwarm,
        "ms_per_step": round(dt * 1e3 / nsteps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded generator, config {cfg})",
        "config": {"workload": workload,
                   "triangles": int(sd.indices.shape[0]), "width": frame.width, "height": frame.height,
                   "spp": frame.spp, "max_bounces": frame.max_bounces, "rays_per_frame": rays // nsteps,
                   "shadow_rays_per_frame": st["shadow_rays"] // nsteps,
                   "schedule": ("per-bounce kernels" if schedule == dmt.SCHEDULE_WAVEFRONT else
                                f"{schedule} per-bounce iteration(s), then one frame-completion launch")},
        "roofline": roof,
        "cpu_baseline": cpu_line,
    }
    scene.close()
    dev.close()
    return line


def bench_batch(np, torch, dmt, nscenes, seed, nsteps, nwarm):
    t0 = time.time()
    dev = dmt.Device(0)
    scenes, frames = [], []
    tris = pixels = 0
    for i in range(nscenes):
        sd = dmt.generate_batch_scene(seed, i)
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

    # plain: each frame is one dmt_render per scene.  Pipelined: the
    # library's own frames in flight (up to four renders overlap on the
    # device).  Sequential: every render runs to completion before the next
    # starts (host sync after each scene's frame).
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
        plain(0, nwarm, each)
        sync()
        for sc in scenes:
            sc.reset_stats()
        sync()
        t0 = time.perf_counter()
        plain(nwarm, nsteps, each)
        sync()
        dt = time.perf_counter() - t0
        rays = 0
        for sc in scenes:
            st = sc.stats()
            rays += st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
        return dt, rays, [sc.readback(f.width, f.height) for sc, f in zip(scenes, frames)]

    dt_pipe, rays_plain, images = timed_plain(None)
    log(f"[bench] plain, pipelined: {nsteps} frames in {dt_pipe * 1e3:.1f} ms "
        f"({rays_plain / dt_pipe / 1e6:.1f} Mrays/s)")
    dt_seq, rays_seq, images_seq = timed_plain(dev.synchronize)
    same_seq = all(np.array_equal(a, b) for a, b in zip(images, images_seq)) and rays_seq == rays_plain
    log(f"[bench] plain, sequential: {nsteps} frames in {dt_seq * 1e3:.1f} ms "
        f"({rays_plain / dt_seq / 1e6:.1f} Mrays/s), images equal to the pipelined renders {same_seq}")

    batch = dmt.Batch(dev, scenes, frames)
    t0 = time.time()
    tuned, tune_ms = batch.autotune(frames_per_trial=2)
    log(f"[bench] autotune {tuned.as_dict()} {tune_ms:.2f} ms/frame ({time.time() - t0:.1f}s)")
    batch.render(0, nwarm)
    sync()
    batch.reset_stats()
    sync()
    t0 = time.perf_counter()
    batch.render(nwarm, nsteps)
    sync()
    dt_batch = time.perf_counter() - t0
    st = batch.stats()
    rays_batch = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
    same = [bool(np.array_equal(sc.readback(f.width, f.height), img)) for sc, f, img in zip(scenes, frames, images)]
    log(f"[bench] batch: {nsteps} frames in {dt_batch * 1e3:.1f} ms ({rays_batch / dt_batch / 1e6:.1f} Mrays/s), "
        f"identical images {sum(same)}/{len(same)}, rays equal {rays_batch == rays_plain}")
    line = {
        "metric": "batched vs plain speedup",
        "value": round(dt_pipe / dt_batch, 3),
        "unit": "x",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": nwarm,
        "ms_per_step": round(dt_batch * 1e3 / nsteps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded batched-workload generator)",
        "config": {"workload": f"{len(scenes)} scenes: resolution uniform in {{128,256,512,1024}}^2, triangles "
                               "log-uniform 10k..2M (icosphere + random tris), odd scenes + 64 spheres; "
                               "1 spp, max depth 6; one persistent launch per frame from a captured hipGraph "
                               "vs 64 per-scene renders issued back to back",
                   "scenes": len(scenes), "pixels": pixels, "triangles": tris, "seed": seed,
                   "params": tuned.as_dict()},
        "plain_ms_per_step": round(dt_pipe * 1e3 / nsteps, 4),
        "batch_mrays": round(rays_batch / dt_batch / 1e6, 2),
        "plain_mrays": round(rays_plain / dt_pipe / 1e6, 2),
        "vs_sequential_plain": round(dt_seq / dt_batch, 3),
        "sequential_plain_ms_per_step": round(dt_seq * 1e3 / nsteps, 4),
        "identical_images": all(same) and rays_batch == rays_plain and same_seq,
        "roofline": None,
        "cpu_baseline": None,
    }
    batch.close()
    for sc in scenes:
        sc.close()
    dev.close()
    return line


def bench_adjoint(np, torch, dmt, seed, nsteps, nwarm):
    sd = dmt.generate(4, seed=seed)
    dev = dmt.Device(0)
    sc = dmt.Scene(dev, sd)
    f = sd.frame.copy()
    f.flags = dmt.DMT_FRAME_NO_RUSSIAN_ROULETTE
    f.reset_accum = 1
    W, H = f.width, f.height
    gt = sc.materials()
    sc.render(f)
    target = sc.readback(W, H)[..., :3].astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    init = gt.copy()
    init[1:] = gt[1:] + rng.normal(0.0, 0.1, size=gt[1:].shape).astype(np.float32)
    init[1:, 0:3] = np.clip(init[1:, 0:3], 0.05, 1.0)
    init[1:, 4:7] = np.maximum(init[1:, 4:7], 0.02)
    sc.set_materials(init.astype(np.float32))

    def sync():
        dev.synchronize()
        torch.cuda.synchronize()

    for _ in range(nwarm):
        sc.render(f)
    sync()
    fwd = []
    for _ in range(nsteps):
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
    for _ in range(nwarm):
        sc.adjoint(f, dimg.data_ptr(), grads.data_ptr())
    sync()
    bwd = []
    for _ in range(nsteps):
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
        "steps": nsteps,
        "warmup": nwarm,
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
        "gradients_finite": bool(np.isfinite(g).all()),
        "roofline": None,
        "cpu_baseline": None,
    }
    sc.close()
    dev.close()
    return line


def bench_features(np, torch, dmt, seed, nsteps, nwarm):
    sd = dmt.generate(1, seed=seed)
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
        for k in range(nwarm):
            f = fr.copy()
            f.frame_index, f.reset_accum = k, 1 if k == 0 else 0
            scene.render(f)
        sync()
        scene.reset_stats()
        sync()
        t0 = time.perf_counter()
        for k in range(nsteps):
            f = fr.copy()
            f.frame_index, f.reset_accum = nwarm + k, 0
            scene.render(f)
        sync()
        dt = time.perf_counter() - t0
        st = scene.stats()
        rays = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
        return rays / dt / 1e6, dt * 1e3 / nsteps, rays / nsteps

    mp, msp, rp = run(plain)
    mf, msf, rf = run(feat)
    log(f"[bench] plain frame {msp:.3f} ms, {mp:.1f} Mrays/s ({rp:.0f} rays); feature frame {msf:.3f} ms, "
        f"{mf:.1f} Mrays/s ({rf:.0f} rays)")
    line = {
        "metric": "Mrays/s with camera-sample features",
        "value": round(mf, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": nwarm,
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
    scene.close()
    dev.close()
    return line


if __name__ == "__main__":
    main()
