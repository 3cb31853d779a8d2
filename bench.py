"""Benchmark of the MI355X path-tracing hot path (one JSON line on stdout).

Workload: the box lit by a 1024x512 sky+sun environment map sampled by the steerable sampler,
GGX dielectric + Lambert spheres, 512x512 pixels, 64 samples per pixel, max path length 10
(-sb 1 -iic 0).  One step = one full render of that image (all 64 iterations of RunIteration).
`value` is Msamples/s with every input already resident in HBM; the roofline object reports the
dominant kernel against MI355X peaks, and `cpu_baseline` times the CPU oracle (a restatement of
the reference's OpenMP renderer) on a bounded slice of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 (vector) peak
TRI_FLOP, SPH_FLOP, SHADE_FLOP = 62, 35, 400
BYTES_PER_SAMPLE, BYTES_PER_SEGMENT = 112, 304
# HBM bytes per launch of the dominant kernel, measured with rocprofv3 (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 KiB, factors from a 512 MiB copy): profiles/render_pt_box_steerable_em_512px_64spp.md
MEASURED_TRAFFIC = {(2, "megakernel"): int((2.0 * 3855186 + 1467896) * 1024)}
# Wave-level VALU instructions per launch of the dominant kernel (rocprofv3 SQ_INSTS_VALU, same
# file) and the issue peak they are priced against: 256 CUs x 2 FP32 wave-64 instructions per
# cycle (four SIMDs) x 2.4 GHz
MEASURED_VALU_INSTS = {(2, "megakernel"): 7047553485}
VALU_ISSUE_PEAK = 256 * 2 * 2.4e9
# HBM bytes per frame of the wavefront pipeline on the default workload (all Wf* launches of a
# frame, same calibration): profiles/wavefront_box_steerable_em_512px_64spp.md
MEASURED_WF_TRAFFIC = int((2.0 * 16145826 + 15518939) * 1024)
# HBM bytes per launch of the split kernel at the reference's defaults (same calibration and file)
MEASURED_SPLIT_TRAFFIC = {"box_area_light_256px_16spp": int((2.0 * 887732 + 696239) * 1024),
                          "box_steerable_em_512px_64spp": int((2.0 * 52301999 + 13773262) * 1024)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=2, help="2: steerable-EM box 512^2 x 64 spp")
    ap.add_argument("--kernel", choices=["megakernel", "wavefront"], default="megakernel")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched-vs-plain measurement")
    ap.add_argument("--batch-tune", type=int, default=5, help="autotuner candidates for the batch (0 = all)")
    ap.add_argument("--no-diff", action="store_true", help="skip the adjoint and inverse-rendering measurements")
    ap.add_argument("--copy-probe", action="store_true", help="only run the HBM-counter calibration copy")
    ap.add_argument("--inverse-steps", type=int, default=200, help="Adam steps of the inverse-rendering run")
    return ap.parse_args()


def roofline_legs(desc, samples, segments, kernel_s):
    """The two algorithmic legs of one kernel (per launch): HBM bytes of the SoA queue model
    (BYTES_PER_SAMPLE per camera sample + BYTES_PER_SEGMENT per path segment) and FP32 FLOP of the
    brute-force intersection + shading model, each against its MI355X peak.  `bound` is the leg
    with the larger fraction of its peak; `frac` = max of the two."""
    ray_flop = TRI_FLOP * desc.num_triangles + SPH_FLOP * desc.num_spheres
    flops = samples * ray_flop + segments * (2 * ray_flop + SHADE_FLOP)
    bytes_alg = BYTES_PER_SAMPLE * samples + BYTES_PER_SEGMENT * segments
    gbs = bytes_alg / kernel_s / 1e9
    tflops = flops / kernel_s / 1e12
    legs = {"hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": int(bytes_alg)},
            "valu": {"achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "flop_per_launch": int(flops),
                     "flop_per_segment": 2 * ray_flop + SHADE_FLOP}}
    bound = max(legs, key=lambda k: legs[k]["frac"])
    return bound, legs


def defaults_measure(ctx):
    """The reference's command-line defaults (-sb 4 -iic 16, Russian roulette) on the C1 and C2
    scenes: first-vertex splitting with indirect clipping (the split path of the fused kernel),
    with the roofline legs of the split kernel."""
    from pg3render_amd import SceneBuild
    out = {}
    for name, wl in (("box_area_light_256px_16spp", 1), ("box_steerable_em_512px_64spp", 2)):
        sb = SceneBuild.workload(wl, seed=1)
        sb.set_params(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0)
        ctx.upload(sb, 1)
        ctx.render(0, 2, 1)
        ctx.synchronize()
        ctx.reset_counters()
        n = sb.spp
        ms = [ctx.render_timed(i * n, n, 1) for i in range(3)]
        seg, smp = ctx.counters()
        ext, shadow = ctx.ray_counters()
        t = sum(ms) / len(ms)
        bound, legs = roofline_legs(sb.desc, smp / 3, seg / 3, t / 1e3)
        out[name] = {"ms_per_frame": round(t, 3), "msamples_per_s": round(smp / 3 / t / 1e3, 1),
                     "msegments_per_s": round(seg / 3 / t / 1e3, 1),
                     "rays_per_s": round((ext + shadow) / 3 / t * 1e3, 1),
                     "roofline": {"bound": bound, "frac": legs[bound]["frac"], "legs": legs,
                                  "traffic": MEASURED_SPLIT_TRAFFIC[name],
                                  "hbm_measured_gbs": round(MEASURED_SPLIT_TRAFFIC[name] / t / 1e6, 1)}}
    return out


def wavefront_measure(ctx, steps=2):
    """The wavefront pipeline (pg3g_set_kernel_mode(WAVEFRONT): ray generation, intersection,
    shading and shadow kernels with path state and ray queues in HBM) on the default workload,
    beside the fused kernel that bench.py times; queue bytes from the fixed SoA record model."""
    from pg3render_amd import SceneBuild, _ffi
    sb = SceneBuild.workload(2, seed=1)
    ctx.set_kernel_mode(_ffi.KERNEL_WAVEFRONT)
    try:
        ctx.upload(sb, 1)
        n = sb.spp
        ctx.render(0, n, 1)
        ctx.synchronize()
        ctx.reset_counters()
        ms = [ctx.render_timed((i + 1) * n, n, 1) for i in range(steps)]
        seg, smp = ctx.counters()
        ext, shadow = ctx.ray_counters()
    finally:
        ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    t = sum(ms) / len(ms) / 1000.0
    qbytes = (BYTES_PER_SAMPLE * smp + BYTES_PER_SEGMENT * seg) / steps
    return {"ms_per_frame": round(t * 1000.0, 3), "msamples_per_s": round(smp / steps / t / 1e6, 1),
            "rays_per_s": round((ext + shadow) / steps / t, 1),
            "queue_model_gbs": round(qbytes / t / 1e9, 1), "queue_model_frac": round(qbytes / t / 1e9 / HBM_PEAK_GBS, 4),
            "hbm_measured_gbs": round(MEASURED_WF_TRAFFIC / t / 1e9, 1),
            "hbm_measured_frac": round(MEASURED_WF_TRAFFIC / t / 1e9 / HBM_PEAK_GBS, 4)}


def host_threads():
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():       # the host share given to this job
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, threads)


def steerable_build_measure():
    """Host build of the steerable tables of the 1024x512 EM of the default workload at the
    reference's default build parameters (levels 4..9, max approximation error 5), on one thread
    and on the host threads (the tables are identical)."""
    from pg3render_amd import SceneBuild, steerable_build
    sb = SceneBuild.workload(2, seed=1)
    sb.desc.em.st_min_subdiv_level = -1
    sb.desc.em.st_max_subdiv_level = -1
    sb.desc.em.st_max_approx_error = -1
    one, _, _ = steerable_build(sb.desc.em, threads=1, leaves=False)
    n = host_threads()
    many, _, _ = steerable_build(sb.desc.em, threads=n, leaves=False)
    return {"em": "1024x512", "levels": "4..9, max error 5 (reference defaults)", "leaves": one["num_leaves"],
            "ms_1_thread": round(one["build_ms"], 1), "threads": n, "ms_threads": round(many["build_ms"], 1),
            "speedup": round(one["build_ms"] / max(many["build_ms"], 1e-6), 2)}


def cpu_baseline(sb, spp, target_s):
    """Times the CPU oracle (all host threads) on a bounded band of image rows at full spp."""
    from oracle import oracle as O
    w, h = sb.resolution
    threads = host_threads()
    orc = O.OracleScene(sb)
    # calibrate on 4 rows, then size the band to about target_s seconds
    probe_rows = 4
    r0 = h // 2 - probe_rows // 2
    _, t = orc.render(0, spp, threads=threads, pixel_begin=r0 * w, pixel_end=(r0 + probe_rows) * w)
    rows = max(probe_rows, min(h, int(probe_rows * target_s / max(t, 1e-3))))
    r0 = max(0, h // 2 - rows // 2)
    _, t = orc.render(0, spp, threads=threads, pixel_begin=r0 * w, pixel_end=(r0 + rows) * w)
    samples = rows * w * spp
    return {"value": samples / t / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"rows {r0}-{r0 + rows - 1} of {h} ({rows * w} pixels) x {spp} spp, {t:.1f} s wall"}


def batch_measure(ctx, tune_candidates):
    """64-scene batch workload: one persistent batched pass vs per-scene rendering, both of the
    same samples (each scene's spp iterations 0..spp-1, same RNG streams), so at equal accuracy.
    Per-scene baselines: one launch per scene rendering all its iterations (`speedup`), the
    RunIteration loop with one launch per iteration (`speedup_vs_run_iteration`), and the same loop
    through pg3g_run_iteration, which launches consecutive iterations in runs of up to 256 (the
    drop-in adapter's path, `queued_run_iteration_msamples_per_s`)."""
    from pg3render_amd import Batch, SceneBuild, _ffi
    scenes = [SceneBuild.workload(3, seed=2024, variant=v) for v in range(64)]
    slots = list(range(1, 65))
    for s, sb in zip(slots, scenes):
        ctx.upload(sb, s)
    spp = [sb.spp for sb in scenes]
    total = sum(sb.resolution[0] * sb.resolution[1] * sb.spp for sb in scenes)
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)

    def plain_pass():
        ctx.synchronize()
        t0 = time.perf_counter()
        for s, n in zip(slots, spp):
            ctx.render(0, n, s)
        ctx.synchronize()
        return time.perf_counter() - t0

    def scene_kernel_ms(n_iter):
        """HIP-event time of each scene's launch rendering n_iter x its spp iterations."""
        out = []
        for s, n in zip(slots, spp):
            out.append(ctx.render_timed(0, n * n_iter, s))
        return out

    def run_iteration_pass():
        # the drop-in renderer as PG3Render's Render() loop drives it: RunIteration, one sample per
        # pixel per call (pg3render.cxx:84-128), each call one launch
        ctx.synchronize()
        t0 = time.perf_counter()
        for s, n in zip(slots, spp):
            for it in range(n):
                ctx.render(it, 1, s)
        ctx.synchronize()
        return time.perf_counter() - t0

    def queued_run_iteration_pass():
        # the same loop through pg3g_run_iteration: the last iteration of each scene launches the
        # run still queued, as the adapter does for the last iteration of a fixed count
        ctx.synchronize()
        t0 = time.perf_counter()
        for s, n in zip(slots, spp):
            for it in range(n):
                ctx.run_iteration(it, s, launch_now=(it == n - 1))
        ctx.synchronize()
        return time.perf_counter() - t0

    plain_pass()                                   # warm-up (code objects, caches)
    plain_s = min(plain_pass() for _ in range(2))
    # What batching can recover at most: a per-scene launch pays its ramp-up and drain (lanes idle
    # while the last paths of the launch finish) once, so a launch of 2 x spp iterations takes
    # t(2n) = 2 t_steady(n) + t_drain and t_steady(n) = t(2n) - t(n); the launch gaps are the
    # wall time of the pass outside its kernels.  A batched pass that removed every gap and drain
    # would take sum t_steady(n).
    k1 = scene_kernel_ms(1)
    k2 = scene_kernel_ms(2)
    steady_ms = sum(max(b - a, 0.0) for a, b in zip(k1, k2))
    ceiling = {"plain_wall_ms": round(plain_s * 1e3, 3), "plain_kernels_ms": round(sum(k1), 3),
               "launch_gaps_ms": round(plain_s * 1e3 - sum(k1), 3),
               "drain_ms": round(sum(2 * a - b for a, b in zip(k1, k2)), 3), "steady_ms": round(steady_ms, 3),
               "max_speedup": round(plain_s * 1e3 / steady_ms, 3)}
    run_it_s = run_iteration_pass()
    queued_s = queued_run_iteration_pass()
    batch = Batch(ctx, slots, spp)
    t0 = time.perf_counter()
    batch.autotune(tune_candidates)
    tune_s = time.perf_counter() - t0
    batch.render(1)                                # warm-up of the captured graph
    ctx.synchronize()
    walls, gpu_ms = [], []
    for _ in range(2):
        t0 = time.perf_counter()
        gpu_ms.append(batch.render_timed(1))
        walls.append(time.perf_counter() - t0)
    batched_s = min(walls)
    cfg = batch.config()
    batch.close()
    return {"workload": "64 seeded box scenes (2 area-light triangles + 256x128 sky/sun EM, steerable), "
                        "side 64..512, spp in {4,16,64,256}, maxpl 8", "samples_per_pass": total,
            "plain_per_scene_msamples_per_s": round(total / plain_s / 1e6, 3),
            "run_iteration_msamples_per_s": round(total / run_it_s / 1e6, 3),
            "queued_run_iteration_msamples_per_s": round(total / queued_s / 1e6, 3),
            "batched_msamples_per_s": round(total / batched_s / 1e6, 3),
            "speedup": round(plain_s / batched_s, 3), "speedup_vs_run_iteration": round(run_it_s / batched_s, 3),
            "batched_gpu_ms_per_pass": round(min(gpu_ms), 3),
            "plain_ms_per_pass": round(plain_s * 1e3, 3), "autotuned_config": cfg, "autotune_s": round(tune_s, 3),
            "ceiling": ceiling}


def diff_measure(ctx):
    """Differentiation workload: box with two GGX-conductor spheres (tint + roughness) and a
    128x64 EM, 256x256, 32 spp; forward (frozen-sampling render) vs path-replay adjoint of an L2
    loss, both timed with HIP events on the library stream."""
    import numpy as np
    from pg3render_amd import SceneBuild
    from pg3render_amd.diff import DiffScene
    sb = SceneBuild.workload(4, seed=2024)
    ctx.upload(sb, 70)
    ds = DiffScene(ctx, sb, [3, 4], slot=70)
    spp = sb.spp
    theta, em = ds.get_params()
    target = ds.render(1 << 20, spp)
    ds.set_params(theta * 1.1, em * 0.9, refresh_sampling=True)
    img = ds.render(0, spp)
    dl = 2.0 * (img - target) / img.size
    ds.adjoint(0, spp, dl)                            # warm-up
    fwd = min(ds.render(0, spp, timed=True)[1] for _ in range(3))
    # the forward keeps each sample's radiance, so the adjoint of the same samples needs one replay
    adj = min(ds.adjoint(0, spp, dl, timed=True)[2] for _ in range(3))
    # standalone adjoint (no stored radiance: the adjoint runs its own forward pass first)
    adj2 = []
    for _ in range(2):
        ds.set_params(theta * 1.1, em * 0.9, refresh_sampling=True)      # drops the stored radiance
        adj2.append(ds.adjoint(0, spp, dl, timed=True)[2])
    adj2 = min(adj2)
    w, h = sb.resolution
    return {"workload": "box + 2 GGX-conductor spheres (tint, roughness) + 128x64 EM (simple sampler), "
                        f"{w}x{h}, {spp} spp, maxpl 8; L2 image loss", "parameters": 8 + int(em.size),
            "forward_ms": round(fwd, 3), "adjoint_ms": round(adj, 3), "adjoint_over_forward": round(adj / fwd, 3),
            "adjoint_two_replays_ms": round(adj2, 3), "adjoint_two_replays_over_forward": round(adj2 / fwd, 3),
            "forward_msamples_per_s": round(w * h * spp / fwd / 1e3, 3)}


def copy_probe():
    """Calibration kernel for the HBM counters: copies 512 MiB device-to-device 4 times (torch's
    copy kernel reads and writes exactly 512 MiB per launch)."""
    import torch
    n = 512 * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_launch": n * 4, "launches": 4}))


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the sharding design is in the docs)", file=sys.stderr)
        sys.exit(2)
    from pg3render_amd import Context, SceneBuild, _ffi

    sb = SceneBuild.workload(args.workload, seed=2024)
    spp = sb.spp
    w, h = sb.resolution
    d = sb.desc
    ctx = Context(0)
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL if args.kernel == "megakernel" else _ffi.KERNEL_WAVEFRONT)
    ctx.upload(sb)

    for i in range(args.warmup):
        ctx.render(i * spp, spp)
    ctx.synchronize()
    ctx.clear(0)
    ctx.reset_counters()

    kernel_ms = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        kernel_ms.append(ctx.render_timed(i * spp, spp))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    segments, samples = ctx.counters()
    ext_rays, shadow_rays = ctx.ray_counters()

    ms_per_step = wall * 1000.0 / args.steps
    total_samples = w * h * spp * args.steps
    msps = total_samples / wall / 1e6
    # algorithmic work of one launch (one path segment = one extension ray + its shadow ray),
    # over the dominant kernel's average HIP-event time on the library stream
    kernel_ms_avg = sum(kernel_ms) / len(kernel_ms)
    bound, legs = roofline_legs(d, samples / args.steps, segments / args.steps, kernel_ms_avg / 1000.0)
    rays_per_s = (ext_rays + shadow_rays) / wall
    traffic = MEASURED_TRAFFIC.get((args.workload, args.kernel))
    measured = ({"hbm_measured_gbs": round(traffic / (kernel_ms_avg / 1000.0) / 1e9, 1),
                 "hbm_measured_frac": round(traffic / (kernel_ms_avg / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)}
                if traffic else {})
    insts = MEASURED_VALU_INSTS.get((args.workload, args.kernel))
    if insts:   # what bounds the fused kernel: VALU issue, with every wavefront of a SIMD waiting otherwise
        rate = insts / (kernel_ms_avg / 1000.0)
        measured["valu_issue_measured"] = {"wave_instructions": insts, "achieved": round(rate / 1e9, 1),
                                           "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave-instr/s",
                                           "frac": round(rate / VALU_ISSUE_PEAK, 4)}
    roofline = {"bound": bound, "achieved": legs[bound]["achieved"], "peak": legs[bound]["peak"],
                "unit": legs[bound]["unit"], "frac": legs[bound]["frac"], "traffic": traffic,
                "legs": legs, "kernel_ms_per_step": round(kernel_ms_avg, 3),
                "segments_per_step": segments // args.steps, **measured}
    out = {"metric": "Msamples/s", "value": round(msps, 3), "unit": "Msamples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "box + 1024x512 sky/sun EM (steerable, levels<=4), GGX dielectric + Lambert "
                                  "spheres, pt -sb 1 -iic 0 -maxpl 10", "resolution": [w, h], "spp": spp,
                      "kernel": args.kernel},
           "rays_per_s": round(rays_per_s, 1),
           "rays": {"extension_per_step": ext_rays // args.steps, "shadow_per_step": shadow_rays // args.steps},
           "roofline": roofline}
    if not args.no_batch:
        out["batched"] = batch_measure(ctx, args.batch_tune)
    out["reference_defaults"] = defaults_measure(ctx)
    out["wavefront_pipeline"] = wavefront_measure(ctx)
    if not args.no_diff:
        out["differentiation"] = diff_measure(ctx)
        if args.inverse_steps > 0:
            from pg3render_amd.inverse import inverse_rendering
            inv = inverse_rendering(ctx, steps=args.inverse_steps)
            inv.pop("loss_history")
            out["inverse_rendering"] = inv
    out["steerable_build"] = steerable_build_measure()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sb, spp, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
