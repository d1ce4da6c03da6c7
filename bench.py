"""Benchmark of the diffuse wavefront frame: Grays/s on the 1280x720 frame of the 118,562-triangle terrain scene
(242x242-vertex fBm heightfield + 200 boxes, Lambert albedo 0.8, sun + sky, next-event estimation), compressed BVH8,
256 samples per pixel in flight (one wavefront pass), 8 bounces, FP32 radiance.

A step is one frame: camera rays for every (pixel, sample), then trace+shade and shadow kernels per bounce until the
queues are empty, then the per-pixel reduction. Scene build and upload happen before the timed region; all path state
lives in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-wavefront-pathtracer_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

# sun direction / irradiance and sky of the reference's published diffuse benchmark (its examples/lighting.json)
LIGHTING = ["--sun_to", "0.608749", "-0.42625", "0.669131", "--sun_rgb", "3", "3", "3", "--sun_angle", "0.526",
            "--sky", "0.3", "0.38", "0.5"]


def workload_args(width, height, batch, bounces, tri, variant, stack, blocks_per_cu=0, bounce_waves=8, fuse=1,
                  persistent=0, shadow_refill=0, sort=0, tail=-1, tail_levels=2):
    return ["--gen", "terrain", "--width", str(width), "--height", str(height), "--batch", str(batch), "--spp", str(batch),
            "--bounces", str(bounces), "--variant", variant, "--tri", tri, "--filter", "box", "--stack", str(stack),
            "--blocks_per_cu", str(blocks_per_cu), "--bounce_waves", str(bounce_waves), "--fuse", str(fuse),
            "--persistent", str(persistent), "--shadow_refill", str(shadow_refill), "--sort", str(sort),
            "--tail", str(tail), "--tail_levels", str(tail_levels)] + LIGHTING


def algorithmic_bytes(st, npaths, sort=0):
    """Compulsory HBM bytes of one frame's trace+shade (bounce) kernels, from the frame's ray counters.

    Per camera ray: 16 B radiance write. Per continuation ray traced: 48 B origin/direction/throughput read at its
    queue slot (the path id rides in the origin's w). Per continuation-ray miss: 32 B radiance read-modify-write. Per
    continuation produced: 48 B ray write at its slot (+ 1 B bucket key with --sort 1; with --sort 2, + 1 B for each
    camera ray's key, written whether or not the ray continues). Per shadow ray produced: 48 B shadow record write."""
    cont_traced = st["rays_closest"] - npaths
    cont_miss = st["rays_miss"] - st["camera_miss"]
    keys = st["rays_cont"] if sort == 1 else (npaths if sort == 2 else 0)
    return (16.0 * npaths + 48.0 * cont_traced + 32.0 * cont_miss + 48.0 * st["rays_cont"] + 1.0 * keys +
            48.0 * st["rays_shadow"])


def measured_traffic(kernel_tag):
    """HBM-side bytes per launch of the dominant kernel from the committed rocprof summary (FETCH_SIZE + WRITE_SIZE,
    each divided by its plain-copy calibration factor), or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_bounce_terrain720p_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_per_launch_bytes " + kernel_tag + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_utilisation(kernel_tag):
    """VALU busy, VALU lane utilisation and L2 hit rate (percent) of the dominant kernel from the committed rocprof
    counter summary, or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "k_bounce_terrain720p_counters.md")
    try:
        for line in open(path):
            if line.startswith("| " + kernel_tag):
                c = [x.strip() for x in line.strip().strip("|").split("|")]
                pct = lambda x: float(x.replace("%", "").strip())
                return {"valu_busy_pct": pct(c[3]), "valu_lane_util_pct": pct(c[4]), "l2_hit_pct": pct(c[5]),
                        "source": "profiles/k_bounce_terrain720p_counters.md (rocprofv3 SQ_ACTIVE_INST_VALU, "
                                  "SQ_THREAD_CYCLES_VALU, TCC_HIT/MISS, GRBM_GUI_ACTIVE)"}
    except (OSError, ValueError, IndexError):
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# The plain method on the CPU: 1,024 seeded boxes + ground (12,290 triangles), Lambert albedo 0.7, sun toward
# (0.3, 0.8, 0.5) with irradiance 5 and sky 0.5, pinhole 256x256, 16 spp, 4 bounces, binned-SAH BVH2 (32 bins, leaf <= 4)
PLAIN_CPU_ARGS = ["--gen", "boxes", "--width", "256", "--height", "256", "--spp", "16", "--batch", "16", "--bounces", "4",
                  "--variant", "bvh2", "--bins", "32", "--sun_to", "0.3", "0.8", "0.5", "--sun_rgb", "5", "5", "5",
                  "--sky", "0.5", "0.5", "0.5", "--filter", "box"]


def cpu_baseline(args):
    """The oracle's multi-threaded CPU renderer (kind "port"): the plain CPU configuration in full, and a bounded sample
    (reduced spp) of the headline terrain frame. Threads: the host cores a GPU of this machine is given (16 per GPU on
    the measurement boxes; the visible core count is reported beside it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cwp
    import pyoracle as po
    visible = os.cpu_count() or 1
    threads = max(1, min(args.cpu_threads, visible))
    s = cwp.Setup(PLAIN_CPU_ARGS)
    sc, _ = po.scene_from_setup(s)
    t0 = time.perf_counter()
    _, st = sc.render(s.params(), 16, batch=16, bounces=4, threads=threads)
    dt1 = time.perf_counter() - t0
    rays1 = float(st[0] + st[1])
    spp = args.cpu_spp
    s = cwp.Setup(workload_args(args.width, args.height, spp, args.bounces, args.tri, args.variant, args.stack))
    sc, _ = po.scene_from_setup(s)
    t0 = time.perf_counter()
    _, st = sc.render(s.params(), spp, batch=spp, bounces=args.bounces, threads=threads)
    dt = time.perf_counter() - t0
    rays = float(st[0] + st[1])
    return {"value": rays / dt / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
            "sample": "oracle CPU renderer, same scene/camera/lighting, %dx%d, %d spp, %d bounces, %d threads: "
                      "%.0f rays in %.1f s" % (args.width, args.height, spp, args.bounces, threads, rays, dt),
            "cpu_model": cpu_model(), "cpus_visible": visible,
            "plain_cpu_config": {"value": rays1 / dt1 / 1e9, "unit": "Grays/s", "frame_s": round(dt1, 3), "rays": rays1,
                                 "workload": "1,024 boxes + ground (12,290 tris), 256x256, 16 spp, 4 bounces, BVH2 "
                                             "32-bin SAH leaf <= 4, albedo 0.7, sun 5 + sky 0.5 (in full)"}}


def calibrate_copy(mib):
    """A plain device copy of known size, to scale rocprof's FETCH_SIZE / WRITE_SIZE (run in its own process: it uses
    torch, whose bundled HIP runtime must not share a process with libcwp.so)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4,
                      "copies": 3}), flush=True)


def bench_instances(args, frames=None, quiet=False):
    """Two-level scene: 5 meshes (1.3k-200k triangles) + ground, 1,000 instances with rigid transforms redrawn every
    frame; per frame the TLAS is rebuilt on the GPU (H-PLOC over the instance boxes, then for BVH8 BLASes the GPU
    collapse to a compressed BVH8 unless --tlas bvh2) and the frame rendered at 1280x720, 64 spp in flight. Reports
    Grays/s over the frames and the TLAS build time per frame."""
    import cwp
    spp = 64
    s = cwp.Setup(["--gen", "instances", "--width", "1280", "--height", "720", "--batch", str(spp), "--spp", str(spp),
                   "--bounces", str(args.bounces), "--variant", args.variant, "--tri", args.tri, "--filter", "box",
                   "--stack", str(args.stack), "--morton_bits", str(args.tlas_morton_bits), "--tlas", args.tlas,
                   "--tail", str(args.tail)] + LIGHTING)
    r = cwp.Renderer(s)
    info = s.info
    print("bench: %d meshes, %d instances, %d BLAS nodes" % (info["num_meshes"], info["num_instances"], info["num_nodes"]),
          file=sys.stderr, flush=True)
    r.render(spp)   # warm-up
    tlas, frame_ms, rays = [], [], 0.0
    frames = args.frames if frames is None else frames
    for f in range(frames):
        tlas.append(r.set_frame(f))
        st = r.render(spp)
        frame_ms.append(st["t_frame_ms"])
        rays += st["rays_closest"] + st["rays_shadow"]
        if f % 10 == 0:
            print("bench: frame %d: TLAS %.3f ms, frame %.2f ms" % (f, tlas[-1], frame_ms[-1]), file=sys.stderr, flush=True)
    tot = sum(frame_ms)
    out = {"metric": "Grays/s and GPU TLAS rebuild ms per frame, 1,000-instance two-level scene", "value": round(rays / (tot * 1e-3) / 1e9, 4),
           "unit": "Grays/s", "n_gpus": 1, "steps": frames, "warmup": 1, "ms_per_step": round(tot / frames, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "instances-1000-1280x720-64spp-depth%d" % args.bounces, "meshes": info["num_meshes"],
                      "instances": info["num_instances"], "bvh": args.variant, "tri": args.tri, "frames": frames,
                      "tlas_morton_bits": args.tlas_morton_bits, "tlas": args.tlas if args.variant == "bvh8" else "bvh2"},
           "tlas_ms_mean": round(float(np.mean(tlas)), 4), "tlas_ms_p50": round(float(np.percentile(tlas, 50)), 4),
           "tlas_ms_max": round(float(np.max(tlas)), 4)}
    r.close()
    if not quiet:
        print(json.dumps(out), flush=True)
    return out


def bench_hploc(args, quiet=False):
    """H-PLOC GPU build over the triangle boxes of heightfield meshes the size of the reference's published builds
    (2.02 M and 2.80 M triangles): device time of one build with the boxes resident (hipEvents, median of 3 builds), and
    the host post-processing to a CWBVH layout (SAH leaf collapse + 8-wide collapse + triangle packing)."""
    import cwp
    rows = []
    for g in (1005, 1184):
        rng = np.random.default_rng(g)
        xs = np.linspace(0.0, 1.0, g, dtype=np.float32)
        X, Y = np.meshgrid(xs, xs, indexing="xy")
        Z = (0.05 * rng.standard_normal((g, g))).astype(np.float32)
        v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
        i = np.arange(g * g, dtype=np.int32).reshape(g, g)
        a, b, c, d = i[:-1, :-1].ravel(), i[:-1, 1:].ravel(), i[1:, :-1].ravel(), i[1:, 1:].ravel()
        t = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)]).astype(np.int32)
        corners = v[t]
        lo, hi = corners.min(1), corners.max(1)
        ms = []
        for _ in range(3):
            nodes, root, gpu_ms = cwp.hploc_build(lo, hi, args.hploc_morton_bits)
            ms.append(gpu_ms)
        t0 = time.perf_counter()
        bv = cwp.bvh_from_hploc(v, t, nodes, root, 8)
        post = (time.perf_counter() - t0) * 1e3
        cms = []
        for _ in range(3):
            wide, order, depth, c_ms = cwp.cwbvh_from_hploc_gpu(nodes, len(t))
            cms.append(c_ms)
        same = bool(np.array_equal(wide.view(np.uint32), bv["nodes"].reshape(-1, 20).view(np.uint32)) and np.array_equal(order, bv["order"]))
        rows.append({"tris": int(len(t)), "gpu_build_ms": round(float(np.median(ms)), 3), "gpu_build_ms_all": [round(x, 3) for x in ms],
                     "gpu_collapse_ms": round(float(np.median(cms)), 3), "gpu_collapse_equals_host": same,
                     "host_post_ms": round(post, 1), "cwbvh_nodes": int(len(bv["nodes"]) // 20 if bv["nodes"].ndim == 1 else len(bv["nodes"])),
                     "wide_depth": int(depth)})
        print("bench: H-PLOC %d tris: %.3f ms on the GPU, CWBVH collapse %.3f ms on the GPU (%s the host's), %.0f ms host "
              "post-processing" % (len(t), rows[-1]["gpu_build_ms"], rows[-1]["gpu_collapse_ms"], "equal to" if same else "DIFFERENT from", post),
              file=sys.stderr, flush=True)
    out = {"metric": "H-PLOC GPU BVH build ms (boxes resident); CWBVH collapse ms on the GPU and on the host", "value": rows[-1]["gpu_build_ms"],
           "unit": "ms", "n_gpus": 1, "steps": 3, "warmup": 1, "ms_per_step": rows[-1]["gpu_build_ms"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "hploc-heightfield-2.02M-2.80M", "builds": rows, "morton_bits": args.hploc_morton_bits},
           "reference_published_h800_ms": {"2.02M": 1.29, "2.80M": 1.74, "host_post_ms": "226-404"}}
    if not quiet:
        print(json.dumps(out), flush=True)
    return out


JOB_RES = (128, 256, 512, 1024)
JOB_SPP = (4, 16, 64)


def bench_reference_config(args):
    """The same frame in the configuration of the reference's published diffuse benchmark (4 bounces, BVH2 and the
    compressed BVH8, Woop triangles, fused kernel, 256 spp in flight at 1280x720): Grays/s of one frame after a
    warm-up, for a like-for-like reading next to the reference's numbers."""
    import cwp
    out = {"bounces": 4, "spp": 256, "width": 1280, "height": 720}
    for variant in ("bvh2", "bvh8"):
        s = cwp.Setup(workload_args(1280, 720, 256, 4, "woop", variant, args.stack, sort=args.sort,
                                    tail=args.tail if variant == "bvh8" else -1))
        r = cwp.Renderer(s)
        r.render(256)
        st = r.render(256)
        out[variant + "_grays"] = round((st["rays_closest"] + st["rays_shadow"]) / (st["t_frame_ms"] * 1e-3) / 1e9, 4)
        out[variant + "_ms"] = round(st["t_frame_ms"], 3)
        r.close()
        print("bench: reference config %s: %.3f Grays/s" % (variant, out[variant + "_grays"]), file=sys.stderr, flush=True)
    return out


def bench_variant_frame(args, width, height, tri):
    """One frame of the terrain workload at another resolution or triangle test (8 bounces, BVH8, 256 spp in flight),
    after a warm-up frame: Grays/s."""
    import cwp
    s = cwp.Setup(workload_args(width, height, 256, args.bounces, tri, "bvh8", args.stack, sort=args.sort, tail=args.tail))
    r = cwp.Renderer(s)
    r.render(256)
    st = r.render(256)
    r.close()
    g = (st["rays_closest"] + st["rays_shadow"]) / (st["t_frame_ms"] * 1e-3) / 1e9
    print("bench: %dx%d %s: %.3f Grays/s" % (width, height, tri, g), file=sys.stderr, flush=True)
    return {"grays": round(g, 4), "ms": round(st["t_frame_ms"], 3), "width": width, "height": height, "tri": tri,
            "bvh": "bvh8", "bounces": args.bounces, "spp": 256}


def bench_full_frame(args, textured):
    """One frame of the terrain with the full integrator (OpenPBR-subset materials on the boxes, 2048 x 1024
    log-normal environment, MIS), 1280x720, 64 spp in flight, 8 bounces, graph replay; with `textured`, generated
    1024 x 1024 base-colour, normal and roughness / metalness maps on every material: Grays/s after a warm-up frame."""
    import cwp
    a = workload_args(1280, 720, 64, args.bounces, "woop", "bvh8", args.stack) + ["--gen_materials", "5", "--gen_env",
                                                                                  "2048,1024,2,1"]
    if textured:
        a += ["--gen_textures", "1024"]
    s = cwp.Setup(a)
    r = cwp.Renderer(s)
    r.render(64)
    st = r.render(64)
    r.close()
    g = (st["rays_closest"] + st["rays_shadow"]) / (st["t_frame_ms"] * 1e-3) / 1e9
    print("bench: full integrator%s: %.3f Grays/s" % (" + textures" if textured else "", g), file=sys.stderr, flush=True)
    return {"grays": round(g, 4), "ms": round(st["t_frame_ms"], 3), "width": 1280, "height": 720, "spp": 64,
            "bounces": args.bounces, "textures": "3 maps of 1024^2 per material" if textured else "none"}


def job_stream_scenes(n_scenes, seed, lo_tris=10_000, hi_tris=2_000_000, full=False):
    """Seeded scenes with triangle counts log-uniform in [lo_tris, hi_tris]: fBm heightfields of (G-1)^2 * 2
    triangles, each with its own terrain seed. With `full`, 64 boxes on each terrain carry OpenPBR-subset materials
    (conductors, rough dielectrics, glossy-diffuse, emitters) and each scene has a 2048 x 1024 log-normal (sigma 2)
    environment map. Returns the setup arguments of each scene."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        t = float(np.exp(rng.uniform(np.log(lo_tris), np.log(hi_tris))))
        g = int(round(np.sqrt(t / 2.0))) + 1
        a = ["--gen", "terrain", "--gen_grid", str(g), "--gen_seed", str(1 + k), "--width", "512", "--height", "512",
             "--bounces", "6", "--filter", "box"] + LIGHTING
        if full:
            a += ["--gen_boxes", "64", "--gen_materials", str(1 + k), "--gen_env", "2048,1024,2,%d" % (1 + k)]
        else:
            a += ["--gen_boxes", "0"]
        out.append(a)
    return out


def job_stream(n_jobs, n_scenes, seed, res=JOB_RES, spp=JOB_SPP):
    """(scene, width, height, spp, bounces, seed) of each job: scene uniform, square resolution uniform over
    {128, 256, 512, 1024}, spp uniform over {4, 16, 64}, depth 6."""
    rng = np.random.default_rng(seed + 1)
    jobs = []
    for j in range(n_jobs):
        r = int(rng.choice(res))
        jobs.append((int(rng.integers(n_scenes)), r, r, int(rng.choice(spp)), 6, 1000 + j))
    return jobs


def make_setups(arg_lists):
    """Host scene setups built in parallel threads (the C ABI releases the interpreter lock)."""
    import cwp
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(8, len(arg_lists))) as ex:
        return list(ex.map(cwp.Setup, arg_lists))


def run_stream(srv, jobs, batched, arrivals=None):
    srv.clear_jobs()
    ids = [srv.submit(*j[:5], seed=j[5], arrival_s=(arrivals[k] if arrivals is not None else 0.0))
           for k, j in enumerate(jobs)]
    st = srv.run(batched=batched)
    return st, ids


def warm_up(srv, jobs):
    """Uploads every scene of the stream and runs a few jobs (kernels loaded, graphs captured)."""
    srv.clear_jobs()
    for j in jobs:
        srv.submit(*j[:5], seed=j[5])
    srv.preload()
    run_stream(srv, jobs[:8], True)


def bench_jobs(args):
    """Mixed job stream: 64 render jobs over 8 seeded scenes (10k-2M triangles), resolutions 128^2-1024^2, 4-64 spp,
    depth 6, all submitted at t = 0. The plain mode renders one job at a time (each job's passes alone on the GPU);
    the batched mode runs every admitted job's paths in one trace+shade and one shadow launch per bounce, from a
    captured graph, under the configuration the autotuner picked. Reports both rays/s, their ratio and job latency."""
    import cwp
    scene_args = job_stream_scenes(args.job_scenes, args.job_seed, full=bool(args.full))
    t0 = time.perf_counter()
    setups = make_setups(scene_args)
    tris = [s.info["num_tris"] for s in setups]
    print("bench: %d scenes (%s tris) set up in %.1f s" % (len(setups), ",".join(str(t) for t in tris),
                                                         time.perf_counter() - t0), file=sys.stderr, flush=True)
    jobs = job_stream(args.jobs, len(setups), args.job_seed, args.job_res, args.job_spp)
    srv = cwp.Server(variant=8, bounces=6, full=int(args.full), tail_lanes=args.tail, tail_levels=args.tail_levels)
    for s in setups:
        srv.add_scene(s)
    warm_up(srv, jobs)
    plain, _ = run_stream(srv, jobs, False)
    print("bench: plain %.3f Grays/s (%.2f s)" % (plain["rays_per_s"] / 1e9, plain["wall_s"]), file=sys.stderr, flush=True)
    untuned, _ = run_stream(srv, jobs, True)
    print("bench: batched (untuned) %.3f Grays/s" % (untuned["rays_per_s"] / 1e9), file=sys.stderr, flush=True)
    trials, chosen = [], srv.get_options()
    if args.autotune:
        srv.clear_jobs()
        for j in jobs[::4]:
            srv.submit(*j[:5], seed=j[5])
        trials, chosen = srv.autotune(compacts=(0, 1))
        print("bench: autotuner chose bvh%d stack %d rays/lane %d compaction %s" % (
            chosen["variant"], chosen["stack_lds"], chosen["rays_per_lane"], ("wave", "lane")[chosen["compact"]]),
            file=sys.stderr, flush=True)
        warm_up(srv, jobs)   # scenes resident in the chosen layout
    batched, ids = run_stream(srv, jobs, True)
    print("bench: batched %.3f Grays/s (%.2f s)" % (batched["rays_per_s"] / 1e9, batched["wall_s"]), file=sys.stderr, flush=True)
    srv.close()
    return {"plain_grays": round(plain["rays_per_s"] / 1e9, 4), "batched_grays": round(batched["rays_per_s"] / 1e9, 4),
            "batched_untuned_grays": round(untuned["rays_per_s"] / 1e9, 4),
            "speedup": round(batched["rays_per_s"] / plain["rays_per_s"], 3),
            "rays": batched["rays"], "plain_wall_s": round(plain["wall_s"], 3), "batched_wall_s": round(batched["wall_s"], 3),
            "latency_s": {"plain_p50": round(plain["latency_p50_s"], 4), "plain_p99": round(plain["latency_p99_s"], 4),
                          "batched_p50": round(batched["latency_p50_s"], 4), "batched_p99": round(batched["latency_p99_s"], 4)},
            "passes": {"plain": plain["passes"], "batched": batched["passes"]},
            "chosen": {"bvh": chosen["variant"], "stack_lds": chosen["stack_lds"], "rays_per_lane": chosen["rays_per_lane"],
                       "compaction": ("wave", "lane")[chosen["compact"]], "tail_lanes": chosen["tail_lanes"],
                       "tail_levels": chosen["tail_levels"]},
            "trials": [{"bvh": t["variant"], "stack": t["stack_lds"], "rpl": t["rays_per_lane"],
                        "compaction": ("wave", "lane")[t["compact"]], "grays": round(t["rays_per_s"] / 1e9, 4)}
                       for t in trials],
            "workload": {"jobs": len(jobs), "scenes": len(setups), "tris": tris, "resolutions": list(args.job_res),
                         "spp": list(args.job_spp), "depth": 6, "arrivals": "all at t=0", "seed": args.job_seed}}


def bench_serve(args):
    """Serving: the same job stream arriving as a Poisson process of rate lambda (jobs/s), continuous batching
    (batched) against one job at a time (plain); rays/s and p50/p99 latency (finish - arrival) per rate."""
    import cwp
    setups = make_setups(job_stream_scenes(args.job_scenes, args.job_seed, full=bool(args.full)))
    jobs = job_stream(args.jobs, len(setups), args.job_seed, args.job_res, args.job_spp)
    srv = cwp.Server(variant=8, bounces=6, full=int(args.full), tail_lanes=args.tail, tail_levels=args.tail_levels)
    for s in setups:
        srv.add_scene(s)
    warm_up(srv, jobs)
    tail = srv.get_options()["tail_lanes"]
    print("bench: serving with hand-off threshold %d lanes" % tail, file=sys.stderr, flush=True)
    rng = np.random.default_rng(args.job_seed + 2)
    gaps = rng.exponential(1.0, len(jobs))
    out = []
    for lam in args.rates:
        arr = np.cumsum(gaps) / lam
        arr -= arr[0]
        row = {"lambda": lam}
        for mode, b in (("batched", True), ("plain", False)):
            st, _ = run_stream(srv, jobs, b, arrivals=arr)
            row[mode] = {"grays": round(st["rays_per_s"] / 1e9, 4), "p50_s": round(st["latency_p50_s"], 4),
                         "p99_s": round(st["latency_p99_s"], 4), "wall_s": round(st["wall_s"], 3)}
        print("bench: lambda %g: batched %.3f Grays/s p50 %.3f s p99 %.3f s | plain %.3f Grays/s p50 %.3f s p99 %.3f s"
              % (lam, row["batched"]["grays"], row["batched"]["p50_s"], row["batched"]["p99_s"], row["plain"]["grays"],
                 row["plain"]["p50_s"], row["plain"]["p99_s"]), file=sys.stderr, flush=True)
        out.append(row)
    srv.close()
    return out, tail


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (rocprof traffic calibration) and exit")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--bounces", type=int, default=8)
    ap.add_argument("--tri", default="woop")
    ap.add_argument("--variant", default="bvh8")
    ap.add_argument("--stack", type=int, default=8)
    ap.add_argument("--bounce-waves", type=int, default=8)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--fuse", type=int, default=1, help="1: fused trace+shade kernel; 0: split trace and shade kernels")
    ap.add_argument("--persistent", type=int, default=0, help="lanes refilled with new rays mid-traversal: 1 split "
                    "trace / shade kernels, 2 the fused kernel")
    ap.add_argument("--shadow-refill", type=int, default=0, help="1: shadow kernel refills idle lanes mid-traversal")
    ap.add_argument("--sort", type=int, default=2, help="continuation order: 0 as appended; 1 every bounce's grouped by "
                    "direction bucket per partition; 2 the first bounce's grouped by elevation within each pixel")
    ap.add_argument("--tail", type=int, default=-1, help="straggler hand-off: a trace+shade or shadow wave hands its "
                    "tracing lanes to a tail kernel once fewer than this many remain (0 off, -1 the library's automatic "
                    "setting); the frame workloads, the job stream and serving")
    ap.add_argument("--tail-levels", type=int, default=2, help="hand-off levels: 1, or 2 (the tail kernels hand their "
                    "own stragglers on)")
    ap.add_argument("--cpu-spp", type=int, default=40)
    ap.add_argument("--workload", default="terrain", choices=["terrain", "instances", "jobs", "serve", "hploc"],
                    help="terrain: the diffuse frame (headline), followed by the mixed job stream unless --no-jobs; "
                         "instances: the two-level scene with the TLAS rebuilt on the GPU every frame (1,000 instances "
                         "of 5 meshes, 720p, 64 spp, 60 frames); jobs: the mixed job stream only, batched vs plain; "
                         "serve: the job stream with Poisson arrivals at each --rates; hploc: GPU BVH build times")
    ap.add_argument("--jobs", type=int, default=64, help="job stream: jobs")
    ap.add_argument("--job-scenes", type=int, default=8, help="job stream: scenes (10k-2M triangles, log-uniform)")
    ap.add_argument("--job-seed", type=int, default=1)
    ap.add_argument("--autotune", type=int, default=1, help="job stream: run the autotuner before the batched run")
    ap.add_argument("--rates", default="10,50,100,200,500", help="serve: Poisson arrival rates (jobs/s)")
    ap.add_argument("--job-res", default="128,256,512,1024", help="job stream: square resolutions (uniform)")
    ap.add_argument("--job-spp", default="4,16,64", help="job stream: samples per pixel (uniform)")
    ap.add_argument("--no-jobs", action="store_true", help="terrain workload without the job stream")
    ap.add_argument("--no-reference-config", action="store_true",
                    help="terrain workload without the frames in the reference's published configuration")
    ap.add_argument("--full", type=int, default=0, help="job stream / serving: the full material integrator "
                    "(OpenPBR-subset materials, emitters, 2048 x 1024 log-normal environment, MIS)")
    ap.add_argument("--frames", type=int, default=60, help="instances workload: animation frames")
    ap.add_argument("--tlas-morton-bits", type=int, default=30, help="two-level scene: Morton keys of the per-frame TLAS "
                    "(30: 10 bits per axis, as the two-level configuration asks; 0: the reference's adaptive keys)")
    ap.add_argument("--tlas", default="bvh8", choices=["bvh2", "bvh8"],
                    help="two-level scene with BVH8 BLASes: the per-frame TLAS as a BVH2 or a compressed BVH8")
    ap.add_argument("--hploc-morton-bits", type=int, default=0, help="hploc workload: Morton keys (0 adaptive, 30)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="CPU baseline threads (the host cores of one GPU)")
    ap.add_argument("--no-extras", action="store_true", help="terrain workload: only the headline frame, the reference "
                    "configuration and the job stream (no resolution / triangle-test variants, serving sweep, H-PLOC "
                    "builds or two-level frames)")
    ap.add_argument("--serve-jobs", type=int, default=32, help="terrain workload: jobs per rate of the serving sweep")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.calibrate_copy_mib > 0:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the renderer runs on a single MI355X (multi-GPU runs are not part of "
              "this build)", file=sys.stderr)
        sys.exit(2)

    args.rates = [float(x) for x in args.rates.split(",") if x]
    args.job_res = tuple(int(x) for x in args.job_res.split(",") if x)
    args.job_spp = tuple(int(x) for x in args.job_spp.split(",") if x)
    import cwp
    if args.workload == "instances":
        bench_instances(args)
        return
    if args.workload == "hploc":
        bench_hploc(args)
        return
    if args.workload in ("jobs", "serve"):
        j = bench_jobs(args)
        out = {"metric": "Grays/s of the mixed job stream, batched vs plain; p50/p99 job latency", "value": j["batched_grays"],
               "unit": "Grays/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(j["batched_wall_s"] * 1e3, 3),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": "job-stream-%d-jobs-%d-scenes-depth6%s" % (args.jobs, args.job_scenes,
                                                                               "-openpbr-env-mis" if args.full else "")},
               "jobs": j}
        if args.workload == "serve":
            out["serve"], out["tail_lanes"] = bench_serve(args)
        print(json.dumps(out), flush=True)
        return
    if args.sort == 2 and (args.fuse == 0 or args.persistent):
        args.sort = 0   # the first-bounce sort runs with the fused kernel only
    if args.tail > 0 and (args.fuse == 0 or args.persistent or args.variant != "bvh8"):
        args.tail = 0   # the hand-off resumes BVH8 traversals of the fused kernel only
    print("bench: setting up scene", file=sys.stderr, flush=True)
    s = cwp.Setup(workload_args(args.width, args.height, args.batch, args.bounces, args.tri, args.variant, args.stack,
                                args.blocks_per_cu, args.bounce_waves, args.fuse, args.persistent, args.shadow_refill,
                                args.sort, args.tail, args.tail_levels))
    r = cwp.Renderer(s)
    info = s.info
    npaths = args.width * args.height * args.batch
    print("bench: %d tris, %d nodes, grid %d workgroups, %.0f MB" % (info["num_tris"], info["num_nodes"], r.grid_blocks,
                                                                    r.vram_mb), file=sys.stderr, flush=True)
    for w in range(args.warmup):
        st = r.render(args.batch)
        print("bench: warmup %d: %.1f ms" % (w, st["t_frame_ms"]), file=sys.stderr, flush=True)
    frames = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        frames.append(r.render(args.batch))
    wall = time.perf_counter() - t0
    for k, st in enumerate(frames):
        print("bench: step %d: %.2f ms, %.0f rays" % (k, st["t_frame_ms"], st["rays_closest"] + st["rays_shadow"]),
              file=sys.stderr, flush=True)
    rays = sum(f["rays_closest"] + f["rays_shadow"] for f in frames)
    frame_ms = sum(f["t_frame_ms"] for f in frames)
    grays = rays / (frame_ms * 1e-3) / 1e9
    bounce_ms = sum(f["t_bounce_ms"] for f in frames)
    launches = sum(f["bounce_launches"] for f in frames)
    bytes_bounce = sum(algorithmic_bytes(f, npaths, args.sort) for f in frames)
    achieved = bytes_bounce / (bounce_ms * 1e-3) / 1e9
    last = frames[-1]
    out = {
        "metric": "Grays/s on 720p 117k-tri diffuse frame; p50/p99 job latency, batched vs plain",
        "value": round(grays, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(frame_ms / args.steps, 3),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "terrain-118k-%dx%d-%dspp-in-flight-depth%d" % (args.width, args.height, args.batch, args.bounces),
                   "scene": "242x242 fBm heightfield + 200 boxes (%d tris), albedo 0.8, sun 3 W/m2 + sky" % info["num_tris"],
                   "width": args.width, "height": args.height, "spp": args.batch, "batch": args.batch,
                   "bounces": args.bounces, "bvh": args.variant, "tri": args.tri, "filter": "box",
                   "stack_lds": args.stack, "bounce_waves": args.bounce_waves, "bvh_nodes": info["num_nodes"],
                   "kernels": {1: "persistent-split", 2: "persistent-fused"}.get(args.persistent, "fused" if args.fuse else "split"),
                   "shadow_refill": args.shadow_refill, "sort": args.sort, "tail": r.frame_json(last)["tail"]},
        "rays_per_frame": last["rays_closest"] + last["rays_shadow"],
        "rays_closest": last["rays_closest"], "rays_shadow": last["rays_shadow"],
        "t_bounce_ms": round(last["t_bounce_ms"], 3), "t_shade_ms": round(last["t_shade_ms"], 3),
        "t_shadow_ms": round(last["t_shadow_ms"], 3), "t_sort_ms": round(last["t_sort_ms"], 3),
        "t_tail_ms": round(last["t_tail_ms"], 3),
        "roofline": {"bound": "hbm", "kernel": ("trace+shade step per depth: k_bounce_first (depth 0, first-bounce sort fused), "
                                                "k_bounce with straggler hand-off + k_bounce_tail (depths 1-9)")
                     if args.fuse and not args.persistent and args.sort == 2 and r.frame_json(last)["tail"] > 0
                     else "k_bounce (fused camera-ray / trace / shade)" if args.fuse and not args.persistent
                     else ("k_bounce_refill (fused, refilled lanes)" if args.persistent == 2 else "k_trace_paths + k_shade (split trace / shade)"),
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic("k_bounce") if args.fuse and not args.persistent else None,
                     "traffic_source": "profiles/k_bounce_terrain720p_traffic.md (rocprofv3 FETCH_SIZE / 0.5 + WRITE_SIZE / 1.0)",
                     "bytes_per_frame": bytes_bounce / args.steps,
                     "utilisation": measured_utilisation("k_bounce family") if args.fuse and not args.persistent
                     and args.variant == "bvh8" else None,
                     "avg_launch_ms": round(bounce_ms / max(1, launches), 4)},
        "latency": None,
        "cpu_baseline": None,
    }
    r.close()
    if not args.no_reference_config:
        out["reference_config"] = bench_reference_config(args)
    if not args.no_extras:
        out["variants"] = {"watertight_1280x720": bench_variant_frame(args, 1280, 720, "watertight"),
                           "woop_512x512": bench_variant_frame(args, 512, 512, "woop"),
                           "full_1280x720": bench_full_frame(args, False),
                           "full_textured_1280x720": bench_full_frame(args, True)}
    if not args.no_jobs:
        print("bench: mixed job stream", file=sys.stderr, flush=True)
        j = bench_jobs(args)
        out["jobs"] = j
        out["latency"] = dict(j["latency_s"], unit="s", speedup_batched_vs_plain=j["speedup"])
    if not args.no_extras:
        print("bench: serving sweep (full integrator)", file=sys.stderr, flush=True)
        sargs = argparse.Namespace(**vars(args))
        sargs.full, sargs.jobs = 1, args.serve_jobs
        rates, tail = bench_serve(sargs)
        out["serve"] = {"integrator": "full (OpenPBR subset, emitters, 2048x1024 log-normal environment, MIS)",
                        "jobs_per_rate": sargs.jobs, "tail_lanes": tail, "rates": rates}
        print("bench: H-PLOC builds", file=sys.stderr, flush=True)
        h = bench_hploc(args, quiet=True)
        out["hploc"] = {"builds": h["config"]["builds"], "morton_bits": args.hploc_morton_bits,
                        "reference_published_h800_ms": h["reference_published_h800_ms"]}
        print("bench: two-level scene", file=sys.stderr, flush=True)
        inst = bench_instances(args, frames=10, quiet=True)
        out["instances"] = {"grays": inst["value"], "ms_per_frame": inst["ms_per_step"], "frames": 10,
                            "tlas_ms_mean": inst["tlas_ms_mean"], "tlas_ms_p50": inst["tlas_ms_p50"],
                            "config": inst["config"]}
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
