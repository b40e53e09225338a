"""Benchmark: incoherent closest-hit Mrays/s on one MI355X.

Workload (a "step"): one closest-hit launch over the intersection
microbench's 4,194,304 incoherent rays (origins uniform in [-1.2,1.2]^3,
directions uniform on S^2) against 1,148,576 triangles (a 1024x512 UV
sphere displaced by Perlin noise, amplitude 0.05, plus 100,000 small random
triangles), FP32, tmax = inf.  Rays, BVH and hit buffers are resident in HBM
before the timed region.

Prints one JSON line (rank 0).  Extra fields: primary-ray Mrays/s, the
roofline of the traversal kernel, and the CPU baseline (the oracle's binary
SAH BVH traversal on host cores) timed on the same ray set.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "closest_incoherent_microbench.md")


def profiled_traffic():
    """HBM bytes per incoherent launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a copy kernel), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None
GUIDING_REF = os.path.join(ROOT, "tests", "golden", "guiding512_ref.npz")
HEADLINE_METRIC = "incoherent Mrays/s (1 GPU); equal-relMSE time, guided vs plain GPU tracer"


def host_info():
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process
    may use, the CPU share the environment grants (OMP_NUM_THREADS: one GPU's
    share of a shared box) and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS")
    share = int(share) if share and share.isdigit() and int(share) > 0 else None
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_share": share, "model": model}


def rel_mse(img, ref):
    """mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2)"""
    img = img.astype(np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def guiding_equal_error(args):
    """Wall time for the plain and the guided GPU tracer to reach relMSE
    args.guiding_target on the slit-lit glossy room (64 GGX spheres, depth 10)
    against a plain-tracer reference of args.guiding_ref_spp samples per pixel.
    Both tracers accumulate chunks of samples into the film; only the
    kp_render_accumulate wall time counts (network training included for the
    guided one), the film reads and error evaluations between chunks do not."""
    from kazen_proto_amd import native, scenes
    res = args.guiding_res
    s = scenes.guiding_scene(width=res, height=res, spp=1, max_depth=10)
    c = native.Context(0)
    c.load_scene(s)
    t0 = time.time()
    fp64 = None
    if res == 512 and os.path.exists(GUIDING_REF) and not args.guiding_gpu_ref:
        z = np.load(GUIDING_REF)
        fp64 = z["ref"].astype(np.float64)
        ref_desc = "FP64 CPU oracle integrator, %d spp, seed %d (%s)" % (int(z["spp"]), int(z["seed"]),
                                                                        os.path.relpath(GUIDING_REF, ROOT))
    c.film_reset(res, res)
    chunk = 256
    done = 0
    ref_sec = 0.0
    paths = rays = 0
    # the plain GPU tracer's reference: the error judge when no FP64 image is
    # there, and the plain throughput figure either way (1,024 spp then)
    ref_spp = args.guiding_ref_spp if fp64 is None else 1024
    while done < ref_spp:
        n = min(chunk, ref_spp - done)
        d = native.make_render_desc(s, spp=n, spp_offset=done)
        d.seed = s.seed + 7919
        c.render_accumulate(d)
        st = c.render_stats()
        ref_sec += st["seconds"]
        paths += st["paths"]
        rays += st["extension_rays"] + st["shadow_rays"]
        done += n
        print("[bench] guiding reference %d/%d spp (%.1fs)" % (done, ref_spp, ref_sec),
              file=sys.stderr, flush=True)
    ref = c.film_read(res, res).astype(np.float64)
    if fp64 is not None:
        gpu_ref_err = rel_mse(ref, fp64)     # the 1,024-spp GPU image against the FP64 one
        ref = fp64
    else:
        ref_desc = "plain GPU tracer, %d spp (FP32 film, Kahan-compensated)" % ref_spp
        gpu_ref_err = None
    plain_rate = {"samples_per_s": round(paths / ref_sec, 1), "mrays_per_s": round(rays / ref_sec / 1e6, 2)}

    def converge(guided):
        if guided:
            c.guiding_init(args.seed_guide)
            c.guiding_configure(bsdf_fraction=args.guiding_bsdf_fraction, min_alpha=args.guiding_min_alpha,
                                lr=args.guiding_lr,
                                train_steps=args.guiding_steps, guide_after=args.guiding_after,
                                max_records=1 << 18, pass_paths=args.guiding_pass_paths,
                                train_passes=args.guiding_train_passes,
                                late_train_steps=args.guiding_late_steps,
                                max_guided_depth=args.guiding_depth)
        c.film_reset(res, res)
        spp = 0
        sec = 0.0
        steps = 0
        curve = []
        step_spp = 4
        while True:
            d = native.make_render_desc(s, spp=step_spp, spp_offset=spp, guided=guided)
            c.render_accumulate(d)
            st = c.render_stats()
            sec += st["seconds"]
            steps += st["train_steps"]
            spp += step_spp
            e = rel_mse(c.film_read(res, res), ref)
            curve.append((spp, sec, e))
            if e <= args.guiding_target or sec > args.guiding_max_seconds:
                break
            if spp >= 64:
                step_spp = 16
            if spp >= 1024:
                step_spp = 64
        t_hit = None
        if curve[-1][2] <= args.guiding_target:
            if len(curve) == 1:
                t_hit = curve[0][1]
            else:
                (_, t1, e1), (_, t2, e2) = curve[-2], curve[-1]
                # log-log interpolation between the two chunks around the target
                f = (np.log(e1) - np.log(args.guiding_target)) / max(np.log(e1) - np.log(e2), 1e-12)
                t_hit = float(np.exp(np.log(t1) + f * (np.log(t2) - np.log(t1))))
        return {"seconds_to_target": None if t_hit is None else round(t_hit, 4), "spp": spp,
                "final_relmse": round(curve[-1][2], 5), "render_seconds": round(sec, 4),
                "train_steps": steps}

    plain = converge(False)
    print("[bench] guiding plain: %s" % plain, file=sys.stderr, flush=True)
    guided = converge(True)
    print("[bench] guiding guided: %s" % guided, file=sys.stderr, flush=True)
    c.close()
    ratio = None
    if plain["seconds_to_target"] and guided["seconds_to_target"]:
        ratio = round(guided["seconds_to_target"] / plain["seconds_to_target"], 4)
    return {"workload": "slit-lit room, 64 GGX spheres, %dx%d, depth 10" % (res, res),
            "target_relmse": args.guiding_target, "reference": ref_desc,
            "plain_1024spp_relmse_vs_fp64": None if gpu_ref_err is None else round(gpu_ref_err, 5),
            "plain": plain, "guided": guided, "time_ratio_guided_over_plain": ratio,
            "plain_throughput": plain_rate, "wall_seconds": round(time.time() - t0, 1)}


def _render_terrain(s, instances):
    from kazen_proto_amd import native
    c = native.Context(0)
    c.load_scene(s, instances=instances)
    st = c.stats()
    c.film_reset(s.width, s.height)
    c.render_accumulate(native.make_render_desc(s, spp=1))      # warm-up (scene upload, path buffers)
    c.film_reset(s.width, s.height)
    c.render_accumulate(native.make_render_desc(s))
    rs = c.render_stats()
    img = c.film_read(s.width, s.height)
    c.close()
    rays = rs["extension_rays"] + rs["shadow_rays"]
    return {"seconds": round(rs["seconds"], 4),
            "samples_per_s": round(rs["paths"] / rs["seconds"], 1),
            "mrays_per_s": round(rays / rs["seconds"] / 1e6, 2),
            "extension_rays": int(rs["extension_rays"]), "shadow_rays": int(rs["shadow_rays"]),
            "passes": int(rs["passes"]), "paths_per_pass": int(rs["paths_per_pass"]),
            "bvh_build_seconds": round(st["build_seconds"], 3), "bvh_nodes": int(st["n_nodes"]),
            "two_level": bool(st["two_level"]), "bvh_bytes": int(st["bvh_bytes"]),
            "image_mean": round(float(img.mean()), 5), "finite": bool(np.isfinite(img).all())}


def cornell_cpu_gpu(threads):
    """The Cornell box workload (256^2, 16 spp, depth 5) rendered by the GPU
    tracer and by the CPU oracle integrator (FP64 film) on `threads` host
    threads: samples/s of both, and the CPU's rays/s from the GPU's ray count
    (both trace the same paths)."""
    import oracle
    from kazen_proto_amd import native, scenes
    s = scenes.cornell_box()
    c = native.Context(0)
    c.load_scene(s)
    c.render(native.make_render_desc(s, spp=1))     # warm-up
    c.render(native.make_render_desc(s))
    rs = c.render_stats()
    c.close()
    rays = rs["extension_rays"] + rs["shadow_rays"]
    sc = oracle.scene_from(s)
    d = oracle.RenderDesc()
    d.width, d.height, d.spp = s.width, s.height, s.spp
    d.max_depth, d.rr_depth, d.seed, d.spp_offset = s.max_depth, s.rr_depth, s.seed, 0
    d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
    d.tan_half_fov = s.tan_half_fov
    sc.render(d, threads=threads)                    # warm-up
    runs = []
    for _ in range(5):
        t0 = time.time()
        sc.render(d, threads=threads)
        runs.append(time.time() - t0)
    sec = float(np.median(runs))
    n = s.width * s.height * s.spp
    return {"workload": "Cornell box %dx%d, %d spp, depth %d" % (s.width, s.height, s.spp, s.max_depth),
            "cpu_samples_per_s": round(n / sec, 1), "cpu_mrays_per_s": round(rays / sec / 1e6, 3),
            "gpu_samples_per_s": round(rs["paths"] / rs["seconds"], 1),
            "gpu_over_cpu": round((rs["paths"] / rs["seconds"]) / (n / sec), 1),
            "sample": "whole image, 1 warm-up then median of 5 (%.3fs each)" % sec}


def terrain_cpu(s, gpu, threads):
    """CPU oracle integrator on every 16th image row of the terrain workload
    (all spp): samples/s, extrapolated whole-frame seconds."""
    import oracle
    t0 = time.time()
    sc = oracle.scene_from(s)
    build = time.time() - t0
    rows = np.arange(0, s.height, 16)
    px = (rows[:, None] * s.width + np.arange(s.width)[None, :]).ravel()
    d = oracle.RenderDesc()
    d.width, d.height, d.spp = s.width, s.height, s.spp
    d.max_depth, d.rr_depth, d.seed, d.spp_offset = s.max_depth, s.rr_depth, s.seed, 0
    d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
    d.tan_half_fov = s.tan_half_fov
    t0 = time.time()
    sc.render_pixels(d, px, threads=threads)
    sec = time.time() - t0
    n = px.size * s.spp
    frac = px.size / float(s.width * s.height)
    return {"cpu_samples_per_s": round(n / sec, 1),
            "gpu_over_cpu": round(gpu["samples_per_s"] / (n / sec), 1),
            "sample": "every 16th row (%d pixels, %.4f of the frame), %d spp, one run of %.2fs"
                      % (px.size, frac, s.spp, sec),
            "cpu_frame_seconds_extrapolated": round(sec / frac, 2),
            "cpu_bvh_build_seconds": round(build, 2)}


def path_trace_leg(args, cpu_threads=None):
    """Plain wavefront path trace of the terrain workload (2048^2 fBm
    heightfield, 256 instanced spheres, three material families, 4 area
    lights) at args.pt_res^2 x args.pt_spp, depth 8: samples/s and rays/s
    over the whole kp_render_accumulate (all bounces, shading, shadow rays,
    film).  The spheres are one prototype placed 256 times: the default build
    bakes the placements into the one-level BVH (they fit the bake budget);
    the same scene with two-level traversal (shared sphere BVH) is rendered
    beside it."""
    from kazen_proto_amd import native, scenes
    t0 = time.time()
    s = scenes.terrain_scene(grid=args.pt_grid, width=args.pt_res, height=args.pt_res, spp=args.pt_spp,
                             instanced=True)
    gen = time.time() - t0
    print("[bench] terrain %d tris (generated in %.1fs)" % (s.n_tris, gen), file=sys.stderr, flush=True)
    out = {"workload": "terrain %dx%d heightfield + 256 instanced spheres, %dx%d, %d spp, depth 8"
                       % (args.pt_grid, args.pt_grid, s.width, s.height, s.spp),
           "triangles": int(s.n_tris)}
    out.update(_render_terrain(s, native.INSTANCES_AUTO))
    out["two_level_traversal"] = _render_terrain(s, native.INSTANCES_TWO_LEVEL)
    if cpu_threads:
        out["cpu"] = dict(terrain_cpu(s, out, cpu_threads), workload=out["workload"])
    print("[bench] path trace: %s" % out, file=sys.stderr, flush=True)
    return out


def batch_leg(args):
    """Batched uneven workloads: 32 terrain-style scenes (seeds 100-131),
    triangle counts log-uniform in [1e4, 2e6], resolutions from {128^2,
    256^2, 512^2, 1024^2}, spp from {4, 16, 64}, depth 8, rendered as one
    batch (one traversal launch per bounce over all scenes, replayed pass
    graph) after autotuning, and for comparison one scene after another."""
    from kazen_proto_amd import native, scenes
    t0 = time.time()
    ss = scenes.batch_scenes(args.batch_scenes, seed=100)
    gen = time.time() - t0
    print("[bench] batch: %d scenes generated (%.1fs)" % (len(ss), gen), file=sys.stderr, flush=True)
    owner = native.Context(0)
    ctxs = []
    for s in ss:
        c = native.Context(0)
        c.load_scene(s)
        ctxs.append(c)
    descs = [native.make_render_desc(s) for s in ss]
    print("[bench] batch: scenes loaded (%.1fs)" % (time.time() - t0), file=sys.stderr, flush=True)
    # one scene after another (each render is its own sequence of launches);
    # each context frees its path state afterwards, so at most one is held
    seq_sec = 0.0
    seq_rays = 0
    films = []
    for c, d in zip(ctxs, descs):
        films.append(c.render(d))
        st1 = c.render_stats()
        seq_sec += st1["seconds"]
        seq_rays += st1["extension_rays"] + st1["shadow_rays"]
        c.release_buffers()
    print("[bench] batch: one by one %.3fs (%.1fs)" % (seq_sec, time.time() - t0), file=sys.stderr, flush=True)
    b = native.Batch(owner, ctxs)
    print("[bench] batch: created (%.1fs)" % (time.time() - t0), file=sys.stderr, flush=True)
    for c in ctxs:
        c.close()
    print("[bench] batch: scene contexts closed (%.1fs)" % (time.time() - t0), file=sys.stderr, flush=True)
    tuned = b.autotune(descs, passes=2).as_dict()
    print("[bench] batch: autotuned %s (%.1fs)" % (tuned, time.time() - t0), file=sys.stderr, flush=True)
    b.render(descs)                 # warm-up with the tuned settings
    b.render(descs)
    st = b.stats()
    # each batch film must equal the scene's own render bit for bit (same
    # paths, same per-pixel summation order)
    films_equal = all(np.array_equal(b.film_read(i, d.width, d.height).view(np.uint32), f.view(np.uint32))
                      for i, (d, f) in enumerate(zip(descs, films)))
    b.close()
    owner.close()
    rays = st["extension_rays"] + st["shadow_rays"]
    out = {"workload": "%d terrain scenes, seeds 100-%d, 1e4-2e6 tris, 128^2-1024^2, 4-64 spp, depth 8"
                       % (len(ss), 99 + len(ss)),
           "triangles": int(sum(s.n_tris for s in ss)), "samples": int(st["paths"]),
           "seconds": round(st["seconds"], 4), "samples_per_s": round(st["paths"] / st["seconds"], 1),
           "mrays_per_s": round(rays / st["seconds"] / 1e6, 2), "passes": int(st["passes"]),
           "one_by_one_seconds": round(seq_sec, 4), "speedup_vs_one_by_one": round(seq_sec / st["seconds"], 3),
           "rays": int(rays), "one_by_one_rays": int(seq_rays), "films_equal_one_by_one": bool(films_equal),
           "tuning": tuned, "setup_seconds": round(time.time() - t0, 1), "generate_seconds": round(gen, 1)}
    print("[bench] batch: %s" % out, file=sys.stderr, flush=True)
    return out


def batch_leg_child(args):
    """The batched-scenes leg in a child process with a time limit, so that a
    stall in it cannot cost the whole benchmark line (runs once stalled in
    batch set-up while the merged BVH copies were not ordered with the kernels
    that rebase them): on a timeout or failure the leg is reported with its
    error and no numbers."""
    cmd = [sys.executable, os.path.abspath(__file__), "--batch-only", "--batch-scenes", str(args.batch_scenes)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, timeout=args.batch_timeout, check=False)
    except subprocess.TimeoutExpired:
        print("[bench] batch leg exceeded %.0f s and was stopped" % args.batch_timeout, file=sys.stderr, flush=True)
        return {"error": "timed out after %.0f s" % args.batch_timeout}
    if r.returncode != 0:
        return {"error": "exit status %d" % r.returncode}
    try:
        return json.loads(r.stdout.decode().strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": "no result line"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = every CPU this process may use, capped by the environment's CPU share "
                         "(OMP_NUM_THREADS) where one is set")
    ap.add_argument("--skip-primary", action="store_true",
                    help="launch only the incoherent set (profiling runs: one kernel population)")
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--refill", type=int, default=40)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device copy kernel of MIB MiB (reference kernel of "
                         "known traffic for HBM counter calibration)")
    ap.add_argument("--no-pathtrace", action="store_true", help="skip the terrain path-trace leg")
    ap.add_argument("--no-width-tune", action="store_true", help="skip the node-width tuner on the microbench")
    ap.add_argument("--launch-scaling", action="store_true",
                    help="also time incoherent launches of 2^16..2^25 rays on the same BVH")
    ap.add_argument("--pt-grid", type=int, default=2048)
    ap.add_argument("--pt-res", type=int, default=1024)
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-scenes leg and print its JSON (the main run starts this)")
    ap.add_argument("--batch-timeout", type=float, default=150.0,
                    help="seconds the batched-scenes leg may take before it is reported as failed")
    ap.add_argument("--batch-scenes", type=int, default=32)
    ap.add_argument("--no-guiding", action="store_true", help="skip the guided-vs-plain equal-error leg")
    ap.add_argument("--guiding-res", type=int, default=512)
    ap.add_argument("--guiding-ref-spp", type=int, default=16384,
                    help="spp of the plain GPU reference when the FP64 oracle image is not used")
    ap.add_argument("--guiding-gpu-ref", action="store_true",
                    help="judge the error against the plain GPU tracer instead of the FP64 oracle image")
    ap.add_argument("--guiding-target", type=float, default=0.01)
    ap.add_argument("--guiding-max-seconds", type=float, default=30.0)
    ap.add_argument("--guiding-lr", type=float, default=1e-3)
    ap.add_argument("--guiding-bsdf-fraction", type=float, default=0.5,
                    help="one-sample MIS probability of sampling the BSDF at a guided vertex")
    ap.add_argument("--guiding-min-alpha", type=float, default=0.5,
                    help="conductors guided from this roughness up (Lambert always)")
    ap.add_argument("--guiding-depth", type=int, default=0,
                    help="guide vertices up to this depth only (0: every vertex)")
    ap.add_argument("--guiding-steps", type=int, default=8, help="Adam steps per pass")
    ap.add_argument("--guiding-pass-paths", type=int, default=1 << 24,
                    help="paths per guided pass (each pass trains once)")
    ap.add_argument("--guiding-train-passes", type=int, default=32,
                    help="passes that take --guiding-steps Adam steps (0 = all)")
    ap.add_argument("--guiding-late-steps", type=int, default=1,
                    help="Adam steps per pass after them (online training continues; 0 = frozen)")
    ap.add_argument("--guiding-after", type=int, default=2, help="training passes before guided sampling")
    ap.add_argument("--seed-guide", type=int, default=17)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2
    if args.batch_only:
        print(json.dumps(batch_leg(args)), flush=True)
        return 0

    from kazen_proto_amd import native, scenes

    t0 = time.time()
    tris = scenes.microbench_triangles(seed=1)
    prim, inc = scenes.microbench_rays(seed=1)
    gen_s = time.time() - t0
    print("[bench] scene %d tris, rays %d+%d (%.1fs)" % (tris.shape[0], prim.shape[0],
                                                         inc.shape[0], gen_s),
          file=sys.stderr, flush=True)

    ctx = native.Context(0)
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        a = native.DeviceBuffer(ctx, (nb // 4,), np.float32)
        b = native.DeviceBuffer(ctx, (nb // 4,), np.float32)
        ctx.device_copy(b, a)
        ctx.synchronize()
        print("[bench] copy calibration kernel: %d bytes read, %d written" % (nb, nb),
              file=sys.stderr, flush=True)
        a.free()
        b.free()
    ctx.set_trace_tuning(args.refill)
    ctx.mesh_add_soup(tris)
    ctx.bvh_build(leaf_size=args.leaf)
    st = ctx.stats()
    print("[bench] bvh8 nodes %d depth %d build %.2fs" % (st["n_nodes"], st["max_depth"],
                                                          st["build_seconds"]),
          file=sys.stderr, flush=True)

    # inputs resident in HBM before the timed region
    d_inc = native.DeviceBuffer.from_array(ctx, inc)
    d_prim = native.DeviceBuffer.from_array(ctx, prim)
    h_inc = native.DeviceBuffer(ctx, (inc.shape[0], 4), np.float32)
    h_prim = native.DeviceBuffer(ctx, (prim.shape[0], 4), np.float32)

    def step(r, h):
        ctx.intersect_closest(r, out=h)

    for _ in range(args.warmup):
        step(d_inc, h_inc)
        if not args.skip_primary:
            step(d_prim, h_prim)
    ctx.synchronize()

    def timed(r, h, k):
        # one process, one GPU: the barrier is the stream synchronisation
        ctx.synchronize()
        ctx.timer_begin()           # HIP event on the launch stream
        for _ in range(k):
            step(r, h)
        ms = ctx.timer_end()        # HIP event, waits for the stream
        ctx.synchronize()           # also surfaces kernel-reported errors
        return ms / k

    ms_inc = timed(d_inc, h_inc, args.steps)
    mrays_inc = inc.shape[0] / (ms_inc * 1e-3) / 1e6
    mrays_prim = None
    if not args.skip_primary:
        ms_prim = timed(d_prim, h_prim, max(1, args.steps // 2))
        mrays_prim = round(prim.shape[0] / (ms_prim * 1e-3) / 1e6, 2)
    print("[bench] incoherent %.3f ms/launch %.1f Mrays/s; primary %s Mrays/s"
          % (ms_inc, mrays_inc, mrays_prim), file=sys.stderr, flush=True)

    # algorithmic bytes per ray (untimed): B_ray = 32 (ray) + 20 (hit) +
    # 128 n_node + 48 n_tri, with the node / triangle counts of the CPU
    # oracle's traversal of the identical BVH8 on a sample of the rays; the
    # kernel's own counts (its leaf postponing visits a few more nodes) beside
    _, nv, nt = ctx.intersect_closest(d_inc, out=h_inc, counted=True)
    n = inc.shape[0]
    k_nodes, k_tris = nv / n, nt / n
    o_nodes, o_tris, o_sample = k_nodes, k_tris, 0
    try:
        import oracle
        nodes8, tris12, _ = native.bvh8_build_host(tris, leaf_size=args.leaf)
        o_sample = min(n, 262144)
        _, onv, ont = oracle.bvh8_closest(nodes8, tris12, inc[:o_sample])
        o_nodes, o_tris = onv / o_sample, ont / o_sample
    except Exception as e:   # the oracle is a checker: without it the kernel's counts stand in
        print("[bench] oracle node counts unavailable: %s" % e, file=sys.stderr, flush=True)
    b_ray = 32 + 20 + 128 * o_nodes + 48 * o_tris
    achieved = b_ray * n / (ms_inc * 1e-3) / 1e9
    traffic = profiled_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/closest_incoherent_microbench.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE)",
                "dram_gbs": None if traffic is None else round(traffic / (ms_inc * 1e-3) / 1e9, 1),
                "bytes_per_ray": round(b_ray, 1), "nodes_per_ray": round(o_nodes, 3),
                "tris_per_ray": round(o_tris, 3), "counts_source": "oracle BVH8 traversal of %d rays" % o_sample,
                "kernel_nodes_per_ray": round(k_nodes, 3), "kernel_tris_per_ray": round(k_tris, 3)}

    # per-launch time against the launch size, same BVH: the incoherent set
    # plus 7 more of the same distribution, traced as prefixes of 2^16..2^25
    launch_scaling = None
    if args.launch_scaling:
        rng = np.random.default_rng(77)
        big = np.concatenate([inc] + [scenes.incoherent_rays(inc.shape[0], rng) for _ in range(7)])
        d_big = native.DeviceBuffer.from_array(ctx, big)
        h_big = native.DeviceBuffer(ctx, (big.shape[0], 4), np.float32)
        del big
        launch_scaling = []
        for lg in range(16, 26):
            k = 1 << lg
            for _ in range(2):
                ctx.intersect_closest(d_big, out=h_big, n=k)
            best = float("inf")
            for _ in range(3):
                ctx.synchronize()
                ctx.timer_begin()
                for _ in range(5):
                    ctx.intersect_closest(d_big, out=h_big, n=k)
                best = min(best, ctx.timer_end() / 5)
            ctx.synchronize()
            row = {"rays": k, "ms": round(best, 4), "mrays_per_s": round(k / best / 1e3, 1),
                   "frac_at_oracle_bytes": round(b_ray * k / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            launch_scaling.append(row)
            print("[bench] launch scaling %s" % row, file=sys.stderr, flush=True)
        d_big.free()
        h_big.free()

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        host = host_info()
        threads = args.cpu_threads or min(host["affinity_cpus"], host["cpu_share"] or host["affinity_cpus"])
        t0 = time.time()
        bvh = oracle.Bvh(tris)
        cpu_build = time.time() - t0
        bvh.closest(inc[:65536], threads=threads)           # warm-up

        def cpu_median(rays):
            runs, hits = [], None
            for _ in range(5):
                hits, sec = bvh.closest_timed(rays, threads=threads)
                runs.append(sec)
            return hits, float(np.median(runs))

        hits_cpu, sec = cpu_median(inc)
        _, sec_p = cpu_median(prim)
        cpu = {"value": round(inc.shape[0] / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": "all %d incoherent rays, oracle binary-SAH BVH + watertight test, "
                         "1 warm-up then median of 5 runs (%.2fs each)" % (inc.shape[0], sec),
               "host": host, "gpu_over_cpu": round(mrays_inc / (inc.shape[0] / sec / 1e6), 1),
               "primary_mrays_per_s": round(prim.shape[0] / sec_p / 1e6, 3),
               "bvh_build_seconds": round(cpu_build, 3)}
        cpu["renders"] = {"cornell": cornell_cpu_gpu(threads)}
        got = h_inc.numpy()
        mm = int(np.count_nonzero(np.any(got.view(np.uint32) != hits_cpu.view(np.uint32), axis=1)))
        ties = bvh.count_near_ties(inc, hits_cpu, threads=threads)
        parity = {"incoherent_mismatches": mm, "rays": int(n), "near_ties": ties,
                  "near_ties_note": "rays whose closest hit has a second primitive within 1 ulp of t "
                                    "(resolved by the lowest primitive id on both sides)"}

    # build-time node-width tuner (KP_NODE_WIDTH_AUTO): widths 4 / 6 / 8 timed
    # on its own probe rays; the benchmark rays then timed at the width kept
    width_tuning = None
    if not args.no_width_tune:
        ctx.bvh_build(leaf_size=args.leaf, node_width=native.NODE_WIDTH_AUTO)
        wst = ctx.stats()
        ms_w = timed(d_inc, h_inc, max(1, args.steps // 2))
        width_tuning = {"probe_ms": {str(4 + 2 * i): round(m, 4) for i, m in enumerate(wst["width_ms"])},
                        "chosen": int(wst["node_width"]),
                        "incoherent_mrays_per_s_at_chosen": round(n / (ms_w * 1e-3) / 1e6, 2),
                        "build_seconds_all_widths": round(wst["build_seconds"], 3)}
        print("[bench] node-width tuner: %s" % width_tuning, file=sys.stderr, flush=True)

    path_trace = None
    if not args.no_pathtrace:
        path_trace = path_trace_leg(args, None if args.no_cpu or cpu is None else cpu["cores"])
        if cpu is not None and path_trace.get("cpu"):
            cpu["renders"]["terrain"] = path_trace.pop("cpu")
    batch = None
    if not args.no_batch:
        batch = batch_leg_child(args)
    guiding = None
    if not args.no_guiding:
        guiding = guiding_equal_error(args)

    line = {
        "metric": HEADLINE_METRIC,
        "value": round(mrays_inc, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_inc, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "intersection microbench, incoherent closest-hit",
                   "triangles": int(tris.shape[0]), "rays_per_step": int(n),
                   "bvh": "BVH8, 128-byte nodes, leaf<=%d" % args.leaf, "tmax": "inf"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "primary_mrays_per_s": mrays_prim,
        "parity": parity,
        "bvh_build_seconds": round(st["build_seconds"], 3),
        "node_width_tuning": width_tuning,
        "launch_scaling": launch_scaling,
        "path_trace": path_trace,
        "batch": batch,
        "guiding": guiding,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
