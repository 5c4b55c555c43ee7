"""Benchmark of the MI355X per-pixel SDF render path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one frame of the 960x540 Cornell GI workload (seed 2: five box walls, three
primitives incl. a smooth union and a twist, emissive bar, sphere light, 16^3 GI probes baked
on the device, soft shadows, AO, one reflection bounce, bloom + tonemap, aa = 2) rendered by
the plain per-frame launch path (`render_only`: render kernel + bloom/tonemap kernels) with
the scene and framebuffers resident in HBM; the render kernel is the scene's generated
kernel (scene -> HIP source -> hipRTC), which the bench waits for before timing.  W untimed frames, then K timed frames between
two device synchronisations.  One JSON line on stdout; progress on stderr.

roofline: the render kernel's algorithmic FP32 work per frame (tests/golden/flops.json,
counted by the instrumented oracle on the same scene and camera) divided by its mean
HIP-event duration over the timed frames, against MI355X's vector FP32 peak.
cpu_baseline: the CPU oracle (a C++ restatement of the reference's multithreaded CPU
renderer) timed on this host on one full frame of the same workload.

More legs ride on the same line (``--legs`` picks them):
  batched: 64 cameras over 8 seeded procedural scenes (20-200 primitives, resolutions from
           {320x180, 640x360, 1280x720, 1920x1080}) rendered once as 64 plain per-frame launches
           (each followed by a wait, as render() does) and once as one autotuned persistent
           tile-queue launch replayed from a hipGraph; Mpix/s of both and their ratio.
  served:  requests over 32 resident scenes (16^3 GI probes each; 854x480 50 %, 1280x720 35 %,
           1920x1080 15 %; 500 by default) arriving as a seeded Poisson stream at 0.3 / 0.5 / 0.7 /
           0.9 of the plain path's measured capacity (the plain FIFO at 0.5 and 0.9), plus a
           burst; p50/p99 latency (arrival -> RGBA8 in
           host memory) and Mpix/s for the continuous-batching server (8 batches of at most 2
           requests in flight) and for a FIFO server that renders each request by a plain
           per-frame launch.
  mesh_bake: the reference mesh demo's torus knot (240 x 36 ring vertices, 17,280 triangles)
           imported from OBJ text and baked to a signed-distance grid with 128 samples along the
           longest axis (124x63x128) on the device; ms per bake (row-parity + nearest-triangle
           kernels, HIP events), Msamples/s, BVH tests per sample, and the CPU oracle's rate.
  frame_gen: c2 at 960x540 as a G-buffer keyframe plus reprojected frames for the camera swinging
           0.01 rad per frame about its target (reproject_hybrid: warp + holes rendered for real), with
           the scene's generated G-buffer / hole kernels; ms per G-buffer frame vs per generated frame
           (both including the RGBA8 readback), holes.
  queued:  c2 compiled with the queued-frame stages (budgeted march passes over compacted ray
           queues, hit-queue shading); the scene autotuner's render-stage ms for the one-kernel
           frame and each queued candidate, its choice, and whether the images are equal.
  stress:  one 3840x2160 frame of the 1000-primitive scene with a 128^3
           sampled volume, 32^3 device-baked GI, 2 bounces and fog; ms/frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "atom-3d-engine_amd"))

METRIC = "Mpix/s at 960x540 GI scene; p50/p99 latency, batched serve vs plain"
PEAK_FP32_TFLOPS = 157.3          # MI355X vector FP32 (dual-issue / packed)
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # one FP32 op per lane per clock
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ir, cam, threads, frames=3):
    """Oracle CPU renderer (row-band threads, the reference's render_with_threads) on full
    frames: 1 untimed warm-up frame, then `frames` timed frames; plus a 1-thread figure on
    four bands of rows spread over the frame (1/8 of its rows)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    s = po.OracleScene(ir, threads=threads)
    s.render_rgba(cam, threads=threads)
    t0 = time.perf_counter()
    for _ in range(frames):
        s.render_rgba(cam, threads=threads)
    dt = (time.perf_counter() - t0) / frames
    band = max(1, s.height // 32)
    rows = 0
    t0 = time.perf_counter()
    for k in range(4):
        y0 = (2 * k + 1) * s.height // 8 - band // 2
        s.render_linear_rows(cam, y0, y0 + band, threads=1)
        rows += band
    dt1 = time.perf_counter() - t0
    return {"mpix_s": s.width * s.height / dt / 1e6, "s_per_frame": dt,
            "mpix_s_1thread": s.width * rows / dt1 / 1e6, "rows_1thread": rows, "s_1thread": dt1}


def flops_record(name):
    with open(os.path.join(ROOT, "tests", "golden", "flops.json")) as f:
        return json.load(f).get(name)


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if xs else None


def batched_leg(hip, configs, r, reps):
    """64 plain per-frame launches vs one persistent tile-queue launch over the same jobs."""
    jobs = configs.c3_jobs()
    scenes = {}
    compiled = []
    t0 = time.perf_counter()
    for sc, cam, (w, h) in jobs:
        key = (id(sc), w, h)
        if key not in scenes:
            scenes[key] = r.compile(sc.to_ir(), w, h)
        compiled.append((scenes[key], cam, (w, h)))
    px = sum(w * h for _, _, (w, h) in jobs)
    for i, g in enumerate(scenes.values()):
        g.codegen_status(wait=True)  # plain launches use each scene's generated kernel
        if i % 8 == 7:
            log(f"batched: {i + 1} of {len(scenes)} (scene, size) pairs have their generated kernels "
                f"({time.perf_counter() - t0:.1f} s)")
    log(f"batched: {len(jobs)} jobs, {px / 1e6:.1f} Mpix, {len(scenes)} compiled (scene, size) pairs (with their "
        f"generated kernels) in {time.perf_counter() - t0:.2f} s")
    # plain: one launch per frame, each waited for (render_only + wait_idle)
    for g, cam, _ in compiled[:4]:
        g.render_only(cam)
        r.wait_idle()
    plain = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for g, cam, _ in compiled:
            g.render_only(cam)
            r.wait_idle()
        plain.append(time.perf_counter() - t0)
    plain_s = min(plain)
    b = hip.Batch(r, compiled)
    t0 = time.perf_counter()
    report = b.autotune(reps=1)
    tune_s = time.perf_counter() - t0
    info = b.info()
    chosen = next((ln[len("chosen: "):] for ln in report.splitlines() if ln.startswith("chosen: ")), "")
    mixed = chosen.startswith("mixed")
    for ln in report.splitlines():
        log("batched autotune: " + ln)
    log(f"batched: autotune {tune_s:.1f} s -> {chosen}")
    b.launch()
    b.wait()
    wall, gpu = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        b.launch()
        gpu.append(b.wait())
        wall.append(time.perf_counter() - t0)
    batched_s = min(wall)
    b.close()
    for g in scenes.values():
        g.close()
    fl = flops_record("c3")
    roof = None
    if fl and fl.get("pixels") == px:
        ach = fl["flops_per_frame"] / (min(gpu) * 1e-3) / 1e12
        kernel = ("each job's generated scene kernel, four streams, then the batch post" if mixed else
                  "persistent batch render (queue reset + render + post)")
        roof = {"bound": "valu", "kernel": kernel,
                "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP32_TFLOPS, 4), "frac_nonpacked_peak": round(ach / PEAK_FP32_NONPACKED_TFLOPS, 4),
                "flops_per_launch": fl["flops_per_frame"], "flops_sample": fl["sample"]}
    out = {"workload": "64 cameras over 8 seeded procedural scenes (20-200 primitives, 30 % smooth, 20 % domain "
                       "ops), per-camera resolution uniform over {320x180, 640x360, 1280x720, 1920x1080}, aa=1, "
                       "1 reflection bounce", "roofline": roof,
           "jobs": len(jobs), "mpix_per_batch": round(px / 1e6, 3),
           "plain_mpix_s": round(px / plain_s / 1e6, 3), "plain_s": round(plain_s, 4),
           "batched_mpix_s": round(px / batched_s / 1e6, 3), "batched_s": round(batched_s, 4),
           "batched_gpu_ms": round(min(gpu), 3), "speedup_vs_plain": round(plain_s / batched_s, 3),
           "mode": chosen, "autotune_s": round(tune_s, 2)}
    if not mixed:
        out["tile"], out["blocks_per_cu"] = list(info["tile"]), info["blocks_per_cu"]
    log(f"batched: plain {out['plain_mpix_s']} Mpix/s, batched {out['batched_mpix_s']} Mpix/s "
        f"({out['speedup_vs_plain']}x)")
    return out


def served_leg(hip, configs, r, n_req, n_scenes=32, inflight=8, batch=2):
    """Poisson request stream: continuous-batching server vs a FIFO server of plain launches."""
    import numpy as np
    t0 = time.perf_counter()
    scs = configs.c4_scenes(n_scenes=n_scenes)
    rng = configs.SplitMix64(11)
    reqs = [configs.c4_request(rng, n_scenes) for _ in range(n_req)]
    need = sorted({(si, w, h) for si, _, (w, h) in reqs})
    plain_sc = {}
    for si, w, h in need:
        plain_sc[(si, w, h)] = r.compile(scs[si].to_ir(), w, h)
    base = [r.compile(sc.to_ir(), 854, 480) for sc in scs]
    log(f"served: {len(plain_sc) + len(base)} (scene, size) pairs compiled in {time.perf_counter() - t0:.1f} s; "
        "waiting for their generated kernels")
    for i, g in enumerate(plain_sc.values()):
        g.codegen_status(wait=True)
        if i % 16 == 15:
            log(f"served: {i + 1} of {len(plain_sc)} plain pairs have their generated kernels "
                f"({time.perf_counter() - t0:.1f} s)")
    log(f"served: {n_scenes} resident scenes + {len(plain_sc)} plain (scene, size) pairs compiled with device GI "
        f"bakes and generated kernels in {time.perf_counter() - t0:.2f} s; {n_req} requests")
    px_req = [w * h for _, _, (w, h) in reqs]
    bufs = {(w, h): np.zeros((h, w, 4), np.uint8) for _, _, (w, h) in reqs}
    for (si, w, h), g in list(plain_sc.items())[:3]:
        g.render_rgba(reqs[0][1])

    def plain_run(arrivals):
        lat = []
        t_start = time.perf_counter()
        for (si, cam, (w, h)), a in zip(reqs, arrivals):
            now = time.perf_counter() - t_start
            if a > now:
                time.sleep(a - now)
            plain_sc[(si, w, h)].render_rgba(cam)
            lat.append((time.perf_counter() - t_start - a) * 1e3)
        span = time.perf_counter() - t_start
        return lat, span

    srv = hip.Server(r, hbm_budget_bytes=8 << 30, max_batch_requests=batch, max_inflight=inflight)
    ids = [srv.add_scene(g) for g in base]

    def served_run(arrivals):
        rids = []
        t_start = time.perf_counter()
        for (si, cam, (w, h)), a in zip(reqs, arrivals):
            now = time.perf_counter() - t_start
            if a > now:
                time.sleep(a - now)
            rids.append(srv.submit(ids[si], cam, w, h, keep_image=False))
        lat = [srv.wait(x)[1] for x in rids]
        span = time.perf_counter() - t_start
        return lat, span

    served_run([0.0] * len(reqs))  # warm the server's pinned staging and slab
    # each request alone (plain launch, readback included): the latency floor no scheduling can beat
    alone = []
    for si, cam, (w, h) in reqs:
        t1 = time.perf_counter()
        plain_sc[(si, w, h)].render_rgba(cam)
        alone.append((time.perf_counter() - t1) * 1e3)
    log(f"served: each request alone: p50/p99/max {pct(alone, 50):.1f}/{pct(alone, 99):.1f}/{max(alone):.1f} ms")
    burst = [0.0] * len(reqs)
    plain_lat, plain_span = plain_run(burst)
    served_lat, served_span = served_run(burst)
    px = sum(px_req)
    plain_cap_qps = len(reqs) / plain_span
    points = []
    prng = np.random.default_rng(11)
    gaps = prng.exponential(1.0, len(reqs))
    for frac in (0.3, 0.5, 0.7, 0.9):
        qps = plain_cap_qps * frac
        arr = list(np.cumsum(gaps / qps) - gaps[0] / qps)
        sl, ss = served_run(arr)
        pt = {"offered_qps": round(qps, 2), "fraction_of_plain_capacity": frac, "requests": len(reqs),
              "served": {"mpix_s": round(px / ss / 1e6, 3), "p50_ms": round(pct(sl, 50), 2),
                         "p99_ms": round(pct(sl, 99), 2), "max_ms": round(max(sl), 2)}}
        msg = f"served: {qps:.1f} req/s offered: served p50/p99 {pct(sl, 50):.1f}/{pct(sl, 99):.1f} ms"
        if frac in (0.5, 0.9):  # the FIFO of plain launches at two of the rates
            pl, ps = plain_run(arr)
            pt["plain"] = {"mpix_s": round(px / ps / 1e6, 3), "p50_ms": round(pct(pl, 50), 2),
                           "p99_ms": round(pct(pl, 99), 2), "max_ms": round(max(pl), 2)}
            msg += f", plain p50/p99 {pct(pl, 50):.1f}/{pct(pl, 99):.1f} ms"
        points.append(pt)
        log(msg)
    st = srv.stats()
    srv.close()
    for g in list(plain_sc.values()) + base:
        g.close()
    out = {"workload": f"{n_req} requests over {n_scenes} resident seeded procedural scenes (20-200 primitives, "
                       "16^3 GI probes baked on device), resolution 854x480 50 % / 1280x720 35 % / 1920x1080 15 %, "
                       "seeded orbit poses; latency = arrival -> RGBA8 in host memory",
           "requests": len(reqs), "mpix_requested": round(px / 1e6, 3),
           "server": {"max_inflight": inflight, "max_batch_requests": batch},
           "burst": {"plain": {"mpix_s": round(px / plain_span / 1e6, 3), "p50_ms": round(pct(plain_lat, 50), 2),
                               "p99_ms": round(pct(plain_lat, 99), 2)},
                     "served": {"mpix_s": round(px / served_span / 1e6, 3),
                                "p50_ms": round(pct(served_lat, 50), 2), "p99_ms": round(pct(served_lat, 99), 2)},
                     "served_vs_plain": round(plain_span / served_span, 3)},
           "alone_ms": {"p50": round(pct(alone, 50), 2), "p99": round(pct(alone, 99), 2), "max": round(max(alone), 2),
                        "over_50ms": sum(1 for a in alone if a > 50.0)},
           "poisson": points, "server_stats": st}
    log(f"served: burst plain {out['burst']['plain']['mpix_s']} Mpix/s, served {out['burst']['served']['mpix_s']} "
        f"Mpix/s ({out['burst']['served_vs_plain']}x)")
    return out


def stress_prepare(hip, configs, r):
    """c5 scene built and compiled (tables, segment grid, device GI bake); its generated kernel
    compiles on the renderer's background threads while the other legs run."""
    t0 = time.perf_counter()
    sc, cam = configs.c5_scene(bake=hip.device_bake(r))  # the star surface's SDF baked on the device
    ir = sc.to_ir()
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = r.compile(ir, sc.width, sc.height)
    compile_s = time.perf_counter() - t0
    log(f"stress: IR {len(ir) / 1e6:.1f} MB built (star mesh baked to an SDF grid on the device) in {build_s:.1f} s; "
        f"compile (tables, grid, 32^3 GI bake) {compile_s:.1f} s; its generated kernel compiles meanwhile")
    return sc, cam, g, compile_s, time.perf_counter()


def stress_leg(hip, configs, r, frames, prep):
    """c5: 3840x2160 frames of a 1000-primitive scene with a 128^3 sampled volume, 32^3 GI
    probes baked on the device, 2 reflection bounces and fog."""
    sc, cam, g, compile_s, t_compiled = prep
    st, msg = g.codegen_status(wait=True)
    codegen_s = time.perf_counter() - t_compiled
    info = g.info()
    log(f"stress: generated kernel state {st} {msg[:200]} {codegen_s:.1f} s after the scene compile; {info}")
    g.render_only(cam)
    r.wait_idle()
    g.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(frames):
        g.render_only(cam)
    r.wait_idle()
    dt = (time.perf_counter() - t0) / frames
    n, render_ms, post_ms = g.timing(reset=True)
    vol = sc.volumes[0]
    out = {"workload": "c5: 3840x2160, 1000 seeded primitives (C3 generator, twist + bend), the signed distance of a "
                       "seeded star-shaped closed mesh (16,128 triangles) baked on the device with 128 samples along "
                       f"its longest axis ({vol.dims[0]}x{vol.dims[1]}x{vol.dims[2]}) and sampled as a volume, 32^3 "
                       "GI probes baked on device, 2 reflection bounces, fog 0.02, aa=1",
           "width": sc.width, "height": sc.height, "frames": frames, "ms_per_frame": round(dt * 1e3, 2),
           "mpix_s": round(sc.width * sc.height / dt / 1e6, 3), "render_kernel": g.kernel_name(),
           "render_kernel_ms": round(render_ms, 2), "post_ms": round(post_ms, 3), "compile_s": round(compile_s, 2),
           "codegen_wait_s": round(codegen_s, 1), "objects": info["objects"], "segments": info["segments"],
           "grid_cells": info["grid_cells"], "roofline": None}
    fl = flops_record("c5")
    if fl and fl.get("width") == sc.width and fl.get("height") == sc.height and n:
        ach = fl["flops_per_frame"] / (render_ms * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu", "kernel": g.kernel_name(), "achieved": round(ach, 3),
                           "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                           "frac_nonpacked_peak": round(ach / PEAK_FP32_NONPACKED_TFLOPS, 4),
                           "flops_per_launch": fl["flops_per_frame"], "flops_sample": fl["sample"]}
    log(f"stress: {out['ms_per_frame']} ms/frame ({out['mpix_s']} Mpix/s) with {out['render_kernel']}")
    g.close()
    return out


# f32 VALU operations of one BVH box test (clamp: 3 max + 3 min; 3 sub; dot: 3 mul + 2 add) and
# of one triangle test (the point-triangle projection's longest, interior path plus the squared
# distance), as the bake kernel computes them
BOX_OPS = 14
TRI_OPS = 79


def mesh_bake_leg(hip, configs, r, reps, resolution=128, padding=0.35, cpu=True):
    """mesh::bake_sdf of the mesh demo's torus knot on the device, timed over `reps` bakes."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    pos0, tri0 = configs.torus_knot_mesh()
    obj = configs.mesh_to_obj(pos0, tri0, "MM3E torus knot")
    t0 = time.perf_counter()
    pos, tri = hip.parse_obj(obj)
    parse_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = hip.MeshBaker(r, pos, tri, resolution, padding)
    create_s = time.perf_counter() - t0
    dims = tuple(b.layout.dims)
    n = int(np.prod(dims))
    b.run()
    b.wait()
    par, near = [], []
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
        a, c = b.wait()
        par.append(a)
        near.append(c)
    wall = (time.perf_counter() - t0) / reps
    vol = b.read()
    info = b.info()
    b.close()
    ms = (sum(par) + sum(near)) / reps
    out = {"workload": f"torus knot p=2 q=3, {len(tri)} triangles from OBJ text, resolution {resolution}, "
                       f"padding {padding}",
           "dims": list(dims), "samples": n, "bvh_depth": info["depth"], "reps": reps,
           "ms_per_bake": round(ms, 4), "parity_ms": round(sum(par) / reps, 4),
           "nearest_ms": round(sum(near) / reps, 4), "wall_ms_per_bake": round(wall * 1e3, 3),
           "msamples_s": round(n / (ms * 1e-3) / 1e6, 1), "parse_obj_ms": round(parse_s * 1e3, 2),
           "create_ms": round(create_s * 1e3, 2), "inside_fraction": round(float((vol < 0).mean()), 5)}
    if cpu:
        import pyoracle as po
        # the BVH work per sample, counted by the oracle on two slabs of grid points
        lay = b.layout
        k = dims[2] // 2
        jj, ii = np.meshgrid(np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
        pts = []
        for kk in (k, k + 1):
            pts.append(np.stack([np.float32(lay.min[0]) + ii.astype(np.float32) * np.float32(lay.cell[0]),
                                 np.float32(lay.min[1]) + jj.astype(np.float32) * np.float32(lay.cell[1]),
                                 np.full(ii.shape, np.float32(lay.min[2]) + np.float32(kk) * np.float32(lay.cell[2]),
                                         np.float32)], -1).reshape(-1, 3))
        pts = np.concatenate(pts)
        _, (boxes, tris) = po.mesh_nearest_sq(pos, tri, pts)
        per_box, per_tri = boxes / len(pts), tris / len(pts)
        ops = (per_box * BOX_OPS + per_tri * TRI_OPS) * n
        out["box_tests_per_sample"] = round(per_box, 1)
        out["tri_tests_per_sample"] = round(per_tri, 1)
        out["nearest_valu_tflops"] = round(ops / (sum(near) / reps * 1e-3) / 1e12, 3)
        out["nearest_valu_frac"] = round(out["nearest_valu_tflops"] / PEAK_FP32_NONPACKED_TFLOPS, 4)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        slabs = min(32, dims[2])
        t0 = time.perf_counter()
        po.mesh_bake(pos, tri, resolution, padding, threads=threads, k0=0, k1=slabs)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(slabs * dims[0] * dims[1] / cdt / 1e6, 3), "unit": "Msamples/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle bake of z-slabs 0..{slabs} ({slabs * dims[0] * dims[1]} samples) "
                                         f"in {cdt:.2f} s"}
    log(f"mesh_bake: {dims} in {out['ms_per_bake']} ms ({out['msamples_s']} Msamples/s)")
    return out


def frame_gen_leg(hip, configs, r_main, frames=8):
    """render_gbuffer keyframe + reproject_hybrid frames on c2 (lib.rs:951-1080), with the scene's
    generated G-buffer / hole kernels."""
    import math
    sc, cam = configs.c2_scene()
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_SYNC)
    r.set_codegen_film(True)
    t0 = time.perf_counter()
    g = r.compile(sc.to_ir(), sc.width, sc.height)
    cg_state, _ = g.codegen_status(wait=True)
    codegen_s = time.perf_counter() - t0
    gf = hip.GFrame(r, sc.width, sc.height)
    # the c2 camera (eye (0, 2, 7.5) looking at (0, 1.6, 0)) swung about the vertical through its target
    cams = [configs.look_at((7.5 * math.sin(0.01 * k), 2.0, 7.5 * math.cos(0.01 * k)), (0.0, 1.6, 0.0),
                            (0.0, 1.0, 0.0), math.radians(50.0)) for k in range(frames + 1)]
    g.render_gbuffer(cams[0], gf)  # warm-up
    t0 = time.perf_counter()
    for _ in range(3):
        g.render_gbuffer(cams[0], gf)
    key_ms = (time.perf_counter() - t0) / 3 * 1e3
    g.reproject_hybrid(gf, cams[1])  # warm-up
    holes, t = [], 0.0
    for c in cams[1:]:
        t0 = time.perf_counter()
        _, n = g.reproject_hybrid(gf, c)
        t += time.perf_counter() - t0
        holes.append(n)
    gen_ms = t / frames * 1e3
    gf.close()
    g.close()
    r.close()
    px = sc.width * sc.height
    out = {"workload": f"c2 960x540, the c2 camera as keyframe, {frames} frames swinging 0.01 rad each about "
                       f"the vertical through its target", "gbuffer_ms": round(key_ms, 3), "hybrid_ms": round(gen_ms, 3),
           "speedup": round(key_ms / gen_ms, 2), "mean_hole_fraction": round(sum(holes) / len(holes) / px, 5),
           "generated_kernels": cg_state == hip.CODEGEN_READY, "compile_and_codegen_s": round(codegen_s, 2)}
    log(f"frame_gen: keyframe {out['gbuffer_ms']} ms, generated {out['hybrid_ms']} ms "
        f"({out['speedup']}x), holes {out['mean_hole_fraction']}")
    return out


def queued_leg(hip, configs, g_plain, reps=5):
    """c2 as queued frames (budgeted march passes over compacted ray queues, hit-queue shading):
    the scene autotuner's timing of the one-kernel frame and each queued candidate, its choice,
    and whether the queued image equals the plain one."""
    sc, cam = configs.c2_scene()
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_SYNC)
    r.set_wavefront(32, 2, 0)  # generated kernels carry the queued stages
    t0 = time.perf_counter()
    g = r.compile(sc.to_ir(), sc.width, sc.height)
    state, _ = g.codegen_status(wait=True)
    codegen_s = time.perf_counter() - t0
    same = bool((g.render_rgba(cam) == g_plain.render_rgba(cam)).all())
    rep = g.autotune(cam, reps=reps)
    log("queued frames (c2):\n" + rep.rstrip())
    cands = {}
    for line in rep.strip().splitlines():
        if line.startswith("chosen:"):
            continue
        k, v = line.rsplit(":", 1)
        cands[k.strip()] = float(v.strip().split()[0])
    chosen = rep.strip().splitlines()[-1]
    g.close()
    r.close()
    return {"codegen_state": state, "compile_and_codegen_s": round(codegen_s, 2), "image_equals_plain": same,
            "render_ms": cands, "chosen": chosen}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="plain,batched,served,stress,mesh_bake,frame_gen,queued",
                    help="comma list of plain (always run), batched, served, stress, mesh_bake, frame_gen, queued")
    ap.add_argument("--bake-reps", type=int, default=20)
    ap.add_argument("--stress-frames", type=int, default=3)
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--served-requests", type=int, default=500)
    ap.add_argument("--served-inflight", type=int, default=8, help="server batches in flight (1..8)")
    ap.add_argument("--served-batch", type=int, default=2, help="requests per server batch at most")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (torch), used to scale rocprof "
                         "FETCH_SIZE/WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py: this render path runs on one GPU; --gpus {args.gpus} is not supported "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from mm3e import configs, hip

    sc, cam = configs.c2_scene()
    ir = sc.to_ir()
    W, H = sc.width, sc.height
    r = hip.Renderer()
    log(f"adapter {r.adapter_name}")
    t0 = time.perf_counter()
    g = r.compile(ir, W, H)
    compile_s = time.perf_counter() - t0
    log(f"compiled c2 (tables + device GI bake) in {compile_s:.3f} s")
    t0 = time.perf_counter()
    cg_state, cg_msg = g.codegen_status(wait=True)
    codegen_s = time.perf_counter() - t0
    log(f"generated kernel: state {cg_state} after {codegen_s:.1f} s {cg_msg[:200]}; rendering with {g.kernel_name()}")
    for _ in range(args.warmup):
        g.render_only(cam)
    r.wait_idle()
    g.timing(reset=True)
    log(f"warmup {args.warmup} frames done; timing {args.steps} frames")
    r.wait_idle()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.render_only(cam)
    r.wait_idle()
    dt = time.perf_counter() - t0
    frames, render_ms, post_ms = g.timing(reset=True)
    ms_per_step = dt / args.steps * 1e3
    value = W * H * args.steps / dt / 1e6
    log(f"{args.steps} frames in {dt:.3f} s: {ms_per_step:.3f} ms/frame, render kernel {render_ms:.3f} ms, "
        f"post {post_ms:.3f} ms ({frames} frames sampled)")

    with open(os.path.join(ROOT, "tests", "golden", "flops.json")) as f:
        fl = json.load(f)["c2"]
    assert fl["width"] == W and fl["height"] == H
    achieved = fl["flops_per_frame"] / (render_ms * 1e-3) / 1e12
    roofline = {
        "bound": "valu", "kernel": g.kernel_name(), "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "frac_nonpacked_peak": round(achieved / PEAK_FP32_NONPACKED_TFLOPS, 4),
        "flops_per_launch": fl["flops_per_frame"], "kernel_ms": round(render_ms, 4),
        "traffic": None,
    }
    prof = os.path.join(ROOT, "profiles", "render_c2_gen_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        if tr.get("kernel") == roofline["kernel"]:  # only a profile of the kernel being timed
            roofline["traffic"] = tr.get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2: 960x540 Cornell room, 3 seeded primitives (smooth k=0.3, twist), 16^3 GI "
                               "probes baked on device (64 requested -> 30 rays/probe), soft shadows, AO, "
                               "1 reflection bounce, bloom+tonemap, aa=2; plain per-frame launch",
                   "width": W, "height": H, "aa": sc.aa, "bounces": sc.bounces, "path": "plain",
                   "compile_s": round(compile_s, 3), "codegen_s": round(codegen_s, 2),
                   "render_kernel": g.kernel_name(), "render_kernel_ms": round(render_ms, 4),
                   "post_ms": round(post_ms, 4)},
        "roofline": roofline,
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        log(f"cpu baseline: oracle on {threads} threads, full frames (1 warm-up + 3 timed), then 1 thread")
        cb = cpu_baseline(ir, cam, threads)
        out["cpu_baseline"] = {"value": round(cb["mpix_s"], 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
                               "sample": f"full c2 960x540 frames (aa=2), 3 timed after 1 warm-up, "
                                         f"{cb['s_per_frame']:.2f} s/frame",
                               "value_1thread": round(cb["mpix_s_1thread"], 4),
                               "sample_1thread": f"{cb['rows_1thread']} rows of the same frame in 4 bands on 1 thread, "
                                                 f"{cb['s_1thread']:.2f} s"}
        log(f"cpu baseline {cb['mpix_s']:.4f} Mpix/s ({cb['s_per_frame']:.2f} s/frame), "
            f"1 thread {cb['mpix_s_1thread']:.4f} Mpix/s")
    legs = set(args.legs.split(","))
    stress_prep = stress_prepare(hip, configs, r) if "stress" in legs else None
    if "batched" in legs:
        out["batched"] = batched_leg(hip, configs, r, args.batched_reps)
    if "served" in legs:
        out["served"] = served_leg(hip, configs, r, args.served_requests, inflight=args.served_inflight,
                                   batch=args.served_batch)
    if "mesh_bake" in legs:
        out["mesh_bake"] = mesh_bake_leg(hip, configs, r, args.bake_reps, cpu=not args.no_cpu)
    if "frame_gen" in legs:
        out["frame_gen"] = frame_gen_leg(hip, configs, r)
    if "queued" in legs:
        out["queued"] = queued_leg(hip, configs, g)
    if "stress" in legs:
        out["stress"] = stress_leg(hip, configs, r, args.stress_frames, stress_prep)
    print(json.dumps(out), flush=True)
    g.close()
    r.close()


if __name__ == "__main__":
    main()
