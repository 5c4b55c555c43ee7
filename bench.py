"""Benchmark of the MI355X per-pixel SDF render path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one frame of the 960x540 Cornell GI workload (seed 2: five box walls, three
primitives incl. a smooth union and a twist, emissive bar, sphere light, 16^3 GI probes baked
on the device, soft shadows, AO, one reflection bounce, bloom + tonemap, aa = 2) rendered by
the plain per-frame launch path (`render_only`: render kernel + bloom/tonemap kernels) with
the scene and framebuffers resident in HBM.  W untimed frames, then K timed frames between
two device synchronisations.  One JSON line on stdout; progress on stderr.

roofline: the render kernel's algorithmic FP32 work per frame (tests/golden/flops.json,
counted by the instrumented oracle on the same scene and camera) divided by its mean
HIP-event duration over the timed frames, against MI355X's vector FP32 peak.
cpu_baseline: the CPU oracle (a C++ restatement of the reference's multithreaded CPU
renderer) timed on this host on one full frame of the same workload.
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


def cpu_baseline(ir, cam, threads, frames=2):
    """Oracle CPU renderer (row-band threads, the reference's render_with_threads) on one full
    frame; 1 untimed warm-up frame, then `frames` timed frames."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    s = po.OracleScene(ir, threads=threads)
    s.render_rgba(cam, threads=threads)
    t0 = time.perf_counter()
    for _ in range(frames):
        s.render_rgba(cam, threads=threads)
    dt = (time.perf_counter() - t0) / frames
    return s.width * s.height / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
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
        "bound": "valu", "kernel": "mm3e_render_table_b1", "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "frac_nonpacked_peak": round(achieved / PEAK_FP32_NONPACKED_TFLOPS, 4),
        "flops_per_launch": fl["flops_per_frame"], "kernel_ms": round(render_ms, 4),
        "traffic": None,
    }
    prof = os.path.join(ROOT, "profiles", "render_c2_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2: 960x540 Cornell room, 3 seeded primitives (smooth k=0.3, twist), 16^3 GI "
                               "probes baked on device (64 requested -> 30 rays/probe), soft shadows, AO, "
                               "1 reflection bounce, bloom+tonemap, aa=2; plain per-frame launch",
                   "width": W, "height": H, "aa": sc.aa, "bounces": sc.bounces, "path": "plain",
                   "compile_s": round(compile_s, 3), "render_kernel_ms": round(render_ms, 4),
                   "post_ms": round(post_ms, 4)},
        "roofline": roofline,
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        log(f"cpu baseline: oracle on {threads} threads, one full frame (1 warm-up + 2 timed)")
        cv, cdt = cpu_baseline(ir, cam, threads)
        out["cpu_baseline"] = {"value": round(cv, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
                               "sample": f"full c2 960x540 frame (aa=2), 2 timed frames after 1 warm-up, "
                                         f"{cdt:.2f} s/frame"}
        log(f"cpu baseline {cv:.4f} Mpix/s ({cdt:.2f} s/frame)")
    print(json.dumps(out), flush=True)
    g.close()
    r.close()


if __name__ == "__main__":
    main()
