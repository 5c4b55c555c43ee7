"""Writes tests/golden/flops.json: algorithmic FP32 work per frame of each benchmark workload,
counted by the instrumented oracle (every add/sub/mul/div/sqrt/min/max/compare = 1 FLOP;
sin/cos/exp/pow/tan listed separately) on the same scene IR and camera the bench renders.

c3 (the 64-camera batch) and c5 (the 3840x2160 stress frame) are counted on row samples and
scaled; each record states its sample.

Run from the repository root after building the oracle:
    python tests/golden/make_flops.py [c1] [c2] [c3] [c5]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "atom-3d-engine_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from mm3e import configs  # noqa: E402


def count(name, sc, cam):
    t0 = time.time()
    s = po.OracleScene(sc.to_ir())
    c = s.count_rows(cam, 0, sc.height)
    px = sc.width * sc.height
    rec = {
        "workload": name, "width": sc.width, "height": sc.height, "aa": sc.aa, "bounces": sc.bounces,
        "flops_per_frame": c["flops"], "trans_per_frame": c["trans"], "sqrt_per_frame": c["sqrt"],
        "div_per_frame": c["div"], "flops_per_pixel": c["flops"] / px,
        "field_evals": c["evals"], "field_evals_per_pixel": sum(c["evals"].values()) / px,
        "object_sdf_evals": c["obj_exact"], "bound_substitutions": c["obj_bound"],
        "count_seconds": round(time.time() - t0, 1),
    }
    print(json.dumps(rec), flush=True)
    return rec


def count_sampled(s, cam, width, height, stride, threads):
    """Counts every `stride`-th row (offset stride // 2) of a frame, one row per call on a pool
    of host threads, and scales the sums by height / rows counted."""
    from concurrent.futures import ThreadPoolExecutor
    rows = list(range(stride // 2, height, stride))
    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(lambda y: s.count_rows(cam, y, y + 1, threads=1), rows))
    scale = height / len(rows)
    tot = {k: sum(p[k] for p in parts) * scale for k in ("flops", "trans", "sqrt", "div", "obj_exact", "obj_bound")}
    tot["evals"] = {k: sum(p["evals"][k] for p in parts) * scale for k in parts[0]["evals"]}
    return tot, len(rows) * width


def count_c3(stride=8, threads=None):
    """The 64-camera batch: each job counted on a row sample of its own size, summed."""
    threads = threads or os.cpu_count() or 1
    t0 = time.time()
    jobs = configs.c3_jobs()
    loaded = {}
    flops = 0.0
    px = 0
    sampled = 0
    evals = 0.0
    per_job = []
    for sc, cam, (w, h) in jobs:
        key = (id(sc), w, h)
        if key not in loaded:
            sc.width, sc.height = w, h
            loaded[key] = po.OracleScene(sc.to_ir())
        tot, n = count_sampled(loaded[key], cam, w, h, stride, threads)
        flops += tot["flops"]
        evals += sum(tot["evals"].values())
        px += w * h
        sampled += n
        per_job.append(round(tot["flops"]))
    rec = {"workload": "c3 batch: 64 cameras over 8 procedural scenes (seeds 3-10) at their own sizes",
           "pixels": px, "flops_per_frame": round(flops), "flops_per_pixel": flops / px,
           "field_evals_per_pixel": evals / px, "flops_per_job": per_job,
           "sample": f"every {stride}th row of every job ({sampled} of {px} pixels), scaled by rows",
           "count_seconds": round(time.time() - t0, 1)}
    print(json.dumps({k: v for k, v in rec.items() if k != "flops_per_job"}), flush=True)
    return rec


def count_c5(stride=32, threads=None):
    """The 3840x2160 stress frame (32^3 GI baked by the oracle at load) on a row sample."""
    threads = threads or os.cpu_count() or 1
    t0 = time.time()
    def bake(pos, tri, res, pad):  # bit-identical to the device bake (tests/test_mesh_bake.py)
        return po.mesh_bake(pos, tri, res, pad, threads=threads)[:4]
    sc, cam = configs.c5_scene(bake=bake)
    s = po.OracleScene(sc.to_ir(), threads=threads)
    t1 = time.time()
    tot, n = count_sampled(s, cam, sc.width, sc.height, stride, threads)
    px = sc.width * sc.height
    rec = {"workload": "c5 stress 3840x2160, 1000 primitives, the star mesh's SDF grid (128 along its longest axis), "
                       "32^3 GI, 2 bounces, fog",
           "width": sc.width, "height": sc.height, "aa": sc.aa, "bounces": sc.bounces,
           "flops_per_frame": round(tot["flops"]), "trans_per_frame": round(tot["trans"]),
           "sqrt_per_frame": round(tot["sqrt"]), "div_per_frame": round(tot["div"]),
           "flops_per_pixel": tot["flops"] / px, "field_evals": {k: round(v) for k, v in tot["evals"].items()},
           "field_evals_per_pixel": sum(tot["evals"].values()) / px,
           "sample": f"every {stride}th row ({n} of {px} pixels), scaled by rows",
           "gi_bake_seconds": round(t1 - t0, 1), "count_seconds": round(time.time() - t1, 1)}
    print(json.dumps(rec), flush=True)
    return rec


def main():
    which = set(sys.argv[1:]) or {"c1", "c2", "c3", "c5"}
    path = os.path.join(ROOT, "tests", "golden", "flops.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    if "c2" in which:
        sc, cam = configs.c2_scene()
        out["c2"] = count("c2 cornell gi 960x540", sc, cam)
    if "c1" in which:
        sc, cam = configs.c1_scene()
        out["c1"] = count("c1 spheres 320x180", sc, cam)
    if "c3" in which:
        out["c3"] = count_c3()
    if "c5" in which:
        out["c5"] = count_c5()
    path = os.path.join(ROOT, "tests", "golden", "flops.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
