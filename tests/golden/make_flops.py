"""Writes tests/golden/flops.json: algorithmic FP32 work per frame of each benchmark workload,
counted by the instrumented oracle (every add/sub/mul/div/sqrt/min/max/compare = 1 FLOP;
sin/cos/exp/pow/tan listed separately) on the same scene IR and camera the bench renders.

Run from the repository root after building the oracle:  python tests/golden/make_flops.py
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


def main():
    out = {}
    sc, cam = configs.c2_scene()
    out["c2"] = count("c2 cornell gi 960x540", sc, cam)
    sc, cam = configs.c1_scene()
    out["c1"] = count("c1 spheres 320x180", sc, cam)
    path = os.path.join(ROOT, "tests", "golden", "flops.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
