"""Image-parity metrics between the HIP render path and the CPU oracle."""
from __future__ import annotations

import numpy as np


def image_metrics(gpu_rgba: np.ndarray, ref_rgba: np.ndarray, tol: int = 2) -> dict:
    """Per-pixel 8-bit comparison: fraction of pixels whose three channels are all within
    `tol`/255, mean absolute error per channel, max error, and the count of pixels off."""
    g = gpu_rgba[..., :3].astype(np.int32)
    r = ref_rgba[..., :3].astype(np.int32)
    d = np.abs(g - r)
    within = (d <= tol).all(axis=-1)
    return {
        "frac_within": float(within.mean()),
        "mae": float(d.mean()),
        "max": int(d.max()) if d.size else 0,
        "pixels_off": int((~within).sum()),
        "pixels": int(within.size),
    }


def aov_metrics(g_hit, g_mat, r_hit, r_mat) -> dict:
    """Hit-mask / material-id agreement; a silhouette mismatch is a pixel whose hit flag
    differs (the ray grazed an edge on one side only)."""
    hit_agree = g_hit == r_hit
    mat_agree = hit_agree & ((g_mat == r_mat) | (r_hit == 0))
    return {
        "hit_agree": float(hit_agree.mean()),
        "mat_agree": float(mat_agree.mean()),
        "silhouette_mismatch": int((~hit_agree).sum()),
        "material_mismatch": int((hit_agree & ~mat_agree).sum()),
    }
