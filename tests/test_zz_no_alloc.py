"""mm3e_scene_render_rgba_noalloc (a frame whose HDR and bloom intermediates live in the
library's fixed device array) against mm3e_scene_render_rgba on the same scene and camera:
equal bit for bit, on two calls in a row (the second a smaller frame), through the C ABI.
Device output buffers come from the HIP runtime the library itself links."""
import ctypes as C

import numpy as np
import pytest

from mm3e import configs, hip

pytestmark = pytest.mark.gpu


def _hip_runtime():
    L = C.CDLL("libamdhip64.so.7")  # the instance libmm3e_hip.so already loaded
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return L


D2H = 2


def _frame(rt, r, g, cam):
    n = g.width * g.height * 4
    ptr = C.c_void_p()
    assert rt.hipMalloc(C.byref(ptr), n) == 0
    try:
        assert rt.hipMemset(ptr, 0x5A, n) == 0
        g.render_rgba_noalloc(cam, ptr.value, n)
        r.wait_idle()
        out = np.zeros((g.height, g.width, 4), np.uint8)
        assert rt.hipMemcpy(out.ctypes.data, ptr, n, D2H) == 0
    finally:
        rt.hipFree(ptr)
    return out


@pytest.mark.parametrize("codegen", [hip.CODEGEN_OFF, hip.CODEGEN_SYNC])
def test_noalloc_frames_equal_render_rgba(codegen):
    rt = _hip_runtime()
    r = hip.Renderer()
    r.set_codegen(codegen)
    try:
        big, cam_b = configs.c2_scene(width=480, height=270, gi_dims=(6, 6, 6), gi_rays=5)
        small, cam_s = configs.c1_scene(width=200, height=112)
        for sc, cam in ((big, cam_b), (small, cam_s)):
            g = r.compile(sc.to_ir(), sc.width, sc.height)
            try:
                if codegen == hip.CODEGEN_SYNC:
                    assert g.codegen_status(wait=True)[0] == hip.CODEGEN_READY
                ref = g.render_rgba(cam)
                got = _frame(rt, r, g, cam)
                assert (got == ref).all(), f"{sc.width}x{sc.height}: noalloc frame differs from render_rgba"
                assert (got[..., 3] == 255).all()
                # the scene's own framebuffers are untouched by the noalloc frame
                assert (g.render_rgba(cam) == ref).all()
            finally:
                g.close()
    finally:
        r.close()


def test_noalloc_rejects_frames_beyond_the_scratch():
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_OFF)
    try:
        sc, cam = configs.c1_scene(width=2048, height=1366)  # 2,797,568 pixels > 2,796,202
        g = r.compile(sc.to_ir(), sc.width, sc.height)
        try:
            with pytest.raises(hip.Mm3eError) as e:
                g.render_rgba_noalloc(cam, 1 << 40, sc.width * sc.height * 4)  # rejected before any use
            assert "frame scratch" in str(e.value) and "2796202 pixels" in str(e.value)
            with pytest.raises(hip.Mm3eError):
                g.render_rgba_noalloc(cam, 1 << 40, 16)  # output too small
        finally:
            g.close()
    finally:
        r.close()
