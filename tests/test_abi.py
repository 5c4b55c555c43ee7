"""C-ABI checks that need no GPU: the library loads, exports every symbol the header declares,
validates scenes with the reference's messages, and its host helpers match the oracle."""
import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

import pyoracle as po
from mm3e import configs, hip
from mm3e.scene import Material, Object, Prim, Scene, Transform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mm3e_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mm3e_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(hip.EXPORTS) <= set(syms)


def test_scene_check_messages_match_reference():
    sc, _ = configs.c1_scene(width=32, height=18)
    ir = sc.to_ir()
    assert hip.scene_check(ir, 32, 18) is None
    assert hip.scene_check(ir, 0, 18) == "GpuRenderer::compile: 0x18 out of range 1..=16384"
    assert hip.scene_check(ir, 16385, 4) == "GpuRenderer::compile: 16385x4 out of range 1..=16384"
    s2 = Scene(8, 8)
    s2.add(Object(Prim.surface(0), Transform(), 0))
    assert hip.scene_check(s2.to_ir(), 8, 8) == \
        "native triangle surfaces are not supported by the GPU backend; use the CPU renderer"
    s3 = Scene(8, 8)
    s3.appearance = 1
    assert hip.scene_check(s3.to_ir(), 8, 8) == "UV textures require the checked CPU renderer"
    bad = bytearray(ir)
    bad[0:1] = b"X"
    assert "magic" in hip.scene_check(bytes(bad), 32, 18)
    assert "length" in hip.scene_check(ir[:-4], 32, 18)
    s4 = Scene(8, 8)
    s4.aa = 256  # 65,536 samples per pixel, 256 at a time
    assert hip.scene_check(s4.to_ir(), 8, 8) is None
    s4.aa = 257
    assert "aa" in hip.scene_check(s4.to_ir(), 8, 8)
    s6 = Scene(8, 8)
    s6.bounces = 16
    assert hip.scene_check(s6.to_ir(), 8, 8) is None
    s6.bounces = 17
    assert "bounces" in hip.scene_check(s6.to_ir(), 8, 8)
    s5 = Scene(8, 8)
    s5.post.bloom_radius = 1 << 20  # wide radii take the cache-read pass; only the weights' size is capped
    assert hip.scene_check(s5.to_ir(), 8, 8) is None
    s5.post.bloom_radius = (1 << 20) + 1
    assert "bloom_radius" in hip.scene_check(s5.to_ir(), 8, 8)


def test_renderer_create_fails_loudly_without_a_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible; this test covers the no-device path")
    except Exception:
        pass
    with pytest.raises(hip.Mm3eError) as e:
        hip.Renderer()
    assert "no GPU adapter" in str(e.value)


def test_camera_helpers_match_oracle_bit_for_bit():
    for eye, tgt, fov in (((0, 2, -8), (0, 0, 0), 1.0471976), ((0, 2, 7.5), (0, 1.6, 0), 0.87266463),
                          ((3.3, -1, 2), (0.5, 0.5, -0.5), 0.3)):
        a = hip.camera_look_at(eye, tgt, (0, 1, 0), fov)
        b = np.zeros(13, np.float32)
        P = lambda x: np.asarray(x, np.float32).ctypes.data_as(C.c_void_p)
        po.lib().orc_camera_look_at(P(eye), P(tgt), P((0, 1, 0)), fov, b.ctypes.data_as(C.c_void_p))
        assert a.tobytes() == b.tobytes()
    o = hip.orbit_camera((0, 1, 0), 10.0, 7.5, 3.0, 0.8)
    assert np.isfinite(o).all() and abs(np.linalg.norm(o[3:6]) - 1) < 1e-6


def test_bmp_matches_oracle_encoder():
    rng = np.random.default_rng(3)
    for w, h in ((1, 1), (3, 2), (5, 7), (64, 9)):
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        img[..., 3] = 255
        assert hip.write_bmp(img) == po.to_bmp(img)


def _patch_header_u32(ir: bytes, field_offset: int, values) -> bytes:
    b = bytearray(ir)
    struct.pack_into("<%dI" % len(values), b, field_offset, *values)
    return bytes(b)


def test_volume_and_gi_dims_that_overflow_are_rejected():
    """Volume dims whose product wraps 64 bits, and GI probe counts beyond 2^32, are rejected at
    validation (volume.rs:44, 77-80 check their products) instead of passing a wrapped bound."""
    import numpy as np
    from mm3e.scene import Object, Prim, Scene, Transform, Volume
    sc = Scene(8, 8)
    vid = sc.volume(Volume((2, 2, 2), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), np.zeros(8, np.float32)))
    sc.add(Object(Prim.volume(vid), Transform(), 0))
    ir = sc.to_ir()
    assert hip.scene_check(ir, 8, 8) is None
    vol_at = _ir_offsets(ir)["volumes"]
    for dims in ((1 << 22, 1 << 22, 1 << 20), (0xFFFFFFFF, 0xFFFFFFFF, 2), (3, 2, 2)):
        bad = _patch_header_u32(ir, vol_at, dims)
        assert hip.scene_check(bad, 8, 8) == "SdfVolume samples lie outside the IR volume data", dims
    # GI: a bake request and baked cubes whose probe count exceeds 2^32
    for mode in (1, 2):
        sc2 = Scene(8, 8)
        sc2.add(Object(Prim.sphere(1.0), Transform(), 0))
        if mode == 2:
            sc2.bake_gi((4, 4, 4), 1)
        else:
            sc2.set_baked_gi((0, 0, 0), (1, 1, 1), (2, 2, 2), np.zeros(2 * 2 * 2 * 18, np.float32))
        ir2 = sc2.to_ir()
        assert hip.scene_check(ir2, 8, 8) is None
        bad = _patch_header_u32(ir2, _ir_offsets(ir2)["gi_dims"], (1 << 22, 1 << 22, 1 << 20))
        assert hip.scene_check(bad, 8, 8) == "GI probe volume dims overflow"
        bad = _patch_header_u32(ir2, _ir_offsets(ir2)["gi_dims"], (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF))
        assert hip.scene_check(bad, 8, 8) == "GI probe volume dims overflow"


def _ir_offsets(ir: bytes) -> dict:
    """Byte offsets inside an IR blob: of the first volume record and of the header's gi_dims."""
    # fixed header words, in Scene.to_ir order: magic(8) version total | 8 I | 4x(3f f) |
    # marcher 12 words | post 8 words | counts 8 words | gi dims 3 + off | gi min 4 | gi size 4 |
    # offsets 8 words
    pos = 8 + 8 + 32 + 64 + 48 + 32 + 32
    gi_dims = pos
    pos += 16 + 16 + 16
    offs = struct.unpack_from("<8I", ir, pos)
    return {"gi_dims": gi_dims, "volumes": offs[3]}
