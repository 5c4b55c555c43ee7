"""The `.mm3e` text scene format (mm3e-orchestrator/src/scene_io.rs) through the C ABI
(mm3e_scene_parse_text / mm3e_scene_to_text).  Ports the reference's own tests
(mm3e-orchestrator/tests/engine.rs:10-80) and checks that a parsed scene renders exactly like
the scene it was written from (oracle, CPU)."""
import math

import numpy as np
import pytest

import pyoracle as po
from mm3e import configs, hip
from mm3e.scene import (MODE_STEPS, Light, Material, Object, Prim, Scene, Transform, normalize, v3)


def demo_scene():
    """engine.rs:10-27 demo_scene."""
    sc = Scene(96, 64)
    sc.aa = 1
    floor = sc.material(Material.solid((1.0, 1.0, 1.0)).checkered().roughness(0.6))
    red = sc.material(Material.solid((0.85, 0.2, 0.2)).roughness(0.3))
    metal = sc.material(Material.solid((0.9, 0.8, 0.4)).metallic(1.0).roughness(0.2).reflective(0.4))
    sc.add(Object(Prim.plane((0.0, 1.0, 0.0), 0.0), Transform(), floor))
    sc.add(Object(Prim.sphere(1.0), Transform.at((-1.2, 1.0, 0.0)), red).round(0.1))
    sc.add(Object(Prim.torus(0.7, 0.25), Transform.at((1.2, 1.0, 0.0)), metal))
    sc.add(Object(Prim.box((0.6, 0.6, 0.6)), Transform.at((0.0, 0.6, 1.6)), red).onion(0.05))
    sc.sun_dir = normalize((0.5, 0.75, 0.4))
    sc.light(Light.directional_light(sc.sun_dir, (2.0, 2.0, 2.0)).soft(0.04))
    sc.light(Light.sphere((-3.0, 4.0, 3.0), (20.0, 20.0, 20.0), 1.0))
    return sc


def cam():
    return configs.orbit_camera((0.0, 0.8, 0.3), 7.0, 0.5, 0.3, math.radians(52.0))


def test_scene_io_roundtrips_stably():  # engine.rs:32-48
    sc = demo_scene()
    sc.mode = MODE_STEPS
    c = cam()
    text = hip.to_text(sc.to_ir(), c)
    ir, _ = hip.parse_text(text)
    again = hip.to_text(ir, c)
    assert text == again
    assert text.count("\nobj ") == 4 and text.count("\nmat ") == 4 and text.count("\nlight ") == 2
    assert "\nmode steps\n" in text


def test_camera_roundtrips():  # engine.rs:50-60
    c = cam()
    _, pc = hip.parse_text(hip.to_text(demo_scene().to_ir(), c))
    for a, b in ((slice(0, 3),) * 2, (slice(3, 6),) * 2, (slice(9, 12),) * 2):
        assert np.linalg.norm(pc[a] - c[b]) < 1e-4
    assert abs(pc[12] - c[12]) < 1e-4


@pytest.mark.parametrize("text", [
    "size 10\n", "obj banana 1 2 3\n", "wat 1 2 3\n", "size 0 100\n", "size 100000 100\n", "sun 0 0 0\n",
    "ambient NaN 0 0\n", "cam 1 2 3 1 2 3 0 1 0 50\n", "cam 0 0 0 0 1 0 0 1 0 50\n", "cam 0 0 5 0 0 0 0 1 0 180\n",
    "marcher 0 120 0.0006 1\n", "marcher 160 120 -1 1\n",
    "obj sphere -1 pos 0 0 0 basis 1 0 0 0 1 0 0 0 1 scale 1 mat 0 combine union\n",
    "obj plane 0 0 0 0 pos 0 0 0 basis 1 0 0 0 1 0 0 0 1 scale 1 mat 0 combine union\n",
    "obj sphere 1 pos 0 0 0 basis 1 0 0 0 1 0 0 0 1 scale 0 mat 0 combine union\n",
    "obj sphere 1 mat 3\n", "mode sideways\n", "light spot 0 1 0 1 1 1 0\n",
])
def test_scene_io_rejects_garbage(text):  # engine.rs:62-80
    with pytest.raises(hip.Mm3eError):
        hip.parse_text(text)


def test_error_messages_name_the_line():
    with pytest.raises(hip.Mm3eError) as e:
        hip.parse_text("# comment\nsize 64 32\nobj banana 1\n")
    assert str(e.value) == "line 3: unknown primitive 'banana'"
    with pytest.raises(hip.Mm3eError) as e:
        hip.parse_text("size 64 32\nobj sphere 1 mat 2\n")
    assert str(e.value) == "object 1: material index 2 out of range (have 1)"


def test_defaults_and_comments():
    ir, c = hip.parse_text("# only a size\nsize 40 30   # trailing comment\n")
    text = hip.to_text(ir, c)
    assert "size 40 30\n" in text and "aa 2\n" in text and "marcher 160 120 0.0006 1\n" in text
    assert "post 1 1 1 0.06 12\n" in text and "\nmat 0.8 0.8 0.8 0 0.35 0 0.5 0 0 0 0\n" in text


def test_parsed_scene_renders_like_the_original():
    """Every field the format carries round-trips bit for bit, so the parsed scene renders the
    same image (oracle) — the reference's scene_file example renders the parsed scene."""
    sc = demo_scene()
    sc.width, sc.height = 48, 32
    c = cam()
    ir0 = sc.to_ir()
    ir1, c1 = hip.parse_text(hip.to_text(ir0, c))
    a = po.OracleScene(ir0, threads=4).render_rgba(c, threads=4)
    b = po.OracleScene(ir1, threads=4).render_rgba(c, threads=4)
    assert np.array_equal(a, b)
    p = configs.procedural_scene(48, 40, 48, 32)  # every analytic kind and modifier the format holds
    p.objects = [o for o in p.objects]
    c2 = configs.orbit_pose(configs.SplitMix64(48))
    ir2, _ = hip.parse_text(hip.to_text(p.to_ir(), c2))
    assert np.array_equal(po.OracleScene(p.to_ir(), threads=4).render_rgba(c2, threads=4),
                          po.OracleScene(ir2, threads=4).render_rgba(c2, threads=4))


# ---- the lossless extension (CSG, volumes, GI, marcher budgets, normal_h) -----------------

def _baked_gi_scene():
    sc, _ = configs.c2_scene(width=40, height=24, gi_dims=(3, 3, 3), gi_rays=1)
    g = po.OracleScene(sc.to_ir()).gi_raw()
    sc.set_baked_gi(g[0], g[1], g[2], g[3])
    return sc


def _ext_cases():
    c1, cam1 = configs.c1_scene()
    c2, cam2 = configs.c2_scene()
    c5, cam5 = configs.c5_scene()
    csg, camc = configs.csg_scene()
    knobs, _ = configs.c1_scene(width=64, height=36)
    knobs.marcher.shadow_steps, knobs.marcher.ao_samples = 17, 3
    knobs.marcher.lod_footprint, knobs.marcher.subitize, knobs.marcher.secant = 0.002, 0.6, True
    knobs.marcher.normal_h = 0.0013
    knobs.sun_dir = (0.3, 0.8, 0.2)  # not unit length: kept as stored
    return {"c1": (c1, cam1), "c2": (c2, cam2), "c5": (c5, cam5), "csg": (csg, camc), "knobs": (knobs, cam1),
            "baked_gi": (_baked_gi_scene(), cam2)}


@pytest.mark.parametrize("name", ["c1", "c2", "c5", "csg", "knobs", "baked_gi"])
def test_ir_text_ir_is_bit_identical(name):
    """IR -> text -> IR returns the same bytes for the benchmark scenes (c2's GI bake request,
    c5's 128^3 volume and GI request), object-local CSG, a baked GI volume and every marcher
    knob; text -> IR -> text is stable too."""
    sc, c = _ext_cases()[name]
    ir0 = sc.to_ir()
    text = hip.to_text(ir0, c)
    ir1, _ = hip.parse_text(text)
    assert ir1 == ir0
    assert hip.to_text(ir1, c) == text


def test_reference_subset_is_written_in_the_reference_format():
    text = hip.to_text(demo_scene().to_ir(), cam())
    for ext in ("sun_raw", "marcher_ext", "\ngi ", "\nvolume ", "csgnode", "csgexpr", "obj csg", "obj volume"):
        assert ext not in text


def test_extension_directives_in_text():
    sc, c = _ext_cases()["c5"]
    text = hip.to_text(sc.to_ir(), c)
    assert "\ngi bake 32 32 32 5\n" in text and "\nvolume sdfv " in text and "\nobj volume 0 " in text
    assert "\nmarcher_ext 64 5 0 0 0 " in text  # normal_h raised by the volume (lib.rs:631-638)
    sc, c = _ext_cases()["csg"]
    text = hip.to_text(sc.to_ir(), c)
    assert text.count("\ncsgexpr ") == len(sc.csg_roots) and text.count("\nobj csg ") == 4


@pytest.mark.parametrize("text,msg", [
    ("gi sideways 1 1 1 1\n", "line 1: unknown gi kind 'sideways'"),
    ("gi baked 1 1 1 0 0 0 1 1 1 AAAA\n", "line 1: gi cubes must be base64 of nx*ny*nz*18 f32"),
    ("volume raw 2 2 2 0 0 0 1 1 1 !!!!\n", "line 1: volume samples must be base64 of nx*ny*nz f32"),
    ("volume sdfv QUJD\n", "line 1: sdfv buffer is shorter than the header"),
    ("volume cube 1\n", "line 1: unknown volume kind 'cube'"),
    ("csgnode 0 0 0 0 1\n", "line 1: expected csg p"),
])
def test_extension_rejects_garbage(text, msg):
    with pytest.raises(hip.Mm3eError) as e:
        hip.parse_text(text)
    assert str(e.value) == msg


def test_extension_ids_are_checked_at_compile():
    """A csg id without a registered expression parses (ids resolve against the whole file) and
    is then refused by the compile-time validation, as for an IR built any other way."""
    ir, _ = hip.parse_text("size 8 8\nobj csg 0 mat 0\n")
    assert hip.scene_check(ir, 8, 8) == "unregistered CSG expression"


def test_text_scene_with_csg_and_volume_renders_like_the_original():
    sc, c = configs.csg_scene(48, 32)
    ir0 = sc.to_ir()
    ir1, _ = hip.parse_text(hip.to_text(ir0, c))
    assert np.array_equal(po.OracleScene(ir0, threads=4).render_rgba(c, threads=4),
                          po.OracleScene(ir1, threads=4).render_rgba(c, threads=4))


@pytest.mark.gpu
def test_server_text_requests_equal_compiled_ir_renders(renderer):
    """Requests given as `.mm3e` text (GI bake request, object-local CSG, a sampled volume) are
    served like requests on a compiled IR: the image equals mm3e_scene_render_rgba of the parsed
    IR and camera, bit for bit; a second request with the same scene and another camera reuses
    the resident compiled scene."""
    c2, cam2 = configs.c2_scene(width=240, height=135, gi_dims=(4, 4, 4), gi_rays=5)
    csg, camc = configs.csg_scene(160, 90)
    c5, cam5 = configs.c5_scene(width=160, height=90, n_prims=60, gi_dims=(4, 4, 4), vol_n=32)
    srv = hip.Server(renderer, hbm_budget_bytes=128 << 20)
    try:
        for sc, c in ((c2, cam2), (csg, camc), (c5, cam5)):
            text = hip.to_text(sc.to_ir(), c)
            ir, pc = hip.parse_text(text)
            assert ir == sc.to_ir()
            g = renderer.compile(ir, sc.width, sc.height)
            try:
                ref = g.render_rgba(pc)
                img, _ = srv.wait(srv.submit_text(text))
                assert (img == ref).all()
                # the same scene from another viewpoint: only the cam line changes
                other = configs.orbit_camera((0.0, 1.0, 0.0), 9.0, 0.7, 0.3, math.radians(50.0))
                text2 = hip.to_text(ir, other)
                ir2, pc2 = hip.parse_text(text2)
                img2, _ = srv.wait(srv.submit_text(text2))
                assert (img2 == g.render_rgba(pc2)).all()
            finally:
                g.close()
        with pytest.raises(hip.Mm3eError) as e:
            srv.submit_text("size 8 8\nobj banana 1\n")
        assert "unknown primitive 'banana'" in str(e.value)
    finally:
        srv.close()
