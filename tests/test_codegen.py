"""Scene -> HIP code generation (the twin of the reference's WGSL generator, mm3e-gpu/src/wgsl.rs).

The generated kernel of a scene folds the same objects through the same device helpers as the
table-driven kernels, with the scene's constants compiled in, so its images must equal theirs
bit for bit; the table-driven kernels are in turn held to the oracle by test_gpu_parity.py.
The CPU tests check the emitted source itself.
"""
import numpy as np
import pytest

from mm3e import configs, hip


def _scenes():
    sc1, cam1 = configs.c1_scene()
    sc2, cam2 = configs.c2_scene(width=480, height=270)
    p60 = configs.procedural_scene(41, 60, 320, 180)
    p40 = configs.procedural_scene(42, 40, 320, 180)
    csg, camc = configs.csg_scene(320, 180)
    refl = configs.procedural_scene(43, 40, 240, 135)
    refl.bounces, refl.fog_density, refl.aa = 2, 0.03, 2
    p20 = configs.procedural_scene(40, 20, 320, 180)  # segment grid, nested generated kernel
    p180 = configs.procedural_scene(42, 180, 160, 90)  # long fold, state machine at 5 waves
    return [("c1", sc1, cam1), ("c2", sc2, cam2),
            ("procedural180", p180, configs.orbit_pose(configs.SplitMix64(42))),
            ("procedural20", p20, configs.orbit_pose(configs.SplitMix64(40))),
            ("procedural60", p60, configs.orbit_pose(configs.SplitMix64(41))),
            ("procedural40", p40, configs.orbit_pose(configs.SplitMix64(42))),
            ("csg", csg, camc), ("bounces2", refl, configs.orbit_pose(configs.SplitMix64(43)))]


def test_generated_kernel_family():
    """Up to 24 objects the generated kernel takes the nested path, segment grid or not,
    register-allocated for 7 waves per SIMD up to 16 objects and 6 above; above 24, the state
    machine, for 4 waves per SIMD (5 above 48 objects, 7 above 200)."""
    c2 = hip.emit_source(configs.c2_scene(width=64, height=36, gi_dims=None)[0].to_ir())
    assert "render_ordered<GenMap, 1, false>" in c2 and "#define MM3E_RENDER_WAVES 7" in c2
    small = hip.emit_source(configs.procedural_scene(40, 20, 64, 36).to_ir())
    assert "grid_select" in small and "render_ordered<GenMap, 1, false>" in small
    assert "#define MM3E_RENDER_WAVES 6" in small
    mid = hip.emit_source(configs.procedural_scene(42, 40, 64, 36).to_ir())
    assert "render_ordered<GenMap, 1, true>" in mid and "#define MM3E_RENDER_WAVES 4" in mid
    big = hip.emit_source(configs.procedural_scene(41, 60, 64, 36).to_ir())
    assert "#define MM3E_RENDER_WAVES 5" in big
    huge = hip.emit_source(configs.procedural_scene(43, 240, 64, 36).to_ir())
    assert "render_ordered<GenMap, 1, true>" in huge and "#define MM3E_RENDER_WAVES 7" in huge


def test_emitted_source_names_every_segment():
    sc = configs.procedural_scene(42, 180, 64, 36)
    src = hip.emit_source(sc.to_ir())
    assert "mm3e_gen_render" in src and "struct GenMap" in src
    assert "__builtin_bit_cast(float" in src  # constants travel as exact bit patterns
    sc1, _ = configs.c1_scene()
    assert hip.emit_source(sc1.to_ir()) != src


@pytest.mark.gpu
def test_generated_kernels_equal_table_kernels():
    """Generated kernels, dispatched in index order and most-expensive-tile first (the order is
    rebuilt from each frame's tile costs, so three frames see three orders), equal the table
    kernels bit for bit."""
    table = hip.Renderer()
    table.set_codegen(hip.CODEGEN_OFF)
    gen = hip.Renderer()
    gen.set_codegen(hip.CODEGEN_SYNC)
    gen.set_tile_order(False)
    ordered = hip.Renderer()
    ordered.set_codegen(hip.CODEGEN_SYNC)
    for name, sc, cam in _scenes():
        ir = sc.to_ir()
        gt = table.compile(ir, sc.width, sc.height)
        gg = gen.compile(ir, sc.width, sc.height)
        go = ordered.compile(ir, sc.width, sc.height)
        for g in (gg, go):
            state, msg = g.codegen_status(wait=True)
            assert state == hip.CODEGEN_READY, msg
        # one generated kernel serves both dispatch orders
        assert gg.kernel_name() == "mm3e_gen_render_o"
        assert go.kernel_name() == "mm3e_gen_render_o"
        a, b = gt.render_rgba(cam), gg.render_rgba(cam)
        frames = [go.render_rgba(cam) for _ in range(3)]
        print(f"\n[codegen {name}] kernel={gt.kernel_name()} equal={bool((a == b).all())} "
              f"ordered_equal={[bool((a == f).all()) for f in frames]}", flush=True)
        assert (a == b).all(), name
        for f in frames:
            assert (a == f).all(), name
        gt.close()
        gg.close()
        go.close()
    table.close()
    gen.close()
    ordered.close()


@pytest.mark.gpu
def test_scenes_over_the_object_cap_keep_the_table_kernels(monkeypatch):
    """MM3E_CODEGEN_MAX_OBJECTS lowers the object cap of code generation (default 2000): a
    180-object scene over a cap of 100 gets no generated kernel and renders with the table walk."""
    monkeypatch.setenv("MM3E_CODEGEN_MAX_OBJECTS", "100")
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_SYNC)
    g = r.compile(configs.procedural_scene(42, 180, 64, 36).to_ir(), 64, 36)
    assert g.codegen_status(wait=True)[0] == hip.CODEGEN_NONE
    assert g.kernel_name().startswith("mm3e_render_")
    g.close()
    r.close()


@pytest.mark.gpu
def test_background_codegen_switches_without_changing_pixels():
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_BACKGROUND)
    sc = configs.procedural_scene(44, 48, 256, 144)
    cam = configs.orbit_pose(configs.SplitMix64(44))
    g = r.compile(sc.to_ir(), 256, 144)
    before = g.render_rgba(cam)          # table kernels while the generated one compiles
    state, msg = g.codegen_status(wait=True)
    assert state == hip.CODEGEN_READY, msg
    after = g.render_rgba(cam)
    assert g.kernel_name() == "mm3e_gen_render_o"
    assert (before == after).all()
    g.close()
    r.close()


@pytest.mark.gpu
def test_generated_film_and_gbuffer_equal_table_kernels():
    """With mm3e_renderer_set_codegen_film, a scene's generated module also holds the film and
    G-buffer / hole kernels; film outputs (pinhole and thin lens), G-buffer frames and hybrid
    reprojections must equal the table-driven kernels' bit for bit."""
    import math
    table = hip.Renderer()
    table.set_codegen(hip.CODEGEN_OFF)
    gen = hip.Renderer()
    gen.set_codegen(hip.CODEGEN_SYNC)
    gen.set_codegen_film(True)
    sc2, cam2 = configs.c2_scene(width=480, height=270)
    sc1, cam1 = configs.c1_scene()
    csg, camc = configs.csg_scene(320, 180)
    for name, sc, cam in [("c1", sc1, cam1), ("c2", sc2, cam2), ("csg", csg, camc)]:
        ir = sc.to_ir()
        gt = table.compile(ir, sc.width, sc.height)
        gg = gen.compile(ir, sc.width, sc.height)
        state, msg = gg.codegen_status(wait=True)
        assert state == hip.CODEGEN_READY, msg
        for radius in (0.0, 0.08):
            ft = gt.render_film(cam, aperture_radius=radius, focus_distance=5.0)
            fg = gg.render_film(cam, aperture_radius=radius, focus_distance=5.0)
            for k in ft:
                same = ft[k].view(np.uint32) == fg[k].view(np.uint32)
                assert same.all(), (name, radius, k, int((~same).sum()))
        frames = []
        for r, g in ((table, gt), (gen, gg)):
            gf = hip.GFrame(r, sc.width, sc.height)
            g.render_gbuffer(cam, gf)
            a = gf.read()
            cam_b = configs.orbit_camera((0.0, 1.0, 0.0), 6.0, 0.2, 0.3, math.radians(50.0))
            rgba, holes = g.reproject_hybrid(gf, cam_b)
            frames.append((a, rgba, holes))
            gf.close()
        (a0, r0, h0), (a1, r1, h1) = frames
        print(f"\n[codegen film/gbuffer {name}] color_equal={bool((a0['color'] == a1['color']).all())} "
              f"hybrid_equal={bool((r0 == r1).all())} holes={h0}/{h1}", flush=True)
        assert (a0["color"] == a1["color"]).all() and (a0["obj"] == a1["obj"]).all()
        assert (a0["depth"].view(np.uint32) == a1["depth"].view(np.uint32)).all()
        assert h0 == h1 and (r0 == r1).all()
        gt.close()
        gg.close()
    table.close()
    gen.close()
