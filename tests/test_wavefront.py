"""Queued frames: the frame split into budgeted march passes over compacted queues of unfinished
rays, shading of the compacted hit queue and a per-pixel resolve (dev_wavefront.h).

A resumed march takes exactly the steps it would have taken in one go, and shading reads the
same hit record, so a queued frame must equal the one-kernel frame bit for bit (HDR and RGBA8)
for every budget, pass count and grid size, on both kernel families (nested loops, state
machine), with table-driven and generated kernels, with dynamic spheres.  The one-kernel frame
is held to the oracle by test_gpu_parity.py.
"""
import numpy as np
import pytest

from mm3e import configs, hip


def _scenes():
    sc1, cam1 = configs.c1_scene()
    sc2, cam2 = configs.c2_scene(width=480, height=270)
    csg, camc = configs.csg_scene(320, 180)
    p60 = configs.procedural_scene(41, 60, 320, 180)
    refl = configs.procedural_scene(43, 40, 240, 135)
    refl.bounces, refl.fog_density, refl.aa = 2, 0.03, 2
    return [("c1", sc1, cam1), ("c2", sc2, cam2), ("csg", csg, camc),
            ("procedural60", p60, configs.orbit_pose(configs.SplitMix64(41))),
            ("bounces2", refl, configs.orbit_pose(configs.SplitMix64(43)))]


# (budget, passes, blocks_per_cu): tiny budgets force many rays through several queues; passes
# 0 leaves only the hit-queue compaction; a small grid makes each lane take several entries.
_SETTINGS = [(1, 7, 0), (8, 2, 0), (24, 2, 2), (48, 1, 8), (1000, 0, 0)]


@pytest.mark.gpu
def test_set_wavefront_rejects_out_of_range():
    r = hip.Renderer()
    with pytest.raises(hip.Mm3eError):
        r.set_wavefront(8, passes=8)
    with pytest.raises(hip.Mm3eError):
        r.set_wavefront(8, blocks_per_cu=65)
    # passes + 1 must not wrap: -1 reaches the library as 0xFFFFFFFF
    for bad in (-1, 0xFFFFFFFF, 0xFFFFFFFE):
        with pytest.raises(hip.Mm3eError):
            r.set_wavefront(8, passes=bad)
    sc, _ = configs.c1_scene(width=16, height=9)
    g = r.compile(sc.to_ir(), 16, 9)
    for bad in (-1, 0xFFFFFFFF, 8):
        with pytest.raises(hip.Mm3eError):
            g.set_wavefront(8, passes=bad)
    g.close()
    r.close()


@pytest.mark.parametrize("value,queued", [("8,1", True), ("8,7", True), ("8,8", False), ("8,-1", False),
                                          ("8,4294967295", False), ("8,1,65", False)])
def test_env_wavefront_range_is_checked(monkeypatch, value, queued):
    """MM3E_WAVEFRONT out of range is ignored (no queued stages in the emitted source); CPU only."""
    sc, _ = configs.c1_scene(width=16, height=9)
    monkeypatch.setenv("MM3E_WAVEFRONT", value)
    assert ("mm3e_gen_wf_" in hip.emit_source(sc.to_ir())) == queued


@pytest.mark.gpu
def test_queued_frames_equal_one_kernel_frames_table_kernels():
    base = hip.Renderer()
    base.set_codegen(hip.CODEGEN_OFF)
    base.set_tile_order(False)
    q = hip.Renderer()
    q.set_codegen(hip.CODEGEN_OFF)
    for name, sc, cam in _scenes():
        ir = sc.to_ir()
        gb = base.compile(ir, sc.width, sc.height)
        gq = q.compile(ir, sc.width, sc.height)
        ref_rgba, ref_hdr = gb.render_rgba(cam), gb.render_linear(cam)
        for budget, passes, bpc in _SETTINGS:
            q.set_wavefront(budget, passes, bpc)
            kn = gq.kernel_name()
            rgba, hdr = gq.render_rgba(cam), gq.render_linear(cam)
            same_hdr = ref_hdr.view(np.uint32) == hdr.view(np.uint32)
            print(f"\n[queued table {name}] {kn} budget={budget} passes={passes} bpc={bpc} "
                  f"hdr_equal={bool(same_hdr.all())} rgba_equal={bool((ref_rgba == rgba).all())}", flush=True)
            assert kn.startswith("mm3e_wf_shade"), kn
            assert same_hdr.all(), (name, budget, passes, bpc, int((~same_hdr).sum()))
            assert (ref_rgba == rgba).all(), (name, budget, passes, bpc)
        q.set_wavefront(0)
        assert (gq.render_rgba(cam) == ref_rgba).all(), name
        gb.close()
        gq.close()
    base.close()
    q.close()


@pytest.mark.gpu
def test_queued_frames_equal_one_kernel_frames_generated_kernels():
    base = hip.Renderer()
    base.set_codegen(hip.CODEGEN_OFF)
    q = hip.Renderer()
    q.set_codegen(hip.CODEGEN_SYNC)
    q.set_wavefront(16, 2, 0)
    for name, sc, cam in _scenes()[:4]:
        ir = sc.to_ir()
        gb = base.compile(ir, sc.width, sc.height)
        gq = q.compile(ir, sc.width, sc.height)
        state, msg = gq.codegen_status(wait=True)
        ref = gb.render_rgba(cam)
        ref_hdr = gb.render_linear(cam)
        if state == hip.CODEGEN_READY:
            assert gq.kernel_name() == "mm3e_gen_wf_shade"
        for budget, passes, bpc in [(16, 2, 0), (4, 5, 4)]:
            q.set_wavefront(budget, passes, bpc)
            rgba, hdr = gq.render_rgba(cam), gq.render_linear(cam)
            same_hdr = ref_hdr.view(np.uint32) == hdr.view(np.uint32)
            print(f"\n[queued gen {name}] {gq.kernel_name()} budget={budget} passes={passes} bpc={bpc} "
                  f"hdr_equal={bool(same_hdr.all())} rgba_equal={bool((ref == rgba).all())}", flush=True)
            assert same_hdr.all(), (name, budget, passes, bpc)
            assert (ref == rgba).all(), (name, budget, passes, bpc)
        gb.close()
        gq.close()
    base.close()
    q.close()


@pytest.mark.gpu
def test_queued_frames_with_dynamic_spheres():
    sc, cam = configs.c1_scene()
    dyn = [(0.5, 0.6, -1.0, 0.4, 0.9, 0.2, 0.1, 0.0), (-1.5, 0.3, 0.5, 0.3, 0.1, 0.3, 0.9, 1.0)]
    base = hip.Renderer()
    base.set_codegen(hip.CODEGEN_OFF)
    q = hip.Renderer()
    q.set_codegen(hip.CODEGEN_OFF)
    q.set_wavefront(6, 3, 0)
    gb = base.compile(sc.to_ir(), sc.width, sc.height)
    gq = q.compile(sc.to_ir(), sc.width, sc.height)
    a = gb.render_rgba(cam, dyn=dyn)
    b = gq.render_rgba(cam, dyn=dyn)
    print(f"\n[queued dyn] equal={bool((a == b).all())}", flush=True)
    assert (a == b).all()
    gb.close()
    gq.close()
    base.close()
    q.close()


@pytest.mark.gpu
def test_scene_autotune_keeps_the_image():
    """The autotuner times one-kernel and queued frames and keeps the fastest as the scene's own
    setting; the image is the same whichever it picks, and a scene setting overrides the
    renderer's."""
    sc, cam = configs.c2_scene(width=480, height=270)
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_OFF)
    g = r.compile(sc.to_ir(), sc.width, sc.height)
    ref = g.render_rgba(cam)
    rep = g.autotune(cam, reps=2)
    print("\n[autotune]\n" + rep, flush=True)
    lines = rep.strip().splitlines()
    assert lines[0].startswith("one kernel:") and lines[-1].startswith("chosen:")
    assert sum(1 for ln in lines if ln.startswith("queued budget")) >= 5
    assert (g.render_rgba(cam) == ref).all()
    g.set_wavefront(8, 3, 0)
    r.set_wavefront(0)
    assert g.kernel_name().startswith("mm3e_wf_shade")
    assert (g.render_rgba(cam) == ref).all()
    g.set_wavefront(None)
    assert not g.kernel_name().startswith("mm3e_wf_shade")
    g.close()
    r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,aa", [(1, 1, 2), (17, 9, 3), (33, 7, 1), (250, 3, 4)])
def test_queued_frames_odd_sizes(w, h, aa):
    """Images whose sizes leave partial tiles (and a single pixel), at several AA factors: the
    queued frame's slots outside the image stay unused and every pixel equals the one-kernel
    frame."""
    sc, cam = configs.c1_scene(width=w, height=h)
    sc.aa = aa
    base = hip.Renderer()
    base.set_codegen(hip.CODEGEN_OFF)
    q = hip.Renderer()
    q.set_codegen(hip.CODEGEN_OFF)
    q.set_wavefront(3, 4, 1)
    gb = base.compile(sc.to_ir(), w, h)
    gq = q.compile(sc.to_ir(), w, h)
    a, b = gb.render_linear(cam), gq.render_linear(cam)
    print(f"\n[queued odd {w}x{h} aa{aa}] equal={bool((a.view(np.uint32) == b.view(np.uint32)).all())}", flush=True)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()
    gb.close()
    gq.close()
    base.close()
    q.close()
