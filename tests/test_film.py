"""Film render through a thin lens (mm3e-orchestrator/src/film.rs, mm3e-kit/src/camera.rs:14-118).

CPU tests pin the oracle with the reference's own film tests (mm3e-orchestrator/tests/film.rs)
and closed forms; GPU tests check mm3e_scene_render_film against the oracle: bit-exact for a
pinhole (and equal to render_linear), within the image tolerance through an open aperture
(the lens map's cos/sin are the device's, a few ulp from glibc's).
"""
import numpy as np
import pytest

import pyoracle as po
from mm3e import configs, hip
from mm3e.scene import Light, Material, Object, Prim, Scene, Transform


def camera():
    return configs.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.8)


def emissive_scene(w, h, emission):
    """film.rs tests: black emissive sphere r = 0.55 at the origin, no lighting terms."""
    sc = Scene(w, h)
    sc.aa = 4
    sc.ao = False
    sc.shadows = False
    sc.bounces = 0
    sc.fog_density = 0.0
    sc.ambient = (0.0, 0.0, 0.0)
    sc.sky_ambient = (0.0, 0.0, 0.0)
    sc.marcher.eps = 0.00001
    sc.marcher.max_steps = 512
    m = sc.material(Material(albedo=(0.0, 0.0, 0.0), emissive_v=tuple(emission)))
    sc.add(Object(Prim.sphere(0.55), Transform(), m))
    return sc


def lit_scene():
    sc = Scene(23, 19)
    sc.aa = 3
    sc.bounces = 2
    m = sc.material(Material.solid((0.9, 0.1, 0.06)).reflective(0.7))
    sc.add(Object(Prim.sphere(0.65), Transform(), m))
    sc.light(Light.directional_light((-1.0, 2.0, 2.0), (3.0, 3.0, 3.0)))
    return sc


def film_of(sc, cam, radius=0.0, focus=1.0, threads=0):
    return po.OracleScene(sc.to_ir()).render_film(cam, radius, focus, threads=threads)


def differing(a, b):
    """Keys whose arrays differ bit for bit, with the count of differing values and the largest
    difference."""
    out = {}
    for k in a:
        x, y = a[k], b[k]
        ne = x.view(np.uint32) != y.view(np.uint32)
        if ne.any():
            fin = np.isfinite(x) & np.isfinite(y) if x.dtype == np.float32 else np.ones_like(ne)
            d = np.abs(x.astype(np.float64) - y.astype(np.float64))[ne & fin]
            out[k] = (int(ne.sum()), float(d.max()) if d.size else None)
    return out


def same(a, b):
    return not differing(a, b)


# ---- oracle against the reference's film tests -------------------------------------------------

def test_pinhole_beauty_is_bitwise_canonical_with_lighting_fog_reflections_and_aa():
    # film.rs:29-42
    sc = lit_scene()
    o = po.OracleScene(sc.to_ir())
    expected = o.render_linear(camera(), threads=2)
    one = o.render_film(camera(), threads=1)
    many = o.render_film(camera(), threads=5)
    assert np.array_equal(one["beauty"].reshape(-1).view(np.uint32), expected.reshape(-1).view(np.uint32))
    assert same(one, many)


def test_premultiplied_foreground_composites_over_independent_physical_background():
    # film.rs:44-70
    emission = np.array([4.0, 0.6, 0.2], np.float32)
    sc = emissive_scene(37, 29, emission)
    film = film_of(sc, camera(), 0.18, 6.0, threads=3)
    a = film["alpha"]
    assert ((a > 0) & (a < 1)).any() and (a == 0).any() and (a == 1).any()
    background = np.array([0.02, 0.08, 0.3], np.float32)
    ref = emissive_scene(37, 29, emission)
    m = ref.material(Material(albedo=(0.0, 0.0, 0.0), emissive_v=tuple(background)))
    ref.add(Object(Prim.plane((0.0, 0.0, 1.0), 3.0), Transform(), m))
    physical = film_of(ref, camera(), 0.18, 6.0, threads=1)
    fg = film["foreground"]
    assert np.linalg.norm(fg - emission * a[..., None], axis=-1).max() < 1e-6
    comp = fg + background * (1.0 - a[..., None])
    assert np.linalg.norm(comp - physical["beauty"], axis=-1).max() < 3e-6
    assert (fg[a == 0] == 0).all()


def test_black_geometry_is_opaque_and_primary_sky_is_uncovered():
    # film.rs:72-86
    film = film_of(emissive_scene(13, 13, (0.0, 0.0, 0.0)), camera(), threads=2)
    assert film["alpha"][6, 6] == 1.0
    assert (film["foreground"][6, 6] == 0).all() and (film["beauty"][6, 6] == 0).all()
    assert film["alpha"][0, 0] == 0.0
    assert np.linalg.norm(film["beauty"][0, 0]) > 0.1
    assert film["depth"][0, 0] == np.inf
    assert (film["normals"][0, 0] == 0).all()
    assert film["material_ids"][0, 0] == 0xFFFFFFFF


def test_geometry_passes_are_raw_center_pinhole_depth_world_normal_and_material():
    # film.rs:88-108
    sc = emissive_scene(3, 3, (1.0, 1.0, 1.0))
    sc.objects.clear()
    sc.add(Object(Prim.plane((0.0, 0.0, 1.0), 0.0), Transform(), 1))
    film = film_of(sc, camera(), 0.9, 1.0, threads=2)
    assert np.abs(film["depth"] - 3.0).max() < 1e-4
    assert np.linalg.norm(film["normals"] - np.array([0, 0, 1.0]), axis=-1).max() < 1e-5
    assert (film["material_ids"] == 1).all()
    sphere = film_of(emissive_scene(3, 3, (1.0, 1.0, 1.0)), camera(), 0.9, 1.0, threads=1)
    assert abs(sphere["depth"][1, 1] - 2.45) < 1e-4
    assert np.linalg.norm(sphere["normals"][1, 1] - np.array([0, 0, 1.0])) < 1e-5


def test_lens_rays_converge_on_camera_forward_focus_plane_even_off_axis():
    # film.rs:110-129
    cam = configs.look_at((2.0, 3.0, 4.0), (-1.0, 0.3, 1.0), (0.0, 1.0, 0.0), 1.1)
    eye, fwd = cam[0:3].astype(np.float64), cam[3:6].astype(np.float64)
    radius, focus_d = 0.2, 5.0
    for image in ([100.5, 20.5], [31.0, 97.5], [80.0, 60.0]):
        po_o, po_d = po.ray_with_lens(cam, image, (160, 120), 0.0, 1.0, (0.5, 0.5))
        expected = po_o + po_d * (focus_d / po_d.astype(np.float64).dot(fwd))
        for ap in ([0.0, 0.0], [0.5, 0.5], [0.2, 0.7], [1.0, 0.8], [0.5, 1.0]):
            o, d = po.ray_with_lens(cam, image, (160, 120), radius, focus_d, ap)
            off = o.astype(np.float64) - eye
            assert abs(off.dot(fwd)) < 1e-6
            assert np.linalg.norm(off) <= radius + 1e-6
            depth = off.dot(fwd)
            at = o + d * ((focus_d - depth) / d.astype(np.float64).dot(fwd))
            assert np.linalg.norm(at - expected) < 3e-6
            zo, zd = po.ray_with_lens(cam, image, (160, 120), 0.0, 1.0, ap)
            assert np.array_equal(zo, po_o) and np.array_equal(zd, po_d)


def test_aperture_changes_defocused_coverage_deterministically_but_not_geometry_aovs():
    # film.rs:131-147
    sc = emissive_scene(31, 25, (1.0, 1.0, 1.0))
    sc.aa = 6
    pin = film_of(sc, camera(), threads=1)
    ap = film_of(sc, camera(), 0.3, 7.0, threads=1)
    assert same(ap, film_of(sc, camera(), 0.3, 7.0, threads=8))
    assert not np.array_equal(pin["alpha"], ap["alpha"])
    for k in ("depth", "normals", "material_ids"):
        assert np.array_equal(pin[k], ap[k])
    partial = lambda a: int(((a > 0) & (a < 1)).sum())
    assert partial(ap["alpha"]) > partial(pin["alpha"])


def test_aperture_samples_cover_every_lens_cell_once():
    # film.rs:200-216: for each pixel the n x n samples land in distinct cells of the lens grid
    for n in (2, 3, 5):
        for pixel in ((0, 0), (7, 3), (123, 77)):
            cells = set()
            for sy in range(n):
                for sx in range(n):
                    a = po.aperture_sample(pixel, (sx, sy), n)
                    assert 0.0 <= a[0] < 1.0 and 0.0 <= a[1] < 1.0
                    cells.add((int(a[0] * n), int(a[1] * n)))
            assert len(cells) == n * n
    assert po.aperture_sample((4, 9), (0, 0), 1).tolist() == [0.5, 0.5]


# ---- device film ---------------------------------------------------------------------------------

def lens_error(renderer, radius, focus):
    sc = emissive_scene(8, 8, (1.0, 1.0, 1.0))
    g = renderer.compile(sc.to_ir(), 8, 8)
    try:
        g.render_film(camera(), radius, focus)
    except hip.Mm3eError as e:
        return str(e)
    finally:
        g.close()
    return None


@pytest.mark.gpu
def test_gpu_film_lens_validation(renderer):
    # camera.rs:25-34 and film.rs:149-157
    assert lens_error(renderer, 0.0, 1.0) is None
    for r in (-0.1, float("nan"), float("inf")):
        assert lens_error(renderer, r, 1.0) == "aperture_radius must be finite and nonnegative"
    for f in (0.0, -1.0, float("nan"), float("inf")):
        assert lens_error(renderer, 0.1, f) == "focus_distance must be finite and positive"


@pytest.mark.gpu
def test_gpu_pinhole_film_is_bit_exact_and_canonical(renderer):
    sc = lit_scene()
    g = renderer.compile(sc.to_ir(), sc.width, sc.height)
    film = g.render_film(camera())
    want = film_of(sc, camera())
    lin = g.render_linear(camera())
    print("film vs oracle:", differing(film, want), "beauty vs render_linear:",
          differing({"b": film["beauty"]}, {"b": lin}), flush=True)
    # the device film is the device's canonical render, bit for bit
    assert np.array_equal(film["beauty"].view(np.uint32), lin.view(np.uint32))
    # against the oracle: the centre-ray AOVs exactly (a dual-normal sphere: every operation
    # correctly rounded); radiance within a few ulp (lighting goes through the device's pow/exp)
    for k in ("alpha", "depth", "normals", "material_ids"):
        assert np.array_equal(film[k], want[k]), k
    for k in ("beauty", "foreground"):
        assert np.allclose(film[k], want[k], rtol=1e-6, atol=1e-7), k
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["emissive", "c1", "c2", "csg"])
def test_gpu_film_through_aperture_matches_oracle(renderer, which):
    """Lens films against the oracle's.  Radiance is compared as displayed: both beauty (and
    both foreground) buffers go through the scene's post pass (the oracle's resolve: bloom,
    exposure, ACES, gamma, 8-bit) and meet the image bar, >= 99.9 % of pixels within 2/255;
    coverage alpha and the centre-ray material ids agree on >= 99.9 % of pixels; depth and
    normals agree to 1e-5 relative / 1e-4."""
    from parity import image_metrics
    if which == "emissive":
        sc = emissive_scene(37, 29, (4.0, 0.6, 0.2))
        cam, lens = camera(), (0.18, 6.0)
    elif which == "c1":
        sc, cam = configs.c1_scene(width=160, height=90)
        sc.aa = 3
        lens = (0.25, 8.0)
    elif which == "c2":
        sc, cam = configs.c2_scene()  # the full 960x540 GI workload
        lens = (0.15, 5.0)
    else:
        sc, cam = configs.csg_scene(96, 64, 5)
        sc.aa = 2
        lens = (0.2, 6.0)
    print(f"film {which} {sc.width}x{sc.height} aa={sc.aa} lens={lens}", flush=True)
    g = renderer.compile(sc.to_ir(), sc.width, sc.height)
    got = g.render_film(cam, *lens)
    g.close()
    o = po.OracleScene(sc.to_ir())
    want = o.render_film(cam, *lens)
    mb = image_metrics(o.resolve(got["beauty"]), o.resolve(want["beauty"]))
    mf = image_metrics(o.resolve(got["foreground"]), o.resolve(want["foreground"]))
    alpha_agree = float((got["alpha"] == want["alpha"]).mean())
    ids_agree = float((got["material_ids"] == want["material_ids"]).mean())
    print(f"alpha_agree={alpha_agree:.5f} ids_agree={ids_agree:.5f} beauty={mb} foreground={mf}", flush=True)
    assert alpha_agree >= 0.999 and ids_agree >= 0.999
    assert mb["frac_within"] >= 0.999 and mf["frac_within"] >= 0.999
    both = np.isfinite(got["depth"]) & np.isfinite(want["depth"])
    assert np.abs(got["depth"][both] - want["depth"][both]).max() <= 1e-5 * max(1.0, float(want["depth"][both].max()))
    same_id = got["material_ids"] == want["material_ids"]
    assert np.abs(got["normals"][same_id] - want["normals"][same_id]).max() <= 1e-4
