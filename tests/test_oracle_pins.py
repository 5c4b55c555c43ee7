"""Closed-form pins for the oracle functions the reference's own tests leave open: the bloom
pass, IBL irradiance, soft shadows, light falloff and shadow sharpness, fog and the reflection
compose.  Every expected value here is derived independently in float64 from the mathematics
(hand-computed tangent frames, analytic ray/plane and ray/sphere distances, explicit Gaussian
sums), so a dropped term, a wrong sign or index or a swapped operand in the oracle fails.

The scenes with a closed-form pixel are also rendered on the device (marked gpu): the HIP path
is then held to the same mathematics, not only to the oracle.

Reference formulas: post.rs:36-115 (bloom), mm3e-orchestrator/src/lib.rs:1379-1398 (IBL),
mm3e-kit/src/march.rs:227-259 (soft shadow), lib.rs:528-556 (lights), shade.rs:113-116 (fog),
lib.rs:1336-1359 (reflection compose, then emissive, then fog).
"""
import math

import numpy as np
import pytest

import pyoracle as po
from mm3e import configs
from mm3e.scene import Light, Material, Object, Prim, Scene, Transform

HORIZON = np.array([0.78, 0.86, 0.96])
ZENITH = np.array([0.20, 0.38, 0.72])
LUMA = np.array([0.2126, 0.7152, 0.0722])
S = math.sqrt(0.5)


def sky64(d, sun, diffuse):
    """shade::sky / sky_diffuse (shade.rs:76-99) in float64."""
    d = np.asarray(d, np.float64)
    t = min(max(0.5 * (d[1] + 1.0), 0.0), 1.0)
    base = HORIZON * (1 - t) + ZENITH * t
    s = max(float(d @ np.asarray(sun, np.float64)), 0.0)
    glow = s ** 8 * 0.3 + (0.0 if diffuse else s ** 220 * 6.0)
    return base + glow


def gauss64(r):
    r = max(r, 1)
    sigma = r / 2.0
    w = np.array([math.exp(-(i * i) / (2 * sigma * sigma)) for i in range(-r, r + 1)])
    return w / w.sum()


# ---- post.rs: gaussian_kernel / bloom_pass / resolve's bloom add ----------------------------

@pytest.mark.parametrize("r", [0, 1, 3, 12])
def test_gaussian_kernel_is_normalised_gaussian(r):
    k = po.gaussian_kernel(r)
    want = gauss64(r)
    assert len(k) == 2 * max(r, 1) + 1
    assert abs(float(k.astype(np.float64).sum()) - 1.0) < 1e-6
    assert np.abs(k - want).max() < 2e-7
    assert (k == k[::-1]).all()  # symmetric


def test_bloom_of_a_uniform_field_is_the_bright_pass():
    """Clamp-to-edge blur of a constant is the constant, so bloom adds intensity * bright-pass."""
    c = np.array([2.0, 1.5, 0.5])
    hdr = np.tile(c, (12, 17, 1)).astype(np.float32)
    l = float(LUMA @ c)
    for thr, inten, r in ((1.0, 0.06, 12), (0.5, 0.8, 3), (1.2, 0.3, 1)):
        out = po.bloom(hdr, thr, inten, r)
        want = c + inten * c * (l - thr) / l
        assert np.abs(out - want).max() < 2e-6 * want.max()
    # luminance below the threshold: nothing added, bit for bit
    out = po.bloom(hdr, l + 0.01, 0.9, 12)
    assert (out == hdr).all()


def test_bloom_of_one_bright_pixel_is_the_kernel_outer_product():
    w = h = 41
    r = 12
    y0, x0 = 20, 20
    c = np.array([3.0, 2.0, 1.0])
    hdr = np.zeros((h, w, 3), np.float32)
    hdr[y0, x0] = c
    l = float(LUMA @ c)
    thr, inten = 1.0, 0.5
    b = c * (l - thr) / l
    k = gauss64(r)
    want = hdr.astype(np.float64).copy()
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            want[y0 + dy, x0 + dx] += inten * k[dy + r] * k[dx + r] * b
    out = po.bloom(hdr, thr, inten, r)
    assert np.abs(out - want).max() < 1e-6
    # nothing outside the (2r+1)^2 footprint
    far = np.ones((h, w), bool)
    far[y0 - r:y0 + r + 1, x0 - r:x0 + r + 1] = False
    assert (out[far] == 0).all()


def test_bloom_clamps_to_the_edge():
    """A bright pixel on the left border: pixel x of its row reads it for every tap with
    x + off <= 0, so it receives b * k[r] (vertical centre tap) * sum_{j <= r - x} k[j]."""
    w, h, r = 30, 31, 6
    y0 = 15
    c = np.array([4.0, 4.0, 4.0])
    hdr = np.zeros((h, w, 3), np.float32)
    hdr[y0, 0] = c
    b = c * (4.0 - 1.0) / 4.0
    k = gauss64(r)
    out = po.bloom(hdr, 1.0, 1.0, r).astype(np.float64)
    for x in range(0, r + 2):
        want = b * k[r] * k[: max(r - x + 1, 0)].sum() + (c if x == 0 else 0)
        assert np.abs(out[y0, x] - want).max() < 2e-6, x


# ---- lib.rs:1379-1398 ibl_irradiance -------------------------------------------------------

def _ibl_taps_plus_y():
    # n = +y: |n.y| > 0.9 picks up = +x; t1 = n x up = (0,0,-1); t2 = n x t1 = (-1,0,0)
    return [((0, 1, 0), 1.0), ((0, S, -S), S), ((0, S, S), S), ((-S, S, 0), S), ((S, S, 0), S)]


def _ibl_taps_plus_x():
    # n = +x: up = +y; t1 = n x up = (0,0,1); t2 = n x t1 = (0,-1,0)
    return [((1, 0, 0), 1.0), ((S, 0, S), S), ((S, 0, -S), S), ((S, -S, 0), S), ((S, S, 0), S)]


@pytest.mark.parametrize("n,taps", [((0, 1, 0), _ibl_taps_plus_y()), ((1, 0, 0), _ibl_taps_plus_x())])
def test_ibl_irradiance_is_five_cosine_weighted_sky_taps(n, taps):
    sc = Scene(8, 8)
    sc.sky_ambient = (0.5, 0.4, 0.3)
    s = po.OracleScene(sc.to_ir())
    acc = np.zeros(3)
    wsum = 0.0
    for d, w in taps:
        acc += w * sky64(d, sc.sun_dir, diffuse=True)
        wsum += w
    want = acc / wsum * np.asarray(sc.sky_ambient, np.float64)
    got = s.ibl_irradiance(n)
    assert np.abs(got - want).max() < 2e-6


def test_ibl_irradiance_of_a_constant_sky_is_the_constant():
    """With the sun far below every tap and all taps at the same height the taps agree,
    so the weighted mean is one sky_diffuse value (a check of the normalisation)."""
    sc = Scene(8, 8)
    sc.sun_dir = (0.0, -1.0, 0.0)
    sc.sky_ambient = (1.0, 1.0, 1.0)
    s = po.OracleScene(sc.to_ir())
    # n = -y: taps at d.y = -1 and -sqrt(1/2) -> t = 0 and 0.146; not constant, so use n = +y
    # and compare only the structure: the zenith tap and the four 45-degree taps
    got = s.ibl_irradiance((0, 1, 0))
    t45 = 0.5 * (1 + S)
    want = (ZENITH * 1.0 + 4 * S * (HORIZON * (1 - t45) + ZENITH * t45)) / (1 + 4 * S)
    assert np.abs(got - want).max() < 2e-6


# ---- march.rs:227-259 soft_shadow ---------------------------------------------------------

def _sphere_scene(center, radius):
    sc = Scene(8, 8)
    sc.add(Object(Prim.sphere(radius), Transform(pos=center), 0))
    return po.OracleScene(sc.to_ir())


def test_soft_shadow_open_sky_is_one():
    sc = Scene(8, 8)
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), 0))
    s = po.OracleScene(sc.to_ir())
    # moving straight away from the plane, h = 1 + t, so k*h/t > 1 for every k >= 1
    for o in ((0.3, 1.0, 0.2), (-5.0, 0.5, 7.0)):
        for k in (1.0, 8.0, 24.0):
            assert s.soft_shadow(o, (0, 1, 0), 50.0, k) == 1.0


def test_soft_shadow_on_axis_blocker_is_zero():
    s = _sphere_scene((0, 5, 0), 1.0)
    assert s.soft_shadow((0, 1, 0), (0, 1, 0), 50.0, 8.0) == 0.0
    # the blocker behind max_t does not shadow: the march stops before reaching it
    assert s.soft_shadow((0, 1, 0), (0, 1, 0), 2.0, 8.0) > 0.0


@pytest.mark.parametrize("offset,k", [(1.3, 4.0), (1.2, 8.0), (2.5, 2.0)])
def test_soft_shadow_penumbra_is_bracketed_by_the_continuous_minimum(offset, k):
    """The shadow is min over the loop's samples t_i of k*h(t_i)/t_i (clamped to [0,1]).  For a
    sphere of radius 1 at lateral offset b, h(t) = sqrt(b^2 + (t - 4)^2) - 1 along the ray from
    (0,1,0) upward.  The samples are a subset of [0.02, inf), so the result is at least the
    continuous minimum; a sample falls within one step (<= max(h, 0.06)) of the minimiser, so
    it is at most the largest value within that window."""
    s = _sphere_scene((offset, 5, 0), 1.0)
    got = s.soft_shadow((0, 1, 0), (0, 1, 0), 60.0, k)
    t = np.linspace(0.02, 60.0, 600001)
    hfun = np.sqrt(offset ** 2 + (t - 4.0) ** 2) - 1.0
    v = k * hfun / t
    i = int(np.argmin(v))
    lo = min(float(v[i]), 1.0)
    step = max(float(hfun[i]), 0.06) + 1e-3
    win = np.abs(t - t[i]) <= step
    hi = min(float(v[win].max()), 1.0)
    assert lo - 1e-5 <= got <= hi + 1e-5, (lo, got, hi)
    assert 0.0 < got < 1.0


# ---- lib.rs:528-556 Light::toward / attenuation / shadow_k --------------------------------

def test_directional_light_terms():
    dvec = np.array([0.6, 0.7, 0.4]) / np.linalg.norm([0.6, 0.7, 0.4])
    for radius, k in ((0.0, 24.0), (0.125, 8.0), (1.0, 2.0), (0.01, 64.0), (0.25, 4.0)):
        d, dist, att, kk = po.light_eval(True, dvec, radius, (3.0, -1.0, 2.0))
        assert np.abs(d - dvec).max() < 1e-7 and math.isinf(dist)
        assert att == 1.0 and kk == np.float32(k), (radius, kk)


def test_point_light_falloff_and_shadow_k():
    pos = (0.0, 3.0, 0.0)
    d, dist, att, k = po.light_eval(False, pos, 0.5, (0.0, 1.0, 0.0))
    assert np.abs(d - (0, 1, 0)).max() == 0 and dist == 2.0
    assert att == 0.25 and k == 4.0              # 1/d^2; clamp(d / r, 2, 64)
    assert po.light_eval(False, pos, 0.0, (0, 1, 0))[3] == 24.0   # hard light
    assert po.light_eval(False, pos, 2.0, (0, 1, 0))[3] == 2.0    # clamped low
    assert po.light_eval(False, pos, 0.01, (0, 1, 0))[3] == 64.0  # clamped high
    # closer than 0.05: the falloff saturates at 1 / 0.05^2
    _, dist, att, _ = po.light_eval(False, pos, 0.0, (0.0, 2.98, 0.0))
    assert abs(dist - 0.02) < 1e-6 and abs(att - 400.0) < 1e-3
    # at the light itself: direction 0 (normalised by 1/max(len, 1e-6)), falloff 400
    d, dist, att, _ = po.light_eval(False, pos, 0.0, pos)
    assert (d == 0).all() and dist == 0.0 and abs(att - 400.0) < 1e-3
    # off axis, a 3-4-5 triangle
    d, dist, att, _ = po.light_eval(False, (3.0, 4.0, 0.0), 0.0, (0.0, 0.0, 0.0))
    assert abs(dist - 5.0) < 1e-6 and np.abs(d - (0.6, 0.8, 0.0)).max() < 1e-7
    assert abs(att - 1 / 25) < 1e-8


# ---- shade.rs:113-116 apply_fog -------------------------------------------------------------

def test_apply_fog_closed_form():
    c = np.array([0.2, 0.5, 0.9])
    fog = np.array([0.62, 0.72, 0.86])
    for t, rho in ((10.0, 0.05), (3.0, 0.012), (200.0, 0.5)):
        f = 1.0 - math.exp(-t * rho)
        want = c * (1 - f) + fog * f
        assert np.abs(po.apply_fog(c, fog, t, rho) - want).max() < 1e-6
    assert (po.apply_fog(c, fog, 10.0, 0.0) == c.astype(np.float32)).all()


# ---- closed-form scenes rendered end to end (oracle, and the device) ----------------------

def _dark_scene(w, h, **kw):
    sc = Scene(w, h)
    sc.aa = 1
    sc.ao = False
    sc.post.bloom = False
    sc.fog_density = 0.0
    sc.ambient = (0.0, 0.0, 0.0)
    sc.sky_ambient = (0.0, 0.0, 0.0)
    for k, v in kw.items():
        setattr(sc, k, v)
    return sc


def _ray_dir(cam, x, y, w, h):
    c = np.asarray(cam, np.float64)
    fwd, right, up, fs = c[3:6], c[6:9], c[9:12], c[12]
    ndx = (2 * (x + 0.5) / w - 1) * (w / h) * fs
    ndy = (1 - 2 * (y + 0.5) / h) * fs
    d = fwd + right * ndx + up * ndy
    return d / np.linalg.norm(d)


def _fog_plane_case(rho):
    """An emissive black plane, no lights, no ambient: radiance = E, then fog over the
    analytic plane distance t = -eye.y / d.y."""
    w, h = 9, 7
    E = np.array([0.3, 0.6, 0.2])
    sc = _dark_scene(w, h, fog_density=rho)
    m = sc.material(Material.solid((0.0, 0.0, 0.0)).emissive(tuple(E)))
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), m))
    cam = configs.look_at((0.0, 2.0, -6.0), (0.0, 0.0, 0.0), (0, 1, 0), math.radians(50))
    fog = np.asarray(sc.fog, np.float64)
    want = np.zeros((h, w, 3))
    tol = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            d = _ray_dir(cam, x, y, w, h)
            if d[1] >= 0 or -2.0 / d[1] > 100.0:
                want[y, x] = np.nan
                continue
            t = -2.0 / d[1]
            f = 1 - math.exp(-t * rho)
            want[y, x] = E * (1 - f) + fog * f
            # the march stops within eps * (1 + 0.5 t) of the surface in field distance, which is
            # that over |d.y| along the ray; f moves by rho * dt
            tol[y, x] = rho * 6e-4 * (1 + 0.5 * t) / -d[1] * 1.5 + 2e-6
    return sc, cam, want, tol


def _mirror_plane_case(rho):
    """A black mirror plane (reflectivity 0.5) with emissive E under an empty sky, 1 bounce:
    radiance = 0*(1-fr) + sky(reflect(d))*fr, then + E, then fog over the hit distance."""
    w, h = 9, 7
    E = np.array([0.1, 0.2, 0.05])
    refl = 0.5
    sc = _dark_scene(w, h, fog_density=rho, bounces=1)
    m = sc.material(Material.solid((0.0, 0.0, 0.0)).reflective(refl).emissive(tuple(E)))
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), m))
    cam = configs.look_at((0.0, 2.0, -6.0), (0.0, 0.0, 0.0), (0, 1, 0), math.radians(50))
    fog = np.asarray(sc.fog, np.float64)
    want = np.zeros((h, w, 3))
    tol = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            d = _ray_dir(cam, x, y, w, h)
            if d[1] >= 0 or -2.0 / d[1] > 100.0:  # the sky: not this closed form
                want[y, x] = np.nan
                continue
            t = -2.0 / d[1]
            c = -d[1]
            fr = refl + (1 - refl) * (1 - c) ** 5
            rdir = np.array([d[0], -d[1], d[2]])
            rad = sky64(rdir, sc.sun_dir, diffuse=False) * fr + E
            f = 1 - math.exp(-t * rho)
            want[y, x] = rad * (1 - f) + fog * f
            tol[y, x] = 2e-5 + rho * 6e-4 * (1 + 0.5 * t) / -d[1] * 1.5
    return sc, cam, want, tol


def _point_light_case(height):
    """A plane lit by one point light straight above the centre pixel's hit point, camera
    looking straight down: n = l = v, so the pixel is brdf(n=l=v) * color / max(d, 0.05)^2."""
    w = h = 9
    alb, rough, spec = 0.6, 0.4, 0.5
    sc = _dark_scene(w, h)
    m = sc.material(Material.solid((alb, alb, alb)).roughness(rough).specular(spec))
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), m))
    col = np.array([5.0, 4.0, 3.0])
    sc.light(Light.point((0.0, height, 0.0), tuple(col)))
    cam = configs.look_at((0.0, 4.0, 0.0), (0.0, 0.0, 0.0), (0, 0, 1), math.radians(40))
    a = max(rough * rough, 1e-3)
    a2 = a * a
    f0 = 0.08 * spec
    brdf = (1 - f0) * alb / math.pi + f0 / (math.pi * a2) * 0.25
    att = 1.0 / max(height, 0.05) ** 2
    want = brdf * col * att
    # the hit point sits up to eps above the plane, moving the light distance by as much
    rel = 2 * 6e-4 * 3 / max(height, 0.05) if height > 0.05 else 1e-6
    return sc, cam, (h // 2, w // 2), want, rel


CASES = {
    "fog": lambda: _fog_plane_case(0.05),
    "mirror": lambda: _mirror_plane_case(0.0),
    "mirror_fog": lambda: _mirror_plane_case(0.03),
}


def _check_case(hdr, want, tol):
    ok = ~np.isnan(want[..., 0])
    assert ok.sum() > 10
    err = np.abs(hdr.astype(np.float64) - want)[ok].max(axis=-1)
    assert (err <= tol[ok]).all(), (err.max(), tol[ok].min())


@pytest.mark.parametrize("name", sorted(CASES))
def test_closed_form_scene_oracle(name):
    sc, cam, want, tol = CASES[name]()
    hdr = po.OracleScene(sc.to_ir()).render_linear(cam)
    _check_case(hdr, want, tol)


@pytest.mark.parametrize("height", [3.0, 0.02])
def test_point_light_scene_oracle(height):
    sc, cam, (y, x), want, rel = _point_light_case(height)
    hdr = po.OracleScene(sc.to_ir()).render_linear(cam)
    assert np.abs(hdr[y, x] - want).max() <= rel * want.max() + 1e-6, (hdr[y, x], want)


def test_reflection_compose_order_is_pinned():
    """Emissive added after the reflection mix (not scaled by 1 - fr) and fog last: the two
    misorderings give values far outside the closed form's tolerance."""
    sc, cam, want, tol = _mirror_plane_case(0.03)
    hdr = po.OracleScene(sc.to_ir()).render_linear(cam).astype(np.float64)
    y, x = 5, 4
    E = np.array([0.1, 0.2, 0.05])
    d = _ray_dir(cam, x, y, 9, 7)
    fr = 0.5 + 0.5 * (1 + d[1]) ** 5
    assert np.abs(hdr[y, x] - want[y, x]).max() <= tol[y, x]
    assert np.abs(hdr[y, x] - (want[y, x] - E * fr * math.exp(-0.03 * -2 / d[1]))).max() > 10 * tol[y, x]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_closed_form_scene_device(renderer, name):
    sc, cam, want, tol = CASES[name]()
    g = renderer.compile(sc.to_ir(), sc.width, sc.height)
    try:
        _check_case(g.render_linear(cam), want, tol)
    finally:
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("height", [3.0, 0.02])
def test_point_light_scene_device(renderer, height):
    sc, cam, (y, x), want, rel = _point_light_case(height)
    g = renderer.compile(sc.to_ir(), sc.width, sc.height)
    try:
        hdr = g.render_linear(cam)
    finally:
        g.close()
    assert np.abs(hdr[y, x] - want).max() <= rel * want.max() + 1e-6, (hdr[y, x], want)


# ---- lib.rs:1223-1243 shade_pixel: the aa x aa grid at ((s + 0.5) / n) --------------------

def test_aa_grid_equals_a_box_filtered_supersampled_frame():
    """Sample (sx, sy) of pixel (x, y) at aa = n is the ray through (x + (sx + 0.5)/n,
    y + (sy + 0.5)/n): the centre of pixel (n x + sx, n y + sy) of the frame n times larger at
    aa = 1.  So the aa = 2 frame is the 2x2 box filter of the aa = 1 frame at twice the size,
    up to the rounding of the ray's screen coordinate (samples off the grid, e.g. at s/n, or a
    mean over the wrong count, miss by far more)."""
    sc, cam = configs.c1_scene(width=40, height=24)
    sc.aa = 2
    sc.post.bloom = False
    small = po.OracleScene(sc.to_ir()).render_linear(cam).astype(np.float64)
    sc.aa = 1
    sc.width, sc.height = 80, 48
    big = po.OracleScene(sc.to_ir()).render_linear(cam).astype(np.float64)
    box = 0.25 * (big[0::2, 0::2] + big[0::2, 1::2] + big[1::2, 0::2] + big[1::2, 1::2])
    err = np.abs(small - box).max(axis=-1)
    assert np.percentile(err, 99) < 1e-5 and np.mean(err < 1e-5) > 0.99, (np.percentile(err, 99), err.max())
