"""Pins the CPU oracle against the reference's own known-answer tests and closed forms.

Each test names the reference test or formula it re-expresses (paths relative to the
reference tree).  The oracle is the checker of the HIP path, so it has to be right first.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from mm3e import configs
from mm3e.scene import (Light, Material, Object, Prim, Scene, Transform, PRIM_BOX, PRIM_CAPSULE, PRIM_CONE,
                        PRIM_CYLINDER, PRIM_ELLIPSOID, PRIM_HEX_PRISM, PRIM_OCTAHEDRON, PRIM_PLANE, PRIM_ROUND_BOX,
                        PRIM_SPHERE, PRIM_TORUS)

F = np.float32


def close(a, b, eps):
    return abs(float(a) - float(b)) <= eps


def scene_of(*objs, **kw):
    sc = Scene(kw.get("w", 8), kw.get("h", 8))
    for _ in range(8):
        sc.material(Material())
    for o in objs:
        sc.add(o)
    for k, v in kw.items():
        if k in ("w", "h"):
            continue
        if k.startswith("m_"):
            setattr(sc.marcher, k[2:], v)
        else:
            setattr(sc, k, v)
    return po.OracleScene(sc.to_ir(), threads=2)


# ---- sdf.rs primitives: mm3e-kit/tests/kit.rs:173-200 -------------------------------------

def test_sphere_sdf_is_signed_distance():
    assert close(po.prim_distance(PRIM_SPHERE, [1.0], [2, 0, 0]), 1.0, 1e-5)
    assert close(po.prim_distance(PRIM_SPHERE, [1.0], [0, 0, 0]), -1.0, 1e-5)


def grad_mag(kind, params, p, h=1e-3):
    p = np.asarray(p, np.float64)
    g = []
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        g.append((po.prim_distance(kind, params, p + e) - po.prim_distance(kind, params, p - e)) / (2 * h))
    return float(np.linalg.norm(g))


def test_box_and_torus_are_lipschitz_outside():
    assert close(po.prim_distance(PRIM_BOX, [1, 1, 1], [2, 0, 0]), 1.0, 1e-5)
    assert close(grad_mag(PRIM_BOX, [1, 1, 1], [2.0, 0.5, 0.3]), 1.0, 2e-2)
    assert close(po.prim_distance(PRIM_TORUS, [1.0, 0.3], [1, 0, 0]), -0.3, 1e-5)
    assert close(grad_mag(PRIM_TORUS, [1.0, 0.3], [2.0, 0.2, 0.0]), 1.0, 2e-2)
    assert close(grad_mag(PRIM_SPHERE, [1.0], [2.0, 0.3, -0.4]), 1.0, 1e-2)


def test_new_primitives_have_correct_sign():  # kit.rs:191-200
    assert po.prim_distance(PRIM_CONE, [0.5, 0.2, 1.0], [0, 0, 0]) < 0
    assert po.prim_distance(PRIM_CONE, [0.5, 0.2, 1.0], [3, 0, 0]) > 0
    assert po.prim_distance(PRIM_ELLIPSOID, [1.0, 2.0, 0.5], [0, 0, 0]) < 0
    assert po.prim_distance(PRIM_OCTAHEDRON, [1.0], [0, 0, 0]) < 0
    assert po.prim_distance(PRIM_OCTAHEDRON, [1.0], [3, 3, 3]) > 0
    assert po.prim_distance(PRIM_HEX_PRISM, [1.0, 1.0], [0, 0, 0]) < 0
    assert po.prim_distance(PRIM_HEX_PRISM, [1.0, 1.0], [3, 0, 0]) > 0


def test_closed_form_distances():
    """Closed forms the formulas reduce to on axis (catch swapped operands / dropped terms)."""
    assert close(po.prim_distance(PRIM_CYLINDER, [1.0, 0.5], [0, 3, 0]), 2.0, 1e-6)      # cap
    assert close(po.prim_distance(PRIM_CYLINDER, [1.0, 0.5], [3, 0, 0]), 2.5, 1e-6)      # side
    assert close(po.prim_distance(PRIM_CAPSULE, [0, -1, 0, 0, 1, 0, 0.5], [0, 3, 0]), 1.5, 1e-6)
    assert close(po.prim_distance(PRIM_CAPSULE, [0, -1, 0, 0, 1, 0, 0.5], [2, 0.3, 0]), 1.5, 1e-6)
    assert close(po.prim_distance(PRIM_ROUND_BOX, [1, 1, 1, 0.25], [3, 0, 0]), 2.0, 1e-6)
    assert close(po.prim_distance(PRIM_PLANE, [0, 1, 0, 0.5], [7, 2, -3]), 2.5, 1e-6)
    assert close(po.prim_distance(PRIM_ELLIPSOID, [2, 1, 1], [4, 0, 0]), 2.0, 1e-5)     # on a principal axis
    s3 = 0.5773503
    assert close(po.prim_distance(PRIM_OCTAHEDRON, [1.0], [2, 0, 0]), 1.0 * s3, 1e-6)
    assert close(po.prim_distance(PRIM_CONE, [0.5, 0.2, 1.0], [0, -2, 0]), 1.5, 1e-6)   # below the base sphere
    assert close(po.prim_distance(PRIM_CONE, [0.5, 0.2, 1.0], [0, 3, 0]), 1.8, 1e-6)    # above the top sphere
    assert close(po.prim_distance(PRIM_HEX_PRISM, [1.0, 1.0], [0, 3, 0]), 2.0, 1e-6)    # cap
    assert close(po.prim_distance(PRIM_HEX_PRISM, [1.0, 1.0], [0, 0, 3]), 2.0, 1e-6)    # flat side (in-radius)


# ---- CSG combinators via a two-object world (kit.rs:204-226) ------------------------------

def test_csg_union_subtract_smooth():
    a = Object(Prim.sphere(1.0), Transform.at((0, 0, 0)), 1)
    b = Object(Prim.sphere(1.0), Transform.at((3, 0, 0)), 2)
    s = scene_of(a, b)
    d, m = s.field([[2.0, 0, 0]])
    assert close(d[0], 0.0, 1e-6) and m[0] == 2  # the nearer surface wins, with its material
    d, m = s.field([[1.5, 0, 0]])
    assert close(d[0], 0.5, 1e-6) and m[0] == 1  # tie-free: a nearer
    sub = scene_of(Object(Prim.sphere(1.0), Transform.at((0, 0, 0)), 1),
                   Object(Prim.sphere(1.0), Transform.at((1.2, 0, 0)), 2).subtract())
    d, m = sub.field([[0.5, 0, 0]])
    assert close(d[0], max(-0.5, 0.3), 1e-6) and m[0] == 1  # max(a, -b) with b = -0.3; keeps a material
    hard = scene_of(a, Object(Prim.sphere(1.0), Transform.at((2.1, 0, 0)), 2))
    soft = scene_of(a, Object(Prim.sphere(1.0), Transform.at((2.1, 0, 0)), 2).smooth(0.3))
    pts = np.array([[1.05, y, 0.0] for y in np.linspace(-0.5, 0.5, 11)], np.float32)
    dh, _ = hard.field(pts)
    ds, _ = soft.field(pts)
    assert (ds <= dh + 1e-6).all() and (ds >= dh - 0.3).all()
    # closed form at the midpoint: da = db = 0.05 -> h = 0.5 -> d = 0.05 - 0.3/4
    assert close(ds[5], 0.05 - 0.3 * 0.25, 1e-5)


def test_domain_ops():  # kit.rs:230-242
    rep = scene_of(Object(Prim.sphere(0.3), Transform(), 1).repeat((1.0, 1.0, 1.0)))
    d, _ = rep.field([[2.6, -3.1, 0.2]])
    q = np.array([2.6 - round(2.6), -3.1 - (-3.0), 0.2], np.float64)  # folds into [-0.5, 0.5]
    assert abs(q[0]) <= 0.5 + 1e-5 and abs(q[1]) <= 0.5 + 1e-5
    assert close(d[0], np.linalg.norm(q) - 0.3, 1e-5)
    rnd = scene_of(Object(Prim.sphere(1.0), Transform(), 1).round(0.2))
    assert close(rnd.field([[2.0, 0, 0]])[0][0], 0.8, 1e-6)
    oni = scene_of(Object(Prim.sphere(1.0), Transform(), 1).onion(0.1))
    assert close(oni.field([[1.0, 0, 0]])[0][0], -0.1, 1e-6)
    mir = scene_of(Object(Prim.capsule((1, 0, 0), (1, 0.001, 0), 0.5), Transform(), 1).mirror(True, False, False))
    assert close(mir.field([[-1.0, 0, 0]])[0][0], -0.5, 1e-5)


def test_twist_bend_elongate_closed_forms():
    """Points inside the prune slack, so the exact object field (not the bounding-sphere lower
    bound that `world()` substitutes far away, accel.rs:196-206) is what comes back."""
    def box(q, half):
        d = np.abs(q) - half
        return np.linalg.norm(np.maximum(d, 0)) + min(d.max(), 0.0)

    half = np.array([0.3, 2.0, 0.6])
    p = np.array([0.4, 1.0, 0.1])
    a = 0.9 * p[1]
    s_, c_ = math.sin(a), math.cos(a)
    tw = scene_of(Object(Prim.box(tuple(half)), Transform(), 1).twist(0.9))
    q = np.array([c_ * p[0] - s_ * p[2], p[1], s_ * p[0] + c_ * p[2]])
    assert close(tw.field([p])[0][0], box(q, half), 1e-5)
    a = 0.5 * p[0]
    s_, c_ = math.sin(a), math.cos(a)
    bd = scene_of(Object(Prim.box(tuple(half)), Transform(), 1).bend(0.5))
    q = np.array([c_ * p[0] - s_ * p[1], s_ * p[0] + c_ * p[1], p[2]])
    assert close(bd.field([p])[0][0], box(q, half), 1e-5)
    # elongate a sphere by (1,0,0): distance from the segment [-1,1]x{0}x{0}
    el = scene_of(Object(Prim.sphere(0.5), Transform(), 1).elongate((1.0, 0.0, 0.0)))
    assert close(el.field([[0.7, 0.6, 0.0]])[0][0], 0.1, 1e-6)
    assert close(el.field([[1.8, 0.0, 0.0]])[0][0], 0.3, 1e-6)
    # far away the world field is the conservative bound |p - c| - r (r = 0.5 + |elongate|)
    assert close(el.field([[0.7, 1.5, 0.0]])[0][0], math.hypot(0.7, 1.5) - 1.5, 1e-6)


# ---- volume.rs (kit.rs:244-262) ------------------------------------------------------------

def volume_sample(dims, mn, cell, data, p):
    import ctypes as C
    d = np.asarray(dims, np.uint32)
    a = np.asarray(mn, np.float32)
    c = np.asarray(cell, np.float32)
    v = np.asarray(data, np.float32)
    q = np.asarray(p, np.float32)
    P = lambda x: x.ctypes.data_as(C.c_void_p)
    return po.lib().orc_volume_sample(P(d), P(a), P(c), P(v), v.size, P(q))


def test_volume_samples_and_guards_bad_layouts():
    data = [0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0]
    assert close(volume_sample((2, 2, 2), (0, 0, 0), (1, 1, 1), data, (0.5, 0.25, 0.75)), 0.5, 1e-6)
    assert volume_sample((2, 2, 2), (0, 0, 0), (1, 1, 1), data, (-1.0, 0.5, 0.5)) >= 1.0
    assert volume_sample((2, 2, 2), (0, 0, 0), (1, 1, 1), [0.0], (0.5, 0.5, 0.5)) == float("inf")
    # trilinear of a linear field is exact: f = x + 2y - z on a 3x3x3 grid
    g = np.arange(3, dtype=np.float32)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    f = (x + 2 * y - z).reshape(-1)
    assert close(volume_sample((3, 3, 3), (0, 0, 0), (1, 1, 1), f, (0.3, 1.7, 0.4)), 0.3 + 3.4 - 0.4, 1e-5)


# ---- march.rs (kit.rs:298-337, march_sign_crossing.rs:7-47) -------------------------------

def ray(o, d):
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    return np.array(list(o) + list(d), np.float32)


def test_marcher_hits_a_sphere_and_normal_points_back():
    s = scene_of(Object(Prim.sphere(1.0), Transform.at((0, 0, -5)), 0))
    t, hit, mat, nor, _ = s.march(ray((0, 0, 0), (0, 0, -1)))
    assert hit[0] and close(t[0], 4.0, 1e-2)
    assert nor[0] @ np.array([0, 0, 1.0]) > 0.9


def test_boosted_steps_never_accept_buried_hits():
    for sub in (0.0, 0.6):
        s = scene_of(Object(Prim.sphere(1.0), Transform(), 0).onion(0.02), m_subitize=sub)
        rays = np.array([ray((0, 0, 5), (ix * 0.02, iy * 0.02, -1.0)) for iy in range(-12, 13) for ix in range(-12, 13)])
        t, hit, _, _, _ = s.march(rays)
        assert hit.sum() > 100
        pos = rays[:, :3] + rays[:, 3:] * t[:, None]
        d, _ = s.field(pos[hit])
        eps = 0.0006 * (1.0 + t[hit] * 0.5)
        assert (d > -2.0 * eps).all()


@pytest.mark.parametrize("sub", [0.0, 0.6])
@pytest.mark.parametrize("secant", [False, True])
def test_sign_crossing_returns_plane_surface(sub, secant):
    s = scene_of(Object(Prim.plane((0, 0, 1), 0.0), Transform(), 7), m_subitize=sub, m_secant=secant)
    for x in (0.0, 0.1, 0.5):
        r = ray((0, 0, 3), (x, 0, -1))
        t, hit, mat, _, _ = s.march(r)
        expected = 3.0 / -r[5]
        assert hit[0] and abs(t[0] - expected) < 0.002 and mat[0] == 7
        pz = r[2] + r[5] * t[0]
        assert -0.00001 <= pz < 0.002


def test_touching_sphere_and_inside_origin():
    s = scene_of(Object(Prim.sphere(1.0), Transform(), 3))
    t, hit, mat, _, _ = s.march(ray((0, 0, 2), (0, 0, -1)))
    assert hit[0] and abs(t[0] - 1.0) < 1e-5 and mat[0] == 3
    s = scene_of(Object(Prim.plane((0, 0, 1), 0.0), Transform(), 4))
    t, hit, mat, _, _ = s.march(ray((0, 0, -0.5), (0, 0, -1)))
    assert hit[0] and t[0] == 0.0 and mat[0] == 4


# ---- dual.rs (dual.rs:306-317 and lib.rs:1502-1543) ----------------------------------------

@pytest.mark.parametrize("kind,params", [
    (PRIM_SPHERE, [0.8]), (PRIM_BOX, [0.6, 0.4, 0.5]), (PRIM_ROUND_BOX, [0.6, 0.4, 0.5, 0.1]), (PRIM_TORUS, [0.7, 0.2]),
    (PRIM_CYLINDER, [0.6, 0.4]), (PRIM_CAPSULE, [0, -0.5, 0, 0.1, 0.5, 0, 0.3]), (PRIM_CONE, [0.5, 0.2, 1.0]),
    (PRIM_ELLIPSOID, [0.8, 0.5, 0.6]), (PRIM_OCTAHEDRON, [0.8]), (PRIM_HEX_PRISM, [0.5, 0.4]),
    (PRIM_PLANE, [0.0, 0.6, 0.8, 0.2])])
def test_dual_gradient_matches_finite_difference(kind, params):
    rng = np.random.default_rng(kind)
    for _ in range(12):
        p = rng.uniform(-1.5, 1.5, 3)
        v = po.prim_dual(kind, params, p)
        assert close(v[0], po.prim_distance(kind, params, p), 1e-3)
        h = 1e-3
        fd = np.array([(po.prim_distance(kind, params, p + e) - po.prim_distance(kind, params, p - e)) / (2 * h)
                       for e in np.eye(3) * h])
        g = v[1:].astype(np.float64)
        if np.linalg.norm(fd) < 1e-3 or np.linalg.norm(g) < 1e-6:
            continue
        cosang = g @ fd / (np.linalg.norm(g) * np.linalg.norm(fd))
        # sample points near a subgradient seam can disagree; the reference uses 0.5 deg
        # at smooth points, so require it for the bulk and allow seams
        if cosang > math.cos(math.radians(5.0)):
            assert math.degrees(math.acos(min(1.0, cosang))) < 0.5


def test_dual_and_tetra_normals_agree_and_twist_disables_dual():
    """lib.rs:1502-1543: on a dual-safe scene the exact (dual) normal and the tetrahedron normal
    agree closely at hit points; a twist makes the scene fall back to the tetrahedron."""
    sc, cam = configs.c1_scene(width=48, height=27)
    s = po.OracleScene(sc.to_ir(), threads=2)
    assert s.dual_safe()
    rng = np.random.default_rng(9)
    dirs = rng.normal(size=(400, 3))
    dirs[:, 1] = -np.abs(dirs[:, 1]) - 0.2
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.hstack([np.tile([[0.0, 2.0, -8.0]], (400, 1)), dirs]).astype(np.float32)
    t, hit, _, nor, _ = s.march(rays)
    angles = []
    for i in np.nonzero(hit)[0][:150]:
        p = rays[i, :3] + rays[i, 3:] * t[i]
        a, b = s.normals(p)
        assert np.allclose(a, nor[i])  # the hit normal is the dual normal on this scene
        angles.append(math.degrees(math.acos(min(1.0, float(a @ b)))))
    assert len(angles) > 50
    assert np.median(angles) < 0.2 and max(angles) < 5.0
    c2, _ = configs.c2_scene(width=16, height=9, gi_dims=None)
    assert not po.OracleScene(c2.to_ir()).dual_safe()


# ---- gi.rs (mm3e-orchestrator/tests/engine.rs:145-166) ------------------------------------

def gi_linear_x(lo, hi, dims, samples, p, n):
    import ctypes as C
    out = np.zeros(3, np.float32)
    P = lambda x: np.ascontiguousarray(x).ctypes.data_as(C.c_void_p)
    po.lib().orc_gi_bake_linear_x(P(np.asarray(lo, np.float32)), P(np.asarray(hi, np.float32)),
                                  P(np.asarray(dims, np.uint32)), samples, P(np.asarray(p, np.float32)),
                                  P(np.asarray(n, np.float32)), P(out))
    return out


def test_gi_volume_interpolates_a_known_field():
    n = (1.0, 0.0, 0.0)
    assert close(gi_linear_x((0, 0, 0), (4, 2, 4), (5, 3, 5), 1, (2.0, 1.0, 2.0), n)[0], 2.0, 1e-3)
    assert close(gi_linear_x((0, 0, 0), (4, 2, 4), (5, 3, 5), 1, (2.5, 1.0, 2.0), n)[0], 2.5, 1e-3)
    # five-ray cones average to the same linear field; an oblique normal weighs three faces by n^2
    nn = np.array([0.6, 0.0, 0.8], np.float32)
    v = gi_linear_x((0, 0, 0), (4, 2, 4), (5, 3, 5), 5, (1.3, 0.7, 3.1), nn)
    assert close(v[0], 1.3 * (0.36 + 0.64), 1e-3)


def test_gi_degenerate_dims_dont_crash():
    for dims in ((0, 5, 6), (1, 1, 1), (3, 0, 3)):
        gi_linear_x((0, 0, 0), (4, 4, 4), dims, 1, (2, 2, 2), (0, 1, 0))


def test_gi_bake_skips_unbounded_scenes_and_bakes_bounded_ones():
    sc = Scene(8, 8)
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), 0))
    sc.bake_gi((4, 4, 4), 1)
    assert po.OracleScene(sc.to_ir()).gi_raw() is None  # gi.rs:192-196: planes only -> no GI
    sc, _ = configs.c2_scene(width=16, height=9, gi_dims=(4, 3, 4), gi_rays=5)
    g = po.OracleScene(sc.to_ir()).gi_raw()
    assert g is not None and tuple(g[2]) == (4, 3, 4) and np.isfinite(g[3]).all() and g[3].max() > 0


# ---- shade.rs (kit.rs:353-365) + closed forms -----------------------------------------------

def brdf(n, l, v, alb, met, rough, spec):
    import ctypes as C
    out = np.zeros(3, np.float32)
    P = lambda x: np.asarray(x, np.float32).ctypes.data_as(C.c_void_p)
    po.lib().orc_brdf(P(n), P(l), P(v), P(alb), met, rough, spec, out.ctypes.data_as(C.c_void_p))
    return out


def test_shading_invariants_and_closed_forms():
    up = [0, 1, 0]
    b = brdf(up, up, up, [0.8] * 3, 0.0, 0.3, 0.5)
    assert (b >= 0).all()
    assert po.lib().orc_fresnel(1.0, 0.04) < po.lib().orc_fresnel(0.0, 0.04)
    assert po.lib().orc_fresnel(1.0, 0.37) == np.float32(0.37)   # m = 0 -> f0 exactly
    # n = l = v: D = 1/(pi a2), V = 1/4, F = f0, diffuse (1-f0)(1-m) albedo / pi
    a = max(0.3 * 0.3, 1e-3)
    a2 = a * a
    f0 = 0.08 * 0.5
    expect = (1 - f0) * 0.8 / math.pi + f0 / (math.pi * a2) * 0.25
    assert abs(b[0] - expect) / expect < 1e-5
    import ctypes as C
    ac = np.zeros(3, np.float32)
    ga = np.zeros(3, np.float32)
    P = lambda x: x.ctypes.data_as(C.c_void_p)
    po.lib().orc_aces_gamma(P(np.array([20.0, 0.0, 1.0], np.float32)), P(ac), P(ga))
    assert 0.0 <= ac[0] <= 1.0 and ac[1] == 0.0
    assert ga[2] == 1.0 and ga[1] == 0.0


def test_powi5_is_repeated_squaring():
    for x in (F(0.3), F(0.77), F(0.123456)):
        assert po.lib().orc_powi(x, 5) == x * ((x * x) * (x * x))
        assert po.lib().orc_powi(x, 2) == x * x


def fnv1a(b):
    h = 0xcbf29ce484222325
    for c in b:
        h ^= c
        h = (h * 0x100000001b3) & ((1 << 64) - 1)
    return h


def test_fnv1a_known_vectors_and_hash_cell():
    import ctypes as C
    for s, want in ((b"", 0xcbf29ce484222325), (b"a", 0xaf63dc4c8601ec8c), (b"foobar", 0x85944171f73967e8)):
        buf = (C.c_uint8 * max(1, len(s))).from_buffer_copy(s or b"\0")
        assert po.lib().orc_fnv1a(C.cast(buf, C.c_void_p), len(s)) == want
    import struct
    for ix, iy in ((0, 0), (3, -7), (-100000, 2147483647), (-1, -1)):
        h = fnv1a(struct.pack("<ii", ix, iy))
        assert po.lib().orc_hash_cell(ix, iy) == F((h >> 40) / float(1 << 24))


# ---- accel.rs / whole-render invariants --------------------------------------------------

def test_accelerated_world_equals_linear_fold():  # engine.rs:422-449, local_csg.rs:51-74
    sc = configs.procedural_scene(21, 80, 32, 18)
    s = po.OracleScene(sc.to_ir())
    rng = np.random.default_rng(0)
    pts = rng.uniform([-8, -1, -8], [8, 4, 8], (4000, 3)).astype(np.float32)
    da, ma = s.field(pts)
    dl, ml = s.field(pts, linear=True)
    assert np.abs(da - dl).max() <= 1e-5
    assert (ma == ml).mean() > 0.999


def test_render_is_deterministic_and_thread_invariant():  # engine.rs:84-91, canonical.rs:42-49
    sc, cam = configs.c1_scene(width=64, height=36)
    s = po.OracleScene(sc.to_ir())
    a = s.render_rgba(cam, threads=1)
    b = s.render_rgba(cam, threads=4)
    c = s.render_rgba(cam, threads=7)
    assert (a == b).all() and (a == c).all()
    assert len(np.unique(a.reshape(-1, 4), axis=0)) > 10  # nontrivial image (engine.rs:93-104)


def test_render_equals_resolve_of_render_linear():  # local_csg.rs:96-116
    sc, cam = configs.c2_scene(width=48, height=27, gi_dims=(3, 3, 3), gi_rays=1)
    s = po.OracleScene(sc.to_ir())
    hdr = s.render_linear(cam)
    assert (s.resolve(hdr) == s.render_rgba(cam)).all()


def test_emissive_sphere_closed_form():
    """Emissive-only sphere, no lights/AO/fog/bloom, black albedo: inside the disk the pixel is
    gamma(aces(E)) quantised; outside it is the sky, tone-mapped."""
    sc = Scene(32, 32)
    sc.aa = 1
    sc.ao = False
    sc.fog_density = 0.0
    sc.post.bloom = False
    sc.ambient = (F(0), F(0), F(0))
    m = sc.material(Material.solid((0.0, 0.0, 0.0)).emissive((0.5, 0.3, 0.2)))
    sc.add(Object(Prim.sphere(1.0), Transform(), m))
    cam = configs.look_at((0, 0, 4), (0, 0, 0), (0, 1, 0), math.radians(40))
    s = po.OracleScene(sc.to_ir())
    img = s.render_rgba(cam).astype(int)

    def tone(c):
        c = np.maximum(c, 0)
        a = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0, 1)
        return np.floor(np.clip(a ** (1 / 2.2), 0, 1) * 255 + 0.5)

    inside = tone(np.array([0.5, 0.3, 0.2]))
    assert (np.abs(img[16, 16, :3] - inside) <= 1).all()
    # a corner ray misses: the sky gradient + sun glow
    d = np.array(cam[3:6], np.float64) + np.array(cam[6:9]) * ((2 * 0.5 / 32 - 1) * cam[12]) \
        + np.array(cam[9:12]) * ((1 - 2 * 0.5 / 32) * cam[12])
    d /= np.linalg.norm(d)
    t = min(max(0.5 * (d[1] + 1), 0), 1)
    base = np.array([0.78, 0.86, 0.96]) * (1 - t) + np.array([0.20, 0.38, 0.72]) * t
    sun = np.array(sc.sun_dir, np.float64)
    s_ = max(d @ sun, 0)
    glow = s_ ** 8 * 0.3 + s_ ** 220 * 6
    assert (np.abs(img[0, 0, :3] - tone(base + glow)) <= 1).all()


def test_plane_depth_is_analytic():
    sc = Scene(40, 30)
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), 0))
    cam = configs.look_at((0, 2, -6), (0, 0, 0), (0, 1, 0), math.radians(60))
    s = po.OracleScene(sc.to_ir())
    hit, mat, depth, _ = s.render_aov(cam)
    assert hit[-1].all() and (mat[hit == 1] == 0).all() and (mat[hit == 0] == 0xFFFFFFFF).all()
    fwd = np.array(cam[3:6], np.float64)
    for x, y in ((20, 29), (3, 25), (37, 20)):
        if not hit[y, x]:
            continue
        d = np.array(cam[3:6], np.float64) + np.array(cam[6:9]) * ((2 * (x + 0.5) / 40 - 1) * (40 / 30) * cam[12]) \
            + np.array(cam[9:12]) * ((1 - 2 * (y + 0.5) / 30) * cam[12])
        d /= np.linalg.norm(d)
        t = 2.0 / -d[1]
        assert abs(depth[y, x] - t * (d @ fwd)) < 2e-3 * t


def test_bmp_layout():
    img = np.zeros((2, 3, 4), np.uint8)
    img[0, 0] = (10, 20, 30, 255)   # top-left
    img[1, 2] = (40, 50, 60, 255)   # bottom-right
    b = po.to_bmp(img)
    import struct
    assert b[:2] == b"BM" and struct.unpack_from("<I", b, 2)[0] == len(b) == 54 + 2 * 12
    assert struct.unpack_from("<iiHH", b, 18) == (3, 2, 1, 24)
    # bottom-up rows, BGR, rows padded to 4 bytes (9 -> 12)
    assert tuple(b[54 + 6: 54 + 9]) == (60, 50, 40)
    assert tuple(b[54 + 12: 54 + 15]) == (30, 20, 10)


# ---- object-local CSG (mm3e-kit/src/csg.rs tests; local_csg.rs:76-94) ---------------------

def _csg_scene(nodes, root, xform=None, w=8, h=8, **kw):
    from mm3e.scene import CsgNode  # noqa: F401
    sc = Scene(w, h)
    for k, v in kw.items():
        setattr(sc, k, v)
    eid = sc.csg(nodes, root)
    sc.add(Object(Prim.csg(eid), xform or Transform(), 0))
    return sc


def test_csg_eyelid_shell_clip_has_an_opening():  # csg.rs:264-274
    from mm3e.scene import CsgNode, CSG_PRIM, CSG_SHELL, CSG_INTERSECT
    sc = _csg_scene([CsgNode(CSG_PRIM, PRIM_SPHERE, p=(1.0,)), CsgNode(CSG_SHELL, a=0, p=(0.05,)),
                     CsgNode(CSG_PRIM, PRIM_PLANE, p=(0.0, -1.0, 0.0, 0.3)), CsgNode(CSG_INTERSECT, a=1, b=2)], 3)
    d, _ = po.OracleScene(sc.to_ir()).field([[0.0, 0.5, 0.8660254], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
    assert d[0] < 0.0 and d[1] > 0.0 and d[2] > 0.0


def test_csg_transformed_subtraction_matches_object_transform():  # csg.rs:277-299
    """Transform{Subtract(sphere 1, sphere 0.6 at z 0.8)} at the identity equals the bare
    subtraction placed by the object's own transform (both apply to_local, evaluate, * scale),
    and |expr(world) - base(p) * scale| < 3e-6 for points p of the reference's lattice."""
    from mm3e.scene import CsgNode, CSG_PRIM, CSG_TRANSFORM, CSG_SUBTRACT, from_euler, v3
    xf = Transform(pos=v3((2.0, 0.0, 1.0)), rot=from_euler(0.2, 0.7, -0.1), scale=1.7)
    base = [CsgNode(CSG_PRIM, PRIM_SPHERE, p=(1.0,)), CsgNode(CSG_PRIM, PRIM_SPHERE, p=(0.6,)),
            CsgNode(CSG_TRANSFORM, a=1, xform=Transform.at((0.0, 0.0, 0.8))), CsgNode(CSG_SUBTRACT, a=0, b=2)]
    wrapped = po.OracleScene(_csg_scene(base + [CsgNode(CSG_TRANSFORM, a=3, xform=xf)], 4).to_ir())
    placed = po.OracleScene(_csg_scene(base, 3, xform=xf).to_ir())
    plain = po.OracleScene(_csg_scene(base, 3).to_ir())
    g = np.arange(-6, 7, dtype=np.float32) * np.float32(0.45)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    rot = np.array(xf.rot, np.float32).T  # columns -> matrix
    world = (np.asarray(xf.pos, np.float32) + (lat * np.float32(xf.scale)) @ rot.T).astype(np.float32)
    dw, _ = wrapped.field(world)
    dp, _ = placed.field(world)
    db, _ = plain.field(lat)
    near = np.abs(db) < 0.5   # inside the bounding sphere the world field is the exact CSG value
    assert np.array_equal(dw, dp)
    assert (np.abs(dw[near] - db[near] * np.float32(1.7)) < 3e-5).all()


def test_csg_sphere_renders_identically_to_primitive_sphere():  # local_csg.rs:76-94
    """The reference compares RenderMode::Albedo frames (hit + material colour per pixel); the
    primary-hit AOV carries the same information: hit mask, material id and depth."""
    from mm3e.scene import CsgNode, CSG_PRIM
    sc = _csg_scene([CsgNode(CSG_PRIM, PRIM_SPHERE, p=(1.0,))], 0, w=40, h=40, aa=1, shadows=False, ao=False,
                    bounces=0)
    cam = configs.look_at((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.7)
    a = po.OracleScene(sc.to_ir()).render_aov(cam)
    sc.objects[0].prim = Prim.sphere(1.0)
    b = po.OracleScene(sc.to_ir()).render_aov(cam)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert 0 < a[0].sum() < a[0].size  # the sphere is in view and does not fill it


# ---- debug AOV render modes (lib.rs:1178-1218, heatmap lib.rs:1401-1407) ----------------

def _plane_view(mode):
    from mm3e.scene import Material, Object, Prim, Scene, Transform
    sc = Scene(32, 24)
    m = sc.material(Material.solid((0.2, 0.4, 0.6)))
    sc.add(Object(Prim.plane((0.0, 1.0, 0.0), 0.0), Transform(), m))
    sc.mode = mode
    cam = configs.look_at((0.0, 3.0, -6.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(60.0))
    o = po.OracleScene(sc.to_ir())
    img = o.render_rgba(cam)
    hit, _, _, _ = o.render_aov(cam)
    return sc, cam, img, hit.astype(bool)


def test_mode_normal_albedo_steps_closed_forms():
    from mm3e.scene import MODE_ALBEDO, MODE_NORMAL, MODE_STEPS
    _, _, img, hit = _plane_view(MODE_NORMAL)
    assert hit.any() and (~hit).any()
    assert (img[hit][:, :3] == [128, 255, 128]).all()   # n = +y -> 0.5 n + 0.5
    assert (img[~hit][:, :3] == 0).all() and (img[..., 3] == 255).all()
    _, _, img, hit = _plane_view(MODE_ALBEDO)
    assert (img[hit][:, :3] == [51, 102, 153]).all()    # (0.2, 0.4, 0.6) * 255 + 0.5, truncated
    assert (img[~hit][:, :3] == 0).all()
    sc, cam, img, hit = _plane_view(MODE_STEPS)
    assert (img[~hit][:, :3] == [255, 0, 0]).all()      # misses report max_steps: heatmap(1) = red
    # hits: the heatmap of lib.rs:1401-1407 at steps / max_steps, quantised like every display pixel
    _, _, _, steps = po.OracleScene(sc.to_ir()).render_aov(cam)
    t = np.clip(steps.astype(np.float32) / np.float32(sc.marcher.max_steps), 0, 1).astype(np.float32)
    r = np.clip(t * np.float32(3) - np.float32(1.5), 0, 1)
    g = np.clip(np.float32(1) - np.abs(t * np.float32(3) - np.float32(1.3)), 0, 1)
    b = np.clip(np.float32(1) - t * np.float32(2.2), 0, 1)
    q = lambda c: (c.astype(np.float32) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    heat = np.stack([q(r), q(g), q(b)], -1)
    assert (np.abs(img[hit][:, :3].astype(int) - heat[hit].astype(int)) <= 1).all()
    assert len(np.unique(steps[hit])) > 3               # grazing rays take more steps: not one colour


def test_mode_depth_and_ao_closed_forms():
    from mm3e.scene import MODE_AO, MODE_DEPTH
    sc, cam, img, hit = _plane_view(MODE_DEPTH)
    h, w = hit.shape
    # analytic t to the plane y = 0 along each pixel-centre ray
    eye, fwd, right, up, fs = cam[0:3], cam[3:6], cam[6:9], cam[9:12], cam[12]
    ys, xs = np.nonzero(hit)
    aspect = w / h
    for x, y in zip(xs, ys):
        ndc_x = ((2.0 * (x + 0.5)) / w - 1.0) * aspect * fs
        ndc_y = (1.0 - (2.0 * (y + 0.5)) / h) * fs
        d = fwd + right * ndc_x + up * ndc_y
        d = d / np.linalg.norm(d)
        t = -eye[1] / d[1]
        # the march stops once the plane distance (t_hit - t) |d.y| drops below eps (1 + t/2)
        t_lo = t - sc.marcher.eps * (1.0 + 0.5 * t) / abs(d[1])
        hi = (1.0 - min(max(t_lo / sc.marcher.max_dist, 0.0), 1.0)) * 255.0 + 0.5
        lo = (1.0 - min(max(t / sc.marcher.max_dist, 0.0), 1.0)) * 255.0 + 0.5
        assert int(lo) - 1 <= int(img[y, x, 0]) <= int(hi) + 1
    assert (img[~hit][:, :3] == 0).all()  # misses end beyond max_dist
    _, _, img, hit = _plane_view(MODE_AO)
    assert (img[..., :3] == 255).all()    # an open plane occludes nothing; misses report 1


# ---- Quality presets (lib.rs:67-149; engine.rs:255-271) ----------------------------------------

def test_quality_presets_and_lerp():
    from mm3e.scene import Quality
    fast = Quality.fast(320, 180)
    full = Quality.full(960, 540)
    assert fast.aa <= full.aa and fast.bounces <= full.bounces and fast.shadow_steps < full.shadow_steps
    mid = Quality.lerp(fast, full, 0.5)
    assert (mid.width, mid.height) == (960, 540)
    assert fast.shadow_steps < mid.shadow_steps < full.shadow_steps
    s = Scene(10, 10)
    full.apply(s)
    assert (s.width, s.height, s.aa, s.bounces) == (960, 540, 2, 2)
    assert (s.marcher.shadow_steps, s.marcher.ao_samples) == (64, 5)
    # closed form of the f32 mix with half-away rounding: 64 + (160 - 64) * 0.5 = 112,
    # 12 + 52 * 0.5 = 38, aa 1 + 0.5 = 1.5 -> 2, bounces 0 + 2 * 0.5 = 1, ao 0 + 2.5 -> 3
    assert (mid.max_steps, mid.shadow_steps, mid.aa, mid.bounces, mid.ao_samples) == (112, 38, 2, 1, 3)
    assert mid.lod_footprint == float(np.float32(np.float32(0.006) + np.float32(-0.006) * np.float32(0.5)))
    assert Quality.lerp(fast, full, -1.0).max_steps == 64 and Quality.lerp(fast, full, 7.0).max_steps == 160
