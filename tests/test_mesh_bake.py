"""Mesh -> SDF bake (mm3e-orchestrator/src/mesh.rs) and the `.sdfv` codec (mm3e-kit/src/volume.rs).

CPU tests pin the oracle (oracle/mesh_oracle.cpp) against the reference's own tests
(mm3e-kit/src/triangle.rs:36-67, mm3e-orchestrator/tests/engine.rs:403-418,
mm3e-kit/tests/kit.rs:264-294) and against closed forms (a box's exact signed distance,
point-triangle projections, brute-force nearest triangles), and check the library's host logic
(OBJ parsing, grid layout, `.sdfv`) against the oracle and the reference's messages.
GPU tests check the device bake against the oracle bit for bit.
"""
import math
import struct

import numpy as np
import pytest

import pyoracle as po
from mm3e import configs, hip


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ---- OBJ import (mesh.rs:57-103) ---------------------------------------------------------------

def test_obj_parser_reference_kat():
    # engine.rs:405-411
    for parse, Err in ((hip.parse_obj, hip.Mm3eError), (po.mesh_parse_obj, po.OracleError)):
        pos, tri = parse("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
        assert pos.shape == (3, 3) and tri.tolist() == [[0, 1, 2]]
        with pytest.raises(Err):
            parse("v 0 0\nf 1 2 3\n")


def test_obj_forms_fans_and_negative_indices():
    text = ("# quad and a pentagon\r\n"
            "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0   # trailing comment\n"
            "vt 0 0\nvn 0 0 1\no thing\ng grp\n"
            "f 1/1/1 2/2/1 3//1 4\n"
            "v +2 -0.5 1e-1\n"
            "f -5 -4 -1 -2 -3\n")
    pos, tri = hip.parse_obj(text)
    assert pos.shape == (5, 3)
    assert pos[4].tolist() == [2.0, -0.5, np.float32(0.1)]
    assert tri.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4], [0, 4, 3], [0, 3, 2]]
    p2, t2 = po.mesh_parse_obj(text)
    assert np.array_equal(bits(p2), bits(pos)) and np.array_equal(t2, tri)


@pytest.mark.parametrize("text,msg", [
    ("", "mesh has no vertices"),
    ("v 0 0 0\nv 1 0 0\n", "mesh has no triangles"),
    ("v 0 0\n", "line 1: missing z"),
    ("v 0\n", "line 1: missing y"),
    ("v\n", "line 1: missing x"),
    ("v 0 0 zz\n", "line 1: bad number for z"),
    ("v 0 0 0\nv inf 0 0\n", "line 2: non-finite vertex"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 x\n", "line 4: bad face index 'x'"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "line 4: face index 4 out of range (have 3 vertices)"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n", "line 4: face index 0 out of range (have 3 vertices)"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -4 1 2\n", "line 4: face index -4 out of range (have 3 vertices)"),
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", "line 3: face needs at least 3 vertices"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf +-1 2 3\n", "line 4: bad face index '+-1'"),
])
def test_obj_errors_carry_reference_messages(text, msg):
    with pytest.raises(hip.Mm3eError) as e:
        hip.parse_obj(text)
    assert str(e.value) == msg
    with pytest.raises(po.OracleError) as e2:
        po.mesh_parse_obj(text)
    assert str(e2.value) == msg


def test_obj_round_trip_of_generated_meshes():
    for pos, tri in (configs.uv_sphere_mesh(12, 18, 1.0), configs.torus_knot_mesh(segments=40, ring_verts=12)):
        p2, t2 = hip.parse_obj(configs.mesh_to_obj(pos, tri))
        assert np.array_equal(bits(p2), bits(pos)) and np.array_equal(t2, tri)


# ---- x_crossing (mm3e-kit/src/triangle.rs) ----------------------------------------------------

def test_x_crossing_shared_diagonal_counted_once():
    # triangle.rs:40-52
    a, b, c, d = (2.0, -1.0, -1.0), (2.0, 1.0, -1.0), (2.0, 1.0, 1.0), (2.0, -1.0, 1.0)
    for tris in ([[a, b, c], [a, c, d]], [[c, b, a], [d, c, a]], [[a, c, b], [a, c, d]]):
        for y in (-0.7, 0.0, 0.4, 0.75):
            hits = [x for x in (po.x_crossing(*t, y, y) for t in tris) if x is not None]
            assert hits == [2.0]


def test_x_crossing_projected_vertices_and_thin_faces():
    # triangle.rs:54-66
    o = (1.0, 0.0, 0.0)
    p = [(1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (1.0, 1.0, 1.0), (1.0, -1.0, 1.0)]
    hits = [po.x_crossing(o, p[i], p[(i + 1) % 4], 0.0, 0.0) for i in range(4)]
    assert sum(h is not None for h in hits) == 1
    assert po.x_crossing(o, (1.0, 1e-7, 0.0), (1.0, 0.0, 1e-7), 2e-8, 2e-8) == 1.0


def test_x_crossing_closed_form_on_a_slanted_triangle():
    # plane x = 1 + 0.5 y + 0.25 z through the three vertices; a line inside hits it there,
    # a line outside misses, a degenerate (edge-on) projection never hits
    a, b, c = (1.0, 0.0, 0.0), (1.5, 1.0, 0.0), (1.25, 0.0, 1.0)
    for y, z in ((0.25, 0.25), (0.5, 0.125), (0.125, 0.5)):
        assert po.x_crossing(a, b, c, y, z) == pytest.approx(1.0 + 0.5 * y + 0.25 * z, abs=1e-7)
        assert po.x_crossing(a, c, b, y, z) == po.x_crossing(a, b, c, y, z)   # winding-free
    assert po.x_crossing(a, b, c, 0.75, 0.5) is None
    assert po.x_crossing((0, 0, 0), (1, 1, 1), (2, 2, 2), 0.5, 0.5) is None


# ---- nearest triangle (mesh.rs:111-260) -------------------------------------------------------

def test_closest_on_triangle_regions_closed_form():
    a, b, c = (0.0, 0.0, 0.0), (4.0, 0.0, 0.0), (0.0, 4.0, 0.0)
    cases = [
        ((1.0, 1.0, 3.0), (1.0, 1.0, 0.0)),      # interior: orthogonal projection
        ((-1.0, -2.0, 1.0), a),                  # vertex regions
        ((6.0, -1.0, 0.0), b),
        ((-1.0, 7.0, 2.0), c),
        ((2.0, -3.0, 1.0), (2.0, 0.0, 0.0)),     # edge ab
        ((-2.0, 1.0, 0.5), (0.0, 1.0, 0.0)),     # edge ac
        ((3.0, 3.0, 0.0), (2.0, 2.0, 0.0)),      # edge bc
    ]
    for p, want in cases:
        assert po.closest_on_triangle(p, a, b, c).tolist() == list(want)


def test_bvh_nearest_equals_brute_force():
    rng = np.random.default_rng(5)
    for pos, tri in (configs.uv_sphere_mesh(12, 18, 1.0), configs.torus_knot_mesh(segments=60, ring_verts=12),
                     configs.box_mesh()):
        pts = rng.uniform(-2.5, 2.5, size=(400, 3)).astype(np.float32)
        fast, (boxes, tris) = po.mesh_nearest_sq(pos, tri, pts)
        slow, _ = po.mesh_nearest_sq(pos, tri, pts, brute=True)
        assert np.array_equal(fast, slow)
        assert tris < 400 * len(tri) / 4   # the tree prunes


# ---- bake (mesh.rs:265-336) --------------------------------------------------------------------

def box_sdf(p, half):
    q = np.abs(p) - half
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside


def grid_points(dims, lo, cell):
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lo = np.asarray(lo, np.float32)
    cell = np.asarray(cell, np.float32)
    # f32, as the bake computes them: lo + index * cell
    return np.stack([lo[0] + i.astype(np.float32) * cell[0], lo[1] + j.astype(np.float32) * cell[1],
                     lo[2] + k.astype(np.float32) * cell[2]], axis=-1)


def test_bake_box_matches_exact_signed_distance():
    half = np.array([1.0, 0.75, 0.5])
    pos, tri = configs.box_mesh(tuple(half))
    dims, lo, cell, d = po.mesh_bake(pos, tri, 25, 0.5)
    # padded size (3, 2.5, 2) with 24 steps along the longest axis: 0.125 everywhere
    assert dims == (25, 21, 17)
    assert lo.tolist() == [-1.5, -1.25, -1.0] and cell.tolist() == [0.125, 0.125, 0.125]
    p = grid_points(dims, lo, cell).astype(np.float64)
    want = box_sdf(p, half)
    on_face = np.isclose(np.abs(want), 0.0, atol=1e-6)
    assert np.abs(d - want)[~on_face].max() < 2e-6
    assert np.all(np.sign(d[~on_face]) == np.sign(want[~on_face]))


def test_bake_sphere_reference_tolerance():
    # engine.rs:413-418: resolution 28, padding 0.25, |sample - (|p| - 1)| < 0.16 at three points
    pos, tri = configs.uv_sphere_mesh(12, 18, 1.0)
    dims, lo, cell, d = po.mesh_bake(pos, tri, 28, 0.25)
    for p in ((0.0, 0.0, 0.0), (1.2, 0.0, 0.0), (0.0, 1.2, 0.0)):
        got = po.volume_sample(dims, lo, cell, d, p)
        assert abs(got - (math.sqrt(sum(c * c for c in p)) - 1.0)) < 0.16


def test_bake_is_one_lipschitz_and_signed_on_a_closed_knot():
    # the mesh demo's Lipschitz report (mesh_demo.rs:72-89): neighbours differ by <= 1.02 cells
    pos, tri = configs.torus_knot_mesh(segments=80, ring_verts=16)
    dims, lo, cell, d = po.mesh_bake(pos, tri, 40, 0.35)
    assert np.abs(np.diff(d, axis=2)).max() <= cell[0] * 1.02
    assert np.abs(np.diff(d, axis=1)).max() <= cell[1] * 1.02
    assert d.min() < 0.0 < d.max()
    # a grid point on the tube's centre line is inside by about the tube radius
    assert -0.23 < d.min() < -0.12


def test_layout_matches_oracle_and_reference_errors():
    for pos, tri, res, pad in ((*configs.uv_sphere_mesh(12, 18, 1.0), 28, 0.25),
                               (*configs.torus_knot_mesh(segments=40, ring_verts=12), 160, 0.35),
                               (*configs.box_mesh((1.0, 0.01, 0.5)), 3, 0.1),
                               (*configs.box_mesh(), 1000, 0.2)):
        lay = hip.sdf_layout(pos, tri, res, pad)
        dims, lo, cell, _ = po.mesh_bake(pos, tri, res, pad, k0=0, k1=1)
        assert tuple(lay.dims) == dims
        assert np.array_equal(bits(lay.min), bits(lo)) and np.array_equal(bits(lay.cell), bits(cell))
    assert tuple(hip.sdf_layout(*configs.box_mesh(), 1000, 0.2).dims) == (512, 512, 512)
    pos, tri = configs.box_mesh()
    for pad in (0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(hip.Mm3eError, match="padding must be positive and finite"):
            hip.sdf_layout(pos, tri, 16, pad)
    with pytest.raises(hip.Mm3eError, match="mesh has no non-degenerate triangles"):
        hip.sdf_layout(np.zeros((3, 3), np.float32), [[0, 1, 2]], 16, 0.1)
    with pytest.raises(hip.Mm3eError, match="mesh triangle 1 references a vertex out of range"):
        hip.sdf_layout(pos, [[0, 1, 2], [0, 1, 8]], 16, 0.1)
    with pytest.raises(hip.Mm3eError, match="mesh vertex 1 is non-finite"):
        hip.sdf_layout([[0, 0, 0], [np.nan, 0, 0], [0, 1, 0]], [[0, 1, 2]], 16, 0.1)
    with pytest.raises(po.OracleError, match="mesh has no non-degenerate triangles"):
        po.mesh_bake(np.zeros((3, 3), np.float32), [[0, 1, 2]], 16, 0.1)


# ---- .sdfv codec (volume.rs:115-187) -------------------------------------------------------------

def sdfv_reference_volume():
    lay = hip.SdfLayout()
    lay.dims[:] = (2, 2, 2)
    lay.min[:] = (-1.0, -2.0, -3.0)
    lay.cell[:] = (0.5, 0.75, 1.25)
    return lay, np.array([0.0, 0.5, 1.0, 1.5, -0.25, 0.25, 0.75, 1.25], np.float32)


def test_sdfv_round_trip_reference_kat():
    # kit.rs:264-294
    lay, data = sdfv_reference_volume()
    blob = hip.sdfv_encode(lay, data)
    assert blob[:8] == b"MM3ESDF1"
    # the documented layout: magic, u32 dims, f32 min, f32 cell, f32 samples, little-endian
    assert blob == b"MM3ESDF1" + struct.pack("<3I3f3f8f", 2, 2, 2, -1.0, -2.0, -3.0, 0.5, 0.75, 1.25, *data)
    lay2, d2 = hip.sdfv_decode(blob)
    assert lay2.as_tuple() == lay.as_tuple()
    assert np.array_equal(d2.reshape(-1), data)
    assert po.volume_sample(lay2.dims, lay2.min, lay2.cell, d2, (-0.75, -1.5, -2.5)) == \
        po.volume_sample(lay.dims, lay.min, lay.cell, data, (-0.75, -1.5, -2.5))
    with pytest.raises(hip.Mm3eError, match="sdfv buffer is shorter than the header"):
        hip.sdfv_decode(blob[:12])
    bad = b"X" + blob[1:]
    with pytest.raises(hip.Mm3eError, match="sdfv magic/version mismatch"):
        hip.sdfv_decode(bad)


def test_sdfv_rejections():
    lay, data = sdfv_reference_volume()
    blob = hip.sdfv_encode(lay, data)
    with pytest.raises(hip.Mm3eError, match="sdfv byte length 68 does not match expected 76"):
        hip.sdfv_decode(blob[:-8])
    nan_sample = blob[:44] + struct.pack("<f", float("nan")) + blob[48:]   # header: 44 bytes
    with pytest.raises(hip.Mm3eError, match="sdfv non-finite f32 while reading sample 0"):
        hip.sdfv_decode(nan_sample)
    inf_min = blob[:20] + struct.pack("<f", float("inf")) + blob[24:]
    with pytest.raises(hip.Mm3eError, match="sdfv non-finite f32 while reading min.x"):
        hip.sdfv_decode(inf_min)
    flat = b"MM3ESDF1" + struct.pack("<3I6f", 1, 2, 2, 0, 0, 0, 1, 1, 1) + b"\0" * 16
    with pytest.raises(hip.Mm3eError, match="SdfVolume dims must all be >= 2"):
        hip.sdfv_decode(flat)
    zero_cell = blob[:36] + struct.pack("<f", 0.0) + blob[40:]
    with pytest.raises(hip.Mm3eError, match="SdfVolume cell sizes must be positive and finite"):
        hip.sdfv_decode(zero_cell)
    with pytest.raises(hip.Mm3eError, match="SdfVolume sample 3 is non-finite"):
        hip.sdfv_encode(lay, np.where(np.arange(8) == 3, np.inf, data).astype(np.float32))
    with pytest.raises(hip.Mm3eError, match="SdfVolume data length 7 does not match dims 2x2x2"):
        hip.sdfv_encode(lay, data[:7])


# ---- the device bake ----------------------------------------------------------------------------

@pytest.mark.gpu
def test_gpu_bake_sphere_and_box_bit_exact(renderer):
    for (pos, tri), res, pad in ((configs.uv_sphere_mesh(12, 18, 1.0), 28, 0.25),
                                 (configs.box_mesh((1.0, 0.75, 0.5)), 25, 0.5),
                                 (configs.uv_sphere_mesh(18, 24, 0.85), 44, 0.3)):
        print(f"bake {len(tri)} triangles at {res}", flush=True)
        lay, got = renderer.bake_sdf(pos, tri, res, pad)
        dims, lo, cell, want = po.mesh_bake(pos, tri, res, pad)
        assert tuple(lay.dims) == dims
        assert np.array_equal(bits(got), bits(want))
    # the reference's tolerance test on the device volume
    pos, tri = configs.uv_sphere_mesh(12, 18, 1.0)
    lay, d = renderer.bake_sdf(pos, tri, 28, 0.25)
    for p in ((0.0, 0.0, 0.0), (1.2, 0.0, 0.0), (0.0, 1.2, 0.0)):
        assert abs(po.volume_sample(lay.dims, lay.min, lay.cell, d, p) - (math.sqrt(sum(c * c for c in p)) - 1.0)) < 0.16


@pytest.mark.gpu
def test_gpu_bake_knot_bit_exact_and_repeatable(renderer):
    pos, tri = configs.torus_knot_mesh()
    print(f"bake torus knot ({len(tri)} triangles) at 64", flush=True)
    b = hip.MeshBaker(renderer, pos, tri, 64, 0.35)
    b.run()
    first = b.read()
    b.run()
    ms = b.wait()
    second = b.read()
    b.close()
    dims, lo, cell, want = po.mesh_bake(pos, tri, 64, 0.35)
    assert np.array_equal(bits(first), bits(want))
    assert np.array_equal(bits(second), bits(want))
    assert ms[0] > 0.0 and ms[1] > 0.0


@pytest.mark.gpu
def test_gpu_bake_knot_full_size_slabs_and_properties(renderer):
    # 128 samples along the longest axis (the stress scene's volume size): slabs against the
    # oracle bit for bit, the whole grid through size-independent properties
    pos, tri = configs.torus_knot_mesh()
    print("bake torus knot at 128", flush=True)
    lay, d = renderer.bake_sdf(pos, tri, 128, 0.35)
    nx, ny, nz = lay.dims
    for k0 in (0, nz // 3, nz // 2, nz - 2):
        _, _, _, want = po.mesh_bake(pos, tri, 128, 0.35, k0=k0, k1=k0 + 2)
        assert np.array_equal(bits(d[k0:k0 + 2]), bits(want)), k0
    cell = np.asarray(lay.cell)
    assert np.abs(np.diff(d, axis=2)).max() <= cell[0] * 1.02
    assert np.abs(np.diff(d, axis=1)).max() <= cell[1] * 1.02
    assert np.abs(np.diff(d, axis=0)).max() <= cell[2] * 1.02
    # the inside samples fill the tube's volume: polygon area (36 sides of radius 0.22) times the
    # knot's arc length
    t = np.linspace(0.0, 2 * np.pi, 20001)
    r = 1.2 + 0.5 * np.cos(3 * t)
    curve = np.stack([r * np.cos(2 * t), 0.5 * np.sin(3 * t), r * np.sin(2 * t)], -1)
    length = np.linalg.norm(np.diff(curve, axis=0), axis=-1).sum()
    tube = 0.5 * 36 * 0.22 ** 2 * math.sin(2 * math.pi / 36) * length
    inside = float((d < 0).sum()) * float(np.prod(cell.astype(np.float64)))
    assert abs(inside / tube - 1.0) < 0.03, (inside, tube)


@pytest.mark.gpu
def test_gpu_bake_single_triangle_and_errors(renderer):
    # a lone triangle: the tree is one leaf; no closed surface, so every sample is unsigned-positive
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    lay, got = renderer.bake_sdf(pos, [[0, 1, 2]], 12, 0.5)
    _, _, _, want = po.mesh_bake(pos, [[0, 1, 2]], 12, 0.5)
    assert np.array_equal(bits(got), bits(want))
    with pytest.raises(hip.Mm3eError, match="padding must be positive and finite"):
        renderer.bake_sdf(pos, [[0, 1, 2]], 12, 0.0)


@pytest.mark.gpu
def test_gpu_baked_mesh_renders_as_a_volume(renderer):
    """The reference's mesh-volume parity scene (mm3e-gpu/examples/gpu_parity.rs:73-85): a UV
    sphere baked on the device at resolution 44 becomes a Prim::Volume over a checkered floor;
    the HIP render of that scene meets the image bars against the oracle's render."""
    from mm3e.scene import Light, Material, Object, Prim, Scene, Transform, Volume
    from test_gpu_parity import image_parity
    pos, tri = configs.uv_sphere_mesh(18, 24, 0.85)
    lay, d = renderer.bake_sdf(pos, tri, 44, 0.3)
    sc = Scene(240, 135)
    sc.aa = 1
    sc.bounces = 1
    sc.post.bloom = False
    floor = sc.material(Material.solid((0.9, 0.9, 0.9)).checkered().roughness(0.55))
    red = sc.material(Material.solid((0.85, 0.2, 0.2)).roughness(0.35))
    vid = sc.volume(Volume(tuple(lay.dims), tuple(lay.min), tuple(lay.cell), d.reshape(-1)))
    sc.add(Object(Prim.plane((0, 1, 0), 0.0), Transform(), floor))
    sc.add(Object(Prim.volume(vid), Transform.at((0.0, 0.9, 0.0)), red))
    sc.sun_dir = tuple(np.array([0.5, 0.75, 0.4]) / np.linalg.norm([0.5, 0.75, 0.4]))
    sc.light(Light.directional_light(sc.sun_dir, (1.35, 1.35, 1.35)).soft(0.04))
    cam = configs.orbit_camera((0.0, 0.75, 0.0), 5.5, 0.5, 0.28, np.radians(50.0))
    image_parity(renderer, sc, cam, "device-baked mesh volume")
