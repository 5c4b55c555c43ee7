"""Seeded synthetic workloads for the per-pixel SDF render path.

Five workloads, each a deterministic function of its seed (SplitMix64 draws, fixed order):

  c1  320x180 CPU-scale frame: ground plane + 8 spheres, GGX materials, one soft directional
      light (k = 8), 5-tap AO, 128 march steps, eps 1e-3, far 100, camera (0,2,-8) -> origin,
      60 deg.  aa = 1, bounces = 1, bloom off (the reference's own 320x180 parity probe,
      mm3e-gpu/examples/gpu_parity.rs:14-20).
  c2  960x540 Cornell-style room (mm3e-orchestrator/examples/gi_demo.rs:14-58): five box
      walls (red/blue sides), three primitives drawn from the 11 analytic kinds (one
      smooth-union k = 0.3, one twist), an emissive ceiling bar, a sphere light, 16^3 GI
      probes baked from the field with 64 requested rays per probe (30 traced: the bake
      takes 1 or 5 rays per cube face, gi.rs:38-50), soft shadows, AO, 1 reflection bounce,
      bloom + tonemap, aa = 2 (Scene::new default).
  c3  procedural scenes of 20..200 primitives (uniform over the 11 kinds, 30 % smooth
      unions, 20 % one domain op among round/onion/repeat/twist/bend/mirror) seen by seeded
      orbit cameras at resolutions drawn from {320x180, 640x360, 1280x720, 1920x1080}.
  c4  request stream: 32 resident procedural scenes with 16^3 GI each, Poisson arrivals,
      resolutions {854x480: .5, 1280x720: .35, 1920x1080: .15}.
  c5  3840x2160 stress frame: 1000 primitives, 2 reflection bounces, exponential fog,
      twist/bend present, one 128^3 sampled SDF volume of a seeded star-shaped surface,
      32^3 GI probes.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .scene import (F, Light, Material, Object, Prim, Scene, Transform, Volume, from_euler, normalize, v3,
                    PRIM_PLANE)


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); `uniform` maps the top 24 bits to [0, 1)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def unit(self) -> float:
        return (self.next_u64() >> 40) / float(1 << 24)

    def uniform(self, a: float, b: float) -> float:
        return float(F(a + (b - a) * self.unit()))

    def randint(self, lo: int, hi_inclusive: int) -> int:
        return lo + (self.next_u64() % (hi_inclusive - lo + 1))

    def bernoulli(self, p: float) -> bool:
        return self.unit() < p


# --------------------------------------------------------------------------------------------
# cameras (mm3e-kit/src/camera.rs:50-55, mm3e-orchestrator/src/lib.rs:1426-1435)
# --------------------------------------------------------------------------------------------

def look_at(eye, target, up, fov_y) -> np.ndarray:
    """Camera::look_at in binary32; returns the 13-float camera record
    (eye, forward, right, up, fov_scale) that the C ABI takes."""
    from .scene import cross, vsub
    e = v3(eye)
    fwd = normalize(vsub(v3(target), e))
    right = normalize(cross(fwd, v3(up)))
    u = cross(right, fwd)
    fov_scale = F(math.tan(float(F(F(fov_y) * F(0.5)))))
    return np.array(list(e) + list(fwd) + list(right) + list(u) + [fov_scale], dtype=np.float32)


def orbit_camera(target, radius, yaw, pitch, fov_y) -> np.ndarray:
    """orbit_camera (lib.rs:1426-1435), evaluated in binary32 on the host."""
    pi = F(math.pi)
    tau = F(2 * math.pi)
    r = F(np.fmod(F(F(yaw) + pi), tau))  # rem_euclid(TAU)
    if r < 0:
        r = F(r + tau)
    yaw = F(r - pi)
    lim = F(F(math.pi / 2) - F(1e-4))
    pitch = F(min(max(F(pitch), -lim), lim))
    radius = F(max(abs(F(radius)), F(1e-4)))
    fov_y = F(min(max(F(fov_y), F(1e-4)), F(pi - F(1e-4))))
    cy, sy = F(math.cos(float(yaw))), F(math.sin(float(yaw)))
    cp, sp = F(math.cos(float(pitch))), F(math.sin(float(pitch)))
    t = v3(target)
    eye = (t[0] + radius * cy * cp, t[1] + radius * sp, t[2] + radius * sy * cp)
    return look_at(eye, t, (0.0, 1.0, 0.0), fov_y)


# --------------------------------------------------------------------------------------------
# c1
# --------------------------------------------------------------------------------------------

def c1_scene(seed: int = 1, width: int = 320, height: int = 180) -> Tuple[Scene, np.ndarray]:
    rng = SplitMix64(seed)
    sc = Scene(width, height)
    sc.aa = 1
    sc.bounces = 1
    sc.post.bloom = False
    sc.marcher.max_steps = 128
    sc.marcher.eps = 1e-3
    sc.marcher.max_dist = 100.0
    sc.marcher.ao_samples = 5
    floor = sc.material(Material.solid((0.9, 0.9, 0.9)).checkered().roughness(0.6))
    sc.add(Object(Prim.plane((0.0, 1.0, 0.0), 0.0), Transform(), floor))
    for _ in range(8):
        x, y, z = rng.uniform(-4, 4), rng.uniform(0.3, 2.0), rng.uniform(-4, 4)
        r = rng.uniform(0.3, 1.0)
        alb = (rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1))
        rough = rng.uniform(0.1, 0.9)
        metal = 1.0 if rng.bernoulli(0.3) else 0.0
        m = sc.material(Material.solid(alb).roughness(rough).metallic(metal))
        sc.add(Object(Prim.sphere(r), Transform.at((x, y, z)), m))
    sc.light(Light.directional_light(sc.sun_dir, (1.5, 1.5, 1.5)).soft(0.125))  # k = 1/0.125 = 8
    cam = look_at((0.0, 2.0, -8.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(60.0))
    return sc, cam


# --------------------------------------------------------------------------------------------
# procedural primitive (shared by c2's three primitives and the c3/c4/c5 generators)
# --------------------------------------------------------------------------------------------

def random_prim(rng: SplitMix64, kind: int, s: float) -> Prim:
    """Parameters for one of the 11 analytic kinds, sized by `s` (about the bounding radius)."""
    if kind == 0:
        return Prim.sphere(s)
    if kind == 1:
        return Prim.box((s * rng.uniform(0.5, 1.0), s * rng.uniform(0.5, 1.0), s * rng.uniform(0.5, 1.0)))
    if kind == 2:
        h = (s * rng.uniform(0.5, 1.0), s * rng.uniform(0.5, 1.0), s * rng.uniform(0.5, 1.0))
        return Prim.round_box(h, min(h) * rng.uniform(0.1, 0.4))
    if kind == 3:
        major = s * rng.uniform(0.6, 0.8)
        return Prim.torus(major, major * rng.uniform(0.2, 0.45))
    if kind == 4:
        return Prim.cylinder(s * rng.uniform(0.5, 1.0), s * rng.uniform(0.3, 0.7))
    if kind == 5:
        a = (0.0, -s * rng.uniform(0.3, 0.7), 0.0)
        b = (0.0, s * rng.uniform(0.3, 0.7), 0.0)
        return Prim.capsule(a, b, s * rng.uniform(0.2, 0.4))
    if kind == 6:
        return Prim.cone(s * rng.uniform(0.3, 0.6), s * rng.uniform(0.05, 0.3), s * rng.uniform(0.8, 1.4))
    if kind == 7:
        return Prim.ellipsoid((s * rng.uniform(0.5, 1.0), s * rng.uniform(0.5, 1.0), s * rng.uniform(0.5, 1.0)))
    if kind == 8:
        return Prim.octahedron(s * rng.uniform(0.7, 1.0))
    if kind == 9:
        return Prim.hex_prism(s * rng.uniform(0.4, 0.7), s * rng.uniform(0.3, 0.7))
    # plane: an extra floor just below the ground (unbounded; never pruned)
    return Prim.plane((0.0, 1.0, 0.0), rng.uniform(0.0, 0.5))


# --------------------------------------------------------------------------------------------
# c2
# --------------------------------------------------------------------------------------------

def c2_scene(seed: int = 2, width: int = 960, height: int = 540, gi_dims=(16, 16, 16), gi_rays=64) \
        -> Tuple[Scene, np.ndarray]:
    rng = SplitMix64(seed)
    sc = Scene(width, height)
    sc.bounces = 1
    sc.fog_density = 0.0
    sc.sky_ambient = v3(0.18, 0.18, 0.18)
    sc.ambient = v3(0.0, 0.0, 0.0)
    white = sc.material(Material.solid((0.85, 0.85, 0.85)).roughness(0.7))
    red = sc.material(Material.solid((0.85, 0.12, 0.12)).roughness(0.7))
    blue = sc.material(Material.solid((0.12, 0.2, 0.85)).roughness(0.7))
    chrome = sc.material(Material.solid((0.9, 0.9, 0.9)).metallic(1.0).roughness(0.12).reflective(0.7))
    accent = sc.material(Material.solid((rng.uniform(0.2, 0.9), rng.uniform(0.2, 0.9), rng.uniform(0.2, 0.9)))
                         .roughness(rng.uniform(0.2, 0.6)))

    def wall(c, half, m):
        return Object(Prim.box(half), Transform.at(c), m)

    sc.add(wall((0.0, -0.1, 0.0), (4.0, 0.1, 4.0), white))
    sc.add(wall((0.0, 4.1, 0.0), (4.0, 0.1, 4.0), white))
    sc.add(wall((0.0, 2.0, -4.1), (4.0, 4.0, 0.1), white))
    sc.add(wall((-4.1, 2.0, 0.0), (0.1, 4.0, 4.0), red))
    sc.add(wall((4.1, 2.0, 0.0), (0.1, 4.0, 4.0), blue))
    spots = [(-1.4, 1.0, 0.3), (1.5, 1.0, -0.7), (0.1, 0.9, 1.3)]
    mats = [white, chrome, accent]
    for i in range(3):
        kind = rng.randint(0, 10)
        s = rng.uniform(0.7, 1.0)
        prim = random_prim(rng, kind, s)
        rot = from_euler(rng.uniform(-0.6, 0.6), rng.uniform(-math.pi, math.pi), rng.uniform(-0.6, 0.6))
        if kind == PRIM_PLANE:  # an infinite plane stays horizontal, just under the floor top
            o = Object(prim, Transform(), mats[i])
        else:
            o = Object(prim, Transform(pos=v3(spots[i]), rot=rot), mats[i])
        if i == 1:
            o.smooth(0.3)
        if i == 2:
            o.twist(rng.uniform(0.6, 1.2))
        sc.add(o)
    lamp = sc.material(Material.solid((1.0, 1.0, 1.0)).emissive((3.0, 3.0, 3.0)))
    sc.add(Object(Prim.box((1.4, 0.06, 1.4)), Transform.at((0.0, 3.9, 0.0)), lamp))
    sc.sun_dir = normalize((0.0, 1.0, 0.2))
    sc.light(Light.sphere((0.0, 3.7, 0.0), (14.0, 14.0, 14.0), 1.2))
    # the twist makes the field non-Lipschitz: under-relax as the reference's examples do
    sc.marcher.step_scale = 0.6
    if gi_dims is not None:
        sc.bake_gi(gi_dims, gi_rays)
    cam = look_at((0.0, 2.0, 7.5), (0.0, 1.6, 0.0), (0.0, 1.0, 0.0), math.radians(50.0))
    return sc, cam


# --------------------------------------------------------------------------------------------
# c3 / c4 / c5 procedural scenes
# --------------------------------------------------------------------------------------------

DOMAIN_OPS = ("round", "onion", "repeat", "twist", "bend", "mirror")


def procedural_scene(seed: int, n_prims: int, width: int, height: int, extent: float = 6.0) -> Scene:
    rng = SplitMix64(seed)
    sc = Scene(width, height)
    sc.aa = 1
    sc.bounces = 1
    mats = []
    floor = sc.material(Material.solid((0.85, 0.85, 0.85)).checkered().roughness(0.6))
    for _ in range(8):
        m = Material.solid((rng.uniform(0.05, 0.95), rng.uniform(0.05, 0.95), rng.uniform(0.05, 0.95)))
        m.roughness(rng.uniform(0.1, 0.9)).metallic(1.0 if rng.bernoulli(0.3) else 0.0)
        if rng.bernoulli(0.2):
            m.reflective(rng.uniform(0.2, 0.6))
        mats.append(sc.material(m))
    sc.add(Object(Prim.plane((0.0, 1.0, 0.0), 0.0), Transform(), floor))
    warped = False
    for _ in range(n_prims):
        kind = rng.randint(0, 10)
        s = rng.uniform(0.25, 0.8)
        prim = random_prim(rng, kind, s)
        pos = (rng.uniform(-extent, extent), rng.uniform(0.2, 2.5), rng.uniform(-extent, extent))
        rot = from_euler(rng.uniform(-math.pi, math.pi), rng.uniform(-math.pi, math.pi), rng.uniform(-math.pi, math.pi))
        if kind == PRIM_PLANE:
            o = Object(prim, Transform(), mats[rng.randint(0, 7)])
        else:
            o = Object(prim, Transform(pos=v3(pos), rot=rot), mats[rng.randint(0, 7)])
        if rng.bernoulli(0.3):
            o.smooth(rng.uniform(0.1, 0.4))
        if rng.bernoulli(0.2):
            op = DOMAIN_OPS[rng.randint(0, 5)]
            if op == "round":
                o.round(rng.uniform(0.02, 0.1))
            elif op == "onion":
                o.onion(rng.uniform(0.02, 0.06))
            elif op == "repeat":
                o.repeat((rng.uniform(2.5, 4.0), 0.0, rng.uniform(2.5, 4.0)))
                warped = True
            elif op == "twist":
                o.twist(rng.uniform(0.3, 1.2))
                warped = True
            elif op == "bend":
                o.bend(rng.uniform(0.2, 0.8))
                warped = True
            else:
                o.mirror(True, False, True)
        sc.add(o)
    sc.sun_dir = normalize((rng.uniform(-0.6, 0.6), 0.8, rng.uniform(-0.6, 0.6)))
    sc.light(Light.directional_light(sc.sun_dir, (1.4, 1.4, 1.4)).soft(0.05))
    sc.light(Light.sphere((rng.uniform(-3, 3), 4.0, rng.uniform(-3, 3)), (6.0, 5.5, 5.0), 0.5))
    if warped:
        sc.marcher.step_scale = 0.6
    return sc


C3_RESOLUTIONS = ((320, 180), (640, 360), (1280, 720), (1920, 1080))
C4_RESOLUTIONS = ((854, 480, 0.50), (1280, 720, 0.35), (1920, 1080, 0.15))


def orbit_pose(rng: SplitMix64) -> np.ndarray:
    return orbit_camera((0.0, 1.0, 0.0), rng.uniform(8.0, 12.0), rng.uniform(-math.pi, math.pi),
                        rng.uniform(0.15, 0.5), math.radians(50.0))


def c3_jobs(seed: int = 3, n_scenes: int = 8, cams_per_scene: int = 8, min_prims=20, max_prims=200):
    """(scene, camera, (w, h)) for 64 cameras over 8 scenes."""
    jobs = []
    for si in range(n_scenes):
        rng = SplitMix64(seed + si)
        n = rng.randint(min_prims, max_prims)
        sc = procedural_scene(seed + si, n, 1920, 1080)
        for _ in range(cams_per_scene):
            w, h = C3_RESOLUTIONS[rng.randint(0, 3)]
            jobs.append((sc, orbit_pose(rng), (w, h)))
    return jobs


def c4_scenes(seed: int = 11, n_scenes: int = 32, min_prims=20, max_prims=200, gi_dims=(16, 16, 16)) -> List[Scene]:
    out = []
    for si in range(n_scenes):
        rng = SplitMix64(seed * 1000 + si)
        sc = procedural_scene(seed * 1000 + si, rng.randint(min_prims, max_prims), 1920, 1080)
        if gi_dims is not None:
            sc.bake_gi(gi_dims, 5)
        out.append(sc)
    return out


def c4_request(rng: SplitMix64, n_scenes: int = 32):
    si = rng.randint(0, n_scenes - 1)
    u = rng.unit()
    acc = 0.0
    w, h = C4_RESOLUTIONS[-1][:2]
    for rw, rh, p in C4_RESOLUTIONS:
        acc += p
        if u < acc:
            w, h = rw, rh
            break
    return si, orbit_pose(rng), (w, h)


def star_volume(seed: int, n: int = 128, r0: float = 0.8, half: float = 1.3) -> Volume:
    """Signed distance samples of a seeded star-shaped closed surface r(th, ph) = r0 (1 + sum a_i Y_i),
    on an n^3 grid over [-half, half]^3.  Distances are radial (|p| - r(dir)) scaled by a
    Lipschitz factor so the field stays a conservative bound."""
    rng = SplitMix64(seed)
    lin = np.linspace(-half, half, n, dtype=np.float64)
    z, y, x = np.meshgrid(lin, lin, lin, indexing="ij")
    rr = np.sqrt(x * x + y * y + z * z) + 1e-9
    th = np.arccos(np.clip(y / rr, -1, 1))
    ph = np.arctan2(z, x)
    shape = np.ones_like(rr)
    amp_total = 0.0
    for _ in range(5):
        l = rng.randint(2, 5)
        m = rng.randint(0, l)
        a = rng.uniform(0.03, 0.08)
        amp_total += a * (l + 1)
        shape += a * np.cos(m * ph) * np.cos(l * th)
    rad = r0 * shape
    d = (rr - rad) / (1.0 + r0 * amp_total)
    cell = 2 * half / (n - 1)
    return Volume((n, n, n), (-half, -half, -half), (cell, cell, cell), d.astype(np.float32))


def real_sh(l: int, m: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Real spherical harmonic Y_l^m (orthonormal; theta polar, phi azimuth) in float64."""
    from scipy.special import lpmv
    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    base = norm * lpmv(am, l, np.cos(theta))
    if m == 0:
        return base
    return math.sqrt(2.0) * base * (np.cos(am * phi) if m > 0 else np.sin(am * phi))


def star_mesh(seed: int, r0: float = 0.8, n_lat: int = 64, n_lon: int = 128):
    """A seeded star-shaped closed triangle mesh r(theta, phi) = r0 (1 + sum_i a_i Y_i(theta, phi))
    over five real spherical harmonics with 2 <= l <= 4 (a_i in [0.05, 0.12], so r stays
    positive): a latitude-longitude grid of n_lat bands and n_lon sectors with one vertex at each
    pole.  Returns (positions[n][3] f32, triangles[t][3] u32)."""
    rng = SplitMix64(seed)
    terms = []
    for _ in range(5):
        l = rng.randint(2, 4)
        m = rng.randint(-l, l)
        terms.append((l, m, rng.uniform(0.05, 0.12)))
    th = np.concatenate([[0.0], np.arange(1, n_lat) * (math.pi / n_lat), [math.pi]])
    ph = np.arange(n_lon) * (2 * math.pi / n_lon)
    T, P = np.meshgrid(th[1:-1], ph, indexing="ij")
    pts_t = np.concatenate([[0.0], T.reshape(-1), [math.pi]])
    pts_p = np.concatenate([[0.0], P.reshape(-1), [0.0]])
    shape = np.ones_like(pts_t)
    for l, m, a in terms:
        shape += a * real_sh(l, m, pts_t, pts_p)
    rad = r0 * shape
    pos = np.stack([rad * np.sin(pts_t) * np.cos(pts_p), rad * np.cos(pts_t), rad * np.sin(pts_t) * np.sin(pts_p)],
                   -1).astype(np.float32)
    ring = lambda i, j: 1 + (i - 1) * n_lon + (j % n_lon)  # noqa: E731  vertex of band ring i (1..n_lat-1)
    south = 1 + (n_lat - 1) * n_lon
    tri = []
    for j in range(n_lon):
        tri.append((0, ring(1, j + 1), ring(1, j)))
        for i in range(1, n_lat - 1):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j), ring(i + 1, j + 1)
            tri.append((a, b, d))
            tri.append((a, d, c))
        tri.append((ring(n_lat - 1, j), ring(n_lat - 1, j + 1), south))
    return pos, np.asarray(tri, np.uint32)


def c5_scene(seed: int = 12, width: int = 3840, height: int = 2160, n_prims: int = 1000, gi_dims=(32, 32, 32),
             vol_n: int = 128, bake=None) -> Tuple[Scene, np.ndarray]:
    """The stress workload.  Its 128^3 grid is the signed distance of a seeded star-shaped closed
    surface: star_mesh baked to an SDF grid with vol_n samples along the longest axis (padding
    0.15, mesh::bake_sdf semantics) by `bake(positions, triangles, resolution, padding)`, which
    returns (dims, min, cell, samples[nz][ny][nx]) — the library's device bake, or any bake with
    the same result.  Without `bake` the grid is star_volume's analytic stand-in of the same
    shape family."""
    sc = procedural_scene(seed, n_prims, width, height, extent=14.0)
    rng = SplitMix64(seed ^ 0x5A5A)
    sc.bounces = 2
    sc.fog_density = 0.02
    # guarantee the warps the workload names
    sc.objects[1].twist(0.8)
    sc.objects[2].bend(0.5)
    sc.marcher.step_scale = 0.6
    if bake is not None:
        mpos, mtri = star_mesh(seed)
        dims, lo, cell, data = bake(mpos, mtri, vol_n, 0.15)
        vol = Volume(tuple(int(d) for d in dims), tuple(float(x) for x in lo), tuple(float(x) for x in cell),
                     np.ascontiguousarray(data, np.float32).reshape(-1))
    else:
        vol = star_volume(seed, vol_n)
    vid = sc.volume(vol)
    vm = sc.material(Material.solid((0.8, 0.3, 0.2)).roughness(0.35))
    sc.add(Object(Prim.volume(vid), Transform.at((0.0, 1.4, 0.0)).scaled(1.5), vm))
    if gi_dims is not None:
        sc.bake_gi(gi_dims, 5)
    cam = orbit_camera((0.0, 1.0, 0.0), rng.uniform(16.0, 20.0), rng.uniform(-math.pi, math.pi), 0.35,
                       math.radians(55.0))
    return sc, cam


# --------------------------------------------------------------------------------------------
# object-local CSG (Prim::Csg, mm3e-kit/src/csg.rs): every node kind, nested transforms
# --------------------------------------------------------------------------------------------

def csg_scene(width: int = 320, height: int = 180, seed: int = 7) -> Tuple[Scene, np.ndarray]:
    """A floor plus four CSG objects covering every expression node: a drilled, smooth-blended
    sphere; a shelled, plane-clipped cap (csg.rs:264-274); a three-level transform chain with
    offset; a balanced union tree (8 leaves) that needs four operand registers."""
    from .scene import CsgNode, CSG_PRIM, CSG_TRANSFORM, CSG_UNION, CSG_INTERSECT, CSG_SUBTRACT
    from .scene import CSG_SMOOTH_UNION, CSG_OFFSET, CSG_SHELL
    from .scene import PRIM_SPHERE, PRIM_BOX, PRIM_TORUS, PRIM_PLANE, PRIM_CAPSULE, PRIM_CYLINDER
    from .scene import PRIM_HEX_PRISM, PRIM_CONE, PRIM_ELLIPSOID, PRIM_OCTAHEDRON, PRIM_ROUND_BOX
    rng = SplitMix64(seed)
    sc = Scene(width, height)
    sc.aa = 1
    floor = sc.material(Material.solid((0.8, 0.8, 0.8)).checkered().roughness(0.7))
    mats = [sc.material(Material.solid((rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9)))
                        .roughness(rng.uniform(0.2, 0.8)).metallic(1.0 if i == 2 else 0.0)) for i in range(4)]
    sc.add(Object(Prim.plane((0.0, 1.0, 0.0), 0.0), Transform(), floor))
    T = lambda pos, rot=(0.0, 0.0, 0.0), s=1.0: Transform(pos=v3(pos), rot=from_euler(*rot), scale=s)
    # 1: (sphere - rotated box) smooth-union torus
    e1 = sc.csg([CsgNode(CSG_PRIM, PRIM_SPHERE, p=(1.0,)),
                 CsgNode(CSG_PRIM, PRIM_BOX, p=(0.55, 0.55, 0.55)),
                 CsgNode(CSG_TRANSFORM, a=1, xform=T((0.6, 0.4, 0.5), (0.3, 0.5, 0.1))),
                 CsgNode(CSG_SUBTRACT, a=0, b=2),
                 CsgNode(CSG_PRIM, PRIM_TORUS, p=(0.7, 0.12)),
                 CsgNode(CSG_TRANSFORM, a=4, xform=T((0.0, -0.6, 0.0), (0.2, 0.0, 0.0), 1.2)),
                 CsgNode(CSG_SMOOTH_UNION, a=3, b=5, p=(0.25,))], root=6)
    sc.add(Object(Prim.csg(e1), T((-2.2, 1.1, 0.0), (0.0, 0.4, 0.0)), mats[0]))
    # 2: shell(sphere) clipped by a plane (an open cap), then offset
    e2 = sc.csg([CsgNode(CSG_PRIM, PRIM_SPHERE, p=(1.0,)),
                 CsgNode(CSG_SHELL, a=0, p=(0.05,)),
                 CsgNode(CSG_PRIM, PRIM_PLANE, p=(0.0, -1.0, 0.0, 0.3)),
                 CsgNode(CSG_INTERSECT, a=1, b=2),
                 CsgNode(CSG_OFFSET, a=3, p=(0.02,))], root=4)
    sc.add(Object(Prim.csg(e2), T((0.0, 1.2, 1.5), (0.4, 0.0, 0.2), 0.9), mats[1]))
    # 3: three nested transforms around a capsule, unioned with a cylinder at the outer level
    e3 = sc.csg([CsgNode(CSG_PRIM, PRIM_CAPSULE, p=(0.0, -0.4, 0.0, 0.0, 0.4, 0.0, 0.2)),
                 CsgNode(CSG_TRANSFORM, a=0, xform=T((0.3, 0.0, 0.0), (0.0, 0.0, 0.7), 0.9)),
                 CsgNode(CSG_TRANSFORM, a=1, xform=T((0.0, 0.2, 0.1), (0.5, 0.0, 0.0), 1.1)),
                 CsgNode(CSG_TRANSFORM, a=2, xform=T((-0.1, 0.0, 0.3), (0.0, 0.9, 0.0), 0.8)),
                 CsgNode(CSG_PRIM, PRIM_CYLINDER, p=(0.6, 0.25)),
                 CsgNode(CSG_UNION, a=3, b=4),
                 CsgNode(CSG_OFFSET, a=5, p=(-0.03,))], root=6)
    sc.add(Object(Prim.csg(e3), T((2.2, 0.9, 0.2), (0.1, 0.2, 0.3)), mats[2]))
    # 4: balanced tree of 8 leaves (all analytic kinds), right child deeper at the root
    leaves = [CsgNode(CSG_PRIM, PRIM_SPHERE, p=(0.3,)), CsgNode(CSG_PRIM, PRIM_BOX, p=(0.25, 0.2, 0.3)),
              CsgNode(CSG_PRIM, PRIM_HEX_PRISM, p=(0.25, 0.2)), CsgNode(CSG_PRIM, PRIM_CONE, p=(0.25, 0.1, 0.5)),
              CsgNode(CSG_PRIM, PRIM_ELLIPSOID, p=(0.3, 0.2, 0.25)), CsgNode(CSG_PRIM, PRIM_OCTAHEDRON, p=(0.3,)),
              CsgNode(CSG_PRIM, PRIM_ROUND_BOX, p=(0.2, 0.2, 0.2, 0.05)), CsgNode(CSG_PRIM, PRIM_TORUS, p=(0.25, 0.08))]
    nodes = list(leaves)
    placed = []
    for i in range(8):
        nodes.append(CsgNode(CSG_TRANSFORM, a=i, xform=T(((i % 4) * 0.45 - 0.7, (i // 4) * 0.5, 0.0))))
        placed.append(len(nodes) - 1)
    lvl = placed
    ops = [CSG_UNION, CSG_SMOOTH_UNION, CSG_UNION, CSG_INTERSECT]
    k = 0
    while len(lvl) > 1:
        nxt = []
        for j in range(0, len(lvl), 2):
            op = ops[k % len(ops)] if len(lvl) > 2 else CSG_UNION
            k += 1
            nodes.append(CsgNode(op, a=lvl[j], b=lvl[j + 1], p=(0.1,)))
            nxt.append(len(nodes) - 1)
        lvl = nxt
    nodes.append(CsgNode(CSG_PRIM, PRIM_SPHERE, p=(0.2,)))
    small = len(nodes) - 1
    nodes.append(CsgNode(CSG_UNION, a=small, b=lvl[0]))   # shallow left, deep right: swapped order
    e4 = sc.csg(nodes, root=len(nodes) - 1)
    sc.add(Object(Prim.csg(e4), T((0.0, 0.6, -1.8), (0.0, 0.3, 0.0), 1.3), mats[3]))
    sc.light(Light.directional_light(normalize((0.5, 0.8, 0.3)), (1.3, 1.3, 1.3)).soft(0.05))
    sc.light(Light.point((2.0, 3.5, 2.0), (4.0, 3.6, 3.2)))
    cam = look_at((0.0, 3.0, -6.5), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), math.radians(55.0))
    return sc, cam


# ---- triangle meshes (inputs of the mesh -> SDF bake) ------------------------------------------

def uv_sphere_mesh(rings: int, segments: int, radius: float) -> Tuple[np.ndarray, np.ndarray]:
    """Closed UV sphere as in the reference's mesh tests (mm3e-orchestrator/tests/engine.rs:370-401):
    a pole vertex, rings-1 rings of `segments` vertices, the other pole; outward winding."""
    pos = [(0.0, radius, 0.0)]
    for r in range(1, rings):
        phi = np.float32(np.pi) * np.float32(r) / np.float32(rings)
        for s in range(segments):
            th = np.float32(2 * np.pi) * np.float32(s) / np.float32(segments)
            pos.append((radius * np.sin(phi) * np.cos(th), radius * np.cos(phi), radius * np.sin(phi) * np.sin(th)))
    pos.append((0.0, -radius, 0.0))

    def ring(r, s):
        return 1 + (r - 1) * segments + (s % segments)

    tri = [(0, ring(1, s + 1), ring(1, s)) for s in range(segments)]
    for r in range(1, rings - 1):
        for s in range(segments):
            a, b, c, d = ring(r, s), ring(r, s + 1), ring(r + 1, s + 1), ring(r + 1, s)
            tri += [(a, b, c), (a, c, d)]
    south = len(pos) - 1
    tri += [(south, ring(rings - 1, s), ring(rings - 1, s + 1)) for s in range(segments)]
    return np.asarray(pos, np.float32), np.asarray(tri, np.uint32)


def box_mesh(half=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0)) -> Tuple[np.ndarray, np.ndarray]:
    """Axis-aligned box as 12 outward-facing triangles (two per face, split on a diagonal)."""
    hx, hy, hz = half
    cx, cy, cz = center
    pos = np.array([(cx + sx * hx, cy + sy * hy, cz + sz * hz)
                    for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)], np.float32)
    # vertex v = x + 2y + 4z (bits of the + corners)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tri = []
    for a, b, c, d in quads:
        tri += [(a, b, c), (a, c, d)]
    return pos, np.asarray(tri, np.uint32)


def torus_knot_mesh(p: float = 2.0, q: float = 3.0, segments: int = 240, ring_verts: int = 36,
                    tube_r: float = 0.22) -> Tuple[np.ndarray, np.ndarray]:
    """Closed (p, q) torus-knot tube, the reference's mesh demo input
    (mm3e-orchestrator/examples/mesh_demo.rs:12-40): curve r = 1.2 + 0.5 cos(q t), a
    parallel-transported frame, `ring_verts` vertices per ring, two triangles per quad."""
    def curve(t):
        r = 1.2 + 0.5 * np.cos(q * t)
        return np.array([r * np.cos(p * t), 0.5 * np.sin(q * t), r * np.sin(p * t)])

    def norm(v):
        n = np.linalg.norm(v)
        return v / n if n > 1e-12 else v * 0.0

    pos = []
    normal = np.array([0.0, 1.0, 0.0])
    for s in range(segments):
        t = 2 * np.pi * s / segments
        tangent = norm(curve(t + 1e-3) - curve(t - 1e-3))
        normal = norm(normal - tangent * normal.dot(tangent))
        binormal = np.cross(tangent, normal)
        center = curve(t)
        for v in range(ring_verts):
            a = 2 * np.pi * v / ring_verts
            pos.append(center + normal * (tube_r * np.cos(a)) + binormal * (tube_r * np.sin(a)))

    def idx(s, v):
        return (s % segments) * ring_verts + (v % ring_verts)

    tri = []
    for s in range(segments):
        for v in range(ring_verts):
            a, b, c, d = idx(s, v), idx(s + 1, v), idx(s + 1, v + 1), idx(s, v + 1)
            tri += [(a, b, c), (a, c, d)]
    return np.asarray(pos, np.float32), np.asarray(tri, np.uint32)


def mesh_to_obj(pos: np.ndarray, tri: np.ndarray, header: str = "MM3E mesh") -> str:
    """Wavefront OBJ text (`v` lines with shortest round-trip floats, 1-based `f` lines)."""
    out = [f"# {header}"]
    out += ["v {} {} {}".format(*(repr(float(np.float32(c))) for c in p)) for p in pos]
    out += ["f {} {} {}".format(*(int(i) + 1 for i in t)) for t in tri]
    return "\n".join(out) + "\n"
