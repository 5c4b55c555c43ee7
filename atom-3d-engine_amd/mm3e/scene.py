"""Scene description mirroring the reference's `mm3e_orchestrator::Scene` builder API, serialised
to the flat binary IR declared in include/mm3e_scene_ir.h.

Reference interface mirrored here (file:line in the reference tree):
  Scene::new defaults ............ mm3e-orchestrator/src/lib.rs:595-620
  Scene::material/volume/add/light lib.rs:623-659 (volume() raises marcher.normal_h, 631-638)
  Scene::bake_gi ................. lib.rs:665-682 (carried as a bake request in the IR)
  Object + builder methods ....... lib.rs:330-377
  Prim kinds ..................... lib.rs:157-212
  Light::{directional,point,sphere,soft} lib.rs:512-527
  Material + builders ............ mm3e-kit/src/color.rs:37-89
  Marcher defaults ............... mm3e-kit/src/march.rs:81-96
  Post defaults .................. mm3e-orchestrator/src/post.rs:25-29
  Mat3::from_euler / axis-angle .. mm3e-kit/src/vec.rs:136-149

All arithmetic that the reference performs while *building* a scene (light-direction
normalisation, default sun direction, Euler rotation) is done in IEEE binary32 with numpy
scalars, one rounding per operation, so the IR holds the same bits a Rust `Scene` would.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

F = np.float32

IR_MAGIC = b"MM3EIR01"
IR_VERSION = 1
HEADER_WORDS = 76
OBJECT_WORDS = 36
MATERIAL_WORDS = 12
LIGHT_WORDS = 8
VOLUME_WORDS = 12
CSG_NODE_WORDS = 28

PRIM_SPHERE, PRIM_BOX, PRIM_ROUND_BOX, PRIM_TORUS, PRIM_CYLINDER, PRIM_CAPSULE = range(6)
PRIM_CONE, PRIM_ELLIPSOID, PRIM_OCTAHEDRON, PRIM_HEX_PRISM, PRIM_PLANE = range(6, 11)
PRIM_CSG, PRIM_VOLUME, PRIM_SURFACE = 11, 12, 13
ANALYTIC_PRIMS = tuple(range(11))
PRIM_NAMES = ("sphere", "box", "round_box", "torus", "cylinder", "capsule", "cone",
              "ellipsoid", "octahedron", "hex_prism", "plane", "csg", "volume", "surface")

COMBINE_UNION, COMBINE_SMOOTH, COMBINE_SUBTRACT = 0, 1, 2
# RenderMode (mm3e-orchestrator/src/lib.rs:49-61), declaration order
MODE_BEAUTY, MODE_NORMAL, MODE_DEPTH, MODE_AO, MODE_STEPS, MODE_ALBEDO = range(6)
GI_NONE, GI_BAKED, GI_BAKE = 0, 1, 2

CSG_PRIM, CSG_TRANSFORM, CSG_UNION, CSG_INTERSECT = 0, 1, 2, 3
CSG_SUBTRACT, CSG_SMOOTH_UNION, CSG_OFFSET, CSG_SHELL = 4, 5, 6, 7


# --------------------------------------------------------------------------------------------
# binary32 vector helpers (mm3e-kit/src/vec.rs)
# --------------------------------------------------------------------------------------------

def v3(x, y=None, z=None) -> Tuple[np.float32, np.float32, np.float32]:
    if y is None:
        x, y, z = x
    return (F(x), F(y), F(z))


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def vscale(a, s):
    s = F(s)
    return (a[0] * s, a[1] * s, a[2] * s)


def vadd(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def vsub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def normalize(a):
    """Vec3::normalize (vec.rs:37-44): scale by 1/len when len > 1e-12, else zero."""
    a = v3(a)
    ln = np.sqrt(dot(a, a))
    if ln > F(1e-12):
        return vscale(a, F(1.0) / ln)
    return (F(0.0), F(0.0), F(0.0))


IDENTITY = ((F(1), F(0), F(0)), (F(0), F(1), F(0)), (F(0), F(0), F(1)))


def mat_mul_vec(cols, v):
    """Mat3::mul_vec (vec.rs:125-127): c0*x + c1*y + c2*z."""
    return vadd(vadd(vscale(cols[0], v[0]), vscale(cols[1], v[1])), vscale(cols[2], v[2]))


def mat_mul(a, b):
    return tuple(mat_mul_vec(a, c) for c in b)


def from_axis_angle(axis, angle):
    """Mat3::from_axis_angle (vec.rs:136-141), Rodrigues' formula."""
    k = normalize(axis)
    s, c = F(np.sin(F(angle))), F(np.cos(F(angle)))

    def col(e):
        e = v3(e)
        return vadd(vadd(vscale(e, c), vscale(cross(k, e), s)), vscale(k, dot(k, e) * (F(1.0) - c)))

    return (col((1, 0, 0)), col((0, 1, 0)), col((0, 0, 1)))


def from_euler(x, y, z):
    """Mat3::from_euler (vec.rs:144-149): rx * ry * rz."""
    rx = from_axis_angle((1, 0, 0), x)
    ry = from_axis_angle((0, 1, 0), y)
    rz = from_axis_angle((0, 0, 1), z)
    return mat_mul(mat_mul(rx, ry), rz)


# --------------------------------------------------------------------------------------------
# Scene pieces
# --------------------------------------------------------------------------------------------

@dataclass
class Transform:
    """vec.rs:164-222 — rigid + uniform scale."""
    pos: Tuple = (0.0, 0.0, 0.0)
    rot: Tuple = IDENTITY
    scale: float = 1.0

    @staticmethod
    def at(pos):
        return Transform(pos=v3(pos))

    def rotated(self, rot):
        return Transform(self.pos, rot, self.scale)

    def scaled(self, s):
        return Transform(self.pos, self.rot, s)


@dataclass
class Prim:
    kind: int
    params: Tuple[float, ...] = ()

    @staticmethod
    def sphere(r):
        return Prim(PRIM_SPHERE, (r,))

    @staticmethod
    def box(half):
        return Prim(PRIM_BOX, tuple(half))

    @staticmethod
    def round_box(half, radius):
        return Prim(PRIM_ROUND_BOX, tuple(half) + (radius,))

    @staticmethod
    def torus(major, minor):
        return Prim(PRIM_TORUS, (major, minor))

    @staticmethod
    def cylinder(h, r):
        return Prim(PRIM_CYLINDER, (h, r))

    @staticmethod
    def capsule(a, b, r):
        return Prim(PRIM_CAPSULE, tuple(a) + tuple(b) + (r,))

    @staticmethod
    def cone(r1, r2, h):
        return Prim(PRIM_CONE, (r1, r2, h))

    @staticmethod
    def ellipsoid(r):
        return Prim(PRIM_ELLIPSOID, tuple(r))

    @staticmethod
    def octahedron(s):
        return Prim(PRIM_OCTAHEDRON, (s,))

    @staticmethod
    def hex_prism(r, h):
        return Prim(PRIM_HEX_PRISM, (r, h))

    @staticmethod
    def plane(n, h):
        return Prim(PRIM_PLANE, tuple(n) + (h,))

    @staticmethod
    def volume(vid):
        return Prim(PRIM_VOLUME, (vid,))

    @staticmethod
    def csg(eid):
        return Prim(PRIM_CSG, (eid,))

    @staticmethod
    def surface(sid):
        return Prim(PRIM_SURFACE, (sid,))


@dataclass
class Object:
    """lib.rs:330-377."""
    prim: Prim
    xform: Transform
    mat: int
    combine: int = COMBINE_UNION
    k: float = 0.0
    mirror_xyz: Tuple[bool, bool, bool] = (False, False, False)
    repeat_v: Tuple = (0.0, 0.0, 0.0)
    twist_k: float = 0.0
    bend_k: float = 0.0
    elongate_v: Tuple = (0.0, 0.0, 0.0)
    round_r: float = 0.0
    onion_t: float = 0.0

    def smooth(self, k):
        self.combine, self.k = COMBINE_SMOOTH, k
        return self

    def subtract(self):
        self.combine = COMBINE_SUBTRACT
        return self

    def mirror(self, x, y, z):
        self.mirror_xyz = (bool(x), bool(y), bool(z))
        return self

    def repeat(self, period):
        self.repeat_v = tuple(period)
        return self

    def twist(self, k):
        self.twist_k = k
        return self

    def bend(self, k):
        self.bend_k = k
        return self

    def elongate(self, h):
        self.elongate_v = tuple(h)
        return self

    def round(self, r):
        self.round_r = r
        return self

    def onion(self, t):
        self.onion_t = t
        return self


@dataclass
class Material:
    """color.rs:37-60 (Default: albedo 0.8, metallic 0, specular 0.5, roughness 0.35)."""
    albedo: Tuple = (0.8, 0.8, 0.8)
    metallic_v: float = 0.0
    specular_v: float = 0.5
    roughness_v: float = 0.35
    reflectivity: float = 0.0
    emissive_v: Tuple = (0.0, 0.0, 0.0)
    checker: bool = False

    @staticmethod
    def solid(albedo):
        return Material(albedo=tuple(albedo))

    def metallic(self, m):
        self.metallic_v = m
        return self

    def specular(self, s):
        self.specular_v = s
        return self

    def roughness(self, r):
        self.roughness_v = r
        return self

    def reflective(self, r):
        self.reflectivity = r
        return self

    def emissive(self, e):
        self.emissive_v = tuple(e)
        return self

    def checkered(self):
        self.checker = True
        return self


@dataclass
class Light:
    """lib.rs:503-527."""
    vec: Tuple
    color: Tuple
    directional: bool
    radius: float = 0.0

    @staticmethod
    def directional_light(toward, color):
        return Light(normalize(toward), v3(color), True, 0.0)

    @staticmethod
    def point(pos, color):
        return Light(v3(pos), v3(color), False, 0.0)

    @staticmethod
    def sphere(pos, color, radius):
        return Light(v3(pos), v3(color), False, radius)

    def soft(self, radius):
        self.radius = radius
        return self


@dataclass
class Marcher:
    """march.rs:81-96."""
    max_steps: int = 160
    max_dist: float = 120.0
    eps: float = 0.0006
    step_scale: float = 1.0
    shadow_steps: int = 64
    ao_samples: int = 5
    lod_footprint: float = 0.0
    subitize: float = 0.0
    secant: bool = False
    normal_h: float = 0.0009


@dataclass
class Quality:
    """Quality presets (mm3e-orchestrator/src/lib.rs:67-149): resolution and the marcher / AA /
    reflection budgets, written into a scene with `apply`."""
    width: int
    height: int
    aa: int
    bounces: int
    max_steps: int
    shadow_steps: int
    ao_samples: int
    lod_footprint: float

    @staticmethod
    def fast(width: int, height: int) -> "Quality":
        """Low resolution preview: no AA, no reflections, lean shadows, AO off, LOD on."""
        return Quality(width, height, 1, 0, 64, 12, 0, float(F(0.006)))

    @staticmethod
    def balanced(width: int, height: int) -> "Quality":
        return Quality(width, height, 1, 1, 110, 28, 3, float(F(0.002)))

    @staticmethod
    def full(width: int, height: int) -> "Quality":
        """Converged stills: LOD off (exact)."""
        return Quality(width, height, 2, 2, 160, 64, 5, 0.0)

    @staticmethod
    def lerp(a: "Quality", b: "Quality", t: float) -> "Quality":
        """Budgets from `a` toward `b` by t in [0, 1] in f32, rounded half away from zero;
        the resolution is b's (lib.rs:117-136)."""
        t = F(min(max(F(t), F(0.0)), F(1.0)))

        def mix(x, y):
            v = F(F(x) + F(F(F(y) - F(x)) * t))
            r = np.floor(np.abs(v) + F(0.5)) * np.sign(v)  # f32::round: half away from zero
            return max(int(r), 0)  # `as u32` saturates at 0

        def mixf(x, y):
            return float(F(F(x) + F(F(F(y) - F(x)) * t)))

        return Quality(b.width, b.height, mix(a.aa, b.aa), mix(a.bounces, b.bounces),
                       mix(a.max_steps, b.max_steps), mix(a.shadow_steps, b.shadow_steps),
                       mix(a.ao_samples, b.ao_samples), mixf(a.lod_footprint, b.lod_footprint))

    def apply(self, scene: "Scene") -> None:
        scene.width, scene.height = self.width, self.height
        scene.aa, scene.bounces = self.aa, self.bounces
        scene.marcher.max_steps = self.max_steps
        scene.marcher.shadow_steps = self.shadow_steps
        scene.marcher.ao_samples = self.ao_samples
        scene.marcher.lod_footprint = self.lod_footprint


@dataclass
class Post:
    """post.rs:25-29."""
    exposure: float = 1.0
    bloom: bool = True
    bloom_threshold: float = 1.0
    bloom_intensity: float = 0.06
    bloom_radius: int = 12


@dataclass
class Volume:
    """SdfVolume (mm3e-kit/src/volume.rs:14-23): data index (k*ny + j)*nx + i."""
    dims: Tuple[int, int, int]
    min: Tuple
    cell: Tuple
    data: np.ndarray

    def recommended_normal_h(self):
        c = [F(x) for x in self.cell]
        return F(max(max(c[0], c[1]), c[2]) * F(0.6))


@dataclass
class CsgNode:
    op: int
    prim: int = 0
    a: int = 0
    b: int = 0
    p: Tuple = ()
    xform: Optional[Transform] = None


class Scene:
    """Scene::new (lib.rs:595-620) with the same defaults."""

    def __init__(self, width: int, height: int):
        self.width, self.height = int(width), int(height)
        self.objects: List[Object] = []
        self.materials: List[Material] = [Material()]
        self.lights: List[Light] = []
        self.sun_dir = normalize((0.6, 0.7, 0.4))
        self.ambient = v3(0.02, 0.02, 0.02)
        self.sky_ambient = v3(0.4, 0.4, 0.4)
        self.fog = v3(0.62, 0.72, 0.86)
        self.fog_density = 0.012
        self.aa = 2
        self.bounces = 1
        self.shadows = True
        self.ao = True
        self.marcher = Marcher()
        self.post = Post()
        self.mode = 0
        self.appearance = 0
        self.volumes: List[Volume] = []
        self.csg_nodes: List[CsgNode] = []
        self.csg_roots: List[int] = []
        self.gi_mode = GI_NONE
        self.gi_dims = (0, 0, 0)
        self.gi_samples = 0
        self.gi_min = (0.0, 0.0, 0.0)
        self.gi_size = (0.0, 0.0, 0.0)
        self.gi_cubes: Optional[np.ndarray] = None

    # builder API ------------------------------------------------------------------------
    def material(self, m: Material) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    def volume(self, v: Volume) -> int:
        """lib.rs:631-638: registering a volume raises normal_h to 0.6 * max cell."""
        sug = v.recommended_normal_h()
        if np.isfinite(sug):
            self.marcher.normal_h = float(max(F(self.marcher.normal_h), sug))
        self.volumes.append(v)
        return len(self.volumes) - 1

    def csg(self, nodes: Sequence[CsgNode], root: int) -> int:
        base = len(self.csg_nodes)
        for n in nodes:
            nn = CsgNode(n.op, n.prim, n.a + base, n.b + base, n.p, n.xform)
            self.csg_nodes.append(nn)
        self.csg_roots.append(root + base)
        return len(self.csg_roots) - 1

    def add(self, obj: Object):
        self.objects.append(obj)

    def light(self, l: Light):
        self.lights.append(l)

    def bake_gi(self, dims, samples):
        """Request the reference's `Scene::bake_gi(dims, samples)`; executed by the consumer
        of the IR (the oracle on the CPU, the library on the device)."""
        self.gi_mode = GI_BAKE
        self.gi_dims = tuple(int(d) for d in dims)
        self.gi_samples = int(samples)

    def set_baked_gi(self, gmin, gsize, dims, cubes):
        self.gi_mode = GI_BAKED
        self.gi_min, self.gi_size, self.gi_dims = tuple(gmin), tuple(gsize), tuple(int(d) for d in dims)
        self.gi_samples = 0
        self.gi_cubes = np.ascontiguousarray(cubes, dtype=np.float32).reshape(-1)

    # serialisation ------------------------------------------------------------------------
    def to_ir(self) -> bytes:
        objs = bytearray()
        for o in self.objects:
            p = list(o.prim.params) + [0.0] * (8 - len(o.prim.params))
            pw = []
            for i, x in enumerate(p):
                if o.prim.kind in (PRIM_CSG, PRIM_VOLUME, PRIM_SURFACE) and i == 0:
                    pw.append(struct.pack("<I", int(x)))
                else:
                    pw.append(struct.pack("<f", F(x)))
            rot = [F(c) for col in o.xform.rot for c in col]
            mirror = (1 if o.mirror_xyz[0] else 0) | (2 if o.mirror_xyz[1] else 0) | (4 if o.mirror_xyz[2] else 0)
            objs += struct.pack("<IIIf", o.prim.kind, o.mat, o.combine, F(o.k))
            objs += b"".join(pw)
            objs += struct.pack("<9f", *rot)
            objs += struct.pack("<3ff", *v3(o.xform.pos), F(o.xform.scale))
            objs += struct.pack("<I3fff3fff", mirror, *v3(o.repeat_v), F(o.twist_k), F(o.bend_k),
                                *v3(o.elongate_v), F(o.round_r), F(o.onion_t))
        assert len(objs) == 4 * OBJECT_WORDS * len(self.objects)
        mats = bytearray()
        for m in self.materials:
            mats += struct.pack("<3fffff3fII", *v3(m.albedo), F(m.metallic_v), F(m.specular_v),
                                F(m.roughness_v), F(m.reflectivity), *v3(m.emissive_v), int(m.checker), 0)
        lights = bytearray()
        for l in self.lights:
            lights += struct.pack("<3ff3fI", *v3(l.vec), F(l.radius), *v3(l.color), int(l.directional))
        vols = bytearray()
        vdata = []
        off = 0
        for v in self.volumes:
            d = np.ascontiguousarray(v.data, dtype=np.float32).reshape(-1)
            vols += struct.pack("<4I3f3f2I", *[int(x) for x in v.dims], off, *v3(v.min), *v3(v.cell), 0, 0)
            vdata.append(d)
            off += d.size
        vdata_b = np.concatenate(vdata).tobytes() if vdata else b""
        csgn = bytearray()
        for n in self.csg_nodes:
            p = list(n.p) + [0.0] * (8 - len(n.p))
            xf = n.xform or Transform()
            rot = [F(c) for col in xf.rot for c in col]
            csgn += struct.pack("<4I8f9f3ff3I", n.op, n.prim, n.a, n.b, *[F(x) for x in p], *rot,
                                *v3(xf.pos), F(xf.scale), 0, 0, 0)
        roots = struct.pack("<%dI" % len(self.csg_roots), *self.csg_roots)
        cubes = b""
        if self.gi_mode == GI_BAKED:
            cubes = self.gi_cubes.tobytes()

        hdr_len = 4 * HEADER_WORDS
        off_objects = hdr_len
        off_materials = off_objects + len(objs)
        off_lights = off_materials + len(mats)
        off_volumes = off_lights + len(lights)
        off_vdata = off_volumes + len(vols)
        off_csg = off_vdata + len(vdata_b)
        off_roots = off_csg + len(csgn)
        off_cubes = off_roots + len(roots)
        total = off_cubes + len(cubes)

        mk = self.marcher
        po = self.post
        hdr = bytearray()
        hdr += IR_MAGIC
        hdr += struct.pack("<II", IR_VERSION, total)
        hdr += struct.pack("<8I", self.width, self.height, self.aa, self.bounces, int(self.shadows),
                           int(self.ao), self.mode, self.appearance)
        hdr += struct.pack("<3ff", *v3(self.sun_dir), F(self.fog_density))
        hdr += struct.pack("<3ff", *v3(self.ambient), 0.0)
        hdr += struct.pack("<3ff", *v3(self.sky_ambient), 0.0)
        hdr += struct.pack("<3ff", *v3(self.fog), 0.0)
        hdr += struct.pack("<IfffIIffIf2I", mk.max_steps, F(mk.max_dist), F(mk.eps), F(mk.step_scale),
                           mk.shadow_steps, mk.ao_samples, F(mk.lod_footprint), F(mk.subitize),
                           int(mk.secant), F(mk.normal_h), 0, 0)
        hdr += struct.pack("<fIffI3I", F(po.exposure), int(po.bloom), F(po.bloom_threshold),
                           F(po.bloom_intensity), po.bloom_radius, 0, 0, 0)
        hdr += struct.pack("<8I", len(self.objects), len(self.materials), len(self.lights), len(self.volumes),
                           len(self.csg_roots), len(self.csg_nodes), self.gi_mode, self.gi_samples)
        hdr += struct.pack("<3II", *[int(d) for d in self.gi_dims], off)
        hdr += struct.pack("<3ff", *v3(self.gi_min), 0.0)
        hdr += struct.pack("<3ff", *v3(self.gi_size), 0.0)
        hdr += struct.pack("<8I", off_objects, off_materials, off_lights, off_volumes, off_vdata, off_csg,
                           off_roots, off_cubes)
        assert len(hdr) == hdr_len, len(hdr)
        blob = bytes(hdr) + bytes(objs) + bytes(mats) + bytes(lights) + bytes(vols) + vdata_b + bytes(csgn) \
            + roots + cubes
        assert len(blob) == total
        return blob

    def is_dual_safe(self) -> bool:
        """lib.rs:426-430 + 811-813."""
        for o in self.objects:
            if o.twist_k != 0.0 or o.bend_k != 0.0 or o.prim.kind in (PRIM_VOLUME, PRIM_CSG, PRIM_SURFACE):
                return False
        return True
