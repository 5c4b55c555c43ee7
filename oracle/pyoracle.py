"""ctypes bindings for the CPU oracle (liboracle_mm3e.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module, as
the checker of the HIP render path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_mm3e.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        fp = C.POINTER(C.c_float)
        L.orc_scene_load.argtypes = [P, C.c_size_t, C.c_uint, C.POINTER(P), C.c_char_p, C.c_size_t]
        L.orc_scene_free.argtypes = [P]
        L.orc_gi_raw.argtypes = [P, P, P, P, P]
        L.orc_gi_raw.restype = C.c_uint64
        L.orc_render_linear.argtypes = [P, P, C.c_uint, P]
        L.orc_render_linear_rows.argtypes = [P, P, C.c_uint, C.c_uint32, C.c_uint32, P]
        L.orc_resolve.argtypes = [P, P, P]
        L.orc_render_rgba.argtypes = [P, P, C.c_uint, P]
        L.orc_render_aov.argtypes = [P, P, C.c_uint, P, P, P, P]
        L.orc_field.argtypes = [P, P, C.c_size_t, C.c_int, P, P]
        L.orc_march.argtypes = [P, P, C.c_size_t, P, P, P, P, P]
        L.orc_prim_distance.argtypes = [C.c_uint32, P, P]
        L.orc_prim_distance.restype = C.c_float
        L.orc_prim_dual.argtypes = [C.c_uint32, P, P, P]
        L.orc_normals.argtypes = [P, P, P, P]
        L.orc_dual_safe.argtypes = [P]
        L.orc_brdf.argtypes = [P, P, P, P, C.c_float, C.c_float, C.c_float, P]
        L.orc_fresnel.argtypes = [C.c_float, C.c_float]
        L.orc_fresnel.restype = C.c_float
        L.orc_aces_gamma.argtypes = [P, P, P]
        L.orc_sky.argtypes = [P, P, C.c_int, P]
        L.orc_hash_cell.argtypes = [C.c_int32, C.c_int32]
        L.orc_hash_cell.restype = C.c_float
        L.orc_fnv1a.argtypes = [P, C.c_size_t]
        L.orc_fnv1a.restype = C.c_uint64
        L.orc_powi.argtypes = [C.c_float, C.c_int]
        L.orc_powi.restype = C.c_float
        L.orc_soft_shadow.argtypes = [P, P, P, C.c_float, C.c_float]
        L.orc_soft_shadow.restype = C.c_float
        L.orc_ao.argtypes = [P, P, P]
        L.orc_ao.restype = C.c_float
        L.orc_gi_sample.argtypes = [P, P, P, P]
        L.orc_gi_bake_linear_x.argtypes = [P, P, P, C.c_uint32, P, P, P]
        L.orc_volume_sample.argtypes = [P, P, P, P, C.c_size_t, P]
        L.orc_volume_sample.restype = C.c_float
        L.orc_camera_look_at.argtypes = [P, P, P, C.c_float, P]
        L.orc_gaussian_kernel.argtypes = [C.c_uint32, P, P]
        L.orc_gaussian_kernel.restype = None
        L.orc_bloom.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint32, P]
        L.orc_bloom.restype = None
        L.orc_ibl_irradiance.argtypes = [P, P, P]
        L.orc_ibl_irradiance.restype = None
        L.orc_light_eval.argtypes = [C.c_int, P, C.c_float, P, P]
        L.orc_light_eval.restype = None
        L.orc_apply_fog.argtypes = [P, P, C.c_float, C.c_float, P]
        L.orc_apply_fog.restype = None
        L.orc_count_rows.argtypes = [P, P, C.c_uint, C.c_uint32, C.c_uint32, P]
        L.orc_phase_count.restype = C.c_int
        L.orc_to_bmp.argtypes = [P, C.c_uint32, C.c_uint32, P]
        L.orc_to_bmp.restype = C.c_size_t
        L.orc_render_film.argtypes = [P, P, C.c_float, C.c_float, C.c_uint, P, P, P, P, P, P]
        L.orc_render_gbuffer.argtypes = [P, P, P, C.c_uint32, C.c_uint, P, P, P]
        L.orc_reproject.argtypes = [C.c_uint32, C.c_uint32, P, P, P, P, P, P, C.c_uint32, C.c_int, P, P]
        L.orc_reproject.restype = None
        L.orc_reproject_hybrid.argtypes = [P, C.c_uint32, C.c_uint32, P, P, P, P, P, P, C.c_uint32, P]
        L.orc_reproject_hybrid.restype = C.c_uint32
        L.orc_ray_with_lens.argtypes = [P, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                        C.c_float, C.c_float, P, P]
        L.orc_ray_with_lens.restype = None
        L.orc_aperture_sample.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P]
        L.orc_aperture_sample.restype = None
        Z = C.c_size_t
        L.orc_mesh_parse_obj.argtypes = [C.c_char_p, Z, P, Z, P, Z, C.POINTER(Z), C.POINTER(Z), C.c_char_p, Z]
        L.orc_x_crossing.argtypes = [P, P, P, C.c_float, C.c_float, P]
        L.orc_closest_on_triangle.argtypes = [P, P, P, P, P]
        L.orc_closest_on_triangle.restype = None
        L.orc_mesh_nearest_sq.argtypes = [P, Z, P, Z, P, Z, C.c_int, P, P]
        L.orc_mesh_bake.argtypes = [P, Z, P, Z, Z, C.c_float, C.c_uint, Z, Z, P, P, P, P, Z, C.c_char_p, Z]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


PHASES = ("march", "refl_march", "normal", "shadow", "ao", "gi", "other")


class OracleScene:
    """A scene loaded into the oracle (bake requests are executed at load, on `threads` cores)."""

    def __init__(self, ir: bytes, threads: int = 0):
        L = lib()
        self._buf = C.create_string_buffer(ir, len(ir))
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        threads = threads or os.cpu_count() or 1
        rc = L.orc_scene_load(C.cast(self._buf, C.c_void_p), len(ir), threads, C.byref(h), err, 512)
        if rc != 0:
            raise ValueError(err.value.decode())
        self.h = h
        import struct
        self.width, self.height = struct.unpack_from("<II", ir, 16)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_scene_free(self.h)
        except Exception:
            pass

    def render_linear(self, cam, threads=0):
        out = np.zeros((self.height, self.width, 3), np.float32)
        lib().orc_render_linear(self.h, _p(f32(cam)), threads or os.cpu_count(), _p(out))
        return out

    def render_linear_rows(self, cam, y0, y1, threads=0):
        out = np.zeros((y1 - y0, self.width, 3), np.float32)
        lib().orc_render_linear_rows(self.h, _p(f32(cam)), threads or os.cpu_count(), y0, y1, _p(out))
        return out

    def resolve(self, hdr):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        lib().orc_resolve(self.h, _p(f32(hdr)), _p(out))
        return out

    def render_rgba(self, cam, threads=0):
        """lib.rs:866-879 `render`: beauty through the HDR buffer and post pass, or, in a debug
        AOV mode, one centre sample per pixel written straight to display."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        lib().orc_render_rgba(self.h, _p(f32(cam)), threads or os.cpu_count(), _p(out))
        return out

    def render_aov(self, cam, threads=0):
        n = self.width * self.height
        hit = np.zeros(n, np.uint8)
        mat = np.zeros(n, np.uint32)
        depth = np.zeros(n, np.float32)
        steps = np.zeros(n, np.uint32)
        lib().orc_render_aov(self.h, _p(f32(cam)), threads or os.cpu_count(), _p(hit), _p(mat), _p(depth), _p(steps))
        sh = (self.height, self.width)
        return hit.reshape(sh), mat.reshape(sh), depth.reshape(sh), steps.reshape(sh)

    def render_film(self, cam, aperture_radius=0.0, focus_distance=1.0, threads=0):
        """render_film_with_threads (film.rs:74-198) -> dict of beauty, foreground, alpha, depth,
        normals, material_ids as [h, w(, 3)] arrays."""
        h, w = self.height, self.width
        out = {"beauty": np.zeros((h, w, 3), np.float32), "foreground": np.zeros((h, w, 3), np.float32),
               "alpha": np.zeros((h, w), np.float32), "depth": np.zeros((h, w), np.float32),
               "normals": np.zeros((h, w, 3), np.float32), "material_ids": np.zeros((h, w), np.uint32)}
        lib().orc_render_film(self.h, _p(f32(cam)), float(aperture_radius), float(focus_distance),
                              threads or os.cpu_count(), *(_p(out[k]) for k in
                                                           ("beauty", "foreground", "alpha", "depth", "normals",
                                                            "material_ids")))
        return out

    def render_gbuffer(self, cam, movers=(), threads=0):
        """render_gbuffer (lib.rs:951-1030) -> dict(color[h,w,4] u8, depth[h,w], obj[h,w] i32, camera)."""
        h, w = self.height, self.width
        mv = f32(np.asarray(movers, np.float32).reshape(-1, 4)) if len(movers) else np.zeros((0, 4), np.float32)
        g = {"color": np.zeros((h, w, 4), np.uint8), "depth": np.zeros((h, w), np.float32),
             "obj": np.zeros((h, w), np.int32), "camera": f32(cam).copy()}
        lib().orc_render_gbuffer(self.h, _p(f32(cam)), _p(mv) if len(mv) else None, len(mv), threads or os.cpu_count(),
                                 _p(g["color"]), _p(g["depth"]), _p(g["obj"]))
        return g

    def reproject_hybrid(self, g, cam, deltas=()):
        """reproject_hybrid (lib.rs:1043-1080) of G-frame g -> (rgba[h,w,4], holes)."""
        h, w = g["depth"].shape
        d = f32(np.asarray(deltas, np.float32).reshape(-1, 3)) if len(deltas) else np.zeros((0, 3), np.float32)
        out = np.zeros((h, w, 4), np.uint8)
        holes = lib().orc_reproject_hybrid(self.h, w, h, _p(f32(g["camera"])), _p(np.ascontiguousarray(g["color"])),
                                           _p(f32(g["depth"])), _p(np.ascontiguousarray(g["obj"], np.int32)),
                                           _p(f32(cam)), _p(d) if len(d) else None, len(d), _p(out))
        return out, int(holes)

    def field(self, pts, linear=False):
        pts = f32(pts).reshape(-1, 3)
        d = np.zeros(len(pts), np.float32)
        m = np.zeros(len(pts), np.uint32)
        lib().orc_field(self.h, _p(pts), len(pts), 1 if linear else 0, _p(d), _p(m))
        return d, m

    def march(self, rays):
        rays = f32(rays).reshape(-1, 6)
        n = len(rays)
        t = np.zeros(n, np.float32)
        hit = np.zeros(n, np.uint8)
        mat = np.zeros(n, np.uint32)
        nor = np.zeros((n, 3), np.float32)
        st = np.zeros(n, np.uint32)
        lib().orc_march(self.h, _p(rays), n, _p(t), _p(hit), _p(mat), _p(nor), _p(st))
        return t, hit.astype(bool), mat, nor, st

    def gi_raw(self):
        L = lib()
        mn = np.zeros(3, np.float32)
        sz = np.zeros(3, np.float32)
        dims = np.zeros(3, np.uint32)
        n = L.orc_gi_raw(self.h, _p(mn), _p(sz), _p(dims), None)
        if n == 0:
            return None
        cubes = np.zeros((int(n), 6, 3), np.float32)
        L.orc_gi_raw(self.h, None, None, None, _p(cubes))
        return mn, sz, dims, cubes

    def normals(self, p):
        a = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        lib().orc_normals(self.h, _p(f32(p)), _p(a), _p(b))
        return a, b

    def dual_safe(self):
        return bool(lib().orc_dual_safe(self.h))

    def soft_shadow(self, o, d, max_t, k):
        return lib().orc_soft_shadow(self.h, _p(f32(o)), _p(f32(d)), max_t, k)

    def ao(self, p, n):
        return lib().orc_ao(self.h, _p(f32(p)), _p(f32(n)))

    def ibl_irradiance(self, n):
        out = np.zeros(3, np.float32)
        lib().orc_ibl_irradiance(self.h, _p(f32(n)), _p(out))
        return out

    def gi_sample(self, p, n):
        out = np.zeros(3, np.float32)
        lib().orc_gi_sample(self.h, _p(f32(p)), _p(f32(n)), _p(out))
        return out

    def count_rows(self, cam, y0, y1, threads=0):
        nph = lib().orc_phase_count()
        c = np.zeros(6 + nph, np.uint64)
        lib().orc_count_rows(self.h, _p(f32(cam)), threads or os.cpu_count(), y0, y1, _p(c))
        out = {"flops": int(c[0]), "trans": int(c[1]), "sqrt": int(c[2]), "div": int(c[3]),
               "obj_exact": int(c[4 + nph]), "obj_bound": int(c[5 + nph])}
        out["evals"] = {PHASES[i]: int(c[4 + i]) for i in range(nph)}
        return out


def prim_distance(kind, params, p):
    pr = np.zeros(8, np.float32)
    pr[: len(params)] = params
    return lib().orc_prim_distance(kind, _p(pr), _p(f32(p)))


def prim_dual(kind, params, p):
    pr = np.zeros(8, np.float32)
    pr[: len(params)] = params
    out = np.zeros(4, np.float32)
    lib().orc_prim_dual(kind, _p(pr), _p(f32(p)), _p(out))
    return out


def to_bmp(rgba):
    h, w = rgba.shape[:2]
    rgba = np.ascontiguousarray(rgba, np.uint8)
    n = lib().orc_to_bmp(_p(rgba), w, h, None)
    out = np.zeros(n, np.uint8)
    lib().orc_to_bmp(_p(rgba), w, h, _p(out))
    return out.tobytes()


# ---- mesh -> SDF bake (mesh_oracle.cpp) -----------------------------------------------------

class OracleError(RuntimeError):
    pass


def _mesh(pos, tri):
    return f32(np.asarray(pos).reshape(-1, 3)), np.ascontiguousarray(np.asarray(tri).reshape(-1, 3), np.uint32)


def mesh_parse_obj(text: str):
    b = text.encode()
    n, t = C.c_size_t(0), C.c_size_t(0)
    err = C.create_string_buffer(1024)
    if lib().orc_mesh_parse_obj(b, len(b), None, 0, None, 0, C.byref(n), C.byref(t), err, 1024) != 0:
        raise OracleError(err.value.decode())
    pos = np.zeros((n.value, 3), np.float32)
    tri = np.zeros((t.value, 3), np.uint32)
    lib().orc_mesh_parse_obj(b, len(b), _p(pos), n.value, _p(tri), t.value, C.byref(n), C.byref(t), err, 1024)
    return pos, tri


def x_crossing(a, b, c, y, z):
    """mm3e-kit triangle::x_crossing: the x of the +x line at (y, z) through the triangle, or None."""
    x = np.zeros(1, np.float32)
    hit = lib().orc_x_crossing(_p(f32(a)), _p(f32(b)), _p(f32(c)), float(y), float(z), _p(x))
    return float(x[0]) if hit else None


def closest_on_triangle(p, a, b, c):
    out = np.zeros(3, np.float32)
    lib().orc_closest_on_triangle(_p(f32(p)), _p(f32(a)), _p(f32(b)), _p(f32(c)), _p(out))
    return out


def mesh_nearest_sq(pos, tri, pts, brute=False):
    """Squared distances to the nearest non-degenerate triangle; with brute=False through the BVH,
    also returning (box tests, triangle tests)."""
    pos, tri = _mesh(pos, tri)
    pts = f32(np.asarray(pts).reshape(-1, 3))
    out = np.zeros(len(pts), np.float32)
    counts = np.zeros(2, np.uint64)
    if lib().orc_mesh_nearest_sq(_p(pos), len(pos), _p(tri), len(tri), _p(pts), len(pts), 1 if brute else 0,
                                 _p(out), _p(counts)) != 0:
        raise OracleError("invalid mesh")
    return out, (int(counts[0]), int(counts[1]))


def mesh_bake(pos, tri, resolution, padding, threads=8, k0=0, k1=0):
    """mesh::bake_sdf (z-slabs k0..k1, all when k1 == 0) -> (dims, lo, cell, samples[k1-k0, ny, nx])."""
    pos, tri = _mesh(pos, tri)
    dims = np.zeros(3, np.uint32)
    lo = np.zeros(3, np.float32)
    cell = np.zeros(3, np.float32)
    err = C.create_string_buffer(1024)
    # layout first (one slab), then the requested slabs
    if lib().orc_mesh_bake(_p(pos), len(pos), _p(tri), len(tri), int(resolution), float(padding), 1, 0, 1,
                           _p(dims), _p(lo), _p(cell), None, 0, err, 1024) != 0:
        raise OracleError(err.value.decode())
    nx, ny, nz = (int(v) for v in dims)
    kk1 = nz if k1 == 0 else min(int(k1), nz)
    out = np.zeros((max(kk1 - int(k0), 0), ny, nx), np.float32)
    if lib().orc_mesh_bake(_p(pos), len(pos), _p(tri), len(tri), int(resolution), float(padding), int(threads),
                           int(k0), kk1, _p(dims), _p(lo), _p(cell), _p(out), out.size, err, 1024) != 0:
        raise OracleError(err.value.decode())
    return tuple(int(v) for v in dims), lo, cell, out


def volume_sample(dims, mn, cell, data, p):
    """SdfVolume::sample (mm3e-kit/src/volume.rs:194-228) at local point p."""
    d = np.asarray(dims, np.uint32)
    a = f32(mn)
    c = f32(cell)
    v = f32(np.asarray(data).reshape(-1))
    return lib().orc_volume_sample(_p(d), _p(a), _p(c), _p(v), v.size, _p(f32(p)))


# ---- thin lens (camera.rs:71-118, film.rs:200-224) ---------------------------------------------

def ray_with_lens(cam, image_sample, image_size, aperture_radius, focus_distance, aperture_sample):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_ray_with_lens(_p(f32(cam)), float(image_sample[0]), float(image_sample[1]), int(image_size[0]),
                            int(image_size[1]), float(aperture_radius), float(focus_distance),
                            float(aperture_sample[0]), float(aperture_sample[1]), _p(o), _p(d))
    return o, d


def aperture_sample(pixel, sample, n):
    out = np.zeros(2, np.float32)
    lib().orc_aperture_sample(int(pixel[0]), int(pixel[1]), int(sample[0]), int(sample[1]), int(n), _p(out))
    return out


def reproject(g, cam, deltas=(), fill=True):
    """reproject::reproject (fill) or reproject::warp (no fill) of G-frame g -> (rgba, filled)."""
    h, w = g["depth"].shape
    d = f32(np.asarray(deltas, np.float32).reshape(-1, 3)) if len(deltas) else np.zeros((0, 3), np.float32)
    out = np.zeros((h, w, 4), np.uint8)
    filled = np.zeros((h, w), np.uint8)
    lib().orc_reproject(w, h, _p(f32(g["camera"])), _p(np.ascontiguousarray(g["color"])), _p(f32(g["depth"])),
                        _p(np.ascontiguousarray(g["obj"], np.int32)), _p(f32(cam)), _p(d) if len(d) else None,
                        len(d), 1 if fill else 0, _p(out), _p(filled))
    return out, filled


def gaussian_kernel(radius):
    n = C.c_uint32(0)
    lib().orc_gaussian_kernel(radius, None, C.byref(n))
    out = np.zeros(n.value, np.float32)
    lib().orc_gaussian_kernel(radius, _p(out), C.byref(n))
    return out


def bloom(hdr, threshold, intensity, radius):
    """post::resolve's HDR after the bloom add (before exposure / tone mapping)."""
    hdr = f32(hdr)
    h, w = hdr.shape[:2]
    out = np.zeros_like(hdr)
    lib().orc_bloom(_p(hdr), w, h, threshold, intensity, radius, _p(out))
    return out


def light_eval(directional, vec, radius, p):
    """(direction, distance, attenuation, shadow k) of Light::toward/attenuation/shadow_k."""
    out = np.zeros(6, np.float32)
    lib().orc_light_eval(1 if directional else 0, _p(f32(vec)), radius, _p(f32(p)), _p(out))
    return out[:3], out[3], out[4], out[5]


def apply_fog(c, fog, t, density):
    out = np.zeros(3, np.float32)
    lib().orc_apply_fog(_p(f32(c)), _p(f32(fog)), t, density, _p(out))
    return out
