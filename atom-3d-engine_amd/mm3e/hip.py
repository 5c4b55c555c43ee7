"""ctypes bindings over the C ABI of libmm3e_hip.so (include/mm3e_hip.h).

This is the Python-side equivalent of the FFI stub a reference-side caller would write: it
mirrors `mm3e_gpu::GpuRenderer` / `GpuScene` (mm3e-gpu/src/lib.rs:57-416) method for method.
The library is the HIP/gfx950 render path; there is no CPU fallback, and constructing a
`Renderer` without a usable MI355X raises.
"""
from __future__ import annotations

import ctypes as C
import struct
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "csrc", "libmm3e_hip.so")
MAX_DYN = 24

EXPORTS = (
    "mm3e_renderer_create", "mm3e_renderer_adapter_name", "mm3e_renderer_wait_idle", "mm3e_renderer_destroy",
    "mm3e_list_adapters", "mm3e_scene_compile", "mm3e_scene_destroy", "mm3e_scene_render_rgba",
    "mm3e_scene_render_only", "mm3e_scene_render_linear", "mm3e_scene_render_aov", "mm3e_scene_gi_raw",
    "mm3e_scene_last_timing", "mm3e_scene_timing", "mm3e_scene_field", "mm3e_scene_march", "mm3e_camera_look_at",
    "mm3e_orbit_camera", "mm3e_write_bmp", "mm3e_scene_check", "mm3e_batch_create", "mm3e_batch_autotune",
    "mm3e_batch_set_camera", "mm3e_batch_launch", "mm3e_batch_wait", "mm3e_batch_read", "mm3e_batch_info",
    "mm3e_batch_destroy", "mm3e_render_batch", "mm3e_server_default_config", "mm3e_server_create",
    "mm3e_server_add_scene", "mm3e_server_submit", "mm3e_server_poll", "mm3e_server_wait", "mm3e_server_stats",
    "mm3e_server_destroy", "mm3e_scene_kernel_name",
    "mm3e_scene_info", "mm3e_renderer_set_codegen", "mm3e_renderer_set_tile_order", "mm3e_renderer_set_wavefront", "mm3e_renderer_set_codegen_film", "mm3e_scene_set_wavefront", "mm3e_scene_autotune", "mm3e_selftest_sqrt", "mm3e_scene_codegen_status", "mm3e_scene_emit_source",
    "mm3e_scene_parse_text", "mm3e_scene_to_text",
    "mm3e_mesh_parse_obj", "mm3e_mesh_sdf_layout", "mm3e_mesh_bake_sdf", "mm3e_mesh_baker_create",
    "mm3e_mesh_baker_layout", "mm3e_mesh_baker_info", "mm3e_mesh_baker_run", "mm3e_mesh_baker_wait",
    "mm3e_mesh_baker_read", "mm3e_mesh_baker_destroy", "mm3e_sdfv_encode", "mm3e_sdfv_decode",
    "mm3e_scene_render_film", "mm3e_gframe_create", "mm3e_gframe_destroy", "mm3e_gframe_upload",
    "mm3e_gframe_read", "mm3e_scene_render_gbuffer", "mm3e_gframe_reproject", "mm3e_scene_reproject_hybrid",
    "mm3e_scene_render_rgba_noalloc", "mm3e_server_submit_text",
)

CODEGEN_OFF, CODEGEN_BACKGROUND, CODEGEN_SYNC = 0, 1, 2
CODEGEN_NONE, CODEGEN_QUEUED, CODEGEN_READY, CODEGEN_FAILED = 0, 1, 2, 3

_lib = None


class Mm3eError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmm3e_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.mm3e_renderer_create.argtypes = [C.c_char_p, C.POINTER(P), C.c_char_p, C.c_size_t]
        L.mm3e_scene_render_rgba_noalloc.argtypes = [P, P, P, C.c_size_t, C.c_char_p, C.c_size_t]
        L.mm3e_server_submit_text.argtypes = [P, C.c_char_p, C.c_size_t, C.c_uint32, C.c_uint32, P,
                                              C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]
        L.mm3e_renderer_adapter_name.argtypes = [P]
        L.mm3e_renderer_adapter_name.restype = C.c_char_p
        L.mm3e_renderer_wait_idle.argtypes = [P]
        L.mm3e_renderer_destroy.argtypes = [P]
        L.mm3e_list_adapters.argtypes = [C.c_char_p, C.c_size_t]
        L.mm3e_scene_compile.argtypes = [P, P, C.c_size_t, C.c_uint32, C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
        L.mm3e_scene_destroy.argtypes = [P]
        L.mm3e_scene_render_rgba.argtypes = [P, P, P, C.c_uint32, P, C.c_size_t]
        L.mm3e_scene_render_only.argtypes = [P, P]
        L.mm3e_scene_render_linear.argtypes = [P, P, P, C.c_size_t]
        L.mm3e_scene_render_aov.argtypes = [P, P, P, P, P, P]
        L.mm3e_scene_gi_raw.argtypes = [P, P, P, P, P, P]
        L.mm3e_scene_last_timing.argtypes = [P, P, P]
        L.mm3e_scene_timing.argtypes = [P, C.c_int, P, P, P]
        L.mm3e_scene_kernel_name.argtypes = [P, C.c_char_p, C.c_size_t]
        L.mm3e_scene_info.argtypes = [P, P, P, P, P, P]
        L.mm3e_renderer_set_codegen.argtypes = [P, C.c_int]
        L.mm3e_renderer_set_tile_order.argtypes = [P, C.c_int]
        L.mm3e_renderer_set_wavefront.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
        L.mm3e_renderer_set_codegen_film.argtypes = [P, C.c_int]
        L.mm3e_scene_set_wavefront.argtypes = [P, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
        L.mm3e_scene_autotune.argtypes = [P, P, C.c_uint32, C.c_char_p, C.c_size_t]
        L.mm3e_selftest_sqrt.argtypes = [P, C.POINTER(C.c_uint64)]
        L.mm3e_scene_codegen_status.argtypes = [P, C.c_int, P, C.c_char_p, C.c_size_t]
        L.mm3e_scene_emit_source.argtypes = [P, C.c_size_t, C.c_char_p, C.POINTER(C.c_size_t)]
        L.mm3e_scene_parse_text.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.POINTER(C.c_size_t), P,
                                            C.c_char_p, C.c_size_t]
        L.mm3e_scene_to_text.argtypes = [P, C.c_size_t, P, C.c_char_p, C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]
        L.mm3e_scene_field.argtypes = [P, P, C.c_uint32, P, P]
        L.mm3e_scene_march.argtypes = [P, P, C.c_uint32, P, P, P, P, P]
        L.mm3e_camera_look_at.argtypes = [P, P, P, C.c_float, P]
        L.mm3e_camera_look_at.restype = None
        L.mm3e_orbit_camera.argtypes = [P, C.c_float, C.c_float, C.c_float, C.c_float, P]
        L.mm3e_orbit_camera.restype = None
        L.mm3e_write_bmp.argtypes = [P, C.c_uint32, C.c_uint32, P, C.POINTER(C.c_size_t)]
        L.mm3e_scene_check.argtypes = [P, C.c_size_t, C.c_uint32, C.c_uint32, C.c_char_p, C.c_size_t]
        L.mm3e_batch_create.argtypes = [P, P, C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
        L.mm3e_batch_autotune.argtypes = [P, C.c_uint32, C.c_char_p, C.c_size_t]
        L.mm3e_batch_set_camera.argtypes = [P, C.c_uint32, P]
        L.mm3e_batch_launch.argtypes = [P]
        L.mm3e_batch_wait.argtypes = [P, P]
        L.mm3e_batch_read.argtypes = [P, C.c_uint32, P, C.c_size_t]
        L.mm3e_batch_info.argtypes = [P, P, P, P, P, P, P]
        L.mm3e_batch_destroy.argtypes = [P]
        L.mm3e_render_batch.argtypes = [P, P, C.c_uint32, C.c_char_p, C.c_size_t]
        L.mm3e_server_default_config.argtypes = [P]
        L.mm3e_server_default_config.restype = None
        L.mm3e_server_create.argtypes = [P, P, C.POINTER(P), C.c_char_p, C.c_size_t]
        L.mm3e_server_add_scene.argtypes = [P, P, C.POINTER(C.c_uint32)]
        L.mm3e_server_submit.argtypes = [P, C.c_uint32, P, C.c_uint32, C.c_uint32, P, C.POINTER(C.c_uint64)]
        L.mm3e_server_poll.argtypes = [P, C.c_uint64, C.POINTER(C.c_int)]
        L.mm3e_server_wait.argtypes = [P, C.c_uint64, C.POINTER(C.c_double)]
        L.mm3e_server_stats.argtypes = [P, P, P, P, P, P]
        L.mm3e_server_destroy.argtypes = [P]
        L.mm3e_server_destroy.restype = None
        Z = C.c_size_t
        L.mm3e_mesh_parse_obj.argtypes = [C.c_char_p, Z, P, Z, P, Z, C.POINTER(Z), C.POINTER(Z), C.c_char_p, Z]
        L.mm3e_mesh_sdf_layout.argtypes = [P, Z, P, Z, C.c_uint32, C.c_float, P, C.c_char_p, Z]
        L.mm3e_mesh_bake_sdf.argtypes = [P, P, Z, P, Z, C.c_uint32, C.c_float, P, Z, P, C.c_char_p, Z]
        L.mm3e_mesh_baker_create.argtypes = [P, P, Z, P, Z, C.c_uint32, C.c_float, C.POINTER(P), C.c_char_p, Z]
        L.mm3e_mesh_baker_layout.argtypes = [P, P]
        L.mm3e_mesh_baker_info.argtypes = [P, P, P, P]
        L.mm3e_mesh_baker_run.argtypes = [P]
        L.mm3e_mesh_baker_wait.argtypes = [P, P, P]
        L.mm3e_mesh_baker_read.argtypes = [P, P, Z, C.c_char_p, Z]
        L.mm3e_mesh_baker_destroy.argtypes = [P]
        L.mm3e_mesh_baker_destroy.restype = None
        L.mm3e_sdfv_encode.argtypes = [P, P, Z, P, C.POINTER(Z), C.c_char_p, Z]
        L.mm3e_sdfv_decode.argtypes = [P, Z, P, P, Z, C.c_char_p, Z]
        L.mm3e_scene_render_film.argtypes = [P, P, P, P, P, P, P, P, P, C.c_char_p, Z]
        L.mm3e_gframe_create.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(P), C.c_char_p, Z]
        L.mm3e_gframe_destroy.argtypes = [P]
        L.mm3e_gframe_destroy.restype = None
        L.mm3e_gframe_upload.argtypes = [P, P, P, P, P]
        L.mm3e_gframe_read.argtypes = [P, P, P, P, P]
        L.mm3e_scene_render_gbuffer.argtypes = [P, P, P, C.c_uint32, P, C.c_char_p, Z]
        L.mm3e_gframe_reproject.argtypes = [P, P, P, C.c_uint32, C.c_int, P, P, C.c_char_p, Z]
        L.mm3e_scene_reproject_hybrid.argtypes = [P, P, P, P, C.c_uint32, P, C.POINTER(C.c_uint32), C.c_char_p, Z]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _cam(cam) -> np.ndarray:
    c = np.ascontiguousarray(cam, dtype=np.float32).reshape(-1)
    if c.size != 13:
        raise ValueError("camera record must hold 13 floats (eye, forward, right, up, fov_scale)")
    return c


def camera_look_at(eye, target, up, fov_y) -> np.ndarray:
    out = np.zeros(13, np.float32)
    lib().mm3e_camera_look_at(_p(np.asarray(eye, np.float32)), _p(np.asarray(target, np.float32)),
                              _p(np.asarray(up, np.float32)), float(fov_y), _p(out))
    return out


def orbit_camera(target, radius, yaw, pitch, fov_y) -> np.ndarray:
    out = np.zeros(13, np.float32)
    lib().mm3e_orbit_camera(_p(np.asarray(target, np.float32)), radius, yaw, pitch, fov_y, _p(out))
    return out


def write_bmp(rgba: np.ndarray) -> bytes:
    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w = rgba.shape[:2]
    n = C.c_size_t(0)
    lib().mm3e_write_bmp(_p(rgba), w, h, None, C.byref(n))
    out = np.zeros(n.value, np.uint8)
    rc = lib().mm3e_write_bmp(_p(rgba), w, h, _p(out), C.byref(n))
    if rc != 0:
        raise Mm3eError(rc, "mm3e_write_bmp failed")
    return out.tobytes()


def scene_check(ir: bytes, width: int, height: int):
    """Device-free half of compile_checked: None if the scene compiles, else the message."""
    err = C.create_string_buffer(1024)
    rc = lib().mm3e_scene_check(ir, len(ir), width, height, err, 1024)
    return None if rc == 0 else err.value.decode()


def emit_source(ir: bytes) -> str:
    """wgsl::build_shader_checked twin: the generated HIP source of a scene IR."""
    n = C.c_size_t(0)
    lib().mm3e_scene_emit_source(ir, len(ir), None, C.byref(n))
    buf = C.create_string_buffer(n.value)
    rc = lib().mm3e_scene_emit_source(ir, len(ir), buf, C.byref(n))
    if rc != 0:
        raise Mm3eError(rc, "scene IR rejected")
    return buf.value.decode()


def parse_text(text: str):
    """scene_io::parse: `.mm3e` text -> (scene IR bytes, camera[13])."""
    b = text.encode()
    n = C.c_size_t(0)
    err = C.create_string_buffer(1024)
    cam = np.zeros(13, np.float32)
    rc = lib().mm3e_scene_parse_text(b, len(b), None, C.byref(n), None, err, 1024)
    if n.value == 0:
        raise Mm3eError(rc, err.value.decode())
    buf = C.create_string_buffer(n.value)
    rc = lib().mm3e_scene_parse_text(b, len(b), buf, C.byref(n), _p(cam), err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    return buf.raw[:n.value], cam


def to_text(ir: bytes, cam) -> str:
    """scene_io::try_serialize: scene IR + camera -> `.mm3e` text."""
    n = C.c_size_t(0)
    err = C.create_string_buffer(1024)
    c = _cam(cam)
    lib().mm3e_scene_to_text(ir, len(ir), _p(c), None, C.byref(n), err, 1024)
    if n.value == 0:
        raise Mm3eError(1, err.value.decode())
    buf = C.create_string_buffer(n.value)
    rc = lib().mm3e_scene_to_text(ir, len(ir), _p(c), buf, C.byref(n), err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    return buf.value.decode()


class SdfLayout(C.Structure):
    """mm3e_sdf_layout: SdfVolume's dims / min / cell (mm3e-kit/src/volume.rs:14-23)."""
    _fields_ = [("dims", C.c_uint32 * 3), ("min", C.c_float * 3), ("cell", C.c_float * 3)]

    def as_tuple(self):
        return tuple(self.dims), tuple(self.min), tuple(self.cell)


def _mesh_arrays(positions, triangles):
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    return pos, tri


def parse_obj(text: str):
    """mesh::parse_obj: OBJ text -> (positions[n,3] f32, triangles[t,3] u32)."""
    b = text.encode()
    n, t = C.c_size_t(0), C.c_size_t(0)
    err = C.create_string_buffer(1024)
    rc = lib().mm3e_mesh_parse_obj(b, len(b), None, 0, None, 0, C.byref(n), C.byref(t), err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    pos = np.zeros((n.value, 3), np.float32)
    tri = np.zeros((t.value, 3), np.uint32)
    rc = lib().mm3e_mesh_parse_obj(b, len(b), _p(pos), n.value, _p(tri), t.value, C.byref(n), C.byref(t), err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    return pos, tri


def sdf_layout(positions, triangles, resolution: int, padding: float) -> SdfLayout:
    """The grid mesh::bake_sdf would produce (host only)."""
    pos, tri = _mesh_arrays(positions, triangles)
    out = SdfLayout()
    err = C.create_string_buffer(1024)
    rc = lib().mm3e_mesh_sdf_layout(_p(pos), len(pos), _p(tri), len(tri), int(resolution), float(padding),
                                    C.byref(out), err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    return out


def sdfv_encode(layout: SdfLayout, data) -> bytes:
    """SdfVolume::to_sdfv_bytes."""
    d = np.ascontiguousarray(data, np.float32).reshape(-1)
    n = C.c_size_t(0)
    err = C.create_string_buffer(1024)
    lib().mm3e_sdfv_encode(C.byref(layout), _p(d), d.size, None, C.byref(n), err, 1024)
    if n.value == 0:
        raise Mm3eError(1, err.value.decode())
    out = np.zeros(n.value, np.uint8)
    rc = lib().mm3e_sdfv_encode(C.byref(layout), _p(d), d.size, _p(out), C.byref(n), err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    return out.tobytes()


def sdfv_decode(blob: bytes):
    """SdfVolume::from_sdfv_bytes -> (layout, samples[nz, ny, nx])."""
    b = np.frombuffer(blob, np.uint8)
    lay = SdfLayout()
    err = C.create_string_buffer(1024)
    rc = lib().mm3e_sdfv_decode(_p(b) if b.size else None, b.size, C.byref(lay), None, 0, err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    nx, ny, nz = lay.dims
    data = np.zeros((nz, ny, nx), np.float32)
    rc = lib().mm3e_sdfv_decode(_p(b), b.size, C.byref(lay), _p(data), data.size, err, 1024)
    if rc != 0:
        raise Mm3eError(rc, err.value.decode())
    return lay, data


def _deltas(deltas):
    d = np.ascontiguousarray(np.asarray(deltas, np.float32).reshape(-1, 3)) if len(deltas) else None
    return d, (0 if d is None else len(d))


class GFrame:
    """A device-resident G-buffer frame (reproject::GFrame, reproject.rs:24-36)."""

    def __init__(self, renderer, width: int, height: int):
        self.h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_gframe_create(renderer.h, int(width), int(height), C.byref(self.h), err, 1024)
        if rc != 0:
            self.h = None
            raise Mm3eError(rc, err.value.decode())
        self.width, self.height = int(width), int(height)

    def upload(self, g: dict):
        """Load a host G-frame dict (color, depth, obj, camera)."""
        rc = lib().mm3e_gframe_upload(self.h, _p(_cam(g["camera"])), _p(np.ascontiguousarray(g["color"], np.uint8)),
                                      _p(np.ascontiguousarray(g["depth"], np.float32)),
                                      _p(np.ascontiguousarray(g["obj"], np.int32)))
        if rc != 0:
            raise Mm3eError(rc, "gframe upload failed")

    def read(self) -> dict:
        sh = (self.height, self.width)
        g = {"color": np.zeros(sh + (4,), np.uint8), "depth": np.zeros(sh, np.float32),
             "obj": np.zeros(sh, np.int32), "camera": np.zeros(13, np.float32)}
        rc = lib().mm3e_gframe_read(self.h, _p(g["camera"]), _p(g["color"]), _p(g["depth"]), _p(g["obj"]))
        if rc != 0:
            raise Mm3eError(rc, "gframe holds no frame")
        return g

    def reproject(self, cam, deltas=(), fill=True):
        """reproject::reproject (fill) / reproject::warp on the device -> (rgba, filled)."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        filled = np.zeros((self.height, self.width), np.uint8)
        d, n = _deltas(deltas)
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_gframe_reproject(self.h, _p(_cam(cam)), _p(d), n, 1 if fill else 0, _p(out), _p(filled), err,
                                         1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        return out, filled

    def close(self):
        if self.h:
            lib().mm3e_gframe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MeshBaker:
    """A resident mesh -> SDF bake (mesh::bake_sdf on the GPU): create once, run / read often."""

    def __init__(self, renderer, positions, triangles, resolution: int, padding: float):
        pos, tri = _mesh_arrays(positions, triangles)
        self.h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_mesh_baker_create(renderer.h, _p(pos), len(pos), _p(tri), len(tri), int(resolution),
                                          float(padding), C.byref(self.h), err, 1024)
        if rc != 0:
            self.h = None
            raise Mm3eError(rc, err.value.decode())
        self.layout = SdfLayout()
        lib().mm3e_mesh_baker_layout(self.h, C.byref(self.layout))

    def info(self):
        t, n, d = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        lib().mm3e_mesh_baker_info(self.h, C.byref(t), C.byref(n), C.byref(d))
        return {"triangles": t.value, "nodes": n.value, "depth": d.value}

    def run(self):
        rc = lib().mm3e_mesh_baker_run(self.h)
        if rc != 0:
            raise Mm3eError(rc, "mesh bake launch failed")

    def wait(self):
        """GPU milliseconds of (parity kernel, nearest kernel) of the last run."""
        a, b = C.c_float(0), C.c_float(0)
        rc = lib().mm3e_mesh_baker_wait(self.h, C.byref(a), C.byref(b))
        if rc != 0:
            raise Mm3eError(rc, "mesh bake failed")
        return a.value, b.value

    def read(self) -> np.ndarray:
        nx, ny, nz = self.layout.dims
        out = np.zeros((nz, ny, nx), np.float32)
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_mesh_baker_read(self.h, _p(out), out.size, err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        return out

    def close(self):
        if self.h:
            lib().mm3e_mesh_baker_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def list_adapters() -> str:
    buf = C.create_string_buffer(4096)
    lib().mm3e_list_adapters(buf, 4096)
    return buf.value.decode()


class Renderer:
    """GpuRenderer (mm3e-gpu/src/lib.rs:57-128)."""

    def __init__(self, adapter: Optional[str] = None):
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_renderer_create(adapter.encode() if adapter else None, C.byref(h), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        self.h = h

    @property
    def adapter_name(self) -> str:
        return lib().mm3e_renderer_adapter_name(self.h).decode()

    def wait_idle(self):
        rc = lib().mm3e_renderer_wait_idle(self.h)
        if rc != 0:
            raise Mm3eError(rc, "wait_idle failed")

    def set_codegen(self, mode: int):
        """CODEGEN_OFF / CODEGEN_BACKGROUND / CODEGEN_SYNC for scenes compiled afterwards."""
        rc = lib().mm3e_renderer_set_codegen(self.h, mode)
        if rc != 0:
            raise Mm3eError(rc, "bad codegen mode")

    def set_tile_order(self, on: bool):
        """Most-expensive-first tile dispatch of generated kernels for scenes compiled afterwards."""
        rc = lib().mm3e_renderer_set_tile_order(self.h, 1 if on else 0)
        if rc != 0:
            raise Mm3eError(rc, "set_tile_order failed")

    def set_codegen_film(self, on: bool):
        """Generated modules of scenes compiled afterwards also hold the film and G-buffer kernels."""
        rc = lib().mm3e_renderer_set_codegen_film(self.h, 1 if on else 0)
        if rc != 0:
            raise Mm3eError(rc, "set_codegen_film failed")

    def set_wavefront(self, budget: int, passes: int = 2, blocks_per_cu: int = 0):
        """Queued frames: march passes of at most `budget` evaluations over compacted queues of
        unfinished rays, then shading of the compacted hit queue (budget 0: one kernel per frame).
        Generated kernels of scenes compiled while this is on carry the queued stages."""
        rc = lib().mm3e_renderer_set_wavefront(self.h, budget, passes, blocks_per_cu)
        if rc != 0:
            raise Mm3eError(rc, "set_wavefront: passes 0..7, blocks_per_cu 0..64")

    def selftest_sqrt(self) -> int:
        """Inputs (of all 2^32) where the kernels' square root differs from IEEE sqrtf."""
        n = C.c_uint64(0)
        rc = lib().mm3e_selftest_sqrt(self.h, C.byref(n))
        if rc != 0:
            raise Mm3eError(rc, "sqrt self-test failed to run")
        return n.value

    def compile(self, ir: bytes, width: int, height: int) -> "GpuScene":
        return GpuScene(self, ir, width, height)

    def close(self):
        if getattr(self, "h", None):
            lib().mm3e_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bake_sdf(self, positions, triangles, resolution: int, padding: float):
        """mesh::bake_sdf on this renderer's GPU -> (layout, samples[nz, ny, nx])."""
        pos, tri = _mesh_arrays(positions, triangles)
        lay = sdf_layout(pos, tri, resolution, padding)
        nx, ny, nz = lay.dims
        out = np.zeros((nz, ny, nx), np.float32)
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_mesh_bake_sdf(self.h, _p(pos), len(pos), _p(tri), len(tri), int(resolution),
                                      float(padding), _p(out), out.size, C.byref(lay), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        return lay, out


def device_bake(renderer):
    """A `bake(positions, triangles, resolution, padding) -> (dims, min, cell, samples)` callable
    running mesh::bake_sdf on the renderer's GPU (what configs.c5_scene takes)."""
    def bake(positions, triangles, resolution, padding):
        lay, out = renderer.bake_sdf(positions, triangles, resolution, padding)
        return tuple(int(d) for d in lay.dims), tuple(lay.min), tuple(lay.cell), out
    return bake


class GpuScene:
    """GpuScene (mm3e-gpu/src/lib.rs:289-416)."""

    def __init__(self, renderer: Renderer, ir: bytes, width: int, height: int):
        self.renderer = renderer
        self.width, self.height = int(width), int(height)
        self._ir = C.create_string_buffer(ir, len(ir))
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_scene_compile(renderer.h, C.cast(self._ir, C.c_void_p), len(ir), self.width, self.height,
                                      C.byref(h), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mm3e_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render_rgba(self, cam, dyn: Optional[Sequence] = None) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.uint8)
        d = None
        n = 0
        if dyn is not None and len(dyn):
            d = np.ascontiguousarray(dyn, np.float32).reshape(-1, 8)
            n = len(d)
        rc = lib().mm3e_scene_render_rgba(self.h, _p(_cam(cam)), _p(d), n, _p(out), out.nbytes)
        if rc != 0:
            raise Mm3eError(rc, "render_rgba failed")
        return out

    def set_wavefront(self, budget=None, passes: int = 2, blocks_per_cu: int = 0):
        """The scene's own queued-frame setting; budget None returns it to the renderer's."""
        own = budget is not None
        rc = lib().mm3e_scene_set_wavefront(self.h, 1 if own else 0, budget or 0, passes, blocks_per_cu)
        if rc != 0:
            raise Mm3eError(rc, "set_wavefront: passes 0..7, blocks_per_cu 0..64")

    def autotune(self, cam, reps: int = 5) -> str:
        """Times one-kernel and queued frames of `cam`; keeps the fastest; returns the report."""
        rep = C.create_string_buffer(2048)
        rc = lib().mm3e_scene_autotune(self.h, _p(_cam(cam)), reps, rep, 2048)
        if rc != 0:
            raise Mm3eError(rc, "scene autotune failed: " + rep.value.decode(errors="replace"))
        return rep.value.decode()

    def render_rgba_noalloc(self, cam, rgba_dev: int, nbytes: int):
        """mm3e_scene_render_rgba_noalloc: enqueue a frame into the device buffer at address
        rgba_dev (w*h*4 bytes); the library allocates nothing.  Does not wait."""
        err = C.create_string_buffer(512)
        rc = lib().mm3e_scene_render_rgba_noalloc(self.h, _p(_cam(cam)), C.c_void_p(rgba_dev), nbytes, err, 512)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())

    def render_only(self, cam):
        rc = lib().mm3e_scene_render_only(self.h, _p(_cam(cam)))
        if rc != 0:
            raise Mm3eError(rc, "render_only failed")

    def render_linear(self, cam) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float32)
        rc = lib().mm3e_scene_render_linear(self.h, _p(_cam(cam)), _p(out), out.size)
        if rc != 0:
            raise Mm3eError(rc, "render_linear failed")
        return out

    def render_aov(self, cam):
        sh = (self.height, self.width)
        hit = np.zeros(sh, np.uint8)
        mat = np.zeros(sh, np.uint32)
        depth = np.zeros(sh, np.float32)
        steps = np.zeros(sh, np.uint32)
        rc = lib().mm3e_scene_render_aov(self.h, _p(_cam(cam)), _p(hit), _p(mat), _p(depth), _p(steps))
        if rc != 0:
            raise Mm3eError(rc, "render_aov failed")
        return hit, mat, depth, steps

    def render_gbuffer(self, cam, gframe: "GFrame", movers=()):
        """render_gbuffer (lib.rs:951-1030) into gframe; movers: (cx, cy, cz, radius) each."""
        mv = np.ascontiguousarray(np.asarray(movers, np.float32).reshape(-1, 4)) if len(movers) else None
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_scene_render_gbuffer(self.h, _p(_cam(cam)), _p(mv), 0 if mv is None else len(mv), gframe.h,
                                             err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())

    def reproject_hybrid(self, gframe: "GFrame", cam, deltas=()):
        """reproject_hybrid (lib.rs:1043-1080) on the device -> (rgba, holes rendered)."""
        out = np.zeros((gframe.height, gframe.width, 4), np.uint8)
        d, n = _deltas(deltas)
        holes = C.c_uint32(0)
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_scene_reproject_hybrid(self.h, gframe.h, _p(_cam(cam)), _p(d), n, _p(out), C.byref(holes), err,
                                               1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        return out, holes.value

    def render_film(self, cam, aperture_radius=0.0, focus_distance=1.0):
        """render_film (film.rs:63-198) on the device -> dict of beauty, foreground, alpha, depth,
        normals, material_ids as [h, w(, 3)] arrays."""
        h, w = self.height, self.width
        out = {"beauty": np.zeros((h, w, 3), np.float32), "foreground": np.zeros((h, w, 3), np.float32),
               "alpha": np.zeros((h, w), np.float32), "depth": np.zeros((h, w), np.float32),
               "normals": np.zeros((h, w, 3), np.float32), "material_ids": np.zeros((h, w), np.uint32)}
        lens = (C.c_float * 2)(aperture_radius, focus_distance)
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_scene_render_film(self.h, _p(_cam(cam)), C.byref(lens),
                                          *(_p(out[k]) for k in ("beauty", "foreground", "alpha", "depth", "normals",
                                                                 "material_ids")), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        return out

    def last_timing(self):
        a, b = C.c_float(0), C.c_float(0)
        rc = lib().mm3e_scene_last_timing(self.h, C.byref(a), C.byref(b))
        if rc != 0:
            raise Mm3eError(rc, "no frame timed yet")
        return a.value, b.value

    def codegen_status(self, wait=False):
        """(state, message): CODEGEN_NONE / QUEUED / READY / FAILED of the scene's generated kernel."""
        st = C.c_int(0)
        err = C.create_string_buffer(4096)
        rc = lib().mm3e_scene_codegen_status(self.h, 1 if wait else 0, C.byref(st), err, 4096)
        if rc != 0:
            raise Mm3eError(rc, "codegen_status failed")
        return st.value, err.value.decode()

    def info(self):
        v = [C.c_uint32() for _ in range(5)]
        rc = lib().mm3e_scene_info(self.h, *[C.byref(x) for x in v])
        if rc != 0:
            raise Mm3eError(rc, "info failed")
        return dict(zip(("objects", "segments", "grid_cells", "grid_words", "stage_bytes"), [x.value for x in v]))

    def kernel_name(self) -> str:
        """Render kernel that render_only / render_rgba dispatch for this scene."""
        buf = C.create_string_buffer(64)
        rc = lib().mm3e_scene_kernel_name(self.h, buf, 64)
        if rc != 0:
            raise Mm3eError(rc, "kernel_name failed")
        return buf.value.decode()

    def timing(self, reset=True):
        """(frames, mean render-kernel ms, mean post ms) over the frames since the last reset."""
        n = C.c_uint64(0)
        a, b = C.c_double(0), C.c_double(0)
        rc = lib().mm3e_scene_timing(self.h, 1 if reset else 0, C.byref(n), C.byref(a), C.byref(b))
        if rc != 0:
            raise Mm3eError(rc, "timing failed")
        k = max(n.value, 1)
        return n.value, a.value / k, b.value / k

    def gi_raw(self):
        mn = np.zeros(3, np.float32)
        sz = np.zeros(3, np.float32)
        dims = np.zeros(3, np.uint32)
        n = C.c_uint64(0)
        lib().mm3e_scene_gi_raw(self.h, _p(mn), _p(sz), _p(dims), None, C.byref(n))
        if n.value == 0:
            return None
        cubes = np.zeros((n.value, 6, 3), np.float32)
        rc = lib().mm3e_scene_gi_raw(self.h, None, None, None, _p(cubes), C.byref(n))
        if rc != 0:
            raise Mm3eError(rc, "gi_raw failed")
        return mn, sz, dims, cubes

    def field(self, pts):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        d = np.zeros(len(pts), np.float32)
        m = np.zeros(len(pts), np.uint32)
        rc = lib().mm3e_scene_field(self.h, _p(pts), len(pts), _p(d), _p(m))
        if rc != 0:
            raise Mm3eError(rc, "field probe failed")
        return d, m

    def march(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = len(rays)
        t = np.zeros(n, np.float32)
        hit = np.zeros(n, np.uint8)
        mat = np.zeros(n, np.uint32)
        nor = np.zeros((n, 3), np.float32)
        st = np.zeros(n, np.uint32)
        rc = lib().mm3e_scene_march(self.h, _p(rays), n, _p(t), _p(hit), _p(mat), _p(nor), _p(st))
        if rc != 0:
            raise Mm3eError(rc, "march probe failed")
        return t, hit.astype(bool), mat, nor, st


class _BatchJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("cam", C.c_float * 13), ("width", C.c_uint32), ("height", C.c_uint32),
                ("out", C.c_void_p)]


class Batch:
    """Many (compiled scene, camera, size) jobs in one persistent tile-queue launch, replayed
    from a captured hipGraph (mm3e_batch_* in include/mm3e_hip.h)."""

    def __init__(self, renderer: Renderer, jobs):
        """jobs: sequence of (GpuScene, camera[13], (width, height))."""
        self.renderer = renderer
        self.sizes = [(int(w), int(h)) for _, _, (w, h) in jobs]
        self._scenes = [g for g, _, _ in jobs]  # keep compiled scenes alive
        arr = (_BatchJob * len(jobs))()
        for i, (g, cam, (w, h)) in enumerate(jobs):
            arr[i].scene = g.h
            arr[i].cam[:] = [float(x) for x in _cam(cam)]
            arr[i].width, arr[i].height = int(w), int(h)
            arr[i].out = None
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_batch_create(renderer.h, C.cast(arr, C.c_void_p), len(jobs), C.byref(h), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        self.h = h

    def autotune(self, reps=2) -> str:
        rep = C.create_string_buffer(4096)
        rc = lib().mm3e_batch_autotune(self.h, reps, rep, 4096)
        if rc != 0:
            raise Mm3eError(rc, "autotune failed")
        return rep.value.decode()

    def set_camera(self, job, cam):
        rc = lib().mm3e_batch_set_camera(self.h, job, _p(_cam(cam)))
        if rc != 0:
            raise Mm3eError(rc, "set_camera failed")

    def launch(self):
        rc = lib().mm3e_batch_launch(self.h)
        if rc != 0:
            raise Mm3eError(rc, "batch launch failed")

    def wait(self) -> float:
        ms = C.c_float(0)
        rc = lib().mm3e_batch_wait(self.h, C.byref(ms))
        if rc != 0:
            raise Mm3eError(rc, "batch wait failed")
        return ms.value

    def read(self, job) -> np.ndarray:
        w, h = self.sizes[job]
        out = np.zeros((h, w, 4), np.uint8)
        rc = lib().mm3e_batch_read(self.h, job, _p(out), out.nbytes)
        if rc != 0:
            raise Mm3eError(rc, "batch read failed")
        return out

    def info(self):
        px, tiles, bpc, tw, th, lds = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        lib().mm3e_batch_info(self.h, C.byref(px), C.byref(tiles), C.byref(bpc), C.byref(tw), C.byref(th), C.byref(lds))
        return {"pixels": px.value, "tiles": tiles.value, "blocks_per_cu": bpc.value, "tile": (tw.value, th.value),
                "lds_bytes": lds.value}

    def close(self):
        if getattr(self, "h", None):
            lib().mm3e_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_requests", C.c_uint32), ("max_inflight", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("blocks_per_cu", C.c_uint32)]


class Server:
    """Continuous-batching render server over resident compiled scenes (mm3e_server_*)."""

    def __init__(self, renderer: Renderer, **cfg):
        c = ServerConfig()
        lib().mm3e_server_default_config(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_server_create(renderer.h, C.byref(c), C.byref(h), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode())
        self.h = h
        self.renderer = renderer
        self._scenes = []
        self._outs = {}

    def add_scene(self, g: GpuScene) -> int:
        sid = C.c_uint32()
        rc = lib().mm3e_server_add_scene(self.h, g.h, C.byref(sid))
        if rc != 0:
            raise Mm3eError(rc, "add_scene failed")
        self._scenes.append(g)
        return sid.value

    def submit(self, scene_id, cam, w, h, keep_image=True) -> int:
        out = np.zeros((h, w, 4), np.uint8) if keep_image else None
        rid = C.c_uint64()
        rc = lib().mm3e_server_submit(self.h, scene_id, _p(_cam(cam)), w, h, _p(out), C.byref(rid))
        if rc != 0:
            raise Mm3eError(rc, "submit refused (queue full)" if rc == 3 else "submit failed")
        self._outs[rid.value] = out
        return rid.value

    def submit_text(self, text: str, w: int = 0, h: int = 0, keep_image=True) -> int:
        """A request as `.mm3e` text (scene + camera); compiled once per distinct scene and kept
        resident.  w = h = 0 renders at the text's `size`."""
        b = text.encode()
        ir, _ = parse_text(text)
        hw = struct.unpack_from("<II", ir, 16)
        ww, hh = (w, h) if (w or h) else hw
        out = np.zeros((hh, ww, 4), np.uint8) if keep_image else None
        rid = C.c_uint64()
        err = C.create_string_buffer(1024)
        rc = lib().mm3e_server_submit_text(self.h, b, len(b), w, h, _p(out), C.byref(rid), err, 1024)
        if rc != 0:
            raise Mm3eError(rc, err.value.decode() or "submit_text failed")
        self._outs[rid.value] = out
        return rid.value

    def poll(self, rid) -> bool:
        d = C.c_int(0)
        lib().mm3e_server_poll(self.h, rid, C.byref(d))
        return bool(d.value)

    def wait(self, rid):
        """-> (image or None, latency_ms)."""
        lat = C.c_double(0)
        rc = lib().mm3e_server_wait(self.h, rid, C.byref(lat))
        out = self._outs.pop(rid, None)
        if rc != 0:
            raise Mm3eError(rc, f"request {rid} failed")
        return out, lat.value

    def stats(self):
        v = [C.c_uint64() for _ in range(5)]
        lib().mm3e_server_stats(self.h, *[C.byref(x) for x in v])
        return dict(zip(("batches", "completed", "rejected", "max_batch", "pixels"), [x.value for x in v]))

    def close(self):
        if getattr(self, "h", None):
            lib().mm3e_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
