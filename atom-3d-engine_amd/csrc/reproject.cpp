// reproject.cpp — G-buffer frames and reprojection (mm3e-orchestrator/src/reproject.rs,
// lib.rs:951-1080) on the device: a frame's display colour, primary depth and mover tags stay
// resident in HBM with its camera; later cameras warp it (z-buffered forward splat), fill holes
// along rows, or re-render just the holes through the scene (the hybrid).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/mm3e_hip.h"
#include "launch.h"
#include "runtime.h"

using namespace mm3e;

struct mm3e_gframe {
  mm3e_renderer* r = nullptr;
  uint32_t w = 0, h = 0;
  mm3e_camera cam{};
  bool valid = false;            // holds a frame (rendered or uploaded)
  uchar4* d_color = nullptr;     // GFrame::color
  float* d_depth = nullptr;      // GFrame::depth
  int32_t* d_obj = nullptr;      // GFrame::obj
  // reprojection scratch: z-buffer keys, output, filled mask, hole list + count, movers/deltas
  unsigned long long* d_keys = nullptr;
  uchar4* d_out = nullptr;
  uint8_t* d_filled = nullptr;
  uint32_t* d_list = nullptr;
  uint32_t* d_count = nullptr;
  float* d_param = nullptr;      // kMaxMovers * 4 floats
};

namespace {

constexpr uint32_t kMaxMovers = 256;

void free_gframe(mm3e_gframe* g) {
  if (!g) return;
  hipSetDevice(g->r->device);
  hipStreamSynchronize(g->r->stream);
  hipFree(g->d_color);
  hipFree(g->d_depth);
  hipFree(g->d_obj);
  hipFree(g->d_keys);
  hipFree(g->d_out);
  hipFree(g->d_filled);
  hipFree(g->d_list);
  hipFree(g->d_count);
  hipFree(g->d_param);
  delete g;
}

// Warp g into `cam` (reproject.rs:82-130) into g->d_out / g->d_filled.
bool warp(mm3e_gframe* g, const mm3e_camera* cam, const float* deltas, uint32_t n_deltas, std::string& msg) {
  if (n_deltas > kMaxMovers) {
    msg = "at most " + std::to_string(kMaxMovers) + " mover deltas";
    return false;
  }
  hipStream_t st = g->r->stream;
  if (n_deltas &&
      !hip_ok(hipMemcpyAsync(g->d_param, deltas, (size_t)n_deltas * 12, hipMemcpyHostToDevice, st), "upload deltas", &msg))
    return false;
  const FrameParams prev = make_frame(g->cam, g->w, g->h);
  const FrameParams next = make_frame(*cam, g->w, g->h);
  return hip_ok(launch_warp(st, g->w, g->h, prev, next, g->d_depth, g->d_obj, g->d_param, n_deltas, g->d_color,
                            g->d_keys, g->d_out, g->d_filled),
                "warp", &msg);
}

bool read_out(mm3e_gframe* g, uint8_t* out, uint8_t* filled, std::string& msg) {
  const size_t n = (size_t)g->w * g->h;
  bool ok = hip_ok(hipStreamSynchronize(g->r->stream), "reproject", &msg);
  if (ok && out) ok = hip_ok(hipMemcpy(out, g->d_out, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok && filled) ok = hip_ok(hipMemcpy(filled, g->d_filled, n, hipMemcpyDeviceToHost), "readback", &msg);
  return ok;
}

}  // namespace

extern "C" {

int mm3e_gframe_create(mm3e_renderer* r, uint32_t width, uint32_t height, mm3e_gframe** out, char* err,
                       size_t errlen) {
  if (!r || !out) return MM3E_ERR_INVALID;
  *out = nullptr;
  if (width == 0 || height == 0 || width > 16384 || height > 16384) {
    set_err(err, errlen, "GFrame: " + std::to_string(width) + "x" + std::to_string(height) + " out of range 1..=16384");
    return MM3E_ERR_INVALID;
  }
  auto* g = new mm3e_gframe();
  g->r = r;
  g->w = width;
  g->h = height;
  const size_t n = (size_t)width * height;
  hipSetDevice(r->device);
  std::string msg;
  bool ok = hip_ok(hipMalloc((void**)&g->d_color, n * 4), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_depth, n * 4), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_obj, n * 4), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_keys, n * 8), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_out, n * 4), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_filled, n), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_list, n * 4), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_count, 4), "hipMalloc(gframe)", &msg) &&
            hip_ok(hipMalloc((void**)&g->d_param, kMaxMovers * 16), "hipMalloc(gframe)", &msg);
  if (!ok) {
    set_err(err, errlen, msg);
    free_gframe(g);
    return MM3E_ERR_DEVICE;
  }
  *out = g;
  return MM3E_OK;
}

void mm3e_gframe_destroy(mm3e_gframe* g) { free_gframe(g); }

int mm3e_gframe_upload(mm3e_gframe* g, const mm3e_camera* cam, const uint8_t* color, const float* depth,
                       const int32_t* obj) {
  if (!g || !cam || !color || !depth || !obj) return MM3E_ERR_INVALID;
  const size_t n = (size_t)g->w * g->h;
  hipSetDevice(g->r->device);
  std::string msg;
  bool ok = hip_ok(hipStreamSynchronize(g->r->stream), "sync", &msg) &&
            hip_ok(hipMemcpy(g->d_color, color, n * 4, hipMemcpyHostToDevice), "upload", &msg) &&
            hip_ok(hipMemcpy(g->d_depth, depth, n * 4, hipMemcpyHostToDevice), "upload", &msg) &&
            hip_ok(hipMemcpy(g->d_obj, obj, n * 4, hipMemcpyHostToDevice), "upload", &msg);
  if (!ok) return MM3E_ERR_DEVICE;
  g->cam = *cam;
  g->valid = true;
  return MM3E_OK;
}

int mm3e_gframe_read(mm3e_gframe* g, mm3e_camera* cam, uint8_t* color, float* depth, int32_t* obj) {
  if (!g || !g->valid) return MM3E_ERR_INVALID;
  const size_t n = (size_t)g->w * g->h;
  hipSetDevice(g->r->device);
  std::string msg;
  bool ok = hip_ok(hipStreamSynchronize(g->r->stream), "sync", &msg);
  if (ok && color) ok = hip_ok(hipMemcpy(color, g->d_color, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok && depth) ok = hip_ok(hipMemcpy(depth, g->d_depth, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok && obj) ok = hip_ok(hipMemcpy(obj, g->d_obj, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (cam) *cam = g->cam;
  return ok ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_scene_render_gbuffer(mm3e_scene* s, const mm3e_camera* cam, const mm3e_mover* movers, uint32_t n_movers,
                              mm3e_gframe* g, char* err, size_t errlen) {
  if (!s || !cam || !g || (!movers && n_movers)) return MM3E_ERR_INVALID;
  if (g->r != s->r || g->w != s->w || g->h != s->h) {
    set_err(err, errlen, "render_gbuffer: the GFrame must match the scene's renderer and size");
    return MM3E_ERR_INVALID;
  }
  if (n_movers > kMaxMovers) {
    set_err(err, errlen, "render_gbuffer: at most " + std::to_string(kMaxMovers) + " movers");
    return MM3E_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  hipStream_t st = s->r->stream;
  std::string msg;
  bool ok = true;
  if (n_movers)
    ok = hip_ok(hipMemcpyAsync(g->d_param, movers, (size_t)n_movers * sizeof(mm3e_mover), hipMemcpyHostToDevice, st),
                "upload movers", &msg);
  const FrameParams fp = make_frame(*cam, s->w, s->h);
  if (ok)
    ok = hip_ok(scene_gbuffer(s, st, fp, reinterpret_cast<const GMover*>(g->d_param), n_movers, nullptr, nullptr, 0,
                              g->d_color, g->d_depth, g->d_obj),
                "gbuffer launch", &msg);
  // the mover table must not change under the kernel: wait before returning
  if (ok) ok = hip_ok(hipStreamSynchronize(st), "gbuffer", &msg);
  if (!ok) {
    set_err(err, errlen, msg);
    return MM3E_ERR_DEVICE;
  }
  g->cam = *cam;
  g->valid = true;
  return MM3E_OK;
}

int mm3e_gframe_reproject(mm3e_gframe* g, const mm3e_camera* cam, const float* mover_deltas, uint32_t n_deltas,
                          int fill_holes, uint8_t* out, uint8_t* filled, char* err, size_t errlen) {
  if (!g || !cam || (!mover_deltas && n_deltas)) return MM3E_ERR_INVALID;
  if (!g->valid) {
    set_err(err, errlen, "reproject: the GFrame holds no frame");
    return MM3E_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g->r->mu);
  hipSetDevice(g->r->device);
  std::string msg;
  bool ok = warp(g, cam, mover_deltas, n_deltas, msg);
  if (ok && fill_holes) ok = hip_ok(launch_row_fill(g->r->stream, g->w, g->h, g->d_out, g->d_filled), "row fill", &msg);
  if (ok) ok = read_out(g, out, filled, msg);
  if (!ok) {
    set_err(err, errlen, msg);
    return MM3E_ERR_DEVICE;
  }
  return MM3E_OK;
}

int mm3e_scene_reproject_hybrid(mm3e_scene* s, mm3e_gframe* g, const mm3e_camera* cam, const float* mover_deltas,
                                uint32_t n_deltas, uint8_t* out, uint32_t* n_holes, char* err, size_t errlen) {
  if (!s || !g || !cam || (!mover_deltas && n_deltas)) return MM3E_ERR_INVALID;
  if (!g->valid || g->r != s->r) {
    set_err(err, errlen, "reproject_hybrid: the GFrame holds no frame of this renderer");
    return MM3E_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g->r->mu);
  hipSetDevice(g->r->device);
  hipStream_t st = g->r->stream;
  std::string msg;
  bool ok = warp(g, cam, mover_deltas, n_deltas, msg);
  uint32_t holes = 0;
  // the hole list: counter cleared, then compaction; the render reads the count on the device
  if (ok) ok = hip_ok(hipMemsetAsync(g->d_count, 0, 4, st), "reset", &msg);
  if (ok) ok = hip_ok(launch_hole_list(st, g->w * g->h, g->d_filled, g->d_list, g->d_count), "hole list", &msg);
  // the holes rendered for real at the new camera, over the frame's size (lib.rs:1059-1079)
  const FrameParams fp = make_frame(*cam, g->w, g->h);
  if (ok)
    ok = hip_ok(scene_gbuffer(s, st, fp, nullptr, 0, g->d_list, g->d_count, g->w * g->h, g->d_out, nullptr, nullptr),
                "hole render", &msg);
  if (ok) ok = hip_ok(hipMemcpyAsync(&holes, g->d_count, 4, hipMemcpyDeviceToHost, st), "hole count", &msg);
  if (ok) ok = read_out(g, out, nullptr, msg);
  if (!ok) {
    set_err(err, errlen, msg);
    return MM3E_ERR_DEVICE;
  }
  if (n_holes) *n_holes = holes;
  return MM3E_OK;
}

}  // extern "C"
