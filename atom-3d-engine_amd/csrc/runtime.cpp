// runtime.cpp — the C ABI of libmm3e_hip.so: renderer / compiled-scene objects, device-side
// GI bake, resident framebuffers and per-frame launches.
//
// Mirrors the reference GPU backend surface (mm3e-gpu/src/lib.rs): GpuRenderer::{new,
// with_adapter, adapter_name, wait_idle, compile_checked} and GpuScene::{render_only,
// render_rgba(_dyn), render}.  A compiled scene owns one HBM allocation with its tables, its
// GI probes, and its framebuffers (HDR, bloom scratch, RGBA8) plus a pinned host staging
// buffer, so a frame is: one camera record by value -> render kernel -> 2 post kernels ->
// one D2H copy.  All work goes to one default-priority stream per renderer.
#include "runtime.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "codegen.h"
#include "launch.h"
#include "rtc.h"

namespace mm3e {

void set_err(char* err, size_t n, const std::string& s) {
  if (err && n) std::snprintf(err, n, "%s", s.c_str());
}

bool hip_ok(hipError_t e, const char* what, std::string* msg) {
  if (e == hipSuccess) return true;
  if (msg) *msg = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

static std::string describe(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
  return std::string(p.name) + " (" + p.gcnArchName + ")";
}

static std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

FrameParams make_frame(const mm3e_camera& c, uint32_t w, uint32_t h) {
  FrameParams fp;
  std::memset(&fp, 0, sizeof fp);
  std::memcpy(fp.eye, c.eye, 12);
  fp.fov_scale = c.fov_scale;
  std::memcpy(fp.fwd, c.forward, 12);
  std::memcpy(fp.right, c.right, 12);
  std::memcpy(fp.up, c.up, 12);
  fp.width = w;
  fp.height = h;
  fp.aspect = (float)w / (float)h;  // camera.rs:60
  return fp;
}

static_assert(GEN_N_OPT == 5, "GenJob::fn_opt holds one entry per GenOpt");

hipFunction_t gen_opt(const mm3e_scene* s, int k) {
  const GenJob* j = s->gen.get();
  if (!j || j->state.load(std::memory_order_acquire) != MM3E_CODEGEN_READY) return nullptr;
  return j->fn_opt[k];
}

hipError_t scene_gbuffer(const mm3e_scene* s, hipStream_t st, const FrameParams& fp, const GMover* movers,
                         uint32_t n_movers, const uint32_t* list, const uint32_t* n_list, uint32_t max_list,
                         uchar4* color, float* depth, int32_t* obj) {
  hipFunction_t fn = gen_opt(s, GEN_GBUFFER);
  if (!fn)
    return launch_gbuffer(s->bounces, st, s->d_scene, s->general, fp, s->pp.exposure, movers, n_movers, list, n_list,
                          max_list, color, depth, obj);
  if (list && max_list == 0) return hipSuccess;
  const uint32_t gx = list ? (max_list + 255) / 256 : (fp.width + 15) / 16;
  const uint32_t gy = list ? 1u : (fp.height + 15) / 16;
  const DevScene* sc = s->d_scene;
  FrameParams f = fp;
  float exposure = s->pp.exposure;
  void* args[] = {&sc, &f, &exposure, &movers, &n_movers, &list, &n_list, &color, &depth, &obj};
  return hipModuleLaunchKernel(fn, gx, gy, 1, 256, 1, 1, 0, st, args, nullptr);
}

// Tile shape for one block of <= 256 lanes: tw*th pixels x aa^2 samples.
void tile_shape(uint32_t aa, uint32_t* tw, uint32_t* th) {
  uint32_t ns = aa * aa;
  uint32_t P = std::max(1u, 256u / ns);
  uint32_t t = (uint32_t)std::floor(std::sqrt((double)P));
  if (t == 0) t = 1;
  *th = t;
  *tw = std::max(1u, P / t);
}

// ---- device GI bake (gi.rs:55-125, lib.rs:665-711) ---------------------------------------
bool bake_gi_device(mm3e_scene* s, std::string& msg) {
  const mm3e_ir_header& h = s->host.hdr;
  float lo[3], hi[3];
  if (!gi_bounds(s->host, lo, hi)) return true;  // nothing bounded: no GI (gi.rs:192-196)
  const uint32_t nx = std::max(h.gi_dims[0], 1u), ny = std::max(h.gi_dims[1], 1u), nz = std::max(h.gi_dims[2], 1u);
  const uint64_t n_probes = (uint64_t)nx * ny * nz;
  GiBakeParams bp;
  std::memset(&bp, 0, sizeof bp);
  float size[3];
  for (int a = 0; a < 3; a++) {
    size[a] = hi[a] - lo[a];
    bp.lo[a] = lo[a];
  }
  const uint32_t dd[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    bp.cell[a] = size[a] / (float)(std::max(dd[a], 2u) - 1);
    bp.dims[a] = dd[a];
  }
  gi_axis_dirs(h.gi_samples, bp.dirs, &bp.rays_per_face);
  const uint32_t R = bp.rays_per_face;
  float* d_gi = nullptr;
  float4* d_rays = nullptr;
  if (!hip_ok(hipMalloc(&d_gi, n_probes * 24 * sizeof(float)), "hipMalloc(gi)", &msg)) return false;
  s->d_gi = d_gi;
  // bounded scratch: at most 2^20 rays in flight per pass
  const uint64_t probes_per_pass = std::max<uint64_t>(1, (1u << 20) / (6 * R));
  const uint64_t pass_probes = std::min<uint64_t>(probes_per_pass, n_probes);
  if (!hip_ok(hipMalloc(&d_rays, pass_probes * 6 * R * sizeof(float4)), "hipMalloc(gi rays)", &msg)) return false;
  hipStream_t st = s->r->stream;
  bool ok = true;
  for (uint64_t p0 = 0; p0 < n_probes && ok; p0 += pass_probes) {
    const uint32_t np = (uint32_t)std::min<uint64_t>(pass_probes, n_probes - p0);
    const uint64_t lanes = (uint64_t)np * 6 * R;
    ok = hip_ok(launch_gi_gather_table(dim3((uint32_t)((lanes + 255) / 256)), dim3(256), st, s->d_scene,
                                       s->stage_bytes, bp, d_rays,
                                       (uint32_t)p0, np),
                "gi gather launch", &msg);
    if (!ok) break;
    ok = hip_ok(launch_gi_reduce(dim3((np * 6 + 255) / 256), dim3(256), st, bp, d_rays,
                                 reinterpret_cast<float4*>(d_gi) + (size_t)p0 * 6, np),
                "gi reduce launch", &msg);
  }
  if (ok) ok = hip_ok(hipStreamSynchronize(st), "gi bake", &msg);
  hipFree(d_rays);
  if (!ok) return false;
  s->gi_dims[0] = nx; s->gi_dims[1] = ny; s->gi_dims[2] = nz;
  for (int a = 0; a < 3; a++) {
    s->gi_min[a] = lo[a];
    s->gi_size[a] = size[a];
  }
  s->has_gi = true;
  return true;
}

static void patch_gi(mm3e_scene* s, DevScene& d) {
  if (!s->has_gi) return;
  d.flags |= SC_HAS_GI;
  d.gi = s->d_gi;
  for (int a = 0; a < 3; a++) {
    d.gi_dims[a] = s->gi_dims[a];
    d.gi_min[a] = s->gi_min[a];
    d.gi_size[a] = s->gi_size[a];
  }
}

bool upload_tables(mm3e_scene* s, const std::vector<mm3e_ir_object>& xo, const std::vector<mm3e_ir_material>& xm,
                   uint8_t** d_buf, size_t* cap, DevScene* hdr_out, std::string& msg) {
  const uint64_t need = blob_size(s->host, xo, xm);
  if (*d_buf == nullptr || *cap < need) {
    if (*d_buf) hipFree(*d_buf);
    *d_buf = nullptr;
    if (!hip_ok(hipMalloc((void**)d_buf, need), "hipMalloc(scene tables)", &msg)) return false;
    *cap = need;
  }
  Blob b;
  build_blob(s->host, xo, xm, s->aa, s->bounces, (uint64_t)(uintptr_t)*d_buf, s->d_gi, false, b);
  patch_gi(s, b.header);
  std::memcpy(b.bytes.data(), &b.header, sizeof b.header);
  if (!hip_ok(hipMemcpyAsync(*d_buf, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice, s->r->stream),
              "upload scene tables", &msg))
    return false;
  if (!hip_ok(hipStreamSynchronize(s->r->stream), "upload scene tables", &msg)) return false;
  if (hdr_out) *hdr_out = b.header;
  const uint32_t stage = b.header.stage_bytes <= s->r->lds_budget ? b.header.stage_bytes : 0;
  if (*d_buf == s->d_tables) {
    s->stage_bytes = stage;
    s->general = scene_is_general(b.header);
  } else {
    s->dyn_stage_bytes = stage;
    s->dyn_general = scene_is_general(b.header);
  }
  return true;
}

// ---- scene code generation ----------------------------------------------------------------
static void gen_run(mm3e::GenJob& j) {
  j.state.store(MM3E_CODEGEN_QUEUED);
  hipSetDevice(j.device);
  hipFunction_t fn = nullptr;
  std::string err;
  hipFunction_t fn_o = nullptr;
  hipFunction_t fn_opt[GEN_N_OPT] = {};
  const bool ok = rtc_kernel(j.src, j.device, &fn, &fn_o, err, fn_opt);
  {
    std::lock_guard<std::mutex> lk(j.mu);
    j.fn = fn;
    j.fn_o = fn_o;
    for (int k = 0; k < GEN_N_OPT; k++) j.fn_opt[k] = fn_opt[k];
    j.err = err;
    j.state.store(ok ? MM3E_CODEGEN_READY : MM3E_CODEGEN_FAILED, std::memory_order_release);
  }
  j.cv.notify_all();
}

static void gen_worker(mm3e_renderer* r) {
  for (;;) {
    std::shared_ptr<GenJob> j;
    {
      std::unique_lock<std::mutex> lk(r->gen_mu);
      r->gen_cv.wait(lk, [&] { return r->gen_stop || !r->gen_queue.empty(); });
      if (r->gen_stop) return;
      j = r->gen_queue.front();
      r->gen_queue.pop_front();
    }
    gen_run(*j);
  }
}

// Emit the scene's generated kernel and compile it (SYNC) or queue it (BACKGROUND).  Scenes
// with more than kCodegenMaxObjects objects keep the table kernels.  Measured on MI355X at
// 1280x720 (ordered tiles, the generated kernel's occupancy picked by codegen), the generated
// fold beats the table walk at every size of the batched load's eight scenes (39 / 72 / 76 / 82
// / 117 / 127 / 174 / 184 objects: 7.4 / 11.8 / 1.9 / 12.1 / 1.8 / 36.1 / 53.4 / 4.2 ms against
// 9.1 / 14.5 / 2.1 / 15.1 / 2.2 / 39.3 / 66.1 / 5.9 ms) and on the 1000-object stress frame
// (777 against 844 ms at 3840x2160), once its grid masks are per lane and its sin/cos is
// inline; the cap only bounds the compile (~45 s for 1000 objects, in a worker process).
static const size_t kCodegenMaxObjects = 2000;
size_t codegen_object_limit() {  // MM3E_CODEGEN_MAX_OBJECTS overrides (tuning)
  if (const char* v = std::getenv("MM3E_CODEGEN_MAX_OBJECTS")) {
    const long n = std::atol(v);
    if (n >= 0) return (size_t)n;
  }
  return kCodegenMaxObjects;
}
hipError_t launch_generated_index(hipFunction_t fn, const DevScene* sc, const FrameParams& fp, float* hdr, uint32_t tw,
                                  uint32_t th, uint32_t aa, uint32_t stage_bytes, hipStream_t st) {
  const uint32_t tiles_x = (fp.width + tw - 1) / tw, tiles_y = (fp.height + th - 1) / th;
  TileOrder o{};
  o.tiles_x = tiles_x;
  const DevScene* sc_arg = sc;
  FrameParams fp_arg = fp;
  float* hdr_arg = hdr;
  void* args[] = {&sc_arg, &fp_arg, &hdr_arg, &tw, &th, &stage_bytes, &o};
  return hipModuleLaunchKernel(fn, tiles_x * tiles_y, 1, 1, std::min(256u, tw * th * aa * aa), 1, 1, stage_bytes, st,
                               args, nullptr);
}

static void gen_schedule(mm3e_scene* s) {
  mm3e_renderer* r = s->r;
  if (s->host.objects.size() > codegen_object_limit()) return;
  Blob b;
  build_blob(s->host, {}, {}, s->aa, s->bounces, (uint64_t)(uintptr_t)s->d_tables, s->d_gi, false, b);
  auto j = std::make_shared<GenJob>();
  j->src = emit_scene_source(b, s->bounces, generated_prefers_sm(b), r->wf_budget > 0, r->gen_film);
  j->device = r->device;
  s->gen_stage = generated_needs_tables(b) ? s->stage_bytes : 0u;
  s->gen = j;
  if (r->codegen == MM3E_CODEGEN_SYNC) {
    gen_run(*j);
    return;
  }
  std::lock_guard<std::mutex> lk(r->gen_mu);
  if (r->gen_threads.empty()) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned n = std::max(1u, std::min(8u, hw / 2));
    for (unsigned i = 0; i < n; i++) r->gen_threads.emplace_back(gen_worker, r);
  }
  r->gen_queue.push_back(j);
  r->gen_cv.notify_one();
}

// The tile order of a scene's generated kernel, set up on its first use: the frame's tiles in
// index order, zero costs, ticket counter at zero.
static bool tile_order_ready(mm3e_scene* s, uint32_t n, std::string& msg) {
  if (s->d_to && s->to_tiles == n) return true;
  if (s->d_to) {
    hipFree(s->d_to);
    s->d_to = nullptr;
  }
  std::vector<uint32_t> init(4 * (size_t)n + 1, 0u);  // [order][cost][start][done][ticket]
  for (uint32_t i = 0; i < n; i++) init[i] = i;
  if (!hip_ok(hipMalloc((void**)&s->d_to, init.size() * 4), "hipMalloc(tile order)", &msg) ||
      !hip_ok(hipMemcpy(s->d_to, init.data(), init.size() * 4, hipMemcpyHostToDevice), "upload tile order", &msg)) {
    hipFree(s->d_to);
    s->d_to = nullptr;
    s->to_on = false;
    return false;
  }
  s->to_tiles = n;
  return true;
}

static TileOrder scene_tile_order(const mm3e_scene* s, uint32_t tiles_x) {
  const size_t n = s->to_tiles;
  uint32_t* d = s->d_to;
  return TileOrder{d, d + n, d + 4 * n, d + 2 * n, d + 3 * n, tiles_x, 0u};
}

// The queued-frame setting in force for a scene: its own, else the renderer's.
// A tile of more than 256 samples (aa > 16) renders in rounds inside one block, which the
// queued frame's one-lane-per-sample layout does not have: such scenes take the one-kernel frame.
static void wf_settings(const mm3e_scene* s, uint32_t* budget, uint32_t* passes, uint32_t* bpc) {
  *budget = s->aa * s->aa > 256u ? 0u : s->wf_own ? s->wf_set[0] : s->r->wf_budget;
  *passes = s->wf_own ? s->wf_set[1] : s->r->wf_passes;
  *bpc = s->wf_own ? s->wf_set[2] : s->r->wf_bpc;
}

// Queued-frame buffers for `cap` sample slots, one allocation:
// [mq0][mq1] WF_STATE_WORDS*cap words each, [mslot0][mslot1][hq 4*cap][hslot][hmat][rad 3*cap][cnt].
static bool wf_ready(mm3e_scene* s, uint32_t cap, std::string& msg) {
  if (s->d_wf && s->wf_cap == cap) return true;
  if (s->d_wf) hipFree(s->d_wf);
  s->d_wf = nullptr;
  s->wf_cap = 0;
  const size_t words = (size_t)cap * (2 * WF_STATE_WORDS + 2 + WF_HIT_WORDS + 2 + 3) + 16;
  if (!hip_ok(hipMalloc((void**)&s->d_wf, words * 4), "hipMalloc(queued frame)", &msg)) {
    s->d_wf = nullptr;
    return false;
  }
  s->wf_cap = cap;
  return true;
}

static WfFrame wf_frame(mm3e_scene* s, uint32_t tiles_x) {
  WfFrame w{};
  const size_t c = s->wf_cap;
  float* f = (float*)s->d_wf;
  w.mq[0] = f;
  w.mq[1] = f + WF_STATE_WORDS * c;
  uint32_t* u = (uint32_t*)(f + 2 * WF_STATE_WORDS * c);
  w.mslot[0] = u;
  w.mslot[1] = u + c;
  w.hq = (float*)(u + 2 * c);
  w.hslot = (uint32_t*)(w.hq + WF_HIT_WORDS * c);
  w.hmat = w.hslot + c;
  w.rad = (float*)(w.hmat + c);
  w.cnt = (uint32_t*)(w.rad + 3 * c);
  w.cap = s->wf_cap;
  w.tw = s->tw;
  w.th = s->th;
  w.tiles_x = tiles_x;
  return w;
}

// A frame as queued stages (dev_wavefront.h) into s->d_hdr: budgeted march passes over
// compacted continuation queues, shading of the compacted hit queue, per-pixel resolve.
static bool launch_queued(mm3e_scene* s, const DevScene* dsc, const FrameParams& fp, dim3 grid, dim3 block,
                          uint32_t stage, bool general, std::string& msg) {
  mm3e_renderer* r = s->r;
  hipStream_t st = r->stream;
  // sample slots are indexed in 32 bits (the queues hold at most every slot once)
  const uint64_t slots = (uint64_t)grid.x * grid.y * block.x;
  if (slots > 0xFFFFFF00ull) {
    msg = "queued frame: " + std::to_string(slots) + " sample slots exceed the 32-bit queue index";
    return false;
  }
  const uint32_t cap = (uint32_t)slots;
  if (!wf_ready(s, cap, msg)) return false;
  WfFrame w = wf_frame(s, grid.x);
  uint32_t budget, passes, bpc;
  wf_settings(s, &budget, &passes, &bpc);
  // grid-stride stages: cus * bpc blocks of 256, or enough for one lane per slot
  const uint32_t full = (cap + 255) / 256;
  const uint32_t gs = bpc ? std::min(full, r->cus * bpc) : full;
  const mm3e::GenJob* gj = s->gen.get();
  const bool gen = dsc == s->d_scene && gj && gj->state.load(std::memory_order_acquire) == MM3E_CODEGEN_READY &&
                   gj->fn_opt[GEN_WF_FIRST] && gj->fn_opt[GEN_WF_CONT] && gj->fn_opt[GEN_WF_SHADE];
  uint32_t gstage = gen ? s->gen_stage : stage;
  const DevScene* sc_arg = dsc;
  FrameParams fp_arg = fp;
  if (!hip_ok(launch_wf_reset(st, w.cnt), "queued frame reset", &msg)) return false;
  for (uint32_t k = 0; k <= passes; k++) {
    uint32_t b = k == passes ? 0u : budget;  // the last pass runs every march to its end
    uint32_t kk = k;
    const dim3 g = k == 0 ? dim3(grid.x * grid.y) : dim3(gs);
    const dim3 bl = k == 0 ? block : dim3(256);
    hipError_t e;
    if (gen) {
      if (k == 0) {
        void* args[] = {&sc_arg, &fp_arg, &w, &b, &gstage};
        e = hipModuleLaunchKernel(gj->fn_opt[GEN_WF_FIRST], g.x, 1, 1, bl.x, 1, 1, gstage, st, args, nullptr);
      } else {
        void* args[] = {&sc_arg, &fp_arg, &w, &kk, &b, &gstage};
        e = hipModuleLaunchKernel(gj->fn_opt[GEN_WF_CONT], g.x, 1, 1, bl.x, 1, 1, gstage, st, args, nullptr);
      }
    } else {
      e = launch_wf_march(st, general, k, g, bl, dsc, stage, fp, w, b);
    }
    if (!hip_ok(e, "queued march launch", &msg)) return false;
  }
  hipError_t e;
  if (gen) {
    void* args[] = {&sc_arg, &fp_arg, &w, &gstage};
    e = hipModuleLaunchKernel(gj->fn_opt[GEN_WF_SHADE], gs, 1, 1, 256, 1, 1, gstage, st, args, nullptr);
  } else {
    e = launch_wf_shade(s->bounces, dim3(gs), st, dsc, stage, general, fp, w);
  }
  if (!hip_ok(e, "queued shade launch", &msg)) return false;
  return hip_ok(launch_wf_resolve(st, fp, w, s->aa, s->d_hdr), "queued resolve launch", &msg);
}

// Where a frame's kernels write (HDR radiance, bloom intermediate, RGBA8) when not into the
// scene's own framebuffers.  Such frames take neither the tile-order nor the queued path, which
// use per-scene buffers allocated on first use: one kernel per frame in tile index order.
struct FrameTargets {
  float* hdr;
  float* tmp;
  uchar4* rgba;
};

bool launch_frame(mm3e_scene* s, const DevScene* dsc, const FrameParams& fp, std::string& msg,
                  const FrameTargets* tg = nullptr) {
  float* const d_hdr = tg ? tg->hdr : s->d_hdr;
  float* const d_tmp = tg ? tg->tmp : s->d_tmp;
  uchar4* const d_rgba = tg ? tg->rgba : s->d_rgba;
  const bool to_on = s->to_on && !tg;
  hipStream_t st = s->r->stream;
  const uint32_t stage = dsc == s->d_scene ? s->stage_bytes : s->dyn_stage_bytes;
  const bool general = dsc == s->d_scene ? s->general : s->dyn_general;
  dim3 grid((fp.width + s->tw - 1) / s->tw, (fp.height + s->th - 1) / s->th);
  dim3 block(std::min(256u, s->tw * s->th * s->aa * s->aa));  // aa > 16: one pixel, samples in rounds
  hipEvent_t* ev = s->ev[s->frames % mm3e_scene::kRing];
  hipEventRecord(ev[0], st);
  if (s->host.hdr.mode != 0) {  // debug render mode: straight to RGBA8, no post (lib.rs:873-878)
    if (!hip_ok(launch_aov_mode_table(st, dsc, fp, d_rgba, s->host.hdr.mode), "render-mode launch", &msg))
      return false;
    hipEventRecord(ev[1], st);
    hipEventRecord(ev[2], st);
    s->frames++;
    s->timing_valid = true;
    return true;
  }
  const mm3e::GenJob* gj = s->gen.get();
  bool ordered = false;
  uint32_t wf_b, wf_n, wf_c;
  wf_settings(s, &wf_b, &wf_n, &wf_c);
  if (tg) wf_b = 0;
  if (wf_b > 0) {
    if (!launch_queued(s, dsc, fp, grid, block, stage, general, msg)) return false;
  } else if (dsc == s->d_scene && gj && gj->state.load(std::memory_order_acquire) == MM3E_CODEGEN_READY) {
    const DevScene* sc_arg = dsc;
    FrameParams fp_arg = fp;
    float* hdr_arg = d_hdr;
    uint32_t tw = s->tw, th = s->th, gstage = s->gen_stage;
    if (gj->fn_o && to_on && tile_order_ready(s, grid.x * grid.y, msg)) {
      TileOrder o = scene_tile_order(s, grid.x);
      void* args[] = {&sc_arg, &fp_arg, &hdr_arg, &tw, &th, &gstage, &o};
      if (!hip_ok(hipModuleLaunchKernel(gj->fn_o, s->to_tiles, 1, 1, block.x, 1, 1, gstage, st, args, nullptr),
                  "generated render launch", &msg))
        return false;
      ordered = true;
    } else if (!hip_ok(launch_generated_index(gj->fn, sc_arg, fp_arg, hdr_arg, tw, th, s->aa, gstage, st),
                       "generated render launch", &msg)) {
      return false;
    }
  } else if (dsc == s->d_scene && to_on && tile_order_ready(s, grid.x * grid.y, msg)) {
    TileOrder o = scene_tile_order(s, grid.x);
    if (!hip_ok(launch_render_ordered(s->bounces, s->to_tiles, block, st, dsc, stage, general, fp, d_hdr, s->tw,
                                      s->th, o),
                "render launch", &msg))
      return false;
    ordered = true;
  } else if (!hip_ok(launch_render_table(s->bounces, grid, block, st, dsc, stage, general, fp, d_hdr, s->tw, s->th),
                     "render launch", &msg)) {
    return false;
  }
  hipEventRecord(ev[1], st);
  if (!hip_ok(launch_post(st, s->pp, s->d_kern, d_hdr, d_tmp, d_rgba), "post launch", &msg)) return false;
  hipEventRecord(ev[2], st);
  if (ordered && !hip_ok(launch_tile_order(st, s->d_to + s->to_tiles, s->to_tiles, s->d_to), "tile order launch", &msg))
    return false;
  s->frames++;
  s->timing_valid = true;
  return true;
}

}  // namespace mm3e

using namespace mm3e;

// ---- frames that allocate nothing -------------------------------------------------------
// A fixed device array holds the HDR radiance and the bloom intermediate of frames rendered
// by mm3e_scene_render_rgba_noalloc (12 + 12 bytes per pixel).  Each GPU the library runs on
// holds its own copy until the process ends; in host code `scratch` is only a handle for
// hipGetSymbolAddress, which yields the address of the copy of the current device.
static constexpr size_t kNoallocScratchBytes = (size_t)MM3E_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Binds the calling thread's current device's copy of `scratch` (once per device).
static bool bind_scratch(int dev, std::string& msg) {
  if (dev < 0 || dev >= kMaxDevices) {
    msg = "device index out of range for the frame scratch";
    return false;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return true;
  void* ptr = nullptr;
  if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)", &msg)) return false;
  if (!ptr) {
    msg = "hipGetSymbolAddress(scratch) returned no address";
    return false;
  }
  g_scratch_ptr[dev] = ptr;
  return true;
}
static void* scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return dev >= 0 && dev < kMaxDevices ? g_scratch_ptr[dev] : nullptr;
}

extern "C" {

// Queued-frame default from MM3E_WAVEFRONT="budget[,passes[,blocks_per_cu]]"; out of range or
// unset leaves the arguments unchanged.
static void env_wavefront(uint32_t* budget, uint32_t* passes, uint32_t* bpc) {
  const char* wf = std::getenv("MM3E_WAVEFRONT");
  if (!wf) return;
  unsigned b = 0, n = *passes, c = *bpc;
  const int got = std::sscanf(wf, "%u,%u,%u", &b, &n, &c);
  if (got >= 1 && n < WF_MAX_PASSES && c <= 64) {
    *budget = b;
    *passes = n;
    *bpc = c;
  }
}

// Default of mm3e_renderer_set_codegen_film: MM3E_CODEGEN_FILM=1 / on.
static bool env_codegen_film() {
  const char* v = std::getenv("MM3E_CODEGEN_FILM");
  return v && (std::strcmp(v, "1") == 0 || lower(v) == "on");
}

int mm3e_list_adapters(char* buf, size_t buflen) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  std::string all;
  for (int i = 0; i < n; i++) all += (i ? " | " : "") + describe(i);
  set_err(buf, buflen, all);
  return n;
}

int mm3e_renderer_create(const char* adapter_filter, mm3e_renderer** out, char* err, size_t errlen) {
  if (!out) return MM3E_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    set_err(err, errlen, "no GPU adapter: no HIP device is visible");
    return MM3E_ERR_DEVICE;
  }
  std::string filter;
  if (adapter_filter) filter = adapter_filter;
  else if (const char* env = std::getenv("MM3E_GPU_ADAPTER")) filter = env;  // lib.rs:81-84
  while (!filter.empty() && std::isspace((unsigned char)filter.back())) filter.pop_back();
  while (!filter.empty() && std::isspace((unsigned char)filter.front())) filter.erase(filter.begin());
  int dev = -1;
  if (!filter.empty()) {
    std::string want = lower(filter), names;
    for (int i = 0; i < n; i++) {
      std::string d = describe(i);
      names += (i ? " | " : "") + d;
      if (dev < 0 && lower(d).find(want) != std::string::npos) dev = i;
    }
    if (dev < 0) {  // lib.rs:100
      set_err(err, errlen, "no GPU adapter matching '" + filter + "'; available: " + names);
      return MM3E_ERR_INVALID;
    }
  } else {
    dev = 0;
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
    set_err(err, errlen, "request_device failed: hipGetDeviceProperties");
    return MM3E_ERR_DEVICE;
  }
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    set_err(err, errlen, std::string("request_device failed: libmm3e_hip is built for gfx950, device is ") + p.gcnArchName);
    return MM3E_ERR_DEVICE;
  }
  auto* r = new mm3e_renderer();
  r->device = dev;
  if (const char* cg = std::getenv("MM3E_CODEGEN")) {
    const std::string m = lower(cg);
    if (m == "off" || m == "0") r->codegen = MM3E_CODEGEN_OFF;
    else if (m == "sync" || m == "2") r->codegen = MM3E_CODEGEN_SYNC;
    else r->codegen = MM3E_CODEGEN_BACKGROUND;
  }
  if (const char* to = std::getenv("MM3E_TILE_ORDER")) r->tile_order = !(lower(to) == "off" || lower(to) == "0");
  env_wavefront(&r->wf_budget, &r->wf_passes, &r->wf_bpc);
  r->gen_film = env_codegen_film();
  r->cus = (uint32_t)std::max(1, p.multiProcessorCount);
  r->name = describe(dev);
  // Tables staged in LDS share the block's LDS with the 3 KiB sample-reduction buffer.  The
  // table-walk kernels run 4 waves per SIMD (4 blocks of 256 lanes per CU); a larger copy would
  // lower that occupancy, which costs more than reading the (divergent) tree walk's nodes from
  // L2 (a 174-object scene at 1280x720: 74 ms staged at 45 KiB, 65 ms from HBM).  Wave-uniform
  // records come from HBM through scalar loads either way.
  r->lds_budget = (uint32_t)std::min<size_t>(36 * 1024, p.sharedMemPerBlock > 4096 ? p.sharedMemPerBlock - 4096 : 0);
  std::string msg;
  if (!hip_ok(hipSetDevice(dev), "hipSetDevice", &msg) || !hip_ok(hipStreamCreate(&r->stream), "hipStreamCreate", &msg) ||
      !bind_scratch(dev, msg)) {
    set_err(err, errlen, "request_device failed: " + msg);
    delete r;
    return MM3E_ERR_DEVICE;
  }
  *out = r;
  return MM3E_OK;
}

const char* mm3e_renderer_adapter_name(const mm3e_renderer* r) { return r ? r->name.c_str() : ""; }

int mm3e_renderer_wait_idle(mm3e_renderer* r) {
  if (!r) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(r->mu);
  hipSetDevice(r->device);
  return hipStreamSynchronize(r->stream) == hipSuccess ? MM3E_OK : MM3E_ERR_DEVICE;
}

void mm3e_renderer_destroy(mm3e_renderer* r) {
  if (!r) return;
  {
    std::lock_guard<std::mutex> lk(r->gen_mu);
    r->gen_stop = true;
  }
  r->gen_cv.notify_all();
  for (auto& t : r->gen_threads) t.join();  // a compile in progress finishes first
  hipSetDevice(r->device);
  hipStreamSynchronize(r->stream);
  hipStreamDestroy(r->stream);
  delete r;
}

int mm3e_scene_compile(mm3e_renderer* r, const void* ir, size_t ir_len, uint32_t w, uint32_t h, mm3e_scene** out,
                       char* err, size_t errlen) {
  if (!r || !out) return MM3E_ERR_INVALID;
  *out = nullptr;
  if (w == 0 || h == 0 || w > 16384 || h > 16384) {  // lib.rs:139-141
    set_err(err, errlen, "GpuRenderer::compile: " + std::to_string(w) + "x" + std::to_string(h) + " out of range 1..=16384");
    return MM3E_ERR_INVALID;
  }
  auto* s = new mm3e_scene();
  std::string msg;
  if (!parse_scene_ir(ir, ir_len, s->host, msg)) {
    set_err(err, errlen, msg);
    delete s;
    return MM3E_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(r->mu);
  hipSetDevice(r->device);
  s->r = r;
  s->w = w;
  s->h = h;
  s->aa = std::max(s->host.hdr.aa, 1u);  // lib.rs:282
  s->to_on = r->tile_order;
  s->bounces = s->host.hdr.bounces;
  tile_shape(s->aa, &s->tw, &s->th);
  bool ok = true;
  if (s->host.hdr.gi_mode == MM3E_GI_BAKED) {
    // the baked cubes travel inside the table blob; the header points at them
    s->has_gi = true;
    for (int a = 0; a < 3; a++) {
      s->gi_dims[a] = s->host.hdr.gi_dims[a];
      s->gi_min[a] = s->host.hdr.gi_min[a];
      s->gi_size[a] = s->host.hdr.gi_size[a];
    }
  }
  // tables first (the GI bake marches this scene, without GI)
  DevScene hdr;
  if (s->host.hdr.gi_mode == MM3E_GI_BAKED) {
    // point d_gi at the in-blob array after upload
    const uint64_t need = blob_size(s->host, {}, {});
    ok = hip_ok(hipMalloc((void**)&s->d_tables, need), "hipMalloc(scene tables)", &msg);
    if (ok) {
      s->tables_cap = need;
      Blob b;
      build_blob(s->host, {}, {}, s->aa, s->bounces, (uint64_t)(uintptr_t)s->d_tables, nullptr, false, b);
      s->d_gi_in_blob = (float*)(s->d_tables + b.gi_offset);
      s->d_gi = s->d_gi_in_blob;
      patch_gi(s, b.header);
      std::memcpy(b.bytes.data(), &b.header, sizeof b.header);
      ok = hip_ok(hipMemcpy(s->d_tables, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice), "upload tables", &msg);
      hdr = b.header;
      s->stage_bytes = b.header.stage_bytes <= r->lds_budget ? b.header.stage_bytes : 0;
      s->general = scene_is_general(b.header);
    }
  } else {
    ok = upload_tables(s, {}, {}, &s->d_tables, &s->tables_cap, &hdr, msg);
    if (ok && s->host.hdr.gi_mode == MM3E_GI_BAKE) {
      s->d_scene = reinterpret_cast<const DevScene*>(s->d_tables);
      ok = bake_gi_device(s, msg);
      if (ok && s->has_gi) ok = upload_tables(s, {}, {}, &s->d_tables, &s->tables_cap, &hdr, msg);
    }
  }
  s->d_scene = reinterpret_cast<const DevScene*>(s->d_tables);
  s->hdr = hdr;
  // post: bloom kernel weights + params (post.rs:36-58)
  const mm3e_ir_header& ih = s->host.hdr;
  std::vector<float> kern = gaussian_kernel(ih.bloom_radius);
  s->pp.width = w;
  s->pp.height = h;
  s->pp.radius = (uint32_t)(kern.size() / 2);
  s->pp.bloom = (ih.bloom && ih.bloom_intensity > 0.0f) ? 1u : 0u;
  s->pp.threshold = ih.bloom_threshold;
  s->pp.intensity = ih.bloom_intensity;
  s->pp.exposure = ih.exposure;
  const size_t npx = (size_t)w * h;
  if (ok) ok = hip_ok(hipMalloc(&s->d_kern, kern.size() * 4), "hipMalloc(kernel)", &msg);
  if (ok) ok = hip_ok(hipMemcpy(s->d_kern, kern.data(), kern.size() * 4, hipMemcpyHostToDevice), "upload kernel", &msg);
  if (ok) ok = hip_ok(hipMalloc(&s->d_hdr, npx * 12), "hipMalloc(hdr)", &msg);
  if (ok) ok = hip_ok(hipMalloc(&s->d_tmp, npx * 12), "hipMalloc(bloom)", &msg);
  if (ok) ok = hip_ok(hipMalloc(&s->d_rgba, npx * 4), "hipMalloc(rgba)", &msg);
  if (ok) ok = hip_ok(hipHostMalloc((void**)&s->h_rgba, npx * 4, hipHostMallocDefault), "hipHostMalloc(rgba)", &msg);
  for (int f = 0; ok && f < mm3e_scene::kRing; f++)
    for (int i = 0; ok && i < 3; i++) ok = hip_ok(hipEventCreate(&s->ev[f][i]), "hipEventCreate", &msg);
  if (!ok) {
    set_err(err, errlen, msg);
    mm3e_scene_destroy(s);
    return MM3E_ERR_DEVICE;
  }
  if (r->codegen != MM3E_CODEGEN_OFF) gen_schedule(s);
  *out = s;
  return MM3E_OK;
}

int mm3e_scene_check(const void* ir, size_t ir_len, uint32_t w, uint32_t h, char* err, size_t errlen) {
  if (w == 0 || h == 0 || w > 16384 || h > 16384) {  // lib.rs:139-141
    set_err(err, errlen, "GpuRenderer::compile: " + std::to_string(w) + "x" + std::to_string(h) + " out of range 1..=16384");
    return MM3E_ERR_INVALID;
  }
  HostScene hs;
  std::string msg;
  if (!parse_scene_ir(ir, ir_len, hs, msg)) {
    set_err(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  return MM3E_OK;
}

void mm3e_scene_destroy(mm3e_scene* s) {
  if (!s) return;
  if (s->r) {
    hipSetDevice(s->r->device);
    hipStreamSynchronize(s->r->stream);
  }
  if (s->d_tables) hipFree(s->d_tables);
  if (s->d_dyn) hipFree(s->d_dyn);
  if (s->d_gi && s->d_gi != s->d_gi_in_blob) hipFree(s->d_gi);
  if (s->d_kern) hipFree(s->d_kern);
  if (s->d_hdr) hipFree(s->d_hdr);
  if (s->d_tmp) hipFree(s->d_tmp);
  if (s->d_rgba) hipFree(s->d_rgba);
  if (s->d_to) hipFree(s->d_to);
  if (s->d_wf) hipFree(s->d_wf);
  if (s->h_rgba) hipHostFree(s->h_rgba);
  for (int f = 0; f < mm3e_scene::kRing; f++)
    for (int i = 0; i < 3; i++)
      if (s->ev[f][i]) hipEventDestroy(s->ev[f][i]);
  delete s;
}

static int validate_cam(const mm3e_camera* c) { return c ? MM3E_OK : MM3E_ERR_INVALID; }

// Dynamic spheres as appended objects with Material{albedo, metallic, ..default}
// (mm3e-cpu-parallel/src/lib.rs:50-84); at most MM3E_MAX_DYN (mm3e-gpu/src/lib.rs:306-307).
static bool dyn_scene(mm3e_scene* s, const mm3e_dyn_sphere* dyn, uint32_t n_dyn, const DevScene** out,
                      std::string& msg) {
  n_dyn = std::min<uint32_t>(n_dyn, MM3E_MAX_DYN);
  std::vector<mm3e_ir_object> xo;
  std::vector<mm3e_ir_material> xm;
  const uint32_t base_mat = (uint32_t)s->host.materials.size();
  for (uint32_t i = 0; i < n_dyn; i++) {
    const mm3e_dyn_sphere& d = dyn[i];
    char buf[160];
    if (!(std::isfinite(d.pos[0]) && std::isfinite(d.pos[1]) && std::isfinite(d.pos[2]))) {
      std::snprintf(buf, sizeof buf, "dynamic sphere %u: position must be finite", i);
      msg = buf;
      return false;
    }
    if (!std::isfinite(d.radius) || d.radius <= 0.0f) {
      std::snprintf(buf, sizeof buf, "dynamic sphere %u: radius must be finite and positive", i);
      msg = buf;
      return false;
    }
    const float v[4] = {d.albedo[0], d.albedo[1], d.albedo[2], d.metallic};
    for (float x : v)
      if (!std::isfinite(x) || x < 0.0f || x > 1.0f) {
        std::snprintf(buf, sizeof buf, "dynamic sphere %u: linear albedo and metallic must be finite and in [0,1]", i);
        msg = buf;
        return false;
      }
    mm3e_ir_object o;
    std::memset(&o, 0, sizeof o);
    o.prim = MM3E_PRIM_SPHERE;
    o.mat = base_mat + i;
    o.combine = MM3E_COMBINE_UNION;
    o.p[0] = d.radius;
    o.rot[0] = o.rot[4] = o.rot[8] = 1.0f;
    std::memcpy(o.pos, d.pos, 12);
    o.scale = 1.0f;
    xo.push_back(o);
    mm3e_ir_material m;
    std::memset(&m, 0, sizeof m);
    std::memcpy(m.albedo, d.albedo, 12);
    m.metallic = d.metallic;
    m.specular = 0.5f;   // color.rs:64-75 Material::default
    m.roughness = 0.35f;
    xm.push_back(m);
  }
  if (!upload_tables(s, xo, xm, &s->d_dyn, &s->dyn_cap, nullptr, msg)) return false;
  *out = reinterpret_cast<const DevScene*>(s->d_dyn);
  return true;
}

int mm3e_scene_render_rgba_noalloc(mm3e_scene* s, const mm3e_camera* cam, void* rgba_dev, size_t rgba_len,
                                   char* err, size_t errlen) {
  if (!s || validate_cam(cam) || !rgba_dev) {
    set_err(err, errlen, "render_rgba_noalloc: null scene, camera or output, or a non-finite camera");
    return MM3E_ERR_INVALID;
  }
  const uint64_t npx = (uint64_t)s->w * s->h;
  if (rgba_len < npx * 4) {
    set_err(err, errlen, "render_rgba_noalloc: output buffer smaller than width*height*4 bytes");
    return MM3E_ERR_INVALID;
  }
  const uint64_t hdr_bytes = (npx * 12 + 255) & ~uint64_t(255);  // the bloom intermediate starts 256-aligned
  if (hdr_bytes + npx * 12 > kNoallocScratchBytes) {
    set_err(err, errlen, "render_rgba_noalloc: " + std::to_string(s->w) + "x" + std::to_string(s->h) +
                             " needs " + std::to_string(npx * 24) + " bytes of frame scratch; the maximum is " +
                             std::to_string(kNoallocScratchBytes) + " (" +
                             std::to_string(kNoallocScratchBytes / 2 / 12) + " pixels)");
    return MM3E_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  uint8_t* base = static_cast<uint8_t*>(scratch_of(s->r->device));
  if (!base) {
    set_err(err, errlen, "render_rgba_noalloc: the frame scratch is not bound on this device");
    return MM3E_ERR_DEVICE;
  }
  FrameTargets tg{reinterpret_cast<float*>(base), reinterpret_cast<float*>(base + hdr_bytes),
                  static_cast<uchar4*>(rgba_dev)};
  std::string msg;
  if (!launch_frame(s, s->d_scene, make_frame(*cam, s->w, s->h), msg, &tg)) {
    set_err(err, errlen, "render_rgba_noalloc: " + msg);
    return MM3E_ERR_DEVICE;
  }
  return MM3E_OK;
}

int mm3e_scene_render_only(mm3e_scene* s, const mm3e_camera* cam) {
  if (!s || validate_cam(cam)) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  std::string msg;
  return launch_frame(s, s->d_scene, make_frame(*cam, s->w, s->h), msg) ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_scene_render_rgba(mm3e_scene* s, const mm3e_camera* cam, const mm3e_dyn_sphere* dyn, uint32_t n_dyn,
                           uint8_t* out, size_t out_len) {
  if (!s || validate_cam(cam) || (n_dyn && !dyn)) return MM3E_ERR_INVALID;
  const size_t need = (size_t)s->w * s->h * 4;
  if (out && out_len < need) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  std::string msg;
  const DevScene* dsc = s->d_scene;
  if (n_dyn && !dyn_scene(s, dyn, n_dyn, &dsc, msg)) return MM3E_ERR_INVALID;
  if (!launch_frame(s, dsc, make_frame(*cam, s->w, s->h), msg)) return MM3E_ERR_DEVICE;
  if (!hip_ok(hipMemcpyAsync(s->h_rgba, s->d_rgba, need, hipMemcpyDeviceToHost, s->r->stream), "readback", &msg))
    return MM3E_ERR_DEVICE;
  if (!hip_ok(hipStreamSynchronize(s->r->stream), "frame", &msg)) return MM3E_ERR_DEVICE;
  if (out) std::memcpy(out, s->h_rgba, need);
  return MM3E_OK;
}

int mm3e_scene_render_linear(mm3e_scene* s, const mm3e_camera* cam, float* out, size_t out_floats) {
  if (!s || validate_cam(cam) || !out) return MM3E_ERR_INVALID;
  const size_t need = (size_t)s->w * s->h * 3;
  if (out_floats < need) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  std::string msg;
  if (!launch_frame(s, s->d_scene, make_frame(*cam, s->w, s->h), msg)) return MM3E_ERR_DEVICE;
  if (!hip_ok(hipMemcpyAsync(out, s->d_hdr, need * 4, hipMemcpyDeviceToHost, s->r->stream), "readback", &msg))
    return MM3E_ERR_DEVICE;
  return hip_ok(hipStreamSynchronize(s->r->stream), "frame", &msg) ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_scene_info(const mm3e_scene* s, uint32_t* n_objects, uint32_t* n_segments, uint32_t* grid_cells,
                    uint32_t* grid_words, uint32_t* stage_bytes) {
  if (!s) return MM3E_ERR_INVALID;
  const mm3e::DevScene& d = s->hdr;
  if (n_objects) *n_objects = d.n_obj;
  if (n_segments) *n_segments = d.n_seg;
  if (grid_cells) *grid_cells = d.grid ? d.grid_dims[0] * d.grid_dims[1] * d.grid_dims[2] : 0u;
  if (grid_words) *grid_words = d.grid ? d.grid_words : 0u;
  if (stage_bytes) *stage_bytes = s->stage_bytes;
  return MM3E_OK;
}

int mm3e_scene_kernel_name(const mm3e_scene* s, char* buf, size_t buflen) {
  if (!s || !buf || buflen == 0) return MM3E_ERR_INVALID;
  const bool gen_ready = s->gen && s->gen->state.load(std::memory_order_acquire) == MM3E_CODEGEN_READY;
  uint32_t wf_b, wf_n, wf_c;
  wf_settings(s, &wf_b, &wf_n, &wf_c);
  if (wf_b > 0) {  // queued frame: name its shading stage
    if (gen_ready && s->gen->fn_opt[GEN_WF_SHADE]) std::snprintf(buf, buflen, "mm3e_gen_wf_shade");
    else std::snprintf(buf, buflen, "%s%u", s->general ? "mm3e_wf_shade_x_b" : "mm3e_wf_shade_b", s->bounces);
    return MM3E_OK;
  }
  if (gen_ready) {  // one kernel for both dispatch orders (tile order or index order)
    std::snprintf(buf, buflen, "mm3e_gen_render_o");
    return MM3E_OK;
  }
  const char* base = s->to_on ? (s->general ? "mm3e_render_ord_x_b" : "mm3e_render_ord_b")
                   : s->general ? (s->stage_bytes ? "mm3e_render_lds_x_b" : "mm3e_render_table_x_b")
                                : (s->stage_bytes ? "mm3e_render_lds_b" : "mm3e_render_table_b");
  std::snprintf(buf, buflen, "%s%u", base, s->bounces);
  return MM3E_OK;
}

int mm3e_scene_last_timing(mm3e_scene* s, float* render_ms, float* post_ms) {
  if (!s || !s->timing_valid || s->frames == 0) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  hipEvent_t* ev = s->ev[(s->frames - 1) % mm3e_scene::kRing];
  if (hipEventSynchronize(ev[2]) != hipSuccess) return MM3E_ERR_DEVICE;
  float a = 0, b = 0;
  hipEventElapsedTime(&a, ev[0], ev[1]);
  hipEventElapsedTime(&b, ev[1], ev[2]);
  if (render_ms) *render_ms = a;
  if (post_ms) *post_ms = b;
  return MM3E_OK;
}

int mm3e_scene_timing(mm3e_scene* s, int reset, uint64_t* frames, double* render_ms_sum, double* post_ms_sum) {
  if (!s) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  const uint64_t n = std::min<uint64_t>(s->frames, mm3e_scene::kRing);
  double ra = 0.0, pa = 0.0;
  if (n) {
    if (hipStreamSynchronize(s->r->stream) != hipSuccess) return MM3E_ERR_DEVICE;
    for (uint64_t k = s->frames - n; k < s->frames; k++) {
      hipEvent_t* ev = s->ev[k % mm3e_scene::kRing];
      float a = 0, b = 0;
      if (hipEventElapsedTime(&a, ev[0], ev[1]) != hipSuccess) return MM3E_ERR_DEVICE;
      if (hipEventElapsedTime(&b, ev[1], ev[2]) != hipSuccess) return MM3E_ERR_DEVICE;
      ra += a;
      pa += b;
    }
  }
  if (frames) *frames = n;
  if (render_ms_sum) *render_ms_sum = ra;
  if (post_ms_sum) *post_ms_sum = pa;
  if (reset) s->frames = 0;
  return MM3E_OK;
}

int mm3e_scene_render_aov(mm3e_scene* s, const mm3e_camera* cam, uint8_t* hit, uint32_t* mat, float* depth,
                          uint32_t* steps) {
  if (!s || validate_cam(cam)) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  const size_t n = (size_t)s->w * s->h;
  uint8_t* d_hit = nullptr;
  uint32_t *d_mat = nullptr, *d_steps = nullptr;
  float* d_depth = nullptr;
  std::string msg;
  bool ok = hip_ok(hipMalloc(&d_hit, n), "hipMalloc", &msg) && hip_ok(hipMalloc(&d_mat, n * 4), "hipMalloc", &msg) &&
            hip_ok(hipMalloc(&d_depth, n * 4), "hipMalloc", &msg) && hip_ok(hipMalloc(&d_steps, n * 4), "hipMalloc", &msg);
  FrameParams fp = make_frame(*cam, s->w, s->h);
  if (ok) ok = hip_ok(launch_primary_aov_table(s->r->stream, s->d_scene, fp, d_hit, d_mat, d_depth, d_steps), "aov", &msg);
  if (ok) ok = hip_ok(hipStreamSynchronize(s->r->stream), "aov", &msg);
  if (ok && hit) ok = hip_ok(hipMemcpy(hit, d_hit, n, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok && mat) ok = hip_ok(hipMemcpy(mat, d_mat, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok && depth) ok = hip_ok(hipMemcpy(depth, d_depth, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok && steps) ok = hip_ok(hipMemcpy(steps, d_steps, n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  hipFree(d_hit);
  hipFree(d_mat);
  hipFree(d_depth);
  hipFree(d_steps);
  return ok ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_scene_render_film(mm3e_scene* s, const mm3e_camera* cam, const mm3e_lens* lens, float* beauty,
                           float* foreground, float* alpha, float* depth, float* normals, uint32_t* material_ids,
                           char* err, size_t errlen) {
  if (!s || validate_cam(cam) || !lens) return MM3E_ERR_INVALID;
  // Lens::validate (camera.rs:25-34)
  if (!std::isfinite(lens->aperture_radius) || lens->aperture_radius < 0.0f) {
    set_err(err, errlen, "aperture_radius must be finite and nonnegative");
    return MM3E_ERR_INVALID;
  }
  if (!std::isfinite(lens->focus_distance) || lens->focus_distance <= 0.0f) {
    set_err(err, errlen, "focus_distance must be finite and positive");
    return MM3E_ERR_INVALID;
  }
  if (s->aa > 16) {  // film_tile keeps one lane per sample: at most 256 samples per pixel
    set_err(err, errlen, "render_film supports aa <= 16 on the MI355X render path");
    return MM3E_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  const size_t n = (size_t)s->w * s->h;
  hipStream_t st = s->r->stream;
  // one allocation: beauty[3n], foreground[3n], alpha[n], depth[n], normals[3n], ids[n], hit[n]
  uint8_t* d = nullptr;
  std::string msg;
  bool ok = hip_ok(hipMalloc(&d, n * (4 * 12 + 1)), "hipMalloc(film)", &msg);
  float* d_beauty = (float*)d;
  float* d_fg = d_beauty + 3 * n;
  float* d_alpha = d_fg + 3 * n;
  float* d_depth = d_alpha + n;
  float* d_nrm = d_depth + n;
  uint32_t* d_ids = (uint32_t*)(d_nrm + 3 * n);
  uint8_t* d_hit = (uint8_t*)(d_ids + n);
  FrameParams fp = make_frame(*cam, s->w, s->h);
  FilmParams lp{lens->aperture_radius, lens->focus_distance, 0u, 0u};
  dim3 grid((s->w + s->tw - 1) / s->tw, (s->h + s->th - 1) / s->th);
  dim3 block(s->tw * s->th * s->aa * s->aa);
  if (ok) {
    hipError_t e;
    if (hipFunction_t fn = gen_opt(s, GEN_FILM)) {
      const DevScene* sc = s->d_scene;
      uint32_t tw = s->tw, th = s->th;
      void* args[] = {&sc, &fp, &lp, &d_beauty, &d_fg, &d_alpha, &tw, &th};
      e = hipModuleLaunchKernel(fn, grid.x, grid.y, 1, block.x, 1, 1, 0, st, args, nullptr);
    } else {
      e = launch_film(s->bounces, grid, block, st, s->d_scene, s->general, fp, lp, d_beauty, d_fg, d_alpha, s->tw,
                      s->th);
    }
    ok = hip_ok(e, "film launch", &msg);
  }
  if (ok) ok = hip_ok(launch_primary_aov_table(st, s->d_scene, fp, d_hit, d_ids, d_depth, nullptr, d_nrm), "aov", &msg);
  if (ok) ok = hip_ok(hipStreamSynchronize(st), "film", &msg);
  struct Out {
    void* host;
    const void* dev;
    size_t bytes;
  } outs[] = {{beauty, d_beauty, n * 12}, {foreground, d_fg, n * 12}, {alpha, d_alpha, n * 4},
              {depth, d_depth, n * 4},   {normals, d_nrm, n * 12},   {material_ids, d_ids, n * 4}};
  for (const Out& o : outs)
    if (ok && o.host) ok = hip_ok(hipMemcpy(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost), "readback", &msg);
  hipFree(d);
  if (!ok) {
    set_err(err, errlen, msg);
    return MM3E_ERR_DEVICE;
  }
  return MM3E_OK;
}

int mm3e_scene_gi_raw(mm3e_scene* s, float* min3, float* size3, uint32_t* dims3, float* cubes, uint64_t* n_probes) {
  if (!s) return MM3E_ERR_INVALID;
  if (!s->has_gi) {
    if (n_probes) *n_probes = 0;
    return MM3E_OK;
  }
  const uint64_t n = (uint64_t)s->gi_dims[0] * s->gi_dims[1] * s->gi_dims[2];
  if (n_probes) *n_probes = n;
  if (min3) std::memcpy(min3, s->gi_min, 12);
  if (size3) std::memcpy(size3, s->gi_size, 12);
  if (dims3) std::memcpy(dims3, s->gi_dims, 12);
  if (cubes) {
    std::lock_guard<std::mutex> lk(s->r->mu);
    hipSetDevice(s->r->device);
    std::vector<float> g4(n * 24);
    if (hipMemcpy(g4.data(), s->d_gi, n * 24 * 4, hipMemcpyDeviceToHost) != hipSuccess) return MM3E_ERR_DEVICE;
    for (uint64_t i = 0; i < n * 6; i++) {
      cubes[3 * i] = g4[4 * i];
      cubes[3 * i + 1] = g4[4 * i + 1];
      cubes[3 * i + 2] = g4[4 * i + 2];
    }
  }
  return MM3E_OK;
}

int mm3e_scene_field(mm3e_scene* s, const float* pts, uint32_t n, float* dist, uint32_t* mat) {
  if (!s || (n && (!pts || !dist || !mat))) return MM3E_ERR_INVALID;
  if (n == 0) return MM3E_OK;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  float *d_p = nullptr, *d_d = nullptr;
  uint32_t* d_m = nullptr;
  std::string msg;
  bool ok = hip_ok(hipMalloc(&d_p, (size_t)n * 12), "hipMalloc", &msg) && hip_ok(hipMalloc(&d_d, (size_t)n * 4), "hipMalloc", &msg) &&
            hip_ok(hipMalloc(&d_m, (size_t)n * 4), "hipMalloc", &msg);
  if (ok) ok = hip_ok(hipMemcpy(d_p, pts, (size_t)n * 12, hipMemcpyHostToDevice), "upload", &msg);
  if (ok) ok = hip_ok(launch_field_probe_table(s->r->stream, s->d_scene, d_p, n, d_d, d_m), "field probe", &msg);
  if (ok) ok = hip_ok(hipStreamSynchronize(s->r->stream), "field probe", &msg);
  if (ok) ok = hip_ok(hipMemcpy(dist, d_d, (size_t)n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok) ok = hip_ok(hipMemcpy(mat, d_m, (size_t)n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  hipFree(d_p);
  hipFree(d_d);
  hipFree(d_m);
  return ok ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_scene_march(mm3e_scene* s, const float* rays, uint32_t n, float* t, uint8_t* hit, uint32_t* mat, float* normal,
                     uint32_t* steps) {
  if (!s || (n && (!rays || !t || !hit || !mat || !normal || !steps))) return MM3E_ERR_INVALID;
  if (n == 0) return MM3E_OK;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  float *d_r = nullptr, *d_t = nullptr, *d_n = nullptr;
  uint8_t* d_h = nullptr;
  uint32_t *d_m = nullptr, *d_s = nullptr;
  std::string msg;
  bool ok = hip_ok(hipMalloc(&d_r, (size_t)n * 24), "hipMalloc", &msg) && hip_ok(hipMalloc(&d_t, (size_t)n * 4), "hipMalloc", &msg) &&
            hip_ok(hipMalloc(&d_n, (size_t)n * 12), "hipMalloc", &msg) && hip_ok(hipMalloc(&d_h, n), "hipMalloc", &msg) &&
            hip_ok(hipMalloc(&d_m, (size_t)n * 4), "hipMalloc", &msg) && hip_ok(hipMalloc(&d_s, (size_t)n * 4), "hipMalloc", &msg);
  if (ok) ok = hip_ok(hipMemcpy(d_r, rays, (size_t)n * 24, hipMemcpyHostToDevice), "upload", &msg);
  if (ok) ok = hip_ok(launch_march_probe_table(s->r->stream, s->d_scene, d_r, n, d_t, d_h, d_m, d_n, d_s), "march probe", &msg);
  if (ok) ok = hip_ok(hipStreamSynchronize(s->r->stream), "march probe", &msg);
  if (ok) ok = hip_ok(hipMemcpy(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok) ok = hip_ok(hipMemcpy(hit, d_h, n, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok) ok = hip_ok(hipMemcpy(mat, d_m, (size_t)n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok) ok = hip_ok(hipMemcpy(normal, d_n, (size_t)n * 12, hipMemcpyDeviceToHost), "readback", &msg);
  if (ok) ok = hip_ok(hipMemcpy(steps, d_s, (size_t)n * 4, hipMemcpyDeviceToHost), "readback", &msg);
  hipFree(d_r);
  hipFree(d_t);
  hipFree(d_n);
  hipFree(d_h);
  hipFree(d_m);
  hipFree(d_s);
  return ok ? MM3E_OK : MM3E_ERR_DEVICE;
}

// ---- host helpers ------------------------------------------------------------------------
void mm3e_camera_look_at(const float eye[3], const float target[3], const float up[3], float fov_y, mm3e_camera* out) {
  // camera.rs:50-55
  auto nrm = [](float v[3]) {
    float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 1e-12f) {
      float inv = 1.0f / l;
      v[0] *= inv; v[1] *= inv; v[2] *= inv;
    } else {
      v[0] = v[1] = v[2] = 0.0f;
    }
  };
  auto cross = [](const float* a, const float* o, float* r) {
    r[0] = a[1] * o[2] - a[2] * o[1];
    r[1] = a[2] * o[0] - a[0] * o[2];
    r[2] = a[0] * o[1] - a[1] * o[0];
  };
  float f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  nrm(f);
  float r[3];
  cross(f, up, r);
  nrm(r);
  float u[3];
  cross(r, f, u);
  std::memcpy(out->eye, eye, 12);
  std::memcpy(out->forward, f, 12);
  std::memcpy(out->right, r, 12);
  std::memcpy(out->up, u, 12);
  out->fov_scale = tanf(fov_y * 0.5f);
}

void mm3e_orbit_camera(const float target[3], float radius, float yaw, float pitch, float fov_y, mm3e_camera* out) {
  // lib.rs:1414-1435
  const float PI = 3.14159265358979323846f, TAU = 6.28318530717958647692f;
  if (std::isfinite(yaw)) {
    float v = std::fmod(yaw + PI, TAU);
    if (v < 0.0f) v += std::fabs(TAU);
    yaw = v - PI;
  } else {
    yaw = 0.0f;
  }
  const float lim = PI / 2.0f - 1.0e-4f;
  pitch = std::isfinite(pitch) ? (pitch < -lim ? -lim : (pitch > lim ? lim : pitch)) : 0.0f;
  radius = std::isfinite(radius) ? std::fmax(std::fabs(radius), 1.0e-4f) : 1.0f;
  if (std::isfinite(fov_y)) {
    const float lo = 1.0e-4f, hi = PI - 1.0e-4f;
    fov_y = fov_y < lo ? lo : (fov_y > hi ? hi : fov_y);
  } else {
    fov_y = 45.0f * (PI / 180.0f);
  }
  float eye[3] = {target[0] + radius * cosf(yaw) * cosf(pitch), target[1] + radius * sinf(pitch),
                  target[2] + radius * sinf(yaw) * cosf(pitch)};
  const float up[3] = {0.0f, 1.0f, 0.0f};
  mm3e_camera_look_at(eye, target, up, fov_y, out);
}

int mm3e_write_bmp(const uint8_t* rgba, uint32_t w, uint32_t h, uint8_t* out, size_t* out_len) {
  // framebuffer.rs:55-98 (an opaque pixel round-trips through the f32 framebuffer exactly)
  if (!out_len || (!rgba && out)) return MM3E_ERR_INVALID;
  const size_t pad = (4 - ((size_t)w * 3) % 4) % 4;
  const size_t row = (size_t)w * 3 + pad, pix = row * h, total = 54 + pix;
  if (total > 0xffffffffull) return MM3E_ERR_INVALID;  // framebuffer.rs:68
  if (!out) {
    *out_len = total;
    return MM3E_OK;
  }
  if (*out_len < total) return MM3E_ERR_INVALID;
  *out_len = total;
  auto u32le = [](uint8_t* p, uint32_t v) { std::memcpy(p, &v, 4); };
  auto u16le = [](uint8_t* p, uint16_t v) { std::memcpy(p, &v, 2); };
  std::memset(out, 0, 54);
  out[0] = 'B';
  out[1] = 'M';
  u32le(out + 2, (uint32_t)total);
  u32le(out + 10, 54);
  u32le(out + 14, 40);
  u32le(out + 18, w);
  u32le(out + 22, h);
  u16le(out + 26, 1);
  u16le(out + 28, 24);
  u32le(out + 34, (uint32_t)pix);
  u32le(out + 38, 2835);
  u32le(out + 42, 2835);
  uint8_t* q = out + 54;
  for (int64_t y = (int64_t)h - 1; y >= 0; y--) {
    for (uint32_t x = 0; x < w; x++) {
      const uint8_t* px = rgba + 4 * ((size_t)y * w + x);
      *q++ = px[2];
      *q++ = px[1];
      *q++ = px[0];
    }
    for (size_t i = 0; i < pad; i++) *q++ = 0;
  }
  return MM3E_OK;
}

int mm3e_selftest_sqrt(mm3e_renderer* r, uint64_t* mismatches) {
  if (!r || !mismatches) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(r->mu);
  hipSetDevice(r->device);
  unsigned long long* d = nullptr;
  std::string msg;
  unsigned long long h = 0;
  bool ok = hip_ok(hipMalloc((void**)&d, 8), "hipMalloc", &msg) &&
            hip_ok(hipMemcpy(d, &h, 8, hipMemcpyHostToDevice), "upload", &msg) &&
            hip_ok(launch_sqrt_selftest(r->stream, d), "self-test launch", &msg) &&
            hip_ok(hipStreamSynchronize(r->stream), "self-test", &msg) &&
            hip_ok(hipMemcpy(&h, d, 8, hipMemcpyDeviceToHost), "readback", &msg);
  hipFree(d);
  *mismatches = h;
  return ok ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_renderer_set_tile_order(mm3e_renderer* r, int on) {
  if (!r) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(r->mu);
  r->tile_order = on != 0;
  return MM3E_OK;
}

int mm3e_renderer_set_wavefront(mm3e_renderer* r, uint32_t budget, uint32_t passes, uint32_t blocks_per_cu) {
  if (!r || passes >= WF_MAX_PASSES || blocks_per_cu > 64) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(r->mu);
  r->wf_budget = budget;
  r->wf_passes = passes;
  r->wf_bpc = blocks_per_cu;
  return MM3E_OK;
}

int mm3e_scene_set_wavefront(mm3e_scene* s, int own, uint32_t budget, uint32_t passes, uint32_t blocks_per_cu) {
  if (!s || passes >= WF_MAX_PASSES || blocks_per_cu > 64) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  s->wf_own = own != 0;
  s->wf_set[0] = budget;
  s->wf_set[1] = passes;
  s->wf_set[2] = blocks_per_cu;
  return MM3E_OK;
}

// Times the scene's frame (render stage, kernel events) under each candidate: one kernel, then
// queued frames over a grid of budgets / passes / grid sizes; keeps the fastest as the scene's
// own setting.
int mm3e_scene_autotune(mm3e_scene* s, const mm3e_camera* cam, uint32_t reps, char* report, size_t report_len) {
  if (!s || validate_cam(cam)) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->r->mu);
  hipSetDevice(s->r->device);
  reps = std::max(1u, std::min(reps, (uint32_t)mm3e_scene::kRing - 1));
  static const uint32_t cand[][3] = {{0, 0, 0},   {1000, 0, 0}, {64, 1, 0}, {32, 2, 0},
                                     {16, 3, 0},  {32, 2, 8},   {64, 1, 4}};
  const FrameParams fp = make_frame(*cam, s->w, s->h);
  const bool own0 = s->wf_own;
  uint32_t set0[3] = {s->wf_set[0], s->wf_set[1], s->wf_set[2]};
  std::string msg, rep;
  float best = 1e30f;
  size_t best_i = 0;
  for (size_t c = 0; c < sizeof cand / sizeof cand[0]; c++) {
    s->wf_own = true;
    for (int k = 0; k < 3; k++) s->wf_set[k] = cand[c][k];
    bool ok = launch_frame(s, s->d_scene, fp, msg);  // warm-up (queued buffers, caches)
    const uint64_t first = s->frames;
    for (uint32_t i = 0; ok && i < reps; i++) ok = launch_frame(s, s->d_scene, fp, msg);
    ok = ok && hip_ok(hipStreamSynchronize(s->r->stream), "autotune frames", &msg);
    if (!ok) {
      s->wf_own = own0;
      for (int k = 0; k < 3; k++) s->wf_set[k] = set0[k];
      set_err(report, report_len, "autotune failed: " + msg);
      return MM3E_ERR_DEVICE;
    }
    float sum = 0.0f;
    for (uint64_t f = first; f < s->frames; f++) {
      float a = 0.0f;
      hipEvent_t* ev = s->ev[f % mm3e_scene::kRing];
      hipEventElapsedTime(&a, ev[0], ev[1]);
      sum += a;
    }
    const float ms = sum / (float)reps;
    char line[128];
    if (cand[c][0] == 0) std::snprintf(line, sizeof line, "one kernel: %.3f ms\n", ms);
    else
      std::snprintf(line, sizeof line, "queued budget %u passes %u blocks/CU %u: %.3f ms\n", cand[c][0], cand[c][1],
                    cand[c][2], ms);
    rep += line;
    if (ms < best) {
      best = ms;
      best_i = c;
    }
  }
  s->wf_own = true;
  for (int k = 0; k < 3; k++) s->wf_set[k] = cand[best_i][k];
  char line[96];
  std::snprintf(line, sizeof line, "chosen: %s (%.3f ms)\n", cand[best_i][0] ? "queued" : "one kernel", best);
  rep += line;
  set_err(report, report_len, rep);
  return MM3E_OK;
}

int mm3e_renderer_set_codegen_film(mm3e_renderer* r, int on) {
  if (!r) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(r->mu);
  r->gen_film = on != 0;
  return MM3E_OK;
}

int mm3e_renderer_set_codegen(mm3e_renderer* r, int mode) {
  if (!r || mode < MM3E_CODEGEN_OFF || mode > MM3E_CODEGEN_SYNC) return MM3E_ERR_INVALID;
  r->codegen = mode;
  return MM3E_OK;
}

int mm3e_scene_codegen_status(mm3e_scene* s, int wait, int* state, char* err, size_t errlen) {
  if (!s || !state) return MM3E_ERR_INVALID;
  mm3e::GenJob* j = s->gen.get();
  if (!j) {
    *state = MM3E_CODEGEN_NONE;
    return MM3E_OK;
  }
  std::unique_lock<std::mutex> lk(j->mu);
  if (wait)
    j->cv.wait(lk, [&] {
      const int st = j->state.load();
      return st == MM3E_CODEGEN_READY || st == MM3E_CODEGEN_FAILED;
    });
  *state = j->state.load();
  if (*state == MM3E_CODEGEN_FAILED) set_err(err, errlen, j->err);
  return MM3E_OK;
}

int mm3e_scene_emit_source(const void* ir, size_t ir_len, char* src, size_t* src_len) {
  if (!src_len) return MM3E_ERR_INVALID;
  HostScene hs;
  std::string msg;
  if (!parse_scene_ir(ir, ir_len, hs, msg)) return MM3E_ERR_INVALID;
  Blob b;
  build_blob(hs, {}, {}, hs.hdr.aa == 0 ? 1u : hs.hdr.aa, hs.hdr.bounces, 0, nullptr, false, b);
  uint32_t wb = 0, wn = 2, wc = 0;
  env_wavefront(&wb, &wn, &wc);
  const std::string text = emit_scene_source(b, hs.hdr.bounces, generated_prefers_sm(b), wb > 0, env_codegen_film());
  const size_t need = text.size() + 1;
  if (!src || *src_len < need) {
    *src_len = need;
    return MM3E_ERR_INVALID;
  }
  std::memcpy(src, text.c_str(), need);
  *src_len = need;
  return MM3E_OK;
}

}  // extern "C"
