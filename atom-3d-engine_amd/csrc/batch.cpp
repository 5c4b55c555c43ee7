// batch.cpp — batched rendering of many (scene, camera, size) jobs in one persistent launch.
// When some jobs' scenes have generated kernels, the autotuner also times a mixed mode: those
// jobs as their own launches spread over four streams, then the persistent launch for the
// rest, and keeps whichever is faster.
//
// A batch owns a job table in HBM, per-job framebuffers (HDR, bloom scratch, RGBA8) carved
// from one arena, and a queue counter.  A frame of the batch is three kernels — queue reset,
// the persistent tile-queue render (batch_kernel.hip), the batched bloom/tonemap over all
// jobs' rows — captured once into a hipGraph and replayed; cameras change between replays by
// rewriting the job table (outside the graph).  The autotuner times the persistent launch
// over blocks-per-CU and, for single-sample jobs, the tile shape, and keeps the fastest.
// Tiling never changes a pixel: each pixel's samples are summed in the same order whatever
// tile or block renders it.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "batch.h"
#include "codegen.h"
#include "launch.h"
#include "rtc.h"
#include "runtime.h"

struct mm3e_batch {
  mm3e_renderer* r = nullptr;
  std::vector<mm3e_batch_job> in;     // caller's jobs (original order)
  std::vector<uint32_t> order;        // table position -> caller job index
  std::vector<uint32_t> pos_of;       // caller job index -> table position
  std::vector<mm3e::BatchJob> jobs;   // host copy of the device job table
  std::vector<mm3e::BatchPost> posts;
  mm3e::BatchJob* d_jobs = nullptr;
  mm3e::BatchPost* d_posts = nullptr;
  uint32_t* d_counter = nullptr;
  uint8_t* d_arena = nullptr;
  size_t arena_bytes = 0;
  uint32_t total_tiles = 0, total_rows = 0, max_w = 0, max_radius = 0, maxb = 0;
  uint32_t lds_bytes = 0;             // > 0: scene tables staged in LDS
  uint32_t blocks_per_cu = 4, cus = 256;
  uint32_t tile_aa1[2] = {16, 16};    // tile shape for single-sample jobs
  bool jobs_dirty = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool launched = false;
  uint64_t total_pixels = 0;
  // Mixed mode: jobs whose scene has a generated kernel run as their own launches, spread over
  // kStreams streams; then the other jobs' tiles go through the persistent table kernel and the
  // batched post.  `gen` holds those jobs' table positions.
  bool mixed = false;
  std::vector<uint32_t> gen;                 // positions (in posts) of generated-kernel jobs
  std::vector<mm3e::BatchJob> gen_jobs;      // their launch records (tile shape, camera, hdr)
  static constexpr int kStreams = 4;
  hipStream_t aux[kStreams - 1] = {};
  hipEvent_t fork = nullptr, join[kStreams - 1] = {};
  // tile order of the persistent launch: [order n][cost n][ticket], n = total_tiles; each replay
  // records tile costs and re-sorts the order for the next one (expensive tiles first)
  uint32_t* d_tq = nullptr;
  uint32_t tq_tiles = 0;
  // Generated batch module: the maps of the batch's scenes that get generated kernels, in one
  // module with the persistent loop (mm3e_gen_batch); built by the autotuner, used when faster.
  std::vector<const mm3e_scene*> gen_scenes;  // map k renders jobs of gen_scenes[k]
  hipFunction_t gen_fn = nullptr;
  bool use_gen = false;
};

namespace mm3e {
namespace {

void job_tiles(const mm3e_batch* b, uint32_t aa, uint32_t* tw, uint32_t* th) {
  if (aa <= 1) {
    *tw = b->tile_aa1[0];
    *th = b->tile_aa1[1];
  } else {
    tile_shape(aa, tw, th);
  }
}

// (Re)build the job and post tables for the current tiling parameters.
// A job's scene has a generated kernel ready (read once per table build).
bool job_generated(const mm3e_batch_job& in) {
  const GenJob* g = in.scene->gen.get();
  return g && g->state.load(std::memory_order_acquire) == MM3E_CODEGEN_READY && g->fn;
}

void build_tables(mm3e_batch* b) {
  b->jobs.clear();
  b->posts.clear();
  b->gen.clear();
  b->gen_jobs.clear();
  uint32_t tile = 0, row = 0;
  b->max_w = 0;
  b->max_radius = 1;
  b->maxb = 0;
  uint32_t stage_max = 0;
  bool all_fit = true;
  size_t off = 0;
  for (uint32_t k = 0; k < b->order.size(); k++) {
    const mm3e_batch_job& in = b->in[b->order[k]];
    mm3e_scene* s = in.scene;
    const uint32_t w = in.width, h = in.height;
    BatchJob J{};
    J.sc = s->d_scene;
    J.fp = make_frame(in.cam, w, h);
    const size_t px = (size_t)w * h;
    J.hdr = (float*)(b->d_arena + off);
    const bool generated = b->mixed && job_generated(in);
    if (generated) {
      // its own launch: the scene's tile shape and table staging, as a plain frame
      J.tw = s->tw;
      J.th = s->th;
      J.tiles_x = (w + J.tw - 1) / J.tw;
      J.stage_bytes = s->gen_stage;
      b->gen.push_back((uint32_t)b->posts.size());
      b->gen_jobs.push_back(J);
    } else {
      job_tiles(b, s->aa, &J.tw, &J.th);
      J.tiles_x = (w + J.tw - 1) / J.tw;
      const uint32_t tiles_y = (h + J.th - 1) / J.th;
      J.tile_base = tile;
      tile += J.tiles_x * tiles_y;
      J.stage_bytes = s->hdr.stage_bytes;
      J.gen = kGenTable;
      for (size_t g = 0; g < b->gen_scenes.size(); g++)
        if (b->gen_scenes[g] == s) J.gen = (uint32_t)g;
      stage_max = std::max(stage_max, J.stage_bytes);
      all_fit = all_fit && J.stage_bytes <= b->r->lds_budget;
    }
    BatchPost P{};
    P.pp = s->pp;
    P.pp.width = w;
    P.pp.height = h;
    P.kern = s->d_kern;
    P.hdr = J.hdr;
    P.tmp = (float*)(b->d_arena + off + px * 12);
    P.rgba = (uchar4*)(b->d_arena + off + px * 24);
    P.row_base = row;
    row += h;
    off += px * 28;
    off = (off + 255) & ~size_t(255);
    b->max_w = std::max(b->max_w, w);
    b->max_radius = std::max(b->max_radius, P.pp.radius);
    if (!generated) {
      b->maxb = std::max(b->maxb, s->bounces);
      b->jobs.push_back(J);
    }
    b->posts.push_back(P);
  }
  b->total_tiles = tile;
  b->total_rows = row;
  b->lds_bytes = all_fit ? stage_max : 0;
}

size_t arena_size(const std::vector<mm3e_batch_job>& in) {
  size_t off = 0;
  for (const auto& j : in) {
    off += (size_t)j.width * j.height * 28;
    off = (off + 255) & ~size_t(255);
  }
  return off;
}

// The tile order for the current table, reset to index order when the tile count changed.
bool ensure_order(mm3e_batch* b, std::string& msg) {
  if (b->d_tq && b->tq_tiles == b->total_tiles) return true;
  hipFree(b->d_tq);
  b->d_tq = nullptr;
  b->tq_tiles = 0;
  if (b->total_tiles == 0) return true;
  std::vector<uint32_t> init(2 * (size_t)b->total_tiles + 1, 0u);
  for (uint32_t i = 0; i < b->total_tiles; i++) init[i] = i;
  if (!hip_ok(hipMalloc((void**)&b->d_tq, init.size() * 4), "hipMalloc(batch tile order)", &msg) ||
      !hip_ok(hipMemcpy(b->d_tq, init.data(), init.size() * 4, hipMemcpyHostToDevice), "upload tile order", &msg))
    return false;
  b->tq_tiles = b->total_tiles;
  return true;
}

TileOrder batch_order(const mm3e_batch* b) {
  if (!b->d_tq) return TileOrder{};
  return TileOrder{b->d_tq, b->d_tq + b->tq_tiles, b->d_tq + 2 * b->tq_tiles, nullptr, nullptr, 0u, 0u};
}

// Persistent render of the table jobs on `ts`, then the next replay's order.
bool enqueue_persistent(mm3e_batch* b, hipStream_t ts, std::string& msg) {
  const TileOrder q = batch_order(b);
  if (!hip_ok(launch_batch_reset(ts, b->d_counter), "batch reset", &msg)) return false;
  if (b->use_gen && b->gen_fn) {
    const BatchJob* jobs = b->d_jobs;
    uint32_t n_jobs = (uint32_t)b->jobs.size(), total = b->total_tiles;
    uint32_t* counter = b->d_counter;
    TileOrder qq = q;
    void* args[] = {&jobs, &n_jobs, &total, &counter, &qq};
    if (!hip_ok(hipModuleLaunchKernel(b->gen_fn, b->cus * b->blocks_per_cu, 1, 1, 256, 1, 1, b->lds_bytes, ts, args,
                                      nullptr),
                "generated batch render", &msg))
      return false;
  } else if (!hip_ok(launch_batch_render(b->maxb, b->cus * b->blocks_per_cu, ts, b->d_jobs,
                                         (uint32_t)b->jobs.size(), b->total_tiles, b->d_counter, b->lds_bytes, q),
                     "batch render", &msg)) {
    return false;
  }
  if (q.order && !hip_ok(launch_tile_order(ts, q.cost, b->tq_tiles, b->d_tq), "batch tile order", &msg)) return false;
  return true;
}

bool enqueue(mm3e_batch* b, std::string& msg) {
  hipStream_t st = b->r->stream;
  if (b->mixed) {
    // fork: the aux streams start after everything already on the renderer's stream
    if (!hip_ok(hipEventRecord(b->fork, st), "fork", &msg)) return false;
    for (int k = 0; k < mm3e_batch::kStreams - 1; k++)
      if (!hip_ok(hipStreamWaitEvent(b->aux[k], b->fork, 0), "fork wait", &msg)) return false;
    // generated jobs round-robin over the renderer's stream and the aux streams
    for (size_t i = 0; i < b->gen_jobs.size(); i++) {
      const BatchJob& J = b->gen_jobs[i];
      const mm3e_scene* s = b->in[b->order[b->gen[i]]].scene;
      const int lane = (int)(i % mm3e_batch::kStreams);
      hipStream_t gs = lane == 0 ? st : b->aux[lane - 1];
      if (!hip_ok(launch_generated_index(s->gen->fn, J.sc, J.fp, J.hdr, J.tw, J.th, s->aa, J.stage_bytes, gs),
                  "generated batch job", &msg))
        return false;
    }
    // join, then the table jobs through the persistent launch on the whole chip
    for (int k = 0; k < mm3e_batch::kStreams - 1; k++) {
      if (!hip_ok(hipEventRecord(b->join[k], b->aux[k]), "join", &msg)) return false;
      if (!hip_ok(hipStreamWaitEvent(st, b->join[k], 0), "join wait", &msg)) return false;
    }
    if (!b->jobs.empty() && !enqueue_persistent(b, st, msg)) return false;
  } else if (!enqueue_persistent(b, st, msg)) {
    return false;
  }
  return hip_ok(launch_batch_post(st, b->d_posts, (uint32_t)b->posts.size(), b->total_rows, b->max_w, b->max_radius),
                "batch post", &msg);
}

bool upload(mm3e_batch* b, std::string& msg) {
  hipStream_t st = b->r->stream;
  if (!ensure_order(b, msg)) return false;
  if (!b->jobs.empty() &&
      !hip_ok(hipMemcpyAsync(b->d_jobs, b->jobs.data(), b->jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st),
              "upload jobs", &msg))
    return false;
  if (!hip_ok(hipMemcpyAsync(b->d_posts, b->posts.data(), b->posts.size() * sizeof(BatchPost), hipMemcpyHostToDevice,
                             st),
              "upload posts", &msg))
    return false;
  b->jobs_dirty = false;
  return hip_ok(hipStreamSynchronize(st), "upload", &msg);
}

bool capture(mm3e_batch* b, std::string& msg) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->mixed) return true;  // mixed batches are enqueued directly on several streams
  hipStream_t st = b->r->stream;
  if (!hip_ok(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture", &msg)) return false;
  std::string m2;
  bool ok = enqueue(b, m2);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (!ok) {
    msg = m2;
    if (g) hipGraphDestroy(g);
    return false;
  }
  if (!hip_ok(e, "end capture", &msg)) return false;
  b->graph = g;
  return hip_ok(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate", &msg);
}

// The generated batch module over the batch's scenes that qualify for generated kernels (code
// generation on, at most codegen_object_limit() objects): their maps and the persistent loop in
// one hipRTC module.  False (and no module) when no scene qualifies or compilation fails.
bool build_gen_module(mm3e_batch* b, std::string& msg) {
  b->gen_scenes.clear();
  b->gen_fn = nullptr;
  if (b->r->codegen == MM3E_CODEGEN_OFF) return false;
  for (const auto& j : b->in) {
    const mm3e_scene* s = j.scene;
    if (s->host.objects.size() > codegen_object_limit()) continue;
    if (std::find(b->gen_scenes.begin(), b->gen_scenes.end(), s) == b->gen_scenes.end()) b->gen_scenes.push_back(s);
  }
  if (b->gen_scenes.empty()) return false;
  std::vector<Blob> blobs(b->gen_scenes.size());
  std::vector<const Blob*> maps;
  for (size_t k = 0; k < b->gen_scenes.size(); k++) {
    const mm3e_scene* s = b->gen_scenes[k];
    build_blob(s->host, {}, {}, s->aa, s->bounces, (uint64_t)(uintptr_t)s->d_tables, s->d_gi, false, blobs[k]);
    maps.push_back(&blobs[k]);
  }
  const std::string src = emit_batch_source(maps, b->maxb, b->lds_bytes > 0);
  if (!rtc_function(src, b->r->device, "mm3e_gen_batch", &b->gen_fn, msg)) {
    b->gen_scenes.clear();
    b->gen_fn = nullptr;
    return false;
  }
  return true;
}

}  // namespace
}  // namespace mm3e

using namespace mm3e;

extern "C" {

void mm3e_batch_destroy(mm3e_batch* b) {
  if (!b) return;
  if (b->r) {
    hipSetDevice(b->r->device);
    hipStreamSynchronize(b->r->stream);
  }
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->d_jobs) hipFree(b->d_jobs);
  if (b->d_posts) hipFree(b->d_posts);
  if (b->d_counter) hipFree(b->d_counter);
  if (b->d_tq) hipFree(b->d_tq);
  if (b->d_arena) hipFree(b->d_arena);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->fork) hipEventDestroy(b->fork);
  for (int k = 0; k < mm3e_batch::kStreams - 1; k++) {
    if (b->join[k]) hipEventDestroy(b->join[k]);
    if (b->aux[k]) {
      hipStreamSynchronize(b->aux[k]);
      hipStreamDestroy(b->aux[k]);
    }
  }
  delete b;
}

int mm3e_batch_create(mm3e_renderer* r, const mm3e_batch_job* jobs, uint32_t n_jobs, mm3e_batch** out, char* err,
                      size_t errlen) {
  if (!r || !out || (n_jobs && !jobs)) return MM3E_ERR_INVALID;
  *out = nullptr;
  if (n_jobs == 0) {
    set_err(err, errlen, "a batch needs at least one job");
    return MM3E_ERR_INVALID;
  }
  for (uint32_t i = 0; i < n_jobs; i++) {
    const mm3e_batch_job& j = jobs[i];
    if (!j.scene || j.scene->r != r) {
      set_err(err, errlen, "batch job " + std::to_string(i) + ": scene is missing or compiled on another renderer");
      return MM3E_ERR_INVALID;
    }
    if (j.scene->host.hdr.mode != 0) {
      set_err(err, errlen, "batch job " + std::to_string(i) +
                               ": debug render modes render through mm3e_scene_render_rgba, not in batches");
      return MM3E_ERR_INVALID;
    }
    if (j.width == 0 || j.height == 0 || j.width > 16384 || j.height > 16384) {
      set_err(err, errlen, "batch job " + std::to_string(i) + ": " + std::to_string(j.width) + "x" +
                               std::to_string(j.height) + " out of range 1..=16384");
      return MM3E_ERR_INVALID;
    }
  }
  std::lock_guard<std::mutex> lk(r->mu);
  hipSetDevice(r->device);
  auto* b = new mm3e_batch();
  b->r = r;
  b->in.assign(jobs, jobs + n_jobs);
  b->order.resize(n_jobs);
  for (uint32_t i = 0; i < n_jobs; i++) b->order[i] = i;
  // group jobs of the same scene so a block rarely re-stages scene tables
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](uint32_t a, uint32_t c) { return (uintptr_t)b->in[a].scene < (uintptr_t)b->in[c].scene; });
  b->pos_of.resize(n_jobs);
  for (uint32_t k = 0; k < n_jobs; k++) b->pos_of[b->order[k]] = k;
  for (const auto& j : b->in) b->total_pixels += (uint64_t)j.width * j.height;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, r->device) == hipSuccess && p.multiProcessorCount > 0) b->cus = p.multiProcessorCount;
  std::string msg;
  b->arena_bytes = arena_size(b->in);
  bool ok = hip_ok(hipMalloc(&b->d_arena, b->arena_bytes), "hipMalloc(batch framebuffers)", &msg) &&
            hip_ok(hipMalloc(&b->d_jobs, n_jobs * sizeof(BatchJob)), "hipMalloc(jobs)", &msg) &&
            hip_ok(hipMalloc(&b->d_posts, n_jobs * sizeof(BatchPost)), "hipMalloc(posts)", &msg) &&
            hip_ok(hipMalloc(&b->d_counter, 256), "hipMalloc(counter)", &msg) &&
            hip_ok(hipEventCreate(&b->ev0), "hipEventCreate", &msg) && hip_ok(hipEventCreate(&b->ev1), "hipEventCreate", &msg) &&
            hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate", &msg);
  for (int k = 0; ok && k < mm3e_batch::kStreams - 1; k++)
    ok = hip_ok(hipStreamCreate(&b->aux[k]), "hipStreamCreate", &msg) &&
         hip_ok(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming), "hipEventCreate", &msg);
  if (ok) {
    build_tables(b);
    ok = upload(b, msg) && capture(b, msg);
  }
  if (!ok) {
    set_err(err, errlen, msg);
    mm3e_batch_destroy(b);
    return MM3E_ERR_DEVICE;
  }
  *out = b;
  return MM3E_OK;
}

int mm3e_batch_set_camera(mm3e_batch* b, uint32_t job, const mm3e_camera* cam) {
  if (!b || !cam || job >= b->in.size()) return MM3E_ERR_INVALID;
  b->in[job].cam = *cam;
  b->jobs_dirty = true;  // the job records are rebuilt from the cameras at the next launch
  return MM3E_OK;
}

int mm3e_batch_launch(mm3e_batch* b) {
  if (!b) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->r->mu);
  hipSetDevice(b->r->device);
  std::string msg;
  if (b->jobs_dirty) {
    build_tables(b);
    if (!upload(b, msg)) return MM3E_ERR_DEVICE;
  }
  hipStream_t st = b->r->stream;
  hipEventRecord(b->ev0, st);
  if (b->mixed) {
    if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;
  } else if (!hip_ok(hipGraphLaunch(b->exec, st), "graph launch", &msg)) {
    return MM3E_ERR_DEVICE;
  }
  hipEventRecord(b->ev1, st);
  b->launched = true;
  return MM3E_OK;
}

int mm3e_batch_wait(mm3e_batch* b, float* gpu_ms) {
  if (!b || !b->launched) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->r->mu);
  hipSetDevice(b->r->device);
  if (hipEventSynchronize(b->ev1) != hipSuccess) return MM3E_ERR_DEVICE;
  if (gpu_ms) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    *gpu_ms = ms;
  }
  return MM3E_OK;
}

int mm3e_batch_read(mm3e_batch* b, uint32_t job, uint8_t* out, size_t out_len) {
  if (!b || !out || job >= b->in.size()) return MM3E_ERR_INVALID;
  const size_t need = (size_t)b->in[job].width * b->in[job].height * 4;
  if (out_len < need) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->r->mu);
  hipSetDevice(b->r->device);
  const BatchPost& P = b->posts[b->pos_of[job]];
  std::string msg;
  if (!hip_ok(hipMemcpyAsync(out, P.rgba, need, hipMemcpyDeviceToHost, b->r->stream), "batch readback", &msg))
    return MM3E_ERR_DEVICE;
  return hipStreamSynchronize(b->r->stream) == hipSuccess ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_batch_info(mm3e_batch* b, uint64_t* total_pixels, uint32_t* total_tiles, uint32_t* blocks_per_cu,
                    uint32_t* tile_w, uint32_t* tile_h, uint32_t* lds_bytes) {
  if (!b) return MM3E_ERR_INVALID;
  if (total_pixels) *total_pixels = b->total_pixels;
  if (total_tiles) *total_tiles = b->total_tiles;
  if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
  if (tile_w) *tile_w = b->tile_aa1[0];
  if (tile_h) *tile_h = b->tile_aa1[1];
  if (lds_bytes) *lds_bytes = b->lds_bytes;
  return MM3E_OK;
}

int mm3e_batch_autotune(mm3e_batch* b, uint32_t reps, char* report, size_t report_len) {
  if (!b) return MM3E_ERR_INVALID;
  std::lock_guard<std::mutex> lk(b->r->mu);
  hipSetDevice(b->r->device);
  reps = std::max(1u, reps);
  b->mixed = false;
  static const uint32_t bpcs[] = {2, 3, 4, 6, 8};
  static const uint32_t tiles[][2] = {{16, 16}, {32, 8}, {64, 4}, {8, 32}};
  bool any_aa1 = false;
  for (const auto& j : b->in) any_aa1 = any_aa1 || j.scene->aa <= 1;
  const size_t n_tiles = any_aa1 ? sizeof(tiles) / sizeof(tiles[0]) : 1;
  std::string msg, rep;
  float best = 1e30f;
  uint32_t best_bpc = b->blocks_per_cu, best_t[2] = {b->tile_aa1[0], b->tile_aa1[1]};
  hipStream_t st = b->r->stream;
  for (size_t ti = 0; ti < n_tiles; ti++) {
    if (any_aa1) {
      b->tile_aa1[0] = tiles[ti][0];
      b->tile_aa1[1] = tiles[ti][1];
    }
    build_tables(b);
    if (!upload(b, msg)) return MM3E_ERR_DEVICE;
    for (uint32_t bpc : bpcs) {
      b->blocks_per_cu = bpc;
      if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;  // warm-up
      hipEventRecord(b->ev0, st);
      for (uint32_t k = 0; k < reps; k++)
        if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;
      hipEventRecord(b->ev1, st);
      if (hipEventSynchronize(b->ev1) != hipSuccess) return MM3E_ERR_DEVICE;
      float ms = 0.0f;
      hipEventElapsedTime(&ms, b->ev0, b->ev1);
      ms /= (float)reps;
      char line[96];
      std::snprintf(line, sizeof line, "tile %ux%u blocks/CU %u: %.3f ms\n", b->tile_aa1[0], b->tile_aa1[1], bpc, ms);
      rep += line;
      if (ms < best) {
        best = ms;
        best_bpc = bpc;
        best_t[0] = b->tile_aa1[0];
        best_t[1] = b->tile_aa1[1];
      }
    }
  }
  b->blocks_per_cu = best_bpc;
  b->tile_aa1[0] = best_t[0];
  b->tile_aa1[1] = best_t[1];
  // Generated batch module: the scenes' own maps inside the persistent loop.  Opt-in
  // (MM3E_BATCH_GEN=auto: keep it when faster; =1: keep it whenever it builds): measured on
  // MI355X it is not faster than the table walk in the loop (a 7-job test batch 19.4 against
  // 18.1-18.5 ms; the 64-job c3 load unchanged) and its hipRTC build of several scene maps in one
  // module takes over a minute.
  const char* gen_env = std::getenv("MM3E_BATCH_GEN");
  const bool gen_force = gen_env && std::strcmp(gen_env, "1") == 0;
  const bool gen_skip = !(gen_force || (gen_env && std::strcmp(gen_env, "auto") == 0));
  b->use_gen = false;
  build_tables(b);
  std::string gmsg;
  if (!gen_skip && build_gen_module(b, gmsg)) {
    build_tables(b);  // jobs now carry their map index
    if (!upload(b, msg)) return MM3E_ERR_DEVICE;
    b->use_gen = true;
    if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;  // warm-up
    hipEventRecord(b->ev0, st);
    for (uint32_t k = 0; k < reps; k++)
      if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;
    hipEventRecord(b->ev1, st);
    if (hipEventSynchronize(b->ev1) != hipSuccess) return MM3E_ERR_DEVICE;
    float gen_ms = 0.0f;
    hipEventElapsedTime(&gen_ms, b->ev0, b->ev1);
    gen_ms /= (float)reps;
    char line[160];
    std::snprintf(line, sizeof line, "generated maps of %zu scenes, tile %ux%u blocks/CU %u: %.3f ms\n",
                  b->gen_scenes.size(), best_t[0], best_t[1], best_bpc, gen_ms);
    rep += line;
    if (gen_ms < best || gen_force) best = gen_ms;
    else b->use_gen = false;
  } else if (!gmsg.empty()) {
    rep += "generated batch module not built: " + gmsg.substr(0, 200) + "\n";
  }
  // mixed mode (generated kernels as their own launches on several streams) when any job's
  // scene has one; kept if faster than the best persistent-only configuration
  bool any_gen = false;
  for (const auto& j : b->in) any_gen = any_gen || job_generated(j);
  float mixed_ms = -1.0f;
  if (any_gen) {
    b->mixed = true;
    build_tables(b);
    if (!upload(b, msg)) return MM3E_ERR_DEVICE;
    if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;  // warm-up
    hipEventRecord(b->ev0, st);
    for (uint32_t k = 0; k < reps; k++)
      if (!enqueue(b, msg)) return MM3E_ERR_DEVICE;
    hipEventRecord(b->ev1, st);
    if (hipEventSynchronize(b->ev1) != hipSuccess) return MM3E_ERR_DEVICE;
    hipEventElapsedTime(&mixed_ms, b->ev0, b->ev1);
    mixed_ms /= (float)reps;
    char line[128];
    std::snprintf(line, sizeof line, "mixed: %zu generated-kernel jobs on %d streams, rest tile %ux%u blocks/CU %u: %.3f ms\n",
                  b->gen_jobs.size(), mm3e_batch::kStreams, best_t[0], best_t[1], best_bpc, mixed_ms);
    rep += line;
    b->mixed = mixed_ms < best;
  }
  build_tables(b);
  if (!upload(b, msg) || !capture(b, msg)) return MM3E_ERR_DEVICE;
  char line[128];
  if (b->mixed)
    std::snprintf(line, sizeof line, "chosen: mixed (%.3f ms)\n", mixed_ms);
  else
    std::snprintf(line, sizeof line, "chosen: %stile %ux%u blocks/CU %u (%.3f ms)\n",
                  b->use_gen ? "generated maps, " : "", best_t[0], best_t[1], best_bpc, best);
  rep += line;
  set_err(report, report_len, rep);
  return MM3E_OK;
}

int mm3e_render_batch(mm3e_renderer* r, const mm3e_batch_job* jobs, uint32_t n_jobs, char* err, size_t errlen) {
  mm3e_batch* b = nullptr;
  int rc = mm3e_batch_create(r, jobs, n_jobs, &b, err, errlen);
  if (rc != MM3E_OK) return rc;
  rc = mm3e_batch_launch(b);
  if (rc == MM3E_OK) rc = mm3e_batch_wait(b, nullptr);
  for (uint32_t i = 0; rc == MM3E_OK && i < n_jobs; i++)
    if (jobs[i].out) rc = mm3e_batch_read(b, i, jobs[i].out, (size_t)jobs[i].width * jobs[i].height * 4);
  if (rc != MM3E_OK) set_err(err, errlen, "batched render failed on the device");
  mm3e_batch_destroy(b);
  return rc;
}

}  // extern "C"
