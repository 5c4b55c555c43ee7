// runtime.h — internal state behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mm3e_hip.h"
#include "dev_render.h"
#include "scene_compile.h"

namespace mm3e {
// One scene's generated kernel: compiled by the renderer's code-generation workers.
struct GenJob {
  std::string src;
  int device = 0;
  std::atomic<int> state{MM3E_CODEGEN_QUEUED};
  // the render kernel mm3e_gen_render_o (both fields name it): tiles in a given order, or in
  // index order with a null TileOrder::order (launch_generated_index)
  hipFunction_t fn = nullptr;
  hipFunction_t fn_o = nullptr;
  // optional kernels (GenOpt order: queued-frame stages, film, G-buffer); null unless the
  // source was generated with them
  hipFunction_t fn_opt[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  std::string err;
  std::mutex mu;
  std::condition_variable cv;
};
}  // namespace mm3e

struct mm3e_renderer {
  int device = 0;
  uint32_t lds_budget = 0;       // largest table copy staged in LDS per block
  std::string name;
  hipStream_t stream = nullptr;  // default priority
  std::mutex mu;
  // scene code generation: mode, queue and worker threads (started on first use)
  int codegen = MM3E_CODEGEN_BACKGROUND;
  bool tile_order = true;        // generated kernels dispatch tiles most expensive first
  // queued frames (dev_wavefront.h): evaluations per march pass (0: one kernel per frame),
  // continuation passes, blocks per CU of the grid-stride stages (0: one lane per entry)
  uint32_t wf_budget = 0, wf_passes = 2, wf_bpc = 0;
  uint32_t cus = 0;              // compute units of the device
  bool gen_film = false;         // generated modules also hold the film and G-buffer kernels
  std::mutex gen_mu;
  std::condition_variable gen_cv;
  std::deque<std::shared_ptr<mm3e::GenJob>> gen_queue;
  std::vector<std::thread> gen_threads;
  bool gen_stop = false;
};

struct mm3e_scene {
  mm3e_renderer* r = nullptr;
  mm3e::HostScene host;
  uint32_t w = 0, h = 0, aa = 1, bounces = 0, tw = 16, th = 16;
  uint8_t* d_tables = nullptr;   // DevScene + tables
  size_t tables_cap = 0;
  uint8_t* d_dyn = nullptr;      // tables with dynamic spheres appended
  size_t dyn_cap = 0;
  const mm3e::DevScene* d_scene = nullptr;
  mm3e::DevScene hdr{};
  float* d_gi = nullptr;         // [probe][6][4]
  float* d_gi_in_blob = nullptr;
  bool has_gi = false;
  uint32_t gi_dims[3] = {0, 0, 0};
  float gi_min[3] = {0, 0, 0}, gi_size[3] = {0, 0, 0};
  mm3e::PostParams pp{};
  float* d_kern = nullptr;
  float* d_hdr = nullptr;
  float* d_tmp = nullptr;
  uchar4* d_rgba = nullptr;
  uint8_t* h_rgba = nullptr;     // pinned staging
  static constexpr int kRing = 128;   // frames whose kernel times are kept
  hipEvent_t ev[kRing][3] = {};
  uint64_t frames = 0;                 // frames launched since the last timing reset
  bool timing_valid = false;
  uint32_t stage_bytes = 0;            // LDS bytes of the staged tables (0: read from HBM)
  uint32_t dyn_stage_bytes = 0;        // the same for the tables with dynamic spheres appended
  bool general = false;                // the scene needs the general kernels (segment grid / CSG)
  bool dyn_general = false;            // the same for the tables with dynamic spheres appended
  std::shared_ptr<mm3e::GenJob> gen;   // generated kernel (null when code generation is off)
  // tile order of the generated kernel: the frame's tiles most expensive first (costs from the
  // previous frame), their costs, the ticket counter
  uint32_t* d_to = nullptr;            // [order n_tiles][cost n_tiles][ticket]
  uint32_t to_tiles = 0;
  bool to_on = false;
  uint32_t gen_stage = 0;              // LDS bytes the generated kernel stages (0: none needed)
  // queued-frame buffers (allocated on the first queued frame, resized with the sample count)
  uint8_t* d_wf = nullptr;
  uint32_t wf_cap = 0;
  // the scene's own queued-frame setting (budget, passes, blocks per CU; from the autotuner or
  // mm3e_scene_set_wavefront) when wf_own, else the renderer's
  bool wf_own = false;
  uint32_t wf_set[3] = {0, 0, 0};
};

namespace mm3e {
void set_err(char* err, size_t n, const std::string& s);
bool hip_ok(hipError_t e, const char* what, std::string* msg);
FrameParams make_frame(const mm3e_camera& c, uint32_t w, uint32_t h);
void tile_shape(uint32_t aa, uint32_t* tw, uint32_t* th);
// Largest scene (objects) that gets a generated kernel.
size_t codegen_object_limit();
// The generated render kernel over a frame's tiles in index order (one block per tile).
hipError_t launch_generated_index(hipFunction_t fn, const DevScene* sc, const FrameParams& fp, float* hdr, uint32_t tw,
                                  uint32_t th, uint32_t aa, uint32_t stage_bytes, hipStream_t st);
// The scene's generated optional kernel `k` (GenOpt order) when its module is ready, else null.
hipFunction_t gen_opt(const mm3e_scene* s, int k);
// G-buffer pass of a scene (its generated kernel when ready, else the table kernels): the
// frame (list == nullptr) or the *n_list listed pixels (at most max_list).
hipError_t scene_gbuffer(const mm3e_scene* s, hipStream_t st, const FrameParams& fp, const GMover* movers,
                         uint32_t n_movers, const uint32_t* list, const uint32_t* n_list, uint32_t max_list,
                         uchar4* color, float* depth, int32_t* obj);
}  // namespace mm3e
