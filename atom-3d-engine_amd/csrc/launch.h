// launch.h — host-side launchers for the ahead-of-time kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_render.h"
#include "dev_wavefront.h"

namespace mm3e {

struct GiBakeParams;
struct PostParams;
struct BatchJob;
struct BatchPost;

#define MM3E_DECL_BATCH(B)                                                                                   \
  hipError_t launch_batch_render_b##B(uint32_t blocks, hipStream_t st, const BatchJob* jobs, uint32_t n_jobs, \
                                      uint32_t total_tiles, uint32_t* counter, uint32_t lds_bytes,            \
                                      const TileOrder& q);
MM3E_DECL_BATCH(1)
MM3E_DECL_BATCH(2)
MM3E_DECL_BATCH(4)
MM3E_DECL_BATCH(16)
#undef MM3E_DECL_BATCH
// Persistent batched render; maxb is the batch's largest `bounces` (variants 1, 2, 4, 16 cover
// 0..16).
// q.order == nullptr: tiles in index order, no cost record.
hipError_t launch_batch_render(uint32_t maxb, uint32_t blocks, hipStream_t st, const BatchJob* jobs, uint32_t n_jobs,
                               uint32_t total_tiles, uint32_t* counter, uint32_t lds_bytes,
                               const TileOrder& q = TileOrder{});
hipError_t launch_batch_post(hipStream_t st, const BatchPost* posts, uint32_t n, uint32_t total_rows, uint32_t max_w,
                             uint32_t max_radius);
hipError_t launch_batch_reset(hipStream_t st, uint32_t* counter);

#define MM3E_DECL_RENDER(B)                                                                          \
  hipError_t launch_render_table_b##B(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, \
                                      uint32_t stage_bytes, bool general, const FrameParams& fp, float* hdr, \
                                      uint32_t tw, uint32_t th);
MM3E_DECL_RENDER(0)
MM3E_DECL_RENDER(1)
MM3E_DECL_RENDER(2)
MM3E_DECL_RENDER(3)
MM3E_DECL_RENDER(4)
MM3E_DECL_RENDER(16)  // 5..16 bounces: the state machine only
#undef MM3E_DECL_RENDER

#define MM3E_DECL_ORD(B)                                                                                    \
  hipError_t launch_render_ordered_b##B(uint32_t n_tiles, dim3 block, hipStream_t st, const DevScene* sc,     \
                                        uint32_t stage_bytes, bool general, const FrameParams& fp, float* hdr, \
                                        uint32_t tw, uint32_t th, const TileOrder& o);
MM3E_DECL_ORD(0)
MM3E_DECL_ORD(1)
MM3E_DECL_ORD(2)
MM3E_DECL_ORD(3)
MM3E_DECL_ORD(4)
MM3E_DECL_ORD(16)  // 5..16 bounces: the state machine only
#undef MM3E_DECL_ORD
// The table kernels with ordered tiles (one block per tile, block b renders o.order[b]).
hipError_t launch_render_ordered(uint32_t maxb, uint32_t n_tiles, dim3 block, hipStream_t st, const DevScene* sc,
                                 uint32_t stage_bytes, bool general, const FrameParams& fp, float* hdr, uint32_t tw,
                                 uint32_t th, const TileOrder& o);
#define MM3E_DECL_WF(B)                                                                                       \
  hipError_t launch_wf_shade_b##B(dim3 grid, hipStream_t st, const DevScene* sc, uint32_t stage_bytes, bool general, \
                                  const FrameParams& fp, const WfFrame& w);
MM3E_DECL_WF(0)
MM3E_DECL_WF(1)
MM3E_DECL_WF(2)
MM3E_DECL_WF(3)
MM3E_DECL_WF(4)
MM3E_DECL_WF(16)  // 5..16 bounces: the state machine only
#undef MM3E_DECL_WF
// The queued frame (dev_wavefront.h): counter reset; march pass k (0: one block per tile of
// tw*th*aa*aa lanes; k >= 1: grid-stride over queue k-1); shading of the hit queue
// (grid-stride, 256 lanes a block); per-pixel resolve into hdr.
hipError_t launch_wf_reset(hipStream_t st, uint32_t* cnt);
hipError_t launch_wf_march(hipStream_t st, bool general, uint32_t k, dim3 grid, dim3 block, const DevScene* sc,
                           uint32_t stage_bytes, const FrameParams& fp, const WfFrame& w, uint32_t budget);
hipError_t launch_wf_shade(uint32_t maxb, dim3 grid, hipStream_t st, const DevScene* sc, uint32_t stage_bytes,
                           bool general, const FrameParams& fp, const WfFrame& w);
hipError_t launch_wf_resolve(hipStream_t st, const FrameParams& fp, const WfFrame& w, uint32_t aa, float* hdr);
// stage_bytes > 0 selects the variant that copies the scene tables into LDS first; general
// selects the kernels with the segment grid and CSG (scene_is_general).
hipError_t launch_render_table(uint32_t maxb, dim3 grid, dim3 block, hipStream_t st, const DevScene* sc,
                               uint32_t stage_bytes, bool general, const FrameParams& fp, float* hdr, uint32_t tw,
                               uint32_t th);
// Whether a compiled scene needs the general kernels: it carries a segment grid or CSG programs.
inline bool scene_is_general(const DevScene& h) { return h.grid != nullptr || h.csg != nullptr; }
hipError_t launch_gi_gather_table(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, uint32_t stage_bytes,
                                  const GiBakeParams& bp, float4* out, uint32_t first_probe, uint32_t n_probes);
hipError_t launch_gi_reduce(dim3 grid, dim3 block, hipStream_t st, const GiBakeParams& bp, const float4* rays,
                            float4* cubes, uint32_t n_probes);
hipError_t launch_post(hipStream_t st, const PostParams& pp, const float* kern, const float* hdr, float* tmp,
                       uchar4* rgba);
hipError_t launch_primary_aov_table(hipStream_t st, const DevScene* sc, const FrameParams& fp, uint8_t* hit,
                                    uint32_t* mat, float* depth, uint32_t* steps, float* nrm = nullptr);

#define MM3E_DECL_FILM(B)                                                                                       \
  hipError_t launch_film_b##B(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, bool general,          \
                              const FrameParams& fp, const FilmParams& lp, float* beauty, float* fg, float* alpha, \
                              uint32_t tw, uint32_t th);
MM3E_DECL_FILM(0)
MM3E_DECL_FILM(1)
MM3E_DECL_FILM(2)
MM3E_DECL_FILM(3)
MM3E_DECL_FILM(4)
MM3E_DECL_FILM(16)
#undef MM3E_DECL_FILM
struct GMover;
#define MM3E_DECL_GBUF(B)                                                                                          \
  hipError_t launch_gbuffer_b##B(hipStream_t st, const DevScene* sc, bool general, const FrameParams& fp,          \
                                 float exposure, const GMover* movers, uint32_t n_movers, const uint32_t* list,    \
                                 const uint32_t* n_list, uint32_t max_list, uchar4* color, float* depth,          \
                                 int32_t* obj);
MM3E_DECL_GBUF(0)
MM3E_DECL_GBUF(1)
MM3E_DECL_GBUF(2)
MM3E_DECL_GBUF(3)
MM3E_DECL_GBUF(4)
MM3E_DECL_GBUF(16)
#undef MM3E_DECL_GBUF
// G-buffer of a frame (list == nullptr) or the display colour of the listed pixels (*n_list of
// them in device memory, at most max_list).
hipError_t launch_gbuffer(uint32_t maxb, hipStream_t st, const DevScene* sc, bool general, const FrameParams& fp,
                          float exposure, const GMover* movers, uint32_t n_movers, const uint32_t* list,
                          const uint32_t* n_list, uint32_t max_list, uchar4* color, float* depth, int32_t* obj);
// Reprojection (reproject.rs:57-170): reset the z-buffer keys, splat the previous frame's
// pixels into the new camera, resolve the winners, optionally fill holes along rows, list holes.
hipError_t launch_warp(hipStream_t st, uint32_t w, uint32_t h, const FrameParams& prev, const FrameParams& next,
                       const float* depth, const int32_t* obj, const float* deltas, uint32_t n_deltas,
                       const uchar4* color, unsigned long long* keys, uchar4* out, uint8_t* filled);
hipError_t launch_row_fill(hipStream_t st, uint32_t w, uint32_t h, uchar4* out, const uint8_t* filled);
hipError_t launch_hole_list(hipStream_t st, uint32_t n, const uint8_t* filled, uint32_t* list, uint32_t* count);
// crsqrt against sqrtf over all 2^32 inputs; *mismatches must start at zero.
hipError_t launch_sqrt_selftest(hipStream_t st, unsigned long long* mismatches);
// Next frame's tile order from this frame's per-tile costs.
hipError_t launch_tile_order(hipStream_t st, const uint32_t* cost, uint32_t n, uint32_t* order);
// render_film's sample pass for a scene with `maxb` bounces (0..4).
hipError_t launch_film(uint32_t maxb, dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, bool general,
                       const FrameParams& fp, const FilmParams& lp, float* beauty, float* fg, float* alpha,
                       uint32_t tw, uint32_t th);
// Debug render mode (lib.rs:1178-1218) straight into RGBA8; mode 1..5.
hipError_t launch_aov_mode_table(hipStream_t st, const DevScene* sc, const FrameParams& fp, uchar4* rgba,
                                 uint32_t mode);
hipError_t launch_field_probe_table(hipStream_t st, const DevScene* sc, const float* pts, uint32_t n, float* dist,
                                    uint32_t* mat);
hipError_t launch_march_probe_table(hipStream_t st, const DevScene* sc, const float* rays, uint32_t n, float* t,
                                    uint8_t* hit, uint32_t* mat, float* nrm, uint32_t* steps);

}  // namespace mm3e
