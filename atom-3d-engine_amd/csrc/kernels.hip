// kernels.hip — ahead-of-time gfx950 kernels over the table-driven world field (GI bake,
// post, parity probes; the render variants live in render_kernel.hip), plus the host-side
// launch helpers the C++ runtime calls.  Built with
//   hipcc --offload-arch=gfx950 -O3 -ffp-contract=off
//         -Xclang -target-feature -Xclang -packed-fp32-ops
#include <hip/hip_runtime.h>

#include "batch.h"
#include "dev_render.h"
#include "launch.h"

using namespace mm3e;

extern "C" __global__ void __launch_bounds__(256)
    mm3e_gi_gather_table(const DevScene* __restrict__ sc, GiBakeParams bp, float4* __restrict__ out,
                         uint32_t first_probe, uint32_t n_probes) {
  gi_gather_rays<TableMap>(sc, global_tables(sc), bp, out, first_probe, n_probes);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_gi_gather_lds(const DevScene* __restrict__ sc, GiBakeParams bp, float4* __restrict__ out,
                       uint32_t first_probe, uint32_t n_probes, uint32_t stage_bytes) {
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_tables(sc, tables_lds, stage_bytes);
  gi_gather_rays<TableMap>(sc, tb, bp, out, first_probe, n_probes);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_gi_reduce(GiBakeParams bp, const float4* __restrict__ rays, float4* __restrict__ cubes, uint32_t n_probes) {
  gi_reduce(bp, rays, cubes, n_probes);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_bloom_h(PostParams pp, const float* __restrict__ kern, const float* __restrict__ hdr, float* __restrict__ tmp) {
  extern __shared__ float lds[];
  bloom_h(pp, kern, hdr, tmp, lds, blockIdx.x * blockDim.x, blockIdx.y);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_bloom_h_wide(PostParams pp, const float* __restrict__ kern, const float* __restrict__ hdr,
                      float* __restrict__ tmp) {
  bloom_h_wide(pp, kern, hdr, tmp, blockIdx.x * blockDim.x, blockIdx.y);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_bloom_v_resolve(PostParams pp, const float* __restrict__ kern, const float* __restrict__ hdr,
                         const float* __restrict__ tmp, uchar4* __restrict__ rgba) {
  bloom_v_resolve(pp, kern, hdr, tmp, rgba, blockIdx.x * blockDim.x, blockIdx.y);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_resolve_only(PostParams pp, const float* __restrict__ hdr, uchar4* __restrict__ rgba) {
  resolve_only(pp, hdr, rgba, blockIdx.x * blockDim.x, blockIdx.y);
}

// Batched post: blockIdx.y enumerates the rows of all jobs (BatchPost::row_base), blockIdx.x
// the 256-pixel segments of the widest job; blocks past a job's width leave at once.
extern "C" __global__ void __launch_bounds__(256)
    mm3e_batch_bloom_h(const BatchPost* __restrict__ posts, uint32_t n) {
  extern __shared__ float lds[];
  const uint32_t row = blockIdx.y;
  const BatchPost& P = posts[find_job(posts, n, row, [](const BatchPost& b) { return b.row_base; })];
  const int x0 = blockIdx.x * blockDim.x;
  if (!P.pp.bloom || x0 >= (int)P.pp.width) return;  // uniform per block
  bloom_h(P.pp, P.kern, P.hdr, P.tmp, lds, x0, (int)(row - P.row_base));
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_batch_bloom_h_wide(const BatchPost* __restrict__ posts, uint32_t n) {
  const uint32_t row = blockIdx.y;
  const BatchPost& P = posts[find_job(posts, n, row, [](const BatchPost& b) { return b.row_base; })];
  const int x0 = blockIdx.x * blockDim.x;
  if (!P.pp.bloom || x0 >= (int)P.pp.width) return;  // uniform per block
  bloom_h_wide(P.pp, P.kern, P.hdr, P.tmp, x0, (int)(row - P.row_base));
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_batch_resolve(const BatchPost* __restrict__ posts, uint32_t n) {
  const uint32_t row = blockIdx.y;
  const BatchPost& P = posts[find_job(posts, n, row, [](const BatchPost& b) { return b.row_base; })];
  const int x0 = blockIdx.x * blockDim.x;
  if (x0 >= (int)P.pp.width) return;
  const int y = (int)(row - P.row_base);
  if (P.pp.bloom) bloom_v_resolve(P.pp, P.kern, P.hdr, P.tmp, P.rgba, x0, y);
  else resolve_only(P.pp, P.hdr, P.rgba, x0, y);
}

// Queue reset as a kernel (no memset node inside captured graphs).
extern "C" __global__ void mm3e_batch_reset(uint32_t* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_primary_aov_table(const DevScene* __restrict__ sc, FrameParams fp, uint8_t* __restrict__ hit,
                           uint32_t* __restrict__ mat, float* __restrict__ depth, uint32_t* __restrict__ steps,
                           float* __restrict__ nrm) {
  primary_aov<TableMap>(sc, fp, hit, mat, depth, steps, nrm);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_field_probe_table(const DevScene* __restrict__ sc, const float* __restrict__ pts, uint32_t n,
                           float* __restrict__ dist, uint32_t* __restrict__ mat) {
  field_probe<TableMap>(sc, pts, n, dist, mat);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_aov_mode_table(const DevScene* __restrict__ sc, FrameParams fp, uchar4* __restrict__ rgba, uint32_t mode) {
  aov_mode_pixel<TableMap>(sc, fp, rgba, mode);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_march_probe_table(const DevScene* __restrict__ sc, const float* __restrict__ rays, uint32_t n,
                           float* __restrict__ t, uint8_t* __restrict__ hit, uint32_t* __restrict__ mat,
                           float* __restrict__ nrm, uint32_t* __restrict__ steps) {
  march_probe<TableMap>(sc, rays, n, t, hit, mat, nrm, steps);
}

// ------------------------------------------------------------------------------------------
// launch helpers (host)
// ------------------------------------------------------------------------------------------
namespace mm3e {

hipError_t launch_render_table(uint32_t maxb, dim3 grid, dim3 block, hipStream_t st, const DevScene* sc,
                               uint32_t stage_bytes, bool general, const FrameParams& fp, float* hdr, uint32_t tw,
                               uint32_t th) {
  switch (maxb) {
    case 0: return launch_render_table_b0(grid, block, st, sc, stage_bytes, general, fp, hdr, tw, th);
    case 1: return launch_render_table_b1(grid, block, st, sc, stage_bytes, general, fp, hdr, tw, th);
    case 2: return launch_render_table_b2(grid, block, st, sc, stage_bytes, general, fp, hdr, tw, th);
    case 3: return launch_render_table_b3(grid, block, st, sc, stage_bytes, general, fp, hdr, tw, th);
    case 4: return launch_render_table_b4(grid, block, st, sc, stage_bytes, general, fp, hdr, tw, th);
    default:
      return maxb <= kMaxBounces ? launch_render_table_b16(grid, block, st, sc, stage_bytes, general, fp, hdr, tw, th)
                                 : hipErrorInvalidValue;
  }
}

hipError_t launch_render_ordered(uint32_t maxb, uint32_t n_tiles, dim3 block, hipStream_t st, const DevScene* sc,
                                 uint32_t stage_bytes, bool general, const FrameParams& fp, float* hdr, uint32_t tw,
                                 uint32_t th, const TileOrder& o) {
#define MM3E_ORD(B) launch_render_ordered_b##B(n_tiles, block, st, sc, stage_bytes, general, fp, hdr, tw, th, o)
  switch (maxb) {
    case 0: return MM3E_ORD(0);
    case 1: return MM3E_ORD(1);
    case 2: return MM3E_ORD(2);
    case 3: return MM3E_ORD(3);
    case 4: return MM3E_ORD(4);
    default: return maxb <= kMaxBounces ? MM3E_ORD(16) : hipErrorInvalidValue;
  }
#undef MM3E_ORD
}

hipError_t launch_film(uint32_t maxb, dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, bool general,
                       const FrameParams& fp, const FilmParams& lp, float* beauty, float* fg, float* alpha,
                       uint32_t tw, uint32_t th) {
  switch (maxb) {
    case 0: return launch_film_b0(grid, block, st, sc, general, fp, lp, beauty, fg, alpha, tw, th);
    case 1: return launch_film_b1(grid, block, st, sc, general, fp, lp, beauty, fg, alpha, tw, th);
    case 2: return launch_film_b2(grid, block, st, sc, general, fp, lp, beauty, fg, alpha, tw, th);
    case 3: return launch_film_b3(grid, block, st, sc, general, fp, lp, beauty, fg, alpha, tw, th);
    case 4: return launch_film_b4(grid, block, st, sc, general, fp, lp, beauty, fg, alpha, tw, th);
    default:
      return maxb <= kMaxBounces ? launch_film_b16(grid, block, st, sc, general, fp, lp, beauty, fg, alpha, tw, th)
                                 : hipErrorInvalidValue;
  }
}

hipError_t launch_gi_gather_table(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, uint32_t stage_bytes,
                                  const GiBakeParams& bp, float4* out, uint32_t first_probe, uint32_t n_probes) {
  if (stage_bytes)
    hipLaunchKernelGGL(mm3e_gi_gather_lds, grid, block, stage_bytes, st, sc, bp, out, first_probe, n_probes, stage_bytes);
  else hipLaunchKernelGGL(mm3e_gi_gather_table, grid, block, 0, st, sc, bp, out, first_probe, n_probes);
  return hipGetLastError();
}

hipError_t launch_gi_reduce(dim3 grid, dim3 block, hipStream_t st, const GiBakeParams& bp, const float4* rays,
                            float4* cubes, uint32_t n_probes) {
  hipLaunchKernelGGL(mm3e_gi_reduce, grid, block, 0, st, bp, rays, cubes, n_probes);
  return hipGetLastError();
}

// LDS of the horizontal bloom pass: a 256-pixel row segment and its halo; wider radii
// (more than 2,602 pixels) read their taps from the cache instead (mm3e_bloom_h_wide).
static const size_t kBloomLdsMax = 64 * 1024;

hipError_t launch_post(hipStream_t st, const PostParams& pp, const float* kern, const float* hdr, float* tmp,
                       uchar4* rgba) {
  const uint32_t bx = 256;
  dim3 grid((pp.width + bx - 1) / bx, pp.height);
  if (pp.bloom) {
    const size_t lds = (size_t)(bx + 2 * pp.radius) * 3 * sizeof(float);
    if (lds <= kBloomLdsMax) hipLaunchKernelGGL(mm3e_bloom_h, grid, dim3(bx), lds, st, pp, kern, hdr, tmp);
    else hipLaunchKernelGGL(mm3e_bloom_h_wide, grid, dim3(bx), 0, st, pp, kern, hdr, tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mm3e_bloom_v_resolve, grid, dim3(bx), 0, st, pp, kern, hdr, tmp, rgba);
  } else {
    hipLaunchKernelGGL(mm3e_resolve_only, grid, dim3(bx), 0, st, pp, hdr, rgba);
  }
  return hipGetLastError();
}

hipError_t launch_batch_post(hipStream_t st, const BatchPost* posts, uint32_t n, uint32_t total_rows, uint32_t max_w,
                             uint32_t max_radius) {
  const uint32_t bx = 256;
  dim3 grid((max_w + bx - 1) / bx, total_rows);
  const size_t lds = (size_t)(bx + 2 * max_radius) * 3 * sizeof(float);
  if (lds <= kBloomLdsMax) hipLaunchKernelGGL(mm3e_batch_bloom_h, grid, dim3(bx), lds, st, posts, n);
  else hipLaunchKernelGGL(mm3e_batch_bloom_h_wide, grid, dim3(bx), 0, st, posts, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mm3e_batch_resolve, grid, dim3(bx), 0, st, posts, n);
  return hipGetLastError();
}

hipError_t launch_batch_reset(hipStream_t st, uint32_t* counter) {
  hipLaunchKernelGGL(mm3e_batch_reset, dim3(1), dim3(64), 0, st, counter);
  return hipGetLastError();
}

hipError_t launch_batch_render(uint32_t maxb, uint32_t blocks, hipStream_t st, const BatchJob* jobs, uint32_t n_jobs,
                               uint32_t total_tiles, uint32_t* counter, uint32_t lds_bytes, const TileOrder& q) {
  switch (maxb) {
    case 0:
    case 1: return launch_batch_render_b1(blocks, st, jobs, n_jobs, total_tiles, counter, lds_bytes, q);
    case 2: return launch_batch_render_b2(blocks, st, jobs, n_jobs, total_tiles, counter, lds_bytes, q);
    case 3:
    case 4: return launch_batch_render_b4(blocks, st, jobs, n_jobs, total_tiles, counter, lds_bytes, q);
    default:
      return maxb <= kMaxBounces ? launch_batch_render_b16(blocks, st, jobs, n_jobs, total_tiles, counter, lds_bytes, q)
                                 : hipErrorInvalidValue;
  }
}

hipError_t launch_primary_aov_table(hipStream_t st, const DevScene* sc, const FrameParams& fp, uint8_t* hit,
                                    uint32_t* mat, float* depth, uint32_t* steps, float* nrm) {
  dim3 grid((fp.width + 255) / 256, fp.height);
  hipLaunchKernelGGL(mm3e_primary_aov_table, grid, dim3(256), 0, st, sc, fp, hit, mat, depth, steps, nrm);
  return hipGetLastError();
}

hipError_t launch_aov_mode_table(hipStream_t st, const DevScene* sc, const FrameParams& fp, uchar4* rgba,
                                 uint32_t mode) {
  hipLaunchKernelGGL(mm3e_aov_mode_table, dim3((fp.width + 255) / 256, fp.height), dim3(256), 0, st, sc, fp, rgba,
                     mode);
  return hipGetLastError();
}

hipError_t launch_field_probe_table(hipStream_t st, const DevScene* sc, const float* pts, uint32_t n, float* dist,
                                    uint32_t* mat) {
  hipLaunchKernelGGL(mm3e_field_probe_table, dim3((n + 255) / 256), dim3(256), 0, st, sc, pts, n, dist, mat);
  return hipGetLastError();
}

hipError_t launch_march_probe_table(hipStream_t st, const DevScene* sc, const float* rays, uint32_t n, float* t,
                                    uint8_t* hit, uint32_t* mat, float* nrm, uint32_t* steps) {
  hipLaunchKernelGGL(mm3e_march_probe_table, dim3((n + 255) / 256), dim3(256), 0, st, sc, rays, n, t, hit, mat, nrm,
                     steps);
  return hipGetLastError();
}

}  // namespace mm3e

// ---------------------------------------------------------------------------------------
// Reprojection of a G-buffer frame into a new camera (reproject.rs:57-170)
// ---------------------------------------------------------------------------------------
namespace mm3e {

// reproject.rs:57-69: screen position and forward depth of a world point, false behind the camera.
MM3E_DEV bool project_px(const FrameParams& c, V3 world, float& sx, float& sy, float& fz) {
  const V3 v = world - ld3(c.eye);
  fz = dot(v, ld3(c.fwd));
  if (fz <= 1e-4f) return false;
  const float aspect = (float)c.width / (float)c.height;
  const float ndc_x = dot(v, ld3(c.right)) / fz;
  const float ndc_y = dot(v, ld3(c.up)) / fz;
  sx = (ndc_x / (aspect * c.fov_scale) + 1.0f) * 0.5f * (float)c.width;
  sy = (1.0f - ndc_y / c.fov_scale) * 0.5f * (float)c.height;
  return true;
}

MM3E_DEV int32_t sat_i32(float f) {  // Rust `f32 as i32`: saturating, NaN -> 0
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f < -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)f;
}

}  // namespace mm3e

extern "C" __global__ void __launch_bounds__(256) mm3e_warp_reset(unsigned long long* __restrict__ keys, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) keys[i] = ~0ull;
}

// One lane per source pixel: its world point (depth, or 1e6 for sky), advanced by its mover's
// delta, projected into the new camera; the target keeps the nearest point, ties to the lowest
// source index (the reference's row-major loop with a strict `<`), as one 64-bit atomic min of
// (forward-depth bits, source index): positive f32 bit patterns order as the values.
extern "C" __global__ void __launch_bounds__(256)
    mm3e_warp_splat(uint32_t w, uint32_t h, FrameParams prev, FrameParams next, const float* __restrict__ depth,
                    const int32_t* __restrict__ obj, const float* __restrict__ deltas, uint32_t n_deltas,
                    unsigned long long* __restrict__ keys) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= w * h) return;
  const uint32_t x = i % w, y = i / w;
  V3 eye, dir;
  camera_ray(prev, (float)x + 0.5f, (float)y + 0.5f, eye, dir);
  const float dp = depth[i];
  const float d = isfinite(dp) ? dp : 1.0e6f;
  V3 world = eye + scale(dir, d);
  const int32_t oi = obj[i];
  if (oi >= 0 && (uint32_t)oi < n_deltas)
    world = world + V3{deltas[3 * oi], deltas[3 * oi + 1], deltas[3 * oi + 2]};
  float sx, sy, fz;
  if (!project_px(next, world, sx, sy, fz)) return;
  const int32_t nx = sat_i32(floorf(sx)), ny = sat_i32(floorf(sy));
  if (nx < 0 || ny < 0 || (uint32_t)nx >= w || (uint32_t)ny >= h) return;
  if (!(fz < __builtin_inff())) return;  // `fz < zbuf[j]` with zbuf starting at +inf
  const size_t j = (size_t)ny * w + (uint32_t)nx;
  atomicMin(&keys[j], ((unsigned long long)__float_as_uint(fz) << 32) | (unsigned long long)i);
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_warp_resolve(uint32_t n, const unsigned long long* __restrict__ keys, const uchar4* __restrict__ color,
                      uchar4* __restrict__ out, uint8_t* __restrict__ filled) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= n) return;
  const unsigned long long k = keys[j];
  const bool f = k != ~0ull;
  out[j] = f ? color[(uint32_t)(k & 0xffffffffull)] : uchar4{0, 0, 0, 0};
  filled[j] = f ? 1 : 0;
}

// reproject.rs:141-168: per row, holes take the last filled pixel to their left, then the
// row's leading holes take the first filled pixel to their right.  After the first sweep every
// pixel from the row's first filled one on is filled, so the second sweep only copies that
// pixel into the leading holes; the warp's coverage mask is left as the warp made it.  One lane
// per row.
extern "C" __global__ void __launch_bounds__(256)
    mm3e_row_fill(uint32_t w, uint32_t h, uchar4* __restrict__ out, const uint8_t* __restrict__ filled) {
  const uint32_t y = blockIdx.x * 256u + threadIdx.x;
  if (y >= h) return;
  uchar4* row = out + (size_t)y * w;
  const uint8_t* fr = filled + (size_t)y * w;
  uint32_t first = w;
  uchar4 last = uchar4{0, 0, 0, 0};
  for (uint32_t x = 0; x < w; x++) {
    if (fr[x]) {
      last = row[x];
      if (first == w) first = x;
    } else if (first < w) {
      row[x] = last;
    }
  }
  if (first < w) {
    const uchar4 lead = row[first];
    for (uint32_t x = 0; x < first; x++) row[x] = lead;
  }
}

extern "C" __global__ void __launch_bounds__(256)
    mm3e_hole_list(uint32_t n, const uint8_t* __restrict__ filled, uint32_t* __restrict__ list,
                   uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n || filled[i]) return;
  list[atomicAdd(count, 1u)] = i;
}

namespace mm3e {

hipError_t launch_warp(hipStream_t st, uint32_t w, uint32_t h, const FrameParams& prev, const FrameParams& next,
                       const float* depth, const int32_t* obj, const float* deltas, uint32_t n_deltas,
                       const uchar4* color, unsigned long long* keys, uchar4* out, uint8_t* filled) {
  const uint32_t n = w * h;
  const dim3 g((n + 255) / 256);
  hipLaunchKernelGGL(mm3e_warp_reset, g, dim3(256), 0, st, keys, n);
  hipLaunchKernelGGL(mm3e_warp_splat, g, dim3(256), 0, st, w, h, prev, next, depth, obj, deltas, n_deltas, keys);
  hipLaunchKernelGGL(mm3e_warp_resolve, g, dim3(256), 0, st, n, keys, color, out, filled);
  return hipGetLastError();
}

hipError_t launch_row_fill(hipStream_t st, uint32_t w, uint32_t h, uchar4* out, const uint8_t* filled) {
  hipLaunchKernelGGL(mm3e_row_fill, dim3((h + 255) / 256), dim3(256), 0, st, w, h, out, filled);
  return hipGetLastError();
}

hipError_t launch_hole_list(hipStream_t st, uint32_t n, const uint8_t* filled, uint32_t* list, uint32_t* count) {
  hipLaunchKernelGGL(mm3e_hole_list, dim3((n + 255) / 256), dim3(256), 0, st, n, filled, list, count);
  return hipGetLastError();
}

hipError_t launch_gbuffer(uint32_t maxb, hipStream_t st, const DevScene* sc, bool general, const FrameParams& fp,
                          float exposure, const GMover* movers, uint32_t n_movers, const uint32_t* list,
                          const uint32_t* n_list, uint32_t max_list, uchar4* color, float* depth, int32_t* obj) {
#define MM3E_GB(B) \
  launch_gbuffer_b##B(st, sc, general, fp, exposure, movers, n_movers, list, n_list, max_list, color, depth, obj)
  switch (maxb) {
    case 0: return MM3E_GB(0);
    case 1: return MM3E_GB(1);
    case 2: return MM3E_GB(2);
    case 3: return MM3E_GB(3);
    case 4: return MM3E_GB(4);
    default: return maxb <= kMaxBounces ? MM3E_GB(16) : hipErrorInvalidValue;
  }
#undef MM3E_GB
}

}  // namespace mm3e

// ---------------------------------------------------------------------------------------
// Tile order for the next frame of an ordered-tiles render: tiles sorted into 128 buckets of
// descending cost (4 per power of two of the ticket count), expensive buckets first; the order
// inside a bucket is arbitrary (the image does not depend on it).  One block.
// ---------------------------------------------------------------------------------------
namespace mm3e {
MM3E_DEV uint32_t cost_bucket(uint32_t c) {
  if (c == 0) return 127u;
  const uint32_t e = 31u - (uint32_t)__builtin_clz(c);
  const uint32_t m = e >= 2 ? (c >> (e - 2)) & 3u : 0u;
  return 127u - (e * 4u + m);
}
}  // namespace mm3e

extern "C" __global__ void __launch_bounds__(1024)
    mm3e_tile_order(const uint32_t* __restrict__ cost, uint32_t n, uint32_t* __restrict__ order) {
  __shared__ uint32_t hist[128];
  const uint32_t t = threadIdx.x;
  if (t < 128) hist[t] = 0;
  __syncthreads();
  for (uint32_t i = t; i < n; i += 1024) atomicAdd(&hist[cost_bucket(cost[i])], 1u);
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (uint32_t b = 0; b < 128; b++) {
      const uint32_t c = hist[b];
      hist[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < n; i += 1024) order[atomicAdd(&hist[cost_bucket(cost[i])], 1u)] = i;
}

namespace mm3e {
hipError_t launch_tile_order(hipStream_t st, const uint32_t* cost, uint32_t n, uint32_t* order) {
  hipLaunchKernelGGL(mm3e_tile_order, dim3(1), dim3(1024), 0, st, cost, n, order);
  return hipGetLastError();
}
}  // namespace mm3e

// ---------------------------------------------------------------------------------------
// Self-test of crsqrt against the compiler's correctly rounded sqrtf over every 32-bit pattern
// (NaNs compare equal to NaNs).  Each lane checks a contiguous run of 256 patterns.
// ---------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256) mm3e_sqrt_selftest(unsigned long long* __restrict__ mismatches) {
  const uint64_t lo = ((uint64_t)blockIdx.x * 256u + threadIdx.x) * 256u;
  uint32_t bad = 0;
  for (uint32_t k = 0; k < 256; k++) {
    const uint64_t u = lo + k;
    if (u > 0xffffffffull) break;
    const float x = __uint_as_float((uint32_t)u);
    const float a = __builtin_sqrtf(x), b = crsqrt(x);
    if (__float_as_uint(a) != __float_as_uint(b) && !(a != a && b != b)) bad++;
  }
  if (bad) atomicAdd(mismatches, (unsigned long long)bad);
}

namespace mm3e {
hipError_t launch_sqrt_selftest(hipStream_t st, unsigned long long* mismatches) {
  hipLaunchKernelGGL(mm3e_sqrt_selftest, dim3(65536), dim3(256), 0, st, mismatches);
  return hipGetLastError();
}
}  // namespace mm3e
