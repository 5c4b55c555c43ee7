// wavefront.hip — the depth-independent stages of the queued frame (dev_wavefront.h): counter
// reset, the budgeted primary march passes for both kernel families, and the per-pixel resolve.
// The shading stage depends on the reflection depth and lives in render_kernel.hip.
#include <hip/hip_runtime.h>

#include "dev_wavefront.h"
#include "launch.h"

using namespace mm3e;

extern "C" __global__ void __launch_bounds__(64) mm3e_wf_reset(uint32_t* __restrict__ cnt) {
  if (threadIdx.x <= WF_MAX_PASSES) cnt[threadIdx.x] = 0u;
}

#define MM3E_WF_MARCH(NAME, MAP)                                                                              \
  extern "C" __global__ void __launch_bounds__(256) NAME##_first(const DevScene* __restrict__ sc, FrameParams fp, \
                                                                 WfFrame w, uint32_t budget, uint32_t stage_bytes) { \
    extern __shared__ uint4 tables_lds[];                                                                     \
    const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);            \
    wf_march_first<MAP>(sc, tb, fp, w, budget);                                                               \
  }                                                                                                           \
  extern "C" __global__ void __launch_bounds__(256) NAME##_cont(const DevScene* __restrict__ sc, FrameParams fp,  \
                                                                WfFrame w, uint32_t k, uint32_t budget,          \
                                                                uint32_t stage_bytes) {                          \
    extern __shared__ uint4 tables_lds[];                                                                     \
    const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);            \
    wf_march_cont<MAP>(sc, tb, fp, w, k, budget);                                                             \
  }

MM3E_WF_MARCH(mm3e_wf_march, SimpleMap)
MM3E_WF_MARCH(mm3e_wf_march_x, TableMap)
#undef MM3E_WF_MARCH

// Pixel (px, py): mean of its sample slots in sample order, as render_tile sums them.
extern "C" __global__ void __launch_bounds__(256) mm3e_wf_resolve(FrameParams fp, WfFrame w, uint32_t aa,
                                                                   float* __restrict__ hdr) {
  const uint32_t px = blockIdx.x * 16u + (threadIdx.x & 15u), py = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (px >= fp.width || py >= fp.height) return;
  const uint32_t ns = aa * aa;
  const uint32_t tile = (py / w.th) * w.tiles_x + px / w.tw;
  const uint32_t pl = (py % w.th) * w.tw + px % w.tw;
  const uint32_t g0 = tile * (w.tw * w.th * ns) + pl * ns;
  V3 acc = V3{0.0f, 0.0f, 0.0f};
  for (uint32_t k = 0; k < ns; k++)
    acc = acc + V3{w.rad[g0 + k], w.rad[(size_t)w.cap + g0 + k], w.rad[2 * (size_t)w.cap + g0 + k]};
  const float inv = 1.0f / (float)ns;
  acc = scale(acc, inv);
  float* o = hdr + 3 * ((size_t)py * fp.width + px);
  o[0] = acc.x;
  o[1] = acc.y;
  o[2] = acc.z;
}

namespace mm3e {

hipError_t launch_wf_reset(hipStream_t st, uint32_t* cnt) {
  hipLaunchKernelGGL(mm3e_wf_reset, dim3(1), dim3(64), 0, st, cnt);
  return hipGetLastError();
}

hipError_t launch_wf_march(hipStream_t st, bool general, uint32_t k, dim3 grid, dim3 block, const DevScene* sc,
                           uint32_t stage_bytes, const FrameParams& fp, const WfFrame& w, uint32_t budget) {
  if (k == 0) {
    if (general) hipLaunchKernelGGL(mm3e_wf_march_x_first, grid, block, stage_bytes, st, sc, fp, w, budget, stage_bytes);
    else hipLaunchKernelGGL(mm3e_wf_march_first, grid, block, stage_bytes, st, sc, fp, w, budget, stage_bytes);
  } else {
    if (general) hipLaunchKernelGGL(mm3e_wf_march_x_cont, grid, block, stage_bytes, st, sc, fp, w, k, budget, stage_bytes);
    else hipLaunchKernelGGL(mm3e_wf_march_cont, grid, block, stage_bytes, st, sc, fp, w, k, budget, stage_bytes);
  }
  return hipGetLastError();
}

hipError_t launch_wf_shade(uint32_t maxb, dim3 grid, hipStream_t st, const DevScene* sc, uint32_t stage_bytes,
                           bool general, const FrameParams& fp, const WfFrame& w) {
#define MM3E_WS(B) launch_wf_shade_b##B(grid, st, sc, stage_bytes, general, fp, w)
  switch (maxb) {
    case 0: return MM3E_WS(0);
    case 1: return MM3E_WS(1);
    case 2: return MM3E_WS(2);
    case 3: return MM3E_WS(3);
    case 4: return MM3E_WS(4);
    default: return maxb <= kMaxBounces ? MM3E_WS(16) : hipErrorInvalidValue;
  }
#undef MM3E_WS
}

hipError_t launch_wf_resolve(hipStream_t st, const FrameParams& fp, const WfFrame& w, uint32_t aa, float* hdr) {
  hipLaunchKernelGGL(mm3e_wf_resolve, dim3((fp.width + 15) / 16, (fp.height + 15) / 16), dim3(256), 0, st, fp, w, aa,
                     hdr);
  return hipGetLastError();
}

}  // namespace mm3e
