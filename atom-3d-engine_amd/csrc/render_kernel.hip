// render_kernel.hip — the per-pixel render kernel over the table-driven world field for one
// reflection depth.  Compiled once per MM3E_MAXB (0..4) into separate objects so the variants
// build in parallel; the runtime picks the variant equal to the scene's `bounces`.
#include <hip/hip_runtime.h>

#include "dev_render.h"
#include "launch.h"

#ifndef MM3E_MAXB
#error "MM3E_MAXB must be defined"
#endif

#define MM3E_CAT2(a, b) a##b
#define MM3E_CAT(a, b) MM3E_CAT2(a, b)
#define KNAME MM3E_CAT(mm3e_render_table_b, MM3E_MAXB)
#define KNAME_LDS MM3E_CAT(mm3e_render_lds_b, MM3E_MAXB)
#define LNAME MM3E_CAT(launch_render_table_b, MM3E_MAXB)

using namespace mm3e;

// Scene tables read from HBM (scenes whose tables exceed the LDS budget).
extern "C" __global__ void __launch_bounds__(256)
    KNAME(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 3];
  render_tile<TableMap, MM3E_MAXB>(sc, global_tables(sc), fp, hdr, tw, th, blockIdx.x, blockIdx.y, smem);
}

// Scene tables staged in LDS by each block (dynamic LDS = DevScene::stage_bytes).
extern "C" __global__ void __launch_bounds__(256)
    KNAME_LDS(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th,
              uint32_t stage_bytes) {
  __shared__ float smem[256 * 3];
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_tables(sc, tables_lds, stage_bytes);
  render_tile<TableMap, MM3E_MAXB>(sc, tb, fp, hdr, tw, th, blockIdx.x, blockIdx.y, smem);
}

namespace mm3e {
hipError_t LNAME(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, uint32_t stage_bytes,
                 const FrameParams& fp, float* hdr, uint32_t tw, uint32_t th) {
  if (stage_bytes) hipLaunchKernelGGL(KNAME_LDS, grid, block, stage_bytes, st, sc, fp, hdr, tw, th, stage_bytes);
  else hipLaunchKernelGGL(KNAME, grid, block, 0, st, sc, fp, hdr, tw, th);
  return hipGetLastError();
}
}  // namespace mm3e
