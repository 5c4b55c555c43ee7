// render_kernel.hip — the per-pixel render kernels for one reflection depth.  Compiled once per
// MM3E_MAXB (0..4, and 16 for 5..16 bounces) into separate objects so the variants build in
// parallel; the runtime picks the variant equal to the scene's `bounces` (16 above 4), and per
// scene one of two families (the deep variant has only the state machine, whose pending
// reflection levels wait in a per-lane stack instead of nested inlined copies):
//   mm3e_render_{table,lds}_bN     scenes without a segment grid or CSG objects: the world fold
//                                  without those features, march/shade as nested loops;
//   mm3e_render_{table,lds}_x_bN   every other scene: the general fold (segment grid, CSG) and
//                                  the single-evaluation-site state machine.
#include <hip/hip_runtime.h>

#include "dev_render.h"
#include "dev_wavefront.h"
#include "launch.h"

#ifndef MM3E_MAXB
#error "MM3E_MAXB must be defined"
#endif

#define MM3E_CAT2(a, b) a##b
#define MM3E_CAT(a, b) MM3E_CAT2(a, b)
#define KNAME MM3E_CAT(mm3e_render_table_b, MM3E_MAXB)
#define KNAME_LDS MM3E_CAT(mm3e_render_lds_b, MM3E_MAXB)
#define KNAME_X MM3E_CAT(mm3e_render_table_x_b, MM3E_MAXB)
#define KNAME_LDS_X MM3E_CAT(mm3e_render_lds_x_b, MM3E_MAXB)
#define LNAME MM3E_CAT(launch_render_table_b, MM3E_MAXB)
#define ONAME MM3E_CAT(mm3e_render_ord_b, MM3E_MAXB)
#define ONAME_X MM3E_CAT(mm3e_render_ord_x_b, MM3E_MAXB)
#define OLNAME MM3E_CAT(launch_render_ordered_b, MM3E_MAXB)
#define WNAME MM3E_CAT(mm3e_wf_shade_b, MM3E_MAXB)
#define WNAME_X MM3E_CAT(mm3e_wf_shade_x_b, MM3E_MAXB)
#define WLNAME MM3E_CAT(launch_wf_shade_b, MM3E_MAXB)

using namespace mm3e;

#if MM3E_MAXB <= 4
// Scene tables read from HBM (scenes whose tables exceed the LDS budget).
extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 3];
  render_tile<SimpleMap, MM3E_MAXB, false>(sc, global_tables(sc), fp, hdr, tw, th, blockIdx.x, blockIdx.y, smem);
}

// Scene tables staged in LDS by each block (dynamic LDS = DevScene::stage_bytes).
extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME_LDS(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th,
              uint32_t stage_bytes) {
  __shared__ float smem[256 * 3];
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_tables(sc, tables_lds, stage_bytes);
  render_tile<SimpleMap, MM3E_MAXB, false>(sc, tb, fp, hdr, tw, th, blockIdx.x, blockIdx.y, smem);
}

#endif

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME_X(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 3];
  render_tile<TableMap, MM3E_MAXB, true>(sc, global_tables(sc), fp, hdr, tw, th, blockIdx.x, blockIdx.y, smem);
}

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME_LDS_X(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th,
                uint32_t stage_bytes) {
  __shared__ float smem[256 * 3];
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_tables(sc, tables_lds, stage_bytes);
  render_tile<TableMap, MM3E_MAXB, true>(sc, tb, fp, hdr, tw, th, blockIdx.x, blockIdx.y, smem);
}

#if MM3E_MAXB <= 4
// Ordered tiles (block b renders tile o.order[b]; see render_ordered), tables staged in LDS
// when stage_bytes > 0, else read from HBM.
extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    ONAME(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th,
          uint32_t stage_bytes, TileOrder o) {
  __shared__ float smem[256 * 3];
  __shared__ uint32_t start;
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);
  render_ordered<SimpleMap, MM3E_MAXB, false>(sc, tb, fp, hdr, tw, th, o, smem, &start);
}

#endif

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    ONAME_X(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw, uint32_t th,
            uint32_t stage_bytes, TileOrder o) {
  __shared__ float smem[256 * 3];
  __shared__ uint32_t start;
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);
  render_ordered<TableMap, MM3E_MAXB, true>(sc, tb, fp, hdr, tw, th, o, smem, &start);
}

#if MM3E_MAXB <= 4
// Shading stage of the queued frame (dev_wavefront.h): the hit queue, nested loops or the
// per-lane state machine.
extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    WNAME(const DevScene* __restrict__ sc, FrameParams fp, WfFrame w, uint32_t stage_bytes) {
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);
  wf_shade<SimpleMap, MM3E_MAXB, false>(sc, tb, fp, w);
}

#endif

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    WNAME_X(const DevScene* __restrict__ sc, FrameParams fp, WfFrame w, uint32_t stage_bytes) {
  extern __shared__ uint4 tables_lds[];
  const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);
  wf_shade<TableMap, MM3E_MAXB, true>(sc, tb, fp, w);
}

namespace mm3e {
hipError_t WLNAME(dim3 grid, hipStream_t st, const DevScene* sc, uint32_t stage_bytes, bool general,
                  const FrameParams& fp, const WfFrame& w) {
#if MM3E_MAXB <= 4
  if (!general) {
    hipLaunchKernelGGL(WNAME, grid, dim3(256), stage_bytes, st, sc, fp, w, stage_bytes);
    return hipGetLastError();
  }
#endif
  hipLaunchKernelGGL(WNAME_X, grid, dim3(256), stage_bytes, st, sc, fp, w, stage_bytes);
  return hipGetLastError();
}

hipError_t OLNAME(uint32_t n_tiles, dim3 block, hipStream_t st, const DevScene* sc, uint32_t stage_bytes, bool general,
                  const FrameParams& fp, float* hdr, uint32_t tw, uint32_t th, const TileOrder& o) {
#if MM3E_MAXB <= 4
  if (!general) {
    hipLaunchKernelGGL(ONAME, dim3(n_tiles), block, stage_bytes, st, sc, fp, hdr, tw, th, stage_bytes, o);
    return hipGetLastError();
  }
#endif
  hipLaunchKernelGGL(ONAME_X, dim3(n_tiles), block, stage_bytes, st, sc, fp, hdr, tw, th, stage_bytes, o);
  return hipGetLastError();
}

hipError_t LNAME(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, uint32_t stage_bytes, bool general,
                 const FrameParams& fp, float* hdr, uint32_t tw, uint32_t th) {
#if MM3E_MAXB <= 4
  if (!general) {
    if (stage_bytes) hipLaunchKernelGGL(KNAME_LDS, grid, block, stage_bytes, st, sc, fp, hdr, tw, th, stage_bytes);
    else hipLaunchKernelGGL(KNAME, grid, block, 0, st, sc, fp, hdr, tw, th);
    return hipGetLastError();
  }
#endif
  if (stage_bytes) hipLaunchKernelGGL(KNAME_LDS_X, grid, block, stage_bytes, st, sc, fp, hdr, tw, th, stage_bytes);
  else hipLaunchKernelGGL(KNAME_X, grid, block, 0, st, sc, fp, hdr, tw, th);
  return hipGetLastError();
}
}  // namespace mm3e
