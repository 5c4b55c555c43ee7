// batch_kernel.hip — persistent batched render: each 256-lane block takes the next tile from
// a global atomic counter until the queue is drained; tiles of all jobs form one index space
// (jobs sorted by scene, so a block rarely switches scene).  With STAGED, a block keeps the
// current scene's tables in LDS and re-stages only when its next tile belongs to another
// scene.  Compiled per MM3E_MAXB (the batch's largest `bounces`).
#include <hip/hip_runtime.h>

#include "batch.h"

#ifndef MM3E_MAXB
#error "MM3E_MAXB must be defined"
#endif
#define MM3E_CAT2(a, b) a##b
#define MM3E_CAT(a, b) MM3E_CAT2(a, b)

using namespace mm3e;

// With q.order, queue position k holds tile q.order[k] (most expensive first, from the previous
// replay's costs) and each tile's cost (tickets drawn meanwhile from q.ticket) goes to q.cost;
// without, tiles are taken in index order.
template <bool STAGED>
__device__ __forceinline__ void batch_render_body(const BatchJob* __restrict__ jobs, uint32_t n_jobs,
                                                  uint32_t total_tiles, uint32_t* __restrict__ counter,
                                                  const TileOrder& q, uint4* tables_lds) {
  __shared__ float smem[256 * 3];
  __shared__ uint32_t s_tile, s_start;
  const DevScene* staged = nullptr;
  for (;;) {
    if (threadIdx.x == 0) {
      const uint32_t k = atomicAdd(counter, 1u);
      s_tile = (k < total_tiles && q.order) ? q.order[k] : k;
      if (q.order) s_start = atomicAdd(q.ticket, 1u);
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    __syncthreads();  // every lane has read s_tile before lane 0 may overwrite it
    if (tile >= total_tiles) break;  // uniform: the whole block leaves together
    const uint32_t j = find_job(jobs, n_jobs, tile, [](const BatchJob& b) { return b.tile_base; });
    const BatchJob& J = jobs[j];
    const DevScene* sc = J.sc;
    const uint32_t local = tile - J.tile_base;
    const uint32_t tx = local % J.tiles_x, ty = local / J.tiles_x;
    if constexpr (STAGED) {
      if (sc != staged) {
        // all lanes are past the previous tile's last LDS read (the loop-top barriers)
        const uint32_t n16 = J.stage_bytes / 16;
        const uint4* src = reinterpret_cast<const uint4*>(sc->obj);
        for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) tables_lds[i] = src[i];
        __syncthreads();
        staged = sc;
      }
      const char* b = reinterpret_cast<const char*>(tables_lds);
      const Tables tb{reinterpret_cast<const DObj*>(b), reinterpret_cast<const DSeg*>(b + sc->st_seg),
                      reinterpret_cast<const DNode*>(b + sc->st_node),
                      reinterpret_cast<const uint32_t*>(b + sc->st_unb)};
      render_tile<TableMap, MM3E_MAXB, true>(sc, tb, J.fp, J.hdr, J.tw, J.th, tx, ty, smem);
    } else {
      render_tile<TableMap, MM3E_MAXB, true>(sc, global_tables(sc), J.fp, J.hdr, J.tw, J.th, tx, ty, smem);
    }
    __syncthreads();  // smem is reused by the next tile
    if (q.order && threadIdx.x == 0) q.cost[tile] = atomicAdd(q.ticket, 1u) - s_start;
  }
}

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    MM3E_CAT(mm3e_batch_render_lds_b, MM3E_MAXB)(const BatchJob* __restrict__ jobs, uint32_t n_jobs,
                                                 uint32_t total_tiles, uint32_t* __restrict__ counter, TileOrder q) {
  extern __shared__ uint4 tables_lds[];
  batch_render_body<true>(jobs, n_jobs, total_tiles, counter, q, tables_lds);
}

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    MM3E_CAT(mm3e_batch_render_table_b, MM3E_MAXB)(const BatchJob* __restrict__ jobs, uint32_t n_jobs,
                                                   uint32_t total_tiles, uint32_t* __restrict__ counter, TileOrder q) {
  batch_render_body<false>(jobs, n_jobs, total_tiles, counter, q, nullptr);
}

namespace mm3e {
hipError_t MM3E_CAT(launch_batch_render_b, MM3E_MAXB)(uint32_t blocks, hipStream_t st, const BatchJob* jobs,
                                                      uint32_t n_jobs, uint32_t total_tiles, uint32_t* counter,
                                                      uint32_t lds_bytes, const TileOrder& q) {
  if (lds_bytes)
    hipLaunchKernelGGL(MM3E_CAT(mm3e_batch_render_lds_b, MM3E_MAXB), dim3(blocks), dim3(256), lds_bytes, st, jobs,
                       n_jobs, total_tiles, counter, q);
  else
    hipLaunchKernelGGL(MM3E_CAT(mm3e_batch_render_table_b, MM3E_MAXB), dim3(blocks), dim3(256), 0, st, jobs, n_jobs,
                       total_tiles, counter, q);
  return hipGetLastError();
}
}  // namespace mm3e
