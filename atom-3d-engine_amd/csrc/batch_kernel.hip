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

// The table-walk tile render of a batch job (general fold, state machine).
struct TableTile {
  MM3E_DEV void operator()(const BatchJob& J, const Tables& tb, uint32_t tx, uint32_t ty, float* smem) const {
    render_tile<TableMap, MM3E_MAXB, true>(J.sc, tb, J.fp, J.hdr, J.tw, J.th, tx, ty, smem);
  }
};

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    MM3E_CAT(mm3e_batch_render_lds_b, MM3E_MAXB)(const BatchJob* __restrict__ jobs, uint32_t n_jobs,
                                                 uint32_t total_tiles, uint32_t* __restrict__ counter, TileOrder q) {
  extern __shared__ uint4 tables_lds[];
  batch_loop<true>(jobs, n_jobs, total_tiles, counter, q, tables_lds, TableTile{});
}

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    MM3E_CAT(mm3e_batch_render_table_b, MM3E_MAXB)(const BatchJob* __restrict__ jobs, uint32_t n_jobs,
                                                   uint32_t total_tiles, uint32_t* __restrict__ counter, TileOrder q) {
  batch_loop<false>(jobs, n_jobs, total_tiles, counter, q, nullptr, TableTile{});
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
