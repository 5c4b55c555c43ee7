// dev_batch.h — device side of the batched render: the job records and the persistent loop in
// which each 256-lane block takes the next tile from a global atomic counter until the queue is
// drained.  Tiles of all jobs form one index space (jobs sorted by scene, so a block rarely
// switches scene); with STAGED a block keeps the current scene's tables in LDS and re-stages
// only when its next tile belongs to another scene.  The per-tile render is a functor, so the
// library's kernels (table walk) and generated batch modules (each scene's own map, switched per
// tile on a uniform job field) share the loop.
#pragma once
#include "dev_render.h"

namespace mm3e {

enum : uint32_t { kGenTable = 0xFFFFFFFFu };

struct alignas(16) BatchJob {
  const DevScene* sc;      // compiled scene (tables in HBM)
  float* hdr;              // w*h*3 scene-linear radiance
  FrameParams fp;          // camera + this job's image size
  uint32_t tile_base;      // first global tile index of this job
  uint32_t tiles_x;        // tiles per row
  uint32_t tw, th;         // tile shape (tw*th*aa^2 <= 256 lanes)
  uint32_t stage_bytes;    // scene table bytes (staged in LDS when the batch stages)
  uint32_t gen;            // generated batch modules: this job's scene map (kGenTable: table walk)
  uint32_t pad[2];
};

struct alignas(16) BatchPost {
  PostParams pp;
  const float* kern;       // bloom kernel weights (2r+1)
  const float* hdr;
  float* tmp;
  uchar4* rgba;
  uint32_t row_base;       // first global row index of this job
  uint32_t pad[3];
};

// Largest i with base(i) <= key over a sorted table (uniform per block).
template <class T, class F> MM3E_DEV uint32_t find_job(const T* tab, uint32_t n, uint32_t key, F base) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base(tab[mid]) <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// With q.order, queue position k holds tile q.order[k] (most expensive first, from the previous
// replay's costs) and each tile's cost (tickets drawn meanwhile from q.ticket) goes to q.cost;
// without, tiles are taken in index order.
// `render(J, tb, tx, ty, smem)` renders tile (tx, ty) of job J with the scene tables tb.
template <bool STAGED, class Render>
MM3E_DEV void batch_loop(const BatchJob* __restrict__ jobs, uint32_t n_jobs, uint32_t total_tiles,
                         uint32_t* __restrict__ counter, const TileOrder& q, uint4* tables_lds, Render render) {
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
    // the same value in every lane: made provably uniform so the job record, the scene pointer
    // and the scene's wave-uniform table records are read with scalar loads
    const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
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
      render(J, tb, tx, ty, smem);
    } else {
      render(J, global_tables(sc), tx, ty, smem);
    }
    __syncthreads();  // smem is reused by the next tile
    if (q.order && threadIdx.x == 0) q.cost[tile] = atomicAdd(q.ticket, 1u) - s_start;
  }
}

}  // namespace mm3e
