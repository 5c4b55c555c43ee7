// batch.h — device records of a batched render: many (scene, camera, size) jobs served by one
// persistent launch that pulls tiles from a global atomic queue.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_render.h"

namespace mm3e {

struct alignas(16) BatchJob {
  const DevScene* sc;      // compiled scene (tables in HBM)
  float* hdr;              // w*h*3 scene-linear radiance
  FrameParams fp;          // camera + this job's image size
  uint32_t tile_base;      // first global tile index of this job
  uint32_t tiles_x;        // tiles per row
  uint32_t tw, th;         // tile shape (tw*th*aa^2 <= 256 lanes)
  uint32_t stage_bytes;    // scene table bytes (staged in LDS when the batch stages)
  uint32_t pad[3];
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

}  // namespace mm3e
