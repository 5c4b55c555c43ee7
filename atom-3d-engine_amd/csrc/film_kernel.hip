// film_kernel.hip — render_film's sample pass (film.rs:63-198) for one reflection depth,
// compiled once per MM3E_MAXB (0..4) like the render kernels: beauty, premultiplied foreground
// and coverage alpha per pixel through a thin lens.  Two families as the render kernels:
//   mm3e_film_bN     scenes without a segment grid or CSG objects (nested march/shade);
//   mm3e_film_x_bN   every other scene (general fold, single-evaluation-site state machine).
// The scene tables are read from HBM (a film render is not the per-frame hot loop).
#include <hip/hip_runtime.h>

#include "dev_render.h"
#include "launch.h"

#ifndef MM3E_MAXB
#error "MM3E_MAXB must be defined"
#endif

#define MM3E_CAT2(a, b) a##b
#define MM3E_CAT(a, b) MM3E_CAT2(a, b)
#define KNAME MM3E_CAT(mm3e_film_b, MM3E_MAXB)
#define KNAME_X MM3E_CAT(mm3e_film_x_b, MM3E_MAXB)
#define LNAME MM3E_CAT(launch_film_b, MM3E_MAXB)

using namespace mm3e;

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME(const DevScene* __restrict__ sc, FrameParams fp, FilmParams lp, float* __restrict__ beauty,
          float* __restrict__ fg, float* __restrict__ alpha, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 4];
  film_tile<SimpleMap, MM3E_MAXB, false>(sc, global_tables(sc), fp, lp, beauty, fg, alpha, tw, th, blockIdx.x,
                                         blockIdx.y, smem);
}

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME_X(const DevScene* __restrict__ sc, FrameParams fp, FilmParams lp, float* __restrict__ beauty,
            float* __restrict__ fg, float* __restrict__ alpha, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 4];
  film_tile<TableMap, MM3E_MAXB, true>(sc, global_tables(sc), fp, lp, beauty, fg, alpha, tw, th, blockIdx.x,
                                       blockIdx.y, smem);
}

namespace mm3e {
hipError_t LNAME(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, bool general, const FrameParams& fp,
                 const FilmParams& lp, float* beauty, float* fg, float* alpha, uint32_t tw, uint32_t th) {
  if (general) hipLaunchKernelGGL(KNAME_X, grid, block, 0, st, sc, fp, lp, beauty, fg, alpha, tw, th);
  else hipLaunchKernelGGL(KNAME, grid, block, 0, st, sc, fp, lp, beauty, fg, alpha, tw, th);
  return hipGetLastError();
}
}  // namespace mm3e
