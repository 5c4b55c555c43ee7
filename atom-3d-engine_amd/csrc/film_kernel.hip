// film_kernel.hip — render_film's sample pass (film.rs:63-198) for one reflection depth,
// compiled once per MM3E_MAXB (0..4, 16) like the render kernels: beauty, premultiplied foreground
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
#define GNAME MM3E_CAT(mm3e_gbuffer_b, MM3E_MAXB)
#define GNAME_X MM3E_CAT(mm3e_gbuffer_x_b, MM3E_MAXB)
#define GLNAME MM3E_CAT(launch_gbuffer_b, MM3E_MAXB)

using namespace mm3e;

#if MM3E_MAXB <= 4
extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME(const DevScene* __restrict__ sc, FrameParams fp, FilmParams lp, float* __restrict__ beauty,
          float* __restrict__ fg, float* __restrict__ alpha, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 4];
  film_tile<SimpleMap, MM3E_MAXB, false>(sc, global_tables(sc), fp, lp, beauty, fg, alpha, tw, th, blockIdx.x,
                                         blockIdx.y, smem);
}

#endif

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    KNAME_X(const DevScene* __restrict__ sc, FrameParams fp, FilmParams lp, float* __restrict__ beauty,
            float* __restrict__ fg, float* __restrict__ alpha, uint32_t tw, uint32_t th) {
  __shared__ float smem[256 * 4];
  film_tile<TableMap, MM3E_MAXB, true>(sc, global_tables(sc), fp, lp, beauty, fg, alpha, tw, th, blockIdx.x,
                                       blockIdx.y, smem);
}

// G-buffer pixels (16x16 tiles) or the pixels of a list (hybrid reprojection's holes).
#if MM3E_MAXB <= 4
extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    GNAME(const DevScene* __restrict__ sc, FrameParams fp, float exposure, const GMover* __restrict__ movers,
          uint32_t n_movers, const uint32_t* __restrict__ list, const uint32_t* __restrict__ n_list,
          uchar4* __restrict__ color,
          float* __restrict__ depth, int32_t* __restrict__ obj) {
  gbuffer_tile<SimpleMap, MM3E_MAXB, false>(sc, fp, exposure, movers, n_movers, list, n_list, color, depth, obj);
}

#endif

extern "C" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY
    GNAME_X(const DevScene* __restrict__ sc, FrameParams fp, float exposure, const GMover* __restrict__ movers,
            uint32_t n_movers, const uint32_t* __restrict__ list, const uint32_t* __restrict__ n_list,
          uchar4* __restrict__ color,
            float* __restrict__ depth, int32_t* __restrict__ obj) {
  gbuffer_tile<TableMap, MM3E_MAXB, true>(sc, fp, exposure, movers, n_movers, list, n_list, color, depth, obj);
}

namespace mm3e {
hipError_t GLNAME(hipStream_t st, const DevScene* sc, bool general, const FrameParams& fp, float exposure,
                  const GMover* movers, uint32_t n_movers, const uint32_t* list, const uint32_t* n_list,
                  uint32_t max_list, uchar4* color, float* depth, int32_t* obj) {
  const dim3 grid = list ? dim3((max_list + 255) / 256) : dim3((fp.width + 15) / 16, (fp.height + 15) / 16);
  if (list && max_list == 0) return hipSuccess;
#if MM3E_MAXB <= 4
  if (!general) {
    hipLaunchKernelGGL(GNAME, grid, dim3(256), 0, st, sc, fp, exposure, movers, n_movers, list, n_list, color, depth,
                       obj);
    return hipGetLastError();
  }
#endif
  hipLaunchKernelGGL(GNAME_X, grid, dim3(256), 0, st, sc, fp, exposure, movers, n_movers, list, n_list, color, depth,
                     obj);
  return hipGetLastError();
}

hipError_t LNAME(dim3 grid, dim3 block, hipStream_t st, const DevScene* sc, bool general, const FrameParams& fp,
                 const FilmParams& lp, float* beauty, float* fg, float* alpha, uint32_t tw, uint32_t th) {
#if MM3E_MAXB <= 4
  if (!general) {
    hipLaunchKernelGGL(KNAME, grid, block, 0, st, sc, fp, lp, beauty, fg, alpha, tw, th);
    return hipGetLastError();
  }
#endif
  hipLaunchKernelGGL(KNAME_X, grid, block, 0, st, sc, fp, lp, beauty, fg, alpha, tw, th);
  return hipGetLastError();
}
}  // namespace mm3e
