// mesh_bake.hip — mesh -> signed-distance volume bake on the GPU (bake_sdf,
// mm3e-orchestrator/src/mesh.rs:265-336).
//
// Two kernels per bake, both bit-exact with the reference's CPU bake:
//   mm3e_mesh_row_parity — one block per 4 grid rows along +x.  Each thread takes one triangle
//     per pass, computes its +x line crossing for the 4 rows in f64 (x_crossing,
//     mm3e-kit/src/triangle.rs:4-34), and appends hits to a per-row LDS list; the threads owning
//     the row's samples then count the hits left of their sample.  inside = odd count
//     (`partition_point(x < p.x) % 2 == 1` over the sorted crossings, mesh.rs:315-325).
//   mm3e_mesh_nearest — one thread per grid point, a wave per 4x4x4 brick of points, walks the
//     nearest-triangle BVH (TriBvh::nearest_sq, mesh.rs:225-260) with a per-lane LDS stack in
//     the reference's visiting order (near child first, a node skipped when its box is not
//     closer than the best so far), and writes +-sqrt(best).
// The device tree keeps both child boxes in the parent (64-byte interior nodes) and the
// triangles in leaf order, so a popped entry already carries the box distance the reference
// computes on pop: the same f32 operations on the same values, one fewer node load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mm3e_hip.h"
#include "mesh.h"
#include "runtime.h"

namespace mm3e {

struct BakeGrid {
  float lo[3];
  float cell[3];
  uint32_t nx, ny, nz;
};

// Interior node: the boxes and references of both children.  A reference c >= 0 is an interior
// node, c < 0 the leaf holding triangle slot -(c + 1).
struct alignas(16) DevBvhNode {
  float lo0[3];
  int32_t c0;
  float hi0[3];
  int32_t c1;
  float lo1[3];
  uint32_t pad0;
  float hi1[3];
  uint32_t pad1;
};
static_assert(sizeof(DevBvhNode) == 64, "DevBvhNode layout");

constexpr uint32_t kParityRows = 4;     // grid rows per block of the parity kernel
constexpr uint32_t kParityItems = 8;    // samples per thread: kParityRows * 512 / 256
constexpr uint32_t kMaxStack = 32;      // LDS stack entries per lane (tree depth + 1): 64 KB

// ---- f64 line crossing (mm3e-kit/src/triangle.rs:4-34) ---------------------------------------
__device__ __forceinline__ double edge_d(float ay, float az, float by, float bz, double y, double z) {
  return ((double)by - (double)ay) * (z - (double)az) - ((double)bz - (double)az) * (y - (double)ay);
}
__device__ __forceinline__ bool owns_boundary(float ay, float az, float by, float bz) {
  return bz > az || (bz == az && by < ay);
}

// ---- f32 nearest point on a triangle (mesh.rs:111-146) ---------------------------------------
struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 axpy3(F3 a, F3 d, float s) { return F3{a.x + d.x * s, a.y + d.y * s, a.z + d.z * s}; }

__device__ __forceinline__ float tri_dist_sq(F3 p, F3 a, F3 b, F3 c) {
  const F3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
  const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  F3 q;
  if (d1 <= 0.0f && d2 <= 0.0f) {
    q = a;
  } else {
    const F3 bp = sub3(p, b);
    const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    const float vc = d1 * d4 - d3 * d2;
    const F3 cp = sub3(p, c);
    const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    const float vb = d5 * d2 - d1 * d6;
    const float va = d3 * d6 - d5 * d4;
    if (d3 >= 0.0f && d4 <= d3) {
      q = b;
    } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
      q = axpy3(a, ab, d1 / (d1 - d3));
    } else if (d6 >= 0.0f && d5 <= d6) {
      q = c;
    } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
      q = axpy3(a, ac, d2 / (d2 - d6));
    } else if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
      q = axpy3(b, sub3(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6)));
    } else {
      const float denom = 1.0f / (va + vb + vc);
      const F3 t = axpy3(a, ab, vb * denom);
      q = axpy3(t, ac, vc * denom);
    }
  }
  const F3 d = sub3(p, q);
  return dot3(d, d);
}

// box_dist_sq (mesh.rs:161-163): |p - p.max(lo).min(hi)|^2
__device__ __forceinline__ float box_dist_sq(F3 p, const float* lo, const float* hi) {
  const F3 c{fminf(fmaxf(p.x, lo[0]), hi[0]), fminf(fmaxf(p.y, lo[1]), hi[1]), fminf(fmaxf(p.z, lo[2]), hi[2])};
  const F3 d = sub3(p, c);
  return dot3(d, d);
}

__device__ __forceinline__ void load_tri(const float4* __restrict__ tv, uint32_t t, F3& a, F3& b, F3& c) {
  const float4 u = tv[3 * (size_t)t], v = tv[3 * (size_t)t + 1], w = tv[3 * (size_t)t + 2];
  a = F3{u.x, u.y, u.z};
  b = F3{u.w, v.x, v.y};
  c = F3{v.z, v.w, w.x};
}

}  // namespace mm3e

using namespace mm3e;

// One block per kParityRows consecutive rows (row r = k*ny + j); rows past the grid idle.
extern "C" __global__ void __launch_bounds__(256)
    mm3e_mesh_row_parity(const float4* __restrict__ tri_all, uint32_t n_tris, BakeGrid g,
                         uint8_t* __restrict__ inside) {
  __shared__ float xs[kParityRows][256];
  __shared__ uint32_t cnt[kParityRows];
  const uint32_t tid = threadIdx.x;
  const uint32_t rows = g.ny * g.nz;
  const uint32_t r0 = blockIdx.x * kParityRows;
  float ry[kParityRows], rz[kParityRows];
  bool rv[kParityRows];
#pragma unroll
  for (uint32_t q = 0; q < kParityRows; q++) {
    const uint32_t r = r0 + q;
    rv[q] = r < rows;
    const uint32_t j = rv[q] ? r % g.ny : 0, k = rv[q] ? r / g.ny : 0;
    ry[q] = g.lo[1] + (float)j * g.cell[1];
    rz[q] = g.lo[2] + (float)k * g.cell[2];
  }
  // samples owned by this thread: s = tid + 256*u -> (row q = s / nx, i = s % nx)
  float px[kParityItems];
  uint32_t pq[kParityItems], hits[kParityItems];
#pragma unroll
  for (uint32_t u = 0; u < kParityItems; u++) {
    const uint32_t s = tid + 256u * u;
    pq[u] = s < kParityRows * g.nx ? s / g.nx : kParityRows;
    px[u] = g.lo[0] + (float)(s % g.nx) * g.cell[0];
    hits[u] = 0;
  }
  if (tid < kParityRows) cnt[tid] = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n_tris; base += 256) {
    const uint32_t t = base + tid;
    if (t < n_tris) {
      F3 a, b, c;
      load_tri(tri_all, t, a, b, c);
      double area = edge_d(a.y, a.z, b.y, b.z, (double)c.y, (double)c.z);
      if (area != 0.0 && isfinite(area)) {
        if (area < 0.0) {
          const F3 sw = b;
          b = c;
          c = sw;
          area = -area;
        }
#pragma unroll
        for (uint32_t q = 0; q < kParityRows; q++) {
          if (!rv[q] || !isfinite(ry[q]) || !isfinite(rz[q])) continue;
          const double y = (double)ry[q], z = (double)rz[q];
          const double w0 = edge_d(b.y, b.z, c.y, c.z, y, z);
          const double w1 = edge_d(c.y, c.z, a.y, a.z, y, z);
          const double w2 = edge_d(a.y, a.z, b.y, b.z, y, z);
          if (w0 < 0.0 || (w0 == 0.0 && !owns_boundary(b.y, b.z, c.y, c.z))) continue;
          if (w1 < 0.0 || (w1 == 0.0 && !owns_boundary(c.y, c.z, a.y, a.z))) continue;
          if (w2 < 0.0 || (w2 == 0.0 && !owns_boundary(a.y, a.z, b.y, b.z))) continue;
          const float x = (float)((w0 * (double)a.x + w1 * (double)b.x + w2 * (double)c.x) / area);
          if (!isfinite(x)) continue;
          const uint32_t slot = atomicAdd(&cnt[q], 1u);
          xs[q][slot] = x;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kParityItems; u++) {
      if (pq[u] >= kParityRows) continue;
      const uint32_t n = cnt[pq[u]];
      for (uint32_t e = 0; e < n; e++) hits[u] += xs[pq[u]][e] < px[u] ? 1u : 0u;
    }
    __syncthreads();
    if (tid < kParityRows) cnt[tid] = 0;
    __syncthreads();
  }
#pragma unroll
  for (uint32_t u = 0; u < kParityItems; u++) {
    if (pq[u] >= kParityRows || !rv[pq[u]]) continue;
    const uint32_t s = tid + 256u * u;
    inside[(size_t)(r0 + pq[u]) * g.nx + s % g.nx] = (uint8_t)(hits[u] & 1u);
  }
}

// One thread per grid point; block = 8x8x4 points as four 4x4x4 wave bricks.  Dynamic LDS:
// `stack` entries of (reference, box distance) per lane.
extern "C" __global__ void __launch_bounds__(256)
    mm3e_mesh_nearest(const DevBvhNode* __restrict__ nodes, int32_t root,
                      const float4* __restrict__ tv, const uint8_t* __restrict__ inside, BakeGrid g,
                      float root_lo_x, float root_lo_y, float root_lo_z, float root_hi_x, float root_hi_y,
                      float root_hi_z, uint32_t stack, float* __restrict__ out) {
  extern __shared__ uint32_t lds_stack[];  // [stack][256] refs, then [stack][256] distances
  uint32_t* sref = lds_stack;
  float* sdist = reinterpret_cast<float*>(lds_stack + stack * 256);
  const uint32_t tid = threadIdx.x;
  const uint32_t w = tid >> 6, l = tid & 63;
  const uint32_t i = blockIdx.x * 8 + (w & 1) * 4 + (l & 3);
  const uint32_t j = blockIdx.y * 8 + (w >> 1) * 4 + ((l >> 2) & 3);
  const uint32_t k = blockIdx.z * 4 + (l >> 4);
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const F3 p{g.lo[0] + (float)i * g.cell[0], g.lo[1] + (float)j * g.cell[1], g.lo[2] + (float)k * g.cell[2]};
  const float rlo[3] = {root_lo_x, root_lo_y, root_lo_z}, rhi[3] = {root_hi_x, root_hi_y, root_hi_z};
  float best = __builtin_inff();
  uint32_t sp = 1;
  sref[tid] = (uint32_t)root;
  sdist[tid] = box_dist_sq(p, rlo, rhi);
  while (sp > 0) {
    sp--;
    const int32_t c = (int32_t)sref[sp * 256 + tid];
    const float d = sdist[sp * 256 + tid];
    if (d >= best) continue;
    if (c < 0) {
      F3 a, b, cc;
      load_tri(tv, (uint32_t)(-(c + 1)), a, b, cc);
      best = fminf(best, tri_dist_sq(p, a, b, cc));
    } else {
      const DevBvhNode& n = nodes[c];
      const float ld = box_dist_sq(p, n.lo0, n.hi0);
      const float rd = box_dist_sq(p, n.lo1, n.hi1);
      const bool left_near = ld <= rd;
      sref[sp * 256 + tid] = (uint32_t)(left_near ? n.c1 : n.c0);
      sdist[sp * 256 + tid] = left_near ? rd : ld;
      sref[(sp + 1) * 256 + tid] = (uint32_t)(left_near ? n.c0 : n.c1);
      sdist[(sp + 1) * 256 + tid] = left_near ? ld : rd;
      sp += 2;
    }
  }
  const float u = sqrtf(best);
  const size_t idx = ((size_t)k * g.ny + j) * g.nx + i;
  out[idx] = inside[idx] ? -u : u;
}

// ---- host side ------------------------------------------------------------------------------

struct mm3e_mesh_baker {
  mm3e_renderer* r = nullptr;
  BakeLayout L{};
  BakeGrid g{};
  uint32_t n_tris = 0, stack = 0, n_nodes = 0, depth = 0;
  int32_t root = 0;
  float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
  DevBvhNode* d_nodes = nullptr;
  float4* d_tv = nullptr;    // triangles in leaf order (nearest search)
  float4* d_all = nullptr;   // every triangle in mesh order (crossings)
  uint8_t* d_inside = nullptr;
  float* d_out = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool launched = false;
};

namespace {

// Re-lay the reference-shaped tree (post-order, one box per node) as parent-held child boxes in
// pre-order, triangles numbered in leaf order.
struct Relayout {
  const TriBvhHost& t;
  const HostMesh& m;
  std::vector<DevBvhNode> out;
  std::vector<float4> tv;
  int32_t ref(uint32_t n) {
    const BakeNode& nd = t.nodes[n];
    if (nd.a < 0) {  // leaf
      const uint32_t tri = (uint32_t)(-(nd.a + 1));
      const uint32_t slot = (uint32_t)(tv.size() / 3);
      float v[12] = {0};
      for (int c = 0; c < 3; c++)
        for (int a = 0; a < 3; a++) v[3 * c + a] = m.pos[3 * (size_t)m.tris[3 * (size_t)tri + c] + a];
      tv.push_back(make_float4(v[0], v[1], v[2], v[3]));
      tv.push_back(make_float4(v[4], v[5], v[6], v[7]));
      tv.push_back(make_float4(v[8], 0.0f, 0.0f, 0.0f));
      return -(int32_t)slot - 1;
    }
    const uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    const BakeNode& L = t.nodes[(uint32_t)nd.a];
    const BakeNode& R = t.nodes[nd.b];
    const int32_t cl = ref((uint32_t)nd.a);
    const int32_t cr = ref(nd.b);
    DevBvhNode& d = out[me];
    std::memset(&d, 0, sizeof d);
    for (int a = 0; a < 3; a++) {
      d.lo0[a] = L.lo[a];
      d.hi0[a] = L.hi[a];
      d.lo1[a] = R.lo[a];
      d.hi1[a] = R.hi[a];
    }
    d.c0 = cl;
    d.c1 = cr;
    return (int32_t)me;
  }
};

void free_baker(mm3e_mesh_baker* b) {
  if (!b) return;
  if (b->r) hipSetDevice(b->r->device);
  if (b->r && b->launched) hipStreamSynchronize(b->r->stream);
  hipFree(b->d_nodes);
  hipFree(b->d_tv);
  hipFree(b->d_all);
  hipFree(b->d_inside);
  hipFree(b->d_out);
  for (hipEvent_t e : b->ev)
    if (e) hipEventDestroy(e);
  delete b;
}

}  // namespace

extern "C" {

int mm3e_mesh_baker_create(mm3e_renderer* r, const float* positions, size_t n_positions, const uint32_t* triangles,
                           size_t n_triangles, uint32_t resolution, float padding, mm3e_mesh_baker** out, char* err,
                           size_t errlen) {
  if (!r || !out || (!positions && n_positions) || (!triangles && n_triangles)) return MM3E_ERR_INVALID;
  *out = nullptr;
  HostMesh m;
  m.pos.assign(positions, positions + n_positions * 3);
  m.tris.assign(triangles, triangles + n_triangles * 3);
  std::string msg;
  BakeLayout L;
  if (!bake_layout(m, resolution, padding, L, msg)) {
    set_err(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  TriBvhHost bvh;
  if (!build_tri_bvh(m, bvh)) {
    set_err(err, errlen, "mesh has no non-degenerate triangles");
    return MM3E_ERR_INVALID;
  }
  if (!check_layout(L, msg)) {
    set_err(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  if (bvh.depth + 1 > kMaxStack || m.n_tris() > 0x7fffffffu) {
    set_err(err, errlen, "mesh too large for the bake (BVH depth " + std::to_string(bvh.depth) + ")");
    return MM3E_ERR_INVALID;
  }
  Relayout rl{bvh, m, {}, {}};
  const int32_t root = rl.ref(bvh.root);
  std::vector<float4> all;
  all.reserve(m.n_tris() * 3);
  for (size_t t = 0; t < m.n_tris(); t++) {
    float v[9];
    for (int c = 0; c < 3; c++)
      for (int a = 0; a < 3; a++) v[3 * c + a] = m.pos[3 * (size_t)m.tris[3 * t + c] + a];
    all.push_back(make_float4(v[0], v[1], v[2], v[3]));
    all.push_back(make_float4(v[4], v[5], v[6], v[7]));
    all.push_back(make_float4(v[8], 0.0f, 0.0f, 0.0f));
  }
  auto* b = new mm3e_mesh_baker();
  b->r = r;
  b->L = L;
  for (int a = 0; a < 3; a++) {
    b->g.lo[a] = L.lo[a];
    b->g.cell[a] = L.cell[a];
    b->root_lo[a] = bvh.nodes[bvh.root].lo[a];
    b->root_hi[a] = bvh.nodes[bvh.root].hi[a];
  }
  b->g.nx = L.dims[0];
  b->g.ny = L.dims[1];
  b->g.nz = L.dims[2];
  b->n_tris = (uint32_t)m.n_tris();
  b->stack = bvh.depth + 1;
  b->depth = bvh.depth;
  b->n_nodes = (uint32_t)rl.out.size();
  b->root = root;
  const size_t n = L.samples();
  hipSetDevice(r->device);
  bool ok = true;
  const size_t node_bytes = std::max<size_t>(rl.out.size(), 1) * sizeof(DevBvhNode);
  ok = ok && hip_ok(hipMalloc((void**)&b->d_nodes, node_bytes), "hipMalloc(bvh)", &msg);
  ok = ok && hip_ok(hipMalloc((void**)&b->d_tv, rl.tv.size() * sizeof(float4)), "hipMalloc(triangles)", &msg);
  ok = ok && hip_ok(hipMalloc((void**)&b->d_all, all.size() * sizeof(float4)), "hipMalloc(triangles)", &msg);
  ok = ok && hip_ok(hipMalloc((void**)&b->d_inside, n), "hipMalloc(parity)", &msg);
  ok = ok && hip_ok(hipMalloc((void**)&b->d_out, n * sizeof(float)), "hipMalloc(volume)", &msg);
  if (ok && !rl.out.empty())
    ok = hip_ok(hipMemcpy(b->d_nodes, rl.out.data(), rl.out.size() * sizeof(DevBvhNode), hipMemcpyHostToDevice),
                "upload bvh", &msg);
  ok = ok && hip_ok(hipMemcpy(b->d_tv, rl.tv.data(), rl.tv.size() * sizeof(float4), hipMemcpyHostToDevice),
                    "upload triangles", &msg);
  ok = ok && hip_ok(hipMemcpy(b->d_all, all.data(), all.size() * sizeof(float4), hipMemcpyHostToDevice),
                    "upload triangles", &msg);
  for (int e = 0; ok && e < 3; e++) ok = hip_ok(hipEventCreate(&b->ev[e]), "hipEventCreate", &msg);
  if (!ok) {
    set_err(err, errlen, msg);
    free_baker(b);
    return MM3E_ERR_DEVICE;
  }
  *out = b;
  return MM3E_OK;
}

int mm3e_mesh_baker_layout(const mm3e_mesh_baker* b, mm3e_sdf_layout* layout) {
  if (!b || !layout) return MM3E_ERR_INVALID;
  for (int a = 0; a < 3; a++) {
    layout->dims[a] = b->L.dims[a];
    layout->min[a] = b->L.lo[a];
    layout->cell[a] = b->L.cell[a];
  }
  return MM3E_OK;
}

int mm3e_mesh_baker_info(const mm3e_mesh_baker* b, uint32_t* n_triangles, uint32_t* n_nodes, uint32_t* depth) {
  if (!b) return MM3E_ERR_INVALID;
  if (n_triangles) *n_triangles = b->n_tris;
  if (n_nodes) *n_nodes = b->n_nodes;
  if (depth) *depth = b->depth;
  return MM3E_OK;
}

int mm3e_mesh_baker_run(mm3e_mesh_baker* b) {
  if (!b) return MM3E_ERR_INVALID;
  hipSetDevice(b->r->device);
  hipStream_t st = b->r->stream;
  const uint32_t rows = b->g.ny * b->g.nz;
  const dim3 gp((rows + kParityRows - 1) / kParityRows);
  const dim3 gn((b->g.nx + 7) / 8, (b->g.ny + 7) / 8, (b->g.nz + 3) / 4);
  const size_t lds = (size_t)b->stack * 256 * 8;
  // grid shapes against what the kernels index: nx <= 512 (kParityRows * 512 samples per block
  // spread over kParityItems per thread), stack entries <= kMaxStack
  if (b->g.nx > 512 || b->g.ny > 512 || b->g.nz > 512 || b->stack > kMaxStack) return MM3E_ERR_INVALID;
  hipEventRecord(b->ev[0], st);
  hipLaunchKernelGGL(mm3e_mesh_row_parity, gp, dim3(256), 0, st, b->d_all, b->n_tris, b->g, b->d_inside);
  hipEventRecord(b->ev[1], st);
  hipLaunchKernelGGL(mm3e_mesh_nearest, gn, dim3(256), lds, st, b->d_nodes, b->root, b->d_tv, b->d_inside, b->g,
                     b->root_lo[0], b->root_lo[1], b->root_lo[2], b->root_hi[0], b->root_hi[1], b->root_hi[2],
                     b->stack, b->d_out);
  hipEventRecord(b->ev[2], st);
  b->launched = true;
  return hipGetLastError() == hipSuccess ? MM3E_OK : MM3E_ERR_DEVICE;
}

int mm3e_mesh_baker_wait(mm3e_mesh_baker* b, float* parity_ms, float* nearest_ms) {
  if (!b) return MM3E_ERR_INVALID;
  if (!b->launched) return MM3E_ERR_INVALID;
  hipSetDevice(b->r->device);
  if (hipEventSynchronize(b->ev[2]) != hipSuccess) return MM3E_ERR_DEVICE;
  float a = 0.0f, c = 0.0f;
  if (hipEventElapsedTime(&a, b->ev[0], b->ev[1]) != hipSuccess) return MM3E_ERR_DEVICE;
  if (hipEventElapsedTime(&c, b->ev[1], b->ev[2]) != hipSuccess) return MM3E_ERR_DEVICE;
  if (parity_ms) *parity_ms = a;
  if (nearest_ms) *nearest_ms = c;
  return MM3E_OK;
}

int mm3e_mesh_baker_read(mm3e_mesh_baker* b, float* data, size_t data_cap, char* err, size_t errlen) {
  if (!b || !data) return MM3E_ERR_INVALID;
  if (!b->launched) {
    set_err(err, errlen, "mm3e_mesh_baker_read: nothing baked yet");
    return MM3E_ERR_INVALID;
  }
  const size_t n = b->L.samples();
  if (data_cap < n) {
    set_err(err, errlen, "mm3e_mesh_baker_read: data buffer too small");
    return MM3E_ERR_INVALID;
  }
  hipSetDevice(b->r->device);
  std::string msg;
  if (!hip_ok(hipStreamSynchronize(b->r->stream), "bake", &msg) ||
      !hip_ok(hipMemcpy(data, b->d_out, n * sizeof(float), hipMemcpyDeviceToHost), "download volume", &msg)) {
    set_err(err, errlen, msg);
    return MM3E_ERR_DEVICE;
  }
  for (size_t i = 0; i < n; i++)  // SdfVolume::new (volume.rs:27-40)
    if (!std::isfinite(data[i])) {
      set_err(err, errlen, "SdfVolume sample " + std::to_string(i) + " is non-finite");
      return MM3E_ERR_INVALID;
    }
  return MM3E_OK;
}

void mm3e_mesh_baker_destroy(mm3e_mesh_baker* b) { free_baker(b); }

int mm3e_mesh_bake_sdf(mm3e_renderer* r, const float* positions, size_t n_positions, const uint32_t* triangles,
                       size_t n_triangles, uint32_t resolution, float padding, float* data, size_t data_cap,
                       mm3e_sdf_layout* layout, char* err, size_t errlen) {
  mm3e_mesh_baker* b = nullptr;
  int rc = mm3e_mesh_baker_create(r, positions, n_positions, triangles, n_triangles, resolution, padding, &b, err,
                                  errlen);
  if (rc != MM3E_OK) return rc;
  if (layout) mm3e_mesh_baker_layout(b, layout);
  rc = mm3e_mesh_baker_run(b);
  if (rc == MM3E_OK) rc = mm3e_mesh_baker_read(b, data, data_cap, err, errlen);
  mm3e_mesh_baker_destroy(b);
  return rc;
}

}  // extern "C"
