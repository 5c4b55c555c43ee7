// scene_compile.cpp — IR -> validated host scene -> device table blob.
//
// Host-side twin of the reference's scene compile step (mm3e-gpu/src/lib.rs:138-285 and
// wgsl.rs:41-102): instead of emitting shader text with the scene baked in, it lays the scene
// out as tables in one HBM allocation (objects with precomputed inverse transforms and world
// bounding spheres, the accel.rs world plan with its bounding-sphere trees, materials,
// lights, volumes, GI probes).  Bounds and the tree are computed exactly as the reference
// does (lib.rs:264-292, 470-496; accel.rs:26-185) so the device walk visits the same nodes.
#include "scene_compile.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace mm3e {

#ifndef MM3E_LINEAR_RUN_MAX
#define MM3E_LINEAR_RUN_MAX 16
#endif
// Longest run folded linearly; longer runs walk their tree (O(log n) per evaluation).
static const size_t kLinearRunMax = MM3E_LINEAR_RUN_MAX;
// Largest AA factor: aa^2 = 65,536 samples per pixel, rendered 256 at a time by one block.
static const uint32_t kMaxAA = 256;
// Largest bloom radius: the post pass keeps its 2r+1 Gaussian weights in HBM (8 MiB at most).
static const uint32_t kMaxBloomRadius = 1u << 20;
// Segment candidate grid: built for scenes with at least this many segments; about
// kGridTargetCells cells; the level-l mask keeps the segments within slack + kGridReach[l] of a
// cell.  The grid spans the bounded segments' spheres plus slack + kGridPadReach; a point
// outside it is farther than that from every bounded segment, so the outside cell holds every
// segment only at the levels that reach farther.
static const size_t kGridMinSegments = 8;
static const double kGridTargetCells = 32768.0;
static const double kGridMaxDim = 256.0;
static const double kGridReach[kGridLevels] = {0.25, 0.75, 1.5, 2.5, 4.0, 6.0, 10.0, 16.0};
static const double kGridPadReach = 2.0;

namespace {

struct F3 {
  float x, y, z;
};
inline F3 f3(const float* a) { return F3{a[0], a[1], a[2]}; }
inline F3 add(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 scl(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len(F3 a) { return std::sqrt(dot(a, a)); }
inline F3 cross(F3 a, F3 o) { return F3{a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x}; }
inline F3 normalize(F3 a) {
  float l = len(a);
  if (l > 1e-12f) return scl(a, 1.0f / l);
  return F3{0.0f, 0.0f, 0.0f};
}
inline F3 vmin(F3 a, F3 b) { return F3{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline F3 vmax(F3 a, F3 b) { return F3{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline F3 mul_vec(const float* rot, F3 v) {  // Mat3::mul_vec (vec.rs:125-127)
  return add(add(scl(f3(rot), v.x), scl(f3(rot + 3), v.y)), scl(f3(rot + 6), v.z));
}
inline bool nonzero3(const float* a) { return !(a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f); }

Sphere enclose(F3 c1, float r1, F3 c2, float r2) {  // accel.rs:26-36
  float d = len(sub(c2, c1));
  if (d + r2 <= r1) return Sphere{true, {c1.x, c1.y, c1.z}, r1};
  if (d + r1 <= r2) return Sphere{true, {c2.x, c2.y, c2.z}, r2};
  float radius = (d + r1 + r2) * 0.5f;
  F3 c = add(c1, scl(sub(c2, c1), (radius - r1) / std::fmax(d, 1e-12f)));
  return Sphere{true, {c.x, c.y, c.z}, radius};
}

}  // namespace

Sphere csg_bound(const HostScene& sc, uint32_t ni) {  // csg.rs:120-150
  const mm3e_ir_csg_node& n = sc.csg_nodes[ni];
  const float* p = n.p;
  auto s0 = [](float r) { return Sphere{true, {0.0f, 0.0f, 0.0f}, r}; };
  const Sphere none{false, {0, 0, 0}, 0};
  switch (n.op) {
    case MM3E_CSG_PRIM:
      switch (n.prim) {
        case MM3E_PRIM_SPHERE: return s0(p[0]);
        case MM3E_PRIM_BOX:
        case MM3E_PRIM_ROUND_BOX: return s0(len(f3(p)));
        case MM3E_PRIM_TORUS: return s0(p[0] + p[1]);
        case MM3E_PRIM_CYLINDER: return s0(std::sqrt(p[0] * p[0] + p[1] * p[1]));
        case MM3E_PRIM_CAPSULE: {
          F3 a = f3(p), b = f3(p + 3);
          F3 c = scl(add(a, b), 0.5f);
          return Sphere{true, {c.x, c.y, c.z}, len(sub(a, b)) * 0.5f + p[6]};
        }
        case MM3E_PRIM_CONE: return Sphere{true, {0.0f, p[2] * 0.5f, 0.0f}, p[2] * 0.5f + std::fmax(p[0], p[1])};
        case MM3E_PRIM_HEX_PRISM: return s0(std::sqrt(4.0f * p[0] * p[0] / 3.0f + p[1] * p[1]));
        default: return none;  // ellipsoid, octahedron, plane: no proven underestimate
      }
    case MM3E_CSG_TRANSFORM: {
      Sphere c = csg_bound(sc, n.a);
      if (!c.some) return c;
      F3 w = add(f3(n.pos), mul_vec(n.rot, scl(f3(c.c), n.scale)));  // vec.rs:190-192 to_world
      return Sphere{true, {w.x, w.y, w.z}, c.r * n.scale};
    }
    case MM3E_CSG_UNION:
    case MM3E_CSG_SMOOTH_UNION: {
      Sphere a = csg_bound(sc, n.a);
      if (!a.some) return none;
      Sphere b = csg_bound(sc, n.b);
      if (!b.some) return none;
      Sphere e = enclose(f3(a.c), a.r, f3(b.c), b.r);
      if (n.op == MM3E_CSG_SMOOTH_UNION) e.r += p[0] * 0.25f;
      return e;
    }
    case MM3E_CSG_INTERSECT: {
      Sphere a = csg_bound(sc, n.a), b = csg_bound(sc, n.b);
      if (a.some && b.some) return a.r <= b.r ? a : b;
      return a.some ? a : b;
    }
    case MM3E_CSG_SUBTRACT: return csg_bound(sc, n.a);
    case MM3E_CSG_OFFSET: {
      Sphere c = csg_bound(sc, n.a);
      if (c.some) c.r = std::fmax(c.r + p[0], 0.0f);
      return c;
    }
    case MM3E_CSG_SHELL: {
      Sphere c = csg_bound(sc, n.a);
      if (c.some) c.r = c.r + p[0];
      return c;
    }
  }
  return none;
}

namespace {
// Registers a subtree needs in postfix order when the deeper child goes first (Sethi-Ullman).
uint32_t csg_need(const HostScene& sc, uint32_t ni) {
  const mm3e_ir_csg_node& n = sc.csg_nodes[ni];
  if (n.op == MM3E_CSG_PRIM) return 1;
  if (n.op == MM3E_CSG_TRANSFORM || n.op == MM3E_CSG_OFFSET || n.op == MM3E_CSG_SHELL) return csg_need(sc, n.a);
  const uint32_t a = csg_need(sc, n.a), b = csg_need(sc, n.b);
  return a == b ? a + 1 : std::max(a, b);
}

DCsgOp csg_xform_op(uint32_t op, const mm3e_ir_csg_node& n) {
  DCsgOp o{};
  o.op = op;
  std::memcpy(o.c0, n.rot, 12);
  std::memcpy(o.c1, n.rot + 3, 12);
  std::memcpy(o.c2, n.rot + 6, 12);
  std::memcpy(o.pos, n.pos, 12);
  o.scale = n.scale;
  o.inv = std::fabs(n.scale) > 1e-12f ? 1.0f / n.scale : 0.0f;  // vec.rs:216-221
  return o;
}

void csg_emit(const HostScene& sc, uint32_t ni, std::vector<uint32_t>& frames, std::vector<DCsgOp>& out) {
  const mm3e_ir_csg_node& n = sc.csg_nodes[ni];
  switch (n.op) {
    case MM3E_CSG_PRIM: {
      DCsgOp o{};
      o.op = CSG_LEAF;
      o.prim = n.prim;
      std::memcpy(o.p, n.p, 32);
      out.push_back(o);
      return;
    }
    case MM3E_CSG_TRANSFORM: {
      out.push_back(csg_xform_op(CSG_ENTER, n));
      frames.push_back(ni);
      csg_emit(sc, n.a, frames, out);
      frames.pop_back();
      out.push_back(csg_xform_op(CSG_LEAVE, n));
      for (uint32_t f : frames) out.push_back(csg_xform_op(CSG_ENTER, sc.csg_nodes[f]));  // back to the parent frame
      return;
    }
    case MM3E_CSG_OFFSET:
    case MM3E_CSG_SHELL: {
      csg_emit(sc, n.a, frames, out);
      DCsgOp o{};
      o.op = n.op == MM3E_CSG_OFFSET ? CSG_OFFSET : CSG_SHELL;
      o.p[0] = n.p[0];
      out.push_back(o);
      return;
    }
    default: {
      const bool swap = csg_need(sc, n.b) > csg_need(sc, n.a);
      csg_emit(sc, swap ? n.b : n.a, frames, out);
      csg_emit(sc, swap ? n.a : n.b, frames, out);
      DCsgOp o{};
      o.op = n.op == MM3E_CSG_UNION ? CSG_UNION
             : n.op == MM3E_CSG_INTERSECT ? CSG_INTERSECT
             : n.op == MM3E_CSG_SUBTRACT ? CSG_SUBTRACT : CSG_SMOOTH;
      o.swap = swap ? 1u : 0u;
      o.p[0] = n.p[0];
      out.push_back(o);
    }
  }
}
}  // namespace

void csg_compile(const HostScene& sc, uint32_t id, std::vector<DCsgOp>& out) {
  std::vector<uint32_t> frames;
  csg_emit(sc, sc.csg_roots[id], frames, out);
}

namespace {

Sphere local_bound(const HostScene& sc, const mm3e_ir_object& o) {  // lib.rs:264-292
  const float* p = o.p;
  auto s0 = [](float r) { return Sphere{true, {0.0f, 0.0f, 0.0f}, r}; };
  switch (o.prim) {
    case MM3E_PRIM_SPHERE: return s0(p[0]);
    case MM3E_PRIM_BOX:
    case MM3E_PRIM_ROUND_BOX: return s0(len(f3(p)));
    case MM3E_PRIM_TORUS: return s0(p[0] + p[1]);
    case MM3E_PRIM_CYLINDER: return s0(std::sqrt(p[0] * p[0] + p[1] * p[1]));
    case MM3E_PRIM_CAPSULE: {
      F3 a = f3(p), b = f3(p + 3);
      F3 c = scl(add(a, b), 0.5f);
      return Sphere{true, {c.x, c.y, c.z}, len(scl(sub(a, b), 0.5f)) + p[6]};
    }
    case MM3E_PRIM_CONE:
      return Sphere{true, {0.0f, p[2] * 0.5f, 0.0f}, std::fmax(std::fmax(p[2] * 0.5f + std::fmax(p[0], p[1]), p[0]), p[1])};
    case MM3E_PRIM_ELLIPSOID: return s0(std::fmax(std::fmax(p[0], p[1]), p[2]));
    case MM3E_PRIM_OCTAHEDRON: return s0(p[0]);
    case MM3E_PRIM_HEX_PRISM: return s0(std::sqrt(p[0] * p[0] + p[1] * p[1]));
    case MM3E_PRIM_CSG: {
      uint32_t id;
      std::memcpy(&id, &o.p[0], 4);
      return csg_bound(sc, sc.csg_roots[id]);
    }
    case MM3E_PRIM_VOLUME: {
      if (sc.volumes.empty()) return Sphere{false, {0, 0, 0}, 0};
      uint32_t vid;
      std::memcpy(&vid, &o.p[0], 4);
      const mm3e_ir_volume& v = sc.volumes[vid % sc.volumes.size()];
      F3 lo = f3(v.min);
      F3 sz{(float)(v.dims[0] - 1) * v.cell[0], (float)(v.dims[1] - 1) * v.cell[1], (float)(v.dims[2] - 1) * v.cell[2]};
      F3 hi = add(lo, sz);
      F3 c = scl(add(lo, hi), 0.5f);
      return Sphere{true, {c.x, c.y, c.z}, len(scl(sub(hi, lo), 0.5f))};
    }
    default: return Sphere{false, {0, 0, 0}, 0};  // plane (unbounded)
  }
}

// accel.rs:39-77 median-split tree; children are pushed before their parent.
struct Item {
  uint32_t i;
  F3 c;
  float r;
};
int32_t bvh_build(Item* items, size_t n, std::vector<DNode>& nodes, uint32_t base) {
  if (n == 1) {
    DNode nd{};
    nd.c[0] = items[0].c.x; nd.c[1] = items[0].c.y; nd.c[2] = items[0].c.z;
    nd.r = items[0].r;
    nd.left = nd.right = 0;
    nd.leaf = (int32_t)items[0].i;
    nd.parent = -1;
    nodes.push_back(nd);
    return (int32_t)(base + nodes.size() - 1);
  }
  F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (size_t k = 0; k < n; k++) {
    lo = vmin(lo, items[k].c);
    hi = vmax(hi, items[k].c);
  }
  F3 ext = sub(hi, lo);
  int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
  auto coord = [axis](const F3& c) { return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); };
  std::sort(items, items + n, [&](const Item& a, const Item& b) {
    float ca = coord(a.c), cb = coord(b.c);
    if (ca < cb) return true;
    if (ca > cb) return false;
    return a.i < b.i;
  });
  size_t mid = n / 2;
  int32_t l = bvh_build(items, mid, nodes, base);
  int32_t r = bvh_build(items + mid, n - mid, nodes, base);
  const DNode& L = nodes[l - base];
  const DNode& R = nodes[r - base];
  Sphere e = enclose(f3(L.c), L.r, f3(R.c), R.r);
  DNode nd{};
  nd.c[0] = e.c[0]; nd.c[1] = e.c[1]; nd.c[2] = e.c[2];
  nd.r = e.r;
  nd.left = l;
  nd.right = r;
  nd.leaf = -1;
  nd.parent = -1;
  nodes.push_back(nd);
  int32_t me = (int32_t)(base + nodes.size() - 1);
  nodes[l - base].parent = me;
  nodes[r - base].parent = me;
  return me;
}

template <class T> void put(std::vector<uint8_t>& b, uint64_t off, const T* src, size_t n) {
  std::memcpy(b.data() + off, src, n * sizeof(T));
}
inline uint64_t align16(uint64_t v) { return (v + 15) & ~uint64_t(15); }

}  // namespace

bool parse_scene_ir(const void* ir, size_t len_, HostScene& sc, std::string& err) {
  const uint8_t* b = (const uint8_t*)ir;
  if (!ir || len_ < sizeof(mm3e_ir_header)) {
    err = "scene IR is shorter than its header";
    return false;
  }
  std::memcpy(&sc.hdr, b, sizeof sc.hdr);
  const mm3e_ir_header& h = sc.hdr;
  if (std::memcmp(h.magic, MM3E_IR_MAGIC, 8) != 0 || h.version != MM3E_IR_VERSION) {
    err = "scene IR magic/version mismatch";
    return false;
  }
  if (h.total_bytes != len_) {
    err = "scene IR length does not match its header";
    return false;
  }
  auto span = [&](uint32_t off, uint64_t bytes) { return (uint64_t)off + bytes <= len_ && (off % 4) == 0; };
  if (!span(h.off_objects, (uint64_t)h.n_objects * sizeof(mm3e_ir_object)) ||
      !span(h.off_materials, (uint64_t)h.n_materials * sizeof(mm3e_ir_material)) ||
      !span(h.off_lights, (uint64_t)h.n_lights * sizeof(mm3e_ir_light)) ||
      !span(h.off_volumes, (uint64_t)h.n_volumes * sizeof(mm3e_ir_volume)) ||
      !span(h.off_volume_data, (uint64_t)h.n_volume_floats * 4) ||
      !span(h.off_csg_nodes, (uint64_t)h.n_csg_nodes * sizeof(mm3e_ir_csg_node)) ||
      !span(h.off_csg_roots, (uint64_t)h.n_csg_exprs * 4)) {
    err = "scene IR array lies outside the blob";
    return false;
  }
  if (h.mode > 5) {
    err = "unknown render mode in scene IR";
    return false;
  }
  if (h.appearance) {  // wgsl.rs:42-44
    err = "UV textures require the checked CPU renderer";
    return false;
  }
  sc.objects.resize(h.n_objects);
  if (h.n_objects) std::memcpy(sc.objects.data(), b + h.off_objects, h.n_objects * sizeof(mm3e_ir_object));
  sc.materials.resize(h.n_materials);
  if (h.n_materials) std::memcpy(sc.materials.data(), b + h.off_materials, h.n_materials * sizeof(mm3e_ir_material));
  sc.lights.resize(h.n_lights);
  if (h.n_lights) std::memcpy(sc.lights.data(), b + h.off_lights, h.n_lights * sizeof(mm3e_ir_light));
  sc.volumes.resize(h.n_volumes);
  if (h.n_volumes) std::memcpy(sc.volumes.data(), b + h.off_volumes, h.n_volumes * sizeof(mm3e_ir_volume));
  sc.volume_data.resize(h.n_volume_floats);
  if (h.n_volume_floats) std::memcpy(sc.volume_data.data(), b + h.off_volume_data, (size_t)h.n_volume_floats * 4);
  sc.csg_nodes.resize(h.n_csg_nodes);
  if (h.n_csg_nodes) std::memcpy(sc.csg_nodes.data(), b + h.off_csg_nodes, h.n_csg_nodes * sizeof(mm3e_ir_csg_node));
  sc.csg_roots.resize(h.n_csg_exprs);
  if (h.n_csg_exprs) std::memcpy(sc.csg_roots.data(), b + h.off_csg_roots, (size_t)h.n_csg_exprs * 4);
  // csg.rs:160-235 budgets: depth <= 64 (root = 1), <= 1024 nodes; children must exist.
  for (uint32_t e = 0; e < h.n_csg_exprs; e++) {
    uint32_t count = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{sc.csg_roots[e], 1u}};
    while (!stack.empty()) {
      auto [ni, depth] = stack.back();
      stack.pop_back();
      if (ni >= h.n_csg_nodes || depth > 64 || ++count > 1024) {
        err = "CSG expression exceeds depth/node budget";
        return false;
      }
      const mm3e_ir_csg_node& n = sc.csg_nodes[ni];
      if (n.op > MM3E_CSG_SHELL || (n.op == MM3E_CSG_PRIM && n.prim > MM3E_PRIM_PLANE)) {
        err = "invalid CSG primitive or modifier parameters";
        return false;
      }
      if (n.op == MM3E_CSG_PRIM) continue;
      stack.push_back({n.a, depth + 1});
      if (n.op >= MM3E_CSG_UNION && n.op <= MM3E_CSG_SMOOTH_UNION) stack.push_back({n.b, depth + 1});
    }
  }
  for (const auto& o : sc.objects) {
    if (o.prim == MM3E_PRIM_SURFACE) {  // wgsl.rs:45-47
      err = "native triangle surfaces are not supported by the GPU backend; use the CPU renderer";
      return false;
    }
    if (o.prim == MM3E_PRIM_CSG) {
      uint32_t id;
      std::memcpy(&id, &o.p[0], 4);
      if (id >= h.n_csg_exprs) {  // lib.rs:234 expect("unregistered CSG expression")
        err = "unregistered CSG expression";
        return false;
      }
    }
    if (o.prim > MM3E_PRIM_SURFACE) {
      err = "unknown primitive kind in scene IR";
      return false;
    }
    if (o.combine > MM3E_COMBINE_SUBTRACT) {
      err = "unknown combine mode in scene IR";
      return false;
    }
  }
  if (sc.materials.empty()) {  // appearance.rs:251-253
    err = "scene needs at least one material";
    return false;
  }
  for (const auto& v : sc.volumes) {
    // products in 128 bits: three u32 dims can exceed 2^64 (volume.rs:44, 77-80 use checked_mul)
    const unsigned __int128 n = (unsigned __int128)v.dims[0] * v.dims[1] * v.dims[2];
    if ((unsigned __int128)v.offset + n > h.n_volume_floats) {
      err = "SdfVolume samples lie outside the IR volume data";
      return false;
    }
  }
  // probe counts above 2^32 (1.5 TiB of cubes) cannot be allocated; rejecting them also keeps
  // every later probe-count product inside 64 bits
  const unsigned __int128 gi_probes = (unsigned __int128)std::max(h.gi_dims[0], 1u) * std::max(h.gi_dims[1], 1u) *
                                      std::max(h.gi_dims[2], 1u);
  if (h.gi_mode != MM3E_GI_NONE && gi_probes > ((unsigned __int128)1 << 32)) {
    err = "GI probe volume dims overflow";
    return false;
  }
  if (h.gi_mode == MM3E_GI_BAKED) {
    uint64_t n = (uint64_t)h.gi_dims[0] * h.gi_dims[1] * h.gi_dims[2] * 18;
    if (n == 0 || !span(h.off_gi_cubes, n * 4)) {
      err = "baked GI cubes lie outside the IR";
      return false;
    }
    sc.gi_cubes.resize(n);
    std::memcpy(sc.gi_cubes.data(), b + h.off_gi_cubes, n * 4);
  } else if (h.gi_mode > MM3E_GI_BAKE) {
    err = "unknown GI mode in scene IR";
    return false;
  }
  uint32_t aa = h.aa == 0 ? 1 : h.aa;
  if (aa > kMaxAA) {
    err = "aa > 256 is not supported by the MI355X render path";
    return false;
  }
  if (h.bounces > kMaxBounces) {
    err = "bounces > 16 is not supported by the MI355X render path";
    return false;
  }
  if (h.bloom && h.bloom_radius > kMaxBloomRadius) {  // its 2r+1 Gaussian weights live in HBM
    err = "bloom_radius > 1048576 is not supported by the MI355X render path";
    return false;
  }
  return true;
}

Sphere world_bound(const HostScene& sc, const mm3e_ir_object& o) {  // lib.rs:470-496
  Sphere s = local_bound(sc, o);
  if (!s.some) return s;
  if (nonzero3(o.repeat) || o.bend != 0.0f) return Sphere{false, {0, 0, 0}, 0};
  F3 center{s.c[0], s.c[1], s.c[2]};
  float radius = s.r;
  if (o.mirror & 7u) {
    radius += len(center);
    center = F3{0.0f, 0.0f, 0.0f};
  }
  if (o.twist != 0.0f) {
    radius += std::sqrt(center.x * center.x + center.z * center.z);
    center.x = 0.0f;
    center.z = 0.0f;
  }
  radius += len(f3(o.elongate));
  F3 wc = add(f3(o.pos), mul_vec(o.rot, scl(center, o.scale)));
  return Sphere{true, {wc.x, wc.y, wc.z}, radius * o.scale + std::fmax(o.round, 0.0f) + std::fabs(o.onion)};
}

// Local bound of a primitive whose SDF is exact: SDF(q) >= |q - c| - r with |c| returned in
// *c.  False for the primitives whose SDF is only a bound (or not exact for these parameters).
static bool exact_prim_bound(const mm3e_ir_object& o, double* c_out, double* r_out) {
  double c = 0.0, r = 0.0;
  const float* p = o.p;
  auto n3 = [](double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); };
  switch (o.prim) {
    case MM3E_PRIM_SPHERE: r = p[0]; break;
    case MM3E_PRIM_BOX: r = n3(p[0], p[1], p[2]); break;
    case MM3E_PRIM_ROUND_BOX:  // sdf.rs:44-46: exact while the radius fits every half extent
      if (!(p[3] >= 0.0f && p[3] <= std::min(p[0], std::min(p[1], p[2])))) return false;
      r = n3(p[0], p[1], p[2]);
      break;
    case MM3E_PRIM_TORUS: r = (double)p[0] + p[1]; break;
    case MM3E_PRIM_CYLINDER: r = n3(p[0], p[1], 0.0); break;
    case MM3E_PRIM_CAPSULE:
      c = n3(0.5 * ((double)p[0] + p[3]), 0.5 * ((double)p[1] + p[4]), 0.5 * ((double)p[2] + p[5]));
      r = 0.5 * n3((double)p[0] - p[3], (double)p[1] - p[4], (double)p[2] - p[5]) + p[6];
      break;
    case MM3E_PRIM_CONE:
      if (!(std::fabs(p[0] - p[1]) < p[2])) return false;
      c = 0.5 * std::fabs((double)p[2]);
      r = 0.5 * std::fabs((double)p[2]) + std::max((double)p[0], (double)p[1]);
      break;
    default: return false;
  }
  *c_out = c;
  *r_out = r;
  return true;
}

// A sphere below the field of an object the reference leaves unbounded because of a bend
// (lib.rs:478-480).  Mirror, twist and bend map the local point by reflections and rotations,
// which keep |l|; elongation moves it by at most |h|; so for a primitive whose SDF is exact
// (SDF(q) >= |q - c| - r for its local bound (c, r)) the object's field is at least
// (|l| - |h| - |c| - r) * scale - max(round, 0) - |onion|, and |l| * scale = |p - pos| for a
// rigid transform with positive uniform scale.  Repetition folds space and has no such bound.
// The radius is computed in double and rounded up; the device's far test adds its own margin
// for the float rounding of the object's evaluation.
Sphere skip_bound(const mm3e_ir_object& o) {
  const Sphere none{false, {0, 0, 0}, 0};
  if (o.bend == 0.0f || nonzero3(o.repeat) || !(o.scale > 0.0f) || !std::isfinite(o.scale)) return none;
  double c = 0.0, r = 0.0;
  if (!exact_prim_bound(o, &c, &r)) return none;
  const double h = std::sqrt((double)o.elongate[0] * o.elongate[0] + (double)o.elongate[1] * o.elongate[1] +
                             (double)o.elongate[2] * o.elongate[2]);
  double rad = (h + c + r) * (double)o.scale + std::max((double)o.round, 0.0) + std::fabs((double)o.onion);
  rad = rad * (1.0 + 1e-6) + 1e-6;
  if (!std::isfinite(rad) || !(rad > 0.0)) return none;
  return Sphere{true, {o.pos[0], o.pos[1], o.pos[2]}, (float)rad};
}

bool gi_bounds(const HostScene& sc, float lo[3], float hi[3]) {
  F3 l{INFINITY, INFINITY, INFINITY}, h{-INFINITY, -INFINITY, -INFINITY};
  bool any = false;
  for (const auto& o : sc.objects) {
    Sphere s = world_bound(sc, o);
    if (!s.some) continue;
    any = true;
    F3 c = f3(s.c);
    l = vmin(l, sub(c, F3{s.r, s.r, s.r}));
    h = vmax(h, add(c, F3{s.r, s.r, s.r}));
  }
  if (!any) return false;
  l = sub(l, F3{1.5f, 1.5f, 1.5f});
  h = add(h, F3{1.5f, 1.5f, 1.5f});
  lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
  hi[0] = h.x; hi[1] = h.y; hi[2] = h.z;
  return true;
}

bool scene_dual_safe(const HostScene& sc) {  // lib.rs:426-430
  for (const auto& o : sc.objects) {
    if (o.twist != 0.0f || o.bend != 0.0f) return false;
    if (o.prim == MM3E_PRIM_VOLUME || o.prim == MM3E_PRIM_CSG || o.prim == MM3E_PRIM_SURFACE) return false;
  }
  return true;
}

void gi_axis_dirs(uint32_t samples, float dirs[6][5][4], uint32_t* rays_per_face) {  // gi.rs:38-50
  static const float axes[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  std::memset(dirs, 0, sizeof(float) * 6 * 5 * 4);
  *rays_per_face = samples > 1 ? 5 : 1;
  for (int a = 0; a < 6; a++) {
    F3 axis = f3(axes[a]);
    F3 up = std::fabs(axis.y) > 0.9f ? F3{1.0f, 0.0f, 0.0f} : F3{0.0f, 1.0f, 0.0f};
    F3 t1 = normalize(cross(axis, up));
    F3 t2 = normalize(cross(axis, t1));
    dirs[a][0][0] = axis.x; dirs[a][0][1] = axis.y; dirs[a][0][2] = axis.z; dirs[a][0][3] = 1.0f;
    if (samples > 1) {
      F3 ts[4] = {t1, t2, scl(t1, -1.0f), scl(t2, -1.0f)};
      for (int k = 0; k < 4; k++) {
        F3 d = normalize(add(scl(axis, 0.72f), scl(ts[k], 0.72f)));
        dirs[a][k + 1][0] = d.x; dirs[a][k + 1][1] = d.y; dirs[a][k + 1][2] = d.z;
        dirs[a][k + 1][3] = std::fmax(dot(d, axis), 0.0f);
      }
    }
  }
}

std::vector<float> gaussian_kernel(uint32_t radius) {  // post.rs:78-92
  int r = (int)std::max(radius, 1u);
  float sigma = std::fmax((float)r / 2.0f, 1e-3f);
  std::vector<float> k;
  float sum = 0.0f;
  for (int i = -r; i <= r; i++) {
    float v = expf((float)(-(i * i)) / (2.0f * sigma * sigma));
    k.push_back(v);
    sum += v;
  }
  for (float& v : k) v /= sum;
  return k;
}

// Segment candidate grid.  Every segment gets a sphere enclosing the bounding spheres of all
// the objects it folds (computed in double and rounded outwards), or none when one of them is
// unbounded (then the segment is a candidate everywhere).  A cell's level-L mask keeps the
// segments whose sphere comes within m[L] of the cell, with the cell grown by a margin that
// covers the float rounding of the device's cell index.  The device may drop the clear bits of
// a lane's cell once max(acc + kmax, slack) < m[L] (after margins): every member of a dropped
// segment then has a bounding-sphere lower bound above the threshold at which folding it is a
// no-op (union: lower > acc; smooth union: lower - acc > k so h == 1; tree walks prune or keep
// acc), and since acc only decreases along a fold without subtract segments it stays so.
void build_seg_grid(const std::vector<DSeg>& segs, const std::vector<std::vector<uint32_t>>& members,
                    const std::vector<Sphere>& bounds, float slack, SegGrid& g) {
  g = SegGrid{};
  const size_t n = segs.size();
  if (n < kGridMinSegments) return;
  double kmax = 0.0;
  for (const DSeg& sg : segs) {
    if (sg.kind == 0 && sg.combine == MM3E_COMBINE_SUBTRACT) return;
    if (sg.kind == 0 && sg.combine == MM3E_COMBINE_SMOOTH && sg.k > 0.0f) kmax = std::max(kmax, (double)sg.k);
  }
  struct S { bool always; double c[3], r; };
  std::vector<S> sph(n);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  bool any_bounded = false;
  for (size_t s = 0; s < n; s++) {
    S& q = sph[s];
    q.always = false;
    double bl[3] = {1e300, 1e300, 1e300}, bh[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i : members[s]) {
      if (!bounds[i].some || !std::isfinite(bounds[i].r)) { q.always = true; break; }
      for (int a = 0; a < 3; a++) {
        bl[a] = std::min(bl[a], (double)bounds[i].c[a] - (double)bounds[i].r);
        bh[a] = std::max(bh[a], (double)bounds[i].c[a] + (double)bounds[i].r);
      }
    }
    if (q.always || members[s].empty()) { q.always = true; continue; }
    for (int a = 0; a < 3; a++) q.c[a] = 0.5 * (bl[a] + bh[a]);
    double r = 0.0;
    for (uint32_t i : members[s]) {
      double dx = q.c[0] - bounds[i].c[0], dy = q.c[1] - bounds[i].c[1], dz = q.c[2] - bounds[i].c[2];
      r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz) + (double)bounds[i].r);
    }
    q.r = r * (1.0 + 1e-9) + 1e-9;
    if (!std::isfinite(q.r)) { q.always = true; continue; }
    any_bounded = true;
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], q.c[a] - q.r);
      hi[a] = std::max(hi[a], q.c[a] + q.r);
    }
  }
  if (!any_bounded) return;
  float in_lo[3], in_hi[3];  // the spheres' box, rounded outwards to float
  for (int a = 0; a < 3; a++) {
    in_lo[a] = std::nextafter((float)lo[a], -INFINITY);
    in_hi[a] = std::nextafter((float)hi[a], INFINITY);
  }
  double mreach[kGridLevels];
  for (uint32_t l = 0; l < kGridLevels; l++) mreach[l] = (double)slack + kGridReach[l];
  const double m1 = (double)slack + kGridPadReach;  // padding of the grid around the spheres
  double ext = 0.0;
  for (int a = 0; a < 3; a++) ext = std::max(ext, hi[a] - lo[a] + 2.0 * m1);
  const double pad = 0.01 + 1e-5 * ext;
  double vol = 1.0;
  for (int a = 0; a < 3; a++) {
    lo[a] -= m1 + pad;
    hi[a] += m1 + pad;
    vol *= hi[a] - lo[a];
  }
  double cs = std::max(std::cbrt(vol / (double)kGridTargetCells), 1e-3 * ext);
  uint32_t dims[3];
  for (int a = 0; a < 3; a++) {
    double d = std::ceil((hi[a] - lo[a]) / cs);
    dims[a] = (uint32_t)std::min(std::max(d, 1.0), (double)kGridMaxDim);
  }
  // widen the cell if a dimension hit the cap so the grid still covers [lo, hi]
  for (int a = 0; a < 3; a++) cs = std::max(cs, (hi[a] - lo[a]) / dims[a]);
  const float csf = (float)cs;
  const float inv = 1.0f / csf;
  const double csd = 1.0 / (double)inv;  // the cell side the device's index actually uses
  const uint64_t n_cells = (uint64_t)dims[0] * dims[1] * dims[2];
  const uint32_t W = (uint32_t)((n + 31) / 32);
  const uint32_t L = kGridLevels;
  // the device indexes the masks with 32-bit word offsets
  if ((n_cells + 1) * L * W > 0xFFFFFFFFull) return;
  g.words.assign((n_cells + 1) * L * W, 0u);
  const double e = 1e-3 * csd + 1e-5 * ext;  // index rounding margin
  float lof[3];
  for (int a = 0; a < 3; a++) lof[a] = (float)lo[a];
  for (size_t s = 0; s < n; s++) {
    const uint32_t word = (uint32_t)(s / 32), bit = 1u << (s % 32);
    if (sph[s].always) {
      for (uint64_t c = 0; c <= n_cells; c++)
        for (uint32_t l = 0; l < L; l++) g.words[(c * L + l) * W + word] |= bit;
      continue;
    }
    // outside the grid: farther than the padding reach from this segment, nearer is unknown
    for (uint32_t l = 0; l < L; l++)
      if (mreach[l] > m1) g.words[(n_cells * L + l) * W + word] |= bit;
    const S& q = sph[s];
    const double mmax = mreach[L - 1];
    int64_t r0[3], r1[3];
    for (int a = 0; a < 3; a++) {
      r0[a] = (int64_t)std::floor((q.c[a] - q.r - mmax - e - (double)lof[a]) / csd) - 1;
      r1[a] = (int64_t)std::floor((q.c[a] + q.r + mmax + e - (double)lof[a]) / csd) + 1;
      r0[a] = std::max<int64_t>(r0[a], 0);
      r1[a] = std::min<int64_t>(r1[a], (int64_t)dims[a] - 1);
    }
    for (int64_t z = r0[2]; z <= r1[2]; z++)
      for (int64_t y = r0[1]; y <= r1[1]; y++)
        for (int64_t x = r0[0]; x <= r1[0]; x++) {
          const int64_t ix[3] = {x, y, z};
          double d2 = 0.0;
          for (int a = 0; a < 3; a++) {
            const double blo = (double)lof[a] + ix[a] * csd - e, bhi = (double)lof[a] + (ix[a] + 1) * csd + e;
            const double t = q.c[a] < blo ? blo - q.c[a] : (q.c[a] > bhi ? q.c[a] - bhi : 0.0);
            d2 += t * t;
          }
          const double tag = std::sqrt(d2) - q.r;
          const uint64_t c = ((uint64_t)z * dims[1] + (uint64_t)y) * dims[0] + (uint64_t)x;
          for (uint32_t l = 0; l < L; l++)
            if (tag <= mreach[l]) g.words[(c * L + l) * W + word] |= bit;
        }
  }
  for (int a = 0; a < 3; a++) {
    g.lo[a] = lof[a];
    g.dims[a] = dims[a];
    g.in_lo[a] = in_lo[a];
    g.in_hi[a] = in_hi[a];
  }
  g.inv = inv;
  g.n_words = W;
  for (uint32_t l = 0; l < L; l++) g.m[l] = (float)mreach[l];
  g.kmax = (float)kmax;
  // absolute margin for the device's float bounding-sphere lower bounds inside the grid
  g.eps = (float)(1e-4 + 4e-6 * (ext + std::fabs(lo[0]) + std::fabs(lo[1]) + std::fabs(lo[2])));
}

// Layout: [DevScene][objects][segments][nodes][unbounded][materials][lights][volumes]
//         [volume data][gi][segment grid][csg programs]
uint64_t blob_size(const HostScene& sc, const std::vector<mm3e_ir_object>& extra_objects,
                   const std::vector<mm3e_ir_material>& extra_materials) {
  Blob b;
  build_blob(sc, extra_objects, extra_materials, 1, 0, 0, nullptr, false, b);
  return b.bytes.size();
}

void build_blob(const HostScene& sc, const std::vector<mm3e_ir_object>& extra_objects,
                const std::vector<mm3e_ir_material>& extra_materials, uint32_t aa, uint32_t bounces,
                uint64_t dev_base, const float* gi_dev, bool with_gi_flag, Blob& out) {
  std::vector<mm3e_ir_object> objs = sc.objects;
  objs.insert(objs.end(), extra_objects.begin(), extra_objects.end());
  std::vector<mm3e_ir_material> mats = sc.materials;
  mats.insert(mats.end(), extra_materials.begin(), extra_materials.end());
  const mm3e_ir_header& h = sc.hdr;
  const uint64_t n_obj = objs.size();

  // CSG expressions -> postfix programs (one per registered expression)
  std::vector<DCsgOp> csg_ops;
  std::vector<uint32_t> csg_start(sc.csg_roots.size()), csg_count(sc.csg_roots.size());
  for (uint32_t e = 0; e < sc.csg_roots.size(); e++) {
    csg_start[e] = (uint32_t)csg_ops.size();
    csg_compile(sc, e, csg_ops);
    csg_count[e] = (uint32_t)csg_ops.size() - csg_start[e];
  }

  // objects + bounds
  std::vector<DObj> dobj(n_obj);
  std::vector<Sphere> bounds(n_obj);       // Object::world_bound (the reference's plan and prune)
  std::vector<Sphere> grid_bounds(n_obj);  // the same, or a skip sphere (candidate grid only)
  float max_k = 0.0f;
  for (uint64_t i = 0; i < n_obj; i++) {
    const mm3e_ir_object& o = objs[i];
    DObj& d = dobj[i];
    std::memset(&d, 0, sizeof d);
    std::memcpy(d.c0, o.rot, 12);
    std::memcpy(d.c1, o.rot + 3, 12);
    std::memcpy(d.c2, o.rot + 6, 12);
    d.inv = std::fabs(o.scale) > 1e-12f ? 1.0f / o.scale : 0.0f;
    d.scale = o.scale;
    d.mat = o.mat;
    std::memcpy(d.pos, o.pos, 12);
    std::memcpy(d.p, o.p, 32);
    if (o.prim == MM3E_PRIM_CSG) {
      uint32_t id;
      std::memcpy(&id, &o.p[0], 4);
      std::memcpy(&d.p[0], &csg_start[id], 4);
      std::memcpy(&d.p[1], &csg_count[id], 4);
    }
    std::memcpy(d.rep, o.repeat, 12);
    d.twist = o.twist;
    std::memcpy(d.elo, o.elongate, 12);
    d.bend = o.bend;
    d.round = o.round;
    d.onion = o.onion;
    d.k = o.k;
    uint32_t f = o.prim & OBJ_KIND_MASK;
    if (o.mirror & 1u) f |= OBJ_MIRROR_X;
    if (o.mirror & 2u) f |= OBJ_MIRROR_Y;
    if (o.mirror & 4u) f |= OBJ_MIRROR_Z;
    if (nonzero3(o.elongate)) f |= OBJ_ELONGATE;
    if (nonzero3(o.repeat)) f |= OBJ_REPEAT;
    if (o.twist != 0.0f) f |= OBJ_TWIST;
    if (o.bend != 0.0f) f |= OBJ_BEND;
    if (o.round != 0.0f) f |= OBJ_ROUND;
    if (o.onion != 0.0f) f |= OBJ_ONION;
    f |= (o.combine & 3u) << OBJ_COMBINE_SHIFT;
    static const float ident[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (std::memcmp(o.rot, ident, sizeof ident) == 0) f |= OBJ_IDENT_ROT;
    if (o.scale == 1.0f) f |= OBJ_UNIT_SCALE;
    bounds[i] = world_bound(sc, o);
    grid_bounds[i] = bounds[i];
    if (bounds[i].some) {
      f |= OBJ_BOUNDED;
      std::memcpy(d.bc, bounds[i].c, 12);
      d.bc_r = bounds[i].r;
    } else {
      Sphere k = skip_bound(o);
      if (k.some) {
        f |= OBJ_SKIP;
        std::memcpy(d.bc, k.c, 12);
        d.bc_r = k.r;
        grid_bounds[i] = k;
      }
    }
    d.flags = f;
    max_k = std::fmax(max_k, o.combine == MM3E_COMBINE_SMOOTH ? o.k : 0.0f);
  }
  const float slack = max_k + 0.1f;  // accel.rs:141-148

  // world plan (accel.rs:134-185)
  std::vector<DSeg> segs;
  std::vector<DNode> nodes;
  std::vector<uint32_t> unb;
  std::vector<uint32_t> run;
  std::vector<std::vector<uint32_t>> seg_members;  // objects folded by each segment
  // An object is "exact-bounded" when its field value never drops below its bounding-sphere
  // lower bound |p - c| - r: an exact Euclidean SDF (sphere, box, rounded box, torus, capped
  // cylinder, capsule, round cone with |r1 - r2| < h) placed by a rigid transform with positive
  // scale, optionally mirrored / elongated / rounded / hollowed (all accounted for in the
  // bound).  Unbounded members are evaluated exactly anyway.  For a run of such objects the
  // reference's tree query returns exactly the index-order minimum.
  auto exact_bounded = [&](uint32_t i) {
    const mm3e_ir_object& o = objs[i];
    if (!bounds[i].some) return true;
    if (!(o.scale > 0.0f) || o.twist != 0.0f) return false;
    switch (o.prim) {
      case MM3E_PRIM_SPHERE: case MM3E_PRIM_BOX: case MM3E_PRIM_TORUS:
      case MM3E_PRIM_CYLINDER: case MM3E_PRIM_CAPSULE:
        return true;
      case MM3E_PRIM_ROUND_BOX:  // sdf.rs:44-46: exact while the radius fits every half extent
        return o.p[3] >= 0.0f && o.p[3] <= std::min(o.p[0], std::min(o.p[1], o.p[2]));
      case MM3E_PRIM_CONE: return std::fabs(o.p[0] - o.p[1]) < o.p[2];
      default: return false;
    }
  };
  auto close_run = [&]() {
    if (run.empty()) return;
    bool linear = run.size() <= kLinearRunMax;
    for (uint32_t i : run) linear = linear && exact_bounded(i);
    if (linear) {
      DSeg s{};
      s.kind = 2;
      s.unb_base = (uint32_t)unb.size();
      for (uint32_t i : run) unb.push_back(i);
      s.unb_count = (uint32_t)run.size();
      segs.push_back(s);
      seg_members.push_back(run);
      run.clear();
      return;
    }
    DSeg s{};
    s.kind = 1;
    s.unb_base = (uint32_t)unb.size();
    std::vector<Item> items;
    for (uint32_t i : run) {
      if (bounds[i].some) items.push_back(Item{i, f3(bounds[i].c), bounds[i].r});
      else unb.push_back(i);
    }
    s.unb_count = (uint32_t)unb.size() - s.unb_base;
    if (!items.empty()) {
      std::vector<DNode> local;
      uint32_t base = (uint32_t)nodes.size();
      int32_t root = bvh_build(items.data(), items.size(), local, base);
      nodes.insert(nodes.end(), local.begin(), local.end());
      s.has_tree = 1;
      s.index = (uint32_t)root;
    }
    segs.push_back(s);
    seg_members.push_back(run);
    run.clear();
  };
  for (uint32_t i = 0; i < n_obj; i++) {
    if (objs[i].combine == MM3E_COMBINE_UNION) {
      run.push_back(i);
    } else {
      close_run();
      DSeg s{};
      s.kind = 0;
      s.index = i;
      s.combine = objs[i].combine;
      s.k = objs[i].k;
      segs.push_back(s);
      seg_members.push_back(std::vector<uint32_t>{i});
    }
  }
  close_run();

  // Segment candidate grid (dynamic-sphere frames rebuild their blob per frame and skip it).
  SegGrid grid;
  if (extra_objects.empty()) build_seg_grid(segs, seg_members, grid_bounds, slack, grid);

  std::vector<DMat> dmat(mats.size());
  for (size_t i = 0; i < mats.size(); i++) {
    const mm3e_ir_material& m = mats[i];
    DMat& d = dmat[i];
    std::memset(&d, 0, sizeof d);
    std::memcpy(d.albedo, m.albedo, 12);
    d.metallic = m.metallic;
    d.specular = m.specular;
    d.roughness = m.roughness;
    d.reflectivity = m.reflectivity;
    d.checker = m.checker ? 1u : 0u;
    std::memcpy(d.emissive, m.emissive, 12);
  }
  std::vector<DLight> dl(sc.lights.size());
  for (size_t i = 0; i < sc.lights.size(); i++) {
    const mm3e_ir_light& l = sc.lights[i];
    std::memcpy(dl[i].vec, l.vec, 12);
    dl[i].radius = l.radius;
    std::memcpy(dl[i].color, l.color, 12);
    dl[i].directional = l.directional ? 1u : 0u;
  }
  std::vector<DVol> dv(sc.volumes.size());
  for (size_t i = 0; i < sc.volumes.size(); i++) {
    const mm3e_ir_volume& v = sc.volumes[i];
    DVol& d = dv[i];
    std::memset(&d, 0, sizeof d);
    std::memcpy(d.dims, v.dims, 12);
    d.offset = v.offset;
    std::memcpy(d.min, v.min, 12);
    std::memcpy(d.cell, v.cell, 12);
    bool ok = v.dims[0] >= 2 && v.dims[1] >= 2 && v.dims[2] >= 2;
    for (int a = 0; a < 3; a++) ok = ok && std::isfinite(v.min[a]) && std::isfinite(v.cell[a]) && v.cell[a] > 0.0f;
    d.valid = ok ? 1u : 0u;
    for (int a = 0; a < 3; a++) d.hi[a] = v.min[a] + (float)(v.dims[a] - 1) * v.cell[a];  // volume.rs:97-103
  }

  // offsets
  uint64_t off = align16(sizeof(DevScene));
  const uint64_t o_obj = off; off += align16(n_obj * sizeof(DObj));
  const uint64_t o_seg = off; off += align16((n_obj + 1) * sizeof(DSeg));
  const uint64_t o_node = off; off += align16((2 * n_obj + 1) * sizeof(DNode));
  const uint64_t o_unb = off; off += align16((n_obj + 1) * sizeof(uint32_t));
  const uint64_t o_mat = off; off += align16(mats.size() * sizeof(DMat));
  const uint64_t o_light = off; off += align16((sc.lights.size() + 1) * sizeof(DLight));
  const uint64_t o_vol = off; off += align16((sc.volumes.size() + 1) * sizeof(DVol));
  const uint64_t o_vdata = off; off += align16((sc.volume_data.size() + 1) * sizeof(float));
  uint64_t o_gi = 0, gi_bytes = 0;
  if (h.gi_mode == MM3E_GI_BAKED) {
    o_gi = off;
    gi_bytes = (uint64_t)h.gi_dims[0] * h.gi_dims[1] * h.gi_dims[2] * 24 * 4;
    off += align16(gi_bytes);
  }
  uint64_t o_grid = 0, o_csg = 0;
  if (!grid.words.empty()) {
    o_grid = off;
    off += align16(grid.words.size() * sizeof(uint32_t));
  }
  if (!csg_ops.empty()) {
    o_csg = off;
    off += align16(csg_ops.size() * sizeof(DCsgOp));
  }
  out.bytes.assign(off, 0);
  put(out.bytes, o_obj, dobj.data(), dobj.size());
  put(out.bytes, o_seg, segs.data(), segs.size());
  put(out.bytes, o_node, nodes.data(), nodes.size());
  put(out.bytes, o_unb, unb.data(), unb.size());
  put(out.bytes, o_mat, dmat.data(), dmat.size());
  put(out.bytes, o_light, dl.data(), dl.size());
  put(out.bytes, o_vol, dv.data(), dv.size());
  put(out.bytes, o_vdata, sc.volume_data.data(), sc.volume_data.size());
  if (o_gi) {
    std::vector<float> g4((size_t)gi_bytes / 4, 0.0f);
    const size_t nf = sc.gi_cubes.size() / 3;
    for (size_t i = 0; i < nf; i++) {
      g4[4 * i + 0] = sc.gi_cubes[3 * i + 0];
      g4[4 * i + 1] = sc.gi_cubes[3 * i + 1];
      g4[4 * i + 2] = sc.gi_cubes[3 * i + 2];
    }
    put(out.bytes, o_gi, g4.data(), g4.size());
  }
  if (o_grid) put(out.bytes, o_grid, grid.words.data(), grid.words.size());
  if (o_csg) put(out.bytes, o_csg, csg_ops.data(), csg_ops.size());
  out.gi_offset = o_gi;
  out.gi_bytes = gi_bytes;
  out.objs = dobj;
  out.segs = segs;
  out.unb = unb;

  DevScene& d = out.header;
  std::memset(&d, 0, sizeof d);
  d.n_obj = (uint32_t)n_obj;
  d.n_seg = (uint32_t)segs.size();
  d.n_mat = (uint32_t)mats.size();
  d.n_light = (uint32_t)sc.lights.size();
  d.max_steps = h.max_steps;
  d.shadow_steps = h.shadow_steps;
  d.ao_samples = h.ao_samples;
  d.bounces = bounces;
  d.n_vol = (uint32_t)sc.volumes.size();
  d.aa = aa;
  uint32_t fl = 0;
  if (h.shadows) fl |= SC_SHADOWS;
  if (h.ao) fl |= SC_AO;
  if (scene_dual_safe(sc) && [&] {
        for (const auto& o : extra_objects)
          if (o.twist != 0.0f || o.bend != 0.0f || o.prim >= MM3E_PRIM_CSG) return false;
        return true;
      }())
    fl |= SC_DUAL_SAFE;
  if (h.secant) fl |= SC_SECANT;
  if (h.step_scale >= 1.0f) fl |= SC_SS1;
  if (with_gi_flag) fl |= SC_HAS_GI;
  d.flags = fl;
  d.max_dist = h.max_dist;
  d.eps = h.eps;
  d.step_scale = h.step_scale;
  d.lod = h.lod_footprint;
  d.subitize = h.subitize;
  d.normal_h = h.normal_h;
  d.slack = slack;
  d.fog_density = h.fog_density;
  std::memcpy(d.sun, h.sun_dir, 12);
  d.exposure = h.exposure;
  std::memcpy(d.ambient, h.ambient, 12);
  d.bloom_threshold = h.bloom_threshold;
  std::memcpy(d.sky_ambient, h.sky_ambient, 12);
  d.bloom_intensity = h.bloom_intensity;
  std::memcpy(d.fog, h.fog, 12);
  d.obj = (const DObj*)(dev_base + o_obj);
  d.seg = (const DSeg*)(dev_base + o_seg);
  d.nodes = (const DNode*)(dev_base + o_node);
  d.unb = (const uint32_t*)(dev_base + o_unb);
  d.mat = (const DMat*)(dev_base + o_mat);
  d.light = (const DLight*)(dev_base + o_light);
  d.vol = (const DVol*)(dev_base + o_vol);
  d.vol_data = (const float*)(dev_base + o_vdata);
  d.gi = gi_dev ? gi_dev : (o_gi ? (const float*)(dev_base + o_gi) : nullptr);
  d.stage_bytes = (uint32_t)(o_mat - o_obj);
  d.st_seg = (uint32_t)(o_seg - o_obj);
  d.st_node = (uint32_t)(o_node - o_obj);
  d.st_unb = (uint32_t)(o_unb - o_obj);
  if (o_grid) {
    for (int a = 0; a < 3; a++) {
      d.grid_lo[a] = grid.lo[a];
      d.grid_dims[a] = grid.dims[a];
      d.grid_in_lo[a] = grid.in_lo[a];
      d.grid_in_hi[a] = grid.in_hi[a];
    }
    d.grid_inv = grid.inv;
    d.grid_words = grid.n_words;
    for (uint32_t l = 0; l < kGridLevels; l++) d.grid_m[l] = grid.m[l];
    d.grid_kmax = grid.kmax;
    d.grid_eps = grid.eps;
    d.grid = (const uint32_t*)(dev_base + o_grid);
  }
  if (o_csg) d.csg = (const DCsgOp*)(dev_base + o_csg);
  std::memcpy(out.bytes.data(), &d, sizeof d);
}

}  // namespace mm3e
