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
#include <cstring>

namespace mm3e {

#ifndef MM3E_LINEAR_RUN_MAX
#define MM3E_LINEAR_RUN_MAX 16
#endif
// Longest run folded linearly; longer runs walk their tree (O(log n) per evaluation).
static const size_t kLinearRunMax = MM3E_LINEAR_RUN_MAX;

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
      !span(h.off_volume_data, (uint64_t)h.n_volume_floats * 4)) {
    err = "scene IR array lies outside the blob";
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
  for (const auto& o : sc.objects) {
    if (o.prim == MM3E_PRIM_SURFACE) {  // wgsl.rs:45-47
      err = "native triangle surfaces are not supported by the GPU backend; use the CPU renderer";
      return false;
    }
    if (o.prim == MM3E_PRIM_CSG) {
      err = "object-local CSG expressions (Prim::Csg) are not supported by the MI355X render path";
      return false;
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
    uint64_t n = (uint64_t)v.dims[0] * v.dims[1] * v.dims[2];
    if ((uint64_t)v.offset + n > h.n_volume_floats) {
      err = "SdfVolume samples lie outside the IR volume data";
      return false;
    }
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
  if (aa > 16) {
    err = "aa > 16 is not supported by the MI355X render path";
    return false;
  }
  if (h.bounces > 4) {
    err = "bounces > 4 is not supported by the MI355X render path";
    return false;
  }
  if (h.bloom && h.bloom_radius > 64) {
    err = "bloom_radius > 64 is not supported by the MI355X render path";
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

// Layout: [DevScene][objects][segments][nodes][unbounded][materials][lights][volumes]
//         [volume data][gi]
uint64_t blob_size(const HostScene& sc, size_t extra_objects, size_t extra_materials) {
  const uint64_t n_obj = sc.objects.size() + extra_objects;
  uint64_t s = align16(sizeof(DevScene));
  s += align16(n_obj * sizeof(DObj));
  s += align16((n_obj + 1) * sizeof(DSeg));
  s += align16((2 * n_obj + 1) * sizeof(DNode));
  s += align16((n_obj + 1) * sizeof(uint32_t));
  s += align16((sc.materials.size() + extra_materials) * sizeof(DMat));
  s += align16((sc.lights.size() + 1) * sizeof(DLight));
  s += align16((sc.volumes.size() + 1) * sizeof(DVol));
  s += align16((sc.volume_data.size() + 1) * sizeof(float));
  if (sc.hdr.gi_mode == MM3E_GI_BAKED) s += align16((uint64_t)sc.hdr.gi_dims[0] * sc.hdr.gi_dims[1] * sc.hdr.gi_dims[2] * 24 * 4);
  return s;
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

  // objects + bounds
  std::vector<DObj> dobj(n_obj);
  std::vector<Sphere> bounds(n_obj);
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
    if (bounds[i].some) {
      f |= OBJ_BOUNDED;
      std::memcpy(d.bc, bounds[i].c, 12);
      d.bc_r = bounds[i].r;
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
      case MM3E_PRIM_SPHERE: case MM3E_PRIM_BOX: case MM3E_PRIM_ROUND_BOX: case MM3E_PRIM_TORUS:
      case MM3E_PRIM_CYLINDER: case MM3E_PRIM_CAPSULE:
        return true;
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
    }
  }
  close_run();

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
  out.gi_offset = o_gi;
  out.gi_bytes = gi_bytes;

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
  std::memcpy(out.bytes.data(), &d, sizeof d);
}

}  // namespace mm3e
