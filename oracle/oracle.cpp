// oracle.cpp — CPU FP32 restatement of the reference MM3E CPU renderer (TEST INFRASTRUCTURE).
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this library
// (liboracle_mm3e.so).  It is the checker for the HIP render path, never part of it.
//
// Restated reference code (paths relative to the reference tree):
//   mm3e-kit/src/{vec,sdf,dual,volume,csg,camera,march,shade,color,framebuffer}.rs, lib.rs:63-78
//   mm3e-orchestrator/src/lib.rs:156-1435, accel.rs, gi.rs, post.rs, appearance.rs:427-431,
//   film.rs:63-225 (render_film, primary-hit AOVs), camera.rs:71-118 (thin lens),
//   lib.rs:951-1080 + reproject.rs:57-170 (G-buffer, warp, reprojection, hybrid),
//   mm3e-cpu-parallel/src/lib.rs:50-84.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../include/mm3e_scene_ir.h"
#include "oracle_core.h"

namespace orc {

thread_local OpCounts* g_ops = nullptr;

enum Phase { PH_MARCH = 0, PH_REFL_MARCH, PH_NORMAL, PH_SHADOW, PH_AO, PH_GI, PH_OTHER, PH_COUNT };
thread_local int g_phase = PH_OTHER;
thread_local uint64_t* g_evals = nullptr;       // [PH_COUNT] field evaluations per phase
thread_local uint64_t* g_obj_exact = nullptr;   // exact object SDF evaluations
thread_local uint64_t* g_obj_bound = nullptr;   // bounding-sphere substitutions

// ---------------------------------------------------------------------------------------
// Parsed scene (float storage; the templated code lifts values to R on use)
// ---------------------------------------------------------------------------------------
struct Obj {
  uint32_t prim, mat, combine;
  float k;
  float p[8];
  uint32_t pid;  // id payload for CSG / VOLUME
  float rot[9], pos[3], scale;
  bool mirror[3];
  float repeat[3], twist, bend, elongate[3], round, onion;
};
struct Mat {
  float albedo[3], metallic, specular, roughness, reflectivity, emissive[3];
  bool checker;
};
struct Lgt {
  float vec[3], radius, color[3];
  bool directional;
};
struct Vol {
  uint32_t dims[3];
  float min[3], cell[3];
  std::vector<float> data;
};
struct CsgNode {
  uint32_t op, prim, a, b;
  float p[8];
  float rot[9], pos[3], scale;
};
struct Gi {
  float min[3], size[3];
  uint32_t dims[3];
  std::vector<float> cubes;  // per probe: 6 faces x RGB
};

struct Scene {
  uint32_t width, height, aa, bounces;
  bool shadows, ao;
  uint32_t mode, appearance;
  float sun_dir[3], ambient[3], sky_ambient[3], fog[3], fog_density;
  // marcher
  uint32_t max_steps, shadow_steps, ao_samples;
  float max_dist, eps, step_scale, lod_footprint, subitize, normal_h;
  bool secant;
  // post
  float exposure, bloom_threshold, bloom_intensity;
  bool bloom;
  uint32_t bloom_radius;
  std::vector<Obj> objects;
  std::vector<Mat> materials;
  std::vector<Lgt> lights;
  std::vector<Vol> volumes;
  std::vector<CsgNode> csg_nodes;
  std::vector<uint32_t> csg_roots;
  uint32_t gi_mode, gi_samples, gi_req_dims[3];
  bool has_gi = false;
  Gi gi;
};

template <class R> inline V3<R> f3(const float* a) { return V3<R>(R(a[0]), R(a[1]), R(a[2])); }
template <class R> inline M3<R> m3(const float* r) {
  M3<R> m;
  m.c[0] = f3<R>(r);
  m.c[1] = f3<R>(r + 3);
  m.c[2] = f3<R>(r + 6);
  return m;
}

// ---------------------------------------------------------------------------------------
// mm3e-kit/src/volume.rs:194-228 — conservative trilinear sample
// ---------------------------------------------------------------------------------------
static bool vol_valid(const Vol& v) {  // volume.rs:47-62
  return v.dims[0] >= 2 && v.dims[1] >= 2 && v.dims[2] >= 2 && std::isfinite(v.min[0]) &&
         std::isfinite(v.min[1]) && std::isfinite(v.min[2]) && std::isfinite(v.cell[0]) &&
         std::isfinite(v.cell[1]) && std::isfinite(v.cell[2]) && v.cell[0] > 0.0f && v.cell[1] > 0.0f &&
         v.cell[2] > 0.0f && (uint64_t)v.dims[0] * v.dims[1] * v.dims[2] == v.data.size();
}
template <class R> static void vol_bounds(const Vol& v, V3<R>& lo, V3<R>& hi) {  // volume.rs:97-103
  R sx = R((float)(v.dims[0] - 1)) * R(v.cell[0]);
  R sy = R((float)(v.dims[1] - 1)) * R(v.cell[1]);
  R sz = R((float)(v.dims[2] - 1)) * R(v.cell[2]);
  lo = f3<R>(v.min);
  hi = lo + V3<R>(sx, sy, sz);
}
template <class R> static R vol_sample(const Vol& v, V3<R> p) {
  if (!vol_valid(v)) return R(INFINITY);
  V3<R> lo, hi;
  vol_bounds(v, lo, hi);
  V3<R> c = clamp_to(p, lo, hi);
  R outside = length(p - c);
  V3<R> g((c.x - R(v.min[0])) / R(v.cell[0]), (c.y - R(v.min[1])) / R(v.cell[1]), (c.z - R(v.min[2])) / R(v.cell[2]));
  const uint64_t nx = v.dims[0], ny = v.dims[1], nz = v.dims[2];
  auto clampi = [](R x, uint64_t n) -> uint64_t { return std::min(to_usize(val(rmax(rfloor(x), R(0.0f)))), n - 2); };
  uint64_t i0 = clampi(g.x, nx), j0 = clampi(g.y, ny), k0 = clampi(g.z, nz);
  R fx = rclamp(g.x - R((float)i0), R(0.0f), R(1.0f));
  R fy = rclamp(g.y - R((float)j0), R(0.0f), R(1.0f));
  R fz = rclamp(g.z - R((float)k0), R(0.0f), R(1.0f));
  auto at = [&](uint64_t i, uint64_t j, uint64_t k) { return R(v.data[(k * ny + j) * nx + i]); };
  auto lerp = [](R a, R b, R t) { return a + (b - a) * t; };
  R x00 = lerp(at(i0, j0, k0), at(i0 + 1, j0, k0), fx);
  R x10 = lerp(at(i0, j0 + 1, k0), at(i0 + 1, j0 + 1, k0), fx);
  R x01 = lerp(at(i0, j0, k0 + 1), at(i0 + 1, j0, k0 + 1), fx);
  R x11 = lerp(at(i0, j0 + 1, k0 + 1), at(i0 + 1, j0 + 1, k0 + 1), fx);
  R d = lerp(lerp(x00, x10, fy), lerp(x01, x11, fy), fz);
  if (!risfinite(d)) return R(INFINITY);
  if (outside > R(0.0f)) return rmax(outside, d - outside);
  return d;
}

// ---------------------------------------------------------------------------------------
// Analytic primitives by kind (lib.rs:215-241) and CSG expressions (csg.rs:95-150)
// ---------------------------------------------------------------------------------------
template <class R> static R prim_distance(uint32_t kind, const float* p, V3<R> l) {
  switch (kind) {
    case MM3E_PRIM_SPHERE: return sd_sphere(l, R(p[0]));
    case MM3E_PRIM_BOX: return sd_box(l, f3<R>(p));
    case MM3E_PRIM_ROUND_BOX: return sd_round_box(l, f3<R>(p), R(p[3]));
    case MM3E_PRIM_TORUS: return sd_torus(l, R(p[0]), R(p[1]));
    case MM3E_PRIM_CYLINDER: return sd_cylinder(l, R(p[0]), R(p[1]));
    case MM3E_PRIM_CAPSULE: return sd_capsule(l, f3<R>(p), f3<R>(p + 3), R(p[6]));
    case MM3E_PRIM_CONE: return sd_round_cone(l, R(p[0]), R(p[1]), R(p[2]));
    case MM3E_PRIM_ELLIPSOID: return sd_ellipsoid(l, f3<R>(p));
    case MM3E_PRIM_OCTAHEDRON: return sd_octahedron(l, R(p[0]));
    case MM3E_PRIM_HEX_PRISM: return sd_hex_prism(l, R(p[0]), R(p[1]));
    case MM3E_PRIM_PLANE: return sd_plane(l, f3<R>(p), R(p[3]));
  }
  return R(INFINITY);
}

template <class R> static R csg_distance(const Scene& sc, uint32_t ni, V3<R> p) {  // csg.rs:95-118
  const CsgNode& n = sc.csg_nodes[ni];
  switch (n.op) {
    case MM3E_CSG_PRIM: return prim_distance<R>(n.prim, n.p, p);
    case MM3E_CSG_TRANSFORM: {
      Xform<R> x{m3<R>(n.rot), f3<R>(n.pos), R(n.scale)};
      return csg_distance<R>(sc, n.a, to_local(x, p)) * R(n.scale);
    }
    case MM3E_CSG_UNION: return rmin(csg_distance<R>(sc, n.a, p), csg_distance<R>(sc, n.b, p));
    case MM3E_CSG_INTERSECT: return rmax(csg_distance<R>(sc, n.a, p), csg_distance<R>(sc, n.b, p));
    case MM3E_CSG_SUBTRACT: return rmax(csg_distance<R>(sc, n.a, p), -csg_distance<R>(sc, n.b, p));
    case MM3E_CSG_SMOOTH_UNION: {
      Field<R> a{csg_distance<R>(sc, n.a, p), 0u}, b{csg_distance<R>(sc, n.b, p), 0u};
      return f_smooth_union(a, b, R(n.p[0])).dist;
    }
    case MM3E_CSG_OFFSET: return csg_distance<R>(sc, n.a, p) - R(n.p[0]);
    case MM3E_CSG_SHELL: return rabs(csg_distance<R>(sc, n.a, p)) - R(n.p[0]);
  }
  return R(INFINITY);
}

struct Sph {
  bool some;
  V3<float> c;
  float r;
};
static Sph enclose(V3<float> c1, float r1, V3<float> c2, float r2) {  // accel.rs:26-36, csg.rs:245-256
  float d = length(c2 - c1);
  if (d + r2 <= r1) return Sph{true, c1, r1};
  if (d + r1 <= r2) return Sph{true, c2, r2};
  float radius = (d + r1 + r2) * 0.5f;
  return Sph{true, c1 + scale(c2 - c1, (radius - r1) / rmax(d, 1e-12f)), radius};
}
static Sph csg_bound(const Scene& sc, uint32_t ni) {  // csg.rs:120-150
  const CsgNode& n = sc.csg_nodes[ni];
  const float* p = n.p;
  switch (n.op) {
    case MM3E_CSG_PRIM:
      switch (n.prim) {
        case MM3E_PRIM_SPHERE: return Sph{true, V3<float>(), p[0]};
        case MM3E_PRIM_BOX:
        case MM3E_PRIM_ROUND_BOX: return Sph{true, V3<float>(), length(f3<float>(p))};
        case MM3E_PRIM_TORUS: return Sph{true, V3<float>(), p[0] + p[1]};
        case MM3E_PRIM_CYLINDER: return Sph{true, V3<float>(), std::sqrt(p[0] * p[0] + p[1] * p[1])};
        case MM3E_PRIM_CAPSULE: {
          V3<float> a = f3<float>(p), b = f3<float>(p + 3);
          return Sph{true, scale(a + b, 0.5f), length(a - b) * 0.5f + p[6]};
        }
        case MM3E_PRIM_CONE: return Sph{true, V3<float>(0.0f, p[2] * 0.5f, 0.0f), p[2] * 0.5f + std::fmax(p[0], p[1])};
        case MM3E_PRIM_HEX_PRISM: return Sph{true, V3<float>(), std::sqrt(4.0f * p[0] * p[0] / 3.0f + p[1] * p[1])};
        default: return Sph{false, V3<float>(), 0.0f};
      }
    case MM3E_CSG_TRANSFORM: {
      Sph s = csg_bound(sc, n.a);
      if (!s.some) return s;
      Xform<float> x{m3<float>(n.rot), f3<float>(n.pos), n.scale};
      return Sph{true, to_world(x, s.c), s.r * n.scale};
    }
    case MM3E_CSG_UNION:
    case MM3E_CSG_SMOOTH_UNION: {
      Sph a = csg_bound(sc, n.a);
      if (!a.some) return a;
      Sph b = csg_bound(sc, n.b);
      if (!b.some) return b;
      Sph e = enclose(a.c, a.r, b.c, b.r);
      if (n.op == MM3E_CSG_SMOOTH_UNION) e.r += n.p[0] * 0.25f;
      return e;
    }
    case MM3E_CSG_INTERSECT: {
      Sph a = csg_bound(sc, n.a), b = csg_bound(sc, n.b);
      if (a.some && b.some) return a.r <= b.r ? a : b;
      return a.some ? a : b;
    }
    case MM3E_CSG_SUBTRACT: return csg_bound(sc, n.a);
    case MM3E_CSG_OFFSET: {
      Sph s = csg_bound(sc, n.a);
      if (s.some) s.r = std::fmax(s.r + n.p[0], 0.0f);
      return s;
    }
    case MM3E_CSG_SHELL: {
      Sph s = csg_bound(sc, n.a);
      if (s.some) s.r = s.r + n.p[0];
      return s;
    }
  }
  return Sph{false, V3<float>(), 0.0f};
}

// ---------------------------------------------------------------------------------------
// Object (lib.rs:380-496)
// ---------------------------------------------------------------------------------------
template <class R> static V3<R> local_point(const Obj& o, V3<R> p) {  // lib.rs:400-419
  Xform<R> x{m3<R>(o.rot), f3<R>(o.pos), R(o.scale)};
  V3<R> l = to_local(x, p);
  if (o.mirror[0] || o.mirror[1] || o.mirror[2]) l = op_mirror(l, o.mirror[0], o.mirror[1], o.mirror[2]);
  if (!(o.elongate[0] == 0.0f && o.elongate[1] == 0.0f && o.elongate[2] == 0.0f)) l = op_elongate(l, f3<R>(o.elongate));
  if (!(o.repeat[0] == 0.0f && o.repeat[1] == 0.0f && o.repeat[2] == 0.0f)) l = op_repeat(l, f3<R>(o.repeat));
  if (o.twist != 0.0f) l = op_twist(l, R(o.twist));
  if (o.bend != 0.0f) l = op_bend(l, R(o.bend));
  return l;
}
template <class R> static Field<R> object_field(const Scene& sc, const Obj& o, V3<R> p) {  // lib.rs:380-397
  if (g_obj_exact) (*g_obj_exact)++;
  V3<R> l = local_point(o, p);
  R d;
  if (o.prim == MM3E_PRIM_CSG) {
    d = csg_distance<R>(sc, sc.csg_roots.at(o.pid), l);
  } else if (o.prim == MM3E_PRIM_VOLUME) {
    d = sc.volumes.empty() ? R(INFINITY) : vol_sample(sc.volumes[o.pid % sc.volumes.size()], l);
  } else {
    d = prim_distance<R>(o.prim, o.p, l);
  }
  d = d * R(o.scale);
  if (o.round != 0.0f) d = op_round(d, R(o.round));
  if (o.onion != 0.0f) d = op_onion(d, R(o.onion));
  return Field<R>{d, o.mat};
}

static Sph local_bound(const Scene& sc, const Obj& o) {  // lib.rs:264-292
  const float* p = o.p;
  switch (o.prim) {
    case MM3E_PRIM_SPHERE: return Sph{true, V3<float>(), p[0]};
    case MM3E_PRIM_BOX:
    case MM3E_PRIM_ROUND_BOX: return Sph{true, V3<float>(), length(f3<float>(p))};
    case MM3E_PRIM_TORUS: return Sph{true, V3<float>(), p[0] + p[1]};
    case MM3E_PRIM_CYLINDER: return Sph{true, V3<float>(), std::sqrt(p[0] * p[0] + p[1] * p[1])};
    case MM3E_PRIM_CAPSULE: {
      V3<float> a = f3<float>(p), b = f3<float>(p + 3);
      return Sph{true, scale(a + b, 0.5f), length(scale(a - b, 0.5f)) + p[6]};
    }
    case MM3E_PRIM_CONE:
      return Sph{true, V3<float>(0.0f, p[2] * 0.5f, 0.0f), std::fmax(std::fmax(p[2] * 0.5f + std::fmax(p[0], p[1]), p[0]), p[1])};
    case MM3E_PRIM_ELLIPSOID: return Sph{true, V3<float>(), std::fmax(std::fmax(p[0], p[1]), p[2])};
    case MM3E_PRIM_OCTAHEDRON: return Sph{true, V3<float>(), p[0]};
    case MM3E_PRIM_HEX_PRISM: return Sph{true, V3<float>(), std::sqrt(p[0] * p[0] + p[1] * p[1])};
    case MM3E_PRIM_PLANE: return Sph{false, V3<float>(), 0.0f};
    case MM3E_PRIM_CSG: return csg_bound(sc, sc.csg_roots.at(o.pid));
    case MM3E_PRIM_VOLUME: {
      if (sc.volumes.empty()) return Sph{false, V3<float>(), 0.0f};
      V3<float> lo, hi;
      vol_bounds(sc.volumes[o.pid % sc.volumes.size()], lo, hi);
      return Sph{true, scale(lo + hi, 0.5f), length(scale(hi - lo, 0.5f))};
    }
  }
  return Sph{false, V3<float>(), 0.0f};
}
static Sph world_bound(const Scene& sc, const Obj& o) {  // lib.rs:470-496
  Sph s = local_bound(sc, o);
  if (!s.some) return s;
  bool rep = !(o.repeat[0] == 0.0f && o.repeat[1] == 0.0f && o.repeat[2] == 0.0f);
  if (rep || o.bend != 0.0f) return Sph{false, V3<float>(), 0.0f};
  V3<float> center = s.c;
  float radius = s.r;
  if (o.mirror[0] || o.mirror[1] || o.mirror[2]) {
    radius += length(center);
    center = V3<float>();
  }
  if (o.twist != 0.0f) {
    radius += std::sqrt(center.x * center.x + center.z * center.z);
    center.x = 0.0f;
    center.z = 0.0f;
  }
  radius += length(f3<float>(o.elongate));
  Xform<float> x{m3<float>(o.rot), f3<float>(o.pos), o.scale};
  V3<float> wc = x.pos + mul_vec(x.rot, scale(center, o.scale));
  return Sph{true, wc, radius * o.scale + std::fmax(o.round, 0.0f) + std::fabs(o.onion)};
}

static bool scene_dual_safe(const Scene& sc) {  // lib.rs:426-430, 811-813
  for (const Obj& o : sc.objects) {
    if (o.twist != 0.0f || o.bend != 0.0f) return false;
    if (o.prim == MM3E_PRIM_VOLUME || o.prim == MM3E_PRIM_CSG || o.prim == MM3E_PRIM_SURFACE) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------------------
// accel.rs — WorldPlan: BVH over maximal runs of plain unions (the reference `world()`)
// ---------------------------------------------------------------------------------------
struct BvhNode {
  V3<float> center;
  float radius;
  uint32_t left, right;
  int32_t leaf;
};
struct Segment {
  bool one;
  uint32_t index;                 // one
  std::vector<BvhNode> nodes;     // run
  uint32_t root = 0;
  bool has_tree = false;
  std::vector<uint32_t> unbounded;
};
struct Plan {
  std::vector<Segment> segments;
  std::vector<Sph> bounds;
  float slack;
};

struct Item {
  uint32_t i;
  V3<float> c;
  float r;
};
static uint32_t bvh_build(Item* items, size_t n, std::vector<BvhNode>& nodes) {  // accel.rs:39-77
  if (n == 1) {
    nodes.push_back(BvhNode{items[0].c, items[0].r, 0, 0, (int32_t)items[0].i});
    return (uint32_t)(nodes.size() - 1);
  }
  V3<float> lo = splat(INFINITY), hi = splat(-INFINITY);
  for (size_t k = 0; k < n; k++) {
    lo = vmin(lo, items[k].c);
    hi = vmax(hi, items[k].c);
  }
  V3<float> ext = hi - lo;
  int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
  auto coord = [axis](const V3<float>& c) { return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); };
  std::sort(items, items + n, [&](const Item& a, const Item& b) {
    float ca = coord(a.c), cb = coord(b.c);
    if (ca < cb) return true;
    if (ca > cb) return false;
    return a.i < b.i;
  });
  size_t mid = n / 2;
  uint32_t left = bvh_build(items, mid, nodes);
  uint32_t right = bvh_build(items + mid, n - mid, nodes);
  Sph e = enclose(nodes[left].center, nodes[left].radius, nodes[right].center, nodes[right].radius);
  nodes.push_back(BvhNode{e.c, e.r, left, right, -1});
  return (uint32_t)(nodes.size() - 1);
}

static Plan plan_build(const Scene& sc) {  // accel.rs:134-185
  Plan pl;
  for (const Obj& o : sc.objects) pl.bounds.push_back(world_bound(sc, o));
  float mk = 0.0f;
  for (const Obj& o : sc.objects) mk = std::fmax(mk, o.combine == MM3E_COMBINE_SMOOTH ? o.k : 0.0f);
  pl.slack = mk + 0.1f;
  std::vector<uint32_t> run;
  auto close_run = [&]() {
    if (run.empty()) return;
    Segment s;
    s.one = false;
    std::vector<Item> items;
    for (uint32_t i : run) {
      if (pl.bounds[i].some) items.push_back(Item{i, pl.bounds[i].c, pl.bounds[i].r});
      else s.unbounded.push_back(i);
    }
    if (!items.empty()) {
      s.root = bvh_build(items.data(), items.size(), s.nodes);
      s.has_tree = true;
    }
    pl.segments.push_back(std::move(s));
    run.clear();
  };
  for (uint32_t i = 0; i < sc.objects.size(); i++) {
    if (sc.objects[i].combine == MM3E_COMBINE_UNION) {
      run.push_back(i);
    } else {
      close_run();
      Segment s;
      s.one = true;
      s.index = i;
      pl.segments.push_back(std::move(s));
    }
  }
  close_run();
  return pl;
}

template <class R> struct World {
  const Scene* sc;
  Plan plan;

  void note_eval() const {
    if (g_evals) g_evals[g_phase]++;
  }
  Field<R> contrib(uint32_t i, V3<R> p) const {  // accel.rs:187-207
    const Sph& b = plan.bounds[i];
    if (b.some) {
      R lower = length(p - f3c(b.c)) - R(b.r);
      if (lower > R(plan.slack)) {
        if (g_obj_bound) (*g_obj_bound)++;
        return Field<R>{lower, sc->objects[i].mat};
      }
      return object_field<R>(*sc, sc->objects[i], p);
    }
    return object_field<R>(*sc, sc->objects[i], p);
  }
  static V3<R> f3c(const V3<float>& c) { return V3<R>(R(c.x), R(c.y), R(c.z)); }

  struct Best {
    R d;
    int64_t idx;
    uint32_t mat;
  };
  void query(const Segment& s, V3<R> p, Best& best) const {  // accel.rs:79-119
    uint32_t stack[64];
    int sp = 0;
    stack[sp++] = s.root;
    while (sp > 0) {
      const BvhNode& node = s.nodes[stack[--sp]];
      R lower = length(p - f3c(node.center)) - R(node.radius);
      if (lower > best.d) continue;
      if (node.leaf >= 0) {
        uint32_t i = (uint32_t)node.leaf;
        Field<R> c;
        if (lower > R(plan.slack)) {
          if (g_obj_bound) (*g_obj_bound)++;
          c = Field<R>{lower, sc->objects[i].mat};
        } else {
          c = object_field<R>(*sc, sc->objects[i], p);
        }
        if (c.dist < best.d || (val(c.dist) == val(best.d) && (int64_t)i < best.idx)) best = Best{c.dist, (int64_t)i, c.mat};
      } else {
        uint32_t l = node.left, r = node.right;
        R ld = length_sq(p - f3c(s.nodes[l].center));
        R rd = length_sq(p - f3c(s.nodes[r].center));
        uint32_t nearc = ld <= rd ? l : r, farc = ld <= rd ? r : l;
        stack[sp] = farc;
        stack[sp + 1] = nearc;
        sp += 2;
      }
    }
  }
  Field<R> operator()(V3<R> p) const {  // accel.rs:209-252
    note_eval();
    Field<R> acc = field_far<R>();
    bool seeded = false;
    for (const Segment& seg : plan.segments) {
      if (seg.one) {
        Field<R> f = contrib(seg.index, p);
        if (!seeded) {
          acc = f;
          seeded = true;
          continue;
        }
        const Obj& o = sc->objects[seg.index];
        if (o.combine == MM3E_COMBINE_UNION) acc = f_union(acc, f);
        else if (o.combine == MM3E_COMBINE_SMOOTH) acc = f_smooth_union(acc, f, R(o.k));
        else acc = f_subtract(acc, f);
      } else {
        Best best = seeded ? Best{acc.dist, -1, acc.mat} : Best{R(INFINITY), INT64_MAX, 0u};
        for (uint32_t i : seg.unbounded) {
          Field<R> f = object_field<R>(*sc, sc->objects[i], p);
          if (f.dist < best.d || (val(f.dist) == val(best.d) && (int64_t)i < best.idx)) best = Best{f.dist, (int64_t)i, f.mat};
        }
        if (seg.has_tree) query(seg, p, best);
        acc = Field<R>{best.d, best.mat};
        seeded = true;
      }
    }
    return acc;
  }
  // lib.rs:766-807 — reference linear fold (validation only)
  Field<R> linear(V3<R> p) const {
    Field<R> acc = field_far<R>();
    bool first = true;
    for (uint32_t i = 0; i < sc->objects.size(); i++) {
      Field<R> f = contrib(i, p);
      if (first) {
        first = false;
        acc = f;
        continue;
      }
      const Obj& o = sc->objects[i];
      if (o.combine == MM3E_COMBINE_UNION) acc = f_union(acc, f);
      else if (o.combine == MM3E_COMBINE_SMOOTH) acc = f_smooth_union(acc, f, R(o.k));
      else acc = f_subtract(acc, f);
    }
    return acc;
  }
};

// lib.rs:435-463 — dual-number twin of Object::field
template <class R> static Dual<R> object_field_dual(const Obj& o, Dual<R> wx, Dual<R> wy, Dual<R> wz) {
  R inv = rabs(R(o.scale)) > R(1e-12f) ? R(1.0f) / R(o.scale) : R(0.0f);
  Dual<R> tx = dadds(wx, -R(o.pos[0]));
  Dual<R> ty = dadds(wy, -R(o.pos[1]));
  Dual<R> tz = dadds(wz, -R(o.pos[2]));
  M3<R> rt = transpose(m3<R>(o.rot));
  Dual<R> lx = dscale(dadd(dadd(dscale(tx, rt.c[0].x), dscale(ty, rt.c[1].x)), dscale(tz, rt.c[2].x)), inv);
  Dual<R> ly = dscale(dadd(dadd(dscale(tx, rt.c[0].y), dscale(ty, rt.c[1].y)), dscale(tz, rt.c[2].y)), inv);
  Dual<R> lz = dscale(dadd(dadd(dscale(tx, rt.c[0].z), dscale(ty, rt.c[1].z)), dscale(tz, rt.c[2].z)), inv);
  if (o.mirror[0] || o.mirror[1] || o.mirror[2]) {  // dual.rs:242-244
    if (o.mirror[0]) lx = dabs(lx);
    if (o.mirror[1]) ly = dabs(ly);
    if (o.mirror[2]) lz = dabs(lz);
  }
  if (!(o.elongate[0] == 0.0f && o.elongate[1] == 0.0f && o.elongate[2] == 0.0f)) {  // dual.rs:246-251
    Dual<R> cx = dclamp(lx, R(-o.elongate[0]), R(o.elongate[0]));
    Dual<R> cy = dclamp(ly, R(-o.elongate[1]), R(o.elongate[1]));
    Dual<R> cz = dclamp(lz, R(-o.elongate[2]), R(o.elongate[2]));
    lx = dsub(lx, cx);
    ly = dsub(ly, cy);
    lz = dsub(lz, cz);
  }
  if (!(o.repeat[0] == 0.0f && o.repeat[1] == 0.0f && o.repeat[2] == 0.0f)) {  // dual.rs:263-265
    lx = drep_axis(lx, R(o.repeat[0]));
    ly = drep_axis(ly, R(o.repeat[1]));
    lz = drep_axis(lz, R(o.repeat[2]));
  }
  const float* p = o.p;
  Dual<R> d;
  switch (o.prim) {  // lib.rs:244-261
    case MM3E_PRIM_SPHERE: d = dsd_sphere(lx, ly, lz, R(p[0])); break;
    case MM3E_PRIM_BOX: d = dsd_box(lx, ly, lz, f3<R>(p)); break;
    case MM3E_PRIM_ROUND_BOX: d = dsd_round_box(lx, ly, lz, f3<R>(p), R(p[3])); break;
    case MM3E_PRIM_TORUS: d = dsd_torus(lx, ly, lz, R(p[0]), R(p[1])); break;
    case MM3E_PRIM_CYLINDER: d = dsd_cylinder(lx, ly, lz, R(p[0]), R(p[1])); break;
    case MM3E_PRIM_CAPSULE: d = dsd_capsule(lx, ly, lz, f3<R>(p), f3<R>(p + 3), R(p[6])); break;
    case MM3E_PRIM_CONE: d = dsd_round_cone(lx, ly, lz, R(p[0]), R(p[1]), R(p[2])); break;
    case MM3E_PRIM_ELLIPSOID: d = dsd_ellipsoid(lx, ly, lz, f3<R>(p)); break;
    case MM3E_PRIM_OCTAHEDRON: d = dsd_octahedron(lx, ly, lz, R(p[0])); break;
    case MM3E_PRIM_HEX_PRISM: d = dsd_hex_prism(lx, ly, lz, R(p[0]), R(p[1])); break;
    case MM3E_PRIM_PLANE: d = dsd_plane(lx, ly, lz, f3<R>(p), R(p[3])); break;
    default: d = dcon(R(INFINITY)); break;
  }
  d = dscale(d, R(o.scale));
  if (o.round != 0.0f) d = dadds(d, -R(o.round));
  if (o.onion != 0.0f) d = dadds(dabs(d), -R(o.onion));
  return d;
}

template <class R> static V3<R> normal_dual(const Scene& sc, V3<R> p) {  // lib.rs:820-835
  Dual<R> x{p.x, V3<R>(R(1.0f), R(0.0f), R(0.0f))};
  Dual<R> y{p.y, V3<R>(R(0.0f), R(1.0f), R(0.0f))};
  Dual<R> z{p.z, V3<R>(R(0.0f), R(0.0f), R(1.0f))};
  bool have = false;
  Dual<R> acc;
  for (const Obj& o : sc.objects) {
    Dual<R> f = object_field_dual<R>(o, x, y, z);
    if (!have) {
      acc = f;
      have = true;
      continue;
    }
    if (o.combine == MM3E_COMBINE_UNION) acc = dunion(acc, f);
    else if (o.combine == MM3E_COMBINE_SMOOTH) acc = dsmooth_union(acc, f, R(o.k));
    else acc = dsubtract(acc, f);
  }
  return have ? normalize(acc.g) : V3<R>();
}

// ---------------------------------------------------------------------------------------
// mm3e-kit/src/march.rs
// ---------------------------------------------------------------------------------------
template <class R> struct Ray {
  V3<R> o, d;
  V3<R> at(R t) const { return o + scale(d, t); }
};
template <class R> struct Hit {
  bool hit;
  R t;
  V3<R> pos, normal;
  uint32_t mat, steps;
};

template <class R> struct Marcher {
  const Scene* sc;
  // march.rs:212-223 tetrahedron normal
  template <class FN> V3<R> normal(const FN& field, V3<R> p) const {
    R h(sc->normal_h);
    V3<R> k0(R(1.0f), R(-1.0f), R(-1.0f)), k1(R(-1.0f), R(-1.0f), R(1.0f)), k2(R(-1.0f), R(1.0f), R(-1.0f)),
        k3(R(1.0f), R(1.0f), R(1.0f));
    V3<R> g = scale(k0, field(p + scale(k0, h)).dist) + scale(k1, field(p + scale(k1, h)).dist) +
              scale(k2, field(p + scale(k2, h)).dist) + scale(k3, field(p + scale(k3, h)).dist);
    return normalize(g);
  }
  // march.rs:120-209 enhanced sphere tracing with the overlap guard
  template <class FN, class NF> Hit<R> march_with(const FN& field, const NF& normal_fn, const Ray<R>& ray) const {
    enum { OMEGA, LEAP, SECANT } boost = OMEGA;
    bool ss1 = sc->step_scale >= 1.0f;
    R omega = ss1 ? R(1.4f) : R(sc->step_scale);
    R subitize = ss1 ? R(sc->subitize) : R(0.0f);
    bool secant_on = sc->secant && ss1;
    R t(0.0f), prev_radius(0.0f), step_len(0.0f), t_prev(0.0f), d_prev(INFINITY);
    const R eps0(sc->eps), lod(sc->lod_footprint), max_dist(sc->max_dist);
    for (uint32_t i = 0; i < sc->max_steps; i++) {
      V3<R> p = ray.at(t);
      Field<R> f = field(p);
      R radius = rabs(f.dist);
      R eps = eps0 * (R(1.0f) + t * R(0.5f)) + lod * t;
      if ((step_len > R(0.0f) && f.dist < R(0.0f)) || step_len > radius + prev_radius) {
        R gap_lo = t_prev + prev_radius;
        R gap_hi = t - radius;
        R mid = R(0.5f) * (gap_lo + gap_hi);
        bool covered = f.dist > R(0.0f) && field(ray.at(mid)).dist >= R(0.5f) * (gap_hi - gap_lo);
        if (!covered) {
          t = gap_lo;
          if (boost == OMEGA) omega = rmin(omega, R(1.0f));
          else if (boost == LEAP) subitize = R(0.0f);
          else secant_on = false;
          boost = OMEGA;
          prev_radius = R(0.0f);
          step_len = R(0.0f);
          d_prev = R(INFINITY);
          continue;
        }
      }
      if (f.dist < eps) {
        int saved = g_phase;
        g_phase = PH_NORMAL;
        V3<R> n = normal_fn(p);
        g_phase = saved;
        return Hit<R>{true, t, p, n, f.mat, i};
      }
      step_len = f.dist * omega;
      boost = OMEGA;
      if (subitize > R(0.0f) && f.dist > eps * R(6.0f)) {
        step_len = step_len * (R(1.0f) + subitize);
        boost = LEAP;
      }
      if (secant_on && f.dist < R(0.08f) && risfinite(d_prev)) {
        R dt = t - t_prev;
        R dd = f.dist - d_prev;
        if (dt > R(1e-6f) && dd < R(-1e-6f)) {
          step_len = rclamp(-f.dist * dt / dd, f.dist, f.dist * R(4.0f));
          boost = SECANT;
        }
      }
      prev_radius = radius;
      d_prev = f.dist;
      t_prev = t;
      t = t + step_len;
      if (t > max_dist) break;
    }
    return Hit<R>{false, t, ray.at(t), V3<R>(), 0u, sc->max_steps};
  }
  // march.rs:227-259
  template <class FN> R soft_shadow(const FN& field, V3<R> origin, V3<R> dir, R max_t, R k) const {
    R res(1.0f);
    R hb = origin.x * R(127.1f) + origin.y * R(311.7f) + origin.z * R(74.7f);
    R sv = rsin(hb) * R(43758.547f);
    R jitter = rabs(sv - rtrunc(sv));  // f32::fract = x - trunc(x)
    R t = R(0.02f) + jitter * R(0.16f);
    for (uint32_t s = 0; s < sc->shadow_steps; s++) {
      R h = field(origin + scale(dir, t)).dist;
      if (h < R(0.0008f)) return R(0.0f);
      res = rmin(res, k * h / t);
      t = t + rclamp(h, R(0.06f), R(0.5f) + R(0.7f) * t);
      if (t > max_t) break;
    }
    return rclamp(res, R(0.0f), R(1.0f));
  }
  // march.rs:264-280
  template <class FN> R ambient_occlusion(const FN& field, V3<R> p, V3<R> n) const {
    uint32_t ns = sc->ao_samples;
    if (ns == 0) return R(1.0f);
    R base = rmax(R(sc->normal_h), R(0.01f));
    R span = R((float)(std::max(ns, 2u) - 1));
    R occ(0.0f), sca(1.0f);
    for (uint32_t i = 0; i < ns; i++) {
      R hr = base + R(0.12f) * R((float)i) / span;
      R d = field(p + scale(n, hr)).dist;
      occ = occ + (hr - d) * sca;
      sca = sca * R(0.92f);
    }
    return rclamp(R(1.0f) - R(2.6f) * occ, R(0.0f), R(1.0f));
  }
};

// ---------------------------------------------------------------------------------------
// mm3e-kit/src/shade.rs
// ---------------------------------------------------------------------------------------
template <class R> inline R lambert(V3<R> n, V3<R> l) { return rmax(dot(n, l), R(0.0f)); }
template <class R> inline R fresnel_schlick(R c, R f0) {  // shade.rs:14-17
  R m = rclamp(R(1.0f) - c, R(0.0f), R(1.0f));
  return f0 + (R(1.0f) - f0) * rpowi(m, 5);
}
template <class R> inline V3<R> fresnel_schlick_vec(R c, V3<R> f0) {  // shade.rs:20-23
  R m = rpowi(rclamp(R(1.0f) - c, R(0.0f), R(1.0f)), 5);
  return f0 + scale(splat(R(1.0f)) - f0, m);
}
template <class R> inline R luminance(V3<R> c) { return c.x * R(0.2126f) + c.y * R(0.7152f) + c.z * R(0.0722f); }
static const float kPi = 3.14159265358979323846f;
template <class R> inline R d_ggx(R noh, R a2) {  // shade.rs:31-36
  R d = noh * noh * (a2 - R(1.0f)) + R(1.0f);
  return a2 / rmax(R(kPi) * d * d, R(1e-12f));
}
template <class R> inline R v_smith(R nol, R nov, R a2) {  // shade.rs:39-43
  R gv = nol * rsqrt_(nov * nov * (R(1.0f) - a2) + a2);
  R gl = nov * rsqrt_(nol * nol * (R(1.0f) - a2) + a2);
  return R(0.5f) / rmax(gv + gl, R(1e-5f));
}
template <class R>
inline V3<R> brdf(V3<R> n, V3<R> l, V3<R> view, V3<R> albedo, R metallic, R roughness, R reflectance) {  // shade.rs:51-72
  R nol = rmax(dot(n, l), R(0.0f));
  if (nol <= R(0.0f)) return V3<R>();
  R nov = rmax(dot(n, view), R(1e-4f));
  V3<R> h = normalize(l + view);
  R noh = rmax(dot(n, h), R(0.0f));
  R voh = rmax(dot(view, h), R(0.0f));
  R a = rmax(roughness * roughness, R(1e-3f));
  R a2 = a * a;
  V3<R> f0 = mix(splat(R(0.08f) * rclamp(reflectance, R(0.0f), R(1.0f))), albedo, metallic);
  V3<R> f = fresnel_schlick_vec(voh, f0);
  V3<R> spec = scale(f, d_ggx(noh, a2) * v_smith(nol, nov, a2));
  V3<R> kd = scale(splat(R(1.0f)) - f, R(1.0f) - metallic);
  V3<R> diff = scale(cmul(kd, albedo), R(1.0f / kPi));
  return scale(diff + spec, nol);
}
template <class R> inline V3<R> sky(V3<R> dir, V3<R> sun_dir) {  // shade.rs:76-84
  R t = rclamp(R(0.5f) * (dir.y + R(1.0f)), R(0.0f), R(1.0f));
  V3<R> base = mix(V3<R>(R(0.78f), R(0.86f), R(0.96f)), V3<R>(R(0.20f), R(0.38f), R(0.72f)), t);
  R sun = rmax(dot(dir, sun_dir), R(0.0f));
  R glow = rpow(sun, R(8.0f)) * R(0.3f) + rpow(sun, R(220.0f)) * R(6.0f);
  return base + splat(glow);
}
template <class R> inline V3<R> sky_diffuse(V3<R> dir, V3<R> sun_dir) {  // shade.rs:88-95
  R t = rclamp(R(0.5f) * (dir.y + R(1.0f)), R(0.0f), R(1.0f));
  V3<R> base = mix(V3<R>(R(0.78f), R(0.86f), R(0.96f)), V3<R>(R(0.20f), R(0.38f), R(0.72f)), t);
  R soft = rpow(rmax(dot(dir, sun_dir), R(0.0f)), R(8.0f)) * R(0.3f);
  return base + splat(soft);
}
template <class R> inline R aces1(R v) {  // shade.rs:98-104
  v = rmax(v, R(0.0f));
  return rclamp((v * (R(2.51f) * v + R(0.03f))) / (v * (R(2.43f) * v + R(0.59f)) + R(0.14f)), R(0.0f), R(1.0f));
}
template <class R> inline V3<R> aces(V3<R> x) { return V3<R>(aces1(x.x), aces1(x.y), aces1(x.z)); }
template <class R> inline V3<R> gamma(V3<R> x) {  // shade.rs:107-110
  const R inv(1.0f / 2.2f);
  return V3<R>(rpow(rmax(x.x, R(0.0f)), inv), rpow(rmax(x.y, R(0.0f)), inv), rpow(rmax(x.z, R(0.0f)), inv));
}
template <class R> inline V3<R> apply_fog(V3<R> c, V3<R> fog, R t, R density) {  // shade.rs:113-116
  R f = R(1.0f) - rexp(-t * density);
  return mix(c, fog, rclamp(f, R(0.0f), R(1.0f)));
}

// mm3e-kit/src/lib.rs:63-78 — FNV-1a hash of a lattice cell
static uint64_t fnv1a(const uint8_t* b, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; i++) {
    h ^= b[i];
    h *= 0x00000100000001b3ull;
  }
  return h;
}
static float hash_cell(int32_t ix, int32_t iy) {
  uint8_t buf[8];
  std::memcpy(buf, &ix, 4);  // x86-64 is little-endian, as `to_le_bytes`
  std::memcpy(buf + 4, &iy, 4);
  return (float)(fnv1a(buf, 8) >> 40) / (float)(1ull << 24);
}

// lib.rs:1365-1375
template <class R> static V3<R> surface_albedo(const Mat& m, V3<R> p) {
  V3<R> alb = f3<R>(m.albedo);
  if (!m.checker) return alb;
  int32_t ix = to_i32(val(rfloor(p.x)));
  int32_t iz = to_i32(val(rfloor(p.z)));
  int32_t s = (int32_t)((uint32_t)ix + (uint32_t)iz);  // release-mode wrapping add
  bool parity = ((s % 2) + 2) % 2 == 0;                 // rem_euclid(2)
  V3<R> base = parity ? splat(R(0.92f)) : splat(R(0.20f));
  R tint = R(hash_cell(ix, iz)) * R(0.06f) - R(0.03f);
  return max_scalar(cmul(base + splat(tint), alb), R(0.0f));
}

// ---------------------------------------------------------------------------------------
// Light (lib.rs:528-556)
// ---------------------------------------------------------------------------------------
template <class R> static void light_toward(const Lgt& l, V3<R> p, V3<R>& dir, R& dist) {
  if (l.directional) {
    dir = f3<R>(l.vec);
    dist = R(INFINITY);
  } else {
    V3<R> d = f3<R>(l.vec) - p;
    R len = length(d);
    dir = scale(d, R(1.0f) / rmax(len, R(1e-6f)));
    dist = len;
  }
}
template <class R> static R light_atten(const Lgt& l, R dist) {
  if (l.directional) return R(1.0f);
  return R(1.0f) / rpowi(rmax(dist, R(0.05f)), 2);
}
template <class R> static R light_shadow_k(const Lgt& l, R dist) {
  if (l.radius <= 0.0f) return R(24.0f);
  if (l.directional) return rclamp(R(1.0f) / R(l.radius), R(2.0f), R(64.0f));
  return rclamp(dist / R(l.radius), R(2.0f), R(64.0f));
}

// ---------------------------------------------------------------------------------------
// gi.rs — ambient-cube probe volume
// ---------------------------------------------------------------------------------------
template <class R> static V3<R> gi_sample(const Gi& g, V3<R> pos, V3<R> n) {  // gi.rs:138-187
  const uint64_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  auto axis = [](R p, R lo, R ext, uint64_t n) -> R {
    if (rabs(ext) > R(1e-6f)) return (p - lo) / ext * R((float)(std::max<uint64_t>(n, 2) - 1));
    return R(0.0f);
  };
  V3<R> gg(axis(pos.x, R(g.min[0]), R(g.size[0]), nx), axis(pos.y, R(g.min[1]), R(g.size[1]), ny),
           axis(pos.z, R(g.min[2]), R(g.size[2]), nz));
  auto clampi = [](R v, uint64_t hi) -> uint64_t {
    return std::min(to_usize(val(rmax(rfloor(v), R(0.0f)))), hi == 0 ? 0 : hi - 1);
  };
  uint64_t i0 = clampi(gg.x, nx), j0 = clampi(gg.y, ny), k0 = clampi(gg.z, nz);
  uint64_t i1 = std::min(i0 + 1, nx - 1), j1 = std::min(j0 + 1, ny - 1), k1 = std::min(k0 + 1, nz - 1);
  R fx = rclamp(gg.x - R((float)i0), R(0.0f), R(1.0f));
  R fy = rclamp(gg.y - R((float)j0), R(0.0f), R(1.0f));
  R fz = rclamp(gg.z - R((float)k0), R(0.0f), R(1.0f));
  auto cube = [&](uint64_t i, uint64_t j, uint64_t k, int f) {
    const float* c = &g.cubes[(((k * ny + j) * nx + i) * 6 + f) * 3];
    return f3<R>(c);
  };
  V3<R> face[6];
  for (int f = 0; f < 6; f++) {
    V3<R> x00 = mix(cube(i0, j0, k0, f), cube(i1, j0, k0, f), fx);
    V3<R> x10 = mix(cube(i0, j1, k0, f), cube(i1, j1, k0, f), fx);
    V3<R> x01 = mix(cube(i0, j0, k1, f), cube(i1, j0, k1, f), fx);
    V3<R> x11 = mix(cube(i0, j1, k1, f), cube(i1, j1, k1, f), fx);
    V3<R> y0 = mix(x00, x10, fy);
    V3<R> y1 = mix(x01, x11, fy);
    face[f] = mix(y0, y1, fz);
  }
  V3<R> nn(n.x * n.x, n.y * n.y, n.z * n.z);
  V3<R> fxf = n.x >= R(0.0f) ? face[0] : face[1];
  V3<R> fyf = n.y >= R(0.0f) ? face[2] : face[3];
  V3<R> fzf = n.z >= R(0.0f) ? face[4] : face[5];
  return scale(fxf, nn.x) + scale(fyf, nn.y) + scale(fzf, nn.z);
}

static const float kAxes[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

struct DirW {
  V3<float> d;
  float w;
};
static std::vector<DirW> axis_dirs(V3<float> axis, uint32_t samples) {  // gi.rs:38-50
  V3<float> up = std::fabs(axis.y) > 0.9f ? V3<float>(1.0f, 0.0f, 0.0f) : V3<float>(0.0f, 1.0f, 0.0f);
  V3<float> t1 = normalize(cross(axis, up));
  V3<float> t2 = normalize(cross(axis, t1));
  std::vector<DirW> dirs{{axis, 1.0f}};
  if (samples > 1) {
    V3<float> ts[4] = {t1, t2, scale(t1, -1.0f), scale(t2, -1.0f)};
    for (auto& t : ts) {
      V3<float> d = normalize(scale(axis, 0.72f) + scale(t, 0.72f));
      dirs.push_back({d, std::fmax(dot(d, axis), 0.0f)});
    }
  }
  return dirs;
}

// ---------------------------------------------------------------------------------------
// Renderer (lib.rs:686-711, 1223-1398)
// ---------------------------------------------------------------------------------------
struct Camera {
  float eye[3], forward[3], right[3], up[3], fov_scale;
};

template <class R> struct Renderer {
  const Scene* sc;
  World<R> world;
  Marcher<R> mr;
  bool dual_safe;

  explicit Renderer(const Scene* s) : sc(s) {
    world.sc = s;
    world.plan = plan_build(*s);
    mr.sc = s;
    dual_safe = scene_dual_safe(*s);
  }
  Hit<R> march(const Ray<R>& ray) const {
    if (dual_safe) return mr.march_with(world, [this](V3<R> p) { return normal_dual<R>(*sc, p); }, ray);
    return mr.march_with(world, [this](V3<R> p) { return mr.normal(world, p); }, ray);
  }
  V3<R> sun() const { return f3<R>(sc->sun_dir); }

  // lib.rs:686-711
  V3<R> gather(V3<R> origin, V3<R> dir) const {
    Ray<R> ray{origin, dir};
    Hit<R> hit = march(ray);
    if (!hit.hit) return sky_diffuse(dir, sun());
    const Mat& m = sc->materials[hit.mat % sc->materials.size()];
    V3<R> albedo = surface_albedo<R>(m, hit.pos);
    V3<R> n = hit.normal;
    V3<R> c = f3<R>(m.emissive);
    for (const Lgt& l : sc->lights) {
      V3<R> ld;
      R dist;
      light_toward<R>(l, hit.pos, ld, dist);
      R ndl = lambert(n, ld);
      if (ndl > R(0.0f)) {
        V3<R> out = brdf(n, ld, scale(dir, R(-1.0f)), albedo, R(m.metallic), R(m.roughness), R(m.specular));
        c = c + scale(cmul(out, f3<R>(l.color)), light_atten<R>(l, dist));
      }
    }
    return c;
  }

  // lib.rs:1379-1398
  V3<R> ibl_irradiance(V3<R> n) const {
    V3<R> up = rabs(n.y) > R(0.9f) ? V3<R>(R(1.0f), R(0.0f), R(0.0f)) : V3<R>(R(0.0f), R(1.0f), R(0.0f));
    V3<R> t1 = normalize(cross(n, up));
    V3<R> t2 = normalize(cross(n, t1));
    V3<R> dirs[5] = {n, normalize(scale(n, R(0.6f)) + scale(t1, R(0.6f))), normalize(scale(n, R(0.6f)) - scale(t1, R(0.6f))),
                     normalize(scale(n, R(0.6f)) + scale(t2, R(0.6f))), normalize(scale(n, R(0.6f)) - scale(t2, R(0.6f)))};
    V3<R> acc;
    R w(0.0f);
    for (auto& d : dirs) {
      R cw = rmax(dot(d, n), R(0.0f));
      acc = acc + scale(sky_diffuse(d, sun()), cw);
      w = w + cw;
    }
    return cmul(scale(acc, R(1.0f) / rmax(w, R(1e-4f))), f3<R>(sc->sky_ambient));
  }

  // lib.rs:1265-1282
  V3<R> trace(const Ray<R>& ray, uint32_t depth) const {
    int saved = g_phase;
    g_phase = depth == 0 ? PH_MARCH : PH_REFL_MARCH;
    Hit<R> hit = march(ray);
    g_phase = saved;
    if (!hit.hit) return sky(ray.d, sun());
    return shade_hit(hit, ray, depth);
  }

  // lib.rs:1286-1360
  V3<R> shade_hit(const Hit<R>& hit, const Ray<R>& ray, uint32_t depth) const {
    const Mat& m = sc->materials[hit.mat % sc->materials.size()];
    V3<R> albedo = surface_albedo<R>(m, hit.pos);
    V3<R> normal = hit.normal;
    V3<R> view = scale(ray.d, R(-1.0f));
    int saved = g_phase;
    R occ(1.0f);
    if (sc->ao) {
      g_phase = PH_AO;
      occ = mr.ambient_occlusion(world, hit.pos, hit.normal);
      g_phase = saved;
    }
    V3<R> gi = sc->has_gi ? gi_sample<R>(sc->gi, hit.pos, normal) : V3<R>();
    V3<R> ambient_irr = ibl_irradiance(normal) + gi + f3<R>(sc->ambient);
    V3<R> radiance = scale(cmul(albedo, ambient_irr), occ);
    for (const Lgt& l : sc->lights) {
      V3<R> ld;
      R dist;
      light_toward<R>(l, hit.pos, ld, dist);
      if (lambert(normal, ld) <= R(0.0f)) continue;
      R shadow(1.0f);
      if (sc->shadows) {
        R max_t = rmin(dist, R(sc->max_dist));
        R lift = rmax(R(sc->normal_h) * R(2.0f), R(0.01f));
        g_phase = PH_SHADOW;
        shadow = mr.soft_shadow(world, hit.pos + scale(hit.normal, lift), ld, max_t, light_shadow_k<R>(l, dist));
        g_phase = saved;
      }
      if (shadow <= R(0.0f)) continue;
      V3<R> surf = brdf(normal, ld, view, albedo, R(m.metallic), R(m.roughness), R(m.specular));
      radiance = radiance + scale(cmul(surf, f3<R>(l.color)), shadow * light_atten<R>(l, dist));
    }
    if (depth < sc->bounces && m.reflectivity > 0.0f) {
      R c = rmax(dot(normal, view), R(0.0f));
      R fr = fresnel_schlick(c, R(m.reflectivity));
      V3<R> rdir = normalize(reflect(ray.d, normal));
      V3<R> rorigin = hit.pos + scale(hit.normal, rmax(R(sc->normal_h) * R(2.0f), R(0.02f)));
      V3<R> refl = trace(Ray<R>{rorigin, rdir}, depth + 1);
      radiance = scale(radiance, R(1.0f) - fr) + scale(refl, fr);
    }
    radiance = radiance + f3<R>(m.emissive);
    return apply_fog(radiance, f3<R>(sc->fog), hit.t, R(sc->fog_density));
  }

  // camera.rs:59-65
  static Ray<R> cam_ray(const Camera& cam, R sx, R sy, uint32_t w, uint32_t h) {
    R aspect = R((float)w) / R((float)h);
    R ndc_x = (R(2.0f) * sx / R((float)w) - R(1.0f)) * aspect * R(cam.fov_scale);
    R ndc_y = (R(1.0f) - R(2.0f) * sy / R((float)h)) * R(cam.fov_scale);
    V3<R> dir = normalize(f3<R>(cam.forward) + scale(f3<R>(cam.right), ndc_x) + scale(f3<R>(cam.up), ndc_y));
    return Ray<R>{f3<R>(cam.eye), dir};
  }

  // lib.rs:1223-1243
  V3<R> shade_pixel(const Camera& cam, uint32_t x, uint32_t y) const {
    uint32_t n = std::max(sc->aa, 1u);
    R inv = R(1.0f) / R((float)(n * n));
    V3<R> acc;
    for (uint32_t sy = 0; sy < n; sy++) {
      for (uint32_t sx = 0; sx < n; sx++) {
        R ox = (R((float)sx) + R(0.5f)) / R((float)n);
        R oy = (R((float)sy) + R(0.5f)) / R((float)n);
        Ray<R> ray = cam_ray(cam, R((float)x) + ox, R((float)y) + oy, sc->width, sc->height);
        acc = acc + trace(ray, 0);
      }
    }
    return scale(acc, inv);
  }
};

// gi.rs:55-125 + lib.rs:668-682 — bake over the scene (threads over z slices)
static void bake_gi(Scene& sc, uint32_t dx, uint32_t dy, uint32_t dz, uint32_t samples, unsigned threads) {
  std::vector<Sph> spheres;
  for (const Obj& o : sc.objects) {
    Sph s = world_bound(sc, o);
    if (s.some) spheres.push_back(s);
  }
  if (spheres.empty()) return;  // gi.rs:192-203 -> None: no GI
  V3<float> lo = splat(INFINITY), hi = splat(-INFINITY);
  for (const Sph& s : spheres) {
    lo = vmin(lo, s.c - splat(s.r));
    hi = vmax(hi, s.c + splat(s.r));
  }
  lo = lo - splat(1.5f);
  hi = hi + splat(1.5f);
  uint64_t nx = std::max(dx, 1u), ny = std::max(dy, 1u), nz = std::max(dz, 1u);
  V3<float> size = hi - lo;
  V3<float> cell(size.x / (float)(std::max<uint64_t>(nx, 2) - 1), size.y / (float)(std::max<uint64_t>(ny, 2) - 1),
                 size.z / (float)(std::max<uint64_t>(nz, 2) - 1));
  std::vector<std::vector<DirW>> sets;
  for (auto& a : kAxes) sets.push_back(axis_dirs(V3<float>(a[0], a[1], a[2]), samples));
  Gi g;
  g.min[0] = lo.x; g.min[1] = lo.y; g.min[2] = lo.z;
  g.size[0] = size.x; g.size[1] = size.y; g.size[2] = size.z;
  g.dims[0] = (uint32_t)nx; g.dims[1] = (uint32_t)ny; g.dims[2] = (uint32_t)nz;
  g.cubes.assign(nx * ny * nz * 18, 0.0f);
  Renderer<float> rd(&sc);  // gather sees the scene without GI (gi is assigned after the bake)
  std::atomic<uint64_t> next{0};
  auto worker = [&]() {
    int saved = g_phase;
    g_phase = PH_GI;
    for (;;) {
      uint64_t idx = next.fetch_add(1);
      if (idx >= nx * ny * nz) break;
      uint64_t i = idx % nx, j = (idx / nx) % ny, k = idx / (nx * ny);
      V3<float> probe = lo + V3<float>((float)i * cell.x, (float)j * cell.y, (float)k * cell.z);
      for (int a = 0; a < 6; a++) {
        V3<float> acc;
        float wsum = 0.0f;
        for (const DirW& dw : sets[a]) {
          acc = acc + scale(rd.gather(probe, dw.d), dw.w);
          wsum += dw.w;
        }
        V3<float> c = scale(acc, 1.0f / std::fmax(wsum, 1e-4f));
        float* o = &g.cubes[(idx * 6 + a) * 3];
        o[0] = c.x; o[1] = c.y; o[2] = c.z;
      }
    }
    g_phase = saved;
  };
  unsigned nt = std::max(1u, threads);
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < nt; t++) ts.emplace_back(worker);
  for (auto& t : ts) t.join();
  sc.gi = std::move(g);
  sc.has_gi = true;
}

// post.rs:36-115
static std::vector<float> gaussian_kernel(uint32_t radius) {
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
static void blur_axis(const std::vector<V3<float>>& src, uint32_t w, uint32_t h, const std::vector<float>& k, bool horiz,
                      std::vector<V3<float>>& out) {
  out.assign(src.size(), V3<float>());
  int r = (int)(k.size() / 2);
  int wi = (int)w, hi = (int)h;
  for (int y = 0; y < hi; y++) {
    for (int x = 0; x < wi; x++) {
      V3<float> acc;
      for (size_t ki = 0; ki < k.size(); ki++) {
        int off = (int)ki - r;
        int sx = horiz ? std::min(std::max(x + off, 0), wi - 1) : x;
        int sy = horiz ? y : std::min(std::max(y + off, 0), hi - 1);
        acc = acc + scale(src[(size_t)sy * w + sx], k[ki]);
      }
      out[(size_t)y * w + x] = acc;
    }
  }
}
// post.rs:41-46 + bloom_pass (post.rs:61-75): color += blur(bright-pass(color)) * intensity.
static void bloom_add(std::vector<V3<float>>& color, uint32_t w, uint32_t h, float threshold, float intensity,
                      uint32_t radius) {
  size_t n = (size_t)w * h;
  std::vector<V3<float>> bright(n);
  for (size_t i = 0; i < n; i++) {
    float l = luminance(color[i]);
    if (l > threshold) bright[i] = scale(color[i], (l - threshold) / std::fmax(l, 1e-4f));
  }
  std::vector<float> k = gaussian_kernel(radius);
  std::vector<V3<float>> hz, bl;
  if (w > 0 && h > 0) {
    blur_axis(bright, w, h, k, true, hz);
    blur_axis(hz, w, h, k, false, bl);
  } else {
    bl.assign(n, V3<float>());
  }
  for (size_t i = 0; i < n; i++) color[i] = color[i] + scale(bl[i], intensity);
}
static void resolve(const Scene& sc, const float* hdr, uint32_t w, uint32_t h, uint8_t* rgba) {
  size_t n = (size_t)w * h;
  std::vector<V3<float>> color(n);
  for (size_t i = 0; i < n; i++) color[i] = V3<float>(hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]);
  if (sc.bloom && sc.bloom_intensity > 0.0f) bloom_add(color, w, h, sc.bloom_threshold, sc.bloom_intensity, sc.bloom_radius);
  for (size_t i = 0; i < n; i++) {
    V3<float> c = clamp01(gamma(aces(scale(color[i], sc.exposure))));  // Rgba::from_vec3 clamps
    // framebuffer.rs:102-113 with an opaque pixel: to_u8 = (c*255 + 0.5) as u8
    rgba[4 * i + 0] = to_u8(rclamp(c.x, 0.0f, 1.0f) * 255.0f + 0.5f);
    rgba[4 * i + 1] = to_u8(rclamp(c.y, 0.0f, 1.0f) * 255.0f + 0.5f);
    rgba[4 * i + 2] = to_u8(rclamp(c.z, 0.0f, 1.0f) * 255.0f + 0.5f);
    rgba[4 * i + 3] = 255;
  }
}

// lib.rs:908-947 row bands; `body(y0, y1)` fills rows [y0, y1)
static void for_bands(uint32_t h, unsigned threads, const std::function<void(uint32_t, uint32_t)>& body) {
  size_t t = std::max<size_t>(1, threads);
  t = std::min<size_t>(t, std::max<uint32_t>(h, 1));
  size_t band = h / t, extra = h % t;
  std::vector<std::thread> ts;
  for (size_t ti = 0; ti < t; ti++) {
    uint32_t y0 = (uint32_t)(ti * band + std::min(ti, extra));
    uint32_t y1 = y0 + (uint32_t)band + (ti < extra ? 1 : 0);
    ts.emplace_back([=, &body]() { body(y0, y1); });
  }
  for (auto& th : ts) th.join();
}

// ---------------------------------------------------------------------------------------
// IR parsing
// ---------------------------------------------------------------------------------------
static bool parse_ir(const uint8_t* b, size_t len, Scene& sc, std::string& err) {
  if (len < sizeof(mm3e_ir_header)) { err = "IR shorter than header"; return false; }
  mm3e_ir_header h;
  std::memcpy(&h, b, sizeof h);
  if (std::memcmp(h.magic, MM3E_IR_MAGIC, 8) != 0 || h.version != MM3E_IR_VERSION) { err = "bad IR magic/version"; return false; }
  if (h.total_bytes != len) { err = "IR length mismatch"; return false; }
  auto span_ok = [&](uint32_t off, uint64_t bytes) { return (uint64_t)off + bytes <= len; };
  sc.width = h.width; sc.height = h.height; sc.aa = h.aa; sc.bounces = h.bounces;
  sc.shadows = h.shadows != 0; sc.ao = h.ao != 0; sc.mode = h.mode; sc.appearance = h.appearance;
  std::memcpy(sc.sun_dir, h.sun_dir, 12); std::memcpy(sc.ambient, h.ambient, 12);
  std::memcpy(sc.sky_ambient, h.sky_ambient, 12); std::memcpy(sc.fog, h.fog, 12);
  sc.fog_density = h.fog_density;
  sc.max_steps = h.max_steps; sc.max_dist = h.max_dist; sc.eps = h.eps; sc.step_scale = h.step_scale;
  sc.shadow_steps = h.shadow_steps; sc.ao_samples = h.ao_samples; sc.lod_footprint = h.lod_footprint;
  sc.subitize = h.subitize; sc.secant = h.secant != 0; sc.normal_h = h.normal_h;
  sc.exposure = h.exposure; sc.bloom = h.bloom != 0; sc.bloom_threshold = h.bloom_threshold;
  sc.bloom_intensity = h.bloom_intensity; sc.bloom_radius = h.bloom_radius;
  if (!span_ok(h.off_objects, (uint64_t)h.n_objects * sizeof(mm3e_ir_object)) ||
      !span_ok(h.off_materials, (uint64_t)h.n_materials * sizeof(mm3e_ir_material)) ||
      !span_ok(h.off_lights, (uint64_t)h.n_lights * sizeof(mm3e_ir_light)) ||
      !span_ok(h.off_volumes, (uint64_t)h.n_volumes * sizeof(mm3e_ir_volume)) ||
      !span_ok(h.off_volume_data, (uint64_t)h.n_volume_floats * 4) ||
      !span_ok(h.off_csg_nodes, (uint64_t)h.n_csg_nodes * sizeof(mm3e_ir_csg_node)) ||
      !span_ok(h.off_csg_roots, (uint64_t)h.n_csg_exprs * 4)) {
    err = "IR array out of bounds";
    return false;
  }
  for (uint32_t i = 0; i < h.n_objects; i++) {
    mm3e_ir_object o;
    std::memcpy(&o, b + h.off_objects + i * sizeof o, sizeof o);
    Obj x;
    x.prim = o.prim; x.mat = o.mat; x.combine = o.combine; x.k = o.k;
    std::memcpy(x.p, o.p, sizeof x.p);
    std::memcpy(&x.pid, &o.p[0], 4);
    std::memcpy(x.rot, o.rot, 36); std::memcpy(x.pos, o.pos, 12); x.scale = o.scale;
    for (int a = 0; a < 3; a++) x.mirror[a] = (o.mirror >> a) & 1u;
    std::memcpy(x.repeat, o.repeat, 12); x.twist = o.twist; x.bend = o.bend;
    std::memcpy(x.elongate, o.elongate, 12); x.round = o.round; x.onion = o.onion;
    sc.objects.push_back(x);
  }
  for (uint32_t i = 0; i < h.n_materials; i++) {
    mm3e_ir_material m;
    std::memcpy(&m, b + h.off_materials + i * sizeof m, sizeof m);
    Mat x;
    std::memcpy(x.albedo, m.albedo, 12); x.metallic = m.metallic; x.specular = m.specular;
    x.roughness = m.roughness; x.reflectivity = m.reflectivity; std::memcpy(x.emissive, m.emissive, 12);
    x.checker = m.checker != 0;
    sc.materials.push_back(x);
  }
  for (uint32_t i = 0; i < h.n_lights; i++) {
    mm3e_ir_light l;
    std::memcpy(&l, b + h.off_lights + i * sizeof l, sizeof l);
    Lgt x;
    std::memcpy(x.vec, l.vec, 12); x.radius = l.radius; std::memcpy(x.color, l.color, 12);
    x.directional = l.directional != 0;
    sc.lights.push_back(x);
  }
  const float* vdata = (const float*)(b + h.off_volume_data);
  for (uint32_t i = 0; i < h.n_volumes; i++) {
    mm3e_ir_volume v;
    std::memcpy(&v, b + h.off_volumes + i * sizeof v, sizeof v);
    Vol x;
    std::memcpy(x.dims, v.dims, 12); std::memcpy(x.min, v.min, 12); std::memcpy(x.cell, v.cell, 12);
    uint64_t n = (uint64_t)v.dims[0] * v.dims[1] * v.dims[2];
    if ((uint64_t)v.offset + n > h.n_volume_floats) { err = "volume data out of bounds"; return false; }
    x.data.resize(n);
    std::memcpy(x.data.data(), vdata + v.offset, n * 4);
    sc.volumes.push_back(std::move(x));
  }
  for (uint32_t i = 0; i < h.n_csg_nodes; i++) {
    mm3e_ir_csg_node c;
    std::memcpy(&c, b + h.off_csg_nodes + i * sizeof c, sizeof c);
    CsgNode x;
    x.op = c.op; x.prim = c.prim; x.a = c.a; x.b = c.b;
    std::memcpy(x.p, c.p, 32); std::memcpy(x.rot, c.rot, 36); std::memcpy(x.pos, c.pos, 12); x.scale = c.scale;
    sc.csg_nodes.push_back(x);
  }
  for (uint32_t i = 0; i < h.n_csg_exprs; i++) {
    uint32_t r;
    std::memcpy(&r, b + h.off_csg_roots + 4 * i, 4);
    sc.csg_roots.push_back(r);
  }
  sc.gi_mode = h.gi_mode; sc.gi_samples = h.gi_samples;
  std::memcpy(sc.gi_req_dims, h.gi_dims, 12);
  if (h.gi_mode == MM3E_GI_BAKED) {
    uint64_t n = (uint64_t)h.gi_dims[0] * h.gi_dims[1] * h.gi_dims[2] * 18;
    if (!span_ok(h.off_gi_cubes, n * 4) || n == 0) { err = "GI cubes out of bounds"; return false; }
    sc.gi.cubes.resize(n);
    std::memcpy(sc.gi.cubes.data(), b + h.off_gi_cubes, n * 4);
    std::memcpy(sc.gi.min, h.gi_min, 12); std::memcpy(sc.gi.size, h.gi_size, 12); std::memcpy(sc.gi.dims, h.gi_dims, 12);
    sc.has_gi = true;
  }
  if (sc.materials.empty()) { err = "scene needs at least one material"; return false; }  // appearance.rs:251-253
  return true;
}

}  // namespace orc

// =========================================================================================
// C ABI used by tests and the bench's CPU-baseline leg
// =========================================================================================
using namespace orc;

struct orc_scene {
  Scene sc;
};

static void set_err(char* err, size_t n, const std::string& s) {
  if (err && n) {
    std::snprintf(err, n, "%s", s.c_str());
  }
}
static Camera to_cam(const float* c) {
  Camera cam;
  std::memcpy(cam.eye, c, 12); std::memcpy(cam.forward, c + 3, 12); std::memcpy(cam.right, c + 6, 12);
  std::memcpy(cam.up, c + 9, 12); cam.fov_scale = c[12];
  return cam;
}

extern "C" {

int orc_scene_load(const void* ir, size_t len, unsigned bake_threads, orc_scene** out, char* err, size_t errlen) {
  auto* s = new orc_scene();
  std::string e;
  if (!parse_ir((const uint8_t*)ir, len, s->sc, e)) {
    set_err(err, errlen, e);
    delete s;
    return 1;
  }
  if (s->sc.gi_mode == MM3E_GI_BAKE) {
    bake_gi(s->sc, s->sc.gi_req_dims[0], s->sc.gi_req_dims[1], s->sc.gi_req_dims[2], s->sc.gi_samples, bake_threads);
  }
  *out = s;
  return 0;
}
void orc_scene_free(orc_scene* s) { delete s; }

// GI raw layout (gi.rs:133-135); returns probe count (0 = no GI).
uint64_t orc_gi_raw(const orc_scene* s, float* min3, float* size3, uint32_t* dims3, float* cubes) {
  if (!s->sc.has_gi) return 0;
  const Gi& g = s->sc.gi;
  if (min3) std::memcpy(min3, g.min, 12);
  if (size3) std::memcpy(size3, g.size, 12);
  if (dims3) std::memcpy(dims3, g.dims, 12);
  if (cubes) std::memcpy(cubes, g.cubes.data(), g.cubes.size() * 4);
  return (uint64_t)g.dims[0] * g.dims[1] * g.dims[2];
}

// render_linear_unchecked (lib.rs:908-947): HDR w*h*3 floats, row-major.
int orc_render_linear(const orc_scene* s, const float* cam13, unsigned threads, float* hdr) {
  const Scene& sc = s->sc;
  Camera cam = to_cam(cam13);
  Renderer<float> rd(&sc);
  for_bands(sc.height, threads, [&](uint32_t y0, uint32_t y1) {
    for (uint32_t y = y0; y < y1; y++)
      for (uint32_t x = 0; x < sc.width; x++) {
        V3<float> c = rd.shade_pixel(cam, x, y);
        float* o = hdr + 3 * ((size_t)y * sc.width + x);
        o[0] = c.x; o[1] = c.y; o[2] = c.z;
      }
  });
  return 0;
}

// Rows [y0, y1) only (bounded CPU-baseline samples).
int orc_render_linear_rows(const orc_scene* s, const float* cam13, unsigned threads, uint32_t y0, uint32_t y1, float* hdr) {
  const Scene& sc = s->sc;
  Camera cam = to_cam(cam13);
  Renderer<float> rd(&sc);
  uint32_t rows = y1 - y0;
  for_bands(rows, threads, [&](uint32_t a, uint32_t b) {
    for (uint32_t y = a; y < b; y++)
      for (uint32_t x = 0; x < sc.width; x++) {
        V3<float> c = rd.shade_pixel(cam, x, y0 + y);
        float* o = hdr + 3 * ((size_t)y * sc.width + x);
        o[0] = c.x; o[1] = c.y; o[2] = c.z;
      }
  });
  return 0;
}

// post::resolve + Framebuffer::to_rgba8 (post.rs:36-58, framebuffer.rs:102-113)
int orc_resolve(const orc_scene* s, const float* hdr, uint8_t* rgba) {
  resolve(s->sc, hdr, s->sc.width, s->sc.height, rgba);
  return 0;
}

// lib.rs:1401-1407 blue -> green -> yellow -> red step heatmap.
static V3<float> heatmap(float t) {
  t = rclamp(t, 0.0f, 1.0f);
  const float r = rclamp(t * 3.0f - 1.5f, 0.0f, 1.0f);
  const float g = rclamp(1.0f - rabs(t * 3.0f - 1.3f), 0.0f, 1.0f);
  const float b = rclamp(1.0f - t * 2.2f, 0.0f, 1.0f);
  return V3<float>(r, g, b);
}

// lib.rs:1178-1218 aov_pixel: one centre sample per pixel, written straight to display.
static V3<float> aov_pixel(const Renderer<float>& rd, const Camera& cam, uint32_t x, uint32_t y) {
  const Scene& sc = *rd.sc;
  Ray<float> ray = Renderer<float>::cam_ray(cam, (float)x + 0.5f, (float)y + 0.5f, sc.width, sc.height);
  Hit<float> hit = rd.march(ray);
  switch (sc.mode) {
    case 1:  // Normal
      return hit.hit ? scale(hit.normal, 0.5f) + V3<float>(0.5f, 0.5f, 0.5f) : V3<float>();
    case 2: {  // Depth
      const float d = rclamp(hit.t / sc.max_dist, 0.0f, 1.0f);
      return V3<float>(1.0f - d, 1.0f - d, 1.0f - d);
    }
    case 3: {  // Ao
      if (!hit.hit) return V3<float>(1.0f, 1.0f, 1.0f);
      const float a = rd.mr.ambient_occlusion(rd.world, hit.pos, hit.normal);
      return V3<float>(a, a, a);
    }
    case 4:  // Steps
      return heatmap((float)hit.steps / (float)sc.max_steps);
    case 5:  // Albedo (appearance empty: material_for_hit = surface_albedo, appearance.rs:427-431)
      return hit.hit ? surface_albedo<float>(sc.materials[hit.mat % sc.materials.size()], hit.pos) : V3<float>();
    default:
      return V3<float>();
  }
}

int orc_render_rgba(const orc_scene* s, const float* cam13, unsigned threads, uint8_t* rgba) {
  const Scene& sc = s->sc;
  if (sc.mode != 0) {  // lib.rs:873-878: debug AOV modes skip the HDR buffer and post pass
    Camera cam = to_cam(cam13);
    Renderer<float> rd(&sc);
    for_bands(sc.height, threads, [&](uint32_t y0, uint32_t y1) {
      for (uint32_t y = y0; y < y1; y++)
        for (uint32_t x = 0; x < sc.width; x++) {
          const V3<float> c = clamp01(aov_pixel(rd, cam, x, y));  // Rgba::from_vec3
          uint8_t* o = rgba + 4 * ((size_t)y * sc.width + x);
          o[0] = to_u8(rclamp(c.x, 0.0f, 1.0f) * 255.0f + 0.5f);
          o[1] = to_u8(rclamp(c.y, 0.0f, 1.0f) * 255.0f + 0.5f);
          o[2] = to_u8(rclamp(c.z, 0.0f, 1.0f) * 255.0f + 0.5f);
          o[3] = 255;
        }
    });
    return 0;
  }
  std::vector<float> hdr((size_t)sc.width * sc.height * 3);
  orc_render_linear(s, cam13, threads, hdr.data());
  return orc_resolve(s, hdr.data(), rgba);
}

// Centre-pixel pinhole primary-hit AOVs (film.rs:114-118, 164-165).
int orc_render_aov(const orc_scene* s, const float* cam13, unsigned threads, uint8_t* hit, uint32_t* mat, float* depth,
                   uint32_t* steps) {
  const Scene& sc = s->sc;
  Camera cam = to_cam(cam13);
  Renderer<float> rd(&sc);
  for_bands(sc.height, threads, [&](uint32_t y0, uint32_t y1) {
    for (uint32_t y = y0; y < y1; y++)
      for (uint32_t x = 0; x < sc.width; x++) {
        Ray<float> ray = Renderer<float>::cam_ray(cam, (float)x + 0.5f, (float)y + 0.5f, sc.width, sc.height);
        Hit<float> hh = rd.march(ray);
        size_t i = (size_t)y * sc.width + x;
        hit[i] = hh.hit ? 1 : 0;
        mat[i] = hh.hit ? hh.mat : 0xffffffffu;
        depth[i] = hh.hit ? dot(hh.pos - f3<float>(cam.eye), f3<float>(cam.forward)) : INFINITY;
        if (steps) steps[i] = hh.steps;
      }
  });
  return 0;
}

// ---- film (film.rs:63-225, camera.rs:71-118) -------------------------------------------

}  // extern "C"

static uint32_t film_hash(uint32_t v) {  // film.rs:218-224
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  return v ^ (v >> 16);
}

static void film_aperture_sample(uint32_t px, uint32_t py, uint32_t sx, uint32_t sy, uint32_t n, float* out) {
  if (n == 1) {  // film.rs:203-216
    out[0] = 0.5f;
    out[1] = 0.5f;
    return;
  }
  const uint32_t pixel_seed = film_hash(px ^ film_hash(py));
  const uint64_t n64 = n;
  const uint32_t cell_x = (uint32_t)(((uint64_t)sx + (uint64_t)sy + (uint64_t)pixel_seed) % n64);
  const uint32_t cell_y = (uint32_t)(((uint64_t)sx + 2 * (uint64_t)sy + (uint64_t)film_hash(pixel_seed)) % n64);
  const uint32_t seed = film_hash(pixel_seed ^ film_hash(sx) ^ film_hash(sy + 0xa511e9b3u));
  auto unit = [](uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); };
  out[0] = ((float)cell_x + unit(seed)) / (float)n;
  out[1] = ((float)cell_y + unit(film_hash(seed))) / (float)n;
}

static Ray<float> film_ray_with_lens(const Camera& cam, float ix, float iy, uint32_t w, uint32_t h, float radius,
                                     float focus, float ax, float ay) {
  Ray<float> pin = Renderer<float>::cam_ray(cam, ix, iy, w, h);  // camera.rs:71-102
  if (radius == 0.0f) return pin;
  float dx = 0.0f, dy = 0.0f;
  {  // concentric_disk, camera.rs:105-118
    const float x = 2.0f * ax - 1.0f, y = 2.0f * ay - 1.0f;
    if (!(x == 0.0f && y == 0.0f)) {
      float r, th;
      if (std::fabs(x) > std::fabs(y)) {
        r = x;
        th = 0.785398163397448309616f * (y / x);
      } else {
        r = y;
        th = 1.57079632679489661923f - 0.785398163397448309616f * (x / y);
      }
      dx = r * cosf(th);
      dy = r * sinf(th);
    }
  }
  const V3<float> eye = f3<float>(cam.eye);
  const V3<float> origin = eye + scale(f3<float>(cam.right), dx * radius) + scale(f3<float>(cam.up), dy * radius);
  const V3<float> foc = eye + scale(pin.d, focus / dot(pin.d, f3<float>(cam.forward)));
  return Ray<float>{origin, normalize(foc - origin)};
}

extern "C" {

void orc_ray_with_lens(const float* cam13, float ix, float iy, uint32_t w, uint32_t h, float radius, float focus,
                       float ax, float ay, float* origin3, float* dir3) {
  const Ray<float> r = film_ray_with_lens(to_cam(cam13), ix, iy, w, h, radius, focus, ax, ay);
  origin3[0] = r.o.x;
  origin3[1] = r.o.y;
  origin3[2] = r.o.z;
  dir3[0] = r.d.x;
  dir3[1] = r.d.y;
  dir3[2] = r.d.z;
}

void orc_aperture_sample(uint32_t px, uint32_t py, uint32_t sx, uint32_t sy, uint32_t n, float* out2) {
  film_aperture_sample(px, py, sx, sy, n, out2);
}

// render_film_with_threads (film.rs:74-198).  Any output may be NULL.
int orc_render_film(const orc_scene* s, const float* cam13, float radius, float focus, unsigned threads, float* beauty,
                    float* fg, float* alpha, float* depth, float* normals, uint32_t* mat) {
  const Scene& sc = s->sc;
  const Camera cam = to_cam(cam13);
  Renderer<float> rd(&sc);
  const uint32_t n = std::max(sc.aa, 1u);
  const float inv = 1.0f / (float)((uint64_t)n * (uint64_t)n);
  for_bands(sc.height, threads, [&](uint32_t y0, uint32_t y1) {
    for (uint32_t y = y0; y < y1; y++)
      for (uint32_t x = 0; x < sc.width; x++) {
        const Ray<float> cray = Renderer<float>::cam_ray(cam, (float)x + 0.5f, (float)y + 0.5f, sc.width, sc.height);
        const Hit<float> center = rd.march(cray);
        V3<float> b, f;
        uint64_t hits = 0;
        for (uint32_t sy = 0; sy < n; sy++)
          for (uint32_t sx = 0; sx < n; sx++) {
            const float ix = (float)x + ((float)sx + 0.5f) / (float)n;
            const float iy = (float)y + ((float)sy + 0.5f) / (float)n;
            float ap[2];
            film_aperture_sample(x, y, sx, sy, n, ap);
            const Ray<float> ray = n == 1 ? cray : film_ray_with_lens(cam, ix, iy, sc.width, sc.height, radius, focus,
                                                                      ap[0], ap[1]);
            const Hit<float> hit = n == 1 ? center : rd.march(ray);
            if (hit.hit) {
              const V3<float> r = rd.shade_hit(hit, ray, 0);
              b = b + r;
              f = f + r;
              hits++;
            } else {
              b = b + sky(ray.d, rd.sun());
            }
          }
        const size_t i = (size_t)y * sc.width + x;
        const V3<float> bo = scale(b, inv), fo = scale(f, inv);
        if (beauty) {
          beauty[3 * i] = bo.x;
          beauty[3 * i + 1] = bo.y;
          beauty[3 * i + 2] = bo.z;
        }
        if (fg) {
          fg[3 * i] = fo.x;
          fg[3 * i + 1] = fo.y;
          fg[3 * i + 2] = fo.z;
        }
        if (alpha) alpha[i] = (float)hits * inv;
        if (depth) depth[i] = center.hit ? dot(center.pos - f3<float>(cam.eye), f3<float>(cam.forward)) : INFINITY;
        if (normals) {
          const V3<float> nn = center.hit ? center.normal : V3<float>();
          normals[3 * i] = nn.x;
          normals[3 * i + 1] = nn.y;
          normals[3 * i + 2] = nn.z;
        }
        if (mat) mat[i] = center.hit ? center.mat : 0xffffffffu;
      }
  });
  return 0;
}

// ---- G-buffer frames and reprojection (lib.rs:951-1080, reproject.rs:57-170) -------------

}  // extern "C"

// lib.rs:981-1016: one centre sample, display colour (no bloom), hit distance, mover tag.
static void gbuffer_pixel(const Renderer<float>& rd, const Camera& cam, uint32_t w, uint32_t h, uint32_t x, uint32_t y,
                          const float* movers, uint32_t n_movers, uint8_t* color, float* depth, int32_t* obj) {
  const Scene& sc = *rd.sc;
  const Ray<float> ray = Renderer<float>::cam_ray(cam, (float)x + 0.5f, (float)y + 0.5f, w, h);
  const Hit<float> hit = rd.march(ray);
  V3<float> hdr;
  float dd = INFINITY;
  int32_t o = -1;
  if (hit.hit) {
    for (uint32_t k = 0; k < n_movers; k++) {
      const V3<float> c(movers[4 * k], movers[4 * k + 1], movers[4 * k + 2]);
      if (length(hit.pos - c) <= movers[4 * k + 3] + 0.05f) {
        o = (int32_t)k;
        break;
      }
    }
    hdr = rd.shade_hit(hit, ray, 0);
    dd = hit.t;
  } else {
    hdr = sky(ray.d, rd.sun());
  }
  const V3<float> d = clamp01(gamma(aces(scale(hdr, sc.exposure))));
  color[0] = (uint8_t)(d.x * 255.0f + 0.5f);  // `(d * 255.0 + 0.5) as u8` on a clamped value
  color[1] = (uint8_t)(d.y * 255.0f + 0.5f);
  color[2] = (uint8_t)(d.z * 255.0f + 0.5f);
  color[3] = 255;
  if (depth) *depth = dd;
  if (obj) *obj = o;
}

// reproject.rs:57-69
static bool project_px(const Camera& c, V3<float> world, uint32_t w, uint32_t h, float& sx, float& sy, float& fz) {
  const V3<float> v = world - f3<float>(c.eye);
  fz = dot(v, f3<float>(c.forward));
  if (fz <= 1e-4f) return false;
  const float aspect = (float)w / (float)h;
  const float ndc_x = dot(v, f3<float>(c.right)) / fz;
  const float ndc_y = dot(v, f3<float>(c.up)) / fz;
  sx = (ndc_x / (aspect * c.fov_scale) + 1.0f) * 0.5f * (float)w;
  sy = (1.0f - ndc_y / c.fov_scale) * 0.5f * (float)h;
  return true;
}

static int32_t as_i32(float f) {  // Rust `f32 as i32`
  if (std::isnan(f)) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f < -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}

// reproject.rs:82-130, the reference's sequential loop and z-buffer
static void warp_frame(uint32_t w, uint32_t h, const Camera& prev, const uint8_t* color, const float* depth,
                       const int32_t* obj, const Camera& next, const float* deltas, uint32_t n_deltas, uint8_t* out,
                       uint8_t* filled) {
  const size_t n = (size_t)w * h;
  std::vector<float> zbuf(n, INFINITY);
  std::memset(out, 0, n * 4);
  std::memset(filled, 0, n);
  for (uint32_t y = 0; y < h; y++)
    for (uint32_t x = 0; x < w; x++) {
      const size_t i = (size_t)y * w + x;
      const V3<float> dir = Renderer<float>::cam_ray(prev, (float)x + 0.5f, (float)y + 0.5f, w, h).d;
      const float d = std::isfinite(depth[i]) ? depth[i] : 1.0e6f;
      V3<float> world = f3<float>(prev.eye) + scale(dir, d);
      const int32_t oi = obj[i];
      if (oi >= 0 && (uint32_t)oi < n_deltas)
        world = world + V3<float>(deltas[3 * oi], deltas[3 * oi + 1], deltas[3 * oi + 2]);
      float sx, sy, fz;
      if (!project_px(next, world, w, h, sx, sy, fz)) continue;
      const int32_t nx = as_i32(std::floor(sx)), ny = as_i32(std::floor(sy));
      if (nx >= 0 && ny >= 0 && (uint32_t)nx < w && (uint32_t)ny < h) {
        const size_t j = (size_t)ny * w + (uint32_t)nx;
        if (fz < zbuf[j]) {
          zbuf[j] = fz;
          std::memcpy(out + 4 * j, color + 4 * i, 4);
          filled[j] = 1;
        }
      }
    }
}

extern "C" {

// render_gbuffer (lib.rs:951-1030); movers: [n][4] centre xyz, radius.
int orc_render_gbuffer(const orc_scene* s, const float* cam13, const float* movers, uint32_t n_movers, unsigned threads,
                       uint8_t* color, float* depth, int32_t* obj) {
  const Scene& sc = s->sc;
  const Camera cam = to_cam(cam13);
  Renderer<float> rd(&sc);
  for_bands(sc.height, threads, [&](uint32_t y0, uint32_t y1) {
    for (uint32_t y = y0; y < y1; y++)
      for (uint32_t x = 0; x < sc.width; x++) {
        const size_t i = (size_t)y * sc.width + x;
        gbuffer_pixel(rd, cam, sc.width, sc.height, x, y, movers, n_movers, color + 4 * i, depth + i, obj + i);
      }
  });
  return 0;
}

// reproject::warp (fill == 0) or reproject::reproject (fill != 0), reproject.rs:82-170.
void orc_reproject(uint32_t w, uint32_t h, const float* prev_cam13, const uint8_t* color, const float* depth,
                   const int32_t* obj, const float* cam13, const float* deltas, uint32_t n_deltas, int fill,
                   uint8_t* out, uint8_t* filled) {
  warp_frame(w, h, to_cam(prev_cam13), color, depth, obj, to_cam(cam13), deltas, n_deltas, out, filled);
  if (!fill) return;
  std::vector<uint8_t> f(filled, filled + (size_t)w * h);
  for (uint32_t y = 0; y < h; y++) {
    bool have = false;
    uint8_t last[4] = {0, 0, 0, 0};
    for (uint32_t x = 0; x < w; x++) {
      const size_t i = (size_t)y * w + x;
      if (f[i]) {
        std::memcpy(last, out + 4 * i, 4);
        have = true;
      } else if (have) {
        std::memcpy(out + 4 * i, last, 4);
        f[i] = 1;
      }
    }
    have = false;
    for (uint32_t x = w; x-- > 0;) {
      const size_t i = (size_t)y * w + x;
      if (f[i]) {
        std::memcpy(last, out + 4 * i, 4);
        have = true;
      } else if (have) {
        std::memcpy(out + 4 * i, last, 4);
      }
    }
  }
}

// reproject_hybrid (lib.rs:1043-1080): the warp, then the holes rendered through the scene.
uint32_t orc_reproject_hybrid(const orc_scene* s, uint32_t w, uint32_t h, const float* prev_cam13, const uint8_t* color,
                              const float* depth, const int32_t* obj, const float* cam13, const float* deltas,
                              uint32_t n_deltas, uint8_t* out) {
  std::vector<uint8_t> filled((size_t)w * h);
  const Camera cam = to_cam(cam13);
  warp_frame(w, h, to_cam(prev_cam13), color, depth, obj, cam, deltas, n_deltas, out, filled.data());
  Renderer<float> rd(&s->sc);
  uint32_t holes = 0;
  for (uint32_t y = 0; y < h; y++)
    for (uint32_t x = 0; x < w; x++) {
      const size_t i = (size_t)y * w + x;
      if (filled[i]) continue;
      gbuffer_pixel(rd, cam, w, h, x, y, nullptr, 0, out + 4 * i, nullptr, nullptr);
      holes++;
    }
  return holes;
}

// World field at points: linear=0 -> accelerated `world()`, 1 -> `world_linear()`.
int orc_field(const orc_scene* s, const float* pts, size_t n, int linear, float* dist, uint32_t* mat) {
  Renderer<float> rd(&s->sc);
  for (size_t i = 0; i < n; i++) {
    V3<float> p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    Field<float> f = linear ? rd.world.linear(p) : rd.world(p);
    dist[i] = f.dist;
    mat[i] = f.mat;
  }
  return 0;
}

// Raw march (film_depth_tests.rs semantics): rays o.xyz d.xyz -> t, hit, mat, normal, steps.
int orc_march(const orc_scene* s, const float* rays, size_t n, float* t, uint8_t* hit, uint32_t* mat, float* normal,
              uint32_t* steps) {
  Renderer<float> rd(&s->sc);
  for (size_t i = 0; i < n; i++) {
    Ray<float> r{V3<float>(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]),
                 V3<float>(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5])};
    Hit<float> h = rd.march(r);
    t[i] = h.t;
    hit[i] = h.hit;
    mat[i] = h.mat;
    normal[3 * i] = h.normal.x; normal[3 * i + 1] = h.normal.y; normal[3 * i + 2] = h.normal.z;
    steps[i] = h.steps;
  }
  return 0;
}

// Single-primitive SDF (sdf.rs) for known-answer tests.
float orc_prim_distance(uint32_t kind, const float* params, const float* p) {
  return prim_distance<float>(kind, params, V3<float>(p[0], p[1], p[2]));
}
// Dual-number gradient of one primitive at p (dual.rs) — value + gradient.
void orc_prim_dual(uint32_t kind, const float* params, const float* p, float* out4) {
  Obj o{};
  o.prim = kind;
  std::memcpy(o.p, params, 32);
  float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::memcpy(o.rot, id, 36);
  o.scale = 1.0f;
  Dual<float> x{p[0], V3<float>(1, 0, 0)}, y{p[1], V3<float>(0, 1, 0)}, z{p[2], V3<float>(0, 0, 1)};
  Dual<float> d = object_field_dual<float>(o, x, y, z);
  out4[0] = d.v; out4[1] = d.g.x; out4[2] = d.g.y; out4[3] = d.g.z;
}
// Scene-level dual normal / tetrahedron normal at p.
void orc_normals(const orc_scene* s, const float* p, float* dual3, float* tetra3) {
  Renderer<float> rd(&s->sc);
  V3<float> q(p[0], p[1], p[2]);
  V3<float> a = normal_dual<float>(s->sc, q), b = rd.mr.normal(rd.world, q);
  dual3[0] = a.x; dual3[1] = a.y; dual3[2] = a.z;
  tetra3[0] = b.x; tetra3[1] = b.y; tetra3[2] = b.z;
}
int orc_dual_safe(const orc_scene* s) { return scene_dual_safe(s->sc) ? 1 : 0; }

// Shading closed forms for known-answer tests.
void orc_brdf(const float* n, const float* l, const float* v, const float* albedo, float metallic, float roughness,
              float reflectance, float* out3) {
  V3<float> r = brdf(f3<float>(n), f3<float>(l), f3<float>(v), f3<float>(albedo), metallic, roughness, reflectance);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float orc_fresnel(float c, float f0) { return fresnel_schlick(c, f0); }
void orc_aces_gamma(const float* c, float* aces3, float* gamma3) {
  V3<float> a = aces(f3<float>(c)), g = gamma(f3<float>(c));
  aces3[0] = a.x; aces3[1] = a.y; aces3[2] = a.z;
  gamma3[0] = g.x; gamma3[1] = g.y; gamma3[2] = g.z;
}
void orc_sky(const float* dir, const float* sun, int diffuse, float* out3) {
  V3<float> r = diffuse ? sky_diffuse(f3<float>(dir), f3<float>(sun)) : sky(f3<float>(dir), f3<float>(sun));
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float orc_hash_cell(int32_t ix, int32_t iy) { return hash_cell(ix, iy); }
uint64_t orc_fnv1a(const uint8_t* b, size_t n) { return fnv1a(b, n); }
float orc_powi(float x, int n) { return rpowi(x, n); }
float orc_soft_shadow(const orc_scene* s, const float* o, const float* d, float max_t, float k) {
  Renderer<float> rd(&s->sc);
  return rd.mr.soft_shadow(rd.world, f3<float>(o), f3<float>(d), max_t, k);
}
float orc_ao(const orc_scene* s, const float* p, const float* n) {
  Renderer<float> rd(&s->sc);
  return rd.mr.ambient_occlusion(rd.world, f3<float>(p), f3<float>(n));
}
void orc_gi_sample(const orc_scene* s, const float* p, const float* n, float* out3) {
  V3<float> r = s->sc.has_gi ? gi_sample<float>(s->sc.gi, f3<float>(p), f3<float>(n)) : V3<float>();
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
// GiVolume::bake over an arbitrary radiance = linear field in x (engine.rs:145-156 KAT).
int orc_gi_bake_linear_x(const float* lo, const float* hi, const uint32_t* dims, uint32_t samples, const float* p,
                         const float* n, float* out3) {
  Gi g;
  uint64_t nx = std::max(dims[0], 1u), ny = std::max(dims[1], 1u), nz = std::max(dims[2], 1u);
  V3<float> mn = f3<float>(lo), size = f3<float>(hi) - mn;
  V3<float> cell(size.x / (float)(std::max<uint64_t>(nx, 2) - 1), size.y / (float)(std::max<uint64_t>(ny, 2) - 1),
                 size.z / (float)(std::max<uint64_t>(nz, 2) - 1));
  g.min[0] = mn.x; g.min[1] = mn.y; g.min[2] = mn.z;
  g.size[0] = size.x; g.size[1] = size.y; g.size[2] = size.z;
  g.dims[0] = (uint32_t)nx; g.dims[1] = (uint32_t)ny; g.dims[2] = (uint32_t)nz;
  g.cubes.assign(nx * ny * nz * 18, 0.0f);
  for (uint64_t k = 0; k < nz; k++)
    for (uint64_t j = 0; j < ny; j++)
      for (uint64_t i = 0; i < nx; i++) {
        V3<float> probe = mn + V3<float>((float)i * cell.x, (float)j * cell.y, (float)k * cell.z);
        for (int a = 0; a < 6; a++) {
          auto dirs = axis_dirs(V3<float>(kAxes[a][0], kAxes[a][1], kAxes[a][2]), samples);
          V3<float> acc;
          float ws = 0;
          for (auto& dw : dirs) {
            acc = acc + scale(splat(probe.x), dw.w);
            ws += dw.w;
          }
          V3<float> c = scale(acc, 1.0f / std::fmax(ws, 1e-4f));
          float* o = &g.cubes[(((k * ny + j) * nx + i) * 6 + a) * 3];
          o[0] = c.x; o[1] = c.y; o[2] = c.z;
        }
      }
  V3<float> r = gi_sample<float>(g, f3<float>(p), f3<float>(n));
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
  return 0;
}
// Volume sample KAT (volume.rs:194-228).
float orc_volume_sample(const uint32_t* dims, const float* mn, const float* cell, const float* data, size_t n, const float* p) {
  Vol v;
  std::memcpy(v.dims, dims, 12); std::memcpy(v.min, mn, 12); std::memcpy(v.cell, cell, 12);
  v.data.assign(data, data + n);
  return vol_sample<float>(v, f3<float>(p));
}
// Camera::look_at (camera.rs:50-55) restated for cross-checking the library's host helper.
void orc_camera_look_at(const float* eye, const float* target, const float* up, float fov_y, float* cam13) {
  V3<float> e = f3<float>(eye);
  V3<float> fwd = normalize(f3<float>(target) - e);
  V3<float> right = normalize(cross(fwd, f3<float>(up)));
  V3<float> u = cross(right, fwd);
  float c[13] = {e.x, e.y, e.z, fwd.x, fwd.y, fwd.z, right.x, right.y, right.z, u.x, u.y, u.z, tanf(fov_y * 0.5f)};
  std::memcpy(cam13, c, sizeof c);
}
// The HDR image after the bloom add of post::resolve (post.rs:41-46), before exposure and
// tone mapping: hdr and out are w*h*3 floats.
void orc_bloom(const float* hdr, uint32_t w, uint32_t h, float threshold, float intensity, uint32_t radius, float* out) {
  size_t n = (size_t)w * h;
  std::vector<V3<float>> color(n);
  for (size_t i = 0; i < n; i++) color[i] = V3<float>(hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]);
  bloom_add(color, w, h, threshold, intensity, radius);
  for (size_t i = 0; i < n; i++) { out[3 * i] = color[i].x; out[3 * i + 1] = color[i].y; out[3 * i + 2] = color[i].z; }
}
// Scene::ibl_irradiance (lib.rs:1379-1398) of the scene's sky for a unit normal n.
void orc_ibl_irradiance(const orc_scene* s, const float* n, float* out3) {
  Renderer<float> rd(&s->sc);
  V3<float> r = rd.ibl_irradiance(f3<float>(n));
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
// Light::toward / attenuation / shadow_k (lib.rs:528-556) of one light seen from p:
// out[0..3] = direction, out[3] = distance, out[4] = attenuation, out[5] = shadow k.
void orc_light_eval(int directional, const float* vec3, float radius, const float* p, float* out6) {
  Lgt l{};
  l.directional = directional != 0;
  std::memcpy(l.vec, vec3, 12);
  l.radius = radius;
  V3<float> d;
  float dist;
  light_toward<float>(l, f3<float>(p), d, dist);
  out6[0] = d.x; out6[1] = d.y; out6[2] = d.z; out6[3] = dist;
  out6[4] = light_atten<float>(l, dist);
  out6[5] = light_shadow_k<float>(l, dist);
}
// shade::apply_fog (shade.rs:113-116).
void orc_apply_fog(const float* c, const float* fog, float t, float density, float* out3) {
  V3<float> r = apply_fog(f3<float>(c), f3<float>(fog), t, density);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
void orc_gaussian_kernel(uint32_t radius, float* out, uint32_t* n) {
  auto k = gaussian_kernel(radius);
  *n = (uint32_t)k.size();
  if (out) std::memcpy(out, k.data(), k.size() * 4);
}

// Instrumented frame: FLOPs (CountF instantiation) + field evaluations per phase.
// counters: [0]=flops [1]=trans [2]=sqrt [3]=div [4..4+PH_COUNT)=evals per phase
//           [4+PH_COUNT]=exact object evals [5+PH_COUNT]=bound substitutions
// Rows [y0, y1) only; threads share nothing but the read-only scene.
int orc_count_rows(const orc_scene* s, const float* cam13, unsigned threads, uint32_t y0, uint32_t y1, uint64_t* counters) {
  const Scene& sc = s->sc;
  Camera cam = to_cam(cam13);
  Renderer<CountF> rd(&sc);
  std::mutex mu;
  uint32_t rows = y1 - y0;
  for (int i = 0; i < 6 + PH_COUNT; i++) counters[i] = 0;
  for_bands(rows, threads, [&](uint32_t a, uint32_t b) {
    OpCounts ops;
    uint64_t ev[PH_COUNT] = {0};
    uint64_t ex = 0, bd = 0;
    g_ops = &ops;
    g_evals = ev;
    g_obj_exact = &ex;
    g_obj_bound = &bd;
    for (uint32_t y = a; y < b; y++)
      for (uint32_t x = 0; x < sc.width; x++) (void)rd.shade_pixel(cam, x, y0 + y);
    g_ops = nullptr;
    g_evals = nullptr;
    g_obj_exact = g_obj_bound = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    counters[0] += ops.flops; counters[1] += ops.trans; counters[2] += ops.sqrt; counters[3] += ops.div;
    for (int i = 0; i < PH_COUNT; i++) counters[4 + i] += ev[i];
    counters[4 + PH_COUNT] += ex;
    counters[5 + PH_COUNT] += bd;
  });
  return 0;
}
int orc_phase_count(void) { return PH_COUNT; }

// BMP encoder (framebuffer.rs:55-98) over opaque RGBA8 pixels.
size_t orc_to_bmp(const uint8_t* rgba, uint32_t w, uint32_t h, uint8_t* out) {
  size_t pad = (4 - (w * 3) % 4) % 4;
  size_t row = (size_t)w * 3 + pad, pix = row * h, total = 54 + pix;
  if (!out) return total;
  auto u32le = [](uint8_t* p, uint32_t v) { std::memcpy(p, &v, 4); };
  auto u16le = [](uint8_t* p, uint16_t v) { std::memcpy(p, &v, 2); };
  std::memset(out, 0, 54);
  out[0] = 'B'; out[1] = 'M';
  u32le(out + 2, (uint32_t)total); u32le(out + 10, 54); u32le(out + 14, 40);
  u32le(out + 18, w); u32le(out + 22, h); u16le(out + 26, 1); u16le(out + 28, 24);
  u32le(out + 34, (uint32_t)pix); u32le(out + 38, 2835); u32le(out + 42, 2835);
  uint8_t* q = out + 54;
  for (int64_t y = (int64_t)h - 1; y >= 0; y--) {
    for (uint32_t x = 0; x < w; x++) {
      const uint8_t* px = rgba + 4 * ((size_t)y * w + x);
      *q++ = px[2]; *q++ = px[1]; *q++ = px[0];
    }
    for (size_t i = 0; i < pad; i++) *q++ = 0;
  }
  return total;
}

}  // extern "C"
