// mesh_oracle.cpp — CPU restatement of the reference's mesh -> SDF bake (TEST INFRASTRUCTURE).
//
// Linked into liboracle_mm3e.so; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load it, as the checker of the GPU bake (libmm3e_hip.so), never as part of it.
//
// Restated reference code (paths relative to the reference tree):
//   mm3e-orchestrator/src/mesh.rs:10-103   Mesh, Mesh::validate, parse_obj
//   mm3e-orchestrator/src/mesh.rs:111-146  closest_on_triangle
//   mm3e-orchestrator/src/mesh.rs:148-260  TriBvh (build, nearest_sq), box_dist_sq
//   mm3e-orchestrator/src/mesh.rs:265-336  bake_sdf (threads over z-slabs)
//   mm3e-kit/src/triangle.rs:4-34          edge, owns_boundary, x_crossing
// Every f32 / f64 operation is written in the reference's order; the library is built with
// -ffp-contract=off, so each rounds once, as Rust's arithmetic does.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace orc_mesh {

struct V {
  float x, y, z;
};
static V operator+(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
static V operator-(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
static V scale(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
static float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V cross(V a, V b) { return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static V vmin(V a, V b) { return V{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
static V vmax(V a, V b) { return V{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
static V splat(float s) { return V{s, s, s}; }

struct Mesh {
  std::vector<V> positions;
  std::vector<uint32_t> tri;  // [t*3 + c]
  size_t ntri() const { return tri.size() / 3; }
};

// mesh.rs:33-52
static bool validate(const Mesh& m, std::string& err) {
  if (m.positions.empty()) return err = "mesh has no vertices", false;
  if (m.tri.empty()) return err = "mesh has no triangles", false;
  for (size_t i = 0; i < m.positions.size(); i++) {
    const V p = m.positions[i];
    if (!(std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z)))
      return err = "mesh vertex " + std::to_string(i) + " is non-finite", false;
  }
  for (size_t t = 0; t < m.ntri(); t++)
    for (int c = 0; c < 3; c++)
      if (m.tri[3 * t + c] >= m.positions.size())
        return err = "mesh triangle " + std::to_string(t) + " references a vertex out of range", false;
  return true;
}

// mesh.rs:57-103.  Numbers: Rust's f32 / i64 `parse` on the token.
static bool num_f32(const std::string& s, float& v) {
  size_t b = (!s.empty() && s[0] == '+') ? 1 : 0;
  if (b >= s.size()) return false;
  auto r = std::from_chars(s.data() + b, s.data() + s.size(), v);
  if (r.ptr != s.data() + s.size()) return false;
  if (r.ec == std::errc::result_out_of_range) v = std::strtof(s.c_str() + b, nullptr);
  else if (r.ec != std::errc()) return false;
  return true;
}
static bool num_i64(const std::string& s, int64_t& v) {
  size_t b = (!s.empty() && s[0] == '+') ? 1 : 0;
  if (b >= s.size() || (b == 1 && s[1] == '-')) return false;
  auto r = std::from_chars(s.data() + b, s.data() + s.size(), v);
  return r.ec == std::errc() && r.ptr == s.data() + s.size();
}

static bool parse_obj(const std::string& src, Mesh& m, std::string& err) {
  size_t ln = 0, pos = 0;
  while (pos < src.size()) {
    size_t nl = src.find('\n', pos);
    if (nl == std::string::npos) nl = src.size();
    std::string raw = src.substr(pos, nl - pos);
    pos = nl + 1;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    ln++;
    std::string line = raw.substr(0, raw.find('#'));
    std::vector<std::string> toks;
    {
      size_t i = 0;
      auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; };
      while (i < line.size()) {
        while (i < line.size() && ws(line[i])) i++;
        size_t j = i;
        while (j < line.size() && !ws(line[j])) j++;
        if (j > i) toks.push_back(line.substr(i, j - i));
        i = j;
      }
    }
    if (toks.empty()) continue;
    const std::string at = "line " + std::to_string(ln) + ": ";
    if (toks[0] == "v") {
      float c[3];
      const char* what[3] = {"x", "y", "z"};
      for (int a = 0; a < 3; a++) {
        if (toks.size() < (size_t)a + 2) return err = at + "missing " + what[a], false;
        if (!num_f32(toks[a + 1], c[a])) return err = at + "bad number for " + what[a], false;
      }
      if (!(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2])))
        return err = at + "non-finite vertex", false;
      m.positions.push_back(V{c[0], c[1], c[2]});
    } else if (toks[0] == "f") {
      std::vector<uint32_t> idx;
      for (size_t t = 1; t < toks.size(); t++) {
        const std::string first = toks[t].substr(0, toks[t].find('/'));
        int64_t raw_i;
        if (!num_i64(first, raw_i)) return err = at + "bad face index '" + toks[t] + "'", false;
        const int64_t n = (int64_t)m.positions.size();
        const int64_t r = raw_i > 0 ? raw_i - 1 : n + raw_i;
        if (r < 0 || r >= n)
          return err = at + "face index " + std::to_string(raw_i) + " out of range (have " + std::to_string(n) +
                       " vertices)",
                 false;
        idx.push_back((uint32_t)r);
      }
      if (idx.size() < 3) return err = at + "face needs at least 3 vertices", false;
      for (size_t t = 1; t + 1 < idx.size(); t++) {
        m.tri.push_back(idx[0]);
        m.tri.push_back(idx[t]);
        m.tri.push_back(idx[t + 1]);
      }
    }
  }
  return validate(m, err);
}

// mm3e-kit/src/triangle.rs:4-34
static double edge(V a, V b, double y, double z) {
  return ((double)b.y - (double)a.y) * (z - (double)a.z) - ((double)b.z - (double)a.z) * (y - (double)a.y);
}
static bool owns_boundary(V a, V b) { return b.z > a.z || (b.z == a.z && b.y < a.y); }
static bool x_crossing(V a, V b, V c, float y, float z, float& x) {
  double area = edge(a, b, (double)c.y, (double)c.z);
  if (area == 0.0 || !std::isfinite(area) || !std::isfinite(y) || !std::isfinite(z)) return false;
  if (area < 0.0) {
    std::swap(b, c);
    area = -area;
  }
  const double yd = y, zd = z;
  const double w[3] = {edge(b, c, yd, zd), edge(c, a, yd, zd), edge(a, b, yd, zd)};
  const V s[3] = {b, c, a}, e[3] = {c, a, b};
  for (int i = 0; i < 3; i++)
    if (w[i] < 0.0 || (w[i] == 0.0 && !owns_boundary(s[i], e[i]))) return false;
  x = (float)((w[0] * (double)a.x + w[1] * (double)b.x + w[2] * (double)c.x) / area);
  return std::isfinite(x);
}

// mesh.rs:111-146
static V closest_on_triangle(V p, V a, V b, V c) {
  const V ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const V bp = p - b;
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + scale(ab, d1 / (d1 - d3));
  const V cp = p - c;
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + scale(ac, d2 / (d2 - d6));
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    return b + scale(c - b, (d4 - d3) / ((d4 - d3) + (d5 - d6)));
  const float denom = 1.0f / (va + vb + vc);
  return a + scale(ab, vb * denom) + scale(ac, vc * denom);
}

// mesh.rs:148-260
struct Node {
  V lo, hi;
  uint32_t left, right;
  int32_t leaf;
};
struct Item {
  size_t i;
  V lo, hi, c;
};
static float box_dist_sq(V p, V lo, V hi) {
  const V d = p - vmin(vmax(p, lo), hi);
  return dot(d, d);
}

struct Bvh {
  std::vector<Node> nodes;
  uint32_t root = 0;

  static uint32_t build_from(std::vector<Item>& items, size_t b, size_t e, std::vector<Node>& nodes) {
    V lo = splat(INFINITY), hi = splat(-INFINITY);
    for (size_t i = b; i < e; i++) {
      lo = vmin(lo, items[i].lo);
      hi = vmax(hi, items[i].hi);
    }
    if (e - b == 1) {
      nodes.push_back(Node{lo, hi, 0, 0, (int32_t)items[b].i});
      return (uint32_t)nodes.size() - 1;
    }
    V clo = splat(INFINITY), chi = splat(-INFINITY);
    for (size_t i = b; i < e; i++) {
      clo = vmin(clo, items[i].c);
      chi = vmax(chi, items[i].c);
    }
    const V ext = chi - clo;
    float (*coord)(V);
    if (ext.x >= ext.y && ext.x >= ext.z) coord = [](V c) { return c.x; };
    else if (ext.y >= ext.z) coord = [](V c) { return c.y; };
    else coord = [](V c) { return c.z; };
    std::sort(items.begin() + b, items.begin() + e, [coord](const Item& x, const Item& y) {
      const float cx = coord(x.c), cy = coord(y.c);
      if (cx < cy) return true;
      if (cy < cx) return false;
      return x.i < y.i;
    });
    const size_t mid = b + (e - b) / 2;
    const uint32_t left = build_from(items, b, mid, nodes);
    const uint32_t right = build_from(items, mid, e, nodes);
    nodes.push_back(Node{lo, hi, left, right, -1});
    return (uint32_t)nodes.size() - 1;
  }

  bool build(const Mesh& m) {
    std::vector<Item> items;
    for (size_t t = 0; t < m.ntri(); t++) {
      const V a = m.positions[m.tri[3 * t]], b = m.positions[m.tri[3 * t + 1]], c = m.positions[m.tri[3 * t + 2]];
      const V n = cross(b - a, c - a);
      if (dot(n, n) < 1e-16f) continue;
      items.push_back(Item{t, vmin(vmin(a, b), c), vmax(vmax(a, b), c), scale(a + b + c, 1.0f / 3.0f)});
    }
    if (items.empty()) return false;
    nodes.reserve(items.size() * 2);
    root = build_from(items, 0, items.size(), nodes);
    return true;
  }

  float nearest_sq(V p, const Mesh& m, uint64_t* boxes, uint64_t* tris) const {
    float best = INFINITY;
    uint32_t stack[64];
    size_t sp = 0;
    stack[sp++] = root;
    if (boxes) (*boxes)++;  // the root's box; every other box is tested once, from its parent
    while (sp > 0) {
      const Node& n = nodes[stack[--sp]];
      if (box_dist_sq(p, n.lo, n.hi) >= best) continue;
      if (n.leaf >= 0) {
        const size_t t = (size_t)n.leaf;
        const V q = closest_on_triangle(p, m.positions[m.tri[3 * t]], m.positions[m.tri[3 * t + 1]],
                                        m.positions[m.tri[3 * t + 2]]);
        if (tris) (*tris)++;
        const V d = p - q;
        best = std::fmin(best, dot(d, d));
      } else {
        const float ld = box_dist_sq(p, nodes[n.left].lo, nodes[n.left].hi);
        const float rd = box_dist_sq(p, nodes[n.right].lo, nodes[n.right].hi);
        if (boxes) (*boxes) += 2;
        const uint32_t near_n = ld <= rd ? n.left : n.right, far_n = ld <= rd ? n.right : n.left;
        stack[sp] = far_n;
        stack[sp + 1] = near_n;
        sp += 2;
      }
    }
    return best;
  }
};

static float brute_nearest_sq(V p, const Mesh& m) {
  float best = INFINITY;
  for (size_t t = 0; t < m.ntri(); t++) {
    const V a = m.positions[m.tri[3 * t]], b = m.positions[m.tri[3 * t + 1]], c = m.positions[m.tri[3 * t + 2]];
    const V n = cross(b - a, c - a);
    if (dot(n, n) < 1e-16f) continue;  // the BVH holds only these
    const V d = p - closest_on_triangle(p, a, b, c);
    best = std::fmin(best, dot(d, d));
  }
  return best;
}

struct Layout {
  size_t nx, ny, nz;
  V lo, cell;
};

// mesh.rs:265-336, slabs k0..k1 of the grid (all of it when k1 == 0)
static bool bake(const Mesh& m, size_t resolution, float padding, unsigned threads, size_t k0, size_t k1,
                 Layout& L, std::vector<float>& data, std::string& err) {
  if (!validate(m, err)) return false;
  V mlo = splat(INFINITY), mhi = splat(-INFINITY);
  for (const V& p : m.positions) {
    mlo = vmin(mlo, p);
    mhi = vmax(mhi, p);
  }
  if (padding <= 0.0f || !std::isfinite(padding)) return err = "padding must be positive and finite", false;
  resolution = std::min<size_t>(std::max<size_t>(resolution, 2), 512);
  const V lo = mlo - splat(padding), hi = mhi + splat(padding);
  const V size = hi - lo;
  const float longest = std::fmax(std::fmax(std::fmax(size.x, size.y), size.z), 1e-6f);
  auto dim = [&](float s) {
    const float r = std::round((s / longest) * (float)(resolution - 1));
    const size_t u = r > 0.0f ? (size_t)r : 0;  // `as usize`
    return std::max<size_t>(u + 1, 2);
  };
  L.nx = dim(size.x);
  L.ny = dim(size.y);
  L.nz = dim(size.z);
  L.lo = lo;
  L.cell = V{size.x / (float)(L.nx - 1), size.y / (float)(L.ny - 1), size.z / (float)(L.nz - 1)};
  Bvh bvh;
  if (!bvh.build(m)) return err = "mesh has no non-degenerate triangles", false;
  if (k1 == 0 || k1 > L.nz) k1 = L.nz;
  if (k0 > k1) k0 = k1;
  const size_t nx = L.nx, ny = L.ny;
  data.assign((k1 - k0) * nx * ny, 0.0f);
  threads = std::max(threads, 1u);
  const size_t span = (k1 - k0 + threads - 1) / threads;
  std::vector<std::thread> pool;
  for (unsigned ti = 0; ti < threads; ti++) {
    const size_t a = k0 + ti * span, b = std::min(k0 + (ti + 1) * span, k1);
    if (a >= b) break;
    pool.emplace_back([&, a, b] {
      std::vector<float> xs;
      for (size_t k = a; k < b; k++) {
        const float z = L.lo.z + (float)k * L.cell.z;
        for (size_t j = 0; j < ny; j++) {
          const float y = L.lo.y + (float)j * L.cell.y;
          xs.clear();
          for (size_t t = 0; t < m.ntri(); t++) {
            float x;
            if (x_crossing(m.positions[m.tri[3 * t]], m.positions[m.tri[3 * t + 1]], m.positions[m.tri[3 * t + 2]], y,
                           z, x))
              xs.push_back(x);
          }
          std::sort(xs.begin(), xs.end());
          for (size_t i = 0; i < nx; i++) {
            const V p{L.lo.x + (float)i * L.cell.x, L.lo.y + (float)j * L.cell.y, L.lo.z + (float)k * L.cell.z};
            const float u = std::sqrt(bvh.nearest_sq(p, m, nullptr, nullptr));
            const size_t below = (size_t)(std::partition_point(xs.begin(), xs.end(), [&](float x) { return x < p.x; }) -
                                          xs.begin());
            data[((k - k0) * ny + j) * nx + i] = (below % 2 == 1) ? -u : u;
          }
        }
      }
    });
  }
  for (auto& t : pool) t.join();
  // SdfVolume::new: finite cells and samples (volume.rs:27-85)
  const float ms[3] = {L.lo.x, L.lo.y, L.lo.z};
  for (float v : ms)
    if (!std::isfinite(v)) return err = "SdfVolume min must be finite", false;
  const float cs[3] = {L.cell.x, L.cell.y, L.cell.z};
  for (float c : cs)
    if (!(std::isfinite(c) && c > 0.0f)) return err = "SdfVolume cell sizes must be positive and finite", false;
  for (size_t i = 0; i < data.size(); i++)
    if (!std::isfinite(data[i])) return err = "SdfVolume sample " + std::to_string(i) + " is non-finite", false;
  return true;
}

static Mesh from_arrays(const float* pos, size_t npos, const uint32_t* tris, size_t ntri) {
  Mesh m;
  for (size_t i = 0; i < npos; i++) m.positions.push_back(V{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
  m.tri.assign(tris, tris + 3 * ntri);
  return m;
}

static void put_err(char* err, size_t n, const std::string& s) {
  if (err && n) std::snprintf(err, n, "%s", s.c_str());
}

}  // namespace orc_mesh

using namespace orc_mesh;

extern "C" {

int orc_mesh_parse_obj(const char* text, size_t len, float* pos, size_t pos_cap, uint32_t* tris, size_t tri_cap,
                       size_t* npos, size_t* ntri, char* err, size_t errlen) {
  Mesh m;
  std::string e;
  if (!parse_obj(std::string(text, len), m, e)) {
    put_err(err, errlen, e);
    return 1;
  }
  *npos = m.positions.size();
  *ntri = m.ntri();
  if (pos && pos_cap >= m.positions.size())
    for (size_t i = 0; i < m.positions.size(); i++) {
      pos[3 * i] = m.positions[i].x;
      pos[3 * i + 1] = m.positions[i].y;
      pos[3 * i + 2] = m.positions[i].z;
    }
  if (tris && tri_cap >= m.ntri()) std::memcpy(tris, m.tri.data(), m.tri.size() * 4);
  return 0;
}

int orc_x_crossing(const float* a, const float* b, const float* c, float y, float z, float* x) {
  return x_crossing(V{a[0], a[1], a[2]}, V{b[0], b[1], b[2]}, V{c[0], c[1], c[2]}, y, z, *x) ? 1 : 0;
}

void orc_closest_on_triangle(const float* p, const float* a, const float* b, const float* c, float* out) {
  const V q = closest_on_triangle(V{p[0], p[1], p[2]}, V{a[0], a[1], a[2]}, V{b[0], b[1], b[2]}, V{c[0], c[1], c[2]});
  out[0] = q.x;
  out[1] = q.y;
  out[2] = q.z;
}

// Squared nearest-triangle distances of n points: through the BVH (brute == 0, also counting
// box and triangle tests into counts[0..1]) or over every triangle (brute != 0).
int orc_mesh_nearest_sq(const float* pos, size_t npos, const uint32_t* tris, size_t ntri, const float* pts, size_t n,
                        int brute, float* out, uint64_t* counts) {
  const Mesh m = from_arrays(pos, npos, tris, ntri);
  std::string e;
  if (!validate(m, e)) return 1;
  Bvh bvh;
  if (!bvh.build(m)) return 1;
  uint64_t boxes = 0, tt = 0;
  for (size_t i = 0; i < n; i++) {
    const V p{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    out[i] = brute ? brute_nearest_sq(p, m) : bvh.nearest_sq(p, m, &boxes, &tt);
  }
  if (counts) {
    counts[0] = boxes;
    counts[1] = tt;
  }
  return 0;
}

// bake_sdf over z-slabs [k0, k1) (k1 == 0: every slab).  dims/lo/cell receive the full grid's
// layout; data receives (k1 - k0) * ny * nx samples when data_cap allows.
int orc_mesh_bake(const float* pos, size_t npos, const uint32_t* tris, size_t ntri, size_t resolution, float padding,
                  unsigned threads, size_t k0, size_t k1, uint32_t* dims, float* lo, float* cell, float* data,
                  size_t data_cap, char* err, size_t errlen) {
  const Mesh m = from_arrays(pos, npos, tris, ntri);
  Layout L{};
  std::vector<float> out;
  std::string e;
  if (!bake(m, resolution, padding, threads, k0, k1, L, out, e)) {
    put_err(err, errlen, e);
    return 1;
  }
  dims[0] = (uint32_t)L.nx;
  dims[1] = (uint32_t)L.ny;
  dims[2] = (uint32_t)L.nz;
  lo[0] = L.lo.x;
  lo[1] = L.lo.y;
  lo[2] = L.lo.z;
  cell[0] = L.cell.x;
  cell[1] = L.cell.y;
  cell[2] = L.cell.z;
  if (data) {
    if (data_cap < out.size()) return 2;
    std::memcpy(data, out.data(), out.size() * 4);
  }
  return 0;
}

}  // extern "C"
