// mesh.cpp — host side of the mesh -> SDF bake: Wavefront OBJ parsing and mesh validation
// (mm3e-orchestrator/src/mesh.rs:10-103), the bake's grid layout (mesh.rs:265-279), the
// nearest-triangle BVH build (mesh.rs:148-222), and the `.sdfv` volume codec
// (mm3e-kit/src/volume.rs:115-187, 229-248).  Every f32 operation that decides a result is
// written in the reference's order (built with -ffp-contract=off), so layouts and trees match.
#include "mesh.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "../../include/mm3e_hip.h"

namespace mm3e {

namespace {

void set_err_local(char* err, size_t n, const std::string& s) {
  if (err && n) std::snprintf(err, n, "%s", s.c_str());
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Rust's `str::parse::<f32>`: optional sign, decimal digits with optional fraction and exponent,
// or inf / infinity / nan; a value beyond the f32 range reads as +-inf (or 0), not an error.
bool parse_f32_tok(const std::string& t, float& v) {
  const char* s = t.c_str();
  const char* e = s + t.size();
  const char* b = s;
  if (b < e && *b == '+') b++;
  if (b == e) return false;
  float x = 0.0f;
  auto r = std::from_chars(b, e, x);
  if (r.ptr != e) return false;
  if (r.ec == std::errc::result_out_of_range) {
    x = std::strtof(b, nullptr);  // same accepted text; strtof saturates to inf / rounds to 0
  } else if (r.ec != std::errc()) {
    return false;
  }
  v = x;
  return true;
}

// Rust's `str::parse::<i64>`: optional sign then at least one decimal digit.
bool parse_i64_tok(const std::string& t, int64_t& v) {
  const char* s = t.c_str();
  const char* e = s + t.size();
  const char* b = s;
  if (b < e && *b == '+') b++;
  if (b == e || (*b == '-' && s != b)) return false;  // "+-1" is not a number
  auto r = std::from_chars(b, e, v);
  return r.ec == std::errc() && r.ptr == e;
}

struct V3f {
  float x, y, z;
};
V3f vmin(V3f a, V3f b) { return V3f{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
V3f vmax(V3f a, V3f b) { return V3f{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
V3f vsub(V3f a, V3f b) { return V3f{a.x - b.x, a.y - b.y, a.z - b.z}; }

V3f vert(const HostMesh& m, uint32_t i) { return V3f{m.pos[3 * (size_t)i], m.pos[3 * (size_t)i + 1], m.pos[3 * (size_t)i + 2]}; }

// Rust's `f32 as usize`: saturating, NaN -> 0.
uint64_t f32_to_usize(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 18446744073709551616.0f) return UINT64_MAX;
  return (uint64_t)x;
}

struct Item {
  uint32_t tri;
  V3f lo, hi, c;
};

uint32_t build_from(Item* items, size_t n, std::vector<BakeNode>& nodes, uint32_t* depth) {
  V3f lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; i++) {
    lo = vmin(lo, items[i].lo);
    hi = vmax(hi, items[i].hi);
  }
  BakeNode nd{};
  nd.lo[0] = lo.x; nd.lo[1] = lo.y; nd.lo[2] = lo.z;
  nd.hi[0] = hi.x; nd.hi[1] = hi.y; nd.hi[2] = hi.z;
  if (n == 1) {
    *depth = 0;
    nd.a = -(int32_t)items[0].tri - 1;
    nd.b = 0;
    nodes.push_back(nd);
    return (uint32_t)(nodes.size() - 1);
  }
  V3f clo{INFINITY, INFINITY, INFINITY}, chi{-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; i++) {
    clo = vmin(clo, items[i].c);
    chi = vmax(chi, items[i].c);
  }
  const V3f ext = vsub(chi, clo);
  const int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
  auto coord = [axis](const Item& it) { return axis == 0 ? it.c.x : (axis == 1 ? it.c.y : it.c.z); };
  // partial_cmp(..).unwrap_or(Equal) then the triangle index
  std::sort(items, items + n, [&](const Item& a, const Item& b) {
    const float ca = coord(a), cb = coord(b);
    if (ca < cb) return true;
    if (cb < ca) return false;
    return a.tri < b.tri;
  });
  const size_t mid = n / 2;
  uint32_t dl = 0, dr = 0;
  const uint32_t left = build_from(items, mid, nodes, &dl);
  const uint32_t right = build_from(items + mid, n - mid, nodes, &dr);
  *depth = 1 + std::max(dl, dr);
  nd.a = (int32_t)left;
  nd.b = right;
  nodes.push_back(nd);
  return (uint32_t)(nodes.size() - 1);
}

}  // namespace

bool validate_mesh(const HostMesh& m, std::string& err) {
  if (m.pos.empty()) {
    err = "mesh has no vertices";
    return false;
  }
  if (m.tris.empty()) {
    err = "mesh has no triangles";
    return false;
  }
  for (size_t i = 0; i < m.n_pos(); i++) {
    if (!(std::isfinite(m.pos[3 * i]) && std::isfinite(m.pos[3 * i + 1]) && std::isfinite(m.pos[3 * i + 2]))) {
      err = "mesh vertex " + std::to_string(i) + " is non-finite";
      return false;
    }
  }
  const uint64_t n = m.n_pos();
  for (size_t i = 0; i < m.n_tris(); i++) {
    for (int c = 0; c < 3; c++) {
      if (m.tris[3 * i + c] >= n) {
        err = "mesh triangle " + std::to_string(i) + " references a vertex out of range";
        return false;
      }
    }
  }
  return true;
}

bool parse_obj(const char* text, size_t len, HostMesh& m, std::string& err) {
  m.pos.clear();
  m.tris.clear();
  size_t at = 0, ln = 0;
  std::vector<std::string> toks;
  std::vector<uint32_t> idx;
  while (at < len) {
    // str::lines: split on '\n', drop one trailing '\r'
    size_t end = at;
    while (end < len && text[end] != '\n') end++;
    size_t stop = end;
    if (stop > at && text[stop - 1] == '\r') stop--;
    const size_t line_no = ++ln;
    // everything before '#', whitespace-split
    size_t cut = at;
    while (cut < stop && text[cut] != '#') cut++;
    toks.clear();
    for (size_t p = at; p < cut;) {
      while (p < cut && is_ws(text[p])) p++;
      size_t q = p;
      while (q < cut && !is_ws(text[q])) q++;
      if (q > p) toks.emplace_back(text + p, q - p);
      p = q;
    }
    at = end < len ? end + 1 : len;
    if (toks.empty()) continue;
    const std::string pre = "line " + std::to_string(line_no) + ": ";
    if (toks[0] == "v") {
      static const char* what[3] = {"x", "y", "z"};
      float v[3];
      for (int c = 0; c < 3; c++) {
        if ((size_t)c + 1 >= toks.size()) {
          err = pre + "missing " + what[c];
          return false;
        }
        if (!parse_f32_tok(toks[c + 1], v[c])) {
          err = pre + "bad number for " + what[c];
          return false;
        }
      }
      if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]))) {
        err = pre + "non-finite vertex";
        return false;
      }
      m.pos.insert(m.pos.end(), v, v + 3);
    } else if (toks[0] == "f") {
      idx.clear();
      for (size_t t = 1; t < toks.size(); t++) {
        const std::string& tok = toks[t];
        const std::string first = tok.substr(0, tok.find('/'));
        int64_t raw = 0;
        if (!parse_i64_tok(first, raw)) {
          err = pre + "bad face index '" + tok + "'";
          return false;
        }
        const int64_t n = (int64_t)m.n_pos();
        const int64_t resolved = raw > 0 ? raw - 1 : n + raw;
        if (resolved < 0 || resolved >= n) {
          err = pre + "face index " + std::to_string(raw) + " out of range (have " + std::to_string(n) + " vertices)";
          return false;
        }
        idx.push_back((uint32_t)resolved);
      }
      if (idx.size() < 3) {
        err = pre + "face needs at least 3 vertices";
        return false;
      }
      for (size_t t = 1; t + 1 < idx.size(); t++) {  // fan triangulation
        m.tris.push_back(idx[0]);
        m.tris.push_back(idx[t]);
        m.tris.push_back(idx[t + 1]);
      }
    }
  }
  return validate_mesh(m, err);
}

bool bake_layout(const HostMesh& m, uint32_t resolution, float padding, BakeLayout& out, std::string& err) {
  if (!validate_mesh(m, err)) return false;
  V3f mlo{INFINITY, INFINITY, INFINITY}, mhi{-INFINITY, -INFINITY, -INFINITY};  // Mesh::bounds
  for (size_t i = 0; i < m.n_pos(); i++) {
    const V3f p = vert(m, (uint32_t)i);
    mlo = vmin(mlo, p);
    mhi = vmax(mhi, p);
  }
  if (padding <= 0.0f || !std::isfinite(padding)) {
    err = "padding must be positive and finite";
    return false;
  }
  const uint32_t res = std::min(std::max(resolution, 2u), 512u);
  const V3f lo{mlo.x - padding, mlo.y - padding, mlo.z - padding};
  const V3f hi{mhi.x + padding, mhi.y + padding, mhi.z + padding};
  const V3f size = vsub(hi, lo);
  const float longest = std::fmax(std::fmax(std::fmax(size.x, size.y), size.z), 1e-6f);
  auto dim = [&](float s) -> uint64_t {
    const uint64_t d = f32_to_usize(std::round((s / longest) * (float)(res - 1))) + 1;
    return std::max<uint64_t>(d, 2);
  };
  const uint64_t n3[3] = {dim(size.x), dim(size.y), dim(size.z)};
  for (int a = 0; a < 3; a++) out.dims[a] = (uint32_t)std::min<uint64_t>(n3[a], 512);  // s <= longest
  out.lo[0] = lo.x; out.lo[1] = lo.y; out.lo[2] = lo.z;
  out.cell[0] = size.x / (float)(out.dims[0] - 1);
  out.cell[1] = size.y / (float)(out.dims[1] - 1);
  out.cell[2] = size.z / (float)(out.dims[2] - 1);
  return true;
}

bool check_layout(const BakeLayout& L, std::string& err) {
  for (int a = 0; a < 3; a++)
    if (!std::isfinite(L.lo[a])) {
      err = "SdfVolume min must be finite";
      return false;
    }
  for (int a = 0; a < 3; a++)
    if (!(std::isfinite(L.cell[a]) && L.cell[a] > 0.0f)) {
      err = "SdfVolume cell sizes must be positive and finite";
      return false;
    }
  return true;
}

bool build_tri_bvh(const HostMesh& m, TriBvhHost& out) {
  std::vector<Item> items;
  items.reserve(m.n_tris());
  for (size_t i = 0; i < m.n_tris(); i++) {
    const V3f a = vert(m, m.tris[3 * i]), b = vert(m, m.tris[3 * i + 1]), c = vert(m, m.tris[3 * i + 2]);
    const V3f ab = vsub(b, a), ac = vsub(c, a);
    const V3f cr{ab.y * ac.z - ab.z * ac.y, ab.z * ac.x - ab.x * ac.z, ab.x * ac.y - ab.y * ac.x};
    if (cr.x * cr.x + cr.y * cr.y + cr.z * cr.z < 1e-16f) continue;
    const float third = 1.0f / 3.0f;
    const V3f sum{(a.x + b.x) + c.x, (a.y + b.y) + c.y, (a.z + b.z) + c.z};
    items.push_back(Item{(uint32_t)i, vmin(vmin(a, b), c), vmax(vmax(a, b), c), V3f{sum.x * third, sum.y * third, sum.z * third}});
  }
  out.nodes.clear();
  if (items.empty()) return false;
  out.nodes.reserve(items.size() * 2);
  out.root = build_from(items.data(), items.size(), out.nodes, &out.depth);
  return true;
}

}  // namespace mm3e

using namespace mm3e;

namespace {
const char kSdfvMagic[8] = {'M', 'M', '3', 'E', 'S', 'D', 'F', '1'};
constexpr size_t kSdfvHeader = 8 + 3 * 4 + 3 * 4 + 3 * 4;

void put_u32(uint8_t* p, uint32_t v) {
  for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get_u32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
uint32_t f_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// SdfVolume::validate (volume.rs:33-85) on a layout plus n samples.
bool validate_volume(const mm3e_sdf_layout& L, const float* data, size_t n, std::string& err) {
  if (L.dims[0] < 2 || L.dims[1] < 2 || L.dims[2] < 2) {
    err = "SdfVolume dims must all be >= 2";
    return false;
  }
  for (int a = 0; a < 3; a++)
    if (!std::isfinite(L.min[a])) {
      err = "SdfVolume min must be finite";
      return false;
    }
  for (int a = 0; a < 3; a++)
    if (!(std::isfinite(L.cell[a]) && L.cell[a] > 0.0f)) {
      err = "SdfVolume cell sizes must be positive and finite";
      return false;
    }
  const uint64_t expected = (uint64_t)L.dims[0] * L.dims[1] * L.dims[2];
  if (n != expected) {
    err = "SdfVolume data length " + std::to_string(n) + " does not match dims " + std::to_string(L.dims[0]) + "x" +
          std::to_string(L.dims[1]) + "x" + std::to_string(L.dims[2]);
    return false;
  }
  for (size_t i = 0; data && i < n; i++)
    if (!std::isfinite(data[i])) {
      err = "SdfVolume sample " + std::to_string(i) + " is non-finite";
      return false;
    }
  return true;
}
}  // namespace

extern "C" {

int mm3e_mesh_parse_obj(const char* text, size_t len, float* positions, size_t pos_cap, uint32_t* triangles,
                        size_t tri_cap, size_t* n_positions, size_t* n_triangles, char* err, size_t errlen) {
  if (!text && len) return MM3E_ERR_INVALID;
  HostMesh m;
  std::string msg;
  if (!parse_obj(text, len, m, msg)) {
    set_err_local(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  if (n_positions) *n_positions = m.n_pos();
  if (n_triangles) *n_triangles = m.n_tris();
  if (!positions && !triangles) return MM3E_OK;  // size query
  if (!positions || !triangles || pos_cap < m.n_pos() || tri_cap < m.n_tris()) {
    set_err_local(err, errlen, "mm3e_mesh_parse_obj: output arrays too small");
    return MM3E_ERR_INVALID;
  }
  std::memcpy(positions, m.pos.data(), m.pos.size() * 4);
  std::memcpy(triangles, m.tris.data(), m.tris.size() * 4);
  return MM3E_OK;
}

int mm3e_mesh_sdf_layout(const float* positions, size_t n_positions, const uint32_t* triangles, size_t n_triangles,
                         uint32_t resolution, float padding, mm3e_sdf_layout* layout, char* err, size_t errlen) {
  if (!layout || (!positions && n_positions) || (!triangles && n_triangles)) return MM3E_ERR_INVALID;
  HostMesh m;
  m.pos.assign(positions, positions + n_positions * 3);
  m.tris.assign(triangles, triangles + n_triangles * 3);
  BakeLayout L;
  TriBvhHost bvh;
  std::string msg;
  if (!bake_layout(m, resolution, padding, L, msg)) {
    set_err_local(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  if (!build_tri_bvh(m, bvh)) {
    set_err_local(err, errlen, "mesh has no non-degenerate triangles");
    return MM3E_ERR_INVALID;
  }
  if (!check_layout(L, msg)) {
    set_err_local(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  for (int a = 0; a < 3; a++) {
    layout->dims[a] = L.dims[a];
    layout->min[a] = L.lo[a];
    layout->cell[a] = L.cell[a];
  }
  return MM3E_OK;
}

int mm3e_sdfv_encode(const mm3e_sdf_layout* layout, const float* data, size_t n, uint8_t* out, size_t* out_len,
                     char* err, size_t errlen) {
  if (!layout || !out_len || (!data && n)) return MM3E_ERR_INVALID;
  std::string msg;
  if (!validate_volume(*layout, data, n, msg)) {
    set_err_local(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  const size_t need = kSdfvHeader + n * 4;
  if (!out || *out_len < need) {
    *out_len = need;
    return MM3E_ERR_INVALID;
  }
  std::memcpy(out, kSdfvMagic, 8);
  uint8_t* p = out + 8;
  for (int a = 0; a < 3; a++, p += 4) put_u32(p, layout->dims[a]);
  for (int a = 0; a < 3; a++, p += 4) put_u32(p, f_bits(layout->min[a]));
  for (int a = 0; a < 3; a++, p += 4) put_u32(p, f_bits(layout->cell[a]));
  for (size_t i = 0; i < n; i++, p += 4) put_u32(p, f_bits(data[i]));
  *out_len = need;
  return MM3E_OK;
}

int mm3e_sdfv_decode(const uint8_t* bytes, size_t len, mm3e_sdf_layout* layout, float* data, size_t data_cap,
                     char* err, size_t errlen) {
  if (!layout || (!bytes && len)) return MM3E_ERR_INVALID;
  auto fail = [&](const std::string& m) {
    set_err_local(err, errlen, m);
    return MM3E_ERR_INVALID;
  };
  if (len < kSdfvHeader) return fail("sdfv buffer is shorter than the header");
  if (std::memcmp(bytes, kSdfvMagic, 8) != 0) return fail("sdfv magic/version mismatch");
  mm3e_sdf_layout L;
  const uint8_t* p = bytes + 8;
  for (int a = 0; a < 3; a++, p += 4) L.dims[a] = get_u32(p);
  static const char* mn[3] = {"min.x", "min.y", "min.z"};
  static const char* cl[3] = {"cell.x", "cell.y", "cell.z"};
  for (int a = 0; a < 3; a++, p += 4) {
    L.min[a] = bits_f(get_u32(p));
    if (!std::isfinite(L.min[a])) return fail(std::string("sdfv non-finite f32 while reading ") + mn[a]);
  }
  for (int a = 0; a < 3; a++, p += 4) {
    L.cell[a] = bits_f(get_u32(p));
    if (!std::isfinite(L.cell[a])) return fail(std::string("sdfv non-finite f32 while reading ") + cl[a]);
  }
  // the usize products of volume.rs:161-170 (checked count, checked byte length, then the
  // header added)
  const unsigned __int128 samples = (unsigned __int128)L.dims[0] * L.dims[1] * L.dims[2];
  if (samples > (unsigned __int128)SIZE_MAX) return fail("sdfv dimensions overflow");
  if (samples * 4 > (unsigned __int128)SIZE_MAX) return fail("sdfv payload length overflows");
  const size_t expected = (size_t)(kSdfvHeader + (uint64_t)samples * 4);
  if (len != expected)
    return fail("sdfv byte length " + std::to_string(len) + " does not match expected " + std::to_string(expected));
  const size_t n = (size_t)samples;
  for (size_t i = 0; i < n; i++) {
    const float v = bits_f(get_u32(p + 4 * i));
    if (!std::isfinite(v)) return fail("sdfv non-finite f32 while reading sample " + std::to_string(i));
  }
  std::string msg;
  if (!validate_volume(L, nullptr, n, msg)) return fail(msg);
  *layout = L;
  if (!data) return MM3E_OK;  // layout query
  if (data_cap < n) return fail("mm3e_sdfv_decode: data buffer too small");
  for (size_t i = 0; i < n; i++) data[i] = bits_f(get_u32(p + 4 * i));
  return MM3E_OK;
}

}  // extern "C"
