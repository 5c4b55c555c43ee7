// scene_text.cpp — the reference's `.mm3e` text scene format (mm3e-orchestrator/src/scene_io.rs)
// as a wire format for this library: parse text into the flat scene IR plus a camera, and
// write a scene IR plus camera back to text.  Line-oriented: a directive keyword followed by
// whitespace-separated values, `#` starts a comment.  Numbers print as the shortest decimal
// that reads back to the same f32 (Rust's `{}` for f32), so text -> IR -> text is stable.
// The reference's format drops CSG, volumes, GI, the shadow/AO budgets, LOD, the march boosts
// and normal_h (scene_io.rs:47-62, 103-105).  This library extends it losslessly with
// directives the reference's parser does not know, written only when a scene needs them (a
// scene inside the reference's subset is written exactly in the reference's format):
//   sun_raw x y z                       sun direction as stored (when re-normalising changes it)
//   marcher_ext shadow_steps ao_samples lod_footprint subitize secant normal_h
//   gi bake nx ny nz samples            Scene::bake_gi request (lib.rs:665-711)
//   gi baked nx ny nz minx miny minz sx sy sz <cubes>   a baked GiVolume, [probe][6][3] f32
//   volume sdfv <bytes>                 an SdfVolume as its `.sdfv` file (volume.rs:115-187)
//   volume raw nx ny nz minx miny minz cx cy cz <samples>   a volume .sdfv would reject
//   csgnode op prim a b p0..p7 r0..r8 px py pz scale   one csg::Expr node (csg.rs:17-91)
//   csgexpr root                        one registered expression (Scene::csg)
//   obj csg <id> ... / obj volume <id> ...   Prim::Csg / Prim::Volume objects
// <bytes>, <cubes>, <samples>: base64 (RFC 4648) of the little-endian bytes.  Text -> IR is
// then bit-exact for every IR the library accepts: IR -> text -> IR returns the same bytes.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mm3e_hip.h"
#include "scene_compile.h"

using namespace mm3e;

namespace {

void set_err(char* err, size_t n, const std::string& s) {
  if (err && n) std::snprintf(err, n, "%s", s.c_str());
}

// ---- numbers ---------------------------------------------------------------------------
std::string fmt(float x) {  // f32 Display: shortest round-trip decimal, never exponent form
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-inf" : "inf";
  char b[128];
  auto r = std::to_chars(b, b + sizeof b, x, std::chars_format::fixed);
  return std::string(b, r.ptr);
}
std::string fmt3(const float* v) { return fmt(v[0]) + " " + fmt(v[1]) + " " + fmt(v[2]); }

bool parse_f32(const std::string& t, float& v) {
  const char* s = t.c_str();
  const char* e = s + t.size();
  if (s < e && *s == '+') s++;  // Rust's f32 parser accepts one leading '+'
  auto r = std::from_chars(s, e, v);
  return r.ec == std::errc() && r.ptr == e && s != e;
}
bool parse_u32(const std::string& t, uint32_t& v) {
  const char* s = t.c_str();
  const char* e = s + t.size();
  if (s < e && *s == '+') s++;
  auto r = std::from_chars(s, e, v);
  return r.ec == std::errc() && r.ptr == e && s != e;
}

// ---- base64 (RFC 4648 with padding) for the binary payloads of the extension ------------
const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
std::string b64_encode(const uint8_t* p, size_t n) {
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
    out += kB64[v >> 18];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += kB64[v & 63];
  }
  if (i < n) {
    const uint32_t v = (uint32_t)p[i] << 16 | (i + 1 < n ? (uint32_t)p[i + 1] << 8 : 0u);
    out += kB64[v >> 18];
    out += kB64[(v >> 12) & 63];
    out += i + 1 < n ? kB64[(v >> 6) & 63] : '=';
    out += '=';
  }
  return out;
}
bool b64_decode(const std::string& s, std::vector<uint8_t>& out) {
  if (s.size() % 4) return false;
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return -1;
  };
  out.clear();
  out.reserve(s.size() / 4 * 3);
  for (size_t i = 0; i < s.size(); i += 4) {
    const bool last = i + 4 == s.size();
    const int a = val(s[i]), b = val(s[i + 1]);
    const int c = (last && s[i + 2] == '=') ? 0 : val(s[i + 2]);
    const int d = (last && s[i + 3] == '=') ? 0 : val(s[i + 3]);
    if (a < 0 || b < 0 || c < 0 || d < 0) return false;
    if (!last && (s[i + 2] == '=' || s[i + 3] == '=')) return false;
    if (last && s[i + 2] == '=' && s[i + 3] != '=') return false;
    const uint32_t v = (uint32_t)a << 18 | (uint32_t)b << 12 | (uint32_t)c << 6 | (uint32_t)d;
    out.push_back((uint8_t)(v >> 16));
    if (!(last && s[i + 2] == '=')) out.push_back((uint8_t)(v >> 8));
    if (!(last && s[i + 3] == '=')) out.push_back((uint8_t)v);
  }
  return true;
}

// ---- Vec3 helpers in f32, in the reference's operation order (vec.rs) ---------------------
struct F3 {
  float x, y, z;
};
F3 normalize(F3 a) {  // vec.rs:37-44
  const float len = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  if (len > 1e-12f) {
    const float k = 1.0f / len;
    return F3{a.x * k, a.y * k, a.z * k};
  }
  return F3{0, 0, 0};
}
bool is_zero(F3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
F3 cross(F3 a, F3 o) { return F3{a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x}; }

// ---- token cursor (scene_io.rs:168-198) --------------------------------------------------
struct Cur {
  std::vector<std::string> toks;
  size_t i = 0;
  size_t line = 0;
  std::string err;
  bool next(const char* what, std::string& out) {
    if (i >= toks.size()) {
      err = "line " + std::to_string(line) + ": expected " + what;
      return false;
    }
    out = toks[i++];
    return true;
  }
  bool f(const char* what, float& v) {
    std::string t;
    if (!next(what, t)) return false;
    if (!parse_f32(t, v)) {
      err = "line " + std::to_string(line) + ": bad number '" + t + "' for " + what;
      return false;
    }
    if (!std::isfinite(v)) {
      err = "line " + std::to_string(line) + ": non-finite number '" + t + "' for " + what;
      return false;
    }
    return true;
  }
  bool u(const char* what, uint32_t& v) {
    std::string t;
    if (!next(what, t)) return false;
    if (!parse_u32(t, v)) {
      err = "line " + std::to_string(line) + ": bad integer '" + t + "' for " + what;
      return false;
    }
    return true;
  }
  bool b(const char* what, uint32_t& v) {
    uint32_t x;
    if (!u(what, x)) return false;
    v = x != 0 ? 1u : 0u;
    return true;
  }
  bool v3(const char* what, float* v) { return f(what, v[0]) && f(what, v[1]) && f(what, v[2]); }
};

const char* kModeNames[6] = {"beauty", "normal", "depth", "ao", "steps", "albedo"};

// Scene::new (lib.rs:595-620) with Marcher::default (march.rs:81-96) and Post::default
// (post.rs:25-29).
void scene_defaults(mm3e_ir_header& h, uint32_t w, uint32_t hgt) {
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, MM3E_IR_MAGIC, 8);
  h.version = MM3E_IR_VERSION;
  h.width = w;
  h.height = hgt;
  h.aa = 2;
  h.bounces = 1;
  h.shadows = 1;
  h.ao = 1;
  const F3 sun = normalize(F3{0.6f, 0.7f, 0.4f});
  h.sun_dir[0] = sun.x;
  h.sun_dir[1] = sun.y;
  h.sun_dir[2] = sun.z;
  for (int a = 0; a < 3; a++) {
    h.ambient[a] = 0.02f;
    h.sky_ambient[a] = 0.4f;
  }
  h.fog[0] = 0.62f;
  h.fog[1] = 0.72f;
  h.fog[2] = 0.86f;
  h.fog_density = 0.012f;
  h.max_steps = 160;
  h.max_dist = 120.0f;
  h.eps = 0.0006f;
  h.step_scale = 1.0f;
  h.shadow_steps = 64;
  h.ao_samples = 5;
  h.normal_h = 0.0009f;
  h.exposure = 1.0f;
  h.bloom = 1;
  h.bloom_threshold = 1.0f;
  h.bloom_intensity = 0.06f;
  h.bloom_radius = 12;
}

mm3e_ir_material default_material() {  // color.rs:47-60
  mm3e_ir_material m;
  std::memset(&m, 0, sizeof m);
  for (int a = 0; a < 3; a++) m.albedo[a] = 0.8f;
  m.specular = 0.5f;
  m.roughness = 0.35f;
  return m;
}

// The parsed scene before it is laid out as IR.
struct TextScene {
  mm3e_ir_header h;
  std::vector<mm3e_ir_object> objs;
  std::vector<mm3e_ir_material> mats;
  std::vector<mm3e_ir_light> lights;
  std::vector<mm3e_ir_volume> vols;
  std::vector<float> vdata;
  std::vector<mm3e_ir_csg_node> csg;
  std::vector<uint32_t> roots;
  std::vector<float> cubes;
};

// IR sections in a fixed order: header, objects, materials, lights, volumes, volume samples,
// CSG nodes, CSG roots, GI cubes.
bool write_ir(TextScene& t, std::vector<uint8_t>& out, std::string& err) {
  mm3e_ir_header& h = t.h;
  h.n_objects = (uint32_t)t.objs.size();
  h.n_materials = (uint32_t)t.mats.size();
  h.n_lights = (uint32_t)t.lights.size();
  h.n_volumes = (uint32_t)t.vols.size();
  h.n_volume_floats = (uint32_t)t.vdata.size();
  h.n_csg_nodes = (uint32_t)t.csg.size();
  h.n_csg_exprs = (uint32_t)t.roots.size();
  uint64_t off = sizeof(mm3e_ir_header);
  auto place = [&](uint32_t& field, uint64_t bytes) {
    field = (uint32_t)off;
    off += bytes;
  };
  place(h.off_objects, t.objs.size() * sizeof(mm3e_ir_object));
  place(h.off_materials, t.mats.size() * sizeof(mm3e_ir_material));
  place(h.off_lights, t.lights.size() * sizeof(mm3e_ir_light));
  place(h.off_volumes, t.vols.size() * sizeof(mm3e_ir_volume));
  place(h.off_volume_data, t.vdata.size() * 4);
  place(h.off_csg_nodes, t.csg.size() * sizeof(mm3e_ir_csg_node));
  place(h.off_csg_roots, t.roots.size() * 4);
  place(h.off_gi_cubes, t.cubes.size() * 4);
  if (off > 0xFFFFFFFFull) {
    err = "scene too large for the IR (more than 4 GiB)";
    return false;
  }
  h.total_bytes = (uint32_t)off;
  out.assign(off, 0);
  std::memcpy(out.data(), &h, sizeof h);
  auto put = [&](uint32_t at, const void* src, size_t bytes) {
    if (bytes) std::memcpy(out.data() + at, src, bytes);
  };
  put(h.off_objects, t.objs.data(), t.objs.size() * sizeof(mm3e_ir_object));
  put(h.off_materials, t.mats.data(), t.mats.size() * sizeof(mm3e_ir_material));
  put(h.off_lights, t.lights.data(), t.lights.size() * sizeof(mm3e_ir_light));
  put(h.off_volumes, t.vols.data(), t.vols.size() * sizeof(mm3e_ir_volume));
  put(h.off_volume_data, t.vdata.data(), t.vdata.size() * 4);
  put(h.off_csg_nodes, t.csg.data(), t.csg.size() * sizeof(mm3e_ir_csg_node));
  put(h.off_csg_roots, t.roots.data(), t.roots.size() * 4);
  put(h.off_gi_cubes, t.cubes.data(), t.cubes.size() * 4);
  return true;
}

bool positive3(const float* v) { return v[0] > 0.0f && v[1] > 0.0f && v[2] > 0.0f; }

// scene_io.rs:334-357 validate_prim
bool validate_prim(size_t line, const mm3e_ir_object& o, std::string& err) {
  const float* p = o.p;
  auto bad = [&](const char* m) {
    err = "line " + std::to_string(line) + ": " + m;
    return false;
  };
  switch (o.prim) {
    case MM3E_PRIM_SPHERE:
      if (p[0] <= 0.0f) return bad("sphere radius must be positive");
      break;
    case MM3E_PRIM_BOX:
      if (!positive3(p)) return bad("box half extents must be positive");
      break;
    case MM3E_PRIM_ROUND_BOX:
      if (!positive3(p) || p[3] < 0.0f) return bad("roundbox half extents must be positive and radius non-negative");
      break;
    case MM3E_PRIM_TORUS:
      if (p[0] <= 0.0f || p[1] <= 0.0f) return bad("torus radii must be positive");
      break;
    case MM3E_PRIM_CYLINDER:
      if (p[0] <= 0.0f || p[1] <= 0.0f) return bad("cylinder height and radius must be positive");
      break;
    case MM3E_PRIM_CAPSULE: {
      const float dx = p[3] - p[0], dy = p[4] - p[1], dz = p[5] - p[2];
      if (p[6] <= 0.0f || std::sqrt(dx * dx + dy * dy + dz * dz) <= 1e-6f)
        return bad("capsule radius must be positive and endpoints must differ");
      break;
    }
    case MM3E_PRIM_CONE:
      if (p[0] < 0.0f || p[1] < 0.0f || p[2] <= 0.0f || (p[0] == 0.0f && p[1] == 0.0f))
        return bad("cone height must be positive and at least one radius must be positive");
      break;
    case MM3E_PRIM_ELLIPSOID:
      if (!positive3(p)) return bad("ellipsoid radii must be positive");
      break;
    case MM3E_PRIM_OCTAHEDRON:
      if (p[0] <= 0.0f) return bad("octahedron size must be positive");
      break;
    case MM3E_PRIM_HEX_PRISM:
      if (p[0] <= 0.0f || p[1] <= 0.0f) return bad("hex prism radius and height must be positive");
      break;
    case MM3E_PRIM_PLANE:
      if (is_zero(normalize(F3{p[0], p[1], p[2]}))) return bad("plane normal must be non-zero");
      break;
    default:
      break;
  }
  return true;
}

// scene_io.rs:359-413 parse_object
bool parse_object(Cur& c, mm3e_ir_object& o) {
  std::memset(&o, 0, sizeof o);
  std::string name;
  if (!c.next("primitive", name)) return false;
  float* p = o.p;
  bool ok = true;
  if (name == "sphere") { o.prim = MM3E_PRIM_SPHERE; ok = c.f("r", p[0]); }
  else if (name == "box") { o.prim = MM3E_PRIM_BOX; ok = c.v3("half", p); }
  else if (name == "roundbox") { o.prim = MM3E_PRIM_ROUND_BOX; ok = c.v3("half", p) && c.f("radius", p[3]); }
  else if (name == "torus") { o.prim = MM3E_PRIM_TORUS; ok = c.f("major", p[0]) && c.f("minor", p[1]); }
  else if (name == "cylinder") { o.prim = MM3E_PRIM_CYLINDER; ok = c.f("h", p[0]) && c.f("r", p[1]); }
  else if (name == "capsule") { o.prim = MM3E_PRIM_CAPSULE; ok = c.v3("a", p) && c.v3("b", p + 3) && c.f("r", p[6]); }
  else if (name == "cone") { o.prim = MM3E_PRIM_CONE; ok = c.f("r1", p[0]) && c.f("r2", p[1]) && c.f("h", p[2]); }
  else if (name == "ellipsoid") { o.prim = MM3E_PRIM_ELLIPSOID; ok = c.v3("r", p); }
  else if (name == "octahedron") { o.prim = MM3E_PRIM_OCTAHEDRON; ok = c.f("s", p[0]); }
  else if (name == "hexprism") { o.prim = MM3E_PRIM_HEX_PRISM; ok = c.f("r", p[0]) && c.f("h", p[1]); }
  else if (name == "plane") { o.prim = MM3E_PRIM_PLANE; ok = c.v3("n", p) && c.f("h", p[3]); }
  else if (name == "csg" || name == "volume") {  // extension: Prim::Csg { id } / Prim::Volume { id }
    uint32_t id = 0;
    o.prim = name == "csg" ? MM3E_PRIM_CSG : MM3E_PRIM_VOLUME;
    ok = c.u(name == "csg" ? "csg id" : "volume id", id);
    std::memcpy(&p[0], &id, 4);
  } else {
    c.err = "line " + std::to_string(c.line) + ": unknown primitive '" + name + "'";
    return false;
  }
  if (!ok) return false;
  if (!validate_prim(c.line, o, c.err)) return false;
  // Transform::IDENTITY, material 0, Combine::Union, no modifiers
  o.rot[0] = o.rot[4] = o.rot[8] = 1.0f;
  o.scale = 1.0f;
  std::string kw;
  while (c.i < c.toks.size()) {
    kw = c.toks[c.i++];
    if (kw == "pos") ok = c.v3("pos", o.pos);
    else if (kw == "basis") ok = c.v3("col0", o.rot) && c.v3("col1", o.rot + 3) && c.v3("col2", o.rot + 6);
    else if (kw == "scale") {
      ok = c.f("scale", o.scale);
      if (ok && o.scale <= 0.0f) {
        c.err = "line " + std::to_string(c.line) + ": object scale must be positive";
        return false;
      }
    } else if (kw == "mat") ok = c.u("mat", o.mat);
    else if (kw == "combine") {
      std::string k;
      if (!c.next("combine kind", k)) return false;
      if (k == "union") o.combine = MM3E_COMBINE_UNION;
      else if (k == "smooth") { o.combine = MM3E_COMBINE_SMOOTH; ok = c.f("k", o.k); }
      else if (k == "subtract") o.combine = MM3E_COMBINE_SUBTRACT;
      else {
        c.err = "line " + std::to_string(c.line) + ": unknown combine '" + k + "'";
        return false;
      }
    } else if (kw == "round") ok = c.f("round", o.round);
    else if (kw == "onion") ok = c.f("onion", o.onion);
    else if (kw == "twist") ok = c.f("twist", o.twist);
    else if (kw == "bend") ok = c.f("bend", o.bend);
    else if (kw == "mirror") {
      uint32_t mx = 0, my = 0, mz = 0;
      ok = c.b("mx", mx) && c.b("my", my) && c.b("mz", mz);
      o.mirror = mx | (my << 1) | (mz << 2);
    } else if (kw == "repeat") ok = c.v3("repeat", o.repeat);
    else if (kw == "elongate") ok = c.v3("elongate", o.elongate);
    else {
      c.err = "line " + std::to_string(c.line) + ": unknown object field '" + kw + "'";
      return false;
    }
    if (!ok) return false;
  }
  return true;
}

// scene_io.rs:200-331 parse
bool parse_text(const std::string& src, std::vector<uint8_t>& ir, mm3e_camera& cam, std::string& err) {
  TextScene ts;
  mm3e_ir_header& h = ts.h;
  scene_defaults(h, 640, 360);
  std::vector<mm3e_ir_object>& objs = ts.objs;
  std::vector<mm3e_ir_material>& mats = ts.mats;  // Scene::new's default material is cleared
  std::vector<mm3e_ir_light>& lights = ts.lights;
  {
    const float eye[3] = {0.0f, 1.0f, 6.0f}, target[3] = {0, 0, 0}, up[3] = {0.0f, 1.0f, 0.0f};
    mm3e_camera_look_at(eye, target, up, 1.0f, &cam);
  }
  size_t pos = 0, lineno = 0;
  while (pos <= src.size()) {
    size_t nl = src.find('\n', pos);
    if (nl == std::string::npos) nl = src.size();
    std::string raw = src.substr(pos, nl - pos);
    pos = nl + 1;
    lineno++;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();  // str::lines drops a trailing \r
    const size_t hash = raw.find('#');
    if (hash != std::string::npos) raw.resize(hash);
    Cur c;
    c.line = lineno;
    {
      size_t i = 0;
      while (i < raw.size()) {
        while (i < raw.size() && std::isspace((unsigned char)raw[i])) i++;
        size_t j = i;
        while (j < raw.size() && !std::isspace((unsigned char)raw[j])) j++;
        if (j > i) c.toks.push_back(raw.substr(i, j - i));
        i = j;
      }
    }
    if (c.toks.empty()) {
      if (nl == src.size()) break;
      continue;
    }
    std::string d;
    c.next("directive", d);
    bool ok = true;
    const std::string L = "line " + std::to_string(lineno) + ": ";
    if (d == "size") {
      ok = c.u("width", h.width) && c.u("height", h.height);
      if (ok && (h.width == 0 || h.height == 0 || h.width > 16384 || h.height > 16384)) {
        err = L + "size " + std::to_string(h.width) + "x" + std::to_string(h.height) + " out of range 1..=16384";
        return false;
      }
    } else if (d == "aa") ok = c.u("aa", h.aa);
    else if (d == "bounces") ok = c.u("bounces", h.bounces);
    else if (d == "shadows") ok = c.b("shadows", h.shadows);
    else if (d == "ao") ok = c.b("ao", h.ao);
    else if (d == "sun") {
      float v[3];
      ok = c.v3("sun", v);
      if (ok) {
        const F3 n = normalize(F3{v[0], v[1], v[2]});
        h.sun_dir[0] = n.x;
        h.sun_dir[1] = n.y;
        h.sun_dir[2] = n.z;
        if (is_zero(n)) {
          err = L + "sun direction must be non-zero";
          return false;
        }
      }
    } else if (d == "ambient") ok = c.v3("ambient", h.ambient);
    else if (d == "sky_ambient") ok = c.v3("sky_ambient", h.sky_ambient);
    else if (d == "fog") ok = c.v3("fog", h.fog) && c.f("fog density", h.fog_density);
    else if (d == "marcher") {
      ok = c.u("max_steps", h.max_steps) && c.f("max_dist", h.max_dist) && c.f("eps", h.eps) &&
           c.f("step_scale", h.step_scale);
      if (ok && (h.max_steps == 0 || h.max_steps > 100000)) {
        err = L + "marcher max_steps out of range 1..=100000";
        return false;
      }
      if (ok && !(h.max_dist > 0.0f && h.eps > 0.0f && h.step_scale > 0.0f)) {
        err = L + "marcher max_dist/eps/step_scale must be positive";
        return false;
      }
    } else if (d == "post") {
      ok = c.f("exposure", h.exposure) && c.b("bloom", h.bloom) && c.f("bloom_threshold", h.bloom_threshold) &&
           c.f("bloom_intensity", h.bloom_intensity) && c.u("bloom_radius", h.bloom_radius);
    } else if (d == "mode") {
      std::string m;
      ok = c.next("mode", m);
      if (ok) {
        int found = -1;
        for (int k = 0; k < 6; k++)
          if (m == kModeNames[k]) found = k;
        if (found < 0) {
          err = L + "unknown render mode '" + m + "'";
          return false;
        }
        h.mode = (uint32_t)found;
      }
    } else if (d == "cam") {
      float eye[3], target[3], up[3], fov_deg;
      ok = c.v3("eye", eye) && c.v3("target", target) && c.v3("up", up) && c.f("fov", fov_deg);
      if (ok) {
        const F3 fwd = normalize(F3{target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]});
        if (is_zero(fwd)) {
          err = L + "cam eye and target must not coincide";
          return false;
        }
        if (is_zero(normalize(cross(fwd, F3{up[0], up[1], up[2]})))) {
          err = L + "cam up must not be parallel to the view direction";
          return false;
        }
        if (!(fov_deg > 0.0f) || fov_deg >= 180.0f) {
          err = L + "cam fov must be in (0, 180) degrees";
          return false;
        }
        // f32::to_radians: x * (PI / 180) with the constant rounded to f32
        mm3e_camera_look_at(eye, target, up, fov_deg * (3.14159265358979323846f / 180.0f), &cam);
      }
    } else if (d == "mat") {
      mm3e_ir_material m;
      std::memset(&m, 0, sizeof m);
      ok = c.v3("albedo", m.albedo) && c.f("metallic", m.metallic) && c.f("roughness", m.roughness) &&
           c.f("reflectivity", m.reflectivity) && c.f("specular", m.specular) && c.v3("emissive", m.emissive) &&
           c.b("checker", m.checker);
      if (ok) mats.push_back(m);
    } else if (d == "light") {
      std::string kind;
      mm3e_ir_light l;
      std::memset(&l, 0, sizeof l);
      ok = c.next("light kind", kind) && c.v3("light vec", l.vec) && c.v3("light color", l.color) &&
           c.f("light radius", l.radius);
      if (ok) {
        if (kind == "dir") l.directional = 1;
        else if (kind == "point") l.directional = 0;
        else {
          err = L + "unknown light kind '" + kind + "'";
          return false;
        }
        lights.push_back(l);
      }
    } else if (d == "obj") {
      mm3e_ir_object o;
      ok = parse_object(c, o);
      if (ok) objs.push_back(o);
    } else if (d == "sun_raw") {
      ok = c.v3("sun", h.sun_dir);
    } else if (d == "marcher_ext") {
      ok = c.u("shadow_steps", h.shadow_steps) && c.u("ao_samples", h.ao_samples) &&
           c.f("lod_footprint", h.lod_footprint) && c.f("subitize", h.subitize) && c.b("secant", h.secant) &&
           c.f("normal_h", h.normal_h);
    } else if (d == "gi") {
      std::string kind;
      ok = c.next("gi kind", kind);
      if (ok && kind == "bake") {
        h.gi_mode = MM3E_GI_BAKE;
        ok = c.u("gi nx", h.gi_dims[0]) && c.u("gi ny", h.gi_dims[1]) && c.u("gi nz", h.gi_dims[2]) &&
             c.u("gi samples", h.gi_samples);
      } else if (ok && kind == "baked") {
        h.gi_mode = MM3E_GI_BAKED;
        h.gi_samples = 0;
        std::string b64;
        std::vector<uint8_t> bytes;
        ok = c.u("gi nx", h.gi_dims[0]) && c.u("gi ny", h.gi_dims[1]) && c.u("gi nz", h.gi_dims[2]) &&
             c.v3("gi min", h.gi_min) && c.v3("gi size", h.gi_size) && c.next("gi cubes", b64);
        if (ok) {
          const unsigned __int128 want = (unsigned __int128)h.gi_dims[0] * h.gi_dims[1] * h.gi_dims[2] * 18 * 4;
          if (!b64_decode(b64, bytes) || (unsigned __int128)bytes.size() != want) {
            err = L + "gi cubes must be base64 of nx*ny*nz*18 f32";
            return false;
          }
          ts.cubes.resize(bytes.size() / 4);
          std::memcpy(ts.cubes.data(), bytes.data(), bytes.size());
        }
      } else if (ok) {
        err = L + "unknown gi kind '" + kind + "'";
        return false;
      }
    } else if (d == "volume") {
      std::string kind, b64;
      std::vector<uint8_t> bytes;
      mm3e_ir_volume v;
      std::memset(&v, 0, sizeof v);
      ok = c.next("volume kind", kind);
      if (ok && kind == "sdfv") {
        ok = c.next("sdfv bytes", b64);
        if (ok) {
          if (!b64_decode(b64, bytes)) {
            err = L + "volume sdfv bytes are not base64";
            return false;
          }
          mm3e_sdf_layout lay;
          char e2[256];
          if (mm3e_sdfv_decode(bytes.data(), bytes.size(), &lay, nullptr, 0, e2, sizeof e2) != MM3E_OK) {
            err = L + e2;
            return false;
          }
          const size_t n = (size_t)lay.dims[0] * lay.dims[1] * lay.dims[2];
          const size_t at = ts.vdata.size();
          ts.vdata.resize(at + n);
          if (mm3e_sdfv_decode(bytes.data(), bytes.size(), &lay, ts.vdata.data() + at, n, e2, sizeof e2) != MM3E_OK) {
            err = L + e2;
            return false;
          }
          std::memcpy(v.dims, lay.dims, 12);
          std::memcpy(v.min, lay.min, 12);
          std::memcpy(v.cell, lay.cell, 12);
          v.offset = (uint32_t)at;
          ts.vols.push_back(v);
        }
      } else if (ok && kind == "raw") {
        ok = c.u("nx", v.dims[0]) && c.u("ny", v.dims[1]) && c.u("nz", v.dims[2]) && c.v3("volume min", v.min) &&
             c.v3("volume cell", v.cell) && c.next("volume samples", b64);
        if (ok) {
          const unsigned __int128 want = (unsigned __int128)v.dims[0] * v.dims[1] * v.dims[2] * 4;
          if (!b64_decode(b64, bytes) || (unsigned __int128)bytes.size() != want) {
            err = L + "volume samples must be base64 of nx*ny*nz f32";
            return false;
          }
          v.offset = (uint32_t)ts.vdata.size();
          ts.vdata.resize(ts.vdata.size() + bytes.size() / 4);
          std::memcpy(ts.vdata.data() + v.offset, bytes.data(), bytes.size());
          ts.vols.push_back(v);
        }
      } else if (ok) {
        err = L + "unknown volume kind '" + kind + "'";
        return false;
      }
    } else if (d == "csgnode") {
      mm3e_ir_csg_node n;
      std::memset(&n, 0, sizeof n);
      ok = c.u("csg op", n.op) && c.u("csg prim", n.prim) && c.u("csg a", n.a) && c.u("csg b", n.b);
      for (int k = 0; ok && k < 8; k++) ok = c.f("csg p", n.p[k]);
      for (int k = 0; ok && k < 9; k++) ok = c.f("csg rot", n.rot[k]);
      ok = ok && c.v3("csg pos", n.pos) && c.f("csg scale", n.scale);
      if (ok) ts.csg.push_back(n);
    } else if (d == "csgexpr") {
      uint32_t root = 0;
      ok = c.u("csg root", root);
      if (ok) ts.roots.push_back(root);
    } else {
      err = L + "unknown directive '" + d + "'";
      return false;
    }
    if (!ok) {
      err = c.err;
      return false;
    }
    if (nl == src.size()) break;
  }
  if (mats.empty()) mats.push_back(default_material());
  for (size_t i = 0; i < objs.size(); i++)
    if (objs[i].mat >= mats.size()) {
      err = "object " + std::to_string(i + 1) + ": material index " + std::to_string(objs[i].mat) +
            " out of range (have " + std::to_string(mats.size()) + ")";
      return false;
    }
  return write_ir(ts, ir, err);
}

// scene_io.rs:32-164 try_serialize
bool write_text(const HostScene& sc, const mm3e_camera& cam, std::string& s, std::string& err) {
  const mm3e_ir_header& h = sc.hdr;
  if (h.appearance != 0) {
    err = "UV textures are not representable in legacy .mm3e; use the native editor format";
    return false;
  }
  for (const auto& o : sc.objects) {
    if (o.prim == MM3E_PRIM_SURFACE) {
      err = "triangle surfaces are not representable in the legacy .mm3e text format; use the durable editor format";
      return false;
    }
  }
  s = "# MM3E scene\n";
  s += "size " + std::to_string(h.width) + " " + std::to_string(h.height) + "\n";
  s += "aa " + std::to_string(h.aa) + "\n";
  s += "bounces " + std::to_string(h.bounces) + "\n";
  s += "shadows " + std::to_string(h.shadows ? 1 : 0) + "\n";
  s += "ao " + std::to_string(h.ao ? 1 : 0) + "\n";
  s += "sun " + fmt3(h.sun_dir) + "\n";
  {  // the parser re-normalises `sun`; keep the stored vector when that would change its bits
    const F3 n = normalize(F3{h.sun_dir[0], h.sun_dir[1], h.sun_dir[2]});
    const float nn[3] = {n.x, n.y, n.z};
    if (std::memcmp(nn, h.sun_dir, 12) != 0) s += "sun_raw " + fmt3(h.sun_dir) + "\n";
  }
  s += "ambient " + fmt3(h.ambient) + "\n";
  s += "sky_ambient " + fmt3(h.sky_ambient) + "\n";
  s += "fog " + fmt3(h.fog) + " " + fmt(h.fog_density) + "\n";
  s += "marcher " + std::to_string(h.max_steps) + " " + fmt(h.max_dist) + " " + fmt(h.eps) + " " + fmt(h.step_scale) + "\n";
  s += "post " + fmt(h.exposure) + " " + std::to_string(h.bloom ? 1 : 0) + " " + fmt(h.bloom_threshold) + " " +
       fmt(h.bloom_intensity) + " " + std::to_string(h.bloom_radius) + "\n";
  s += std::string("mode ") + kModeNames[h.mode < 6 ? h.mode : 0] + "\n";
  {  // marcher budgets the reference's `marcher` line leaves out (Marcher::default otherwise)
    mm3e_ir_header d;
    scene_defaults(d, 1, 1);
    const bool same = h.shadow_steps == d.shadow_steps && h.ao_samples == d.ao_samples &&
                      std::memcmp(&h.lod_footprint, &d.lod_footprint, 4) == 0 &&
                      std::memcmp(&h.subitize, &d.subitize, 4) == 0 && (h.secant != 0) == (d.secant != 0) &&
                      std::memcmp(&h.normal_h, &d.normal_h, 4) == 0;
    if (!same)
      s += "marcher_ext " + std::to_string(h.shadow_steps) + " " + std::to_string(h.ao_samples) + " " +
           fmt(h.lod_footprint) + " " + fmt(h.subitize) + " " + std::to_string(h.secant ? 1 : 0) + " " +
           fmt(h.normal_h) + "\n";
  }
  if (h.gi_mode == MM3E_GI_BAKE) {
    s += "gi bake " + std::to_string(h.gi_dims[0]) + " " + std::to_string(h.gi_dims[1]) + " " +
         std::to_string(h.gi_dims[2]) + " " + std::to_string(h.gi_samples) + "\n";
  } else if (h.gi_mode == MM3E_GI_BAKED) {
    s += "gi baked " + std::to_string(h.gi_dims[0]) + " " + std::to_string(h.gi_dims[1]) + " " +
         std::to_string(h.gi_dims[2]) + " " + fmt3(h.gi_min) + " " + fmt3(h.gi_size) + " " +
         b64_encode(reinterpret_cast<const uint8_t*>(sc.gi_cubes.data()), sc.gi_cubes.size() * 4) + "\n";
  }
  // eye, target = eye + forward, up, vertical FOV in degrees: (atan(fov_scale) * 2).to_degrees()
  const float target[3] = {cam.eye[0] + cam.forward[0], cam.eye[1] + cam.forward[1], cam.eye[2] + cam.forward[2]};
  const float fov_deg = (std::atan(cam.fov_scale) * 2.0f) * 57.2957795130823208767981548141051703f;
  s += "cam " + fmt3(cam.eye) + " " + fmt3(target) + " " + fmt3(cam.up) + " " + fmt(fov_deg) + "\n";
  for (const auto& m : sc.materials)
    s += "mat " + fmt3(m.albedo) + " " + fmt(m.metallic) + " " + fmt(m.roughness) + " " + fmt(m.reflectivity) + " " +
         fmt(m.specular) + " " + fmt3(m.emissive) + " " + std::to_string(m.checker ? 1 : 0) + "\n";
  for (const auto& l : sc.lights)
    s += std::string("light ") + (l.directional ? "dir" : "point") + " " + fmt3(l.vec) + " " + fmt3(l.color) + " " +
         fmt(l.radius) + "\n";
  for (const auto& v : sc.volumes) {  // SdfVolume as its .sdfv bytes when the codec takes it
    const size_t n = (size_t)v.dims[0] * v.dims[1] * v.dims[2];
    const float* data = sc.volume_data.data() + v.offset;
    mm3e_sdf_layout lay;
    std::memcpy(lay.dims, v.dims, 12);
    std::memcpy(lay.min, v.min, 12);
    std::memcpy(lay.cell, v.cell, 12);
    size_t len = 0;
    mm3e_sdfv_encode(&lay, data, n, nullptr, &len, nullptr, 0);
    std::vector<uint8_t> bytes(len);
    if (len && mm3e_sdfv_encode(&lay, data, n, bytes.data(), &len, nullptr, 0) == MM3E_OK) {
      s += "volume sdfv " + b64_encode(bytes.data(), len) + "\n";
    } else {
      s += "volume raw " + std::to_string(v.dims[0]) + " " + std::to_string(v.dims[1]) + " " +
           std::to_string(v.dims[2]) + " " + fmt3(v.min) + " " + fmt3(v.cell) + " " +
           b64_encode(reinterpret_cast<const uint8_t*>(data), n * 4) + "\n";
    }
  }
  for (const auto& n : sc.csg_nodes) {
    std::string t = "csgnode " + std::to_string(n.op) + " " + std::to_string(n.prim) + " " + std::to_string(n.a) +
                    " " + std::to_string(n.b);
    for (int k = 0; k < 8; k++) t += " " + fmt(n.p[k]);
    for (int k = 0; k < 9; k++) t += " " + fmt(n.rot[k]);
    s += t + " " + fmt3(n.pos) + " " + fmt(n.scale) + "\n";
  }
  for (uint32_t r : sc.csg_roots) s += "csgexpr " + std::to_string(r) + "\n";
  static const char* names[13] = {"sphere", "box", "roundbox", "torus", "cylinder", "capsule", "cone",
                                  "ellipsoid", "octahedron", "hexprism", "plane", "csg", "volume"};
  for (const auto& o : sc.objects) {
    const float* p = o.p;
    std::string t = std::string("obj ") + names[o.prim] + " ";
    switch (o.prim) {
      case MM3E_PRIM_CSG: case MM3E_PRIM_VOLUME: {
        uint32_t id;
        std::memcpy(&id, &p[0], 4);
        t += std::to_string(id);
        break;
      }
      case MM3E_PRIM_SPHERE: case MM3E_PRIM_OCTAHEDRON: t += fmt(p[0]); break;
      case MM3E_PRIM_BOX: case MM3E_PRIM_ELLIPSOID: t += fmt3(p); break;
      case MM3E_PRIM_ROUND_BOX: t += fmt3(p) + " " + fmt(p[3]); break;
      case MM3E_PRIM_TORUS: case MM3E_PRIM_CYLINDER: case MM3E_PRIM_HEX_PRISM: t += fmt(p[0]) + " " + fmt(p[1]); break;
      case MM3E_PRIM_CAPSULE: t += fmt3(p) + " " + fmt3(p + 3) + " " + fmt(p[6]); break;
      case MM3E_PRIM_CONE: t += fmt(p[0]) + " " + fmt(p[1]) + " " + fmt(p[2]); break;
      case MM3E_PRIM_PLANE: t += fmt3(p) + " " + fmt(p[3]); break;
    }
    t += " pos " + fmt3(o.pos) + " basis " + fmt3(o.rot) + " " + fmt3(o.rot + 3) + " " + fmt3(o.rot + 6) + " scale " +
         fmt(o.scale) + " mat " + std::to_string(o.mat);
    if (o.combine == MM3E_COMBINE_UNION) t += " combine union";
    else if (o.combine == MM3E_COMBINE_SMOOTH) t += " combine smooth " + fmt(o.k);
    else t += " combine subtract";
    if (o.round != 0.0f) t += " round " + fmt(o.round);
    if (o.onion != 0.0f) t += " onion " + fmt(o.onion);
    if (o.twist != 0.0f) t += " twist " + fmt(o.twist);
    if (o.bend != 0.0f) t += " bend " + fmt(o.bend);
    if (o.mirror & 7u)
      t += " mirror " + std::to_string(o.mirror & 1u) + " " + std::to_string((o.mirror >> 1) & 1u) + " " +
           std::to_string((o.mirror >> 2) & 1u);
    if (!(o.repeat[0] == 0.0f && o.repeat[1] == 0.0f && o.repeat[2] == 0.0f)) t += " repeat " + fmt3(o.repeat);
    if (!(o.elongate[0] == 0.0f && o.elongate[1] == 0.0f && o.elongate[2] == 0.0f)) t += " elongate " + fmt3(o.elongate);
    s += t + "\n";
  }
  return true;
}

}  // namespace

extern "C" {

int mm3e_scene_parse_text(const char* text, size_t text_len, uint8_t* ir, size_t* ir_len, mm3e_camera* cam,
                          char* err, size_t errlen) {
  if (!text || !ir_len) return MM3E_ERR_INVALID;
  std::vector<uint8_t> blob;
  mm3e_camera c;
  std::string msg;
  if (!parse_text(std::string(text, text_len), blob, c, msg)) {
    set_err(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  if (!ir || *ir_len < blob.size()) {
    *ir_len = blob.size();
    set_err(err, errlen, "IR buffer too small");
    return MM3E_ERR_INVALID;
  }
  std::memcpy(ir, blob.data(), blob.size());
  *ir_len = blob.size();
  if (cam) *cam = c;
  return MM3E_OK;
}

int mm3e_scene_to_text(const void* ir, size_t ir_len, const mm3e_camera* cam, char* text, size_t* text_len, char* err,
                       size_t errlen) {
  if (!ir || !cam || !text_len) return MM3E_ERR_INVALID;
  HostScene hs;
  std::string msg, out;
  if (!parse_scene_ir(ir, ir_len, hs, msg) || !write_text(hs, *cam, out, msg)) {
    set_err(err, errlen, msg);
    return MM3E_ERR_INVALID;
  }
  if (!text || *text_len < out.size() + 1) {
    *text_len = out.size() + 1;
    set_err(err, errlen, "text buffer too small");
    return MM3E_ERR_INVALID;
  }
  std::memcpy(text, out.c_str(), out.size() + 1);
  *text_len = out.size() + 1;
  return MM3E_OK;
}

}  // extern "C"
