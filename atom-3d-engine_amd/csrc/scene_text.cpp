// scene_text.cpp — the reference's `.mm3e` text scene format (mm3e-orchestrator/src/scene_io.rs)
// as a wire format for this library: parse text into the flat scene IR plus a camera, and
// write a scene IR plus camera back to text.  Line-oriented: a directive keyword followed by
// whitespace-separated values, `#` starts a comment.  Numbers print as the shortest decimal
// that reads back to the same f32 (Rust's `{}` for f32), so text -> IR -> text is stable.
// The format carries no CSG, volumes, GI, shadow/AO budgets or normal_h (scene_io.rs:47-62,
// 103-105); parsed scenes take Scene::new's values for those.
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

std::vector<uint8_t> write_ir(mm3e_ir_header h, const std::vector<mm3e_ir_object>& objs,
                              const std::vector<mm3e_ir_material>& mats, const std::vector<mm3e_ir_light>& lights) {
  h.n_objects = (uint32_t)objs.size();
  h.n_materials = (uint32_t)mats.size();
  h.n_lights = (uint32_t)lights.size();
  uint32_t off = sizeof(mm3e_ir_header);
  h.off_objects = off;
  off += (uint32_t)(objs.size() * sizeof(mm3e_ir_object));
  h.off_materials = off;
  off += (uint32_t)(mats.size() * sizeof(mm3e_ir_material));
  h.off_lights = off;
  off += (uint32_t)(lights.size() * sizeof(mm3e_ir_light));
  h.off_volumes = h.off_volume_data = h.off_csg_nodes = h.off_csg_roots = h.off_gi_cubes = off;
  h.total_bytes = off;
  std::vector<uint8_t> out(off);
  std::memcpy(out.data(), &h, sizeof h);
  if (!objs.empty()) std::memcpy(out.data() + h.off_objects, objs.data(), objs.size() * sizeof(mm3e_ir_object));
  if (!mats.empty()) std::memcpy(out.data() + h.off_materials, mats.data(), mats.size() * sizeof(mm3e_ir_material));
  if (!lights.empty()) std::memcpy(out.data() + h.off_lights, lights.data(), lights.size() * sizeof(mm3e_ir_light));
  return out;
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
  else {
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
  mm3e_ir_header h;
  scene_defaults(h, 640, 360);
  std::vector<mm3e_ir_object> objs;
  std::vector<mm3e_ir_material> mats;  // Scene::new's default material is cleared
  std::vector<mm3e_ir_light> lights;
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
  ir = write_ir(h, objs, mats, lights);
  return true;
}

// scene_io.rs:32-164 try_serialize
bool write_text(const HostScene& sc, const mm3e_camera& cam, std::string& s, std::string& err) {
  const mm3e_ir_header& h = sc.hdr;
  for (const auto& o : sc.objects) {
    if (o.prim == MM3E_PRIM_SURFACE) {
      err = "triangle surfaces are not representable in the legacy .mm3e text format; use the durable editor format";
      return false;
    }
    if (o.prim == MM3E_PRIM_CSG) {
      err = "object-local CSG is not representable in the legacy .mm3e text format";
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
  s += "ambient " + fmt3(h.ambient) + "\n";
  s += "sky_ambient " + fmt3(h.sky_ambient) + "\n";
  s += "fog " + fmt3(h.fog) + " " + fmt(h.fog_density) + "\n";
  s += "marcher " + std::to_string(h.max_steps) + " " + fmt(h.max_dist) + " " + fmt(h.eps) + " " + fmt(h.step_scale) + "\n";
  s += "post " + fmt(h.exposure) + " " + std::to_string(h.bloom ? 1 : 0) + " " + fmt(h.bloom_threshold) + " " +
       fmt(h.bloom_intensity) + " " + std::to_string(h.bloom_radius) + "\n";
  s += std::string("mode ") + kModeNames[h.mode < 6 ? h.mode : 0] + "\n";
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
  static const char* names[11] = {"sphere", "box", "roundbox", "torus", "cylinder", "capsule",
                                  "cone", "ellipsoid", "octahedron", "hexprism", "plane"};
  for (const auto& o : sc.objects) {
    if (o.prim == MM3E_PRIM_VOLUME) {
      uint32_t id;
      std::memcpy(&id, &o.p[0], 4);
      s += "# volume object omitted (baked SDF volume id " + std::to_string(id) + " is not text-serializable)\n";
      continue;
    }
    const float* p = o.p;
    std::string t = std::string("obj ") + names[o.prim] + " ";
    switch (o.prim) {
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
