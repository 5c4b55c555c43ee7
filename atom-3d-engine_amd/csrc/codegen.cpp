// codegen.cpp — scene -> HIP source.
//
// The reference compiles each Scene into WGSL text with the scene baked in
// (mm3e-gpu/src/wgsl.rs:34-362: constants, CSG helpers, material switch, ordered `map`).  The
// twin here emits the world fold of WorldPlan::eval (accel.rs:209-252) for one compiled scene:
// every segment in plan order, each object handed to the same device helpers the table walk
// uses (seg_one / run_member / contrib / object_distance) as a compile-time constant `DObj`, so
// the compiler folds the transform, the primitive switch, the modifier branches and the
// bounding sphere into immediates while every floating-point operation stays the one the table
// path (and the reference) performs, in the same order.  Union runs that the reference walks
// through a bounding-sphere tree keep the table walk (seg_step on that segment index).  The
// segment stepping, with or without the segment grid, is world_eval's.
// Constants are written as bit patterns, so the text round-trips every float exactly.
#include "codegen.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace mm3e {

namespace {

void put_f(std::string& o, float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  char b[48];
  std::snprintf(b, sizeof b, "__builtin_bit_cast(float, 0x%08xu)", u);
  o += b;
}
void put_u(std::string& o, uint32_t u) {
  char b[24];
  std::snprintf(b, sizeof b, "0x%08xu", u);
  o += b;
}
void put_fa(std::string& o, const float* a, int n) {
  o += "{";
  for (int i = 0; i < n; i++) {
    if (i) o += ", ";
    put_f(o, a[i]);
  }
  o += "}";
}

// Aggregate initialiser in DObj member order (dev_scene.h).
void put_obj(std::string& o, const DObj& d) {
  o += "{";
  put_fa(o, d.c0, 3); o += ", "; put_f(o, d.inv); o += ", ";
  put_fa(o, d.c1, 3); o += ", "; put_f(o, d.scale); o += ", ";
  put_fa(o, d.c2, 3); o += ", "; put_u(o, d.mat); o += ", ";
  put_fa(o, d.pos, 3); o += ", "; put_u(o, d.flags); o += ", ";
  put_fa(o, d.p, 8); o += ", ";
  put_fa(o, d.rep, 3); o += ", "; put_f(o, d.twist); o += ", ";
  put_fa(o, d.elo, 3); o += ", "; put_f(o, d.bend); o += ", ";
  put_f(o, d.round); o += ", "; put_f(o, d.onion); o += ", "; put_f(o, d.bc_r); o += ", "; put_f(o, d.k); o += ", ";
  put_fa(o, d.bc, 3); o += ", 0u}";
}

bool has_csg(const Blob& b) { return b.header.csg != nullptr; }

// The scene's world fold as a map functor `name` (TableMap's interface) inside namespace mm3e.
void emit_map(std::string& o, const Blob& b, const std::string& name) {
  const char* csg = has_csg(b) ? "true" : "false";
  o += "struct " + name + " {\n  const DevScene* sc;\n  Tables tb;\n  MM3E_DEV Field operator()(V3 p) const;\n};\n";
  // The fold is the loop of world_eval with the per-segment body specialised: a uniform switch
  // on the segment index holds each One / linear-run segment's code once; tree runs share the
  // table walk in the default case.  The stepping (every segment, or the set bits of the
  // wave's grid masks) is world_eval's.
  o += "MM3E_DEV Field " + name + "::operator()(V3 p) const {\n  const DevScene& S = *sc;\n";
  o += "  Field acc{__builtin_inff(), 0u};\n  bool seeded = false;\n";
  const size_t n = b.segs.size();
  if (n == 0) {
    o += "  return acc;\n}\n";
  } else {
    o += "  const uint32_t nseg = " + std::to_string(n) + "u;\n";
    // The empty asm makes the point look modified on every iteration: otherwise the compiler
    // hoists every segment's point-dependent work (p - centre, ...) out of the loop and keeps
    // hundreds of values live across it.
    o += "  GridSel g;\n  g.on = false;\n  uint32_t w = 0, bits = 0, si = 0;\n  for (;;) {\n"
         "    V3 q = p;\n    asm volatile(\"\" : \"+v\"(q.x), \"+v\"(q.y), \"+v\"(q.z));\n    switch (si) {\n";
    for (size_t si = 0; si < n; si++) {
      const DSeg& g = b.segs[si];
      if (g.kind == 1) continue;
      o += "      case " + std::to_string(si) + "u: {\n";
      if (g.kind == 0) {
        o += "        constexpr DObj o = ";
        put_obj(o, b.objs[g.index]);
        o += ";\n        seg_one<";
        o += csg;
        o += ">(S, o, " + std::to_string(g.combine) + "u, ";
        put_f(o, g.k);
        o += ", q, acc, seeded);\n";
      } else {
        o += "        float bd;\n        uint32_t bm;\n        bool any;\n        run_begin(acc, seeded, bd, bm, any);\n";
        for (uint32_t u = 0; u < g.unb_count; u++) {
          o += "        {\n          constexpr DObj o = ";
          put_obj(o, b.objs[b.unb[g.unb_base + u]]);
          o += ";\n          run_member<";
          o += csg;
          o += ">(S, o, q, bd, bm, any);\n        }\n";
        }
        o += "        run_end(acc, seeded, bd, bm);\n";
      }
      o += "        break;\n      }\n";
    }
    o += "      default:\n        seg_step<";
    o += csg;
    o += ">(S, tb, q, si, acc, seeded);\n    }\n";
    if (b.header.grid != nullptr) {
      o += "    if (si == 0) {\n      g = grid_select(S, q, acc);\n      if (g.on) bits = grid_word(g, 0) & ~1u;\n    }\n"
           "    if (g.on) {\n      while (bits == 0) {\n        if (++w >= g.W) return acc;\n"
           "        bits = grid_word(g, w);\n      }\n      si = w * 32u + (uint32_t)__builtin_ctz(bits);\n"
           "      bits &= bits - 1u;\n      continue;\n    }\n";
    }
    o += "    if (++si >= nseg) return acc;\n  }\n}\n";
  }
}

}  // namespace

bool generated_needs_tables(const Blob& b) {
  for (const DSeg& s : b.segs)
    if (s.kind == 1) return true;
  return false;
}

// The state machine above 24 objects, nested loops below, with or without a segment grid:
// 1280x720 procedural scenes of 13 / 21 / 41 objects take 2.12 / 3.74 / 7.08 ms nested (6
// waves) against 2.08 / 5.40 / 6.32 ms as a state machine (4 waves); the nested module takes
// 15-24 s to compile instead of 4.  MM3E_GEN_SM=0/1 overrides.
bool generated_prefers_sm(const Blob& b) {
  if (const char* v = std::getenv("MM3E_GEN_SM")) return std::strcmp(v, "1") == 0;
  return b.objs.size() > 24;
}

std::string emit_scene_source(const Blob& b, uint32_t maxb, bool sm, bool queued, bool film) {
  // more than 4 reflection levels: the state machine (pending levels in a per-lane stack), not
  // one nested inlined copy of march and shading per level
  if (maxb > 4) sm = true;
  std::string o;
  o.reserve(4096 + b.objs.size() * 900);
  o += "// generated: world fold of one compiled scene (objects: " + std::to_string(b.objs.size()) +
       ", segments: " + std::to_string(b.segs.size()) + ")\n";
  // Waves per SIMD the generated kernel is register-allocated for.  Nested path: 6 (at most 80
  // VGPRs); it is bound by dependency latency more than by issue (c2 960x540 at 4 / 5 / 6 / 7
  // waves: 1.59 / 1.38 / 1.32 / 1.34 ms; c1 and a CSG scene at 1280x720 also fastest at 6).
  // Since the kernel holds one trace call site, 7 (at most 72) wins for short folds and loses for
  // longer ones (alternating on one box, 6 against 7: c2's 9 objects 1.306 / 1.294 ms, a
  // 13-object scene at 1280x720 2.111 / 2.084 ms, c1 level; 20 / 21 objects 2.18 / 2.27 and
  // 4.98 / 5.11 ms): 7 up to 16 objects.
  // State-machine path: 4 (at most 128 VGPRs) -- its per-lane path state spills in the hot loop
  // below that (1280x720 procedural scenes of 13 / 21 / 41 objects at 4 / 5 / 6 waves:
  // 2.08 / 4.06 / 6.83, 5.40 / 6.38 / 7.61, 6.37 / 8.15 / 10.05 ms) -- except for long folds,
  // where latency hiding wins again (61 objects: 22.2 / 20.6 / 20.4 ms): 5 above 48 objects,
  // and 7 above 200: the 1000-object stress frame, alternating builds on one box each, at 5 / 6
  // waves 708.5 / 649.9 and 709.1 / 653.6 ms, at 6 / 7 668.0 / 634.0 and 668.9 / 626.1 ms, at 7 /
  // 8 631.7 / 666.1 and 638.2 / 667.3 ms (at 6 the 127-object scene is 4 % slower, 174 level).
  // MM3E_GEN_WAVES overrides.
  int waves = sm ? (b.objs.size() > 200 ? 7 : b.objs.size() > 48 ? 5 : 4) : (b.objs.size() <= 16 ? 7 : 6);
  if (const char* wv = std::getenv("MM3E_GEN_WAVES")) {
    const int n = std::atoi(wv);
    if (n >= 1 && n <= 8) waves = n;
  }
  o += "#define MM3E_RENDER_WAVES " + std::to_string(waves) + "\n";
  o += "#ifndef __HIPCC_RTC__\n#include <hip/hip_runtime.h>\n#endif\n#include \"dev_render.h\"\n";
  if (queued) o += "#include \"dev_wavefront.h\"\n";
  o += "namespace mm3e {\n";
  emit_map(o, b, "GenMap");
  o += "}  // namespace mm3e\n";
  o += "using namespace mm3e;\n";
  // The scene's one render kernel: block b renders tile o.order[b] (most expensive first), or
  // tile b when o.order is null.  One kernel for both dispatch orders halves the run-time
  // compile, which grows with every inlined copy of the fold (and hipRTC compiles one program
  // at a time per process).
  o += "extern \"C\" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY\n"
       "    mm3e_gen_render_o(const DevScene* __restrict__ sc, FrameParams fp, float* __restrict__ hdr, uint32_t tw,\n"
       "                      uint32_t th, uint32_t stage_bytes, TileOrder o) {\n"
       "  __shared__ float smem[256 * 3];\n"
       "  __shared__ uint32_t start;\n"
       "  extern __shared__ uint4 tables_lds[];\n"
       "  const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);\n";
  o += "  render_ordered<GenMap, " + std::to_string(maxb) + ", " + (sm ? "true" : "false") +
       ">(sc, tb, fp, hdr, tw, th, o, smem, &start);\n}\n";
  if (film) {  // render_film's sample pass and the G-buffer / hole pass (tables read from HBM)
    const std::string tmpl = std::to_string(maxb) + ", " + (sm ? "true" : "false");
    o += "extern \"C\" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY\n"
         "    mm3e_gen_film(const DevScene* __restrict__ sc, FrameParams fp, FilmParams lp, float* __restrict__ beauty,\n"
         "                  float* __restrict__ fg, float* __restrict__ alpha, uint32_t tw, uint32_t th) {\n"
         "  __shared__ float smem[256 * 4];\n"
         "  film_tile<GenMap, " + tmpl + ">(sc, global_tables(sc), fp, lp, beauty, fg, alpha, tw, th, blockIdx.x,\n"
         "                                  blockIdx.y, smem);\n}\n";
    o += "extern \"C\" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY\n"
         "    mm3e_gen_gbuffer(const DevScene* __restrict__ sc, FrameParams fp, float exposure,\n"
         "                     const GMover* __restrict__ movers, uint32_t n_movers, const uint32_t* __restrict__ list,\n"
         "                     const uint32_t* __restrict__ n_list, uchar4* __restrict__ color, float* __restrict__ depth,\n"
         "                     int32_t* __restrict__ obj) {\n"
         "  gbuffer_tile<GenMap, " + tmpl + ">(sc, fp, exposure, movers, n_movers, list, n_list, color, depth, obj);\n}\n";
  }
  if (queued) {
    const std::string tbl = "  extern __shared__ uint4 tables_lds[];\n"
                            "  const Tables tb = stage_bytes ? stage_tables(sc, tables_lds, stage_bytes) : global_tables(sc);\n";
    o += "extern \"C\" __global__ void __launch_bounds__(256)\n"
         "    mm3e_gen_wf_first(const DevScene* __restrict__ sc, FrameParams fp, WfFrame w, uint32_t budget,\n"
         "                      uint32_t stage_bytes) {\n" + tbl + "  wf_march_first<GenMap>(sc, tb, fp, w, budget);\n}\n";
    o += "extern \"C\" __global__ void __launch_bounds__(256)\n"
         "    mm3e_gen_wf_cont(const DevScene* __restrict__ sc, FrameParams fp, WfFrame w, uint32_t k, uint32_t budget,\n"
         "                     uint32_t stage_bytes) {\n" + tbl + "  wf_march_cont<GenMap>(sc, tb, fp, w, k, budget);\n}\n";
    o += "extern \"C\" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY\n"
         "    mm3e_gen_wf_shade(const DevScene* __restrict__ sc, FrameParams fp, WfFrame w, uint32_t stage_bytes) {\n" +
         tbl + "  wf_shade<GenMap, " + std::to_string(maxb) + ", " + (sm ? "true" : "false") + ">(sc, tb, fp, w);\n}\n";
  }
  return o;
}

std::string emit_batch_source(const std::vector<const Blob*>& maps, uint32_t maxb, bool staged) {
  std::string o;
  o += "// generated: batch module with " + std::to_string(maps.size()) + " scene maps\n";
  // every map runs the state machine, as the table walk does; the occupancy of a single scene's
  // state-machine kernel for the longest map (MM3E_GEN_WAVES overrides)
  size_t most = 0;
  for (const Blob* m : maps) most = std::max(most, m->objs.size());
  int waves = most > 200 ? 7 : most > 48 ? 5 : 4;
  if (const char* wv = std::getenv("MM3E_GEN_WAVES")) {
    const int n = std::atoi(wv);
    if (n >= 1 && n <= 8) waves = n;
  }
  o += "#define MM3E_RENDER_WAVES " + std::to_string(waves) + "\n";
  o += "#ifndef __HIPCC_RTC__\n#include <hip/hip_runtime.h>\n#endif\n#include \"dev_batch.h\"\nnamespace mm3e {\n";
  for (size_t k = 0; k < maps.size(); k++) emit_map(o, *maps[k], "GenMap" + std::to_string(k));
  // One map for the whole module, switching on the job's scene map (uniform per tile) inside the
  // fold: the kernel holds one inlined copy of the march and shading and one fold per scene.
  o += "struct MultiMap {\n  const DevScene* sc;\n  Tables tb;\n  uint32_t k;  // the job's scene map\n"
       "  MM3E_DEV Field operator()(V3 p) const {\n    switch (k) {\n";
  for (size_t k = 0; k < maps.size(); k++)
    o += "      case " + std::to_string(k) + "u: return GenMap" + std::to_string(k) + "{sc, tb}(p);\n";
  o += "      default: return TableMap{sc, tb}(p);\n    }\n  }\n};\n";
  o += "}  // namespace mm3e\nusing namespace mm3e;\n";
  const std::string mb = std::to_string(maxb);
  o += "struct GenTiles {\n"
       "  MM3E_DEV void operator()(const BatchJob& J, const Tables& tb, uint32_t tx, uint32_t ty, float* smem) const {\n"
       "    render_tile_map<MultiMap, " + mb + ", true>(MultiMap{J.sc, tb, J.gen}, J.sc, J.fp, J.hdr, J.tw, J.th, tx, ty, smem);\n"
       "  }\n};\n";
  o += "extern \"C\" __global__ void __launch_bounds__(256) MM3E_RENDER_OCCUPANCY\n"
       "    mm3e_gen_batch(const BatchJob* __restrict__ jobs, uint32_t n_jobs, uint32_t total_tiles,\n"
       "                   uint32_t* __restrict__ counter, TileOrder q) {\n"
       "  extern __shared__ uint4 tables_lds[];\n";
  o += std::string("  batch_loop<") + (staged ? "true" : "false") +
       ">(jobs, n_jobs, total_tiles, counter, q, tables_lds, GenTiles{});\n}\n";
  return o;
}

}  // namespace mm3e
