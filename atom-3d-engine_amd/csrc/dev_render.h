// dev_render.h — kernel bodies of the MI355X render path, templated on the world-field
// implementation (`TableMap` walks the compiled scene tables; generated maps specialise it).
//
//   render_tile      one lane per (pixel, AA sample): march + shade + reflection chain
//                    (lib.rs:1223-1360), nested loops or a per-lane state machine; the first
//                    lane of each pixel sums its samples in the reference's sy-major order, by
//                    cross-lane reads when aa^2 divides 64, else through LDS.
//   gather_rays      GI bake rays (lib.rs:686-711): one lane per (probe, face, ray).
//   gi_reduce        weighted mean per cube face (gi.rs:327-341).
//   bloom_h / bloom_v_resolve / resolve_only   post.rs:36-115 + framebuffer.rs:102-113.
//   film_tile        render_film's samples through a thin lens: beauty, premultiplied
//                    foreground, coverage alpha (film.rs:63-225).
//   primary_aov      centre-pixel primary hit: mask, material, depth, normal (film.rs:114-122).
//   aov_mode_pixel   debug render modes Normal / Depth / Ao / Steps / Albedo (lib.rs:1178-1218).
//   field_probe / march_probe   map / raw-march parity probes (wgsl.rs:641-646 twin).
#pragma once
#include "dev_math.h"

// Waves per SIMD the render kernels are register-allocated for: 4 keeps them at <= 128 VGPRs
// (16 waves per CU), trading a few SGPR spills into VGPR lanes for a quarter more latency
// hiding than the 3 waves the unconstrained allocation reaches.
#ifndef MM3E_RENDER_WAVES
#define MM3E_RENDER_WAVES 4
#endif
#define MM3E_RENDER_OCCUPANCY __attribute__((amdgpu_waves_per_eu(MM3E_RENDER_WAVES)))

namespace mm3e {

MM3E_DEV void camera_ray(const FrameParams& fp, float sx, float sy, V3& ro, V3& rd) {  // camera.rs:59-65
  const float w = (float)fp.width, h = (float)fp.height;
  float ndc_x = (2.0f * sx / w - 1.0f) * fp.aspect * fp.fov_scale;
  float ndc_y = (1.0f - 2.0f * sy / h) * fp.fov_scale;
  rd = normalize(ld3(fp.fwd) + scale(ld3(fp.right), ndc_x) + scale(ld3(fp.up), ndc_y));
  ro = ld3(fp.eye);
}

// One surface interaction: everything of shade_hit (lib.rs:1286-1360) except the reflection
// recursion, which the caller chains.  Returns the pre-reflection radiance `base`.
template <class Map>
MM3E_DEV V3 shade_surface(const DevScene& sc, const Map& field, const Hit& h, V3 n, V3 rd, const DMat& m) {
  V3 albedo = surface_albedo(m, h.pos);
  V3 view = scale(rd, -1.0f);
  float occ = 1.0f;
  if (sc.flags & SC_AO) occ = ambient_occlusion(sc, field, h.pos, n);
  V3 gi = V3{0.0f, 0.0f, 0.0f};
  if (sc.flags & SC_HAS_GI) gi = gi_sample(sc, h.pos, n);
  V3 amb = ibl_irradiance(sc, n) + gi + ld3(sc.ambient);
  V3 radiance = scale(cmul(albedo, amb), occ);
  const uint32_t nl = sc.n_light;
  for (uint32_t li = 0; li < nl; li++) {
    const DLight& l = sc.light[li];
    V3 ld;
    float dist;
    light_toward(l, h.pos, ld, dist);
    if (lambert(n, ld) <= 0.0f) continue;
    float shadow = 1.0f;
    if (sc.flags & SC_SHADOWS) {
      float max_t = rmin(dist, sc.max_dist);
      float lift = rmax(sc.normal_h * 2.0f, 0.01f);
      shadow = soft_shadow(sc, field, h.pos + scale(n, lift), ld, max_t, light_shadow_k(l, dist));
    }
    if (shadow <= 0.0f) continue;
    V3 surf = brdf(n, ld, view, albedo, m.metallic, m.roughness, m.specular);
    radiance = radiance + scale(cmul(surf, ld3(l.color)), shadow * light_atten(l, dist));
  }
  return radiance;
}

// trace (lib.rs:1265-1282) as nested loops: the kernel for scenes whose world fold is short.
// The reflection recursion is bounded at compile time by MAXB (the scene's `bounces`), so each
// level is a separate inlined copy and no per-level state goes to scratch:
// out = fog(base*(1-fr) + trace(reflected)*fr + emissive).
template <class Map, int D, int MAXB>
MM3E_DEV V3 shade_level(const DevScene& sc, const Map& field, const Hit& h, V3 rd);

template <class Map, int D, int MAXB>
MM3E_DEV V3 trace_level(const DevScene& sc, const Map& field, V3 ro, V3 rd) {
  Hit h = march(sc, field, ro, rd);
  if (!h.hit) return sky(rd, ld3(sc.sun));
  return shade_level<Map, D, MAXB>(sc, field, h, rd);
}

// shade_hit (lib.rs:1286-1360) for a confirmed hit at recursion depth D.
template <class Map, int D, int MAXB>
MM3E_DEV V3 shade_level(const DevScene& sc, const Map& field, const Hit& h, V3 rd) {
  V3 n = hit_normal(sc, field, h.pos);
  const DMat& m = sc.mat[h.mat % sc.n_mat];
  V3 radiance = shade_surface(sc, field, h, n, rd, m);
  if constexpr (D < MAXB) {
    if ((uint32_t)D < sc.bounces && m.reflectivity > 0.0f) {
      V3 view = scale(rd, -1.0f);
      float c = rmax(dot(n, view), 0.0f);
      float fr = fresnel_schlick(c, m.reflectivity);
      V3 rdir = normalize(reflect(rd, n));
      V3 rorg = h.pos + scale(n, rmax(sc.normal_h * 2.0f, 0.02f));
      V3 refl = trace_level<Map, D + 1, MAXB>(sc, field, rorg, rdir);
      radiance = scale(radiance, 1.0f - fr) + scale(refl, fr);
    }
  }
  radiance = radiance + ld3(m.emissive);
  return apply_fog(radiance, ld3(sc.fog), h.t, sc.fog_density);
}

// ---------------------------------------------------------------------------------------
// trace / shade_hit (lib.rs:1247-1360) as one per-lane state machine: the kernel for scenes
// whose world fold is long (segment grid, CSG objects).
//
// Every field evaluation of a sample's path — march steps and the overlap-guard probe, the
// tetrahedron normal, AO taps, soft-shadow steps, the reflected rays' marches — goes through
// the single `field(ps.q)` site of trace_path, so the kernel holds one inlined copy of the world
// fold, and the lanes of a wave share each evaluation whichever search they are in.  Per lane
// the sequence of operations is exactly the recursive reference:
//   trace(ray, d): miss -> sky(dir); hit -> radiance = shade (AO, GI, IBL, lights with soft
//   shadows), then if d < bounces and reflectivity > 0: radiance = radiance*(1-fr) +
//   trace(reflected, d+1)*fr; then + emissive; then fog(t).
// The pending levels of that recursion (radiance, fr, t, material) wait in a per-lane stack.
// ---------------------------------------------------------------------------------------
enum : uint32_t { PH_MARCH = 0, PH_NORMAL = 1, PH_AO = 2, PH_SHADOW = 3, PH_DONE = 4 };

// The searches never overlap in time, so they share storage.
union SearchState {
  MarchState m;
  TetraState tn;
  AoState ao;
  ShadowState sh;
};

struct PathState {
  uint32_t phase, mst, level, li;
  V3 q;         // next point to evaluate
  V3 rd;        // current ray direction
  Hit h;        // current level's hit
  V3 n;         // its normal
  V3 radiance;  // current level's radiance so far; the sample's radiance once PH_DONE
  float value;  // finished AO / shadow search result
  SearchState x;
};

template <int MAXB>
struct LevelStack {
  // base radiance xyz, fr, hit t and the material index per pending level; volatile keeps it
  // in per-lane scratch memory rather than in registers that every iteration would carry
  volatile float v[MAXB > 0 ? MAXB : 1][5];
  volatile uint32_t mat[MAXB > 0 ? MAXB : 1];
};

// Fold finished levels back up the recursion: out = fog(base*(1-fr) + inner*fr + emissive).
template <int MAXB>
MM3E_DEV void path_unwind(const DevScene& sc, PathState& ps, LevelStack<MAXB>& stk, V3 r) {
  while (ps.level > 0) {
    ps.level--;
    const volatile float* e = stk.v[ps.level];
    const V3 base = V3{e[0], e[1], e[2]};
    const float fr = e[3], t = e[4];
    const uint32_t mat = stk.mat[ps.level];
    r = scale(base, 1.0f - fr) + scale(r, fr);
    r = r + ld3(sc.mat[mat].emissive);
    r = apply_fog(r, ld3(sc.fog), t, sc.fog_density);
  }
  ps.radiance = r;
  ps.phase = PH_DONE;
}

template <int MAXB>
MM3E_DEV void path_march(const DevScene& sc, PathState& ps, V3 ro, V3 rd) {
  ps.rd = rd;
  ps.phase = PH_MARCH;
  ps.mst = march_init(sc, ps.x.m, ro, rd, ps.h, ps.q);
}

// Lights from ps.li on (lib.rs:1333-1352); returns true when a shadow search needs evaluating.
template <int MAXB>
MM3E_DEV bool path_lights(const DevScene& sc, PathState& ps, float shadow_in, bool resume) {
  const DMat& m = sc.mat[ps.h.mat % sc.n_mat];
  const V3 albedo = surface_albedo(m, ps.h.pos);
  const V3 view = scale(ps.rd, -1.0f);
  for (;;) {
    if (ps.li >= sc.n_light) return false;
    const DLight& l = sc.light[ps.li];
    V3 ld;
    float dist;
    light_toward(l, ps.h.pos, ld, dist);
    float shadow = 1.0f;
    if (resume) {
      shadow = shadow_in;
      resume = false;
    } else {
      if (lambert(ps.n, ld) <= 0.0f) {
        ps.li++;
        continue;
      }
      if (sc.flags & SC_SHADOWS) {
        const float max_t = rmin(dist, sc.max_dist);
        const float lift = rmax(sc.normal_h * 2.0f, 0.01f);
        if (shadow_init(sc, ps.x.sh, ps.h.pos + scale(ps.n, lift), ld, max_t, light_shadow_k(l, dist), ps.q, shadow)) {
          ps.phase = PH_SHADOW;
          return true;
        }
      }
    }
    if (shadow > 0.0f) {
      const V3 surf = brdf(ps.n, ld, view, albedo, m.metallic, m.roughness, m.specular);
      ps.radiance = ps.radiance + scale(cmul(surf, ld3(l.color)), shadow * light_atten(l, dist));
    }
    ps.li++;
  }
}

// After the lights: reflect (push this level) or finish the level and unwind.
template <int MAXB>
MM3E_DEV void path_level_end(const DevScene& sc, PathState& ps, LevelStack<MAXB>& stk) {
  const uint32_t mi = ps.h.mat % sc.n_mat;
  const DMat& m = sc.mat[mi];
  if (MAXB > 0 && ps.level < (uint32_t)MAXB && ps.level < sc.bounces && m.reflectivity > 0.0f) {
    const V3 view = scale(ps.rd, -1.0f);
    const float c = rmax(dot(ps.n, view), 0.0f);
    const float fr = fresnel_schlick(c, m.reflectivity);
    const V3 rdir = normalize(reflect(ps.rd, ps.n));
    const V3 rorg = ps.h.pos + scale(ps.n, rmax(sc.normal_h * 2.0f, 0.02f));
    volatile float* e = stk.v[ps.level];
    e[0] = ps.radiance.x;
    e[1] = ps.radiance.y;
    e[2] = ps.radiance.z;
    e[3] = fr;
    e[4] = ps.h.t;
    stk.mat[ps.level] = mi;
    ps.level++;
    path_march<MAXB>(sc, ps, rorg, rdir);
    return;
  }
  V3 r = ps.radiance + ld3(m.emissive);
  r = apply_fog(r, ld3(sc.fog), ps.h.t, sc.fog_density);
  path_unwind<MAXB>(sc, ps, stk, r);
}

// Surface shading up to the lights (lib.rs:1286-1332): albedo, AO, GI, IBL, ambient.
template <int MAXB>
MM3E_DEV void path_shade_after_ao(const DevScene& sc, PathState& ps, LevelStack<MAXB>& stk, float occ) {
  const DMat& m = sc.mat[ps.h.mat % sc.n_mat];
  const V3 albedo = surface_albedo(m, ps.h.pos);
  V3 gi = V3{0.0f, 0.0f, 0.0f};
  if (sc.flags & SC_HAS_GI) gi = gi_sample(sc, ps.h.pos, ps.n);
  const V3 amb = ibl_irradiance(sc, ps.n) + gi + ld3(sc.ambient);
  ps.radiance = scale(cmul(albedo, amb), occ);
  ps.li = 0;
  if (path_lights<MAXB>(sc, ps, 0.0f, false)) return;
  path_level_end<MAXB>(sc, ps, stk);
}

template <int MAXB>
MM3E_DEV void path_shade(const DevScene& sc, PathState& ps, LevelStack<MAXB>& stk) {
  float occ = 1.0f;
  if (sc.flags & SC_AO) {
    if (ao_init(sc, ps.x.ao, ps.h.pos, ps.n, ps.q, occ)) {
      ps.phase = PH_AO;
      return;
    }
  }
  path_shade_after_ao<MAXB>(sc, ps, stk, occ);
}

template <int MAXB>
MM3E_DEV void path_hit(const DevScene& sc, const Tables& tb, PathState& ps, LevelStack<MAXB>& stk) {
  if (!ps.h.hit) {
    path_unwind<MAXB>(sc, ps, stk, sky(ps.rd, ld3(sc.sun)));
    return;
  }
  if (sc.flags & SC_DUAL_SAFE) {
    ps.n = normal_dual(sc, tb, ps.h.pos);
    path_shade<MAXB>(sc, ps, stk);
    return;
  }
  ps.q = tetra_init(ps.x.tn, ps.h.pos, sc.normal_h);
  ps.phase = PH_NORMAL;
}

// One step of the lane's current search with the field value at ps.q.  Returns true when
// the search has finished; its result then waits in the search state (march: ps.h; normal:
// ps.x.tn.g; AO / shadow: ps.occ_or_shadow) until the lane's transition runs.
template <int MAXB>
MM3E_DEV bool path_search_step(const DevScene& sc, PathState& ps, Field f) {
  switch (ps.phase) {
    case PH_MARCH:
      ps.mst = march_step(sc, ps.x.m, ps.mst, f, ps.h, ps.q);
      return ps.mst == MARCH_DONE;
    case PH_NORMAL:
      return tetra_step(ps.x.tn, f.dist, ps.q);
    case PH_AO:
      return ao_step(sc, ps.x.ao, f.dist, ps.q, ps.value);
    default:  // PH_SHADOW
      return shadow_step(sc, ps.x.sh, f.dist, ps.q, ps.value);
  }
}

// The work between searches (shading, lights, reflection set-up, unwinding), run for a lane
// whose search has finished; it ends by starting the next search or with PH_DONE.
template <int MAXB>
MM3E_DEV void path_transition(const DevScene& sc, const Tables& tb, PathState& ps, LevelStack<MAXB>& stk) {
  switch (ps.phase) {
    case PH_MARCH:
      path_hit<MAXB>(sc, tb, ps, stk);
      return;
    case PH_NORMAL:
      ps.n = normalize(ps.x.tn.g);
      path_shade<MAXB>(sc, ps, stk);
      return;
    case PH_AO:
      path_shade_after_ao<MAXB>(sc, ps, stk, ps.value);
      return;
    default:  // PH_SHADOW
      if (path_lights<MAXB>(sc, ps, ps.value, true)) return;
      path_level_end<MAXB>(sc, ps, stk);
      return;
  }
}

// Lanes whose search finished wait until at least this many lanes of the wave are waiting (or
// no lane is searching any more) and then run their transitions together: the shading code
// executes a few times per wave instead of once per finishing lane, while every field
// evaluation is still shared by all searching lanes whatever search they are in.
#ifndef MM3E_PATH_GATE
#define MM3E_PATH_GATE 1
#endif

// Runs a lane's path to PH_DONE from its current state (`waiting`: its search has finished and
// its transition is due).  HIT0 also returns the primary (depth-0) march's hit record in *hit0.
template <class Map, int MAXB, bool HIT0>
MM3E_DEV void path_run(const DevScene& sc, const Map& field, PathState& ps, LevelStack<MAXB>& stk, bool waiting,
                       Hit* hit0) {
  for (;;) {
    const bool searching = ps.phase != PH_DONE && !waiting;
    if (searching) waiting = path_search_step<MAXB>(sc, ps, field(ps.q));
    const uint64_t wait_mask = __builtin_amdgcn_ballot_w64(waiting);
    const uint64_t search_mask = __builtin_amdgcn_ballot_w64(ps.phase != PH_DONE && !waiting);
    if (wait_mask != 0 && (search_mask == 0 || __builtin_popcountll(wait_mask) >= MM3E_PATH_GATE)) {
      if (waiting) {
        if constexpr (HIT0) {
          if (ps.phase == PH_MARCH && ps.level == 0) *hit0 = ps.h;
        }
        path_transition<MAXB>(sc, field.tb, ps, stk);
        // the only search that can finish before its first evaluation is a march with
        // max_steps == 0; every other search either asks for a point or completes inside
        // the transition
        waiting = ps.phase == PH_MARCH && ps.mst == MARCH_DONE;
      }
    }
    if (__builtin_amdgcn_ballot_w64(ps.phase != PH_DONE) == 0) break;
  }
}

template <class Map, int MAXB, bool HIT0 = false>
MM3E_DEV V3 trace_sm(const DevScene& sc, const Map& field, V3 ro, V3 rd, Hit* hit0 = nullptr) {
  PathState ps;
  LevelStack<MAXB> stk;
  ps.level = 0;
  path_march<MAXB>(sc, ps, ro, rd);
  path_run<Map, MAXB, HIT0>(sc, field, ps, stk, ps.mst == MARCH_DONE, hit0);
  return ps.radiance;
}

// Render one tile: blockDim.x = tw*th*aa*aa lanes (at most 256; above that the tile is one pixel
// and its samples run in rounds); lane -> (pixel in tile, sample).
// SM selects the state-machine path (long world folds) over the nested one.
template <class Map, int MAXB, bool SM>
MM3E_DEV void render_tile_map(const Map& field, const DevScene* __restrict__ scp, const FrameParams& fp,
                              float* __restrict__ hdr, uint32_t tw, uint32_t th, uint32_t tile_x, uint32_t tile_y,
                              float* smem) {
  const DevScene& sc = *scp;
  const uint32_t aa = sc.aa;
  const uint32_t ns = aa * aa;
  const uint32_t t = threadIdx.x;
  // More samples than lanes (aa > 16): the tile is one pixel and the block runs its samples in
  // rounds of blockDim.x; the first lane adds each round's results in sample order, so the sum
  // is the reference's sy-major sequence of additions.  Both forms share one trace call site,
  // so the kernel holds a single inlined copy of the march and the shading.
  const bool rounds = ns > blockDim.x;
  const uint32_t n_rounds = rounds ? (ns + blockDim.x - 1u) / blockDim.x : 1u;
  V3 sum = V3{0.0f, 0.0f, 0.0f};
  for (uint32_t r = 0; r < n_rounds; r++) {
    const uint32_t pl = rounds ? 0u : t / ns;
    const uint32_t s = rounds ? r * blockDim.x + t : t - pl * ns;
    const uint32_t px = tile_x * tw + pl % tw, py = tile_y * th + pl / tw;
    const bool valid = pl < tw * th && px < fp.width && py < fp.height && s < ns;
    V3 rad = V3{0.0f, 0.0f, 0.0f};
    if (valid) {
      const uint32_t sy = s / aa, sx = s - sy * aa;
      const float ox = ((float)sx + 0.5f) / (float)aa;
      const float oy = ((float)sy + 0.5f) / (float)aa;
      V3 ro, rd;
      camera_ray(fp, (float)px + ox, (float)py + oy, ro, rd);
      if constexpr (SM) rad = trace_sm<Map, MAXB>(sc, field, ro, rd);
      else rad = trace_level<Map, 0, MAXB>(sc, field, ro, rd);
    }
    if (rounds) {
      smem[3 * t + 0] = rad.x;
      smem[3 * t + 1] = rad.y;
      smem[3 * t + 2] = rad.z;
      __syncthreads();
      if (t == 0) {
        const uint32_t n = ns - r * blockDim.x < blockDim.x ? ns - r * blockDim.x : blockDim.x;
        for (uint32_t k = 0; k < n; k++) sum = sum + V3{smem[3 * k], smem[3 * k + 1], smem[3 * k + 2]};
      }
      __syncthreads();
      continue;
    }
    if (ns == 1) {
      if (valid) {
        float* o = hdr + 3 * ((size_t)py * fp.width + px);
        // acc = ZERO + s0; acc.scale(1/1)
        V3 acc = V3{0.0f, 0.0f, 0.0f} + rad;
        o[0] = acc.x * 1.0f; o[1] = acc.y * 1.0f; o[2] = acc.z * 1.0f;
      }
      return;
    }
    if ((64u % ns) == 0u) {
      // A pixel's samples sit in adjacent lanes of one wave (lane t = pixel * ns + sample with
      // ns | 64): its first lane sums them by cross-lane reads in sample order, the same additions
      // as below, and no wave waits for the rest of the block.
      V3 acc = V3{0.0f, 0.0f, 0.0f} + rad;
      for (uint32_t k = 1; k < ns; k++)
        acc = acc + V3{__shfl_down(rad.x, k), __shfl_down(rad.y, k), __shfl_down(rad.z, k)};
      if (valid && s == 0) {
        const float inv = 1.0f / (float)ns;
        acc = scale(acc, inv);
        float* o = hdr + 3 * ((size_t)py * fp.width + px);
        o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
      }
      return;
    }
    smem[3 * t + 0] = rad.x;
    smem[3 * t + 1] = rad.y;
    smem[3 * t + 2] = rad.z;
    __syncthreads();
    if (valid && s == 0) {
      V3 acc = V3{0.0f, 0.0f, 0.0f};
      for (uint32_t k = 0; k < ns; k++) {
        const float* q = smem + 3 * (t + k);
        acc = acc + V3{q[0], q[1], q[2]};
      }
      const float inv = 1.0f / (float)ns;
      acc = scale(acc, inv);
      float* o = hdr + 3 * ((size_t)py * fp.width + px);
      o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
    }
    return;
  }
  // rounds: the block's one pixel
  const uint32_t px = tile_x * tw, py = tile_y * th;
  if (t == 0 && px < fp.width && py < fp.height) {
    const float inv = 1.0f / (float)ns;
    sum = scale(sum, inv);
    float* o = hdr + 3 * ((size_t)py * fp.width + px);
    o[0] = sum.x; o[1] = sum.y; o[2] = sum.z;
  }
}

// render_tile_map with the map built from the scene and its tables.
template <class Map, int MAXB, bool SM>
MM3E_DEV void render_tile(const DevScene* __restrict__ scp, const Tables& tb, const FrameParams& fp,
                          float* __restrict__ hdr, uint32_t tw, uint32_t th, uint32_t tile_x, uint32_t tile_y,
                          float* smem) {
  render_tile_map<Map, MAXB, SM>(Map{scp, tb}, scp, fp, hdr, tw, th, tile_x, tile_y, smem);
}

// ---------------------------------------------------------------------------------------
// Ordered tiles: a launch of one block per tile whose block b renders tile order[b], so the
// tiles the previous frame found most expensive are dispatched first.  A frame whose cost sits
// in a few long tiles otherwise ends with those tiles running alone on an idle chip; started
// first, they overlap the cheap ones.  The block's first wave draws a ticket from one global
// counter when it starts, and the tile's last wave to finish draws another; the difference (how
// many tiles started or finished meanwhile) is the tile's cost for the next frame's order.  The
// waves of a tile finish independently (a per-tile count of finished waves, reset by the last
// one), so no wave waits at a barrier for the tile's slowest.  Every tile is rendered exactly
// as render_tile renders it, whatever the order.
// ---------------------------------------------------------------------------------------
struct TileOrder {
  const uint32_t* order;  // tile indices (ty * tiles_x + tx), most expensive first; null: index order
  uint32_t* cost;         // per tile: tickets drawn by other tiles while it ran
  uint32_t* ticket;       // running ticket counter (wraps; only differences are used)
  uint32_t* start;        // per tile: its start ticket (plain launches only)
  uint32_t* done;         // per tile: waves finished this frame (plain launches only; 0 between frames)
  uint32_t tiles_x, pad;
};

template <class Map, int MAXB, bool SM>
MM3E_DEV void render_ordered(const DevScene* __restrict__ scp, const Tables& tb, const FrameParams& fp,
                             float* __restrict__ hdr, uint32_t tw, uint32_t th, const TileOrder& o, float* smem,
                             uint32_t* /*unused*/) {
  // o.order null: tile index order, no cost bookkeeping.  One call site of render_tile either
  // way, so the kernel holds one inlined copy of the march and the shading.
  const bool ordered = o.order != nullptr;
  const uint32_t tile = ordered ? o.order[blockIdx.x] : blockIdx.x;
  // Device-scope atomics only: a __threadfence here writes back the XCD's L2 and costs more
  // than the frame (measured: c2 1.4 -> 2.2 ms).  The cost is a heuristic for the next order,
  // so the start ticket is published by an atomic exchange and read back by an atomic add of 0.
  if (ordered && threadIdx.x == 0) atomicExch(o.start + tile, atomicAdd(o.ticket, 1u));
  render_tile<Map, MAXB, SM>(scp, tb, fp, hdr, tw, th, tile % o.tiles_x, tile / o.tiles_x, smem);
  if (ordered && __lane_id() == 0) {  // every lane of the wave is here: render_tile returns convergent
    const uint32_t waves = (blockDim.x + 63u) / 64u;
    if (atomicAdd(o.done + tile, 1u) == waves - 1u) {  // the tile's last wave
      const uint32_t t0 = atomicAdd(o.start + tile, 0u);
      o.cost[tile] = atomicAdd(o.ticket, 1u) - t0;
      atomicExch(o.done + tile, 0u);
    }
  }
}

// ---------------------------------------------------------------------------------------
// Film: beauty, premultiplied foreground and coverage alpha through a thin lens
// (film.rs:63-225, camera.rs:71-102)
// ---------------------------------------------------------------------------------------
MM3E_DEV uint32_t film_hash(uint32_t v) {  // film.rs:218-224
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  return v ^ (v >> 16);
}

// One jittered aperture sample per grid cell, cells permuted per pixel (film.rs:200-216).
MM3E_DEV void aperture_sample(uint32_t px, uint32_t py, uint32_t sx, uint32_t sy, uint32_t n, float& ax, float& ay) {
  if (n == 1) {
    ax = 0.5f;
    ay = 0.5f;
    return;
  }
  const uint32_t pixel_seed = film_hash(px ^ film_hash(py));
  const uint64_t n64 = n;
  const uint32_t cx = (uint32_t)(((uint64_t)sx + (uint64_t)sy + (uint64_t)pixel_seed) % n64);
  const uint32_t cy = (uint32_t)(((uint64_t)sx + 2ull * (uint64_t)sy + (uint64_t)film_hash(pixel_seed)) % n64);
  const uint32_t seed = film_hash(pixel_seed ^ film_hash(sx) ^ film_hash(sy + 0xa511e9b3u));
  const float unit = 1.0f / 16777216.0f;
  ax = ((float)cx + (float)(seed >> 8) * unit) / (float)n;
  ay = ((float)cy + (float)(film_hash(seed) >> 8) * unit) / (float)n;
}

// Equal-area concentric square-to-disk map (camera.rs:105-118).
MM3E_DEV void concentric_disk(float s0, float s1, float& dx, float& dy) {
  const float x = 2.0f * s0 - 1.0f, y = 2.0f * s1 - 1.0f;
  if (x == 0.0f && y == 0.0f) {
    dx = 0.0f;
    dy = 0.0f;
    return;
  }
  const float quarter = 0.785398163397448309616f, half = 1.57079632679489661923f;
  float r, th;
  if (fabsf(x) > fabsf(y)) {
    r = x;
    th = quarter * (y / x);
  } else {
    r = y;
    th = half - quarter * (x / y);
  }
  // cos and sin of an f32 angle, computed in f64 and rounded once: the correctly rounded f32
  // value but for rare double roundings, as the reference's libm cosf / sinf return it
  dx = r * (float)cos((double)th);
  dy = r * (float)sin((double)th);
}

// Camera::ray_with_lens (camera.rs:71-102): the pinhole ray when the radius is zero, else a ray
// from a point of the lens disk through the pinhole ray's point on the focus plane.
MM3E_DEV void lens_ray(const FrameParams& fp, const FilmParams& lp, float sx, float sy, float ax, float ay, V3& ro,
                       V3& rd) {
  camera_ray(fp, sx, sy, ro, rd);
  if (lp.radius == 0.0f) return;
  float dx, dy;
  concentric_disk(ax, ay, dx, dy);
  const V3 eye = ld3(fp.eye);
  const V3 origin = eye + scale(ld3(fp.right), dx * lp.radius) + scale(ld3(fp.up), dy * lp.radius);
  const V3 focus = eye + scale(rd, lp.focus / dot(rd, ld3(fp.fwd)));
  ro = origin;
  rd = normalize(focus - origin);
}

// trace (lib.rs:1265-1282) that also returns the primary march's hit record.
template <class Map, int MAXB>
MM3E_DEV V3 trace_hit(const DevScene& sc, const Map& field, V3 ro, V3 rd, Hit& h) {
  h = march(sc, field, ro, rd);
  if (!h.hit) return sky(rd, ld3(sc.sun));
  return shade_level<Map, 0, MAXB>(sc, field, h, rd);
}

// sample_pixel (film.rs:140-198) over one tile: lane -> (pixel, sample) as render_tile; the
// first lane of each pixel sums its samples in sy-major order: beauty over every sample,
// foreground over the samples whose primary ray hit, alpha = hits / n^2.
template <class Map, int MAXB, bool SM>
MM3E_DEV void film_tile(const DevScene* __restrict__ scp, const Tables& tb, const FrameParams& fp,
                        const FilmParams& lp, float* __restrict__ beauty, float* __restrict__ fg,
                        float* __restrict__ alpha, uint32_t tw, uint32_t th, uint32_t tile_x, uint32_t tile_y,
                        float* smem) {
  const DevScene& sc = *scp;
  Map field{scp, tb};
  const uint32_t aa = sc.aa;
  const uint32_t ns = aa * aa;
  const uint32_t t = threadIdx.x;
  const uint32_t pl = t / ns, s = t - pl * ns;
  const uint32_t px = tile_x * tw + pl % tw, py = tile_y * th + pl / tw;
  const bool valid = pl < tw * th && px < fp.width && py < fp.height;
  V3 rad = V3{0.0f, 0.0f, 0.0f};
  bool hit = false;
  if (valid) {
    V3 ro, rd;
    if (ns == 1) {
      camera_ray(fp, (float)px + 0.5f, (float)py + 0.5f, ro, rd);  // the centre ray
    } else {
      const uint32_t sy = s / aa, sx = s - sy * aa;
      const float ix = (float)px + ((float)sx + 0.5f) / (float)aa;
      const float iy = (float)py + ((float)sy + 0.5f) / (float)aa;
      float ax, ay;
      aperture_sample(px, py, sx, sy, aa, ax, ay);
      lens_ray(fp, lp, ix, iy, ax, ay, ro, rd);
    }
    Hit h0;
    if constexpr (SM) rad = trace_sm<Map, MAXB, true>(sc, field, ro, rd, &h0);
    else rad = trace_hit<Map, MAXB>(sc, field, ro, rd, h0);
    hit = h0.hit;
  }
  smem[4 * t + 0] = rad.x;
  smem[4 * t + 1] = rad.y;
  smem[4 * t + 2] = rad.z;
  smem[4 * t + 3] = hit ? 1.0f : 0.0f;
  __syncthreads();
  if (valid && s == 0) {
    V3 b = V3{0.0f, 0.0f, 0.0f}, f = V3{0.0f, 0.0f, 0.0f};
    uint32_t hits = 0;
    for (uint32_t k = 0; k < ns; k++) {
      const float* q = smem + 4 * (t + k);
      const V3 r = V3{q[0], q[1], q[2]};
      b = b + r;
      if (q[3] != 0.0f) {
        f = f + r;
        hits++;
      }
    }
    const float inv = 1.0f / (float)ns;
    const size_t i = (size_t)py * fp.width + px;
    const V3 bo = scale(b, inv), fo = scale(f, inv);
    beauty[3 * i] = bo.x;
    beauty[3 * i + 1] = bo.y;
    beauty[3 * i + 2] = bo.z;
    fg[3 * i] = fo.x;
    fg[3 * i + 1] = fo.y;
    fg[3 * i + 2] = fo.z;
    alpha[i] = (float)hits * inv;
  }
}

// ---------------------------------------------------------------------------------------
// G-buffer for reprojection (lib.rs:951-1030) and the re-rendered holes of the hybrid
// reprojection (lib.rs:1043-1080): one centre sample per pixel straight to display colour
// (exposure, ACES, gamma, 8 bit; no bloom), the primary hit distance and the moving object the
// hit belongs to.
// ---------------------------------------------------------------------------------------
struct GMover {  // (centre, radius) of a tagged moving object
  float c[3], r;
};

MM3E_DEV uint8_t to_u8(float c);

MM3E_DEV uchar4 display_px(V3 c, float exposure) {  // shade::gamma(shade::aces(c * exposure)).clamp01()
  c = scale(c, exposure);
  V3 g = V3{gamma1(aces1(c.x)), gamma1(aces1(c.y)), gamma1(aces1(c.z))};
  g = clamp01(g);
  return uchar4{to_u8(g.x), to_u8(g.y), to_u8(g.z), (uint8_t)255};
}

// Lanes map to pixels of a 16x16 tile (list == nullptr) or to list[blockIdx.x*256 + lane]
// (pixel indices of the frame; *n_list of them, read on the device so the list's producer needs
// no round trip to the host).  depth / obj may be nullptr (hole re-rendering).
template <class Map, int MAXB, bool SM>
MM3E_DEV void gbuffer_tile(const DevScene* __restrict__ scp, const FrameParams& fp, float exposure,
                           const GMover* __restrict__ movers, uint32_t n_movers, const uint32_t* __restrict__ list,
                           const uint32_t* __restrict__ n_list, uchar4* __restrict__ color,
                           float* __restrict__ depth, int32_t* __restrict__ obj) {
  const DevScene& sc = *scp;
  Map field{scp, global_tables(scp)};
  const uint32_t t = threadIdx.x;
  uint32_t x, y;
  bool valid;
  if (list) {
    const uint32_t k = blockIdx.x * 256u + t;
    valid = k < *n_list;
    const uint32_t p = valid ? list[k] : 0u;
    x = p % fp.width;
    y = p / fp.width;
  } else {
    x = blockIdx.x * 16u + (t & 15u);
    y = blockIdx.y * 16u + (t >> 4);
    valid = x < fp.width && y < fp.height;
  }
  if (!valid) return;
  V3 ro, rd;
  camera_ray(fp, (float)x + 0.5f, (float)y + 0.5f, ro, rd);
  Hit h;
  V3 hdr;
  if constexpr (SM) hdr = trace_sm<Map, MAXB, true>(sc, field, ro, rd, &h);
  else hdr = trace_hit<Map, MAXB>(sc, field, ro, rd, h);
  const size_t i = (size_t)y * fp.width + x;
  color[i] = display_px(hdr, exposure);
  if (depth) depth[i] = h.hit ? h.t : __builtin_inff();
  if (obj) {
    int32_t o = -1;
    if (h.hit) {
      for (uint32_t k = 0; k < n_movers; k++) {
        const GMover& m = movers[k];
        if (length(h.pos - V3{m.c[0], m.c[1], m.c[2]}) <= m.r + 0.05f) {
          o = (int32_t)k;
          break;
        }
      }
    }
    obj[i] = o;
  }
}

// ---------------------------------------------------------------------------------------
// GI bake (gi.rs:55-125 with the lib.rs:686-711 gather)
// ---------------------------------------------------------------------------------------
struct GiBakeParams {
  float lo[3], pad0;
  float cell[3], pad1;
  uint32_t dims[3];
  uint32_t rays_per_face;  // 1 or 5
  float dirs[6][5][4];     // axis_dirs (gi.rs:38-50): xyz + weight, computed on the host
};

template <class Map>
MM3E_DEV V3 gather(const DevScene& sc, const Map& field, V3 o, V3 d) {
  Hit h = march(sc, field, o, d);
  V3 sun = ld3(sc.sun);
  if (!h.hit) return sky_diffuse(d, sun);
  V3 n = hit_normal(sc, field, h.pos);
  const DMat& m = sc.mat[h.mat % sc.n_mat];
  V3 albedo = surface_albedo(m, h.pos);
  V3 c = ld3(m.emissive);
  const uint32_t nl = sc.n_light;
  for (uint32_t li = 0; li < nl; li++) {
    const DLight& l = sc.light[li];
    V3 ld;
    float dist;
    light_toward(l, h.pos, ld, dist);
    float ndl = lambert(n, ld);
    if (ndl > 0.0f) {
      V3 outg = brdf(n, ld, scale(d, -1.0f), albedo, m.metallic, m.roughness, m.specular);
      c = c + scale(cmul(outg, ld3(l.color)), light_atten(l, dist));
    }
  }
  return c;
}

// One lane per (probe, face, ray) of probes [first_probe, first_probe + n_probes); radiance is
// written to out[((probe - first_probe)*6 + face)*R + ray].
template <class Map>
MM3E_DEV void gi_gather_rays(const DevScene* __restrict__ scp, const Tables& tb, const GiBakeParams& bp,
                             float4* __restrict__ out, uint32_t first_probe, uint32_t n_probes) {
  const uint32_t R = bp.rays_per_face;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t per_probe = 6 * R;
  const uint32_t local = gid / per_probe;
  if (local >= n_probes) return;
  const uint32_t probe = first_probe + local;
  const uint32_t rem = gid - local * per_probe;
  const uint32_t face = rem / R, ray = rem - face * R;
  const uint32_t nx = bp.dims[0], ny = bp.dims[1];
  const uint32_t i = probe % nx, j = (probe / nx) % ny, k = probe / (nx * ny);
  V3 o = V3{bp.lo[0], bp.lo[1], bp.lo[2]} + V3{(float)i * bp.cell[0], (float)j * bp.cell[1], (float)k * bp.cell[2]};
  V3 d = V3{bp.dirs[face][ray][0], bp.dirs[face][ray][1], bp.dirs[face][ray][2]};
  Map field{scp, tb};
  V3 c = gather(*scp, field, o, d);
  out[((size_t)local * 6 + face) * R + ray] = float4{c.x, c.y, c.z, 0.0f};
}

MM3E_DEV void gi_reduce(const GiBakeParams& bp, const float4* __restrict__ rays, float4* __restrict__ cubes,
                        uint32_t n_probes) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n_probes * 6) return;
  const uint32_t face = gid % 6;
  const uint32_t R = bp.rays_per_face;
  V3 acc = V3{0.0f, 0.0f, 0.0f};
  float wsum = 0.0f;
  for (uint32_t r = 0; r < R; r++) {
    float4 v = rays[(size_t)gid * R + r];
    float w = bp.dirs[face][r][3];
    acc = acc + scale(V3{v.x, v.y, v.z}, w);
    wsum = wsum + w;
  }
  V3 c = scale(acc, 1.0f / rmax(wsum, 1e-4f));
  cubes[gid] = float4{c.x, c.y, c.z, 0.0f};
}

// ---------------------------------------------------------------------------------------
// Post (post.rs:36-115): bright-pass fused into the horizontal blur, vertical blur fused
// with bloom add + exposure + ACES + gamma + 8-bit quantisation.
// ---------------------------------------------------------------------------------------
struct PostParams {
  uint32_t width, height, radius, bloom;
  float threshold, intensity, exposure, pad;
};

MM3E_DEV V3 bright_pass(V3 c, float thr) {
  float l = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
  if (l > thr) return scale(c, (l - thr) / rmax(l, 1e-4f));
  return V3{0.0f, 0.0f, 0.0f};
}

// Block = 256 lanes along x of one row; LDS holds the bright-pass of the row segment + halo.
MM3E_DEV void bloom_h(const PostParams& pp, const float* __restrict__ kern, const float* __restrict__ hdr,
                      float* __restrict__ tmp, float* lds, int x0, int y) {
  const int W = (int)pp.width, r = (int)pp.radius;
  const int span = (int)blockDim.x + 2 * r;
  for (int i = threadIdx.x; i < span; i += blockDim.x) {
    int sx = x0 - r + i;
    sx = sx < 0 ? 0 : (sx > W - 1 ? W - 1 : sx);
    const float* s = hdr + 3 * ((size_t)y * W + sx);
    V3 b = bright_pass(V3{s[0], s[1], s[2]}, pp.threshold);
    lds[3 * i + 0] = b.x; lds[3 * i + 1] = b.y; lds[3 * i + 2] = b.z;
  }
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= W) return;
  V3 acc = V3{0.0f, 0.0f, 0.0f};
  const int n = 2 * r + 1;
  for (int k = 0; k < n; k++) {
    // LDS slot i holds the bright-pass of column clamp(x0 - r + i), so tap k of column x
    // (source column clamp(x + k - r)) is slot x - x0 + k
    const float* q = lds + 3 * (x - x0 + k);
    acc = acc + scale(V3{q[0], q[1], q[2]}, kern[k]);
  }
  float* o = tmp + 3 * ((size_t)y * W + x);
  o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
}

// bloom_h for radii whose LDS row segment + halo would not fit a block's LDS: every lane reads
// its 2r+1 taps from the HDR image through the cache and takes their bright-pass itself (the
// same operations in the same order, so the same sums).
MM3E_DEV void bloom_h_wide(const PostParams& pp, const float* __restrict__ kern, const float* __restrict__ hdr,
                           float* __restrict__ tmp, int x0, int y) {
  const int W = (int)pp.width, r = (int)pp.radius;
  const int x = x0 + threadIdx.x;
  if (x >= W) return;
  V3 acc = V3{0.0f, 0.0f, 0.0f};
  const int n = 2 * r + 1;
  for (int k = 0; k < n; k++) {
    int sx = x + k - r;
    sx = sx < 0 ? 0 : (sx > W - 1 ? W - 1 : sx);
    const float* s = hdr + 3 * ((size_t)y * W + sx);
    acc = acc + scale(bright_pass(V3{s[0], s[1], s[2]}, pp.threshold), kern[k]);
  }
  float* o = tmp + 3 * ((size_t)y * W + x);
  o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
}

MM3E_DEV uint8_t to_u8(float c) {  // framebuffer.rs:102-113: (clamp(c,0,1)*255 + 0.5) as u8
  float v = rclamp(c, 0.0f, 1.0f) * 255.0f + 0.5f;
  return (uint8_t)(v >= 255.0f ? 255.0f : v);
}

MM3E_DEV uchar4 tonemap_px(const PostParams& pp, V3 c) {
  c = scale(c, pp.exposure);
  V3 g = V3{gamma1(aces1(c.x)), gamma1(aces1(c.y)), gamma1(aces1(c.z))};
  g = clamp01(g);
  return uchar4{to_u8(g.x), to_u8(g.y), to_u8(g.z), (uint8_t)255};
}

// One lane per pixel; the vertical taps read the horizontally blurred image through L2.
MM3E_DEV void bloom_v_resolve(const PostParams& pp, const float* __restrict__ kern, const float* __restrict__ hdr,
                              const float* __restrict__ tmp, uchar4* __restrict__ rgba, int x0, int y) {
  const int W = (int)pp.width, H = (int)pp.height, r = (int)pp.radius;
  const int x = x0 + (int)threadIdx.x;
  if (x >= W) return;
  V3 acc = V3{0.0f, 0.0f, 0.0f};
  const int n = 2 * r + 1;
  for (int k = 0; k < n; k++) {
    int sy = y + k - r;
    sy = sy < 0 ? 0 : (sy > H - 1 ? H - 1 : sy);
    const float* q = tmp + 3 * ((size_t)sy * W + x);
    acc = acc + scale(V3{q[0], q[1], q[2]}, kern[k]);
  }
  const float* s = hdr + 3 * ((size_t)y * W + x);
  V3 c = V3{s[0], s[1], s[2]} + scale(acc, pp.intensity);
  rgba[(size_t)y * W + x] = tonemap_px(pp, c);
}

MM3E_DEV void resolve_only(const PostParams& pp, const float* __restrict__ hdr, uchar4* __restrict__ rgba, int x0,
                           int y) {
  const int W = (int)pp.width;
  const int x = x0 + (int)threadIdx.x;
  if (x >= W) return;
  const float* s = hdr + 3 * ((size_t)y * W + x);
  rgba[(size_t)y * W + x] = tonemap_px(pp, V3{s[0], s[1], s[2]});
}

// ---------------------------------------------------------------------------------------
// parity probes
// ---------------------------------------------------------------------------------------
// lib.rs:1401-1407 blue -> green -> yellow -> red heatmap of the march-step count.
MM3E_DEV V3 heatmap(float t) {
  t = rclamp(t, 0.0f, 1.0f);
  return V3{rclamp(t * 3.0f - 1.5f, 0.0f, 1.0f), rclamp(1.0f - rabs(t * 3.0f - 1.3f), 0.0f, 1.0f),
            rclamp(1.0f - t * 2.2f, 0.0f, 1.0f)};
}

// lib.rs:1178-1218 aov_pixel: a debug render mode (1 Normal, 2 Depth, 3 Ao, 4 Steps,
// 5 Albedo), one centre sample per pixel written straight to RGBA8 (no HDR buffer, no post).
template <class Map>
MM3E_DEV void aov_mode_pixel(const DevScene* __restrict__ scp, const FrameParams& fp, uchar4* __restrict__ rgba,
                             uint32_t mode) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= fp.width) return;
  const DevScene& sc = *scp;
  Map field{scp, global_tables(scp)};
  V3 ro, rd;
  camera_ray(fp, (float)x + 0.5f, (float)y + 0.5f, ro, rd);
  const Hit h = march(sc, field, ro, rd);
  V3 v = V3{0.0f, 0.0f, 0.0f};
  if (mode == 1) {
    if (h.hit) v = scale(hit_normal(sc, field, h.pos), 0.5f) + V3{0.5f, 0.5f, 0.5f};
  } else if (mode == 2) {
    const float d = rclamp(h.t / sc.max_dist, 0.0f, 1.0f);
    v = V3{1.0f - d, 1.0f - d, 1.0f - d};
  } else if (mode == 3) {
    float a = 1.0f;
    if (h.hit) a = ambient_occlusion(sc, field, h.pos, hit_normal(sc, field, h.pos));
    v = V3{a, a, a};
  } else if (mode == 4) {
    v = heatmap((float)h.steps / (float)sc.max_steps);
  } else if (mode == 5) {
    if (h.hit) v = surface_albedo(sc.mat[h.mat % sc.n_mat], h.pos);
  }
  rgba[(size_t)y * fp.width + x] = uchar4{to_u8(v.x), to_u8(v.y), to_u8(v.z), (uint8_t)255};
}

template <class Map>
MM3E_DEV void primary_aov(const DevScene* __restrict__ scp, const FrameParams& fp, uint8_t* __restrict__ hit,
                          uint32_t* __restrict__ mat, float* __restrict__ depth, uint32_t* __restrict__ steps,
                          float* __restrict__ nrm) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= fp.width) return;
  Map field{scp, global_tables(scp)};
  V3 ro, rd;
  camera_ray(fp, (float)x + 0.5f, (float)y + 0.5f, ro, rd);
  Hit h = march(*scp, field, ro, rd);
  const size_t i = (size_t)y * fp.width + x;
  hit[i] = h.hit ? 1 : 0;
  mat[i] = h.hit ? h.mat : 0xffffffffu;
  depth[i] = h.hit ? dot(h.pos - ld3(fp.eye), ld3(fp.fwd)) : __builtin_inff();
  if (steps) steps[i] = h.steps;
  if (nrm) {  // the hit's world normal, zero on a miss (film.rs:34-35, 119-122)
    const V3 n = h.hit ? hit_normal(*scp, field, h.pos) : V3{0.0f, 0.0f, 0.0f};
    nrm[3 * i] = n.x;
    nrm[3 * i + 1] = n.y;
    nrm[3 * i + 2] = n.z;
  }
}

template <class Map>
MM3E_DEV void field_probe(const DevScene* __restrict__ scp, const float* __restrict__ pts, uint32_t n,
                          float* __restrict__ dist, uint32_t* __restrict__ mat) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Map field{scp, global_tables(scp)};
  Field f = field(V3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
  dist[i] = f.dist;
  mat[i] = f.mat;
}

template <class Map>
MM3E_DEV void march_probe(const DevScene* __restrict__ scp, const float* __restrict__ rays, uint32_t n,
                          float* __restrict__ t, uint8_t* __restrict__ hit, uint32_t* __restrict__ mat,
                          float* __restrict__ nrm, uint32_t* __restrict__ steps) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Map field{scp, global_tables(scp)};
  const float* r = rays + 6 * i;
  Hit h = march(*scp, field, V3{r[0], r[1], r[2]}, V3{r[3], r[4], r[5]});
  V3 nn = h.hit ? hit_normal(*scp, field, h.pos) : V3{0.0f, 0.0f, 0.0f};
  t[i] = h.t;
  hit[i] = h.hit ? 1 : 0;
  mat[i] = h.mat;
  nrm[3 * i] = nn.x; nrm[3 * i + 1] = nn.y; nrm[3 * i + 2] = nn.z;
  steps[i] = h.steps;
}

}  // namespace mm3e
