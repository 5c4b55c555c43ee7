// dev_wavefront.h — a frame as stages over compacted ray queues, against step-count divergence.
//
// In the one-lane-per-sample kernels a wave lives as long as its slowest ray: lanes whose march
// ended idle while a grazing ray takes its 160 steps, and miss lanes idle through their
// neighbours' shading.  Here the frame is split:
//   wf_march_first   one lane per (pixel, AA sample) in the tile layout of render_tile; each
//                    march runs at most `budget` field evaluations.  Misses write sky(dir) to
//                    their sample slot, hits go to the hit queue, unfinished marches are pushed
//                    with their march state onto a continuation queue.
//   wf_march_cont    continuation pass k: the rays of queue k-1, packed 64 to a wave, run
//                    another `budget` evaluations (the last pass: to the end) and again split
//                    into sky / hit queue / queue k.
//   wf_shade         the hit queue, packed: shade_hit with its reflection chain (lib.rs:1286-1360),
//                    nested loops or the per-lane state machine, into the sample slot.
//   wf_resolve       per pixel the mean of its samples in the reference's sy-major order.
// A march resumed from its saved state takes exactly the steps it would have taken in one go
// (every field of MarchState the ray does not fix is saved bit for bit; the ray is recomputed
// from the slot), and shading reads the same hit record, so every sample equals the one-kernel
// render bit for bit; queue order (atomics) decides only which lane does the work.
// Queue pushes are wave-aggregated: one atomic per wave and push.
#pragma once
#include "dev_render.h"

namespace mm3e {

enum : uint32_t { WF_MAX_PASSES = 8, WF_STATE_WORDS = 15, WF_HIT_WORDS = 4 };

struct WfFrame {
  float* mq[2];        // continuation queues (pass k writes mq[k & 1]): WF_STATE_WORDS x cap, SoA
  uint32_t* mslot[2];  // their sample slots
  float* hq;           // hit queue: t, pos.x, pos.y, pos.z (WF_HIT_WORDS x cap, SoA)
  uint32_t* hslot;     // hit queue sample slots
  uint32_t* hmat;      // hit queue materials
  float* rad;          // per sample slot radiance: 3 x cap, SoA
  uint32_t* cnt;       // cnt[k]: entries continuation pass k pushed (k = 0: the first pass);
                       // cnt[WF_MAX_PASSES]: hit queue entries
  uint32_t cap;        // sample slots = tiles * tw * th * aa * aa
  uint32_t tw, th, tiles_x;
};

// Wave-aggregated append: every active lane with `want` gets a distinct index in
// [old, old + count) of the queue whose length is *counter; one atomic per wave.
MM3E_DEV uint32_t wave_push(uint32_t* counter, bool want) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(want);
  if (m == 0) return 0u;
  const uint32_t before =
      __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0u;
  if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__builtin_popcountll(m));
  return (uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader) + before;
}

// Sample slot g = tile * (tw*th*aa*aa) + lane of render_tile's block for that tile.
MM3E_DEV bool wf_slot_ray(const FrameParams& fp, const WfFrame& w, uint32_t aa, uint32_t g, V3& ro, V3& rd) {
  const uint32_t ns = aa * aa;
  const uint32_t bs = w.tw * w.th * ns;
  const uint32_t tile = g / bs, t = g - tile * bs;
  const uint32_t pl = t / ns, s = t - pl * ns;
  const uint32_t px = (tile % w.tiles_x) * w.tw + pl % w.tw, py = (tile / w.tiles_x) * w.th + pl / w.tw;
  if (px >= fp.width || py >= fp.height) return false;
  const uint32_t sy = s / aa, sx = s - sy * aa;
  const float ox = ((float)sx + 0.5f) / (float)aa;
  const float oy = ((float)sy + 0.5f) / (float)aa;
  camera_ray(fp, (float)px + ox, (float)py + oy, ro, rd);
  return true;
}

MM3E_DEV void wf_save(float* q, uint32_t cap, uint32_t i, const MarchState& m, uint32_t mst) {
  const bool mid = mst == MARCH_EVAL_MID;  // the guard's fields are live only during its probe
  const float v[WF_STATE_WORDS] = {m.t,
                                   m.prev_radius,
                                   m.step_len,
                                   m.t_prev,
                                   m.d_prev,
                                   m.omega,
                                   m.subitize,
                                   mid ? m.fm.dist : 0.0f,
                                   mid ? m.radius_m : 0.0f,
                                   mid ? m.eps_m : 0.0f,
                                   mid ? m.gap_lo : 0.0f,
                                   mid ? m.gap_hi : 0.0f,
                                   __uint_as_float(m.i),
                                   __uint_as_float(mid ? m.fm.mat : 0u),
                                   __uint_as_float(mst | (m.boost << 4) | ((m.secant_on ? 1u : 0u) << 8))};
#pragma unroll
  for (uint32_t k = 0; k < WF_STATE_WORDS; k++) q[(size_t)k * cap + i] = v[k];
}

// Restores the state and the point its next step evaluates (march_next / the guard's midpoint).
MM3E_DEV uint32_t wf_load(const float* q, uint32_t cap, uint32_t i, V3 ro, V3 rd, MarchState& m, V3& p) {
  float v[WF_STATE_WORDS];
#pragma unroll
  for (uint32_t k = 0; k < WF_STATE_WORDS; k++) v[k] = q[(size_t)k * cap + i];
  m.ro = ro;
  m.rd = rd;
  m.t = v[0];
  m.prev_radius = v[1];
  m.step_len = v[2];
  m.t_prev = v[3];
  m.d_prev = v[4];
  m.omega = v[5];
  m.subitize = v[6];
  m.fm.dist = v[7];
  m.radius_m = v[8];
  m.eps_m = v[9];
  m.gap_lo = v[10];
  m.gap_hi = v[11];
  m.i = __float_as_uint(v[12]);
  m.fm.mat = __float_as_uint(v[13]);
  const uint32_t bits = __float_as_uint(v[14]);
  m.boost = (bits >> 4) & 15u;
  m.secant_on = ((bits >> 8) & 1u) != 0u;
  const uint32_t mst = bits & 15u;
  p = m.ro + scale(m.rd, mst == MARCH_EVAL_MID ? 0.5f * (m.gap_lo + m.gap_hi) : m.t);
  return mst;
}

// Runs at most `budget` evaluations (0: to the end) of a march, then files the sample: miss ->
// sky into its slot, hit -> hit queue, unfinished -> continuation queue `out`.  Called by
// every lane of the wave (`live` false for lanes without a sample).
template <class Map>
MM3E_DEV void wf_march_run(const DevScene& sc, const Map& field, const WfFrame& w, bool live, uint32_t g, V3 rd,
                           MarchState& m, uint32_t mst, V3 q, Hit& h, uint32_t budget, float* out,
                           uint32_t* out_slot, uint32_t* out_cnt) {
  if (live) {
    uint32_t n = 0;
    while (mst != MARCH_DONE && (budget == 0u || n < budget)) {
      mst = march_step(sc, m, mst, field(q), h, q);
      n++;
    }
  }
  const bool done = live && mst == MARCH_DONE;
  const bool is_hit = done && h.hit;
  if (done && !h.hit) {
    const V3 c = sky(rd, ld3(sc.sun));
    w.rad[g] = c.x;
    w.rad[(size_t)w.cap + g] = c.y;
    w.rad[2 * (size_t)w.cap + g] = c.z;
  }
  const uint32_t hi = wave_push(w.cnt + WF_MAX_PASSES, is_hit);
  if (is_hit) {
    w.hq[hi] = h.t;
    w.hq[(size_t)w.cap + hi] = h.pos.x;
    w.hq[2 * (size_t)w.cap + hi] = h.pos.y;
    w.hq[3 * (size_t)w.cap + hi] = h.pos.z;
    w.hslot[hi] = g;
    w.hmat[hi] = h.mat;
  }
  const bool pend = live && mst != MARCH_DONE;
  const uint32_t qi = wave_push(out_cnt, pend);
  if (pend) {
    wf_save(out, w.cap, qi, m, mst);
    out_slot[qi] = g;
  }
}

// First pass: block = one tile of render_tile's layout (blockDim.x = tw*th*aa*aa).
template <class Map>
MM3E_DEV void wf_march_first(const DevScene* __restrict__ scp, const Tables& tb, const FrameParams& fp,
                             const WfFrame& w, uint32_t budget) {
  const DevScene& sc = *scp;
  Map field{scp, tb};
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  V3 ro, rd;
  const bool live = wf_slot_ray(fp, w, sc.aa, g, ro, rd);
  MarchState m;
  Hit h;
  V3 q;
  uint32_t mst = MARCH_DONE;
  if (live) mst = march_init(sc, m, ro, rd, h, q);
  wf_march_run(sc, field, w, live, g, rd, m, mst, q, h, budget, w.mq[0], w.mslot[0], w.cnt);
}

// Continuation pass k >= 1 over queue k-1; grid-stride over its entries, 64 to a wave.
template <class Map>
MM3E_DEV void wf_march_cont(const DevScene* __restrict__ scp, const Tables& tb, const FrameParams& fp,
                            const WfFrame& w, uint32_t k, uint32_t budget) {
  const DevScene& sc = *scp;
  Map field{scp, tb};
  const uint32_t n = w.cnt[k - 1];
  const float* in = w.mq[(k - 1) & 1];
  const uint32_t* in_slot = w.mslot[(k - 1) & 1];
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = (uint64_t)wave * 64u; base < n; base += (uint64_t)n_waves * 64u) {
    const uint32_t i = (uint32_t)base + lane;
    const bool live = i < n;
    uint32_t g = 0, mst = MARCH_DONE;
    MarchState m;
    Hit h;
    V3 ro, rd, q;
    if (live) {
      g = in_slot[i];
      wf_slot_ray(fp, w, sc.aa, g, ro, rd);
      mst = wf_load(in, w.cap, i, ro, rd, m, q);
    }
    wf_march_run(sc, field, w, live, g, rd, m, mst, q, h, budget, w.mq[k & 1], w.mslot[k & 1], w.cnt + k);
  }
}

// shade_hit from a finished primary march, through the per-lane state machine.
template <class Map, int MAXB>
MM3E_DEV V3 shade_sm(const DevScene& sc, const Map& field, const Hit& h, V3 rd) {
  PathState ps;
  LevelStack<MAXB> stk;
  ps.level = 0;
  ps.rd = rd;
  ps.phase = PH_MARCH;
  ps.mst = MARCH_DONE;
  ps.h = h;
  path_run<Map, MAXB, false>(sc, field, ps, stk, true, nullptr);
  return ps.radiance;
}

// Shades the hit queue; grid-stride over its entries, 64 to a wave.
template <class Map, int MAXB, bool SM>
MM3E_DEV void wf_shade(const DevScene* __restrict__ scp, const Tables& tb, const FrameParams& fp, const WfFrame& w) {
  const DevScene& sc = *scp;
  Map field{scp, tb};
  const uint32_t n = w.cnt[WF_MAX_PASSES];
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = (uint64_t)wave * 64u; base < n; base += (uint64_t)n_waves * 64u) {
    const uint32_t i = (uint32_t)base + lane;
    if (i < n) {
      const uint32_t g = w.hslot[i];
      Hit h;
      h.hit = true;
      h.t = w.hq[i];
      h.pos = V3{w.hq[(size_t)w.cap + i], w.hq[2 * (size_t)w.cap + i], w.hq[3 * (size_t)w.cap + i]};
      h.mat = w.hmat[i];
      h.steps = 0;
      V3 ro, rd;
      wf_slot_ray(fp, w, sc.aa, g, ro, rd);
      V3 c;
      if constexpr (SM) c = shade_sm<Map, MAXB>(sc, field, h, rd);
      else c = shade_level<Map, 0, MAXB>(sc, field, h, rd);
      w.rad[g] = c.x;
      w.rad[(size_t)w.cap + g] = c.y;
      w.rad[2 * (size_t)w.cap + g] = c.z;
    }
  }
}

}  // namespace mm3e
