// dev_math.h — device-side vector math, SDF primitives, the world field, the sphere tracer
// and the shading model of the MM3E render path, for gfx950.
//
// Semantics follow the reference CPU renderer (Rust f32), which the oracle restates:
//   vec.rs, sdf.rs, dual.rs, volume.rs, march.rs, shade.rs, color.rs (mm3e-kit/src)
//   lib.rs:215-557, 820-835, 1286-1398, accel.rs, gi.rs:138-187 (mm3e-orchestrator/src)
// All code is compiled with -ffp-contract=off so each operation rounds once, in source
// order; sqrt and division are IEEE correctly rounded (the HIP defaults).  Rust's
// f32::min/max are IEEE minNum/maxNum (v_min_f32/v_max_f32); f32::clamp is spelled out as
// compare-select so NaN and signed zero behave as in the reference.
#pragma once
#include "cr_trig.h"
#include "dev_scene.h"

#ifndef MM3E_DEV
#define MM3E_DEV __device__ __forceinline__
#endif

namespace mm3e {

// ---------------------------------------------------------------------------------------
// vec.rs
// ---------------------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
MM3E_DEV V3 mk(float x, float y, float z) { return V3{x, y, z}; }
MM3E_DEV V3 ld3(const float* a) { return V3{a[0], a[1], a[2]}; }
MM3E_DEV V3 splat(float v) { return V3{v, v, v}; }
MM3E_DEV V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MM3E_DEV V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MM3E_DEV V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
MM3E_DEV V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
MM3E_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MM3E_DEV V3 cross(V3 a, V3 o) { return V3{a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x}; }
MM3E_DEV float length_sq(V3 a) { return dot(a, a); }
// Correctly rounded f32 square root, bit-identical to __builtin_sqrtf for every input: the
// hardware estimate (within one ulp) replaced by a neighbour when that neighbour's residual
// x - n*s says so — the compiler's IEEE sequence without its tiny-input rescaling (taken only
// below 2^-96, where the residuals would lose bits; zero and negative inputs go there too) and
// without its +-0 / +inf class fix-up, which the two comparisons already leave right (a NaN
// residual compares false).  Six fewer VALU instructions per root on the common path.  An
// argument known at compile time (a generated kernel's object constants) takes __builtin_sqrtf,
// which the compiler folds to the correctly rounded constant; the hardware estimate would
// instead be computed at kernel start and held in a register, or spilled, for the whole kernel.
MM3E_DEV float crsqrt(float x) {
  if (__builtin_constant_p(x)) return __builtin_sqrtf(x);
  if (__builtin_expect(!(x >= 0x1p-96f), 0)) return __builtin_sqrtf(x);  // tiny, zero, negative, NaN
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sdn = __uint_as_float(__float_as_uint(s) - 1u);
  const float sup = __uint_as_float(__float_as_uint(s) + 1u);
  const float rdn = __builtin_fmaf(-sdn, s, x);
  const float rup = __builtin_fmaf(-sup, s, x);
  const float r = rdn <= 0.0f ? sdn : s;
  return rup > 0.0f ? sup : r;
}

MM3E_DEV float length(V3 a) { return crsqrt(dot(a, a)); }
MM3E_DEV V3 normalize(V3 a) {  // vec.rs:37-44
  float len = length(a);
  if (len > 1e-12f) return scale(a, 1.0f / len);
  return V3{0.0f, 0.0f, 0.0f};
}
MM3E_DEV V3 cmul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
MM3E_DEV float rmin(float a, float b) { return __builtin_fminf(a, b); }
MM3E_DEV float rmax(float a, float b) { return __builtin_fmaxf(a, b); }
MM3E_DEV float rclamp(float x, float lo, float hi) {
  x = x < lo ? lo : x;
  return x > hi ? hi : x;
}
MM3E_DEV float rabs(float x) { return __builtin_fabsf(x); }
MM3E_DEV V3 vabs(V3 a) { return V3{rabs(a.x), rabs(a.y), rabs(a.z)}; }
MM3E_DEV V3 vmin(V3 a, V3 b) { return V3{rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)}; }
MM3E_DEV V3 vmax(V3 a, V3 b) { return V3{rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)}; }
MM3E_DEV V3 max_scalar(V3 a, float m) { return V3{rmax(a.x, m), rmax(a.y, m), rmax(a.z, m)}; }
MM3E_DEV float max_element(V3 a) { return rmax(rmax(a.x, a.y), a.z); }
MM3E_DEV V3 clamp01(V3 a) { return V3{rclamp(a.x, 0.0f, 1.0f), rclamp(a.y, 0.0f, 1.0f), rclamp(a.z, 0.0f, 1.0f)}; }
MM3E_DEV V3 mix(V3 a, V3 b, float t) { return scale(a, 1.0f - t) + scale(b, t); }  // vec.rs:75
MM3E_DEV V3 reflect(V3 v, V3 n) { return v - scale(n, 2.0f * dot(v, n)); }         // vec.rs:79
MM3E_DEV float powi2(float x) { return x * x; }
MM3E_DEV float powi5(float x) {  // f32::powi(5) as LLVM expands it: x * ((x*x)*(x*x))
  float x2 = x * x;
  return x * (x2 * x2);
}
// f32::signum: +-1 by the sign bit (NaN passes through).
MM3E_DEV float signum(float x) {
  if (x != x) return x;
  return __builtin_signbit(x) ? -1.0f : 1.0f;
}

// ---------------------------------------------------------------------------------------
// sdf.rs primitives (local space)
// ---------------------------------------------------------------------------------------
struct Field {
  float dist;
  uint32_t mat;
};

MM3E_DEV float sd_sphere(V3 p, float r) { return length(p) - r; }
MM3E_DEV float sd_box(V3 p, V3 half) {
  V3 q = vabs(p) - half;
  return length(max_scalar(q, 0.0f)) + rmin(max_element(q), 0.0f);
}
MM3E_DEV float sd_round_box(V3 p, V3 half, float radius) { return sd_box(p, half - splat(radius)) - radius; }
MM3E_DEV float sd_plane(V3 p, V3 n, float h) { return dot(p, n) + h; }
MM3E_DEV float sd_torus(V3 p, float major, float minor) {
  float q = crsqrt(p.x * p.x + p.z * p.z) - major;
  return crsqrt(q * q + p.y * p.y) - minor;
}
MM3E_DEV float sd_cylinder(V3 p, float h, float r) {
  float dx = crsqrt(p.x * p.x + p.z * p.z) - r;
  float dy = rabs(p.y) - h;
  float ox = rmax(dx, 0.0f), oy = rmax(dy, 0.0f);
  return rmin(rmax(dx, dy), 0.0f) + crsqrt(ox * ox + oy * oy);
}
MM3E_DEV float sd_capsule(V3 p, V3 a, V3 b, float r) {
  V3 pa = p - a, ba = b - a;
  float h = rclamp(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
  return length(pa - scale(ba, h)) - r;
}
MM3E_DEV float sd_round_cone(V3 p, float r1, float r2, float h) {
  float qx = crsqrt(p.x * p.x + p.z * p.z);
  float qy = p.y;
  float b = (r1 - r2) / rmax(h, 1e-6f);
  float a = crsqrt(rmax(1.0f - b * b, 0.0f));
  float k = qx * -b + qy * a;
  if (k < 0.0f) return crsqrt(qx * qx + qy * qy) - r1;
  if (k > a * h) return crsqrt(qx * qx + (qy - h) * (qy - h)) - r2;
  return qx * a + qy * b - r1;
}
MM3E_DEV float sd_ellipsoid(V3 p, V3 r) {
  float k0 = length(V3{p.x / r.x, p.y / r.y, p.z / r.z});
  if (k0 < 1e-6f) return -rmin(rmin(r.x, r.y), r.z);
  float k1 = length(V3{p.x / (r.x * r.x), p.y / (r.y * r.y), p.z / (r.z * r.z)});
  return k0 * (k0 - 1.0f) / rmax(k1, 1e-9f);
}
MM3E_DEV float sd_octahedron(V3 p, float s) {
  V3 q = vabs(p);
  return (q.x + q.y + q.z - s) * 0.5773503f;
}
MM3E_DEV float sd_hex_prism(V3 p, float r, float h) {
  const float kx = -0.8660254f, ky = 0.5f, kz = 0.57735f;
  V3 a = vabs(p);
  float t = 2.0f * rmin(kx * a.x + ky * a.z, 0.0f);
  a.x = a.x - t * kx;
  a.z = a.z - t * ky;
  float cl = rclamp(a.x, -kz * r, kz * r);
  float dx = crsqrt((a.x - cl) * (a.x - cl) + (a.z - r) * (a.z - r)) * signum(a.z - r);
  float dy = a.y - h;
  return rmin(rmax(dx, dy), 0.0f) + crsqrt(powi2(rmax(dx, 0.0f)) + powi2(rmax(dy, 0.0f)));
}

MM3E_DEV Field f_union(Field a, Field b) { return a.dist <= b.dist ? a : b; }
MM3E_DEV Field f_subtract(Field a, Field b) { return Field{rmax(a.dist, -b.dist), a.mat}; }
MM3E_DEV Field f_smooth_union(Field a, Field b, float k) {  // sdf.rs:152-160
  if (k <= 0.0f) return f_union(a, b);
  float h = rclamp(0.5f + 0.5f * (b.dist - a.dist) / k, 0.0f, 1.0f);
  float dist = b.dist * (1.0f - h) + a.dist * h - k * h * (1.0f - h);
  return Field{dist, h > 0.5f ? a.mat : b.mat};
}

// Prim::distance over the analytic kinds (lib.rs:222-233).
MM3E_DEV float prim_distance(uint32_t kind, const float* p, V3 l) {
  switch (kind) {
    case 0: return sd_sphere(l, p[0]);
    case 1: return sd_box(l, ld3(p));
    case 2: return sd_round_box(l, ld3(p), p[3]);
    case 3: return sd_torus(l, p[0], p[1]);
    case 4: return sd_cylinder(l, p[0], p[1]);
    case 5: return sd_capsule(l, ld3(p), ld3(p + 3), p[6]);
    case 6: return sd_round_cone(l, p[0], p[1], p[2]);
    case 7: return sd_ellipsoid(l, ld3(p));
    case 8: return sd_octahedron(l, p[0]);
    case 9: return sd_hex_prism(l, p[0], p[1]);
    case 10: return sd_plane(l, ld3(p), p[3]);
    default: return __builtin_inff();
  }
}

// volume.rs:194-228 conservative trilinear sample.
MM3E_DEV float vol_sample(const DVol& v, const float* __restrict__ data, V3 p) {
  if (!v.valid) return __builtin_inff();
  V3 lo = ld3(v.min), hi = ld3(v.hi);
  V3 c = vmin(vmax(p, lo), hi);
  float outside = length(p - c);
  V3 g = V3{(c.x - v.min[0]) / v.cell[0], (c.y - v.min[1]) / v.cell[1], (c.z - v.min[2]) / v.cell[2]};
  const uint32_t nx = v.dims[0], ny = v.dims[1], nz = v.dims[2];
  // (v.floor().max(0.0) as usize).min(n - 2): the float is >= 0 and, after the clamp, the
  // index fits u32 for any grid that fits in device memory.
  auto clampi = [](float x, uint32_t n) -> uint32_t {
    float f = rmax(__builtin_floorf(x), 0.0f);
    float lim = (float)(n - 2);
    return f >= lim ? n - 2 : (uint32_t)f;
  };
  uint32_t i0 = clampi(g.x, nx), j0 = clampi(g.y, ny), k0 = clampi(g.z, nz);
  float fx = rclamp(g.x - (float)i0, 0.0f, 1.0f);
  float fy = rclamp(g.y - (float)j0, 0.0f, 1.0f);
  float fz = rclamp(g.z - (float)k0, 0.0f, 1.0f);
  const float* d = data + v.offset;
  const uint64_t row = nx, slab = (uint64_t)nx * ny;
  const uint64_t b = (uint64_t)k0 * slab + (uint64_t)j0 * row + i0;
  float c000 = d[b], c100 = d[b + 1], c010 = d[b + row], c110 = d[b + row + 1];
  float c001 = d[b + slab], c101 = d[b + slab + 1], c011 = d[b + slab + row], c111 = d[b + slab + row + 1];
  float x00 = c000 + (c100 - c000) * fx;
  float x10 = c010 + (c110 - c010) * fx;
  float x01 = c001 + (c101 - c001) * fx;
  float x11 = c011 + (c111 - c011) * fx;
  float y0 = x00 + (x10 - x00) * fy;
  float y1 = x01 + (x11 - x01) * fy;
  float dd = y0 + (y1 - y0) * fz;
  if (!__builtin_isfinite(dd)) return __builtin_inff();
  if (outside > 0.0f) return rmax(outside, dd - outside);
  return dd;
}

// lib.rs:400-419 object-local point.
MM3E_DEV V3 local_point(const DObj& o, V3 p) {
  V3 v = p - ld3(o.pos);
  const uint32_t f = o.flags;
  // Rᵀ·v: component i is dot(column_i, v) (vec.rs:125-133, 216-221).  For the identity,
  // (1·x + 0·y) + 0·z == x for finite y, z (only the sign of a zero x can differ), and a unit
  // scale multiplies by exactly 1: both are skipped.
  V3 l = v;
  if (!(f & OBJ_IDENT_ROT))
    l = V3{(o.c0[0] * v.x + o.c0[1] * v.y) + o.c0[2] * v.z, (o.c1[0] * v.x + o.c1[1] * v.y) + o.c1[2] * v.z,
           (o.c2[0] * v.x + o.c2[1] * v.y) + o.c2[2] * v.z};
  if (!(f & OBJ_UNIT_SCALE)) l = scale(l, o.inv);
  if (f & (OBJ_MIRROR_X | OBJ_MIRROR_Y | OBJ_MIRROR_Z)) {
    if (f & OBJ_MIRROR_X) l.x = rabs(l.x);
    if (f & OBJ_MIRROR_Y) l.y = rabs(l.y);
    if (f & OBJ_MIRROR_Z) l.z = rabs(l.z);
  }
  if (f & OBJ_ELONGATE) {  // sdf.rs:202-204: p - p.clamp_to(-h, h)
    V3 h = ld3(o.elo);
    l = l - vmin(vmax(l, scale(h, -1.0f)), h);
  }
  if (f & OBJ_REPEAT) {  // sdf.rs:206-217 (Rust round = half away from zero)
    if (o.rep[0] > 0.0f) l.x = l.x - o.rep[0] * __builtin_roundf(l.x / o.rep[0]);
    if (o.rep[1] > 0.0f) l.y = l.y - o.rep[1] * __builtin_roundf(l.y / o.rep[1]);
    if (o.rep[2] > 0.0f) l.z = l.z - o.rep[2] * __builtin_roundf(l.z / o.rep[2]);
  }
  if (f & OBJ_TWIST) {  // sdf.rs:220-224
    float a = o.twist * l.y;
    float s, c;
    cr_sincos(a, s, c);
    l = V3{c * l.x - s * l.z, l.y, s * l.x + c * l.z};
  }
  if (f & OBJ_BEND) {  // sdf.rs:227-231
    float a = o.bend * l.x;
    float s, c;
    cr_sincos(a, s, c);
    l = V3{c * l.x - s * l.y, s * l.x + c * l.y, l.z};
  }
  return l;
}

// csg.rs:95-118 Expr::distance over the postfix program ops[0..n) at the object-local point l.
// The operand stack lives in per-lane scratch memory (volatile keeps the compiler from
// promoting it to registers, which would raise the register count of every kernel that can
// reach this code): CSG objects are rare and the common path stays unaffected.
MM3E_DEV float csg_distance(const DCsgOp* __restrict__ ops, uint32_t n, V3 l) {
  volatile float st[kCsgStack];
  uint32_t sp = 0;
  V3 pt = l;
  for (uint32_t i = 0; i < n; i++) {
    const DCsgOp& o = ops[i];
    switch (o.op) {
      case CSG_LEAF:
        st[sp++] = prim_distance(o.prim, o.p, pt);
        break;
      case CSG_ENTER: {  // vec.rs:216-221 Transform::to_local
        V3 v = pt - ld3(o.pos);
        pt = scale(V3{(o.c0[0] * v.x + o.c0[1] * v.y) + o.c0[2] * v.z, (o.c1[0] * v.x + o.c1[1] * v.y) + o.c1[2] * v.z,
                      (o.c2[0] * v.x + o.c2[1] * v.y) + o.c2[2] * v.z},
                   o.inv);
        break;
      }
      case CSG_LEAVE:
        st[sp - 1] = st[sp - 1] * o.scale;
        pt = l;
        break;
      case CSG_OFFSET:
        st[sp - 1] = st[sp - 1] - o.p[0];
        break;
      case CSG_SHELL:
        st[sp - 1] = rabs(st[sp - 1]) - o.p[0];
        break;
      default: {
        const float x = st[--sp], y = st[sp - 1];
        const float a = o.swap ? x : y, b = o.swap ? y : x;
        float r;
        if (o.op == CSG_UNION) r = rmin(a, b);
        else if (o.op == CSG_INTERSECT) r = rmax(a, b);
        else if (o.op == CSG_SUBTRACT) r = rmax(a, -b);
        else r = f_smooth_union(Field{a, 0u}, Field{b, 0u}, o.p[0]).dist;
        st[sp - 1] = r;
      }
    }
  }
  return st[0];
}

// lib.rs:380-397 Object::field.
// CSG: whether Prim::Csg objects can occur (kernels for scenes without them leave the
// evaluator out).
template <bool CSG>
MM3E_DEV float object_distance(const DevScene& sc, const DObj& o, V3 p) {
  V3 l = local_point(o, p);
  const uint32_t kind = o.flags & OBJ_KIND_MASK;
  float d;
  if (CSG && kind == 11) {
    d = csg_distance(sc.csg + __builtin_bit_cast(uint32_t, o.p[0]), __builtin_bit_cast(uint32_t, o.p[1]), l);
  } else if (kind == 12) {
    uint32_t vid = __builtin_bit_cast(uint32_t, o.p[0]);
    d = sc.n_vol == 0 ? __builtin_inff() : vol_sample(sc.vol[vid % sc.n_vol], sc.vol_data, l);
  } else {
    d = prim_distance(kind, o.p, l);
  }
  if (!(o.flags & OBJ_UNIT_SCALE)) d = d * o.scale;
  if (o.flags & OBJ_ROUND) d = d - o.round;
  if (o.flags & OBJ_ONION) d = rabs(d) - o.onion;
  return d;
}

// Provable no-op test for a far object.  When |p - c|^2 exceeds t^2 with
// t = (thr + r) * (1 + 2e-5) + 1e-6 > 0, the object's bounding-sphere lower bound
// lower = |p - c| - r (what `contrib` returns once lower > slack) satisfies lower >= thr with
// a relative margin far above the rounding of the squared test, so folding it in leaves the
// accumulator bit-identical.  The caller picks thr per combine mode:
//   union:        thr = max(acc, slack)          (lower >= acc keeps acc; ties keep acc)
//   smooth, k>0:  thr = max(acc + k, slack)      ((lower - acc)/k >= 1 -> h = 1 -> acc)
//   subtract:     thr = max(-acc, slack)         (max(acc, -lower) = acc)
// No sqrt and no smooth-union division are spent on objects that cannot matter.
MM3E_DEV bool far_noop(const DObj& o, V3 p, float thr) {
  const float t = (thr + o.bc_r) * 1.00002f + 1e-6f;
  V3 d = p - ld3(o.bc);
  return t > 0.0f && dot(d, d) > t * t;
}

// accel.rs:187-207 WorldPlan::contrib.  The bound's root is skipped where its test must fail:
// with t = (slack + r) * (1 - 2e-5) - 1e-6 > 0, |p - c|^2 < t^2 puts |p - c| - r below slack by
// a margin far above the rounding of the root and the subtraction, so the reference's
// `lower > slack` is false there and the exact field follows either way.
template <bool CSG>
MM3E_DEV Field contrib(const DevScene& sc, const DObj& o, V3 p) {
  if (o.flags & OBJ_BOUNDED) {
    const V3 d = p - ld3(o.bc);
    const float d2 = dot(d, d);
    const float t = (sc.slack + o.bc_r) * 0.99998f - 1e-6f;
    if (!(t > 0.0f && d2 < t * t)) {
      const float lower = crsqrt(d2) - o.bc_r;
      if (lower > sc.slack) return Field{lower, o.mat};
    }
  }
  return Field{object_distance<CSG>(sc, o, p), o.mat};
}

#ifdef MM3E_MAP_NOINLINE
#define MM3E_MAP_ATTR __device__ __noinline__
#else
#define MM3E_MAP_ATTR MM3E_DEV
#endif

// Segment::One (accel.rs:209-235): fold object o with its combine into the accumulator.
template <bool CSG>
MM3E_DEV void seg_one(const DevScene& sc, const DObj& o, uint32_t combine, float k, V3 p, Field& acc, bool& seeded) {
  if (seeded && (o.flags & (OBJ_BOUNDED | OBJ_SKIP))) {
    float thr = combine == 2 ? -acc.dist : (combine == 1 && k > 0.0f ? acc.dist + k : acc.dist);
    if (far_noop(o, p, rmax(thr, sc.slack))) return;
  }
  Field f = contrib<CSG>(sc, o, p);
  if (!seeded) {
    acc = f;
    seeded = true;
    return;
  }
  if (combine == 0) acc = f_union(acc, f);
  else if (combine == 1) acc = f_smooth_union(acc, f, k);
  else acc = f_subtract(acc, f);
}

// A union run whose members all have an exact SDF inside a true bounding sphere
// (SDF >= |p - c| - r): the tree walk can only skip objects that would not have won, so the
// run's result is the plain index-order minimum (strict < keeps the earlier object, and the
// seeded accumulator wins ties, as in accel.rs:236-247).
MM3E_DEV void run_begin(const Field& acc, bool seeded, float& bd, uint32_t& bm, bool& any) {
  bd = seeded ? acc.dist : __builtin_inff();
  bm = seeded ? acc.mat : 0u;
  any = seeded;
}
template <bool CSG>
MM3E_DEV void run_member(const DevScene& sc, const DObj& o, V3 p, float& bd, uint32_t& bm, bool& any) {
  if (any && (o.flags & (OBJ_BOUNDED | OBJ_SKIP)) && far_noop(o, p, rmax(bd, sc.slack))) return;
  Field f = contrib<CSG>(sc, o, p);
  if (!any || f.dist < bd) {
    bd = f.dist;
    bm = f.mat;
    any = true;
  }
}
MM3E_DEV void run_end(Field& acc, bool& seeded, float bd, uint32_t bm) {
  acc = Field{bd, bm};
  seeded = true;
}

// A wave-uniform record of the scene's HBM tables (the segment a fold step visits, the objects
// of a One segment, of a linear run and a run's unbounded members are the same for every lane)
// read with scalar loads: the fields land in SGPRs, which VALU instructions read directly, and
// only the fields a path uses are loaded.  `i` must be the same in every active lane.
template <class T> MM3E_DEV T load_uniform(const T* table, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(4))) const T ConstT;
  // the base may reach here in VGPRs (a scene pointer read from a batch job record): take it
  // through readfirstlane as well, so the load's address is scalar.  The builtin returns a
  // signed int, so each half is widened through uint32_t: a sign-extended low word would set
  // the high word of every address whose bit 31 is set.
  const uint64_t a = (uint64_t)(uintptr_t)table;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  const uint64_t base = (uint64_t)lo | ((uint64_t)hi << 32);
  return ((ConstT*)(uintptr_t)base)[(uint32_t)__builtin_amdgcn_readfirstlane(i)];
#else
  return table[i];  // the host pass only parses device code
#endif
}

// One segment of WorldPlan::eval (accel.rs:209-252) folded into `acc`.  Runs of plain unions
// walk the same bounding-sphere tree in the same near-first order with the same `lower > best`
// cut and earliest-index tie break as Bvh::query (accel.rs:79-119); instead of a 64-entry
// stack each lane keeps one bit per tree level saying which child it visited first, and
// climbs back through parent links, so traversal state is four registers.
template <bool CSG>
MM3E_DEV void seg_step(const DevScene& sc, const Tables& tb, V3 p, uint32_t si, Field& acc, bool& seeded) {
  {
    const DSeg seg = load_uniform(sc.seg, si);
    if (seg.kind == 0) {
      const DObj o = load_uniform(sc.obj, seg.index);
      seg_one<CSG>(sc, o, seg.combine, seg.k, p, acc, seeded);
      return;
    }
    if (seg.kind == 2) {
      float bd;
      uint32_t bm;
      bool any;
      run_begin(acc, seeded, bd, bm, any);
      for (uint32_t u = 0; u < seg.unb_count; u++) {
        const DObj o = load_uniform(sc.obj, load_uniform(sc.unb, seg.unb_base + u));
        run_member<CSG>(sc, o, p, bd, bm, any);
      }
      run_end(acc, seeded, bd, bm);
      return;
    }
    float bd;
    int64_t bi;
    uint32_t bm;
    if (seeded) {
      bd = acc.dist; bi = -1; bm = acc.mat;
    } else {
      bd = __builtin_inff(); bi = INT64_MAX; bm = 0u;
    }
    for (uint32_t u = 0; u < seg.unb_count; u++) {
      const uint32_t i = load_uniform(sc.unb, seg.unb_base + u);
      const DObj o = load_uniform(sc.obj, i);
      // a skip sphere proves d > bd, so neither the < nor the tie test can fire
      if ((o.flags & OBJ_SKIP) && far_noop(o, p, rmax(bd, sc.slack))) continue;
      float d = object_distance<CSG>(sc, o, p);
      if (d < bd || (d == bd && (int64_t)i < bi)) { bd = d; bi = i; bm = o.mat; }
    }
    if (seg.has_tree) {
      const DNode* nodes = tb.nodes;
      const int32_t root = (int32_t)seg.index;
      int32_t node = root;
      uint32_t depth = 0, nearbits = 0;
      for (;;) {
        // descend into `node`; the squared pre-test proves `lower > bd` without a sqrt
        const DNode& n = nodes[node];
        bool climb = true;
        const float tn = (bd + n.r) * 1.00002f + 1e-6f;
        V3 dn = p - ld3(n.c);
        const float dn2 = dot(dn, dn);
        float lower = 0.0f;
        const bool pruned = tn > 0.0f && dn2 > tn * tn;
        if (!pruned) lower = crsqrt(dn2) - n.r;
        if (!pruned && !(lower > bd)) {
          if (n.leaf >= 0) {
            const uint32_t i = (uint32_t)n.leaf;
            const DObj& o = tb.obj[i];
            float d;
            if (lower > sc.slack) d = lower;
            else d = object_distance<CSG>(sc, o, p);
            if (d < bd || (d == bd && (int64_t)i < bi)) { bd = d; bi = i; bm = o.mat; }
          } else {
            const DNode& L = nodes[n.left];
            const DNode& R = nodes[n.right];
            float ldist = length_sq(p - ld3(L.c));
            float rdist = length_sq(p - ld3(R.c));
            bool near_left = ldist <= rdist;
            nearbits = near_left ? (nearbits | (1u << depth)) : (nearbits & ~(1u << depth));
            node = near_left ? n.left : n.right;
            depth++;
            climb = false;
          }
        }
        if (!climb) continue;
        // climb until an unvisited far child exists
        bool done = true;
        while (node != root) {
          const int32_t parent = nodes[node].parent;
          depth--;
          const DNode& P = nodes[parent];
          const bool near_left = (nearbits >> depth) & 1u;
          const int32_t nearc = near_left ? P.left : P.right;
          if (node == nearc) {
            node = near_left ? P.right : P.left;
            depth++;
            done = false;
            break;
          }
          node = parent;
        }
        if (done) break;
      }
    }
    acc = Field{bd, bm};
    seeded = true;
  }
}

// Segment-grid cell of p (index nx*ny*nz for points outside the grid, NaN included).
MM3E_DEV uint32_t grid_cell(const DevScene& sc, V3 p) {
  const float fx = (p.x - sc.grid_lo[0]) * sc.grid_inv;
  const float fy = (p.y - sc.grid_lo[1]) * sc.grid_inv;
  const float fz = (p.z - sc.grid_lo[2]) * sc.grid_inv;
  const uint32_t nx = sc.grid_dims[0], ny = sc.grid_dims[1], nz = sc.grid_dims[2];
  const bool in = fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)nx && fy < (float)ny && fz < (float)nz;
  if (!in) return nx * ny * nz;
  // fx < (float)nx with fx >= 0 truncates to at most nx - 1
  const uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy, iz = (uint32_t)fz;
  return (iz * ny + iy) * nx + ix;
}

// OR of v over the active lanes of the wave, uniform: each pass takes the bits of the first
// lane that still adds any, so the loop runs once per distinct contribution (lanes in the same
// cell and level share theirs) and ends after at most 32 passes.
MM3E_DEV uint32_t wave_or(uint32_t v) {
  uint32_t acc = 0u;
  uint64_t rem = __builtin_amdgcn_ballot_w64(true);
  while (rem != 0) {
    acc |= (uint32_t)__builtin_amdgcn_readlane(v, (uint32_t)__builtin_ctzll(rem));
    rem = __builtin_amdgcn_ballot_w64((v & ~acc) != 0u);
  }
  return acc;
}

// The wave's segment masks once the first segment has seeded `acc` (see world_eval): `on` is
// false (fold every segment) when some active lane's threshold lies beyond the grid's widest
// level.  Otherwise each lane takes its own cell at the smallest level its own threshold is
// below, and word w of the wave's mask is the OR of those lanes' words (`grid_word`): the
// segments any lane may need, whatever the number of distinct cells the wave's points touch.
struct GridSel {
  bool on;
  uint32_t W;
  uint32_t base;  // per lane: first word of the lane's (cell, level) mask
  const uint32_t* m;
};
MM3E_DEV GridSel grid_select(const DevScene& sc, V3 p, const Field& acc) {
  GridSel g;
  g.on = false;
  const float thr = rmax(acc.dist + sc.grid_kmax, sc.slack) * 1.0001f + sc.grid_eps;
  const uint32_t cell = grid_cell(sc, p);
  // A point outside the grid whose distance to the box of the bounded segments' spheres is above
  // its threshold (with margins for the f32 arithmetic) has every bounded member farther than the
  // threshold: it folds only the unbounded segments (the outside cell's level 0), whatever its
  // threshold — rays far beyond the scene, toward the horizon or the sun.
  bool far = false;
  if (cell == sc.grid_dims[0] * sc.grid_dims[1] * sc.grid_dims[2]) {
    const float dx = rmax(rmax(sc.grid_in_lo[0] - p.x, p.x - sc.grid_in_hi[0]), 0.0f);
    const float dy = rmax(rmax(sc.grid_in_lo[1] - p.y, p.y - sc.grid_in_hi[1]), 0.0f);
    const float dz = rmax(rmax(sc.grid_in_lo[2] - p.z, p.z - sc.grid_in_hi[2]), 0.0f);
    const float tf = thr * 1.0002f + sc.grid_eps;
    far = dx * dx + dy * dy + dz * dz > tf * tf;
  }
  if (__builtin_amdgcn_ballot_w64(!far && !(thr < sc.grid_m[kGridLevels - 1])) != 0) return g;
  uint32_t level = kGridLevels - 1;
  for (int l = (int)kGridLevels - 2; l >= 0; l--)
    if (thr < sc.grid_m[l]) level = (uint32_t)l;
  if (far) level = 0;
  g.on = true;
  g.W = sc.grid_words;
  g.base = (cell * kGridLevels + level) * g.W;
  g.m = sc.grid;
  return g;
}
MM3E_DEV uint32_t grid_word(const GridSel& g, uint32_t w) { return wave_or(g.m[g.base + w]); }

// accel.rs:209-252 WorldPlan::eval — the reference `world()` field: the ordered fold of
// seg_step over every segment.  With a segment grid, after the first segment seeds the
// accumulator each lane checks that max(acc + kmax, slack) (plus rounding margins) is below a
// grid level's reach; if every active lane passes, each lane picks its cell's mask at its own
// level, the wave ORs those masks word by word and folds only the segments whose bits are set,
// in index order.  A clear bit is a segment whose every member
// is a provable no-op for every point of that cell at that threshold, so the fold is the same.
// GRID: whether the scene may carry a segment grid (kernels for scenes without one keep the
// plain fold).
template <bool GRID, bool CSG>
MM3E_MAP_ATTR Field world_eval(const DevScene& sc, const Tables& tb, V3 p) {
  Field acc{__builtin_inff(), 0u};
  bool seeded = false;
  const uint32_t nseg = sc.n_seg;
  if constexpr (!GRID) {
    for (uint32_t si = 0; si < nseg; si++) seg_step<CSG>(sc, tb, p, si, acc, seeded);
    return acc;
  }
  if (nseg == 0) return acc;
  // One call site for seg_step (it carries the tree walk and is inlined): the loop either
  // steps si through every segment or, once the wave switches to the grid, through the set
  // bits of the OR of its cells' masks (word w, remaining bits `bits`).
  GridSel g;
  g.on = false;
  uint32_t w = 0, bits = 0;
  uint32_t si = 0;
  for (;;) {
    seg_step<CSG>(sc, tb, p, si, acc, seeded);
    if (si == 0 && sc.grid != nullptr) {
      g = grid_select(sc, p, acc);
      if (g.on) bits = grid_word(g, 0) & ~1u;
    }
    if (g.on) {
      while (bits == 0) {
        if (++w >= g.W) return acc;
        bits = grid_word(g, w);
      }
      si = w * 32u + (uint32_t)__builtin_ctz(bits);
      bits &= bits - 1u;
    } else if (++si >= nseg) {
      return acc;
    }
  }
}

// ---------------------------------------------------------------------------------------
// dual.rs — exact gradients for dual-safe scenes (lib.rs:820-835)
// ---------------------------------------------------------------------------------------
struct Dual {
  float v;
  V3 g;
};
MM3E_DEV Dual dcon(float v) { return Dual{v, V3{0.0f, 0.0f, 0.0f}}; }
MM3E_DEV Dual dscale(Dual a, float s) { return Dual{a.v * s, scale(a.g, s)}; }
MM3E_DEV Dual dadds(Dual a, float s) { return Dual{a.v + s, a.g}; }
MM3E_DEV Dual dneg(Dual a) { return Dual{-a.v, -a.g}; }
MM3E_DEV Dual dabs(Dual a) { return a.v >= 0.0f ? a : dneg(a); }
MM3E_DEV Dual dsqrt(Dual a) {
  float v = crsqrt(rmax(a.v, 0.0f));
  float k = v > 1e-9f ? 0.5f / v : 0.0f;
  return Dual{v, scale(a.g, k)};
}
MM3E_DEV Dual dmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
MM3E_DEV Dual dmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
MM3E_DEV Dual dmax0(Dual a) { return dmax(a, dcon(0.0f)); }
MM3E_DEV Dual dmaxc(Dual a, float c) { return a.v >= c ? a : dcon(c); }
MM3E_DEV Dual dclamp(Dual a, float lo, float hi) { return dmin(dmax(a, dcon(lo)), dcon(hi)); }
MM3E_DEV Dual dadd(Dual a, Dual b) { return Dual{a.v + b.v, a.g + b.g}; }
MM3E_DEV Dual dsub(Dual a, Dual b) { return Dual{a.v - b.v, a.g - b.g}; }
MM3E_DEV Dual dmul(Dual a, Dual b) { return Dual{a.v * b.v, scale(a.g, b.v) + scale(b.g, a.v)}; }
MM3E_DEV Dual ddiv(Dual a, Dual b) {
  float inv2 = 1.0f / (b.v * b.v);
  return Dual{a.v / b.v, scale(scale(a.g, b.v) - scale(b.g, a.v), inv2)};
}

MM3E_DEV Dual dprim(uint32_t kind, const float* p, Dual x, Dual y, Dual z) {
  switch (kind) {
    case 0: return dadds(dsqrt(dadd(dadd(dmul(x, x), dmul(y, y)), dmul(z, z))), -p[0]);
    case 1:
    case 2: {
      V3 half = ld3(p);
      float rr = 0.0f;
      if (kind == 2) {
        rr = p[3];
        half = half - splat(rr);
      }
      Dual qx = dadds(dabs(x), -half.x), qy = dadds(dabs(y), -half.y), qz = dadds(dabs(z), -half.z);
      Dual out = dsqrt(dadd(dadd(dmul(dmax0(qx), dmax0(qx)), dmul(dmax0(qy), dmax0(qy))), dmul(dmax0(qz), dmax0(qz))));
      Dual in = dmin(dmax(qx, dmax(qy, qz)), dcon(0.0f));
      Dual d = dadd(out, in);
      return kind == 2 ? dadds(d, -rr) : d;
    }
    case 3: {
      Dual q = dadds(dsqrt(dadd(dmul(x, x), dmul(z, z))), -p[0]);
      return dadds(dsqrt(dadd(dmul(q, q), dmul(y, y))), -p[1]);
    }
    case 4: {
      Dual dx = dadds(dsqrt(dadd(dmul(x, x), dmul(z, z))), -p[1]);
      Dual dy = dadds(dabs(y), -p[0]);
      Dual in = dmin(dmax(dx, dy), dcon(0.0f));
      Dual ox = dmax0(dx), oy = dmax0(dy);
      return dadd(in, dsqrt(dadd(dmul(ox, ox), dmul(oy, oy))));
    }
    case 5: {
      V3 a = ld3(p), b = ld3(p + 3);
      Dual pax = dadds(x, -a.x), pay = dadds(y, -a.y), paz = dadds(z, -a.z);
      V3 ba = b - a;
      float dbb = rmax(dot(ba, ba), 1e-12f);
      Dual dpb = dadd(dadd(dscale(pax, ba.x), dscale(pay, ba.y)), dscale(paz, ba.z));
      Dual hh = dclamp(dscale(dpb, 1.0f / dbb), 0.0f, 1.0f);
      Dual rx = dsub(pax, dscale(hh, ba.x)), ry = dsub(pay, dscale(hh, ba.y)), rz = dsub(paz, dscale(hh, ba.z));
      return dadds(dsqrt(dadd(dadd(dmul(rx, rx), dmul(ry, ry)), dmul(rz, rz))), -p[6]);
    }
    case 6: {
      const float r1 = p[0], r2 = p[1], h = p[2];
      Dual qx = dsqrt(dadd(dmul(x, x), dmul(z, z)));
      Dual qy = y;
      float b = (r1 - r2) / rmax(h, 1e-6f);
      float a = crsqrt(rmax(1.0f - b * b, 0.0f));
      Dual k = dadd(dscale(qx, -b), dscale(qy, a));
      if (k.v < 0.0f) return dadds(dsqrt(dadd(dmul(qx, qx), dmul(qy, qy))), -r1);
      if (k.v > a * h) {
        Dual qyh = dadds(qy, -h);
        return dadds(dsqrt(dadd(dmul(qx, qx), dmul(qyh, qyh))), -r2);
      }
      return dadds(dadd(dscale(qx, a), dscale(qy, b)), -r1);
    }
    case 7: {
      V3 r = ld3(p);
      Dual k0x = dscale(x, 1.0f / r.x), k0y = dscale(y, 1.0f / r.y), k0z = dscale(z, 1.0f / r.z);
      Dual k0 = dsqrt(dadd(dadd(dmul(k0x, k0x), dmul(k0y, k0y)), dmul(k0z, k0z)));
      if (k0.v < 1e-6f) return dcon(-rmin(rmin(r.x, r.y), r.z));
      Dual k1x = dscale(x, 1.0f / (r.x * r.x)), k1y = dscale(y, 1.0f / (r.y * r.y)), k1z = dscale(z, 1.0f / (r.z * r.z));
      Dual k1 = dmaxc(dsqrt(dadd(dadd(dmul(k1x, k1x), dmul(k1y, k1y)), dmul(k1z, k1z))), 1e-9f);
      return ddiv(dmul(k0, dadds(k0, -1.0f)), k1);
    }
    case 8: return dscale(dadds(dadd(dadd(dabs(x), dabs(y)), dabs(z)), -p[0]), 0.5773503f);
    case 9: {
      const float kx = -0.8660254f, ky = 0.5f, kz = 0.57735f, r = p[0], h = p[1];
      Dual ax0 = dabs(x), az0 = dabs(z);
      Dual ay = dadds(dabs(y), -h);
      Dual t = dscale(dmin(dadd(dscale(ax0, kx), dscale(az0, ky)), dcon(0.0f)), 2.0f);
      Dual ax = dsub(ax0, dscale(t, kx)), az = dsub(az0, dscale(t, ky));
      Dual cl = dclamp(ax, -kz * r, kz * r);
      Dual diffx = dsub(ax, cl), diffz = dadds(az, -r);
      float sign = diffz.v >= 0.0f ? 1.0f : -1.0f;
      Dual dx = dscale(dsqrt(dadd(dmul(diffx, diffx), dmul(diffz, diffz))), sign);
      Dual in = dmin(dmax(dx, ay), dcon(0.0f));
      Dual ox = dmax0(dx), oy = dmax0(ay);
      return dadd(in, dsqrt(dadd(dmul(ox, ox), dmul(oy, oy))));
    }
    case 10: return dadds(dadd(dadd(dscale(x, p[0]), dscale(y, p[1])), dscale(z, p[2])), p[3]);
    default: return dcon(__builtin_inff());
  }
}

// lib.rs:435-463 Object::field_dual.
MM3E_DEV Dual object_dual(const DObj& o, V3 w) {
  Dual tx = Dual{w.x + -o.pos[0], V3{1.0f, 0.0f, 0.0f}};
  Dual ty = Dual{w.y + -o.pos[1], V3{0.0f, 1.0f, 0.0f}};
  Dual tz = Dual{w.z + -o.pos[2], V3{0.0f, 0.0f, 1.0f}};
  // rt.cols[j].i == rot.cols[i].j
  Dual lx = dscale(dadd(dadd(dscale(tx, o.c0[0]), dscale(ty, o.c0[1])), dscale(tz, o.c0[2])), o.inv);
  Dual ly = dscale(dadd(dadd(dscale(tx, o.c1[0]), dscale(ty, o.c1[1])), dscale(tz, o.c1[2])), o.inv);
  Dual lz = dscale(dadd(dadd(dscale(tx, o.c2[0]), dscale(ty, o.c2[1])), dscale(tz, o.c2[2])), o.inv);
  const uint32_t f = o.flags;
  if (f & OBJ_MIRROR_X) lx = dabs(lx);
  if (f & OBJ_MIRROR_Y) ly = dabs(ly);
  if (f & OBJ_MIRROR_Z) lz = dabs(lz);
  if (f & OBJ_ELONGATE) {
    lx = dsub(lx, dclamp(lx, -o.elo[0], o.elo[0]));
    ly = dsub(ly, dclamp(ly, -o.elo[1], o.elo[1]));
    lz = dsub(lz, dclamp(lz, -o.elo[2], o.elo[2]));
  }
  if (f & OBJ_REPEAT) {  // dual.rs:253-265
    if (o.rep[0] > 0.0f) lx = dadds(lx, -__builtin_roundf(lx.v / o.rep[0]) * o.rep[0]);
    if (o.rep[1] > 0.0f) ly = dadds(ly, -__builtin_roundf(ly.v / o.rep[1]) * o.rep[1]);
    if (o.rep[2] > 0.0f) lz = dadds(lz, -__builtin_roundf(lz.v / o.rep[2]) * o.rep[2]);
  }
  Dual d = dscale(dprim(f & OBJ_KIND_MASK, o.p, lx, ly, lz), o.scale);
  if (f & OBJ_ROUND) d = dadds(d, -o.round);
  if (f & OBJ_ONION) d = dadds(dabs(d), -o.onion);
  return d;
}

MM3E_DEV V3 normal_dual(const DevScene& sc, const Tables& tb, V3 p) {
  const uint32_t n = sc.n_obj;
  if (n == 0) return V3{0.0f, 0.0f, 0.0f};
  Dual acc = object_dual(tb.obj[0], p);
  for (uint32_t i = 1; i < n; i++) {
    const DObj& o = tb.obj[i];
    Dual f = object_dual(o, p);
    const uint32_t comb = (o.flags >> OBJ_COMBINE_SHIFT) & 3u;
    if (comb == 0) acc = dmin(acc, f);
    else if (comb == 1) {
      const float k = o.k;
      if (k <= 0.0f) acc = dmin(acc, f);
      else {
        Dual h = dclamp(dadds(dscale(dsub(f, acc), 0.5f / k), 0.5f), 0.0f, 1.0f);
        Dual omh = dsub(dcon(1.0f), h);
        acc = dsub(dadd(dmul(f, omh), dmul(acc, h)), dscale(dmul(h, omh), k));
      }
    } else acc = dmax(acc, dneg(f));
  }
  return normalize(acc.g);
}

// ---------------------------------------------------------------------------------------
// march.rs
// ---------------------------------------------------------------------------------------
struct Hit {
  bool hit;
  float t;
  V3 pos;
  uint32_t mat, steps;
};

// The searches below (march, tetrahedron normal, soft shadow, AO) are written as state
// records with init/step functions that take one field value per step, so a kernel can run
// all of them from a single field-evaluation site (one inlined copy of the world fold) and
// the lanes of a wave share every evaluation whatever search each is in.  The arithmetic of
// each step is exactly the reference loop body it restates.

// march.rs:212-223 tetrahedron gradient: corner k of (+--, --+, -+-, +++) at p + k*h.
MM3E_DEV V3 tetra_corner(uint32_t k) {
  return k == 0 ? V3{1.0f, -1.0f, -1.0f} : k == 1 ? V3{-1.0f, -1.0f, 1.0f} : k == 2 ? V3{-1.0f, 1.0f, -1.0f}
                                                                                    : V3{1.0f, 1.0f, 1.0f};
}
struct TetraState {
  V3 p, g;
  float h;
  uint32_t k;
};
MM3E_DEV V3 tetra_init(TetraState& s, V3 p, float h) {
  s.p = p;
  s.h = h;
  s.k = 0;
  s.g = V3{0.0f, 0.0f, 0.0f};
  return p + scale(tetra_corner(0), h);
}
// Folds corner k's value in left-to-right order; true when the gradient is complete.
MM3E_DEV bool tetra_step(TetraState& s, float d, V3& q) {
  const V3 term = scale(tetra_corner(s.k), d);
  s.g = s.k == 0 ? term : s.g + term;
  if (++s.k == 4) return true;
  q = s.p + scale(tetra_corner(s.k), s.h);
  return false;
}

// march.rs:120-209 — enhanced sphere tracing with the unconditional overlap guard.
enum : uint32_t { MARCH_EVAL = 0, MARCH_EVAL_MID = 1, MARCH_DONE = 2 };
struct MarchState {
  V3 ro, rd;
  float t, prev_radius, step_len, t_prev, d_prev, omega, subitize;
  uint32_t i, boost;  // boost: 0 omega, 1 leap, 2 secant
  bool secant_on;
  // the main evaluation while the guard's midpoint probe is in flight
  Field fm;
  float radius_m, eps_m, gap_lo, gap_hi;
};
MM3E_DEV uint32_t march_miss(const DevScene& sc, const MarchState& m, Hit& h) {
  h.hit = false;
  h.t = m.t;
  h.pos = m.ro + scale(m.rd, m.t);
  h.mat = 0;
  h.steps = sc.max_steps;
  return MARCH_DONE;
}
MM3E_DEV uint32_t march_next(const DevScene& sc, MarchState& m, Hit& h, V3& q) {  // loop increment + test
  if (++m.i >= sc.max_steps) return march_miss(sc, m, h);
  q = m.ro + scale(m.rd, m.t);
  return MARCH_EVAL;
}
MM3E_DEV uint32_t march_init(const DevScene& sc, MarchState& m, V3 ro, V3 rd, Hit& h, V3& q) {
  const bool ss1 = (sc.flags & SC_SS1) != 0;
  m.ro = ro;
  m.rd = rd;
  m.omega = ss1 ? 1.4f : sc.step_scale;
  m.subitize = ss1 ? sc.subitize : 0.0f;
  m.secant_on = ((sc.flags & SC_SECANT) != 0) && ss1;
  m.boost = 0;
  m.t = 0.0f;
  m.prev_radius = 0.0f;
  m.step_len = 0.0f;
  m.t_prev = 0.0f;
  m.d_prev = __builtin_inff();
  m.i = 0;
  if (sc.max_steps == 0) return march_miss(sc, m, h);
  q = m.ro + scale(m.rd, m.t);
  return MARCH_EVAL;
}
MM3E_DEV uint32_t march_retreat(const DevScene& sc, MarchState& m, Hit& h, V3& q) {
  m.t = m.gap_lo;
  if (m.boost == 0) m.omega = rmin(m.omega, 1.0f);
  else if (m.boost == 1) m.subitize = 0.0f;
  else m.secant_on = false;
  m.boost = 0;
  m.prev_radius = 0.0f;
  m.step_len = 0.0f;
  m.d_prev = __builtin_inff();
  return march_next(sc, m, h, q);
}
MM3E_DEV uint32_t march_finish(const DevScene& sc, MarchState& m, Field f, float radius, float eps, Hit& h, V3& q) {
  if (f.dist < eps) {
    h.hit = true;
    h.t = m.t;
    h.pos = m.ro + scale(m.rd, m.t);
    h.mat = f.mat;
    h.steps = m.i;
    return MARCH_DONE;
  }
  float step_len = f.dist * m.omega;
  m.boost = 0;
  if (m.subitize > 0.0f && f.dist > eps * 6.0f) {
    step_len = step_len * (1.0f + m.subitize);
    m.boost = 1;
  }
  if (m.secant_on && f.dist < 0.08f && __builtin_isfinite(m.d_prev)) {
    float dt = m.t - m.t_prev;
    float dd = f.dist - m.d_prev;
    if (dt > 1e-6f && dd < -1e-6f) {
      step_len = rclamp(-f.dist * dt / dd, f.dist, f.dist * 4.0f);
      m.boost = 2;
    }
  }
  m.step_len = step_len;
  m.prev_radius = radius;
  m.d_prev = f.dist;
  m.t_prev = m.t;
  m.t = m.t + step_len;
  if (m.t > sc.max_dist) return march_miss(sc, m, h);
  return march_next(sc, m, h, q);
}
// One step given the field at the point march_init / march_step last asked for (`st` is the
// status that request came with).  Returns the next status; q is the next point to evaluate.
MM3E_DEV uint32_t march_step(const DevScene& sc, MarchState& m, uint32_t st, Field f, Hit& h, V3& q) {
  if (st == MARCH_EVAL_MID) {
    const bool covered = f.dist >= 0.5f * (m.gap_hi - m.gap_lo);
    if (!covered) return march_retreat(sc, m, h, q);
    return march_finish(sc, m, m.fm, m.radius_m, m.eps_m, h, q);
  }
  const float radius = rabs(f.dist);
  const float eps = sc.eps * (1.0f + m.t * 0.5f) + sc.lod * m.t;
  if ((m.step_len > 0.0f && f.dist < 0.0f) || m.step_len > radius + m.prev_radius) {
    m.gap_lo = m.t_prev + m.prev_radius;
    m.gap_hi = m.t - radius;
    if (f.dist > 0.0f) {
      const float mid = 0.5f * (m.gap_lo + m.gap_hi);
      m.fm = f;
      m.radius_m = radius;
      m.eps_m = eps;
      q = m.ro + scale(m.rd, mid);
      return MARCH_EVAL_MID;
    }
    return march_retreat(sc, m, h, q);
  }
  return march_finish(sc, m, f, radius, eps, h, q);
}

template <class Map> MM3E_DEV Hit march(const DevScene& sc, const Map& field, V3 ro, V3 rd) {
  MarchState m;
  Hit h;
  V3 q;
  uint32_t st = march_init(sc, m, ro, rd, h, q);
  while (st != MARCH_DONE) st = march_step(sc, m, st, field(q), h, q);
  return h;
}

template <class Map> MM3E_DEV V3 tetra_normal(const Map& field, float h, V3 p) {  // march.rs:212-223
  TetraState s;
  V3 q = tetra_init(s, p, h);
  while (!tetra_step(s, field(q).dist, q)) {
  }
  return normalize(s.g);
}

// Hit normal: dual-number gradient on dual-safe scenes, tetrahedron otherwise (lib.rs:1273-1277).
template <class Map> MM3E_DEV V3 hit_normal(const DevScene& sc, const Map& field, V3 p) {
  if (sc.flags & SC_DUAL_SAFE) return normal_dual(sc, field.tb, p);
  return tetra_normal(field, sc.normal_h, p);
}

// march.rs:227-259 soft shadow.
struct ShadowState {
  V3 o, d;
  float res, t, max_t, k;
  uint32_t s;
};
// false: finished without evaluating (`result` set); true: evaluate q.
MM3E_DEV bool shadow_init(const DevScene& sc, ShadowState& s, V3 o, V3 d, float max_t, float k, V3& q, float& result) {
  s.o = o;
  s.d = d;
  s.max_t = max_t;
  s.k = k;
  s.res = 1.0f;
  float hb = o.x * 127.1f + o.y * 311.7f + o.z * 74.7f;
  float sv = cr_sin(hb) * 43758.547f;
  float jitter = rabs(sv - __builtin_truncf(sv));  // f32::fract
  s.t = 0.02f + jitter * 0.16f;
  s.s = 0;
  if (sc.shadow_steps == 0) {
    result = rclamp(s.res, 0.0f, 1.0f);
    return false;
  }
  q = s.o + scale(s.d, s.t);
  return true;
}
// true: finished (`result` set); false: evaluate q next.
MM3E_DEV bool shadow_step(const DevScene& sc, ShadowState& s, float h, V3& q, float& result) {
  if (h < 0.0008f) {
    result = 0.0f;
    return true;
  }
  s.res = rmin(s.res, s.k * h / s.t);
  s.t = s.t + rclamp(h, 0.06f, 0.5f + 0.7f * s.t);
  if (s.t > s.max_t || ++s.s >= sc.shadow_steps) {
    result = rclamp(s.res, 0.0f, 1.0f);
    return true;
  }
  q = s.o + scale(s.d, s.t);
  return false;
}
template <class Map> MM3E_DEV float soft_shadow(const DevScene& sc, const Map& field, V3 o, V3 d, float max_t, float k) {
  ShadowState s;
  V3 q;
  float r;
  if (!shadow_init(sc, s, o, d, max_t, k, q, r)) return r;
  while (!shadow_step(sc, s, field(q).dist, q, r)) {
  }
  return r;
}

// march.rs:264-280 ambient occlusion.
struct AoState {
  V3 p, n;
  float occ, sca;
  uint32_t i;
};
MM3E_DEV float ao_radius(const DevScene& sc, uint32_t i) {
  const uint32_t ns = sc.ao_samples;
  const float base = rmax(sc.normal_h, 0.01f);
  const float span = (float)((ns > 2u ? ns : 2u) - 1u);
  return base + 0.12f * (float)i / span;
}
MM3E_DEV bool ao_init(const DevScene& sc, AoState& a, V3 p, V3 n, V3& q, float& result) {
  a.p = p;
  a.n = n;
  a.occ = 0.0f;
  a.sca = 1.0f;
  a.i = 0;
  if (sc.ao_samples == 0) {
    result = 1.0f;
    return false;
  }
  q = p + scale(n, ao_radius(sc, 0));
  return true;
}
MM3E_DEV bool ao_step(const DevScene& sc, AoState& a, float d, V3& q, float& result) {
  const float hr = ao_radius(sc, a.i);
  a.occ = a.occ + (hr - d) * a.sca;
  a.sca = a.sca * 0.92f;
  if (++a.i >= sc.ao_samples) {
    result = rclamp(1.0f - 2.6f * a.occ, 0.0f, 1.0f);
    return true;
  }
  q = a.p + scale(a.n, ao_radius(sc, a.i));
  return false;
}
template <class Map> MM3E_DEV float ambient_occlusion(const DevScene& sc, const Map& field, V3 p, V3 n) {
  AoState a;
  V3 q;
  float r;
  if (!ao_init(sc, a, p, n, q, r)) return r;
  while (!ao_step(sc, a, field(q).dist, q, r)) {
  }
  return r;
}

// ---------------------------------------------------------------------------------------
// shade.rs
// ---------------------------------------------------------------------------------------
MM3E_DEV float lambert(V3 n, V3 l) { return rmax(dot(n, l), 0.0f); }
MM3E_DEV float fresnel_schlick(float c, float f0) {
  float m = rclamp(1.0f - c, 0.0f, 1.0f);
  return f0 + (1.0f - f0) * powi5(m);
}
MM3E_DEV V3 brdf(V3 n, V3 l, V3 view, V3 albedo, float metallic, float roughness, float reflectance) {  // shade.rs:51-72
  float nol = rmax(dot(n, l), 0.0f);
  if (nol <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
  float nov = rmax(dot(n, view), 1e-4f);
  V3 h = normalize(l + view);
  float noh = rmax(dot(n, h), 0.0f);
  float voh = rmax(dot(view, h), 0.0f);
  float a = rmax(roughness * roughness, 1e-3f);
  float a2 = a * a;
  V3 f0 = mix(splat(0.08f * rclamp(reflectance, 0.0f, 1.0f)), albedo, metallic);
  float m5 = powi5(rclamp(1.0f - voh, 0.0f, 1.0f));
  V3 f = f0 + scale(splat(1.0f) - f0, m5);
  float d = noh * noh * (a2 - 1.0f) + 1.0f;
  float dg = a2 / rmax(3.14159265358979323846f * d * d, 1e-12f);
  float gv = nol * crsqrt(nov * nov * (1.0f - a2) + a2);
  float gl = nov * crsqrt(nol * nol * (1.0f - a2) + a2);
  float vis = 0.5f / rmax(gv + gl, 1e-5f);
  V3 spec = scale(f, dg * vis);
  V3 kd = scale(splat(1.0f) - f, 1.0f - metallic);
  V3 diff = scale(cmul(kd, albedo), 1.0f / 3.14159265358979323846f);
  return scale(diff + spec, nol);
}
MM3E_DEV V3 sky_base(V3 dir) {
  float t = rclamp(0.5f * (dir.y + 1.0f), 0.0f, 1.0f);
  return mix(V3{0.78f, 0.86f, 0.96f}, V3{0.20f, 0.38f, 0.72f}, t);
}
MM3E_DEV V3 sky(V3 dir, V3 sun) {  // shade.rs:76-84
  V3 base = sky_base(dir);
  float s = rmax(dot(dir, sun), 0.0f);
  float glow = powf(s, 8.0f) * 0.3f + powf(s, 220.0f) * 6.0f;
  return base + splat(glow);
}
MM3E_DEV V3 sky_diffuse(V3 dir, V3 sun) {  // shade.rs:88-95
  V3 base = sky_base(dir);
  float soft = powf(rmax(dot(dir, sun), 0.0f), 8.0f) * 0.3f;
  return base + splat(soft);
}
MM3E_DEV float aces1(float v) {  // shade.rs:98-104
  v = rmax(v, 0.0f);
  return rclamp((v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f), 0.0f, 1.0f);
}
MM3E_DEV float gamma1(float v) { return powf(rmax(v, 0.0f), 1.0f / 2.2f); }  // shade.rs:107-110
MM3E_DEV V3 apply_fog(V3 c, V3 fog, float t, float density) {              // shade.rs:113-116
  float f = 1.0f - expf(-t * density);
  return mix(c, fog, rclamp(f, 0.0f, 1.0f));
}

// mm3e-kit/src/lib.rs:63-78 FNV-1a lattice hash; lib.rs:1365-1375 checker albedo.
MM3E_DEV float hash_cell(int32_t ix, int32_t iy) {
  uint64_t h = 0xcbf29ce484222325ull;
  uint32_t w[2] = {(uint32_t)ix, (uint32_t)iy};
#pragma unroll
  for (int k = 0; k < 2; k++) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      h ^= (uint64_t)((w[k] >> (8 * b)) & 0xffu);
      h *= 0x00000100000001b3ull;
    }
  }
  return (float)(uint32_t)(h >> 40) / 16777216.0f;
}
MM3E_DEV int32_t f2i_sat(float x) {  // `as i32`: saturating, NaN -> 0
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}
MM3E_DEV V3 surface_albedo(const DMat& m, V3 p) {
  V3 alb = ld3(m.albedo);
  if (!m.checker) return alb;
  int32_t ix = f2i_sat(__builtin_floorf(p.x));
  int32_t iz = f2i_sat(__builtin_floorf(p.z));
  int32_t s = (int32_t)((uint32_t)ix + (uint32_t)iz);
  bool parity = (s & 1) == 0;  // rem_euclid(2) == 0  <=> even
  V3 base = parity ? splat(0.92f) : splat(0.20f);
  float tint = hash_cell(ix, iz) * 0.06f - 0.03f;
  return max_scalar(cmul(base + splat(tint), alb), 0.0f);
}

// lib.rs:528-556 Light
MM3E_DEV void light_toward(const DLight& l, V3 p, V3& dir, float& dist) {
  if (l.directional) {
    dir = ld3(l.vec);
    dist = __builtin_inff();
  } else {
    V3 d = ld3(l.vec) - p;
    float len = length(d);
    dir = scale(d, 1.0f / rmax(len, 1e-6f));
    dist = len;
  }
}
MM3E_DEV float light_atten(const DLight& l, float dist) {
  if (l.directional) return 1.0f;
  return 1.0f / powi2(rmax(dist, 0.05f));
}
MM3E_DEV float light_shadow_k(const DLight& l, float dist) {
  if (l.radius <= 0.0f) return 24.0f;
  if (l.directional) return rclamp(1.0f / l.radius, 2.0f, 64.0f);
  return rclamp(dist / l.radius, 2.0f, 64.0f);
}

// gi.rs:138-187 ambient-cube sample (probe layout [probe][face][4]).
MM3E_DEV V3 gi_sample(const DevScene& sc, V3 pos, V3 n) {
  const uint32_t nx = sc.gi_dims[0], ny = sc.gi_dims[1], nz = sc.gi_dims[2];
  auto axis = [](float p, float lo, float ext, uint32_t nn) -> float {
    if (rabs(ext) > 1e-6f) return (p - lo) / ext * (float)((nn > 2u ? nn : 2u) - 1u);
    return 0.0f;
  };
  float gx = axis(pos.x, sc.gi_min[0], sc.gi_size[0], nx);
  float gy = axis(pos.y, sc.gi_min[1], sc.gi_size[1], ny);
  float gz = axis(pos.z, sc.gi_min[2], sc.gi_size[2], nz);
  auto clampi = [](float v, uint32_t hi) -> uint32_t {
    float f = rmax(__builtin_floorf(v), 0.0f);
    uint32_t lim = hi - 1u;
    return f >= (float)lim ? lim : (uint32_t)f;
  };
  uint32_t i0 = clampi(gx, nx), j0 = clampi(gy, ny), k0 = clampi(gz, nz);
  uint32_t i1 = i0 + 1 < nx ? i0 + 1 : nx - 1, j1 = j0 + 1 < ny ? j0 + 1 : ny - 1, k1 = k0 + 1 < nz ? k0 + 1 : nz - 1;
  float fx = rclamp(gx - (float)i0, 0.0f, 1.0f), fy = rclamp(gy - (float)j0, 0.0f, 1.0f), fz = rclamp(gz - (float)k0, 0.0f, 1.0f);
  // only the three faces the normal selects are needed (gi.rs:435-439 picks one per axis)
  const int fxs = n.x >= 0.0f ? 0 : 1, fys = n.y >= 0.0f ? 2 : 3, fzs = n.z >= 0.0f ? 4 : 5;
  const float4* g = reinterpret_cast<const float4*>(sc.gi);
  auto probe = [&](uint32_t i, uint32_t j, uint32_t k) -> uint32_t { return (k * ny + j) * nx + i; };
  const uint32_t p000 = probe(i0, j0, k0), p100 = probe(i1, j0, k0), p010 = probe(i0, j1, k0), p110 = probe(i1, j1, k0);
  const uint32_t p001 = probe(i0, j0, k1), p101 = probe(i1, j0, k1), p011 = probe(i0, j1, k1), p111 = probe(i1, j1, k1);
  auto face = [&](int f) -> V3 {
    auto c = [&](uint32_t pi) -> V3 {
      float4 v = g[pi * 6 + f];
      return V3{v.x, v.y, v.z};
    };
    V3 x00 = mix(c(p000), c(p100), fx);
    V3 x10 = mix(c(p010), c(p110), fx);
    V3 x01 = mix(c(p001), c(p101), fx);
    V3 x11 = mix(c(p011), c(p111), fx);
    return mix(mix(x00, x10, fy), mix(x01, x11, fy), fz);
  };
  V3 nn{n.x * n.x, n.y * n.y, n.z * n.z};
  return scale(face(fxs), nn.x) + scale(face(fys), nn.y) + scale(face(fzs), nn.z);
}

// lib.rs:1379-1398
MM3E_DEV V3 ibl_irradiance(const DevScene& sc, V3 n) {
  V3 up = rabs(n.y) > 0.9f ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 1.0f, 0.0f};
  V3 t1 = normalize(cross(n, up));
  V3 t2 = normalize(cross(n, t1));
  V3 sun = ld3(sc.sun);
  V3 n6 = scale(n, 0.6f);
  V3 dirs[5] = {n, normalize(n6 + scale(t1, 0.6f)), normalize(n6 - scale(t1, 0.6f)), normalize(n6 + scale(t2, 0.6f)),
                normalize(n6 - scale(t2, 0.6f))};
  V3 acc{0.0f, 0.0f, 0.0f};
  float w = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    float cw = rmax(dot(dirs[i], n), 0.0f);
    acc = acc + scale(sky_diffuse(dirs[i], sun), cw);
    w = w + cw;
  }
  return cmul(scale(acc, 1.0f / rmax(w, 1e-4f)), ld3(sc.sky_ambient));
}

// Table-driven world field over the compiled scene (tables in HBM or staged in LDS).
template <bool GRID, bool CSG>
struct TableMapT {
  const DevScene* sc;
  Tables tb;
  MM3E_DEV Field operator()(V3 p) const { return world_eval<GRID, CSG>(*sc, tb, p); }
};
// Every feature (segment grid, CSG objects): the general map.
using TableMap = TableMapT<true, true>;
// Scenes with neither: the fold without the grid gather and without the CSG evaluator.
using SimpleMap = TableMapT<false, false>;

MM3E_DEV Tables global_tables(const DevScene* sc) { return Tables{sc->obj, sc->seg, sc->nodes, sc->unb}; }

// Copy the object/segment/node/run-list tables (one contiguous range of `bytes`, which the
// launch sizes the dynamic LDS for) into LDS with the whole block, 16 bytes per lane per
// step, and return a view onto the copy.
MM3E_DEV Tables stage_tables(const DevScene* sc, uint4* lds, uint32_t bytes) {
  const uint32_t n16 = bytes / 16;
  const uint4* src = reinterpret_cast<const uint4*>(sc->obj);
  for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) lds[i] = src[i];
  __syncthreads();
  const char* b = reinterpret_cast<const char*>(lds);
  return Tables{reinterpret_cast<const DObj*>(b), reinterpret_cast<const DSeg*>(b + sc->st_seg),
                reinterpret_cast<const DNode*>(b + sc->st_node), reinterpret_cast<const uint32_t*>(b + sc->st_unb)};
}

}  // namespace mm3e
