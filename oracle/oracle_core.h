// oracle_core.h — CPU FP32 restatement of the reference MM3E CPU renderer.
//
// TEST INFRASTRUCTURE ONLY.  This is the parity oracle (and the CPU baseline) for the
// MI355X render path; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
// may load it.  The product library never links or calls it.
//
// Every function cites the reference function it restates (paths relative to the reference
// tree).  Operation order follows the Rust source so that the float instantiation rounds at
// the same points (host build: -O2 -ffp-contract=off, no fast-math).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "oracle_real.h"

namespace orc {

// ---------------------------------------------------------------------------------------
// mm3e-kit/src/vec.rs
// ---------------------------------------------------------------------------------------
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0.0f), y(0.0f), z(0.0f) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
};
template <class R> inline V3<R> splat(R v) { return V3<R>(v, v, v); }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }
template <class R> inline V3<R> scale(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }   // vec.rs:45
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }    // vec.rs:24
template <class R> inline V3<R> cross(V3<R> a, V3<R> o) {                                          // vec.rs:27
  return V3<R>(a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x);
}
template <class R> inline R length_sq(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return rsqrt_(length_sq(a)); }
template <class R> inline V3<R> normalize(V3<R> a) {  // vec.rs:37-44
  R len = length(a);
  if (len > R(1e-12f)) return scale(a, R(1.0f) / len);
  return V3<R>();
}
template <class R> inline V3<R> cmul(V3<R> a, V3<R> o) { return V3<R>(a.x * o.x, a.y * o.y, a.z * o.z); }
template <class R> inline V3<R> vabs(V3<R> a) { return V3<R>(rabs(a.x), rabs(a.y), rabs(a.z)); }
template <class R> inline V3<R> vmin(V3<R> a, V3<R> o) { return V3<R>(rmin(a.x, o.x), rmin(a.y, o.y), rmin(a.z, o.z)); }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> o) { return V3<R>(rmax(a.x, o.x), rmax(a.y, o.y), rmax(a.z, o.z)); }
template <class R> inline V3<R> max_scalar(V3<R> a, R m) { return V3<R>(rmax(a.x, m), rmax(a.y, m), rmax(a.z, m)); }
template <class R> inline R max_element(V3<R> a) { return rmax(rmax(a.x, a.y), a.z); }
template <class R> inline V3<R> clamp01(V3<R> a) {
  return V3<R>(rclamp(a.x, R(0.0f), R(1.0f)), rclamp(a.y, R(0.0f), R(1.0f)), rclamp(a.z, R(0.0f), R(1.0f)));
}
template <class R> inline V3<R> clamp_to(V3<R> a, V3<R> lo, V3<R> hi) { return vmin(vmax(a, lo), hi); }  // vec.rs:71
template <class R> inline V3<R> mix(V3<R> a, V3<R> o, R t) { return scale(a, R(1.0f) - t) + scale(o, t); }  // vec.rs:75
template <class R> inline V3<R> reflect(V3<R> v, V3<R> n) { return v - scale(n, R(2.0f) * dot(v, n)); }   // vec.rs:79
template <class R> inline bool is_zero(V3<R> a) { return val(a.x) == 0.0f && val(a.y) == 0.0f && val(a.z) == 0.0f; }

template <class R> struct M3 {  // vec.rs:112-114 (columns)
  V3<R> c[3];
};
template <class R> inline V3<R> mul_vec(const M3<R>& m, V3<R> v) {  // vec.rs:125-127
  return scale(m.c[0], v.x) + scale(m.c[1], v.y) + scale(m.c[2], v.z);
}
template <class R> inline M3<R> transpose(const M3<R>& m) {  // vec.rs:130-133
  M3<R> t;
  t.c[0] = V3<R>(m.c[0].x, m.c[1].x, m.c[2].x);
  t.c[1] = V3<R>(m.c[0].y, m.c[1].y, m.c[2].y);
  t.c[2] = V3<R>(m.c[0].z, m.c[1].z, m.c[2].z);
  return t;
}

template <class R> struct Xform {  // vec.rs:165-169
  M3<R> rot;
  V3<R> pos;
  R scale_;
};
template <class R> inline V3<R> to_local(const Xform<R>& x, V3<R> p) {  // vec.rs:216-221
  R inv = rabs(x.scale_) > R(1e-12f) ? R(1.0f) / x.scale_ : R(0.0f);
  return scale(mul_vec(transpose(x.rot), p - x.pos), inv);
}
template <class R> inline V3<R> to_world(const Xform<R>& x, V3<R> p) {  // vec.rs:190-192
  return x.pos + mul_vec(x.rot, scale(p, x.scale_));
}

// ---------------------------------------------------------------------------------------
// mm3e-kit/src/sdf.rs
// ---------------------------------------------------------------------------------------
template <class R> struct Field {  // sdf.rs:16-26
  R dist;
  uint32_t mat;
};
template <class R> inline Field<R> field_far() { return Field<R>{R(INFINITY), 0u}; }

template <class R> inline R sd_sphere(V3<R> p, R r) { return length(p) - r; }  // sdf.rs:33-35
template <class R> inline R sd_box(V3<R> p, V3<R> half) {                        // sdf.rs:38-41
  V3<R> q = vabs(p) - half;
  return length(max_scalar(q, R(0.0f))) + rmin(max_element(q), R(0.0f));
}
template <class R> inline R sd_round_box(V3<R> p, V3<R> half, R radius) {  // sdf.rs:44-46
  return sd_box(p, half - splat(radius)) - radius;
}
template <class R> inline R sd_plane(V3<R> p, V3<R> n, R h) { return dot(p, n) + h; }  // sdf.rs:49-51
template <class R> inline R sd_torus(V3<R> p, R major, R minor) {                     // sdf.rs:54-57
  R q = rsqrt_(p.x * p.x + p.z * p.z) - major;
  return rsqrt_(q * q + p.y * p.y) - minor;
}
template <class R> inline R sd_cylinder(V3<R> p, R h, R r) {  // sdf.rs:60-66
  R dx = rsqrt_(p.x * p.x + p.z * p.z) - r;
  R dy = rabs(p.y) - h;
  R ox = rmax(dx, R(0.0f));
  R oy = rmax(dy, R(0.0f));
  return rmin(rmax(dx, dy), R(0.0f)) + rsqrt_(ox * ox + oy * oy);
}
template <class R> inline R sd_capsule(V3<R> p, V3<R> a, V3<R> b, R r) {  // sdf.rs:69-74
  V3<R> pa = p - a;
  V3<R> ba = b - a;
  R h = rclamp(dot(pa, ba) / dot(ba, ba), R(0.0f), R(1.0f));
  return length(pa - scale(ba, h)) - r;
}
template <class R> inline R sd_round_cone(V3<R> p, R r1, R r2, R h) {  // sdf.rs:78-91
  R qx = rsqrt_(p.x * p.x + p.z * p.z);
  R qy = p.y;
  R b = (r1 - r2) / rmax(h, R(1e-6f));
  R a = rsqrt_(rmax(R(1.0f) - b * b, R(0.0f)));
  R k = qx * -b + qy * a;
  if (k < R(0.0f)) return rsqrt_(qx * qx + qy * qy) - r1;
  if (k > a * h) return rsqrt_(qx * qx + (qy - h) * (qy - h)) - r2;
  return qx * a + qy * b - r1;
}
template <class R> inline R sd_ellipsoid(V3<R> p, V3<R> r) {  // sdf.rs:94-101
  R k0 = length(V3<R>(p.x / r.x, p.y / r.y, p.z / r.z));
  if (k0 < R(1e-6f)) return -rmin(rmin(r.x, r.y), r.z);
  R k1 = length(V3<R>(p.x / (r.x * r.x), p.y / (r.y * r.y), p.z / (r.z * r.z)));
  return k0 * (k0 - R(1.0f)) / rmax(k1, R(1e-9f));
}
template <class R> inline R sd_octahedron(V3<R> p, R s) {  // sdf.rs:104-107
  V3<R> q = vabs(p);
  return (q.x + q.y + q.z - s) * R(0.5773503f);
}
template <class R> inline R signum(R x) {  // f32::signum: 1.0 for +0.0/positive, -1.0 for -0.0/negative
  float v = val(x);
  if (std::isnan(v)) return x;
  return std::signbit(v) ? R(-1.0f) : R(1.0f);
}
template <class R> inline R sd_hex_prism(V3<R> p, R r, R h) {  // sdf.rs:110-120
  const R kx(-0.8660254f), ky(0.5f), kz(0.57735f);
  V3<R> a = vabs(p);
  R t = R(2.0f) * rmin(kx * a.x + ky * a.z, R(0.0f));
  a.x = a.x - t * kx;
  a.z = a.z - t * ky;
  R clamped = rclamp(a.x, -kz * r, kz * r);
  R dx = rsqrt_((a.x - clamped) * (a.x - clamped) + (a.z - r) * (a.z - r)) * signum(a.z - r);
  R dy = a.y - h;
  return rmin(rmax(dx, dy), R(0.0f)) + rsqrt_(rpowi(rmax(dx, R(0.0f)), 2) + rpowi(rmax(dy, R(0.0f)), 2));
}

template <class R> inline Field<R> f_union(Field<R> a, Field<R> b) { return a.dist <= b.dist ? a : b; }      // sdf.rs:127
template <class R> inline Field<R> f_intersect(Field<R> a, Field<R> b) { return a.dist >= b.dist ? a : b; }  // sdf.rs:136
template <class R> inline Field<R> f_subtract(Field<R> a, Field<R> b) {                                      // sdf.rs:145
  return Field<R>{rmax(a.dist, -b.dist), a.mat};
}
template <class R> inline Field<R> f_smooth_union(Field<R> a, Field<R> b, R k) {  // sdf.rs:152-160
  if (k <= R(0.0f)) return f_union(a, b);
  R h = rclamp(R(0.5f) + R(0.5f) * (b.dist - a.dist) / k, R(0.0f), R(1.0f));
  R dist = b.dist * (R(1.0f) - h) + a.dist * h - k * h * (R(1.0f) - h);
  uint32_t mat = h > R(0.5f) ? a.mat : b.mat;
  return Field<R>{dist, mat};
}

template <class R> inline R op_round(R d, R r) { return d - r; }           // sdf.rs:192
template <class R> inline R op_onion(R d, R t) { return rabs(d) - t; }     // sdf.rs:197
template <class R> inline V3<R> op_elongate(V3<R> p, V3<R> h) {            // sdf.rs:202-204
  return p - clamp_to(p, scale(h, R(-1.0f)), h);
}
template <class R> inline R rep_axis(R v, R period) {  // sdf.rs:206-212
  if (period > R(0.0f)) return v - period * rround(v / period);
  return v;
}
template <class R> inline V3<R> op_repeat(V3<R> p, V3<R> period) {  // sdf.rs:215-217
  return V3<R>(rep_axis(p.x, period.x), rep_axis(p.y, period.y), rep_axis(p.z, period.z));
}
template <class R> inline V3<R> op_twist(V3<R> p, R k) {  // sdf.rs:220-224
  R a = k * p.y;
  R s = rsin(a), c = rcos(a);
  return V3<R>(c * p.x - s * p.z, p.y, s * p.x + c * p.z);
}
template <class R> inline V3<R> op_bend(V3<R> p, R k) {  // sdf.rs:227-231
  R a = k * p.x;
  R s = rsin(a), c = rcos(a);
  return V3<R>(c * p.x - s * p.y, s * p.x + c * p.y, p.z);
}
template <class R> inline V3<R> op_mirror(V3<R> p, bool x, bool y, bool z) {  // sdf.rs:234-236
  return V3<R>(x ? rabs(p.x) : p.x, y ? rabs(p.y) : p.y, z ? rabs(p.z) : p.z);
}

// ---------------------------------------------------------------------------------------
// mm3e-kit/src/dual.rs — forward-mode duals (exact gradients for dual-safe scenes)
// ---------------------------------------------------------------------------------------
template <class R> struct Dual {  // dual.rs:20-23
  R v;
  V3<R> g;
};
template <class R> inline Dual<R> dcon(R v) { return Dual<R>{v, V3<R>()}; }
template <class R> inline Dual<R> dscale(Dual<R> a, R s) { return Dual<R>{a.v * s, scale(a.g, s)}; }
template <class R> inline Dual<R> dadds(Dual<R> a, R s) { return Dual<R>{a.v + s, a.g}; }
template <class R> inline Dual<R> dneg(Dual<R> a) { return Dual<R>{-a.v, -a.g}; }
template <class R> inline Dual<R> dabs(Dual<R> a) { return a.v >= R(0.0f) ? a : dneg(a); }  // dual.rs:36-42
template <class R> inline Dual<R> dsqrt(Dual<R> a) {                                         // dual.rs:43-47
  R v = rsqrt_(rmax(a.v, R(0.0f)));
  R k = v > R(1e-9f) ? R(0.5f) / v : R(0.0f);
  return Dual<R>{v, scale(a.g, k)};
}
template <class R> inline Dual<R> dmin(Dual<R> a, Dual<R> o) { return a.v <= o.v ? a : o; }  // dual.rs:49
template <class R> inline Dual<R> dmax(Dual<R> a, Dual<R> o) { return a.v >= o.v ? a : o; }  // dual.rs:56
template <class R> inline Dual<R> dmax0(Dual<R> a) { return dmax(a, dcon(R(0.0f))); }
template <class R> inline Dual<R> dmaxc(Dual<R> a, R c) { return a.v >= c ? a : dcon(c); }  // dual.rs:67-73
template <class R> inline Dual<R> dclamp(Dual<R> a, R lo, R hi) { return dmin(dmax(a, dcon(lo)), dcon(hi)); }
template <class R> inline Dual<R> dadd(Dual<R> a, Dual<R> o) { return Dual<R>{a.v + o.v, a.g + o.g}; }
template <class R> inline Dual<R> dsub(Dual<R> a, Dual<R> o) { return Dual<R>{a.v - o.v, a.g - o.g}; }
template <class R> inline Dual<R> dmul(Dual<R> a, Dual<R> o) {  // dual.rs:98-103
  return Dual<R>{a.v * o.v, scale(a.g, o.v) + scale(o.g, a.v)};
}
template <class R> inline Dual<R> ddiv(Dual<R> a, Dual<R> o) {  // dual.rs:106-112
  R inv2 = R(1.0f) / (o.v * o.v);
  return Dual<R>{a.v / o.v, scale(scale(a.g, o.v) - scale(o.g, a.v), inv2)};
}

template <class R> inline Dual<R> dsd_sphere(Dual<R> x, Dual<R> y, Dual<R> z, R r) {  // dual.rs:129-131
  return dadds(dsqrt(dadd(dadd(dmul(x, x), dmul(y, y)), dmul(z, z))), -r);
}
template <class R> inline Dual<R> dsd_box(Dual<R> x, Dual<R> y, Dual<R> z, V3<R> half) {  // dual.rs:133-140
  Dual<R> qx = dadds(dabs(x), -half.x);
  Dual<R> qy = dadds(dabs(y), -half.y);
  Dual<R> qz = dadds(dabs(z), -half.z);
  Dual<R> outside = dsqrt(dadd(dadd(dmul(dmax0(qx), dmax0(qx)), dmul(dmax0(qy), dmax0(qy))), dmul(dmax0(qz), dmax0(qz))));
  Dual<R> inside = dmin(dmax(qx, dmax(qy, qz)), dcon(R(0.0f)));
  return dadd(outside, inside);
}
template <class R> inline Dual<R> dsd_round_box(Dual<R> x, Dual<R> y, Dual<R> z, V3<R> half, R radius) {
  return dadds(dsd_box(x, y, z, half - splat(radius)), -radius);  // dual.rs:142-144
}
template <class R> inline Dual<R> dsd_plane(Dual<R> x, Dual<R> y, Dual<R> z, V3<R> n, R h) {  // dual.rs:146-148
  return dadds(dadd(dadd(dscale(x, n.x), dscale(y, n.y)), dscale(z, n.z)), h);
}
template <class R> inline Dual<R> dsd_torus(Dual<R> x, Dual<R> y, Dual<R> z, R major, R minor) {  // dual.rs:150-153
  Dual<R> q = dadds(dsqrt(dadd(dmul(x, x), dmul(z, z))), -major);
  return dadds(dsqrt(dadd(dmul(q, q), dmul(y, y))), -minor);
}
template <class R> inline Dual<R> dsd_cylinder(Dual<R> x, Dual<R> y, Dual<R> z, R h, R r) {  // dual.rs:155-163
  Dual<R> dx = dadds(dsqrt(dadd(dmul(x, x), dmul(z, z))), -r);
  Dual<R> dy = dadds(dabs(y), -h);
  Dual<R> inside = dmin(dmax(dx, dy), dcon(R(0.0f)));
  Dual<R> ox = dmax0(dx), oy = dmax0(dy);
  Dual<R> outside = dsqrt(dadd(dmul(ox, ox), dmul(oy, oy)));
  return dadd(inside, outside);
}
template <class R> inline Dual<R> dsd_capsule(Dual<R> x, Dual<R> y, Dual<R> z, V3<R> a, V3<R> b, R r) {  // dual.rs:165-175
  Dual<R> pax = dadds(x, -a.x), pay = dadds(y, -a.y), paz = dadds(z, -a.z);
  V3<R> ba = b - a;
  R dbb = rmax(dot(ba, ba), R(1e-12f));
  Dual<R> dpb = dadd(dadd(dscale(pax, ba.x), dscale(pay, ba.y)), dscale(paz, ba.z));
  Dual<R> hh = dclamp(dscale(dpb, R(1.0f) / dbb), R(0.0f), R(1.0f));
  Dual<R> rx = dsub(pax, dscale(hh, ba.x));
  Dual<R> ry = dsub(pay, dscale(hh, ba.y));
  Dual<R> rz = dsub(paz, dscale(hh, ba.z));
  return dadds(dsqrt(dadd(dadd(dmul(rx, rx), dmul(ry, ry)), dmul(rz, rz))), -r);
}
template <class R> inline Dual<R> dsd_round_cone(Dual<R> x, Dual<R> y, Dual<R> z, R r1, R r2, R h) {  // dual.rs:177-191
  Dual<R> qx = dsqrt(dadd(dmul(x, x), dmul(z, z)));
  Dual<R> qy = y;
  R b = (r1 - r2) / rmax(h, R(1e-6f));
  R a = rsqrt_(rmax(R(1.0f) - b * b, R(0.0f)));
  Dual<R> k = dadd(dscale(qx, -b), dscale(qy, a));
  if (k.v < R(0.0f)) return dadds(dsqrt(dadd(dmul(qx, qx), dmul(qy, qy))), -r1);
  if (k.v > a * h) {
    Dual<R> qyh = dadds(qy, -h);
    return dadds(dsqrt(dadd(dmul(qx, qx), dmul(qyh, qyh))), -r2);
  }
  return dadds(dadd(dscale(qx, a), dscale(qy, b)), -r1);
}
template <class R> inline Dual<R> dsd_ellipsoid(Dual<R> x, Dual<R> y, Dual<R> z, V3<R> r) {  // dual.rs:193-206
  Dual<R> k0x = dscale(x, R(1.0f) / r.x), k0y = dscale(y, R(1.0f) / r.y), k0z = dscale(z, R(1.0f) / r.z);
  Dual<R> k0 = dsqrt(dadd(dadd(dmul(k0x, k0x), dmul(k0y, k0y)), dmul(k0z, k0z)));
  if (k0.v < R(1e-6f)) return dcon(-rmin(rmin(r.x, r.y), r.z));
  Dual<R> k1x = dscale(x, R(1.0f) / (r.x * r.x)), k1y = dscale(y, R(1.0f) / (r.y * r.y)),
          k1z = dscale(z, R(1.0f) / (r.z * r.z));
  Dual<R> k1 = dmaxc(dsqrt(dadd(dadd(dmul(k1x, k1x), dmul(k1y, k1y)), dmul(k1z, k1z))), R(1e-9f));
  return ddiv(dmul(k0, dadds(k0, R(-1.0f))), k1);
}
template <class R> inline Dual<R> dsd_octahedron(Dual<R> x, Dual<R> y, Dual<R> z, R s) {  // dual.rs:208-210
  return dscale(dadds(dadd(dadd(dabs(x), dabs(y)), dabs(z)), -s), R(0.5773503f));
}
template <class R> inline Dual<R> dsd_hex_prism(Dual<R> x, Dual<R> y, Dual<R> z, R r, R h) {  // dual.rs:212-232
  const R kx(-0.8660254f), ky(0.5f), kz(0.57735f);
  Dual<R> ax0 = dabs(x), az0 = dabs(z);
  Dual<R> ay = dadds(dabs(y), -h);
  Dual<R> tpre = dadd(dscale(ax0, kx), dscale(az0, ky));
  Dual<R> t = dscale(dmin(tpre, dcon(R(0.0f))), R(2.0f));
  Dual<R> ax = dsub(ax0, dscale(t, kx));
  Dual<R> az = dsub(az0, dscale(t, ky));
  Dual<R> clamped = dclamp(ax, -kz * r, kz * r);
  Dual<R> diffx = dsub(ax, clamped);
  Dual<R> diffz = dadds(az, -r);
  R sign = diffz.v >= R(0.0f) ? R(1.0f) : R(-1.0f);
  Dual<R> dx = dscale(dsqrt(dadd(dmul(diffx, diffx), dmul(diffz, diffz))), sign);
  Dual<R> dy = ay;
  Dual<R> inside = dmin(dmax(dx, dy), dcon(R(0.0f)));
  Dual<R> ox = dmax0(dx), oy = dmax0(dy);
  Dual<R> outside = dsqrt(dadd(dmul(ox, ox), dmul(oy, oy)));
  return dadd(inside, outside);
}
template <class R> inline Dual<R> dunion(Dual<R> a, Dual<R> b) { return dmin(a, b); }          // dual.rs:278
template <class R> inline Dual<R> dsubtract(Dual<R> a, Dual<R> b) { return dmax(a, dneg(b)); }  // dual.rs:281
template <class R> inline Dual<R> dsmooth_union(Dual<R> a, Dual<R> b, R k) {                    // dual.rs:284-291
  if (k <= R(0.0f)) return dunion(a, b);
  Dual<R> h = dclamp(dadds(dscale(dsub(b, a), R(0.5f) / k), R(0.5f)), R(0.0f), R(1.0f));
  Dual<R> omh = dsub(dcon(R(1.0f)), h);
  return dsub(dadd(dmul(b, omh), dmul(a, h)), dscale(dmul(h, omh), k));
}
template <class R> inline Dual<R> drep_axis(Dual<R> v, R period) {  // dual.rs:253-262
  if (period > R(0.0f)) return dadds(v, -rround(v.v / period) * period);
  return v;
}

}  // namespace orc
