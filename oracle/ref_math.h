// Oracle: CPU restatement of the reference's math layer.
//
// TEST INFRASTRUCTURE ONLY. Nothing under natsel_amd/ may include, link or call this
// code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
// as the checker.
//
// Every helper keeps the reference's operand order and association exactly
// (compiled with -ffp-contract=off, no fast-math), because the reference engine is
// bit-deterministic and the golden state hashes pin these bits.
//
// Sources: src/math/functions.rs, src/math/types.rs, src/collision/internal_math.rs,
// src/base/constants.rs.
#pragma once
#include "ref_plant.h"
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace nslref {

// src/base/constants.rs:6-49, :81-113 (length units fixed at 1.0)
constexpr float kPi = 3.14159265359f;
constexpr int kNullIndex = -1;
constexpr int kGraphColorCount = 24;
constexpr float kMinFrictionWeight = 1e-10f;
constexpr float kMaxRotation = 0.25f * kPi;
constexpr float kContactRecycleAngularDistance = 0.99240388f;
constexpr float kAabbMarginFraction = 0.125f;
constexpr float kTimeToSleep = 0.5f;
constexpr int kMaxManifoldPoints = 4;
constexpr int kGyroscopicIterations = 1;
constexpr float kParallelEdgeTolerance = 0.005f;
constexpr int kShapePower = 22;
constexpr int kChildPower = 64 - 2 * kShapePower;
constexpr int kMaxShapes = 1 << kShapePower;
constexpr long long kMaxChildShapes = 1LL << kChildPower;
inline float huge_() { return 1.0e5f * 1.0f; }
inline float linear_slop() { return 0.005f * 1.0f; }
inline float speculative_distance() { return 4.0f * linear_slop(); }
inline float contact_recycle_distance() { return 10.0f * linear_slop(); }
inline float max_aabb_margin() { return 0.05f * 1.0f; }
inline float min_capsule_length() { return linear_slop(); }

struct Vec2 { float x, y; };
struct Vec3 { float x, y, z; };
struct Quat { Vec3 v; float s; };
struct Xf { Vec3 p; Quat q; };            // Transform (types.rs:73-86)
struct Pos { double x, y, z; };           // f64 world position (types.rs:88-112)
struct WXf { Pos p; Quat q; };            // WorldTransform (types.rs:114-135)
struct Mat3 { Vec3 cx, cy, cz; };         // column major (types.rs:137-155)
struct Mat2 { Vec2 cx, cy; };             // internal_math.rs:8-20
struct AABB { Vec3 lo, hi; };
struct Plane { Vec3 n; float off; };
struct SegDist { Vec3 p1; float f1; Vec3 p2; float f2; };

constexpr Vec3 kZero3 = {0.f, 0.f, 0.f};
constexpr Quat kQuatIdentity = {{0.f, 0.f, 0.f}, 1.f};

// functions.rs:15-66: ties return the other operand.
inline float minf(float a, float b) { return a < b ? a : b; }
inline float maxf(float a, float b) { return a > b ? a : b; }
inline int mini(int a, int b) { return a < b ? a : b; }
inline int maxi(int a, int b) { return a > b ? a : b; }
inline float absf(float a) { return a < 0.f ? -a : a; }
inline float clampf(float a, float lo, float hi) { return a < lo ? lo : (hi < a ? hi : a); }

inline Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
inline Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 mulv(Vec3 a, Vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Vec3 scale(float s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline Vec3 neg(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(Vec3 v) { return std::sqrt(dot(v, v)); }
inline float length_sq(Vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float distance_sq(Vec3 a, Vec3 b) {
  Vec3 d = {b.x - a.x, b.y - a.y, b.z - a.z};
  return d.x * d.x + d.y * d.y + d.z * d.z;
}
inline float distance(Vec3 a, Vec3 b) { return length(Vec3{b.x - a.x, b.y - a.y, b.z - a.z}); }
// functions.rs:266-275
inline Vec3 normalize(Vec3 a) {
  float ls = a.x * a.x + a.y * a.y + a.z * a.z;
  if (ls > 1000.0f * FLT_MIN) {
    float s = 1.0f / std::sqrt(ls);
    return {s * a.x, s * a.y, s * a.z};
  }
  return kZero3;
}
// functions.rs:290-297
inline Vec3 perp(Vec3 a) {
  Vec3 p = (a.x < -0.5f || 0.5f < a.x) ? Vec3{a.y, -a.x, 0.0f} : Vec3{0.0f, a.z, -a.y};
  return normalize(p);
}
inline Vec3 mul_add(Vec3 a, float s, Vec3 b) { return {a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }
inline Vec3 mul_sub(Vec3 a, float s, Vec3 b) { return {a.x - s * b.x, a.y - s * b.y, a.z - s * b.z}; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Vec3 lerp(Vec3 a, Vec3 b, float t) {
  return {(1.0f - t) * a.x + t * b.x, (1.0f - t) * a.y + t * b.y, (1.0f - t) * a.z + t * b.z};
}
inline Vec3 blend2(float s, Vec3 a, float t, Vec3 b) {
  return {s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z};
}
inline Vec3 abs3(Vec3 a) { return {absf(a.x), absf(a.y), absf(a.z)}; }
inline Vec3 min3(Vec3 a, Vec3 b) { return {minf(a.x, b.x), minf(a.y, b.y), minf(a.z, b.z)}; }
inline Vec3 max3(Vec3 a, Vec3 b) { return {maxf(a.x, b.x), maxf(a.y, b.y), maxf(a.z, b.z)}; }

// quaternion (functions.rs:402-466)
inline Vec3 rotate(Quat q, Vec3 v) {
  Vec3 t1 = cross(q.v, v);
  Vec3 t2 = mul_add(t1, q.s, v);
  Vec3 t3 = cross(q.v, t2);
  return mul_add(v, 2.0f, t3);
}
inline Vec3 inv_rotate(Quat q, Vec3 v) {
  Vec3 t1 = cross(q.v, v);
  Vec3 t2 = mul_sub(t1, q.s, v);
  Vec3 t3 = cross(q.v, t2);
  return mul_add(v, 2.0f, t3);
}
inline float dot_quat(Quat a, Quat b) { return a.v.x * b.v.x + a.v.y * b.v.y + a.v.z * b.v.z + a.s * b.s; }
inline Quat mul_quat(Quat q1, Quat q2) {
  Vec3 t1 = cross(q1.v, q2.v);
  Vec3 t2 = mul_add(t1, q1.s, q2.v);
  Vec3 t3 = mul_add(t2, q2.s, q1.v);
  return Quat{t3, q1.s * q2.s - dot(q1.v, q2.v)};
}
inline Quat inv_mul_quat(Quat q1, Quat q2) {
  Vec3 t1 = cross(q2.v, q1.v);
  Vec3 t2 = mul_add(t1, q1.s, q2.v);
  Vec3 t3 = mul_sub(t2, q2.s, q1.v);
  return Quat{t3, q1.s * q2.s + dot(q1.v, q2.v)};
}
inline Quat conjugate(Quat q) { return Quat{{-q.v.x, -q.v.y, -q.v.z}, q.s}; }
inline Quat normalize_quat(Quat q) {
  float ls = dot_quat(q, q);
  if (ls > 1000.0f * FLT_MIN) {
    float s = 1.0f / std::sqrt(ls);
    return Quat{{s * q.v.x, s * q.v.y, s * q.v.z}, s * q.s};
  }
  return kQuatIdentity;
}
// functions.rs:511-520: normalised linear blend along the shorter arc
inline Quat nlerp(Quat q1, Quat q2, float t) {
  if (dot_quat(q1, q2) < 0.0f) q1 = Quat{{-q1.v.x, -q1.v.y, -q1.v.z}, -q1.s};
  Vec3 v = lerp(q1.v, q2.v, t);
  float s = (1.0f - t) * q1.s + t * q2.s;
  return normalize_quat(Quat{v, s});
}
// internal_math.rs:322-325
inline Quat integrate_rotation(Quat q1, Vec3 dr) {
  Quat qd = mul_quat(Quat{scale(0.5f, dr), 0.0f}, q1);
  return normalize_quat(Quat{add(q1.v, qd.v), qd.s + q1.s});
}

// transforms (functions.rs:484-516)
inline Xf inv_mul_xf(Xf a, Xf b) { return Xf{inv_rotate(a.q, sub(b.p, a.p)), inv_mul_quat(a.q, b.q)}; }
inline Xf invert_xf(Xf t) { return Xf{inv_rotate(t.q, neg(t.p)), conjugate(t.q)}; }
inline Vec3 xf_point(Xf t, Vec3 v) { return add(rotate(t.q, v), t.p); }
inline Vec3 inv_xf_point(Xf t, Vec3 v) { return inv_rotate(t.q, sub(v, t.p)); }

// f64 world positions (functions.rs:520-620)
inline Vec3 sub_pos(Pos a, Pos b) {
  return {(float)(a.x - b.x), (float)(a.y - b.y), (float)(a.z - b.z)};
}
inline Pos offset_pos(Pos p, Vec3 d) { return Pos{p.x + (double)d.x, p.y + (double)d.y, p.z + (double)d.z}; }
inline float round_down_float(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafterf(f, -INFINITY);
  return f;
}
inline float round_up_float(double x) {
  float f = (float)x;
  if ((double)f < x) f = std::nextafterf(f, INFINITY);
  return f;
}
inline Pos xf_world_point(WXf t, Vec3 p) { return offset_pos(t.p, rotate(t.q, p)); }
inline Vec3 inv_xf_world_point(WXf t, Pos p) { return inv_rotate(t.q, sub_pos(p, t.p)); }
inline Xf inv_mul_wxf(WXf a, WXf b) { return Xf{inv_xf_world_point(a, b.p), inv_mul_quat(a.q, b.q)}; }
inline AABB offset_aabb(AABB b, Pos o) {
  return AABB{{round_down_float(o.x + (double)b.lo.x), round_down_float(o.y + (double)b.lo.y),
               round_down_float(o.z + (double)b.lo.z)},
              {round_up_float(o.x + (double)b.hi.x), round_up_float(o.y + (double)b.hi.y),
               round_up_float(o.z + (double)b.hi.z)}};
}

// matrices (functions.rs:677-836)
inline Mat3 mat3(Vec3 cx, Vec3 cy, Vec3 cz) { return Mat3{cx, cy, cz}; }
constexpr Mat3 kMat3Zero = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
inline float det(Mat3 m) { return dot(m.cx, cross(m.cy, m.cz)); }
inline Vec3 mul_mv(Mat3 m, Vec3 a) {
  return {m.cx.x * a.x + m.cy.x * a.y + m.cz.x * a.z, m.cx.y * a.x + m.cy.y * a.y + m.cz.y * a.z,
          m.cx.z * a.x + m.cy.z * a.y + m.cz.z * a.z};
}
inline Mat3 add_mm(Mat3 a, Mat3 b) { return {add(a.cx, b.cx), add(a.cy, b.cy), add(a.cz, b.cz)}; }
inline Mat3 sub_mm(Mat3 a, Mat3 b) { return {sub(a.cx, b.cx), sub(a.cy, b.cy), sub(a.cz, b.cz)}; }
inline Mat3 mul_sm(float s, Mat3 a) { return {scale(s, a.cx), scale(s, a.cy), scale(s, a.cz)}; }
inline Mat3 mul_mm(Mat3 a, Mat3 b) { return {mul_mv(a, b.cx), mul_mv(a, b.cy), mul_mv(a, b.cz)}; }
inline Mat3 transpose(Mat3 m) {
  return {{m.cx.x, m.cy.x, m.cz.x}, {m.cx.y, m.cy.y, m.cz.y}, {m.cx.z, m.cy.z, m.cz.z}};
}
inline Mat3 invert_matrix(Mat3 m) {
  float d = det(m);
  if (absf(d) > 1000.0f * FLT_MIN) {
    float inv_d = 1.0f / d;
    Mat3 out = {scale(inv_d, cross(m.cy, m.cz)), scale(inv_d, cross(m.cz, m.cx)), scale(inv_d, cross(m.cx, m.cy))};
    return transpose(out);
  }
  return kMat3Zero;
}
inline Vec3 solve3(Mat3 m, Vec3 a) {
  float d = det(m);
  if (absf(d) > 1000.0f * FLT_MIN) {
    float inv_d = 1.0f / d;
    Vec3 sx = cross(m.cy, m.cz), sy = cross(m.cz, m.cx), sz = cross(m.cx, m.cy);
    return {inv_d * dot(sx, a), inv_d * dot(sy, a), inv_d * dot(sz, a)};
  }
  return kZero3;
}
inline Mat3 invert_t(Mat3 m) {
  float d = det(m);
  if (absf(d) > 1000.0f * FLT_MIN) {
    float inv_d = 1.0f / d;
    return {scale(inv_d, cross(m.cy, m.cz)), scale(inv_d, cross(m.cz, m.cx)), scale(inv_d, cross(m.cx, m.cy))};
  }
  return kMat3Zero;
}
inline Mat3 abs_mat3(Mat3 m) { return {abs3(m.cx), abs3(m.cy), abs3(m.cz)}; }
inline Mat3 matrix_from_quat(Quat q) {
  float xx = q.v.x * q.v.x, yy = q.v.y * q.v.y, zz = q.v.z * q.v.z;
  float xy = q.v.x * q.v.y, xz = q.v.x * q.v.z, xw = q.v.x * q.s;
  float yz = q.v.y * q.v.z, yw = q.v.y * q.s, zw = q.v.z * q.s;
  return {{1.0f - 2.0f * (yy + zz), 2.0f * (xy + zw), 2.0f * (xz - yw)},
          {2.0f * (xy - zw), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + xw)},
          {2.0f * (xz + yw), 2.0f * (yz - xw), 1.0f - 2.0f * (xx + yy)}};
}
inline Mat3 diag3(float a, float b, float c) { return {{a, 0, 0}, {0, b, 0}, {0, 0, c}}; }
// functions.rs:1190-1211
inline Mat3 steiner(float mass, Vec3 o) {
  float ixx = mass * (o.y * o.y + o.z * o.z);
  float iyy = mass * (o.x * o.x + o.z * o.z);
  float izz = mass * (o.x * o.x + o.y * o.y);
  float ixy = -mass * o.x * o.y;
  float ixz = -mass * o.x * o.z;
  float iyz = -mass * o.y * o.z;
  Mat3 out;
  out.cx = {ixx, ixy, ixz};
  out.cy = {ixy, iyy, iyz};
  out.cz = {ixz, iyz, izz};
  return out;
}
// internal_math.rs:256-262
inline Vec3 modified_cross(Vec3 a, Vec3 b) {
  return {a.y * b.z + a.z * b.y, a.z * b.x + a.x * b.z, a.x * b.y + a.y * b.x};
}
inline float scalar_triple(Vec3 a, Vec3 b, Vec3 c) {
  Vec3 d = {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
  return a.x * d.x + a.y * d.y + a.z * d.z;
}
inline Mat3 rotate_inertia(Quat q, Mat3 ci) {
  Mat3 r = matrix_from_quat(q);
  return mul_mm(r, mul_mm(ci, transpose(r)));
}

// AABB (functions.rs:840-935)
inline Vec3 aabb_center(AABB a) { return scale(0.5f, add(a.hi, a.lo)); }
inline Vec3 aabb_extents(AABB a) { return scale(0.5f, sub(a.hi, a.lo)); }
inline AABB aabb_union(AABB a, AABB b) { return {min3(a.lo, b.lo), max3(a.hi, b.hi)}; }
inline AABB aabb_inflate(AABB a, float e) {
  Vec3 r = {e, e, e};
  return {sub(a.lo, r), add(a.hi, r)};
}
inline bool aabb_overlaps(AABB a, AABB b) {
  if (a.hi.x < b.lo.x || a.lo.x > b.hi.x) return false;
  if (a.hi.y < b.lo.y || a.lo.y > b.hi.y) return false;
  if (a.hi.z < b.lo.z || a.lo.z > b.hi.z) return false;
  return true;
}
inline bool aabb_contains(AABB a, AABB b) {
  if (a.lo.x > b.lo.x || b.hi.x > a.hi.x) return false;
  if (a.lo.y > b.lo.y || b.hi.y > a.hi.y) return false;
  if (a.lo.z > b.lo.z || b.hi.z > a.hi.z) return false;
  return true;
}
inline AABB aabb_transform(Xf t, AABB a) {
  Vec3 c = xf_point(t, aabb_center(a));
  Mat3 m = matrix_from_quat(t.q);
  Vec3 e = mul_mv(abs_mat3(m), aabb_extents(a));
  return {sub(c, e), add(c, e)};
}

// planes (internal_math.rs:332-380)
inline Plane plane_from_normal_point(Vec3 n, Vec3 p) { return Plane{n, dot(n, p)}; }
inline Plane transform_plane(Xf t, Plane p) {
  Vec3 n = rotate(t.q, p.n);
  return Plane{n, p.off + dot(n, t.p)};
}
inline float plane_sep(Plane p, Vec3 x) { return dot(p.n, x) - p.off; }

// internal_math.rs:263-281
inline Vec3 arbitrary_perp(Vec3 v) {
  Vec3 p;
  if (v.x < -0.5f || v.x > 0.5f) {
    const float a = 0.67f, b = -0.42f;
    p = {a * v.y + b * v.z, -a * v.x, -b * v.x};
  } else if (v.y < -0.5f || v.y > 0.5f) {
    const float a = 0.67f, c = -0.42f;
    p = {a * v.y, -a * v.x + c * v.z, -c * v.y};
  } else {
    const float a = 0.67f, b = -0.42f;
    p = {a * v.z, b * v.z, -a * v.x - b * v.y};
  }
  return normalize(p);
}

// internal_math.rs (2-D helpers)
inline Vec2 mul_mv2(Mat2 m, Vec2 a) { return {m.cx.x * a.x + m.cy.x * a.y, m.cx.y * a.x + m.cy.y * a.y}; }
inline float det2(Mat2 m) { return m.cx.x * m.cy.y - m.cx.y * m.cy.x; }
inline Mat2 invert2(Mat2 m) {
  float d = det2(m);
  if (absf(d) > 1000.0f * FLT_MIN) {
    float inv = 1.0f / d;
    return Mat2{{inv * m.cy.y, -inv * m.cx.y}, {-inv * m.cy.x, inv * m.cx.x}};
  }
  return Mat2{{0, 0}, {0, 0}};
}

// functions.rs:1069-1101
inline SegDist line_distance(Vec3 p1, Vec3 d1, Vec3 p2, Vec3 d2) {
  SegDist r{};
  float a11 = dot(d1, d1);
  float a12 = -dot(d1, d2);
  float a21 = dot(d2, d1);
  float a22 = -dot(d2, d2);
  Vec3 w = sub(p1, p2);
  float b1 = -dot(d1, w);
  float b2 = -dot(d2, w);
  float d = a11 * a22 - a12 * a21;
  if (d * d < 1000.0f * FLT_MIN) {
    float s1 = dot(sub(p2, p1), d1) / dot(d1, d1);
    r.p1 = mul_add(p1, s1, d1);
    r.f1 = s1;
    r.p2 = p2;
    r.f2 = 0.0f;
    return r;
  }
  float s1 = (a22 * b1 - a12 * b2) / d;
  float s2 = (a11 * b2 - a21 * b1) / d;
  r.p1 = mul_add(p1, s1, d1);
  r.f1 = s1;
  r.p2 = mul_add(p2, s2, d2);
  r.f2 = s2;
  return r;
}

// functions.rs:1105-1165
inline SegDist segment_distance(Vec3 p1, Vec3 q1, Vec3 p2, Vec3 q2) {
  SegDist r{};
  Vec3 d1 = sub(q1, p1), d2 = sub(q2, p2), rr = sub(p1, p2);
  float a = dot(d1, d1), b = dot(d1, d2), c = dot(d1, rr), e = dot(d2, d2), f = dot(d2, rr);
  const float eps = 100.0f * FLT_EPSILON;
  if (a < eps && e < eps) {
    r.p1 = p1; r.f1 = 0.0f; r.p2 = p2; r.f2 = 0.0f;
    return r;
  }
  if (a < eps) {
    float s2 = clampf(f / e, 0.0f, 1.0f);
    r.p1 = p1; r.f1 = 0.0f; r.p2 = mul_add(p2, s2, d2); r.f2 = s2;
    return r;
  }
  if (e < eps) {
    float s1 = clampf(-c / a, 0.0f, 1.0f);
    r.p1 = mul_add(p1, s1, d1); r.f1 = s1; r.p2 = p2; r.f2 = 0.0f;
    return r;
  }
  float denom = a * e - b * b;
  float s1 = denom > 1000.0f * FLT_MIN ? clampf((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
  float s2 = (b * s1 + f) / e;
  if (s2 < 0.0f) {
    s1 = clampf(-c / a, 0.0f, 1.0f);
    s2 = 0.0f;
  } else if (s2 > 1.0f) {
    s1 = clampf(PLANTED(11) ? (b + c) / a : (b - c) / a, 0.0f, 1.0f);
    s2 = 1.0f;
  }
  r.p1 = mul_add(p1, s1, d1); r.f1 = s1; r.p2 = mul_add(p2, s2, d2); r.f2 = s2;
  return r;
}

// functions.rs:1169-1186
inline Vec3 point_to_segment(Vec3 a, Vec3 b, Vec3 q) {
  Vec3 ab = sub(b, a), aq = sub(q, a);
  float alpha = dot(ab, aq);
  if (alpha <= 0.0f) return a;
  float den = dot(ab, ab);
  if (alpha > den) return PLANTED(12) ? a : b;
  alpha /= den;
  return mul_add(a, alpha, ab);
}

inline bool is_normalized(Vec3 a) { return absf(1.0f - dot(a, a)) < 100.0f * FLT_EPSILON; }

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint64_t d2u(double f) { uint64_t u; std::memcpy(&u, &f, 8); return u; }

}  // namespace nslref
