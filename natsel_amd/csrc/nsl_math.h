// Rigid-body math shared by the host world code and the gfx950 kernels.
//
// Every helper keeps the reference engine's operand order and association, because
// the engine is bit-deterministic: the GPU step must reproduce the CPU step's bits.
// All code including this header is compiled with -ffp-contract=off, correctly
// rounded f32 division and square root, and f32 denormals preserved.
//
// Sources: src/math/functions.rs, src/math/types.rs, src/collision/internal_math.rs,
// src/base/constants.rs, src/simulation/contact_solver.rs:19-115 (SymMatrix3).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NSL_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define NSL_HD inline
#endif

namespace nsl {

constexpr float kPi = 3.14159265359f;
constexpr int kNullIndex = -1;
constexpr int kColorCount = 24;              // constants.rs:23 GRAPH_COLOR_COUNT
constexpr int kOverflowColor = kColorCount - 1;
constexpr int kDynamicColorCount = kColorCount - 4;
constexpr float kMinFrictionWeight = 1e-10f;
constexpr float kMaxRotation = 0.25f * kPi;
constexpr float kRecycleAngularDistance = 0.99240388f;
constexpr float kAabbMarginFraction = 0.125f;
constexpr int kGyroIterations = 1;
constexpr float kParallelEdgeTolerance = 0.005f;
constexpr float kLinearSlop = 0.005f;
constexpr float kSpeculativeDistance = 4.0f * kLinearSlop;
constexpr float kRecycleDistance = 10.0f * kLinearSlop;
constexpr float kMaxAabbMargin = 0.05f;
constexpr float kHuge = 1.0e5f;
constexpr float kFltMin = 1.17549435e-38f;
constexpr float kFltMax = 3.40282347e+38f;
constexpr float kFltEps = 1.19209290e-07f;

struct Vec2 { float x, y; };
struct Vec3 { float x, y, z; };
struct Quat { Vec3 v; float s; };
struct Xf { Vec3 p; Quat q; };
struct Pos { double x, y, z; };
struct WXf { Pos p; Quat q; };
struct Mat3 { Vec3 cx, cy, cz; };
struct Mat2 { Vec2 cx, cy; };
struct Aabb { Vec3 lo, hi; };
struct Plane { Vec3 n; float off; };
struct SymMat3 { float cxx, cxy, cxz, cyy, cyz, czz; };
struct SegDist { Vec3 p1; float f1; Vec3 p2; float f2; };

NSL_HD uint32_t bits_of(float f) { return __builtin_bit_cast(uint32_t, f); }
NSL_HD float float_of(uint32_t u) { return __builtin_bit_cast(float, u); }
NSL_HD uint64_t bits_of(double f) { return __builtin_bit_cast(uint64_t, f); }

NSL_HD float minf(float a, float b) { return a < b ? a : b; }
NSL_HD float maxf(float a, float b) { return a > b ? a : b; }
NSL_HD int mini(int a, int b) { return a < b ? a : b; }
NSL_HD int maxi(int a, int b) { return a > b ? a : b; }
NSL_HD float absf(float a) { return a < 0.0f ? -a : a; }
NSL_HD float clampf(float a, float lo, float hi) { return a < lo ? lo : (hi < a ? hi : a); }
NSL_HD float sqrtf_(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(x);
#else
  return std::sqrt(x);
#endif
}

NSL_HD Vec3 zero3() { return Vec3{0.0f, 0.0f, 0.0f}; }
NSL_HD Quat quat_identity() { return Quat{{0.0f, 0.0f, 0.0f}, 1.0f}; }
NSL_HD Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
NSL_HD Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
NSL_HD Vec3 mulv(Vec3 a, Vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
NSL_HD Vec3 scale(float s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
NSL_HD Vec3 neg(Vec3 a) { return {-a.x, -a.y, -a.z}; }
NSL_HD float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
NSL_HD float length(Vec3 v) { return sqrtf_(dot(v, v)); }
NSL_HD float length_sq(Vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
NSL_HD float distance_sq(Vec3 a, Vec3 b) {
  Vec3 d = {b.x - a.x, b.y - a.y, b.z - a.z};
  return d.x * d.x + d.y * d.y + d.z * d.z;
}
NSL_HD float distance(Vec3 a, Vec3 b) { return length(Vec3{b.x - a.x, b.y - a.y, b.z - a.z}); }
NSL_HD Vec3 normalize(Vec3 a) {
  float ls = a.x * a.x + a.y * a.y + a.z * a.z;
  if (ls > 1000.0f * kFltMin) {
    float s = 1.0f / sqrtf_(ls);
    return {s * a.x, s * a.y, s * a.z};
  }
  return zero3();
}
NSL_HD Vec3 perp(Vec3 a) {
  Vec3 p = (a.x < -0.5f || 0.5f < a.x) ? Vec3{a.y, -a.x, 0.0f} : Vec3{0.0f, a.z, -a.y};
  return normalize(p);
}
NSL_HD Vec3 mul_add(Vec3 a, float s, Vec3 b) { return {a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }
NSL_HD Vec3 mul_sub(Vec3 a, float s, Vec3 b) { return {a.x - s * b.x, a.y - s * b.y, a.z - s * b.z}; }
NSL_HD Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
NSL_HD Vec3 lerp(Vec3 a, Vec3 b, float t) {
  return {(1.0f - t) * a.x + t * b.x, (1.0f - t) * a.y + t * b.y, (1.0f - t) * a.z + t * b.z};
}
NSL_HD Vec3 blend2(float s, Vec3 a, float t, Vec3 b) { return {s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z}; }
NSL_HD Vec3 abs3(Vec3 a) { return {absf(a.x), absf(a.y), absf(a.z)}; }
NSL_HD Vec3 min3(Vec3 a, Vec3 b) { return {minf(a.x, b.x), minf(a.y, b.y), minf(a.z, b.z)}; }
NSL_HD Vec3 max3(Vec3 a, Vec3 b) { return {maxf(a.x, b.x), maxf(a.y, b.y), maxf(a.z, b.z)}; }

NSL_HD Vec3 rotate(Quat q, Vec3 v) {
  Vec3 t1 = cross(q.v, v);
  Vec3 t2 = mul_add(t1, q.s, v);
  Vec3 t3 = cross(q.v, t2);
  return mul_add(v, 2.0f, t3);
}
NSL_HD Vec3 inv_rotate(Quat q, Vec3 v) {
  Vec3 t1 = cross(q.v, v);
  Vec3 t2 = mul_sub(t1, q.s, v);
  Vec3 t3 = cross(q.v, t2);
  return mul_add(v, 2.0f, t3);
}
// contact_solver.rs:291-300
NSL_HD Vec3 inv_rotate_wide(Quat q, Vec3 a) {
  Vec3 t0 = cross(q.v, a);
  Vec3 t = {t0.x + t0.x, t0.y + t0.y, t0.z + t0.z};
  Vec3 u = cross(q.v, t);
  return {(a.x - q.s * t.x) + u.x, (a.y - q.s * t.y) + u.y, (a.z - q.s * t.z) + u.z};
}
NSL_HD float dot_quat(Quat a, Quat b) { return a.v.x * b.v.x + a.v.y * b.v.y + a.v.z * b.v.z + a.s * b.s; }
NSL_HD Quat mul_quat(Quat q1, Quat q2) {
  Vec3 t1 = cross(q1.v, q2.v);
  Vec3 t2 = mul_add(t1, q1.s, q2.v);
  Vec3 t3 = mul_add(t2, q2.s, q1.v);
  return Quat{t3, q1.s * q2.s - dot(q1.v, q2.v)};
}
NSL_HD Quat inv_mul_quat(Quat q1, Quat q2) {
  Vec3 t1 = cross(q2.v, q1.v);
  Vec3 t2 = mul_add(t1, q1.s, q2.v);
  Vec3 t3 = mul_sub(t2, q2.s, q1.v);
  return Quat{t3, q1.s * q2.s + dot(q1.v, q2.v)};
}
NSL_HD Quat conjugate(Quat q) { return Quat{{-q.v.x, -q.v.y, -q.v.z}, q.s}; }
NSL_HD Quat normalize_quat(Quat q) {
  float ls = dot_quat(q, q);
  if (ls > 1000.0f * kFltMin) {
    float s = 1.0f / sqrtf_(ls);
    return Quat{{s * q.v.x, s * q.v.y, s * q.v.z}, s * q.s};
  }
  return quat_identity();
}
// functions.rs:511-520: normalised linear blend along the shorter arc
NSL_HD Quat nlerp(Quat q1, Quat q2, float t) {
  if (dot_quat(q1, q2) < 0.0f) q1 = Quat{{-q1.v.x, -q1.v.y, -q1.v.z}, -q1.s};
  Vec3 v = lerp(q1.v, q2.v, t);
  float s = (1.0f - t) * q1.s + t * q2.s;
  return normalize_quat(Quat{v, s});
}
NSL_HD Quat integrate_rotation(Quat q1, Vec3 dr) {
  Quat qd = mul_quat(Quat{scale(0.5f, dr), 0.0f}, q1);
  return normalize_quat(Quat{add(q1.v, qd.v), qd.s + q1.s});
}

NSL_HD Xf inv_mul_xf(Xf a, Xf b) { return Xf{inv_rotate(a.q, sub(b.p, a.p)), inv_mul_quat(a.q, b.q)}; }
NSL_HD Xf invert_xf(Xf t) { return Xf{inv_rotate(t.q, neg(t.p)), conjugate(t.q)}; }
NSL_HD Vec3 xf_point(Xf t, Vec3 v) { return add(rotate(t.q, v), t.p); }
NSL_HD Vec3 inv_xf_point(Xf t, Vec3 v) { return inv_rotate(t.q, sub(v, t.p)); }

NSL_HD Vec3 sub_pos(Pos a, Pos b) { return {(float)(a.x - b.x), (float)(a.y - b.y), (float)(a.z - b.z)}; }
NSL_HD Pos offset_pos(Pos p, Vec3 d) { return Pos{p.x + (double)d.x, p.y + (double)d.y, p.z + (double)d.z}; }
// Neighbouring floats toward -inf / +inf for finite inputs (f32::next_down/next_up).
NSL_HD float next_down(float f) {
  uint32_t u = bits_of(f);
  if (f != f || u == 0xff800000u) return f;
  if ((u & 0x7fffffffu) == 0) return float_of(0x80000001u);
  return float_of((u & 0x80000000u) ? u + 1 : u - 1);
}
NSL_HD float next_up(float f) {
  uint32_t u = bits_of(f);
  if (f != f || u == 0x7f800000u) return f;
  if ((u & 0x7fffffffu) == 0) return float_of(0x00000001u);
  return float_of((u & 0x80000000u) ? u - 1 : u + 1);
}
NSL_HD float round_down_float(double x) {
  float f = (float)x;
  return ((double)f > x) ? next_down(f) : f;
}
NSL_HD float round_up_float(double x) {
  float f = (float)x;
  return ((double)f < x) ? next_up(f) : f;
}
NSL_HD Pos xf_world_point(WXf t, Vec3 p) { return offset_pos(t.p, rotate(t.q, p)); }
NSL_HD Xf inv_mul_wxf(WXf a, WXf b) { return Xf{inv_rotate(a.q, sub_pos(b.p, a.p)), inv_mul_quat(a.q, b.q)}; }
NSL_HD Aabb offset_aabb(Aabb b, Pos o) {
  return Aabb{{round_down_float(o.x + (double)b.lo.x), round_down_float(o.y + (double)b.lo.y),
               round_down_float(o.z + (double)b.lo.z)},
              {round_up_float(o.x + (double)b.hi.x), round_up_float(o.y + (double)b.hi.y),
               round_up_float(o.z + (double)b.hi.z)}};
}

NSL_HD Mat3 mat3_zero() { return Mat3{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}; }
NSL_HD float det(Mat3 m) { return dot(m.cx, cross(m.cy, m.cz)); }
NSL_HD Vec3 mul_mv(Mat3 m, Vec3 a) {
  return {m.cx.x * a.x + m.cy.x * a.y + m.cz.x * a.z, m.cx.y * a.x + m.cy.y * a.y + m.cz.y * a.z,
          m.cx.z * a.x + m.cy.z * a.y + m.cz.z * a.z};
}
NSL_HD Mat3 add_mm(Mat3 a, Mat3 b) { return {add(a.cx, b.cx), add(a.cy, b.cy), add(a.cz, b.cz)}; }
NSL_HD Mat3 mul_sm(float s, Mat3 a) { return {scale(s, a.cx), scale(s, a.cy), scale(s, a.cz)}; }
NSL_HD Mat3 mul_mm(Mat3 a, Mat3 b) { return {mul_mv(a, b.cx), mul_mv(a, b.cy), mul_mv(a, b.cz)}; }
NSL_HD Mat3 transpose(Mat3 m) { return {{m.cx.x, m.cy.x, m.cz.x}, {m.cx.y, m.cy.y, m.cz.y}, {m.cx.z, m.cy.z, m.cz.z}}; }
NSL_HD Mat3 invert_matrix(Mat3 m) {
  float d = det(m);
  if (absf(d) > 1000.0f * kFltMin) {
    float inv = 1.0f / d;
    return transpose(Mat3{scale(inv, cross(m.cy, m.cz)), scale(inv, cross(m.cz, m.cx)), scale(inv, cross(m.cx, m.cy))});
  }
  return mat3_zero();
}
NSL_HD Mat3 invert_t(Mat3 m) {
  float d = det(m);
  if (absf(d) > 1000.0f * kFltMin) {
    float inv = 1.0f / d;
    return Mat3{scale(inv, cross(m.cy, m.cz)), scale(inv, cross(m.cz, m.cx)), scale(inv, cross(m.cx, m.cy))};
  }
  return mat3_zero();
}
NSL_HD Vec3 solve3(Mat3 m, Vec3 a) {
  float d = det(m);
  if (absf(d) > 1000.0f * kFltMin) {
    float inv = 1.0f / d;
    Vec3 sx = cross(m.cy, m.cz), sy = cross(m.cz, m.cx), sz = cross(m.cx, m.cy);
    return {inv * dot(sx, a), inv * dot(sy, a), inv * dot(sz, a)};
  }
  return zero3();
}
NSL_HD Mat3 abs_mat3(Mat3 m) { return {abs3(m.cx), abs3(m.cy), abs3(m.cz)}; }
NSL_HD Mat3 matrix_from_quat(Quat q) {
  float xx = q.v.x * q.v.x, yy = q.v.y * q.v.y, zz = q.v.z * q.v.z;
  float xy = q.v.x * q.v.y, xz = q.v.x * q.v.z, xw = q.v.x * q.s;
  float yz = q.v.y * q.v.z, yw = q.v.y * q.s, zw = q.v.z * q.s;
  return {{1.0f - 2.0f * (yy + zz), 2.0f * (xy + zw), 2.0f * (xz - yw)},
          {2.0f * (xy - zw), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + xw)},
          {2.0f * (xz + yw), 2.0f * (yz - xw), 1.0f - 2.0f * (xx + yy)}};
}
NSL_HD Mat3 diag3(float a, float b, float c) { return {{a, 0, 0}, {0, b, 0}, {0, 0, c}}; }
NSL_HD Mat3 steiner(float mass, Vec3 o) {
  float ixx = mass * (o.y * o.y + o.z * o.z);
  float iyy = mass * (o.x * o.x + o.z * o.z);
  float izz = mass * (o.x * o.x + o.y * o.y);
  float ixy = -mass * o.x * o.y;
  float ixz = -mass * o.x * o.z;
  float iyz = -mass * o.y * o.z;
  return Mat3{{ixx, ixy, ixz}, {ixy, iyy, iyz}, {ixz, iyz, izz}};
}
NSL_HD Vec3 modified_cross(Vec3 a, Vec3 b) { return {a.y * b.z + a.z * b.y, a.z * b.x + a.x * b.z, a.x * b.y + a.y * b.x}; }
NSL_HD Mat3 rotate_inertia(Quat q, Mat3 ci) {
  Mat3 r = matrix_from_quat(q);
  return mul_mm(r, mul_mm(ci, transpose(r)));
}

NSL_HD SymMat3 sym_from(Mat3 m) { return {m.cx.x, m.cx.y, m.cx.z, m.cy.y, m.cy.z, m.cz.z}; }
NSL_HD Vec3 sym_mv(SymMat3 s, Vec3 a) {
  return {s.cxx * a.x + (s.cxy * a.y + s.cxz * a.z), s.cxy * a.x + (s.cyy * a.y + s.cyz * a.z),
          s.cxz * a.x + (s.cyz * a.y + s.czz * a.z)};
}
NSL_HD SymMat3 add_sym(SymMat3 a, SymMat3 b) {
  return {a.cxx + b.cxx, a.cxy + b.cxy, a.cxz + b.cxz, a.cyy + b.cyy, a.cyz + b.cyz, a.czz + b.czz};
}
NSL_HD SymMat3 invert_sym(SymMat3 m) {
  float cxx = m.cyy * m.czz - m.cyz * m.cyz;
  float cxy = m.cxz * m.cyz - m.cxy * m.czz;
  float cxz = m.cxy * m.cyz - m.cxz * m.cyy;
  float cyy = m.cxx * m.czz - m.cxz * m.cxz;
  float cyz = m.cxy * m.cxz - m.cxx * m.cyz;
  float czz = m.cxx * m.cyy - m.cxy * m.cxy;
  float d = m.cxx * cxx + (m.cxy * cxy + m.cxz * cxz);
  float inv = absf(d) > 1000.0f * kFltMin ? 1.0f / d : 0.0f;
  return {inv * cxx, inv * cxy, inv * cxz, inv * cyy, inv * cyz, inv * czz};
}
NSL_HD Vec2 mul_mv2(Mat2 m, Vec2 a) { return {m.cx.x * a.x + m.cy.x * a.y, m.cx.y * a.x + m.cy.y * a.y}; }
NSL_HD Mat2 invert2(Mat2 m) {
  float d = m.cx.x * m.cy.y - m.cx.y * m.cy.x;
  if (absf(d) > 1000.0f * kFltMin) {
    float inv = 1.0f / d;
    return Mat2{{inv * m.cy.y, -inv * m.cx.y}, {-inv * m.cy.x, inv * m.cx.x}};
  }
  return Mat2{{0, 0}, {0, 0}};
}

NSL_HD Vec3 aabb_center(Aabb a) { return scale(0.5f, add(a.hi, a.lo)); }
NSL_HD Vec3 aabb_extents(Aabb a) { return scale(0.5f, sub(a.hi, a.lo)); }
NSL_HD Aabb aabb_inflate(Aabb a, float e) {
  Vec3 r = {e, e, e};
  return {sub(a.lo, r), add(a.hi, r)};
}
NSL_HD Aabb aabb_union(Aabb a, Aabb b) { return {min3(a.lo, b.lo), max3(a.hi, b.hi)}; }
NSL_HD bool aabb_overlaps(Aabb a, Aabb b) {
  if (a.hi.x < b.lo.x || a.lo.x > b.hi.x) return false;
  if (a.hi.y < b.lo.y || a.lo.y > b.hi.y) return false;
  if (a.hi.z < b.lo.z || a.lo.z > b.hi.z) return false;
  return true;
}
NSL_HD bool aabb_contains(Aabb a, Aabb b) {
  if (a.lo.x > b.lo.x || b.hi.x > a.hi.x) return false;
  if (a.lo.y > b.lo.y || b.hi.y > a.hi.y) return false;
  if (a.lo.z > b.lo.z || b.hi.z > a.hi.z) return false;
  return true;
}
NSL_HD Aabb aabb_transform(Xf t, Aabb a) {
  Vec3 c = xf_point(t, aabb_center(a));
  Mat3 m = matrix_from_quat(t.q);
  Vec3 e = mul_mv(abs_mat3(m), aabb_extents(a));
  return {sub(c, e), add(c, e)};
}
NSL_HD Plane plane_from_normal_point(Vec3 n, Vec3 p) { return Plane{n, dot(n, p)}; }
NSL_HD Plane transform_plane(Xf t, Plane p) {
  Vec3 n = rotate(t.q, p.n);
  return Plane{n, p.off + dot(n, t.p)};
}
NSL_HD float plane_sep(Plane p, Vec3 x) { return dot(p.n, x) - p.off; }
NSL_HD Vec3 arbitrary_perp(Vec3 v) {
  Vec3 p;
  if (v.x < -0.5f || v.x > 0.5f) {
    p = {0.67f * v.y + -0.42f * v.z, -0.67f * v.x, -(-0.42f) * v.x};
  } else if (v.y < -0.5f || v.y > 0.5f) {
    p = {0.67f * v.y, -0.67f * v.x + -0.42f * v.z, -(-0.42f) * v.y};
  } else {
    p = {0.67f * v.z, -0.42f * v.z, -0.67f * v.x - -0.42f * v.y};
  }
  return normalize(p);
}
NSL_HD SegDist line_distance(Vec3 p1, Vec3 d1, Vec3 p2, Vec3 d2) {
  SegDist r{};
  float a11 = dot(d1, d1), a12 = -dot(d1, d2), a21 = dot(d2, d1), a22 = -dot(d2, d2);
  Vec3 w = sub(p1, p2);
  float b1 = -dot(d1, w), b2 = -dot(d2, w);
  float d = a11 * a22 - a12 * a21;
  if (d * d < 1000.0f * kFltMin) {
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

// solver_integration.rs:72-87
struct Softness { float bias_rate, mass_scale, impulse_scale; };
NSL_HD Softness make_soft(float hertz, float zeta, float h) {
  if (hertz == 0.0f) return Softness{0, 0, 0};
  float omega = 2.0f * kPi * hertz;
  float a1 = 2.0f * zeta + h * omega;
  float a2 = h * omega * a1;
  float a3 = 1.0f / (1.0f + a2);
  return Softness{omega / a1, a2 * a3, a3};
}

}  // namespace nsl
