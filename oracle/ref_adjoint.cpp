// Oracle: reference gradients of the differentiable rollout model.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// The reference has no differentiable path. The GPU library differentiates a recorded
// batched rollout with the solver's active set frozen (the recorded contacts, their
// geometry and effective masses, the colour order and every clamp branch held fixed); this
// file is the CPU check of those gradients, built independently of the GPU code:
//
//  1. the oracle World steps the scene and, after each step's prepare, records a tape:
//     the prepared contact constraints in solve order (colour-major), body states and f64
//     centres at the step's start, the shapes of each contact and the link of each
//     constraint to the previous step's constraint of the same contact, with its points
//     matched by feature id (the warm-start carry-over, contact_update.rs);
//  2. the model runs forward from the tape with a generic scalar. In float it repeats the
//     oracle's stage arithmetic (ref_solver.cpp: integrate_velocities, warm_start_color,
//     push_color, integrate_positions, relax_color; contact_solver.rs:697-1097,
//     solver_integration.rs:263-426) operation for operation, so each step must end on the
//     recorded next state bit for bit; that pins the restatement, and the float run records
//     the branch every clamp took;
//  3. the model runs again over a reverse-mode automatic-differentiation scalar in double
//     precision with those branches frozen, and the tape of operations is swept backwards
//     for the gradient of the loss |com(T) - target|^2 with respect to every dynamic body's
//     initial linear and angular velocity and its shape's friction.
//
// Model (the same definition the GPU adjoint differentiates): per step, the recorded
// constraints fix normals, tangents, anchors, friction centres, effective masses, lever
// arms and inverse masses and inertias; the parameters act through body velocities and
// substep positions and rotations, each point's separation (recorded base separation plus
// n . (dc_b - dc_a), dc the body centre's offset from the recorded run: the recycle refresh
// of a separation, collide_pass.rs:641-662), the friction coefficient sqrt(f_a f_b)
// (contact_update.rs mixing), and the warm-start impulses carried from the previous step's
// final impulses by contact and feature id, friction re-projected onto the new tangents as
// store and prepare do (contact_solver.rs:1258-1299, :445-662). Steps with overflow
// constraints, restitution, rolling resistance, tangent velocity or gyroscopic bodies lie
// outside the model (status 1).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "ref_world.h"

namespace nslref {
namespace adj {

// ---------------------------------------------------------------------------------------
// reverse-mode automatic differentiation over double

struct Node {
  int32_t a, b;
  double da, db;
};
struct OpTape {
  std::vector<Node> nodes;
};
static thread_local OpTape* g_ops = nullptr;

struct Ad {
  double v = 0.0;
  int32_t id = -1;  // -1: a constant
  Ad() = default;
  Ad(double x) : v(x), id(-1) {}  // NOLINT: constants mix freely with variables
};
inline Ad make(double v, int32_t a, double da, int32_t b, double db) {
  Ad r(v);
  if (a < 0 && b < 0) return r;
  r.id = (int32_t)g_ops->nodes.size();
  g_ops->nodes.push_back(Node{a, b, da, db});
  return r;
}
inline Ad variable(double v) { return make(v, (int32_t)g_ops->nodes.size(), 0.0, -1, 0.0); }
inline Ad operator+(Ad x, Ad y) { return make(x.v + y.v, x.id, 1.0, y.id, 1.0); }
inline Ad operator-(Ad x, Ad y) { return make(x.v - y.v, x.id, 1.0, y.id, -1.0); }
inline Ad operator*(Ad x, Ad y) { return make(x.v * y.v, x.id, y.v, y.id, x.v); }
inline Ad operator/(Ad x, Ad y) {
  const double q = x.v / y.v;
  return make(q, x.id, 1.0 / y.v, y.id, -q / y.v);
}
inline Ad operator-(Ad x) { return make(-x.v, x.id, -1.0, -1, 0.0); }
inline Ad& operator+=(Ad& x, Ad y) { return x = x + y; }
inline bool operator>(Ad x, Ad y) { return x.v > y.v; }
inline bool operator<(Ad x, Ad y) { return x.v < y.v; }
inline Ad sqrt_(Ad x) {
  const double s = std::sqrt(x.v);
  return make(s, x.id, 0.5 / s, -1, 0.0);
}
inline float sqrt_(float x) { return std::sqrt(x); }
inline double sqrt_(double x) { return std::sqrt(x); }
inline double val(Ad a) { return a.v; }
inline double val(double a) { return a; }
inline double val(float a) { return a; }

// adjoints of every node for d out / d node
std::vector<double> sweep(const OpTape& t, int32_t out) {
  std::vector<double> g(t.nodes.size(), 0.0);
  if (out < 0) return g;
  g[out] = 1.0;
  for (int32_t i = out; i >= 0; --i) {
    const double gi = g[i];
    if (gi == 0.0) continue;
    const Node& n = t.nodes[i];
    if (n.a >= 0 && n.a != i) g[n.a] += gi * n.da;
    if (n.b >= 0) g[n.b] += gi * n.db;
  }
  return g;
}

// ---------------------------------------------------------------------------------------
// generic vector helpers, operand order as in ref_math.h

template <class T> struct V3 { T x, y, z; };
template <class T> struct Q4 { V3<T> v; T s; };
template <class T> V3<T> mk(Vec3 a) { return {T(a.x), T(a.y), T(a.z)}; }
template <class T> V3<T> zero3() { return {T(0.0f), T(0.0f), T(0.0f)}; }
template <class T> V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V3<T> scale(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> V3<T> mul_add(V3<T> a, T s, V3<T> b) { return {a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }
template <class T> V3<T> sym_mv(const SymMat3& m, V3<T> a) {
  return {T(m.cxx) * a.x + (T(m.cxy) * a.y + T(m.cxz) * a.z), T(m.cxy) * a.x + (T(m.cyy) * a.y + T(m.cyz) * a.z),
          T(m.cxz) * a.x + (T(m.cyz) * a.y + T(m.czz) * a.z)};
}
template <class T> T max_(T a, T b) { return a > b ? a : b; }
template <class T> T min_(T a, T b) { return a < b ? a : b; }
// inv_rotate_vector_wide (contact_solver.rs:291-300)
template <class T> V3<T> inv_rotate_wide(Q4<T> q, V3<T> a) {
  const V3<T> t0 = cross(q.v, a);
  const V3<T> t = {t0.x + t0.x, t0.y + t0.y, t0.z + t0.z};
  const V3<T> u = cross(q.v, t);
  return {(a.x - q.s * t.x) + u.x, (a.y - q.s * t.y) + u.y, (a.z - q.s * t.z) + u.z};
}
// integrate_rotation (internal_math.rs:322-325) with normalize_quat (functions.rs:450-458)
template <class T> Q4<T> integrate_rotation(Q4<T> q1, V3<T> dr) {
  const Q4<T> a = {scale(T(0.5f), dr), T(0.0f)};
  const V3<T> t1 = cross(a.v, q1.v);
  const V3<T> t2 = mul_add(t1, a.s, q1.v);
  const V3<T> t3 = mul_add(t2, q1.s, a.v);
  const Q4<T> qd = {t3, a.s * q1.s - dot(a.v, q1.v)};
  const Q4<T> q = {add(q1.v, qd.v), qd.s + q1.s};
  const T ls = q.v.x * q.v.x + q.v.y * q.v.y + q.v.z * q.v.z + q.s * q.s;
  if (ls > T(1000.0f * FLT_MIN)) {
    const T s = T(1.0f) / sqrt_(ls);
    return {{s * q.v.x, s * q.v.y, s * q.v.z}, s * q.s};
  }
  return {zero3<T>(), T(1.0f)};
}

// ---------------------------------------------------------------------------------------
// tape of a rollout (recorded by the oracle World)

struct Con {
  ContactConstraint c;
  ManifoldConstraint m;
  int32_t shape_a, shape_b;
  uint32_t features[4];
  int32_t link;       // the previous step's index of the same contact, or -1
  int8_t pmap[4];     // the matching previous point per point, or -1
};
struct Step {
  std::vector<Con> cons;               // solve order
  std::vector<int32_t> colour_end;     // end index of each colour in cons (colours in order)
  std::vector<BodyState> states;       // awake states at the step's start
  std::vector<Pos> centers;            // awake centres at the step's start
  bool ok = true;
};
struct Rollout {
  std::vector<Step> steps;
  std::vector<BodyState> final_states;
  std::vector<Pos> final_centers;
  std::vector<BodySim> sims;           // awake sims (masses, gravity scale, damping)
  std::vector<float> friction;         // per awake body: its first shape's friction
  std::vector<float> shape_friction;   // per shape id
  std::vector<int32_t> shape_body;     // per shape id: awake index of its body, or -1
  StepContext ctx;
  float contact_speed = 3.0f;
  Vec3 gravity{};
};

void record_step(World& w, Rollout& r) {
  Step s;
  s.states = w.awake_states;
  for (const BodySim& sim : w.awake_sims) s.centers.push_back(sim.center);
  s.ok = w.overflow_constraints.empty();
  for (const BodySim& sim : w.awake_sims)
    if (sim.flags & BF::ENABLE_GYROSCOPIC_TORQUE) s.ok = false;
  for (int ci = 0; ci < kGraphColorCount; ++ci) {
    for (const ContactConstraint& c : w.color_constraints[ci]) {
      Con k{};
      k.c = c;
      k.m = w.manifold_pool[c.manifold_start];
      const Contact& ct = w.contacts[c.contact_id];
      k.shape_a = ct.shape_id_a;
      k.shape_b = ct.shape_id_b;
      for (int i = 0; i < 4; ++i) k.features[i] = i < ct.manifold.point_count ? ct.manifold.points[i].feature_id : 0xffffffffu;
      if (c.manifold_count != 1 || c.rolling_resistance > 0.0f || c.restitution > 0.0f || k.m.tangent_velocity1 != 0.0f ||
          k.m.tangent_velocity2 != 0.0f)
        s.ok = false;
      k.link = -1;
      for (int i = 0; i < 4; ++i) k.pmap[i] = -1;
      s.cons.push_back(k);
    }
    if (!w.color_constraints[ci].empty()) s.colour_end.push_back((int32_t)s.cons.size());
  }
  // links to the previous step's constraint of the same contact (same id and shapes), points
  // matched by feature id
  if (!r.steps.empty()) {
    const Step& p = r.steps.back();
    for (Con& k : s.cons)
      for (int j = 0; j < (int)p.cons.size(); ++j) {
        const Con& q = p.cons[j];
        if (q.c.contact_id != k.c.contact_id || q.shape_a != k.shape_a || q.shape_b != k.shape_b) continue;
        k.link = j;
        uint32_t old[4];
        for (int i = 0; i < 4; ++i) old[i] = i < q.m.point_count ? q.features[i] : 0xffffffffu;
        for (int i = 0; i < k.m.point_count; ++i)
          for (int o = 0; o < q.m.point_count; ++o)
            if (old[o] == k.features[i]) {
              k.pmap[i] = (int8_t)o;
              old[o] = 0xffffffffu;
              break;
            }
        break;
      }
  }
  r.steps.push_back(std::move(s));
}

// ---------------------------------------------------------------------------------------
// the model

template <class T> struct Body {
  V3<T> v, w, dp;
  Q4<T> dq;
  uint32_t flags;
};
template <class T> struct Imp {
  T ni[4];
  T fx, fy, ti;
};
template <class T> Body<T> body_at(const std::vector<Body<T>>& s, int32_t i) {
  if (i >= 0) return s[i];
  Body<T> b{zero3<T>(), zero3<T>(), zero3<T>(), {zero3<T>(), T(1.0f)}, 0u};
  return b;
}
template <class T> void set_body(std::vector<Body<T>>& s, int32_t i, const Body<T>& b) {
  if (i >= 0 && (b.flags & BF::DYNAMIC_FLAG) != 0) s[i] = b;
}

// branch record per constraint and substep: push points 0..3 (bit0 separated, bit1 bias at
// -contact_speed, bit2 impulse at -ni), relax points 0..3 (bit0 s > 0, bit2 at -ni), twist
// (0 interior, 1 upper, 2 lower), friction cone (1 active)
constexpr int kBranchBytes = 10;

// warm_start_color (contact_solver.rs:697-763)
template <class T> void warm_start(const Con& k, const Imp<T>& im, std::vector<Body<T>>& s) {
  const ContactConstraint& c = k.c;
  const ManifoldConstraint& cm = k.m;
  V3<T> linear = zero3<T>(), ang_a = zero3<T>(), ang_b = zero3<T>();
  T total = T(0.0f);
  V3<T> moment_a = zero3<T>(), moment_b = zero3<T>();
  for (int p = 0; p < cm.point_count; ++p) {
    total += im.ni[p];
    moment_a = mul_add(moment_a, im.ni[p], mk<T>(cm.points[p].r_a));
    moment_b = mul_add(moment_b, im.ni[p], mk<T>(cm.points[p].r_b));
  }
  const V3<T> normal = mk<T>(cm.normal);
  V3<T> fi = scale(im.fx, mk<T>(cm.tangent1));
  fi = mul_add(fi, im.fy, mk<T>(cm.tangent2));
  const V3<T> lin = mul_add(fi, total, normal);
  const V3<T> twist = scale(im.ti, normal);
  const V3<T> a_a = add(add(cross(moment_a, normal), cross(mk<T>(cm.center_a), fi)), twist);
  const V3<T> a_b = add(add(cross(moment_b, normal), cross(mk<T>(cm.center_b), fi)), twist);
  linear = add(linear, lin);
  ang_a = add(ang_a, a_a);
  ang_b = add(ang_b, a_b);
  Body<T> ba = body_at(s, c.index_a), bb = body_at(s, c.index_b);
  ba.w = sub(ba.w, sym_mv(c.inv_ia, ang_a));
  ba.v = sub(ba.v, scale(T(c.inv_mass_a), linear));
  bb.w = add(bb.w, sym_mv(c.inv_ib, ang_b));
  bb.v = add(bb.v, scale(T(c.inv_mass_b), linear));
  set_body(s, c.index_a, ba);
  set_body(s, c.index_b, bb);
}

// push_color (contact_solver.rs:796-894); `bs` the points' base separations; `rec` records
// the branches (forced from `force` when given)
template <class T>
void push(const Con& k, const T* bs, Imp<T>& im, std::vector<Body<T>>& s, const StepContext& ctx, float contact_speed, uint8_t* rec,
          const uint8_t* force) {
  const ContactConstraint& c = k.c;
  const ManifoldConstraint& cm = k.m;
  const Softness& sf = c.is_static ? ctx.static_softness : ctx.contact_softness;
  const T inv_h = T(ctx.inv_h), speed_neg = T(-contact_speed);
  const T bias_rate = T(sf.mass_scale * sf.bias_rate), mass_scale = T(sf.mass_scale), impulse_scale = T(sf.impulse_scale);
  Body<T> ba = body_at(s, c.index_a), bb = body_at(s, c.index_b);
  const V3<T> dp = sub(bb.dp, ba.dp);
  const T inv_mass = T(c.inv_mass_a) + T(c.inv_mass_b);
  const V3<T> normal = mk<T>(cm.normal);
  const T normal_sep = dot(normal, dp);
  const V3<T> n_a = inv_rotate_wide(ba.dq, normal), n_b = inv_rotate_wide(bb.dq, normal);
  T lnv = dot(normal, sub(bb.v, ba.v));
  T manifold_impulse = T(0.0f);
  for (int p = 0; p < cm.point_count; ++p) {
    const V3<T> r_a = mk<T>(cm.points[p].r_a), r_b = mk<T>(cm.points[p].r_b);
    const T sep = (normal_sep + (dot(n_b, r_b) - dot(n_a, r_a))) + bs[p];
    const bool separated = force ? (force[p] & 1) != 0 : sep > T(0.0f);
    const T overlap = bias_rate * sep;
    const bool clamped = force ? (force[p] & 2) != 0 : !(overlap > speed_neg);
    const T velocity_bias = separated ? sep * inv_h : (force ? (clamped ? speed_neg : overlap) : max_(overlap, speed_neg));
    const T pm = separated ? T(1.0f) : mass_scale;
    const T pi = separated ? T(0.0f) : impulse_scale;
    const V3<T> u_a = cross(r_a, normal), u_b = cross(r_b, normal);
    const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
    const T v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
    const T ni = im.ni[p];
    const T neg_impulse = T(cm.points[p].normal_mass) * (pm * v_n + velocity_bias) + pi * ni;
    const bool at_floor = force ? (force[p] & 4) != 0 : !(-neg_impulse > -ni);
    const T impulse = force ? (at_floor ? -ni : -neg_impulse) : max_(-neg_impulse, -ni);
    im.ni[p] = ni + impulse;
    lnv += inv_mass * impulse;
    manifold_impulse += impulse;
    ba.w = sub(ba.w, scale(impulse, g_a));
    bb.w = add(bb.w, scale(impulse, g_b));
    if (rec) rec[p] = (uint8_t)((separated ? 1 : 0) | (!separated && clamped ? 2 : 0) | (at_floor ? 4 : 0));
  }
  const V3<T> pp = scale(manifold_impulse, normal);
  ba.v = sub(ba.v, scale(T(c.inv_mass_a), pp));
  bb.v = add(bb.v, scale(T(c.inv_mass_b), pp));
  set_body(s, c.index_a, ba);
  set_body(s, c.index_b, bb);
}

// relax_color (contact_solver.rs:927-1097) without rolling resistance; `mu` the friction
template <class T>
void relax(const Con& k, const T* bs, T mu, Imp<T>& im, std::vector<Body<T>>& s, const StepContext& ctx, uint8_t* rec,
           const uint8_t* force) {
  const ContactConstraint& c = k.c;
  const ManifoldConstraint& cm = k.m;
  const T inv_h = T(ctx.inv_h), eps = T(FLT_EPSILON);
  Body<T> ba = body_at(s, c.index_a), bb = body_at(s, c.index_b);
  const V3<T> dp = sub(bb.dp, ba.dp);
  const T inv_mass = T(c.inv_mass_a) + T(c.inv_mass_b);
  const V3<T> normal = mk<T>(cm.normal), t1 = mk<T>(cm.tangent1), t2 = mk<T>(cm.tangent2);
  const T normal_sep = dot(normal, dp);
  const V3<T> n_a = inv_rotate_wide(ba.dq, normal), n_b = inv_rotate_wide(bb.dq, normal);
  const V3<T> dv = sub(bb.v, ba.v);
  T lnv = dot(normal, dv);
  const T ltv1 = dot(t1, dv) - T(cm.tangent_velocity1);
  const T ltv2 = dot(t2, dv) - T(cm.tangent_velocity2);
  T manifold_impulse = T(0.0f), total_normal = T(0.0f), total_twist_limit = T(0.0f);
  for (int p = 0; p < cm.point_count; ++p) {
    const V3<T> r_a = mk<T>(cm.points[p].r_a), r_b = mk<T>(cm.points[p].r_b);
    const T sep = (normal_sep + (dot(n_b, r_b) - dot(n_a, r_a))) + bs[p];
    const bool positive = force ? (force[p] & 1) != 0 : T(0.0f) < sep * inv_h;
    const T velocity_bias = force ? (positive ? sep * inv_h : T(0.0f)) : max_(T(0.0f), sep * inv_h);
    const V3<T> u_a = cross(r_a, normal), u_b = cross(r_b, normal);
    const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
    const T v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
    const T ni = im.ni[p];
    const T neg_impulse = T(cm.points[p].normal_mass) * (v_n + velocity_bias);
    const bool at_floor = force ? (force[p] & 4) != 0 : !(-neg_impulse > -ni);
    const T impulse = force ? (at_floor ? -ni : -neg_impulse) : max_(-neg_impulse, -ni);
    const T new_impulse = ni + impulse;
    im.ni[p] = new_impulse;
    total_normal += new_impulse;
    total_twist_limit += T(cm.points[p].lever_arm) * new_impulse;
    lnv += inv_mass * impulse;
    manifold_impulse += impulse;
    ba.w = sub(ba.w, scale(impulse, g_a));
    bb.w = add(bb.w, scale(impulse, g_b));
    if (rec) rec[p] = (uint8_t)((positive ? 1 : 0) | (at_floor ? 4 : 0));
  }
  {
    const T twist_speed = dot(normal, sub(bb.w, ba.w));
    const T max_lambda = mu * total_twist_limit;
    const T old = im.ti;
    const T raw = -(T(cm.twist_mass) * twist_speed);
    const T hi = min_(raw, max_lambda - old);
    const uint8_t branch = !(-max_lambda - old < hi) ? 2 : (!(raw < max_lambda - old) ? 1 : 0);
    const uint8_t b = force ? force[4] : branch;
    const T impulse = b == 2 ? -max_lambda - old : (b == 1 ? max_lambda - old : raw);
    if (rec) rec[4] = branch;
    im.ti = old + impulse;
    ba.w = sub(ba.w, scale(impulse, sym_mv(c.inv_ia, normal)));
    bb.w = add(bb.w, scale(impulse, sym_mv(c.inv_ib, normal)));
  }
  {
    const V3<T> r_a = mk<T>(cm.center_a), r_b = mk<T>(cm.center_b);
    const V3<T> arv = sub(cross(bb.w, r_b), cross(ba.w, r_a));
    const T vt1 = ltv1 + dot(arv, t1), vt2 = ltv2 + dot(arv, t2);
    const Mat2& tm = cm.tangent_mass;
    T dix = T(tm.cx.x) * vt1 + T(tm.cy.x) * vt2;
    T diy = T(tm.cx.y) * vt1 + T(tm.cy.y) * vt2;
    dix = -dix;
    diy = -diy;
    const T fx_old = im.fx, fy_old = im.fy;
    T nix = fx_old + dix, niy = fy_old + diy;
    const T max_impulse = mu * total_normal;
    const T lsq = nix * nix + niy * niy;
    const bool cone = lsq > max_impulse * max_impulse;
    const bool mask = force ? force[5] != 0 : cone;
    const T nrm = max_impulse / (sqrt_(lsq) + eps);
    const T sc = mask ? nrm : T(1.0f);
    nix = sc * nix;
    niy = sc * niy;
    if (rec) rec[5] = cone ? 1 : 0;
    dix = nix - fx_old;
    diy = niy - fy_old;
    im.fx = nix;
    im.fy = niy;
    V3<T> pp = scale(dix, t1);
    pp = add(pp, scale(diy, t2));
    ba.w = sub(ba.w, sym_mv(c.inv_ia, cross(r_a, pp)));
    bb.w = add(bb.w, sym_mv(c.inv_ib, cross(r_b, pp)));
    pp = add(pp, scale(manifold_impulse, normal));
    ba.v = sub(ba.v, scale(T(c.inv_mass_a), pp));
    bb.v = add(bb.v, scale(T(c.inv_mass_b), pp));
  }
  set_body(s, c.index_a, ba);
  set_body(s, c.index_b, bb);
}

// integrate_velocities (solver_integration.rs:263-358) with no applied force or torque and
// no gyroscopic term
template <class T> void integrate_velocities(Body<T>& b, const BodySim& sim, float h, Vec3 g) {
  const float linear_damping = 1.0f / (1.0f + h * sim.linear_damping);
  const float angular_damping = 1.0f / (1.0f + h * sim.angular_damping);
  const float gravity_scale = sim.inv_mass > 0.0f ? sim.gravity_scale : 0.0f;
  const T hm = T(h * sim.inv_mass), hg = T(h * gravity_scale);
  const V3<T> dv = {hm * T(0.0f) + hg * T(g.x), hm * T(0.0f) + hg * T(g.y), hm * T(0.0f) + hg * T(g.z)};
  b.v = mul_add(dv, T(linear_damping), b.v);
  const V3<T> dw = {T(h) * T(0.0f), T(h) * T(0.0f), T(h) * T(0.0f)};
  b.w = mul_add(dw, T(angular_damping), b.w);
}
// integrate_positions (solver_integration.rs:377-426) below the speed caps
template <class T> void integrate_positions(Body<T>& b, float h) {
  b.dp = mul_add(b.dp, T(h), b.v);
  b.dq = integrate_rotation(b.dq, scale(T(h), b.w));
}

// The model forward over a rollout. Parameters: x[7 * i + {0..2}] linear velocity,
// {3..5} angular velocity, 6 friction of awake body i. Returns the loss; with `branches`
// empty it records every branch, else takes them from it. `exact` (float runs) reports
// whether every step ended on the recorded next state bit for bit.
template <class T>
T model(const Rollout& r, const std::vector<T>& x, const float* target, std::vector<uint8_t>& branches, bool record,
        bool* exact) {
  const size_t nb = r.sims.size();
  std::vector<Body<T>> S(nb);
  std::vector<double> dcx(nb, 0.0), dcy(nb, 0.0), dcz(nb, 0.0);  // float runs: offsets stay 0
  std::vector<T> DX(nb, T(0.0f)), DY(nb, T(0.0f)), DZ(nb, T(0.0f));
  for (size_t i = 0; i < nb; ++i) {
    const BodyState& st = r.steps[0].states[i];
    S[i] = Body<T>{{x[7 * i + 0], x[7 * i + 1], x[7 * i + 2]}, {x[7 * i + 3], x[7 * i + 4], x[7 * i + 5]}, zero3<T>(),
                   {zero3<T>(), T(1.0f)}, st.flags};
  }
  auto friction_of = [&](int32_t shape) -> T {
    const int32_t b = r.shape_body[shape];
    return b >= 0 ? x[7 * b + 6] : T(r.shape_friction[shape]);
  };
  std::vector<Imp<T>> F;  // final impulses of the previous step
  const int subs = r.ctx.sub_step_count;
  size_t br = 0;
  if (exact) *exact = true;
  for (size_t k = 0; k < r.steps.size(); ++k) {
    const Step& st = r.steps[k];
    const size_t n = st.cons.size();
    std::vector<Imp<T>> I(n);
    std::vector<T> BS(4 * n), MU(n);
    for (size_t j = 0; j < n; ++j) {
      const Con& kc = st.cons[j];
      const ManifoldConstraint& cm = kc.m;
      Imp<T>& im = I[j];
      for (int i = 0; i < 4; ++i) {
        T ni = T(i < cm.point_count ? cm.points[i].normal_impulse : 0.0f);
        if (kc.link >= 0 && i < cm.point_count && kc.pmap[i] >= 0) ni = F[kc.link].ni[kc.pmap[i]];
        im.ni[i] = ni;
      }
      if (kc.link >= 0) {
        const Con& q = r.steps[k - 1].cons[kc.link];
        const V3<T> fv = add(scale(F[kc.link].fx, mk<T>(q.m.tangent1)), scale(F[kc.link].fy, mk<T>(q.m.tangent2)));
        im.fx = dot(fv, mk<T>(cm.tangent1));
        im.fy = dot(fv, mk<T>(cm.tangent2));
        im.ti = F[kc.link].ti;
      } else {
        im.fx = T(cm.friction_impulse.x);
        im.fy = T(cm.friction_impulse.y);
        im.ti = T(cm.twist_impulse);
      }
      const V3<T> nn = mk<T>(cm.normal);
      V3<T> dca = zero3<T>(), dcb = zero3<T>();
      if (kc.c.index_a >= 0) dca = {DX[kc.c.index_a], DY[kc.c.index_a], DZ[kc.c.index_a]};
      if (kc.c.index_b >= 0) dcb = {DX[kc.c.index_b], DY[kc.c.index_b], DZ[kc.c.index_b]};
      const T shift = dot(nn, sub(dcb, dca));
      for (int i = 0; i < 4; ++i) {
        const T base = T(i < cm.point_count ? cm.points[i].base_separation : 0.0f);
        // float runs repeat the recorded step: the offsets are zero there by definition
        BS[4 * j + i] = std::is_same<T, float>::value ? base : base + shift;
      }
      MU[j] = sqrt_(friction_of(kc.shape_a) * friction_of(kc.shape_b));
    }
    for (size_t i = 0; i < nb; ++i) {
      S[i].dp = zero3<T>();
      S[i].dq = {zero3<T>(), T(1.0f)};
    }
    const size_t colours = st.colour_end.size();
    for (int sub_ = 0; sub_ < subs; ++sub_) {
      if (record) branches.resize(br + n * kBranchBytes, 0);
      uint8_t* B = branches.data() + br;
      for (size_t i = 0; i < nb; ++i) integrate_velocities(S[i], r.sims[i], r.ctx.h, r.gravity);
      for (size_t a = 0, b0 = 0; a < colours; b0 = st.colour_end[a++])
        for (size_t j = b0; j < (size_t)st.colour_end[a]; ++j) warm_start(st.cons[j], I[j], S);
      for (size_t a = 0, b0 = 0; a < colours; b0 = st.colour_end[a++])
        for (size_t j = b0; j < (size_t)st.colour_end[a]; ++j)
          push(st.cons[j], &BS[4 * j], I[j], S, r.ctx, r.contact_speed, record ? B + j * kBranchBytes : nullptr,
               record ? nullptr : B + j * kBranchBytes);
      for (size_t i = 0; i < nb; ++i) integrate_positions(S[i], r.ctx.h);
      for (size_t a = 0, b0 = 0; a < colours; b0 = st.colour_end[a++])
        for (size_t j = b0; j < (size_t)st.colour_end[a]; ++j)
          relax(st.cons[j], &BS[4 * j], MU[j], I[j], S, r.ctx, record ? B + j * kBranchBytes + 4 : nullptr,
                record ? nullptr : B + j * kBranchBytes + 4);
      br += n * kBranchBytes;
    }
    F = I;
    // the body centre's offset from the recording: centre_{k+1} = centre_k + dp
    const std::vector<Pos>& c0 = st.centers;
    const std::vector<Pos>& c1 = k + 1 < r.steps.size() ? r.steps[k + 1].centers : r.final_centers;
    const std::vector<BodyState>& next = k + 1 < r.steps.size() ? r.steps[k + 1].states : r.final_states;
    for (size_t i = 0; i < nb; ++i) {
      DX[i] += S[i].dp.x - T(c1[i].x - c0[i].x);
      DY[i] += S[i].dp.y - T(c1[i].y - c0[i].y);
      DZ[i] += S[i].dp.z - T(c1[i].z - c0[i].z);
      if (exact) {
        const float got[6] = {(float)val(S[i].v.x), (float)val(S[i].v.y), (float)val(S[i].v.z), (float)val(S[i].w.x),
                              (float)val(S[i].w.y), (float)val(S[i].w.z)};
        const float want[6] = {next[i].v.x, next[i].v.y, next[i].v.z, next[i].w.x, next[i].w.y, next[i].w.z};
        if (std::memcmp(got, want, sizeof got) != 0) *exact = false;
      }
    }
  }
  // loss at the final centres
  double m_sum = 0.0;
  T mx = T(0.0), my = T(0.0), mz = T(0.0);
  for (size_t i = 0; i < nb; ++i) {
    const double m = r.sims[i].inv_mass > 0.0f ? 1.0 / (double)r.sims[i].inv_mass : 0.0;
    m_sum += m;
    const Pos c = r.final_centers[i];
    mx += T(m) * (T(c.x) + DX[i]);
    my += T(m) * (T(c.y) + DY[i]);
    mz += T(m) * (T(c.z) + DZ[i]);
  }
  const T ex = mx / T(m_sum) - T((double)target[0]);
  const T ey = my / T(m_sum) - T((double)target[1]);
  const T ez = mz / T(m_sum) - T((double)target[2]);
  return ex * ex + ey * ey + ez * ez;
}

}  // namespace adj
}  // namespace nslref

// ---------------------------------------------------------------------------------------
// C entry points (ctypes)

using namespace nslref;
using namespace nslref::adj;

extern "C" {

// Steps `world` (built from the scene, not yet stepped) `steps` steps, recording the tape;
// returns a rollout handle (free with nslo_rollout_destroy).
void* nslo_rollout_record(void* wp, float dt, int sub_steps, int steps) {
  World* w = static_cast<World*>(wp);
  Rollout* r = new Rollout();
  w->after_prepare = [r](World& ww, const StepContext& ctx) {
    r->ctx = ctx;
    record_step(ww, *r);
  };
  for (int k = 0; k < steps; ++k) w->step(dt, sub_steps);
  w->after_prepare = nullptr;
  r->final_states = w->awake_states;
  for (const BodySim& sim : w->awake_sims) r->final_centers.push_back(sim.center);
  r->sims = w->awake_sims;
  r->gravity = w->gravity;
  r->contact_speed = w->contact_speed;
  r->shape_friction.assign(w->shapes.size(), 0.0f);
  r->shape_body.assign(w->shapes.size(), -1);
  r->friction.assign(w->awake_sims.size(), 0.0f);
  std::vector<bool> seen(w->awake_sims.size(), false);
  for (size_t s = 0; s < w->shapes.size(); ++s) {
    const Shape& sh = w->shapes[s];
    if (sh.id < 0) continue;
    r->shape_friction[s] = sh.friction;
    const BodySim& sim0 = w->body_sim(sh.body_id);
    for (size_t i = 0; i < w->awake_sims.size(); ++i)
      if (w->awake_sims[i].body_id == sim0.body_id && (w->awake_states[i].flags & BF::DYNAMIC_FLAG)) {
        r->shape_body[s] = (int32_t)i;
        if (!seen[i]) r->friction[i] = sh.friction;
        seen[i] = true;
      }
  }
  return r;
}

void nslo_rollout_destroy(void* r) { delete static_cast<Rollout*>(r); }

int nslo_rollout_bodies(void* rp) { return (int)static_cast<Rollout*>(rp)->sims.size(); }

// 0: every recorded step lies inside the model; 1: some step does not
int nslo_rollout_status(void* rp) {
  for (const Step& s : static_cast<Rollout*>(rp)->steps)
    if (!s.ok) return 1;
  return 0;
}

// The recorded parameters: x[7 i + 0..2] = v, 3..5 = w, 6 = friction of awake body i.
static std::vector<double> recorded_params(const Rollout& r) {
  std::vector<double> x(7 * r.sims.size());
  for (size_t i = 0; i < r.sims.size(); ++i) {
    const BodyState& s = r.steps[0].states[i];
    const float p[7] = {s.v.x, s.v.y, s.v.z, s.w.x, s.w.y, s.w.z, r.friction[i]};
    for (int j = 0; j < 7; ++j) x[7 * i + j] = (double)p[j];
  }
  return x;
}

// Runs the model in float at the recorded parameters (records the branch of every clamp;
// `exact` = every step ends on the recorded next state bit for bit), then in double over the
// differentiation scalar with those branches frozen: the loss and its gradient with respect
// to every awake body's initial velocity (grad_v, grad_w: [bodies][3]) and friction.
int nslo_rollout_gradient(void* rp, const float* target, double* loss, double* grad_v, double* grad_w, double* grad_f,
                          int* exact) {
  Rollout& r = *static_cast<Rollout*>(rp);
  if (r.steps.empty()) return 1;
  const std::vector<double> x0 = recorded_params(r);
  std::vector<float> xf(x0.begin(), x0.end());
  std::vector<uint8_t> branches;
  bool ex = false;
  model<float>(r, xf, target, branches, true, &ex);
  *exact = ex ? 1 : 0;
  OpTape ops;
  ops.nodes.reserve(1u << 22);
  g_ops = &ops;
  std::vector<Ad> x(x0.size());
  for (size_t i = 0; i < x0.size(); ++i) x[i] = variable(x0[i]);
  const Ad l = model<Ad>(r, x, target, branches, false, nullptr);
  const std::vector<double> g = sweep(ops, l.id);
  g_ops = nullptr;
  *loss = l.v;
  for (size_t i = 0; i < r.sims.size(); ++i) {
    for (int j = 0; j < 3; ++j) {
      grad_v[3 * i + j] = g[x[7 * i + j].id];
      grad_w[3 * i + j] = g[x[7 * i + 3 + j].id];
    }
    grad_f[i] = g[x[7 * i + 6].id];
  }
  return nslo_rollout_status(rp);
}

// The model's loss in double precision with parameter `param` (7 * body + component) moved by
// `delta` from its recorded value; `frozen` takes the branches from the float run at the
// recorded parameters (the model the gradient differentiates), else the model decides its
// own branches.
double nslo_rollout_loss(void* rp, const float* target, int param, double delta, int frozen) {
  Rollout& r = *static_cast<Rollout*>(rp);
  std::vector<double> x = recorded_params(r);
  std::vector<uint8_t> branches;
  if (frozen) {
    std::vector<float> xf(x.begin(), x.end());
    model<float>(r, xf, target, branches, true, nullptr);
  }
  if (param >= 0 && param < (int)x.size()) x[param] += delta;
  std::vector<uint8_t> own;
  return frozen ? model<double>(r, x, target, branches, false, nullptr) : model<double>(r, x, target, own, true, nullptr);
}

}  // extern "C"
