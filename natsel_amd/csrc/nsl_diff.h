// Differentiable model of a recorded batched rollout: the contact solver of each step with
// its active set frozen, templated on the scalar type (float: the adjoint's recompute,
// bit-identical to the step kernels' arithmetic; double: the finite-difference check).
//
// Per step k the model takes from the tape (the forward run's prepared constraints):
// normal, tangents, anchors, friction centres, effective masses, lever arms, inverse
// masses and inertias, the colour order, and the branch every clamp took. Functions of
// the parameters are: body velocities, sub-step positions and rotations, the body
// centres' offsets from the recorded run (the separation of each point follows them:
// separation = recorded + n . (dc_b - dc_a), the reference's recycle refresh,
// collide_pass.rs:641-662), the friction coefficient sqrt(f_a f_b) (contact_update.rs
// mixing), and the warm-start impulses, carried from the previous step's final impulses
// by feature id (contact_update.rs impulse carry-over) with friction re-projected onto the
// new tangents (contact_solver.rs:1258-1299 store, :445-662 prepare).
//
// Stage arithmetic follows the step kernels (nsl_stages.h), which follow:
//   integrate_velocities   solver_integration.rs:263-358
//   integrate_positions    solver_integration.rs:377-426, internal_math.rs:322-325
//   warm start             contact_solver.rs:697-763
//   push (biased solve)    contact_solver.rs:796-894
//   relax                  contact_solver.rs:927-1097
#pragma once
#include "nsl_types.h"

namespace nsl {
namespace diff {

template <class T> struct V3 { T x, y, z; };
template <class T> struct Q4 { V3<T> v; T s; };

template <class T> __host__ __device__ __forceinline__ V3<T> mk(Vec3 a) { return {(T)a.x, (T)a.y, (T)a.z}; }
template <class T> __host__ __device__ __forceinline__ Q4<T> mkq(Quat q) { return {mk<T>(q.v), (T)q.s}; }
template <class T> __host__ __device__ __forceinline__ V3<T> zero() { return {(T)0, (T)0, (T)0}; }
template <class T> __host__ __device__ __forceinline__ V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> scale(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __host__ __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __host__ __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> __host__ __device__ __forceinline__ V3<T> mul_add(V3<T> a, T s, V3<T> b) {
  return {a.x + s * b.x, a.y + s * b.y, a.z + s * b.z};
}
template <class T> __host__ __device__ __forceinline__ V3<T> sym_mv(const SymMat3& m, V3<T> a) {
  return {(T)m.cxx * a.x + ((T)m.cxy * a.y + (T)m.cxz * a.z), (T)m.cxy * a.x + ((T)m.cyy * a.y + (T)m.cyz * a.z),
          (T)m.cxz * a.x + ((T)m.cyz * a.y + (T)m.czz * a.z)};
}
template <class T> __host__ __device__ __forceinline__ T maxT(T a, T b) { return a > b ? a : b; }
template <class T> __host__ __device__ __forceinline__ T minT(T a, T b) { return a < b ? a : b; }
template <class T> __host__ __device__ __forceinline__ T sqrtT(T x) { return sqrt(x); }
template <> __host__ __device__ __forceinline__ float sqrtT<float>(float x) { return sqrtf_(x); }

// contact_solver.rs:291-300
template <class T> __host__ __device__ __forceinline__ V3<T> inv_rotate_wide(Q4<T> q, V3<T> a) {
  const V3<T> t0 = cross(q.v, a);
  const V3<T> t = {t0.x + t0.x, t0.y + t0.y, t0.z + t0.z};
  const V3<T> u = cross(q.v, t);
  return {(a.x - q.s * t.x) + u.x, (a.y - q.s * t.y) + u.y, (a.z - q.s * t.z) + u.z};
}
// adjoint of inv_rotate_wide w.r.t. q, accumulated into qb
template <class T> __host__ __device__ __forceinline__ void inv_rotate_wide_adj(Q4<T> q, V3<T> a, V3<T> ob, Q4<T>& qb) {
  const V3<T> t0 = cross(q.v, a);
  const V3<T> t = {t0.x + t0.x, t0.y + t0.y, t0.z + t0.z};
  // out = (a - s t) + (v x t)
  qb.s -= dot(ob, t);
  V3<T> tb = sub(cross(ob, q.v), scale(q.s, ob));  // from -s t and from v x t (tb = ob x v)
  qb.v = add(qb.v, cross(t, ob));                   // v x t: vb += t x ob
  const V3<T> t0b = add(tb, tb);
  qb.v = add(qb.v, cross(a, t0b));                  // t0 = v x a: vb += a x t0b
}

// internal_math.rs:322-325 with normalize_quat (functions.rs:450-458)
template <class T> __host__ __device__ __forceinline__ Q4<T> integrate_rotation(Q4<T> q1, V3<T> dr) {
  const V3<T> hd = scale((T)0.5, dr);
  const Q4<T> a = {hd, (T)0};
  // mul_quat(a, q1)
  const V3<T> t1 = cross(a.v, q1.v);
  const V3<T> t2 = mul_add(t1, a.s, q1.v);
  const V3<T> t3 = mul_add(t2, q1.s, a.v);
  const Q4<T> qd = {t3, a.s * q1.s - dot(a.v, q1.v)};
  const Q4<T> q2 = {add(q1.v, qd.v), qd.s + q1.s};
  const T ls = dot(q2.v, q2.v) + q2.s * q2.s;
  const T lim = (T)(1000.0f * kFltMin);
  if (ls > lim) {
    const T s = (T)1 / sqrtT(ls);
    return {scale(s, q2.v), s * q2.s};
  }
  return {zero<T>(), (T)1};
}
template <class T>
__host__ __device__ __forceinline__ void integrate_rotation_adj(Q4<T> q1, V3<T> dr, Q4<T> ob, Q4<T>& q1b, V3<T>& drb) {
  const V3<T> hd = scale((T)0.5, dr);
  const V3<T> t1 = cross(hd, q1.v);
  const V3<T> t3 = add(add(t1, scale((T)0, q1.v)), scale(q1.s, hd));
  const T qds = (T)0 * q1.s - dot(hd, q1.v);
  const Q4<T> q2 = {add(q1.v, t3), qds + q1.s};
  const T ls = dot(q2.v, q2.v) + q2.s * q2.s;
  if (!(ls > (T)(1000.0f * kFltMin))) return;
  const T s = (T)1 / sqrtT(ls);
  // out = s q2 with s = ls^-1/2
  const T sb = dot(ob.v, q2.v) + ob.s * q2.s;
  const T lsb = sb * ((T)-0.5 * s * s * s);
  Q4<T> q2b = {add(scale(s, ob.v), scale(lsb + lsb, q2.v)), s * ob.s + (lsb + lsb) * q2.s};
  // q2.v = q1.v + t3, q2.s = qds + q1.s
  q1b.v = add(q1b.v, q2b.v);
  q1b.s += q2b.s;
  // t3 = hd x q1.v + 0 q1.v + q1.s hd
  const V3<T> t3b = q2b.v;
  V3<T> hdb = add(cross(q1.v, t3b), scale(q1.s, t3b));  // hd x v: hdb += v x t3b
  q1b.v = add(q1b.v, cross(t3b, hd));                    // hd x v: vb += t3b x hd ... (y = hd x v -> vb = ... )
  q1b.s += dot(t3b, hd);
  // qds = 0 s - dot(hd, q1.v)
  hdb = sub(hdb, scale(q2b.s, q1.v));
  q1b.v = sub(q1b.v, scale(q2b.s, hd));
  drb = add(drb, scale((T)0.5, hdb));
}

// One body in the model: velocities and the step's accumulated position and rotation deltas.
template <class T> struct Body {
  V3<T> v, w, dp;
  Q4<T> dq;
  uint32_t flags;
};

// One constraint's impulses (the solver's mutable state).
template <class T> struct Impulses {
  T ni[4];
  T fx, fy, ti;
};

// Per-constraint parameters of the model that are functions of the optimisation
// variables: base separations (through the centre offsets) and friction.
template <class T> struct ConParams {
  T bs[4];
  T mu;
};

template <class T> __host__ __device__ __forceinline__ Body<T> load_body(const Body<T>* s, int32_t i) {
  if (i >= 0) return s[i];
  Body<T> b;
  b.v = b.w = b.dp = zero<T>();
  b.dq = {zero<T>(), (T)1};
  b.flags = 0;
  return b;
}
template <class T> __host__ __device__ __forceinline__ void store_body(Body<T>* s, int32_t i, const Body<T>& b) {
  if (i >= 0 && (b.flags & bflag::kDynamic)) s[i] = b;
}

// ---------------------------------------------------------------------------------------
// forward stages

template <class T> __host__ __device__ __forceinline__ void warm_start(const Constraint& c, const Impulses<T>& im, Body<T>* s) {
  Body<T> ba = load_body(s, c.index_a), bb = load_body(s, c.index_b);
  V3<T> linear = zero<T>(), ang_a = zero<T>(), ang_b = zero<T>();
  if (c.manifold_count > 0) {
    T total = (T)0;
    V3<T> moment_a = zero<T>(), moment_b = zero<T>();
    for (int i = 0; i < c.point_count; ++i) {
      const T ni = im.ni[i];
      total += ni;
      moment_a = mul_add(moment_a, ni, mk<T>(c.points[i].r_a));
      moment_b = mul_add(moment_b, ni, mk<T>(c.points[i].r_b));
    }
    const V3<T> n = mk<T>(c.normal), t1 = mk<T>(c.tangent1), t2 = mk<T>(c.tangent2);
    V3<T> fi = scale(im.fx, t1);
    fi = mul_add(fi, im.fy, t2);
    const V3<T> lin = mul_add(fi, total, n);
    const V3<T> twist = scale(im.ti, n);
    const V3<T> a_a = add(add(cross(moment_a, n), cross(mk<T>(c.center_a), fi)), twist);
    const V3<T> a_b = add(add(cross(moment_b, n), cross(mk<T>(c.center_b), fi)), twist);
    linear = add(linear, lin);
    ang_a = add(ang_a, a_a);
    ang_b = add(ang_b, a_b);
  }
  ba.w = sub(ba.w, sym_mv(c.inv_ia, ang_a));
  ba.v = sub(ba.v, scale((T)c.inv_mass_a, linear));
  bb.w = add(bb.w, sym_mv(c.inv_ib, ang_b));
  bb.v = add(bb.v, scale((T)c.inv_mass_b, linear));
  store_body(s, c.index_a, ba);
  store_body(s, c.index_b, bb);
}

// Softness rows of push for a constraint (static or dynamic contact softness).
struct PushRows {
  float bias_rate, mass_scale, impulse_scale, inv_h, speed_neg;
};

// Forward push of one constraint; `rec` (optional) receives the branch record per point:
// bit0 separated, bit1 bias clamped at -contact_speed, bit2 impulse clamped at -ni. With
// `force` the branches are taken from that record instead of the comparisons (the model
// with its active set frozen).
template <class T>
__host__ __device__ __forceinline__ void push(const Constraint& c, const ConParams<T>& cp, Impulses<T>& im, Body<T>* s, const PushRows& r,
                                              uint8_t* rec, const uint8_t* force = nullptr) {
  Body<T> ba = load_body(s, c.index_a), bb = load_body(s, c.index_b);
  const T inv_h = (T)r.inv_h, speed_neg = (T)r.speed_neg, bias_rate = (T)r.bias_rate;
  const V3<T> dp = sub(bb.dp, ba.dp);
  const T inv_mass = (T)c.inv_mass_a + (T)c.inv_mass_b;
  if (c.manifold_count > 0) {
    const V3<T> n = mk<T>(c.normal);
    const T normal_sep = dot(n, dp);
    const V3<T> n_a = inv_rotate_wide(ba.dq, n), n_b = inv_rotate_wide(bb.dq, n);
    T lnv = dot(n, sub(bb.v, ba.v));
    T manifold_impulse = (T)0;
    for (int i = 0; i < c.point_count; ++i) {
      const V3<T> r_a = mk<T>(c.points[i].r_a), r_b = mk<T>(c.points[i].r_b);
      const T sep = (normal_sep + (dot(n_b, r_b) - dot(n_a, r_a))) + cp.bs[i];
      const bool separated = force ? (force[i] & 1) != 0 : sep > (T)0;
      const T soft = bias_rate * sep;
      const bool clamped = force ? (force[i] & 2) != 0 : (!separated && !(soft > speed_neg));
      const T velocity_bias = separated ? sep * inv_h : (force ? (clamped ? speed_neg : soft) : maxT(soft, speed_neg));
      const T pm = separated ? (T)1 : (T)r.mass_scale;
      const T pi = separated ? (T)0 : (T)r.impulse_scale;
      const V3<T> u_a = cross(r_a, n), u_b = cross(r_b, n);
      const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
      const T v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
      const T ni = im.ni[i];
      const T neg_impulse = (T)c.points[i].normal_mass * (pm * v_n + velocity_bias) + pi * ni;
      const bool at_floor = force ? (force[i] & 4) != 0 : !(-neg_impulse > -ni);
      const T impulse = force ? (at_floor ? -ni : -neg_impulse) : maxT(-neg_impulse, -ni);
      im.ni[i] = ni + impulse;
      lnv += inv_mass * impulse;
      manifold_impulse += impulse;
      ba.w = sub(ba.w, scale(impulse, g_a));
      bb.w = add(bb.w, scale(impulse, g_b));
      if (rec) rec[i] = (uint8_t)((separated ? 1 : 0) | (clamped ? 2 : 0) | (at_floor ? 4 : 0));
    }
    const V3<T> pp = scale(manifold_impulse, n);
    ba.v = sub(ba.v, scale((T)c.inv_mass_a, pp));
    bb.v = add(bb.v, scale((T)c.inv_mass_b, pp));
  }
  store_body(s, c.index_a, ba);
  store_body(s, c.index_b, bb);
}

// Forward relax; `rec` per point: bit0 s > 0, bit2 impulse clamped at -ni; rec[4]: twist
// branch (0 interior, 1 upper, 2 lower); rec[5]: friction cone active. `force` as in push.
template <class T>
__host__ __device__ __forceinline__ void relax(const Constraint& c, const ConParams<T>& cp, Impulses<T>& im, Body<T>* s, float inv_h_f,
                                               uint8_t* rec, const uint8_t* force = nullptr) {
  Body<T> ba = load_body(s, c.index_a), bb = load_body(s, c.index_b);
  const T inv_h = (T)inv_h_f, eps = (T)kFltEps;
  const V3<T> dp = sub(bb.dp, ba.dp);
  const T inv_mass = (T)c.inv_mass_a + (T)c.inv_mass_b;
  if (c.manifold_count > 0) {
    const V3<T> n = mk<T>(c.normal), t1 = mk<T>(c.tangent1), t2 = mk<T>(c.tangent2);
    const T normal_sep = dot(n, dp);
    const V3<T> n_a = inv_rotate_wide(ba.dq, n), n_b = inv_rotate_wide(bb.dq, n);
    const V3<T> dv = sub(bb.v, ba.v);
    T lnv = dot(n, dv);
    const T ltv1 = dot(t1, dv) - (T)c.tangent_velocity1;
    const T ltv2 = dot(t2, dv) - (T)c.tangent_velocity2;
    T manifold_impulse = (T)0, total_normal = (T)0, total_twist_limit = (T)0;
    for (int i = 0; i < c.point_count; ++i) {
      const V3<T> r_a = mk<T>(c.points[i].r_a), r_b = mk<T>(c.points[i].r_b);
      const T sep = (normal_sep + (dot(n_b, r_b) - dot(n_a, r_a))) + cp.bs[i];
      const T velocity_bias = force ? ((force[i] & 1) ? sep * inv_h : (T)0) : maxT((T)0, sep * inv_h);
      const V3<T> u_a = cross(r_a, n), u_b = cross(r_b, n);
      const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
      const T v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
      const T ni = im.ni[i];
      const T neg_impulse = (T)c.points[i].normal_mass * (v_n + velocity_bias);
      const bool at_floor = force ? (force[i] & 4) != 0 : !(-neg_impulse > -ni);
      const T impulse = force ? (at_floor ? -ni : -neg_impulse) : maxT(-neg_impulse, -ni);
      const T new_impulse = ni + impulse;
      im.ni[i] = new_impulse;
      total_normal += new_impulse;
      total_twist_limit += (T)c.points[i].lever_arm * new_impulse;
      lnv += inv_mass * impulse;
      manifold_impulse += impulse;
      ba.w = sub(ba.w, scale(impulse, g_a));
      bb.w = add(bb.w, scale(impulse, g_b));
      if (rec) rec[i] = (uint8_t)(((T)0 < sep * inv_h ? 1 : 0) | (at_floor ? 4 : 0));
    }
    {
      const T twist_speed = dot(n, sub(bb.w, ba.w));
      const T max_lambda = cp.mu * total_twist_limit;
      const T old = im.ti;
      const T raw = -((T)c.twist_mass * twist_speed);
      const T hi = minT(raw, max_lambda - old);
      T impulse = maxT(-max_lambda - old, hi);
      if (force) impulse = force[4] == 2 ? -max_lambda - old : (force[4] == 1 ? max_lambda - old : raw);
      if (rec) rec[4] = (uint8_t)(!(-max_lambda - old < hi) ? 2 : (!(raw < max_lambda - old) ? 1 : 0));
      im.ti = old + impulse;
      ba.w = sub(ba.w, scale(impulse, sym_mv(c.inv_ia, n)));
      bb.w = add(bb.w, scale(impulse, sym_mv(c.inv_ib, n)));
    }
    {
      const V3<T> r_a = mk<T>(c.center_a), r_b = mk<T>(c.center_b);
      const V3<T> arv = sub(cross(bb.w, r_b), cross(ba.w, r_a));
      const T vt1 = ltv1 + dot(arv, t1), vt2 = ltv2 + dot(arv, t2);
      const Mat2& tm = c.tangent_mass;
      T dix = (T)tm.cx.x * vt1 + (T)tm.cy.x * vt2;
      T diy = (T)tm.cx.y * vt1 + (T)tm.cy.y * vt2;
      dix = -dix;
      diy = -diy;
      const T fx_old = im.fx, fy_old = im.fy;
      T nix = fx_old + dix, niy = fy_old + diy;
      const T max_impulse = cp.mu * total_normal;
      const T lsq = nix * nix + niy * niy;
      const bool mask = force ? force[5] != 0 : lsq > max_impulse * max_impulse;
      const T nrm = max_impulse / (sqrtT(lsq) + eps);
      const T sc = mask ? nrm : (T)1;
      nix = sc * nix;
      niy = sc * niy;
      if (rec) rec[5] = mask ? 1 : 0;
      dix = nix - fx_old;
      diy = niy - fy_old;
      im.fx = nix;
      im.fy = niy;
      V3<T> pp = scale(dix, t1);
      pp = add(pp, scale(diy, t2));
      ba.w = sub(ba.w, sym_mv(c.inv_ia, cross(r_a, pp)));
      bb.w = add(bb.w, sym_mv(c.inv_ib, cross(r_b, pp)));
      pp = add(pp, scale(manifold_impulse, n));
      ba.v = sub(ba.v, scale((T)c.inv_mass_a, pp));
      bb.v = add(bb.v, scale((T)c.inv_mass_b, pp));
    }
  }
  store_body(s, c.index_a, ba);
  store_body(s, c.index_b, bb);
}

// integrate_velocities for a body with no applied force or torque and no gyroscopic term:
// v += dv * damping with dv = h * gravity_scale * gravity.
template <class T>
__host__ __device__ __forceinline__ void integrate_velocities(Body<T>& b, float h, float inv_mass, float gravity_scale, float lin_damp,
                                                              float ang_damp, Vec3 g) {
  const float ld = 1.0f / (1.0f + h * lin_damp);
  const float ad = 1.0f / (1.0f + h * ang_damp);
  const float gs = inv_mass > 0.0f ? gravity_scale : 0.0f;
  const float hm = h * inv_mass, hg = h * gs;
  const V3<T> dv = {(T)hm * (T)0 + (T)hg * (T)g.x, (T)hm * (T)0 + (T)hg * (T)g.y, (T)hm * (T)0 + (T)hg * (T)g.z};
  b.v = mul_add(dv, (T)ld, b.v);
  const V3<T> dw = {(T)h * (T)0, (T)h * (T)0, (T)h * (T)0};
  b.w = mul_add(dw, (T)ad, b.w);
}

template <class T> __host__ __device__ __forceinline__ void integrate_positions(Body<T>& b, float h) {
  b.dp = mul_add(b.dp, (T)h, b.v);
  b.dq = integrate_rotation(b.dq, scale((T)h, b.w));
}

// ---------------------------------------------------------------------------------------
// adjoints (reverse mode). Each takes the stage's inputs as the forward saw them, recomputes
// the branch record, and turns the adjoints of the outputs (in place in `sb`, `ib`) into
// adjoints of the inputs; parameter adjoints (base separations, friction) accumulate in
// `cb`. Only dynamic bodies carry adjoints: a non-dynamic body's state is never written by
// a constraint and does not depend on the parameters.

template <class T> struct BodyAdj {
  V3<T> v, w, dp;
  Q4<T> dq;
};
template <class T> struct ConParamsAdj {
  T bs[4];
  T mu;
};

template <class T> __host__ __device__ __forceinline__ bool adj_dynamic(const Body<T>* s, int32_t i) {
  return i >= 0 && (s[i].flags & bflag::kDynamic) != 0;
}

// Warm start: outputs v, w of both bodies; impulses are read only.
template <class T>
__host__ __device__ __forceinline__ void warm_start_adj(const Constraint& c, const Body<T>* s, BodyAdj<T>* sb, Impulses<T>& ib) {
  if (c.manifold_count <= 0) return;
  const bool da = adj_dynamic(s, c.index_a), db = adj_dynamic(s, c.index_b);
  const V3<T> wab = da ? sb[c.index_a].w : zero<T>(), vab = da ? sb[c.index_a].v : zero<T>();
  const V3<T> wbb = db ? sb[c.index_b].w : zero<T>(), vbb = db ? sb[c.index_b].v : zero<T>();
  const V3<T> n = mk<T>(c.normal), t1 = mk<T>(c.tangent1), t2 = mk<T>(c.tangent2);
  const V3<T> lin_b = sub(scale((T)c.inv_mass_b, vbb), scale((T)c.inv_mass_a, vab));
  const V3<T> aa_b = scale((T)-1, sym_mv(c.inv_ia, wab));
  const V3<T> ab_b = sym_mv(c.inv_ib, wbb);
  ib.ti += dot(add(aa_b, ab_b), n);
  const V3<T> fi_b = add(add(lin_b, cross(aa_b, mk<T>(c.center_a))), cross(ab_b, mk<T>(c.center_b)));
  ib.fx += dot(fi_b, t1);
  ib.fy += dot(fi_b, t2);
  const T total_b = dot(lin_b, n);
  const V3<T> ma_b = cross(n, aa_b), mb_b = cross(n, ab_b);
  for (int i = 0; i < c.point_count; ++i)
    ib.ni[i] += total_b + dot(ma_b, mk<T>(c.points[i].r_a)) + dot(mb_b, mk<T>(c.points[i].r_b));
}

// Push: `pre` bodies and impulses as the forward saw them.
template <class T>
__host__ __device__ __forceinline__ void push_adj(const Constraint& c, const ConParams<T>& cp, const Impulses<T>& im_pre,
                                                  const Body<T>* pre, BodyAdj<T>* sb, Impulses<T>& ib, ConParamsAdj<T>& cb,
                                                  const PushRows& r) {
  if (c.manifold_count <= 0) return;
  uint8_t rec[4];
  {
    Body<T> tmp[2] = {load_body(pre, c.index_a), load_body(pre, c.index_b)};
    Constraint cl = c;
    cl.index_a = 0;
    cl.index_b = 1;
    tmp[0].flags = tmp[1].flags = bflag::kDynamic;
    Impulses<T> im = im_pre;
    push(cl, cp, im, tmp, r, rec);
  }
  const Body<T> ba = load_body(pre, c.index_a), bb = load_body(pre, c.index_b);
  const bool da = adj_dynamic(pre, c.index_a), db = adj_dynamic(pre, c.index_b);
  const T inv_h = (T)r.inv_h, bias_rate = (T)r.bias_rate;
  const T inv_mass = (T)c.inv_mass_a + (T)c.inv_mass_b;
  const V3<T> n = mk<T>(c.normal);
  V3<T> vab = da ? sb[c.index_a].v : zero<T>(), wab = da ? sb[c.index_a].w : zero<T>();
  V3<T> vbb = db ? sb[c.index_b].v : zero<T>(), wbb = db ? sb[c.index_b].w : zero<T>();
  const T m_b = dot(n, sub(scale((T)c.inv_mass_b, vbb), scale((T)c.inv_mass_a, vab)));
  T lnv_b = (T)0, nsep_b = (T)0;
  V3<T> na_b = zero<T>(), nb_b = zero<T>();
  for (int i = c.point_count - 1; i >= 0; --i) {
    const V3<T> r_a = mk<T>(c.points[i].r_a), r_b = mk<T>(c.points[i].r_b);
    const V3<T> u_a = cross(r_a, n), u_b = cross(r_b, n);
    const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
    const bool separated = rec[i] & 1, clamped = rec[i] & 2, at_floor = rec[i] & 4;
    const T pm = separated ? (T)1 : (T)r.mass_scale;
    const T pi = separated ? (T)0 : (T)r.impulse_scale;
    const T imp_b = ib.ni[i] + inv_mass * lnv_b + m_b - dot(g_a, wab) + dot(g_b, wbb);
    T ni_b = ib.ni[i];
    T neg_b = (T)0;
    if (at_floor) ni_b -= imp_b;
    else neg_b = -imp_b;
    const T nm = (T)c.points[i].normal_mass;
    const T vn_b = neg_b * nm * pm;
    const T vbias_b = neg_b * nm;
    ni_b += neg_b * pi;
    const T s_b = separated ? vbias_b * inv_h : (clamped ? (T)0 : vbias_b * bias_rate);
    lnv_b += vn_b;
    wbb = add(wbb, scale(vn_b, u_b));
    wab = sub(wab, scale(vn_b, u_a));
    nsep_b += s_b;
    nb_b = add(nb_b, scale(s_b, r_b));
    na_b = sub(na_b, scale(s_b, r_a));
    cb.bs[i] += s_b;
    ib.ni[i] = ni_b;
  }
  vbb = add(vbb, scale(lnv_b, n));
  vab = sub(vab, scale(lnv_b, n));
  if (da) {
    BodyAdj<T>& a = sb[c.index_a];
    a.v = vab;
    a.w = wab;
    a.dp = sub(a.dp, scale(nsep_b, n));
    inv_rotate_wide_adj(ba.dq, n, na_b, a.dq);
  }
  if (db) {
    BodyAdj<T>& b = sb[c.index_b];
    b.v = vbb;
    b.w = wbb;
    b.dp = add(b.dp, scale(nsep_b, n));
    inv_rotate_wide_adj(bb.dq, n, nb_b, b.dq);
  }
}

template <class T>
__host__ __device__ __forceinline__ void relax_adj(const Constraint& c, const ConParams<T>& cp, const Impulses<T>& im_pre,
                                                   const Body<T>* pre, BodyAdj<T>* sb, Impulses<T>& ib, ConParamsAdj<T>& cb,
                                                   float inv_h_f) {
  if (c.manifold_count <= 0) return;
  const Body<T> ba = load_body(pre, c.index_a), bb = load_body(pre, c.index_b);
  const bool da = adj_dynamic(pre, c.index_a), db = adj_dynamic(pre, c.index_b);
  const T inv_h = (T)inv_h_f, eps = (T)kFltEps;
  const T inv_mass = (T)c.inv_mass_a + (T)c.inv_mass_b;
  const V3<T> n = mk<T>(c.normal), t1 = mk<T>(c.tangent1), t2 = mk<T>(c.tangent2);
  const V3<T> ca = mk<T>(c.center_a), cbv = mk<T>(c.center_b);
  // forward recompute of the values the adjoint needs
  uint8_t rec[6];
  T total_normal = (T)0, total_twist_limit = (T)0;
  V3<T> wa_t, wb_t;  // angular velocities after the normal rows (twist input)
  V3<T> wa_f, wb_f;  // after the twist row (friction input)
  T nix = (T)0, niy = (T)0, lsq = (T)0, max_impulse = (T)0;
  {
    Body<T> x = ba, y = bb;
    const V3<T> dp = sub(y.dp, x.dp);
    const T normal_sep = dot(n, dp);
    const V3<T> n_a = inv_rotate_wide(x.dq, n), n_b = inv_rotate_wide(y.dq, n);
    T lnv = dot(n, sub(y.v, x.v));
    for (int i = 0; i < c.point_count; ++i) {
      const V3<T> r_a = mk<T>(c.points[i].r_a), r_b = mk<T>(c.points[i].r_b);
      const T sep = (normal_sep + (dot(n_b, r_b) - dot(n_a, r_a))) + cp.bs[i];
      const T velocity_bias = maxT((T)0, sep * inv_h);
      const V3<T> u_a = cross(r_a, n), u_b = cross(r_b, n);
      const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
      const T v_n = lnv + (dot(u_b, y.w) - dot(u_a, x.w));
      const T ni = im_pre.ni[i];
      const T neg_impulse = (T)c.points[i].normal_mass * (v_n + velocity_bias);
      const bool at_floor = !(-neg_impulse > -ni);
      const T impulse = maxT(-neg_impulse, -ni);
      const T new_impulse = ni + impulse;
      total_normal += new_impulse;
      total_twist_limit += (T)c.points[i].lever_arm * new_impulse;
      lnv += inv_mass * impulse;
      x.w = sub(x.w, scale(impulse, g_a));
      y.w = add(y.w, scale(impulse, g_b));
      rec[i] = (uint8_t)(((T)0 < sep * inv_h ? 1 : 0) | (at_floor ? 4 : 0));
    }
    wa_t = x.w;
    wb_t = y.w;
    const T twist_speed = dot(n, sub(y.w, x.w));
    const T max_lambda = cp.mu * total_twist_limit;
    const T old = im_pre.ti;
    const T raw = -((T)c.twist_mass * twist_speed);
    const T hi = minT(raw, max_lambda - old);
    const T impulse = maxT(-max_lambda - old, hi);
    rec[4] = (uint8_t)(!(-max_lambda - old < hi) ? 2 : (!(raw < max_lambda - old) ? 1 : 0));
    x.w = sub(x.w, scale(impulse, sym_mv(c.inv_ia, n)));
    y.w = add(y.w, scale(impulse, sym_mv(c.inv_ib, n)));
    wa_f = x.w;
    wb_f = y.w;
    const V3<T> dv = sub(y.v, x.v);
    const T ltv1 = dot(t1, dv) - (T)c.tangent_velocity1;
    const T ltv2 = dot(t2, dv) - (T)c.tangent_velocity2;
    const V3<T> arv = sub(cross(y.w, cbv), cross(x.w, ca));
    const T vt1 = ltv1 + dot(arv, t1), vt2 = ltv2 + dot(arv, t2);
    const Mat2& tm = c.tangent_mass;
    const T dix = -((T)tm.cx.x * vt1 + (T)tm.cy.x * vt2);
    const T diy = -((T)tm.cx.y * vt1 + (T)tm.cy.y * vt2);
    nix = im_pre.fx + dix;
    niy = im_pre.fy + diy;
    max_impulse = cp.mu * total_normal;
    lsq = nix * nix + niy * niy;
    rec[5] = lsq > max_impulse * max_impulse ? 1 : 0;
  }
  (void)wa_t;
  (void)wb_t;
  (void)wa_f;
  (void)wb_f;
  V3<T> vab = da ? sb[c.index_a].v : zero<T>(), wab = da ? sb[c.index_a].w : zero<T>();
  V3<T> vbb = db ? sb[c.index_b].v : zero<T>(), wbb = db ? sb[c.index_b].w : zero<T>();
  // final linear update: pp2 = pp + M n; va -= ma pp2; vb += mb pp2
  const V3<T> pp2_b = sub(scale((T)c.inv_mass_b, vbb), scale((T)c.inv_mass_a, vab));
  const T m_b = dot(pp2_b, n);
  V3<T> pp_b = pp2_b;
  // friction angular updates: wa -= Ia (ca x pp); wb += Ib (cb x pp)
  pp_b = add(pp_b, cross(scale((T)-1, sym_mv(c.inv_ia, wab)), ca));
  pp_b = add(pp_b, cross(sym_mv(c.inv_ib, wbb), cbv));
  // pp = dix t1 + diy t2 (dix = nix' - fx_old after the cone scaling)
  const T di_xb = dot(pp_b, t1), di_yb = dot(pp_b, t2);
  T ns_xb = di_xb + ib.fx, ns_yb = di_yb + ib.fy;  // adjoint of the scaled impulse (incl. its output)
  T fold_xb = -di_xb, fold_yb = -di_yb;
  T nr_xb, nr_yb, maxi_b = (T)0;
  if (rec[5]) {
    const T L = sqrtT(lsq);
    const T den = L + eps;
    const T sc = max_impulse / den;
    const T dotp = ns_xb * nix + ns_yb * niy;
    const T k = (L > (T)0) ? -max_impulse / (den * den * L) : (T)0;
    nr_xb = sc * ns_xb + dotp * k * nix;
    nr_yb = sc * ns_yb + dotp * k * niy;
    maxi_b = dotp / den;
  } else {
    nr_xb = ns_xb;
    nr_yb = ns_yb;
  }
  T mu_b = maxi_b * total_normal;
  T total_normal_b = maxi_b * cp.mu;
  // ni_raw = f_old + di_raw
  fold_xb += nr_xb;
  fold_yb += nr_yb;
  // di_raw = -(TM vt)
  const Mat2& tm = c.tangent_mass;
  const T vt1_b = -((T)tm.cx.x * nr_xb + (T)tm.cx.y * nr_yb);
  const T vt2_b = -((T)tm.cy.x * nr_xb + (T)tm.cy.y * nr_yb);
  T ltv1_b = vt1_b, ltv2_b = vt2_b;
  const V3<T> arv_b = add(scale(vt1_b, t1), scale(vt2_b, t2));
  // arv = wb x cb - wa x ca (w after the twist row)
  wbb = add(wbb, cross(cbv, arv_b));
  wab = sub(wab, cross(ca, arv_b));
  ib.fx = fold_xb;
  ib.fy = fold_yb;
  // twist row: wa_f = wa_t - imp Ia n; wb_f = wb_t + imp Ib n
  {
    const V3<T> ian = sym_mv(c.inv_ia, n), ibn = sym_mv(c.inv_ib, n);
    const T imp_b = ib.ti - dot(ian, wab) + dot(ibn, wbb);
    T old_b = ib.ti;
    T raw_b = (T)0, ml_b = (T)0;
    if (rec[4] == 0) raw_b = imp_b;
    else if (rec[4] == 1) { ml_b += imp_b; old_b -= imp_b; }
    else { ml_b -= imp_b; old_b -= imp_b; }
    const T ts_b = -(T)c.twist_mass * raw_b;
    wbb = add(wbb, scale(ts_b, n));
    wab = sub(wab, scale(ts_b, n));
    mu_b += ml_b * total_twist_limit;
    const T ttl_b = ml_b * cp.mu;
    ib.ti = old_b;
    // normal rows, reversed
    T lnv_b = (T)0, nsep_b = (T)0;
    V3<T> na_b = zero<T>(), nb_b = zero<T>();
    for (int i = c.point_count - 1; i >= 0; --i) {
      const V3<T> r_a = mk<T>(c.points[i].r_a), r_b = mk<T>(c.points[i].r_b);
      const V3<T> u_a = cross(r_a, n), u_b = cross(r_b, n);
      const V3<T> g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
      const T new_b = total_normal_b + (T)c.points[i].lever_arm * ttl_b + ib.ni[i];
      const T imp_bi = new_b + inv_mass * lnv_b + m_b - dot(g_a, wab) + dot(g_b, wbb);
      T ni_b = new_b;
      T neg_b = (T)0;
      if (rec[i] & 4) ni_b -= imp_bi;
      else neg_b = -imp_bi;
      const T nm = (T)c.points[i].normal_mass;
      const T vn_b = neg_b * nm;
      const T s_b = (rec[i] & 1) ? neg_b * nm * inv_h : (T)0;
      lnv_b += vn_b;
      wbb = add(wbb, scale(vn_b, u_b));
      wab = sub(wab, scale(vn_b, u_a));
      nsep_b += s_b;
      nb_b = add(nb_b, scale(s_b, r_b));
      na_b = sub(na_b, scale(s_b, r_a));
      cb.bs[i] += s_b;
      ib.ni[i] = ni_b;
    }
    // lnv0 = n . dv, ltv = t . dv with dv = vb - va
    const V3<T> dv_b = add(add(scale(lnv_b, n), scale(ltv1_b, t1)), scale(ltv2_b, t2));
    vbb = add(vbb, dv_b);
    vab = sub(vab, dv_b);
    cb.mu += mu_b;
    if (da) {
      BodyAdj<T>& a = sb[c.index_a];
      a.v = vab;
      a.w = wab;
      a.dp = sub(a.dp, scale(nsep_b, n));
      inv_rotate_wide_adj(ba.dq, n, na_b, a.dq);
    }
    if (db) {
      BodyAdj<T>& b = sb[c.index_b];
      b.v = vbb;
      b.w = wbb;
      b.dp = add(b.dp, scale(nsep_b, n));
      inv_rotate_wide_adj(bb.dq, n, nb_b, b.dq);
    }
  }
}

// integrate_positions: dp += h v; dq = integrate_rotation(dq, h w).
template <class T> __host__ __device__ __forceinline__ void integrate_positions_adj(const Body<T>& pre, BodyAdj<T>& a, float h) {
  a.v = add(a.v, scale((T)h, a.dp));
  Q4<T> dq_b = {zero<T>(), (T)0};
  V3<T> dr_b = zero<T>();
  integrate_rotation_adj(pre.dq, scale((T)h, pre.w), a.dq, dq_b, dr_b);
  a.dq = dq_b;
  a.w = add(a.w, scale((T)h, dr_b));
}

// integrate_velocities: v = dv + ld v, w = dw + ad w.
template <class T>
__host__ __device__ __forceinline__ void integrate_velocities_adj(BodyAdj<T>& a, float h, float lin_damp, float ang_damp) {
  const float ld = 1.0f / (1.0f + h * lin_damp);
  const float ad = 1.0f / (1.0f + h * ang_damp);
  a.v = scale((T)ld, a.v);
  a.w = scale((T)ad, a.w);
}

}  // namespace diff
}  // namespace nsl
