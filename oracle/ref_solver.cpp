// Oracle: contact preparation, the colored soft-step stages, the sequential overflow
// solver, integration and impulse storage.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// Sources: src/simulation/solver_prepare.rs:269-474, contact_solver.rs:101-1299,
// overflow_contact_solver.rs:92-722, solver_integration.rs:263-426,
// solver_stages.rs:543-589. The reference's 4-wide path is bit-identical to its
// scalar path (src/tests/simulation/wide_equivalence_tests.rs), so the scalar
// formulas are restated.
#include "ref_world.h"

namespace nslref {

SymMat3 invert_sym(SymMat3 m) {
  float cxx = m.cyy * m.czz - m.cyz * m.cyz;
  float cxy = m.cxz * m.cyz - m.cxy * m.czz;
  float cxz = m.cxy * m.cyz - m.cxz * m.cyy;
  float cyy = m.cxx * m.czz - m.cxz * m.cxz;
  float cyz = m.cxy * m.cxz - m.cxx * m.cyz;
  float czz = m.cxx * m.cyy - m.cxy * m.cxy;
  float d = m.cxx * cxx + (m.cxy * cxy + m.cxz * cxz);
  bool valid = absf(d) > 1000.0f * FLT_MIN;
  float inv = valid ? 1.0f / d : 0.0f;
  return {inv * cxx, inv * cxy, inv * cxz, inv * cyy, inv * cyz, inv * czz};
}

static inline BodyState identity_state() {
  BodyState s{};
  s.dq = kQuatIdentity;
  return s;
}
static inline BodyState state_at(std::vector<BodyState>& states, int32_t index) {
  return index >= 0 ? states[index] : identity_state();
}
static inline void set_state_at(std::vector<BodyState>& states, int32_t index, const BodyState& s) {
  if (index >= 0 && (s.flags & BF::DYNAMIC_FLAG) != 0) states[index] = s;
}

// contact_solver.rs:395-662
void World::prepare_contact_constraint(int32_t contact_id, ContactConstraint& cc, int32_t manifold_start, Softness softness,
                                       Softness static_softness, float warm_start_scale) {
  const float inv_tau = 1.0f / speculative_distance();
  const Contact& contact = contacts[contact_id];
  Vec3 tangent_velocity = contact.tangent_velocity;
  cc = ContactConstraint{};
  cc.contact_id = contact_id;
  cc.manifold_start = manifold_start;
  cc.manifold_count = 0;
  int32_t ia = contact.encoded_a >= 0 ? contact.encoded_a : kNullIndex;
  int32_t ib = contact.encoded_b >= 0 ? contact.encoded_b : kNullIndex;
  cc.index_a = ia;
  cc.index_b = ib;
  if (ia >= 0) { cc.inv_mass_a = awake_sims[ia].inv_mass; cc.inv_ia = sym_from(awake_sims[ia].inv_inertia_world); }
  if (ib >= 0) { cc.inv_mass_b = awake_sims[ib].inv_mass; cc.inv_ib = sym_from(awake_sims[ib].inv_inertia_world); }
  cc.friction = contact.friction;
  cc.restitution = contact.restitution;
  cc.rolling_resistance = contact.rolling_resistance;
  cc.is_static = (contact.flags & CF::STATIC_FLAG) != 0;
  cc.softness = cc.is_static ? static_softness : softness;
  SymMat3 inv_iab = add_sym(cc.inv_ia, cc.inv_ib);
  cc.rolling_mass = cc.rolling_resistance > 0.0f ? invert_sym(inv_iab) : SymMat3{};

  for (int mi = 0; mi < contact.manifold_count; ++mi) {
    const Manifold& m = contact.manifold;
    ManifoldConstraint mc{};
    int pc = m.point_count;
    mc.point_count = pc;
    mc.normal = m.normal;
    mc.tangent1 = perp(m.normal);
    mc.tangent2 = cross(mc.tangent1, m.normal);
    mc.tangent_velocity1 = dot(tangent_velocity, mc.tangent1);
    mc.tangent_velocity2 = dot(tangent_velocity, mc.tangent2);
    mc.twist_impulse = warm_start_scale * m.twist_impulse;
    mc.rolling_impulse = cc.rolling_resistance > 0.0f ? scale(warm_start_scale, m.rolling_impulse) : kZero3;
    mc.friction_impulse = Vec2{warm_start_scale * dot(m.friction_impulse, mc.tangent1),
                               warm_start_scale * dot(m.friction_impulse, mc.tangent2)};
    Vec3 normal = mc.normal, t1 = mc.tangent1, t2 = mc.tangent2;
    Vec3 center_a = kZero3, center_b = kZero3;
    float total_weight = 0.0f;
    for (int p = 0; p < pc; ++p) {
      const ManifoldPoint& mp = m.points[p];
      ConstraintPoint cp{};
      cp.r_a = mp.anchor_a;
      cp.r_b = mp.anchor_b;
      cp.base_separation = mp.separation - dot(sub(cp.r_b, cp.r_a), normal);
      cp.normal_impulse = warm_start_scale * mp.normal_impulse;
      float weight = clampf(2.0f - mp.separation * inv_tau, kMinFrictionWeight, 1.0f);
      center_a = add(center_a, scale(weight, cp.r_a));
      center_b = add(center_b, scale(weight, cp.r_b));
      total_weight += weight;
      Vec3 rn_a = cross(cp.r_a, normal), rn_b = cross(cp.r_b, normal);
      float k = cc.inv_mass_a + cc.inv_mass_b;
      k += dot(rn_a, sym_mv(cc.inv_ia, rn_a));
      k += dot(rn_b, sym_mv(cc.inv_ib, rn_b));
      cp.normal_mass = k > 0.0f ? 1.0f / k : 0.0f;
      mc.points[p] = cp;
    }
    if (total_weight > 0.0f) {
      mc.center_a = scale(1.0f / total_weight, center_a);
      mc.center_b = scale(1.0f / total_weight, center_b);
    } else {
      mc.center_a = kZero3;
      mc.center_b = kZero3;
    }
    for (int p = 0; p < pc; ++p) {
      Vec3 d = sub(mc.points[p].r_a, mc.center_a);
      mc.points[p].lever_arm = std::sqrt(dot(d, d));
    }
    {
      Vec3 rt_a1 = cross(mc.center_a, t1), rt_a2 = cross(mc.center_a, t2);
      Vec3 rt_b1 = cross(mc.center_b, t1), rt_b2 = cross(mc.center_b, t2);
      Vec3 ia1 = sym_mv(cc.inv_ia, rt_a1), ia2 = sym_mv(cc.inv_ia, rt_a2);
      Vec3 ib1 = sym_mv(cc.inv_ib, rt_b1), ib2 = sym_mv(cc.inv_ib, rt_b2);
      float m_sum = cc.inv_mass_a + cc.inv_mass_b;
      float kxx = (m_sum + dot(rt_a1, ia1)) + dot(rt_b1, ib1);
      float kyy = (m_sum + dot(rt_a2, ia2)) + dot(rt_b2, ib2);
      float kxy = dot(rt_a1, ia2) + dot(rt_b1, ib2);
      float d = kxx * kyy - kxy * kxy;
      bool valid = absf(d) > 1000.0f * FLT_MIN && total_weight > 0.0f;
      float inv_det = valid ? 1.0f / d : 0.0f;
      mc.tangent_mass.cx = Vec2{inv_det * kyy, -(inv_det * kxy)};
      mc.tangent_mass.cy = Vec2{-(inv_det * kxy), inv_det * kxx};
    }
    float k_twist = dot(normal, sym_mv(inv_iab, normal));
    mc.twist_mass = k_twist > 0.0f ? 1.0f / k_twist : 0.0f;
    bool have_restitution = contact.restitution > 0.0f;
    bool hit_events = (contact.flags & CF::SIM_ENABLE_HIT_EVENT) != 0;
    if (have_restitution || hit_events) {
      BodyState sa = ia >= 0 ? awake_states[ia] : identity_state();
      BodyState sb = ib >= 0 ? awake_states[ib] : identity_state();
      for (int p = 0; p < pc; ++p) {
        ConstraintPoint& cp = mc.points[p];
        Vec3 vr_a = add(sa.v, cross(sa.w, cp.r_a));
        Vec3 vr_b = add(sb.v, cross(sb.w, cp.r_b));
        cp.relative_velocity = dot(normal, sub(vr_b, vr_a));
      }
    }
    manifold_pool[manifold_start + cc.manifold_count] = mc;
    cc.manifold_count += 1;
  }
}

// overflow_contact_solver.rs:92-257
void World::prepare_overflow_contacts(Softness softness, Softness static_softness) {
  overflow_constraints.clear();
  overflow_manifolds.clear();
  const std::vector<ContactSpec>& specs = colors[kGraphColorCount - 1].contacts;
  if (specs.empty()) return;
  const float warm = enable_warm_starting ? 1.0f : 0.0f;
  const float inv_tau = 1.0f / speculative_distance();
  bool any_restitution = false;
  for (const ContactSpec& spec : specs) {
    Contact& contact = contacts[spec.contact_id];
    int32_t ia = contact.encoded_a >= 0 ? contact.encoded_a : kNullIndex;
    int32_t ib = contact.encoded_b >= 0 ? contact.encoded_b : kNullIndex;
    OverflowConstraint c{};
    c.contact_id = spec.contact_id;
    c.index_a = ia;
    c.index_b = ib;
    if (ia >= 0) { c.inv_mass_a = awake_sims[ia].inv_mass; c.inv_ia = awake_sims[ia].inv_inertia_world; }
    if (ib >= 0) { c.inv_mass_b = awake_sims[ib].inv_mass; c.inv_ib = awake_sims[ib].inv_inertia_world; }
    float mA = c.inv_mass_a, mB = c.inv_mass_b;
    Mat3 i_a = c.inv_ia, i_b = c.inv_ib;
    c.rolling_mass = invert_matrix(add_mm(i_a, i_b));
    c.softness = (contact.flags & CF::STATIC_FLAG) ? static_softness : softness;
    c.friction = contact.friction;
    c.restitution = contact.restitution;
    c.rolling_resistance = contact.rolling_resistance;
    bool have_restitution = contact.restitution > 0.0f;
    bool hit_events = (contact.flags & CF::SIM_ENABLE_HIT_EVENT) != 0;
    bool sample_velocity = have_restitution || hit_events;
    any_restitution = any_restitution || have_restitution;
    Vec3 v_a = kZero3, w_a = kZero3, v_b = kZero3, w_b = kZero3;
    if (sample_velocity) {
      if (ia >= 0) { v_a = awake_states[ia].v; w_a = awake_states[ia].w; }
      if (ib >= 0) { v_b = awake_states[ib].v; w_b = awake_states[ib].w; }
    }
    c.manifold_start = (int32_t)overflow_manifolds.size();
    c.manifold_count = contact.manifold_count;
    for (int mi = 0; mi < contact.manifold_count; ++mi) {
      Manifold& m = contact.manifold;
      ManifoldConstraint mc{};
      int pc = m.point_count;
      Vec3 normal = m.normal;
      Vec3 t1 = perp(normal);
      Vec3 t2 = cross(t1, normal);
      mc.point_count = pc;
      mc.normal = normal;
      mc.tangent1 = t1;
      mc.tangent2 = t2;
      mc.tangent_velocity1 = dot(contact.tangent_velocity, t1);
      mc.tangent_velocity2 = dot(contact.tangent_velocity, t2);
      Vec3 center_a = kZero3, center_b = kZero3;
      float total_w = 0.0f;
      for (int p = 0; p < pc; ++p) {
        ConstraintPoint cp{};
        ManifoldPoint mp = m.points[p];
        cp.r_a = mp.anchor_a;
        cp.r_b = mp.anchor_b;
        float s = mp.separation;
        cp.base_separation = s - dot(sub(cp.r_b, cp.r_a), normal);
        cp.normal_impulse = warm * mp.normal_impulse;
        Vec3 rn_a = cross(cp.r_a, normal), rn_b = cross(cp.r_b, normal);
        float k = mA + mB + dot(rn_a, mul_mv(i_a, rn_a)) + dot(rn_b, mul_mv(i_b, rn_b));
        cp.normal_mass = k > 0.0f ? 1.0f / k : 0.0f;
        if (sample_velocity) {
          Vec3 vr_a = add(v_a, cross(w_a, cp.r_a));
          Vec3 vr_b = add(v_b, cross(w_b, cp.r_b));
          float vn = dot(normal, sub(vr_b, vr_a));
          cp.relative_velocity = vn;
          m.points[p].normal_velocity = hit_events ? vn : 0.0f;
        } else {
          cp.relative_velocity = 0.0f;
          m.points[p].normal_velocity = 0.0f;
        }
        float weight = clampf(2.0f - s * inv_tau, kMinFrictionWeight, 1.0f);
        center_a = mul_add(center_a, weight, cp.r_a);
        center_b = mul_add(center_b, weight, cp.r_b);
        total_w += weight;
        mc.points[p] = cp;
      }
      float inv_w = 1.0f / total_w;
      center_a = scale(inv_w, center_a);
      center_b = scale(inv_w, center_b);
      mc.center_a = center_a;
      mc.center_b = center_b;
      for (int p = 0; p < pc; ++p) mc.points[p].lever_arm = distance(mc.points[p].r_a, center_a);
      Vec3 rt_a1 = cross(center_a, t1), rt_a2 = cross(center_a, t2);
      Vec3 rt_b1 = cross(center_b, t1), rt_b2 = cross(center_b, t2);
      Mat2 k{};
      k.cx.x = mA + mB + dot(rt_a1, mul_mv(i_a, rt_a1)) + dot(rt_b1, mul_mv(i_b, rt_b1));
      k.cy.y = mA + mB + dot(rt_a2, mul_mv(i_a, rt_a2)) + dot(rt_b2, mul_mv(i_b, rt_b2));
      k.cx.y = dot(rt_a1, mul_mv(i_a, rt_a2)) + dot(rt_b1, mul_mv(i_b, rt_b2));
      k.cy.x = k.cx.y;
      mc.tangent_mass = invert2(k);
      mc.friction_impulse = Vec2{warm * dot(m.friction_impulse, t1), warm * dot(m.friction_impulse, t2)};
      float k_twist = dot(normal, mul_mv(add_mm(i_a, i_b), normal));
      mc.twist_mass = k_twist > 0.0f ? 1.0f / k_twist : 0.0f;
      mc.twist_impulse = warm * m.twist_impulse;
      mc.rolling_impulse = scale(warm, m.rolling_impulse);
      overflow_manifolds.push_back(mc);
    }
    overflow_constraints.push_back(c);
  }
  stats.overflow_constraints += (int64_t)overflow_constraints.size();
  if (any_restitution) solve_any_restitution = true;
}

// solver_prepare.rs:269-474 (scalar plan; no joints, no mesh contacts)
void World::prepare_contacts(const StepContext& ctx) {
  for (int i = 0; i < kGraphColorCount; ++i) color_constraints[i].clear();
  solve_any_restitution = false;
  float contact_hz = minf(contact_hertz, 0.125f * ctx.inv_h);
  Softness soft = make_soft(contact_hz, contact_damping_ratio, ctx.h);
  Softness static_soft = make_soft(2.0f * contact_hz, 0.5f * contact_damping_ratio, ctx.h);
  prepare_overflow_contacts(soft, static_soft);
  int32_t manifold_total = 0;
  bool any_restitution = false;
  for (int ci = 0; ci < kGraphColorCount - 1; ++ci)
    for (int32_t id : colors[ci].convex_contacts) {
      manifold_total += contacts[id].manifold_count;
      any_restitution = any_restitution || contacts[id].restitution > 0.0f;
    }
  if (any_restitution) solve_any_restitution = true;
  if ((int32_t)manifold_pool.size() < manifold_total) manifold_pool.resize(manifold_total);
  const float warm = enable_warm_starting ? 1.0f : 0.0f;
  // manifold slots are assigned in colour order first; the constraints are then independent
  struct Item { int32_t ci, k, start; };
  std::vector<Item> items;
  int32_t start = 0;
  for (int ci = 0; ci < kGraphColorCount - 1; ++ci) {
    color_constraints[ci].resize(colors[ci].convex_contacts.size());
    for (size_t k = 0; k < colors[ci].convex_contacts.size(); ++k) {
      items.push_back(Item{ci, (int32_t)k, start});
      start += contacts[colors[ci].convex_contacts[k]].manifold_count;
    }
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t t = 0; t < items.size(); ++t) {
    const Item& it = items[t];
    prepare_contact_constraint(colors[it.ci].convex_contacts[it.k], color_constraints[it.ci][it.k], it.start, soft, static_soft, warm);
  }
}

// solver_integration.rs:263-358
void World::integrate_velocities(float h) {
  Vec3 g = gravity;
#pragma omp for schedule(static)
  for (size_t i = 0; i < awake_sims.size(); ++i) {
    const BodySim& sim = awake_sims[i];
    BodyState& state = awake_states[i];
    Vec3 v = state.v, w = state.w;
    float linear_damping = 1.0f / (1.0f + h * sim.linear_damping);
    float angular_damping = 1.0f / (1.0f + h * sim.angular_damping);
    float gravity_scale = sim.inv_mass > 0.0f ? sim.gravity_scale : 0.0f;
    Vec3 dv = blend2(h * sim.inv_mass, sim.force, h * gravity_scale, g);
    v = mul_add(dv, linear_damping, v);
    Vec3 dw = scale(h, mul_mv(sim.inv_inertia_world, sim.torque));
    w = mul_add(dw, angular_damping, w);
    if (sim.flags & BF::ENABLE_GYROSCOPIC_TORQUE) {
      Quat q0 = sim.transform.q;
      Quat q = mul_quat(state.dq, q0);
      Vec3 omega1 = inv_rotate(q, w);
      Vec3 omega2 = omega1;
      float i00 = sim.inertia.cx.x, i01 = sim.inertia.cy.x, i02 = sim.inertia.cz.x;
      float i11 = sim.inertia.cy.y, i12 = sim.inertia.cz.y, i22 = sim.inertia.cz.z;
      for (int it = 0; it < kGyroscopicIterations; ++it) {
        float w1 = omega2.x, w2 = omega2.y, w3 = omega2.z;
        float iw1 = i00 * w1 + i01 * w2 + i02 * w3;
        float iw2 = i01 * w1 + i11 * w2 + i12 * w3;
        float iw3 = i02 * w1 + i12 * w2 + i22 * w3;
        Vec3 dw_ = sub(omega2, omega1);
        Vec3 b = {i00 * dw_.x + i01 * dw_.y + i02 * dw_.z + h * (w2 * iw3 - w3 * iw2),
                  i01 * dw_.x + i11 * dw_.y + i12 * dw_.z + h * (w3 * iw1 - w1 * iw3),
                  i02 * dw_.x + i12 * dw_.y + i22 * dw_.z + h * (w1 * iw2 - w2 * iw1)};
        Mat3 j = {{i00 + h * (w2 * i02 - w3 * i01), i01 + h * (w3 * i00 - w1 * i02 - iw3), i02 + h * (w1 * i01 - w2 * i00 + iw2)},
                  {i01 + h * (w2 * i12 - w3 * i11 + iw3), i11 + h * (w3 * i01 - w1 * i12), i12 + h * (w1 * i11 - w2 * i01 - iw1)},
                  {i02 + h * (w2 * i22 - w3 * i12 - iw2), i12 + h * (w3 * i02 - w1 * i22 + iw1), i22 + h * (w1 * i12 - w2 * i02)}};
        omega2 = sub(omega2, solve3(j, b));
      }
      w = rotate(q, omega2);
    }
    state.v = v;
    state.w = w;
  }
}

// solver_integration.rs:377-426
void World::integrate_positions(float h, float inv_dt) {
  float max_lin = max_linear_speed;
  float max_ang = kMaxRotation * inv_dt;
  float max_lin_sq = max_lin * max_lin, max_ang_sq = max_ang * max_ang;
#pragma omp for schedule(static)
  for (size_t i = 0; i < awake_states.size(); ++i) {
    BodyState& st = awake_states[i];
    Vec3 v = st.v, w = st.w;
    if (st.flags & BF::LOCK_LINEAR_X) v.x = 0.0f;
    if (st.flags & BF::LOCK_LINEAR_Y) v.y = 0.0f;
    if (st.flags & BF::LOCK_LINEAR_Z) v.z = 0.0f;
    if (st.flags & BF::LOCK_ANGULAR_X) w.x = 0.0f;
    if (st.flags & BF::LOCK_ANGULAR_Y) w.y = 0.0f;
    if (st.flags & BF::LOCK_ANGULAR_Z) w.z = 0.0f;
    if (dot(v, v) > max_lin_sq) {
      float ratio = max_lin / length(v);
      v = scale(ratio, v);
      st.flags |= BF::IS_SPEED_CAPPED;
    }
    if (dot(w, w) > max_ang_sq && (st.flags & BF::ALLOW_FAST_ROTATION) == 0) {
      float ratio = max_ang / length(w);
      w = scale(ratio, w);
      st.flags |= BF::IS_SPEED_CAPPED;
    }
    st.v = v;
    st.w = w;
    st.dp = mul_add(st.dp, h, v);
    st.dq = integrate_rotation(st.dq, scale(h, w));
  }
}

// contact_solver.rs:697-763
void World::warm_start_color(int ci) {
  const std::vector<ContactConstraint>& list_ = color_constraints[ci];
#pragma omp for schedule(static)
  for (size_t k_ = 0; k_ < list_.size(); ++k_) {
    const ContactConstraint& c = list_[k_];
    Vec3 linear = kZero3, ang_a = kZero3, ang_b = kZero3;
    for (int mo = 0; mo < c.manifold_count; ++mo) {
      ManifoldConstraint& cm = manifold_pool[c.manifold_start + mo];
      float total = 0.0f;
      Vec3 moment_a = kZero3, moment_b = kZero3;
      for (int p = 0; p < cm.point_count; ++p) {
        const ConstraintPoint& cp = cm.points[p];
        total += cp.normal_impulse;
        moment_a = mul_add(moment_a, cp.normal_impulse, cp.r_a);
        moment_b = mul_add(moment_b, cp.normal_impulse, cp.r_b);
      }
      Vec3 normal = cm.normal;
      Vec3 fi = scale(cm.friction_impulse.x, cm.tangent1);
      fi = mul_add(fi, cm.friction_impulse.y, cm.tangent2);
      Vec3 lin = mul_add(fi, total, normal);
      Vec3 twist = scale(cm.twist_impulse, normal);
      Vec3 a_a = add(add(cross(moment_a, normal), cross(cm.center_a, fi)), twist);
      Vec3 a_b = add(add(cross(moment_b, normal), cross(cm.center_b, fi)), twist);
      if (c.rolling_resistance > 0.0f) {
        a_a = add(a_a, cm.rolling_impulse);
        a_b = add(a_b, cm.rolling_impulse);
      }
      linear = add(linear, lin);
      ang_a = add(ang_a, a_a);
      ang_b = add(ang_b, a_b);
    }
    BodyState ba = state_at(awake_states, c.index_a);
    BodyState bb = state_at(awake_states, c.index_b);
    ba.w = sub(ba.w, sym_mv(c.inv_ia, ang_a));
    ba.v = sub(ba.v, scale(c.inv_mass_a, linear));
    bb.w = add(bb.w, sym_mv(c.inv_ib, ang_b));
    bb.v = add(bb.v, scale(c.inv_mass_b, linear));
    set_state_at(awake_states, c.index_a, ba);
    set_state_at(awake_states, c.index_b, bb);
  }
}

// contact_solver.rs:796-894
void World::push_color(int ci, const StepContext& ctx) {
  const float inv_h = ctx.inv_h;
  const float contact_speed_neg = -contact_speed;
  const float dyn_bias = ctx.contact_softness.mass_scale * ctx.contact_softness.bias_rate;
  const float dyn_mass = ctx.contact_softness.mass_scale, dyn_imp = ctx.contact_softness.impulse_scale;
  const float st_bias = ctx.static_softness.mass_scale * ctx.static_softness.bias_rate;
  const float st_mass = ctx.static_softness.mass_scale, st_imp = ctx.static_softness.impulse_scale;
  const std::vector<ContactConstraint>& list_ = color_constraints[ci];
#pragma omp for schedule(static)
  for (size_t k_ = 0; k_ < list_.size(); ++k_) {
    const ContactConstraint& c = list_[k_];
    BodyState ba = state_at(awake_states, c.index_a);
    BodyState bb = state_at(awake_states, c.index_b);
    float bias_rate = c.is_static ? st_bias : dyn_bias;
    float mass_scale = c.is_static ? st_mass : dyn_mass;
    float impulse_scale = c.is_static ? st_imp : dyn_imp;
    Vec3 dp = sub(bb.dp, ba.dp);
    float inv_mass = c.inv_mass_a + c.inv_mass_b;
    for (int mo = 0; mo < c.manifold_count; ++mo) {
      ManifoldConstraint& cm = manifold_pool[c.manifold_start + mo];
      Vec3 normal = cm.normal;
      float normal_sep = dot(normal, dp);
      Vec3 n_a = inv_rotate_vector_wide(ba.dq, normal);
      Vec3 n_b = inv_rotate_vector_wide(bb.dq, normal);
      float lnv = dot(normal, sub(bb.v, ba.v));
      float manifold_impulse = 0.0f;
      for (int p = 0; p < cm.point_count; ++p) {
        ConstraintPoint& cp = cm.points[p];
        float s = (normal_sep + (dot(n_b, cp.r_b) - dot(n_a, cp.r_a))) + cp.base_separation;
        bool separated = s > 0.0f;
        float spec_bias = s * inv_h;
        float overlap_bias = maxf(bias_rate * s, contact_speed_neg);
        float velocity_bias = separated ? spec_bias : overlap_bias;
        float pm = separated ? 1.0f : mass_scale;
        float pi = separated ? 0.0f : impulse_scale;
        Vec3 u_a = cross(cp.r_a, normal), u_b = cross(cp.r_b, normal);
        Vec3 g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
        float v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
        float neg_impulse = cp.normal_mass * (pm * v_n + velocity_bias) + pi * cp.normal_impulse;
        float impulse = maxf(-neg_impulse, -cp.normal_impulse);
        cp.normal_impulse = cp.normal_impulse + impulse;
        lnv += inv_mass * impulse;
        manifold_impulse += impulse;
        ba.w = sub(ba.w, scale(impulse, g_a));
        bb.w = add(bb.w, scale(impulse, g_b));
      }
      Vec3 pp = scale(manifold_impulse, normal);
      ba.v = sub(ba.v, scale(c.inv_mass_a, pp));
      bb.v = add(bb.v, scale(c.inv_mass_b, pp));
    }
    set_state_at(awake_states, c.index_a, ba);
    set_state_at(awake_states, c.index_b, bb);
  }
}

// contact_solver.rs:927-1097
void World::relax_color(int ci, const StepContext& ctx) {
  const float inv_h = ctx.inv_h;
  const float eps = FLT_EPSILON;
  const std::vector<ContactConstraint>& list_ = color_constraints[ci];
#pragma omp for schedule(static)
  for (size_t k_ = 0; k_ < list_.size(); ++k_) {
    const ContactConstraint& c = list_[k_];
    BodyState ba = state_at(awake_states, c.index_a);
    BodyState bb = state_at(awake_states, c.index_b);
    bool any_rolling = c.rolling_resistance > 0.0f;
    Vec3 dp = sub(bb.dp, ba.dp);
    float inv_mass = c.inv_mass_a + c.inv_mass_b;
    for (int mo = 0; mo < c.manifold_count; ++mo) {
      ManifoldConstraint& cm = manifold_pool[c.manifold_start + mo];
      Vec3 normal = cm.normal;
      float normal_sep = dot(normal, dp);
      Vec3 n_a = inv_rotate_vector_wide(ba.dq, normal);
      Vec3 n_b = inv_rotate_vector_wide(bb.dq, normal);
      Vec3 dv = sub(bb.v, ba.v);
      float lnv = dot(normal, dv);
      float ltv1 = dot(cm.tangent1, dv) - cm.tangent_velocity1;
      float ltv2 = dot(cm.tangent2, dv) - cm.tangent_velocity2;
      float manifold_impulse = 0.0f, total_normal = 0.0f, total_twist_limit = 0.0f;
      for (int p = 0; p < cm.point_count; ++p) {
        ConstraintPoint& cp = cm.points[p];
        float s = (normal_sep + (dot(n_b, cp.r_b) - dot(n_a, cp.r_a))) + cp.base_separation;
        float velocity_bias = maxf(0.0f, s * inv_h);
        Vec3 u_a = cross(cp.r_a, normal), u_b = cross(cp.r_b, normal);
        Vec3 g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
        float v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
        float neg_impulse = cp.normal_mass * (v_n + velocity_bias);
        float impulse = maxf(-neg_impulse, -cp.normal_impulse);
        float new_impulse = cp.normal_impulse + impulse;
        cp.normal_impulse = new_impulse;
        cp.total_normal_impulse = cp.total_normal_impulse + new_impulse;
        total_normal += new_impulse;
        total_twist_limit += cp.lever_arm * new_impulse;
        lnv += inv_mass * impulse;
        manifold_impulse += impulse;
        ba.w = sub(ba.w, scale(impulse, g_a));
        bb.w = add(bb.w, scale(impulse, g_b));
      }
      if (any_rolling) {
        Vec3 delta = sym_mv(c.rolling_mass, sub(ba.w, bb.w));
        Vec3 old = cm.rolling_impulse;
        cm.rolling_impulse = add(old, delta);
        float max_impulse = c.rolling_resistance * total_normal;
        float lsq = dot(cm.rolling_impulse, cm.rolling_impulse);
        bool mask = lsq > eps + max_impulse * max_impulse;
        float nrm = max_impulse / (std::sqrt(lsq) + eps);
        float sc = mask ? nrm : 1.0f;
        Vec3 scaled = scale(sc, cm.rolling_impulse);
        if (c.rolling_resistance == 0.0f) scaled = kZero3;
        cm.rolling_impulse = scaled;
        delta = sub(cm.rolling_impulse, old);
        ba.w = sub(ba.w, sym_mv(c.inv_ia, delta));
        bb.w = add(bb.w, sym_mv(c.inv_ib, delta));
      }
      {
        float twist_speed = dot(normal, sub(bb.w, ba.w));
        float max_lambda = c.friction * total_twist_limit;
        float old = cm.twist_impulse;
        float impulse = -(cm.twist_mass * twist_speed);
        impulse = maxf(-max_lambda - old, minf(impulse, max_lambda - old));
        cm.twist_impulse = old + impulse;
        ba.w = sub(ba.w, scale(impulse, sym_mv(c.inv_ia, normal)));
        bb.w = add(bb.w, scale(impulse, sym_mv(c.inv_ib, normal)));
      }
      {
        Vec3 t1 = cm.tangent1, t2 = cm.tangent2;
        Vec3 r_a = cm.center_a, r_b = cm.center_b;
        Vec3 arv = sub(cross(bb.w, r_b), cross(ba.w, r_a));
        Vec2 vt = {ltv1 + dot(arv, t1), ltv2 + dot(arv, t2)};
        Vec2 di = mul_mv2(cm.tangent_mass, vt);
        di = Vec2{-di.x, -di.y};
        Vec2 ni = {cm.friction_impulse.x + di.x, cm.friction_impulse.y + di.y};
        float max_impulse = c.friction * total_normal;
        float lsq = ni.x * ni.x + ni.y * ni.y;
        bool mask = lsq > max_impulse * max_impulse;
        float nrm = max_impulse / (std::sqrt(lsq) + eps);
        float sc = mask ? nrm : 1.0f;
        ni = Vec2{sc * ni.x, sc * ni.y};
        di = Vec2{ni.x - cm.friction_impulse.x, ni.y - cm.friction_impulse.y};
        cm.friction_impulse = ni;
        Vec3 p = scale(di.x, t1);
        p = add(p, scale(di.y, t2));
        ba.w = sub(ba.w, sym_mv(c.inv_ia, cross(r_a, p)));
        bb.w = add(bb.w, sym_mv(c.inv_ib, cross(r_b, p)));
        p = add(p, scale(manifold_impulse, normal));
        ba.v = sub(ba.v, scale(c.inv_mass_a, p));
        bb.v = add(bb.v, scale(c.inv_mass_b, p));
      }
    }
    set_state_at(awake_states, c.index_a, ba);
    set_state_at(awake_states, c.index_b, bb);
  }
}

// The residual sweep (no reference counterpart; the checker of the GPU's per-colour
// residuals): relax_color once more over each colour's contact constraints, in colour order,
// from the state the solver stages end in; the impulse changes it makes are measured and
// then the constraints and body states are put back. Per constraint: the sum (in point
// order) and the largest |normal impulse change| and the length of the friction impulse
// change. The GPU reduces 64 constraints at a time with a lane butterfly (32, 16, 8, 4, 2,
// 1) and adds the partials in double; the same grouping is reproduced here.
void World::colour_residuals(const StepContext& ctx) {
  const std::vector<BodyState> states = awake_states;
  for (int ci = 0; ci < kGraphColorCount - 1; ++ci) {
    const std::vector<ContactConstraint>& list = color_constraints[ci];
    std::vector<ManifoldConstraint> before;
    for (const ContactConstraint& c : list)
      for (int mo = 0; mo < c.manifold_count; ++mo) before.push_back(manifold_pool[c.manifold_start + mo]);
    relax_color(ci, ctx);
    ColourResidual r;
    r.constraints = (int32_t)list.size();
    size_t b = 0;
    for (size_t w0 = 0; w0 < list.size(); w0 += 64) {
      float sum[64];
      for (int l = 0; l < 64; ++l) {
        sum[l] = 0.0f;
        if (w0 + l >= list.size()) continue;
        const ContactConstraint& c = list[w0 + l];
        for (int mo = 0; mo < c.manifold_count; ++mo, ++b) {
          const ManifoldConstraint& old = before[b];
          const ManifoldConstraint& now = manifold_pool[c.manifold_start + mo];
          for (int p = 0; p < now.point_count; ++p) {
            const float d = std::fabs(now.points[p].normal_impulse - old.points[p].normal_impulse);
            sum[l] += d;
            r.max_normal = std::max(r.max_normal, d);
          }
          const float dx = now.friction_impulse.x - old.friction_impulse.x;
          const float dy = now.friction_impulse.y - old.friction_impulse.y;
          r.max_friction = std::max(r.max_friction, std::sqrt(dx * dx + dy * dy));
        }
      }
      for (int off = 32; off >= 1; off >>= 1) {
        float next[64];
        for (int l = 0; l < 64; ++l) next[l] = sum[l] + sum[l ^ off];
        std::copy(next, next + 64, sum);
      }
      r.sum_normal += (double)sum[0];
    }
    b = 0;
    for (const ContactConstraint& c : list)
      for (int mo = 0; mo < c.manifold_count; ++mo) manifold_pool[c.manifold_start + mo] = before[b++];
    residuals[ci] = r;
  }
  awake_states = states;
}

// contact_solver.rs:1135-1229
void World::restitution_color(int ci, const StepContext& ctx) {
  const float inv_h = ctx.inv_h;
  const float neg_threshold = -restitution_threshold;
  const bool propagate = enable_restitution_propagation;
  const std::vector<ContactConstraint>& list_ = color_constraints[ci];
#pragma omp for schedule(static)
  for (size_t k_ = 0; k_ < list_.size(); ++k_) {
    const ContactConstraint& c = list_[k_];
    if (!propagate && c.restitution == 0.0f) continue;
    BodyState ba = state_at(awake_states, c.index_a);
    BodyState bb = state_at(awake_states, c.index_b);
    bool rmask = c.restitution > 0.0f;
    Vec3 dp = sub(bb.dp, ba.dp);
    for (int mo = 0; mo < c.manifold_count; ++mo) {
      ManifoldConstraint& cm = manifold_pool[c.manifold_start + mo];
      Vec3 normal = cm.normal;
      float normal_sep = dot(normal, dp);
      Vec3 n_a = inv_rotate_vector_wide(ba.dq, normal);
      Vec3 n_b = inv_rotate_vector_wide(bb.dq, normal);
      for (int p = 0; p < cm.point_count; ++p) {
        ConstraintPoint& cp = cm.points[p];
        float s = (normal_sep + (dot(n_b, cp.r_b) - dot(n_a, cp.r_a))) + cp.base_separation;
        float normal_mass = propagate ? cp.normal_mass : (rmask ? cp.normal_mass : 0.0f);
        float compression = cp.total_normal_impulse - cp.restitution_impulse;
        bool armed = rmask && cp.relative_velocity < neg_threshold && compression > 0.0f;
        float spec_bias = maxf(0.0f, s * inv_h);
        float velocity_bias = armed ? c.restitution * cp.relative_velocity : spec_bias;
        Vec3 vr_a = add(ba.v, cross(ba.w, cp.r_a));
        Vec3 vr_b = add(bb.v, cross(bb.w, cp.r_b));
        float v_n = dot(sub(vr_b, vr_a), normal);
        float neg_impulse = normal_mass * (v_n + velocity_bias);
        float new_impulse = maxf(cp.normal_impulse - neg_impulse, 0.0f);
        float impulse = new_impulse - cp.normal_impulse;
        float approach = minf(maxf(-normal_mass * v_n, 0.0f), maxf(impulse, 0.0f));
        float allowance = c.restitution * (compression + approach) - cp.restitution_impulse;
        float max_impulse = approach + maxf(allowance, 0.0f);
        impulse = armed ? minf(impulse, max_impulse) : impulse;
        cp.normal_impulse = cp.normal_impulse + impulse;
        cp.restitution_impulse = cp.restitution_impulse + (impulse - approach);
        cp.total_normal_impulse = cp.total_normal_impulse + impulse;
        Vec3 pp = scale(impulse, normal);
        ba.w = sub(ba.w, sym_mv(c.inv_ia, cross(cp.r_a, pp)));
        ba.v = sub(ba.v, scale(c.inv_mass_a, pp));
        bb.w = add(bb.w, sym_mv(c.inv_ib, cross(cp.r_b, pp)));
        bb.v = add(bb.v, scale(c.inv_mass_b, pp));
      }
    }
    set_state_at(awake_states, c.index_a, ba);
    set_state_at(awake_states, c.index_b, bb);
  }
}

// contact_solver.rs:1258-1299
void World::store_color(int ci) {
  const std::vector<ContactConstraint>& list_ = color_constraints[ci];
#pragma omp for schedule(static)
  for (size_t k_ = 0; k_ < list_.size(); ++k_) {
    const ContactConstraint& c = list_[k_];
    Contact& contact = contacts[c.contact_id];
    if (contact.manifold_count <= 0) continue;
    bool hit_events = (contact.flags & CF::SIM_ENABLE_HIT_EVENT) != 0;
    for (int mo = 0; mo < c.manifold_count; ++mo) {
      const ManifoldConstraint& cm = manifold_pool[c.manifold_start + mo];
      Manifold& out = contact.manifold;
      int count = mini(out.point_count, cm.point_count);
      for (int p = 0; p < count; ++p) {
        out.points[p].normal_impulse = cm.points[p].normal_impulse;
        out.points[p].total_normal_impulse = cm.points[p].total_normal_impulse;
        if (hit_events) out.points[p].normal_velocity = cm.points[p].relative_velocity;
      }
      out.twist_impulse = cm.twist_impulse;
      out.friction_impulse = add(scale(cm.friction_impulse.x, cm.tangent1), scale(cm.friction_impulse.y, cm.tangent2));
      out.rolling_impulse = cm.rolling_impulse;
    }
  }
}

static inline void store_overflow_velocity(std::vector<BodyState>& states, int32_t index, uint32_t flags, Vec3 v, Vec3 w) {
  if (index >= 0 && (flags & BF::DYNAMIC_FLAG) != 0) {
    states[index].v = v;
    states[index].w = w;
  }
}

// overflow_contact_solver.rs:278-322
void World::warm_start_overflow() {
  for (const OverflowConstraint& c : overflow_constraints) {
    Vec3 linear = kZero3, ang_a = kZero3, ang_b = kZero3;
    for (int mi = 0; mi < c.manifold_count; ++mi) {
      const ManifoldConstraint& mc = overflow_manifolds[c.manifold_start + mi];
      Vec3 normal = mc.normal;
      float total = 0.0f;
      Vec3 moment_a = kZero3, moment_b = kZero3;
      for (int j = 0; j < mc.point_count; ++j) {
        const ConstraintPoint& cp = mc.points[j];
        total += cp.normal_impulse;
        moment_a = mul_add(moment_a, cp.normal_impulse, cp.r_a);
        moment_b = mul_add(moment_b, cp.normal_impulse, cp.r_b);
      }
      Vec3 fi = scale(mc.friction_impulse.x, mc.tangent1);
      fi = mul_add(fi, mc.friction_impulse.y, mc.tangent2);
      linear = add(linear, mul_add(fi, total, normal));
      Vec3 twist = scale(mc.twist_impulse, normal);
      Vec3 a_a = add(add(cross(moment_a, normal), cross(mc.center_a, fi)), twist);
      Vec3 a_b = add(add(cross(moment_b, normal), cross(mc.center_b, fi)), twist);
      ang_a = add(ang_a, add(a_a, mc.rolling_impulse));
      ang_b = add(ang_b, add(a_b, mc.rolling_impulse));
    }
    if (c.index_a >= 0 && (awake_states[c.index_a].flags & BF::DYNAMIC_FLAG)) {
      BodyState& s = awake_states[c.index_a];
      s.v = mul_sub(s.v, c.inv_mass_a, linear);
      s.w = sub(s.w, mul_mv(c.inv_ia, ang_a));
    }
    if (c.index_b >= 0 && (awake_states[c.index_b].flags & BF::DYNAMIC_FLAG)) {
      BodyState& s = awake_states[c.index_b];
      s.v = mul_add(s.v, c.inv_mass_b, linear);
      s.w = add(s.w, mul_mv(c.inv_ib, ang_b));
    }
  }
}

// overflow_contact_solver.rs:324-399
void World::push_overflow(const StepContext& ctx) {
  const float inv_h = ctx.inv_h;
  for (const OverflowConstraint& c : overflow_constraints) {
    float mA = c.inv_mass_a, mB = c.inv_mass_b;
    Mat3 i_a = c.inv_ia, i_b = c.inv_ib;
    BodyState sa = state_at(awake_states, c.index_a), sb = state_at(awake_states, c.index_b);
    Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;
    Quat dq_a = sa.dq, dq_b = sb.dq;
    Vec3 dp = sub(sb.dp, sa.dp);
    Softness soft = c.softness;
    for (int j = 0; j < c.manifold_count; ++j) {
      ManifoldConstraint& mc = overflow_manifolds[c.manifold_start + j];
      Vec3 normal = mc.normal;
      for (int p = 0; p < mc.point_count; ++p) {
        Vec3 r_a = mc.points[p].r_a, r_b = mc.points[p].r_b;
        Vec3 ds = add(dp, sub(rotate(dq_b, r_b), rotate(dq_a, r_a)));
        float s = dot(ds, normal) + mc.points[p].base_separation;
        float velocity_bias, mass_scale, impulse_scale;
        if (s > 0.0f) {
          velocity_bias = s * inv_h;
          mass_scale = 1.0f;
          impulse_scale = 0.0f;
        } else {
          velocity_bias = maxf(soft.mass_scale * soft.bias_rate * s, -contact_speed);
          mass_scale = soft.mass_scale;
          impulse_scale = soft.impulse_scale;
        }
        Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
        float vn = dot(sub(vr_b, vr_a), normal);
        float ni = mc.points[p].normal_impulse;
        float delta = -mc.points[p].normal_mass * (mass_scale * vn + velocity_bias) - impulse_scale * ni;
        float new_impulse = maxf(ni + delta, 0.0f);
        delta = new_impulse - ni;
        mc.points[p].normal_impulse = new_impulse;
        Vec3 pp = scale(delta, normal);
        v_a = mul_sub(v_a, mA, pp);
        w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
        v_b = mul_add(v_b, mB, pp);
        w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
      }
    }
    store_overflow_velocity(awake_states, c.index_a, sa.flags, v_a, w_a);
    store_overflow_velocity(awake_states, c.index_b, sb.flags, v_b, w_b);
  }
}

// overflow_contact_solver.rs:401-527
void World::relax_overflow(const StepContext& ctx) {
  const float inv_h = ctx.inv_h;
  for (const OverflowConstraint& c : overflow_constraints) {
    float mA = c.inv_mass_a, mB = c.inv_mass_b;
    Mat3 i_a = c.inv_ia, i_b = c.inv_ib;
    BodyState sa = state_at(awake_states, c.index_a), sb = state_at(awake_states, c.index_b);
    Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;
    Quat dq_a = sa.dq, dq_b = sb.dq;
    Vec3 dp = sub(sb.dp, sa.dp);
    float friction = c.friction, rolling = c.rolling_resistance;
    for (int j = 0; j < c.manifold_count; ++j) {
      ManifoldConstraint& mc = overflow_manifolds[c.manifold_start + j];
      Vec3 normal = mc.normal;
      float total_normal = 0.0f, total_twist_limit = 0.0f;
      for (int p = 0; p < mc.point_count; ++p) {
        Vec3 r_a = mc.points[p].r_a, r_b = mc.points[p].r_b;
        Vec3 ds = add(dp, sub(rotate(dq_b, r_b), rotate(dq_a, r_a)));
        float s = dot(ds, normal) + mc.points[p].base_separation;
        float velocity_bias = s > 0.0f ? s * inv_h : 0.0f;
        Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
        float vn = dot(sub(vr_b, vr_a), normal);
        float ni = mc.points[p].normal_impulse;
        float delta = -mc.points[p].normal_mass * (vn + velocity_bias);
        float new_impulse = maxf(ni + delta, 0.0f);
        delta = new_impulse - ni;
        mc.points[p].normal_impulse = new_impulse;
        mc.points[p].total_normal_impulse += new_impulse;
        total_normal += new_impulse;
        total_twist_limit += mc.points[p].lever_arm * new_impulse;
        Vec3 pp = scale(delta, normal);
        v_a = mul_sub(v_a, mA, pp);
        w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
        v_b = mul_add(v_b, mB, pp);
        w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
      }
      {
        float twist_speed = dot(normal, sub(w_b, w_a));
        float max_impulse = friction * total_twist_limit;
        float delta = -mc.twist_mass * twist_speed;
        float old = mc.twist_impulse;
        mc.twist_impulse = clampf(old + delta, -max_impulse, max_impulse);
        delta = mc.twist_impulse - old;
        w_a = sub(w_a, mul_mv(i_a, scale(delta, normal)));
        w_b = add(w_b, mul_mv(i_b, scale(delta, normal)));
      }
      if (rolling > 0.0f) {
        Vec3 delta = neg(mul_mv(c.rolling_mass, sub(w_b, w_a)));
        Vec3 old = mc.rolling_impulse;
        mc.rolling_impulse = add(old, delta);
        float max_impulse = rolling * total_normal;
        float mag_sq = dot(mc.rolling_impulse, mc.rolling_impulse);
        if (mag_sq > max_impulse * max_impulse + FLT_EPSILON)
          mc.rolling_impulse = scale(max_impulse / std::sqrt(mag_sq), mc.rolling_impulse);
        delta = sub(mc.rolling_impulse, old);
        w_a = sub(w_a, mul_mv(i_a, delta));
        w_b = add(w_b, mul_mv(i_b, delta));
      }
      {
        Vec3 t1 = mc.tangent1, t2 = mc.tangent2, r_a = mc.center_a, r_b = mc.center_b;
        Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
        Vec3 vr = sub(vr_b, vr_a);
        Vec2 vt = {dot(vr, t1) - mc.tangent_velocity1, dot(vr, t2) - mc.tangent_velocity2};
        Vec2 tm = mul_mv2(mc.tangent_mass, vt);
        Vec2 di = {-tm.x, -tm.y};
        Vec2 ni = {mc.friction_impulse.x + di.x, mc.friction_impulse.y + di.y};
        float max_impulse = friction * total_normal;
        float lsq = ni.x * ni.x + ni.y * ni.y;
        if (lsq > max_impulse * max_impulse) {
          float sc = max_impulse / std::sqrt(lsq);
          ni.x *= sc;
          ni.y *= sc;
        }
        di = Vec2{ni.x - mc.friction_impulse.x, ni.y - mc.friction_impulse.y};
        mc.friction_impulse = ni;
        Vec3 pp = blend2(di.x, t1, di.y, t2);
        v_a = mul_sub(v_a, mA, pp);
        w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
        v_b = mul_add(v_b, mB, pp);
        w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
      }
    }
    store_overflow_velocity(awake_states, c.index_a, sa.flags, v_a, w_a);
    store_overflow_velocity(awake_states, c.index_b, sb.flags, v_b, w_b);
  }
}

// overflow_contact_solver.rs:529-630
void World::restitution_overflow(const StepContext& ctx) {
  const float threshold = restitution_threshold, inv_h = ctx.inv_h;
  const bool propagate = enable_restitution_propagation;
  for (const OverflowConstraint& c : overflow_constraints) {
    float restitution = c.restitution;
    if (!propagate && restitution == 0.0f) continue;
    float mA = c.inv_mass_a, mB = c.inv_mass_b;
    Mat3 i_a = c.inv_ia, i_b = c.inv_ib;
    BodyState sa = state_at(awake_states, c.index_a), sb = state_at(awake_states, c.index_b);
    Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;
    Quat dq_a = sa.dq, dq_b = sb.dq;
    Vec3 dp = sub(sb.dp, sa.dp);
    for (int j = 0; j < c.manifold_count; ++j) {
      ManifoldConstraint& mc = overflow_manifolds[c.manifold_start + j];
      Vec3 normal = mc.normal;
      for (int p = 0; p < mc.point_count; ++p) {
        ConstraintPoint cp = mc.points[p];
        Vec3 r_a = cp.r_a, r_b = cp.r_b;
        float compression = cp.total_normal_impulse - cp.restitution_impulse;
        bool armed = restitution > 0.0f && cp.relative_velocity < -threshold && compression > 0.0f;
        float velocity_bias;
        if (armed) {
          velocity_bias = restitution * cp.relative_velocity;
        } else {
          Vec3 ds = add(dp, sub(rotate(dq_b, r_b), rotate(dq_a, r_a)));
          float s = dot(ds, normal) + cp.base_separation;
          velocity_bias = s > 0.0f ? s * inv_h : 0.0f;
        }
        Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
        float vn = dot(sub(vr_b, vr_a), normal);
        float impulse = -cp.normal_mass * (vn + velocity_bias);
        float new_impulse = maxf(cp.normal_impulse + impulse, 0.0f);
        impulse = new_impulse - cp.normal_impulse;
        float approach = minf(maxf(-cp.normal_mass * vn, 0.0f), maxf(impulse, 0.0f));
        if (armed) {
          float allowance = restitution * (compression + approach) - cp.restitution_impulse;
          impulse = minf(impulse, approach + maxf(allowance, 0.0f));
        }
        mc.points[p].normal_impulse += impulse;
        mc.points[p].restitution_impulse += impulse - approach;
        mc.points[p].total_normal_impulse += impulse;
        Vec3 pp = scale(impulse, normal);
        v_a = mul_sub(v_a, mA, pp);
        w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
        v_b = mul_add(v_b, mB, pp);
        w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
      }
    }
    store_overflow_velocity(awake_states, c.index_a, sa.flags, v_a, w_a);
    store_overflow_velocity(awake_states, c.index_b, sb.flags, v_b, w_b);
  }
}

// overflow_contact_solver.rs:632-663
void World::store_overflow() {
  for (const OverflowConstraint& c : overflow_constraints) {
    if (c.manifold_count == 0) continue;
    Contact& contact = contacts[c.contact_id];
    for (int mi = 0; mi < c.manifold_count; ++mi) {
      Manifold& m = contact.manifold;
      const ManifoldConstraint& mc = overflow_manifolds[c.manifold_start + mi];
      m.twist_impulse = mc.twist_impulse;
      m.friction_impulse = blend2(mc.friction_impulse.x, mc.tangent1, mc.friction_impulse.y, mc.tangent2);
      m.rolling_impulse = mc.rolling_impulse;
      for (int p = 0; p < mc.point_count; ++p) {
        m.points[p].normal_impulse = mc.points[p].normal_impulse;
        m.points[p].total_normal_impulse = mc.points[p].total_normal_impulse;
      }
    }
  }
}

// solver_stages.rs:543-589
// One parallel region: every thread walks the schedule; each stage's loop over one
// colour (or over bodies) is shared out with an implicit barrier at its end, and the
// overflow list runs on one thread.
// solver_stages.rs:541-589: per substep, overflow joints then overflow contacts, then each
// colour's joints and contacts (disjoint bodies within a colour).
void World::solve_stages(const StepContext& ctx) {
#pragma omp parallel
  {
    for (int sub_i = 0; sub_i < ctx.sub_step_count; ++sub_i) {
      integrate_velocities(ctx.h);
#pragma omp single
      {
        warm_start_overflow_joints();
        warm_start_overflow();
      }
      for (int ci = 0; ci < kGraphColorCount - 1; ++ci) {
        warm_start_joints(ci);
        warm_start_color(ci);
      }
#pragma omp single
      {
        solve_overflow_joints(ctx, true);
        push_overflow(ctx);
      }
      for (int ci = 0; ci < kGraphColorCount - 1; ++ci) {
        solve_joints(ci, ctx, true);
        push_color(ci, ctx);
      }
      integrate_positions(ctx.h, ctx.inv_dt);
#pragma omp single
      {
        solve_overflow_joints(ctx, false);
        relax_overflow(ctx);
      }
      for (int ci = 0; ci < kGraphColorCount - 1; ++ci) {
        solve_joints(ci, ctx, false);
        relax_color(ci, ctx);
      }
    }
    if (solve_any_restitution && restitution_iterations > 0) {
      for (int it = 0; it < restitution_iterations; ++it) {
#pragma omp single
        restitution_overflow(ctx);
        for (int ci = 0; ci < kGraphColorCount - 1; ++ci) restitution_color(ci, ctx);
      }
    }
#pragma omp single
    store_overflow();
    for (int ci = 0; ci < kGraphColorCount - 1; ++ci) store_color(ci);
  }
}

}  // namespace nslref
