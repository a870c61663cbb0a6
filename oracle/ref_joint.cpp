// Oracle: revolute joints (creation, graph colour, prepare, warm start, solve).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Nothing under natsel_amd/
// includes or links this code.
//
// Follows:
//   create_joint                joint.rs:164-375 (awake and static sets; no sleeping sets)
//   destroy_contacts_between    joint.rs:130-161
//   should_bodies_collide       contact_creation.rs:110-143
//   assign_joint_color          constraint_graph.rs:256-298
//   create_joint_in_graph       constraint_graph.rs:301-323
//   prepare_joint               distance_joint_solver.rs:76-101
//   prepare_revolute_joint      revolute_joint.rs:176-261
//   warm_start_revolute_joint   revolute_joint.rs:263-311
//   solve_revolute_joint        revolute_joint.rs:313-552
//   solve_point_constraint      revolute_joint.rs:100-145
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "ref_world.h"

namespace nslref {

namespace {

// functions.rs:68-97
float atan2_approx(float y, float x) {
  if (x == 0.0f && y == 0.0f) return 0.0f;
  float ax = absf(x), ay = absf(y);
  float mx = maxf(ay, ax), mn = minf(ay, ax);
  float a = mn / mx;
  float s = a * a;
  float c = s * a;
  float q = s * s;
  float r = 0.024840285f * q + 0.18681418f;
  float t = -0.094097948f * q - 0.33213072f;
  r = r * s + t;
  r = r * c + a;
  if (ay > ax) r = 1.57079637f - r;
  if (x < 0.0f) r = 3.14159274f - r;
  if (y < 0.0f) r = -r;
  return r;
}

// functions.rs:490-499
float get_twist_angle(Quat q) {
  float twist = q.s < 0.0f ? atan2_approx(-q.v.z, -q.s) : atan2_approx(q.v.z, q.s);
  return twist * 2.0f;
}

// internal_math.rs:191-202
Vec2 solve2(Mat2 m, Vec2 b) {
  float d = det2(m);
  if (d > 1000.0f * FLT_MIN) {
    float inv_det = 1.0f / d;
    return Vec2{inv_det * m.cy.y * b.x - inv_det * m.cy.x * b.y, -inv_det * m.cx.y * b.x + inv_det * m.cx.x * b.y};
  }
  return Vec2{0.0f, 0.0f};
}

Quat negate_quat(Quat q) { return Quat{{-q.v.x, -q.v.y, -q.v.z}, -q.s}; }

// revolute_joint.rs:100-145
Vec3 solve_point_constraint(float m_a, float m_b, Mat3 i_a, Mat3 i_b, Vec3 r_a, Vec3 r_b, Vec3 rhs) {
  float axy = r_a.z * i_a.cy.y - r_a.y * i_a.cz.y;
  float axz = r_a.z * i_a.cy.z - r_a.y * i_a.cz.z;
  Vec3 ay = sub(scale(r_a.x, i_a.cz), scale(r_a.z, i_a.cx));
  Vec3 az = sub(scale(r_a.y, i_a.cx), scale(r_a.x, i_a.cy));

  float bxy = r_b.z * i_b.cy.y - r_b.y * i_b.cz.y;
  float bxz = r_b.z * i_b.cy.z - r_b.y * i_b.cz.z;
  Vec3 by = sub(scale(r_b.x, i_b.cz), scale(r_b.z, i_b.cx));
  Vec3 bz = sub(scale(r_b.y, i_b.cx), scale(r_b.x, i_b.cy));

  float m = m_a + m_b;
  float kxx = ((axy * r_a.z - axz * r_a.y) + (bxy * r_b.z - bxz * r_b.y)) + m;
  float kyy = ((ay.z * r_a.x - ay.x * r_a.z) + (by.z * r_b.x - by.x * r_b.z)) + m;
  float kzz = ((az.x * r_a.y - az.y * r_a.x) + (bz.x * r_b.y - bz.y * r_b.x)) + m;
  float kxy = (ay.y * r_a.z - ay.z * r_a.y) + (by.y * r_b.z - by.z * r_b.y);
  float kxz = (az.y * r_a.z - az.z * r_a.y) + (bz.y * r_b.z - bz.z * r_b.y);
  float kyz = (az.z * r_a.x - az.x * r_a.z) + (bz.z * r_b.x - bz.x * r_b.z);

  float cxx = kyy * kzz - kyz * kyz;
  float cyy = kxx * kzz - kxz * kxz;
  float czz = kxx * kyy - kxy * kxy;
  float cxy = kxz * kyz - kxy * kzz;
  float cxz = kxy * kyz - kyy * kxz;
  float cyz = kxy * kxz - kxx * kyz;

  float determinant = (kxx * cxx + kxy * cxy) + kxz * cxz;
  if (absf(determinant) > 1000.0f * FLT_MIN) {
    float inv_det = 1.0f / determinant;
    return Vec3{inv_det * ((cxx * rhs.x + cxy * rhs.y) + cxz * rhs.z), inv_det * ((cxy * rhs.x + cyy * rhs.y) + cyz * rhs.z),
                inv_det * ((cxz * rhs.x + cyz * rhs.y) + czz * rhs.z)};
  }
  return kZero3;
}

BodyState identity_state() {
  BodyState s{};
  s.dq = kQuatIdentity;
  return s;
}
BodyState state_at(const std::vector<BodyState>& states, int32_t index) {
  return index >= 0 ? states[index] : identity_state();
}
void set_state_at(std::vector<BodyState>& states, int32_t index, const BodyState& s) {
  if (index >= 0 && (s.flags & BF::DYNAMIC_FLAG)) states[index] = s;
}

// 0.5 * rotate_vector(q, rel.s * e + cross(rel.v, e)) for a unit axis e (revolute_joint.rs:239-250)
Vec3 perp_axis(Quat q, Quat rel, Vec3 e) { return scale(0.5f, rotate(q, add(scale(rel.s, e), cross(rel.v, e)))); }

constexpr Vec3 kAxisX{1.0f, 0.0f, 0.0f}, kAxisY{0.0f, 1.0f, 0.0f}, kAxisZ{0.0f, 0.0f, 1.0f};

}  // namespace

// contact_creation.rs:110-143
bool World::should_bodies_collide(const Body& a, const Body& b) const {
  if (a.type != kDynamic && b.type != kDynamic) return false;
  int32_t key, other;
  if (a.joint_count < b.joint_count) {
    key = a.head_joint_key;
    other = b.id;
  } else {
    key = b.head_joint_key;
    other = a.id;
  }
  while (key != kNullIndex) {
    int32_t jid = key >> 1, edge = key & 1;
    const Joint& j = joints[jid];
    if (!j.collide_connected && j.edges[edge ^ 1].body_id == other) return false;
    key = j.edges[edge].next_key;
  }
  return true;
}

// joint.rs:130-161. The reference walks the contact list of the body with fewer contacts;
// that list is head-inserted, so the contacts between the two bodies go newest first.
void World::destroy_contacts_between_bodies(int32_t body_a, int32_t body_b) {
  std::vector<std::pair<uint64_t, int32_t>> between;
  for (const Contact& c : contacts) {
    if (c.contact_id == kNullIndex) continue;
    int32_t a = c.edges[0].body_id, b = c.edges[1].body_id;
    if ((a == body_a && b == body_b) || (a == body_b && b == body_a)) between.push_back({c.created, c.contact_id});
  }
  std::sort(between.begin(), between.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  for (const auto& e : between) destroy_contact(e.second);
}

// constraint_graph.rs:256-298
int32_t World::assign_joint_color(int32_t a, int32_t b, BodyType ta, BodyType tb) {
  const int kDynamicColorCount = kGraphColorCount - 4;
  const int kOverflow = kGraphColorCount - 1;
  if (ta == kDynamic && tb == kDynamic) {
    for (int i = 0; i < kDynamicColorCount; ++i) {
      if (colors[i].body_set.get(a) || colors[i].body_set.get(b)) continue;
      colors[i].body_set.set_grow(a);
      colors[i].body_set.set_grow(b);
      return i;
    }
  } else if (ta == kDynamic) {
    for (int i = kOverflow - 1; i >= 1; --i)
      if (!colors[i].body_set.get(a)) { colors[i].body_set.set_grow(a); return i; }
  } else if (tb == kDynamic) {
    for (int i = kOverflow - 1; i >= 1; --i)
      if (!colors[i].body_set.get(b)) { colors[i].body_set.set_grow(b); return i; }
  }
  return kOverflow;
}

// joint.rs:164-375 + revolute_joint.rs:149-174
int32_t World::create_revolute_joint(const RevoluteJointDef& def) {
  const int32_t a = def.body_a, b = def.body_b;
  if (a == b || a < 0 || b < 0 || a >= (int32_t)bodies.size() || b >= (int32_t)bodies.size()) return kNullIndex;
  // a joint with no awake body but a sleeping one stays asleep in the reference (joint.rs:294-
  // 340, merging sleeping sets); not restated: refused
  if (bodies[a].set_index != 2 && bodies[b].set_index != 2 && (bodies[a].set_index >= 3 || bodies[b].set_index >= 3))
    return kNullIndex;
  if (!def.collide_connected) destroy_contacts_between_bodies(a, b);

  int32_t jid = joint_pool.alloc();
  if (jid == (int32_t)joints.size()) joints.push_back(Joint{});
  Joint j;
  j.joint_id = jid;
  j.collide_connected = def.collide_connected;
  Body& ba = bodies[a];
  j.edges[0] = JointEdge{a, kNullIndex, ba.head_joint_key};
  const int32_t key_a = jid << 1;
  if (ba.head_joint_key != kNullIndex) joints[ba.head_joint_key >> 1].edges[ba.head_joint_key & 1].prev_key = key_a;
  ba.head_joint_key = key_a;
  ba.joint_count += 1;
  Body& bb = bodies[b];
  j.edges[1] = JointEdge{b, kNullIndex, bb.head_joint_key};
  const int32_t key_b = (jid << 1) | 1;
  if (bb.head_joint_key != kNullIndex) joints[bb.head_joint_key >> 1].edges[bb.head_joint_key & 1].prev_key = key_b;
  bb.head_joint_key = key_b;
  bb.joint_count += 1;

  JointSim sim;
  sim.joint_id = jid;
  sim.body_id_a = a;
  sim.body_id_b = b;
  sim.local_frame_a = def.local_frame_a;
  sim.local_frame_b = def.local_frame_b;
  sim.constraint_hertz = def.constraint_hertz;
  sim.constraint_damping_ratio = def.constraint_damping_ratio;
  sim.constraint_softness = Softness{0.0f, 1.0f, 0.0f};
  sim.hertz = def.hertz;
  sim.damping_ratio = def.damping_ratio;
  sim.max_motor_torque = def.max_motor_torque;
  sim.motor_speed = def.motor_speed;
  sim.target_angle = def.target_angle;
  sim.lower_angle = def.lower_angle;
  sim.upper_angle = def.upper_angle;
  sim.enable_spring = def.enable_spring;
  sim.enable_motor = def.enable_motor;
  sim.enable_limit = def.enable_limit;

  if (bodies[a].type != kDynamic && bodies[b].type != kDynamic) {
    j.set_index = 0;  // static set
    j.local_index = (int32_t)static_joint_sims.size();
    static_joint_sims.push_back(sim);
  } else {
    // joint.rs:272-293: wake a sleeping partner, then into the graph
    const int32_t max_set = std::max(bodies[a].set_index, bodies[b].set_index);
    joints[jid] = j;
    if (max_set >= 3) wake_solver_set(max_set);
    j.set_index = 2;
    int32_t color = assign_joint_color(a, b, bodies[a].type, bodies[b].type);
    j.color_index = color;
    j.local_index = (int32_t)colors[color].joint_sims.size();
    colors[color].joint_sims.push_back(sim);
  }
  joints[jid] = j;
  if (j.set_index > 1) link_joint(jid);  // joint.rs:366-370
  return jid;
}

const JointSim& World::joint_sim(int32_t jid) const {
  const Joint& j = joints[jid];
  if (j.set_index == 2) return colors[j.color_index].joint_sims[j.local_index];
  if (j.set_index >= 3) return sleep_sets[(size_t)j.set_index].joint_sims[(size_t)j.local_index];
  return static_joint_sims[j.local_index];
}

// revolute_joint.rs:176-261 (body data from solver_prepare.rs:217-241)
void World::prepare_revolute_joint(JointSim& base, const StepContext& ctx) {
  const Body& body_a = bodies[base.body_id_a];
  const Body& body_b = bodies[base.body_id_b];
  const BodySim& sim_a = sim_of(body_a);
  const BodySim& sim_b = sim_of(body_b);

  base.inv_mass_a = sim_a.inv_mass;
  base.inv_mass_b = sim_b.inv_mass;
  base.inv_ia = sim_a.inv_inertia_world;
  base.inv_ib = sim_b.inv_inertia_world;

  Mat3 inv_inertia_sum = add_mm(base.inv_ia, base.inv_ib);
  base.fixed_rotation = det(inv_inertia_sum) < 1000.0f * FLT_MIN;

  JointSim& j = base;
  j.index_a = body_a.set_index == 2 ? body_a.local_index : kNullIndex;
  j.index_b = body_b.set_index == 2 ? body_b.local_index : kNullIndex;

  j.frame_a.q = mul_quat(sim_a.transform.q, base.local_frame_a.q);
  j.frame_a.p = rotate(sim_a.transform.q, sub(base.local_frame_a.p, sim_a.local_center));
  j.frame_b.q = mul_quat(sim_b.transform.q, base.local_frame_b.q);
  j.frame_b.p = rotate(sim_b.transform.q, sub(base.local_frame_b.p, sim_b.local_center));

  j.delta_center = sub_pos(sim_b.center, sim_a.center);

  Vec3 axis = rotate(j.frame_a.q, kAxisZ);
  float k = dot(axis, mul_mv(inv_inertia_sum, axis));
  j.axial_mass = k > 0.0f ? 1.0f / k : 0.0f;
  j.rotation_axis_z = axis;

  Quat rel_q = inv_mul_quat(j.frame_a.q, j.frame_b.q);
  j.perp_axis_x = perp_axis(j.frame_a.q, rel_q, kAxisX);
  j.perp_axis_y = perp_axis(j.frame_a.q, rel_q, kAxisY);

  j.spring_softness = make_soft(j.hertz, j.damping_ratio, ctx.h);

  if (!ctx.enable_warm_starting) {
    j.linear_impulse = kZero3;
    j.perp_impulse = Vec2{0.0f, 0.0f};
    j.motor_impulse = 0.0f;
    j.spring_impulse = 0.0f;
    j.lower_impulse = 0.0f;
    j.upper_impulse = 0.0f;
  }
}

// solver_prepare.rs:178-212 + distance_joint_solver.rs:76-101: every colour including overflow
void World::prepare_joints(const StepContext& ctx) {
  for (int ci = 0; ci < kGraphColorCount; ++ci) {
    for (JointSim& j : colors[ci].joint_sims) {
      float hertz = minf(j.constraint_hertz, 0.25f * ctx.inv_h);
      j.constraint_softness = make_soft(hertz, j.constraint_damping_ratio, ctx.h);
      prepare_revolute_joint(j, ctx);
    }
  }
}

// revolute_joint.rs:263-311
void World::warm_start_joint(JointSim& j) {
  {
    const float m_a = j.inv_mass_a, m_b = j.inv_mass_b;
    const Mat3 i_a = j.inv_ia, i_b = j.inv_ib;
    BodyState sa = state_at(awake_states, j.index_a);
    BodyState sb = state_at(awake_states, j.index_b);
    Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;

    Vec3 r_a = rotate(sa.dq, j.frame_a.p);
    Vec3 r_b = rotate(sb.dq, j.frame_b.p);

    float axial_impulse = ((j.spring_impulse + j.motor_impulse) + j.lower_impulse) - j.upper_impulse;
    Vec3 angular_impulse = add(scale(j.perp_impulse.x, j.perp_axis_x), scale(j.perp_impulse.y, j.perp_axis_y));
    angular_impulse = mul_add(angular_impulse, axial_impulse, j.rotation_axis_z);

    v_a = sub(v_a, scale(m_a, j.linear_impulse));
    w_a = sub(w_a, mul_mv(i_a, add(cross(r_a, j.linear_impulse), angular_impulse)));
    v_b = add(v_b, scale(m_b, j.linear_impulse));
    w_b = add(w_b, mul_mv(i_b, add(cross(r_b, j.linear_impulse), angular_impulse)));

    if (sa.flags & BF::DYNAMIC_FLAG) { sa.v = v_a; sa.w = w_a; }
    if (sb.flags & BF::DYNAMIC_FLAG) { sb.v = v_b; sb.w = w_b; }
    set_state_at(awake_states, j.index_a, sa);
    set_state_at(awake_states, j.index_b, sb);
  }
}

// revolute_joint.rs:313-552
void World::solve_joint(JointSim& j, const StepContext& ctx, bool use_bias) {
  {
    const float m_a = j.inv_mass_a, m_b = j.inv_mass_b;
    const Mat3 i_a = j.inv_ia, i_b = j.inv_ib;
    BodyState sa = state_at(awake_states, j.index_a);
    BodyState sb = state_at(awake_states, j.index_b);
    Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;

    const bool fixed_rotation = j.fixed_rotation;
    Quat quat_a = mul_quat(sa.dq, j.frame_a.q);
    Quat quat_b = mul_quat(sb.dq, j.frame_b.q);
    if (dot_quat(quat_a, quat_b) < 0.0f) quat_b = negate_quat(quat_b);
    Quat rel_q = inv_mul_quat(quat_a, quat_b);

    if (!fixed_rotation && (j.enable_spring || j.enable_motor || j.enable_limit)) {
      Vec3 axis = j.rotation_axis_z;
      Vec3 g_a = mul_mv(i_a, axis);
      Vec3 g_b = mul_mv(i_b, axis);
      float k = dot(axis, add(g_a, g_b));
      float axial_c_dot = dot(sub(w_b, w_a), axis);
      float axial_impulse = 0.0f;

      if (j.enable_spring) {
        float angle = get_twist_angle(rel_q);
        float c = angle - j.target_angle;
        float bias = j.spring_softness.bias_rate * c;
        float mass_scale = j.spring_softness.mass_scale;
        float impulse_scale = j.spring_softness.impulse_scale;
        float delta = -mass_scale * j.axial_mass * (axial_c_dot + bias) - impulse_scale * j.spring_impulse;
        j.spring_impulse += delta;
        axial_c_dot += k * delta;
        axial_impulse += delta;
      }

      if (j.enable_motor) {
        float cdot = axial_c_dot - j.motor_speed;
        float old_impulse = j.motor_impulse;
        float max_impulse = j.max_motor_torque * ctx.h;
        float delta = clampf(-j.axial_mass * cdot, -max_impulse - old_impulse, max_impulse - old_impulse);
        j.motor_impulse = old_impulse + delta;
        axial_c_dot += k * delta;
        axial_impulse += delta;
      }

      if (j.enable_limit) {
        float angle = get_twist_angle(rel_q);
        {
          float c = angle - j.lower_angle;
          float bias = 0.0f, mass_scale = 1.0f, impulse_scale = 0.0f;
          if (c > 0.0f) {
            bias = c * ctx.inv_h;
          } else if (use_bias) {
            bias = j.constraint_softness.bias_rate * c;
            mass_scale = j.constraint_softness.mass_scale;
            impulse_scale = j.constraint_softness.impulse_scale;
          }
          float old_impulse = j.lower_impulse;
          float delta = maxf(-mass_scale * j.axial_mass * (axial_c_dot + bias) - impulse_scale * old_impulse, -old_impulse);
          j.lower_impulse = old_impulse + delta;
          axial_c_dot += k * delta;
          axial_impulse += delta;
        }
        {
          float c = j.upper_angle - angle;
          float bias = 0.0f, mass_scale = 1.0f, impulse_scale = 0.0f;
          if (c > 0.0f) {
            bias = c * ctx.inv_h;
          } else if (use_bias) {
            bias = j.constraint_softness.bias_rate * c;
            mass_scale = j.constraint_softness.mass_scale;
            impulse_scale = j.constraint_softness.impulse_scale;
          }
          float cdot = -axial_c_dot;
          float old_impulse = j.upper_impulse;
          float delta = maxf(-mass_scale * j.axial_mass * (cdot + bias) - impulse_scale * old_impulse, -old_impulse);
          j.upper_impulse = old_impulse + delta;
          axial_impulse -= delta;
        }
      }

      w_a = sub(w_a, scale(axial_impulse, g_a));
      w_b = add(w_b, scale(axial_impulse, g_b));
    }

    if (!fixed_rotation) {
      Vec2 bias{0.0f, 0.0f};
      float mass_scale = 1.0f, impulse_scale = 0.0f;
      if (use_bias) {
        Vec2 c{rel_q.v.x, rel_q.v.y};
        bias = Vec2{j.constraint_softness.bias_rate * c.x, j.constraint_softness.bias_rate * c.y};
        mass_scale = j.constraint_softness.mass_scale;
        impulse_scale = j.constraint_softness.impulse_scale;
      }
      Vec3 pax = perp_axis(quat_a, rel_q, kAxisX);
      Vec3 pay = perp_axis(quat_a, rel_q, kAxisY);
      j.perp_axis_x = pax;
      j.perp_axis_y = pay;

      Mat3 inv_inertia_sum = add_mm(i_a, i_b);
      float kxx = dot(pax, mul_mv(inv_inertia_sum, pax));
      float kyy = dot(pay, mul_mv(inv_inertia_sum, pay));
      float kxy = dot(pax, mul_mv(inv_inertia_sum, pay));
      Mat2 k{{kxx, kxy}, {kxy, kyy}};

      Vec3 w_rel = sub(w_b, w_a);
      Vec2 cdot{dot(w_rel, pax), dot(w_rel, pay)};
      Vec2 old_impulse = j.perp_impulse;
      Vec2 sol = solve2(k, Vec2{cdot.x + bias.x, cdot.y + bias.y});
      Vec2 delta{(-mass_scale) * sol.x - impulse_scale * old_impulse.x, (-mass_scale) * sol.y - impulse_scale * old_impulse.y};
      j.perp_impulse = Vec2{j.perp_impulse.x + delta.x, j.perp_impulse.y + delta.y};

      Vec3 angular_impulse = add(scale(delta.x, pax), scale(delta.y, pay));
      w_a = sub(w_a, mul_mv(i_a, angular_impulse));
      w_b = add(w_b, mul_mv(i_b, angular_impulse));
    }

    {
      Vec3 r_a = rotate(sa.dq, j.frame_a.p);
      Vec3 r_b = rotate(sb.dq, j.frame_b.p);
      Vec3 cdot = sub(sub(add(v_b, cross(w_b, r_b)), v_a), cross(w_a, r_a));

      Vec3 bias = kZero3;
      float mass_scale = 1.0f, impulse_scale = 0.0f;
      if (use_bias) {
        Vec3 dc_a = sa.dp, dc_b = sb.dp;
        Vec3 separation = add(add(sub(dc_b, dc_a), sub(r_b, r_a)), j.delta_center);
        bias = scale(j.constraint_softness.bias_rate, separation);
        mass_scale = j.constraint_softness.mass_scale;
        impulse_scale = j.constraint_softness.impulse_scale;
      }
      Vec3 b = solve_point_constraint(m_a, m_b, j.inv_ia, j.inv_ib, r_a, r_b, add(cdot, bias));
      Vec3 impulse = sub(scale(-mass_scale, b), scale(impulse_scale, j.linear_impulse));
      j.linear_impulse = add(j.linear_impulse, impulse);

      v_a = sub(v_a, scale(m_a, impulse));
      w_a = sub(w_a, mul_mv(i_a, cross(r_a, impulse)));
      v_b = add(v_b, scale(m_b, impulse));
      w_b = add(w_b, mul_mv(i_b, cross(r_b, impulse)));
    }

    if (sa.flags & BF::DYNAMIC_FLAG) { sa.v = v_a; sa.w = w_a; }
    if (sb.flags & BF::DYNAMIC_FLAG) { sb.v = v_b; sb.w = w_b; }
    set_state_at(awake_states, j.index_a, sa);
    set_state_at(awake_states, j.index_b, sb);
  }
}

// distance_joint_solver.rs:232-323: a colour's joints in list order
void World::warm_start_joints(int ci) {
  std::vector<JointSim>& list = colors[ci].joint_sims;
#pragma omp for schedule(static)
  for (size_t t = 0; t < list.size(); ++t) warm_start_joint(list[t]);
}

void World::solve_joints(int ci, const StepContext& ctx, bool use_bias) {
  std::vector<JointSim>& list = colors[ci].joint_sims;
#pragma omp for schedule(static)
  for (size_t t = 0; t < list.size(); ++t) solve_joint(list[t], ctx, use_bias);
}

// the overflow colour's joints run serially (solver_stages.rs:553, :559, :568)
void World::warm_start_overflow_joints() {
  for (JointSim& j : colors[kGraphColorCount - 1].joint_sims) warm_start_joint(j);
}

void World::solve_overflow_joints(const StepContext& ctx, bool use_bias) {
  for (JointSim& j : colors[kGraphColorCount - 1].joint_sims) solve_joint(j, ctx, use_bias);
}

}  // namespace nslref
