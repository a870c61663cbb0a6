// Per-joint device functions of the step kernel: the revolute joint's prepare, warm start
// and solve (point, collinearity, spring, motor and limit rows).
//
// Every expression keeps the reference's operand order and association (the step is
// bit-deterministic and the GPU must reproduce it):
//   prepare_joint               distance_joint_solver.rs:76-101
//   prepare_revolute_joint      revolute_joint.rs:176-261
//   warm_start_revolute_joint   revolute_joint.rs:263-311
//   solve_revolute_joint        revolute_joint.rs:313-552
//   solve_point_constraint      revolute_joint.rs:100-145
//   atan2_approx, get_twist_angle  functions.rs:68-97, :490-499
//   solve2                      internal_math.rs:191-202
#pragma once
#include "nsl_device.h"

namespace nsl {
namespace {

__device__ __forceinline__ float atan2_approx(float y, float x) {
  if (x == 0.0f && y == 0.0f) return 0.0f;
  const float ax = absf(x), ay = absf(y);
  const float mx = maxf(ay, ax), mn = minf(ay, ax);
  const float a = mn / mx;
  const float s = a * a;
  const float c = s * a;
  const float q = s * s;
  float r = 0.024840285f * q + 0.18681418f;
  const float t = -0.094097948f * q - 0.33213072f;
  r = r * s + t;
  r = r * c + a;
  if (ay > ax) r = 1.57079637f - r;
  if (x < 0.0f) r = 3.14159274f - r;
  if (y < 0.0f) r = -r;
  return r;
}

__device__ __forceinline__ float get_twist_angle(Quat q) {
  const float twist = q.s < 0.0f ? atan2_approx(-q.v.z, -q.s) : atan2_approx(q.v.z, q.s);
  return twist * 2.0f;
}

__device__ __forceinline__ Vec2 solve2(Mat2 m, Vec2 b) {
  const float d = m.cx.x * m.cy.y - m.cx.y * m.cy.x;
  if (d > 1000.0f * kFltMin) {
    const float inv_det = 1.0f / d;
    return Vec2{inv_det * m.cy.y * b.x - inv_det * m.cy.x * b.y, -inv_det * m.cx.y * b.x + inv_det * m.cx.x * b.y};
  }
  return Vec2{0.0f, 0.0f};
}

// The inverse of the 3x3 point-constraint mass applied to rhs, by cofactors.
__device__ __forceinline__ Vec3 solve_point_constraint(float m_a, float m_b, const Mat3& i_a, const Mat3& i_b, Vec3 r_a, Vec3 r_b,
                                                       Vec3 rhs) {
  const float axy = r_a.z * i_a.cy.y - r_a.y * i_a.cz.y;
  const float axz = r_a.z * i_a.cy.z - r_a.y * i_a.cz.z;
  const Vec3 ay = sub(scale(r_a.x, i_a.cz), scale(r_a.z, i_a.cx));
  const Vec3 az = sub(scale(r_a.y, i_a.cx), scale(r_a.x, i_a.cy));

  const float bxy = r_b.z * i_b.cy.y - r_b.y * i_b.cz.y;
  const float bxz = r_b.z * i_b.cy.z - r_b.y * i_b.cz.z;
  const Vec3 by = sub(scale(r_b.x, i_b.cz), scale(r_b.z, i_b.cx));
  const Vec3 bz = sub(scale(r_b.y, i_b.cx), scale(r_b.x, i_b.cy));

  const float m = m_a + m_b;
  const float kxx = ((axy * r_a.z - axz * r_a.y) + (bxy * r_b.z - bxz * r_b.y)) + m;
  const float kyy = ((ay.z * r_a.x - ay.x * r_a.z) + (by.z * r_b.x - by.x * r_b.z)) + m;
  const float kzz = ((az.x * r_a.y - az.y * r_a.x) + (bz.x * r_b.y - bz.y * r_b.x)) + m;
  const float kxy = (ay.y * r_a.z - ay.z * r_a.y) + (by.y * r_b.z - by.z * r_b.y);
  const float kxz = (az.y * r_a.z - az.z * r_a.y) + (bz.y * r_b.z - bz.z * r_b.y);
  const float kyz = (az.z * r_a.x - az.x * r_a.z) + (bz.z * r_b.x - bz.x * r_b.z);

  const float cxx = kyy * kzz - kyz * kyz;
  const float cyy = kxx * kzz - kxz * kxz;
  const float czz = kxx * kyy - kxy * kxy;
  const float cxy = kxz * kyz - kxy * kzz;
  const float cxz = kxy * kyz - kyy * kxz;
  const float cyz = kxy * kxz - kxx * kyz;

  const float determinant = (kxx * cxx + kxy * cxy) + kxz * cxz;
  if (absf(determinant) > 1000.0f * kFltMin) {
    const float inv_det = 1.0f / determinant;
    return Vec3{inv_det * ((cxx * rhs.x + cxy * rhs.y) + cxz * rhs.z), inv_det * ((cxy * rhs.x + cyy * rhs.y) + cyz * rhs.z),
                inv_det * ((cxz * rhs.x + cyz * rhs.y) + czz * rhs.z)};
  }
  return zero3();
}

// 0.5 * rotate_vector(q, rel.s * e + cross(rel.v, e)) for a unit axis e
__device__ __forceinline__ Vec3 joint_perp_axis(Quat q, Quat rel, Vec3 e) {
  return scale(0.5f, rotate(q, add(scale(rel.s, e), cross(rel.v, e))));
}

__device__ __forceinline__ BodyState joint_state(const DeviceWorld& d, int32_t index) {
  if (index >= 0) return d.states[index];
  BodyState s{};
  s.dq = quat_identity();
  return s;
}

__device__ __forceinline__ void joint_store(const DeviceWorld& d, int32_t index, const BodyState& s) {
  if (index >= 0 && (s.flags & bflag::kDynamic)) d.states[index] = s;
}

// Prepare one joint: softness from the time step, body masses and world anchor frames.
__device__ inline void prepare_joint_item(const DeviceWorld& d, int32_t jid, const StepParams& p) {
  JointSim j = d.joints[jid];
  const float hertz = minf(j.constraint_hertz, 0.25f * p.inv_h);
  j.constraint_softness = make_soft(hertz, j.constraint_damping_ratio, p.h);

  const BodySim& sim_a = j.enc_a >= 0 ? d.sims[j.enc_a] : d.static_sims[-j.enc_a - 2];
  const BodySim& sim_b = j.enc_b >= 0 ? d.sims[j.enc_b] : d.static_sims[-j.enc_b - 2];
  j.inv_mass_a = sim_a.inv_mass;
  j.inv_mass_b = sim_b.inv_mass;
  j.inv_ia = sim_a.inv_inertia_world;
  j.inv_ib = sim_b.inv_inertia_world;
  const Mat3 inv_inertia_sum = add_mm(j.inv_ia, j.inv_ib);
  j.fixed_rotation = det(inv_inertia_sum) < 1000.0f * kFltMin ? 1 : 0;
  j.index_a = j.enc_a >= 0 ? j.enc_a : kNullIndex;
  j.index_b = j.enc_b >= 0 ? j.enc_b : kNullIndex;

  j.frame_a.q = mul_quat(sim_a.transform.q, j.local_frame_a.q);
  j.frame_a.p = rotate(sim_a.transform.q, sub(j.local_frame_a.p, sim_a.local_center));
  j.frame_b.q = mul_quat(sim_b.transform.q, j.local_frame_b.q);
  j.frame_b.p = rotate(sim_b.transform.q, sub(j.local_frame_b.p, sim_b.local_center));
  j.delta_center = sub_pos(sim_b.center, sim_a.center);

  const Vec3 axis = rotate(j.frame_a.q, Vec3{0.0f, 0.0f, 1.0f});
  const float k = dot(axis, mul_mv(inv_inertia_sum, axis));
  j.axial_mass = k > 0.0f ? 1.0f / k : 0.0f;
  j.rotation_axis_z = axis;

  const Quat rel_q = inv_mul_quat(j.frame_a.q, j.frame_b.q);
  j.perp_axis_x = joint_perp_axis(j.frame_a.q, rel_q, Vec3{1.0f, 0.0f, 0.0f});
  j.perp_axis_y = joint_perp_axis(j.frame_a.q, rel_q, Vec3{0.0f, 1.0f, 0.0f});
  j.spring_softness = make_soft(j.hertz, j.damping_ratio, p.h);

  if (p.warm_start_scale == 0.0f) {
    j.linear_impulse = zero3();
    j.perp_impulse = Vec2{0.0f, 0.0f};
    j.motor_impulse = 0.0f;
    j.spring_impulse = 0.0f;
    j.lower_impulse = 0.0f;
    j.upper_impulse = 0.0f;
  }
  d.joints[jid] = j;
}

// A joint's warm-start velocity changes for its two bodies (warm_start_joint_item's
// operations; the bodies' dq are not changed by warm starting).
__device__ inline void warm_delta_joint(const DeviceWorld& d, const JointSim& j, WarmDelta* out) {
  const BodyState sa = joint_state(d, j.index_a);
  const BodyState sb = joint_state(d, j.index_b);
  const Vec3 r_a = rotate(sa.dq, j.frame_a.p);
  const Vec3 r_b = rotate(sb.dq, j.frame_b.p);
  const float axial_impulse = ((j.spring_impulse + j.motor_impulse) + j.lower_impulse) - j.upper_impulse;
  Vec3 angular_impulse = add(scale(j.perp_impulse.x, j.perp_axis_x), scale(j.perp_impulse.y, j.perp_axis_y));
  angular_impulse = mul_add(angular_impulse, axial_impulse, j.rotation_axis_z);
  out[0] = WarmDelta{scale(j.inv_mass_a, j.linear_impulse), mul_mv(j.inv_ia, add(cross(r_a, j.linear_impulse), angular_impulse))};
  out[1] = WarmDelta{scale(j.inv_mass_b, j.linear_impulse), mul_mv(j.inv_ib, add(cross(r_b, j.linear_impulse), angular_impulse))};
}

__device__ inline void warm_start_joint_item(const DeviceWorld& d, int32_t jid) {
  const JointSim& j = d.joints[jid];
  BodyState sa = joint_state(d, j.index_a);
  BodyState sb = joint_state(d, j.index_b);
  Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;

  const Vec3 r_a = rotate(sa.dq, j.frame_a.p);
  const Vec3 r_b = rotate(sb.dq, j.frame_b.p);
  const float axial_impulse = ((j.spring_impulse + j.motor_impulse) + j.lower_impulse) - j.upper_impulse;
  Vec3 angular_impulse = add(scale(j.perp_impulse.x, j.perp_axis_x), scale(j.perp_impulse.y, j.perp_axis_y));
  angular_impulse = mul_add(angular_impulse, axial_impulse, j.rotation_axis_z);

  v_a = sub(v_a, scale(j.inv_mass_a, j.linear_impulse));
  w_a = sub(w_a, mul_mv(j.inv_ia, add(cross(r_a, j.linear_impulse), angular_impulse)));
  v_b = add(v_b, scale(j.inv_mass_b, j.linear_impulse));
  w_b = add(w_b, mul_mv(j.inv_ib, add(cross(r_b, j.linear_impulse), angular_impulse)));

  if (sa.flags & bflag::kDynamic) { sa.v = v_a; sa.w = w_a; }
  if (sb.flags & bflag::kDynamic) { sb.v = v_b; sb.w = w_b; }
  joint_store(d, j.index_a, sa);
  joint_store(d, j.index_b, sb);
}

__device__ inline void solve_joint_item(const DeviceWorld& d, int32_t jid, const StepParams& p, bool use_bias) {
  JointSim j = d.joints[jid];
  const float m_a = j.inv_mass_a, m_b = j.inv_mass_b;
  BodyState sa = joint_state(d, j.index_a);
  BodyState sb = joint_state(d, j.index_b);
  Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;

  const bool fixed_rotation = j.fixed_rotation != 0;
  const Quat quat_a = mul_quat(sa.dq, j.frame_a.q);
  Quat quat_b = mul_quat(sb.dq, j.frame_b.q);
  if (dot_quat(quat_a, quat_b) < 0.0f) quat_b = Quat{{-quat_b.v.x, -quat_b.v.y, -quat_b.v.z}, -quat_b.s};
  const Quat rel_q = inv_mul_quat(quat_a, quat_b);

  if (!fixed_rotation && (j.flags & (jflag::kSpring | jflag::kMotor | jflag::kLimit))) {
    const Vec3 axis = j.rotation_axis_z;
    const Vec3 g_a = mul_mv(j.inv_ia, axis);
    const Vec3 g_b = mul_mv(j.inv_ib, axis);
    const float k = dot(axis, add(g_a, g_b));
    float axial_c_dot = dot(sub(w_b, w_a), axis);
    float axial_impulse = 0.0f;

    if (j.flags & jflag::kSpring) {
      const float angle = get_twist_angle(rel_q);
      const float c = angle - j.target_angle;
      const float bias = j.spring_softness.bias_rate * c;
      const float mass_scale = j.spring_softness.mass_scale;
      const float impulse_scale = j.spring_softness.impulse_scale;
      const float delta = -mass_scale * j.axial_mass * (axial_c_dot + bias) - impulse_scale * j.spring_impulse;
      j.spring_impulse += delta;
      axial_c_dot += k * delta;
      axial_impulse += delta;
    }

    if (j.flags & jflag::kMotor) {
      const float cdot = axial_c_dot - j.motor_speed;
      const float old_impulse = j.motor_impulse;
      const float max_impulse = j.max_motor_torque * p.h;
      const float delta = clampf(-j.axial_mass * cdot, -max_impulse - old_impulse, max_impulse - old_impulse);
      j.motor_impulse = old_impulse + delta;
      axial_c_dot += k * delta;
      axial_impulse += delta;
    }

    if (j.flags & jflag::kLimit) {
      const float angle = get_twist_angle(rel_q);
      {
        const float c = angle - j.lower_angle;
        float bias = 0.0f, mass_scale = 1.0f, impulse_scale = 0.0f;
        if (c > 0.0f) {
          bias = c * p.inv_h;
        } else if (use_bias) {
          bias = j.constraint_softness.bias_rate * c;
          mass_scale = j.constraint_softness.mass_scale;
          impulse_scale = j.constraint_softness.impulse_scale;
        }
        const float old_impulse = j.lower_impulse;
        const float delta = maxf(-mass_scale * j.axial_mass * (axial_c_dot + bias) - impulse_scale * old_impulse, -old_impulse);
        j.lower_impulse = old_impulse + delta;
        axial_c_dot += k * delta;
        axial_impulse += delta;
      }
      {
        const float c = j.upper_angle - angle;
        float bias = 0.0f, mass_scale = 1.0f, impulse_scale = 0.0f;
        if (c > 0.0f) {
          bias = c * p.inv_h;
        } else if (use_bias) {
          bias = j.constraint_softness.bias_rate * c;
          mass_scale = j.constraint_softness.mass_scale;
          impulse_scale = j.constraint_softness.impulse_scale;
        }
        const float cdot = -axial_c_dot;
        const float old_impulse = j.upper_impulse;
        const float delta = maxf(-mass_scale * j.axial_mass * (cdot + bias) - impulse_scale * old_impulse, -old_impulse);
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
      bias = Vec2{j.constraint_softness.bias_rate * rel_q.v.x, j.constraint_softness.bias_rate * rel_q.v.y};
      mass_scale = j.constraint_softness.mass_scale;
      impulse_scale = j.constraint_softness.impulse_scale;
    }
    const Vec3 pax = joint_perp_axis(quat_a, rel_q, Vec3{1.0f, 0.0f, 0.0f});
    const Vec3 pay = joint_perp_axis(quat_a, rel_q, Vec3{0.0f, 1.0f, 0.0f});
    j.perp_axis_x = pax;
    j.perp_axis_y = pay;

    const Mat3 inv_inertia_sum = add_mm(j.inv_ia, j.inv_ib);
    const float kxx = dot(pax, mul_mv(inv_inertia_sum, pax));
    const float kyy = dot(pay, mul_mv(inv_inertia_sum, pay));
    const float kxy = dot(pax, mul_mv(inv_inertia_sum, pay));
    const Mat2 k{{kxx, kxy}, {kxy, kyy}};

    const Vec3 w_rel = sub(w_b, w_a);
    const Vec2 cdot{dot(w_rel, pax), dot(w_rel, pay)};
    const Vec2 old_impulse = j.perp_impulse;
    const Vec2 sol = solve2(k, Vec2{cdot.x + bias.x, cdot.y + bias.y});
    const Vec2 delta{(-mass_scale) * sol.x - impulse_scale * old_impulse.x, (-mass_scale) * sol.y - impulse_scale * old_impulse.y};
    j.perp_impulse = Vec2{j.perp_impulse.x + delta.x, j.perp_impulse.y + delta.y};

    const Vec3 angular_impulse = add(scale(delta.x, pax), scale(delta.y, pay));
    w_a = sub(w_a, mul_mv(j.inv_ia, angular_impulse));
    w_b = add(w_b, mul_mv(j.inv_ib, angular_impulse));
  }

  {
    const Vec3 r_a = rotate(sa.dq, j.frame_a.p);
    const Vec3 r_b = rotate(sb.dq, j.frame_b.p);
    const Vec3 cdot = sub(sub(add(v_b, cross(w_b, r_b)), v_a), cross(w_a, r_a));
    Vec3 bias = zero3();
    float mass_scale = 1.0f, impulse_scale = 0.0f;
    if (use_bias) {
      const Vec3 separation = add(add(sub(sb.dp, sa.dp), sub(r_b, r_a)), j.delta_center);
      bias = scale(j.constraint_softness.bias_rate, separation);
      mass_scale = j.constraint_softness.mass_scale;
      impulse_scale = j.constraint_softness.impulse_scale;
    }
    const Vec3 b = solve_point_constraint(m_a, m_b, j.inv_ia, j.inv_ib, r_a, r_b, add(cdot, bias));
    const Vec3 impulse = sub(scale(-mass_scale, b), scale(impulse_scale, j.linear_impulse));
    j.linear_impulse = add(j.linear_impulse, impulse);

    v_a = sub(v_a, scale(m_a, impulse));
    w_a = sub(w_a, mul_mv(j.inv_ia, cross(r_a, impulse)));
    v_b = add(v_b, scale(m_b, impulse));
    w_b = add(w_b, mul_mv(j.inv_ib, cross(r_b, impulse)));
  }

  if (sa.flags & bflag::kDynamic) { sa.v = v_a; sa.w = w_a; }
  if (sb.flags & bflag::kDynamic) { sb.v = v_b; sb.w = w_b; }
  joint_store(d, j.index_a, sa);
  joint_store(d, j.index_b, sb);
  // the solve writes the impulses, the perpendicular axes and nothing else of the joint
  JointSim& out = d.joints[jid];
  out.linear_impulse = j.linear_impulse;
  out.perp_impulse = j.perp_impulse;
  out.spring_impulse = j.spring_impulse;
  out.motor_impulse = j.motor_impulse;
  out.lower_impulse = j.lower_impulse;
  out.upper_impulse = j.upper_impulse;
  out.perp_axis_x = j.perp_axis_x;
  out.perp_axis_y = j.perp_axis_y;
}

}  // namespace
}  // namespace nsl
