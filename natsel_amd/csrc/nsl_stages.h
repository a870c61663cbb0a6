// Per-item device functions of the rigid-body step, shared by the single-world step
// kernels (nsl_kernels.hip) and the batched stepper (nsl_batch.hip): narrow phase,
// constraint preparation, integration, the coloured soft-step stages, the overflow
// colour, impulse storage and body finalisation.
//
// Each function restates the reference's per-item formula (file:line at each function)
// with the same operand order, so one GPU step equals one CPU step bit for bit. Items of
// one colour touch disjoint dynamic bodies (constraint_graph.rs:113-197), so one lane
// per constraint within a colour reproduces the CPU's Gauss-Seidel order exactly.
// `DeviceWorld` pointers may address global memory or LDS (body states in the batched
// stepper); the functions only index them.
#pragma once
#include <hip/hip_runtime.h>

#include "nsl_device.h"
#include "nsl_round_collide.h"

namespace nsl {
namespace {

__device__ __forceinline__ BodyState identity_state() {
  BodyState s;
  s.v = zero3(); s.pad1 = 0.0f; s.w = zero3(); s.pad2 = 0.0f; s.dp = zero3(); s.flags = 0; s.dq = quat_identity();
  return s;
}
__device__ __forceinline__ BodyState load_state(const DeviceWorld& d, int32_t i) {
  return i >= 0 ? d.states[i] : identity_state();
}
__device__ __forceinline__ void store_state(const DeviceWorld& d, int32_t i, const BodyState& s) {
  if (i >= 0 && (s.flags & bflag::kDynamic) != 0) d.states[i] = s;
}

// ------------------------------------------------------------------------------
// narrow phase (collide_pass.rs:541-742, contact_update.rs:43-468)

struct BodyView {
  WXf transform;
  Pos center;
  Vec3 local_center;
  Vec3 max_extent;
  uint32_t flags;
  bool is_static;
};

// The sim behind an encoding; a sleeping body (-1) is found through its shape's body
// (collide_pass.rs:105-154).
__device__ __forceinline__ const BodySim* sim_of(const DeviceWorld& d, int32_t enc, int32_t shape) {
  if (enc >= 0) return &d.sims[enc];
  if (enc < kNullIndex) return &d.static_sims[-enc - 2];
  return &d.sleep_sims[d.shapes[shape].body_id];
}

__device__ __forceinline__ BodyView body_view(const DeviceWorld& d, int32_t enc, int32_t shape) {
  BodyView v;
  const BodySim* s = sim_of(d, enc, shape);
  v.transform = s->transform;
  v.center = s->center;
  v.local_center = s->local_center;
  v.max_extent = s->max_extent;
  v.flags = s->flags;
  v.is_static = enc < kNullIndex;
  return v;
}

// compute_convex_manifold (contact_update.rs:43-237) for a box pair.
__device__ bool compute_manifold(const DeviceWorld& d, Contact& c, const Shape& sa, const Shape& sb, WXf xf_a, WXf xf_b) {
  LocalManifold lm;
  lm.point_count = 0;
  Xf b_to_a = inv_mul_wxf(xf_a, xf_b);
  if (sa.kind == kShapeSphere) {
    collide_spheres(lm, sa.c1, sa.radius, sb.c1, sb.radius, b_to_a);
  } else if (sa.kind == kShapeCapsule) {
    if (sb.kind == kShapeSphere) collide_capsule_sphere(lm, sa.c1, sa.c2, sa.radius, sb.c1, sb.radius, b_to_a);
    else collide_capsules(lm, sa.c1, sa.c2, sa.radius, sb.c1, sb.c2, sb.radius, b_to_a);
  } else if (sb.kind == kShapeSphere) {
    SimplexCache cache = c.simplex;
    collide_box_sphere(lm, d.hulls[sa.hull], sb.c1, sb.radius, b_to_a, cache);
    c.simplex = cache;
  } else if (sb.kind == kShapeCapsule) {
    SimplexCache cache = c.simplex;
    collide_box_capsule(lm, d.hulls[sa.hull], sb.c1, sb.c2, sb.radius, b_to_a, cache);
    c.simplex = cache;
  } else {
    SatCache cache = c.sat;
    collide_boxes(lm, d.hulls[sa.hull], d.hulls[sb.hull], b_to_a, cache);
    c.sat = cache;
    atomicAdd(&d.counters[1], 1);
    if (cache.hit) atomicAdd(&d.counters[2], 1);
  }
  if (lm.point_count == 0) {
    if (c.manifold_count > 0) {
      c.point_count = 0;
      c.manifold_count = 0;
    }
    return false;
  }
  ManifoldPoint old_points[4];
  int old_count = 0;
  if (c.manifold_count == 0) {
    c.normal = zero3();
    c.twist_impulse = 0.0f;
    c.friction_impulse = zero3();
    c.rolling_impulse = zero3();
    c.manifold_count = 1;
  } else {
    old_count = c.point_count;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < old_count) old_points[i] = c.points[i];
  }
  c.point_count = lm.point_count;
  Mat3 ma = matrix_from_quat(xf_a.q);
  c.normal = mul_mv(ma, lm.normal);
  Vec3 dpos = sub_pos(xf_a.p, xf_b.p);
  // Point loops run over all four slots with a guard so every index is a constant and
  // the contact, the local manifold and the old points stay in registers.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= lm.point_count) break;
    ManifoldPoint t;
    t.anchor_a = mul_mv(ma, lm.pts[i].point);
    t.anchor_b = add(t.anchor_a, dpos);
    t.separation = lm.pts[i].separation;
    t.feature_id = feature_id(lm.pts[i].pair);
    t.normal_velocity = 0.0f;
    t.total_normal_impulse = 0.0f;
    t.base_separation = 0.0f;
    bool persisted = false;
    t.normal_impulse = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= old_count) break;
      if (t.feature_id == old_points[j].feature_id) {
        t.normal_impulse = old_points[j].normal_impulse;
        persisted = true;
        old_points[j].feature_id = 0xffffffffu;
        break;
      }
    }
    (void)persisted;
    c.points[i] = t;
  }
  return true;
}

// One contact of the collide pass: the fat-AABB gate, recycling, or a full narrow phase;
// returns its ContactChange.
__device__ uint8_t collide_contact(const DeviceWorld& d, int32_t id) {
  Contact c = d.contacts[id];
  uint8_t change = kChangeNone;
  if (!aabb_overlaps(d.fat_aabbs[c.shape_a], d.fat_aabbs[c.shape_b])) {
    c.flags &= ~cflag::kSimTouching;
    d.contacts[id].flags = c.flags;
    return kChangeDisjoint;
  }
  BodyView ba = body_view(d, c.enc_a, c.shape_a), bb = body_view(d, c.enc_b, c.shape_b);
  const bool was_touching = (c.flags & cflag::kSimTouching) != 0;
  const WXf xf_a = ba.transform, xf_b = bb.transform;
  const bool is_fast = ((ba.flags | bb.flags) & bflag::kIsFast) != 0;
  const float tolerance = was_touching ? kRecycleDistance : minf(kRecycleDistance, kSpeculativeDistance);
  if (!is_fast && (c.flags & cflag::kRelativeTransformValid) && (c.flags & cflag::kRecycle)) {
    float angle_a = dot_quat(xf_a.q, c.cached_rot_a);
    float angle_b = dot_quat(xf_b.q, c.cached_rot_b);
    float angular_distance = minf(angle_a * angle_a, angle_b * angle_b);
    Xf xf = inv_mul_wxf(xf_a, xf_b);
    Xf xfc = c.cached_pose;
    Vec3 max_extent = max3(ba.is_static ? zero3() : ba.max_extent, bb.is_static ? zero3() : bb.max_extent);
    float dist_sq = distance_sq(xf.p, xfc.p);
    if (angular_distance > kRecycleAngularDistance && dist_sq < tolerance * tolerance) {
      float slack = tolerance - sqrtf_(dist_sq);
      Quat qr = inv_mul_quat(xfc.q, xf.q);
      Vec3 arc = modified_cross(abs3(qr.v), max_extent);
      if (4.0f * length_sq(arc) < slack * slack) {
        Mat3 m_a = matrix_from_quat(mul_quat(xf_a.q, conjugate(c.cached_rot_a)));
        Mat3 m_b = matrix_from_quat(mul_quat(xf_b.q, conjugate(c.cached_rot_b)));
        Vec3 dc = sub_pos(bb.center, ba.center);
        if (c.manifold_count > 0) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (i >= c.point_count) break;
            Vec3 r_a = mul_mv(m_a, c.points[i].anchor_a);
            Vec3 r_b = mul_mv(m_b, c.points[i].anchor_b);
            Vec3 dp = add(dc, sub(r_b, r_a));
            d.contacts[id].points[i].separation = c.points[i].base_separation + dot(dp, c.normal);
            d.contacts[id].points[i].normal_velocity = 0.0f;
          }
        }
        atomicAdd(&d.counters[3], 1);
        return kChangeNone;
      }
    }
  }
  c.cached_rot_a = xf_a.q;
  c.cached_rot_b = xf_b.q;
  c.cached_pose = inv_mul_wxf(xf_a, xf_b);
  c.flags |= cflag::kRelativeTransformValid;

  const Shape sa = d.shapes[c.shape_a];
  const Shape sb = d.shapes[c.shape_b];
  bool touching = compute_manifold(d, c, sa, sb, xf_a, xf_b);
  if (touching) {
    // update_convex_contact (contact_update.rs:243-360)
    c.friction = sqrtf_(sa.friction * sb.friction);
    c.restitution = maxf(sa.restitution, sb.restitution);
    if (sa.rolling_resistance > 0.0f || sb.rolling_resistance > 0.0f) {
      float ra = sa.kind == kShapeHull ? 0.25f * d.hulls[sa.hull].inner_radius : sa.radius;
      float rb = sb.kind == kShapeHull ? 0.25f * d.hulls[sb.hull].inner_radius : sb.radius;
      c.rolling_resistance = maxf(sa.rolling_resistance, sb.rolling_resistance) * maxf(ra, rb);
    } else {
      c.rolling_resistance = 0.0f;
    }
    c.tangent_velocity = sub(rotate(xf_a.q, sa.tangent_velocity), rotate(xf_b.q, sb.tangent_velocity));
    if (sa.enable_hit_events || sb.enable_hit_events) c.flags |= cflag::kHitEvent;
    else c.flags &= ~cflag::kHitEvent;
    Vec3 ca = rotate(xf_a.q, ba.local_center), cb = rotate(xf_b.q, bb.local_center);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= c.point_count) break;
      c.points[j].anchor_a = sub(c.points[j].anchor_a, ca);
      c.points[j].anchor_b = sub(c.points[j].anchor_b, cb);
      c.points[j].base_separation = c.points[j].separation;
    }
    c.flags |= cflag::kSimTouching;
  } else {
    c.flags &= ~cflag::kSimTouching;
  }
  if (touching && !was_touching) change = kChangeBegan;
  else if (!touching && was_touching) change = kChangeEnded;
  d.contacts[id] = c;
  return change;
}

// ------------------------------------------------------------------------------
// constraint record stores

// A record moves in kRecPieces pieces of 16 bytes.
constexpr int kRecPieces = (int)(sizeof(Constraint) / 16);
static_assert(sizeof(Constraint) % 16 == 0, "constraint records move in 16-byte pieces");
static_assert(offsetof(Constraint, index_a) == 4 && offsetof(Constraint, index_b) == 8,
              "body indices sit in a record's first piece");

// Constraint records (the single world's, and each batched world's segment, which starts on a
// block) in blocks of kRecBlock slots (nsl_types.h), piece-major inside a block: piece j of
// slot s at p[((s / kRecBlock) * kRecPieces + j) * kRecBlock + s % kRecBlock].
// A wavefront touching consecutive slots reads or writes one contiguous KiB per piece (a record
// in slot order would spread each such access over 64 cache lines). Capacity: whole blocks.
struct BlockedCons {
  float4* p;
  __device__ __forceinline__ float4* piece(int s, int j) const {
    return p + ((size_t)(s / kRecBlock) * kRecPieces + j) * kRecBlock + (s % kRecBlock);
  }
  __device__ __forceinline__ Constraint load(int s) const {
    Constraint c;
    float4* q = reinterpret_cast<float4*>(&c);  // 16-byte aligned, kRecPieces pieces
#pragma unroll
    for (int j = 0; j < kRecPieces; ++j) q[j] = *piece(s, j);
    return c;
  }
  __device__ __forceinline__ void store(int s, const Constraint& c) const {
    const float4* q = reinterpret_cast<const float4*>(&c);
#pragma unroll
    for (int j = 0; j < kRecPieces; ++j) *piece(s, j) = q[j];
  }
  __device__ __forceinline__ int2 indices(int s) const {
    const float4 h = *piece(s, 0);
    return make_int2(__float_as_int(h.y), __float_as_int(h.z));
  }
  // a field of T (4-byte words) at byte offset `off` of the record
  template <class T>
  __device__ __forceinline__ void put(int s, size_t off, T v) const {
    static_assert(sizeof(T) % 4 == 0, "fields are 4-byte words");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (size_t k = 0; k < sizeof(T) / 4; ++k) {
      const size_t o = off + 4 * k;
      reinterpret_cast<uint32_t*>(piece(s, (int)(o / 16)))[(o % 16) / 4] = w[k];
    }
  }
};
__device__ __forceinline__ BlockedCons blocked_cons(const DeviceWorld& d) { return BlockedCons{reinterpret_cast<float4*>(d.constraints)}; }

// The solver stages' writes to one slot's record.
template <class Cons>
struct ConOut {
  Cons cons;
  int slot;
  static constexpr size_t point_off(int i) { return offsetof(Constraint, points) + (size_t)i * sizeof(ConstraintPoint); }
  __device__ __forceinline__ void normal_impulse(int i, float v) const {
    cons.put(slot, point_off(i) + offsetof(ConstraintPoint, normal_impulse), v);
  }
  __device__ __forceinline__ void total_normal_impulse(int i, float v) const {
    cons.put(slot, point_off(i) + offsetof(ConstraintPoint, total_normal_impulse), v);
  }
  __device__ __forceinline__ void restitution_impulse(int i, float v) const {
    cons.put(slot, point_off(i) + offsetof(ConstraintPoint, restitution_impulse), v);
  }
  __device__ __forceinline__ void twist_impulse(float v) const { cons.put(slot, offsetof(Constraint, twist_impulse), v); }
  __device__ __forceinline__ void friction_impulse(Vec2 v) const { cons.put(slot, offsetof(Constraint, friction_impulse), v); }
  __device__ __forceinline__ void rolling_impulse(Vec3 v) const { cons.put(slot, offsetof(Constraint, rolling_impulse), v); }
};
template <class Cons>
__device__ __forceinline__ ConOut<Cons> con_out(Cons cons, int slot) { return ConOut<Cons>{cons, slot}; }

// ------------------------------------------------------------------------------
// constraint preparation (contact_solver.rs:395-662)

struct NoPost {
  __device__ void operator()(const Constraint&) const {}
};

// Returns the number of manifold points prepared; the record goes to `cons`, and `post` sees it
// while it is still in registers.
template <class Cons, class Post = NoPost>
__device__ __forceinline__ int prepare_contact(const DeviceWorld& d, Cons cons, int slot, int32_t id, const StepParams& p,
                                               float* restitution = nullptr, Post post = Post{}) {
  const Contact& contact = d.contacts[id];
  const float inv_tau = 1.0f / kSpeculativeDistance;
  Constraint cc;
  cc.contact_id = id;
  int32_t ia = contact.enc_a >= 0 ? contact.enc_a : kNullIndex;
  int32_t ib = contact.enc_b >= 0 ? contact.enc_b : kNullIndex;
  cc.index_a = ia;
  cc.index_b = ib;
  cc.inv_mass_a = 0.0f; cc.inv_mass_b = 0.0f;
  cc.inv_ia = SymMat3{0, 0, 0, 0, 0, 0};
  cc.inv_ib = cc.inv_ia;
  if (ia >= 0) { cc.inv_mass_a = d.sims[ia].inv_mass; cc.inv_ia = sym_from(d.sims[ia].inv_inertia_world); }
  if (ib >= 0) { cc.inv_mass_b = d.sims[ib].inv_mass; cc.inv_ib = sym_from(d.sims[ib].inv_inertia_world); }
  cc.friction = contact.friction;
  cc.restitution = contact.restitution;
  cc.rolling_resistance = contact.rolling_resistance;
  cc.is_static = (contact.flags & cflag::kStatic) != 0;
  SymMat3 inv_iab = add_sym(cc.inv_ia, cc.inv_ib);
  cc.rolling_mass = cc.rolling_resistance > 0.0f ? invert_sym(inv_iab) : SymMat3{0, 0, 0, 0, 0, 0};
  cc.manifold_count = contact.manifold_count;
  cc.point_count = contact.manifold_count > 0 ? contact.point_count : 0;
  const float warm = p.warm_start_scale;
  Vec3 normal = contact.normal;
  cc.normal = normal;
  cc.tangent1 = perp(normal);
  cc.tangent2 = cross(cc.tangent1, normal);
  cc.tangent_velocity1 = dot(contact.tangent_velocity, cc.tangent1);
  cc.tangent_velocity2 = dot(contact.tangent_velocity, cc.tangent2);
  cc.twist_impulse = warm * contact.twist_impulse;
  cc.rolling_impulse = cc.rolling_resistance > 0.0f ? scale(warm, contact.rolling_impulse) : zero3();
  cc.friction_impulse = Vec2{warm * dot(contact.friction_impulse, cc.tangent1), warm * dot(contact.friction_impulse, cc.tangent2)};
  Vec3 center_a = zero3(), center_b = zero3();
  float total_weight = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= cc.point_count) break;
    const ManifoldPoint& mp = contact.points[i];
    ConstraintPoint cp;
    cp.r_a = mp.anchor_a;
    cp.r_b = mp.anchor_b;
    cp.base_separation = mp.separation - dot(sub(cp.r_b, cp.r_a), normal);
    cp.normal_impulse = warm * mp.normal_impulse;
    cp.total_normal_impulse = 0.0f;
    cp.relative_velocity = 0.0f;
    cp.restitution_impulse = 0.0f;
    float weight = clampf(2.0f - mp.separation * inv_tau, kMinFrictionWeight, 1.0f);
    center_a = add(center_a, scale(weight, cp.r_a));
    center_b = add(center_b, scale(weight, cp.r_b));
    total_weight += weight;
    Vec3 rn_a = cross(cp.r_a, normal), rn_b = cross(cp.r_b, normal);
    float k = cc.inv_mass_a + cc.inv_mass_b;
    k += dot(rn_a, sym_mv(cc.inv_ia, rn_a));
    k += dot(rn_b, sym_mv(cc.inv_ib, rn_b));
    cp.normal_mass = k > 0.0f ? 1.0f / k : 0.0f;
    cp.lever_arm = 0.0f;
    cc.points[i] = cp;
  }
  if (total_weight > 0.0f) {
    cc.center_a = scale(1.0f / total_weight, center_a);
    cc.center_b = scale(1.0f / total_weight, center_b);
  } else {
    cc.center_a = zero3();
    cc.center_b = zero3();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= cc.point_count) break;
    Vec3 dd = sub(cc.points[i].r_a, cc.center_a);
    cc.points[i].lever_arm = sqrtf_(dot(dd, dd));
  }
  {
    Vec3 rt_a1 = cross(cc.center_a, cc.tangent1), rt_a2 = cross(cc.center_a, cc.tangent2);
    Vec3 rt_b1 = cross(cc.center_b, cc.tangent1), rt_b2 = cross(cc.center_b, cc.tangent2);
    Vec3 ia1 = sym_mv(cc.inv_ia, rt_a1), ia2 = sym_mv(cc.inv_ia, rt_a2);
    Vec3 ib1 = sym_mv(cc.inv_ib, rt_b1), ib2 = sym_mv(cc.inv_ib, rt_b2);
    float m_sum = cc.inv_mass_a + cc.inv_mass_b;
    float kxx = (m_sum + dot(rt_a1, ia1)) + dot(rt_b1, ib1);
    float kyy = (m_sum + dot(rt_a2, ia2)) + dot(rt_b2, ib2);
    float kxy = dot(rt_a1, ia2) + dot(rt_b1, ib2);
    float det_k = kxx * kyy - kxy * kxy;
    bool valid = absf(det_k) > 1000.0f * kFltMin && total_weight > 0.0f;
    float inv_det = valid ? 1.0f / det_k : 0.0f;
    cc.tangent_mass.cx = Vec2{inv_det * kyy, -(inv_det * kxy)};
    cc.tangent_mass.cy = Vec2{-(inv_det * kxy), inv_det * kxx};
  }
  float k_twist = dot(normal, sym_mv(inv_iab, normal));
  cc.twist_mass = k_twist > 0.0f ? 1.0f / k_twist : 0.0f;
  if (contact.restitution > 0.0f || (contact.flags & cflag::kHitEvent) != 0) {
    BodyState sa = load_state(d, ia), sb = load_state(d, ib);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= cc.point_count) break;
      Vec3 vr_a = add(sa.v, cross(sa.w, cc.points[i].r_a));
      Vec3 vr_b = add(sb.v, cross(sb.w, cc.points[i].r_b));
      cc.points[i].relative_velocity = dot(normal, sub(vr_b, vr_a));
    }
  }
  cons.store(slot, cc);
  if (restitution) *restitution = cc.restitution;
  post(cc);
  return cc.point_count;
}

// The single world's coloured slot `slot`.
template <class Post>
__device__ __forceinline__ int prepare_item(const DeviceWorld& d, int slot, const StepParams& p, Post post) {
  return prepare_contact(d, blocked_cons(d), slot, d.solve_ids[slot], p, nullptr, post);
}

// overflow_contact_solver.rs:92-257
__device__ __forceinline__ int prepare_overflow_item(const DeviceWorld& d, int k, const StepParams& p) {
  const int32_t id = d.overflow_ids[k];
  Contact& contact = d.contacts[id];
  const float warm = p.warm_start_scale;
  const float inv_tau = 1.0f / kSpeculativeDistance;
  OverflowConstraint c;
  int32_t ia = contact.enc_a >= 0 ? contact.enc_a : kNullIndex;
  int32_t ib = contact.enc_b >= 0 ? contact.enc_b : kNullIndex;
  c.contact_id = id;
  c.index_a = ia;
  c.index_b = ib;
  c.inv_mass_a = 0.0f; c.inv_mass_b = 0.0f;
  c.inv_ia = mat3_zero(); c.inv_ib = mat3_zero();
  if (ia >= 0) { c.inv_mass_a = d.sims[ia].inv_mass; c.inv_ia = d.sims[ia].inv_inertia_world; }
  if (ib >= 0) { c.inv_mass_b = d.sims[ib].inv_mass; c.inv_ib = d.sims[ib].inv_inertia_world; }
  float mA = c.inv_mass_a, mB = c.inv_mass_b;
  Mat3 i_a = c.inv_ia, i_b = c.inv_ib;
  c.rolling_mass = invert_matrix(add_mm(i_a, i_b));
  c.softness = (contact.flags & cflag::kStatic) ? p.static_softness : p.contact_softness;
  c.friction = contact.friction;
  c.restitution = contact.restitution;
  c.rolling_resistance = contact.rolling_resistance;
  bool hit_events = (contact.flags & cflag::kHitEvent) != 0;
  bool sample_velocity = contact.restitution > 0.0f || hit_events;
  Vec3 v_a = zero3(), w_a = zero3(), v_b = zero3(), w_b = zero3();
  if (sample_velocity) {
    if (ia >= 0) { v_a = d.states[ia].v; w_a = d.states[ia].w; }
    if (ib >= 0) { v_b = d.states[ib].v; w_b = d.states[ib].w; }
  }
  c.manifold_count = contact.manifold_count;
  c.point_count = contact.manifold_count > 0 ? contact.point_count : 0;
  Vec3 normal = contact.normal;
  Vec3 t1 = perp(normal), t2 = cross(t1, normal);
  c.normal = normal; c.tangent1 = t1; c.tangent2 = t2;
  c.tangent_velocity1 = dot(contact.tangent_velocity, t1);
  c.tangent_velocity2 = dot(contact.tangent_velocity, t2);
  Vec3 center_a = zero3(), center_b = zero3();
  float total_w = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= c.point_count) break;
    ManifoldPoint mp = contact.points[i];
    ConstraintPoint cp;
    cp.r_a = mp.anchor_a;
    cp.r_b = mp.anchor_b;
    float s = mp.separation;
    cp.base_separation = s - dot(sub(cp.r_b, cp.r_a), normal);
    cp.normal_impulse = warm * mp.normal_impulse;
    cp.total_normal_impulse = 0.0f;
    cp.restitution_impulse = 0.0f;
    Vec3 rn_a = cross(cp.r_a, normal), rn_b = cross(cp.r_b, normal);
    float k2 = mA + mB + dot(rn_a, mul_mv(i_a, rn_a)) + dot(rn_b, mul_mv(i_b, rn_b));
    cp.normal_mass = k2 > 0.0f ? 1.0f / k2 : 0.0f;
    if (sample_velocity) {
      Vec3 vr_a = add(v_a, cross(w_a, cp.r_a)), vr_b = add(v_b, cross(w_b, cp.r_b));
      float vn = dot(normal, sub(vr_b, vr_a));
      cp.relative_velocity = vn;
      contact.points[i].normal_velocity = hit_events ? vn : 0.0f;
    } else {
      cp.relative_velocity = 0.0f;
      contact.points[i].normal_velocity = 0.0f;
    }
    float weight = clampf(2.0f - s * inv_tau, kMinFrictionWeight, 1.0f);
    center_a = mul_add(center_a, weight, cp.r_a);
    center_b = mul_add(center_b, weight, cp.r_b);
    total_w += weight;
    cp.lever_arm = 0.0f;
    c.points[i] = cp;
  }
  if (c.point_count > 0) {
    float inv_w = 1.0f / total_w;
    center_a = scale(inv_w, center_a);
    center_b = scale(inv_w, center_b);
  }
  c.center_a = center_a;
  c.center_b = center_b;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < c.point_count) c.points[i].lever_arm = distance(c.points[i].r_a, center_a);
  Vec3 rt_a1 = cross(center_a, t1), rt_a2 = cross(center_a, t2);
  Vec3 rt_b1 = cross(center_b, t1), rt_b2 = cross(center_b, t2);
  Mat2 km;
  km.cx.x = mA + mB + dot(rt_a1, mul_mv(i_a, rt_a1)) + dot(rt_b1, mul_mv(i_b, rt_b1));
  km.cy.y = mA + mB + dot(rt_a2, mul_mv(i_a, rt_a2)) + dot(rt_b2, mul_mv(i_b, rt_b2));
  km.cx.y = dot(rt_a1, mul_mv(i_a, rt_a2)) + dot(rt_b1, mul_mv(i_b, rt_b2));
  km.cy.x = km.cx.y;
  c.tangent_mass = invert2(km);
  c.friction_impulse = Vec2{warm * dot(contact.friction_impulse, t1), warm * dot(contact.friction_impulse, t2)};
  float k_twist = dot(normal, mul_mv(add_mm(i_a, i_b), normal));
  c.twist_mass = k_twist > 0.0f ? 1.0f / k_twist : 0.0f;
  c.twist_impulse = warm * contact.twist_impulse;
  c.rolling_impulse = scale(warm, contact.rolling_impulse);
  d.overflow[k] = c;
  return c.point_count;
}

// ------------------------------------------------------------------------------
// integration (solver_integration.rs:263-426)

__device__ __forceinline__ void integrate_velocities_item(const DeviceWorld& d, int i, const StepParams& p) {
  const float h = p.h;
  const BodySim& sim = d.sims[i];
  BodyState state = d.states[i];
  Vec3 v = state.v, w = state.w;
  float linear_damping = 1.0f / (1.0f + h * sim.linear_damping);
  float angular_damping = 1.0f / (1.0f + h * sim.angular_damping);
  float gravity_scale = sim.inv_mass > 0.0f ? sim.gravity_scale : 0.0f;
  Vec3 dv = blend2(h * sim.inv_mass, sim.force, h * gravity_scale, p.gravity);
  v = mul_add(dv, linear_damping, v);
  Vec3 dw = scale(h, mul_mv(sim.inv_inertia_world, sim.torque));
  w = mul_add(dw, angular_damping, w);
  if (sim.flags & bflag::kGyroscopic) {
    Quat q = mul_quat(state.dq, sim.transform.q);
    Vec3 omega1 = inv_rotate(q, w);
    Vec3 omega2 = omega1;
    float i00 = sim.inertia.cx.x, i01 = sim.inertia.cy.x, i02 = sim.inertia.cz.x;
    float i11 = sim.inertia.cy.y, i12 = sim.inertia.cz.y, i22 = sim.inertia.cz.z;
    for (int it = 0; it < kGyroIterations; ++it) {
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
  d.states[i].v = v;
  d.states[i].w = w;
}

__device__ __forceinline__ void integrate_positions_item(const DeviceWorld& d, int i, const StepParams& p) {
  BodyState st = d.states[i];
  const float max_lin = p.max_linear_speed;
  const float max_ang = kMaxRotation * p.inv_dt;
  Vec3 v = st.v, w = st.w;
  if (st.flags & bflag::kLockLinX) v.x = 0.0f;
  if (st.flags & bflag::kLockLinY) v.y = 0.0f;
  if (st.flags & bflag::kLockLinZ) v.z = 0.0f;
  if (st.flags & bflag::kLockAngX) w.x = 0.0f;
  if (st.flags & bflag::kLockAngY) w.y = 0.0f;
  if (st.flags & bflag::kLockAngZ) w.z = 0.0f;
  if (dot(v, v) > max_lin * max_lin) {
    v = scale(max_lin / length(v), v);
    st.flags |= bflag::kSpeedCapped;
  }
  if (dot(w, w) > max_ang * max_ang && (st.flags & bflag::kAllowFastRotation) == 0) {
    w = scale(max_ang / length(w), w);
    st.flags |= bflag::kSpeedCapped;
  }
  st.v = v;
  st.w = w;
  st.dp = mul_add(st.dp, p.h, v);
  st.dq = integrate_rotation(st.dq, scale(p.h, w));
  d.states[i] = st;
}

// ------------------------------------------------------------------------------
// coloured stages: one lane per constraint slot of one colour; the colour's slots touch
// distinct dynamic bodies, so they run in any order

// Each item takes its constraint's body indices from the caller, which reads them before
// the barrier that opens the phase (they do not change during the solve): the body
// states and the constraint's fields are then all requested at once. The manifold points
// are read unconditionally (the slot always holds four) and used up to point_count.

// contact_solver.rs:697-763
// The impulse a contact constraint's warm start applies: linear, and angular for each body.
struct WarmImpulse {
  Vec3 linear, ang_a, ang_b;
};
__device__ __forceinline__ WarmImpulse warm_impulse(const Constraint& c) {
  const int pc = c.point_count;
  ConstraintPoint pt[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pt[i] = c.points[i];
  Vec3 linear = zero3(), ang_a = zero3(), ang_b = zero3();
  if (c.manifold_count > 0) {
    float total = 0.0f;
    Vec3 moment_a = zero3(), moment_b = zero3();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= pc) break;
      float ni = pt[i].normal_impulse;
      total += ni;
      moment_a = mul_add(moment_a, ni, pt[i].r_a);
      moment_b = mul_add(moment_b, ni, pt[i].r_b);
    }
    Vec3 normal = c.normal;
    Vec3 fi = scale(c.friction_impulse.x, c.tangent1);
    fi = mul_add(fi, c.friction_impulse.y, c.tangent2);
    Vec3 lin = mul_add(fi, total, normal);
    Vec3 twist = scale(c.twist_impulse, normal);
    Vec3 a_a = add(add(cross(moment_a, normal), cross(c.center_a, fi)), twist);
    Vec3 a_b = add(add(cross(moment_b, normal), cross(c.center_b, fi)), twist);
    if (c.rolling_resistance > 0.0f) {
      a_a = add(a_a, c.rolling_impulse);
      a_b = add(a_b, c.rolling_impulse);
    }
    linear = add(linear, lin);
    ang_a = add(ang_a, a_a);
    ang_b = add(ang_b, a_b);
  }
  return WarmImpulse{linear, ang_a, ang_b};
}

__device__ __forceinline__ void warm_start_item(const DeviceWorld& d, int slot, int32_t ia, int32_t ib) {
  BodyState ba = load_state(d, ia), bb = load_state(d, ib);
  const Constraint c = blocked_cons(d).load(slot);
  const WarmImpulse wi = warm_impulse(c);
  ba.w = sub(ba.w, sym_mv(c.inv_ia, wi.ang_a));
  ba.v = sub(ba.v, scale(c.inv_mass_a, wi.linear));
  bb.w = add(bb.w, sym_mv(c.inv_ib, wi.ang_b));
  bb.v = add(bb.v, scale(c.inv_mass_b, wi.linear));
  store_state(d, ia, ba);
  store_state(d, ib, bb);
}

// Warm start by body: an item's velocity changes for each of its two bodies, computed with
// the operations warm_start_item / warm_start_joint_item use and applied by the body in
// colour order (side a subtracts, side b adds; contacts change w then v, joints v then w).
__device__ __forceinline__ void warm_delta_contact(const Constraint& c, WarmDelta* out) {
  const WarmImpulse wi = warm_impulse(c);
  out[0] = WarmDelta{scale(c.inv_mass_a, wi.linear), sym_mv(c.inv_ia, wi.ang_a)};
  out[1] = WarmDelta{scale(c.inv_mass_b, wi.linear), sym_mv(c.inv_ib, wi.ang_b)};
}

// A constraint record staged in LDS by the step kernel: 16-byte piece j of the record of
// lane `lane` at rec[j * lanes + lane] (written by global_load_lds, one piece per lane per
// instruction, so consecutive lanes hold consecutive pieces).

// Waits for this wave's outstanding loads (the item's body states and staged-record reads)
// and keeps the compiler from moving memory operations across the wait.
__device__ __forceinline__ void wait_before_copy() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ Constraint staged_record(const float4* rec, int lanes, int lane) {
  float4 piece[kRecPieces];
#pragma unroll
  for (int j = 0; j < kRecPieces; ++j) piece[j] = rec[j * lanes + lane];
  Constraint c;
  __builtin_memcpy(&c, piece, sizeof(Constraint));
  return c;
}

// contact_solver.rs:796-894. Reads the record from `c` (a copy) and writes its impulses
// through `out` (ConOut).
template <class Out>
__device__ __forceinline__ void push_core(const Constraint& c, const Out& out, const StepParams& p, BodyState& ba,
                                          BodyState& bb) {
  const int pc = c.point_count;
  ConstraintPoint pt[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pt[i] = c.points[i];
  const Softness sf = c.is_static ? p.static_softness : p.contact_softness;
  const float bias_rate = sf.mass_scale * sf.bias_rate;
  const float mass_scale = sf.mass_scale, impulse_scale = sf.impulse_scale;
  const float inv_h = p.inv_h, speed_neg = -p.contact_speed;
  Vec3 dp = sub(bb.dp, ba.dp);
  float inv_mass = c.inv_mass_a + c.inv_mass_b;
  if (c.manifold_count > 0) {
    Vec3 normal = c.normal;
    float normal_sep = dot(normal, dp);
    Vec3 n_a = inv_rotate_wide(ba.dq, normal), n_b = inv_rotate_wide(bb.dq, normal);
    float lnv = dot(normal, sub(bb.v, ba.v));
    float manifold_impulse = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= pc) break;
      const ConstraintPoint& cp = pt[i];
      float s = (normal_sep + (dot(n_b, cp.r_b) - dot(n_a, cp.r_a))) + cp.base_separation;
      bool separated = s > 0.0f;
      float velocity_bias = separated ? s * inv_h : maxf(bias_rate * s, speed_neg);
      float pm = separated ? 1.0f : mass_scale;
      float pi = separated ? 0.0f : impulse_scale;
      Vec3 u_a = cross(cp.r_a, normal), u_b = cross(cp.r_b, normal);
      Vec3 g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
      float v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
      float neg_impulse = cp.normal_mass * (pm * v_n + velocity_bias) + pi * cp.normal_impulse;
      float impulse = maxf(-neg_impulse, -cp.normal_impulse);
      out.normal_impulse(i, cp.normal_impulse + impulse);
      lnv += inv_mass * impulse;
      manifold_impulse += impulse;
      ba.w = sub(ba.w, scale(impulse, g_a));
      bb.w = add(bb.w, scale(impulse, g_b));
    }
    Vec3 pp = scale(manifold_impulse, normal);
    ba.v = sub(ba.v, scale(c.inv_mass_a, pp));
    bb.v = add(bb.v, scale(c.inv_mass_b, pp));
  }
}

struct NoHook {
  __device__ void operator()() const {}
};

// `staged`: the record's copy in LDS (rec, lanes, lane) or null to read it from HBM. `hook`
// runs once the body states have arrived and the staged copy has been read out of LDS (the
// step kernel issues the next colour's record copy there, so that copy overlaps this item's
// work). The body states are waited for before the copy is issued: with a copy to LDS in
// flight, any wait for a load waits for every load, the copy included.
template <class Hook = NoHook>
__device__ __forceinline__ void push_item(const DeviceWorld& d, int slot, const StepParams& p, int32_t ia, int32_t ib,
                                          const float4* staged = nullptr, int lanes = 0, Hook hook = Hook{}) {
  BodyState ba = load_state(d, ia), bb = load_state(d, ib);
  const BlockedCons cons = blocked_cons(d);
  if (staged != nullptr) {
    const Constraint c = staged_record(staged, lanes, (int)threadIdx.x);
    wait_before_copy();
    hook();
    push_core(c, con_out(cons, slot), p, ba, bb);
  } else {
    hook();
    push_core(cons.load(slot), con_out(cons, slot), p, ba, bb);
  }
  store_state(d, ia, ba);
  store_state(d, ib, bb);
}

// contact_solver.rs:927-1097 (reads `c`, writes through `out` as push_core)
template <class Out>
__device__ __forceinline__ void relax_core(const Constraint& c, const Out& out, const StepParams& p, BodyState& ba,
                                           BodyState& bb) {
  const int pc = c.point_count;
  ConstraintPoint pt[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pt[i] = c.points[i];
  const float inv_h = p.inv_h, eps = kFltEps;
  Vec3 dp = sub(bb.dp, ba.dp);
  float inv_mass = c.inv_mass_a + c.inv_mass_b;
  if (c.manifold_count > 0) {
    Vec3 normal = c.normal;
    float normal_sep = dot(normal, dp);
    Vec3 n_a = inv_rotate_wide(ba.dq, normal), n_b = inv_rotate_wide(bb.dq, normal);
    Vec3 dv = sub(bb.v, ba.v);
    float lnv = dot(normal, dv);
    float ltv1 = dot(c.tangent1, dv) - c.tangent_velocity1;
    float ltv2 = dot(c.tangent2, dv) - c.tangent_velocity2;
    float manifold_impulse = 0.0f, total_normal = 0.0f, total_twist_limit = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= pc) break;
      const ConstraintPoint& cp = pt[i];
      float s = (normal_sep + (dot(n_b, cp.r_b) - dot(n_a, cp.r_a))) + cp.base_separation;
      float velocity_bias = maxf(0.0f, s * inv_h);
      Vec3 u_a = cross(cp.r_a, normal), u_b = cross(cp.r_b, normal);
      Vec3 g_a = sym_mv(c.inv_ia, u_a), g_b = sym_mv(c.inv_ib, u_b);
      float v_n = lnv + (dot(u_b, bb.w) - dot(u_a, ba.w));
      float neg_impulse = cp.normal_mass * (v_n + velocity_bias);
      float impulse = maxf(-neg_impulse, -cp.normal_impulse);
      float new_impulse = cp.normal_impulse + impulse;
      out.normal_impulse(i, new_impulse);
      out.total_normal_impulse(i, cp.total_normal_impulse + new_impulse);
      total_normal += new_impulse;
      total_twist_limit += cp.lever_arm * new_impulse;
      lnv += inv_mass * impulse;
      manifold_impulse += impulse;
      ba.w = sub(ba.w, scale(impulse, g_a));
      bb.w = add(bb.w, scale(impulse, g_b));
    }
    if (c.rolling_resistance > 0.0f) {
      Vec3 delta = sym_mv(c.rolling_mass, sub(ba.w, bb.w));
      Vec3 old = c.rolling_impulse;
      Vec3 ri = add(old, delta);
      float max_impulse = c.rolling_resistance * total_normal;
      float lsq = dot(ri, ri);
      bool mask = lsq > eps + max_impulse * max_impulse;
      float nrm = max_impulse / (sqrtf_(lsq) + eps);
      ri = scale(mask ? nrm : 1.0f, ri);
      out.rolling_impulse(ri);
      delta = sub(ri, old);
      ba.w = sub(ba.w, sym_mv(c.inv_ia, delta));
      bb.w = add(bb.w, sym_mv(c.inv_ib, delta));
    }
    {
      float twist_speed = dot(normal, sub(bb.w, ba.w));
      float max_lambda = c.friction * total_twist_limit;
      float old = c.twist_impulse;
      float impulse = -(c.twist_mass * twist_speed);
      impulse = maxf(-max_lambda - old, minf(impulse, max_lambda - old));
      out.twist_impulse(old + impulse);
      ba.w = sub(ba.w, scale(impulse, sym_mv(c.inv_ia, normal)));
      bb.w = add(bb.w, scale(impulse, sym_mv(c.inv_ib, normal)));
    }
    {
      Vec3 t1 = c.tangent1, t2 = c.tangent2, r_a = c.center_a, r_b = c.center_b;
      Vec3 arv = sub(cross(bb.w, r_b), cross(ba.w, r_a));
      Vec2 vt = {ltv1 + dot(arv, t1), ltv2 + dot(arv, t2)};
      Vec2 di = mul_mv2(c.tangent_mass, vt);
      di = Vec2{-di.x, -di.y};
      const Vec2 fi_old = c.friction_impulse;
      Vec2 ni = {fi_old.x + di.x, fi_old.y + di.y};
      float max_impulse = c.friction * total_normal;
      float lsq = ni.x * ni.x + ni.y * ni.y;
      bool mask = lsq > max_impulse * max_impulse;
      float nrm = max_impulse / (sqrtf_(lsq) + eps);
      float sc = mask ? nrm : 1.0f;
      ni = Vec2{sc * ni.x, sc * ni.y};
      di = Vec2{ni.x - fi_old.x, ni.y - fi_old.y};
      out.friction_impulse(ni);
      Vec3 pp = scale(di.x, t1);
      pp = add(pp, scale(di.y, t2));
      ba.w = sub(ba.w, sym_mv(c.inv_ia, cross(r_a, pp)));
      bb.w = add(bb.w, sym_mv(c.inv_ib, cross(r_b, pp)));
      pp = add(pp, scale(manifold_impulse, normal));
      ba.v = sub(ba.v, scale(c.inv_mass_a, pp));
      bb.v = add(bb.v, scale(c.inv_mass_b, pp));
    }
  }
}

template <class Hook = NoHook>
__device__ __forceinline__ void relax_item(const DeviceWorld& d, int slot, const StepParams& p, int32_t ia, int32_t ib,
                                           const float4* staged = nullptr, int lanes = 0, Hook hook = Hook{}) {
  BodyState ba = load_state(d, ia), bb = load_state(d, ib);
  const BlockedCons cons = blocked_cons(d);
  if (staged != nullptr) {
    const Constraint c = staged_record(staged, lanes, (int)threadIdx.x);
    wait_before_copy();
    hook();
    relax_core(c, con_out(cons, slot), p, ba, bb);
  } else {
    hook();
    relax_core(cons.load(slot), con_out(cons, slot), p, ba, bb);
  }
  store_state(d, ia, ba);
  store_state(d, ib, bb);
}

// The residual sweep's output: relax_core's impulse writes are not kept; each is compared
// with the record's impulse instead. Per constraint: the sum and the largest |change| of
// its points' normal impulses and the length of its friction impulse's change.
struct ResidualOut {
  const Constraint* c;
  float* sum_normal;
  float* max_normal;
  float* friction;
  __device__ __forceinline__ void normal_impulse(int i, float v) const {
    const float r = fabsf(v - c->points[i].normal_impulse);
    *sum_normal += r;
    *max_normal = fmaxf(*max_normal, r);
  }
  __device__ __forceinline__ void total_normal_impulse(int, float) const {}
  __device__ __forceinline__ void twist_impulse(float) const {}
  __device__ __forceinline__ void friction_impulse(Vec2 v) const {
    const float dx = v.x - c->friction_impulse.x, dy = v.y - c->friction_impulse.y;
    *friction = sqrtf_(dx * dx + dy * dy);
  }
  __device__ __forceinline__ void rolling_impulse(Vec3) const {}
};

struct Residual {
  float sum_normal = 0.0f, max_normal = 0.0f, friction = 0.0f;
};

// One constraint of the residual sweep: relax (contact_solver.rs:927-1097) from the record's
// final impulses on the body states of `d` (a copy), which it updates; the record is not
// written.
__device__ __forceinline__ Residual residual_item(const DeviceWorld& d, int slot, const StepParams& p) {
  const Constraint c = blocked_cons(d).load(slot);
  BodyState ba = load_state(d, c.index_a), bb = load_state(d, c.index_b);
  Residual r;
  relax_core(c, ResidualOut{&c, &r.sum_normal, &r.max_normal, &r.friction}, p, ba, bb);
  store_state(d, c.index_a, ba);
  store_state(d, c.index_b, bb);
  return r;
}

// The three solve stages on a record the caller already holds (a copy of slot's record,
// loaded early); impulses are written to the slot's record.
__device__ __forceinline__ void warm_start_record(const DeviceWorld& d, const Constraint& c) {
  BodyState ba = load_state(d, c.index_a), bb = load_state(d, c.index_b);
  const WarmImpulse wi = warm_impulse(c);
  ba.w = sub(ba.w, sym_mv(c.inv_ia, wi.ang_a));
  ba.v = sub(ba.v, scale(c.inv_mass_a, wi.linear));
  bb.w = add(bb.w, sym_mv(c.inv_ib, wi.ang_b));
  bb.v = add(bb.v, scale(c.inv_mass_b, wi.linear));
  store_state(d, c.index_a, ba);
  store_state(d, c.index_b, bb);
}
__device__ __forceinline__ void push_record(const DeviceWorld& d, int slot, const StepParams& p, const Constraint& c) {
  BodyState ba = load_state(d, c.index_a), bb = load_state(d, c.index_b);
  push_core(c, con_out(blocked_cons(d), slot), p, ba, bb);
  store_state(d, c.index_a, ba);
  store_state(d, c.index_b, bb);
}
__device__ __forceinline__ void relax_record(const DeviceWorld& d, int slot, const StepParams& p, const Constraint& c) {
  BodyState ba = load_state(d, c.index_a), bb = load_state(d, c.index_b);
  relax_core(c, con_out(blocked_cons(d), slot), p, ba, bb);
  store_state(d, c.index_a, ba);
  store_state(d, c.index_b, bb);
}

// contact_solver.rs:1135-1229
template <class Cons>
__device__ __forceinline__ void restitution_item(const DeviceWorld& d, Cons cons, int slot, const StepParams& p, int32_t ia,
                                                 int32_t ib) {
  const Constraint c = cons.load(slot);
  const ConOut<Cons> out = con_out(cons, slot);
  const bool propagate = p.enable_restitution_propagation != 0;
  if (!propagate && c.restitution == 0.0f) return;
  BodyState ba = load_state(d, ia), bb = load_state(d, ib);
  const int pc = c.point_count;
  ConstraintPoint pt[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pt[i] = c.points[i];
  const bool rmask = c.restitution > 0.0f;
  const float neg_threshold = -p.restitution_threshold;
  Vec3 dp = sub(bb.dp, ba.dp);
  if (c.manifold_count > 0) {
    Vec3 normal = c.normal;
    float normal_sep = dot(normal, dp);
    Vec3 n_a = inv_rotate_wide(ba.dq, normal), n_b = inv_rotate_wide(bb.dq, normal);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= pc) break;
      const ConstraintPoint& cp = pt[i];
      float s = (normal_sep + (dot(n_b, cp.r_b) - dot(n_a, cp.r_a))) + cp.base_separation;
      float normal_mass = propagate ? cp.normal_mass : (rmask ? cp.normal_mass : 0.0f);
      float compression = cp.total_normal_impulse - cp.restitution_impulse;
      bool armed = rmask && cp.relative_velocity < neg_threshold && compression > 0.0f;
      float velocity_bias = armed ? c.restitution * cp.relative_velocity : maxf(0.0f, s * p.inv_h);
      Vec3 vr_a = add(ba.v, cross(ba.w, cp.r_a)), vr_b = add(bb.v, cross(bb.w, cp.r_b));
      float v_n = dot(sub(vr_b, vr_a), normal);
      float neg_impulse = normal_mass * (v_n + velocity_bias);
      float new_impulse = maxf(cp.normal_impulse - neg_impulse, 0.0f);
      float impulse = new_impulse - cp.normal_impulse;
      float approach = minf(maxf(-normal_mass * v_n, 0.0f), maxf(impulse, 0.0f));
      float allowance = c.restitution * (compression + approach) - cp.restitution_impulse;
      float max_impulse = approach + maxf(allowance, 0.0f);
      impulse = armed ? minf(impulse, max_impulse) : impulse;
      out.normal_impulse(i, cp.normal_impulse + impulse);
      out.restitution_impulse(i, cp.restitution_impulse + (impulse - approach));
      out.total_normal_impulse(i, cp.total_normal_impulse + impulse);
      Vec3 pp = scale(impulse, normal);
      ba.w = sub(ba.w, sym_mv(c.inv_ia, cross(cp.r_a, pp)));
      ba.v = sub(ba.v, scale(c.inv_mass_a, pp));
      bb.w = add(bb.w, sym_mv(c.inv_ib, cross(cp.r_b, pp)));
      bb.v = add(bb.v, scale(c.inv_mass_b, pp));
    }
  }
  store_state(d, ia, ba);
  store_state(d, ib, bb);
}

// contact_solver.rs:1258-1299
template <class Cons>
__device__ __forceinline__ void store_item(const DeviceWorld& d, Cons cons, int slot) {
  const Constraint c = cons.load(slot);
  Contact& contact = d.contacts[c.contact_id];
  if (contact.manifold_count <= 0 || c.manifold_count <= 0) return;
  bool hit = (contact.flags & cflag::kHitEvent) != 0;
  int count = mini(contact.point_count, c.point_count);
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // constant indices: the record stays in registers
    if (i >= count) break;
    contact.points[i].normal_impulse = c.points[i].normal_impulse;
    contact.points[i].total_normal_impulse = c.points[i].total_normal_impulse;
    if (hit) contact.points[i].normal_velocity = c.points[i].relative_velocity;
  }
  contact.twist_impulse = c.twist_impulse;
  contact.friction_impulse = add(scale(c.friction_impulse.x, c.tangent1), scale(c.friction_impulse.y, c.tangent2));
  contact.rolling_impulse = c.rolling_impulse;
}

// ------------------------------------------------------------------------------
// overflow colour: sequential in graph order (overflow_contact_solver.rs:259-722).
// One lane walks the list; these constraints are rare.

__device__ __forceinline__ void overflow_store_velocity(const DeviceWorld& d, int32_t i, uint32_t flags, Vec3 v, Vec3 w) {
  if (i >= 0 && (flags & bflag::kDynamic) != 0) {
    d.states[i].v = v;
    d.states[i].w = w;
  }
}

// Warm start of overflow constraint k.
__device__ inline void overflow_warm_item(const DeviceWorld& d, int k) {
  {
    const OverflowConstraint& c = d.overflow[k];
    Vec3 linear = zero3(), ang_a = zero3(), ang_b = zero3();
    if (c.manifold_count > 0) {
      Vec3 normal = c.normal;
      float total = 0.0f;
      Vec3 moment_a = zero3(), moment_b = zero3();
      for (int j = 0; j < c.point_count; ++j) {
        total += c.points[j].normal_impulse;
        moment_a = mul_add(moment_a, c.points[j].normal_impulse, c.points[j].r_a);
        moment_b = mul_add(moment_b, c.points[j].normal_impulse, c.points[j].r_b);
      }
      Vec3 fi = scale(c.friction_impulse.x, c.tangent1);
      fi = mul_add(fi, c.friction_impulse.y, c.tangent2);
      linear = add(linear, mul_add(fi, total, normal));
      Vec3 twist = scale(c.twist_impulse, normal);
      Vec3 a_a = add(add(cross(moment_a, normal), cross(c.center_a, fi)), twist);
      Vec3 a_b = add(add(cross(moment_b, normal), cross(c.center_b, fi)), twist);
      ang_a = add(ang_a, add(a_a, c.rolling_impulse));
      ang_b = add(ang_b, add(a_b, c.rolling_impulse));
    }
    if (c.index_a >= 0 && (d.states[c.index_a].flags & bflag::kDynamic)) {
      BodyState& s = d.states[c.index_a];
      s.v = mul_sub(s.v, c.inv_mass_a, linear);
      s.w = sub(s.w, mul_mv(c.inv_ia, ang_a));
    }
    if (c.index_b >= 0 && (d.states[c.index_b].flags & bflag::kDynamic)) {
      BodyState& s = d.states[c.index_b];
      s.v = mul_add(s.v, c.inv_mass_b, linear);
      s.w = add(s.w, mul_mv(c.inv_ib, ang_b));
    }
  }
}

// Overflow constraint k; mode 0: push (biased), 1: relax, 2: restitution.
__device__ inline void overflow_solve_item(const DeviceWorld& d, int k, const StepParams& p, int mode) {
  const float inv_h = p.inv_h;
  {
    OverflowConstraint& c = d.overflow[k];
    if (mode == 2 && p.enable_restitution_propagation == 0 && c.restitution == 0.0f) return;
    float mA = c.inv_mass_a, mB = c.inv_mass_b;
    Mat3 i_a = c.inv_ia, i_b = c.inv_ib;
    BodyState sa = load_state(d, c.index_a), sb = load_state(d, c.index_b);
    Vec3 v_a = sa.v, w_a = sa.w, v_b = sb.v, w_b = sb.w;
    Quat dq_a = sa.dq, dq_b = sb.dq;
    Vec3 dp = sub(sb.dp, sa.dp);
    if (c.manifold_count > 0) {
      Vec3 normal = c.normal;
      float total_normal = 0.0f, total_twist_limit = 0.0f;
      for (int j = 0; j < c.point_count; ++j) {
        ConstraintPoint& cp = c.points[j];
        Vec3 r_a = cp.r_a, r_b = cp.r_b;
        if (mode == 0) {
          Vec3 ds = add(dp, sub(rotate(dq_b, r_b), rotate(dq_a, r_a)));
          float s = dot(ds, normal) + cp.base_separation;
          float velocity_bias, mass_scale, impulse_scale;
          if (s > 0.0f) {
            velocity_bias = s * inv_h; mass_scale = 1.0f; impulse_scale = 0.0f;
          } else {
            velocity_bias = maxf(c.softness.mass_scale * c.softness.bias_rate * s, -p.contact_speed);
            mass_scale = c.softness.mass_scale;
            impulse_scale = c.softness.impulse_scale;
          }
          Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
          float vn = dot(sub(vr_b, vr_a), normal);
          float ni = cp.normal_impulse;
          float delta = -cp.normal_mass * (mass_scale * vn + velocity_bias) - impulse_scale * ni;
          float new_impulse = maxf(ni + delta, 0.0f);
          delta = new_impulse - ni;
          cp.normal_impulse = new_impulse;
          Vec3 pp = scale(delta, normal);
          v_a = mul_sub(v_a, mA, pp);
          w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
          v_b = mul_add(v_b, mB, pp);
          w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
        } else if (mode == 1) {
          Vec3 ds = add(dp, sub(rotate(dq_b, r_b), rotate(dq_a, r_a)));
          float s = dot(ds, normal) + cp.base_separation;
          float velocity_bias = s > 0.0f ? s * inv_h : 0.0f;
          Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
          float vn = dot(sub(vr_b, vr_a), normal);
          float ni = cp.normal_impulse;
          float delta = -cp.normal_mass * (vn + velocity_bias);
          float new_impulse = maxf(ni + delta, 0.0f);
          delta = new_impulse - ni;
          cp.normal_impulse = new_impulse;
          cp.total_normal_impulse += new_impulse;
          total_normal += new_impulse;
          total_twist_limit += cp.lever_arm * new_impulse;
          Vec3 pp = scale(delta, normal);
          v_a = mul_sub(v_a, mA, pp);
          w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
          v_b = mul_add(v_b, mB, pp);
          w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
        } else {
          float restitution = c.restitution;
          float compression = cp.total_normal_impulse - cp.restitution_impulse;
          bool armed = restitution > 0.0f && cp.relative_velocity < -p.restitution_threshold && compression > 0.0f;
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
          cp.normal_impulse += impulse;
          cp.restitution_impulse += impulse - approach;
          cp.total_normal_impulse += impulse;
          Vec3 pp = scale(impulse, normal);
          v_a = mul_sub(v_a, mA, pp);
          w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
          v_b = mul_add(v_b, mB, pp);
          w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
        }
      }
      if (mode == 1) {
        {
          float twist_speed = dot(normal, sub(w_b, w_a));
          float max_impulse = c.friction * total_twist_limit;
          float delta = -c.twist_mass * twist_speed;
          float old = c.twist_impulse;
          c.twist_impulse = clampf(old + delta, -max_impulse, max_impulse);
          delta = c.twist_impulse - old;
          w_a = sub(w_a, mul_mv(i_a, scale(delta, normal)));
          w_b = add(w_b, mul_mv(i_b, scale(delta, normal)));
        }
        if (c.rolling_resistance > 0.0f) {
          Vec3 delta = neg(mul_mv(c.rolling_mass, sub(w_b, w_a)));
          Vec3 old = c.rolling_impulse;
          c.rolling_impulse = add(old, delta);
          float max_impulse = c.rolling_resistance * total_normal;
          float mag_sq = dot(c.rolling_impulse, c.rolling_impulse);
          if (mag_sq > max_impulse * max_impulse + kFltEps) c.rolling_impulse = scale(max_impulse / sqrtf_(mag_sq), c.rolling_impulse);
          delta = sub(c.rolling_impulse, old);
          w_a = sub(w_a, mul_mv(i_a, delta));
          w_b = add(w_b, mul_mv(i_b, delta));
        }
        {
          Vec3 t1 = c.tangent1, t2 = c.tangent2, r_a = c.center_a, r_b = c.center_b;
          Vec3 vr_a = add(v_a, cross(w_a, r_a)), vr_b = add(v_b, cross(w_b, r_b));
          Vec3 vr = sub(vr_b, vr_a);
          Vec2 vt = {dot(vr, t1) - c.tangent_velocity1, dot(vr, t2) - c.tangent_velocity2};
          Vec2 tm = mul_mv2(c.tangent_mass, vt);
          Vec2 ni = {c.friction_impulse.x + -tm.x, c.friction_impulse.y + -tm.y};
          float max_impulse = c.friction * total_normal;
          float lsq = ni.x * ni.x + ni.y * ni.y;
          if (lsq > max_impulse * max_impulse) {
            float sc = max_impulse / sqrtf_(lsq);
            ni.x *= sc;
            ni.y *= sc;
          }
          Vec2 di = {ni.x - c.friction_impulse.x, ni.y - c.friction_impulse.y};
          c.friction_impulse = ni;
          Vec3 pp = blend2(di.x, t1, di.y, t2);
          v_a = mul_sub(v_a, mA, pp);
          w_a = sub(w_a, mul_mv(i_a, cross(r_a, pp)));
          v_b = mul_add(v_b, mB, pp);
          w_b = add(w_b, mul_mv(i_b, cross(r_b, pp)));
        }
      }
    }
    overflow_store_velocity(d, c.index_a, sa.flags, v_a, w_a);
    overflow_store_velocity(d, c.index_b, sb.flags, v_b, w_b);
  }
}

// The overflow list in order by one lane (overflow_contact_solver.rs: sequential).
__device__ void overflow_warm_all(const DeviceWorld& d, int n) {
  for (int k = 0; k < n; ++k) overflow_warm_item(d, k);
}

__device__ void overflow_solve_all(const DeviceWorld& d, int n, const StepParams& p, int mode) {
  for (int k = 0; k < n; ++k) overflow_solve_item(d, k, p, mode);
}

__device__ __forceinline__ void overflow_store_item(const DeviceWorld& d, int k) {
  const OverflowConstraint& c = d.overflow[k];
  if (c.manifold_count == 0) return;
  Contact& m = d.contacts[c.contact_id];
  m.twist_impulse = c.twist_impulse;
  m.friction_impulse = blend2(c.friction_impulse.x, c.tangent1, c.friction_impulse.y, c.tangent2);
  m.rolling_impulse = c.rolling_impulse;
  for (int i = 0; i < c.point_count; ++i) {
    m.points[i].normal_impulse = c.points[i].normal_impulse;
    m.points[i].total_normal_impulse = c.points[i].total_normal_impulse;
  }
}

// ------------------------------------------------------------------------------
// finalize (solver_integration.rs:587-799)

// compute_shape_aabb (shape_runtime.rs:370-387) at a transform
__device__ Aabb shape_aabb_at(const DeviceWorld& d, const Shape& s, Xf t) {
  if (s.kind == kShapeHull) return aabb_transform(t, d.hulls[s.hull].aabb);
  float r = s.radius;
  Vec3 e = {r, r, r};
  if (s.kind == kShapeSphere) {
    Vec3 c = xf_point(t, s.c1);
    return Aabb{sub(c, e), add(c, e)};
  }
  Vec3 c1 = xf_point(t, s.c1), c2 = xf_point(t, s.c2);
  return Aabb{sub(min3(c1, c2), e), add(max3(c1, c2), e)};
}
__device__ __forceinline__ Aabb shape_local_aabb(const DeviceWorld& d, const Shape& s, Quat q) {
  return shape_aabb_at(d, s, Xf{zero3(), q});
}

__device__ __forceinline__ void finalize_item(const DeviceWorld& d, int i, const StepParams& p) {
  BodyState state = d.states[i];
  BodySim sim = d.sims[i];
  const float time_step = p.dt, inv_time_step = p.inv_dt;
  Vec3 v = state.v, w = state.w;
  Vec3 local_omega = inv_rotate(sim.transform.q, w);
  Vec3 local_dr = inv_rotate(sim.transform.q, state.dq.v);
  sim.center = offset_pos(sim.center, state.dp);
  sim.transform.q = normalize_quat(mul_quat(state.dq, sim.transform.q));
  float max_velocity = length(v) + length(modified_cross(abs3(local_omega), sim.max_extent));
  float max_delta_position = length(state.dp) + 2.0f * length(modified_cross(abs3(local_dr), sim.max_extent));
  float sleep_velocity = maxf(max_velocity, 0.5f * inv_time_step * max_delta_position);
  state.dp = zero3();
  state.dq = quat_identity();
  sim.transform.p = offset_pos(sim.center, neg(rotate(sim.transform.q, sim.local_center)));
  sim.sleep_velocity = sleep_velocity;
  sim.force = zero3();
  sim.torque = zero3();
  sim.body_flags &= ~bflag::kTransient;
  sim.flags &= ~(bflag::kIsFast | bflag::kTransient);
  sim.body_flags |= state.flags & bflag::kSpeedCapped;
  state.flags &= ~bflag::kTransient;
  if (!p.enable_sleep || (sim.body_flags & bflag::kEnableSleep) == 0 || sleep_velocity > sim.sleep_threshold) {
    sim.sleep_time = 0.0f;
    float max_motion = maxf(max_delta_position, max_velocity * time_step);
    if (sim.body_type == kBodyDynamic && p.enable_continuous && max_motion > sim.safety_factor * sim.min_extent) {
      sim.flags |= bflag::kIsFast;
      atomicAdd(&d.counters[kCntFastBodies], 1);
      // the continuous pass sweeps it and updates its bounds (solver_integration.rs:696-710)
      d.fast_ids[atomicAdd(&d.counters[kCntFastList], 1)] = d.fast_base + i;
    } else {
      sim.center0 = sim.center;
      sim.rotation0 = sim.transform.q;
    }
  } else {
    sim.center0 = sim.center;
    sim.rotation0 = sim.transform.q;
    sim.sleep_time += time_step;
  }
  if (d.sleep_times) d.sleep_times[i] = sim.sleep_time;
  Mat3 rm = matrix_from_quat(sim.transform.q);
  sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
  if ((sim.flags & bflag::kIsFast) == 0) {
    for (int32_t sid = sim.head_shape; sid != kNullIndex; sid = d.shapes[sid].next_shape) {
      const Shape& s = d.shapes[sid];
      Aabb aabb = offset_aabb(aabb_inflate(shape_local_aabb(d, s, sim.transform.q), kSpeculativeDistance), sim.transform.p);
      d.shape_aabbs[sid] = aabb;
      if (!aabb_contains(d.fat_aabbs[sid], aabb)) {
        d.fat_aabbs[sid] = aabb_inflate(aabb, s.aabb_margin);
        d.shape_moved[sid] = 1;
        atomicAdd(&d.counters[kCntMovedProxies], 1);
      }
    }
  }
  d.states[i] = state;
  d.sims[i] = sim;
}

}  // namespace
}  // namespace nsl
