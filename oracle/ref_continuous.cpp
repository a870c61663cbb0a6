// Oracle: continuous collision of fast bodies (time of impact against other shapes).
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// Restates src/collision/time_of_impact.rs:1-598 (separation functions over two sweeps
// and the push-back root finder), src/collision/distance.rs:817-832 (sweep transforms),
// src/simulation/continuous.rs:121-621 (is_shape_fast, the swept query callback and
// solve_continuous) and the two places the step runs it: fast bodies in body order inside
// finalize (solver_integration.rs:538-566), then bullets (solver_integration.rs:948-995).
//
// Candidate order: the reference visits the shapes its broad-phase trees return for the
// swept box; here they are visited in shape id order (static shapes, then for bullets the
// kinematic and dynamic ones). A candidate's query is bounded by the earliest impact found
// so far, so when two candidates both hit within one step the last bits of the fraction
// can depend on that order. With one candidate hit per step the results do not.
#include "ref_world.h"

namespace nslref {

namespace {

// distance.rs:817-832
Xf sweep_xf(const Sweep& s, float t) {
  Quat q = nlerp(s.q1, s.q2, t);
  return Xf{sub(lerp(s.c1, s.c2, t), rotate(q, s.local_center)), q};
}
Xf sweep_final_xf(const Sweep& s) { return Xf{sub(s.c2, rotate(s.q2, s.local_center)), s.q2}; }

enum SepKind : int { kSepUnknown = 0, kSepVertices, kSepEdges, kSepFaceA, kSepFaceB };

struct SepFn {
  const ToiProxy* pa;
  const ToiProxy* pb;
  Sweep sa, sb;
  Vec3 w1, w2;  // axis and point (faces), local edges (edges) or world axis (vertices)
  int kind;
};

// time_of_impact.rs:36-63
int unique_count(int n, const int* v) {
  if (n == 1) return 1;
  if (n == 2) return v[0] != v[1] ? 2 : 1;
  if (v[0] != v[1] && v[0] != v[2] && v[1] != v[2]) return 3;
  if (v[0] == v[1] && v[0] == v[2] && v[1] == v[2]) return 1;
  return 2;
}

// time_of_impact.rs:65-77: the edge cross product flips direction over the sweep
bool edges_flip(Xf xa, Vec3 local_edge_a, Xf xb, Vec3 local_edge_b, Vec3 axis0) {
  Vec3 axis = cross(rotate(xa.q, local_edge_a), rotate(xb.q, local_edge_b));
  return dot(axis, axis0) < 0.0f;
}

// Edge pair axis, or the fallbacks to a vertex axis (time_of_impact.rs:133-166, :213-258).
void edge_pair_axis(SepFn& f, Vec3 va1, Vec3 va2, Vec3 vb1, Vec3 vb2, Quat qa, Quat qb, Vec3 dp, Vec3 world_normal,
                    float tolerance_sq) {
  Vec3 local_edge_a = normalize(sub(va2, va1));
  Vec3 edge_a = rotate(qa, local_edge_a);
  Vec3 local_edge_b = normalize(sub(vb2, vb1));
  Vec3 edge_b = rotate(qb, local_edge_b);
  Vec3 axis = cross(edge_a, edge_b);
  if (length_sq(axis) < tolerance_sq) {
    f.kind = kSepVertices;
    f.w1 = world_normal;
    return;
  }
  Vec3 delta = add(sub(rotate(qb, vb1), rotate(qa, va1)), dp);
  if (dot(delta, axis) < 0.0f) {
    axis = neg(axis);
    local_edge_b = neg(local_edge_b);
  }
  if (edges_flip(sweep_final_xf(f.sa), local_edge_a, sweep_final_xf(f.sb), local_edge_b, axis)) {
    f.kind = kSepVertices;
    f.w1 = normalize(axis);
  } else {
    f.kind = kSepEdges;
    f.w1 = local_edge_a;
    f.w2 = local_edge_b;
  }
}

// time_of_impact.rs:79-263
SepFn make_separation(const SimplexCache& cache, const ToiProxy& pa, const Sweep& sa, const ToiProxy& pb,
                      const Sweep& sb, Vec3 world_normal, float t1) {
  SepFn f{&pa, &pb, sa, sb, kZero3, kZero3, kSepUnknown};
  int ia[3] = {cache.index_a[0], cache.index_a[1], cache.index_a[2]};
  int ib[3] = {cache.index_b[0], cache.index_b[1], cache.index_b[2]};
  const int ua = unique_count(cache.count, ia), ub = unique_count(cache.count, ib);
  Xf xa = sweep_xf(sa, t1), xb = sweep_xf(sb, t1);
  const Quat qa = xa.q, qb = xb.q;
  const Vec3 dp = sub(xb.p, xa.p);
  const Vec3* A = pa.points;
  const Vec3* B = pb.points;
  if (cache.count == 1) {
    f.kind = kSepVertices;
    f.w1 = world_normal;
  } else if (cache.count == 2) {
    if (ua == 2 && ub == 2) {
      edge_pair_axis(f, A[ia[0]], A[ia[1]], B[ib[0]], B[ib[1]], qa, qb, dp, world_normal, 0.05f * 0.05f);
    } else {
      f.kind = kSepVertices;
      f.w1 = world_normal;
    }
  } else if (ua == 3) {
    Vec3 va1 = A[ia[0]], va2 = A[ia[1]], va3 = A[ia[2]];
    Vec3 local_axis = normalize(cross(sub(va2, va1), sub(va3, va1)));
    Vec3 axis = rotate(qa, local_axis);
    Vec3 local_point_a = scale(1.0f / 3.0f, add(add(va1, va2), va3));
    Vec3 local_point_b = B[ib[0]];
    Vec3 delta = add(sub(rotate(qb, local_point_b), rotate(qa, local_point_a)), dp);
    if (dot(delta, axis) < 0.0f) local_axis = neg(local_axis);
    f.kind = kSepFaceA;
    f.w1 = local_axis;
    f.w2 = local_point_a;
  } else if (ub == 3) {
    Vec3 vb1 = B[ib[0]], vb2 = B[ib[1]], vb3 = B[ib[2]];
    Vec3 local_axis = normalize(cross(sub(vb2, vb1), sub(vb3, vb1)));
    Vec3 axis = rotate(qb, local_axis);
    Vec3 local_point_a = A[ia[0]];
    Vec3 local_point_b = scale(1.0f / 3.0f, add(add(vb1, vb2), vb3));
    Vec3 delta = sub(sub(rotate(qa, local_point_a), rotate(qb, local_point_b)), dp);
    if (dot(delta, axis) < 0.0f) local_axis = neg(local_axis);
    f.kind = kSepFaceB;
    f.w1 = local_axis;
    f.w2 = local_point_b;
  } else {
    if (ia[0] == ia[1]) ia[1] = ia[2];
    if (ib[0] == ib[1]) ib[1] = ib[2];
    edge_pair_axis(f, A[ia[0]], A[ia[1]], B[ib[0]], B[ib[1]], qa, qb, dp, world_normal, 0.005f * 0.005f);
  }
  return f;
}

// time_of_impact.rs:266-344: the deepest points along the axis at time t
float find_min_separation(const SepFn& f, int* ia, int* ib, float t) {
  Xf xa = sweep_xf(f.sa, t), xb = sweep_xf(f.sb, t);
  const Vec3* A = f.pa->points;
  const Vec3* B = f.pb->points;
  switch (f.kind) {
    case kSepVertices: {
      Vec3 axis = f.w1;
      *ia = gjk_point_support(A, f.pa->count, inv_rotate(xa.q, axis));
      *ib = gjk_point_support(B, f.pb->count, inv_rotate(xb.q, neg(axis)));
      Vec3 dp = sub(xb.p, xa.p);
      Vec3 delta = add(sub(rotate(xb.q, B[*ib]), rotate(xa.q, A[*ia])), dp);
      return dot(delta, axis);
    }
    case kSepEdges: {
      Vec3 axis = normalize(cross(rotate(xa.q, f.w1), rotate(xb.q, f.w2)));
      *ia = gjk_point_support(A, f.pa->count, inv_rotate(xa.q, axis));
      *ib = gjk_point_support(B, f.pb->count, neg(inv_rotate(xb.q, axis)));
      Vec3 dp = sub(xb.p, xa.p);
      Vec3 delta = add(sub(rotate(xb.q, B[*ib]), rotate(xa.q, A[*ia])), dp);
      return dot(delta, axis);
    }
    case kSepFaceA: {
      Vec3 normal = rotate(xa.q, f.w1);
      *ia = -1;
      Vec3 point_a = xf_point(xa, f.w2);
      *ib = gjk_point_support(B, f.pb->count, neg(inv_rotate(xb.q, normal)));
      Vec3 point_b = xf_point(xb, B[*ib]);
      return dot(sub(point_b, point_a), normal);
    }
    case kSepFaceB: {
      Vec3 normal = rotate(xb.q, f.w1);
      *ia = gjk_point_support(A, f.pa->count, neg(inv_rotate(xa.q, normal)));
      Vec3 point_a = xf_point(xa, A[*ia]);
      *ib = -1;
      Vec3 point_b = xf_point(xb, f.w2);
      return dot(sub(point_a, point_b), normal);
    }
    default:
      return 0.0f;
  }
}

// time_of_impact.rs:347-386: separation of fixed witness points at time beta
float evaluate_separation(const SepFn& f, int i1, int i2, float beta) {
  Xf x1 = sweep_xf(f.sa, beta), x2 = sweep_xf(f.sb, beta);
  const Vec3* A = f.pa->points;
  const Vec3* B = f.pb->points;
  switch (f.kind) {
    case kSepVertices:
      return dot(sub(xf_point(x2, B[i2]), xf_point(x1, A[i1])), f.w1);
    case kSepEdges: {
      Vec3 axis = normalize(cross(rotate(x1.q, f.w1), rotate(x2.q, f.w2)));
      return dot(sub(xf_point(x2, B[i2]), xf_point(x1, A[i1])), axis);
    }
    case kSepFaceA: {
      Vec3 axis = rotate(x1.q, f.w1);
      return dot(sub(xf_point(x2, B[i2]), xf_point(x1, f.w2)), axis);
    }
    case kSepFaceB: {
      Vec3 axis = rotate(x2.q, f.w1);
      return dot(sub(xf_point(x1, A[i1]), xf_point(x2, f.w2)), axis);
    }
    default:
      return 0.0f;
  }
}

// time_of_impact.rs:389-403
void force_fixed_axis(SepFn& f, float beta) {
  Xf x1 = sweep_xf(f.sa, beta), x2 = sweep_xf(f.sb, beta);
  f.kind = kSepVertices;
  f.w1 = normalize(cross(rotate(x1.q, f.w1), rotate(x2.q, f.w2)));
  f.w2 = kZero3;
}

}  // namespace

// time_of_impact.rs:406-598. Conservative advancement: at each GJK distance, build a
// separating axis from the simplex and push t forward to where the deepest points along it
// reach the target separation (linear slop short of touching).
ToiOutput time_of_impact(ToiProxy proxy_a, Sweep sweep_a, ToiProxy proxy_b, Sweep sweep_b, float max_fraction) {
  ToiOutput out;
  const Vec3 origin = sweep_a.c1;
  sweep_a.c1 = kZero3;
  sweep_a.c2 = sub(sweep_a.c2, origin);
  sweep_b.c1 = sub(sweep_b.c1, origin);
  sweep_b.c2 = sub(sweep_b.c2, origin);
  const int max_push_back = proxy_a.count + proxy_b.count;
  const float t_max = max_fraction;
  const float slop = linear_slop();
  const float total_radius = proxy_a.radius + proxy_b.radius;
  const float target = maxf(slop, total_radius - slop);
  const float tolerance = 0.25f * slop;
  float t1 = 0.0f;
  const int max_iterations = 25;
  int distance_iterations = 0;
  SimplexCache cache{};
  for (;;) {
    Xf xa = sweep_xf(sweep_a, t1), xb = sweep_xf(sweep_b, t1);
    DistanceOutput d = gjk_distance(proxy_a.points, proxy_a.count, proxy_b.points, proxy_b.count, inv_mul_xf(xa, xb), cache);
    out.distance = d.distance;
    const Vec3 world_normal = rotate(xa.q, d.normal);
    const Vec3 world_point_a = xf_point(xa, d.point_a);
    const Vec3 world_point_b = xf_point(xa, d.point_b);
    out.distance_iterations += 1;
    distance_iterations += 1;
    auto finish = [&](int state, float fraction) {
      out.state = state;
      out.fraction = fraction;
      Vec3 p_a = mul_add(world_point_a, proxy_a.radius, world_normal);
      Vec3 p_b = mul_add(world_point_b, -proxy_b.radius, world_normal);
      out.point = add(lerp(p_a, p_b, 0.5f), origin);
      out.normal = world_normal;
    };
    if (d.distance <= 0.0f) {
      finish(kToiOverlapped, 0.0f);
      break;
    }
    if (d.distance <= target + tolerance) {
      finish(kToiHit, t1);
      break;
    }
    if (distance_iterations == max_iterations) {
      finish(kToiFailed, t1);
      break;
    }
    SepFn f = make_separation(cache, proxy_a, sweep_a, proxy_b, sweep_b, world_normal, t1);
    bool done = false;
    int done_state = kToiUnknown;
    float done_fraction = 0.0f;
    float t2 = t_max;
    int push_back = 0;
    for (;;) {
      int ia = 0, ib = 0;
      float s2 = find_min_separation(f, &ia, &ib, t2);
      if (s2 - target > tolerance) {  // separated at the end of the bracket
        done = true;
        done_state = kToiSeparated;
        done_fraction = max_fraction;
        break;
      }
      if (s2 >= target - tolerance) {  // advance to t2
        t1 = t2;
        break;
      }
      float s1 = evaluate_separation(f, ia, ib, t1);
      if (s1 < target - tolerance) {
        done = true;
        done_state = kToiFailed;
        done_fraction = t1;
        break;
      }
      if (s1 <= target + tolerance) {
        done = true;
        done_state = kToiHit;
        done_fraction = t1;
        break;
      }
      // root of s(t) = target in [t1, t2]: bisection alternating with the secant rule
      int root_count = 0;
      const int max_root = 50;
      float a1 = t1, a2 = t2;
      for (;;) {
        float t = (root_count & 1) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : 0.5f * (a1 + a2);
        out.root_iterations += 1;
        root_count += 1;
        float s = evaluate_separation(f, ia, ib, t);
        if (absf(s - target) <= tolerance) {
          t2 = t;
          break;
        }
        if (s > target) {
          a1 = t;
          s1 = s;
        } else {
          a2 = t;
          s2 = s;
        }
        if (root_count == max_root) break;
      }
      if (root_count == max_root - 1 && f.kind == kSepEdges) {
        t2 = max_fraction;
        force_fixed_axis(f, t1);
      }
      out.push_back_iterations += 1;
      push_back += 1;
      if (push_back == max_push_back) break;
    }
    if (done) {
      finish(done_state, done_fraction);
      break;
    }
  }
  return out;
}

// ---------------------------------------------------------------------------------
// continuous.rs:121-621 on the world

// shape_runtime.rs:643-658
ToiProxy World::shape_proxy(const Shape& s) const {
  ToiProxy p{};
  if (s.type == kSphere) {
    p.points[0] = s.sphere.center;
    p.count = 1;
    p.radius = s.sphere.radius;
  } else if (s.type == kCapsule) {
    p.points[0] = s.capsule.c1;
    p.points[1] = s.capsule.c2;
    p.count = 2;
    p.radius = s.capsule.radius;
  } else {
    const Hull& h = hulls[s.hull];
    for (int i = 0; i < h.vertex_count; ++i) p.points[i] = h.points[i];
    p.count = h.vertex_count;
    p.radius = 0.0f;
  }
  return p;
}

// shape_runtime.rs:434-450
Vec3 World::shape_centroid(const Shape& s) const {
  if (s.type == kSphere) return s.sphere.center;
  if (s.type == kCapsule) return lerp(s.capsule.c1, s.capsule.c2, 0.5f);
  return hulls[s.hull].center;
}

// continuous.rs:121-151 and compute_shape_min_extent (shape_runtime.rs:547-554)
bool World::is_shape_fast(const Shape& s, Vec3 centroid1, Vec3 centroid2, float rotation_chord, float safety_factor) const {
  const Vec3 lc = shape_centroid(s);
  float radius, min_extent;
  if (s.type == kSphere) {
    radius = s.sphere.radius;
    min_extent = s.sphere.radius;
  } else if (s.type == kCapsule) {
    radius = 0.5f * distance(s.capsule.c1, s.capsule.c2) + s.capsule.radius;
    min_extent = s.capsule.radius;
  } else {
    const Hull& h = hulls[s.hull];
    // internal_math.rs:426-432
    Vec3 far{(lc.x - h.aabb.lo.x) > (h.aabb.hi.x - lc.x) ? h.aabb.lo.x : h.aabb.hi.x,
             (lc.y - h.aabb.lo.y) > (h.aabb.hi.y - lc.y) ? h.aabb.lo.y : h.aabb.hi.y,
             (lc.z - h.aabb.lo.z) > (h.aabb.hi.z - lc.z) ? h.aabb.lo.z : h.aabb.hi.z};
    radius = distance(far, lc);
    min_extent = h.inner_radius;
  }
  float max_motion = distance(centroid1, centroid2) + rotation_chord * radius;
  return max_motion > safety_factor * min_extent;
}

// continuous.rs:337-621 for the awake body at sim_index (a fast body or a bullet).
void World::solve_continuous(int32_t sim_index, float dt) {
  BodySim fast = awake_sims[(size_t)sim_index];
  const Body& fast_body = bodies[fast.body_id];
  const Pos base = fast.center0;
  // world_runtime.rs:929-937 (make_relative_sweep)
  const Sweep sweep{fast.local_center, sub_pos(fast.center0, base), sub_pos(fast.center, base), fast.rotation0,
                    fast.transform.q};
  const Xf xf1{sub(sweep.c1, rotate(sweep.q1, sweep.local_center)), sweep.q1};
  const Xf xf2{sub(sweep.c2, rotate(sweep.q2, sweep.local_center)), sweep.q2};
  const bool is_bullet = (fast.flags & BF::IS_BULLET) != 0;
  const float safety_factor = fast_body.safety_factor;
  const float rotation_chord = 2.0f * length(inv_mul_quat(sweep.q1, sweep.q2).v);
  float fraction = 1.0f;
  for (int32_t sid = fast_body.head_shape_id; sid != kNullIndex;) {
    Shape& fs = shapes[sid];
    sid = fs.next_shape_id;
    const Vec3 lc = shape_centroid(fs);
    const Vec3 centroid1 = xf_point(xf1, lc), centroid2 = xf_point(xf2, lc);
    const AABB box1 = fs.aabb;
    const AABB box2 = offset_aabb(compute_shape_aabb(fs, xf2), base);
    fs.aabb = box2;
    if (fs.is_sensor) continue;  // no continuous collision for sensors (continuous.rs:398-401)
    if (!is_bullet && !is_shape_fast(fs, centroid1, centroid2, rotation_chord, safety_factor)) continue;
    const AABB swept = aabb_union(box1, box2);
    const ToiProxy fast_proxy = shape_proxy(fs);
    // the static tree; bullets also query the kinematic and dynamic trees (continuous.rs:424-445)
    const int tree_count = is_bullet ? 3 : 1;
    for (int tree = 0; tree < tree_count; ++tree) {
      for (const Shape& s : shapes) {
        if (s.id == kNullIndex || s.proxy_type != tree) continue;
        if (!aabb_overlaps(fat_aabbs[s.id], swept)) continue;
        // continuous_query_callback_shapes (continuous.rs:218-298)
        if (s.id == fs.id || s.body_id == fs.body_id) continue;
        // a sensor candidate reports only sensor events, which are not produced (continuous.rs:239-246)
        if (s.is_sensor) continue;
        if (s.filter.group == fs.filter.group && s.filter.group != 0) {
          if (!(fs.filter.group > 0)) continue;
        } else if (!((fs.filter.mask & s.filter.category) != 0 && (fs.filter.category & s.filter.mask) != 0)) {
          continue;
        }
        const Body& other = bodies[s.body_id];
        if ((other.flags & BF::IS_BULLET) != 0) continue;
        if (!should_bodies_collide(fast_body, other)) continue;
        const BodySim& os = sim_of(other);
        const Sweep other_sweep{os.local_center, sub_pos(os.center0, base), sub_pos(os.center, base), os.rotation0,
                                os.transform.q};
        ToiOutput toi = time_of_impact(shape_proxy(s), other_sweep, fast_proxy, sweep, fraction);
        stats.toi_calls += 1;
        if (0.0f < toi.fraction && toi.fraction < fraction) fraction = toi.fraction;
      }
    }
  }
  const float spec = speculative_distance();
  if (fraction < 1.0f) {
    const Quat q = nlerp(sweep.q1, sweep.q2, fraction);
    const Vec3 c = lerp(sweep.c1, sweep.c2, fraction);
    const Vec3 origin = sub(c, rotate(q, sweep.local_center));
    const WXf transform{offset_pos(base, origin), q};
    const Pos center = offset_pos(base, c);
    fast.transform = transform;
    fast.center = center;
    fast.rotation0 = q;
    fast.center0 = center;
    fast.flags |= BF::HAD_TOI;
    // the gravity of the lost time (continuous.rs:475-500)
    BodyState& state = awake_states[(size_t)sim_index];
    const float time_loss = (1.0f - fraction) * dt;
    Vec3 dv = scale(-time_loss * fast.gravity_scale, gravity);
    if (state.flags & BF::LOCK_LINEAR_X) dv.x = 0.0f;
    if (state.flags & BF::LOCK_LINEAR_Y) dv.y = 0.0f;
    if (state.flags & BF::LOCK_LINEAR_Z) dv.z = 0.0f;
    state.v = add(state.v, dv);
    stats.toi_hits += 1;
    for (int32_t sid = fast_body.head_shape_id; sid != kNullIndex; sid = shapes[sid].next_shape_id) {
      Shape& s = shapes[sid];
      const AABB aabb = compute_fat_shape_aabb(s, transform, spec);
      s.aabb = aabb;
      if (!aabb_contains(fat_aabbs[sid], aabb)) {
        fat_aabbs[sid] = aabb_inflate(aabb, s.aabb_margin);
        s.proxy_moved = true;  // mark_proxy_moved, or the bullet's enlarge_proxy
      }
    }
  } else {
    fast.rotation0 = fast.transform.q;
    fast.center0 = fast.center;
    for (int32_t sid = fast_body.head_shape_id; sid != kNullIndex; sid = shapes[sid].next_shape_id) {
      Shape& s = shapes[sid];
      if (!aabb_contains(fat_aabbs[sid], s.aabb)) {
        fat_aabbs[sid] = aabb_inflate(s.aabb, s.aabb_margin);
        s.proxy_moved = true;
      }
    }
  }
  awake_sims[(size_t)sim_index] = fast;
}

}  // namespace nslref
