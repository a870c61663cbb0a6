// Continuous collision on gfx950: the time of impact of a fast body's swept shapes against
// the shapes their swept bounds overlap, and the move back to the earliest impact.
//
// Follows src/collision/time_of_impact.rs:1-598, src/collision/distance.rs:817-832 and
// src/simulation/continuous.rs:121-621 operation for operation, so a body moved back to an
// impact lands on the same bits as the CPU step. One lane per fast body: fast bodies are
// few, and each sweep is a short serial root search over a few support points.
//
// Candidates are visited in shape id order (static shapes; bullets then kinematic and
// dynamic ones). The reference visits them in its broad-phase trees' order; the earliest
// impact does not depend on the order unless two candidates hit within the same step.
#pragma once
#include "nsl_ccd.h"
#include "nsl_stages.h"

namespace nsl {

// ShapeProxy (queries.rs): support points and a radius.
struct ToiProxy {
  const Vec3* points;
  int32_t count;
  float radius;
};

namespace toi {

// distance.rs:817-832
__device__ __forceinline__ Xf sweep_xf(const Sweep& s, float t) {
  Quat q = nlerp(s.q1, s.q2, t);
  return Xf{sub(lerp(s.c1, s.c2, t), rotate(q, s.local_center)), q};
}
__device__ __forceinline__ Xf sweep_final_xf(const Sweep& s) { return Xf{sub(s.c2, rotate(s.q2, s.local_center)), s.q2}; }

enum SepKind : int32_t { kSepUnknown = 0, kSepVertices, kSepEdges, kSepFaceA, kSepFaceB };

struct SepFn {
  Vec3 w1, w2;  // world axis (vertices), local edges (edges), local axis and point (faces)
  int32_t kind;
};

// time_of_impact.rs:36-63
__device__ __forceinline__ int unique_count(int n, const int* v) {
  if (n == 1) return 1;
  if (n == 2) return v[0] != v[1] ? 2 : 1;
  if (v[0] != v[1] && v[0] != v[2] && v[1] != v[2]) return 3;
  if (v[0] == v[1] && v[0] == v[2] && v[1] == v[2]) return 1;
  return 2;
}

// Edge pair axis or its vertex fallbacks (time_of_impact.rs:133-166, :213-258).
__device__ void edge_pair_axis(SepFn& f, const Sweep& sa, const Sweep& sb, Vec3 va1, Vec3 va2, Vec3 vb1, Vec3 vb2,
                               Quat qa, Quat qb, Vec3 dp, Vec3 world_normal, float tolerance_sq) {
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
  // time_of_impact.rs:65-77: the edges' cross product turns over during the sweep
  Xf xa2 = sweep_final_xf(sa), xb2 = sweep_final_xf(sb);
  Vec3 axis2 = cross(rotate(xa2.q, local_edge_a), rotate(xb2.q, local_edge_b));
  if (dot(axis2, axis) < 0.0f) {
    f.kind = kSepVertices;
    f.w1 = normalize(axis);
  } else {
    f.kind = kSepEdges;
    f.w1 = local_edge_a;
    f.w2 = local_edge_b;
  }
}

// time_of_impact.rs:79-263
__device__ SepFn make_separation(const SimplexCache& cache, const ToiProxy& pa, const Sweep& sa, const ToiProxy& pb,
                                 const Sweep& sb, Vec3 world_normal, float t1) {
  SepFn f{zero3(), zero3(), kSepUnknown};
  int ia[3] = {cache.index_a[0], cache.index_a[1], cache.index_a[2]};
  int ib[3] = {cache.index_b[0], cache.index_b[1], cache.index_b[2]};
  const int ua = unique_count(cache.count, ia), ub = unique_count(cache.count, ib);
  const Xf xa = sweep_xf(sa, t1), xb = sweep_xf(sb, t1);
  const Quat qa = xa.q, qb = xb.q;
  const Vec3 dp = sub(xb.p, xa.p);
  const Vec3* A = pa.points;
  const Vec3* B = pb.points;
  if (cache.count == 1) {
    f.kind = kSepVertices;
    f.w1 = world_normal;
  } else if (cache.count == 2) {
    if (ua == 2 && ub == 2) {
      edge_pair_axis(f, sa, sb, A[ia[0]], A[ia[1]], B[ib[0]], B[ib[1]], qa, qb, dp, world_normal, 0.05f * 0.05f);
    } else {
      f.kind = kSepVertices;
      f.w1 = world_normal;
    }
  } else if (ua == 3) {
    Vec3 va1 = A[ia[0]], va2 = A[ia[1]], va3 = A[ia[2]];
    Vec3 local_axis = normalize(cross(sub(va2, va1), sub(va3, va1)));
    Vec3 axis = rotate(qa, local_axis);
    Vec3 local_point_a = scale(1.0f / 3.0f, add(add(va1, va2), va3));
    Vec3 delta = add(sub(rotate(qb, B[ib[0]]), rotate(qa, local_point_a)), dp);
    if (dot(delta, axis) < 0.0f) local_axis = neg(local_axis);
    f.kind = kSepFaceA;
    f.w1 = local_axis;
    f.w2 = local_point_a;
  } else if (ub == 3) {
    Vec3 vb1 = B[ib[0]], vb2 = B[ib[1]], vb3 = B[ib[2]];
    Vec3 local_axis = normalize(cross(sub(vb2, vb1), sub(vb3, vb1)));
    Vec3 axis = rotate(qb, local_axis);
    Vec3 local_point_b = scale(1.0f / 3.0f, add(add(vb1, vb2), vb3));
    Vec3 delta = sub(sub(rotate(qa, A[ia[0]]), rotate(qb, local_point_b)), dp);
    if (dot(delta, axis) < 0.0f) local_axis = neg(local_axis);
    f.kind = kSepFaceB;
    f.w1 = local_axis;
    f.w2 = local_point_b;
  } else {
    if (ia[0] == ia[1]) ia[1] = ia[2];
    if (ib[0] == ib[1]) ib[1] = ib[2];
    edge_pair_axis(f, sa, sb, A[ia[0]], A[ia[1]], B[ib[0]], B[ib[1]], qa, qb, dp, world_normal, 0.005f * 0.005f);
  }
  return f;
}

// time_of_impact.rs:266-344: the deepest points along the axis at time t
__device__ float find_min_separation(const SepFn& f, const ToiProxy& pa, const Sweep& sa, const ToiProxy& pb,
                                     const Sweep& sb, int* ia, int* ib, float t) {
  const Xf xa = sweep_xf(sa, t), xb = sweep_xf(sb, t);
  const Vec3* A = pa.points;
  const Vec3* B = pb.points;
  if (f.kind == kSepVertices) {
    Vec3 axis = f.w1;
    *ia = gjk::point_support(A, pa.count, inv_rotate(xa.q, axis));
    *ib = gjk::point_support(B, pb.count, inv_rotate(xb.q, neg(axis)));
    Vec3 delta = add(sub(rotate(xb.q, B[*ib]), rotate(xa.q, A[*ia])), sub(xb.p, xa.p));
    return dot(delta, axis);
  }
  if (f.kind == kSepEdges) {
    Vec3 axis = normalize(cross(rotate(xa.q, f.w1), rotate(xb.q, f.w2)));
    *ia = gjk::point_support(A, pa.count, inv_rotate(xa.q, axis));
    *ib = gjk::point_support(B, pb.count, neg(inv_rotate(xb.q, axis)));
    Vec3 delta = add(sub(rotate(xb.q, B[*ib]), rotate(xa.q, A[*ia])), sub(xb.p, xa.p));
    return dot(delta, axis);
  }
  if (f.kind == kSepFaceA) {
    Vec3 normal = rotate(xa.q, f.w1);
    *ia = -1;
    Vec3 point_a = xf_point(xa, f.w2);
    *ib = gjk::point_support(B, pb.count, neg(inv_rotate(xb.q, normal)));
    return dot(sub(xf_point(xb, B[*ib]), point_a), normal);
  }
  if (f.kind == kSepFaceB) {
    Vec3 normal = rotate(xb.q, f.w1);
    *ia = gjk::point_support(A, pa.count, neg(inv_rotate(xa.q, normal)));
    Vec3 point_a = xf_point(xa, A[*ia]);
    *ib = -1;
    return dot(sub(point_a, xf_point(xb, f.w2)), normal);
  }
  return 0.0f;
}

// time_of_impact.rs:347-386: separation of fixed witness points at time beta
__device__ float evaluate_separation(const SepFn& f, const ToiProxy& pa, const Sweep& sa, const ToiProxy& pb,
                                     const Sweep& sb, int i1, int i2, float beta) {
  const Xf x1 = sweep_xf(sa, beta), x2 = sweep_xf(sb, beta);
  if (f.kind == kSepVertices) return dot(sub(xf_point(x2, pb.points[i2]), xf_point(x1, pa.points[i1])), f.w1);
  if (f.kind == kSepEdges) {
    Vec3 axis = normalize(cross(rotate(x1.q, f.w1), rotate(x2.q, f.w2)));
    return dot(sub(xf_point(x2, pb.points[i2]), xf_point(x1, pa.points[i1])), axis);
  }
  if (f.kind == kSepFaceA) return dot(sub(xf_point(x2, pb.points[i2]), xf_point(x1, f.w2)), rotate(x1.q, f.w1));
  if (f.kind == kSepFaceB) return dot(sub(xf_point(x1, pa.points[i1]), xf_point(x2, f.w2)), rotate(x2.q, f.w1));
  return 0.0f;
}

}  // namespace toi

// time_of_impact.rs:406-598. Conservative advancement: at each GJK distance build a
// separating axis from the simplex and push t forward to where the deepest points along it
// reach the target separation (one linear slop short of touching).
__device__ ToiOutput time_of_impact(const ToiProxy& pa, Sweep sa, const ToiProxy& pb, Sweep sb, float max_fraction) {
  using namespace toi;
  ToiOutput out;
  out.state = kToiUnknown;
  out.point = zero3();
  out.normal = zero3();
  out.fraction = -1.0f;
  out.distance = 0.0f;
  out.distance_iterations = out.push_back_iterations = out.root_iterations = 0;
  const Vec3 origin = sa.c1;
  sa.c1 = zero3();
  sa.c2 = sub(sa.c2, origin);
  sb.c1 = sub(sb.c1, origin);
  sb.c2 = sub(sb.c2, origin);
  const int max_push_back = pa.count + pb.count;
  const float t_max = max_fraction;
  const float target = maxf(kLinearSlop, pa.radius + pb.radius - kLinearSlop);
  const float tolerance = 0.25f * kLinearSlop;
  float t1 = 0.0f;
  const int max_iterations = 25;
  int distance_iterations = 0;
  SimplexCache cache;
  cache.metric = 0.0f;
  cache.count = 0;
  for (int i = 0; i < 4; ++i) cache.index_a[i] = cache.index_b[i] = 0;
  for (;;) {
    const Xf xa = sweep_xf(sa, t1), xb = sweep_xf(sb, t1);
    const gjk::Output d = gjk::distance(pa.points, pa.count, pb.points, pb.count, inv_mul_xf(xa, xb), cache);
    out.distance = d.distance;
    const Vec3 world_normal = rotate(xa.q, d.normal);
    const Vec3 world_point_a = xf_point(xa, d.point_a);
    const Vec3 world_point_b = xf_point(xa, d.point_b);
    out.distance_iterations += 1;
    distance_iterations += 1;
    int state = kToiUnknown;
    float fraction = 0.0f;
    if (d.distance <= 0.0f) {
      state = kToiOverlapped;
    } else if (d.distance <= target + tolerance) {
      state = kToiHit;
      fraction = t1;
    } else if (distance_iterations == max_iterations) {
      state = kToiFailed;
      fraction = t1;
    } else {
      SepFn f = make_separation(cache, pa, sa, pb, sb, world_normal, t1);
      float t2 = t_max;
      int push_back = 0;
      for (;;) {
        int ia = 0, ib = 0;
        float s2 = find_min_separation(f, pa, sa, pb, sb, &ia, &ib, t2);
        if (s2 - target > tolerance) {  // separated at the end of the bracket
          state = kToiSeparated;
          fraction = max_fraction;
          break;
        }
        if (s2 >= target - tolerance) {  // advance to t2
          t1 = t2;
          break;
        }
        float s1 = evaluate_separation(f, pa, sa, pb, sb, ia, ib, t1);
        if (s1 < target - tolerance) {
          state = kToiFailed;
          fraction = t1;
          break;
        }
        if (s1 <= target + tolerance) {
          state = kToiHit;
          fraction = t1;
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
          float s = evaluate_separation(f, pa, sa, pb, sb, ia, ib, t);
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
        if (root_count == max_root - 1 && f.kind == kSepEdges) {  // time_of_impact.rs:389-403
          t2 = max_fraction;
          const Xf x1 = sweep_xf(sa, t1), x2 = sweep_xf(sb, t1);
          f.w1 = normalize(cross(rotate(x1.q, f.w1), rotate(x2.q, f.w2)));
          f.w2 = zero3();
          f.kind = kSepVertices;
        }
        out.push_back_iterations += 1;
        push_back += 1;
        if (push_back == max_push_back) break;
      }
    }
    if (state != kToiUnknown) {
      out.state = state;
      out.fraction = fraction;
      Vec3 p_a = mul_add(world_point_a, pa.radius, world_normal);
      Vec3 p_b = mul_add(world_point_b, -pb.radius, world_normal);
      out.point = add(lerp(p_a, p_b, 0.5f), origin);
      out.normal = world_normal;
      return out;
    }
  }
}

// shape_runtime.rs:643-658: the shape's support points; `store` holds a sphere's centre or a
// capsule's segment.
__device__ __forceinline__ ToiProxy shape_toi_proxy(const DeviceWorld& d, const Shape& s, Vec3* store) {
  if (s.kind == kShapeHull) return ToiProxy{d.hulls[s.hull].points, 8, 0.0f};
  store[0] = s.c1;
  store[1] = s.c2;
  return ToiProxy{store, s.kind == kShapeSphere ? 1 : 2, s.radius};
}

// shape_runtime.rs:434-450
__device__ __forceinline__ Vec3 shape_centroid(const DeviceWorld& d, const Shape& s) {
  if (s.kind == kShapeSphere) return s.c1;
  if (s.kind == kShapeCapsule) return lerp(s.c1, s.c2, 0.5f);
  return d.hulls[s.hull].center;
}

// continuous.rs:121-151 with compute_shape_min_extent (shape_runtime.rs:547-554)
__device__ bool is_shape_fast(const DeviceWorld& d, const Shape& s, Vec3 centroid1, Vec3 centroid2, float rotation_chord,
                              float safety_factor) {
  float radius, min_extent;
  if (s.kind == kShapeSphere) {
    radius = s.radius;
    min_extent = s.radius;
  } else if (s.kind == kShapeCapsule) {
    radius = 0.5f * distance(s.c1, s.c2) + s.radius;
    min_extent = s.radius;
  } else {
    const BoxHull& h = d.hulls[s.hull];
    const Vec3 lc = h.center;
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

// world_runtime.rs:929-937
__device__ __forceinline__ Sweep relative_sweep(const BodySim& sim, Pos base) {
  return Sweep{sim.local_center, sub_pos(sim.center0, base), sub_pos(sim.center, base), sim.rotation0, sim.transform.q};
}

// solve_continuous (continuous.rs:337-621) for the awake body i, flagged fast by finalize.
// Reads the candidates' bodies and writes only body i, its state and its shapes' bounds.
// Returns whether the body was moved back to an impact.
__device__ bool solve_continuous_body(const DeviceWorld& d, const ContinuousArgs& a, int i, const StepParams& p) {
  BodySim fast = d.sims[i];
  const Pos base = fast.center0;
  const Sweep sweep = relative_sweep(fast, base);
  const Xf xf1{sub(sweep.c1, rotate(sweep.q1, sweep.local_center)), sweep.q1};
  const Xf xf2{sub(sweep.c2, rotate(sweep.q2, sweep.local_center)), sweep.q2};
  const bool is_bullet = (fast.flags & bflag::kIsBullet) != 0;
  const float rotation_chord = 2.0f * length(inv_mul_quat(sweep.q1, sweep.q2).v);
  const int n_candidates = is_bullet ? a.n_static + a.n_kinematic + a.n_dynamic : a.n_static;
  float fraction = 1.0f;
  int toi_calls = 0;
  for (int32_t sid = fast.head_shape; sid != kNullIndex;) {
    const Shape fs = d.shapes[sid];
    const int32_t fast_id = sid;
    sid = fs.next_shape;
    const Vec3 lc = shape_centroid(d, fs);
    const Aabb box1 = d.shape_aabbs[fast_id];
    const Aabb box2 = offset_aabb(shape_aabb_at(d, fs, xf2), base);
    d.shape_aabbs[fast_id] = box2;
    const Proxy fp = a.proxies[fast_id];
    if (fp.flags & kProxySensor) continue;  // no continuous collision for sensors (continuous.rs:398-401)
    if (!is_bullet && !is_shape_fast(d, fs, xf_point(xf1, lc), xf_point(xf2, lc), rotation_chord, fast.safety_factor))
      continue;
    const Aabb swept = aabb_union(box1, box2);
    Vec3 fast_store[2];
    const ToiProxy fast_proxy = shape_toi_proxy(d, fs, fast_store);
    for (int k = 0; k < n_candidates; ++k) {
      const int32_t cid = a.candidates[k];
      if (!aabb_overlaps(d.fat_aabbs[cid], swept)) continue;
      // continuous_query_callback_shapes (continuous.rs:218-298)
      const Proxy cp = a.proxies[cid];
      if (cid == fast_id || cp.body == fp.body) continue;
      // a sensor candidate reports only sensor events, which are not produced (continuous.rs:239-246)
      if (cp.flags & kProxySensor) continue;
      if (cp.group == fp.group && cp.group != 0) {  // shape_runtime.rs:285-292
        if (!(fp.group > 0)) continue;
      } else if (!((fp.mask & cp.category) != 0 && (fp.category & cp.mask) != 0)) {
        continue;
      }
      const Shape cs = d.shapes[cid];
      const BodySim& os = *sim_of(d, cs.body_enc, cid);
      if (os.flags & bflag::kIsBullet) continue;
      if (d.joint_pairs.slots != nullptr && pairset_contains(d.joint_pairs, pair_key(fp.body, cp.body))) continue;
      Vec3 store[2];
      const ToiProxy proxy = shape_toi_proxy(d, cs, store);
      const ToiOutput out = time_of_impact(proxy, relative_sweep(os, base), fast_proxy, sweep, fraction);
      toi_calls += 1;
      if (0.0f < out.fraction && out.fraction < fraction) fraction = out.fraction;
    }
  }
  if (toi_calls) atomicAdd(&d.counters[kCntToiCalls], toi_calls);
  int moved = 0;
  const bool hit = fraction < 1.0f;
  if (hit) {
    const Quat q = nlerp(sweep.q1, sweep.q2, fraction);
    const Vec3 c = lerp(sweep.c1, sweep.c2, fraction);
    const Vec3 origin = sub(c, rotate(q, sweep.local_center));
    fast.transform = WXf{offset_pos(base, origin), q};
    fast.center = offset_pos(base, c);
    fast.rotation0 = q;
    fast.center0 = fast.center;
    fast.flags |= bflag::kHadToi;
    // the gravity of the lost time (continuous.rs:475-500)
    BodyState& state = d.states[i];
    const float time_loss = (1.0f - fraction) * p.dt;
    Vec3 dv = scale(-time_loss * fast.gravity_scale, p.gravity);
    if (state.flags & bflag::kLockLinX) dv.x = 0.0f;
    if (state.flags & bflag::kLockLinY) dv.y = 0.0f;
    if (state.flags & bflag::kLockLinZ) dv.z = 0.0f;
    state.v = add(state.v, dv);
    for (int32_t sid = fast.head_shape; sid != kNullIndex; sid = d.shapes[sid].next_shape) {
      const Shape& s = d.shapes[sid];
      const Aabb aabb = offset_aabb(aabb_inflate(shape_aabb_at(d, s, Xf{zero3(), q}), kSpeculativeDistance), fast.transform.p);
      d.shape_aabbs[sid] = aabb;
      if (!aabb_contains(d.fat_aabbs[sid], aabb)) {
        d.fat_aabbs[sid] = aabb_inflate(aabb, s.aabb_margin);
        d.shape_moved[sid] = 1;
        ++moved;
      }
    }
  } else {
    fast.rotation0 = fast.transform.q;
    fast.center0 = fast.center;
    for (int32_t sid = fast.head_shape; sid != kNullIndex; sid = d.shapes[sid].next_shape) {
      const Aabb aabb = d.shape_aabbs[sid];
      if (!aabb_contains(d.fat_aabbs[sid], aabb)) {
        d.fat_aabbs[sid] = aabb_inflate(aabb, d.shapes[sid].aabb_margin);
        d.shape_moved[sid] = 1;
        ++moved;
      }
    }
  }
  if (moved) atomicAdd(&d.counters[kCntMovedProxies], moved);
  d.sims[i] = fast;
  return hit;
}

}  // namespace nsl
