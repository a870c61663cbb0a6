// Sphere and capsule narrow phase for gfx950 (one lane per contact), with the GJK
// distance query the hull-sphere and hull-capsule pairs use.
//
// Follows src/collision/convex_manifold_shapes.rs:21-797,
// src/collision/convex_manifold.rs:137-200 and src/collision/distance.rs:11-709
// operation for operation; the GJK simplex cache lives in the contact record so warm
// starts match the CPU step.
#pragma once
#include "nsl_box_collide.h"

namespace nsl {

namespace gjk {

constexpr int kMaxIterations = 32;

struct Vertex { Vec3 wa, wb, w; float a; int32_t index_a, index_b; };
struct Simplex { Vertex v[4]; int32_t count; };
struct Output { Vec3 point_a, point_b, normal; float distance; };
struct Bary2 { float u, v, d; };
struct Bary3 { float u, v, w, d; };

__device__ __forceinline__ Vec3 blend3(float s, Vec3 a, float t, Vec3 b, float u, Vec3 c) {
  return {s * a.x + t * b.x + u * c.x, s * a.y + t * b.y + u * c.y, s * a.z + t * b.z + u * c.z};
}
__device__ __forceinline__ float triple(Vec3 a, Vec3 b, Vec3 c) {
  Vec3 d = {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
  return a.x * d.x + a.y * d.y + a.z * d.z;
}
__device__ __forceinline__ int point_support(const Vec3* pts, int n, Vec3 axis) {
  Vec3 origin = pts[0];
  int best = 0;
  float best_proj = 0.0f;
  for (int i = 1; i < n; ++i) {
    float p = dot(axis, sub(pts[i], origin));
    if (p > best_proj) { best = i; best_proj = p; }
  }
  return best;
}
__device__ __forceinline__ Bary2 edge(Vec3 a, Vec3 b) {
  Vec3 ab = sub(b, a);
  return Bary2{dot(b, ab), -dot(a, ab), dot(ab, ab)};
}
__device__ __forceinline__ Bary3 tri(Vec3 a, Vec3 b, Vec3 c) {
  Vec3 ab = sub(b, a), ac = sub(c, a);
  Vec3 n = cross(ab, ac);
  return Bary3{dot(cross(b, c), n), dot(cross(c, a), n), dot(cross(a, b), n), dot(n, n)};
}
__device__ __forceinline__ float metric(const Simplex& s) {
  switch (s.count) {
    case 2: return distance(s.v[0].w, s.v[1].w);
    case 3: return length(cross(sub(s.v[1].w, s.v[0].w), sub(s.v[2].w, s.v[0].w))) / 2.0f;
    case 4: return triple(sub(s.v[1].w, s.v[0].w), sub(s.v[2].w, s.v[0].w), sub(s.v[3].w, s.v[0].w)) / 6.0f;
    default: return 0.0f;
  }
}
__device__ __forceinline__ void set_vertex(Simplex& s, const Vertex& x) {
  s.count = 1;
  s.v[0] = x;
  s.v[0].a = 1.0f;
}
__device__ __forceinline__ bool set_edge(Simplex& s, const Vertex& x, const Vertex& y, Bary2 w) {
  s.count = 2;
  s.v[0] = x;
  s.v[1] = y;
  if (w.d <= 0.0f) return false;
  s.v[0].a = w.u / w.d;
  s.v[1].a = w.v / w.d;
  return true;
}
__device__ __forceinline__ bool set_tri(Simplex& s, const Vertex& x, const Vertex& y, const Vertex& z, Bary3 w) {
  s.count = 3;
  s.v[0] = x;
  s.v[1] = y;
  s.v[2] = z;
  if (w.d <= 0.0f) return false;
  s.v[0].a = w.u / w.d;
  s.v[1].a = w.v / w.d;
  s.v[2].a = w.w / w.d;
  return true;
}
// distance.rs:122-158
__device__ bool solve2(Simplex& s) {
  Vec3 a = s.v[0].w, b = s.v[1].w;
  Vec3 ab = sub(b, a);
  float divisor = dot(ab, ab), u = dot(b, ab), v = -dot(a, ab);
  if (v <= 0.0f) { s.count = 1; s.v[0].a = 1.0f; return true; }
  if (u <= 0.0f) { s.count = 1; s.v[0] = s.v[1]; s.v[0].a = 1.0f; return true; }
  if (divisor <= 0.0f) return false;
  float den = 1.0f / divisor;
  s.v[0].a = den * u;
  s.v[1].a = den * v;
  return true;
}
// distance.rs:160-240
__device__ bool solve3(Simplex& s) {
  Vertex v1 = s.v[0], v2 = s.v[1], v3 = s.v[2];
  Bary2 ab = edge(v1.w, v2.w), bc = edge(v2.w, v3.w), ca = edge(v3.w, v1.w);
  if (ab.v <= 0.0f && ca.u <= 0.0f) { set_vertex(s, v1); return true; }
  if (bc.v <= 0.0f && ab.u <= 0.0f) { set_vertex(s, v2); return true; }
  if (ca.v <= 0.0f && bc.u <= 0.0f) { set_vertex(s, v3); return true; }
  Bary3 abc = tri(v1.w, v2.w, v3.w);
  if (abc.w <= 0.0f && ab.u > 0.0f && ab.v > 0.0f) return set_edge(s, v1, v2, ab);
  if (abc.u <= 0.0f && bc.u > 0.0f && bc.v > 0.0f) return set_edge(s, v2, v3, bc);
  if (abc.v <= 0.0f && ca.u > 0.0f && ca.v > 0.0f) return set_edge(s, v3, v1, ca);
  if (abc.d <= 0.0f) return false;
  s.v[0].a = abc.u / abc.d;
  s.v[1].a = abc.v / abc.d;
  s.v[2].a = abc.w / abc.d;
  return true;
}
// distance.rs:242-418
__device__ bool solve4(Simplex& s) {
  Vertex A = s.v[0], B = s.v[1], C = s.v[2], D = s.v[3];
  Bary2 ab = edge(A.w, B.w), ac = edge(A.w, C.w), ad = edge(A.w, D.w);
  Bary2 bc = edge(B.w, C.w), cd = edge(C.w, D.w), db = edge(D.w, B.w);
  if (ab.v <= 0.0f && ac.v <= 0.0f && ad.v <= 0.0f) { set_vertex(s, A); return true; }
  if (ab.u <= 0.0f && db.u <= 0.0f && bc.v <= 0.0f) { set_vertex(s, B); return true; }
  if (ac.u <= 0.0f && bc.u <= 0.0f && cd.v <= 0.0f) { set_vertex(s, C); return true; }
  if (ad.u <= 0.0f && cd.u <= 0.0f && db.v <= 0.0f) { set_vertex(s, D); return true; }
  Bary3 acb = tri(A.w, C.w, B.w), abd = tri(A.w, B.w, D.w), adc = tri(A.w, D.w, C.w), bcd = tri(B.w, C.w, D.w);
  if (abd.w <= 0.0f && acb.v <= 0.0f && ab.u > 0.0f && ab.v > 0.0f) return set_edge(s, A, B, ab);
  if (acb.w <= 0.0f && adc.v <= 0.0f && ac.u > 0.0f && ac.v > 0.0f) return set_edge(s, A, C, ac);
  if (adc.w <= 0.0f && abd.v <= 0.0f && ad.u > 0.0f && ad.v > 0.0f) return set_edge(s, A, D, ad);
  if (acb.u <= 0.0f && bcd.w <= 0.0f && bc.u > 0.0f && bc.v > 0.0f) return set_edge(s, B, C, bc);
  if (adc.u <= 0.0f && bcd.u <= 0.0f && cd.u > 0.0f && cd.v > 0.0f) return set_edge(s, C, D, cd);
  if (abd.u <= 0.0f && bcd.v <= 0.0f && db.u > 0.0f && db.v > 0.0f) return set_edge(s, D, B, db);
  float divisor = triple(sub(B.w, A.w), sub(C.w, A.w), sub(D.w, A.w));
  float sign = divisor < 0.0f ? -1.0f : 1.0f;
  float tu = sign * triple(B.w, C.w, D.w);
  float tv = sign * triple(A.w, D.w, C.w);
  float tw = sign * triple(A.w, B.w, D.w);
  float tx = sign * triple(A.w, C.w, B.w);
  float td = sign * divisor;
  if (tx < 0.0f && acb.u > 0.0f && acb.v > 0.0f && acb.w > 0.0f) return set_tri(s, A, C, B, acb);
  if (tw < 0.0f && abd.u > 0.0f && abd.v > 0.0f && abd.w > 0.0f) return set_tri(s, A, B, D, abd);
  if (tv < 0.0f && adc.u > 0.0f && adc.v > 0.0f && adc.w > 0.0f) return set_tri(s, A, D, C, adc);
  if (tu < 0.0f && bcd.u > 0.0f && bcd.v > 0.0f && bcd.w > 0.0f) return set_tri(s, B, C, D, bcd);
  if (td <= 0.0f) return false;
  s.v[0].a = tu / td;
  s.v[1].a = tv / td;
  s.v[2].a = tw / td;
  s.v[3].a = tx / td;
  return true;
}
__device__ __forceinline__ void witness(const Simplex& s, Vec3* a, Vec3* b) {
  const Vertex* v = s.v;
  if (s.count == 1) {
    *a = v[0].wa;
    *b = v[0].wb;
  } else if (s.count == 2) {
    *a = blend2(v[0].a, v[0].wa, v[1].a, v[1].wa);
    *b = blend2(v[0].a, v[0].wb, v[1].a, v[1].wb);
  } else if (s.count == 3) {
    *a = blend3(v[0].a, v[0].wa, v[1].a, v[1].wa, v[2].a, v[2].wa);
    *b = blend3(v[0].a, v[0].wb, v[1].a, v[1].wb, v[2].a, v[2].wb);
  } else {
    Vec3 sum = add(blend2(v[0].a, v[0].wa, v[1].a, v[1].wa), blend2(v[2].a, v[2].wa, v[3].a, v[3].wa));
    *a = sum;
    *b = sum;
  }
}
__device__ __forceinline__ Vec3 closest(const Simplex& s) {
  const Vertex* v = s.v;
  if (s.count == 1) return v[0].w;
  if (s.count == 2) return blend2(v[0].a, v[0].w, v[1].a, v[1].w);
  if (s.count == 3) return blend3(v[0].a, v[0].w, v[1].a, v[1].w, v[2].a, v[2].w);
  return add(blend2(v[0].a, v[0].w, v[1].a, v[1].w), blend2(v[2].a, v[2].w, v[3].a, v[3].w));
}

// distance.rs:483-709 with use_radii = false
__device__ Output distance(const Vec3* pa, int na, const Vec3* pb, int nb, Xf xf, SimplexCache& cache) {
  Mat3 m = matrix_from_quat(xf.q);
  Mat3 mt = transpose(m);
  Simplex s;
  s.count = cache.count;
  for (int i = 0; i < cache.count; ++i) {
    int i1 = cache.index_a[i], i2 = cache.index_b[i];
    Vec3 v1 = pa[i1], v2 = add(mul_mv(m, pb[i2]), xf.p);
    s.v[i] = Vertex{v1, v2, sub(v2, v1), 0.0f, i1, i2};
  }
  if (s.count > 0) {
    float m1 = cache.metric, m2 = metric(s);
    if (2.0f * m1 < m2 || m2 < 0.5f * m1 || m2 < kFltEps) s.count = 0;
  }
  if (s.count == 0) {
    Vec3 v1 = pa[0], v2 = add(mul_mv(m, pb[0]), xf.p);
    s.count = 1;
    s.v[0] = Vertex{v1, v2, sub(v2, v1), 0.0f, 0, 0};
  }
  Simplex backup;
  backup.count = 0;
  Output out;
  out.normal = zero3();
  out.distance = 0.0f;
  float dist_sq = kFltMax;
  Vec3 normal = zero3();
  for (int iteration = 0; iteration < kMaxIterations; ++iteration) {
    bool solved;
    if (s.count == 1) {
      s.v[0].a = 1.0f;
      solved = true;
    } else if (s.count == 2) {
      solved = solve2(s);
    } else if (s.count == 3) {
      solved = solve3(s);
    } else {
      solved = solve4(s);
    }
    if (!solved) { s = backup; break; }
    if (s.count == 4) {
      witness(s, &out.point_a, &out.point_b);
      out.normal = zero3();
      out.distance = 0.0f;
      return out;
    }
    float old = dist_sq;
    Vec3 cp = closest(s);
    dist_sq = dot(cp, cp);
    if (dist_sq >= old) { s = backup; break; }
    Vec3 dir;
    if (s.count == 1) {
      dir = neg(s.v[0].w);
    } else if (s.count == 2) {
      Vec3 a = s.v[0].w, b = s.v[1].w;
      Vec3 ab = sub(b, a);
      dir = cross(cross(ab, neg(a)), ab);
    } else {
      Vec3 a = s.v[0].w;
      Vec3 n = cross(sub(s.v[1].w, a), sub(s.v[2].w, a));
      dir = dot(n, a) < 0.0f ? n : neg(n);
    }
    if (length_sq(dir) < 1000.0f * kFltMin) {
      witness(s, &out.point_a, &out.point_b);
      out.normal = zero3();
      out.distance = 0.0f;
      return out;
    }
    normal = neg(dir);
    int ia = point_support(pa, na, neg(dir));
    int ib = point_support(pb, nb, mul_mv(mt, dir));
    Vec3 sa = pa[ia], sb = add(mul_mv(m, pb[ib]), xf.p);
    backup = s;
    bool duplicate = false;
    for (int i = 0; i < s.count; ++i)
      if (s.v[i].index_a == ia && s.v[i].index_b == ib) { duplicate = true; break; }
    if (duplicate) break;
    s.v[s.count] = Vertex{sa, sb, sub(sb, sa), 0.0f, ia, ib};
    s.count += 1;
  }
  witness(s, &out.point_a, &out.point_b);
  normal = normalize(normal);
  if (!(absf(1.0f - dot(normal, normal)) < 100.0f * kFltEps)) {
    out.distance = 0.0f;
    out.normal = zero3();
    return out;
  }
  out.distance = nsl::distance(out.point_a, out.point_b);
  out.normal = normal;
  cache.metric = metric(s);
  cache.count = (uint16_t)s.count;
  for (int i = 0; i < s.count; ++i) {
    cache.index_a[i] = (uint8_t)s.v[i].index_a;
    cache.index_b[i] = (uint8_t)s.v[i].index_b;
  }
  return out;
}

}  // namespace gjk

__device__ __forceinline__ void set_point(LocalManifold& m, int i, Vec3 p, float sep, FeaturePair pair) {
  m.pts[i].point = p;
  m.pts[i].separation = sep;
  m.pts[i].pair = pair;
}

// convex_manifold.rs:137-170
__device__ __forceinline__ int clip_segment(ClipVertex* seg, Plane plane) {
  int n = 0;
  ClipVertex v1 = seg[0], v2 = seg[1];
  float d1 = plane_sep(plane, v1.position), d2 = plane_sep(plane, v2.position);
  if (d1 <= 0.0f) seg[n++] = v1;
  if (d2 <= 0.0f) seg[n++] = v2;
  if (d1 * d2 < 0.0f) {
    float t = d1 / (d1 - d2);
    seg[n].position = add(scale(1.0f - t, v1.position), scale(t, v2.position));
    seg[n].pair = d1 > 0.0f ? v1.pair : v2.pair;
    n += 1;
  }
  return n;
}

// convex_manifold.rs:172-200 on a box face
__device__ int clip_segment_to_box_face(ClipVertex* seg, const BoxHull& h, int ref_face) {
  Plane ref_plane = h.planes[ref_face];
  int start = c_box_face_edge[ref_face];
  int e = start;
  for (;;) {
    int nx = c_box_edge[e][0];
    Vec3 p1 = h.points[c_box_edge[e][2]], p2 = h.points[c_box_edge[nx][2]];
    Vec3 binormal = cross(normalize(sub(p2, p1)), ref_plane.n);
    if (clip_segment(seg, plane_from_normal_point(binormal, p1)) < 2) return 0;
    e = nx;
    if (e == start) break;
  }
  return 2;
}

// convex_manifold_shapes.rs:21-71
__device__ void collide_spheres(LocalManifold& m, Vec3 ca, float ra, Vec3 cb, float rb, Xf b_to_a) {
  m.point_count = 0;
  Vec3 c1 = ca, c2 = xf_point(b_to_a, cb);
  float total = ra + rb;
  float max_d = total + kSpeculativeDistance;
  Vec3 offset = sub(c2, c1);
  float dsq = length_sq(offset);
  if (dsq > max_d * max_d) return;
  Vec3 normal = {0.0f, 1.0f, 0.0f};
  float d = sqrtf_(dsq);
  if (d * d > 1000.0f * kFltMin) normal = scale(1.0f / d, offset);
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, scale(0.5f, mul_sub(add(mul_add(c1, ra, normal), c2), rb, normal)), d - total, make_pair(0, 0, 0, 0));
}

// convex_manifold_shapes.rs:74-129
__device__ void collide_capsule_sphere(LocalManifold& m, Vec3 a1, Vec3 a2, float ra, Vec3 cb, float rb, Xf b_to_a) {
  m.point_count = 0;
  Vec3 center = xf_point(b_to_a, cb);
  float total = rb + ra;
  float max_d = total + kSpeculativeDistance;
  // point_to_segment_distance (functions.rs:1169-1186)
  Vec3 closest;
  {
    Vec3 ab = sub(a2, a1), aq = sub(center, a1);
    float alpha = dot(ab, aq);
    if (alpha <= 0.0f) {
      closest = a1;
    } else {
      float den = dot(ab, ab);
      closest = alpha > den ? a2 : mul_add(a1, alpha / den, ab);
    }
  }
  Vec3 offset = sub(center, closest);
  float dsq = length_sq(offset);
  if (dsq > max_d * max_d) return;
  Vec3 normal = {0.0f, 1.0f, 0.0f};
  float d = sqrtf_(dsq);
  if (d * d > 1000.0f * kFltMin) normal = scale(1.0f / d, offset);
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, scale(0.5f, mul_add(add(mul_sub(center, rb, normal), closest), ra, normal)), d - total, make_pair(0, 0, 0, 0));
}

__device__ __forceinline__ Vec3 point_to_segment(Vec3 a, Vec3 b, Vec3 q) {
  Vec3 ab = sub(b, a), aq = sub(q, a);
  float alpha = dot(ab, aq);
  if (alpha <= 0.0f) return a;
  float den = dot(ab, ab);
  if (alpha > den) return b;
  alpha /= den;
  return mul_add(a, alpha, ab);
}

// functions.rs:1105-1165
__device__ SegDist segment_distance(Vec3 p1, Vec3 q1, Vec3 p2, Vec3 q2) {
  SegDist r;
  Vec3 d1 = sub(q1, p1), d2 = sub(q2, p2), rr = sub(p1, p2);
  float a = dot(d1, d1), b = dot(d1, d2), c = dot(d1, rr), e = dot(d2, d2), f = dot(d2, rr);
  const float eps = 100.0f * kFltEps;
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
  float s1 = denom > 1000.0f * kFltMin ? clampf((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
  float s2 = (b * s1 + f) / e;
  if (s2 < 0.0f) {
    s1 = clampf(-c / a, 0.0f, 1.0f);
    s2 = 0.0f;
  } else if (s2 > 1.0f) {
    s1 = clampf((b - c) / a, 0.0f, 1.0f);
    s2 = 1.0f;
  }
  r.p1 = mul_add(p1, s1, d1); r.f1 = s1; r.p2 = mul_add(p2, s2, d2); r.f2 = s2;
  return r;
}

__device__ __forceinline__ void length_normalize(Vec3 v, float* len, Vec3* n) {
  float ls = v.x * v.x + v.y * v.y + v.z * v.z;
  if (ls > 1000.0f * kFltMin) {
    float l = sqrtf_(ls);
    float s = 1.0f / l;
    *len = l;
    *n = Vec3{s * v.x, s * v.y, s * v.z};
  } else {
    *len = 0.0f;
    *n = zero3();
  }
}

// convex_manifold_shapes.rs:245-394
__device__ void collide_capsules(LocalManifold& m, Vec3 a1, Vec3 a2, float ra, Vec3 bc1, Vec3 bc2, float rb, Xf b_to_a) {
  m.point_count = 0;
  Vec3 b1 = xf_point(b_to_a, bc1), b2 = xf_point(b_to_a, bc2);
  float radius = ra + rb;
  float max_d = radius + kSpeculativeDistance;
  SegDist r = segment_distance(a1, a2, b1, b2);
  Vec3 offset = sub(r.p2, r.p1);
  float dsq = length_sq(offset);
  const float min_d = 0.01f * kLinearSlop;
  if (dsq > max_d * max_d || dsq < min_d * min_d) return;
  float len_a, len_b;
  Vec3 edge_a, edge_b;
  length_normalize(sub(a2, a1), &len_a, &edge_a);
  if (len_a < kLinearSlop) return;
  length_normalize(sub(b2, b1), &len_b, &edge_b);
  if (len_b < kLinearSlop) return;
  Vec3 axis = cross(edge_a, edge_b);
  if (length_sq(axis) < 0.05f * 0.05f) {
    Plane pa0 = Plane{neg(edge_a), -dot(edge_a, a1)}, pa1 = Plane{edge_a, dot(edge_a, a2)};
    ClipVertex vb[2];
    vb[0] = ClipVertex{b1, 0.0f, make_pair(0, 0, 0, 0)};
    vb[1] = ClipVertex{b2, 0.0f, make_pair(0, 1, 0, 1)};
    int count = clip_segment(vb, pa0);
    if (count == 2) count = clip_segment(vb, pa1);
    if (count == 2) {
      Vec3 cp1 = point_to_segment(a1, a2, vb[0].position), cp2 = point_to_segment(a1, a2, vb[1].position);
      float d1 = distance(cp1, vb[0].position), d2 = distance(cp2, vb[1].position);
      if (d1 <= radius && d2 <= radius) {
        if (d1 < min_d || d2 < min_d) return;
        Vec3 n1 = scale(1.0f / d1, sub(vb[0].position, cp1));
        Vec3 n2 = scale(1.0f / d2, sub(vb[1].position, cp2));
        Vec3 normal = normalize(add(n1, n2));
        m.normal = normal;
        m.point_count = 2;
        set_point(m, 0, scale(0.5f, mul_sub(add(mul_add(vb[0].position, ra, n1), cp1), rb, normal)), d1 - radius, vb[0].pair);
        set_point(m, 1, scale(0.5f, mul_sub(add(mul_add(vb[1].position, ra, n2), cp2), rb, normal)), d2 - radius, vb[1].pair);
        return;
      }
    }
  }
  float dist_v;
  Vec3 normal;
  length_normalize(offset, &dist_v, &normal);
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, scale(0.5f, mul_sub(add(mul_add(r.p1, ra, normal), r.p2), rb, normal)), dist_v - radius, make_pair(0, 0, 0, 0));
}

// convex_manifold_shapes.rs:132-240
__device__ void collide_box_sphere(LocalManifold& m, const BoxHull& h, Vec3 cb, float rb, Xf b_to_a, SimplexCache& cache) {
  m.point_count = 0;
  Vec3 center = xf_point(b_to_a, cb);
  gjk::Output out = gjk::distance(h.points, 8, &center, 1, Xf{zero3(), quat_identity()}, cache);
  const float radius_a = 0.0f;
  float radius = radius_a + rb;
  if (out.distance > radius + kSpeculativeDistance) {
    cache = SimplexCache{};
    return;
  }
  Vec3 normal;
  float sep;
  if (out.distance > 100.0f * kFltEps) {
    normal = normalize(sub(out.point_b, out.point_a));
    sep = out.distance - radius;
  } else {
    int best = -1;
    float best_d = -kFltMax;
    for (int i = 0; i < 6; ++i) {
      float d = plane_sep(h.planes[i], center);
      if (d > best_d) { best = i; best_d = d; }
    }
    normal = h.planes[best].n;
    sep = best_d - radius;
  }
  Vec3 c_a = mul_add(center, radius_a - dot(sub(center, out.point_a), normal), normal);
  Vec3 c_b = mul_sub(center, rb, normal);
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, lerp(c_a, c_b, 0.5f), sep, make_pair(0, 0, 0, 0));
}

// convex_manifold_shapes.rs:399-640 (queries and builders) and :643-797
__device__ void collide_box_capsule(LocalManifold& m, const BoxHull& h, Vec3 c1, Vec3 c2, float rc, Xf b_to_a, SimplexCache& cache) {
  m.point_count = 0;
  Vec3 cap[2] = {c1, c2};
  gjk::Output out = gjk::distance(h.points, 8, cap, 2, b_to_a, cache);
  if (out.distance > rc + kSpeculativeDistance) {
    cache = SimplexCache{};
    return;
  }
  if (out.distance > 100.0f * kFltEps) {
    Vec3 delta = out.normal;
    int ref_face = -1;
    float best = -kFltMax;
    for (int i = 0; i < 6; ++i) {
      float d = dot(h.planes[i].n, delta);
      if (d > best) { best = d; ref_face = i; }
    }
    Plane ref_plane = h.planes[ref_face];
    if (absf(dot(ref_plane.n, delta)) > 0.998f) {
      ClipVertex vb[2];
      vb[0] = ClipVertex{xf_point(b_to_a, c1), 0.0f, make_pair(0, 0, 0, 0)};
      vb[1] = ClipVertex{xf_point(b_to_a, c2), 0.0f, make_pair(0, 1, 0, 1)};
      if (clip_segment_to_box_face(vb, h, ref_face) == 2) {
        float d1 = plane_sep(ref_plane, vb[0].position), d2 = plane_sep(ref_plane, vb[1].position);
        if (d1 <= rc + kSpeculativeDistance || d2 <= rc + kSpeculativeDistance) {
          Vec3 normal = ref_plane.n;
          m.normal = normal;
          m.point_count = 2;
          set_point(m, 0, mul_sub(vb[0].position, 0.5f * (rc + d1), normal), d1 - rc, vb[0].pair);
          set_point(m, 1, mul_sub(vb[1].position, 0.5f * (rc + d2), normal), d2 - rc, vb[1].pair);
          return;
        }
      }
    }
    m.normal = delta;
    m.point_count = 1;
    set_point(m, 0, scale(0.5f, add(mul_sub(out.point_a, rc, delta), out.point_b)), out.distance - rc, make_pair(0, 0, 0, 0));
    return;
  }
  // deep overlap: face query then edge query (convex_manifold_shapes.rs:399-490)
  Vec3 pts[2] = {xf_point(b_to_a, c1), xf_point(b_to_a, c2)};
  int face = -1, face_vertex = -1;
  float face_sep = -kFltMax;
  for (int f = 0; f < 6; ++f) {
    Plane pl = h.planes[f];
    int vi = gjk::point_support(pts, 2, neg(pl.n));
    float sep = plane_sep(pl, pts[vi]);
    if (sep > face_sep) { face_vertex = vi; face = f; face_sep = sep; }
  }
  (void)face_vertex;
  if (face_sep > rc) return;
  Vec3 e_normal = zero3();
  float e_sep = -kFltMax;
  int e_a = kNullIndex, e_b = kNullIndex;
  {
    Vec3 p_a = pts[0], q_a = pts[1];
    Vec3 ea = sub(q_a, p_a);
    const float sq_tol = kParallelEdgeTolerance * kParallelEdgeTolerance;
    for (int i = 0; i < 24; i += 2) {
      Vec3 q_b = h.points[c_box_edge[i + 1][2]];
      Vec3 u_b = h.planes[c_box_edge[i][3]].n, v_b = h.planes[c_box_edge[i + 1][3]].n;
      float cba = dot(u_b, ea), dba = dot(v_b, ea);
      if (cba * dba < 0.0f) {
        if (maxf(cba * cba, dba * dba) < sq_tol * length_sq(ea)) continue;
        float t = cba / (cba - dba);
        Vec3 axis = normalize(lerp(u_b, v_b, t));
        float sep = dot(axis, sub(q_a, q_b));
        if (sep > e_sep) { e_normal = axis; e_sep = sep; e_a = 0; e_b = i; }
      }
    }
  }
  if (e_sep > rc) return;
  float face_separation = face_sep - rc;
  {
    Plane ref_plane = h.planes[face];
    ClipVertex seg[2];
    seg[0] = ClipVertex{pts[0], 0.0f, make_pair(0, 0, 0, 0)};
    seg[1] = ClipVertex{pts[1], 0.0f, make_pair(0, 1, 0, 1)};
    if (clip_segment_to_box_face(seg, h, face) == 2) {
      float d1 = plane_sep(ref_plane, seg[0].position), d2 = plane_sep(ref_plane, seg[1].position);
      if (d1 <= kSpeculativeDistance || d2 <= kSpeculativeDistance) {
        Vec3 normal = ref_plane.n;
        m.normal = normal;
        m.point_count = 2;
        set_point(m, 0, mul_sub(seg[0].position, 0.5f * (d1 + rc), normal), d1 - rc, seg[0].pair);
        set_point(m, 1, mul_sub(seg[1].position, 0.5f * (d2 + rc), normal), d2 - rc, seg[1].pair);
      }
    }
  }
  if (m.point_count == 2) face_separation = minf(m.pts[0].separation, m.pts[1].separation);
  if (e_a == kNullIndex) return;
  float edge_separation = e_sep - rc;
  if (m.point_count == 0 || edge_separation > face_separation + kLinearSlop) {
    Vec3 pc = pts[0], qc = pts[1];
    Vec3 ph = h.points[c_box_edge[e_b][2]], qh = h.points[c_box_edge[c_box_edge[e_b][1]][2]];
    SegDist r = line_distance(ph, sub(qh, ph), pc, sub(qc, pc));
    if (!((0.0f <= r.f1 && r.f1 <= 1.0f) && (0.0f <= r.f2 && r.f2 <= 1.0f))) return;
    m.normal = e_normal;
    m.point_count = 1;
    set_point(m, 0, scale(0.5f, add(mul_sub(r.p1, rc, e_normal), r.p2)), dot(e_normal, sub(r.p2, r.p1)) - rc,
              make_pair(0, e_a, 1, e_b));
  }
}

}  // namespace nsl
