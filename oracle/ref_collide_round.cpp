// Oracle: sphere and capsule narrow phase and the GJK distance they use.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// Restates src/collision/convex_manifold_shapes.rs:21-797 (collide_spheres,
// collide_capsule_and_sphere, collide_hull_and_sphere, collide_capsules, the
// hull-capsule queries and builders, collide_hull_and_capsule),
// src/collision/convex_manifold.rs:137-200 (clip_segment, clip_segment_to_hull_face)
// and src/collision/distance.rs:11-709 (GJK shape_distance with its simplex cache).
#include "ref_plant.h"
#include "ref_world.h"

namespace nslref {

namespace {

constexpr int kMaxGjkIterations = 32;  // distance.rs:11

struct ClipVertex { Vec3 position; float separation; FeaturePair pair; };
inline FeaturePair fp(int o1, int i1, int o2, int i2) { return FeaturePair{(uint8_t)o1, (uint8_t)i1, (uint8_t)o2, (uint8_t)i2}; }
constexpr FeaturePair kPairSingle = {0, 0, 0, 0};

inline Vec3 blend3(float s, Vec3 a, float t, Vec3 b, float u, Vec3 c) {
  return {s * a.x + t * b.x + u * c.x, s * a.y + t * b.y + u * c.y, s * a.z + t * b.z + u * c.z};
}

// distance.rs:34-50
int point_support(const Vec3* pts, int count, Vec3 axis) {
  Vec3 origin = pts[0];
  int best = 0;
  float best_proj = 0.0f;
  for (int i = 1; i < count; ++i) {
    float p = dot(axis, sub(pts[i], origin));
    if (p > best_proj) { best = i; best_proj = p; }
  }
  return best;
}

struct SimplexVertex { Vec3 wa, wb, w; float a; int32_t index_a, index_b; };
struct Simplex { SimplexVertex v[4]; int32_t count; };

struct Bary2 { float u, v, d; };
struct Bary3 { float u, v, w, d; };
inline Bary2 bary_edge(Vec3 a, Vec3 b) {
  Vec3 ab = sub(b, a);
  return Bary2{dot(b, ab), -dot(a, ab), dot(ab, ab)};
}
inline Bary3 bary_tri(Vec3 a, Vec3 b, Vec3 c) {
  Vec3 ab = sub(b, a), ac = sub(c, a);
  Vec3 bxc = cross(b, c), cxa = cross(c, a), axb = cross(a, b), n = cross(ab, ac);
  return Bary3{dot(bxc, n), dot(cxa, n), dot(axb, n), dot(n, n)};
}

float get_metric(const Simplex& s) {
  switch (s.count) {
    case 1: return 0.0f;
    case 2: return distance(s.v[0].w, s.v[1].w);
    case 3: return length(cross(sub(s.v[1].w, s.v[0].w), sub(s.v[2].w, s.v[0].w))) / 2.0f;
    case 4: return scalar_triple(sub(s.v[1].w, s.v[0].w), sub(s.v[2].w, s.v[0].w), sub(s.v[3].w, s.v[0].w)) / 6.0f;
    default: return 0.0f;
  }
}

void write_cache(SimplexCache& c, const Simplex& s) {
  c.metric = get_metric(s);
  c.count = (uint16_t)s.count;
  for (int i = 0; i < s.count; ++i) {
    c.index_a[i] = (uint8_t)s.v[i].index_a;
    c.index_b[i] = (uint8_t)s.v[i].index_b;
  }
}

bool solve2(Simplex& s) {
  Vec3 a = s.v[0].w, b = s.v[1].w;
  Vec3 ab = sub(b, a);
  float divisor = dot(ab, ab);
  float u = dot(b, ab), v = -dot(a, ab);
  if (v <= 0.0f) { s.count = 1; s.v[0].a = 1.0f; return true; }
  if (u <= 0.0f) { s.count = 1; s.v[0] = s.v[1]; s.v[0].a = 1.0f; return true; }
  if (divisor <= 0.0f) return false;
  float den = 1.0f / divisor;
  s.v[0].a = den * u;
  s.v[1].a = den * v;
  return true;
}

bool set_edge(Simplex& s, SimplexVertex x, SimplexVertex y, Bary2 w) {
  s.count = 2;
  s.v[0] = x;
  s.v[1] = y;
  if (w.d <= 0.0f) return false;
  s.v[0].a = w.u / w.d;
  s.v[1].a = w.v / w.d;
  return true;
}
bool set_tri(Simplex& s, SimplexVertex x, SimplexVertex y, SimplexVertex z, Bary3 w) {
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
void set_vertex(Simplex& s, SimplexVertex x) {
  s.count = 1;
  s.v[0] = x;
  s.v[0].a = 1.0f;
}

// distance.rs:160-240
bool solve3(Simplex& s) {
  SimplexVertex v1 = s.v[0], v2 = s.v[1], v3 = s.v[2];
  Bary2 ab = bary_edge(v1.w, v2.w), bc = bary_edge(v2.w, v3.w), ca = bary_edge(v3.w, v1.w);
  if (ab.v <= 0.0f && ca.u <= 0.0f) { set_vertex(s, v1); return true; }
  if (bc.v <= 0.0f && ab.u <= 0.0f) { set_vertex(s, v2); return true; }
  if (ca.v <= 0.0f && bc.u <= 0.0f) { set_vertex(s, v3); return true; }
  Bary3 abc = bary_tri(v1.w, v2.w, v3.w);
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
bool solve4(Simplex& s) {
  SimplexVertex A = s.v[0], B = s.v[1], C = s.v[2], D = s.v[3];
  Bary2 ab = bary_edge(A.w, B.w), ac = bary_edge(A.w, C.w), ad = bary_edge(A.w, D.w);
  Bary2 bc = bary_edge(B.w, C.w), cd = bary_edge(C.w, D.w), db = bary_edge(D.w, B.w);
  if (ab.v <= 0.0f && ac.v <= 0.0f && ad.v <= 0.0f) { set_vertex(s, A); return true; }
  if (ab.u <= 0.0f && db.u <= 0.0f && bc.v <= 0.0f) { set_vertex(s, B); return true; }
  if (ac.u <= 0.0f && bc.u <= 0.0f && cd.v <= 0.0f) { set_vertex(s, C); return true; }
  if (ad.u <= 0.0f && cd.u <= 0.0f && db.v <= 0.0f) { set_vertex(s, D); return true; }
  Bary3 acb = bary_tri(A.w, C.w, B.w), abd = bary_tri(A.w, B.w, D.w), adc = bary_tri(A.w, D.w, C.w), bcd = bary_tri(B.w, C.w, D.w);
  if (abd.w <= 0.0f && acb.v <= 0.0f && ab.u > 0.0f && ab.v > 0.0f) return set_edge(s, A, B, ab);
  if (acb.w <= 0.0f && adc.v <= 0.0f && ac.u > 0.0f && ac.v > 0.0f) return set_edge(s, A, C, ac);
  if (adc.w <= 0.0f && abd.v <= 0.0f && ad.u > 0.0f && ad.v > 0.0f) return set_edge(s, A, D, ad);
  if (acb.u <= 0.0f && bcd.w <= 0.0f && bc.u > 0.0f && bc.v > 0.0f) return set_edge(s, B, C, bc);
  if (adc.u <= 0.0f && bcd.u <= 0.0f && cd.u > 0.0f && cd.v > 0.0f) return set_edge(s, C, D, cd);
  if (abd.u <= 0.0f && bcd.v <= 0.0f && db.u > 0.0f && db.v > 0.0f) return set_edge(s, D, B, db);
  // barycentric_coords_tet
  Vec3 e1 = sub(B.w, A.w), e2 = sub(C.w, A.w), e3 = sub(D.w, A.w);
  float divisor = scalar_triple(e1, e2, e3);
  float sign = divisor < 0.0f ? -1.0f : 1.0f;
  float tu = sign * scalar_triple(B.w, C.w, D.w);
  float tv = sign * scalar_triple(A.w, D.w, C.w);
  float tw = sign * scalar_triple(A.w, B.w, D.w);
  float tx = sign * scalar_triple(A.w, C.w, B.w);
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

void witness_points(const Simplex& s, Vec3* a, Vec3* b) {
  const SimplexVertex* v = s.v;
  switch (s.count) {
    case 1: *a = v[0].wa; *b = v[0].wb; break;
    case 2:
      if (PLANTED(4)) {
        *a = blend2(v[1].a, v[0].wa, v[0].a, v[1].wa);
        *b = blend2(v[1].a, v[0].wb, v[0].a, v[1].wb);
        break;
      }
      *a = blend2(v[0].a, v[0].wa, v[1].a, v[1].wa);
      *b = blend2(v[0].a, v[0].wb, v[1].a, v[1].wb);
      break;
    case 3:
      *a = blend3(v[0].a, v[0].wa, v[1].a, v[1].wa, v[2].a, v[2].wa);
      *b = blend3(v[0].a, v[0].wb, v[1].a, v[1].wb, v[2].a, v[2].wb);
      break;
    case 4: {
      Vec3 sum = add(blend2(v[0].a, v[0].wa, v[1].a, v[1].wa), blend2(v[2].a, v[2].wa, v[3].a, v[3].wa));
      *a = sum;
      *b = sum;
      break;
    }
    default: break;
  }
}

Vec3 closest_point(const Simplex& s) {
  const SimplexVertex* v = s.v;
  switch (s.count) {
    case 1: return v[0].w;
    case 2: return blend2(v[0].a, v[0].w, v[1].a, v[1].w);
    case 3: return blend3(v[0].a, v[0].w, v[1].a, v[1].w, v[2].a, v[2].w);
    case 4: return add(blend2(v[0].a, v[0].w, v[1].a, v[1].w), blend2(v[2].a, v[2].w, v[3].a, v[3].w));
    default: return kZero3;
  }
}

// distance.rs:483-709 (use_radii is false at every narrow-phase call site)
DistanceOutput shape_distance(const Vec3* pa, int na, const Vec3* pb, int nb, Xf xf, SimplexCache& cache) {
  Mat3 m = matrix_from_quat(xf.q);
  Mat3 mt = transpose(m);
  Simplex s{};
  s.count = cache.count;
  for (int i = 0; i < cache.count; ++i) {
    int i1 = cache.index_a[i], i2 = cache.index_b[i];
    Vec3 v1 = pa[i1];
    Vec3 v2 = add(mul_mv(m, pb[i2]), xf.p);
    s.v[i] = SimplexVertex{v1, v2, sub(v2, v1), 0.0f, i1, i2};
  }
  if (s.count > 0) {
    float m1 = cache.metric, m2 = get_metric(s);
    if (2.0f * m1 < m2 || m2 < 0.5f * m1 || m2 < FLT_EPSILON) s.count = 0;
  }
  if (s.count == 0) {
    Vec3 v1 = pa[0];
    Vec3 v2 = add(mul_mv(m, pb[0]), xf.p);
    s.count = 1;
    s.v[0] = SimplexVertex{v1, v2, sub(v2, v1), 0.0f, 0, 0};
  }
  Simplex backup{};
  DistanceOutput out{};
  float dist_sq = FLT_MAX;
  Vec3 normal = kZero3;
  int iteration = 0;
  while (iteration < kMaxGjkIterations) {
    bool solved;
    switch (s.count) {
      case 1: s.v[0].a = 1.0f; solved = true; break;
      case 2: solved = solve2(s); break;
      case 3: solved = solve3(s); break;
      case 4: solved = solve4(s); break;
      default: solved = false; break;
    }
    if (!solved) { s = backup; break; }
    if (s.count == 4) {
      witness_points(s, &out.point_a, &out.point_b);
      out.normal = kZero3;
      out.distance = 0.0f;
      out.iterations = iteration;
      return out;
    }
    float old = dist_sq;
    Vec3 cp = closest_point(s);
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
      Vec3 a = s.v[0].w, b = s.v[1].w, c = s.v[2].w;
      Vec3 n = cross(sub(b, a), sub(c, a));
      dir = dot(n, a) < 0.0f ? n : neg(n);
    }
    if (length_sq(dir) < 1000.0f * FLT_MIN) {
      witness_points(s, &out.point_a, &out.point_b);
      out.normal = kZero3;
      out.distance = 0.0f;
      out.iterations = iteration;
      return out;
    }
    normal = neg(dir);
    int ia = point_support(pa, na, neg(dir));
    Vec3 sa = pa[ia];
    Vec3 dir2 = mul_mv(mt, dir);
    int ib = point_support(pb, nb, dir2);
    Vec3 sb = add(mul_mv(m, pb[ib]), xf.p);
    backup = s;
    bool duplicate = false;
    for (int i = 0; i < s.count; ++i)
      if (s.v[i].index_a == ia && s.v[i].index_b == ib) { duplicate = true; break; }
    if (duplicate) break;
    s.v[s.count] = SimplexVertex{sa, sb, sub(sb, sa), 0.0f, ia, ib};
    s.count += 1;
    iteration += 1;
  }
  witness_points(s, &out.point_a, &out.point_b);
  out.iterations = iteration;
  normal = normalize(normal);
  if (!is_normalized(normal)) {
    out.distance = 0.0f;
    out.normal = kZero3;
    return out;
  }
  out.distance = distance(out.point_a, out.point_b);
  out.normal = normal;
  write_cache(cache, s);
  return out;
}

// convex_manifold.rs:137-170
int clip_segment(ClipVertex* seg, Plane plane) {
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

// convex_manifold.rs:172-200
int clip_segment_to_hull_face(ClipVertex* seg, const Hull& h, int ref_face) {
  Plane ref_plane = h.planes[ref_face];
  int start = h.face_edge[ref_face];
  int e = start;
  for (;;) {
    int nx = h.edges[e].next;
    Vec3 p1 = h.points[h.edges[e].origin], p2 = h.points[h.edges[nx].origin];
    Vec3 binormal = cross(normalize(sub(p2, p1)), ref_plane.n);
    if (clip_segment(seg, plane_from_normal_point(binormal, p1)) < 2) return 0;
    e = nx;
    if (e == start) break;
  }
  return 2;
}

inline void set_point(LocalManifold& m, int i, Vec3 p, float sep, FeaturePair pair) {
  m.pts[i].point = p;
  m.pts[i].separation = sep;
  m.pts[i].pair = pair;
  m.pts[i].triangle_index = 0;
}

}  // namespace

// Entry points of the GJK restatement for the time-of-impact query.
int gjk_point_support(const Vec3* pts, int count, Vec3 axis) { return point_support(pts, count, axis); }
DistanceOutput gjk_distance(const Vec3* pa, int na, const Vec3* pb, int nb, Xf xf, SimplexCache& cache) {
  return shape_distance(pa, na, pb, nb, xf, cache);
}

// convex_manifold_shapes.rs:21-71
void collide_spheres(LocalManifold& m, Sphere a, Sphere b, Xf b_to_a) {
  Vec3 c1 = a.center, c2 = xf_point(b_to_a, b.center);
  float total = a.radius + b.radius;
  float max_d = total + speculative_distance();
  Vec3 offset = sub(c2, c1);
  float dsq = length_sq(offset);
  if (dsq > max_d * max_d) return;
  Vec3 normal = {0.0f, 1.0f, 0.0f};
  float d = std::sqrt(dsq);
  if (d * d > 1000.0f * FLT_MIN) normal = scale(1.0f / d, offset);
  Vec3 point = scale(0.5f, mul_sub(add(PLANTED(1) ? mul_sub(c1, a.radius, normal) : mul_add(c1, a.radius, normal), c2),
                                    b.radius, normal));
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, point, d - total, kPairSingle);
}

// convex_manifold_shapes.rs:74-129
void collide_capsule_and_sphere(LocalManifold& m, Capsule a, Sphere b, Xf b_to_a) {
  m.point_count = 0;
  Vec3 center = xf_point(b_to_a, b.center);
  float total = b.radius + a.radius;
  float max_d = total + speculative_distance();
  Vec3 closest = point_to_segment(a.c1, a.c2, center);
  Vec3 offset = sub(center, closest);
  float dsq = length_sq(offset);
  if (dsq > max_d * max_d) return;
  Vec3 normal = {0.0f, 1.0f, 0.0f};
  float d = std::sqrt(dsq);
  if (d * d > 1000.0f * FLT_MIN) normal = scale(1.0f / d, offset);
  Vec3 point = scale(0.5f, mul_add(add(mul_sub(center, b.radius, normal), closest), a.radius, normal));
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, point, PLANTED(2) ? d - b.radius : d - total, kPairSingle);
}

// convex_manifold_shapes.rs:132-240
void collide_hull_and_sphere(LocalManifold& m, const Hull& a, Sphere b, Xf b_to_a, SimplexCache& cache) {
  m.point_count = 0;
  Vec3 center = xf_point(b_to_a, b.center);
  const float spec = speculative_distance();
  DistanceOutput out = shape_distance(a.points, a.vertex_count, &center, 1, Xf{kZero3, kQuatIdentity}, cache);
  const float radius_a = 0.0f, radius_b = b.radius;
  float radius = radius_a + radius_b;
  if (out.distance > radius + spec) {
    cache = SimplexCache{};
    return;
  }
  if (out.distance > 100.0f * FLT_EPSILON) {
    Vec3 normal = normalize(sub(out.point_b, out.point_a));
    Vec3 c_a = mul_add(center, radius_a - dot(sub(center, out.point_a), normal), normal);
    Vec3 c_b = mul_sub(center, radius_b, normal);
    m.normal = normal;
    m.point_count = 1;
    set_point(m, 0, lerp(c_a, c_b, 0.5f), out.distance - radius, kPairSingle);
  } else {
    int best = -1;
    float best_d = -FLT_MAX;
    for (int i = 0; i < a.face_count; ++i) {
      float d = plane_sep(a.planes[i], center);
      if (d > best_d) { best = i; best_d = d; }
    }
    Vec3 normal = a.planes[best].n;
    Vec3 c_a = mul_add(center, radius_a - dot(sub(center, out.point_a), normal), normal);
    Vec3 c_b = mul_sub(center, radius_b, normal);
    m.normal = normal;
    m.point_count = 1;
    set_point(m, 0, lerp(c_a, c_b, 0.5f), PLANTED(5) ? best_d + radius : best_d - radius, kPairSingle);
  }
}

// convex_manifold_shapes.rs:245-394
void collide_capsules(LocalManifold& m, Capsule a, Capsule b, Xf b_to_a) {
  m.point_count = 0;
  Vec3 a1 = a.c1, a2 = a.c2;
  Vec3 b1 = xf_point(b_to_a, b.c1), b2 = xf_point(b_to_a, b.c2);
  float radius = a.radius + b.radius;
  float max_d = radius + speculative_distance();
  SegDist r = segment_distance(a1, a2, b1, b2);
  Vec3 offset = sub(r.p2, r.p1);
  float dsq = length_sq(offset);
  const float min_d = 0.01f * linear_slop();
  if (dsq > max_d * max_d || dsq < min_d * min_d) return;
  // get_length_and_normalize (functions.rs:277-287)
  auto len_norm = [](Vec3 v, float* len, Vec3* n) {
    float ls = v.x * v.x + v.y * v.y + v.z * v.z;
    if (ls > 1000.0f * FLT_MIN) {
      float l = std::sqrt(ls);
      float s = 1.0f / l;
      *len = l;
      *n = Vec3{s * v.x, s * v.y, s * v.z};
    } else {
      *len = 0.0f;
      *n = kZero3;
    }
  };
  float len_a, len_b;
  Vec3 edge_a, edge_b;
  len_norm(sub(a2, a1), &len_a, &edge_a);
  if (len_a < min_capsule_length()) return;
  len_norm(sub(b2, b1), &len_b, &edge_b);
  if (len_b < min_capsule_length()) return;
  const float alpha_tol = 0.05f;
  Vec3 axis = cross(edge_a, edge_b);
  if (length_sq(axis) < alpha_tol * alpha_tol) {
    Plane planes_a[2] = {Plane{neg(edge_a), -dot(edge_a, a.c1)}, Plane{edge_a, dot(edge_a, a.c2)}};
    ClipVertex vb[2];
    vb[0] = ClipVertex{b1, 0.0f, fp(0, 0, 0, 0)};
    vb[1] = ClipVertex{b2, 0.0f, fp(0, 1, 0, 1)};
    int count = clip_segment(vb, planes_a[0]);
    if (count == 2) count = clip_segment(vb, planes_a[1]);
    if (count == 2) {
      Vec3 cp1 = point_to_segment(a1, a2, vb[0].position);
      Vec3 cp2 = point_to_segment(a1, a2, vb[1].position);
      float d1 = distance(cp1, vb[0].position), d2 = distance(cp2, vb[1].position);
      if (d1 <= radius && d2 <= radius) {
        if (d1 < min_d || d2 < min_d) return;
        Vec3 n1 = scale(1.0f / d1, PLANTED(3) ? sub(cp1, vb[0].position) : sub(vb[0].position, cp1));
        Vec3 n2 = scale(1.0f / d2, sub(vb[1].position, cp2));
        Vec3 normal = normalize(add(n1, n2));
        Vec3 p1 = scale(0.5f, mul_sub(add(mul_add(vb[0].position, a.radius, n1), cp1), b.radius, normal));
        Vec3 p2 = scale(0.5f, mul_sub(add(mul_add(vb[1].position, a.radius, n2), cp2), b.radius, normal));
        m.normal = normal;
        m.point_count = 2;
        set_point(m, 0, p1, d1 - radius, vb[0].pair);
        set_point(m, 1, p2, d2 - radius, vb[1].pair);
        return;
      }
    }
  }
  float dist_v;
  Vec3 normal;
  len_norm(offset, &dist_v, &normal);
  Vec3 point = scale(0.5f, mul_sub(add(mul_add(r.p1, a.radius, normal), r.p2), b.radius, normal));
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, point, dist_v - radius, kPairSingle);
}

namespace {
struct CapsuleAxis { Vec3 normal; float separation; int32_t index_a, index_b; };

// convex_manifold_shapes.rs:399-432
CapsuleAxis query_face_hull_capsule(const Hull& h, Capsule c, Xf xf) {
  int max_face = -1, max_vertex = -1;
  float max_sep = -FLT_MAX;
  Vec3 pts[2] = {xf_point(xf, c.c1), xf_point(xf, c.c2)};
  for (int f = 0; f < h.face_count; ++f) {
    Plane pl = h.planes[f];
    int vi = point_support(pts, 2, neg(pl.n));
    float sep = plane_sep(pl, pts[vi]);
    if (sep > max_sep) { max_vertex = vi; max_face = f; max_sep = sep; }
  }
  return CapsuleAxis{h.planes[max_face].n, max_sep, max_face, max_vertex};
}

// convex_manifold_shapes.rs:435-490
CapsuleAxis query_edge_hull_capsule(const Hull& h, Capsule c, Xf xf) {
  Vec3 max_normal = kZero3;
  float max_sep = -FLT_MAX;
  int max_a = kNullIndex, max_b = kNullIndex;
  Vec3 p_a = xf_point(xf, c.c1), q_a = xf_point(xf, c.c2);
  Vec3 e_a = sub(q_a, p_a);
  const float sq_tol = kParallelEdgeTolerance * kParallelEdgeTolerance;
  for (int i = 0; i < h.edge_count; i += 2) {
    HalfEdge e = h.edges[i], t = h.edges[i + 1];
    Vec3 q_b = h.points[t.origin];
    Vec3 u_b = h.planes[e.face].n, v_b = h.planes[t.face].n;
    float cba = dot(u_b, e_a), dba = dot(v_b, e_a);
    if (cba * dba < 0.0f) {
      if (maxf(cba * cba, dba * dba) < sq_tol * length_sq(e_a)) continue;
      float tt = cba / (cba - dba);
      Vec3 axis = normalize(lerp(u_b, v_b, tt));
      float sep = dot(axis, sub(q_a, q_b));
      if (sep > max_sep) { max_normal = axis; max_sep = sep; max_a = 0; max_b = i; }
    }
  }
  return CapsuleAxis{max_normal, max_sep, max_a, max_b};
}

// convex_manifold_shapes.rs:493-545
bool hull_face_capsule_contact(LocalManifold& m, const Hull& h, Capsule c, Xf xf, CapsuleAxis q) {
  int ref_face = q.index_a;
  Plane ref_plane = h.planes[ref_face];
  ClipVertex seg[2];
  seg[0] = ClipVertex{xf_point(xf, c.c1), 0.0f, fp(0, 0, 0, 0)};
  seg[1] = ClipVertex{xf_point(xf, c.c2), 0.0f, fp(0, 1, 0, 1)};
  if (clip_segment_to_hull_face(seg, h, ref_face) < 2) return false;
  float d1 = plane_sep(ref_plane, seg[0].position), d2 = plane_sep(ref_plane, seg[1].position);
  const float spec = speculative_distance();
  if (d1 <= spec || d2 <= spec) {
    Vec3 normal = ref_plane.n;
    m.normal = normal;
    m.point_count = 2;
    set_point(m, 0, mul_sub(seg[0].position, 0.5f * (d1 + c.radius), normal), d1 - c.radius, seg[0].pair);
    set_point(m, 1, mul_sub(seg[1].position, 0.5f * (d2 + c.radius), normal), d2 - c.radius, seg[1].pair);
    return true;
  }
  return false;
}

// convex_manifold_shapes.rs:548-640
bool hull_capsule_edge_contact(LocalManifold& m, const Hull& h, Capsule c, Xf xf, CapsuleAxis q) {
  Vec3 pc = xf_point(xf, c.c1), qc = xf_point(xf, c.c2);
  Vec3 ec = sub(qc, pc);
  HalfEdge e2 = h.edges[q.index_b], t2 = h.edges[e2.twin];
  Vec3 ph = h.points[e2.origin], qh = h.points[t2.origin];
  Vec3 eh = sub(qh, ph);
  Vec3 normal = q.normal;
  SegDist r = line_distance(ph, eh, pc, ec);
  if (!((0.0f <= r.f1 && r.f1 <= 1.0f) && (0.0f <= r.f2 && r.f2 <= 1.0f))) return false;
  Vec3 point = scale(0.5f, add(mul_sub(r.p1, c.radius, normal), r.p2));
  float sep = dot(normal, sub(r.p2, r.p1));
  m.normal = normal;
  m.point_count = 1;
  set_point(m, 0, point, sep - c.radius, fp(0, q.index_a, 1, q.index_b));
  return true;
}
}  // namespace

// convex_manifold_shapes.rs:643-797
void collide_hull_and_capsule(LocalManifold& m, const Hull& h, Capsule c, Xf b_to_a, SimplexCache& cache) {
  m.point_count = 0;
  Vec3 cap_pts[2] = {c.c1, c.c2};
  DistanceOutput out = shape_distance(h.points, h.vertex_count, cap_pts, 2, b_to_a, cache);
  const float spec = speculative_distance();
  if (out.distance > c.radius + spec) {
    cache = SimplexCache{};
    return;
  }
  if (out.distance > 100.0f * FLT_EPSILON) {
    Vec3 delta = out.normal;
    // find_hull_support_face (hull.rs:504-520)
    int ref_face = -1;
    float best = -FLT_MAX;
    for (int i = 0; i < h.face_count; ++i) {
      float d = dot(h.planes[i].n, delta);
      if (d > best) { best = d; ref_face = i; }
    }
    Plane ref_plane = h.planes[ref_face];
    if (absf(dot(ref_plane.n, delta)) > 0.998f) {
      ClipVertex vb[2];
      vb[0] = ClipVertex{xf_point(b_to_a, c.c1), 0.0f, fp(0, 0, 0, 0)};
      vb[1] = ClipVertex{xf_point(b_to_a, c.c2), 0.0f, fp(0, 1, 0, 1)};
      if (clip_segment_to_hull_face(vb, h, ref_face) == 2) {
        float d1 = plane_sep(ref_plane, vb[0].position), d2 = plane_sep(ref_plane, vb[1].position);
        if (d1 <= c.radius + spec || d2 <= c.radius + spec) {
          Vec3 normal = ref_plane.n;
          m.normal = normal;
          m.point_count = 2;
          set_point(m, 0, mul_sub(vb[0].position, 0.5f * (PLANTED(6) ? c.radius - d1 : c.radius + d1), normal), d1 - c.radius,
                    vb[0].pair);
          set_point(m, 1, mul_sub(vb[1].position, 0.5f * (PLANTED(6) ? c.radius - d2 : c.radius + d2), normal), d2 - c.radius,
                    vb[1].pair);
          return;
        }
      }
    }
    Vec3 point = scale(0.5f, add(mul_sub(out.point_a, c.radius, delta), out.point_b));
    m.normal = delta;
    m.point_count = 1;
    set_point(m, 0, point, out.distance - c.radius, kPairSingle);
    return;
  }
  CapsuleAxis fq = query_face_hull_capsule(h, c, b_to_a);
  if (fq.separation > c.radius) return;
  CapsuleAxis eq = query_edge_hull_capsule(h, c, b_to_a);
  if (eq.separation > c.radius) return;
  float face_sep = fq.separation - c.radius;
  hull_face_capsule_contact(m, h, c, b_to_a, fq);
  if (m.point_count == 2) face_sep = minf(m.pts[0].separation, m.pts[1].separation);
  if (eq.index_a == kNullIndex) return;
  float edge_sep = eq.separation - c.radius;
  if (m.point_count == 0 || edge_sep > face_sep + linear_slop()) hull_capsule_edge_contact(m, h, c, b_to_a, eq);
}

}  // namespace nslref
