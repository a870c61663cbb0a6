// Oracle: hull-vs-hull narrow phase (boxes are hulls).
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// Restates src/collision/convex_manifold.rs (separating-axis test with the cached
// axis, face clipping, edge contacts, manifold reduction) and the clipping helpers of
// src/collision/manifold.rs, operation for operation.
#include "ref_plant.h"
#include "ref_world.h"

namespace nslref {

namespace {

constexpr float kParallelTol = 1e-4f;      // convex_manifold.rs PARALLEL_TOL
constexpr int kEdgeProbeMinTests = 10;     // EDGE_PROBE_MIN_TESTS
constexpr uint32_t kHullBitCount = 7;      // HULL_BIT_COUNT
constexpr int kMaxClipPoints = 64;         // manifold.rs MAX_CLIP_POINTS

struct ClipVertex { Vec3 position; float separation; FeaturePair pair; };

inline FeaturePair make_pair(int o1, int i1, int o2, int i2) {
  return FeaturePair{(uint8_t)o1, (uint8_t)i1, (uint8_t)o2, (uint8_t)i2};
}
// manifold.rs:82-92
inline FeaturePair flip_pair(FeaturePair p) {
  return FeaturePair{(uint8_t)(1 - p.owner2), p.index2, (uint8_t)(1 - p.owner1), p.index1};
}

// convex_manifold.rs:212-231
bool test_edge_candidate_sorted(float a1, float a2, float c, float bound) {
  float hi = maxf(a1, a2);
  float lo = minf(a1, a2);
  bool exterior = hi >= bound;
  float u = lo - c * hi;
  bool max_is_interior = u >= 0.0f;
  float s = 1.0f - c * c;
  float bound_term = (bound - hi) * (bound + hi);
  float lhs = u * u;
  float rhs = bound_term * s;
  bool max_beats_bound = lhs >= rhs;
  bool nearly_parallel = s < kParallelTol;
  bool interior = max_is_interior && (max_beats_bound || nearly_parallel);
  return exterior || interior;
}

inline int round_up4(int n) { return (n + 3) & ~3; }

void get_face_dots(const Hull& h, Vec3 d, float* dots) {
  for (int i = 0; i < h.face_count; ++i) dots[i] = d.x * h.nx[i] + d.y * h.ny[i] + d.z * h.nz[i];
}
void get_face_plane_separations(const Hull& h, Vec3 c, float* seps) {
  get_face_dots(h, c, seps);
  for (int i = 0; i < h.face_count; ++i) seps[i] -= h.planes[i].off;
}
int filter_edge_candidates(const Hull& h, const float* seps, float bound, int32_t* idx, int count) {
  int kept = 0;
  for (int k = 0; k < count; ++k) {
    int i = idx[k];
    int i1 = h.edges[i].face, i2 = h.edges[i + 1].face;
    float c = h.edge_cosines[i >> 1];
    idx[kept] = i;
    if (test_edge_candidate_sorted(seps[i1], seps[i2], c, bound)) kept += 1;
  }
  return kept;
}

// convex_manifold.rs:293-322: vertex index carried in the low mantissa bits.
float get_support_wide(Vec3 n, const float* vx, const float* vy, const float* vz, int count, float bias, int* out_index) {
  const uint32_t mask = (1u << kHullBitCount) - 1;
  float min_value = INFINITY;
  for (int i = 0; i < count; ++i) {
    float d = n.x * vx[i] + n.y * vy[i] + n.z * vz[i];
    float value = bias - d;
    float augmented = u2f((f2u(value) & ~mask) | ((uint32_t)i & mask));
    if (augmented < min_value) min_value = augmented;
  }
  int vi = (int)(f2u(min_value) & mask);
  *out_index = vi;
  return n.x * vx[vi] + n.y * vy[vi] + n.z * vz[vi];
}

float get_face_separation(Vec3 dir, float plane_separation, const Hull& h, int count, Vec3 center, Vec3 extents, int* vertex_index) {
  float bias = dot(dir, center) + 1.0625f * dot(abs3(dir), extents);
  float support = get_support_wide(dir, h.vx, h.vy, h.vz, count, bias, vertex_index);
  return plane_separation - support;
}

void negative_transform_soa(Mat3 r, Vec3 p, const float* ix, const float* iy, const float* iz, int n, float* ox, float* oy,
                            float* oz, bool is_point) {
  Vec3 r0 = {r.cx.x, r.cy.x, r.cz.x}, r1 = {r.cx.y, r.cy.y, r.cz.y}, r2 = {r.cx.z, r.cy.z, r.cz.z};
  for (int i = 0; i < n; ++i) {
    Vec3 v = {ix[i], iy[i], iz[i]};
    Vec3 out = {dot(r0, v), dot(r1, v), dot(r2, v)};
    if (is_point) out = add(out, p);
    ox[i] = -out.x;
    oy[i] = -out.y;
    oz[i] = -out.z;
  }
}

// convex_manifold.rs:392-899
AxisQuery compute_separating_axis(const Hull& ha, const Hull& hb, Xf xf_b, bool early_return) {
  float dot_a[8] = {}, dot_b[8] = {}, plane_dot_a[8] = {}, plane_dot_b[8] = {};
  int32_t edge_idx_a[24], edge_idx_b[24];
  Mat3 r = matrix_from_quat(xf_b.q);
  Mat3 inv_r = transpose(r);
  const float spec = speculative_distance();

  AxisQuery res;
  res.face_a = SepAxis{kZero3, -INFINITY, kNullIndex, kNullIndex, kFaceAxisA};
  res.face_b = SepAxis{kZero3, -INFINITY, kNullIndex, kNullIndex, kFaceAxisB};
  res.edge = SepAxis{kZero3, -INFINITY, kNullIndex, kNullIndex, kEdgePairAxis};
  res.separated = kInvalidAxis;

  const int wide_vc_b = round_up4(hb.vertex_count);
  Vec3 c_b = aabb_center(hb.aabb), h_b = aabb_extents(hb.aabb);
  Vec3 delta_center = sub(add(mul_mv(r, hb.center), xf_b.p), ha.center);
  float center_distance = length(delta_center);
  float radius = ha.inner_radius + hb.inner_radius;
  float radius_bound = radius - (linear_slop() + 0.001f * (center_distance + radius));

  get_face_dots(ha, delta_center, dot_a);
  int seed_a = 0;
  float max_dot_a = dot_a[0];
  for (int i = 1; i < ha.face_count; ++i)
    if (dot_a[i] > max_dot_a) { max_dot_a = dot_a[i]; seed_a = i; }

  float floor_a = -INFINITY, seed_sep_a = -INFINITY;
  int seed_vertex_b = 0;
  if (early_return) {
    Plane pl = ha.planes[seed_a];
    Vec3 dir = neg(mul_mv(inv_r, pl.n));
    float ps = dot(pl.n, xf_b.p) - pl.off;
    seed_sep_a = get_face_separation(dir, ps, hb, wide_vc_b, c_b, h_b, &seed_vertex_b);
    floor_a = minf(seed_sep_a, spec);
  }
  for (int i = 0; i < ha.face_count; ++i) {
    if (dot_a[i] - radius_bound < maxf(floor_a, res.face_a.separation)) continue;
    Plane pl = ha.planes[i];
    int vertex_index = seed_vertex_b;
    float separation = seed_sep_a;
    if (!early_return || i != seed_a) {
      Vec3 dir = neg(mul_mv(inv_r, pl.n));
      float ps = dot(pl.n, xf_b.p) - pl.off;
      separation = get_face_separation(dir, ps, hb, wide_vc_b, c_b, h_b, &vertex_index);
    }
    if (separation > res.face_a.separation) {
      res.face_a.normal = pl.n;
      res.face_a.separation = separation;
      res.face_a.index_a = i;
      res.face_a.index_b = vertex_index;
      if (separation > spec && early_return) {
        res.separated = kFaceAxisA;
        return res;
      }
    }
  }

  const int wide_vc_a = round_up4(ha.vertex_count);
  Vec3 c_a = aabb_center(ha.aabb), h_a = aabb_extents(ha.aabb);
  get_face_dots(hb, neg(mul_mv(inv_r, delta_center)), dot_b);
  int seed_b = 0;
  float max_dot_b = dot_b[0];
  for (int i = 1; i < hb.face_count; ++i)
    if (dot_b[i] > max_dot_b) { max_dot_b = dot_b[i]; seed_b = i; }

  float floor_b = -INFINITY, seed_sep_b = -INFINITY;
  int seed_vertex_a = 0;
  if (early_return) {
    Plane pl = hb.planes[seed_b];
    Vec3 dir = neg(mul_mv(r, pl.n));
    float ps = dot(dir, xf_b.p) - pl.off;
    seed_sep_b = get_face_separation(dir, ps, ha, wide_vc_a, c_a, h_a, &seed_vertex_a);
    floor_b = maxf(seed_sep_b, res.face_a.separation);
    floor_b = minf(floor_b, spec);
  }
  for (int i = 0; i < hb.face_count; ++i) {
    if (dot_b[i] - radius_bound < maxf(floor_b, res.face_b.separation)) continue;
    Plane pl = hb.planes[i];
    Vec3 dir = neg(mul_mv(r, pl.n));
    int vertex_index = seed_vertex_a;
    float separation = seed_sep_b;
    if (!early_return || i != seed_b) {
      float ps = dot(dir, xf_b.p) - pl.off;
      separation = get_face_separation(dir, ps, ha, wide_vc_a, c_a, h_a, &vertex_index);
    }
    if (separation > res.face_b.separation) {
      res.face_b.normal = dir;
      res.face_b.separation = separation;
      res.face_b.index_a = vertex_index;
      res.face_b.index_b = i;
      if (separation > spec && early_return) {
        res.separated = kFaceAxisB;
        return res;
      }
    }
  }

  int na = 0, nb = 0;
  float max_face_sep = maxf(res.face_a.separation, res.face_b.separation);
  float bound_slack = radius - radius_bound;
  float threshold_a = early_return ? max_face_sep + hb.inner_radius - bound_slack : -INFINITY;
  float threshold_b = early_return ? max_face_sep + ha.inner_radius - bound_slack : -INFINITY;
  Vec3 center_b_in_a = add(mul_mv(r, hb.center), xf_b.p);
  Vec3 center_a_in_b = mul_mv(inv_r, sub(ha.center, xf_b.p));
  get_face_plane_separations(ha, center_b_in_a, plane_dot_a);
  get_face_plane_separations(hb, center_a_in_b, plane_dot_b);

  for (int i = 0; i < ha.edge_count; i += 2) {
    int i1 = ha.edges[i].face, i2 = ha.edges[i + 1].face;
    float c = ha.edge_cosines[i >> 1];
    edge_idx_a[na] = i;
    if (test_edge_candidate_sorted(plane_dot_a[i1], plane_dot_a[i2], c, threshold_a)) na += 1;
  }
  for (int i = 0; i < hb.edge_count; i += 2) {
    int i1 = hb.edges[i].face, i2 = hb.edges[i + 1].face;
    float c = hb.edge_cosines[i >> 1];
    edge_idx_b[nb] = i;
    if (test_edge_candidate_sorted(plane_dot_b[i1], plane_dot_b[i2], c, threshold_b)) nb += 1;
  }

  if (early_return && nb * ((na + 3) >> 2) >= kEdgeProbeMinTests) {
    float probe_bound = max_face_sep - (0.1f * linear_slop() + 0.001f * absf(center_distance + radius));
    int vb = res.face_a.index_b != kNullIndex ? res.face_a.index_b : seed_vertex_b;
    Vec3 probe_b = {hb.vx[vb], hb.vy[vb], hb.vz[vb]};
    get_face_plane_separations(ha, add(mul_mv(r, probe_b), xf_b.p), plane_dot_a);
    na = filter_edge_candidates(ha, plane_dot_a, probe_bound, edge_idx_a, na);
    int va = res.face_b.index_a != kNullIndex ? res.face_b.index_a : seed_vertex_a;
    Vec3 probe_a = {ha.vx[va], ha.vy[va], ha.vz[va]};
    get_face_plane_separations(hb, mul_mv(inv_r, sub(probe_a, xf_b.p)), plane_dot_b);
    nb = filter_edge_candidates(hb, plane_dot_b, probe_bound, edge_idx_b, nb);
  }
  if (na == 0 || nb == 0) return res;

  float b_fn_x[8], b_fn_y[8], b_fn_z[8], b_w_x[8], b_w_y[8], b_w_z[8];
  negative_transform_soa(r, xf_b.p, hb.nx, hb.ny, hb.nz, round_up4(hb.face_count), b_fn_x, b_fn_y, b_fn_z, false);
  negative_transform_soa(r, xf_b.p, hb.vx, hb.vy, hb.vz, wide_vc_b, b_w_x, b_w_y, b_w_z, true);

  float a_n0[12][3], a_n1[12][3], a_d[12][3], a_v0[12][3], a_tol[12];
  const float squared_tol = kParallelEdgeTolerance * kParallelEdgeTolerance;
  for (int k = 0; k < na; ++k) {
    HalfEdge e = ha.edges[edge_idx_a[k]], t = ha.edges[edge_idx_a[k] + 1];
    Vec3 a = ha.planes[e.face].n, b = ha.planes[t.face].n;
    a_n0[k][0] = a.x; a_n0[k][1] = a.y; a_n0[k][2] = a.z;
    a_n1[k][0] = b.x; a_n1[k][1] = b.y; a_n1[k][2] = b.z;
    int v0 = e.origin, v1 = t.origin;
    a_d[k][0] = ha.vx[v1] - ha.vx[v0];
    a_d[k][1] = ha.vy[v1] - ha.vy[v0];
    a_d[k][2] = ha.vz[v1] - ha.vz[v0];
    a_v0[k][0] = ha.vx[v0]; a_v0[k][1] = ha.vy[v0]; a_v0[k][2] = ha.vz[v0];
    a_tol[k] = squared_tol * (a_d[k][0] * a_d[k][0] + a_d[k][1] * a_d[k][1] + a_d[k][2] * a_d[k][2]);
  }
  const float ls = linear_slop();
  const float eps = -ls * ls;
  for (int j = 0; j < nb; ++j) {
    HalfEdge e = hb.edges[edge_idx_b[j]], t = hb.edges[edge_idx_b[j] + 1];
    int f0 = e.face, f1 = t.face, v0 = e.origin, v1 = t.origin;
    Vec3 c = {b_fn_x[f0], b_fn_y[f0], b_fn_z[f0]};
    Vec3 dn = {b_fn_x[f1], b_fn_y[f1], b_fn_z[f1]};
    Vec3 bv0 = {b_w_x[v0], b_w_y[v0], b_w_z[v0]};
    Vec3 bv1 = {b_w_x[v1], b_w_y[v1], b_w_z[v1]};
    Vec3 dc = sub(bv1, bv0);
    for (int i = 0; i < na; ++i) {
      Vec3 d = {a_d[i][0], a_d[i][1], a_d[i][2]};
      float cba = dot(c, d);
      float dba = dot(dn, d);
      if (cba * dba >= eps) continue;
      Vec3 n0 = {a_n0[i][0], a_n0[i][1], a_n0[i][2]};
      Vec3 n1 = {a_n1[i][0], a_n1[i][1], a_n1[i][2]};
      float adc = dot(n0, dc);
      float bdc = dot(n1, dc);
      if (adc * bdc >= eps || cba * bdc >= eps) continue;
      float max_cd = maxf(cba * cba, dba * dba);
      if (max_cd <= a_tol[i]) continue;
      float t_ = PLANTED(8) ? cba / (dba - cba) : -cba / (dba - cba);
      Vec3 n = mul_add(c, t_, sub(dn, c));
      float len2 = dot(n, n);
      float inv = 1.0f / std::sqrt(len2);
      n = scale(inv, n);
      Vec3 av0 = {a_v0[i][0], a_v0[i][1], a_v0[i][2]};
      float separation = -dot(add(av0, bv0), n);
      if (separation > res.edge.separation) {
        res.edge.normal = n;
        res.edge.separation = separation;
        res.edge.index_a = edge_idx_a[i];
        res.edge.index_b = edge_idx_b[j];
        if (separation > spec && early_return) {
          res.separated = kEdgePairAxis;
          return res;
        }
      }
    }
  }
  return res;
}

// convex_manifold.rs:905-1062
void reduce_manifold_points(LocalManifold& m, LocalPoint* points, int count) {
  if (count <= 4) {
    for (int i = 0; i < count; ++i) m.pts[i] = points[i];
    m.point_count = count;
    return;
  }
  Vec3 normal = m.normal;
  const float spec = speculative_distance();
  const float tol_sqr = spec * spec;
  const float bias = 0.95f;
  int best_index = kNullIndex;
  float best_score = -FLT_MAX;
  Vec3 search = arbitrary_perp(normal);
  for (int i = 0; i < count; ++i) {
    const LocalPoint& pt = points[i];
    if (pt.separation > spec) continue;
    float score = -pt.separation + dot(search, pt.point);
    if (bias * score > best_score) { best_index = i; best_score = score; }
  }
  if (best_index == kNullIndex) { m.point_count = 0; return; }
  m.pts[0] = points[best_index];
  m.point_count = 1;
  points[best_index] = points[count - 1];
  count -= 1;
  Vec3 a = m.pts[0].point;

  best_score = 0.0f;
  best_index = kNullIndex;
  for (int i = 0; i < count; ++i) {
    Vec3 p = points[i].point;
    Vec3 d = sub(p, a);
    Vec3 v = mul_sub(d, dot(d, normal), normal);
    float dsq = length_sq(v);
    float separation = maxf(0.0f, -points[i].separation);
    float score = dsq + 4.0f * separation * separation;
    if (bias * score > best_score) { best_score = score; best_index = i; }
  }
  if (best_score < tol_sqr) return;
  m.pts[1] = points[best_index];
  m.point_count = 2;
  points[best_index] = points[count - 1];
  count -= 1;
  Vec3 b = m.pts[1].point;

  best_score = tol_sqr;
  best_index = kNullIndex;
  float best_signed_area = 0.0f;
  Vec3 ba = sub(b, a);
  for (int i = 0; i < count; ++i) {
    Vec3 p = points[i].point;
    float signed_area = dot(normal, cross(ba, sub(p, a)));
    float score = absf(signed_area);
    if (bias * score >= best_score) { best_score = score; best_index = i; best_signed_area = signed_area; }
  }
  if (best_index == kNullIndex) return;
  m.pts[2] = points[best_index];
  m.point_count = 3;
  points[best_index] = points[count - 1];
  count -= 1;
  Vec3 c = m.pts[2].point;

  best_score = tol_sqr;
  best_index = kNullIndex;
  float sign = (best_signed_area < 0.0f) != PLANTED(9) ? -1.0f : 1.0f;
  for (int i = 0; i < count; ++i) {
    Vec3 p = points[i].point;
    float u1 = sign * dot(normal, cross(sub(p, a), ba));
    float u2 = sign * dot(normal, cross(sub(p, b), sub(c, b)));
    float u3 = sign * dot(normal, cross(sub(p, c), sub(a, c)));
    float score = maxf(u1, maxf(u2, u3));
    if (bias * score > best_score) { best_score = score; best_index = i; }
  }
  if (best_index != kNullIndex) {
    m.pts[m.point_count] = points[best_index];
    m.point_count += 1;
  }
}

// manifold.rs:117-170
int find_incident_face(const Hull& h, Vec3 ref_normal, int vertex_index) {
  int min_edge = -1;
  float min_proj = FLT_MAX;
  int edge_index = h.vertex_edge[vertex_index];
  const int start = edge_index;
  Vec3 origin = h.points[h.edges[edge_index].origin];
  for (;;) {
    HalfEdge e = h.edges[edge_index];
    HalfEdge tw = h.edges[e.twin];
    Vec3 axis = normalize(sub(h.points[tw.origin], origin));
    float proj = absf(dot(axis, ref_normal));
    if (proj < min_proj) { min_edge = edge_index; min_proj = proj; }
    edge_index = tw.next;
    if (edge_index == start) break;
  }
  HalfEdge me = h.edges[min_edge];
  int f1 = me.face;
  int f2 = h.edges[me.twin].face;
  if (dot(h.planes[f1].n, ref_normal) < dot(h.planes[f2].n, ref_normal)) return f1;
  return f2;
}

// manifold.rs:173-231
int clip_polygon(ClipVertex* out, const ClipVertex* poly, int count, Plane clip, int edge, Plane ref_plane) {
  ClipVertex v1 = poly[count - 1];
  float d1 = plane_sep(clip, v1.position);
  int n = 0;
  for (int i = 0; i < count; ++i) {
    ClipVertex v2 = poly[i];
    float d2 = plane_sep(clip, v2.position);
    if (d1 <= 0.0f && d2 <= 0.0f) {
      out[n++] = v2;
    } else if (d1 <= 0.0f && d2 > 0.0f) {
      float f = d1 / (d1 - d2);
      Vec3 p = mul_add(v1.position, f, sub(v2.position, v1.position));
      ClipVertex v{p, plane_sep(ref_plane, p), v2.pair};
      v.pair.owner2 = 0;
      v.pair.index2 = (uint8_t)edge;
      out[n++] = v;
    } else if (d2 <= 0.0f && d1 > 0.0f) {
      float f = d1 / (d1 - d2);
      Vec3 p = mul_add(v1.position, f, sub(v2.position, v1.position));
      ClipVertex v{p, plane_sep(ref_plane, p), v1.pair};
      v.pair.owner1 = 0;
      v.pair.index1 = (uint8_t)edge;
      out[n++] = v;
      out[n++] = v2;
    }
    v1 = v2;
    d1 = d2;
  }
  return n;
}

// convex_manifold.rs:1071-1119
int build_polygon(ClipVertex* out, Xf xf, const Hull& h, int inc_face, Plane ref_plane) {
  int edge_index = h.face_edge[inc_face];
  int n = 0;
  Mat3 m = matrix_from_quat(xf.q);
  for (;;) {
    HalfEdge e = h.edges[edge_index];
    int next_index = e.next;
    HalfEdge nx = h.edges[next_index];
    ClipVertex v;
    v.position = add(mul_mv(m, h.points[nx.origin]), xf.p);
    v.separation = plane_sep(ref_plane, v.position);
    v.pair = make_pair(1, edge_index, 1, next_index);
    out[n++] = v;
    edge_index = next_index;
    if (edge_index == h.face_edge[inc_face] || n >= kMaxClipPoints) break;
  }
  return n;
}

float collect_face_points(const ClipVertex* in, LocalPoint* pts, int count, Vec3 normal) {
  float min_sep = FLT_MAX;
  for (int i = 0; i < count; ++i) {
    ClipVertex cp = in[i];
    pts[i] = LocalPoint{mul_sub(cp.position, 0.5f * cp.separation, normal), cp.separation, cp.pair, 0};
    min_sep = minf(min_sep, cp.separation);
  }
  return min_sep;
}

// convex_manifold.rs:1150-1271
bool build_face_a_contact(LocalManifold& m, const Hull& ha, const Hull& hb, Xf b_to_a, SepAxis q, SATCache& cache) {
  ClipVertex poly1[kMaxClipPoints + 8], poly2[kMaxClipPoints + 8];
  LocalPoint pts[kMaxClipPoints + 8];
  int ref_face = q.index_a;
  Plane ref_plane = ha.planes[ref_face];
  Vec3 ref_normal_in_b = inv_rotate(b_to_a.q, ref_plane.n);
  int inc_face = find_incident_face(hb, ref_normal_in_b, q.index_b);
  int count = build_polygon(poly1, b_to_a, hb, inc_face, ref_plane);
  int edge_index = ha.face_edge[ref_face];
  bool use1 = true;
  for (;;) {
    HalfEdge e = ha.edges[edge_index];
    int next_index = e.next;
    HalfEdge nx = ha.edges[next_index];
    Vec3 p1 = ha.points[e.origin], p2 = ha.points[nx.origin];
    Vec3 tangent = normalize(sub(p2, p1));
    Vec3 binormal = cross(tangent, ref_plane.n);
    Plane clip = plane_from_normal_point(binormal, p1);
    if (use1) count = clip_polygon(poly2, poly1, count, clip, edge_index, ref_plane);
    else count = clip_polygon(poly1, poly2, count, clip, edge_index, ref_plane);
    use1 = !use1;
    if (count < 3) { cache = SATCache{}; return false; }
    edge_index = next_index;
    if (edge_index == ha.face_edge[ref_face]) break;
  }
  count = mini(count, kMaxClipPoints);
  m.normal = ref_plane.n;
  float min_sep = collect_face_points(use1 ? poly1 : poly2, pts, count, ref_plane.n);
  if (min_sep >= speculative_distance()) { cache = SATCache{}; return false; }
  reduce_manifold_points(m, pts, count);
  cache.separation = min_sep;
  cache.type = kFaceAxisA;
  cache.index_a = (uint8_t)q.index_a;
  cache.index_b = (uint8_t)q.index_b;
  return true;
}

// convex_manifold.rs:1276-1336
bool build_face_b_contact(LocalManifold& m, const Hull& ha, const Hull& hb, Xf b_to_a, SepAxis q, SATCache& cache) {
  Xf a_to_b = invert_xf(b_to_a);
  SepAxis fq{neg(q.normal), q.separation, q.index_b, q.index_a, kFaceAxisA};
  if (!build_face_a_contact(m, hb, ha, a_to_b, fq, cache)) { cache = SATCache{}; return false; }
  Mat3 mat = matrix_from_quat(b_to_a.q);
  m.normal = neg(mul_mv(mat, m.normal));
  for (int i = 0; i < m.point_count; ++i) {
    m.pts[i].point = add(mul_mv(mat, m.pts[i].point), b_to_a.p);
    m.pts[i].pair = flip_pair(m.pts[i].pair);
  }
  cache.type = kFaceAxisB;
  cache.index_a = (uint8_t)q.index_a;
  cache.index_b = (uint8_t)q.index_b;
  return true;
}

inline bool within_segments(const SegDist& r) {
  return (0.0f <= r.f1 && r.f1 <= 1.0f) && (0.0f <= r.f2 && r.f2 <= 1.0f);
}

// convex_manifold.rs:1340-1407
bool build_edge_contact(LocalManifold& m, const Hull& ha, const Hull& hb, Xf b_to_a, SepAxis q, SATCache& cache) {
  HalfEdge ea = ha.edges[q.index_a], ta = ha.edges[ea.twin];
  Vec3 p_a = ha.points[ea.origin], q_a = ha.points[ta.origin];
  Vec3 e_a = sub(q_a, p_a);
  HalfEdge eb = hb.edges[q.index_b], tb = hb.edges[eb.twin];
  Vec3 p_b = xf_point(b_to_a, hb.points[eb.origin]);
  Vec3 q_b = xf_point(b_to_a, hb.points[tb.origin]);
  Vec3 e_b = sub(q_b, p_b);
  SegDist r = line_distance(p_a, e_a, p_b, e_b);
  if (!within_segments(r)) { cache = SATCache{}; return false; }
  float separation = dot(q.normal, sub(r.p2, r.p1));
  Vec3 point = PLANTED(7) ? r.p1 : scale(0.5f, add(r.p1, r.p2));
  m.normal = q.normal;
  m.point_count = 1;
  m.pts[0].point = point;
  m.pts[0].separation = separation;
  m.pts[0].pair = make_pair(0, q.index_a, 1, q.index_b);
  cache.separation = separation;
  cache.type = kEdgePairAxis;
  cache.index_a = (uint8_t)q.index_a;
  cache.index_b = (uint8_t)q.index_b;
  return true;
}

int find_support_vertex(const Hull& h, Vec3 d) {
  int best = -1;
  float best_dot = -FLT_MAX;
  for (int i = 0; i < h.vertex_count; ++i) {
    float v = d.x * h.vx[i] + d.y * h.vy[i] + d.z * h.vz[i];
    if (v > best_dot) { best_dot = v; best = i; }
  }
  return best;
}

}  // namespace

AxisQuery separating_axis_query(const Hull& ha, const Hull& hb, Xf xf_b, bool early_return) {
  return compute_separating_axis(ha, hb, xf_b, early_return);
}

uint32_t make_feature_id(FeaturePair p) {
  return ((uint32_t)p.owner1 << 24) | ((uint32_t)p.index1 << 16) | ((uint32_t)p.owner2 << 8) | (uint32_t)p.index2;
}

// convex_manifold.rs:1445-1845
void collide_hulls(LocalManifold& m, const Hull& ha, const Hull& hb, Xf b_to_a, SATCache& cache, int64_t* sat_calls, int64_t* sat_hits) {
  m.point_count = 0;
  const float spec = speculative_distance();
  const float ls = linear_slop();
  cache.hit = 0;
  if (sat_calls) __atomic_fetch_add(sat_calls, 1, __ATOMIC_RELAXED);

  switch (cache.type) {
    case kInvalidAxis:
      break;
    case kFaceAxisA: {
      Plane pl = ha.planes[cache.index_a];
      Vec3 dir = neg(inv_rotate(b_to_a.q, pl.n));
      int vi = find_support_vertex(hb, dir);
      Vec3 support = xf_point(b_to_a, hb.points[vi]);
      float separation = plane_sep(pl, support);
      if (separation >= spec) { cache.hit = 1; if (sat_hits) __atomic_fetch_add(sat_hits, 1, __ATOMIC_RELAXED); return; }
      SepAxis fq{pl.n, 0.0f, cache.index_a, vi, kFaceAxisA};
      SATCache local{};
      bool touching = build_face_a_contact(m, ha, hb, b_to_a, fq, local);
      if (touching && absf(cache.separation - local.separation) < ls) { cache.hit = 1; if (sat_hits) __atomic_fetch_add(sat_hits, 1, __ATOMIC_RELAXED); return; }
      break;
    }
    case kFaceAxisB: {
      Plane pl = hb.planes[cache.index_b];
      Vec3 dir = neg(rotate(b_to_a.q, pl.n));
      int vi = find_support_vertex(ha, dir);
      Vec3 support = inv_xf_point(b_to_a, ha.points[vi]);
      float separation = plane_sep(pl, support);
      if (separation >= spec) { cache.hit = 1; if (sat_hits) __atomic_fetch_add(sat_hits, 1, __ATOMIC_RELAXED); return; }
      SepAxis fq{neg(pl.n), 0.0f, vi, cache.index_b, kFaceAxisB};
      SATCache local{};
      bool touching = build_face_b_contact(m, ha, hb, b_to_a, fq, local);
      if (touching && absf(cache.separation - local.separation) < ls) { cache.hit = 1; if (sat_hits) __atomic_fetch_add(sat_hits, 1, __ATOMIC_RELAXED); return; }
      break;
    }
    case kEdgePairAxis: {
      int ia = cache.index_a;
      HalfEdge e1 = ha.edges[ia], t1 = ha.edges[ia + 1];
      Vec3 p_a = ha.points[e1.origin], q_a = ha.points[t1.origin];
      Vec3 e_a = sub(q_a, p_a);
      Vec3 u_a = ha.planes[e1.face].n, v_a = ha.planes[t1.face].n;
      int ib = cache.index_b;
      HalfEdge e2 = hb.edges[ib], t2 = hb.edges[ib + 1];
      Vec3 p_b = xf_point(b_to_a, hb.points[e2.origin]);
      Vec3 q_b = xf_point(b_to_a, hb.points[t2.origin]);
      Vec3 e_b = sub(q_b, p_b);
      Vec3 u_b = rotate(b_to_a.q, hb.planes[e2.face].n);
      Vec3 v_b = rotate(b_to_a.q, hb.planes[t2.face].n);
      float cba = dot(u_b, e_a);
      float dba = dot(v_b, e_a);
      float adc = -dot(u_a, e_b);
      float bdc = -dot(v_a, e_b);
      if (cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f) {
        const float sq_tol = kParallelEdgeTolerance * kParallelEdgeTolerance;
        if (maxf(cba * cba, dba * dba) >= sq_tol * length_sq(e_a)) {
          float t = cba / (cba - dba);
          Vec3 axis = normalize(lerp(u_b, v_b, t));
          float separation = dot(axis, sub(q_a, q_b));
          if (separation > spec) { cache.hit = 1; if (sat_hits) __atomic_fetch_add(sat_hits, 1, __ATOMIC_RELAXED); return; }
          SepAxis eq{neg(axis), 0.0f, cache.index_a, cache.index_b, kEdgePairAxis};
          SATCache local{};
          bool touching = build_edge_contact(m, ha, hb, b_to_a, eq, local);
          if (touching && (PLANTED(10) ? absf(cache.separation - local.separation) > ls : absf(cache.separation - local.separation) < ls)) {
            cache.hit = 1;
            if (sat_hits) __atomic_fetch_add(sat_hits, 1, __ATOMIC_RELAXED);
            return;
          }
        }
      }
      break;
    }
    default:
      break;
  }

  m.point_count = 0;
  cache = SATCache{};
  AxisQuery q = compute_separating_axis(ha, hb, b_to_a, true);
  if (q.separated != kInvalidAxis) {
    cache.type = q.separated;
    const SepAxis& s = q.separated == kFaceAxisA ? q.face_a : (q.separated == kFaceAxisB ? q.face_b : q.edge);
    cache.separation = s.separation;
    cache.index_a = (uint8_t)s.index_a;
    cache.index_b = (uint8_t)s.index_b;
    return;
  }
  if (q.face_a.separation > q.face_b.separation) {
    build_face_a_contact(m, ha, hb, b_to_a, q.face_a, cache);
  } else {
    build_face_b_contact(m, ha, hb, b_to_a, q.face_b, cache);
  }
  SepAxis eq = q.edge;
  if (eq.index_a == kNullIndex) return;
  float face_sep = maxf(q.face_a.separation, q.face_b.separation);
  float clip_sep = cache.separation;
  if ((m.point_count == 0 && eq.separation > face_sep) || eq.separation > clip_sep + ls) {
    LocalManifold em{};
    em.point_count = 0;
    SATCache ec{};
    build_edge_contact(em, ha, hb, b_to_a, eq, ec);
    if (em.point_count == 1) {
      m.normal = em.normal;
      m.point_count = 1;
      m.pts[0] = em.pts[0];
      cache = ec;
    }
  }
}

}  // namespace nslref
