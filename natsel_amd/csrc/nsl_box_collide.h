// Box-vs-box narrow phase for gfx950: separating-axis test with the cached axis,
// incident-face clipping, edge contacts and 4-point manifold reduction. One lane
// handles one contact; clip polygons live in registers/private memory.
//
// Follows src/collision/convex_manifold.rs:137-1845 and src/collision/manifold.rs
// operation for operation so feature ids and points match the CPU step bit for bit.
// Boxes are the reference's half-edge box hulls (hull.rs:736-852): 8 vertices,
// 24 half-edges, 6 faces, zero edge cosines.
#pragma once
#include "nsl_types.h"

namespace nsl {

// hull.rs:736-750, (next, twin, origin, face) per half-edge. Feature ids index into
// this table, so it is the reference's box topology.
static __constant__ uint8_t c_box_edge[24][4] = {
    {2, 1, 2, 0},   {17, 0, 1, 5},  {4, 3, 1, 0},   {20, 2, 5, 3},  {6, 5, 5, 0},   {23, 4, 6, 4},
    {0, 7, 6, 0},   {18, 6, 2, 2},  {10, 9, 0, 1},  {21, 8, 3, 5},  {12, 11, 3, 1}, {16, 10, 7, 2},
    {14, 13, 7, 1}, {19, 12, 4, 4}, {8, 15, 4, 1},  {22, 14, 0, 3}, {7, 17, 3, 2},  {9, 16, 2, 5},
    {11, 19, 6, 2}, {5, 18, 7, 4},  {15, 21, 1, 3}, {1, 20, 0, 5},  {3, 23, 4, 3},  {13, 22, 5, 4},
};
static __constant__ uint8_t c_box_vertex_edge[8] = {8, 1, 0, 9, 13, 3, 5, 11};
static __constant__ uint8_t c_box_face_edge[6] = {0, 8, 16, 20, 19, 21};

#define E_NEXT(i) c_box_edge[i][0]
#define E_TWIN(i) c_box_edge[i][1]
#define E_ORIGIN(i) c_box_edge[i][2]
#define E_FACE(i) c_box_edge[i][3]

// Sutherland-Hodgman on a box face: a quad clipped by four half-planes has at most
// 8 vertices (each clip adds at most one), so 16 slots never overflow.
constexpr int kClipSlots = 16;

struct FeaturePair { uint8_t owner1, index1, owner2, index2; };
struct ClipVertex { Vec3 position; float separation; FeaturePair pair; };
struct LocalPoint { Vec3 point; float separation; FeaturePair pair; };
struct LocalManifold { Vec3 normal; LocalPoint pts[4]; int32_t point_count; };
struct SepAxis { Vec3 normal; float separation; int32_t index_a, index_b; };
struct AxisQuery { SepAxis face_a, face_b, edge; uint8_t separated; };

enum : uint8_t { kAxisInvalid = 0, kAxisFaceA = 2, kAxisFaceB = 3, kAxisEdge = 4 };

__device__ __forceinline__ uint32_t feature_id(FeaturePair p) {
  return ((uint32_t)p.owner1 << 24) | ((uint32_t)p.index1 << 16) | ((uint32_t)p.owner2 << 8) | (uint32_t)p.index2;
}
__device__ __forceinline__ FeaturePair make_pair(int o1, int i1, int o2, int i2) {
  return FeaturePair{(uint8_t)o1, (uint8_t)i1, (uint8_t)o2, (uint8_t)i2};
}
__device__ __forceinline__ FeaturePair flip_pair(FeaturePair p) {
  return FeaturePair{(uint8_t)(1 - p.owner2), p.index2, (uint8_t)(1 - p.owner1), p.index1};
}

__device__ __forceinline__ bool edge_candidate(float a1, float a2, float c, float bound) {
  float hi = maxf(a1, a2), lo = minf(a1, a2);
  bool exterior = hi >= bound;
  float u = lo - c * hi;
  bool max_is_interior = u >= 0.0f;
  float s = 1.0f - c * c;
  float bound_term = (bound - hi) * (bound + hi);
  bool max_beats_bound = u * u >= bound_term * s;
  bool nearly_parallel = s < 1e-4f;
  return exterior || (max_is_interior && (max_beats_bound || nearly_parallel));
}

__device__ __forceinline__ float face_dot(const BoxHull& h, int i, Vec3 d) {
  Vec3 n = h.planes[i].n;
  return d.x * n.x + d.y * n.y + d.z * n.z;
}

// Support with the vertex index carried in the low 7 mantissa bits
// (convex_manifold.rs:293-322).
__device__ __forceinline__ float support_wide(Vec3 n, const BoxHull& h, float bias, int* out_index) {
  float min_value = __builtin_inff();
  for (int i = 0; i < 8; ++i) {
    Vec3 p = h.points[i];
    float d = n.x * p.x + n.y * p.y + n.z * p.z;
    float value = bias - d;
    float aug = float_of((bits_of(value) & ~0x7fu) | (uint32_t)i);
    if (aug < min_value) min_value = aug;
  }
  int vi = (int)(bits_of(min_value) & 0x7fu);
  *out_index = vi;
  Vec3 p = h.points[vi];
  return n.x * p.x + n.y * p.y + n.z * p.z;
}

__device__ __forceinline__ float face_separation(Vec3 dir, float plane_separation, const BoxHull& h, Vec3 c, Vec3 e, int* vi) {
  float bias = dot(dir, c) + 1.0625f * dot(abs3(dir), e);
  return plane_separation - support_wide(dir, h, bias, vi);
}

__device__ __forceinline__ Vec3 neg_xf(Mat3 r, Vec3 p, Vec3 v, bool is_point) {
  Vec3 r0 = {r.cx.x, r.cy.x, r.cz.x}, r1 = {r.cx.y, r.cy.y, r.cz.y}, r2 = {r.cx.z, r.cy.z, r.cz.z};
  Vec3 out = {dot(r0, v), dot(r1, v), dot(r2, v)};
  if (is_point) out = add(out, p);
  return neg(out);
}

// convex_manifold.rs:392-899 specialised to two boxes.
__device__ AxisQuery separating_axis(const BoxHull& ha, const BoxHull& hb, Xf xf_b) {
  Mat3 r = matrix_from_quat(xf_b.q);
  Mat3 inv_r = transpose(r);
  const float spec = kSpeculativeDistance;
  AxisQuery res;
  res.face_a = SepAxis{zero3(), -__builtin_inff(), kNullIndex, kNullIndex};
  res.face_b = res.face_a;
  res.edge = res.face_a;
  res.separated = kAxisInvalid;

  Vec3 c_b = aabb_center(hb.aabb), h_b = aabb_extents(hb.aabb);
  Vec3 delta_center = sub(add(mul_mv(r, hb.center), xf_b.p), ha.center);
  float center_distance = length(delta_center);
  float radius = ha.inner_radius + hb.inner_radius;
  float radius_bound = radius - (kLinearSlop + 0.001f * (center_distance + radius));

  float dot_a[6];
  for (int i = 0; i < 6; ++i) dot_a[i] = face_dot(ha, i, delta_center);
  int seed_a = 0;
  float max_dot_a = dot_a[0];
  for (int i = 1; i < 6; ++i)
    if (dot_a[i] > max_dot_a) { max_dot_a = dot_a[i]; seed_a = i; }
  int seed_vertex_b = 0;
  float seed_sep_a;
  {
    Plane pl = ha.planes[seed_a];
    Vec3 dir = neg(mul_mv(inv_r, pl.n));
    seed_sep_a = face_separation(dir, dot(pl.n, xf_b.p) - pl.off, hb, c_b, h_b, &seed_vertex_b);
  }
  float floor_a = minf(seed_sep_a, spec);
  for (int i = 0; i < 6; ++i) {
    if (dot_a[i] - radius_bound < maxf(floor_a, res.face_a.separation)) continue;
    Plane pl = ha.planes[i];
    int vi = seed_vertex_b;
    float sep = seed_sep_a;
    if (i != seed_a) {
      Vec3 dir = neg(mul_mv(inv_r, pl.n));
      sep = face_separation(dir, dot(pl.n, xf_b.p) - pl.off, hb, c_b, h_b, &vi);
    }
    if (sep > res.face_a.separation) {
      res.face_a = SepAxis{pl.n, sep, i, vi};
      if (sep > spec) { res.separated = kAxisFaceA; return res; }
    }
  }

  Vec3 c_a = aabb_center(ha.aabb), h_a = aabb_extents(ha.aabb);
  Vec3 db = neg(mul_mv(inv_r, delta_center));
  float dot_b[6];
  for (int i = 0; i < 6; ++i) dot_b[i] = face_dot(hb, i, db);
  int seed_b = 0;
  float max_dot_b = dot_b[0];
  for (int i = 1; i < 6; ++i)
    if (dot_b[i] > max_dot_b) { max_dot_b = dot_b[i]; seed_b = i; }
  int seed_vertex_a = 0;
  float seed_sep_b;
  {
    Plane pl = hb.planes[seed_b];
    Vec3 dir = neg(mul_mv(r, pl.n));
    seed_sep_b = face_separation(dir, dot(dir, xf_b.p) - pl.off, ha, c_a, h_a, &seed_vertex_a);
  }
  float floor_b = minf(maxf(seed_sep_b, res.face_a.separation), spec);
  for (int i = 0; i < 6; ++i) {
    if (dot_b[i] - radius_bound < maxf(floor_b, res.face_b.separation)) continue;
    Plane pl = hb.planes[i];
    Vec3 dir = neg(mul_mv(r, pl.n));
    int vi = seed_vertex_a;
    float sep = seed_sep_b;
    if (i != seed_b) sep = face_separation(dir, dot(dir, xf_b.p) - pl.off, ha, c_a, h_a, &vi);
    if (sep > res.face_b.separation) {
      res.face_b = SepAxis{dir, sep, vi, i};
      if (sep > spec) { res.separated = kAxisFaceB; return res; }
    }
  }

  float max_face_sep = maxf(res.face_a.separation, res.face_b.separation);
  float bound_slack = radius - radius_bound;
  float threshold_a = max_face_sep + hb.inner_radius - bound_slack;
  float threshold_b = max_face_sep + ha.inner_radius - bound_slack;
  Vec3 center_b_in_a = add(mul_mv(r, hb.center), xf_b.p);
  Vec3 center_a_in_b = mul_mv(inv_r, sub(ha.center, xf_b.p));
  float pd_a[6], pd_b[6];
  for (int i = 0; i < 6; ++i) pd_a[i] = face_dot(ha, i, center_b_in_a) - ha.planes[i].off;
  for (int i = 0; i < 6; ++i) pd_b[i] = face_dot(hb, i, center_a_in_b) - hb.planes[i].off;
  int8_t idx_a[12], idx_b[12];
  int na = 0, nb = 0;
  for (int i = 0; i < 24; i += 2) {
    idx_a[na] = (int8_t)i;
    if (edge_candidate(pd_a[E_FACE(i)], pd_a[E_FACE(i + 1)], 0.0f, threshold_a)) na += 1;
  }
  for (int i = 0; i < 24; i += 2) {
    idx_b[nb] = (int8_t)i;
    if (edge_candidate(pd_b[E_FACE(i)], pd_b[E_FACE(i + 1)], 0.0f, threshold_b)) nb += 1;
  }
  if (nb * ((na + 3) >> 2) >= 10) {
    float probe_bound = max_face_sep - (0.1f * kLinearSlop + 0.001f * absf(center_distance + radius));
    int vb = res.face_a.index_b != kNullIndex ? res.face_a.index_b : seed_vertex_b;
    Vec3 probe_b = add(mul_mv(r, hb.points[vb]), xf_b.p);
    for (int i = 0; i < 6; ++i) pd_a[i] = face_dot(ha, i, probe_b) - ha.planes[i].off;
    int kept = 0;
    for (int k = 0; k < na; ++k) {
      int i = idx_a[k];
      idx_a[kept] = (int8_t)i;
      if (edge_candidate(pd_a[E_FACE(i)], pd_a[E_FACE(i + 1)], 0.0f, probe_bound)) kept += 1;
    }
    na = kept;
    int va = res.face_b.index_a != kNullIndex ? res.face_b.index_a : seed_vertex_a;
    Vec3 probe_a = mul_mv(inv_r, sub(ha.points[va], xf_b.p));
    for (int i = 0; i < 6; ++i) pd_b[i] = face_dot(hb, i, probe_a) - hb.planes[i].off;
    kept = 0;
    for (int k = 0; k < nb; ++k) {
      int i = idx_b[k];
      idx_b[kept] = (int8_t)i;
      if (edge_candidate(pd_b[E_FACE(i)], pd_b[E_FACE(i + 1)], 0.0f, probe_bound)) kept += 1;
    }
    nb = kept;
  }
  if (na == 0 || nb == 0) return res;

  const float sq_tol = kParallelEdgeTolerance * kParallelEdgeTolerance;
  const float eps = -kLinearSlop * kLinearSlop;
  for (int j = 0; j < nb; ++j) {
    int eb = idx_b[j];
    Vec3 c = neg_xf(r, xf_b.p, hb.planes[E_FACE(eb)].n, false);
    Vec3 dn = neg_xf(r, xf_b.p, hb.planes[E_FACE(eb + 1)].n, false);
    Vec3 bv0 = neg_xf(r, xf_b.p, hb.points[E_ORIGIN(eb)], true);
    Vec3 bv1 = neg_xf(r, xf_b.p, hb.points[E_ORIGIN(eb + 1)], true);
    Vec3 dc = sub(bv1, bv0);
    for (int i = 0; i < na; ++i) {
      int ea = idx_a[i];
      Vec3 av0 = ha.points[E_ORIGIN(ea)];
      Vec3 d = sub(ha.points[E_ORIGIN(ea + 1)], av0);
      float cba = dot(c, d), dba = dot(dn, d);
      if (cba * dba >= eps) continue;
      float adc = dot(ha.planes[E_FACE(ea)].n, dc), bdc = dot(ha.planes[E_FACE(ea + 1)].n, dc);
      if (adc * bdc >= eps || cba * bdc >= eps) continue;
      float tol = sq_tol * (d.x * d.x + d.y * d.y + d.z * d.z);
      if (maxf(cba * cba, dba * dba) <= tol) continue;
      float t = -cba / (dba - cba);
      Vec3 n = mul_add(c, t, sub(dn, c));
      n = scale(1.0f / sqrtf_(dot(n, n)), n);
      float sep = -dot(add(av0, bv0), n);
      if (sep > res.edge.separation) {
        res.edge = SepAxis{n, sep, ea, eb};
        if (sep > spec) { res.separated = kAxisEdge; return res; }
      }
    }
  }
  return res;
}

__device__ __forceinline__ LocalPoint local_point(const ClipVertex& v) { return LocalPoint{v.position, v.separation, v.pair}; }

// convex_manifold.rs:905-1062. The candidate points arrive in a clip buffer whose
// positions already lie midway between the two surfaces.
__device__ void reduce_points(LocalManifold& m, ClipVertex* pts, int count) {
  if (count <= 4) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < count) m.pts[i] = local_point(pts[i]);
    m.point_count = count;
    return;
  }
  Vec3 normal = m.normal;
  const float spec = kSpeculativeDistance, tol_sqr = spec * spec, bias = 0.95f;
  int best = kNullIndex;
  float best_score = -kFltMax;
  Vec3 search = arbitrary_perp(normal);
  for (int i = 0; i < count; ++i) {
    if (pts[i].separation > spec) continue;
    float score = -pts[i].separation + dot(search, pts[i].position);
    if (bias * score > best_score) { best = i; best_score = score; }
  }
  if (best == kNullIndex) { m.point_count = 0; return; }
  m.pts[0] = local_point(pts[best]);
  m.point_count = 1;
  pts[best] = pts[count - 1];
  count -= 1;
  Vec3 a = m.pts[0].point;
  best_score = 0.0f;
  best = kNullIndex;
  for (int i = 0; i < count; ++i) {
    Vec3 d = sub(pts[i].position, a);
    Vec3 v = mul_sub(d, dot(d, normal), normal);
    float s = maxf(0.0f, -pts[i].separation);
    float score = length_sq(v) + 4.0f * s * s;
    if (bias * score > best_score) { best_score = score; best = i; }
  }
  if (best_score < tol_sqr) return;
  m.pts[1] = local_point(pts[best]);
  m.point_count = 2;
  pts[best] = pts[count - 1];
  count -= 1;
  Vec3 b = m.pts[1].point;
  best_score = tol_sqr;
  best = kNullIndex;
  float best_area = 0.0f;
  Vec3 ba = sub(b, a);
  for (int i = 0; i < count; ++i) {
    float area = dot(normal, cross(ba, sub(pts[i].position, a)));
    float score = absf(area);
    if (bias * score >= best_score) { best_score = score; best = i; best_area = area; }
  }
  if (best == kNullIndex) return;
  m.pts[2] = local_point(pts[best]);
  m.point_count = 3;
  pts[best] = pts[count - 1];
  count -= 1;
  Vec3 c = m.pts[2].point;
  best_score = tol_sqr;
  best = kNullIndex;
  float sign = best_area < 0.0f ? -1.0f : 1.0f;
  for (int i = 0; i < count; ++i) {
    Vec3 p = pts[i].position;
    float u1 = sign * dot(normal, cross(sub(p, a), ba));
    float u2 = sign * dot(normal, cross(sub(p, b), sub(c, b)));
    float u3 = sign * dot(normal, cross(sub(p, c), sub(a, c)));
    float score = maxf(u1, maxf(u2, u3));
    if (bias * score > best_score) { best_score = score; best = i; }
  }
  if (best != kNullIndex) {
    m.pts[3] = local_point(pts[best]);
    m.point_count = 4;
  }
}

// manifold.rs:117-170
__device__ int incident_face(const BoxHull& h, Vec3 ref_normal, int vertex_index) {
  int min_edge = -1;
  float min_proj = kFltMax;
  int e = c_box_vertex_edge[vertex_index];
  const int start = e;
  Vec3 origin = h.points[E_ORIGIN(e)];
  for (;;) {
    int tw = E_TWIN(e);
    Vec3 axis = normalize(sub(h.points[E_ORIGIN(tw)], origin));
    float proj = absf(dot(axis, ref_normal));
    if (proj < min_proj) { min_edge = e; min_proj = proj; }
    e = E_NEXT(tw);
    if (e == start) break;
  }
  int f1 = E_FACE(min_edge), f2 = E_FACE(E_TWIN(min_edge));
  return dot(h.planes[f1].n, ref_normal) < dot(h.planes[f2].n, ref_normal) ? f1 : f2;
}

// manifold.rs:173-231
__device__ int clip_polygon(ClipVertex* out, const ClipVertex* poly, int count, Plane clip, int edge, Plane ref_plane) {
  ClipVertex v1 = poly[count - 1];
  float d1 = plane_sep(clip, v1.position);
  int n = 0;
  for (int i = 0; i < count; ++i) {
    ClipVertex v2 = poly[i];
    float d2 = plane_sep(clip, v2.position);
    if (d1 <= 0.0f && d2 <= 0.0f) {
      if (n < kClipSlots) out[n] = v2;
      n += 1;
    } else if (d1 <= 0.0f && d2 > 0.0f) {
      float f = d1 / (d1 - d2);
      Vec3 p = mul_add(v1.position, f, sub(v2.position, v1.position));
      ClipVertex v{p, plane_sep(ref_plane, p), v2.pair};
      v.pair.owner2 = 0;
      v.pair.index2 = (uint8_t)edge;
      if (n < kClipSlots) out[n] = v;
      n += 1;
    } else if (d2 <= 0.0f && d1 > 0.0f) {
      float f = d1 / (d1 - d2);
      Vec3 p = mul_add(v1.position, f, sub(v2.position, v1.position));
      ClipVertex v{p, plane_sep(ref_plane, p), v1.pair};
      v.pair.owner1 = 0;
      v.pair.index1 = (uint8_t)edge;
      if (n < kClipSlots) out[n] = v;
      n += 1;
      if (n < kClipSlots) out[n] = v2;
      n += 1;
    }
    v1 = v2;
    d1 = d2;
  }
  return n < kClipSlots ? n : kClipSlots;
}

// convex_manifold.rs:1150-1271
__device__ bool face_a_contact(LocalManifold& m, const BoxHull& ha, const BoxHull& hb, Xf b_to_a, int ref_face, int vertex_b,
                               SatCache& cache) {
  ClipVertex poly1[kClipSlots], poly2[kClipSlots];
  Plane ref_plane = ha.planes[ref_face];
  Vec3 ref_normal_in_b = inv_rotate(b_to_a.q, ref_plane.n);
  int inc = incident_face(hb, ref_normal_in_b, vertex_b);
  // build_polygon (convex_manifold.rs:1071-1119)
  Mat3 mb = matrix_from_quat(b_to_a.q);
  int count = 0;
  {
    int e = c_box_face_edge[inc];
    for (;;) {
      int nx = E_NEXT(e);
      ClipVertex v;
      v.position = add(mul_mv(mb, hb.points[E_ORIGIN(nx)]), b_to_a.p);
      v.separation = plane_sep(ref_plane, v.position);
      v.pair = make_pair(1, e, 1, nx);
      poly1[count++] = v;
      e = nx;
      if (e == c_box_face_edge[inc]) break;
    }
  }
  int e = c_box_face_edge[ref_face];
  bool use1 = true;
  for (;;) {
    int nx = E_NEXT(e);
    Vec3 p1 = ha.points[E_ORIGIN(e)], p2 = ha.points[E_ORIGIN(nx)];
    Vec3 binormal = cross(normalize(sub(p2, p1)), ref_plane.n);
    Plane clip = plane_from_normal_point(binormal, p1);
    count = use1 ? clip_polygon(poly2, poly1, count, clip, e, ref_plane) : clip_polygon(poly1, poly2, count, clip, e, ref_plane);
    use1 = !use1;
    if (count < 3) { cache = SatCache{}; return false; }
    e = nx;
    if (e == c_box_face_edge[ref_face]) break;
  }
  m.normal = ref_plane.n;
  ClipVertex* pts = use1 ? poly1 : poly2;  // becomes the candidate points in place
  float min_sep = kFltMax;
  for (int i = 0; i < count; ++i) {
    pts[i].position = mul_sub(pts[i].position, 0.5f * pts[i].separation, ref_plane.n);
    min_sep = minf(min_sep, pts[i].separation);
  }
  if (min_sep >= kSpeculativeDistance) { cache = SatCache{}; return false; }
  reduce_points(m, pts, count);
  cache.separation = min_sep;
  cache.type = kAxisFaceA;
  cache.index_a = (uint8_t)ref_face;
  cache.index_b = (uint8_t)vertex_b;
  return true;
}

// The reference face on A (convex_manifold.rs:1150-1271) or, with on_b, on B
// (convex_manifold.rs:1276-1336: A and B swapped, the result mapped back into A's
// frame). ref_face belongs to the reference hull, vertex to the other one.
__device__ __forceinline__ bool face_contact(LocalManifold& m, const BoxHull& ha, const BoxHull& hb, Xf b_to_a, bool on_b,
                                             int ref_face, int vertex, SatCache& cache) {
  const BoxHull& ref = on_b ? hb : ha;
  const BoxHull& inc = on_b ? ha : hb;
  if (!face_a_contact(m, ref, inc, on_b ? invert_xf(b_to_a) : b_to_a, ref_face, vertex, cache)) return false;
  if (!on_b) return true;
  Mat3 mat = matrix_from_quat(b_to_a.q);
  m.normal = neg(mul_mv(mat, m.normal));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= m.point_count) break;
    m.pts[i].point = add(mul_mv(mat, m.pts[i].point), b_to_a.p);
    m.pts[i].pair = flip_pair(m.pts[i].pair);
  }
  cache.type = kAxisFaceB;
  cache.index_a = (uint8_t)vertex;
  cache.index_b = (uint8_t)ref_face;
  return true;
}

// convex_manifold.rs:1340-1407
__device__ bool edge_contact(LocalManifold& m, const BoxHull& ha, const BoxHull& hb, Xf b_to_a, Vec3 normal, int ia, int ib,
                             SatCache& cache) {
  Vec3 p_a = ha.points[E_ORIGIN(ia)], q_a = ha.points[E_ORIGIN(E_TWIN(ia))];
  Vec3 p_b = xf_point(b_to_a, hb.points[E_ORIGIN(ib)]), q_b = xf_point(b_to_a, hb.points[E_ORIGIN(E_TWIN(ib))]);
  SegDist r = line_distance(p_a, sub(q_a, p_a), p_b, sub(q_b, p_b));
  if (!((0.0f <= r.f1 && r.f1 <= 1.0f) && (0.0f <= r.f2 && r.f2 <= 1.0f))) { cache = SatCache{}; return false; }
  float sep = dot(normal, sub(r.p2, r.p1));
  m.normal = normal;
  m.point_count = 1;
  m.pts[0] = LocalPoint{scale(0.5f, add(r.p1, r.p2)), sep, make_pair(0, ia, 1, ib)};
  cache.separation = sep;
  cache.type = kAxisEdge;
  cache.index_a = (uint8_t)ia;
  cache.index_b = (uint8_t)ib;
  return true;
}

__device__ __forceinline__ int support_vertex(const BoxHull& h, Vec3 d) {
  int best = -1;
  float best_dot = -kFltMax;
  for (int i = 0; i < 8; ++i) {
    Vec3 p = h.points[i];
    float v = d.x * p.x + d.y * p.y + d.z * p.z;
    if (v > best_dot) { best_dot = v; best = i; }
  }
  return best;
}

// convex_manifold.rs:1445-1845 (collide_hulls_with_scratch). The face clip runs at
// one site: first (when the cache names a face axis) to test the cached axis, then,
// if that axis no longer holds, for the full query's best face. One site keeps a
// single copy of the clip buffers in the compiled kernel.
__device__ void collide_boxes(LocalManifold& m, const BoxHull& ha, const BoxHull& hb, Xf b_to_a, SatCache& cache) {
  m.point_count = 0;
  const float spec = kSpeculativeDistance;
  cache.hit = 0;
  bool cached_face = false, on_b = false;
  int ref_face = 0, vertex = 0;
  if (cache.type == kAxisFaceA) {
    Plane pl = ha.planes[cache.index_a];
    int vi = support_vertex(hb, neg(inv_rotate(b_to_a.q, pl.n)));
    float sep = plane_sep(pl, xf_point(b_to_a, hb.points[vi]));
    if (sep >= spec) { cache.hit = 1; return; }
    cached_face = true;
    ref_face = cache.index_a;
    vertex = vi;
  } else if (cache.type == kAxisFaceB) {
    Plane pl = hb.planes[cache.index_b];
    int vi = support_vertex(ha, neg(rotate(b_to_a.q, pl.n)));
    float sep = plane_sep(pl, inv_xf_point(b_to_a, ha.points[vi]));
    if (sep >= spec) { cache.hit = 1; return; }
    cached_face = true;
    on_b = true;
    ref_face = cache.index_b;
    vertex = vi;
  } else if (cache.type == kAxisEdge) {
    int ia = cache.index_a, ib = cache.index_b;
    Vec3 p_a = ha.points[E_ORIGIN(ia)], q_a = ha.points[E_ORIGIN(ia + 1)];
    Vec3 e_a = sub(q_a, p_a);
    Vec3 u_a = ha.planes[E_FACE(ia)].n, v_a = ha.planes[E_FACE(ia + 1)].n;
    Vec3 p_b = xf_point(b_to_a, hb.points[E_ORIGIN(ib)]), q_b = xf_point(b_to_a, hb.points[E_ORIGIN(ib + 1)]);
    Vec3 e_b = sub(q_b, p_b);
    Vec3 u_b = rotate(b_to_a.q, hb.planes[E_FACE(ib)].n), v_b = rotate(b_to_a.q, hb.planes[E_FACE(ib + 1)].n);
    float cba = dot(u_b, e_a), dba = dot(v_b, e_a);
    float adc = -dot(u_a, e_b), bdc = -dot(v_a, e_b);
    if (cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f) {
      const float sq_tol = kParallelEdgeTolerance * kParallelEdgeTolerance;
      if (maxf(cba * cba, dba * dba) >= sq_tol * length_sq(e_a)) {
        float t = cba / (cba - dba);
        Vec3 axis = normalize(lerp(u_b, v_b, t));
        float sep = dot(axis, sub(q_a, q_b));
        if (sep > spec) { cache.hit = 1; return; }
        SatCache local{};
        bool touching = edge_contact(m, ha, hb, b_to_a, neg(axis), ia, ib, local);
        if (touching && absf(cache.separation - local.separation) < kLinearSlop) { cache.hit = 1; return; }
      }
    }
  }

  AxisQuery q;
  for (;;) {
    if (!cached_face) {
      m.point_count = 0;
      cache = SatCache{};
      q = separating_axis(ha, hb, b_to_a);
      if (q.separated != kAxisInvalid) {
        SepAxis s = q.edge;
        if (q.separated == kAxisFaceA) s = q.face_a;
        else if (q.separated == kAxisFaceB) s = q.face_b;
        cache.type = q.separated;
        cache.separation = s.separation;
        cache.index_a = (uint8_t)s.index_a;
        cache.index_b = (uint8_t)s.index_b;
        return;
      }
      on_b = !(q.face_a.separation > q.face_b.separation);
      ref_face = on_b ? q.face_b.index_b : q.face_a.index_a;
      vertex = on_b ? q.face_b.index_a : q.face_a.index_b;
    }
    SatCache local{};
    bool touching = face_contact(m, ha, hb, b_to_a, on_b, ref_face, vertex, local);
    if (cached_face) {
      if (touching && absf(cache.separation - local.separation) < kLinearSlop) { cache.hit = 1; return; }
      cached_face = false;  // the cached axis no longer holds: run the full query
      continue;
    }
    cache = local;
    break;
  }
  if (q.edge.index_a == kNullIndex) return;
  float face_sep = maxf(q.face_a.separation, q.face_b.separation);
  if ((m.point_count == 0 && q.edge.separation > face_sep) || q.edge.separation > cache.separation + kLinearSlop) {
    LocalManifold em;
    em.point_count = 0;
    SatCache ec{};
    edge_contact(em, ha, hb, b_to_a, q.edge.normal, q.edge.index_a, q.edge.index_b, ec);
    if (em.point_count == 1) {
      m.normal = em.normal;
      m.point_count = 1;
      m.pts[0] = em.pts[0];
      cache = ec;
    }
  }
}

#undef E_NEXT
#undef E_TWIN
#undef E_ORIGIN
#undef E_FACE

}  // namespace nsl
