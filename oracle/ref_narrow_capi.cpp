// C entry points of the oracle's narrow phase for the known-answer tests.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// Each call runs one routine of the restated narrow phase on plain arrays, so the
// reference's own collision tests (distance_tests.rs, shape_collision_tests.rs,
// convex_manifold_tests.rs, collision/manifold_robustness_tests.rs,
// collision/separating_axis_oracle_tests.rs) can be replayed from Python.
//
// Shape arrays:  box hull  [hx, hy, hz, px, py, pz, qx, qy, qz, qs]
//                          (make_transformed_box_hull, hull.rs:754-852, the hull placed by p, q)
//                sphere    [cx, cy, cz, r]
//                capsule   [c1x, c1y, c1z, c2x, c2y, c2z, r]
// Transform:     [px, py, pz, qx, qy, qz, qs] (B in A's frame)
// Caches:        cache_f[0] = SATCache.separation or SimplexCache.metric;
//                SAT cache_i = {type, index_a, index_b, hit};
//                simplex cache_i = {count, index_a[4], index_b[4]}
// Manifold out:  out_f[19] = normal xyz, then per point x, y, z, separation;
//                out_i[21] = point count, per point (owner1, index1, owner2, index2),
//                then per point the feature id (manifold.rs:58-94).
#include "ref_world.h"

namespace nslref {
namespace {

enum { kKindBox = 0, kKindSphere = 1, kKindCapsule = 2 };

Xf xf_of(const float* x) { return Xf{Vec3{x[0], x[1], x[2]}, Quat{Vec3{x[3], x[4], x[5]}, x[6]}}; }
Hull hull_of(const float* a) { return make_transformed_box_hull(a[0], a[1], a[2], xf_of(a + 3)); }
Sphere sphere_of(const float* a) { return Sphere{Vec3{a[0], a[1], a[2]}, a[3]}; }
Capsule capsule_of(const float* a) { return Capsule{Vec3{a[0], a[1], a[2]}, Vec3{a[3], a[4], a[5]}, a[6]}; }

SATCache sat_in(const float* f, const int32_t* i) {
  return SATCache{f[0], (uint8_t)i[0], (uint8_t)i[1], (uint8_t)i[2], (uint8_t)i[3]};
}
void sat_out(const SATCache& c, float* f, int32_t* i) {
  f[0] = c.separation;
  i[0] = c.type; i[1] = c.index_a; i[2] = c.index_b; i[3] = c.hit;
}
SimplexCache simplex_in(const float* f, const int32_t* i) {
  SimplexCache c{};
  c.metric = f[0];
  c.count = (uint16_t)i[0];
  for (int k = 0; k < 4; ++k) { c.index_a[k] = (uint8_t)i[1 + k]; c.index_b[k] = (uint8_t)i[5 + k]; }
  return c;
}
void simplex_out(const SimplexCache& c, float* f, int32_t* i) {
  f[0] = c.metric;
  i[0] = c.count;
  for (int k = 0; k < 4; ++k) { i[1 + k] = c.index_a[k]; i[5 + k] = c.index_b[k]; }
}
void manifold_out(const LocalManifold& m, float* f, int32_t* i) {
  f[0] = m.normal.x; f[1] = m.normal.y; f[2] = m.normal.z;
  i[0] = m.point_count;
  for (int k = 0; k < 4; ++k) {
    const LocalPoint& p = m.pts[k];
    const bool live = k < m.point_count;
    f[3 + 4 * k] = live ? p.point.x : 0.0f;
    f[4 + 4 * k] = live ? p.point.y : 0.0f;
    f[5 + 4 * k] = live ? p.point.z : 0.0f;
    f[6 + 4 * k] = live ? p.separation : 0.0f;
    i[1 + 4 * k] = live ? p.pair.owner1 : 0;
    i[2 + 4 * k] = live ? p.pair.index1 : 0;
    i[3 + 4 * k] = live ? p.pair.owner2 : 0;
    i[4 + 4 * k] = live ? p.pair.index2 : 0;
    i[17 + k] = live ? (int32_t)make_feature_id(p.pair) : 0;
  }
}

}  // namespace
}  // namespace nslref

using namespace nslref;

extern "C" {

// One collide routine of the register table (contact_creation.rs:88-102): A is the
// primary shape. Pairs: box-box collide_hulls, box-sphere collide_hull_and_sphere,
// box-capsule collide_hull_and_capsule, capsule-sphere collide_capsule_and_sphere,
// capsule-capsule collide_capsules, sphere-sphere collide_spheres. Returns -1 for any
// other order.
int nslo_collide_pair(int kind_a, const float* a, int kind_b, const float* b, const float* xf, float* cache_f,
                      int32_t* cache_i, float* out_f, int32_t* out_i) {
  LocalManifold m{};
  m.point_count = 0;
  const Xf t = xf_of(xf);
  if (kind_a == kKindBox && kind_b == kKindBox) {
    Hull ha = hull_of(a), hb = hull_of(b);
    SATCache c = sat_in(cache_f, cache_i);
    collide_hulls(m, ha, hb, t, c, nullptr, nullptr);
    sat_out(c, cache_f, cache_i);
  } else if (kind_a == kKindBox && kind_b == kKindSphere) {
    Hull ha = hull_of(a);
    SimplexCache c = simplex_in(cache_f, cache_i);
    collide_hull_and_sphere(m, ha, sphere_of(b), t, c);
    simplex_out(c, cache_f, cache_i);
  } else if (kind_a == kKindBox && kind_b == kKindCapsule) {
    Hull ha = hull_of(a);
    SimplexCache c = simplex_in(cache_f, cache_i);
    collide_hull_and_capsule(m, ha, capsule_of(b), t, c);
    simplex_out(c, cache_f, cache_i);
  } else if (kind_a == kKindCapsule && kind_b == kKindSphere) {
    collide_capsule_and_sphere(m, capsule_of(a), sphere_of(b), t);
  } else if (kind_a == kKindCapsule && kind_b == kKindCapsule) {
    collide_capsules(m, capsule_of(a), capsule_of(b), t);
  } else if (kind_a == kKindSphere && kind_b == kKindSphere) {
    collide_spheres(m, sphere_of(a), sphere_of(b), t);
  } else {
    return -1;
  }
  manifold_out(m, out_f, out_i);
  return 0;
}

// compute_separating_axis (convex_manifold.rs:392-899) and get_best_axis (manifold.rs:97-113).
// out_f = best normal xyz, best separation, face_a, face_b and edge separations;
// out_i = best kind, best index_a, best index_b, separated (0 = not separated).
void nslo_separating_axis(const float* a, const float* b, const float* xf, int early_return, float* out_f, int32_t* out_i) {
  Hull ha = hull_of(a), hb = hull_of(b);
  AxisQuery q = separating_axis_query(ha, hb, xf_of(xf), early_return != 0);
  const SepAxis* best;
  if (q.face_a.separation > q.face_b.separation)
    best = q.edge.separation > q.face_a.separation ? &q.edge : &q.face_a;
  else
    best = q.edge.separation > q.face_b.separation ? &q.edge : &q.face_b;
  out_f[0] = best->normal.x; out_f[1] = best->normal.y; out_f[2] = best->normal.z;
  out_f[3] = best->separation;
  out_f[4] = q.face_a.separation; out_f[5] = q.face_b.separation; out_f[6] = q.edge.separation;
  out_i[0] = best->kind; out_i[1] = best->index_a; out_i[2] = best->index_b; out_i[3] = q.separated;
}

// shape_distance (distance.rs:483-709) with use_radii false: na, nb points, B placed by xf.
// out_f = point_a xyz, point_b xyz, normal xyz, distance; out_i[0] = iterations.
void nslo_shape_distance(const float* pa, int na, const float* pb, int nb, const float* xf, float* cache_f, int32_t* cache_i,
                         float* out_f, int32_t* out_i) {
  Vec3 va[8], vb[8];
  for (int k = 0; k < na && k < 8; ++k) va[k] = Vec3{pa[3 * k], pa[3 * k + 1], pa[3 * k + 2]};
  for (int k = 0; k < nb && k < 8; ++k) vb[k] = Vec3{pb[3 * k], pb[3 * k + 1], pb[3 * k + 2]};
  SimplexCache c = simplex_in(cache_f, cache_i);
  DistanceOutput o = gjk_distance(va, na, vb, nb, xf_of(xf), c);
  simplex_out(c, cache_f, cache_i);
  const Vec3 v[3] = {o.point_a, o.point_b, o.normal};
  for (int k = 0; k < 3; ++k) { out_f[3 * k] = v[k].x; out_f[3 * k + 1] = v[k].y; out_f[3 * k + 2] = v[k].z; }
  out_f[9] = o.distance;
  out_i[0] = o.iterations;
}

// The box hull a shape array describes: points[8][3], planes[6][4] (normal, offset),
// half-edges[24][4] (next, twin, origin, face), centre[3].
void nslo_box_hull(const float* a, float* points, float* planes, int32_t* edges, float* center) {
  Hull h = hull_of(a);
  for (int k = 0; k < 8; ++k) { points[3 * k] = h.points[k].x; points[3 * k + 1] = h.points[k].y; points[3 * k + 2] = h.points[k].z; }
  for (int k = 0; k < 6; ++k) {
    planes[4 * k] = h.planes[k].n.x; planes[4 * k + 1] = h.planes[k].n.y; planes[4 * k + 2] = h.planes[k].n.z;
    planes[4 * k + 3] = h.planes[k].off;
  }
  for (int k = 0; k < 24; ++k) {
    edges[4 * k] = h.edges[k].next; edges[4 * k + 1] = h.edges[k].twin;
    edges[4 * k + 2] = h.edges[k].origin; edges[4 * k + 3] = h.edges[k].face;
  }
  center[0] = h.center.x; center[1] = h.center.y; center[2] = h.center.z;
}

}  // extern "C"
