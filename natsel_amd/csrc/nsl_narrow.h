// The narrow phase on its own (nsl_collide, include/natsel_gpu.h): host validation of the
// queries and the launch of one collide call per lane.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/natsel_gpu.h"
#include "nsl_types.h"

namespace nsl {

// make_transformed_box_hull (hull.rs:754-852) for the fields the box routines read: the box
// of half extents `half` placed by `t` in its shape frame.
NSL_HD BoxHull make_box_hull(Vec3 half, Xf t) {
  const float min_h = 0.2f * kLinearSlop;
  const Vec3 h = max3(Vec3{min_h, min_h, min_h}, half);
  BoxHull hull;
  const Vec3 lower = neg(h), upper = h;
  hull.aabb = aabb_transform(t, Aabb{lower, upper});
  hull.inner_radius = minf(h.x, minf(h.y, h.z));
  const Vec3 ax{1, 0, 0}, ay{0, 1, 0}, az{0, 0, 1};
  hull.planes[0] = transform_plane(t, plane_from_normal_point(neg(ax), lower));
  hull.planes[1] = transform_plane(t, plane_from_normal_point(ax, upper));
  hull.planes[2] = transform_plane(t, plane_from_normal_point(neg(ay), lower));
  hull.planes[3] = transform_plane(t, plane_from_normal_point(ay, upper));
  hull.planes[4] = transform_plane(t, plane_from_normal_point(neg(az), lower));
  hull.planes[5] = transform_plane(t, plane_from_normal_point(az, upper));
  const Vec3 corners[8] = {{h.x, h.y, h.z},  {-h.x, h.y, h.z},  {-h.x, -h.y, h.z},  {h.x, -h.y, h.z},
                           {h.x, h.y, -h.z}, {-h.x, h.y, -h.z}, {-h.x, -h.y, -h.z}, {h.x, -h.y, -h.z}};
  for (int i = 0; i < 8; ++i) hull.points[i] = xf_point(t, corners[i]);
  hull.center = t.p;
  return hull;
}

// nullptr when every query is well formed; else the reason, with *bad the query's index.
const char* validate_collide_queries(const nsl_collide_query* q, int32_t n, int32_t* bad);
void launch_collide_queries(const nsl_collide_query* q, nsl_local_manifold* out, int n, hipStream_t s);

// nsl_collide_noalloc: at most this many queries per call, staged in a fixed device array
// (about 2.2 MiB per device, held until the process ends).
constexpr int32_t kNoAllocMaxQueries = 8192;
// One call through the fixed array on `device` (current). nullptr on success, else the reason.
const char* collide_queries_noalloc(int device, const nsl_collide_query* queries, nsl_local_manifold* out, int32_t n);

}  // namespace nsl
