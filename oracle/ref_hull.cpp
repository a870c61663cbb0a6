// Oracle: box hull construction and shape mass properties.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// make_transformed_box_hull follows src/collision/hull.rs:736-852 (topology tables,
// planes, points, bulk properties); box_inertia follows
// src/collision/geometry_queries.rs:81-87.
#include "ref_world.h"

namespace nslref {

// hull.rs:736-750: per-vertex first edge, (next, twin, origin, face) per half-edge,
// first edge per face. The table is the box topology itself, shared with the
// reference because feature ids index into it.
static const uint8_t kBoxVertexEdges[8] = {8, 1, 0, 9, 13, 3, 5, 11};
static const uint8_t kBoxEdgeData[24][4] = {
    {2, 1, 2, 0},   {17, 0, 1, 5},  {4, 3, 1, 0},   {20, 2, 5, 3},  {6, 5, 5, 0},   {23, 4, 6, 4},
    {0, 7, 6, 0},   {18, 6, 2, 2},  {10, 9, 0, 1},  {21, 8, 3, 5},  {12, 11, 3, 1}, {16, 10, 7, 2},
    {14, 13, 7, 1}, {19, 12, 4, 4}, {8, 15, 4, 1},  {22, 14, 0, 3}, {7, 17, 3, 2},  {9, 16, 2, 5},
    {11, 19, 6, 2}, {5, 18, 7, 4},  {15, 21, 1, 3}, {1, 20, 0, 5},  {3, 23, 4, 3},  {13, 22, 5, 4},
};
static const uint8_t kBoxFaceEdges[6] = {0, 8, 16, 20, 19, 21};

static Mat3 box_inertia(float mass, Vec3 mn, Vec3 mx) {
  Vec3 d = sub(mx, mn);
  float ixx = mass * (d.y * d.y + d.z * d.z) / 12.0f;
  float iyy = mass * (d.x * d.x + d.z * d.z) / 12.0f;
  float izz = mass * (d.x * d.x + d.y * d.y) / 12.0f;
  return diag3(ixx, iyy, izz);
}

Hull make_transformed_box_hull(float hx, float hy, float hz, Xf t) {
  Hull hull{};
  float min_h = 0.2f * linear_slop();
  Vec3 h = max3(Vec3{min_h, min_h, min_h}, Vec3{hx, hy, hz});
  for (int i = 0; i < 8; ++i) hull.vertex_edge[i] = kBoxVertexEdges[i];
  for (int i = 0; i < 24; ++i)
    hull.edges[i] = HalfEdge{kBoxEdgeData[i][0], kBoxEdgeData[i][1], kBoxEdgeData[i][2], kBoxEdgeData[i][3]};
  for (int i = 0; i < 6; ++i) hull.face_edge[i] = kBoxFaceEdges[i];

  Vec3 lower = neg(h), upper = h;
  hull.aabb = aabb_transform(t, AABB{lower, upper});
  hull.surface_area = 8.0f * (h.x * h.y + h.x * h.z + h.y * h.z);
  hull.volume = 8.0f * h.x * h.y * h.z;
  hull.inner_radius = minf(h.x, minf(h.y, h.z));
  hull.central_inertia = rotate_inertia(t.q, box_inertia(hull.volume, lower, upper));

  const Vec3 ax = {1, 0, 0}, ay = {0, 1, 0}, az = {0, 0, 1};
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

  for (int i = 0; i < 8; ++i) {
    hull.vx[i] = hull.points[i].x;
    hull.vy[i] = hull.points[i].y;
    hull.vz[i] = hull.points[i].z;
    hull.nx[i] = i < 6 ? hull.planes[i].n.x : 0.0f;
    hull.ny[i] = i < 6 ? hull.planes[i].n.y : 0.0f;
    hull.nz[i] = i < 6 ? hull.planes[i].n.z : 0.0f;
  }
  // The box builder leaves the edge-cosine array zero-initialised (hull.rs:258-282).
  for (int i = 0; i < 12; ++i) hull.edge_cosines[i] = 0.0f;
  return hull;
}

}  // namespace nslref
