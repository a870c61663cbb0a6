// The narrow phase on its own (nsl_collide): one collide call per lane, with the same
// device routines the step's collide kernel uses, so a query's manifold and cache equal the
// step's bit for bit.
#include <mutex>

#include "../../include/natsel_gpu.h"
#include "nsl_box_collide.h"
#include "nsl_narrow.h"
#include "nsl_round_collide.h"

namespace nsl {

namespace {

constexpr int kNarrowBlock = 64;  // one wavefront: the routines diverge by pair type

__device__ __forceinline__ Vec3 v3(const float* p) { return Vec3{p[0], p[1], p[2]}; }
__device__ __forceinline__ Xf xf7(const float* p) { return Xf{Vec3{p[0], p[1], p[2]}, Quat{Vec3{p[3], p[4], p[5]}, p[6]}}; }

__device__ void collide_query(const nsl_collide_query& q, nsl_local_manifold& out) {
  LocalManifold m;
  m.point_count = 0;
  m.normal = zero3();
  const Xf b_to_a = xf7(q.transform_b_to_a);
  const nsl_shape_geometry& a = q.a;
  const nsl_shape_geometry& b = q.b;
  nsl_collide_cache c = q.cache;
  c.hit = 0;
  if (a.type == NSL_SHAPE_HULL) {
    const BoxHull ha = make_box_hull(v3(a.params), xf7(a.frame));
    if (b.type == NSL_SHAPE_HULL) {
      const BoxHull hb = make_box_hull(v3(b.params), xf7(b.frame));
      SatCache sc{c.value, (uint8_t)c.kind, c.index_a[0], c.index_b[0], (uint8_t)c.hit};
      collide_boxes(m, ha, hb, b_to_a, sc);
      c.value = sc.separation;
      c.kind = sc.type;
      c.index_a[0] = sc.index_a;
      c.index_b[0] = sc.index_b;
      c.hit = sc.hit;
    } else {
      SimplexCache sc;
      sc.metric = c.value;
      sc.count = (uint16_t)c.kind;
      for (int k = 0; k < 4; ++k) {
        sc.index_a[k] = c.index_a[k];
        sc.index_b[k] = c.index_b[k];
      }
      if (b.type == NSL_SHAPE_SPHERE) collide_box_sphere(m, ha, v3(b.params), b.params[3], b_to_a, sc);
      else collide_box_capsule(m, ha, v3(b.params), v3(b.params + 3), b.params[6], b_to_a, sc);
      c.value = sc.metric;
      c.kind = sc.count;
      for (int k = 0; k < 4; ++k) {
        c.index_a[k] = sc.index_a[k];
        c.index_b[k] = sc.index_b[k];
      }
    }
  } else if (a.type == NSL_SHAPE_CAPSULE) {
    if (b.type == NSL_SHAPE_SPHERE)
      collide_capsule_sphere(m, v3(a.params), v3(a.params + 3), a.params[6], v3(b.params), b.params[3], b_to_a);
    else
      collide_capsules(m, v3(a.params), v3(a.params + 3), a.params[6], v3(b.params), v3(b.params + 3), b.params[6], b_to_a);
  } else {
    collide_spheres(m, v3(a.params), a.params[3], v3(b.params), b.params[3], b_to_a);
  }
  out.normal[0] = m.normal.x;
  out.normal[1] = m.normal.y;
  out.normal[2] = m.normal.z;
  out.point_count = m.point_count;
  for (int k = 0; k < 4; ++k) {
    const bool live = k < m.point_count;
    out.point[k][0] = live ? m.pts[k].point.x : 0.0f;
    out.point[k][1] = live ? m.pts[k].point.y : 0.0f;
    out.point[k][2] = live ? m.pts[k].point.z : 0.0f;
    out.separation[k] = live ? m.pts[k].separation : 0.0f;
    out.feature_id[k] = live ? feature_id(m.pts[k].pair) : 0u;
  }
  out.cache = c;
}

__global__ __launch_bounds__(kNarrowBlock) void k_collide_queries(const nsl_collide_query* q, nsl_local_manifold* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  nsl_collide_query query = q[i];
  nsl_local_manifold m;
  collide_query(query, m);
  out[i] = m;
}

int support_points(int32_t type) { return type == NSL_SHAPE_HULL ? 8 : (type == NSL_SHAPE_CAPSULE ? 2 : 1); }

}  // namespace

const char* validate_collide_queries(const nsl_collide_query* q, int32_t n, int32_t* bad) {
  for (int32_t i = 0; i < n; ++i) {
    *bad = i;
    const int32_t ta = q[i].a.type, tb = q[i].b.type;
    const bool ordered = (ta == NSL_SHAPE_HULL && (tb == NSL_SHAPE_HULL || tb == NSL_SHAPE_SPHERE || tb == NSL_SHAPE_CAPSULE)) ||
                         (ta == NSL_SHAPE_CAPSULE && (tb == NSL_SHAPE_SPHERE || tb == NSL_SHAPE_CAPSULE)) ||
                         (ta == NSL_SHAPE_SPHERE && tb == NSL_SHAPE_SPHERE);
    if (!ordered) return "shape pair outside the register table's order (hull, capsule, sphere)";
    const nsl_collide_cache& c = q[i].cache;
    if (ta != NSL_SHAPE_HULL) continue;
    if (tb == NSL_SHAPE_HULL) {
      if (c.kind == kAxisFaceA && c.index_a[0] > 5) return "SAT cache: face index of A above 5";
      if (c.kind == kAxisFaceB && c.index_b[0] > 5) return "SAT cache: face index of B above 5";
      if (c.kind == kAxisEdge && (c.index_a[0] > 22 || c.index_b[0] > 22 || (c.index_a[0] & 1) || (c.index_b[0] & 1)))
        return "SAT cache: edge index odd or above 22";
    } else {
      if (c.kind < 0 || c.kind > 3) return "simplex cache: count outside 0..3";
      for (int k = 0; k < c.kind; ++k)
        if (c.index_a[k] >= support_points(ta) || c.index_b[k] >= support_points(tb))
          return "simplex cache: vertex index beyond the shape's support points";
    }
  }
  *bad = -1;
  return nullptr;
}

void launch_collide_queries(const nsl_collide_query* q, nsl_local_manifold* out, int n, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_collide_queries, dim3((n + kNarrowBlock - 1) / kNarrowBlock), dim3(kNarrowBlock), 0, s, q, out, n);
}

// ---------------------------------------------------------------------------------
// The staging memory of nsl_collide_noalloc: the queries and the manifolds of one call, in
// a fixed device array instead of buffers allocated per call. Each device holds its own
// copy of the array; its address on a device comes from hipGetSymbolAddress, once per
// device (the array's name in host code is not a device address).
namespace {

constexpr size_t kQueryBytes = ((size_t)kNoAllocMaxQueries * sizeof(nsl_collide_query) + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = kQueryBytes + (size_t)kNoAllocMaxQueries * sizeof(nsl_local_manifold);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_lock;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

const char* no_alloc_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return "device ordinal out of range";
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (g_scratch_ptr[device] == nullptr) {
    // with `device` current (the caller set it): this device's copy of the array
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return "hipGetSymbolAddress failed";
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return nullptr;
}

const char* collide_queries_noalloc(int device, const nsl_collide_query* queries, nsl_local_manifold* out, int32_t n) {
  void* base = nullptr;
  if (const char* why = no_alloc_scratch(device, &base)) return why;
  nsl_collide_query* dq = static_cast<nsl_collide_query*>(base);
  nsl_local_manifold* dout = reinterpret_cast<nsl_local_manifold*>(static_cast<unsigned char*>(base) + kQueryBytes);
  if (hipMemcpy(dq, queries, (size_t)n * sizeof(nsl_collide_query), hipMemcpyHostToDevice) != hipSuccess)
    return "copy of the queries failed";
  launch_collide_queries(dq, dout, n, nullptr);
  if (hipGetLastError() != hipSuccess) return "collide launch failed";
  if (hipMemcpy(out, dout, (size_t)n * sizeof(nsl_local_manifold), hipMemcpyDeviceToHost) != hipSuccess)
    return "copy of the manifolds failed";
  return nullptr;
}

}  // namespace nsl
