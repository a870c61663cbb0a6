// Host side of the GPU stepper (see nsl_world.h).
#include "nsl_world.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace nsl {

#define NSL_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// hull.rs:754-852 for the box; geometry_queries.rs:81-87 for its inertia.
BoxHull make_box(Vec3 half, Mat3* central_inertia, float* volume) {
  const float min_h = 0.2f * kLinearSlop;
  Vec3 h = max3(Vec3{min_h, min_h, min_h}, half);
  BoxHull hull;
  Vec3 lower = neg(h), upper = h;
  hull.aabb = aabb_transform(Xf{zero3(), quat_identity()}, Aabb{lower, upper});
  *volume = 8.0f * h.x * h.y * h.z;
  hull.inner_radius = minf(h.x, minf(h.y, h.z));
  Vec3 dd = sub(upper, lower);
  float m = *volume;
  Mat3 box_i = diag3(m * (dd.y * dd.y + dd.z * dd.z) / 12.0f, m * (dd.x * dd.x + dd.z * dd.z) / 12.0f,
                     m * (dd.x * dd.x + dd.y * dd.y) / 12.0f);
  *central_inertia = rotate_inertia(quat_identity(), box_i);
  const Xf id{zero3(), quat_identity()};
  const Vec3 ax{1, 0, 0}, ay{0, 1, 0}, az{0, 0, 1};
  hull.planes[0] = transform_plane(id, plane_from_normal_point(neg(ax), lower));
  hull.planes[1] = transform_plane(id, plane_from_normal_point(ax, upper));
  hull.planes[2] = transform_plane(id, plane_from_normal_point(neg(ay), lower));
  hull.planes[3] = transform_plane(id, plane_from_normal_point(ay, upper));
  hull.planes[4] = transform_plane(id, plane_from_normal_point(neg(az), lower));
  hull.planes[5] = transform_plane(id, plane_from_normal_point(az, upper));
  const Vec3 corners[8] = {{h.x, h.y, h.z},  {-h.x, h.y, h.z},  {-h.x, -h.y, h.z},  {h.x, -h.y, h.z},
                           {h.x, h.y, -h.z}, {-h.x, h.y, -h.z}, {-h.x, -h.y, -h.z}, {h.x, -h.y, -h.z}};
  for (int i = 0; i < 8; ++i) hull.points[i] = xf_point(id, corners[i]);
  hull.center = zero3();
  return hull;
}

}  // namespace

GpuWorld::GpuWorld(int device) : device_(device) {
  NSL_CHECK(hipSetDevice(device_));
  NSL_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  timer_.capacity = 64;
  timer_.events = new hipEvent_t[2 * timer_.capacity];
  for (int i = 0; i < 2 * timer_.capacity; ++i) NSL_CHECK(hipEventCreate(&timer_.events[i]));
  NSL_CHECK(hipMalloc(&d_.counters, 8 * sizeof(int32_t)));
  NSL_CHECK(hipMemset(d_.counters, 0, 8 * sizeof(int32_t)));
  NSL_CHECK(hipMalloc(&d_barrier_, sizeof(GridBarrier)));
  NSL_CHECK(hipMemset(d_barrier_, 0, sizeof(GridBarrier)));
  NSL_CHECK(hipMalloc(&d_phase_log_, 2 * kMaxPhaseLog * sizeof(uint64_t)));
  phase_log_host_.resize(2 * kMaxPhaseLog);
  max_blocks_ = solve_step_max_blocks(device_);
  if (max_blocks_ <= 0) throw std::runtime_error("step kernel does not fit on the device");
  int khz = 0;
  NSL_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
  wall_clock_khz_ = (double)khz;
  finder_ = std::make_unique<PairFinder>(stream_);
  ensure_pair_set();
}

GpuWorld::~GpuWorld() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  void* ptrs[] = {d_.sims, d_.states, d_.static_sims, d_.shapes, d_.fat_aabbs, d_.shape_moved, d_.hulls, d_.contacts,
                  d_.contact_change, d_.constraints, d_.overflow, d_.collide_ids, d_.solve_ids, d_.overflow_ids,
                  d_.counters, d_init_ids_, d_init_records_, d_barrier_, d_phase_log_, d_proxies_,
                  d_rebuild_keys_, d_.pairs.slots};
  finder_.reset();
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int i = 0; i < 2 * timer_.capacity; ++i) (void)hipEventDestroy(timer_.events[i]);
  delete[] timer_.events;
  if (stream_) (void)hipStreamDestroy(stream_);
}

template <class T>
void GpuWorld::ensure(T*& ptr, size_t& cap, size_t n, bool keep) {
  if (n <= cap && ptr) return;
  size_t new_cap = std::max<size_t>(n + n / 2, 64);
  T* p = nullptr;
  NSL_CHECK(hipMalloc(&p, new_cap * sizeof(T)));
  if (ptr) {
    if (keep && cap > 0) NSL_CHECK(hipMemcpyAsync(p, ptr, cap * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
    NSL_CHECK(hipFree(ptr));
  }
  ptr = p;
  cap = new_cap;
}

// ---------------------------------------------------------------------------------
// creation (world_runtime.rs:1506-1649, shape_creation.rs:286-559, body_api.rs:611-802)

int32_t GpuWorld::create_body(const BodyDesc& desc) {
  if (device_newer_) pull_bodies();
  int32_t id = (int32_t)bodies_.size();
  BodySim sim{};
  sim.transform = WXf{desc.position, desc.rotation};
  sim.center = desc.position;
  sim.rotation0 = desc.rotation;
  sim.center0 = desc.position;
  sim.min_extent = kHuge;
  sim.linear_damping = desc.linear_damping;
  sim.angular_damping = desc.angular_damping;
  sim.gravity_scale = desc.gravity_scale;
  sim.body_id = id;
  uint32_t flags = desc.lock_flags;
  flags |= desc.is_bullet ? bflag::kIsBullet : 0;
  flags |= desc.allow_fast_rotation ? bflag::kAllowFastRotation : 0;
  flags |= desc.type == kBodyDynamic ? bflag::kDynamic : 0;
  flags |= desc.enable_sleep ? bflag::kEnableSleep : 0;
  flags |= desc.enable_contact_recycling ? bflag::kEnableRecycling : 0;
  sim.flags = flags;
  sim.body_flags = flags;
  sim.sleep_threshold = desc.sleep_threshold;
  sim.safety_factor = desc.safety_factor;
  sim.head_shape = kNullIndex;
  sim.body_type = desc.type;
  HBody hb;
  hb.type = desc.type;
  hb.flags = flags;
  if (desc.type == kBodyStatic) {
    hb.set = 0;
    hb.local = (int32_t)static_host_.size();
    static_host_.push_back(sim);
  } else {
    hb.set = 2;
    hb.local = (int32_t)awake_host_.size();
    awake_host_.push_back(sim);
    BodyState st{};
    st.v = desc.linear_velocity;
    st.w = desc.angular_velocity;
    st.dq = quat_identity();
    st.flags = flags;
    states_host_.push_back(st);
  }
  bodies_.push_back(hb);
  bodies_dirty_ = true;
  return id;
}

Aabb GpuWorld::shape_local_aabb(const Shape& s, Quat q) const {
  Xf t{zero3(), q};
  if (s.kind == kShapeHull) return aabb_transform(t, hulls_[s.hull].aabb);
  Vec3 e{s.radius, s.radius, s.radius};
  if (s.kind == kShapeSphere) {
    Vec3 c = xf_point(t, s.c1);
    return Aabb{sub(c, e), add(c, e)};
  }
  Vec3 c1 = xf_point(t, s.c1), c2 = xf_point(t, s.c2);
  return Aabb{sub(min3(c1, c2), e), add(max3(c1, c2), e)};
}

int32_t GpuWorld::add_shape(int32_t body, Shape shape, const Material& m, float radius) {
  if (device_newer_) pull_bodies();
  HBody& b = bodies_[body];
  int32_t id = (int32_t)shapes_.size();
  shape.body_id = body;
  shape.body_enc = encode(b);
  shape.friction = m.friction;
  shape.restitution = m.restitution;
  shape.rolling_resistance = m.rolling_resistance;
  shape.tangent_velocity = zero3();
  shape.enable_hit_events = m.enable_hit_events ? 1u : 0u;
  // shape_creation.rs:196-231
  float margin = 0.0f;
  if (shape.kind == kShapeSphere) margin = shape.radius;
  else if (shape.kind == kShapeCapsule) margin = 0.5f * distance(shape.c2, shape.c1) + shape.radius;
  else {
    const BoxHull& h = hulls_[shape.hull];
    float mx = 0.0f;
    for (int i = 0; i < 8; ++i) mx = maxf(mx, distance_sq(h.points[i], h.center));
    margin = sqrtf_(mx);
  }
  shape.aabb_margin = minf(kMaxAabbMargin, kAabbMarginFraction * margin);
  // shape_creation.rs:237-283 (proxy creation)
  BodySim& sim = b.set == 2 ? awake_host_[b.local] : static_host_[b.local];
  Aabb aabb = offset_aabb(aabb_inflate(shape_local_aabb(shape, sim.transform.q), kSpeculativeDistance), sim.transform.p);
  HShape hs;
  hs.id = id;
  hs.body = body;
  hs.kind = shape.kind;
  hs.hull = shape.hull;
  hs.category = m.category_bits;
  hs.mask = m.mask_bits;
  hs.group = m.group_index;
  hs.density = m.density;
  hs.restitution = m.restitution;
  hs.rolling_resistance = m.rolling_resistance;
  hs.radius = radius;
  hs.speculative = m.enable_speculative;
  hs.fat = aabb_inflate(aabb, b.type == kBodyStatic ? kSpeculativeDistance : shape.aabb_margin);
  hs.proxy_moved = true;  // dynamic proxies always; static ones via invoke_contact_creation
  shape.next_shape = b.head_shape;
  b.head_shape = id;
  sim.head_shape = id;
  shapes_.push_back(shape);
  hshapes_.push_back(hs);
  shapes_dirty_ = true;
  update_mass(body);
  bodies_dirty_ = true;
  return id;
}

int32_t GpuWorld::create_box(int32_t body, Vec3 half, const Material& m) {
  Shape s{};
  s.kind = kShapeHull;
  Mat3 ci;
  float vol;
  BoxHull h = make_box(half, &ci, &vol);
  hulls_.push_back(h);
  hull_inertia_.push_back(ci);
  hull_volume_.push_back(vol);
  s.hull = (int32_t)hulls_.size() - 1;
  s.radius = 0.0f;
  return add_shape(body, s, m, 0.0f);
}

int32_t GpuWorld::create_sphere(int32_t body, Vec3 center, float radius, const Material& m) {
  Shape s{};
  s.kind = kShapeSphere;
  s.hull = -1;
  s.c1 = center;
  s.c2 = center;
  s.radius = radius;
  return add_shape(body, s, m, radius);
}

int32_t GpuWorld::create_capsule(int32_t body, Vec3 c1, Vec3 c2, float radius, const Material& m) {
  if (distance(c1, c2) < kLinearSlop) return create_sphere(body, scale(0.5f, add(c1, c2)), radius, m);
  Shape s{};
  s.kind = kShapeCapsule;
  s.hull = -1;
  s.c1 = c1;
  s.c2 = c2;
  s.radius = radius;
  return add_shape(body, s, m, radius);
}

namespace {
struct MassData { float mass; Vec3 center; Mat3 inertia; };

Quat quat_between_unit_vectors(Vec3 v1, Vec3 v2) {  // functions.rs:1030-1058
  Quat out{};
  Vec3 m = lerp(v1, v2, 0.5f);
  float tol = 100.0f * kFltEps;
  if (length_sq(m) > tol * tol) {
    out.v = cross(v1, m);
    out.s = dot(v1, m);
  } else {
    out.v = absf(v1.x) > 0.5f ? Vec3{v1.y, -v1.x, 0.0f} : Vec3{0.0f, v1.z, -v1.y};
    out.s = 0.0f;
  }
  return normalize_quat(out);
}
}  // namespace

// body_api.rs:642-802 plus update_isotropic_inertia (body_api.rs:611-640)
void GpuWorld::update_mass(int32_t body) {
  HBody& b = bodies_[body];
  BodySim& sim = b.set == 2 ? awake_host_[b.local] : static_host_[b.local];
  b.flags &= ~(bflag::kDirtyMass | bflag::kGyroscopic);
  sim.inertia = mat3_zero();
  sim.inv_mass = 0.0f;
  sim.inv_inertia_local = mat3_zero();
  sim.inv_inertia_world = mat3_zero();
  sim.local_center = zero3();
  sim.min_extent = kHuge;
  sim.max_extent = zero3();
  auto extent = [&](const Shape& s, Vec3 lc, float* mn, Vec3* mx) {
    if (s.kind == kShapeCapsule) {
      *mn = s.radius;
      *mx = add(max3(abs3(sub(s.c1, lc)), abs3(sub(s.c2, lc))), Vec3{s.radius, s.radius, s.radius});
    } else if (s.kind == kShapeSphere) {
      *mn = s.radius;
      *mx = add(abs3(sub(s.c1, lc)), Vec3{s.radius, s.radius, s.radius});
    } else {
      const BoxHull& h = hulls_[s.hull];
      *mn = h.inner_radius;
      Vec3 e = zero3();
      for (int i = 0; i < 8; ++i) e = max3(e, abs3(sub(h.points[i], lc)));
      *mx = e;
    }
  };
  if (b.type != kBodyDynamic) {
    sim.center = sim.transform.p;
    sim.center0 = sim.center;
    if (b.type == kBodyKinematic) {
      for (int32_t sid = b.head_shape; sid != kNullIndex; sid = shapes_[sid].next_shape) {
        float mn;
        Vec3 mx;
        extent(shapes_[sid], zero3(), &mn, &mx);
        sim.min_extent = minf(sim.min_extent, mn);
        sim.max_extent = max3(sim.max_extent, mx);
      }
    }
    sim.flags = (sim.flags & (bflag::kIsFast | bflag::kHadToi)) | (b.flags & ~bflag::kTransient);
    sim.body_flags = b.flags;
    return;
  }
  std::vector<MassData> masses;
  float total = 0.0f;
  Vec3 lc = zero3();
  for (int32_t sid = b.head_shape; sid != kNullIndex; sid = shapes_[sid].next_shape) {
    const Shape& s = shapes_[sid];
    const HShape& hs = hshapes_[sid];
    if (hs.density == 0.0f) { masses.push_back(MassData{}); continue; }
    MassData md{};
    if (s.kind == kShapeHull) {  // shape_runtime.rs:312-320
      md.mass = hs.density * hull_volume_[s.hull];
      md.center = hulls_[s.hull].center;
      md.inertia = mul_sm(hs.density, hull_inertia_[s.hull]);
    } else if (s.kind == kShapeSphere) {  // sphere.rs:13-24
      float r = s.radius;
      float mass = 4.0f / 3.0f * kPi * r * r * r * hs.density;
      float ixx = 0.4f * mass * r * r;
      md = MassData{mass, s.c1, diag3(ixx, ixx, ixx)};
    } else {  // capsule.rs:27-63
      float r = s.radius;
      float ch = distance(s.c1, s.c2);
      float cyl_mass = kPi * r * r * ch * hs.density;
      float sph_mass = (4.0f / 3.0f) * kPi * r * r * r * hs.density;
      float rr = r * r, hh = ch * ch;
      Mat3 cyl = diag3(cyl_mass * (3.0f * rr + hh) / 12.0f, 0.5f * cyl_mass * rr, cyl_mass * (3.0f * rr + hh) / 12.0f);
      float si = 0.4f * sph_mass * r * r;
      Mat3 inertia = add_mm(cyl, diag3(si, si, si));
      float st = 0.125f * sph_mass * (3.0f * r + 2.0f * ch) * ch;
      inertia.cx.x += st;
      inertia.cz.z += st;
      Mat3 rot = diag3(1.0f, 1.0f, 1.0f);
      if (ch * ch > 1000.0f * kFltMin) rot = matrix_from_quat(quat_between_unit_vectors(Vec3{0, 1, 0}, normalize(sub(s.c2, s.c1))));
      md = MassData{sph_mass + cyl_mass, scale(0.5f, add(s.c1, s.c2)), mul_mm(rot, mul_mm(inertia, transpose(rot)))};
    }
    total += md.mass;
    lc = mul_add(lc, md.mass, md.center);
    masses.push_back(md);
  }
  if (total > 0.0f) {
    sim.inv_mass = 1.0f / total;
    lc = scale(sim.inv_mass, lc);
  }
  for (const MassData& md : masses) {
    if (md.mass == 0.0f) continue;
    sim.inertia = add_mm(sim.inertia, add_mm(md.inertia, steiner(md.mass, sub(lc, md.center))));
  }
  if (det(sim.inertia) > 0.0f) {
    sim.inv_inertia_local = invert_t(sim.inertia);
    Mat3 rm = matrix_from_quat(sim.transform.q);
    sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
  }
  Pos old_center = sim.center;
  sim.local_center = lc;
  sim.center = xf_world_point(sim.transform, lc);
  sim.center0 = sim.center;
  BodyState& st = states_host_[b.local];
  st.v = add(st.v, cross(st.w, sub_pos(sim.center, old_center)));
  for (int32_t sid = b.head_shape; sid != kNullIndex; sid = shapes_[sid].next_shape) {
    float mn;
    Vec3 mx;
    extent(shapes_[sid], lc, &mn, &mx);
    sim.min_extent = minf(sim.min_extent, mn);
    sim.max_extent = max3(sim.max_extent, mx);
  }
  if ((sim.flags & bflag::kFixedRotation) == bflag::kFixedRotation) {
    sim.inertia = mat3_zero();
    sim.inv_inertia_local = mat3_zero();
    sim.inv_inertia_world = mat3_zero();
  }
  Mat3 in = sim.inertia;
  float mean = (in.cx.x + in.cy.y + in.cz.z) / 3.0f;
  float dx = in.cx.x - mean, dy = in.cy.y - mean, dz = in.cz.z - mean;
  float off = in.cy.x * in.cy.x + in.cz.x * in.cz.x + in.cz.y * in.cz.y;
  float deviation = dx * dx + dy * dy + dz * dz + 2.0f * off;
  float tol = options.isotropic_inertia_tolerance;
  if (2.0f * deviation > 3.0f * tol * tol * mean * mean) b.flags |= bflag::kGyroscopic;
  else b.flags &= ~bflag::kGyroscopic;
  sim.flags = (sim.flags & (bflag::kIsFast | bflag::kHadToi)) | (b.flags & ~bflag::kTransient);
  sim.body_flags = b.flags;
  st.flags = b.flags & ~bflag::kTransient;
}

// ---------------------------------------------------------------------------------
// device residency of bodies and shapes

void GpuWorld::upload_bodies() {
  const size_t na = awake_host_.size(), ns = static_host_.size();
  ensure(d_.sims, cap_sims_, std::max<size_t>(na, 1), false);
  ensure(d_.states, cap_states_, std::max<size_t>(na, 1), false);
  ensure(d_.static_sims, cap_static_, std::max<size_t>(ns, 1), false);
  if (na) {
    NSL_CHECK(hipMemcpyAsync(d_.sims, awake_host_.data(), na * sizeof(BodySim), hipMemcpyHostToDevice, stream_));
    NSL_CHECK(hipMemcpyAsync(d_.states, states_host_.data(), na * sizeof(BodyState), hipMemcpyHostToDevice, stream_));
  }
  if (ns) NSL_CHECK(hipMemcpyAsync(d_.static_sims, static_host_.data(), ns * sizeof(BodySim), hipMemcpyHostToDevice, stream_));
  d_.n_awake = (int32_t)na;
  if (shapes_dirty_) {
    const size_t nsh = shapes_.size();
    ensure(d_.shapes, cap_shapes_, std::max<size_t>(nsh, 1), false);
    ensure(d_.fat_aabbs, cap_fat_, std::max<size_t>(nsh, 1), false);
    ensure(d_.shape_moved, cap_moved_, std::max<size_t>(nsh, 1), false);
    ensure(d_proxies_, cap_proxies_, std::max<size_t>(nsh, 1), false);
    ensure(d_.hulls, cap_hulls_, std::max<size_t>(hulls_.size(), 1), false);
    fat_host_.resize(nsh);
    moved_host_.resize(nsh);
    proxies_host_.resize(nsh);
    for (size_t i = 0; i < nsh; ++i) {
      const HShape& hs = hshapes_[i];
      fat_host_[i] = hs.fat;
      moved_host_[i] = hs.proxy_moved ? 1 : 0;
      pairs_pending_ = pairs_pending_ || hs.proxy_moved;
      proxies_host_[i] = Proxy{hs.body, bodies_[hs.body].type, hs.group, 0, hs.category, hs.mask};
    }
    if (nsh) {
      NSL_CHECK(hipMemcpyAsync(d_.shapes, shapes_.data(), nsh * sizeof(Shape), hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_.fat_aabbs, fat_host_.data(), nsh * sizeof(Aabb), hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_.shape_moved, moved_host_.data(), nsh, hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_proxies_, proxies_host_.data(), nsh * sizeof(Proxy), hipMemcpyHostToDevice, stream_));
    }
    if (!hulls_.empty())
      NSL_CHECK(hipMemcpyAsync(d_.hulls, hulls_.data(), hulls_.size() * sizeof(BoxHull), hipMemcpyHostToDevice, stream_));
    // the staging vectors are read by the copies above
    NSL_CHECK(hipStreamSynchronize(stream_));
    n_device_shapes_ = nsh;
    shapes_dirty_ = false;
  }
  bodies_dirty_ = false;
  device_newer_ = false;
}

void GpuWorld::pull_bodies() {
  const size_t na = awake_host_.size();
  if (na && d_.sims) {
    NSL_CHECK(hipMemcpyAsync(awake_host_.data(), d_.sims, na * sizeof(BodySim), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipMemcpyAsync(states_host_.data(), d_.states, na * sizeof(BodyState), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
  }
  if (n_device_shapes_) {  // finalize moves fat AABBs and marks proxies on the device
    fat_host_.resize(n_device_shapes_);
    moved_host_.resize(n_device_shapes_);
    NSL_CHECK(hipMemcpyAsync(fat_host_.data(), d_.fat_aabbs, n_device_shapes_ * sizeof(Aabb), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipMemcpyAsync(moved_host_.data(), d_.shape_moved, n_device_shapes_, hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n_device_shapes_; ++i) {
      hshapes_[i].fat = fat_host_[i];
      hshapes_[i].proxy_moved = moved_host_[i] != 0;
    }
  }
  for (size_t i = 0; i < awake_host_.size(); ++i) bodies_[awake_host_[i].body_id].flags = awake_host_[i].body_flags;
  device_newer_ = false;
}

// ---------------------------------------------------------------------------------
// pair update: contact_creation.rs:901-1042. New pairs come from the device pair finder
// (nsl_broadphase.h) as sorted keys; contacts are created in that order.

void GpuWorld::update_pairs() {
  if (!pairs_pending_ && counters_host_[6] == 0) return;
  finder_->find(d_, d_proxies_, (int)n_device_shapes_, new_keys_);
  for (HShape& s : hshapes_) s.proxy_moved = false;
  pairs_pending_ = false;
  kstats_.pair_tests += finder_->last_tests();
  for (uint64_t key : new_keys_)
    create_contact((int32_t)((key >> 42) & ((1u << 22) - 1)), (int32_t)((key >> 20) & ((1u << 22) - 1)));
}

// Keeps the device pair-key set at a load factor of at most 1/2 (live keys plus
// tombstones); rebuilds it from the live contacts when that would be exceeded.
void GpuWorld::ensure_pair_set() {
  const int64_t live = contact_count();
  const uint32_t cap = d_.pairs.slots ? d_.pairs.mask + 1 : 0;
  if (cap && 2 * (live + pair_set_tombs_) <= (int64_t)cap) return;
  uint32_t want = 1024;
  while ((int64_t)want < 4 * live) want <<= 1;
  if (want != cap) {
    if (d_.pairs.slots) {
      NSL_CHECK(hipStreamSynchronize(stream_));
      NSL_CHECK(hipFree(d_.pairs.slots));
    }
    NSL_CHECK(hipMalloc(&d_.pairs.slots, (size_t)want * sizeof(uint64_t)));
    d_.pairs.mask = want - 1;
  }
  NSL_CHECK(hipMemsetAsync(d_.pairs.slots, 0xFF, (size_t)want * sizeof(uint64_t), stream_));  // all kPairEmpty
  std::vector<uint64_t> keys;
  keys.reserve((size_t)live);
  for (const HContact& c : hcontacts_)
    if (c.alive) keys.push_back(pair_key(c.shape_a, c.shape_b));
  if (!keys.empty()) {
    ensure(d_rebuild_keys_, cap_rebuild_keys_, keys.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_rebuild_keys_, keys.data(), keys.size() * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
    launch_pairset_insert(d_.pairs, d_rebuild_keys_, (int)keys.size(), stream_);
    NSL_CHECK(hipStreamSynchronize(stream_));  // `keys` is read by the copy
  }
  pair_set_tombs_ = 0;
}

// contact_creation.rs:219-425; the primary shape order is hull > capsule > sphere.
void GpuWorld::create_contact(int32_t a, int32_t b) {
  auto rank = [](int32_t kind) { return kind == kShapeHull ? 2 : (kind == kShapeCapsule ? 1 : 0); };
  if (rank(hshapes_[a].kind) < rank(hshapes_[b].kind)) std::swap(a, b);
  const HShape& sa = hshapes_[a];
  const HShape& sb = hshapes_[b];
  const HBody& ba = bodies_[sa.body];
  const HBody& bb = bodies_[sb.body];
  int32_t id;
  if (!contact_free_.empty()) {
    id = contact_free_.back();
    contact_free_.pop_back();
  } else {
    id = contact_pool_next_++;
  }
  if (id >= (int32_t)hcontacts_.size()) hcontacts_.resize(id + 1);
  HContact& hc = hcontacts_[id];
  hc.shape_a = a;
  hc.shape_b = b;
  hc.color = kNullIndex;
  hc.set = 2;
  hc.alive = 1;
  hc.restitution = maxf(sa.restitution, sb.restitution);
  hc.local = (int32_t)non_touching_.size();
  non_touching_.push_back(id);
  Contact c{};
  c.shape_a = a;
  c.shape_b = b;
  c.enc_a = encode(ba);
  c.enc_b = encode(bb);
  if (ba.type == kBodyStatic || bb.type == kBodyStatic) c.flags |= cflag::kStatic;
  if (options.enable_speculative && sa.speculative && sb.speculative) c.flags |= cflag::kSpeculative;
  if ((ba.flags & bflag::kEnableRecycling) && (bb.flags & bflag::kEnableRecycling) && (c.flags & cflag::kSpeculative))
    c.flags |= cflag::kRecycle;
  c.rolling_resistance = maxf(sa.rolling_resistance, sb.rolling_resistance) * maxf(sa.radius, sb.radius);
  c.cached_rot_a = quat_identity();
  c.cached_rot_b = quat_identity();
  init_ids_.push_back(id);
  init_records_.push_back(c);
}

template <class T>
static int32_t swap_remove(std::vector<T>& v, size_t index) {
  size_t last = v.size() - 1;
  if (index != last) {
    v[index] = v[last];
    v.pop_back();
    return (int32_t)last;
  }
  v.pop_back();
  return kNullIndex;
}

// contact_creation.rs:428-552
void GpuWorld::destroy_contact(int32_t id) {
  HContact& c = hcontacts_[id];
  if (!c.alive) return;
  pair_set_tombs_ += 1;  // the collide kernel erased the key
  if (c.color == kNullIndex) {
    int32_t local = c.local;
    if (swap_remove(non_touching_, (size_t)local) != kNullIndex) hcontacts_[non_touching_[local]].local = local;
  } else {
    remove_from_graph(hshapes_[c.shape_a].body, hshapes_[c.shape_b].body, c.color, c.local);
  }
  HContact& c2 = hcontacts_[id];
  c2.alive = 0;
  c2.color = kNullIndex;
  c2.local = kNullIndex;
  c2.set = kNullIndex;
  if (id == contact_pool_next_ - 1) contact_pool_next_ -= 1;  // id_pool.rs:36-45 (never true: ids below next)
  else contact_free_.push_back(id);
}

// constraint_graph.rs:113-197
void GpuWorld::add_to_graph(int32_t id) {
  HContact& c = hcontacts_[id];
  int32_t a = hshapes_[c.shape_a].body, b = hshapes_[c.shape_b].body;
  int32_t ta = bodies_[a].type, tb = bodies_[b].type;
  int32_t color = kOverflowColor;
  if (ta == kBodyDynamic && tb == kBodyDynamic) {
    for (int i = 0; i < kDynamicColorCount; ++i) {
      if (colors_[i].get(a) || colors_[i].get(b)) continue;
      colors_[i].set(a);
      colors_[i].set(b);
      color = i;
      break;
    }
  } else if (ta == kBodyDynamic || tb == kBodyDynamic) {
    int32_t dyn = ta == kBodyDynamic ? a : b;
    for (int i = kOverflowColor - 1; i >= 1; --i) {
      if (!colors_[i].get(dyn)) {
        colors_[i].set(dyn);
        color = i;
        break;
      }
    }
  }
  c.color = color;
  if (color == kOverflowColor) {
    c.local = (int32_t)colors_[color].overflow.size();
    colors_[color].overflow.push_back(id);
  } else {
    c.local = (int32_t)colors_[color].convex.size();
    colors_[color].convex.push_back(id);
  }
}

// constraint_graph.rs:200-251
void GpuWorld::remove_from_graph(int32_t body_a, int32_t body_b, int32_t color, int32_t local) {
  if (color != kOverflowColor) {
    colors_[color].clear(body_a);
    colors_[color].clear(body_b);
    if (swap_remove(colors_[color].convex, (size_t)local) != kNullIndex) hcontacts_[colors_[color].convex[local]].local = local;
  } else {
    if (swap_remove(colors_[color].overflow, (size_t)local) != kNullIndex) hcontacts_[colors_[color].overflow[local]].local = local;
  }
}

// collide_pass.rs:386-478: transitions in contact-id order.
void GpuWorld::state_pass() {
  for (size_t id = 0; id < hcontacts_.size(); ++id) {
    uint8_t ch = change_host_[id];
    if (ch == kChangeNone || !hcontacts_[id].alive) continue;
    if (ch == kChangeDisjoint) {
      destroy_contact((int32_t)id);
    } else if (ch == kChangeBegan) {
      int32_t old_local = hcontacts_[id].local;
      add_to_graph((int32_t)id);
      if (swap_remove(non_touching_, (size_t)old_local) != kNullIndex) hcontacts_[non_touching_[old_local]].local = old_local;
    } else if (ch == kChangeEnded) {
      HContact& c = hcontacts_[id];
      int32_t color = c.color, local = c.local;
      c.color = kNullIndex;
      c.local = (int32_t)non_touching_.size();
      non_touching_.push_back((int32_t)id);
      remove_from_graph(hshapes_[c.shape_a].body, hshapes_[c.shape_b].body, color, local);
    }
  }
}

// ---------------------------------------------------------------------------------
// one step: solver_integration.rs:807-1015

void GpuWorld::step(float dt, int32_t sub_steps) {
  NSL_CHECK(hipSetDevice(device_));
  const double t0 = now_ms();
  if (bodies_dirty_ || shapes_dirty_) upload_bodies();

  update_pairs();
  const double t1 = now_ms();

  StepParams p{};
  p.dt = dt;
  p.sub_steps = maxi(1, sub_steps);
  if (dt > 0.0f) {
    p.inv_dt = 1.0f / dt;
    p.h = dt / (float)p.sub_steps;
    p.inv_h = (float)p.sub_steps * p.inv_dt;
  }
  float contact_hz = minf(options.contact_hertz, 0.125f * p.inv_h);
  p.contact_softness = make_soft(contact_hz, options.contact_damping_ratio, p.h);
  p.static_softness = make_soft(2.0f * contact_hz, 0.5f * options.contact_damping_ratio, p.h);
  p.contact_speed = options.contact_speed;
  p.max_linear_speed = options.max_linear_speed;
  p.restitution_threshold = options.restitution_threshold;
  p.gravity = options.gravity;
  p.warm_start_scale = options.enable_warm_starting ? 1.0f : 0.0f;
  p.enable_sleep = options.enable_sleep;
  p.enable_continuous = options.enable_continuous;
  p.enable_restitution_propagation = options.enable_restitution_propagation;

  const size_t ncap = hcontacts_.size();
  ensure(d_.contacts, cap_contacts_, std::max<size_t>(ncap, 1), true);
  ensure(d_.contact_change, cap_change_, std::max<size_t>(ncap, 1), false);
  if (!init_ids_.empty()) {
    ensure_pair_set();
    ensure(d_init_ids_, cap_init_ids_, init_ids_.size(), false);
    ensure(d_init_records_, cap_init_rec_, init_records_.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_init_ids_, init_ids_.data(), init_ids_.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    NSL_CHECK(hipMemcpyAsync(d_init_records_, init_records_.data(), init_records_.size() * sizeof(Contact), hipMemcpyHostToDevice, stream_));
    launch_init_contacts(d_, d_init_ids_, d_init_records_, (int)init_ids_.size(), stream_);
    init_ids_.clear();
    init_records_.clear();
  }
  if (dt <= 0.0f) return;

  // collide list: colours' contacts then the non-touching contacts (collide_pass.rs:175-190)
  std::vector<int32_t> ids;
  ids.reserve(ncap);
  for (int ci = 0; ci < kColorCount; ++ci) {
    ids.insert(ids.end(), colors_[ci].convex.begin(), colors_[ci].convex.end());
    ids.insert(ids.end(), colors_[ci].overflow.begin(), colors_[ci].overflow.end());
  }
  ids.insert(ids.end(), non_touching_.begin(), non_touching_.end());
  change_host_.assign(ncap, 0);
  if (!ids.empty()) {
    ensure(d_.collide_ids, cap_collide_, ids.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_.collide_ids, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    NSL_CHECK(hipMemsetAsync(d_.contact_change, 0, ncap, stream_));
    launch_collide(d_, (int)ids.size(), stream_, timing_ ? &timer_ : nullptr);
    NSL_CHECK(hipMemcpyAsync(change_host_.data(), d_.contact_change, ncap, hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
  }
  const double t2 = now_ms();
  state_pass();
  const double t3 = now_ms();

  if (d_.n_awake == 0) return;

  // solve lists: colour-major slots (solver_prepare.rs:335-474)
  std::vector<int32_t> solve;
  ColorRanges cr{};
  bool any_restitution = false;
  int colors_used = 0;
  for (int ci = 0; ci < kOverflowColor; ++ci) {
    cr.begin[ci] = (int32_t)solve.size();
    for (int32_t id : colors_[ci].convex) {
      solve.push_back(id);
      any_restitution = any_restitution || hcontacts_[id].restitution > 0.0f;
    }
    cr.end[ci] = (int32_t)solve.size();
    colors_used += cr.end[ci] > cr.begin[ci];
  }
  const std::vector<int32_t>& ovf = colors_[kOverflowColor].overflow;
  for (int32_t id : ovf) any_restitution = any_restitution || hcontacts_[id].restitution > 0.0f;
  cr.begin[kOverflowColor] = cr.end[kOverflowColor] = (int32_t)solve.size();
  if (!solve.empty()) {
    ensure(d_.solve_ids, cap_solve_, solve.size(), false);
    ensure(d_.constraints, cap_constraints_, solve.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_.solve_ids, solve.data(), solve.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  }
  if (!ovf.empty()) {
    ensure(d_.overflow_ids, cap_overflow_ids_, ovf.size(), false);
    ensure(d_.overflow, cap_overflow_, ovf.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_.overflow_ids, ovf.data(), ovf.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  }
  KernelTimer* tm = timing_ ? &timer_ : nullptr;
  SolvePlan plan{};
  plan.cr = cr;
  plan.n_constraints = (int32_t)solve.size();
  plan.n_overflow = (int32_t)ovf.size();
  plan.any_restitution = any_restitution ? 1 : 0;
  plan.restitution_iterations = options.restitution_iterations;
  plan.barrier = d_barrier_;
  plan.phase_log = timing_ ? d_phase_log_ : nullptr;
  int widest = (d_.n_awake + 3) / 4;
  for (int ci = 0; ci < kOverflowColor; ++ci) widest = std::max(widest, cr.end[ci] - cr.begin[ci]);
  int blocks = options.solver_blocks > 0 ? options.solver_blocks : (widest + kSolveBlock - 1) / kSolveBlock;
  blocks = std::max(1, std::min(blocks, std::min(max_blocks_, 128)));
  launch_solve_step(d_, plan, p, blocks, stream_, tm);
  device_newer_ = true;

  // counters[6] (proxies moved by finalize) tells the next pair update whether to run
  NSL_CHECK(hipMemcpyAsync(counters_host_, d_.counters, sizeof(counters_host_), hipMemcpyDeviceToHost, stream_));
  NSL_CHECK(hipMemcpyAsync(&barrier_host_, d_barrier_, sizeof(GridBarrier), hipMemcpyDeviceToHost, stream_));
  if (timing_)
    NSL_CHECK(hipMemcpyAsync(phase_log_host_.data(), d_phase_log_, phase_log_host_.size() * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, stream_));
  NSL_CHECK(hipStreamSynchronize(stream_));
  if (barrier_host_.timed_out) {
    barrier_host_ = GridBarrier{};
    NSL_CHECK(hipMemset(d_barrier_, 0, sizeof(GridBarrier)));
    throw std::runtime_error("step kernel: grid barrier wait exceeded its limit (workgroups not co-resident?)");
  }
  if (timing_) {
    for (int i = 0; i < timer_.used; ++i) {
      float ms = 0.0f;
      NSL_CHECK(hipEventElapsedTime(&ms, timer_.events[2 * i], timer_.events[2 * i + 1]));
      kstats_.ms[timer_.classes[i]] += ms;
      kstats_.launches[timer_.classes[i]] += 1;
    }
    timer_.used = 0;
    // marks: (class of the phase that just ended, wall clock); the first is the kernel start
    int marks = 0;
    for (int i = 0; i < kMaxPhaseLog; ++i, ++marks)
      if (i > 0 && phase_log_host_[2 * i + 1] < phase_log_host_[2 * (i - 1) + 1]) break;
    for (int i = 1; i < marks; ++i) {
      const int cls = (int)phase_log_host_[2 * i];
      if (cls < 0 || cls >= kPhaseClassCount) continue;
      kstats_.phase_ms[cls] += (double)(phase_log_host_[2 * i + 1] - phase_log_host_[2 * (i - 1) + 1]) / wall_clock_khz_;
      kstats_.phase_count[cls] += 1;
    }
  }
  kstats_.steps += 1;
  kstats_.bodies += (int64_t)d_.n_awake;
  kstats_.collided += (int64_t)ids.size();
  kstats_.constraints += (int64_t)solve.size();
  kstats_.fast_bodies = counters_host_[0] - counters_base_[0];
  kstats_.sat_calls = counters_host_[1] - counters_base_[1];
  kstats_.sat_hits = counters_host_[2] - counters_base_[2];
  kstats_.recycled = counters_host_[3] - counters_base_[3];
  kstats_.points = counters_host_[4] - counters_base_[4];
  kstats_.overflow_points = counters_host_[5] - counters_base_[5];
  const double t4 = now_ms();
  kstats_.host_ms[0] += t1 - t0;
  kstats_.host_ms[1] += t2 - t1;
  kstats_.host_ms[2] += t3 - t2;
  kstats_.host_ms[3] += t4 - t3;
  kstats_.solver_blocks = blocks;
  {
    const int n_active = colors_used, ovf_phase = ovf.empty() ? 0 : 1;
    int64_t barriers = 2 + (int64_t)p.sub_steps * (2 + 3 * (n_active + ovf_phase));
    if (any_restitution) barriers += (int64_t)options.restitution_iterations * (n_active + ovf_phase);
    kstats_.barriers += barriers;
  }
  profile_.pairs = (float)(t1 - t0);
  profile_.collide = (float)(t2 - t1);
  profile_.state_pass = (float)(t3 - t2);
  profile_.solve = (float)(t4 - t3);
  profile_.step = (float)(t4 - t0);
  profile_.contacts = contact_count();
  profile_.touching = (int32_t)solve.size() + (int32_t)ovf.size();
  profile_.colors_used = colors_used;
  profile_.overflow = (int32_t)ovf.size();
  profile_.fast_bodies = counters_host_[0];
  profile_.sat_calls = counters_host_[1];
  profile_.sat_hits = counters_host_[2];
  profile_.recycled = counters_host_[3];
}

void GpuWorld::set_kernel_timing(bool on) {
  timing_ = on;
  timer_.used = 0;
}

void GpuWorld::reset_kernel_stats() {
  kstats_ = KernelStats{};
  NSL_CHECK(hipStreamSynchronize(stream_));
  NSL_CHECK(hipMemcpy(counters_base_, d_.counters, sizeof(counters_base_), hipMemcpyDeviceToHost));
}

void GpuWorld::get_bodies(double* pos, float* rot, float* v, float* w) {
  if (device_newer_) pull_bodies();
  for (size_t i = 0; i < bodies_.size(); ++i) {
    const HBody& b = bodies_[i];
    const BodySim& s = b.set == 2 ? awake_host_[b.local] : static_host_[b.local];
    pos[3 * i] = s.transform.p.x;
    pos[3 * i + 1] = s.transform.p.y;
    pos[3 * i + 2] = s.transform.p.z;
    rot[4 * i] = s.transform.q.v.x;
    rot[4 * i + 1] = s.transform.q.v.y;
    rot[4 * i + 2] = s.transform.q.v.z;
    rot[4 * i + 3] = s.transform.q.s;
    Vec3 lv = b.set == 2 ? states_host_[b.local].v : zero3();
    Vec3 av = b.set == 2 ? states_host_[b.local].w : zero3();
    v[3 * i] = lv.x; v[3 * i + 1] = lv.y; v[3 * i + 2] = lv.z;
    w[3 * i] = av.x; w[3 * i + 1] = av.y; w[3 * i + 2] = av.z;
  }
}

// body_queries.rs:432-480 (FNV-1a over transforms and awake velocities)
uint64_t GpuWorld::state_hash() {
  if (device_newer_) pull_bodies();
  const uint64_t prime = 1099511628211ull;
  uint64_t h = 14695981039346656037ull;
  auto mix32 = [&](float f) { h = (h ^ (uint64_t)bits_of(f)) * prime; };
  auto mix64 = [&](double f) { h = (h ^ bits_of(f)) * prime; };
  for (size_t i = 0; i < bodies_.size(); ++i) {
    const HBody& b = bodies_[i];
    const BodySim& s = b.set == 2 ? awake_host_[b.local] : static_host_[b.local];
    mix64(s.transform.p.x);
    mix64(s.transform.p.y);
    mix64(s.transform.p.z);
    mix32(s.transform.q.v.x);
    mix32(s.transform.q.v.y);
    mix32(s.transform.q.v.z);
    mix32(s.transform.q.s);
    if (b.set == 2) {
      const BodyState& st = states_host_[b.local];
      mix32(st.v.x); mix32(st.v.y); mix32(st.v.z);
      mix32(st.w.x); mix32(st.w.y); mix32(st.w.z);
    }
  }
  return h;
}

void GpuWorld::get_contact_colors(int32_t* color, int32_t* local, int32_t* alive) const {
  for (size_t i = 0; i < hcontacts_.size(); ++i) {
    color[i] = hcontacts_[i].color;
    local[i] = hcontacts_[i].local;
    alive[i] = hcontacts_[i].alive;
  }
}

}  // namespace nsl
