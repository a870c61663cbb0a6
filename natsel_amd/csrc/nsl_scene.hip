// Host scene records shared by the single-world and batched steppers (see nsl_scene.h).
#include "nsl_scene.h"

#include <vector>

namespace nsl {

namespace {

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

// ---------------------------------------------------------------------------------
// creation (world_runtime.rs:1506-1649, shape_creation.rs:286-559, body_api.rs:611-802)

int32_t HostScene::create_body(const BodyDesc& desc) {
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
  return id;
}

Aabb HostScene::shape_local_aabb(const Shape& s, Quat q) const {
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

int32_t HostScene::add_shape(int32_t body, Shape shape, const Material& m, float radius) {
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
  hs.sensor = m.is_sensor;
  hs.aabb = aabb;
  hs.fat = aabb_inflate(aabb, b.type == kBodyStatic ? kSpeculativeDistance : shape.aabb_margin);
  hs.proxy_moved = true;  // dynamic proxies always; static ones via invoke_contact_creation
  shape.next_shape = b.head_shape;
  b.head_shape = id;
  sim.head_shape = id;
  shapes_.push_back(shape);
  hshapes_.push_back(hs);
  update_mass(body);
  return id;
}

int32_t HostScene::create_box(int32_t body, Vec3 half, const Material& m) {
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

int32_t HostScene::create_sphere(int32_t body, Vec3 center, float radius, const Material& m) {
  Shape s{};
  s.kind = kShapeSphere;
  s.hull = -1;
  s.c1 = center;
  s.c2 = center;
  s.radius = radius;
  return add_shape(body, s, m, radius);
}

int32_t HostScene::create_capsule(int32_t body, Vec3 c1, Vec3 c2, float radius, const Material& m) {
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
void HostScene::update_mass(int32_t body) {
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
  float tol = isotropic_inertia_tolerance;
  if (2.0f * deviation > 3.0f * tol * tol * mean * mean) b.flags |= bflag::kGyroscopic;
  else b.flags &= ~bflag::kGyroscopic;
  sim.flags = (sim.flags & (bflag::kIsFast | bflag::kHadToi)) | (b.flags & ~bflag::kTransient);
  sim.body_flags = b.flags;
  st.flags = b.flags & ~bflag::kTransient;
}

void HostScene::get_bodies(double* pos, float* rot, float* v, float* w) const {
  for (size_t i = 0; i < bodies_.size(); ++i) {
    const HBody& b = bodies_[i];
    const BodySim& s = sim_of((int32_t)i);
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
uint64_t HostScene::state_hash() const {
  const uint64_t prime = 1099511628211ull;
  uint64_t h = 14695981039346656037ull;
  auto mix32 = [&](float f) { h = (h ^ (uint64_t)bits_of(f)) * prime; };
  auto mix64 = [&](double f) { h = (h ^ bits_of(f)) * prime; };
  for (size_t i = 0; i < bodies_.size(); ++i) {
    const HBody& b = bodies_[i];
    const BodySim& s = sim_of((int32_t)i);
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

}  // namespace nsl
