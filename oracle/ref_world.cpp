// Oracle: world data model, creation, broad-phase pair update, contact lifecycle,
// constraint-graph coloring, collide pass, finalize and the step driver.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
#include <algorithm>
#include <cstdio>

#include <omp.h>

#include <algorithm>

#include "ref_world.h"

namespace nslref {

// solver_integration.rs:72-87
Softness make_soft(float hertz, float zeta, float h) {
  if (hertz == 0.0f) return Softness{0, 0, 0};
  float omega = 2.0f * kPi * hertz;
  float a1 = 2.0f * zeta + h * omega;
  float a2 = h * omega * a1;
  float a3 = 1.0f / (1.0f + a2);
  return Softness{omega / a1, a2 * a3, a3};
}

World::World() { set_pool.next_index = 3; }  // 0 static, 1 disabled, 2 awake

BodySim& World::sim_of(const Body& b) {
  if (b.set_index == 0) return static_sims[b.local_index];
  if (b.set_index == 2) return awake_sims[b.local_index];
  return sleep_sets[(size_t)b.set_index].body_sims[b.local_index];
}
const BodySim& World::sim_of(const Body& b) const {
  if (b.set_index == 0) return static_sims[b.local_index];
  if (b.set_index == 2) return awake_sims[b.local_index];
  return sleep_sets[(size_t)b.set_index].body_sims[b.local_index];
}
const BodySim& World::body_sim(int32_t id) const { return sim_of(bodies[id]); }
BodyState World::body_state(int32_t id) const {
  const Body& b = bodies[id];
  if (b.set_index == 2) return awake_states[b.local_index];
  BodyState s{};
  s.dq = kQuatIdentity;
  return s;
}

// world_runtime.rs:852-864 (static set 0, awake set 2)
int32_t World::encode_body_sim_index(const Body& b) const {
  if (b.set_index == 2) return b.local_index;
  if (b.set_index == 0) return -(b.local_index + 2);
  return kNullIndex;
}

// world_runtime.rs:1506-1649
int32_t World::create_body(const BodyDef& def) {
  int32_t set_id = def.type == kStatic ? 0 : 2;
  int32_t id = body_pool.alloc();
  BodySim sim{};
  sim.transform.p = def.position;
  sim.transform.q = def.rotation;
  sim.center = def.position;
  sim.rotation0 = sim.transform.q;
  sim.center0 = sim.center;
  sim.min_extent = huge_();
  sim.linear_damping = def.linear_damping;
  sim.angular_damping = def.angular_damping;
  sim.gravity_scale = def.gravity_scale;
  sim.body_id = id;
  sim.flags = def.lock_flags;
  sim.flags |= def.is_bullet ? BF::IS_BULLET : 0;
  sim.flags |= def.allow_fast_rotation ? BF::ALLOW_FAST_ROTATION : 0;
  sim.flags |= def.type == kDynamic ? BF::DYNAMIC_FLAG : 0;
  sim.flags |= def.enable_sleep ? BF::ENABLE_SLEEP : 0;
  sim.flags |= def.enable_contact_recycling ? BF::ENABLE_CONTACT_RECYCLING : 0;
  int32_t local_index;
  if (set_id == 0) {
    static_sims.push_back(sim);
    local_index = (int32_t)static_sims.size() - 1;
  } else {
    awake_sims.push_back(sim);
    local_index = (int32_t)awake_sims.size() - 1;
    BodyState st{};
    st.v = def.linear_velocity;
    st.w = def.angular_velocity;
    st.dq = kQuatIdentity;
    st.flags = sim.flags;
    awake_states.push_back(st);
  }
  if (id == (int32_t)bodies.size()) bodies.push_back(Body{});
  Body& b = bodies[id];
  b = Body{};
  b.set_index = set_id;
  b.local_index = local_index;
  b.id = id;
  b.sleep_threshold = def.sleep_threshold;
  b.safety_factor = def.safety_factor;
  b.type = def.type;
  b.flags = sim.flags;
  // world_runtime.rs:1626-1629: kinematic and dynamic bodies get an island
  if (set_id >= 2) create_island_for_body(set_id, id);
  return id;
}

// world_runtime.rs:1479-1500
void World::sync_body_flags(int32_t body_id) {
  Body& b = bodies[body_id];
  BodySim& sim = sim_of(b);
  sim.flags = (sim.flags & (BF::IS_FAST | BF::HAD_TOI)) | (b.flags & ~BF::TRANSIENT);
  if (b.set_index == 2) awake_states[b.local_index].flags = b.flags & ~BF::TRANSIENT;
}

// shape_runtime.rs:370-408 and the per-type AABB helpers
AABB World::compute_shape_aabb(const Shape& s, Xf t) const {
  switch (s.type) {
    case kHull:
      return aabb_transform(t, hulls[s.hull].aabb);
    case kSphere: {
      Vec3 c = xf_point(t, s.sphere.center);
      float r = s.sphere.radius;
      Vec3 e = {r, r, r};
      return AABB{sub(c, e), add(c, e)};
    }
    case kCapsule: {
      float r = s.capsule.radius;
      Vec3 c1 = xf_point(t, s.capsule.c1), c2 = xf_point(t, s.capsule.c2);
      Vec3 e = {r, r, r};
      return AABB{sub(min3(c1, c2), e), add(max3(c1, c2), e)};
    }
    default:
      return AABB{t.p, t.p};
  }
}
AABB World::compute_fat_shape_aabb(const Shape& s, WXf xf, float extra) const {
  AABB local = aabb_inflate(compute_shape_aabb(s, Xf{kZero3, xf.q}), extra);
  return offset_aabb(local, xf.p);
}
// shape_creation.rs:196-231
float World::compute_shape_margin(const Shape& s) const {
  float margin = 0.0f;
  if (s.type == kSphere) {
    margin = s.sphere.radius;
  } else if (s.type == kCapsule) {
    margin = 0.5f * distance(s.capsule.c2, s.capsule.c1) + s.capsule.radius;
  } else if (s.type == kHull) {
    const Hull& h = hulls[s.hull];
    float mx = 0.0f;
    for (int i = 0; i < h.vertex_count; ++i) mx = maxf(mx, distance_sq(h.points[i], h.center));
    margin = std::sqrt(mx);
  }
  return minf(max_aabb_margin(), kAabbMarginFraction * margin);
}

// shape_creation.rs:286-446
int32_t World::create_shape_internal(int32_t body_id, const ShapeDef& def, Shape shape) {
  int32_t id = shape_pool.alloc();
  if (id == (int32_t)shapes.size()) {
    shapes.push_back(Shape{});
    fat_aabbs.push_back(AABB{});
  }
  Body& b = bodies[body_id];
  shape.id = id;
  shape.body_id = body_id;
  shape.density = def.density;
  shape.friction = def.friction;
  shape.restitution = def.restitution;
  shape.rolling_resistance = def.rolling_resistance;
  shape.filter = def.filter;
  shape.enable_speculative = def.enable_speculative;
  shape.enable_hit_events = def.enable_hit_events;
  shape.is_sensor = def.is_sensor;
  shape.aabb_margin = compute_shape_margin(shape);
  // create_shape_proxy / update_shape_aabbs (shape_creation.rs:237-283)
  WXf xf = sim_of(b).transform;
  const float spec = speculative_distance();
  AABB aabb = compute_fat_shape_aabb(shape, xf, spec);
  shape.aabb = aabb;
  float margin = b.type == kStatic ? spec : shape.aabb_margin;
  fat_aabbs[id] = aabb_inflate(aabb, margin);
  bool force_pair_creation = def.invoke_contact_creation;
  shape.proxy_type = b.type;
  shape.proxy_moved = b.type != kStatic || force_pair_creation;
  shape.prev_shape_id = kNullIndex;
  shape.next_shape_id = b.head_shape_id;
  if (b.head_shape_id != kNullIndex) shapes[b.head_shape_id].prev_shape_id = id;
  b.head_shape_id = id;
  b.shape_count += 1;
  shapes[id] = shape;
  update_body_mass_data(body_id);
  return id;
}

int32_t World::create_hull_shape(int32_t body, const ShapeDef& def, const Hull& hull) {
  Shape s;
  s.type = kHull;
  hulls.push_back(hull);
  s.hull = (int32_t)hulls.size() - 1;
  return create_shape_internal(body, def, s);
}
int32_t World::create_sphere_shape(int32_t body, const ShapeDef& def, Sphere sp) {
  Shape s;
  s.type = kSphere;
  s.sphere = sp;
  return create_shape_internal(body, def, s);
}
// shape_creation.rs:523-557: a capsule shorter than the minimum length becomes a sphere.
int32_t World::create_capsule_shape(int32_t body, const ShapeDef& def, Capsule c) {
  if (distance(c.c1, c.c2) < min_capsule_length()) {
    Sphere sp{scale(0.5f, add(c.c1, c.c2)), c.radius};
    return create_sphere_shape(body, def, sp);
  }
  Shape s;
  s.type = kCapsule;
  s.capsule = c;
  return create_shape_internal(body, def, s);
}

struct MassData { float mass; Vec3 center; Mat3 inertia; };

static Mat3 cylinder_inertia(float mass, float radius, float height) {
  // src/collision/geometry_queries.rs (cylinder about its y axis)
  float rr = radius * radius;
  float hh = height * height;
  float ixx = mass * (3.0f * rr + hh) / 12.0f;
  float iyy = 0.5f * mass * rr;
  return diag3(ixx, iyy, ixx);
}
static Mat3 sphere_inertia(float mass, float radius) {
  float i = 0.4f * mass * radius * radius;
  return diag3(i, i, i);
}
static Quat quat_between_unit_vectors(Vec3 v1, Vec3 v2) {
  // functions.rs:1030-1058
  Quat out{};
  Vec3 m = lerp(v1, v2, 0.5f);
  float tol = 100.0f * FLT_EPSILON;
  if (length_sq(m) > tol * tol) {
    out.v = cross(v1, m);
    out.s = dot(v1, m);
  } else {
    if (absf(v1.x) > 0.5f) out.v = {v1.y, -v1.x, 0.0f};
    else out.v = {0.0f, v1.z, -v1.y};
    out.s = 0.0f;
  }
  return normalize_quat(out);
}

static MassData compute_shape_mass(const World& w, const Shape& s) {
  MassData md{};
  if (s.type == kHull) {
    const Hull& h = w.hulls[s.hull];
    md.mass = s.density * h.volume;
    md.center = h.center;
    md.inertia = mul_sm(s.density, h.central_inertia);
  } else if (s.type == kSphere) {
    // sphere.rs:13-24
    float r = s.sphere.radius;
    float volume = 4.0f / 3.0f * kPi * r * r * r;
    float mass = volume * s.density;
    float ixx = 0.4f * mass * r * r;
    md.mass = mass;
    md.center = s.sphere.center;
    md.inertia = diag3(ixx, ixx, ixx);
  } else if (s.type == kCapsule) {
    // capsule.rs:27-63
    Vec3 c1 = s.capsule.c1, c2 = s.capsule.c2;
    float r = s.capsule.radius;
    float ch = distance(c1, c2);
    float cyl_volume = kPi * r * r * ch;
    float cyl_mass = cyl_volume * s.density;
    float sph_volume = (4.0f / 3.0f) * kPi * r * r * r;
    float sph_mass = sph_volume * s.density;
    Mat3 inertia = add_mm(cylinder_inertia(cyl_mass, r, ch), sphere_inertia(sph_mass, r));
    float st = 0.125f * sph_mass * (3.0f * r + 2.0f * ch) * ch;
    inertia.cx.x += st;
    inertia.cz.z += st;
    Mat3 rot = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (ch * ch > 1000.0f * FLT_MIN) {
      Vec3 dir = normalize(sub(c2, c1));
      Quat q = quat_between_unit_vectors(Vec3{0, 1, 0}, dir);
      rot = matrix_from_quat(q);
    }
    md.mass = sph_mass + cyl_mass;
    md.center = scale(0.5f, add(c1, c2));
    md.inertia = mul_mm(rot, mul_mm(inertia, transpose(rot)));
  }
  return md;
}

static void compute_shape_extent(const World& w, const Shape& s, Vec3 lc, float* min_extent, Vec3* max_extent) {
  if (s.type == kCapsule) {
    float r = s.capsule.radius;
    *min_extent = r;
    Vec3 c1 = sub(s.capsule.c1, lc), c2 = sub(s.capsule.c2, lc);
    *max_extent = add(max3(abs3(c1), abs3(c2)), Vec3{r, r, r});
  } else if (s.type == kSphere) {
    float r = s.sphere.radius;
    *min_extent = r;
    *max_extent = add(abs3(sub(s.sphere.center, lc)), Vec3{r, r, r});
  } else {
    const Hull& h = w.hulls[s.hull];
    *min_extent = h.inner_radius;
    Vec3 mx = kZero3;
    for (int i = 0; i < h.vertex_count; ++i) mx = max3(mx, abs3(sub(h.points[i], lc)));
    *max_extent = mx;
  }
}

// body_api.rs:642-802 and update_isotropic_inertia body_api.rs:611-640
void World::update_body_mass_data(int32_t body_id) {
  Body& b = bodies[body_id];
  b.flags &= ~BF::DIRTY_MASS;
  b.flags &= ~BF::ENABLE_GYROSCOPIC_TORQUE;
  sync_body_flags(body_id);
  b.mass = 0.0f;
  BodySim& sim = sim_of(b);
  sim.inertia = kMat3Zero;
  sim.inv_mass = 0.0f;
  sim.inv_inertia_local = kMat3Zero;
  sim.inv_inertia_world = kMat3Zero;
  sim.local_center = kZero3;
  sim.min_extent = huge_();
  sim.max_extent = kZero3;
  if (b.head_shape_id == kNullIndex) return;
  if (b.type != kDynamic) {
    sim.center = sim.transform.p;
    sim.center0 = sim.center;
    if (b.type == kKinematic) {
      for (int32_t sid = b.head_shape_id; sid != kNullIndex; sid = shapes[sid].next_shape_id) {
        float mn; Vec3 mx;
        compute_shape_extent(*this, shapes[sid], kZero3, &mn, &mx);
        sim.min_extent = minf(sim.min_extent, mn);
        sim.max_extent = max3(sim.max_extent, mx);
      }
    }
    return;
  }
  std::vector<MassData> masses;
  Vec3 lc = kZero3;
  for (int32_t sid = b.head_shape_id; sid != kNullIndex; sid = shapes[sid].next_shape_id) {
    const Shape& s = shapes[sid];
    if (s.density == 0.0f) { masses.push_back(MassData{}); continue; }
    MassData md = compute_shape_mass(*this, s);
    b.mass += md.mass;
    lc = mul_add(lc, md.mass, md.center);
    masses.push_back(md);
  }
  if (b.mass > 0.0f) {
    sim.inv_mass = 1.0f / b.mass;
    lc = scale(sim.inv_mass, lc);
  }
  for (const MassData& md : masses) {
    if (md.mass == 0.0f) continue;
    Vec3 offset = sub(lc, md.center);
    Mat3 inertia = add_mm(md.inertia, steiner(md.mass, offset));
    sim.inertia = add_mm(sim.inertia, inertia);
  }
  float determinant = det(sim.inertia);
  if (determinant > 0.0f) {
    sim.inv_inertia_local = invert_t(sim.inertia);
    Mat3 rm = matrix_from_quat(sim.transform.q);
    sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
  }
  Pos old_center = sim.center;
  sim.local_center = lc;
  sim.center = xf_world_point(sim.transform, sim.local_center);
  sim.center0 = sim.center;
  if (b.set_index == 2) {
    BodyState& st = awake_states[b.local_index];
    Vec3 dl = cross(st.w, sub_pos(sim.center, old_center));
    st.v = add(st.v, dl);
  }
  for (int32_t sid = b.head_shape_id; sid != kNullIndex; sid = shapes[sid].next_shape_id) {
    float mn; Vec3 mx;
    compute_shape_extent(*this, shapes[sid], lc, &mn, &mx);
    sim.min_extent = minf(sim.min_extent, mn);
    sim.max_extent = max3(sim.max_extent, mx);
  }
  if ((sim.flags & BF::FIXED_ROTATION) == BF::FIXED_ROTATION) {
    sim.inertia = kMat3Zero;
    sim.inv_inertia_local = kMat3Zero;
    sim.inv_inertia_world = kMat3Zero;
  }
  // update_isotropic_inertia
  Mat3 in = sim.inertia;
  float mean = (in.cx.x + in.cy.y + in.cz.z) / 3.0f;
  float dx = in.cx.x - mean, dy = in.cy.y - mean, dz = in.cz.z - mean;
  float off = in.cy.x * in.cy.x + in.cz.x * in.cz.x + in.cz.y * in.cz.y;
  float deviation = dx * dx + dy * dy + dz * dz + 2.0f * off;
  float tol = isotropic_inertia_tolerance;
  if (2.0f * deviation > 3.0f * tol * tol * mean * mean) b.flags |= BF::ENABLE_GYROSCOPIC_TORQUE;
  else b.flags &= ~BF::ENABLE_GYROSCOPIC_TORQUE;
  sync_body_flags(body_id);
}

// ---------------------------------------------------------------------------------
// Broad phase. The reference (contact_creation.rs:901-1042) traverses three dynamic
// AABB trees; the resulting key set is: every pair of proxies whose fat AABBs overlap,
// at least one of which is marked moved, where the pair is dynamic-vs-{static,
// kinematic, dynamic}; filtered by same-body, existing pair, shape filter; then sorted.
// The trees only prune that set, so a sweep over the same AABBs yields the same keys.

// hash_set.rs:25-40
static uint64_t shape_pair_key(int32_t s1, int32_t s2, int32_t c) {
  const uint64_t shape_mask = (uint64_t)(kMaxShapes - 1);
  const uint64_t child_mask = (uint64_t)(kMaxChildShapes - 1);
  uint64_t child = (uint64_t)(int64_t)c & child_mask;
  uint64_t a = shape_mask & (uint64_t)(int64_t)s1;
  uint64_t b = shape_mask & (uint64_t)(int64_t)s2;
  if (s1 < s2) return (a << (64 - kShapePower)) | (b << (64 - 2 * kShapePower)) | child;
  return (b << (64 - kShapePower)) | (a << (64 - 2 * kShapePower)) | child;
}

bool World::should_create_pair(int32_t ia, int32_t ib) const {
  const Shape& a = shapes[ia];
  const Shape& b = shapes[ib];
  if (a.body_id == b.body_id) return false;
  if (a.is_sensor || b.is_sensor) return false;  // contact_creation.rs:187
  // shape_runtime.rs:285-292
  if (a.filter.group == b.filter.group && a.filter.group != 0) {
    if (!(a.filter.group > 0)) return false;
  } else if (!((a.filter.mask & b.filter.category) != 0 && (a.filter.category & b.filter.mask) != 0)) {
    return false;
  }
  // contact_creation.rs:110-143: a joint without collide_connected blocks the pair
  return should_bodies_collide(bodies[a.body_id], bodies[b.body_id]);
}

void World::update_broad_phase_pairs() {
  last_pair_keys.clear();
  struct Item { float lo; int32_t shape; };
  std::vector<Item> items;
  bool any_moved = false;
  for (const Shape& s : shapes) {
    if (s.id == kNullIndex || s.proxy_type < 0) continue;
    items.push_back(Item{fat_aabbs[s.id].lo.x, s.id});
    any_moved = any_moved || s.proxy_moved;
  }
  if (!any_moved) return;
  std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.lo < y.lo; });
  // Sweep along x: the later proxies that overlap proxy p along x are the positions after
  // p whose lower x does not exceed p's upper x. Positions are independent, so they are
  // shared out between threads; the keys are sorted afterwards.
  const int n = (int)items.size();
  std::vector<float> los((size_t)n);
  for (int i = 0; i < n; ++i) los[i] = items[i].lo;
  std::vector<std::vector<uint64_t>> found((size_t)omp_get_max_threads());
#pragma omp parallel
  {
    std::vector<uint64_t>& mine = found[(size_t)omp_get_thread_num()];
#pragma omp for schedule(dynamic, 32)
    for (int p = 0; p < n; ++p) {
      const int32_t sp = items[p].shape;
      const AABB& box = fat_aabbs[sp];
      const Shape& sa = shapes[sp];
      const int end = (int)(std::upper_bound(los.begin() + p + 1, los.end(), box.hi.x) - los.begin());
      for (int q = p + 1; q < end; ++q) {
        const int32_t o = items[q].shape;
        const Shape& sb = shapes[o];
        if (!(sa.proxy_moved || sb.proxy_moved)) continue;
        if (sa.proxy_type != kDynamic && sb.proxy_type != kDynamic) continue;
        if (!aabb_overlaps(box, fat_aabbs[o])) continue;
        int32_t lo = mini(sp, o), hi = maxi(sp, o);
        if (sa.body_id == sb.body_id) continue;
        uint64_t key = shape_pair_key(lo, hi, 0);
        if (pair_set.count(key)) continue;
        if (should_create_pair(lo, hi)) mine.push_back(key);
      }
    }
  }
  std::vector<uint64_t> keys;
  for (const std::vector<uint64_t>& f : found) keys.insert(keys.end(), f.begin(), f.end());
  for (Shape& s : shapes) s.proxy_moved = false;
  std::sort(keys.begin(), keys.end());
  for (size_t i = 0; i < keys.size(); ++i) {
    uint64_t key = keys[i];
    if (i > 0 && key == keys[i - 1]) continue;
    int32_t a = (int32_t)((key >> (64 - kShapePower)) & (uint64_t)(kMaxShapes - 1));
    int32_t b = (int32_t)((key >> (64 - 2 * kShapePower)) & (uint64_t)(kMaxShapes - 1));
    int32_t child = (int32_t)(key & (uint64_t)(kMaxChildShapes - 1));
    create_contact(a, b, child);
    last_pair_keys.push_back(key);
  }
  stats.pairs_created += (int64_t)last_pair_keys.size();
}

// contact_creation.rs:53-100: primary order of the supported convex pairs.
static bool register_primary(ShapeType a, ShapeType b) {
  if (a == b) return true;
  if (a == kCapsule && b == kSphere) return true;
  if (a == kHull && (b == kSphere || b == kCapsule)) return true;
  return false;
}

// contact_creation.rs:219-425
void World::create_contact(int32_t ia, int32_t ib, int32_t child) {
  if (!register_primary(shapes[ia].type, shapes[ib].type)) std::swap(ia, ib);
  const Shape& sa = shapes[ia];
  const Shape& sb = shapes[ib];
  const Body& ba = bodies[sa.body_id];
  const Body& bb = bodies[sb.body_id];
  int32_t set_index = (ba.set_index == 2 || bb.set_index == 2) ? 2 : 1;
  int32_t id = contact_pool.alloc();
  if (id == (int32_t)contacts.size()) contacts.push_back(Contact{});
  Contact c{};
  c.contact_id = id;
  c.set_index = set_index;
  c.color_index = kNullIndex;
  c.local_index = set_index == 2 ? (int32_t)awake_contact_indices.size() : (int32_t)disabled_contact_indices.size();
  c.shape_id_a = ia;
  c.shape_id_b = ib;
  c.child_index = child;
  c.encoded_a = encode_body_sim_index(ba);
  c.encoded_b = encode_body_sim_index(bb);
  if (ba.type == kStatic || bb.type == kStatic) c.flags |= CF::STATIC_FLAG;
  if (enable_speculative && sa.enable_speculative && sb.enable_speculative) c.flags |= CF::ENABLE_SPECULATIVE_POINTS;
  if ((ba.flags & BF::ENABLE_CONTACT_RECYCLING) && (bb.flags & BF::ENABLE_CONTACT_RECYCLING) &&
      (c.flags & CF::ENABLE_SPECULATIVE_POINTS))
    c.flags |= CF::RECYCLE;
  // contact_creation.rs:357-384: head of each body's contact list
  for (int e = 0; e < 2; ++e) {
    Body& body = bodies[e == 0 ? sa.body_id : sb.body_id];
    const int32_t key = (id << 1) | e;
    c.edges[e] = ContactEdge{e == 0 ? sa.body_id : sb.body_id, kNullIndex, body.head_contact_key};
    if (body.head_contact_key != kNullIndex)
      contacts[(size_t)(body.head_contact_key >> 1)].edges[body.head_contact_key & 1].prev_key = key;
    body.head_contact_key = key;
    body.contact_count += 1;
  }
  pair_set.insert(shape_pair_key(sa.id, sb.id, child));
  if (set_index == 2) awake_contact_indices.push_back(id);
  else disabled_contact_indices.push_back(id);
  float ra = sa.type == kSphere ? sa.sphere.radius : (sa.type == kCapsule ? sa.capsule.radius : 0.0f);
  float rb = sb.type == kSphere ? sb.sphere.radius : (sb.type == kCapsule ? sb.capsule.radius : 0.0f);
  c.rolling_resistance = maxf(sa.rolling_resistance, sb.rolling_resistance) * maxf(ra, rb);
  c.created = ++contact_serial;
  contacts[id] = c;
}

template <class T>
static int32_t remove_swap(std::vector<T>& v, size_t index) {
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
void World::destroy_contact(int32_t id) {
  Contact& c = contacts[id];
  if (c.contact_id == kNullIndex) return;
  pair_set.erase(shape_pair_key(c.shape_id_a, c.shape_id_b, c.child_index));
  // contact_creation.rs:456-494: out of both bodies' contact lists
  for (int e = 0; e < 2; ++e) {
    const ContactEdge edge = contacts[(size_t)id].edges[e];
    if (edge.prev_key != kNullIndex) contacts[(size_t)(edge.prev_key >> 1)].edges[edge.prev_key & 1].next_key = edge.next_key;
    if (edge.next_key != kNullIndex) contacts[(size_t)(edge.next_key >> 1)].edges[edge.next_key & 1].prev_key = edge.prev_key;
    Body& body = bodies[(size_t)edge.body_id];
    if (body.head_contact_key == ((id << 1) | e)) body.head_contact_key = edge.next_key;
    body.contact_count -= 1;
  }
  c.manifold_count = 0;
  c.manifold.point_count = 0;
  if (c.island_id != kNullIndex) unlink_contact(id);  // contact_creation.rs:501-508
  if (c.color_index == kNullIndex) {
    std::vector<int32_t>& list = contact_list(c.set_index);
    int32_t local = c.local_index;
    int32_t moved = remove_swap(list, (size_t)local);
    if (moved != kNullIndex) contacts[list[local]].local_index = local;
  } else {
    remove_contact_from_graph(c.edges[0].body_id, c.edges[1].body_id, c.color_index, c.local_index,
                              (c.flags & CF::SIM_MESH_CONTACT) != 0);
  }
  Contact& c2 = contacts[id];
  c2.contact_id = kNullIndex;
  c2.set_index = kNullIndex;
  c2.color_index = kNullIndex;
  c2.local_index = kNullIndex;
  contact_pool.free(id);
  stats.contacts_destroyed += 1;
}

// constraint_graph.rs:113-197
void World::add_contact_to_graph(int32_t id) {
  Contact& c = contacts[id];
  int32_t color = kGraphColorCount - 1;  // overflow
  int32_t a = c.edges[0].body_id, b = c.edges[1].body_id;
  const Body& ba = bodies[a];
  const Body& bb = bodies[b];
  const int kDynamicColorCount = kGraphColorCount - 4;
  if (ba.type == kDynamic && bb.type == kDynamic) {
    for (int i = 0; i < kDynamicColorCount; ++i) {
      if (colors[i].body_set.get(a) || colors[i].body_set.get(b)) continue;
      colors[i].body_set.set_grow(a);
      colors[i].body_set.set_grow(b);
      color = i;
      break;
    }
  } else if (ba.type == kDynamic) {
    for (int i = kGraphColorCount - 2; i >= 1; --i) {
      if (!colors[i].body_set.get(a)) { colors[i].body_set.set_grow(a); color = i; break; }
    }
  } else if (bb.type == kDynamic) {
    for (int i = kGraphColorCount - 2; i >= 1; --i) {
      if (!colors[i].body_set.get(b)) { colors[i].body_set.set_grow(b); color = i; break; }
    }
  }
  bool is_scalar = (c.flags & CF::SIM_MESH_CONTACT) != 0 || color == kGraphColorCount - 1;
  c.color_index = color;
  c.local_index = is_scalar ? (int32_t)colors[color].contacts.size() : (int32_t)colors[color].convex_contacts.size();
  c.encoded_a = encode_body_sim_index(ba);
  c.encoded_b = encode_body_sim_index(bb);
  if (is_scalar) colors[color].contacts.push_back(ContactSpec{c.contact_id, 0, (uint16_t)c.manifold_count});
  else colors[color].convex_contacts.push_back(c.contact_id);
}

// constraint_graph.rs:200-251
void World::remove_contact_from_graph(int32_t a, int32_t b, int32_t color, int32_t local, bool mesh) {
  if (color != kGraphColorCount - 1) {
    colors[color].body_set.clear(a);
    colors[color].body_set.clear(b);
  }
  if (mesh || color == kGraphColorCount - 1) {
    int32_t moved = remove_swap(colors[color].contacts, (size_t)local);
    if (moved != kNullIndex) contacts[colors[color].contacts[local].contact_id].local_index = local;
  } else {
    int32_t moved = remove_swap(colors[color].convex_contacts, (size_t)local);
    if (moved != kNullIndex) contacts[colors[color].convex_contacts[local]].local_index = local;
  }
}

// collide_pass.rs:743-774
void World::add_non_touching_contact(int32_t id) {
  Contact& c = contacts[id];
  c.color_index = kNullIndex;
  c.local_index = (int32_t)awake_contact_indices.size();
  awake_contact_indices.push_back(id);
}
void World::remove_non_touching_contact(int32_t local) {
  int32_t moved = remove_swap(awake_contact_indices, (size_t)local);
  if (moved != kNullIndex) contacts[awake_contact_indices[local]].local_index = local;
}

// ---------------------------------------------------------------------------------
// Collide pass (collide_pass.rs:164-742, contact_update.rs:43-468)

bool World::compute_convex_manifold(Contact& c, const Shape& sa, WXf xf_a, const Shape& sb, WXf xf_b) {
  LocalManifold lm{};
  lm.point_count = 0;
  Xf b_to_a = inv_mul_wxf(xf_a, xf_b);
  if (sa.type == kSphere) {
    collide_spheres(lm, sa.sphere, sb.sphere, b_to_a);
  } else if (sa.type == kCapsule) {
    if (sb.type == kSphere) collide_capsule_and_sphere(lm, sa.capsule, sb.sphere, b_to_a);
    else collide_capsules(lm, sa.capsule, sb.capsule, b_to_a);
  } else {
    const Hull& ha = hulls[sa.hull];
    if (sb.type == kSphere) collide_hull_and_sphere(lm, ha, sb.sphere, b_to_a, c.simplex_cache);
    else if (sb.type == kCapsule) collide_hull_and_capsule(lm, ha, sb.capsule, b_to_a, c.simplex_cache);
    else collide_hulls(lm, ha, hulls[sb.hull], b_to_a, c.sat_cache, &stats.sat_calls, &stats.sat_hits);
  }
  if (lm.point_count == 0) {
    if (c.manifold_count > 0) {
      c.manifold.point_count = 0;
      c.manifold_count = 0;
    }
    return false;
  }
  ManifoldPoint old_points[4];
  int old_count = 0;
  if (c.manifold_count == 0) {
    c.manifold = Manifold{};
    c.manifold_count = 1;
  } else {
    old_count = c.manifold.point_count;
    for (int i = 0; i < old_count; ++i) old_points[i] = c.manifold.points[i];
  }
  Manifold& m = c.manifold;
  m.point_count = lm.point_count;
  Mat3 ma = matrix_from_quat(xf_a.q);
  m.normal = mul_mv(ma, lm.normal);
  for (int i = 0; i < lm.point_count; ++i) {
    const LocalPoint& src = lm.pts[i];
    ManifoldPoint t{};
    t.anchor_a = mul_mv(ma, src.point);
    t.anchor_b = add(t.anchor_a, sub_pos(xf_a.p, xf_b.p));
    t.separation = src.separation;
    t.feature_id = make_feature_id(src.pair);
    t.triangle_index = kNullIndex;
    t.normal_velocity = 0.0f;
    m.points[i] = t;
  }
  for (int i = 0; i < lm.point_count; ++i) {
    ManifoldPoint p2 = m.points[i];
    p2.total_normal_impulse = 0.0f;
    p2.persisted = false;
    for (int j = 0; j < old_count; ++j) {
      if (p2.feature_id == old_points[j].feature_id) {
        p2.normal_impulse = old_points[j].normal_impulse;
        p2.persisted = true;
        old_points[j].feature_id = 0xffffffffu;
        break;
      }
    }
    if (!p2.persisted) p2.normal_impulse = 0.0f;
    m.points[i] = p2;
  }
  return true;
}

bool World::update_contact(int32_t id, Vec3 lc_a, WXf xf_a, Vec3 lc_b, WXf xf_b) {
  Contact& c = contacts[id];
  const Shape& sa = shapes[c.shape_id_a];
  const Shape& sb = shapes[c.shape_id_b];
  bool touching = compute_convex_manifold(c, sa, xf_a, sb, xf_b);
  if (touching) {
    // update_convex_contact (contact_update.rs:243-360)
    c.friction = std::sqrt(sa.friction * sb.friction);
    c.restitution = maxf(sa.restitution, sb.restitution);
    if (sa.rolling_resistance > 0.0f || sb.rolling_resistance > 0.0f) {
      auto radius = [&](const Shape& s) -> float {
        if (s.type == kSphere) return s.sphere.radius;
        if (s.type == kCapsule) return s.capsule.radius;
        if (s.type == kHull) return 0.25f * hulls[s.hull].inner_radius;
        return 0.0f;
      };
      c.rolling_resistance = maxf(sa.rolling_resistance, sb.rolling_resistance) * maxf(radius(sa), radius(sb));
    } else {
      c.rolling_resistance = 0.0f;
    }
    c.tangent_velocity = sub(rotate(xf_a.q, sa.tangent_velocity), rotate(xf_b.q, sb.tangent_velocity));
    if (sa.enable_hit_events || sb.enable_hit_events) c.flags |= CF::SIM_ENABLE_HIT_EVENT;
    else c.flags &= ~CF::SIM_ENABLE_HIT_EVENT;
    Vec3 ca = rotate(xf_a.q, lc_a), cb = rotate(xf_b.q, lc_b);
    for (int j = 0; j < c.manifold.point_count; ++j) {
      c.manifold.points[j].anchor_a = sub(c.manifold.points[j].anchor_a, ca);
      c.manifold.points[j].anchor_b = sub(c.manifold.points[j].anchor_b, cb);
    }
    c.flags |= CF::SIM_TOUCHING;
  } else {
    c.flags &= ~CF::SIM_TOUCHING;
  }
  return touching;
}

void World::collide_one_contact(int32_t id, std::vector<uint8_t>& state_bits) {
  Contact& c = contacts[id];
  if (!aabb_overlaps(fat_aabbs[c.shape_id_a], fat_aabbs[c.shape_id_b])) {
    c.flags |= CF::SIM_DISJOINT;
    c.flags &= ~CF::SIM_TOUCHING;
    state_bits[id] = 1;
    return;
  }
  // collide_pass.rs:105-154: an index of kNullIndex (a sleeping body) is found through the body
  auto resolve = [&](int32_t enc, int32_t body_id, bool* is_static) -> const BodySim& {
    *is_static = enc < kNullIndex;
    if (enc >= 0) return awake_sims[enc];
    if (enc < kNullIndex) return static_sims[-enc - 2];
    return sim_of(bodies[body_id]);
  };
  bool st_a, st_b;
  const BodySim& sim_a = resolve(c.encoded_a, c.edges[0].body_id, &st_a);
  const BodySim& sim_b = resolve(c.encoded_b, c.edges[1].body_id, &st_b);
  bool was_touching = (c.flags & CF::SIM_TOUCHING) != 0;
  WXf xf_a = sim_a.transform, xf_b = sim_b.transform;
  bool is_fast = ((sim_a.flags | sim_b.flags) & BF::IS_FAST) != 0;
  float recycle_distance = contact_recycle_distance_;
  float spec = speculative_distance();
  float recycle_non_touching = minf(recycle_distance, spec);
  float tolerance = was_touching ? recycle_distance : recycle_non_touching;
  if (!is_fast && recycle_distance > 0.0f && (c.flags & CF::RELATIVE_TRANSFORM_VALID) && (c.flags & CF::RECYCLE)) {
    float angle_a = dot_quat(xf_a.q, c.cached_rotation_a);
    float angle_b = dot_quat(xf_b.q, c.cached_rotation_b);
    float angular_distance = minf(angle_a * angle_a, angle_b * angle_b);
    Xf xf = inv_mul_wxf(xf_a, xf_b);
    Xf xfc = c.cached_relative_pose;
    Vec3 me_a = st_a ? kZero3 : sim_a.max_extent;
    Vec3 me_b = st_b ? kZero3 : sim_b.max_extent;
    Vec3 max_extent = max3(me_a, me_b);
    float dist_sq = distance_sq(xf.p, xfc.p);
    if (angular_distance > kContactRecycleAngularDistance && dist_sq < tolerance * tolerance) {
      float dist = std::sqrt(dist_sq);
      float slack = tolerance - dist;
      Quat qr = inv_mul_quat(xfc.q, xf.q);
      Vec3 arc = modified_cross(abs3(qr.v), max_extent);
      float arc_sq = 4.0f * length_sq(arc);
      if (arc_sq < slack * slack) {
        Quat dq_a = mul_quat(xf_a.q, conjugate(c.cached_rotation_a));
        Quat dq_b = mul_quat(xf_b.q, conjugate(c.cached_rotation_b));
        Mat3 m_a = matrix_from_quat(dq_a), m_b = matrix_from_quat(dq_b);
        Vec3 dc = sub_pos(sim_b.center, sim_a.center);
        if (c.manifold_count > 0) {
          Manifold& m = c.manifold;
          for (int i = 0; i < m.point_count; ++i) {
            Vec3 r_a = mul_mv(m_a, m.points[i].anchor_a);
            Vec3 r_b = mul_mv(m_b, m.points[i].anchor_b);
            Vec3 dp = add(dc, sub(r_b, r_a));
            m.points[i].separation = m.points[i].base_separation + dot(dp, m.normal);
            m.points[i].normal_velocity = 0.0f;
            m.points[i].persisted = true;
          }
        }
        __atomic_fetch_add(&stats.recycled, 1, __ATOMIC_RELAXED);
        return;
      }
    }
  }
  c.cached_rotation_a = xf_a.q;
  c.cached_rotation_b = xf_b.q;
  c.cached_relative_pose = inv_mul_wxf(xf_a, xf_b);
  c.flags |= CF::RELATIVE_TRANSFORM_VALID;

  bool touching = update_contact(id, sim_a.local_center, xf_a, sim_b.local_center, xf_b);
  Contact& c2 = contacts[id];
  if (touching && !was_touching) {
    c2.flags |= CF::SIM_STARTED_TOUCHING;
    state_bits[id] = 1;
  } else if (!touching && was_touching) {
    c2.flags |= CF::SIM_STOPPED_TOUCHING;
    state_bits[id] = 1;
  }
  if (touching) {
    for (int i = 0; i < c2.manifold.point_count; ++i) c2.manifold.points[i].base_separation = c2.manifold.points[i].separation;
  }
}

void World::collide() {
  std::vector<int32_t> ids;
  for (int i = 0; i < kGraphColorCount; ++i) {
    ids.insert(ids.end(), colors[i].convex_contacts.begin(), colors[i].convex_contacts.end());
    for (const ContactSpec& s : colors[i].contacts) ids.push_back(s.contact_id);
  }
  ids.insert(ids.end(), awake_contact_indices.begin(), awake_contact_indices.end());
  if (ids.empty()) return;
  std::vector<uint8_t> state_bits(std::max<size_t>(contacts.size(), 8), 0);
  // each item reads shared state and writes only its own contact and state byte
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t k = 0; k < ids.size(); ++k) collide_one_contact(ids[k], state_bits);
  // collide_pass.rs:386-478: state transitions in contact-id order
  for (size_t id = 0; id < state_bits.size(); ++id) {
    if (!state_bits[id]) continue;
    uint32_t flags = contacts[id].flags;
    if (flags == 0) continue;
    if (flags & CF::SIM_DISJOINT) {
      destroy_contact((int32_t)id);
    } else if (flags & CF::SIM_STARTED_TOUCHING) {
      Contact& c = contacts[id];
      c.flags &= ~CF::SIM_STARTED_TOUCHING;
      c.flags |= CF::TOUCHING;
      // link first: this wakes a sleeping partner so the body sims are in place
      link_contact((int32_t)id);
      int32_t old_local = contacts[id].local_index;
      add_contact_to_graph((int32_t)id);
      remove_non_touching_contact(old_local);
      stats.begin_touch += 1;
    } else if (flags & CF::SIM_STOPPED_TOUCHING) {
      Contact& c = contacts[id];
      c.flags &= ~CF::SIM_STOPPED_TOUCHING;
      c.flags &= ~CF::TOUCHING;
      int32_t color = c.color_index, local = c.local_index;
      unlink_contact((int32_t)id);
      add_non_touching_contact((int32_t)id);
      remove_contact_from_graph(c.edges[0].body_id, c.edges[1].body_id, color, local, (c.flags & CF::SIM_MESH_CONTACT) != 0);
      stats.end_touch += 1;
    }
  }
}

// ---------------------------------------------------------------------------------
// Finalize (solver_integration.rs:435-799)
void World::finalize_bodies(const StepContext& ctx) {
  const float spec = speculative_distance();
  const float time_step = ctx.dt, inv_time_step = ctx.inv_dt;
  std::vector<uint8_t> moved_flags(awake_sims.size(), 0);
  // 1: fast body for the continuous pass in body order, 2: bullet (solver_integration.rs:696-710)
  std::vector<uint8_t> continuous(awake_sims.size(), 0);
  // each item writes its own body, its own shapes and fat AABBs
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < awake_sims.size(); ++i) {
    BodyState& state = awake_states[i];
    BodySim& sim = awake_sims[i];
    Vec3 v = state.v, w = state.w;
    Vec3 local_omega = inv_rotate(sim.transform.q, w);
    Vec3 local_dr = inv_rotate(sim.transform.q, state.dq.v);
    sim.center = offset_pos(sim.center, state.dp);
    sim.transform.q = normalize_quat(mul_quat(state.dq, sim.transform.q));
    Vec3 velocity_arc = modified_cross(abs3(local_omega), sim.max_extent);
    float max_velocity = length(v) + length(velocity_arc);
    Vec3 rotation_arc = modified_cross(abs3(local_dr), sim.max_extent);
    float max_delta_position = length(state.dp) + 2.0f * length(rotation_arc);
    const float position_sleep_factor = 0.5f;
    float sleep_velocity = maxf(max_velocity, position_sleep_factor * inv_time_step * max_delta_position);
    state.dp = kZero3;
    state.dq = kQuatIdentity;
    sim.transform.p = offset_pos(sim.center, neg(rotate(sim.transform.q, sim.local_center)));
    Body& body = bodies[sim.body_id];
    body.sleep_velocity = sleep_velocity;
    sim.force = kZero3;
    sim.torque = kZero3;
    body.flags &= ~BF::TRANSIENT;
    sim.flags &= ~(BF::IS_FAST | BF::TRANSIENT);
    body.flags |= state.flags & BF::IS_SPEED_CAPPED;
    state.flags &= ~BF::TRANSIENT;
    if (!enable_sleep || (body.flags & BF::ENABLE_SLEEP) == 0 || sleep_velocity > body.sleep_threshold) {
      body.sleep_time = 0.0f;
      float max_motion = maxf(max_delta_position, max_velocity * time_step);
      if (body.type == kDynamic && enable_continuous && max_motion > body.safety_factor * sim.min_extent) {
        sim.flags |= BF::IS_FAST;
        continuous[i] = (sim.flags & BF::IS_BULLET) ? 2 : 1;
        __atomic_fetch_add(&stats.fast_bodies, 1, __ATOMIC_RELAXED);
      } else {
        sim.center0 = sim.center;
        sim.rotation0 = sim.transform.q;
      }
    } else {
      sim.center0 = sim.center;
      sim.rotation0 = sim.transform.q;
      body.sleep_time += time_step;
    }
    Mat3 rm = matrix_from_quat(sim.transform.q);
    sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
    if ((sim.flags & BF::IS_FAST) == 0) {
      for (int32_t sid = body.head_shape_id; sid != kNullIndex; sid = shapes[sid].next_shape_id) {
        Shape& s = shapes[sid];
        AABB aabb = compute_fat_shape_aabb(s, sim.transform, spec);
        s.aabb = aabb;
        if (!aabb_contains(fat_aabbs[sid], aabb)) {
          fat_aabbs[sid] = aabb_inflate(aabb, s.aabb_margin);
          moved_flags[i] = 1;
          s.proxy_moved = true;  // mark_proxy_moved in the serial tail
        }
      }
    }
  }
  // The serial tail in body order: fast bodies sweep against the static shapes, and the
  // world inverse inertia follows a rotation moved back to the impact
  // (solver_integration.rs:538-566). Bullets run afterwards and see the finished bodies
  // (solver_integration.rs:948-995).
  for (size_t i = 0; i < awake_sims.size(); ++i) {
    if (continuous[i] != 1) continue;
    solve_continuous((int32_t)i, time_step);
    BodySim& sim = awake_sims[i];
    Mat3 rm = matrix_from_quat(sim.transform.q);
    sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
  }
  for (size_t i = 0; i < awake_sims.size(); ++i)
    if (continuous[i] == 2) solve_continuous((int32_t)i, time_step);
}

// The island part of finalize (solver_integration.rs:735-752, merged over threads at
// :522-538): a body still short of the time to sleep keeps its island awake; a sleepy body
// in an island with removed constraints nominates it for splitting, the sleepiest (then the
// largest island id) winning.
void World::mark_awake_islands() {
  awake_island_bits.assign(awake_islands.size(), 0);
  split_candidate = kNullIndex;
  split_candidate_time = 0.0f;
  for (const BodySim& sim : awake_sims) {
    const Body& body = bodies[(size_t)sim.body_id];
    const Island& island = islands[(size_t)body.island_id];
    if (body.sleep_time < 0.5f) {
      awake_island_bits[(size_t)island.local_index] = 1;
    } else if (island.constraint_remove_count > 0) {
      if (body.sleep_time > split_candidate_time ||
          (body.sleep_time == split_candidate_time && body.island_id > split_candidate)) {
        split_candidate = body.island_id;
        split_candidate_time = body.sleep_time;
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// Step driver (solver_integration.rs:807-1015)
void World::step(float time_step, int sub_step_count) {
  double t0 = omp_get_wtime();
  update_broad_phase_pairs();
  double t1 = omp_get_wtime();
  stats.ms_pairs += 1e3 * (t1 - t0);
  StepContext ctx;
  ctx.dt = time_step;
  ctx.sub_step_count = maxi(1, sub_step_count);
  if (time_step > 0.0f) {
    ctx.inv_dt = 1.0f / time_step;
    ctx.h = time_step / (float)ctx.sub_step_count;
    ctx.inv_h = (float)ctx.sub_step_count * ctx.inv_dt;
  }
  float contact_hz = minf(contact_hertz, 0.125f * ctx.inv_h);
  ctx.contact_softness = make_soft(contact_hz, contact_damping_ratio, ctx.h);
  ctx.static_softness = make_soft(2.0f * contact_hz, 0.5f * contact_damping_ratio, ctx.h);
  ctx.max_linear_velocity = max_linear_speed;
  ctx.enable_warm_starting = enable_warm_starting;
  if (time_step > 0.0f) collide();
  double t2 = omp_get_wtime();
  stats.ms_collide += 1e3 * (t2 - t1);
  if (time_step > 0.0f) {
    if (awake_sims.empty()) return;
    // a split nominated by the last step (solver_integration.rs:879-884)
    if (split_island_id != kNullIndex) {
      split_island(split_island_id);
      split_island_id = kNullIndex;
    }
    prepare_joints(ctx);
    prepare_contacts(ctx);
    if (after_prepare) after_prepare(*this, ctx);
    double t3 = omp_get_wtime();
    solve_stages(ctx);
    if (residuals_on) colour_residuals(ctx);
    double t4 = omp_get_wtime();
    finalize_bodies(ctx);
    mark_awake_islands();
    if (enable_sleep) process_sleep();  // solver_integration.rs:1003-1005
    double t5 = omp_get_wtime();
    stats.ms_prepare += 1e3 * (t3 - t2);
    stats.ms_solve += 1e3 * (t4 - t3);
    stats.ms_finalize += 1e3 * (t5 - t4);
  }
}

// body_queries.rs:432-480
uint64_t World::state_hash() const {
  const uint64_t prime = 1099511628211ull;
  uint64_t hash = 14695981039346656037ull;
  auto mix32 = [&](float v) { hash = (hash ^ (uint64_t)f2u(v)) * prime; };
  auto mix64 = [&](double v) { hash = (hash ^ d2u(v)) * prime; };
  for (size_t i = 0; i < bodies.size(); ++i) {
    const Body& b = bodies[i];
    if (b.id != (int32_t)i) continue;
    const BodySim& sim = sim_of(b);
    mix64(sim.transform.p.x);
    mix64(sim.transform.p.y);
    mix64(sim.transform.p.z);
    mix32(sim.transform.q.v.x);
    mix32(sim.transform.q.v.y);
    mix32(sim.transform.q.v.z);
    mix32(sim.transform.q.s);
    if (b.set_index == 2) {
      const BodyState& st = awake_states[b.local_index];
      mix32(st.v.x);
      mix32(st.v.y);
      mix32(st.v.z);
      mix32(st.w.x);
      mix32(st.w.y);
      mix32(st.w.z);
    }
  }
  return hash;
}

}  // namespace nslref
