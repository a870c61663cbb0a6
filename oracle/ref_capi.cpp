// Oracle: C entry points used by the Python checker (ctypes).
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
#include <cstdint>
#include <cstring>

#include <omp.h>

#include "ref_world.h"

using namespace nslref;

extern "C" {

// Threads for the OpenMP loops (results do not depend on it).
void nslo_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int nslo_max_threads(void) { return omp_get_max_threads(); }

void* nslo_create(void) { return new World(); }
void nslo_destroy(void* w) { delete static_cast<World*>(w); }

void nslo_set_options(void* wp, int enable_sleep, int enable_continuous, int enable_warm_starting, int enable_speculative) {
  World* w = static_cast<World*>(wp);
  w->enable_sleep = enable_sleep != 0;
  w->enable_continuous = enable_continuous != 0;
  w->enable_warm_starting = enable_warm_starting != 0;
  w->enable_speculative = enable_speculative != 0;
}

void nslo_set_gravity(void* wp, float x, float y, float z) { static_cast<World*>(wp)->gravity = Vec3{x, y, z}; }
void nslo_set_restitution(void* wp, int iterations, float threshold, int propagate) {
  World* w = static_cast<World*>(wp);
  w->restitution_iterations = iterations;
  w->restitution_threshold = threshold;
  w->enable_restitution_propagation = propagate != 0;
}

int nslo_create_body(void* wp, int type, const double* pos, const float* rot, const float* v, const float* av,
                     int enable_sleep, float linear_damping, float angular_damping, float gravity_scale,
                     int is_bullet, float safety_factor, float sleep_threshold, int lock_flags, int allow_fast_rotation,
                     int enable_contact_recycling) {
  BodyDef d;
  d.is_bullet = is_bullet != 0;
  d.safety_factor = safety_factor;
  d.sleep_threshold = sleep_threshold;
  d.lock_flags = (uint32_t)lock_flags & BF::FIXED_ROTATION_AND_LINEAR;
  d.allow_fast_rotation = allow_fast_rotation != 0;
  d.enable_contact_recycling = enable_contact_recycling != 0;
  d.type = (BodyType)type;
  d.position = Pos{pos[0], pos[1], pos[2]};
  d.rotation = Quat{{rot[0], rot[1], rot[2]}, rot[3]};
  d.linear_velocity = Vec3{v[0], v[1], v[2]};
  d.angular_velocity = Vec3{av[0], av[1], av[2]};
  d.enable_sleep = enable_sleep != 0;
  d.linear_damping = linear_damping;
  d.angular_damping = angular_damping;
  d.gravity_scale = gravity_scale;
  return static_cast<World*>(wp)->create_body(d);
}

// ShapeDef (shape_def.rs:90-135): mat = density, friction, restitution, rolling resistance;
// opt = filter category, mask and group, is_sensor, enable_speculative_contact, enable_hit_events.
static ShapeDef shape_def(const float* mat, const int64_t* opt) {
  ShapeDef s;
  s.density = mat[0];
  s.friction = mat[1];
  s.restitution = mat[2];
  s.rolling_resistance = mat[3];
  s.filter.category = (uint64_t)opt[0];
  s.filter.mask = (uint64_t)opt[1];
  s.filter.group = (int32_t)opt[2];
  s.is_sensor = opt[3] != 0;
  s.enable_speculative = opt[4] != 0;
  s.enable_hit_events = opt[5] != 0;
  return s;
}

int nslo_create_box(void* wp, int body, float hx, float hy, float hz, const float* mat, const int64_t* opt) {
  return static_cast<World*>(wp)->create_hull_shape(body, shape_def(mat, opt), make_box_hull(hx, hy, hz));
}
int nslo_create_sphere(void* wp, int body, const float* c, float r, const float* mat, const int64_t* opt) {
  return static_cast<World*>(wp)->create_sphere_shape(body, shape_def(mat, opt), Sphere{{c[0], c[1], c[2]}, r});
}
int nslo_create_capsule(void* wp, int body, const float* c1, const float* c2, float r, const float* mat, const int64_t* opt) {
  return static_cast<World*>(wp)->create_capsule_shape(body, shape_def(mat, opt),
                                                       Capsule{{c1[0], c1[1], c1[2]}, {c2[0], c2[1], c2[2]}, r});
}

// Revolute joint (types/joint_defs.rs:30-58, :154-167): frames are (px, py, pz, qx, qy, qz, qs).
int nslo_create_revolute_joint(void* wp, int body_a, int body_b, const float* fa, const float* fb, float constraint_hertz,
                               float constraint_damping_ratio, int collide_connected, int enable_spring, float hertz,
                               float damping_ratio, float target_angle, int enable_limit, float lower_angle,
                               float upper_angle, int enable_motor, float max_motor_torque, float motor_speed) {
  RevoluteJointDef d;
  d.body_a = body_a;
  d.body_b = body_b;
  d.local_frame_a = Xf{{fa[0], fa[1], fa[2]}, {{fa[3], fa[4], fa[5]}, fa[6]}};
  d.local_frame_b = Xf{{fb[0], fb[1], fb[2]}, {{fb[3], fb[4], fb[5]}, fb[6]}};
  d.constraint_hertz = constraint_hertz;
  d.constraint_damping_ratio = constraint_damping_ratio;
  d.collide_connected = collide_connected != 0;
  d.enable_spring = enable_spring != 0;
  d.hertz = hertz;
  d.damping_ratio = damping_ratio;
  d.target_angle = target_angle;
  d.enable_limit = enable_limit != 0;
  d.lower_angle = lower_angle;
  d.upper_angle = upper_angle;
  d.enable_motor = enable_motor != 0;
  d.max_motor_torque = max_motor_torque;
  d.motor_speed = motor_speed;
  return static_cast<World*>(wp)->create_revolute_joint(d);
}

// Accumulated impulses of a joint: linear (3), perpendicular (2), spring, motor, lower,
// upper; returns its graph colour (-1 when it is not in the graph).
int nslo_get_joint(void* wp, int joint_id, float* out) {
  World* w = static_cast<World*>(wp);
  const JointSim& j = w->joint_sim(joint_id);
  const float vals[9] = {j.linear_impulse.x, j.linear_impulse.y, j.linear_impulse.z, j.perp_impulse.x, j.perp_impulse.y,
                         j.spring_impulse, j.motor_impulse, j.lower_impulse, j.upper_impulse};
  std::memcpy(out, vals, sizeof(vals));
  return w->joints[joint_id].color_index;
}

// World::save_state / restore_state (world_state.rs:340, :439): the oracle's World is plain
// values, so a snapshot is a copy of it.
void* nslo_save_state(void* wp) { return new World(*static_cast<World*>(wp)); }
void nslo_restore_state(void* wp, void* sp) { *static_cast<World*>(wp) = *static_cast<World*>(sp); }

void nslo_step(void* wp, float dt, int substeps) { static_cast<World*>(wp)->step(dt, substeps); }

// The per-colour residual sweep (World::colour_residuals) after every step while on; the
// last step's results per colour 0..22: sum (double), largest normal and friction changes,
// constraints.
void nslo_set_residuals(void* wp, int on) { static_cast<World*>(wp)->residuals_on = on != 0; }
void nslo_get_residuals(void* wp, double* sum, float* max_normal, float* max_friction, int32_t* constraints) {
  const World* w = static_cast<World*>(wp);
  for (int ci = 0; ci < 23; ++ci) {
    sum[ci] = w->residuals[ci].sum_normal;
    max_normal[ci] = w->residuals[ci].max_normal;
    max_friction[ci] = w->residuals[ci].max_friction;
    constraints[ci] = w->residuals[ci].constraints;
  }
}
uint64_t nslo_state_hash(void* wp) { return static_cast<World*>(wp)->state_hash(); }
int nslo_body_capacity(void* wp) { return (int)static_cast<World*>(wp)->bodies.size(); }
int nslo_contact_count(void* wp) { return static_cast<World*>(wp)->contact_count(); }

// Per body id: origin position (f64 x3), rotation (x,y,z,s), linear and angular velocity.
void nslo_get_bodies(void* wp, double* pos, float* rot, float* v, float* av) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    const BodySim& s = w->body_sim((int32_t)i);
    BodyState st = w->body_state((int32_t)i);
    pos[3 * i + 0] = s.transform.p.x;
    pos[3 * i + 1] = s.transform.p.y;
    pos[3 * i + 2] = s.transform.p.z;
    rot[4 * i + 0] = s.transform.q.v.x;
    rot[4 * i + 1] = s.transform.q.v.y;
    rot[4 * i + 2] = s.transform.q.v.z;
    rot[4 * i + 3] = s.transform.q.s;
    v[3 * i + 0] = st.v.x; v[3 * i + 1] = st.v.y; v[3 * i + 2] = st.v.z;
    av[3 * i + 0] = st.w.x; av[3 * i + 1] = st.w.y; av[3 * i + 2] = st.w.z;
  }
}

// [pairs_created, destroyed, begin_touch, end_touch, recycled, sat_calls, sat_hits, fast_bodies, overflow,
//  toi_calls, toi_hits]
void nslo_get_stats(void* wp, int64_t* out) {
  const StepStats& s = static_cast<World*>(wp)->stats;
  out[0] = s.pairs_created; out[1] = s.contacts_destroyed; out[2] = s.begin_touch; out[3] = s.end_touch;
  out[4] = s.recycled; out[5] = s.sat_calls; out[6] = s.sat_hits; out[7] = s.fast_bodies; out[8] = s.overflow_constraints;
  out[9] = s.toi_calls; out[10] = s.toi_hits;
}

// Where a body lives: solver set (0 static, 2 awake, 3+ sleeping), index in it, island id.
void nslo_body_location(void* wp, int body, int32_t* out) {
  const World* w = static_cast<World*>(wp);
  const Body& b = w->bodies[(size_t)body];
  out[0] = b.set_index;
  out[1] = b.local_index;
  out[2] = b.island_id;
}

// [island slots, live islands, awake islands, live sleeping sets, pending split island]
// Bodies in the awake set (world_events.rs:36 awake_body_count).
int nslo_awake_body_count(void* wp) { return (int)static_cast<World*>(wp)->awake_sims.size(); }

void nslo_island_summary(void* wp, int64_t* out) {
  const World* w = static_cast<World*>(wp);
  int64_t live = 0, sets = 0;
  for (const Island& is : w->islands) live += is.island_id != kNullIndex;
  for (const SleepSet& s : w->sleep_sets) sets += s.set_index != kNullIndex;
  out[0] = (int64_t)w->islands.size();
  out[1] = live;
  out[2] = (int64_t)w->awake_islands.size();
  out[3] = sets;
  out[4] = w->split_island_id;
}

// island_split_tests.rs:27-44: FNV-1a over every live island's id, set, body ids and
// contact ids, then every live body's island id and index.
uint64_t nslo_island_hash(void* wp, int with_bodies) {
  const World* w = static_cast<World*>(wp);
  uint64_t h = 14695981039346656037ull;
  auto mix = [&](int32_t v) {
    for (int k = 0; k < 4; ++k) {
      h ^= (uint64_t)((uint32_t)v >> (8 * k) & 0xFF);
      h *= 1099511628211ull;
    }
  };
  for (const Island& is : w->islands) {
    if (is.island_id == kNullIndex) continue;
    mix(is.island_id);
    mix(is.set_index);
    for (int32_t b : is.bodies) mix(b);
    for (const ContactLink& l : is.contacts) mix(l.contact_id);
  }
  if (with_bodies)
    for (const Body& b : w->bodies) {
      if (b.id == kNullIndex) continue;
      mix(b.island_id);
      mix(b.island_index);
    }
  return h;
}

// Cross-checks of the island and solver-set records; 0 when consistent, else the number of
// the first failed check. Also checks that every awake-set contact's body encodings equal
// its bodies' current encodings.
int nslo_check_islands(void* wp) {
  World* w = static_cast<World*>(wp);
  for (size_t id = 0; id < w->islands.size(); ++id) {
    const Island& is = w->islands[id];
    if (is.island_id == kNullIndex) continue;
    if (is.island_id != (int32_t)id) return 1;
    const std::vector<int32_t>& list = w->island_list(is.set_index);
    if (is.local_index < 0 || (size_t)is.local_index >= list.size() || list[(size_t)is.local_index] != (int32_t)id) return 2;
    for (size_t k = 0; k < is.bodies.size(); ++k) {
      const Body& b = w->bodies[(size_t)is.bodies[k]];
      if (b.island_id != (int32_t)id || b.island_index != (int32_t)k) return 3;
      if (b.set_index != is.set_index) return 4;
    }
    for (size_t k = 0; k < is.contacts.size(); ++k) {
      const Contact& c = w->contacts[(size_t)is.contacts[k].contact_id];
      if (c.island_id != (int32_t)id || c.island_index != (int32_t)k) return 5;
    }
    for (size_t k = 0; k < is.joints.size(); ++k) {
      const Joint& j = w->joints[(size_t)is.joints[k].joint_id];
      if (j.island_id != (int32_t)id || j.island_index != (int32_t)k) return 6;
    }
  }
  for (const Body& b : w->bodies) {
    if (b.id == kNullIndex) continue;
    if (b.set_index == 2 && w->awake_sims[(size_t)b.local_index].body_id != b.id) return 7;
    if (b.set_index >= 3 && w->sleep_sets[(size_t)b.set_index].body_sims[(size_t)b.local_index].body_id != b.id) return 8;
    if (b.type != kStatic && b.island_id == kNullIndex) return 9;
  }
  for (const Contact& c : w->contacts) {
    if (c.contact_id == kNullIndex) continue;
    if (c.color_index == kNullIndex) {
      const std::vector<int32_t>& list = w->contact_list(c.set_index);
      if (list[(size_t)c.local_index] != c.contact_id) return 10;
    }
    if (c.set_index == 2) {
      const int32_t ea = w->encode_body_sim_index(w->bodies[(size_t)c.edges[0].body_id]);
      const int32_t eb = w->encode_body_sim_index(w->bodies[(size_t)c.edges[1].body_id]);
      if (c.encoded_a != ea || c.encoded_b != eb) return 11;
    }
  }
  return 0;
}

// The first touching contact in id order and its island (island_split_tests.rs:84-97).
void nslo_first_island_contact(void* wp, int32_t* out) {
  const World* w = static_cast<World*>(wp);
  out[0] = out[1] = kNullIndex;
  for (const Contact& c : w->contacts)
    if (c.contact_id != kNullIndex && c.island_id != kNullIndex) {
      out[0] = c.contact_id;
      out[1] = c.island_id;
      return;
    }
}

// island_split_tests.rs:99-100: unlink a contact from its island, then split the island.
void nslo_unlink_and_split(void* wp, int contact, int island) {
  World* w = static_cast<World*>(wp);
  w->unlink_contact(contact);
  w->split_island(island);
}

// One time-of-impact query (time_of_impact.rs:406-598). Sweeps are 17 floats: local
// centre, c1, c2, q1 (x, y, z, s), q2. fout: point xyz, normal xyz, fraction, distance;
// iout: state, distance, push-back and root iterations.
static Sweep sweep_of(const float* s) {
  return Sweep{{s[0], s[1], s[2]}, {s[3], s[4], s[5]}, {s[6], s[7], s[8]}, {{s[9], s[10], s[11]}, s[12]},
               {{s[13], s[14], s[15]}, s[16]}};
}
void nslo_time_of_impact(const float* points_a, int count_a, float radius_a, const float* sweep_a,
                         const float* points_b, int count_b, float radius_b, const float* sweep_b,
                         float max_fraction, float* fout, int32_t* iout) {
  ToiProxy pa{}, pb{};
  for (int i = 0; i < count_a && i < 8; ++i) pa.points[i] = Vec3{points_a[3 * i], points_a[3 * i + 1], points_a[3 * i + 2]};
  for (int i = 0; i < count_b && i < 8; ++i) pb.points[i] = Vec3{points_b[3 * i], points_b[3 * i + 1], points_b[3 * i + 2]};
  pa.count = count_a < 8 ? count_a : 8;
  pb.count = count_b < 8 ? count_b : 8;
  pa.radius = radius_a;
  pb.radius = radius_b;
  ToiOutput o = time_of_impact(pa, sweep_of(sweep_a), pb, sweep_of(sweep_b), max_fraction);
  fout[0] = o.point.x; fout[1] = o.point.y; fout[2] = o.point.z;
  fout[3] = o.normal.x; fout[4] = o.normal.y; fout[5] = o.normal.z;
  fout[6] = o.fraction; fout[7] = o.distance;
  iout[0] = o.state; iout[1] = o.distance_iterations; iout[2] = o.push_back_iterations; iout[3] = o.root_iterations;
}

// Cumulative phase wall times (ms): pairs, collide, prepare, solve, finalize.
void nslo_get_phase_ms(void* wp, double* out) {
  const StepStats& s = static_cast<World*>(wp)->stats;
  out[0] = s.ms_pairs; out[1] = s.ms_collide; out[2] = s.ms_prepare; out[3] = s.ms_solve; out[4] = s.ms_finalize;
}

// Per shape id: fat AABB (lower xyz, upper xyz) and the moved flag.
void nslo_get_shape_bounds(void* wp, float* fat, uint8_t* moved) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->shapes.size(); ++i) {
    const AABB& b = w->fat_aabbs[i];
    const float v[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
    for (int k = 0; k < 6; ++k) fat[6 * i + k] = v[k];
    moved[i] = w->shapes[i].proxy_moved ? 1 : 0;
  }
}

// Per contact id: its two shape ids (-1 when the id is free).
void nslo_get_contact_shapes(void* wp, int32_t* a, int32_t* b) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->contacts.size(); ++i) {
    const bool live = w->contacts[i].contact_id != kNullIndex;
    a[i] = live ? w->contacts[i].shape_id_a : kNullIndex;
    b[i] = live ? w->contacts[i].shape_id_b : kNullIndex;
  }
}

// Graph coloring snapshot: per contact id, its color (-1 when not in the graph) and local index.
int nslo_contact_capacity(void* wp) { return (int)static_cast<World*>(wp)->contacts.size(); }
void nslo_get_contact_colors(void* wp, int32_t* color, int32_t* local, int32_t* alive) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->contacts.size(); ++i) {
    color[i] = w->contacts[i].color_index;
    local[i] = w->contacts[i].local_index;
    alive[i] = w->contacts[i].contact_id != kNullIndex;
  }
}

}  // extern "C"
