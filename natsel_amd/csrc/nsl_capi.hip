// extern "C" boundary of the GPU stepper (declared in include/natsel_gpu.h).
#include <exception>
#include <new>
#include <cstddef>
#include <cstring>
#include <string>

#include <stdexcept>

#include "../../include/natsel_gpu.h"
#include "nsl_batch.h"
#include "nsl_narrow.h"
#include "nsl_world.h"

struct nsl_world {
  nsl::GpuWorld* impl;
};

struct nsl_batch {
  nsl::BatchStepper* impl;
};

namespace {
thread_local std::string g_error;

int fail(int code, const char* what) {
  g_error = what;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return NSL_OK;
  } catch (const std::bad_alloc&) {
    return fail(NSL_ERR_DEVICE, "out of memory");
  } catch (const std::invalid_argument& e) {
    return fail(NSL_ERR_INVALID, e.what());
  } catch (const std::exception& e) {
    return fail(NSL_ERR_DEVICE, e.what());
  } catch (...) {
    return fail(NSL_ERR_INTERNAL, "unknown failure");
  }
}

void apply(nsl::WorldOptions& o, const nsl_options& in) {
  o.gravity = nsl::Vec3{in.gravity[0], in.gravity[1], in.gravity[2]};
  o.contact_hertz = in.contact_hertz;
  o.contact_damping_ratio = in.contact_damping_ratio;
  o.contact_speed = in.contact_speed;
  o.max_linear_speed = in.max_linear_speed;
  o.restitution_threshold = in.restitution_threshold;
  o.restitution_iterations = in.restitution_iterations;
  o.enable_sleep = in.enable_sleep != 0;
  o.enable_continuous = in.enable_continuous != 0;
  o.enable_warm_starting = in.enable_warm_starting != 0;
  o.enable_speculative = in.enable_speculative != 0;
  o.solver_blocks = in.solver_blocks;
  o.enable_restitution_propagation = in.enable_restitution_propagation != 0;
  o.small_colour_items = in.small_colour_items;
  o.stage_records = in.stage_records != 0 ? 1 : 0;
  o.plain_step_launch = in.plain_step_launch != 0 ? 1 : 0;
}

nsl::Material material(const nsl_material* m) {
  nsl::Material out;
  if (m) {
    out.density = m->density;
    out.friction = m->friction;
    out.restitution = m->restitution;
    out.rolling_resistance = m->rolling_resistance;
    out.category_bits = m->filter_category_bits;
    out.mask_bits = m->filter_mask_bits;
    out.group_index = m->filter_group_index;
    out.is_sensor = m->is_sensor != 0;
    out.enable_speculative = m->enable_speculative_contact != 0;
    out.enable_hit_events = m->enable_hit_events != 0;
  }
  return out;
}

bool valid_body(nsl_world* w, int32_t body) { return w && body >= 0 && body < w->impl->body_capacity(); }

nsl::BodyDesc body_desc(const nsl_body_def* def) {
  nsl::BodyDesc d;
  d.type = def->type;
  d.position = nsl::Pos{def->position[0], def->position[1], def->position[2]};
  d.rotation = nsl::Quat{{def->rotation[0], def->rotation[1], def->rotation[2]}, def->rotation[3]};
  d.linear_velocity = nsl::Vec3{def->linear_velocity[0], def->linear_velocity[1], def->linear_velocity[2]};
  d.angular_velocity = nsl::Vec3{def->angular_velocity[0], def->angular_velocity[1], def->angular_velocity[2]};
  d.linear_damping = def->linear_damping;
  d.angular_damping = def->angular_damping;
  d.gravity_scale = def->gravity_scale;
  d.enable_sleep = def->enable_sleep != 0;
  d.is_bullet = def->is_bullet != 0;
  d.safety_factor = def->safety_factor;
  d.sleep_threshold = def->sleep_threshold;
  d.lock_flags = (uint32_t)def->motion_locks & 0x3Fu;
  d.allow_fast_rotation = def->allow_fast_rotation != 0;
  d.enable_contact_recycling = def->enable_contact_recycling != 0;
  // world_runtime.rs:1517-1540: a body that is not enabled goes to the disabled set, one that
  // is not awake (and may sleep) to a sleeping set of its own; neither is built here
  if (!def->is_enabled) throw std::invalid_argument("a body must be enabled (is_enabled = 0 is not supported)");
  if (def->type != NSL_BODY_STATIC && !def->is_awake && def->enable_sleep)
    throw std::invalid_argument("a body must start awake (is_awake = 0 with enable_sleep is not supported)");
  return d;
}

// World `w` of a batch with body `body` (or no body check when body < -1 is passed as -2).
nsl::HostScene* batch_scene(nsl_batch* b, int32_t w, int32_t body) {
  if (!b || w < 0 || w >= b->impl->world_count()) return nullptr;
  nsl::HostScene* sc = &b->impl->scene(w);
  if (body != -2 && (body < 0 || body >= (int32_t)sc->bodies_.size())) return nullptr;
  return sc;
}
}  // namespace

extern "C" {

const char* nsl_last_error(void) { return g_error.c_str(); }

void nsl_default_options(nsl_options* out) {
  if (!out) return;
  nsl::WorldOptions o;
  out->gravity[0] = o.gravity.x;
  out->gravity[1] = o.gravity.y;
  out->gravity[2] = o.gravity.z;
  out->contact_hertz = o.contact_hertz;
  out->contact_damping_ratio = o.contact_damping_ratio;
  out->contact_speed = o.contact_speed;
  out->max_linear_speed = o.max_linear_speed;
  out->restitution_threshold = o.restitution_threshold;
  out->restitution_iterations = o.restitution_iterations;
  out->enable_sleep = o.enable_sleep;
  out->enable_continuous = o.enable_continuous;
  out->enable_warm_starting = o.enable_warm_starting;
  out->enable_speculative = o.enable_speculative;
  out->solver_blocks = o.solver_blocks;
  out->enable_restitution_propagation = o.enable_restitution_propagation;
  out->small_colour_items = o.small_colour_items;
  out->stage_records = o.stage_records;
  out->plain_step_launch = o.plain_step_launch;
}

int nsl_world_create(int device, const nsl_options* options, nsl_world** out) {
  if (!out) return fail(NSL_ERR_INVALID, "null output handle");
  *out = nullptr;
  return guarded([&] {
    nsl_world* w = new nsl_world{new nsl::GpuWorld(device)};
    if (options) apply(w->impl->options, *options);
    *out = w;
  });
}

int nsl_world_destroy(nsl_world* world) {
  if (!world) return fail(NSL_ERR_INVALID, "null world");
  return guarded([&] {
    delete world->impl;
    delete world;
  });
}

int nsl_world_set_options(nsl_world* world, const nsl_options* options) {
  if (!world || !options) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { apply(world->impl->options, *options); });
}

int nsl_create_body(nsl_world* world, const nsl_body_def* def, int32_t* out_body) {
  if (!world || !def || !out_body) return fail(NSL_ERR_INVALID, "null argument");
  if (def->type < NSL_BODY_STATIC || def->type > NSL_BODY_DYNAMIC) return fail(NSL_ERR_INVALID, "bad body type");
  return guarded([&] { *out_body = world->impl->create_body(body_desc(def)); });
}

int nsl_create_box(nsl_world* world, int32_t body, const float h[3], const nsl_material* m, int32_t* out_shape) {
  if (!valid_body(world, body) || !h || !out_shape) return fail(NSL_ERR_INVALID, "bad body or null argument");
  return guarded([&] { *out_shape = world->impl->create_box(body, nsl::Vec3{h[0], h[1], h[2]}, material(m)); });
}

int nsl_create_sphere(nsl_world* world, int32_t body, const float c[3], float radius, const nsl_material* m,
                      int32_t* out_shape) {
  if (!valid_body(world, body) || !c || !out_shape) return fail(NSL_ERR_INVALID, "bad body or null argument");
  return guarded([&] { *out_shape = world->impl->create_sphere(body, nsl::Vec3{c[0], c[1], c[2]}, radius, material(m)); });
}

int nsl_create_capsule(nsl_world* world, int32_t body, const float c1[3], const float c2[3], float radius,
                       const nsl_material* m, int32_t* out_shape) {
  if (!valid_body(world, body) || !c1 || !c2 || !out_shape) return fail(NSL_ERR_INVALID, "bad body or null argument");
  return guarded([&] {
    *out_shape = world->impl->create_capsule(body, nsl::Vec3{c1[0], c1[1], c1[2]}, nsl::Vec3{c2[0], c2[1], c2[2]}, radius,
                                             material(m));
  });
}

int nsl_world_step(nsl_world* world, float dt, int32_t sub_steps) {
  if (!world || !(dt >= 0.0f)) return fail(NSL_ERR_INVALID, "null world or negative time step");
  return guarded([&] { world->impl->step(dt, sub_steps); });
}

void nsl_default_revolute_joint_def(nsl_revolute_joint_def* out) {
  if (!out) return;
  nsl::JointDesc d;
  std::memset(out, 0, sizeof(*out));
  out->body_a = d.body_a;
  out->body_b = d.body_b;
  out->local_frame_a[6] = 1.0f;
  out->local_frame_b[6] = 1.0f;
  out->constraint_hertz = d.constraint_hertz;
  out->constraint_damping_ratio = d.constraint_damping_ratio;
}

void nsl_default_body_def(nsl_body_def* out) {
  if (!out) return;
  nsl::BodyDesc d;
  std::memset(out, 0, sizeof(*out));
  out->type = d.type;
  out->rotation[3] = 1.0f;
  out->linear_damping = d.linear_damping;
  out->angular_damping = d.angular_damping;
  out->gravity_scale = d.gravity_scale;
  out->enable_sleep = d.enable_sleep ? 1 : 0;
  out->is_bullet = d.is_bullet ? 1 : 0;
  out->safety_factor = d.safety_factor;
  out->sleep_threshold = d.sleep_threshold;
  out->motion_locks = (int32_t)d.lock_flags;
  out->is_awake = 1;
  out->is_enabled = 1;
  out->allow_fast_rotation = d.allow_fast_rotation ? 1 : 0;
  out->enable_contact_recycling = d.enable_contact_recycling ? 1 : 0;
}

void nsl_default_material(nsl_material* out) {
  if (!out) return;
  const nsl::Material m;
  std::memset(out, 0, sizeof(*out));
  out->density = m.density;
  out->friction = m.friction;
  out->restitution = m.restitution;
  out->rolling_resistance = m.rolling_resistance;
  out->filter_category_bits = m.category_bits;
  out->filter_mask_bits = m.mask_bits;
  out->filter_group_index = m.group_index;
  out->is_sensor = m.is_sensor ? 1 : 0;
  out->enable_speculative_contact = m.enable_speculative ? 1 : 0;
  out->enable_hit_events = m.enable_hit_events ? 1 : 0;
}

static_assert(sizeof(nsl_toi_input) == sizeof(nsl::ToiQuery), "toi input layout");
static_assert(offsetof(nsl_toi_input, sweep_a) == offsetof(nsl::ToiQuery, sweep_a), "toi input layout");
static_assert(offsetof(nsl_toi_input, max_fraction) == offsetof(nsl::ToiQuery, max_fraction), "toi input layout");
static_assert(sizeof(nsl_toi_output) == sizeof(nsl::ToiOutput), "toi output layout");
static_assert(offsetof(nsl_toi_output, root_iterations) == offsetof(nsl::ToiOutput, root_iterations), "toi output layout");

int nsl_collide(int device, const nsl_collide_query* queries, nsl_local_manifold* out, int32_t n) {
  if (n < 0 || (n > 0 && (!queries || !out))) return fail(NSL_ERR_INVALID, "null argument or negative count");
  int32_t bad = -1;
  if (const char* why = nsl::validate_collide_queries(queries, n, &bad)) {
    g_error = "query " + std::to_string(bad) + ": " + why;
    return NSL_ERR_INVALID;
  }
  if (n == 0) return NSL_OK;
  return guarded([&] {
    auto check = [](hipError_t e) {
      if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    };
    check(hipSetDevice(device));
    nsl_collide_query* dq = nullptr;
    nsl_local_manifold* dout = nullptr;
    check(hipMalloc(&dq, (size_t)n * sizeof(nsl_collide_query)));
    hipError_t e = hipMalloc(&dout, (size_t)n * sizeof(nsl_local_manifold));
    if (e == hipSuccess) e = hipMemcpy(dq, queries, (size_t)n * sizeof(nsl_collide_query), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      nsl::launch_collide_queries(dq, dout, n, nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * sizeof(nsl_local_manifold), hipMemcpyDeviceToHost);
    (void)hipFree(dq);
    if (dout) (void)hipFree(dout);
    check(e);
  });
}

int nsl_collide_noalloc(int device, const nsl_collide_query* queries, nsl_local_manifold* out, int32_t n) {
  if (n < 0 || (n > 0 && (!queries || !out))) return fail(NSL_ERR_INVALID, "null argument or negative count");
  if (n > nsl::kNoAllocMaxQueries) return fail(NSL_ERR_INVALID, "more queries than nsl_collide_noalloc stages (8192)");
  int32_t bad = -1;
  if (const char* why = nsl::validate_collide_queries(queries, n, &bad)) return fail(NSL_ERR_INVALID, why);
  if (n == 0) return NSL_OK;
  if (hipSetDevice(device) != hipSuccess) return fail(NSL_ERR_DEVICE, "hipSetDevice failed");
  if (const char* why = nsl::collide_queries_noalloc(device, queries, out, n)) return fail(NSL_ERR_DEVICE, why);
  return NSL_OK;
}

int nsl_time_of_impact(int device, const nsl_toi_input* in, nsl_toi_output* out, int32_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return fail(NSL_ERR_INVALID, "null argument or negative count");
  for (int32_t k = 0; k < n; ++k)
    if (in[k].count_a < 1 || in[k].count_a > 8 || in[k].count_b < 1 || in[k].count_b > 8)
      return fail(NSL_ERR_INVALID, "a proxy needs 1 to 8 points");
  if (n == 0) return NSL_OK;
  return guarded([&] {
    auto check = [](hipError_t e) {
      if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    };
    check(hipSetDevice(device));
    nsl::ToiQuery* dq = nullptr;
    nsl::ToiOutput* dout = nullptr;
    check(hipMalloc(&dq, (size_t)n * sizeof(nsl::ToiQuery)));
    hipError_t e = hipMalloc(&dout, (size_t)n * sizeof(nsl::ToiOutput));
    if (e == hipSuccess) e = hipMemcpy(dq, in, (size_t)n * sizeof(nsl::ToiQuery), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      nsl::launch_time_of_impact(n, dq, dout, nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * sizeof(nsl::ToiOutput), hipMemcpyDeviceToHost);
    (void)hipFree(dq);
    if (dout) (void)hipFree(dout);
    check(e);
  });
}

int nsl_create_revolute_joint(nsl_world* world, const nsl_revolute_joint_def* def, int32_t* out_joint) {
  if (!world || !def || !out_joint) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] {
    nsl::JointDesc d;
    d.body_a = def->body_a;
    d.body_b = def->body_b;
    const float* fa = def->local_frame_a;
    const float* fb = def->local_frame_b;
    d.local_frame_a = nsl::Xf{{fa[0], fa[1], fa[2]}, {{fa[3], fa[4], fa[5]}, fa[6]}};
    d.local_frame_b = nsl::Xf{{fb[0], fb[1], fb[2]}, {{fb[3], fb[4], fb[5]}, fb[6]}};
    d.constraint_hertz = def->constraint_hertz;
    d.constraint_damping_ratio = def->constraint_damping_ratio;
    d.collide_connected = def->collide_connected != 0;
    d.enable_spring = def->enable_spring != 0;
    d.hertz = def->hertz;
    d.damping_ratio = def->damping_ratio;
    d.target_angle = def->target_angle;
    d.enable_limit = def->enable_limit != 0;
    d.lower_angle = def->lower_angle;
    d.upper_angle = def->upper_angle;
    d.enable_motor = def->enable_motor != 0;
    d.max_motor_torque = def->max_motor_torque;
    d.motor_speed = def->motor_speed;
    *out_joint = world->impl->create_revolute_joint(d);
  });
}

int nsl_world_get_joint(nsl_world* world, int32_t joint, float impulses[9], int32_t* color) {
  if (!world || !impulses || !color) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { world->impl->get_joint(joint, impulses, color); });
}

int nsl_world_joint_count(nsl_world* world, int32_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  *out = world->impl->joint_count();
  return NSL_OK;
}

struct nsl_snapshot {
  nsl::GpuWorld::Snapshot* impl = nullptr;
  ~nsl_snapshot() { nsl::GpuWorld::destroy_snapshot(impl); }
};

int nsl_world_save_state(nsl_world* world, nsl_snapshot** out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] {
    auto s = std::make_unique<nsl_snapshot>();
    s->impl = world->impl->save_state();
    *out = s.release();
  });
}

int nsl_world_restore_state(nsl_world* world, const nsl_snapshot* snapshot) {
  if (!world || !snapshot) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { world->impl->restore_state(*snapshot->impl); });
}

int nsl_snapshot_destroy(nsl_snapshot* snapshot) {
  delete snapshot;
  return NSL_OK;
}

int nsl_world_contact_count(nsl_world* world, int32_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  *out = world->impl->contact_count();
  return NSL_OK;
}

int nsl_world_step_n(nsl_world* world, float dt, int32_t sub_steps, int32_t steps) {
  if (!world || !(dt >= 0.0f) || steps < 0) return fail(NSL_ERR_INVALID, "bad argument");
  return guarded([&] {
    for (int32_t i = 0; i < steps; ++i) world->impl->step(dt, sub_steps);
  });
}

int nsl_world_body_capacity(nsl_world* world, int32_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  *out = world->impl->body_capacity();
  return NSL_OK;
}

int nsl_world_get_bodies(nsl_world* world, double* pos, float* rot, float* v, float* w) {
  if (!world || !pos || !rot || !v || !w) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { world->impl->get_bodies(pos, rot, v, w); });
}

int nsl_world_state_hash(nsl_world* world, uint64_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { *out = world->impl->state_hash(); });
}

int nsl_world_contact_capacity(nsl_world* world, int32_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  *out = world->impl->contact_capacity();
  return NSL_OK;
}

int nsl_world_get_contact_colors(nsl_world* world, int32_t* color, int32_t* local, int32_t* alive) {
  if (!world || !color || !local || !alive) return fail(NSL_ERR_INVALID, "null argument");
  world->impl->get_contact_colors(color, local, alive);
  return NSL_OK;
}

int nsl_world_get_contact_shapes(nsl_world* world, int32_t* shape_a, int32_t* shape_b) {
  if (!world || !shape_a || !shape_b) return fail(NSL_ERR_INVALID, "null argument");
  world->impl->get_contact_shapes(shape_a, shape_b);
  return NSL_OK;
}

int nsl_world_body_location(nsl_world* world, int32_t body, int32_t* set, int32_t* local, int32_t* island) {
  if (!world || !set || !local || !island) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { world->impl->body_location(body, set, local, island); });
}

int nsl_world_island_summary(nsl_world* world, int64_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  world->impl->island_summary(out);
  return NSL_OK;
}

int nsl_world_check_islands(nsl_world* world, int32_t* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  *out = world->impl->check_islands();
  return NSL_OK;
}

int nsl_world_get_profile(nsl_world* world, nsl_profile* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  const nsl::StepProfile& p = world->impl->profile();
  out->pairs_ms = p.pairs;
  out->collide_ms = p.collide;
  out->state_pass_ms = p.state_pass;
  out->solve_ms = p.solve;
  out->finalize_ms = p.finalize;
  out->step_ms = p.step;
  out->contacts = p.contacts;
  out->touching = p.touching;
  out->colors_used = p.colors_used;
  out->overflow = p.overflow;
  out->fast_bodies = p.fast_bodies;
  out->sat_calls = p.sat_calls;
  out->sat_hits = p.sat_hits;
  out->recycled = p.recycled;
  return NSL_OK;
}

int nsl_world_set_kernel_timing(nsl_world* world, int32_t on) {
  if (!world) return fail(NSL_ERR_INVALID, "null world");
  return guarded([&] { world->impl->set_kernel_timing((on & 1) != 0, (on & 2) != 0); });
}

int nsl_world_autotune(nsl_world* world, float dt, int32_t sub_steps, int32_t steps, nsl_tuning* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  if (steps < 1 || steps > 1000 || sub_steps < 1 || !(dt > 0.0f)) return fail(NSL_ERR_INVALID, "bad time step or step count");
  return guarded([&] {
    const nsl::GpuWorld::Tuning t = world->impl->autotune(dt, sub_steps, steps);
    out->solver_blocks = t.blocks;
    out->small_colour_items = t.small_colour_items;
    out->stage_records = t.stage_records;
    out->step_kernel_us = t.best_us;
    out->previous_us = t.previous_us;
    out->candidates = t.candidates;
  });
}

int nsl_world_set_residuals(nsl_world* world, int32_t on) {
  if (!world) return fail(NSL_ERR_INVALID, "null world");
  return guarded([&] { world->impl->set_residuals(on != 0); });
}

int nsl_world_get_residuals(nsl_world* world, nsl_colour_residual* out, int32_t n) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  if (n < 0 || n > NSL_GRAPH_COLOURS) return fail(NSL_ERR_INVALID, "more colours than the graph has");
  return guarded([&] {
    const nsl::GpuWorld::ColourResidual* r = world->impl->residuals();
    for (int32_t i = 0; i < n; ++i) {
      out[i].sum_normal = r[i].sum_normal;
      out[i].max_normal = r[i].max_normal;
      out[i].max_friction = r[i].max_friction;
      out[i].constraints = r[i].constraints;
      out[i].reserved = 0;
    }
  });
}

int nsl_world_reset_kernel_stats(nsl_world* world) {
  if (!world) return fail(NSL_ERR_INVALID, "null world");
  return guarded([&] { world->impl->reset_kernel_stats(); });
}

int nsl_world_get_kernel_stats(nsl_world* world, nsl_kernel_stats* out) {
  if (!world || !out) return fail(NSL_ERR_INVALID, "null argument");
  static_assert((int)NSL_K_CLASS_COUNT == (int)nsl::kKernelClassCount, "kernel class count");
  const auto& k = world->impl->kernel_stats();
  for (int i = 0; i < NSL_K_CLASS_COUNT; ++i) {
    out->ms[i] = k.ms[i];
    out->launches[i] = k.launches[i];
  }
  out->steps = k.steps;
  out->bodies = k.bodies;
  out->collided = k.collided;
  out->constraints = k.constraints;
  out->points = k.points;
  out->overflow_points = k.overflow_points;
  out->recycled = k.recycled;
  out->sat_calls = k.sat_calls;
  out->sat_hits = k.sat_hits;
  out->fast_bodies = k.fast_bodies;
  static_assert((int)NSL_P_CLASS_COUNT == (int)nsl::kPhaseClassCount, "phase class count");
  for (int i = 0; i < NSL_P_CLASS_COUNT; ++i) {
    out->phase_ms[i] = k.phase_ms[i];
    out->phase_count[i] = k.phase_count[i];
  }
  out->barriers = k.barriers;
  out->solver_blocks = k.solver_blocks;
  out->pair_tests = k.pair_tests;
  for (int i = 0; i < 4; ++i) out->host_ms[i] = k.host_ms[i];
  out->joints = k.joints;
  out->toi_calls = k.toi_calls;
  out->toi_hits = k.toi_hits;
  return NSL_OK;
}

// ---------------------------------------------------------------------------------
// batches of independent worlds

int nsl_batch_create(int device, const nsl_options* options, nsl_batch** out) {
  if (!out) return fail(NSL_ERR_INVALID, "null output handle");
  *out = nullptr;
  return guarded([&] {
    auto* b = new nsl_batch{new nsl::BatchStepper(device)};
    if (options) apply(b->impl->options, *options);
    *out = b;
  });
}

int nsl_batch_destroy(nsl_batch* batch) {
  if (!batch) return fail(NSL_ERR_INVALID, "null batch");
  return guarded([&] {
    delete batch->impl;
    delete batch;
  });
}

int nsl_batch_add_world(nsl_batch* batch, int32_t* out_world) {
  if (!batch || !out_world) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { *out_world = batch->impl->add_world(); });
}

int nsl_batch_create_body(nsl_batch* batch, int32_t world, const nsl_body_def* def, int32_t* out_body) {
  if (!def || !out_body) return fail(NSL_ERR_INVALID, "null argument");
  if (def->type < NSL_BODY_STATIC || def->type > NSL_BODY_DYNAMIC) return fail(NSL_ERR_INVALID, "bad body type");
  return guarded([&] {
    nsl::HostScene* sc = batch_scene(batch, world, -2);
    if (!sc) throw std::invalid_argument("bad batch or world");
    *out_body = sc->create_body(body_desc(def));
  });
}

int nsl_batch_create_box(nsl_batch* batch, int32_t world, int32_t body, const float h[3], const nsl_material* m,
                         int32_t* out_shape) {
  if (!h || !out_shape) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] {
    nsl::HostScene* sc = batch_scene(batch, world, body);
    if (!sc) throw std::invalid_argument("bad batch, world or body");
    *out_shape = sc->create_box(body, nsl::Vec3{h[0], h[1], h[2]}, material(m));
  });
}

int nsl_batch_create_sphere(nsl_batch* batch, int32_t world, int32_t body, const float c[3], float radius,
                            const nsl_material* m, int32_t* out_shape) {
  if (!c || !out_shape) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] {
    nsl::HostScene* sc = batch_scene(batch, world, body);
    if (!sc) throw std::invalid_argument("bad batch, world or body");
    *out_shape = sc->create_sphere(body, nsl::Vec3{c[0], c[1], c[2]}, radius, material(m));
  });
}

int nsl_batch_create_capsule(nsl_batch* batch, int32_t world, int32_t body, const float c1[3], const float c2[3], float radius,
                             const nsl_material* m, int32_t* out_shape) {
  if (!c1 || !c2 || !out_shape) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] {
    nsl::HostScene* sc = batch_scene(batch, world, body);
    if (!sc) throw std::invalid_argument("bad batch, world or body");
    *out_shape = sc->create_capsule(body, nsl::Vec3{c1[0], c1[1], c1[2]}, nsl::Vec3{c2[0], c2[1], c2[2]}, radius, material(m));
  });
}

int nsl_batch_step(nsl_batch* batch, float dt, int32_t sub_steps, int32_t steps) {
  if (!batch || !(dt > 0.0f) || steps < 0) return fail(NSL_ERR_INVALID, "null batch, non-positive time step or negative steps");
  return guarded([&] { batch->impl->step(dt, sub_steps, steps); });
}

int nsl_batch_world_count(nsl_batch* batch, int32_t* out) {
  if (!batch || !out) return fail(NSL_ERR_INVALID, "null argument");
  *out = batch->impl->world_count();
  return NSL_OK;
}

int nsl_batch_body_count(nsl_batch* batch, int32_t world, int32_t* out) {
  if (!batch || !out || world < 0 || world >= batch->impl->world_count()) return fail(NSL_ERR_INVALID, "bad argument");
  return guarded([&] { *out = batch->impl->body_count(world); });
}

int nsl_batch_state_hashes(nsl_batch* batch, uint64_t* out) {
  if (!batch || !out) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { batch->impl->state_hashes(out); });
}

int nsl_batch_get_bodies(nsl_batch* batch, int32_t world, double* pos, float* rot, float* v, float* w) {
  if (!batch || !pos || !rot || !v || !w) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { batch->impl->get_bodies(world, pos, rot, v, w); });
}

int nsl_batch_contact_capacity(nsl_batch* batch, int32_t world, int32_t* out) {
  if (!batch || !out || world < 0 || world >= batch->impl->world_count()) return fail(NSL_ERR_INVALID, "bad argument");
  return guarded([&] { *out = batch->impl->contact_slots(world); });
}

int nsl_batch_get_contact_colors(nsl_batch* batch, int32_t world, int32_t* color, int32_t* local, int32_t* alive) {
  if (!batch || !color || !local || !alive || world < 0 || world >= batch->impl->world_count())
    return fail(NSL_ERR_INVALID, "bad argument");
  return guarded([&] { batch->impl->get_contact_colors(world, color, local, alive); });
}

int nsl_batch_get_contact_shapes(nsl_batch* batch, int32_t world, int32_t* shape_a, int32_t* shape_b) {
  if (!batch || !shape_a || !shape_b || world < 0 || world >= batch->impl->world_count()) return fail(NSL_ERR_INVALID, "bad argument");
  return guarded([&] { batch->impl->get_contact_shapes(world, shape_a, shape_b); });
}

int nsl_batch_get_shape_bounds(nsl_batch* batch, int32_t world, float* fat, uint8_t* moved) {
  if (!batch || !fat || !moved || world < 0 || world >= batch->impl->world_count()) return fail(NSL_ERR_INVALID, "bad argument");
  return guarded([&] { batch->impl->get_shape_bounds(world, fat, moved); });
}

int nsl_batch_record(nsl_batch* batch, float dt, int32_t sub_steps, int32_t steps, int32_t max_constraints) {
  if (!batch || !(dt > 0.0f) || steps <= 0) return fail(NSL_ERR_INVALID, "null batch, non-positive time step or step count");
  return guarded([&] { batch->impl->record(dt, sub_steps, steps, max_constraints); });
}

int nsl_batch_adjoint(nsl_batch* batch, const float* targets, float* loss, float* grad_v, float* grad_w, float* grad_friction,
                      int32_t* status) {
  if (!batch || !targets) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { batch->impl->adjoint(targets, loss, grad_v, grad_w, grad_friction, status); });
}

int nsl_batch_model_losses(nsl_batch* batch, const int32_t* worlds, const int32_t* params, const double* deltas, int32_t n,
                           const float* targets, int32_t frozen, double* out) {
  if (!batch || n < 0 || (n > 0 && (!worlds || !params || !deltas || !targets || !out))) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { batch->impl->model_losses(worlds, params, deltas, n, targets, out, frozen != 0); });
}

int nsl_batch_rollout_times(nsl_batch* batch, double* record_ms, double* adjoint_ms) {
  if (!batch || !record_ms || !adjoint_ms) return fail(NSL_ERR_INVALID, "null argument");
  *record_ms = batch->impl->last_record_ms();
  *adjoint_ms = batch->impl->last_adjoint_ms();
  return NSL_OK;
}

static void tuning_out(const nsl::BatchTuning& t, nsl_batch_tuning* out) {
  out->collide_blocks = t.collide_blocks;
  out->use_graph = t.use_graph;
  out->candidates = t.candidates;
  out->reserved = 0;
  out->step_ms = t.step_ms;
  out->previous_ms = t.previous_ms;
}

int nsl_batch_set_tuning(nsl_batch* batch, const nsl_batch_tuning* tuning) {
  if (!batch || !tuning) return fail(NSL_ERR_INVALID, "null argument");
  if (tuning->collide_blocks < 0 || tuning->collide_blocks > (1 << 20)) return fail(NSL_ERR_INVALID, "collide_blocks out of range");
  return guarded([&] {
    nsl::BatchTuning t;
    t.collide_blocks = tuning->collide_blocks;
    t.use_graph = tuning->use_graph;
    batch->impl->set_tuning(t);
  });
}

int nsl_batch_get_tuning(nsl_batch* batch, nsl_batch_tuning* out) {
  if (!batch || !out) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] { tuning_out(batch->impl->tuning(), out); });
}

int nsl_batch_autotune(nsl_batch* batch, float dt, int32_t sub_steps, int32_t steps, nsl_batch_tuning* out) {
  if (!batch || !out) return fail(NSL_ERR_INVALID, "null argument");
  if (steps < 1 || steps > 1000 || sub_steps < 1 || !(dt > 0.0f)) return fail(NSL_ERR_INVALID, "bad time step or step count");
  return guarded([&] { tuning_out(batch->impl->autotune(dt, sub_steps, steps), out); });
}

int nsl_batch_get_stats(nsl_batch* batch, nsl_batch_stats* out) {
  if (!batch || !out) return fail(NSL_ERR_INVALID, "null argument");
  return guarded([&] {
    out->last_step_ms = batch->impl->last_step_ms();
    out->body_steps = batch->impl->total_body_steps();
    out->bodies = batch->impl->total_bodies();
    out->contact_capacity = batch->impl->contact_capacity();
    out->error_worlds = batch->impl->error_worlds(&out->first_error);
    int64_t w[4];
    batch->impl->work_counts(w);
    out->contacts_collided = w[0];
    out->constraints = w[1];
    out->points = w[2];
    out->overflow_constraints = w[3];
  });
}

}  // extern "C"
