// Host records of one world's bodies and shapes: creation, mass properties and the
// initial fat AABBs (world_runtime.rs:1506-1649, shape_creation.rs:196-559,
// body_api.rs:611-802). The single-world stepper and the batched stepper both build
// their worlds with it and copy the records to the device.
#pragma once
#include <cstdint>
#include <vector>

#include "nsl_types.h"

namespace nsl {

struct BodyDesc {
  int32_t type = kBodyStatic;
  Pos position{0, 0, 0};
  Quat rotation{{0, 0, 0}, 1};
  Vec3 linear_velocity{0, 0, 0}, angular_velocity{0, 0, 0};
  float linear_damping = 0, angular_damping = 0, gravity_scale = 1.0f;
  float sleep_threshold = 0.05f, safety_factor = 0.5f;
  uint32_t lock_flags = 0;
  bool enable_sleep = true, is_bullet = false, allow_fast_rotation = false, enable_contact_recycling = true;
};

struct Material {
  float density = 1000.0f, friction = 0.6f, restitution = 0.0f, rolling_resistance = 0.0f;
  bool enable_speculative = true, enable_hit_events = false;
  bool is_sensor = false;  // no contacts, no continuous collision (contact_creation.rs:187, continuous.rs:398-401)
  uint64_t category_bits = 1, mask_bits = ~0ull;
  int32_t group_index = 0;
};

// RevoluteJointDef (types/joint_defs.rs:30-58, :154-167). Frames are in the bodies' own
// frames; the hinge axis is frame A's z axis.
struct JointDesc {
  int32_t body_a = -1, body_b = -1;
  Xf local_frame_a{{0, 0, 0}, {{0, 0, 0}, 1}}, local_frame_b{{0, 0, 0}, {{0, 0, 0}, 1}};
  float constraint_hertz = 60.0f, constraint_damping_ratio = 2.0f;
  bool collide_connected = false;
  bool enable_spring = false;
  float hertz = 0.0f, damping_ratio = 0.0f, target_angle = 0.0f;
  bool enable_limit = false;
  float lower_angle = 0.0f, upper_angle = 0.0f;
  bool enable_motor = false;
  float max_motor_torque = 0.0f, motor_speed = 0.0f;
};

// World settings (world_runtime.rs:1243-1304), shared by the steppers.
struct WorldOptions {
  Vec3 gravity{0.0f, -10.0f, 0.0f};
  float restitution_threshold = 1.0f;
  int32_t restitution_iterations = 2;
  bool enable_restitution_propagation = false;
  float max_linear_speed = 400.0f;
  float contact_speed = 3.0f, contact_hertz = 30.0f, contact_damping_ratio = 10.0f;
  float isotropic_inertia_tolerance = 0.2f;
  bool enable_sleep = true, enable_warm_starting = true, enable_continuous = true, enable_speculative = true;
  int32_t solver_blocks = 0;  // workgroups of the step kernel; 0 = sized from the largest phase
  // consecutive colours of at most this many items (joints and contacts) run on one
  // workgroup with workgroup barriers between them; 0 = every colour behind a grid barrier
  int32_t small_colour_items = 128;
  // 1: the step kernel copies each lane's constraint record of the next colour to LDS while
  // the current colour runs (results do not change)
  int32_t stage_records = 0;
  // 1: the step kernel is launched without the cooperative API (same grid and results)
  int32_t plain_step_launch = 0;
};


class HostScene {
 public:
  virtual ~HostScene() = default;

  virtual int32_t create_body(const BodyDesc& desc);
  int32_t create_box(int32_t body, Vec3 half_extents, const Material& m);
  int32_t create_sphere(int32_t body, Vec3 center, float radius, const Material& m);
  int32_t create_capsule(int32_t body, Vec3 c1, Vec3 c2, float radius, const Material& m);

  float isotropic_inertia_tolerance = 0.2f;  // world_def default

  // From the host records: per body id origin position, rotation (x, y, z, s), linear and
  // angular velocity; and the reference state hash (body_queries.rs:432-480).
  void get_bodies(double* pos, float* rot, float* v, float* w) const;
  uint64_t state_hash() const;

  // set: 0 static, 2 awake, 3 and up a sleeping set (local: index in that set's bodies)
  struct HBody {
    int32_t type = kBodyStatic, set = 0, local = -1, head_shape = -1;
    uint32_t flags = 0;
    int32_t head_contact = -1;               // contact_id << 1 | edge of the newest contact
    int32_t island = -1, island_index = -1;  // kinematic and dynamic bodies
  };
  struct HShape {
    int32_t id = -1, body = -1, kind = kShapeHull, hull = -1;
    uint64_t category = 1, mask = ~0ull;
    int32_t group = 0;
    float density = 0, restitution = 0, rolling_resistance = 0, radius = 0;
    bool speculative = true;
    bool sensor = false;
    bool proxy_moved = false;
    Aabb fat{};
    Aabb aabb{};  // the shape's bounds at its body's last finalize or sweep
  };
  // Encoded body sim index (world_runtime.rs:903-925): awake index, static -(index + 2), and
  // -1 for a sleeping body.
  int32_t encode(const HBody& b) const { return b.set == 2 ? b.local : (b.set == 0 ? -(b.local + 2) : kNullIndex); }
  const BodySim& sim_of(int32_t body_id) const {
    const HBody& b = bodies_[(size_t)body_id];
    return b.set == 2 ? awake_host_[(size_t)b.local] : (b.set == 0 ? static_host_[(size_t)b.local] : sleep_host_[(size_t)body_id]);
  }

  std::vector<HBody> bodies_;
  std::vector<BodySim> awake_host_, static_host_;
  std::vector<BodySim> sleep_host_;  // by body id: the sims of sleeping bodies
  std::vector<BodyState> states_host_;
  std::vector<Shape> shapes_;
  std::vector<HShape> hshapes_;
  std::vector<BoxHull> hulls_;
  std::vector<Mat3> hull_inertia_;
  std::vector<float> hull_volume_;

 protected:
  virtual int32_t add_shape(int32_t body, Shape shape, const Material& m, float radius);
  void update_mass(int32_t body);
  Aabb shape_local_aabb(const Shape& s, Quat q) const;
};

}  // namespace nsl
