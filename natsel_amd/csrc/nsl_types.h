// Device-resident layouts of the rigid-body step (HBM), shared by the host world
// and the gfx950 kernels.
//
// Body state keeps the reference's 64-byte BodyState (world_runtime.rs:209-226) so a
// 64-lane wavefront streams 4 KiB of contiguous state per body pass. Contacts,
// constraints and shapes are indexed by id; per-colour constraint slots are packed
// back to back so one launch (or one grid phase) covers one colour.
#pragma once
#include "nsl_math.h"

namespace nsl {

enum ShapeKind : int32_t { kShapeCapsule = 0, kShapeHull = 3, kShapeSphere = 5 };
enum BodyKind : int32_t { kBodyStatic = 0, kBodyKinematic = 1, kBodyDynamic = 2 };

// world_runtime.rs:81-118
namespace bflag {
constexpr uint32_t kLockLinX = 0x1, kLockLinY = 0x2, kLockLinZ = 0x4;
constexpr uint32_t kLockAngX = 0x8, kLockAngY = 0x10, kLockAngZ = 0x20;
constexpr uint32_t kIsFast = 0x40, kIsBullet = 0x80, kSpeedCapped = 0x100, kHadToi = 0x200;
constexpr uint32_t kAllowFastRotation = 0x400, kDynamic = 0x1000, kEnableSleep = 0x2000;
constexpr uint32_t kEnableRecycling = 0x4000, kDirtyMass = 0x8000, kGyroscopic = 0x10000;
constexpr uint32_t kFixedRotation = kLockAngX | kLockAngY | kLockAngZ;
constexpr uint32_t kTransient = kSpeedCapped | kHadToi;
}  // namespace bflag

// body_api.rs:22-40 (the subset the device reads or writes)
namespace cflag {
constexpr uint32_t kStatic = 0x8, kRecycle = 0x10;
constexpr uint32_t kSimTouching = 0x10000, kHitEvent = 0x100000;
constexpr uint32_t kRelativeTransformValid = 0x800000, kSpeculative = 0x1000000;
}  // namespace cflag

// Transitions reported by the collide kernel, one byte per contact id.
enum ContactChange : uint8_t { kChangeNone = 0, kChangeDisjoint = 1, kChangeBegan = 2, kChangeEnded = 3 };

struct BodyState {  // 64 B, world_runtime.rs:209-226
  Vec3 v; float pad1;
  Vec3 w; float pad2;
  Vec3 dp; uint32_t flags;
  Quat dq;
};
static_assert(sizeof(BodyState) == 64, "BodyState is 64 bytes");

// The solver-facing part of BodySim (world_runtime.rs:261-299).
struct BodySim {
  WXf transform;
  Pos center;
  Pos center0;
  Quat rotation0;
  Vec3 local_center;
  Vec3 max_extent;
  float min_extent;
  uint32_t flags;
  Vec3 force, torque;
  float inv_mass;
  Mat3 inertia;
  Mat3 inv_inertia_local;
  Mat3 inv_inertia_world;
  float linear_damping, angular_damping, gravity_scale;
  int32_t body_id;
  // body-level fields finalize needs (world_runtime.rs:301-343)
  float sleep_threshold, safety_factor, sleep_time, sleep_velocity;
  int32_t head_shape;   // first shape id of the body
  int32_t body_type;
  uint32_t body_flags;
  int32_t pad;
};

// Box hull geometry in local body space (hull.rs:736-852). Topology is the shared
// box table; only points, planes and bulk properties vary per hull.
struct BoxHull {
  Vec3 points[8];
  Plane planes[6];
  Aabb aabb;
  Vec3 center;
  float inner_radius;
};

struct Shape {
  int32_t kind;
  int32_t body_id;
  int32_t body_enc;       // awake index >= 0, static -(i+2)
  int32_t next_shape;
  int32_t hull;           // index into the hull table, or -1
  float radius;           // sphere / capsule radius
  Vec3 c1, c2;            // sphere centre (c1) or capsule segment
  float friction, restitution, rolling_resistance, aabb_margin;
  Vec3 tangent_velocity;
  uint32_t enable_hit_events;
};

struct SatCache { float separation; uint8_t type, index_a, index_b, hit; };
struct SimplexCache { float metric; uint16_t count; uint8_t index_a[4], index_b[4]; };

struct ManifoldPoint {  // collision_types.rs:5-17 (fields the step reads)
  Vec3 anchor_a, anchor_b;
  float separation, normal_impulse, total_normal_impulse, base_separation;
  uint32_t feature_id;
  float normal_velocity;
};

// One contact's device record (world_runtime.rs:517-555 plus its single manifold).
struct Contact {
  int32_t shape_a, shape_b;
  int32_t enc_a, enc_b;
  uint32_t flags;
  int32_t manifold_count;
  int32_t point_count;
  Vec3 normal;
  float twist_impulse;
  Vec3 friction_impulse;
  Vec3 rolling_impulse;
  ManifoldPoint points[4];
  Quat cached_rot_a, cached_rot_b;
  Xf cached_pose;
  SatCache sat;
  SimplexCache simplex;
  float friction, restitution, rolling_resistance;
  Vec3 tangent_velocity;
};

struct ConstraintPoint {  // contact_solver.rs:118-147
  Vec3 r_a, r_b;
  float base_separation, relative_velocity, normal_impulse, total_normal_impulse, restitution_impulse;
  float normal_mass, lever_arm;
};

// ContactConstraint + its single ManifoldConstraint (contact_solver.rs:149-240). 16-byte
// aligned with a size that is a multiple of 16: the step kernel copies a record to LDS in
// 16-byte pieces.
struct alignas(16) Constraint {
  int32_t contact_id, index_a, index_b, point_count;
  int32_t manifold_count;
  uint32_t is_static;
  float inv_mass_a, inv_mass_b;
  SymMat3 inv_ia, inv_ib, rolling_mass;
  float friction, restitution, rolling_resistance;
  Vec3 normal, tangent1, tangent2, center_a, center_b;
  float twist_mass, twist_impulse;
  Mat2 tangent_mass;
  Vec2 friction_impulse;
  Vec3 rolling_impulse;
  float tangent_velocity1, tangent_velocity2;
  ConstraintPoint points[4];
};

// Slots per block of the single world's blocked constraint-record layout (one wavefront).
constexpr int kRecBlock = 64;

// Sequential overflow-colour constraint (overflow_contact_solver.rs:17-33): full
// matrices, different arithmetic from the coloured path.
struct OverflowConstraint {
  int32_t contact_id, index_a, index_b, manifold_count;
  float inv_mass_a, inv_mass_b;
  Mat3 inv_ia, inv_ib, rolling_mass;
  Softness softness;
  float friction, restitution, rolling_resistance;
  int32_t point_count;
  Vec3 normal, tangent1, tangent2, center_a, center_b;
  float twist_mass, twist_impulse;
  Mat2 tangent_mass;
  Vec2 friction_impulse;
  Vec3 rolling_impulse;
  float tangent_velocity1, tangent_velocity2;
  ConstraintPoint points[4];
};

// A revolute joint (JointSim + RevoluteJointSim, world_runtime.rs / revolute_joint.rs:27-57),
// indexed by joint id. The definition fields are written once by the host; prepare fills
// the per-step fields; the impulses persist across steps for warm starting.
namespace jflag {
constexpr uint32_t kSpring = 0x1, kMotor = 0x2, kLimit = 0x4;
}  // namespace jflag

struct JointSim {
  // definition (types/joint_defs.rs:30-58, :154-167)
  int32_t enc_a, enc_b;  // encoded body sim index: awake >= 0, static -(index + 2)
  Xf local_frame_a, local_frame_b;
  float constraint_hertz, constraint_damping_ratio;
  float hertz, damping_ratio, target_angle, lower_angle, upper_angle, max_motor_torque, motor_speed;
  uint32_t flags;  // jflag
  // prepared each step
  Softness constraint_softness, spring_softness;
  float inv_mass_a, inv_mass_b;
  Mat3 inv_ia, inv_ib;
  int32_t index_a, index_b;
  int32_t fixed_rotation;
  float axial_mass;
  Xf frame_a, frame_b;
  Vec3 rotation_axis_z, perp_axis_x, perp_axis_y, delta_center;
  // accumulated impulses
  Vec3 linear_impulse;
  Vec2 perp_impulse;
  float spring_impulse, motor_impulse, lower_impulse, upper_impulse;
};

struct StepParams {
  float dt, inv_dt, h, inv_h;
  int32_t sub_steps;
  Softness contact_softness, static_softness;
  float contact_speed, max_linear_speed, restitution_threshold;
  Vec3 gravity;
  float warm_start_scale;
  uint32_t enable_sleep, enable_continuous, enable_restitution_propagation;
};

}  // namespace nsl
