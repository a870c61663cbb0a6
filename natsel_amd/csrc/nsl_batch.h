// Batched stepper: thousands of independent worlds of up to 256 dynamic bodies each,
// stepped together on the device with no host work between steps.
//
// The reference steps one World at a time (World::step, solver_integration.rs:807); a
// batch holds many such worlds and steps each exactly as the reference would step it
// alone: the same pair order, contact ids, colours and float operations, so every
// world's state hash equals its standalone hash. A step is three launches over all
// worlds, one workgroup per world, with no host work in between: the pair update and
// contact creation (one wavefront), the collide pass, and the contact-state transitions
// and graph colouring (one lane, in contact-id order) followed by preparation, the
// substeps with the world's body states in LDS, impulse storage and finalisation (one
// wavefront). Worlds never synchronise with each other.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "nsl_broadphase.h"
#include "nsl_device.h"
#include "nsl_scene.h"

namespace nsl {

constexpr int kBatchMaxAwake = 256;    // dynamic bodies per world (their states sit in LDS)
constexpr int kBatchMaxShapes = 4096;  // shapes per world (pair keys pack two 16-bit ids)
constexpr int kBatchMaxNewPairs = 2048;  // new pairs per world and step (sorted in LDS)
constexpr int kBatchPairsBlock = 256;    // pair update: four wavefronts per world
constexpr int kPairsLdsShapes = 512;     // shapes per world whose fat AABBs the pair update keeps in LDS
constexpr int kBatchLanes = 64;        // one wavefront per world (pair update, solve)
constexpr int kBatchCollideBlock = 128;  // collide pass: one workgroup per world

// Where one world's records live in the concatenated device arrays.
struct BatchWorldDesc {
  int32_t awake_off, n_awake;
  int32_t static_off, n_static;
  int32_t body_off, n_bodies;
  int32_t shape_off, n_shapes;
  int32_t hull_off;
  int32_t contact_off, contact_cap;    // contacts, constraints, solve ids, free ids, non-touching list
  int32_t overflow_off, overflow_cap;  // overflow constraints and list
  int32_t color_off, color_cap;        // lists of colours 0..22: [23][color_cap] contact ids
  int32_t bits_off, bits_words;        // body bits of colours 0..22: [23][bits_words] u32
  int32_t pair_off;                    // pair-exists bits: n_shapes * n_shapes bits (u32 words)
  int32_t n_ccd_static, n_ccd_kinematic, n_ccd_dynamic;  // continuous-collision candidates at shape_off
};

// The integer bookkeeping of one world (what the single-world stepper keeps on the host).
struct BatchWorldState {
  int32_t next_id;                      // contact id pool (id_pool.rs:28-46)
  int32_t n_free;
  int32_t n_non_touching;               // awake non-touching contacts (collide_pass.rs:743-774)
  int32_t color_count[kOverflowColor];  // contacts in colours 0..22
  int32_t n_overflow;
  int32_t error;                        // non-zero: a capacity was exceeded; the world stops
  int32_t steps;                        // steps taken
};

enum BatchError : int32_t { kBatchOk = 0, kBatchContactsFull = 1, kBatchOverflowFull = 2, kBatchPairsFull = 3 };

struct BodyMeta {  // per body id: type, solver-set encoding, flags
  int32_t type, enc;
  uint32_t flags;
};

struct ContactMeta {  // per contact id: shapes, graph colour and index, liveness
  int32_t shape_a, shape_b, color, local;
  int32_t alive;
  float restitution;
  int32_t pair_type;  // the collide routine: see kPairTypes
  int32_t body_a, body_b;  // the shapes' bodies (world body ids)
  int32_t dynamic;         // bit 0: body_a is dynamic, bit 1: body_b is dynamic
};

// Collide routines by primary shape (hull > capsule > sphere, contact_creation.rs:88-102):
// 0 hull-hull, 1 hull-sphere, 2 hull-capsule, 3 capsule-sphere, 4 capsule-capsule,
// 5 sphere-sphere. The collide launch takes each routine's contacts together.
constexpr int kPairTypes = 6;

// Recording of a batched rollout for its adjoint (nsl_adjoint.hip): per step and world
// the prepared constraints in solve order (warm-start impulses included), their points'
// feature ids and shape pair, the links to the previous step's constraints that carry
// impulses, the colour ranges, and the body states and centres at the step's start.
constexpr int kTapeColourWords = 1 + 2 * kOverflowColor;
constexpr int kTapeBranchBytes = 10;  // push points 0..3, relax points 0..3, twist, cone
constexpr uint8_t kTapeNoPoint = 0xFF;
struct BatchTape {
  Constraint* cons;     // [steps][worlds][max_con]
  uint32_t* features;   // [steps][worlds][max_con][4]
  uint32_t* pairs;      // [steps][worlds][max_con]: (shape_a << 16) | shape_b
  int32_t* links;       // [steps][worlds][max_con]: same contact's slot in the previous step, or -1
  uint8_t* pmap;        // [steps][worlds][max_con][4]: the matching previous point per point (kTapeNoPoint: none)
  int32_t* count;       // [steps][worlds]: constraints; -1: the step is outside the model
  int32_t* colours;     // [steps][worlds][kTapeColourWords]: n, begin[n], end[n] of the non-empty colours
  BodyState* states;    // [steps + 1][worlds][max_bodies]
  Pos* centers;         // [steps + 1][worlds][max_bodies]
  uint8_t* branches;    // [steps][worlds][sub_steps][max_con][kTapeBranchBytes]: the adjoint's recompute's branch record
  int32_t max_con, max_bodies, steps, n_worlds, sub_steps;
};

// Device arrays of a batch, passed by value to the kernel.
struct BatchDevice {
  BodySim* sims;
  BodyState* states;
  BodySim* static_sims;
  BodyMeta* body_meta;
  Shape* shapes;
  Aabb* fat_aabbs;
  Aabb* shape_aabbs;   // per shape: bounds from its body's last finalize or sweep
  int32_t* candidates; // per world at shape_off: continuous-collision candidates (static, kinematic, dynamic)
  int32_t* fast_ids;   // this step's fast bodies (awake index into the batch arrays), counters[kCntFastList] long
  int32_t* awake_world;  // per awake body of the batch: its world
  uint8_t* shape_moved;
  Proxy* proxies;  // pad holds the shape's enable_speculative flag
  BoxHull* hulls;
  Contact* contacts;
  uint8_t* contact_change;
  ContactMeta* contact_meta;
  Constraint* constraints;
  int32_t* solve_ids;
  OverflowConstraint* overflow;
  int32_t* overflow_ids;
  int32_t* free_ids;
  int32_t* non_touching;
  int32_t* overflow_list;
  int32_t* color_lists;
  uint32_t* color_bits;
  uint32_t* pair_bits;
  BatchWorldDesc* worlds;
  BatchWorldState* wstate;
  int32_t* order;     // worlds in launch order (most bodies first)
  int2* collide_items;     // this step's live contacts as (world, id): kPairTypes lists of collide_type_cap
  int32_t* collide_counts; // [kPairTypes] list lengths (written by the pair launch, cleared by the solve launch)
  int32_t collide_type_cap;
  int32_t* counters;  // as DeviceWorld::counters
  unsigned long long* work;  // [0] contacts collided, [1] constraints prepared, [2] their points, [3] overflow constraints
  int32_t n_worlds;
  int32_t enable_speculative;
  int32_t restitution_iterations;
  int32_t enable_continuous;
  int32_t any_bullets;  // a dynamic bullet body in some world: the bullet pass runs
  int32_t collide_blocks;  // workgroups of the collide launch (as many as are resident at once)
  BatchTape tape;
  int32_t tape_step;  // step of the tape to record this launch (-1: not recording)
};

void launch_batch_step(const BatchDevice& b, const StepParams& p, int steps, hipStream_t s);

// Launch shape of a batch step (results do not depend on it): the collide launch's grid
// (1..16 times the resident workgroups; the launch strides over the live contacts),
// and whether the steps of a call replay a captured HIP graph of kBatchGraphSteps steps.
struct BatchTuning {
  int32_t collide_blocks = 0;  // 0: as many as are resident at once
  int32_t use_graph = 0;
  // filled by autotune(): candidates timed, the best's and the starting shape's device
  // time per step (ms)
  int32_t candidates = 0;
  double step_ms = 0.0, previous_ms = 0.0;
};
constexpr int kBatchGraphSteps = 8;

class BatchStepper {
 public:
  explicit BatchStepper(int device);
  ~BatchStepper();
  BatchStepper(const BatchStepper&) = delete;
  BatchStepper& operator=(const BatchStepper&) = delete;

  WorldOptions options;  // shared by every world

  // Worlds and their bodies and shapes can be added until the first step.
  int32_t add_world();
  HostScene& scene(int32_t w);
  int32_t world_count() const { return (int32_t)scenes_.size(); }
  int32_t body_count(int32_t w) const { return (int32_t)scenes_.at(w)->bodies_.size(); }

  void step(float dt, int32_t sub_steps, int32_t steps);

  // Launch shape: set, read, or chosen by timing `steps` steps per candidate from a
  // snapshot of every device array of the batch, which is restored afterwards (so the
  // batch's state and results are as before the call; every candidate must reach the same
  // world hashes).
  void set_tuning(const BatchTuning& t);
  BatchTuning tuning() const { return tuning_; }
  BatchTuning autotune(float dt, int32_t sub_steps, int32_t steps);

  // World w's reference state hash (body_queries.rs:432), for every world.
  void state_hashes(uint64_t* out);
  void get_bodies(int32_t w, double* pos, float* rot, float* v, float* wv);
  // Worlds stopped by a full capacity, and the first such world's error code.
  int32_t error_worlds(int32_t* first_error);
  // World w's contact slots (its contact capacity) and, per slot, graph colour (-1: not in
  // the graph), index in its list and liveness.
  int32_t contact_slots(int32_t w) const;
  void get_contact_colors(int32_t w, int32_t* color, int32_t* local, int32_t* alive);
  // Per shape of world w: fat AABB (lower x, y, z, upper x, y, z) and the moved flag.
  void get_shape_bounds(int32_t w, float* fat, uint8_t* moved);
  // Per contact slot: its two shape ids (-1 when the slot is free).
  void get_contact_shapes(int32_t w, int32_t* shape_a, int32_t* shape_b);
  // Device time of the last step call (HIP events), and totals since creation.
  float last_step_ms() const { return last_ms_; }
  int64_t total_body_steps() const { return body_steps_; }
  int64_t total_bodies() const { return total_awake_; }
  int64_t contact_capacity() const { return total_contact_cap_; }
  // Work since creation: contacts collided, constraints prepared, their manifold points,
  // overflow constraints.
  void work_counts(int64_t out[4]);

  // Differentiable rollout (nsl_adjoint.hip). record() steps like step() while writing the
  // tape (at most max_constraints constraints and kBatchLanes dynamic bodies per world).
  // adjoint() returns per world the loss |com - target|^2 of the dynamic bodies' centre of
  // mass after the recorded steps, and its gradient with respect to each dynamic body's
  // velocity and angular velocity at the first recorded step and its shape's friction; a
  // world's status is 0, or 1 when a recorded step lies outside the differentiable model,
  // 2 when the recompute departed from the recording. model_losses() runs the model in
  // double precision from the tape with one parameter perturbed per case (parameter
  // index = awake body * 7 + {0..2: v, 3..5: w, 6: friction}), for finite differences.
  void record(float dt, int32_t sub_steps, int32_t steps, int32_t max_constraints);
  void adjoint(const float* targets, float* loss, float* grad_v, float* grad_w, float* grad_friction, int32_t* status);
  void model_losses(const int32_t* worlds, const int32_t* params, const double* deltas, int32_t n, const float* targets,
                    double* out, bool frozen);
  float last_record_ms() const { return record_ms_; }
  float last_adjoint_ms() const { return adjoint_ms_; }
  int32_t tape_steps() const { return tape_.steps; }

 private:
  void upload();
  void pull();
  template <class T> T* alloc(size_t n, bool zero);
  void launch_steps(const StepParams& p, int32_t steps);
  void drop_graph();
  std::vector<void*> snapshot();                   // device copies of every allocation
  void restore(const std::vector<void*>& copies);  // and back (the copies are kept)
  StepParams step_params(float dt, int32_t sub_steps) const;
  void free_tape();

  int device_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::vector<std::unique_ptr<HostScene>> scenes_;
  std::vector<BatchWorldDesc> desc_;
  BatchDevice d_{};
  std::vector<void*> allocations_;
  std::vector<size_t> allocation_bytes_;
  int32_t resident_collide_blocks_ = 0;
  BatchTuning tuning_{};
  hipGraphExec_t graph_exec_ = nullptr;
  StepParams graph_params_{};
  int32_t graph_steps_ = 0, graph_collide_blocks_ = 0;
  bool uploaded_ = false;
  bool device_newer_ = false;
  float last_ms_ = 0.0f;
  int64_t body_steps_ = 0;
  int64_t total_awake_ = 0;
  int64_t total_contact_cap_ = 0;
  BatchTape tape_{};
  StepParams tape_params_{};
  std::vector<void*> tape_allocs_;
  int32_t tape_max_colours_ = 0;
  bool tape_branches_ = false;  // the branch record is filled (by adjoint())
  float record_ms_ = 0.0f, adjoint_ms_ = 0.0f;
};

}  // namespace nsl
