// Device buffers of one world and the kernel launchers the host step drives.
#pragma once
#include <hip/hip_runtime.h>

#include "nsl_pairset.h"
#include "nsl_types.h"

namespace nsl {

// Raw device pointers, passed by value to every kernel.
struct DeviceWorld {
  BodySim* sims;            // awake bodies, indexed by awake index
  BodyState* states;        // awake bodies
  BodySim* static_sims;     // static bodies
  Shape* shapes;            // by shape id
  Aabb* fat_aabbs;          // by shape id
  uint8_t* shape_moved;     // by shape id: fat AABB enlarged this step
  BoxHull* hulls;
  Contact* contacts;        // by contact id
  uint8_t* contact_change;  // by contact id: ContactChange of this step
  Constraint* constraints;  // coloured slots, colour-major
  OverflowConstraint* overflow;
  int32_t* collide_ids;     // unused by the single world (its contacts carry contact_active)
  uint8_t* contact_active;  // by contact id: 1 for a contact of the awake set (collided each step)
  int32_t* solve_ids;       // coloured contact ids, colour-major (slot order)
  int32_t* overflow_ids;    // overflow contact ids in graph order
  int32_t* counters;        // kCounterCount work counters, see kCnt*
  PairSet pairs;            // keys of the shape pairs that have a contact
  JointSim* joints;         // by joint id
  int32_t* joint_ids;       // joints in the graph, colour-major; the overflow colour's last
  PairSet joint_pairs;      // body-pair keys of joints without collide_connected (no contact allowed)
  Aabb* shape_aabbs;        // by shape id: the shape's bounds from its body's last finalize or sweep
  int32_t* fast_ids;        // this step's fast bodies as fast_base + awake index (count: counters[kCntFastList])
  int32_t fast_base;        // 0 for a world; the world's offset into a batch's awake bodies
  int32_t n_awake;
  int32_t* change_list;     // contacts whose state changed in k_collide: id << 2 | ContactChange
  float* sleep_times;       // per awake index: the body's sleep time after finalize (null: not kept)
  BodySim* sleep_sims;      // by body id: sims of sleeping bodies (encoding -1)
};

// Device work counters (DeviceWorld::counters).
enum : int {
  kCntFastBodies = 0,   // bodies flagged fast by finalize (cumulative)
  kCntSatCalls = 1,
  kCntSatHits = 2,
  kCntRecycled = 3,
  kCntPoints = 4,       // manifold points prepared (coloured)
  kCntOverflowPoints = 5,
  kCntMovedProxies = 6, // proxies moved (fat AABB enlarged) since the last pair update
  kCntFastList = 7,     // length of this step's fast_ids list (reset by the step kernel)
  kCntToiCalls = 8,     // time-of-impact queries (cumulative)
  kCntToiHits = 9,      // bodies moved back to an impact (cumulative)
  kCntChangeList = 10,  // length of k_collide's change_list (reset by the host before it)
  kCounterCount = 16
};

// Launches timed with HIP events when kernel timing is on.
enum KernelClass : int { kKInit = 0, kKCollide, kKSolve, kKContinuous, kKPrepare, kKFinalize, kKernelClassCount };

// Phases inside the step kernel; timed with the device wall clock at each grid barrier.
enum PhaseClass : int {
  kPPrepare = 0, kPIntegrateVelocities, kPWarmStart, kPPush, kPIntegratePositions, kPRelax, kPRestitution,
  kPOverflow, kPStoreFinalize, kPhaseClassCount
};

// Event pairs bracketing launches on one stream; read back after the step's sync.
struct KernelTimer {
  hipEvent_t* events = nullptr;   // pool, 2 per mark
  int capacity = 0;               // marks
  int used = 0;
  int classes[64];
  void begin(int cls, hipStream_t s) {
    if (used < capacity) {
      classes[used] = cls;
      (void)hipEventRecord(events[2 * used], s);
    }
  }
  void end(hipStream_t s) {
    if (used < capacity) {
      (void)hipEventRecord(events[2 * used + 1], s);
      used += 1;
    }
  }
};

// Grid barrier state of the step kernel (zeroed at creation; count holds the arrivals of
// the current launch, zeroed by k_prepare before each).
struct GridBarrier {
  uint32_t count;
  uint32_t gen;  // unused
  uint32_t timed_out;
  uint32_t pad;
};

struct ColorRanges {
  int32_t begin[kColorCount];
  int32_t end[kColorCount];
};

// One item's warm-start velocity change for one of its bodies.
struct WarmDelta {
  Vec3 dv, dw;
};

// Everything the step kernel needs beyond the world pointers.
struct SolvePlan {
  ColorRanges cr;                 // coloured slots per colour (colour-major)
  ColorRanges jr;                 // joint_ids per colour, overflow colour included
  int32_t n_joints;               // joints in the graph (prepared each step)
  // The overflow colour's joints and contacts in dependency levels: an item's level is one
  // more than the highest level of the earlier items (joints first, then contacts, in list
  // order) that share a dynamic body with it. Items of a level touch disjoint bodies, so
  // running the levels in order with the items of each in parallel reproduces the
  // sequential walk exactly. Items: contact k >= 0 (overflow[k]), joint id -(id + 1).
  const int32_t* ovf_items;
  const int32_t* ovf_level_begin; // [n_ovf_levels + 1]
  int32_t n_ovf_levels;
  int32_t n_constraints;          // coloured slots in total
  int32_t n_overflow;             // overflow constraints, solved in list order by one lane
  int32_t any_restitution;
  int32_t restitution_iterations;
  int32_t small_colour_items;     // runs of colours this small go to workgroup 0 (0: off)
  int32_t stage_records;          // push and relax read their records from LDS copies made a colour ahead
  int32_t n_work;                 // the step kernel's workgroups (the barrier's arrivals per barrier)
  // Warm start by body (every awake body dynamic): ws_table[c * n_awake + i] is the item of
  // colour c (0..22) that touches awake body i, with the body's side s (0: a, 1: b): a
  // contact slot as (slot << 1 | s), a joint as -((joint id << 1 | s) + 2), or -1. k_prepare
  // fills it (the host sets it to -1 first). ws_delta holds each item's velocity change per
  // side: contact slots first (2 * slot + s), then joints by id (2 * (n_constraints + id) + s).
  // Null: warm start colour by colour.
  int32_t* ws_table;
  WarmDelta* ws_delta;
  // one zero WarmDelta after the items' (k_prepare clears it): the body phase loads it for a
  // colour the body has no item in, so all of a body's loads issue together
  WarmDelta* ws_zero;
  // With warm start by body and an overflow colour: per awake body, -1 for a body outside the
  // overflow colour (its velocity is integrated by any lane), else the last substep + 1 whose
  // velocity workgroup 0 integrated (0: none yet). The host sets -1; k_prepare marks.
  int32_t* ovf_done;
  // The overflow colour's dynamic bodies (awake indices, each once): with warm start by body,
  // workgroup 0 does their per-body work before the overflow push and relax levels.
  const int32_t* ovf_bodies;
  int32_t n_ovf_bodies;
  GridBarrier* barrier;
  uint64_t* phase_log;            // null: no phase timing; else [2 * kMaxPhaseLog]: (class, wall clock)
};
constexpr int kMaxPhaseLog = 1024;

// Location patches after bodies changed solver sets (nsl_kernels.hip): `items` holds
// n_shapes (shape id, encoding) pairs, then n_contacts (contact id, enc_a, enc_b, awake-set
// flag) quads and n_joints (joint id, enc_a, enc_b) triples; `sims` holds n_sims sims copied
// to sleep_sims[sim.body_id].
void launch_patch_locations(const DeviceWorld& d, const int32_t* items, int n_shapes, int n_contacts, int n_joints,
                            const BodySim* sims, int n_sims, hipStream_t s);

// kernels (nsl_kernels.hip)
// `timer` may be null (no timing).
void launch_init_contacts(const DeviceWorld& d, const int32_t* ids, const Contact* records, int n, hipStream_t s);
void launch_collide(const DeviceWorld& d, int n, hipStream_t s, KernelTimer* timer);
void launch_pairset_insert(PairSet set, const uint64_t* keys, int n, hipStream_t s);
void launch_pairset_erase(PairSet set, const uint64_t* keys, int n, hipStream_t s);
// Prepare of joints, contact and overflow constraints, one lane per item (before the step kernel).
void launch_prepare(const DeviceWorld& d, const SolvePlan& plan, const StepParams& p, hipStream_t s, KernelTimer* timer);
// One launch (all workgroups resident): the substep schedule and restitution.
void launch_solve_step(const DeviceWorld& d, const SolvePlan& plan, const StepParams& p, bool cooperative, hipStream_t s,
                       KernelTimer* timer);
// One wavefront's part of a colour's residual (k_colour_residual): the sum and the largest
// of its points' normal impulse changes and the largest friction impulse change.
struct ResidualPartial {
  float sum_normal, max_normal, max_friction, pad;
};
// The residual sweep of one colour's coloured slots [begin, begin + count) on the body states
// of `d` (a copy): one partial per 64 slots, in slot order, to `out`.
void launch_colour_residual(const DeviceWorld& d, int begin, int count, const StepParams& p, ResidualPartial* out,
                            hipStream_t s);
// Store of contact and overflow impulses and finalize of the bodies, one lane per item (after
// the step kernel; finalize lists the fast bodies for the continuous pass).
void launch_store_finalize(const DeviceWorld& d, const SolvePlan& plan, const StepParams& p, hipStream_t s,
                           KernelTimer* timer);
// Workgroups of the step kernel that fit on the device at once.
int solve_step_max_blocks(int device);
constexpr int kSolveBlock = 128;  // measured: 128 lanes beat 64, 256 and 512 on the 10k pile

}  // namespace nsl
