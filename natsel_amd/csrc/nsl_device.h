// Device buffers of one world and the kernel launchers the host step drives.
#pragma once
#include <hip/hip_runtime.h>

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
  int32_t* collide_ids;     // contacts to collide this step
  int32_t* solve_ids;       // coloured contact ids, colour-major (slot order)
  int32_t* overflow_ids;    // overflow contact ids in graph order
  int32_t* counters;        // [0] fast bodies flagged, [1] sat calls, [2] sat cache hits, [3] recycled,
                            // [4] manifold points prepared (coloured), [5] overflow points
  int32_t n_awake;
};

// Kernel classes timed with HIP events when kernel timing is on.
enum KernelClass : int {
  kKCollide = 0, kKPrepare, kKIntegrateVelocities, kKWarmStart, kKPush, kKIntegratePositions, kKRelax,
  kKRestitution, kKStore, kKOverflow, kKFinalize, kKernelClassCount
};

// Event pairs bracketing launches on one stream; read back after the step's sync.
struct KernelTimer {
  hipEvent_t* events = nullptr;   // pool, 2 per mark
  int capacity = 0;               // marks
  int used = 0;
  int classes[4096];
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

struct ColorRanges {
  int32_t begin[kColorCount];
  int32_t end[kColorCount];
};

// kernels (nsl_kernels.hip)
// `timer` may be null (no timing).
void launch_init_contacts(const DeviceWorld& d, const int32_t* ids, const Contact* records, int n, hipStream_t s);
void launch_collide(const DeviceWorld& d, int n, hipStream_t s, KernelTimer* timer);
void launch_prepare(const DeviceWorld& d, int n, int n_overflow, const StepParams& p, hipStream_t s, KernelTimer* timer);
void launch_solve(const DeviceWorld& d, const ColorRanges& cr, int n_overflow, bool any_restitution, int restitution_iterations,
                  const StepParams& p, hipStream_t s, KernelTimer* timer);
void launch_finalize(const DeviceWorld& d, const StepParams& p, hipStream_t s, KernelTimer* timer);

}  // namespace nsl
