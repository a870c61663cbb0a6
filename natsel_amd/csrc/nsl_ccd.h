// Continuous collision records shared by the host world and the device code: sweeps,
// time-of-impact queries and results, the candidate lists of the swept pass, and the
// launchers (nsl_continuous.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "nsl_broadphase.h"
#include "nsl_device.h"

namespace nsl {

// Sweep (queries.rs:179-185): centre of mass from c1 to c2, rotation from q1 to q2.
struct Sweep {
  Vec3 local_center, c1, c2;
  Quat q1, q2;
};

enum ToiState : int32_t { kToiUnknown = 0, kToiFailed, kToiOverlapped, kToiHit, kToiSeparated };

// TOIOutput (queries.rs:224-234)
struct ToiOutput {
  int32_t state;
  Vec3 point, normal;
  float fraction, distance;
  int32_t distance_iterations, push_back_iterations, root_iterations;
};

// One query of the batched time-of-impact entry point: two point proxies and their sweeps.
struct ToiQuery {
  Vec3 points_a[8], points_b[8];
  int32_t count_a, count_b;
  float radius_a, radius_b;
  Sweep sweep_a, sweep_b;
  float max_fraction;
};

// The candidate shapes of the sweep and the per-shape filter data.
struct ContinuousArgs {
  const int32_t* candidates;  // static shape ids ascending, then kinematic, then dynamic
  int32_t n_static, n_kinematic, n_dynamic;
  const Proxy* proxies;       // by shape id
};

// Sweeps this step's fast bodies (the step kernel's fast_ids list) against the candidate
// shapes; `bullets` selects the bullet pass, which runs after the others. At most
// `max_fast` bodies are listed.
void launch_continuous(const DeviceWorld& d, const ContinuousArgs& a, const StepParams& p, bool bullets, int max_fast,
                       hipStream_t s, KernelTimer* timer);
// One lane per query.
void launch_time_of_impact(int n, const ToiQuery* queries, ToiOutput* out, hipStream_t s);

}  // namespace nsl
