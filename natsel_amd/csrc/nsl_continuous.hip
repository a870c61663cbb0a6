// Continuous collision launches (see nsl_continuous.h): the swept pass of one world's fast
// bodies after the step kernel, and batched time-of-impact queries.
#include "nsl_continuous.h"

namespace nsl {

namespace {

constexpr int kContinuousBlock = 64;  // one wavefront: the sweeps diverge lane by lane

// The fast bodies finalize listed this step. Without `bullets`: non-bullet fast bodies
// against the static shapes, each followed by its world inverse inertia at the (possibly
// moved back) rotation, as the finalize tail does in body order (solver_integration.rs:538-
// 566); each body reads only static shapes and writes only itself, so the order is free.
// With `bullets`: the bullets, against every shape type, after the other bodies are final
// (solver_integration.rs:948-995); bullets skip each other.
__global__ __launch_bounds__(kContinuousBlock) void k_continuous(DeviceWorld d, ContinuousArgs a, StepParams p,
                                                                 int bullets) {
  const int n = d.counters[kCntFastList];
  int hits = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = d.fast_ids[k];
    const bool is_bullet = (d.sims[i].flags & bflag::kIsBullet) != 0;
    if (is_bullet != (bullets != 0)) continue;
    hits += solve_continuous_body(d, a, i, p) ? 1 : 0;
    if (!bullets) {
      BodySim& sim = d.sims[i];
      Mat3 rm = matrix_from_quat(sim.transform.q);
      sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
    }
  }
  if (hits) atomicAdd(&d.counters[kCntToiHits], hits);
}

// One query per lane (time_of_impact.rs:406-598).
__global__ __launch_bounds__(kContinuousBlock) void k_time_of_impact(int n, const ToiQuery* q, ToiOutput* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const ToiQuery& x = q[k];
  const ToiProxy pa{x.points_a, x.count_a, x.radius_a};
  const ToiProxy pb{x.points_b, x.count_b, x.radius_b};
  // the proxies read their points from the query record in global memory
  out[k] = time_of_impact(pa, x.sweep_a, pb, x.sweep_b, x.max_fraction);
}

}  // namespace

void launch_continuous(const DeviceWorld& d, const ContinuousArgs& a, const StepParams& p, bool bullets, int max_fast,
                       hipStream_t s, KernelTimer* t) {
  if (max_fast <= 0) return;
  const int blocks = std::min((max_fast + kContinuousBlock - 1) / kContinuousBlock, 1024);
  if (t) t->begin(kKContinuous, s);
  hipLaunchKernelGGL(k_continuous, dim3(blocks), dim3(kContinuousBlock), 0, s, d, a, p, bullets ? 1 : 0);
  if (t) t->end(s);
}

void launch_time_of_impact(int n, const ToiQuery* queries, ToiOutput* out, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_time_of_impact, dim3((n + kContinuousBlock - 1) / kContinuousBlock), dim3(kContinuousBlock), 0,
                       s, n, queries, out);
}

}  // namespace nsl
