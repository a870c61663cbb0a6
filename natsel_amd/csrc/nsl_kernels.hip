// gfx950 kernels of the single-world rigid-body step: contact initialisation, the
// narrow phase (collide), preparation, the step kernel (the coloured soft-step stages,
// the overflow colour's levels, integration, restitution) and impulse storage with
// finalisation.
// The per-item functions, with their reference citations, are in nsl_stages.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "nsl_device.h"
#include "nsl_joint.h"
#include "nsl_pairset.h"
#include "nsl_stages.h"

namespace nsl {

namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------------------------
// contact records created by the host's pair update

__global__ void k_init_contacts(DeviceWorld d, const int32_t* ids, const Contact* records, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const Contact& c = records[t];
  d.contacts[ids[t]] = c;
  d.contact_active[ids[t]] = 1;  // a contact made in the disabled set is switched off by the host's next patch
  pairset_insert(d.pairs, pair_key(c.shape_a, c.shape_b));
}

// One item per thread: shape encodings, contact encodings with the awake-set flag, joint
// encodings, sleeping sims.
__global__ void k_patch_locations(DeviceWorld d, const int32_t* items, int n_shapes, int n_contacts, int n_joints,
                                  const BodySim* sims, int n_sims) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_shapes) {
    d.shapes[items[2 * t]].body_enc = items[2 * t + 1];
    return;
  }
  t -= n_shapes;
  const int32_t* quad = items + 2 * n_shapes;
  if (t < n_contacts) {
    const int32_t id = quad[4 * t];
    Contact& c = d.contacts[id];
    c.enc_a = quad[4 * t + 1];
    c.enc_b = quad[4 * t + 2];
    d.contact_active[id] = (uint8_t)quad[4 * t + 3];
    return;
  }
  t -= n_contacts;
  const int32_t* tri = quad + 4 * n_contacts;
  if (t < n_joints) {
    JointSim& j = d.joints[tri[3 * t]];
    j.enc_a = tri[3 * t + 1];
    j.enc_b = tri[3 * t + 2];
    return;
  }
  t -= n_joints;
  if (t < n_sims) d.sleep_sims[sims[t].body_id] = sims[t];
}

__global__ void k_pairset_insert(PairSet set, const uint64_t* keys, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) pairset_insert(set, keys[t]);
}

__global__ void k_pairset_erase(PairSet set, const uint64_t* keys, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) pairset_erase(set, keys[t]);
}

// Every contact id below n whose awake-set flag is on: the awake set's contacts, touching or
// not (collide_pass.rs:175-190; each is independent, so the order of the reference's list
// does not matter). A disjoint contact turns its own flag off: the host destroys it.
__device__ __forceinline__ void collide_item(const DeviceWorld& d, int n) {
  const int32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n || !d.contact_active[id]) return;
  const uint8_t change = collide_contact(d, id);
  if (change != kChangeNone) d.change_list[atomicAdd(&d.counters[kCntChangeList], 1)] = (id << 2) | change;
  if (change == kChangeDisjoint) {
    d.contact_active[id] = 0;
    const Contact& c = d.contacts[id];
    pairset_erase(d.pairs, pair_key(c.shape_a, c.shape_b));  // the host destroys this contact
  }
}
__global__ __launch_bounds__(kBlock) void k_collide(DeviceWorld d, int n) { collide_item(d, n); }
// The same with the register budget cut to two waves per SIMD (spills, twice the resident
// waves): for large contact lists.
__global__ __launch_bounds__(kBlock, 2) void k_collide_wide(DeviceWorld d, int n) { collide_item(d, n); }

// ------------------------------------------------------------------------------
// the step kernel: one launch per step, all workgroups resident

// Polls of a barrier wait before the kernel gives up (a few seconds; far above any phase).
constexpr uint32_t kBarrierSpinLimit = 1u << 22;

// The workgroups of a step-kernel launch: their number (n) and this workgroup's (wg).
struct StepGrid {
  GridBarrier* bar;
  uint32_t n;
  int wg;
};

// Grid-wide barrier over a launch whose workgroups are all resident. Arrivals count up
// for the whole launch (k_prepare zeroes the count before it): barrier k of the launch is
// complete when the count reaches (k + 1) * n, which each workgroup derives from the count
// it saw on arrival, so the last arrival's own increment releases the waiters. Each
// workgroup's agent-scope release before it arrives and acquire after the barrier make every
// store issued before the barrier visible to every load after it, across the XCDs' L2
// caches. (Keeping the grid on one XCD with only an L1 invalidation measured faster but
// lost updates: an invalidation at workgroup scope leaves the CU's L1 lines in place; with
// an agent-scope acquire it was slower than the grid over all XCDs.) A wait that exceeds
// the spin limit raises `timed_out` and every workgroup leaves the kernel at its next
// barrier; returns false then.
__device__ __forceinline__ bool grid_sync(const StepGrid& g) {
  __shared__ int s_ok;
  GridBarrier* b = g.bar;
  // every store of this wave is issued before the wait (the wait alone does not keep the
  // compiler from sinking stores below it), and has reached L2 after it
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const uint32_t arrived = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t target = (arrived / g.n + 1u) * g.n;
    if (arrived + 1u != target) {
      uint32_t spins = 0;
      while (__hip_atomic_load(&b->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        ++spins;
        if (spins > kBarrierSpinLimit ||
            ((spins & 1023u) == 0 && __hip_atomic_load(&b->timed_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          __hip_atomic_store(&b->timed_out, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Sum of one value over the workgroup, added to *dst by one lane.
__device__ void block_add(int32_t* dst, int v) {
  __shared__ int s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v != 0) atomicAdd(&s_sum, v);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum != 0) atomicAdd(dst, s_sum);
}

struct PhaseClock {
  uint64_t* log;
  int n;
  bool lead;  // working workgroup 0
  __device__ __forceinline__ void mark(int cls) {
    if (log != nullptr && lead && threadIdx.x == 0 && n < kMaxPhaseLog) {
      log[2 * n] = (uint64_t)cls;
      log[2 * n + 1] = wall_clock64();
    }
    ++n;
  }
};

// The non-empty colours in order with their contact-slot and joint ranges, kept in LDS (a
// table indexed at run time in private memory would cost a memory round trip per phase).
struct ColourTable {
  int n;
  int begin[kOverflowColor];
  int end[kOverflowColor];
  int jbegin[kOverflowColor];
  int jend[kOverflowColor];
  // group_end[a] > a + 1: colours a..group_end[a]-1 each have at most the plan's
  // small_colour_items items and run on workgroup 0, a workgroup barrier between them
  int group_end[kOverflowColor];
};

// Groups runs of two or more consecutive small colours (items counted with joints).
__device__ void group_small_colours(ColourTable& t, int small_items) {
  for (int a = 0; a < t.n;) {
    int b = a;
    while (b < t.n && (t.end[b] - t.begin[b]) + (t.jend[b] - t.jbegin[b]) <= small_items) ++b;
    if (b - a >= 2) {
      for (int c = a; c < b; ++c) t.group_end[c] = b;
      a = b;
    } else {
      t.group_end[a] = a + 1;
      a += 1;
    }
  }
}

// A colour's items are its joints (in joint_ids order) and then its contact slots: the
// reference's unit order within a colour (solver_stages.rs:303-345). The two touch
// disjoint bodies, so lanes may take them in any mix.
__device__ __forceinline__ int joints_in(const ColourTable& t, int a, bool with_joints) {
  return with_joints ? t.jend[a] - t.jbegin[a] : 0;
}

// Body indices of a lane's first contact slot in active colour `a` ({-1, -1} when its
// first item is a joint or lies past the end).
__device__ __forceinline__ int2 first_indices(const DeviceWorld& d, const ColourTable& t, int a, int tid, bool with_joints) {
  const int nj = joints_in(t, a, with_joints);
  const int slot = t.begin[a] + tid - nj;
  if (tid < nj || slot >= t.end[a]) return make_int2(kNullIndex, kNullIndex);
  return blocked_cons(d).indices(slot);
}

// Copies the record of slot `slot` into `rec` (kRecPieces pieces of 16 bytes, piece j of this
// lane at rec[j * kSolveBlock + threadIdx.x]) with global_load_lds: the copy goes from HBM to
// LDS without registers and completes by the next grid barrier (its wait for all memory
// operations), so it overlaps the rest of the phase. Lanes with consecutive slots read one
// contiguous KiB per piece (the blocked record layout).
__device__ __forceinline__ void stage_record(const BlockedCons& cons, int slot, float4* rec) {
  const int wave_base = (int)threadIdx.x & ~63;
#pragma unroll
  for (int j = 0; j < kRecPieces; ++j)
    __builtin_amdgcn_global_load_lds(cons.piece(slot, j), rec + j * kSolveBlock + wave_base, 16, 0, 0);
}

// One stage over the active colours in order, a grid barrier after each colour or group
// of small colours. A lane's body indices for its first contact slot in the next colour are
// read before the barrier, so each phase opens with its body-state loads. `joint` runs the
// colour's joints when `with_joints` (warm start, push and relax; restitution has no joint
// rows). A group of small colours runs on workgroup 0 alone, in colour order with a
// workgroup barrier between colours (the workgroup shares one L1, so its stores are visible
// to its loads after the barrier): the same items in the same order, fewer grid barriers.
// With kStage (`rec`: two LDS buffers of kRecPieces * kSolveBlock pieces; the world has
// contact slots), a lane's first contact item of a colour reads its record from LDS: the
// lane copied it there during the previous colour's phase (a lane owns the same slot all
// launch and is the only writer of its record, so the copy is current). `item(slot, ia, ib,
// staged, hook)`: `staged` is the LDS copy or null, `hook` issues the next colour's copy
// after the item's body-state loads. Every lane issues the copy (a lane with no item in the
// next colour copies slot 0 into its unused buffer): with no branch around it the item's
// wait for its body states does not also wait for the copy.
template <bool kStage, class JointItem, class Item>
__device__ __forceinline__ bool run_colours(const DeviceWorld& d, const ColourTable& t, int tid, int nthreads, const StepGrid& grid, PhaseClock& clock,
                            int cls, bool with_joints, JointItem joint, Item item, float4* rec = nullptr) {
  const int n_active = t.n;
  if (n_active == 0) return true;
  int2 pf = first_indices(d, t, 0, tid, with_joints);
  int cur = 0;             // rec buffer of this colour's copy
  bool staged = false;     // the lane's first contact item of colour a is in rec[cur]
  for (int a = 0; a < n_active;) {
    const int g = t.group_end[a];
    if (g > a + 1) {
      if (grid.wg == 0) {
        for (int c = a; c < g; ++c) {
          const int nj = joints_in(t, c, with_joints);
          const int jb = t.jbegin[c], cb = t.begin[c] - nj;
          const int n_items = nj + (t.end[c] - t.begin[c]);
          for (int k = (int)threadIdx.x; k < n_items; k += (int)blockDim.x) {
            if (k < nj) {
              joint(d.joint_ids[jb + k]);
            } else {
              const int slot = cb + k;
              {
                const int2 ix = blocked_cons(d).indices(slot);
                item(slot, ix.x, ix.y, nullptr, NoHook{});
              }
            }
          }
          if (c + 1 < g) __syncthreads();
        }
      }
      a = g;
      staged = false;
    } else {
      const int nj = joints_in(t, a, with_joints);
      const int jb = t.jbegin[a], cb = t.begin[a] - nj;
      const int n_items = nj + (t.end[a] - t.begin[a]);
      // the lane's first contact slot of the next colour when that colour runs on its own
      int next_slot = -1;
      if (kStage && a + 1 < n_active && t.group_end[a + 1] == a + 2) {
        const int nj2 = joints_in(t, a + 1, with_joints);
        const int s2 = t.begin[a + 1] + tid - nj2;
        if (tid >= nj2 && s2 < t.end[a + 1]) next_slot = s2;
      }
      float4* next_rec = kStage ? rec + (cur ^ 1) * (kRecPieces * kSolveBlock) : nullptr;
      auto hook = [&]() {
        if constexpr (kStage) stage_record(blocked_cons(d), next_slot >= 0 ? next_slot : 0, next_rec);
      };
      const float4* mine = kStage && staged ? rec + cur * (kRecPieces * kSolveBlock) : nullptr;
      if (mine != nullptr) {  // body indices of the staged record (its first piece: id, index_a, index_b, count)
        const float4 head = mine[threadIdx.x];
        pf = make_int2(__float_as_int(head.y), __float_as_int(head.z));
      }
      int k = tid;
      if (k < nj) {
        joint(d.joint_ids[jb + k]);
        hook();
      } else if (k < n_items) {
        item(cb + k, pf.x, pf.y, mine, hook);
      } else {
        hook();
      }
      for (k += nthreads; k < n_items; k += nthreads) {
        if (k < nj) {
          joint(d.joint_ids[jb + k]);
        } else {
          const int slot = cb + k;
          const int2 ix = blocked_cons(d).indices(slot);
          item(slot, ix.x, ix.y, nullptr, NoHook{});
        }
      }
      staged = next_slot >= 0;
      cur ^= 1;
      a += 1;
    }
    // the next colour's first body indices: from HBM now, or (staged) from LDS after the barrier
    if (a < n_active && !staged) pf = first_indices(d, t, a, tid, with_joints);
    if (!grid_sync(grid)) return false;
    clock.mark(cls);
  }
  return true;
}

// The overflow colour's levels, run by workgroup 0 (one workgroup barrier per level; the
// workgroup shares one L1, so its stores are visible to its loads after the barrier).
template <class JointItem, class Item>
__device__ __forceinline__ void run_overflow_levels(const SolvePlan& plan, int wg, JointItem joint, Item item) {
  if (wg != 0) return;
  for (int l = 0; l < plan.n_ovf_levels; ++l) {
    const int b = plan.ovf_level_begin[l], e = plan.ovf_level_begin[l + 1];
    for (int i = b + (int)threadIdx.x; i < e; i += (int)blockDim.x) {
      const int32_t it = plan.ovf_items[i];
      if (it >= 0) item(it);
      else joint(-it - 1);
    }
    __syncthreads();
  }
}

// Stage schedule of one step (solver_stages.rs:543-589, solver_integration.rs:807-1015)
// between k_prepare and k_store_finalize: per substep, integrate velocities, warm start,
// push, integrate positions, relax, each stage running the overflow colour's levels first
// and then colours 0..22 in order; then the restitution iterations. A grid barrier
// separates dependent phases.
__global__ __launch_bounds__(kSolveBlock) void k_solve_step(DeviceWorld d, SolvePlan plan, StepParams p) {
  const int wg = (int)blockIdx.x;
  const StepGrid grid{plan.barrier, (uint32_t)plan.n_work, wg};
  const int tid = wg * blockDim.x + threadIdx.x;
  const int nthreads = plan.n_work * blockDim.x;
  const int n_ovf = plan.n_overflow;
  PhaseClock clock{plan.phase_log, 0, wg == 0};
  clock.mark(-1);

  __shared__ ColourTable colours;
  // the staged constraint records of push and relax (two buffers, one per colour in flight)
  __shared__ float4 s_rec[2 * kRecPieces * kSolveBlock];
  float4* rec = plan.stage_records ? s_rec : nullptr;  // set only when the world has contact slots
  if (threadIdx.x == 0) {
    int n = 0;
    for (int ci = 0; ci < kOverflowColor; ++ci)
      if (plan.cr.end[ci] > plan.cr.begin[ci] || plan.jr.end[ci] > plan.jr.begin[ci]) {
        colours.begin[n] = plan.cr.begin[ci];
        colours.end[n] = plan.cr.end[ci];
        colours.jbegin[n] = plan.jr.begin[ci];
        colours.jend[n] = plan.jr.end[ci];
        ++n;
      }
    colours.n = n;
    group_small_colours(colours, plan.small_colour_items);
  }
  __syncthreads();
  const bool any_ovf = plan.n_ovf_levels > 0;
  // With warm start by body, the overflow colour's push and relax levels run on workgroup 0
  // at the end of the phase before them (the body phase of warm start, the position
  // integration), after workgroup 0 has done those phases' per-body work for the overflow
  // colour's bodies itself (the other lanes skip them): the overflow items see the body
  // states the sequential walk gives them, without a grid barrier of their own.
  const bool ovf_merged = any_ovf && plan.ws_table != nullptr;
  // In those two phases workgroup 0 does only the overflow colour's work when other
  // workgroups can take the rest: its chain runs beside their per-body work.
  const bool wg0_ovf_only = ovf_merged && plan.n_work > 1;
  const int body_tid = wg0_ovf_only ? tid - (int)blockDim.x : tid;
  const int body_threads = wg0_ovf_only ? nthreads - (int)blockDim.x : nthreads;
  auto no_joint = [](int32_t) {};

  for (int sub = 0; sub < p.sub_steps; ++sub) {
    // with warm start by body, velocities are integrated in its first phase
    if (plan.ws_table == nullptr) {
      for (int i = tid; i < d.n_awake; i += nthreads) integrate_velocities_item(d, i, p);
      if (!grid_sync(grid)) return;
      clock.mark(kPIntegrateVelocities);
    }

    if (plan.ws_table != nullptr) {
      // Warm starting only adds stored impulses to velocities. In a first phase every
      // coloured item computes the velocity change it gives each of its bodies (k_prepare
      // has done so for the first substep; these read impulses and body rotations, which
      // warm starting does not change), velocities are integrated and the overflow colour's
      // levels run; then each body applies its items' changes in colour order with the
      // operations the colour-by-colour walk applies to it: the same bits in two phases.
      // (Computing the changes at the end of each relax item, or on the lanes a relax phase
      // leaves without an item, measured slower: the relax phases grew more than this phase
      // shrank.)
      // Velocities are integrated here too (it writes only v and w, which this phase does not
      // otherwise read): workgroup 0 integrates a body of the overflow colour just before its
      // first overflow item in this substep, every lane integrates the other bodies.
      if (any_ovf) {
        const int32_t stamp = sub + 1;
        auto integrate_once = [&](int32_t b) {
          if (b >= 0 && plan.ovf_done[b] != stamp) {
            integrate_velocities_item(d, b, p);
            plan.ovf_done[b] = stamp;
          }
        };
        run_overflow_levels(
            plan, wg,
            [&](int32_t jid) {
              integrate_once(d.joints[jid].index_a);
              integrate_once(d.joints[jid].index_b);
              warm_start_joint_item(d, jid);
            },
            [&](int k) {
              integrate_once(d.overflow[k].index_a);
              integrate_once(d.overflow[k].index_b);
              overflow_warm_item(d, k);
            });
      }
      for (int i = tid; i < d.n_awake; i += nthreads)
        if (!any_ovf || plan.ovf_done[i] < 0) integrate_velocities_item(d, i, p);
      WarmDelta* jd = plan.ws_delta + 2 * plan.n_constraints;
      if (sub > 0) {
        const int n_cj = plan.jr.begin[kOverflowColor];
        for (int k = tid; k < n_cj; k += nthreads) {
          const int32_t jid = d.joint_ids[k];
          warm_delta_joint(d, d.joints[jid], jd + 2 * jid);
        }
        // four constraints per lane at a time, their records loaded before any is computed
        // (the phase is a chain of memory round trips otherwise)
        constexpr int kU = 4;
        int slot = tid;
        for (; slot + (kU - 1) * nthreads < plan.n_constraints; slot += kU * nthreads) {
          Constraint c[kU];
#pragma unroll
          for (int u = 0; u < kU; ++u) c[u] = blocked_cons(d).load(slot + u * nthreads);
#pragma unroll
          for (int u = 0; u < kU; ++u) warm_delta_contact(c[u], plan.ws_delta + 2 * (slot + u * nthreads));
        }
        for (; slot < plan.n_constraints; slot += nthreads) warm_delta_contact(blocked_cons(d).load(slot), plan.ws_delta + 2 * slot);
      }
      if (!grid_sync(grid)) return;
      clock.mark(kPWarmStart);
      auto ws_body = [&](int i) {
        int32_t e[kOverflowColor];
#pragma unroll
        for (int c = 0; c < kOverflowColor; ++c) e[c] = plan.ws_table[c * d.n_awake + i];
        BodyState s = d.states[i];
        // every colour's delta loaded at once (the zero record for a colour without an item)
        WarmDelta wd[kOverflowColor];
#pragma unroll
        for (int c = 0; c < kOverflowColor; ++c) {
          const int32_t k = e[c] >= 0 ? e[c] : -e[c] - 2;  // (item << 1) | side
          const WarmDelta* src = e[c] == kNullIndex ? plan.ws_zero : (e[c] >= 0 ? plan.ws_delta + k : jd + k);
          wd[c] = *src;
        }
#pragma unroll
        for (int c = 0; c < kOverflowColor; ++c) {
          if (e[c] == kNullIndex) continue;
          const int32_t k = e[c] >= 0 ? e[c] : -e[c] - 2;
          const WarmDelta w = wd[c];
          if (k & 1) {
            s.w = add(s.w, w.dw);
            s.v = add(s.v, w.dv);
          } else {
            s.w = nsl::sub(s.w, w.dw);  // `sub` is the substep counter here
            s.v = nsl::sub(s.v, w.dv);
          }
        }
        d.states[i] = s;
      };
      if (body_tid >= 0)
        for (int i = body_tid; i < d.n_awake; i += body_threads)
          if (!ovf_merged || plan.ovf_done[i] < 0) ws_body(i);
      if (ovf_merged && wg == 0) {
        __syncthreads();
        for (int k = (int)threadIdx.x; k < plan.n_ovf_bodies; k += (int)blockDim.x) ws_body(plan.ovf_bodies[k]);
        __syncthreads();
        run_overflow_levels(plan, wg, [&](int32_t jid) { solve_joint_item(d, jid, p, true); },
                            [&](int k) { overflow_solve_item(d, k, p, 0); });
      }
      if (!grid_sync(grid)) return;
      clock.mark(kPWarmStart);
    } else {
      if (any_ovf) {
        run_overflow_levels(plan, wg, [&](int32_t jid) { warm_start_joint_item(d, jid); }, [&](int k) { overflow_warm_item(d, k); });
        if (!grid_sync(grid)) return;
        clock.mark(kPOverflow);
      }
      if (!run_colours<false>(d, colours, tid, nthreads, grid, clock, kPWarmStart, true,
                       [&](int32_t jid) { warm_start_joint_item(d, jid); },
                       [&](int slot, int32_t ia, int32_t ib, const float4*, auto) { warm_start_item(d, slot, ia, ib); }))
        return;
    }

    if (any_ovf && !ovf_merged) {
      run_overflow_levels(plan, wg, [&](int32_t jid) { solve_joint_item(d, jid, p, true); },
                          [&](int k) { overflow_solve_item(d, k, p, 0); });
      if (!grid_sync(grid)) return;
      clock.mark(kPOverflow);
    }
    {
      auto joint = [&](int32_t jid) { solve_joint_item(d, jid, p, true); };
      auto contact = [&](int slot, int32_t ia, int32_t ib, const float4* staged, auto hook) {
        push_item(d, slot, p, ia, ib, staged, kSolveBlock, hook);
      };
      if (!(rec != nullptr ? run_colours<true>(d, colours, tid, nthreads, grid, clock, kPPush, true, joint, contact, rec)
                           : run_colours<false>(d, colours, tid, nthreads, grid, clock, kPPush, true, joint, contact)))
        return;
    }

    if (body_tid >= 0)
      for (int i = body_tid; i < d.n_awake; i += body_threads)
        if (!ovf_merged || plan.ovf_done[i] < 0) integrate_positions_item(d, i, p);
    if (ovf_merged && wg == 0) {
      __syncthreads();
      for (int k = (int)threadIdx.x; k < plan.n_ovf_bodies; k += (int)blockDim.x)
        integrate_positions_item(d, plan.ovf_bodies[k], p);
      __syncthreads();
      run_overflow_levels(plan, wg, [&](int32_t jid) { solve_joint_item(d, jid, p, false); },
                          [&](int k) { overflow_solve_item(d, k, p, 1); });
    }
    if (!grid_sync(grid)) return;
    clock.mark(kPIntegratePositions);

    if (any_ovf && !ovf_merged) {
      run_overflow_levels(plan, wg, [&](int32_t jid) { solve_joint_item(d, jid, p, false); },
                          [&](int k) { overflow_solve_item(d, k, p, 1); });
      if (!grid_sync(grid)) return;
      clock.mark(kPOverflow);
    }
    {
      auto joint = [&](int32_t jid) { solve_joint_item(d, jid, p, false); };
      auto contact = [&](int slot, int32_t ia, int32_t ib, const float4* staged, auto hook) {
        relax_item(d, slot, p, ia, ib, staged, kSolveBlock, hook);
      };
      if (!(rec != nullptr ? run_colours<true>(d, colours, tid, nthreads, grid, clock, kPRelax, true, joint, contact, rec)
                           : run_colours<false>(d, colours, tid, nthreads, grid, clock, kPRelax, true, joint, contact)))
        return;
    }
  }

  if (plan.any_restitution) {
    for (int it = 0; it < plan.restitution_iterations; ++it) {
      if (n_ovf > 0) {
        run_overflow_levels(plan, wg, no_joint, [&](int k) { overflow_solve_item(d, k, p, 2); });
        if (!grid_sync(grid)) return;
        clock.mark(kPOverflow);
      }
      if (!run_colours<false>(d, colours, tid, nthreads, grid, clock, kPRestitution, false, no_joint,
                       [&](int slot, int32_t ia, int32_t ib, const float4*, auto) { restitution_item(d, blocked_cons(d), slot, p, ia, ib); }))
        return;
    }
  }

}

// Prepare (contact_solver.rs:395-662, overflow_contact_solver.rs, revolute_joint.rs:176-261):
// every joint, contact constraint and overflow constraint is independent, so this runs at the
// width of the work, before the step kernel. It also restarts this step's fast-body list.
__global__ __launch_bounds__(kBlock) void k_prepare(DeviceWorld d, SolvePlan plan, StepParams p) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nthreads = gridDim.x * blockDim.x;
  int points = 0, ovf_points = 0;
  if (tid == 0) {
    d.counters[kCntFastList] = 0;  // finalize lists this step's fast bodies
    // the step kernel's barrier counts its arrivals from zero (stream order: the previous
    // step kernel has finished, the next starts after this kernel)
    __hip_atomic_store(&plan.barrier->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (plan.ws_zero != nullptr) *plan.ws_zero = WarmDelta{zero3(), zero3()};
  }
  for (int k = tid; k < plan.n_joints; k += nthreads) {
    const int32_t jid = d.joint_ids[k];
    prepare_joint_item(d, jid, p);
    if (plan.ws_table != nullptr) {
      int c = 0;
      while (c < kOverflowColor && k >= plan.jr.end[c]) ++c;  // the overflow colour's joints are last
      if (c == kOverflowColor && plan.ovf_done != nullptr) {
        const JointSim& j = d.joints[jid];
        if (j.index_a >= 0) plan.ovf_done[j.index_a] = 0;
        if (j.index_b >= 0) plan.ovf_done[j.index_b] = 0;
      }
      if (c < kOverflowColor) {
        const JointSim& j = d.joints[jid];
        if (j.index_a >= 0) plan.ws_table[c * d.n_awake + j.index_a] = -((jid << 1) + 2);
        if (j.index_b >= 0) plan.ws_table[c * d.n_awake + j.index_b] = -((jid << 1 | 1) + 2);
        warm_delta_joint(d, j, plan.ws_delta + 2 * (plan.n_constraints + jid));  // the first substep's
      }
    }
  }
  for (int slot = tid; slot < plan.n_constraints; slot += nthreads) {
    points += prepare_item(d, slot, p, [&](const Constraint& cc) {
      if (plan.ws_table != nullptr) {
        int c = 0;
        while (c < kOverflowColor - 1 && slot >= plan.cr.end[c]) ++c;  // slots are colour-major
        if (cc.index_a >= 0) plan.ws_table[c * d.n_awake + cc.index_a] = slot << 1;
        if (cc.index_b >= 0) plan.ws_table[c * d.n_awake + cc.index_b] = slot << 1 | 1;
        warm_delta_contact(cc, plan.ws_delta + 2 * slot);  // the first substep's
      }
    });
  }
  for (int k = tid; k < plan.n_overflow; k += nthreads) {
    ovf_points += prepare_overflow_item(d, k, p);
    if (plan.ovf_done != nullptr) {  // bodies of the overflow colour: integrated by workgroup 0
      if (d.overflow[k].index_a >= 0) plan.ovf_done[d.overflow[k].index_a] = 0;
      if (d.overflow[k].index_b >= 0) plan.ovf_done[d.overflow[k].index_b] = 0;
    }
  }
  block_add(&d.counters[kCntPoints], points);
  block_add(&d.counters[kCntOverflowPoints], ovf_points);
}

// Store impulses to the contacts (contact_solver.rs:1258-1299) and finalize the bodies
// (solver_integration.rs:587-799): disjoint data, after the step kernel, at full width.
__global__ __launch_bounds__(kBlock) void k_store_finalize(DeviceWorld d, SolvePlan plan, StepParams p) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nthreads = gridDim.x * blockDim.x;
  for (int k = tid; k < plan.n_overflow; k += nthreads) overflow_store_item(d, k);
  for (int slot = tid; slot < plan.n_constraints; slot += nthreads) store_item(d, blocked_cons(d), slot);
  for (int i = tid; i < d.n_awake; i += nthreads) finalize_item(d, i, p);
}

// The residual sweep of one colour: one lane per coloured slot of [begin, begin + count),
// `d.states` a copy of the body states. Each wavefront reduces its lanes' residuals with
// lane shuffles (sum by the butterfly 32, 16, 8, 4, 2, 1; maxima) and lane 0 writes the
// wavefront's partial (ceil(count / 64) of them); lanes past the colour contribute zeros.
__global__ __launch_bounds__(kBlock) void k_colour_residual(DeviceWorld d, int begin, int count, StepParams p,
                                                            ResidualPartial* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  Residual r;
  if (i < count) r = residual_item(d, begin + i, p);
  float sum = r.sum_normal, mx = r.max_normal, fr = r.friction;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    sum += __shfl_xor(sum, off, 64);
    mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    fr = fmaxf(fr, __shfl_xor(fr, off, 64));
  }
  // lane 0's index is its wavefront's first: only wavefronts that hold a slot write
  if ((threadIdx.x & 63) == 0 && i < count) out[i >> 6] = ResidualPartial{sum, mx, fr, 0.0f};
}

inline int blocks_for(int n) { return (n + kBlock - 1) / kBlock; }

// Launch wrapped in a timer mark when timing is on.
#define NSL_TIMED(timer, cls, stream, launch) \
  do {                                        \
    if (timer) (timer)->begin(cls, stream);   \
    launch;                                   \
    if (timer) (timer)->end(stream);          \
  } while (0)

}  // namespace

void launch_init_contacts(const DeviceWorld& d, const int32_t* ids, const Contact* records, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_init_contacts, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, ids, records, n);
}

void launch_patch_locations(const DeviceWorld& d, const int32_t* items, int n_shapes, int n_contacts, int n_joints,
                            const BodySim* sims, int n_sims, hipStream_t s) {
  const int n = n_shapes + n_contacts + n_joints + n_sims;
  if (n > 0)
    hipLaunchKernelGGL(k_patch_locations, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, items, n_shapes, n_contacts, n_joints,
                       sims, n_sims);
}

void launch_pairset_insert(PairSet set, const uint64_t* keys, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_pairset_insert, dim3(blocks_for(n)), dim3(kBlock), 0, s, set, keys, n);
}

void launch_pairset_erase(PairSet set, const uint64_t* keys, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_pairset_erase, dim3(blocks_for(n)), dim3(kBlock), 0, s, set, keys, n);
}

void launch_collide(const DeviceWorld& d, int n, hipStream_t s, KernelTimer* t) {
  // measured: two waves per SIMD took the 132k-body world's collide from 0.69 to 0.59 ms per
  // step, one wave per SIMD was faster on the 10k pile's 86k contacts
  constexpr int kTwoWavesFrom = 1 << 18;
  if (n >= kTwoWavesFrom)
    NSL_TIMED(t, kKCollide, s, hipLaunchKernelGGL(k_collide_wide, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, n));
  else if (n > 0)
    NSL_TIMED(t, kKCollide, s, hipLaunchKernelGGL(k_collide, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, n));
}

int solve_step_max_blocks(int device) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_solve_step), kSolveBlock, 0) !=
      hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * cus;
}

void launch_prepare(const DeviceWorld& d, const SolvePlan& plan, const StepParams& p, hipStream_t s, KernelTimer* t) {
  const int n = std::max(std::max(plan.n_joints, plan.n_constraints), std::max(plan.n_overflow, 1));
  NSL_TIMED(t, kKPrepare, s, hipLaunchKernelGGL(k_prepare, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, plan, p));
}

void launch_store_finalize(const DeviceWorld& d, const SolvePlan& plan, const StepParams& p, hipStream_t s, KernelTimer* t) {
  const int n = std::max(std::max(plan.n_constraints, d.n_awake), std::max(plan.n_overflow, 1));
  NSL_TIMED(t, kKFinalize, s, hipLaunchKernelGGL(k_store_finalize, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, plan, p));
}

void launch_colour_residual(const DeviceWorld& d, int begin, int count, const StepParams& p, ResidualPartial* out,
                            hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_colour_residual, dim3(blocks_for(count)), dim3(kBlock), 0, s, d, begin, count, p, out);
}

void launch_solve_step(const DeviceWorld& d, const SolvePlan& plan, const StepParams& p, bool cooperative, hipStream_t s,
                       KernelTimer* t) {
  const int blocks = plan.n_work;
  // A cooperative launch: the runtime starts the grid only when every workgroup can be
  // resident at once, which the grid barriers need (`blocks` is clamped by the caller to the
  // occupancy limit); the barriers' bounded wait still turns a stalled grid into an error.
  // The plain launch of the same grid relies on that clamp alone.
  void* args[] = {const_cast<DeviceWorld*>(&d), const_cast<SolvePlan*>(&plan), const_cast<StepParams*>(&p)};
  hipError_t launch = hipSuccess;
  if (cooperative) {
    NSL_TIMED(t, kKSolve, s,
              launch = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_solve_step), dim3(blocks),
                                                  dim3(kSolveBlock), args, 0, s));
  } else {
    NSL_TIMED(t, kKSolve, s, hipLaunchKernelGGL(k_solve_step, dim3(blocks), dim3(kSolveBlock), 0, s, d, plan, p));
  }
  if (launch != hipSuccess) throw std::runtime_error(std::string("step kernel launch: ") + hipGetErrorString(launch));
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) throw std::runtime_error(std::string("step kernel launch: ") + hipGetErrorString(err));
}

}  // namespace nsl
