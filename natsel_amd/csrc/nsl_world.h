// Host side of the GPU stepper: the world/body/shape/step API of the reference's
// World (src/simulation/world_runtime.rs:942), its integer bookkeeping (broad-phase
// pair list, contact lifecycle, graph colouring) and the launch schedule of one step.
// Float work of the step runs on the device (nsl_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nsl_broadphase.h"
#include "nsl_ccd.h"
#include "nsl_device.h"
#include "nsl_scene.h"

namespace nsl {

// Wall-clock phases of the last step, in milliseconds (types/world_def.rs:158-184).
struct StepProfile {
  float pairs, collide, state_pass, solve, finalize, step;
  int32_t contacts, touching, colors_used, overflow;
  int32_t fast_bodies, sat_calls, sat_hits, recycled;
};

class GpuWorld : public HostScene {
 public:
  explicit GpuWorld(int device);
  ~GpuWorld();
  GpuWorld(const GpuWorld&) = delete;
  GpuWorld& operator=(const GpuWorld&) = delete;

  WorldOptions options;

  int32_t create_body(const BodyDesc& desc) override;

  void step(float dt, int32_t sub_steps);

  // A revolute joint between two bodies (joint.rs:164-375); returns its id. Without
  // collide_connected the bodies' contacts are destroyed and no new one forms.
  int32_t create_revolute_joint(const JointDesc& desc);
  int32_t joint_count() const { return (int32_t)hjoints_.size(); }
  // Accumulated impulses (linear xyz, perpendicular xy, spring, motor, lower, upper) and
  // graph colour of a joint (-1: not in the graph, both bodies non-dynamic).
  void get_joint(int32_t id, float* impulses, int32_t* color);

  int32_t body_capacity() const { return (int32_t)bodies_.size(); }
  int32_t contact_count() const { return contact_pool_next_ - (int32_t)contact_free_.size(); }
  // Per body id: origin position, rotation (x, y, z, s), linear and angular velocity.
  void get_bodies(double* pos, float* rot, float* v, float* w);
  uint64_t state_hash();
  // Per contact id: colour (-1 when not in the graph), local index and liveness.
  int32_t contact_capacity() const { return (int32_t)hcontacts_.size(); }
  void get_contact_colors(int32_t* color, int32_t* local, int32_t* alive) const;
  // Per contact id: its shapes, -1 when the id is free.
  void get_contact_shapes(int32_t* shape_a, int32_t* shape_b) const;
  // A body's solver set (0 static, 2 awake, 3 and up sleeping), index in it and island id.
  void body_location(int32_t id, int32_t* set, int32_t* local, int32_t* island) const;
  // Island slots, live islands, awake islands, sleeping sets, the island to split next step.
  void island_summary(int64_t out[5]) const;
  // 0 when the island, solver-set and encoding records agree with each other; else the
  // number of the first failed check (the same checks as the oracle's).
  int32_t check_islands() const;
  const StepProfile& profile() const { return profile_; }

  // Device time per launch class (HIP events on the world's stream) and per phase of the
  // step kernel (device wall clock at its grid barriers), only while timing is on; device
  // work counters always; since the last reset.
  struct KernelStats {
    double ms[kKernelClassCount];
    int64_t launches[kKernelClassCount];
    double phase_ms[kPhaseClassCount];
    int64_t phase_count[kPhaseClassCount];
    int64_t steps, bodies, collided, constraints, points, overflow_points, recycled, sat_calls, sat_hits, fast_bodies;
    int64_t toi_calls, toi_hits;  // continuous collision: time-of-impact queries, bodies moved back
    int64_t barriers, solver_blocks, pair_tests;
    double host_ms[4];  // host wall time: pair update, collide (with its wait), state pass, solve+finalize (with its wait)
    int64_t joints;
  };
  // Launch events when `on`; the step kernel's phase clock (a device wall-clock read at
  // each grid barrier and a download of the marks per step) also when `phases`.
  void set_kernel_timing(bool on, bool phases = true);
  const KernelStats& kernel_stats() const { return kstats_; }
  void reset_kernel_stats();

  // Per-colour residuals of the coloured contact solver: while on, each step ends with one
  // more relax sweep over the coloured contact constraints, colour by colour, from the state
  // the solver ends in, on a copy of the body states (the world's state is unchanged). Per
  // colour: the sum over points of |normal impulse change| (wavefront partials added in slot
  // order in double), the largest such change, the largest |friction impulse change| and
  // the number of constraints. No reference counterpart: a convergence diagnostic.
  struct ColourResidual {
    double sum_normal = 0.0;
    float max_normal = 0.0f, max_friction = 0.0f;
    int32_t constraints = 0, reserved = 0;
  };
  void set_residuals(bool on) { residuals_on_ = on; }
  const ColourResidual* residuals() const { return residuals_; }  // [kOverflowColor], of the last step

  // Everything that decides later steps (world_state.rs:340-438 save_state): body states
  // and sims, fat AABBs and moved flags, contacts with their manifolds and caches, the
  // pair-key set, joints, the id pool, the non-touching list, the colour lists and body
  // bits, the work counters. Device arrays are copied device to device.
  struct Snapshot;
  Snapshot* save_state();  // owned by the caller; free with destroy_snapshot
  struct Tuning {
    int blocks = 0, small_colour_items = 0, stage_records = 0, candidates = 0;
    double best_us = 0.0, previous_us = 0.0;
  };
  // Times candidate launch shapes of the step kernel from a snapshot and keeps the fastest
  // in `options`; the world's state is unchanged.
  Tuning autotune(float dt, int32_t sub_steps, int32_t steps);
  static void destroy_snapshot(Snapshot* s);
  // world_state.rs:439-560 restore_state. The world must hold the bodies, shapes and joints
  // it held when the snapshot was taken.
  void restore_state(const Snapshot& s);

 private:
  // set: 1 disabled, 2 awake, 3 and up sleeping; local: index in the set's list or colour
  struct HContact {
    int32_t shape_a = -1, shape_b = -1, color = -1, local = -1, set = -1;
    int32_t alive = 0;
    float restitution = 0;
    uint64_t created = 0;  // creation order (a body's contact list is newest first)
    // edges of the bodies' contact lists (contact_creation.rs:357-384): body, prev and next
    // key (contact_id << 1 | edge)
    int32_t body[2] = {-1, -1}, prev[2] = {-1, -1}, next[2] = {-1, -1};
    int32_t island = -1, island_index = -1;
  };
  // set: 0 between non-dynamic bodies (never solved), 2 in the graph, 3 and up sleeping
  struct HJoint {
    int32_t body_a = -1, body_b = -1, color = -1, local = -1, set = 0;
    int32_t island = -1, island_index = -1;
    int32_t enc_a = -1, enc_b = -1;  // encodings held by the device record
  };
  // islands (island.rs) and sleeping solver sets (solver_set.rs)
  struct ContactLink { int32_t contact, body_a, body_b; };
  struct JointLink { int32_t joint, body_a, body_b; };
  struct Island {
    int32_t id = -1, set = -1, local = -1, remove_count = 0;
    std::vector<int32_t> bodies;
    std::vector<ContactLink> contacts;
    std::vector<JointLink> joints;
  };
  struct SleepSet {
    std::vector<int32_t> bodies, contacts, joints, islands;  // ids, in set order
  };
  struct IdPool {  // id_pool.rs: LIFO free list
    std::vector<int32_t> free;
    int32_t next = 0;
    int32_t alloc() {
      if (free.empty()) return next++;
      const int32_t id = free.back();
      free.pop_back();
      return id;
    }
    void release(int32_t id) { free.push_back(id); }  // no live id equals next
  };
  struct ColorSet {
    std::vector<int32_t> convex;
    std::vector<int32_t> overflow;
    std::vector<int32_t> joints;  // joint ids in graph order (appended, swap-removed)
  };
  // The colours' body bits (constraint_graph.rs: each colour's body set), kept per body: bit
  // c of body_colours_[b] is set when colour c holds body b, so a colour search is one mask.
  std::vector<uint32_t> body_colours_;
  uint32_t colours_of(int32_t b) const { return (size_t)b < body_colours_.size() ? body_colours_[(size_t)b] : 0u; }
  void colour_set(int c, int32_t b) {
    if ((size_t)b >= body_colours_.size()) body_colours_.resize(std::max((size_t)b + 1, bodies_.size()), 0u);
    body_colours_[(size_t)b] |= 1u << c;
  }
  void colour_clear(int c, int32_t b) {
    if ((size_t)b < body_colours_.size()) body_colours_[(size_t)b] &= ~(1u << c);
  }
  // The colour the reference's search picks for two dynamic bodies (the first colour below
  // kDynamicColorCount holding neither) or for one dynamic body beside a static one (the
  // last colour from kOverflowColor - 1 down to 1 not holding it); kOverflowColor if none.
  int32_t free_colour_dynamic(int32_t a, int32_t b) const {
    const uint32_t avail = ~(colours_of(a) | colours_of(b)) & ((1u << kDynamicColorCount) - 1u);
    return avail != 0u ? __builtin_ctz(avail) : kOverflowColor;
  }
  int32_t free_colour_static(int32_t dyn) const {
    const uint32_t avail = ~colours_of(dyn) & (((1u << kOverflowColor) - 1u) & ~1u);
    return avail != 0u ? 31 - __builtin_clz(avail) : kOverflowColor;
  }

  int32_t add_shape(int32_t body, Shape shape, const Material& m, float radius) override;

  void upload_bodies();
  void pull_bodies();
  template <class T> void ensure(T*& ptr, size_t& cap, size_t n, bool keep);

  void update_pairs();
  void ensure_pair_set();
  void create_contact(int32_t a, int32_t b);
  void destroy_contact(int32_t id, bool device_off = false);
  void mark_contact_dirty(int32_t id);
  void add_to_graph(int32_t id);
  void remove_from_graph(int32_t body_a, int32_t body_b, int32_t color, int32_t local);
  void state_pass();
  int32_t assign_joint_color(int32_t a, int32_t b);

  // islands and sleep (nsl_islands.hip)
  std::vector<int32_t>& island_list(int32_t set);
  std::vector<int32_t>& contact_list(int32_t set);
  int32_t create_island(int32_t set);
  void destroy_island(int32_t id);
  int32_t merge_islands(int32_t a, int32_t b);
  void link_contact(int32_t id);
  void unlink_contact(int32_t id);
  void link_joint(int32_t id);
  void link_body_contact(int32_t id);
  void unlink_body_contact(int32_t id);
  void split_island(int32_t id);
  void mark_awake_islands();
  void process_sleep();
  void try_sleep_island(int32_t id);
  void wake_solver_set(int32_t set);
  void add_joint_to_graph(int32_t id);
  void remove_joint_from_graph(int32_t id);
  void mark_body_moved(int32_t body);
  // Brings the device up to the host's body locations: awake sims and states, sleeping sims,
  // the encodings held by shapes, awake-set contacts and graph joints.
  void flush_locations();
  void upload_joint_pairs();

  int device_;
  hipStream_t stream_ = nullptr;
  KernelTimer timer_{};
  bool timing_ = false, phase_timing_ = false;
  struct StepReadback {  // device layout: the work counters, then the grid barrier
    int32_t counters[kCounterCount];
    GridBarrier barrier;
  };
  StepReadback* readback_ = nullptr;  // pinned
  GridBarrier* d_barrier_ = nullptr;
  uint64_t* d_phase_log_ = nullptr;
  std::vector<uint64_t> phase_log_host_;
  GridBarrier barrier_host_{};
  int max_blocks_ = 0;
  // set when a grid barrier wait timed out mid-step; cleared by restore_state
  bool poisoned_ = false;
  void check_usable() const;
  double wall_clock_khz_ = 0.0;
  KernelStats kstats_{};
  int32_t counters_host_[kCounterCount] = {};
  int32_t counters_base_[kCounterCount] = {};

  std::vector<HContact> hcontacts_;
  std::vector<int32_t> contact_free_;
  int32_t contact_pool_next_ = 0;
  std::vector<int32_t> non_touching_;
  ColorSet colors_[kColorCount];
  std::vector<int32_t> init_ids_;
  std::vector<Contact> init_records_;
  bool bodies_dirty_ = true;   // host copies are newer than the device arrays
  bool device_newer_ = false;  // device arrays are newer than the host copies
  bool shapes_dirty_ = true;

  DeviceWorld d_{};
  size_t cap_sims_ = 0, cap_states_ = 0, cap_static_ = 0, cap_shapes_ = 0, cap_fat_ = 0, cap_moved_ = 0, cap_hulls_ = 0;
  size_t cap_contacts_ = 0, cap_change_ = 0, cap_constraints_ = 0, cap_overflow_ = 0, cap_collide_ = 0, cap_solve_ = 0;
  size_t cap_overflow_ids_ = 0, cap_init_ids_ = 0, cap_init_rec_ = 0;
  int32_t* d_init_ids_ = nullptr;
  Contact* d_init_records_ = nullptr;
  // A growable int list in pinned host memory: the staging of a per-step upload.
  struct PinnedInts {
    int32_t* ptr = nullptr;
    size_t n = 0, cap = 0;
    ~PinnedInts() { if (ptr) (void)hipHostFree(ptr); }
    void clear() { n = 0; }
    bool empty() const { return n == 0; }
    size_t size() const { return n; }
    const int32_t* data() const { return ptr; }
    void append(const std::vector<int32_t>& v) {
      if (n + v.size() > cap) grow(n + v.size());
      if (!v.empty()) std::memcpy(ptr + n, v.data(), v.size() * sizeof(int32_t));
      n += v.size();
    }
    void grow(size_t want);  // keeps the contents; the caller's stream has no copy from it in flight
  };
  PinnedInts solve_stage_;
  int32_t graph_restitution_ = 0;  // graph contacts with restitution > 0
  // contacts whose state changed in the last collide: id << 2 | ContactChange, ascending
  std::vector<int32_t> changes_;
  std::vector<uint64_t> change_bits_;  // sort_changes: contact ids present
  std::vector<uint8_t> change_code_;   // sort_changes: the change of each present id
  void sort_changes(const int32_t* list, size_t m, size_t ncap);
  int32_t* changes_pinned_ = nullptr;  // download staging: [0] count, then the list
  size_t cap_changes_pinned_ = 0, cap_change_list_ = 0, last_changes_ = 0;
  void ensure_pinned_changes(size_t n);
  // the residual sweep after the step kernel, and its per-colour results after the step's sync
  void launch_residual_sweep(const DeviceWorld& dw, const ColorRanges& cr, const StepParams& p);
  void collect_residuals(const ColorRanges& cr);
  std::vector<uint8_t> moved_host_;  // staging for uploads and pulls
  std::vector<Aabb> fat_host_;
  std::vector<Aabb> aabb_host_;
  size_t cap_aabbs_ = 0;
  // continuous collision: candidate shape ids (static, kinematic, dynamic; each ascending)
  // and the fast-body list the step kernel fills
  int32_t* d_candidates_ = nullptr;
  size_t cap_candidates_ = 0, cap_fast_ = 0;
  ContinuousArgs ccd_{};
  int32_t bullet_bodies_ = 0;  // dynamic bullet bodies (the bullet pass runs only if any)
  int32_t kinematic_awake_ = 0;  // awake bodies that are not dynamic (warm start then runs by colour)
  int32_t* d_ws_table_ = nullptr;  // warm start by body: the item of each colour per awake body
  WarmDelta* d_ws_delta_ = nullptr;  // and each item's velocity change per side
  int32_t* d_ovf_done_ = nullptr;  // overflow bodies' velocity integration marks
  int32_t* d_ovf_bodies_ = nullptr;  // the overflow colour's dynamic bodies, each once
  size_t cap_ovf_bodies_ = 0;
  std::vector<int32_t> ovf_bodies_host_;
  // residual sweep: the body-state copy, the wavefront partials (device and pinned) and the
  // last step's per-colour results
  bool residuals_on_ = false;
  BodyState* d_res_states_ = nullptr;
  ResidualPartial* d_res_partials_ = nullptr;
  ResidualPartial* res_partials_host_ = nullptr;
  size_t cap_res_states_ = 0, cap_res_partials_ = 0, cap_res_partials_host_ = 0;
  int32_t res_wave_begin_[kOverflowColor + 1] = {};
  ColourResidual residuals_[kOverflowColor] = {};
  size_t cap_ws_table_ = 0, cap_ws_delta_ = 0, cap_ovf_done_ = 0;
  // device pair finder and the pair-key set it checks (new contacts insert, disjoint ones erase)
  std::unique_ptr<PairFinder> finder_;
  Proxy* d_proxies_ = nullptr;
  size_t cap_proxies_ = 0;
  std::vector<Proxy> proxies_host_;
  std::vector<uint64_t> new_keys_;
  uint64_t* d_rebuild_keys_ = nullptr;
  size_t cap_rebuild_keys_ = 0;
  size_t n_device_shapes_ = 0;  // shapes whose fat AABB and moved flag live on the device
  int64_t pair_set_tombs_ = 0;
  bool pairs_pending_ = true;   // proxies marked moved on the host since the last pair update
  StepProfile profile_{};

  std::vector<HJoint> hjoints_;
  std::vector<Island> islands_;
  IdPool island_pool_;
  std::vector<int32_t> awake_islands_;
  std::vector<SleepSet> sleep_sets_;  // by set id (3 and up)
  IdPool set_pool_{{}, 3};
  std::vector<int32_t> disabled_;     // contacts of the disabled set
  int32_t split_island_ = -1;         // split at the start of the next solve
  std::vector<uint8_t> island_awake_bits_;
  int32_t split_candidate_ = -1;
  float split_candidate_time_ = 0.0f;
  std::vector<float> sleep_times_host_;  // per awake index, from the last finalize
  size_t cap_sleep_times_ = 0, cap_sleep_sims_ = 0;
  // bodies whose location changed since the last flush, and what follows from it
  std::vector<uint8_t> body_moved_;
  std::vector<int32_t> moved_bodies_;
  std::vector<int32_t> slept_bodies_;  // sims to copy to the device's sleeping sims
  std::vector<int32_t> dirty_contacts_;  // contacts whose device flag or encodings need a patch
  size_t cap_active_ = 0;
  bool locations_dirty_ = false;
  std::vector<uint8_t> contact_mark_;
  int32_t* d_patch_ = nullptr;
  BodySim* d_patch_sims_ = nullptr;
  size_t cap_patch_ = 0, cap_patch_sims_ = 0;
  uint64_t contact_serial_ = 0;
  size_t cap_joints_ = 0, cap_joint_ids_ = 0, cap_erase_keys_ = 0;
  uint64_t* d_erase_keys_ = nullptr;
  std::vector<uint64_t> blocked_keys_;  // body pairs of joints without collide_connected
  std::vector<int32_t> joint_ids_host_;
  ColorRanges joint_ranges_{};
  bool joints_dirty_ = false;
  // overflow dependency levels (rebuilt each step)
  std::vector<int32_t> ovf_items_host_, ovf_level_host_, body_level_;
  int32_t* d_ovf_items_ = nullptr;
  int32_t* d_ovf_levels_ = nullptr;
  size_t cap_ovf_items_ = 0, cap_ovf_levels_ = 0;
  int32_t build_overflow_levels();
};

}  // namespace nsl
