// Host side of the GPU stepper (see nsl_world.h).
#include "nsl_world.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace nsl {

#define NSL_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

GpuWorld::GpuWorld(int device) : device_(device) {
  NSL_CHECK(hipSetDevice(device_));
  NSL_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  timer_.capacity = 64;
  timer_.events = new hipEvent_t[2 * timer_.capacity];
  for (int i = 0; i < 2 * timer_.capacity; ++i) NSL_CHECK(hipEventCreate(&timer_.events[i]));
  // the work counters and the grid barrier in one block, read back with one copy per step
  static_assert(sizeof(StepReadback) == kCounterCount * sizeof(int32_t) + sizeof(GridBarrier), "readback layout");
  NSL_CHECK(hipMalloc(&d_.counters, sizeof(StepReadback)));
  NSL_CHECK(hipMemset(d_.counters, 0, sizeof(StepReadback)));
  d_barrier_ = reinterpret_cast<GridBarrier*>(d_.counters + kCounterCount);
  NSL_CHECK(hipHostMalloc((void**)&readback_, sizeof(StepReadback), hipHostMallocDefault));
  NSL_CHECK(hipMalloc(&d_phase_log_, 2 * kMaxPhaseLog * sizeof(uint64_t)));
  // zeroed: the host reads marks up to the first time stamp that goes backwards
  NSL_CHECK(hipMemset(d_phase_log_, 0, 2 * kMaxPhaseLog * sizeof(uint64_t)));
  phase_log_host_.resize(2 * kMaxPhaseLog);
  max_blocks_ = solve_step_max_blocks(device_);
  if (max_blocks_ <= 0) throw std::runtime_error("step kernel does not fit on the device");
  int khz = 0;
  NSL_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
  wall_clock_khz_ = (double)khz;
  finder_ = std::make_unique<PairFinder>(stream_);
  ensure_pair_set();
}

GpuWorld::~GpuWorld() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  void* ptrs[] = {d_.sims, d_.states, d_.static_sims, d_.shapes, d_.fat_aabbs, d_.shape_moved, d_.hulls, d_.contacts,
                  d_.contact_change, d_.constraints, d_.overflow, d_.collide_ids, d_.solve_ids, d_.overflow_ids,
                  d_.counters, d_init_ids_, d_init_records_, d_phase_log_, d_proxies_,
                  d_rebuild_keys_, d_.pairs.slots, d_.joints, d_.joint_ids, d_.joint_pairs.slots, d_erase_keys_,
                  d_ovf_items_, d_ovf_levels_, d_.shape_aabbs, d_.fast_ids, d_candidates_, d_.sleep_times,
                  d_.sleep_sims, d_patch_, d_patch_sims_, d_.change_list, d_.contact_active, d_ws_table_, d_ws_delta_, d_ovf_done_,
                  d_res_states_, d_res_partials_, d_ovf_bodies_};
  finder_.reset();
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (changes_pinned_) (void)hipHostFree(changes_pinned_);
  if (readback_) (void)hipHostFree(readback_);
  if (res_partials_host_) (void)hipHostFree(res_partials_host_);
  for (int i = 0; i < 2 * timer_.capacity; ++i) (void)hipEventDestroy(timer_.events[i]);
  delete[] timer_.events;
  if (stream_) (void)hipStreamDestroy(stream_);
}

template <class T>
void GpuWorld::ensure(T*& ptr, size_t& cap, size_t n, bool keep) {
  if (n <= cap && ptr) return;
  size_t new_cap = std::max<size_t>(n + n / 2, 64);
  T* p = nullptr;
  NSL_CHECK(hipMalloc(&p, new_cap * sizeof(T)));
  if (ptr) {
    if (keep && cap > 0) NSL_CHECK(hipMemcpyAsync(p, ptr, cap * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
    NSL_CHECK(hipFree(ptr));
  }
  ptr = p;
  cap = new_cap;
}

// ---------------------------------------------------------------------------------
// creation: the host scene builds the records; the device copies follow at the next step

int32_t GpuWorld::create_body(const BodyDesc& desc) {
  if (device_newer_) pull_bodies();
  const int32_t id = HostScene::create_body(desc);
  bodies_dirty_ = true;
  // world_runtime.rs:1626-1629: kinematic and dynamic bodies get an island
  if (bodies_[(size_t)id].set == 2) {
    const int32_t island = create_island(2);
    islands_[(size_t)island].bodies.push_back(id);
    bodies_[(size_t)id].island = island;
    bodies_[(size_t)id].island_index = 0;
  }
  return id;
}

int32_t GpuWorld::add_shape(int32_t body, Shape shape, const Material& m, float radius) {
  if (body >= 0 && (size_t)body < bodies_.size() && bodies_[(size_t)body].set >= 3)
    throw std::invalid_argument("shapes cannot be added to a sleeping body");
  if (device_newer_) pull_bodies();
  isotropic_inertia_tolerance = options.isotropic_inertia_tolerance;
  const int32_t id = HostScene::add_shape(body, shape, m, radius);
  shapes_dirty_ = true;
  bodies_dirty_ = true;
  return id;
}

// ---------------------------------------------------------------------------------
// device residency of bodies and shapes

void GpuWorld::upload_bodies() {
  const size_t na = awake_host_.size(), ns = static_host_.size();
  ensure(d_.sims, cap_sims_, std::max<size_t>(na, 1), false);
  ensure(d_.states, cap_states_, std::max<size_t>(na, 1), false);
  ensure(d_.static_sims, cap_static_, std::max<size_t>(ns, 1), false);
  ensure(d_.fast_ids, cap_fast_, std::max<size_t>(na, 1), false);
  if (na) {
    NSL_CHECK(hipMemcpyAsync(d_.sims, awake_host_.data(), na * sizeof(BodySim), hipMemcpyHostToDevice, stream_));
    NSL_CHECK(hipMemcpyAsync(d_.states, states_host_.data(), na * sizeof(BodyState), hipMemcpyHostToDevice, stream_));
  }
  if (ns) NSL_CHECK(hipMemcpyAsync(d_.static_sims, static_host_.data(), ns * sizeof(BodySim), hipMemcpyHostToDevice, stream_));
  d_.n_awake = (int32_t)na;
  if (shapes_dirty_) {
    const size_t nsh = shapes_.size();
    ensure(d_.shapes, cap_shapes_, std::max<size_t>(nsh, 1), false);
    ensure(d_.fat_aabbs, cap_fat_, std::max<size_t>(nsh, 1), false);
    ensure(d_.shape_moved, cap_moved_, std::max<size_t>(nsh, 1), false);
    ensure(d_proxies_, cap_proxies_, std::max<size_t>(nsh, 1), false);
    ensure(d_.shape_aabbs, cap_aabbs_, std::max<size_t>(nsh, 1), false);
    ensure(d_candidates_, cap_candidates_, std::max<size_t>(nsh, 1), false);
    ensure(d_.hulls, cap_hulls_, std::max<size_t>(hulls_.size(), 1), false);
    fat_host_.resize(nsh);
    aabb_host_.resize(nsh);
    moved_host_.resize(nsh);
    proxies_host_.resize(nsh);
    std::vector<int32_t> candidates;  // static, kinematic, dynamic shapes (continuous.rs:424-445)
    int32_t per_type[3] = {0, 0, 0};
    candidates.reserve(nsh);
    for (int type = kBodyStatic; type <= kBodyDynamic; ++type)
      for (size_t i = 0; i < nsh; ++i)
        if (bodies_[hshapes_[i].body].type == type) {
          candidates.push_back((int32_t)i);
          per_type[type] += 1;
        }
    for (size_t i = 0; i < nsh; ++i) {
      const HShape& hs = hshapes_[i];
      fat_host_[i] = hs.fat;
      aabb_host_[i] = hs.aabb;
      moved_host_[i] = hs.proxy_moved ? 1 : 0;
      pairs_pending_ = pairs_pending_ || hs.proxy_moved;
      proxies_host_[i] = Proxy{hs.body, bodies_[hs.body].type, hs.group,
                               (hs.speculative ? kProxySpeculative : 0) | (hs.sensor ? kProxySensor : 0), hs.category, hs.mask};
    }
    if (nsh) {
      NSL_CHECK(hipMemcpyAsync(d_.shapes, shapes_.data(), nsh * sizeof(Shape), hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_.fat_aabbs, fat_host_.data(), nsh * sizeof(Aabb), hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_.shape_aabbs, aabb_host_.data(), nsh * sizeof(Aabb), hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_candidates_, candidates.data(), nsh * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_.shape_moved, moved_host_.data(), nsh, hipMemcpyHostToDevice, stream_));
      NSL_CHECK(hipMemcpyAsync(d_proxies_, proxies_host_.data(), nsh * sizeof(Proxy), hipMemcpyHostToDevice, stream_));
    }
    if (!hulls_.empty())
      NSL_CHECK(hipMemcpyAsync(d_.hulls, hulls_.data(), hulls_.size() * sizeof(BoxHull), hipMemcpyHostToDevice, stream_));
    // the staging vectors are read by the copies above
    NSL_CHECK(hipStreamSynchronize(stream_));
    n_device_shapes_ = nsh;
    shapes_dirty_ = false;
    ccd_ = ContinuousArgs{d_candidates_, per_type[kBodyStatic], per_type[kBodyKinematic], per_type[kBodyDynamic], d_proxies_};
  }
  bullet_bodies_ = 0;
  kinematic_awake_ = 0;
  for (const BodySim& s : awake_host_) {
    if (s.body_type == kBodyDynamic && (s.flags & bflag::kIsBullet)) bullet_bodies_ += 1;
    kinematic_awake_ += s.body_type != kBodyDynamic;
  }
  bodies_dirty_ = false;
  device_newer_ = false;
}

void GpuWorld::pull_bodies() {
  const size_t na = awake_host_.size();
  if (na && d_.sims) {
    NSL_CHECK(hipMemcpyAsync(awake_host_.data(), d_.sims, na * sizeof(BodySim), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipMemcpyAsync(states_host_.data(), d_.states, na * sizeof(BodyState), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
  }
  if (n_device_shapes_) {  // finalize moves fat AABBs and marks proxies on the device
    fat_host_.resize(n_device_shapes_);
    aabb_host_.resize(n_device_shapes_);
    moved_host_.resize(n_device_shapes_);
    NSL_CHECK(hipMemcpyAsync(fat_host_.data(), d_.fat_aabbs, n_device_shapes_ * sizeof(Aabb), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipMemcpyAsync(aabb_host_.data(), d_.shape_aabbs, n_device_shapes_ * sizeof(Aabb), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipMemcpyAsync(moved_host_.data(), d_.shape_moved, n_device_shapes_, hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n_device_shapes_; ++i) {
      hshapes_[i].fat = fat_host_[i];
      hshapes_[i].aabb = aabb_host_[i];
      hshapes_[i].proxy_moved = moved_host_[i] != 0;
    }
  }
  for (size_t i = 0; i < awake_host_.size(); ++i) bodies_[awake_host_[i].body_id].flags = awake_host_[i].body_flags;
  device_newer_ = false;
}

// ---------------------------------------------------------------------------------
// pair update: contact_creation.rs:901-1042. New pairs come from the device pair finder
// (nsl_broadphase.h) as sorted keys; contacts are created in that order.

void GpuWorld::update_pairs() {
  if (!pairs_pending_ && counters_host_[6] == 0) return;
  // the last step's finalize counted the proxies it moved; other moves (new shapes) leave
  // the count to the finder
  finder_->find(d_, d_proxies_, (int)n_device_shapes_, new_keys_, pairs_pending_ ? -1 : (int64_t)counters_host_[6]);
  for (HShape& s : hshapes_) s.proxy_moved = false;
  pairs_pending_ = false;
  kstats_.pair_tests += finder_->last_tests();
  for (uint64_t key : new_keys_)
    create_contact((int32_t)((key >> 42) & ((1u << 22) - 1)), (int32_t)((key >> 20) & ((1u << 22) - 1)));
}

// Keeps the device pair-key set at a load factor of at most 1/2 (live keys plus
// tombstones); rebuilds it from the live contacts when that would be exceeded.
void GpuWorld::ensure_pair_set() {
  const int64_t live = contact_count();
  const uint32_t cap = d_.pairs.slots ? d_.pairs.mask + 1 : 0;
  if (cap && 2 * (live + pair_set_tombs_) <= (int64_t)cap) return;
  uint32_t want = 1024;
  while ((int64_t)want < 4 * live) want <<= 1;
  if (want != cap) {
    if (d_.pairs.slots) {
      NSL_CHECK(hipStreamSynchronize(stream_));
      NSL_CHECK(hipFree(d_.pairs.slots));
    }
    NSL_CHECK(hipMalloc(&d_.pairs.slots, (size_t)want * sizeof(uint64_t)));
    d_.pairs.mask = want - 1;
  }
  NSL_CHECK(hipMemsetAsync(d_.pairs.slots, 0xFF, (size_t)want * sizeof(uint64_t), stream_));  // all kPairEmpty
  std::vector<uint64_t> keys;
  keys.reserve((size_t)live);
  for (const HContact& c : hcontacts_)
    if (c.alive) keys.push_back(pair_key(c.shape_a, c.shape_b));
  if (!keys.empty()) {
    ensure(d_rebuild_keys_, cap_rebuild_keys_, keys.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_rebuild_keys_, keys.data(), keys.size() * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
    launch_pairset_insert(d_.pairs, d_rebuild_keys_, (int)keys.size(), stream_);
    NSL_CHECK(hipStreamSynchronize(stream_));  // `keys` is read by the copy
  }
  pair_set_tombs_ = 0;
}

// contact_creation.rs:219-425; the primary shape order is hull > capsule > sphere.
void GpuWorld::create_contact(int32_t a, int32_t b) {
  auto rank = [](int32_t kind) { return kind == kShapeHull ? 2 : (kind == kShapeCapsule ? 1 : 0); };
  if (rank(hshapes_[a].kind) < rank(hshapes_[b].kind)) std::swap(a, b);
  const HShape& sa = hshapes_[a];
  const HShape& sb = hshapes_[b];
  const HBody& ba = bodies_[sa.body];
  const HBody& bb = bodies_[sb.body];
  int32_t id;
  if (!contact_free_.empty()) {
    id = contact_free_.back();
    contact_free_.pop_back();
  } else {
    id = contact_pool_next_++;
  }
  if (id >= (int32_t)hcontacts_.size()) hcontacts_.resize(id + 1);
  hcontacts_[id] = HContact{};
  HContact& hc = hcontacts_[id];
  hc.shape_a = a;
  hc.shape_b = b;
  hc.color = kNullIndex;
  // the awake list when a body is awake, else the disabled set
  hc.set = (ba.set == 2 || bb.set == 2) ? 2 : 1;
  hc.alive = 1;
  hc.restitution = maxf(sa.restitution, sb.restitution);
  std::vector<int32_t>& list = hc.set == 2 ? non_touching_ : disabled_;
  hc.local = (int32_t)list.size();
  list.push_back(id);
  hc.body[0] = sa.body;
  hc.body[1] = sb.body;
  link_body_contact(id);
  if (hc.set != 2) mark_contact_dirty(id);  // the device turns new contacts on
  init_ids_.push_back(id);
  Contact& c = init_records_.emplace_back();  // zeroed in place: the record the device starts from
  c.shape_a = a;
  c.shape_b = b;
  c.enc_a = encode(ba);
  c.enc_b = encode(bb);
  if (ba.type == kBodyStatic || bb.type == kBodyStatic) c.flags |= cflag::kStatic;
  if (options.enable_speculative && sa.speculative && sb.speculative) c.flags |= cflag::kSpeculative;
  if ((ba.flags & bflag::kEnableRecycling) && (bb.flags & bflag::kEnableRecycling) && (c.flags & cflag::kSpeculative))
    c.flags |= cflag::kRecycle;
  c.rolling_resistance = maxf(sa.rolling_resistance, sb.rolling_resistance) * maxf(sa.radius, sb.radius);
  c.cached_rot_a = quat_identity();
  c.cached_rot_b = quat_identity();
  hc.created = ++contact_serial_;
}

template <class T>
static int32_t swap_remove(std::vector<T>& v, size_t index) {
  size_t last = v.size() - 1;
  if (index != last) {
    v[index] = v[last];
    v.pop_back();
    return (int32_t)last;
  }
  v.pop_back();
  return kNullIndex;
}

// contact_creation.rs:428-552. `device_off`: the collide kernel already turned the
// contact's flag off (it found it disjoint).
void GpuWorld::destroy_contact(int32_t id, bool device_off) {
  HContact& c = hcontacts_[id];
  if (!c.alive) return;
  if (!device_off) mark_contact_dirty(id);
  pair_set_tombs_ += 1;  // the collide kernel erased the key
  unlink_body_contact(id);
  if (c.island != kNullIndex) unlink_contact(id);  // contact_creation.rs:501-508
  if (c.color == kNullIndex) {
    int32_t local = c.local;
    std::vector<int32_t>& list = contact_list(c.set);
    if (swap_remove(list, (size_t)local) != kNullIndex) hcontacts_[list[local]].local = local;
  } else {
    remove_from_graph(hshapes_[c.shape_a].body, hshapes_[c.shape_b].body, c.color, c.local);
  }
  HContact& c2 = hcontacts_[id];
  c2.alive = 0;
  c2.color = kNullIndex;
  c2.local = kNullIndex;
  c2.set = kNullIndex;
  // id_pool.rs:36-45 shrinks the pool only for id == next_index, which no live id equals:
  // every freed id goes to the free list
  contact_free_.push_back(id);
}

// constraint_graph.rs:113-197
void GpuWorld::add_to_graph(int32_t id) {
  HContact& c = hcontacts_[id];
  int32_t a = hshapes_[c.shape_a].body, b = hshapes_[c.shape_b].body;
  int32_t ta = bodies_[a].type, tb = bodies_[b].type;
  int32_t color = kOverflowColor;
  if (ta == kBodyDynamic && tb == kBodyDynamic) {
    color = free_colour_dynamic(a, b);
    if (color != kOverflowColor) {
      colour_set(color, a);
      colour_set(color, b);
    }
  } else if (ta == kBodyDynamic || tb == kBodyDynamic) {
    const int32_t dyn = ta == kBodyDynamic ? a : b;
    color = free_colour_static(dyn);
    if (color != kOverflowColor) colour_set(color, dyn);
  }
  c.color = color;
  graph_restitution_ += c.restitution > 0.0f;
  if (color == kOverflowColor) {
    c.local = (int32_t)colors_[color].overflow.size();
    colors_[color].overflow.push_back(id);
  } else {
    c.local = (int32_t)colors_[color].convex.size();
    colors_[color].convex.push_back(id);
  }
}

// constraint_graph.rs:200-251
void GpuWorld::remove_from_graph(int32_t body_a, int32_t body_b, int32_t color, int32_t local) {
  const int32_t id = color == kOverflowColor ? colors_[color].overflow[(size_t)local] : colors_[color].convex[(size_t)local];
  graph_restitution_ -= hcontacts_[(size_t)id].restitution > 0.0f;
  if (color != kOverflowColor) {
    colour_clear(color, body_a);
    colour_clear(color, body_b);
    if (swap_remove(colors_[color].convex, (size_t)local) != kNullIndex) hcontacts_[colors_[color].convex[local]].local = local;
  } else {
    if (swap_remove(colors_[color].overflow, (size_t)local) != kNullIndex) hcontacts_[colors_[color].overflow[local]].local = local;
  }
}

// collide_pass.rs:386-478: transitions in contact-id order.
void GpuWorld::state_pass() {
  for (const int32_t e : changes_) {
    const size_t id = (size_t)(e >> 2);
    const uint8_t ch = (uint8_t)(e & 3);
    if (!hcontacts_[id].alive) continue;
    if (ch == kChangeDisjoint) {
      destroy_contact((int32_t)id, true);
    } else if (ch == kChangeBegan) {
      // linking first wakes a sleeping partner (island.rs:126-171)
      link_contact((int32_t)id);
      int32_t old_local = hcontacts_[id].local;
      add_to_graph((int32_t)id);
      if (swap_remove(non_touching_, (size_t)old_local) != kNullIndex) hcontacts_[non_touching_[old_local]].local = old_local;
    } else if (ch == kChangeEnded) {
      unlink_contact((int32_t)id);
      HContact& c = hcontacts_[id];
      int32_t color = c.color, local = c.local;
      c.color = kNullIndex;
      c.local = (int32_t)non_touching_.size();
      non_touching_.push_back((int32_t)id);
      remove_from_graph(hshapes_[c.shape_a].body, hshapes_[c.shape_b].body, color, local);
    }
  }
}

// ---------------------------------------------------------------------------------
// joints: joint.rs:130-375, constraint_graph.rs:256-323

// constraint_graph.rs:256-298: the contact colouring's rules on the same body bits.
int32_t GpuWorld::assign_joint_color(int32_t a, int32_t b) {
  const int32_t ta = bodies_[a].type, tb = bodies_[b].type;
  if (ta == kBodyDynamic && tb == kBodyDynamic) {
    const int32_t i = free_colour_dynamic(a, b);
    if (i != kOverflowColor) {
      colour_set(i, a);
      colour_set(i, b);
    }
    return i;
  } else if (ta == kBodyDynamic || tb == kBodyDynamic) {
    const int32_t dyn = ta == kBodyDynamic ? a : b;
    const int32_t i = free_colour_static(dyn);
    if (i != kOverflowColor) colour_set(i, dyn);
    return i;
  }
  return kOverflowColor;
}

// The device set of body pairs that may not touch, rebuilt whole (joints are few and
// created rarely): power-of-two table at most a quarter full.
void GpuWorld::upload_joint_pairs() {
  if (blocked_keys_.empty()) return;
  uint32_t want = 64;
  while ((size_t)want < 4 * blocked_keys_.size()) want <<= 1;
  if (!d_.joint_pairs.slots || d_.joint_pairs.mask + 1 != want) {
    if (d_.joint_pairs.slots) {
      NSL_CHECK(hipStreamSynchronize(stream_));
      NSL_CHECK(hipFree(d_.joint_pairs.slots));
    }
    NSL_CHECK(hipMalloc(&d_.joint_pairs.slots, (size_t)want * sizeof(uint64_t)));
    d_.joint_pairs.mask = want - 1;
  }
  NSL_CHECK(hipMemsetAsync(d_.joint_pairs.slots, 0xFF, (size_t)want * sizeof(uint64_t), stream_));
  ensure(d_erase_keys_, cap_erase_keys_, blocked_keys_.size(), false);
  NSL_CHECK(hipMemcpyAsync(d_erase_keys_, blocked_keys_.data(), blocked_keys_.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                           stream_));
  launch_pairset_insert(d_.joint_pairs, d_erase_keys_, (int)blocked_keys_.size(), stream_);
  NSL_CHECK(hipStreamSynchronize(stream_));
}

// joint.rs:164-375 + revolute_joint.rs:149-174: a joint with a dynamic body wakes a
// sleeping one, goes into the graph and joins the bodies' islands; one between non-dynamic
// bodies is kept but never solved.
int32_t GpuWorld::create_revolute_joint(const JointDesc& desc) {
  const int32_t a = desc.body_a, b = desc.body_b;
  if (a == b || a < 0 || b < 0 || a >= (int32_t)bodies_.size() || b >= (int32_t)bodies_.size())
    throw std::invalid_argument("joint bodies must be two distinct existing bodies");
  // the reference keeps such a joint asleep by merging sleeping sets (joint.rs:294-340)
  if (bodies_[a].set != 2 && bodies_[b].set != 2 && (bodies_[a].set >= 3 || bodies_[b].set >= 3))
    throw std::invalid_argument("a joint between a sleeping body and no awake body is not supported");
  NSL_CHECK(hipSetDevice(device_));

  if (!desc.collide_connected) {
    // joint.rs:130-161: the contacts between the bodies, newest first, are destroyed now
    std::vector<std::pair<uint64_t, int32_t>> between;
    for (size_t id = 0; id < hcontacts_.size(); ++id) {
      const HContact& c = hcontacts_[id];
      if (!c.alive) continue;
      const int32_t ba = hshapes_[c.shape_a].body, bb = hshapes_[c.shape_b].body;
      if ((ba == a && bb == b) || (ba == b && bb == a)) between.push_back({c.created, (int32_t)id});
    }
    std::sort(between.begin(), between.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<uint64_t> erase;
    for (const auto& e : between) {
      const HContact& c = hcontacts_[e.second];
      erase.push_back(pair_key(c.shape_a, c.shape_b));
      destroy_contact(e.second);
    }
    if (!erase.empty()) {
      ensure(d_erase_keys_, cap_erase_keys_, erase.size(), false);
      NSL_CHECK(hipMemcpyAsync(d_erase_keys_, erase.data(), erase.size() * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
      launch_pairset_erase(d_.pairs, d_erase_keys_, (int)erase.size(), stream_);
      NSL_CHECK(hipStreamSynchronize(stream_));
    }
    blocked_keys_.push_back(pair_key(a, b));
    upload_joint_pairs();
  }

  const int32_t id = (int32_t)hjoints_.size();
  HJoint hj;
  hj.body_a = a;
  hj.body_b = b;
  hjoints_.push_back(hj);
  if (bodies_[a].type == kBodyDynamic || bodies_[b].type == kBodyDynamic) {
    const int32_t max_set = std::max(bodies_[a].set, bodies_[b].set);
    if (max_set >= 3) wake_solver_set(max_set);
    hjoints_[id].set = 2;
    add_joint_to_graph(id);
  }

  JointSim j{};
  j.enc_a = hjoints_[id].enc_a = encode(bodies_[a]);
  j.enc_b = hjoints_[id].enc_b = encode(bodies_[b]);
  j.local_frame_a = desc.local_frame_a;
  j.local_frame_b = desc.local_frame_b;
  j.constraint_hertz = desc.constraint_hertz;
  j.constraint_damping_ratio = desc.constraint_damping_ratio;
  j.constraint_softness = Softness{0.0f, 1.0f, 0.0f};
  j.hertz = desc.hertz;
  j.damping_ratio = desc.damping_ratio;
  j.target_angle = desc.target_angle;
  j.lower_angle = desc.lower_angle;
  j.upper_angle = desc.upper_angle;
  j.max_motor_torque = desc.max_motor_torque;
  j.motor_speed = desc.motor_speed;
  j.flags = (desc.enable_spring ? jflag::kSpring : 0u) | (desc.enable_motor ? jflag::kMotor : 0u) |
            (desc.enable_limit ? jflag::kLimit : 0u);
  j.index_a = j.index_b = kNullIndex;
  ensure(d_.joints, cap_joints_, hjoints_.size(), true);
  NSL_CHECK(hipMemcpyAsync(&d_.joints[id], &j, sizeof(JointSim), hipMemcpyHostToDevice, stream_));
  NSL_CHECK(hipStreamSynchronize(stream_));
  joints_dirty_ = true;
  if (hjoints_[id].set > 1) link_joint(id);  // joint.rs:366-370
  return id;
}

void GpuWorld::get_joint(int32_t id, float* impulses, int32_t* color) {
  if (id < 0 || id >= (int32_t)hjoints_.size()) throw std::invalid_argument("no such joint");
  NSL_CHECK(hipSetDevice(device_));
  JointSim j{};
  NSL_CHECK(hipMemcpyAsync(&j, &d_.joints[id], sizeof(JointSim), hipMemcpyDeviceToHost, stream_));
  NSL_CHECK(hipStreamSynchronize(stream_));
  const float vals[9] = {j.linear_impulse.x, j.linear_impulse.y, j.linear_impulse.z, j.perp_impulse.x, j.perp_impulse.y,
                         j.spring_impulse, j.motor_impulse, j.lower_impulse, j.upper_impulse};
  if (impulses) std::memcpy(impulses, vals, sizeof(vals));
  if (color) *color = hjoints_[id].color;
}

// The overflow colour's joints (first) and contacts in list order, grouped into dependency
// levels through their dynamic bodies; uploads items and level offsets, returns the count
// of levels.
int32_t GpuWorld::build_overflow_levels() {
  const std::vector<int32_t>& ovf_joints = colors_[kOverflowColor].joints;
  const std::vector<int32_t>& ovf = colors_[kOverflowColor].overflow;
  const size_t n = ovf_joints.size() + ovf.size();
  ovf_bodies_host_.clear();
  if (n == 0) return 0;
  if (body_level_.size() < bodies_.size()) body_level_.resize(bodies_.size(), 0);
  std::vector<int32_t> level(n);
  std::vector<int32_t> touched;
  int32_t n_levels = 0;
  for (size_t i = 0; i < n; ++i) {
    int32_t a, b;
    if (i < ovf_joints.size()) {
      a = hjoints_[ovf_joints[i]].body_a;
      b = hjoints_[ovf_joints[i]].body_b;
    } else {
      const HContact& c = hcontacts_[ovf[i - ovf_joints.size()]];
      a = hshapes_[c.shape_a].body;
      b = hshapes_[c.shape_b].body;
    }
    const bool da = bodies_[a].type == kBodyDynamic, db = bodies_[b].type == kBodyDynamic;
    const int32_t l = std::max(da ? body_level_[a] : 0, db ? body_level_[b] : 0);  // body_level_: last level + 1
    level[i] = l;
    n_levels = std::max(n_levels, l + 1);
    if (da) { body_level_[a] = l + 1; touched.push_back(a); }
    if (db) { body_level_[b] = l + 1; touched.push_back(b); }
  }
  // reset the levels, listing each dynamic body once (its awake index)
  ovf_bodies_host_.clear();
  for (int32_t t : touched)
    if (body_level_[t] != 0) {
      body_level_[t] = 0;
      if (bodies_[t].set == 2) ovf_bodies_host_.push_back(bodies_[t].local);
    }
  ovf_level_host_.assign(n_levels + 1, 0);
  for (size_t i = 0; i < n; ++i) ovf_level_host_[level[i] + 1] += 1;
  for (int32_t l = 0; l < n_levels; ++l) ovf_level_host_[l + 1] += ovf_level_host_[l];
  ovf_items_host_.assign(n, 0);
  std::vector<int32_t> fill(ovf_level_host_.begin(), ovf_level_host_.end() - 1);
  for (size_t i = 0; i < n; ++i)
    ovf_items_host_[fill[level[i]]++] = i < ovf_joints.size() ? -(ovf_joints[i] + 1) : (int32_t)(i - ovf_joints.size());
  ensure(d_ovf_items_, cap_ovf_items_, n, false);
  ensure(d_ovf_levels_, cap_ovf_levels_, ovf_level_host_.size(), false);
  NSL_CHECK(hipMemcpyAsync(d_ovf_items_, ovf_items_host_.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  NSL_CHECK(hipMemcpyAsync(d_ovf_levels_, ovf_level_host_.data(), ovf_level_host_.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, stream_));
  if (!ovf_bodies_host_.empty()) {
    ensure(d_ovf_bodies_, cap_ovf_bodies_, ovf_bodies_host_.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_ovf_bodies_, ovf_bodies_host_.data(), ovf_bodies_host_.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, stream_));
  }
  return n_levels;
}

// ---------------------------------------------------------------------------------
// snapshots: world_state.rs:340-560

struct GpuWorld::Snapshot {
  int device = 0;
  size_t n_bodies = 0, n_shapes = 0, n_joints = 0, n_awake = 0, n_device_shapes = 0, n_contacts = 0;
  std::vector<void*> buffers;
  BodySim* sims = nullptr;
  BodyState* states = nullptr;
  Aabb* fat = nullptr;
  Aabb* aabbs = nullptr;
  uint8_t* moved = nullptr;
  Contact* contacts = nullptr;
  uint64_t* pair_slots = nullptr;
  uint32_t pair_mask = 0;
  JointSim* joints = nullptr;
  BodySim* sleep_sims = nullptr;
  size_t n_sleep_sims = 0;
  int32_t counters[kCounterCount] = {};
  std::vector<HBody> bodies;
  std::vector<HJoint> hjoints;
  std::vector<BodySim> sleep_host;
  std::vector<Island> islands;
  IdPool island_pool, set_pool;
  std::vector<int32_t> awake_islands, disabled;
  std::vector<SleepSet> sleep_sets;
  int32_t split_island = -1;
  std::vector<BodySim> awake_host;
  std::vector<BodyState> states_host;
  std::vector<HShape> hshapes;
  std::vector<HContact> hcontacts;
  std::vector<int32_t> contact_free, non_touching;
  int32_t contact_pool_next = 0;
  ColorSet colors[kColorCount];
  std::vector<uint32_t> body_colours;
  int32_t counters_host[kCounterCount] = {}, counters_base[kCounterCount] = {};
  int64_t pair_set_tombs = 0;
  bool pairs_pending = false;
  uint64_t contact_serial = 0;
  ~Snapshot() {
    if (!buffers.empty()) (void)hipSetDevice(device);
    for (void* p : buffers) (void)hipFree(p);
  }
  template <class T> T* copy_of(const T* src, size_t n, hipStream_t stream) {
    if (n == 0 || src == nullptr) return nullptr;
    T* p = nullptr;
    NSL_CHECK(hipMalloc(&p, n * sizeof(T)));
    buffers.push_back(p);
    NSL_CHECK(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
    return p;
  }
};

void GpuWorld::destroy_snapshot(Snapshot* s) { delete s; }

GpuWorld::Tuning GpuWorld::autotune(float dt, int32_t sub_steps, int32_t steps) {
  check_usable();
  std::unique_ptr<Snapshot> snap(save_state());
  const WorldOptions saved = options;
  const bool was_timing = timing_, was_phase = phase_timing_;
  // steps of one candidate from the snapshot: average step-kernel time, end hash, workgroups
  auto trial = [&](const WorldOptions& o, uint64_t* hash, int* blocks) {
    options = o;
    restore_state(*snap);
    set_kernel_timing(true, false);
    kstats_ = KernelStats{};
    for (int k = 0; k < steps; ++k) step(dt, sub_steps);
    *hash = state_hash();
    *blocks = (int)kstats_.solver_blocks;
    return 1e3 * kstats_.ms[kKSolve] / (double)std::max<int64_t>(kstats_.launches[kKSolve], 1);
  };
  Tuning t;
  uint64_t h0 = 0;
  int d0 = 0;
  try {
    t.previous_us = trial(saved, &h0, &d0);
    const int cap = std::max(1, std::min(max_blocks_, 96));
    std::vector<int> sizes;
    for (int num : {2, 3, 4, 5, 6}) {  // d0 x 2/4 .. 6/4
      const int b = std::max(1, std::min(cap, (d0 * num + 2) / 4));
      if (std::find(sizes.begin(), sizes.end(), b) == sizes.end()) sizes.push_back(b);
    }
    t.best_us = t.previous_us;
    t.blocks = d0;
    t.small_colour_items = saved.small_colour_items;
    t.stage_records = saved.stage_records;
    for (int b : sizes)
      for (int small : {0, 128, 256})
        for (int stage : {0, 1}) {
          WorldOptions o = saved;
          o.solver_blocks = b;
          o.small_colour_items = small;
          o.stage_records = stage;
          uint64_t h = 0;
          int used = 0;
          const double us = trial(o, &h, &used);
          t.candidates += 1;
          if (h != h0) throw std::runtime_error("autotune: a launch shape changed the results (it must not)");
          if (us < t.best_us) {
            t.best_us = us;
            t.blocks = b;
            t.small_colour_items = small;
            t.stage_records = stage;
          }
        }
  } catch (...) {
    options = saved;
    restore_state(*snap);
    set_kernel_timing(was_timing, was_phase);
    throw;
  }
  options = saved;
  options.solver_blocks = t.blocks;
  options.small_colour_items = t.small_colour_items;
  options.stage_records = t.stage_records;
  restore_state(*snap);
  set_kernel_timing(was_timing, was_phase);
  reset_kernel_stats();
  return t;
}

GpuWorld::Snapshot* GpuWorld::save_state() {
  check_usable();
  NSL_CHECK(hipSetDevice(device_));
  flush_locations();
  if (device_newer_) pull_bodies();
  auto s = std::make_unique<Snapshot>();
  s->device = device_;
  s->n_bodies = bodies_.size();
  s->n_shapes = hshapes_.size();
  s->n_joints = hjoints_.size();
  s->n_awake = awake_host_.size();
  s->n_device_shapes = n_device_shapes_;
  s->n_contacts = hcontacts_.size();
  s->sims = s->copy_of(d_.sims, s->n_awake, stream_);
  s->states = s->copy_of(d_.states, s->n_awake, stream_);
  s->fat = s->copy_of(d_.fat_aabbs, n_device_shapes_, stream_);
  s->aabbs = s->copy_of(d_.shape_aabbs, n_device_shapes_, stream_);
  s->moved = s->copy_of(d_.shape_moved, n_device_shapes_, stream_);
  s->contacts = s->copy_of(d_.contacts, s->n_contacts, stream_);
  if (d_.pairs.slots) {
    s->pair_mask = d_.pairs.mask;
    s->pair_slots = s->copy_of(d_.pairs.slots, (size_t)d_.pairs.mask + 1, stream_);
  }
  s->joints = s->copy_of(d_.joints, hjoints_.size(), stream_);
  if (d_.sleep_sims) {
    s->n_sleep_sims = std::min(cap_sleep_sims_, bodies_.size());
    s->sleep_sims = s->copy_of(d_.sleep_sims, s->n_sleep_sims, stream_);
  }
  NSL_CHECK(hipMemcpyAsync(s->counters, d_.counters, sizeof(s->counters), hipMemcpyDeviceToHost, stream_));
  NSL_CHECK(hipStreamSynchronize(stream_));
  s->bodies = bodies_;
  s->awake_host = awake_host_;
  s->states_host = states_host_;
  s->hshapes = hshapes_;
  s->hcontacts = hcontacts_;
  s->contact_free = contact_free_;
  s->non_touching = non_touching_;
  s->contact_pool_next = contact_pool_next_;
  for (int c = 0; c < kColorCount; ++c) s->colors[c] = colors_[c];
  s->body_colours = body_colours_;
  std::memcpy(s->counters_host, counters_host_, sizeof(counters_host_));
  std::memcpy(s->counters_base, counters_base_, sizeof(counters_base_));
  s->pair_set_tombs = pair_set_tombs_;
  s->pairs_pending = pairs_pending_;
  s->contact_serial = contact_serial_;
  s->hjoints = hjoints_;
  s->sleep_host = sleep_host_;
  s->islands = islands_;
  s->island_pool = island_pool_;
  s->set_pool = set_pool_;
  s->awake_islands = awake_islands_;
  s->disabled = disabled_;
  s->sleep_sets = sleep_sets_;
  s->split_island = split_island_;
  return s.release();
}

void GpuWorld::restore_state(const Snapshot& s) {
  if (s.n_bodies != bodies_.size() || s.n_shapes != hshapes_.size() || s.n_joints != hjoints_.size())
    throw std::invalid_argument("restore_state: the world's bodies, shapes or joints changed since the snapshot");
  NSL_CHECK(hipSetDevice(device_));
  if (bodies_dirty_ || shapes_dirty_) upload_bodies();
  NSL_CHECK(hipStreamSynchronize(stream_));
  // the awake set's size and order come back with the snapshot's host records
  if (s.n_awake > cap_sims_ || s.n_awake > cap_states_) {
    ensure(d_.sims, cap_sims_, s.n_awake, false);
    ensure(d_.states, cap_states_, s.n_awake, false);
  }
  ensure(d_.fast_ids, cap_fast_, std::max<size_t>(s.n_awake, 1), false);
  if (s.sleep_sims) ensure(d_.sleep_sims, cap_sleep_sims_, s.n_sleep_sims, true);
  auto back = [&](auto* dst, const auto* src, size_t n) {
    if (n) NSL_CHECK(hipMemcpyAsync(dst, src, n * sizeof(*src), hipMemcpyDeviceToDevice, stream_));
  };
  back(d_.sims, s.sims, s.n_awake);
  back(d_.states, s.states, s.n_awake);
  back(d_.fat_aabbs, s.fat, s.n_device_shapes);
  back(d_.shape_aabbs, s.aabbs, s.n_device_shapes);
  back(d_.shape_moved, s.moved, s.n_device_shapes);
  ensure(d_.contacts, cap_contacts_, std::max<size_t>(s.n_contacts, 1), false);
  back(d_.contacts, s.contacts, s.n_contacts);
  if (s.pair_slots) {
    if (!d_.pairs.slots || d_.pairs.mask != s.pair_mask) {
      if (d_.pairs.slots) NSL_CHECK(hipFree(d_.pairs.slots));
      NSL_CHECK(hipMalloc(&d_.pairs.slots, ((size_t)s.pair_mask + 1) * sizeof(uint64_t)));
      d_.pairs.mask = s.pair_mask;
    }
    back(d_.pairs.slots, s.pair_slots, (size_t)s.pair_mask + 1);
  }
  back(d_.joints, s.joints, s.n_joints);
  if (s.sleep_sims) back(d_.sleep_sims, s.sleep_sims, s.n_sleep_sims);
  NSL_CHECK(hipMemcpyAsync(d_.counters, s.counters, sizeof(s.counters), hipMemcpyHostToDevice, stream_));
  NSL_CHECK(hipStreamSynchronize(stream_));
  bodies_ = s.bodies;
  awake_host_ = s.awake_host;
  states_host_ = s.states_host;
  hshapes_ = s.hshapes;
  hcontacts_ = s.hcontacts;
  contact_free_ = s.contact_free;
  non_touching_ = s.non_touching;
  contact_pool_next_ = s.contact_pool_next;
  for (int c = 0; c < kColorCount; ++c) colors_[c] = s.colors[c];
  body_colours_ = s.body_colours;
  graph_restitution_ = 0;
  for (int c = 0; c < kColorCount; ++c) {
    for (int32_t id : colors_[c].convex) graph_restitution_ += hcontacts_[(size_t)id].restitution > 0.0f;
    for (int32_t id : colors_[c].overflow) graph_restitution_ += hcontacts_[(size_t)id].restitution > 0.0f;
  }
  std::memcpy(counters_host_, s.counters_host, sizeof(counters_host_));
  std::memcpy(counters_base_, s.counters_base, sizeof(counters_base_));
  pair_set_tombs_ = s.pair_set_tombs;
  pairs_pending_ = s.pairs_pending;
  contact_serial_ = s.contact_serial;
  hjoints_ = s.hjoints;
  sleep_host_ = s.sleep_host;
  islands_ = s.islands;
  island_pool_ = s.island_pool;
  set_pool_ = s.set_pool;
  awake_islands_ = s.awake_islands;
  disabled_ = s.disabled;
  sleep_sets_ = s.sleep_sets;
  split_island_ = s.split_island;
  d_.n_awake = (int32_t)s.n_awake;
  bullet_bodies_ = 0;
  kinematic_awake_ = 0;
  for (const BodySim& b : awake_host_) {
    if (b.body_type == kBodyDynamic && (b.flags & bflag::kIsBullet)) bullet_bodies_ += 1;
    kinematic_awake_ += b.body_type != kBodyDynamic;
  }
  device_newer_ = false;
  joints_dirty_ = true;  // the colour lists of joints are rebuilt from colors_
  // the awake-set flags of every contact id, from the restored records
  {
    std::vector<uint8_t> active(hcontacts_.size(), 0);
    for (size_t i = 0; i < hcontacts_.size(); ++i) active[i] = hcontacts_[i].alive && hcontacts_[i].set == 2;
    ensure(d_.contact_active, cap_active_, std::max<size_t>(active.size(), 1), false);
    if (!active.empty())
      NSL_CHECK(hipMemcpyAsync(d_.contact_active, active.data(), active.size(), hipMemcpyHostToDevice, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
  }
  dirty_contacts_.clear();
  // shapes hold their bodies' encodings: bring them to the restored locations
  moved_bodies_.clear();
  slept_bodies_.clear();
  body_moved_.assign(bodies_.size(), 0);
  for (size_t b = 0; b < bodies_.size(); ++b) mark_body_moved((int32_t)b);
  flush_locations();
  poisoned_ = false;
}

// ---------------------------------------------------------------------------------
// one step: solver_integration.rs:807-1015

void GpuWorld::check_usable() const {
  if (poisoned_)
    throw std::runtime_error("the world was left part-way through a step by a grid barrier timeout: restore a snapshot "
                             "or create the world again");
}

void GpuWorld::step(float dt, int32_t sub_steps) {
  check_usable();
  NSL_CHECK(hipSetDevice(device_));
  const double t0 = now_ms();
  if (bodies_dirty_ || shapes_dirty_) upload_bodies();

  update_pairs();
  const double t1 = now_ms();

  StepParams p{};
  p.dt = dt;
  p.sub_steps = maxi(1, sub_steps);
  if (dt > 0.0f) {
    p.inv_dt = 1.0f / dt;
    p.h = dt / (float)p.sub_steps;
    p.inv_h = (float)p.sub_steps * p.inv_dt;
  }
  float contact_hz = minf(options.contact_hertz, 0.125f * p.inv_h);
  p.contact_softness = make_soft(contact_hz, options.contact_damping_ratio, p.h);
  p.static_softness = make_soft(2.0f * contact_hz, 0.5f * options.contact_damping_ratio, p.h);
  p.contact_speed = options.contact_speed;
  p.max_linear_speed = options.max_linear_speed;
  p.restitution_threshold = options.restitution_threshold;
  p.gravity = options.gravity;
  p.warm_start_scale = options.enable_warm_starting ? 1.0f : 0.0f;
  p.enable_sleep = options.enable_sleep;
  p.enable_continuous = options.enable_continuous;
  p.enable_restitution_propagation = options.enable_restitution_propagation;

  const size_t ncap = hcontacts_.size();
  ensure(d_.contacts, cap_contacts_, std::max<size_t>(ncap, 1), true);
  ensure(d_.contact_active, cap_active_, std::max<size_t>(ncap, 1), true);
  if (!init_ids_.empty()) {
    ensure_pair_set();
    ensure(d_init_ids_, cap_init_ids_, init_ids_.size(), false);
    ensure(d_init_records_, cap_init_rec_, init_records_.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_init_ids_, init_ids_.data(), init_ids_.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    NSL_CHECK(hipMemcpyAsync(d_init_records_, init_records_.data(), init_records_.size() * sizeof(Contact), hipMemcpyHostToDevice, stream_));
    launch_init_contacts(d_, d_init_ids_, d_init_records_, (int)init_ids_.size(), stream_);
    init_ids_.clear();
    init_records_.clear();
  }
  flush_locations();  // bodies put to sleep by the last step, bodies woken by joints
  if (dt <= 0.0f) return;

  // collide: every contact of the awake set, graph and non-touching (collide_pass.rs:175-190),
  // found on the device by its flag
  size_t awake_contacts = non_touching_.size();
  for (int ci = 0; ci < kColorCount; ++ci) awake_contacts += colors_[ci].convex.size() + colors_[ci].overflow.size();
  changes_.clear();
  if (awake_contacts > 0) {
    ensure(d_.change_list, cap_change_list_, ncap, false);  // at most one entry per contact id
    NSL_CHECK(hipMemsetAsync(d_.counters + kCntChangeList, 0, sizeof(int32_t), stream_));
    launch_collide(d_, (int)ncap, stream_, timing_ ? &timer_ : nullptr);
    // the list's length and a first part of it in one download; the rest only if longer
    const size_t guess = std::min(awake_contacts, std::max<size_t>(1024, 2 * last_changes_));
    ensure_pinned_changes(guess + 1);
    NSL_CHECK(hipMemcpyAsync(changes_pinned_, d_.counters + kCntChangeList, sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipMemcpyAsync(changes_pinned_ + 1, d_.change_list, guess * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    NSL_CHECK(hipStreamSynchronize(stream_));
    const size_t m = (size_t)changes_pinned_[0];
    if (m > guess) {
      ensure_pinned_changes(m + 1);
      NSL_CHECK(hipMemcpyAsync(changes_pinned_ + 1, d_.change_list, m * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
      NSL_CHECK(hipStreamSynchronize(stream_));
    }
    sort_changes(changes_pinned_ + 1, m, ncap);  // contact-id order (id in the high bits)
    last_changes_ = m;
  }
  const double t2 = now_ms();
  state_pass();
  flush_locations();  // bodies woken by contacts that began touching
  const double t3 = now_ms();

  if (d_.n_awake == 0) return;
  // an island nominated by the last step is split before the solve (solver_integration.rs:879-884)
  if (split_island_ != kNullIndex) {
    split_island(split_island_);
    split_island_ = kNullIndex;
  }
  if (options.enable_sleep) {
    ensure(d_.sleep_times, cap_sleep_times_, (size_t)d_.n_awake, false);
    sleep_times_host_.resize((size_t)d_.n_awake);
  }

  // solve lists: colour-major slots (solver_prepare.rs:335-474), staged in pinned memory
  PinnedInts& solve = solve_stage_;
  solve.clear();
  ColorRanges cr{};
  const bool any_restitution = graph_restitution_ > 0;
  int colors_used = 0;
  for (int ci = 0; ci < kOverflowColor; ++ci) {
    cr.begin[ci] = (int32_t)solve.size();
    solve.append(colors_[ci].convex);
    cr.end[ci] = (int32_t)solve.size();
    colors_used += cr.end[ci] > cr.begin[ci];
  }
  const std::vector<int32_t>& ovf = colors_[kOverflowColor].overflow;
  cr.begin[kOverflowColor] = cr.end[kOverflowColor] = (int32_t)solve.size();
  if (!solve.empty()) {
    ensure(d_.solve_ids, cap_solve_, solve.size(), false);
    // the step kernels keep records in blocks of kRecBlock slots
    ensure(d_.constraints, cap_constraints_, (solve.size() + kRecBlock - 1) / kRecBlock * kRecBlock, false);
    NSL_CHECK(hipMemcpyAsync(d_.solve_ids, solve.data(), solve.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  }
  if (!ovf.empty()) {
    ensure(d_.overflow_ids, cap_overflow_ids_, ovf.size(), false);
    ensure(d_.overflow, cap_overflow_, ovf.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_.overflow_ids, ovf.data(), ovf.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  }
  // joint lists: colour-major ids, the overflow colour's last (solver_prepare.rs:178-212)
  if (joints_dirty_) {
    joint_ids_host_.clear();
    for (int ci = 0; ci < kColorCount; ++ci) {
      joint_ranges_.begin[ci] = (int32_t)joint_ids_host_.size();
      joint_ids_host_.insert(joint_ids_host_.end(), colors_[ci].joints.begin(), colors_[ci].joints.end());
      joint_ranges_.end[ci] = (int32_t)joint_ids_host_.size();
    }
    if (!joint_ids_host_.empty()) {
      ensure(d_.joint_ids, cap_joint_ids_, joint_ids_host_.size(), false);
      NSL_CHECK(hipMemcpyAsync(d_.joint_ids, joint_ids_host_.data(), joint_ids_host_.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, stream_));
    }
    joints_dirty_ = false;
  }
  KernelTimer* tm = timing_ ? &timer_ : nullptr;
  SolvePlan plan{};
  plan.cr = cr;
  plan.jr = joint_ranges_;
  plan.n_joints = (int32_t)joint_ids_host_.size();
  plan.n_ovf_levels = build_overflow_levels();
  plan.ovf_items = d_ovf_items_;
  plan.ovf_level_begin = d_ovf_levels_;
  plan.ovf_bodies = d_ovf_bodies_;
  plan.n_ovf_bodies = (int32_t)ovf_bodies_host_.size();
  plan.n_constraints = (int32_t)solve.size();
  plan.n_overflow = (int32_t)ovf.size();
  plan.any_restitution = any_restitution ? 1 : 0;
  plan.restitution_iterations = options.restitution_iterations;
  plan.small_colour_items = options.small_colour_items;
  plan.stage_records = options.stage_records && !solve.empty() ? 1 : 0;  // the copy reads slot 0 when a lane has no item
  // warm start by body needs every awake body dynamic (a kinematic body can sit in several
  // items of one colour)
  plan.ws_table = nullptr;
  plan.ws_delta = nullptr;
  plan.ws_zero = nullptr;
  if (kinematic_awake_ == 0) {
    ensure(d_ws_table_, cap_ws_table_, (size_t)kOverflowColor * (size_t)d_.n_awake, false);
    ensure(d_ws_delta_, cap_ws_delta_, 2 * (solve.size() + hjoints_.size()) + 2, false);
    NSL_CHECK(hipMemsetAsync(d_ws_table_, 0xFF, (size_t)kOverflowColor * (size_t)d_.n_awake * sizeof(int32_t), stream_));
    plan.ws_table = d_ws_table_;
    plan.ws_delta = d_ws_delta_;
    plan.ws_zero = d_ws_delta_ + 2 * (solve.size() + hjoints_.size());
  }
  plan.ovf_done = nullptr;
  if (plan.ws_table != nullptr && plan.n_ovf_levels > 0) {
    ensure(d_ovf_done_, cap_ovf_done_, (size_t)d_.n_awake, false);
    NSL_CHECK(hipMemsetAsync(d_ovf_done_, 0xFF, (size_t)d_.n_awake * sizeof(int32_t), stream_));
    plan.ovf_done = d_ovf_done_;
  }
  plan.barrier = d_barrier_;
  plan.phase_log = phase_timing_ ? d_phase_log_ : nullptr;
  int widest = (d_.n_awake + 3) / 4;
  for (int ci = 0; ci < kOverflowColor; ++ci)
    widest = std::max(widest, (cr.end[ci] - cr.begin[ci]) + (joint_ranges_.end[ci] - joint_ranges_.begin[ci]));
  int blocks = options.solver_blocks > 0 ? options.solver_blocks : (widest + kSolveBlock - 1) / kSolveBlock;
  // more workgroups add grid-barrier arrivals faster than they add useful lanes: on the
  // settled 132k-body world 96 workgroups of 128 lanes beat 48..80 and 112..256
  blocks = std::max(1, std::min(blocks, std::min(max_blocks_, 96)));
  plan.n_work = blocks;
  // taken after every buffer above is in place: the kernels get the pointers of this step
  DeviceWorld dw = d_;
  if (!options.enable_sleep) dw.sleep_times = nullptr;
  launch_prepare(dw, plan, p, stream_, tm);
  launch_solve_step(dw, plan, p, options.plain_step_launch == 0, stream_, tm);
  if (residuals_on_) launch_residual_sweep(dw, cr, p);
  launch_store_finalize(dw, plan, p, stream_, tm);
  // continuous collision of the bodies finalize flagged fast, then of the bullets
  if (options.enable_continuous) {
    launch_continuous(dw, ccd_, p, false, d_.n_awake, stream_, tm);
    if (bullet_bodies_ > 0) launch_continuous(dw, ccd_, p, true, d_.n_awake, stream_, tm);
  }
  device_newer_ = true;
  if (options.enable_sleep)
    NSL_CHECK(hipMemcpyAsync(sleep_times_host_.data(), d_.sleep_times, sleep_times_host_.size() * sizeof(float),
                             hipMemcpyDeviceToHost, stream_));

  // counters[6] (proxies moved by finalize) tells the next pair update whether to run
  NSL_CHECK(hipMemcpyAsync(readback_, d_.counters, sizeof(StepReadback), hipMemcpyDeviceToHost, stream_));
  if (phase_timing_)
    NSL_CHECK(hipMemcpyAsync(phase_log_host_.data(), d_phase_log_, phase_log_host_.size() * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, stream_));
  NSL_CHECK(hipStreamSynchronize(stream_));
  std::memcpy(counters_host_, readback_->counters, sizeof(counters_host_));
  barrier_host_ = readback_->barrier;
  if (residuals_on_) collect_residuals(cr);
  if (barrier_host_.timed_out) {
    // the step kernel stopped part-way and store/finalize ran on its partial results: every
    // later call on this world fails until a snapshot is restored
    barrier_host_ = GridBarrier{};
    NSL_CHECK(hipMemset(d_barrier_, 0, sizeof(GridBarrier)));
    poisoned_ = true;
    throw std::runtime_error("step kernel: grid barrier wait exceeded its limit (workgroups not co-resident?)");
  }
  // solver_integration.rs:1003-1005
  if (options.enable_sleep) {
    mark_awake_islands();
    process_sleep();
  }
  if (timing_) {
    for (int i = 0; i < timer_.used; ++i) {
      float ms = 0.0f;
      NSL_CHECK(hipEventElapsedTime(&ms, timer_.events[2 * i], timer_.events[2 * i + 1]));
      kstats_.ms[timer_.classes[i]] += ms;
      kstats_.launches[timer_.classes[i]] += 1;
    }
    timer_.used = 0;
  }
  if (phase_timing_) {
    // marks: (class of the phase that just ended, wall clock); the first is the kernel start
    int marks = 0;
    for (int i = 0; i < kMaxPhaseLog; ++i, ++marks)
      if (i > 0 && phase_log_host_[2 * i + 1] < phase_log_host_[2 * (i - 1) + 1]) break;
    for (int i = 1; i < marks; ++i) {
      const int cls = (int)phase_log_host_[2 * i];
      if (cls < 0 || cls >= kPhaseClassCount) continue;
      kstats_.phase_ms[cls] += (double)(phase_log_host_[2 * i + 1] - phase_log_host_[2 * (i - 1) + 1]) / wall_clock_khz_;
      kstats_.phase_count[cls] += 1;
    }
  }
  kstats_.steps += 1;
  kstats_.bodies += (int64_t)d_.n_awake;
  kstats_.collided += (int64_t)awake_contacts;
  kstats_.constraints += (int64_t)solve.size();
  kstats_.joints += (int64_t)joint_ids_host_.size();
  kstats_.fast_bodies = counters_host_[0] - counters_base_[0];
  kstats_.sat_calls = counters_host_[1] - counters_base_[1];
  kstats_.sat_hits = counters_host_[2] - counters_base_[2];
  kstats_.recycled = counters_host_[3] - counters_base_[3];
  kstats_.points = counters_host_[4] - counters_base_[4];
  kstats_.overflow_points = counters_host_[5] - counters_base_[5];
  kstats_.toi_calls = counters_host_[kCntToiCalls] - counters_base_[kCntToiCalls];
  kstats_.toi_hits = counters_host_[kCntToiHits] - counters_base_[kCntToiHits];
  const double t4 = now_ms();
  kstats_.host_ms[0] += t1 - t0;
  kstats_.host_ms[1] += t2 - t1;
  kstats_.host_ms[2] += t3 - t2;
  kstats_.host_ms[3] += t4 - t3;
  kstats_.solver_blocks = blocks;
  {
    // grid barriers of one stage: one per colour, or per run of two or more small colours
    // (run_colours in nsl_kernels.hip)
    int phases = 0, run = 0;
    for (int ci = 0; ci < kOverflowColor; ++ci) {
      const int items = (cr.end[ci] - cr.begin[ci]) + (joint_ranges_.end[ci] - joint_ranges_.begin[ci]);
      if (items == 0) continue;
      if (items <= options.small_colour_items) {
        run += 1;
      } else {
        phases += (run >= 2 ? 1 : run) + 1;
        run = 0;
      }
    }
    phases += run >= 2 ? 1 : run;
    const int ovf_phase = plan.n_ovf_levels > 0 ? 1 : 0;
    // warm start by body: the delta phase (with the overflow levels and the velocity
    // integration) and the body phase
    const int iv_ws_phases = plan.ws_table != nullptr ? 2 : 1 + phases + ovf_phase;
    // (with warm start by body the overflow push and relax levels end the phases before them)
    const int ovf_solve_phases = plan.ws_table != nullptr ? 0 : ovf_phase;
    int64_t barriers = (int64_t)p.sub_steps * (1 + iv_ws_phases + 2 * (phases + ovf_solve_phases));
    if (any_restitution) barriers += (int64_t)options.restitution_iterations * (phases + (ovf.empty() ? 0 : 1));
    kstats_.barriers += barriers;
  }
  profile_.pairs = (float)(t1 - t0);
  profile_.collide = (float)(t2 - t1);
  profile_.state_pass = (float)(t3 - t2);
  profile_.solve = (float)(t4 - t3);
  profile_.step = (float)(t4 - t0);
  profile_.contacts = contact_count();
  profile_.touching = (int32_t)solve.size() + (int32_t)ovf.size();
  profile_.colors_used = colors_used;
  profile_.overflow = (int32_t)ovf.size();
  profile_.fast_bodies = counters_host_[0];
  profile_.sat_calls = counters_host_[1];
  profile_.sat_hits = counters_host_[2];
  profile_.recycled = counters_host_[3];
}

void GpuWorld::PinnedInts::grow(size_t want) {
  const size_t c = std::max<size_t>(2 * want, 4096);
  int32_t* p = nullptr;
  NSL_CHECK(hipHostMalloc((void**)&p, c * sizeof(int32_t), hipHostMallocDefault));
  if (n) std::memcpy(p, ptr, n * sizeof(int32_t));
  if (ptr) NSL_CHECK(hipHostFree(ptr));
  ptr = p;
  cap = c;
}

void GpuWorld::ensure_pinned_changes(size_t n) {
  if (n <= cap_changes_pinned_) return;
  if (changes_pinned_) {
    NSL_CHECK(hipStreamSynchronize(stream_));
    NSL_CHECK(hipHostFree(changes_pinned_));
  }
  cap_changes_pinned_ = std::max<size_t>(2 * n, 4096);
  NSL_CHECK(hipHostMalloc((void**)&changes_pinned_, cap_changes_pinned_ * sizeof(int32_t), hipHostMallocDefault));
}

void GpuWorld::launch_residual_sweep(const DeviceWorld& dw, const ColorRanges& cr, const StepParams& p) {
  int waves = 0;
  for (int ci = 0; ci < kOverflowColor; ++ci) {
    res_wave_begin_[ci] = waves;
    waves += (std::max(cr.end[ci] - cr.begin[ci], 0) + 63) / 64;
  }
  res_wave_begin_[kOverflowColor] = waves;
  if (waves == 0) return;
  ensure(d_res_states_, cap_res_states_, (size_t)std::max(d_.n_awake, 1), false);
  ensure(d_res_partials_, cap_res_partials_, (size_t)waves, false);
  if ((size_t)waves > cap_res_partials_host_) {
    if (res_partials_host_) {
      NSL_CHECK(hipStreamSynchronize(stream_));
      NSL_CHECK(hipHostFree(res_partials_host_));
    }
    cap_res_partials_host_ = std::max<size_t>(2 * (size_t)waves, 256);
    NSL_CHECK(hipHostMalloc((void**)&res_partials_host_, cap_res_partials_host_ * sizeof(ResidualPartial),
                            hipHostMallocDefault));
  }
  NSL_CHECK(hipMemcpyAsync(d_res_states_, d_.states, (size_t)d_.n_awake * sizeof(BodyState), hipMemcpyDeviceToDevice,
                           stream_));
  DeviceWorld copy = dw;
  copy.states = d_res_states_;
  for (int ci = 0; ci < kOverflowColor; ++ci)
    launch_colour_residual(copy, cr.begin[ci], cr.end[ci] - cr.begin[ci], p, d_res_partials_ + res_wave_begin_[ci],
                           stream_);
  NSL_CHECK(hipGetLastError());
  NSL_CHECK(hipMemcpyAsync(res_partials_host_, d_res_partials_, (size_t)waves * sizeof(ResidualPartial),
                           hipMemcpyDeviceToHost, stream_));
}

void GpuWorld::collect_residuals(const ColorRanges& cr) {
  for (int ci = 0; ci < kOverflowColor; ++ci) {
    ColourResidual r;
    r.constraints = std::max(cr.end[ci] - cr.begin[ci], 0);
    for (int w = res_wave_begin_[ci]; w < res_wave_begin_[ci + 1]; ++w) {
      const ResidualPartial& q = res_partials_host_[w];
      r.sum_normal += (double)q.sum_normal;
      r.max_normal = std::max(r.max_normal, q.max_normal);
      r.max_friction = std::max(r.max_friction, q.max_friction);
    }
    residuals_[ci] = r;
  }
}

// The collide kernel's change list in contact-id order. Each contact id appears at most once
// and lies below `ncap`: long lists go through a bitmap of ids scanned in order (linear in
// ncap / 64 + m), short ones through a comparison sort.
void GpuWorld::sort_changes(const int32_t* list, size_t m, size_t ncap) {
  changes_.assign(list, list + m);
  if (m < 4096) {
    std::sort(changes_.begin(), changes_.end());
    return;
  }
  const size_t words = (ncap + 63) / 64;
  change_bits_.assign(words, 0ull);
  if (change_code_.size() < ncap) change_code_.resize(ncap);
  for (const int32_t v : changes_) {
    const uint32_t id = (uint32_t)v >> 2;
    change_bits_[id >> 6] |= 1ull << (id & 63);
    change_code_[id] = (uint8_t)(v & 3);
  }
  size_t k = 0;
  for (size_t w = 0; w < words; ++w)
    for (uint64_t bits = change_bits_[w]; bits != 0; bits &= bits - 1) {
      const uint32_t id = (uint32_t)(w * 64 + (size_t)__builtin_ctzll(bits));
      changes_[k++] = (int32_t)(id << 2 | change_code_[id]);
    }
}

void GpuWorld::set_kernel_timing(bool on, bool phases) {
  timing_ = on;
  phase_timing_ = on && phases;
  timer_.used = 0;
  if (phase_timing_) {  // no marks of an earlier timed step may be read as this step's
    NSL_CHECK(hipStreamSynchronize(stream_));
    NSL_CHECK(hipMemset(d_phase_log_, 0, 2 * kMaxPhaseLog * sizeof(uint64_t)));
  }
}

void GpuWorld::reset_kernel_stats() {
  kstats_ = KernelStats{};
  NSL_CHECK(hipStreamSynchronize(stream_));
  NSL_CHECK(hipMemcpy(counters_base_, d_.counters, sizeof(counters_base_), hipMemcpyDeviceToHost));
}

void GpuWorld::get_bodies(double* pos, float* rot, float* v, float* w) {
  check_usable();
  if (device_newer_) pull_bodies();
  HostScene::get_bodies(pos, rot, v, w);
}

uint64_t GpuWorld::state_hash() {
  check_usable();
  if (device_newer_) pull_bodies();
  return HostScene::state_hash();
}

void GpuWorld::get_contact_colors(int32_t* color, int32_t* local, int32_t* alive) const {
  for (size_t i = 0; i < hcontacts_.size(); ++i) {
    color[i] = hcontacts_[i].color;
    local[i] = hcontacts_[i].local;
    alive[i] = hcontacts_[i].alive;
  }
}

void GpuWorld::get_contact_shapes(int32_t* shape_a, int32_t* shape_b) const {
  for (size_t i = 0; i < hcontacts_.size(); ++i) {
    shape_a[i] = hcontacts_[i].alive ? hcontacts_[i].shape_a : kNullIndex;
    shape_b[i] = hcontacts_[i].alive ? hcontacts_[i].shape_b : kNullIndex;
  }
}

}  // namespace nsl
