// Host bookkeeping of islands, island splitting and the sleeping and waking of solver sets
// for the GPU stepper: island.rs:25-300 (merge, link and unlink of contacts and joints),
// sleep_island.rs:20-633 (union-find split, try_sleep_island, process_sleep),
// solver_set_ops.rs:21-160 (wake_solver_set) and the island pool helpers of
// world_runtime.rs:1419-1475, 1634-1644. Set ids: 0 static, 1 disabled, 2 awake, 3 and up
// sleeping.
//
// The device holds the awake bodies (by awake index), the sleeping bodies' sims (by body
// id) and, in every shape, awake-set contact and graph joint, the encoding of its bodies
// (awake index, static -(index + 2), sleeping -1). Host changes of body locations are
// recorded per body and applied to the device by flush_locations before the next kernel
// that reads them.
#include "nsl_world.h"

#include <algorithm>
#include <stdexcept>

namespace nsl {

#define NSL_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int32_t kDisabledSet = 1, kAwakeSet = 2, kFirstSleepingSet = 3;
constexpr float kTimeToSleep = 0.5f;  // constants.rs:31

// Removes v[index] by moving the last element into its place; returns the old index of the
// moved element, or kNullIndex when the removed element was the last.
template <class T>
int32_t swap_out(std::vector<T>& v, size_t index) {
  const size_t last = v.size() - 1;
  if (index != last) v[index] = v[last];
  v.pop_back();
  return index != last ? (int32_t)last : kNullIndex;
}

// union find with path halving and union by rank (sleep_island.rs:20-58); a union adds
// the child's constraint counts to the root.
struct Components {
  std::vector<int32_t> parent, rank, contacts, joints;
  explicit Components(size_t n) : parent(n), rank(n, 0), contacts(n, 0), joints(n, 0) {
    for (size_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
  }
  int32_t root(int32_t x) {
    while (parent[(size_t)x] != x) {
      parent[(size_t)x] = parent[(size_t)parent[(size_t)x]];
      x = parent[(size_t)x];
    }
    return x;
  }
  void join(int32_t x, int32_t y) {
    int32_t rx = root(x), ry = root(y);
    if (rx == ry) return;
    if (rank[(size_t)rx] < rank[(size_t)ry]) std::swap(rx, ry);
    else if (rank[(size_t)rx] == rank[(size_t)ry]) rank[(size_t)rx] += 1;
    parent[(size_t)ry] = rx;
    contacts[(size_t)rx] += contacts[(size_t)ry];
    joints[(size_t)rx] += joints[(size_t)ry];
  }
};

}  // namespace

std::vector<int32_t>& GpuWorld::island_list(int32_t set) {
  return set == kAwakeSet ? awake_islands_ : sleep_sets_[(size_t)set].islands;
}

std::vector<int32_t>& GpuWorld::contact_list(int32_t set) {
  if (set == kAwakeSet) return non_touching_;
  if (set == kDisabledSet) return disabled_;
  return sleep_sets_[(size_t)set].contacts;
}

// world_runtime.rs:1419-1445
int32_t GpuWorld::create_island(int32_t set) {
  const int32_t id = island_pool_.alloc();
  if ((size_t)id >= islands_.size()) islands_.resize((size_t)id + 1);
  std::vector<int32_t>& list = island_list(set);
  Island is;
  is.id = id;
  is.set = set;
  is.local = (int32_t)list.size();
  islands_[(size_t)id] = std::move(is);
  list.push_back(id);
  return id;
}

// world_runtime.rs:1447-1475
void GpuWorld::destroy_island(int32_t id) {
  if (split_island_ == id) split_island_ = kNullIndex;
  Island& is = islands_[(size_t)id];
  std::vector<int32_t>& list = island_list(is.set);
  const int32_t local = is.local;
  if (swap_out(list, (size_t)local) != kNullIndex) islands_[(size_t)list[(size_t)local]].local = local;
  islands_[(size_t)id] = Island{};
  island_pool_.release(id);
}

// island.rs:32-106: the island with more bodies (a on a tie) takes the other's bodies,
// contacts and joints in their order and its count of removed constraints.
int32_t GpuWorld::merge_islands(int32_t a, int32_t b) {
  if (a == b || b == kNullIndex) return a;
  if (a == kNullIndex) return b;
  const int32_t keep = islands_[(size_t)a].bodies.size() >= islands_[(size_t)b].bodies.size() ? a : b;
  const int32_t gone = keep == a ? b : a;
  Island& dst = islands_[(size_t)keep];
  Island& src = islands_[(size_t)gone];
  for (int32_t body : src.bodies) {
    bodies_[(size_t)body].island = keep;
    bodies_[(size_t)body].island_index = (int32_t)dst.bodies.size();
    dst.bodies.push_back(body);
  }
  for (const ContactLink& l : src.contacts) {
    hcontacts_[(size_t)l.contact].island = keep;
    hcontacts_[(size_t)l.contact].island_index = (int32_t)dst.contacts.size();
    dst.contacts.push_back(l);
  }
  for (const JointLink& l : src.joints) {
    hjoints_[(size_t)l.joint].island = keep;
    hjoints_[(size_t)l.joint].island_index = (int32_t)dst.joints.size();
    dst.joints.push_back(l);
  }
  dst.remove_count += src.remove_count;
  destroy_island(gone);
  return keep;
}

// island.rs:126-171: a contact that starts touching wakes a sleeping body touched by an
// awake one, then joins the two islands.
void GpuWorld::link_contact(int32_t id) {
  const int32_t a = hcontacts_[(size_t)id].body[0], b = hcontacts_[(size_t)id].body[1];
  const int32_t set_a = bodies_[(size_t)a].set, set_b = bodies_[(size_t)b].set;
  if (set_a == kAwakeSet && set_b >= kFirstSleepingSet) wake_solver_set(set_b);
  if (set_b == kAwakeSet && set_a >= kFirstSleepingSet) wake_solver_set(set_a);
  const int32_t island = merge_islands(bodies_[(size_t)a].island, bodies_[(size_t)b].island);
  HContact& c = hcontacts_[(size_t)id];
  c.island = island;
  c.island_index = (int32_t)islands_[(size_t)island].contacts.size();
  islands_[(size_t)island].contacts.push_back(ContactLink{id, a, b});
}

// island.rs:174-206
void GpuWorld::unlink_contact(int32_t id) {
  HContact& c = hcontacts_[(size_t)id];
  Island& is = islands_[(size_t)c.island];
  const int32_t index = c.island_index;
  if (swap_out(is.contacts, (size_t)index) != kNullIndex) hcontacts_[(size_t)is.contacts[(size_t)index].contact].island_index = index;
  is.remove_count += 1;
  c.island = kNullIndex;
  c.island_index = kNullIndex;
}

// island.rs:226-261
void GpuWorld::link_joint(int32_t id) {
  const int32_t a = hjoints_[(size_t)id].body_a, b = hjoints_[(size_t)id].body_b;
  const int32_t set_a = bodies_[(size_t)a].set, set_b = bodies_[(size_t)b].set;
  if (set_a == kAwakeSet && set_b >= kFirstSleepingSet) wake_solver_set(set_b);
  else if (set_b == kAwakeSet && set_a >= kFirstSleepingSet) wake_solver_set(set_a);
  const int32_t island = merge_islands(bodies_[(size_t)a].island, bodies_[(size_t)b].island);
  HJoint& j = hjoints_[(size_t)id];
  j.island = island;
  j.island_index = (int32_t)islands_[(size_t)island].joints.size();
  islands_[(size_t)island].joints.push_back(JointLink{id, a, b});
}

// contact_creation.rs:357-384: the contact becomes the head of both bodies' contact lists.
void GpuWorld::link_body_contact(int32_t id) {
  HContact& c = hcontacts_[(size_t)id];
  for (int e = 0; e < 2; ++e) {
    HBody& body = bodies_[(size_t)c.body[e]];
    const int32_t key = (id << 1) | e;
    c.prev[e] = kNullIndex;
    c.next[e] = body.head_contact;
    if (body.head_contact != kNullIndex) hcontacts_[(size_t)(body.head_contact >> 1)].prev[body.head_contact & 1] = key;
    body.head_contact = key;
  }
}

// contact_creation.rs:456-494
void GpuWorld::unlink_body_contact(int32_t id) {
  HContact& c = hcontacts_[(size_t)id];
  for (int e = 0; e < 2; ++e) {
    if (c.prev[e] != kNullIndex) hcontacts_[(size_t)(c.prev[e] >> 1)].next[c.prev[e] & 1] = c.next[e];
    if (c.next[e] != kNullIndex) hcontacts_[(size_t)(c.next[e] >> 1)].prev[c.next[e] & 1] = c.prev[e];
    HBody& body = bodies_[(size_t)c.body[e]];
    if (body.head_contact == ((id << 1) | e)) body.head_contact = c.next[e];
  }
}

void GpuWorld::mark_body_moved(int32_t body) {
  if (body_moved_.size() < bodies_.size()) body_moved_.resize(bodies_.size(), 0);
  if (!body_moved_[(size_t)body]) {
    body_moved_[(size_t)body] = 1;
    moved_bodies_.push_back(body);
  }
  locations_dirty_ = true;
}

// sleep_island.rs:85-300: with more than one connected component the island is replaced by
// new awake islands in component order, which take the bodies (in island order), then the
// contacts and joints (in link order).
void GpuWorld::split_island(int32_t base_id) {
  const Island base = islands_[(size_t)base_id];
  const size_t n = base.bodies.size();
  Components cc(n);
  auto connect = [&](int32_t a, int32_t b, bool joint) {
    const int32_t ka = bodies_[(size_t)a].island_index, kb = bodies_[(size_t)b].island_index;
    const bool va = ka >= 0 && (size_t)ka < n, vb = kb >= 0 && (size_t)kb < n;
    if (!va && !vb) return;
    if (va && vb) cc.join(ka, kb);
    const int32_t r = cc.root(va ? ka : kb);
    (joint ? cc.joints : cc.contacts)[(size_t)r] += 1;
  };
  for (const ContactLink& l : base.contacts) connect(l.body_a, l.body_b, false);
  for (const JointLink& l : base.joints) connect(l.body_a, l.body_b, true);
  size_t components = 0;
  for (size_t i = 0; i < n; ++i) {
    cc.parent[i] = cc.root((int32_t)i);
    components += cc.parent[i] == (int32_t)i;
  }
  if (components == 1) {
    islands_[(size_t)base_id].remove_count = 0;
    return;
  }
  std::vector<int32_t> slot(n, kNullIndex);
  std::vector<int32_t> ids;
  for (size_t i = 0; i < n; ++i) {
    const int32_t r = cc.parent[i];
    if (slot[(size_t)r] == kNullIndex) {
      slot[(size_t)r] = (int32_t)ids.size();
      ids.push_back(kNullIndex);
    }
  }
  islands_[(size_t)base_id].bodies.clear();
  islands_[(size_t)base_id].contacts.clear();
  islands_[(size_t)base_id].joints.clear();
  for (int32_t& id : ids) id = create_island(kAwakeSet);
  for (size_t i = 0; i < n; ++i) {
    const int32_t body = base.bodies[i];
    const int32_t island = ids[(size_t)slot[(size_t)cc.root((int32_t)i)]];
    bodies_[(size_t)body].island = island;
    bodies_[(size_t)body].island_index = (int32_t)islands_[(size_t)island].bodies.size();
    islands_[(size_t)island].bodies.push_back(body);
  }
  for (const ContactLink& l : base.contacts) {
    const int32_t ia = bodies_[(size_t)l.body_a].island;
    const int32_t island = ia != kNullIndex ? ia : bodies_[(size_t)l.body_b].island;
    hcontacts_[(size_t)l.contact].island = island;
    hcontacts_[(size_t)l.contact].island_index = (int32_t)islands_[(size_t)island].contacts.size();
    islands_[(size_t)island].contacts.push_back(l);
  }
  for (const JointLink& l : base.joints) {
    const int32_t ia = bodies_[(size_t)l.body_a].island;
    const int32_t island = ia != kNullIndex ? ia : bodies_[(size_t)l.body_b].island;
    hjoints_[(size_t)l.joint].island = island;
    hjoints_[(size_t)l.joint].island_index = (int32_t)islands_[(size_t)island].joints.size();
    islands_[(size_t)island].joints.push_back(l);
  }
  destroy_island(base_id);
}

// The island part of finalize (solver_integration.rs:735-752, merged over threads at
// :522-538) from the sleep times finalize left on the device: a body short of the time to
// sleep keeps its island awake; a sleepy body in an island with removed constraints
// nominates it for splitting, the sleepiest (then the largest island id) winning.
void GpuWorld::mark_awake_islands() {
  island_awake_bits_.assign(awake_islands_.size(), 0);
  split_candidate_ = kNullIndex;
  split_candidate_time_ = 0.0f;
  for (size_t i = 0; i < awake_host_.size(); ++i) {
    const HBody& body = bodies_[(size_t)awake_host_[i].body_id];
    const Island& island = islands_[(size_t)body.island];
    const float t = sleep_times_host_[i];
    if (t < kTimeToSleep) {
      island_awake_bits_[(size_t)island.local] = 1;
    } else if (island.remove_count > 0) {
      if (t > split_candidate_time_ || (t == split_candidate_time_ && body.island > split_candidate_)) {
        split_candidate_ = body.island;
        split_candidate_time_ = t;
      }
    }
  }
}

// sleep_island.rs:563-633: the split candidate is split at the next step; every awake
// island no body kept awake tries to sleep, the last island first.
void GpuWorld::process_sleep() {
  if (split_candidate_ != kNullIndex) split_island_ = split_candidate_;
  const std::vector<int32_t> order = awake_islands_;
  const std::vector<uint8_t> bits = island_awake_bits_;
  for (size_t k = order.size(); k-- > 0;) {
    if (k < bits.size() && bits[k]) continue;
    try_sleep_island(order[k]);
  }
}

// sleep_island.rs:300-560: the island's bodies leave the awake set by swap removal, their
// non-touching contacts with no awake partner go to the disabled set, their touching
// contacts and joints leave the graph for the new sleeping set, and the island moves with
// them.
void GpuWorld::try_sleep_island(int32_t island_id) {
  Island island = islands_[(size_t)island_id];
  if (island.remove_count > 0 && island.bodies.size() > 1) return;
  if (device_newer_) pull_bodies();
  const int32_t set_id = set_pool_.alloc();
  if ((size_t)set_id >= sleep_sets_.size()) sleep_sets_.resize((size_t)set_id + 1);
  if (sleep_host_.size() < bodies_.size()) sleep_host_.resize(bodies_.size());
  SleepSet ss;
  for (int32_t body_id : island.bodies) {
    HBody& body = bodies_[(size_t)body_id];
    const int32_t awake_index = body.local;
    sleep_host_[(size_t)body_id] = awake_host_[(size_t)awake_index];
    ss.bodies.push_back(body_id);
    if (swap_out(awake_host_, (size_t)awake_index) != kNullIndex) {
      const int32_t moved = awake_host_[(size_t)awake_index].body_id;
      bodies_[(size_t)moved].local = awake_index;
      mark_body_moved(moved);
    }
    swap_out(states_host_, (size_t)awake_index);
    body.set = set_id;
    body.local = (int32_t)ss.bodies.size() - 1;
    mark_body_moved(body_id);
    slept_bodies_.push_back(body_id);
    for (int32_t key = body.head_contact; key != kNullIndex;) {
      const int32_t cid = key >> 1, edge = key & 1;
      HContact& c = hcontacts_[(size_t)cid];
      key = c.next[edge];
      if (c.color != kNullIndex) continue;  // touching: leaves with the island below
      if (bodies_[(size_t)c.body[edge ^ 1]].set == kAwakeSet) continue;
      const int32_t local = c.local;
      c.set = kDisabledSet;
      c.local = (int32_t)disabled_.size();
      disabled_.push_back(cid);
      if (swap_out(non_touching_, (size_t)local) != kNullIndex) hcontacts_[(size_t)non_touching_[(size_t)local]].local = local;
    }
  }
  for (const ContactLink& link : island.contacts) {
    HContact& c = hcontacts_[(size_t)link.contact];
    const int32_t color = c.color, local = c.local;
    ss.contacts.push_back(link.contact);
    remove_from_graph(c.body[0], c.body[1], color, local);
    HContact& c2 = hcontacts_[(size_t)link.contact];
    c2.set = set_id;
    c2.color = kNullIndex;
    c2.local = (int32_t)ss.contacts.size() - 1;
  }
  for (const JointLink& link : island.joints) {
    remove_joint_from_graph(link.joint);
    ss.joints.push_back(link.joint);
    HJoint& j = hjoints_[(size_t)link.joint];
    j.set = set_id;
    j.local = (int32_t)ss.joints.size() - 1;
  }
  ss.islands.push_back(island_id);
  if (swap_out(awake_islands_, (size_t)island.local) != kNullIndex)
    islands_[(size_t)awake_islands_[(size_t)island.local]].local = island.local;
  island.set = set_id;
  island.local = 0;
  islands_[(size_t)island_id] = std::move(island);
  if (split_island_ == island_id) split_island_ = kNullIndex;
  sleep_sets_[(size_t)set_id] = std::move(ss);
  bodies_dirty_ = true;
}

// constraint_graph.rs:326-331 (colour rules of constraint_graph.rs:256-298)
void GpuWorld::add_joint_to_graph(int32_t id) {
  HJoint& j = hjoints_[(size_t)id];
  j.color = assign_joint_color(j.body_a, j.body_b);
  j.local = (int32_t)colors_[j.color].joints.size();
  colors_[j.color].joints.push_back(id);
  joints_dirty_ = true;
}

// constraint_graph.rs:334-366
void GpuWorld::remove_joint_from_graph(int32_t id) {
  HJoint& j = hjoints_[(size_t)id];
  const int32_t color = j.color, local = j.local;
  if (color != kOverflowColor) {
    colour_clear(color, j.body_a);
    colour_clear(color, j.body_b);
  }
  std::vector<int32_t>& list = colors_[color].joints;
  if (swap_out(list, (size_t)local) != kNullIndex) hjoints_[(size_t)list[(size_t)local]].local = local;
  j.color = kNullIndex;
  j.local = kNullIndex;
  joints_dirty_ = true;
}

// solver_set_ops.rs:21-160: the bodies return to the awake set at rest with their sleep
// timers reset, their disabled contacts return to the awake list, the touching contacts
// and joints re-enter the graph in sleeping-set order and the islands rejoin the awake set.
void GpuWorld::wake_solver_set(int32_t set) {
  if (device_newer_) pull_bodies();
  SleepSet ss = std::move(sleep_sets_[(size_t)set]);
  sleep_sets_[(size_t)set] = SleepSet{};
  for (int32_t body_id : ss.bodies) {
    HBody& body = bodies_[(size_t)body_id];
    body.set = kAwakeSet;
    body.local = (int32_t)awake_host_.size();
    BodySim sim = sleep_host_[(size_t)body_id];
    sim.sleep_time = 0.0f;
    awake_host_.push_back(sim);
    BodyState st{};
    st.dq = quat_identity();
    st.flags = body.flags;
    states_host_.push_back(st);
    mark_body_moved(body_id);
    for (int32_t key = body.head_contact; key != kNullIndex;) {
      const int32_t cid = key >> 1, edge = key & 1;
      HContact& c = hcontacts_[(size_t)cid];
      key = c.next[edge];
      if (c.set != kDisabledSet) continue;
      const int32_t local = c.local;
      c.set = kAwakeSet;
      c.local = (int32_t)non_touching_.size();
      non_touching_.push_back(cid);
      if (swap_out(disabled_, (size_t)local) != kNullIndex) hcontacts_[(size_t)disabled_[(size_t)local]].local = local;
    }
  }
  for (int32_t cid : ss.contacts) {
    add_to_graph(cid);
    hcontacts_[(size_t)cid].set = kAwakeSet;
  }
  for (int32_t jid : ss.joints) {
    add_joint_to_graph(jid);
    hjoints_[(size_t)jid].set = kAwakeSet;
  }
  for (int32_t id : ss.islands) {
    islands_[(size_t)id].set = kAwakeSet;
    islands_[(size_t)id].local = (int32_t)awake_islands_.size();
    awake_islands_.push_back(id);
  }
  set_pool_.release(set);
  bodies_dirty_ = true;
}

void GpuWorld::mark_contact_dirty(int32_t id) {
  dirty_contacts_.push_back(id);
  locations_dirty_ = true;
}

void GpuWorld::flush_locations() {
  if (bodies_dirty_ || shapes_dirty_) upload_bodies();
  if (!locations_dirty_) return;
  std::vector<int32_t> shape_items, contact_items, joint_items;
  if (contact_mark_.size() < hcontacts_.size()) contact_mark_.resize(hcontacts_.size(), 0);
  std::vector<int32_t> marked;
  // a contact's encodings and awake-set flag, from its current record
  auto patch_contact = [&](int32_t cid) {
    if (contact_mark_[(size_t)cid]) return;
    contact_mark_[(size_t)cid] = 1;
    marked.push_back(cid);
    const HContact& c = hcontacts_[(size_t)cid];
    contact_items.push_back(cid);
    if (c.alive) {
      contact_items.push_back(encode(bodies_[(size_t)c.body[0]]));
      contact_items.push_back(encode(bodies_[(size_t)c.body[1]]));
      contact_items.push_back(c.set == kAwakeSet ? 1 : 0);
    } else {
      contact_items.push_back(kNullIndex);
      contact_items.push_back(kNullIndex);
      contact_items.push_back(0);
    }
  };
  for (int32_t cid : dirty_contacts_) patch_contact(cid);
  dirty_contacts_.clear();
  for (int32_t b : moved_bodies_) {
    const HBody& body = bodies_[(size_t)b];
    const int32_t enc = encode(body);
    for (int32_t sid = body.head_shape; sid != kNullIndex; sid = shapes_[(size_t)sid].next_shape) {
      shapes_[(size_t)sid].body_enc = enc;
      shape_items.push_back(sid);
      shape_items.push_back(enc);
    }
    for (int32_t key = body.head_contact; key != kNullIndex;) {
      const int32_t cid = key >> 1;
      key = hcontacts_[(size_t)cid].next[key & 1];
      patch_contact(cid);  // every set: a contact leaving or joining the awake set changes its flag
    }
    body_moved_[(size_t)b] = 0;
  }
  for (int32_t cid : marked) contact_mark_[(size_t)cid] = 0;
  for (size_t id = 0; id < hjoints_.size(); ++id) {
    HJoint& j = hjoints_[id];
    if (j.set != kAwakeSet) continue;
    const int32_t ea = encode(bodies_[(size_t)j.body_a]), eb = encode(bodies_[(size_t)j.body_b]);
    if (ea == j.enc_a && eb == j.enc_b) continue;
    j.enc_a = ea;
    j.enc_b = eb;
    joint_items.push_back((int32_t)id);
    joint_items.push_back(ea);
    joint_items.push_back(eb);
  }
  std::vector<BodySim> sims;
  for (int32_t b : slept_bodies_)
    if (bodies_[(size_t)b].set >= kFirstSleepingSet) sims.push_back(sleep_host_[(size_t)b]);
  moved_bodies_.clear();
  slept_bodies_.clear();
  locations_dirty_ = false;

  std::vector<int32_t> items;
  items.reserve(shape_items.size() + contact_items.size() + joint_items.size());
  items.insert(items.end(), shape_items.begin(), shape_items.end());
  items.insert(items.end(), contact_items.begin(), contact_items.end());
  items.insert(items.end(), joint_items.begin(), joint_items.end());
  if (items.empty() && sims.empty()) return;
  if (!sims.empty()) {
    ensure(d_.sleep_sims, cap_sleep_sims_, bodies_.size(), true);
    ensure(d_patch_sims_, cap_patch_sims_, sims.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_patch_sims_, sims.data(), sims.size() * sizeof(BodySim), hipMemcpyHostToDevice, stream_));
  }
  if (!items.empty()) {
    ensure(d_patch_, cap_patch_, items.size(), false);
    NSL_CHECK(hipMemcpyAsync(d_patch_, items.data(), items.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  }
  launch_patch_locations(d_, d_patch_, (int)(shape_items.size() / 2), (int)(contact_items.size() / 4),
                         (int)(joint_items.size() / 3), d_patch_sims_, (int)sims.size(), stream_);
  NSL_CHECK(hipStreamSynchronize(stream_));  // the staging vectors are read by the copies above
}

void GpuWorld::body_location(int32_t id, int32_t* set, int32_t* local, int32_t* island) const {
  if (id < 0 || (size_t)id >= bodies_.size()) throw std::invalid_argument("no such body");
  const HBody& b = bodies_[(size_t)id];
  *set = b.set;
  *local = b.local;
  *island = b.island;
}

void GpuWorld::island_summary(int64_t out[5]) const {
  int64_t live = 0, sets = 0;
  for (const Island& is : islands_) live += is.id != kNullIndex;
  for (size_t s = kFirstSleepingSet; s < sleep_sets_.size(); ++s) sets += !sleep_sets_[s].bodies.empty();
  out[0] = (int64_t)islands_.size();
  out[1] = live;
  out[2] = (int64_t)awake_islands_.size();
  out[3] = sets;
  out[4] = split_island_;
}

int32_t GpuWorld::check_islands() const {
  auto list_of = [&](int32_t set) -> const std::vector<int32_t>& {
    return set == kAwakeSet ? awake_islands_ : sleep_sets_[(size_t)set].islands;
  };
  for (size_t id = 0; id < islands_.size(); ++id) {
    const Island& is = islands_[id];
    if (is.id == kNullIndex) continue;
    if (is.id != (int32_t)id) return 1;
    const std::vector<int32_t>& list = list_of(is.set);
    if (is.local < 0 || (size_t)is.local >= list.size() || list[(size_t)is.local] != (int32_t)id) return 2;
    for (size_t k = 0; k < is.bodies.size(); ++k) {
      const HBody& b = bodies_[(size_t)is.bodies[k]];
      if (b.island != (int32_t)id || b.island_index != (int32_t)k) return 3;
      if (b.set != is.set) return 4;
    }
    for (size_t k = 0; k < is.contacts.size(); ++k) {
      const HContact& c = hcontacts_[(size_t)is.contacts[k].contact];
      if (c.island != (int32_t)id || c.island_index != (int32_t)k) return 5;
    }
    for (size_t k = 0; k < is.joints.size(); ++k) {
      const HJoint& j = hjoints_[(size_t)is.joints[k].joint];
      if (j.island != (int32_t)id || j.island_index != (int32_t)k) return 6;
    }
  }
  for (size_t i = 0; i < bodies_.size(); ++i) {
    const HBody& b = bodies_[i];
    if (b.set == kAwakeSet && awake_host_[(size_t)b.local].body_id != (int32_t)i) return 7;
    if (b.set >= kFirstSleepingSet && sleep_sets_[(size_t)b.set].bodies[(size_t)b.local] != (int32_t)i) return 8;
    if (b.type != kBodyStatic && b.island == kNullIndex) return 9;
  }
  for (size_t id = 0; id < hcontacts_.size(); ++id) {
    const HContact& c = hcontacts_[id];
    if (!c.alive) continue;
    if (c.color == kNullIndex) {
      const std::vector<int32_t>& list =
          c.set == kAwakeSet ? non_touching_ : (c.set == kDisabledSet ? disabled_ : sleep_sets_[(size_t)c.set].contacts);
      if (list[(size_t)c.local] != (int32_t)id) return 10;
    }
  }
  return 0;
}

}  // namespace nsl
