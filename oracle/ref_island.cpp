// Oracle: islands, island splitting, sleeping and waking solver sets.
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
//
// Restates src/simulation/island.rs:25-300 (merge, link and unlink of contacts and joints),
// src/simulation/sleep_island.rs:20-633 (union-find split, try_sleep_island, process_sleep),
// src/simulation/solver_set_ops.rs:21-160 (wake_solver_set) and the island pool helpers of
// src/simulation/world_runtime.rs:1419-1475, 1634-1644. Set ids: 0 static, 1 disabled,
// 2 awake, 3 and up sleeping.
#include "ref_world.h"

namespace nslref {

namespace {

constexpr int32_t kAwakeSet = 2, kDisabledSet = 1, kFirstSleepingSet = 3;
constexpr float kTimeToSleep = 0.5f;  // constants.rs:31

template <class T>
int32_t swap_remove(std::vector<T>& v, size_t index) {
  const size_t last = v.size() - 1;
  if (index != last) {
    v[index] = v[last];
    v.pop_back();
    return (int32_t)last;
  }
  v.pop_back();
  return kNullIndex;
}

// union find (sleep_island.rs:20-58)
int32_t find_root(std::vector<int32_t>& parents, int32_t node) {
  while (parents[node] != node) {
    const int32_t grand = parents[parents[node]];
    parents[node] = grand;
    node = grand;
  }
  return node;
}

void unite(std::vector<int32_t>& parents, std::vector<int32_t>& ranks, int32_t n1, int32_t n2,
           std::vector<int32_t>& contact_counts, std::vector<int32_t>& joint_counts) {
  const int32_t r1 = find_root(parents, n1), r2 = find_root(parents, n2);
  if (r1 == r2) return;
  int32_t child = r2, root = r1;
  if (ranks[r1] < ranks[r2]) {
    child = r1;
    root = r2;
  } else if (ranks[r1] == ranks[r2]) {
    ranks[r1] += 1;
  }
  parents[child] = root;
  contact_counts[root] += contact_counts[child];
  joint_counts[root] += joint_counts[child];
}

}  // namespace

std::vector<int32_t>& World::island_list(int32_t set_index) {
  return set_index == kAwakeSet ? awake_islands : sleep_sets[(size_t)set_index].island_ids;
}

std::vector<int32_t>& World::contact_list(int32_t set_index) {
  if (set_index == kAwakeSet) return awake_contact_indices;
  if (set_index == kDisabledSet) return disabled_contact_indices;
  return sleep_sets[(size_t)set_index].contact_indices;
}

// world_runtime.rs:1419-1445
int32_t World::create_island(int32_t set_index) {
  const int32_t id = island_pool.alloc();
  if (id == (int32_t)islands.size()) islands.push_back(Island{});
  Island is;
  is.set_index = set_index;
  is.local_index = (int32_t)island_list(set_index).size();
  is.island_id = id;
  islands[(size_t)id] = is;
  island_list(set_index).push_back(id);
  return id;
}

// world_runtime.rs:1447-1475
void World::destroy_island(int32_t island_id) {
  if (split_island_id == island_id) split_island_id = kNullIndex;
  Island& is = islands[(size_t)island_id];
  std::vector<int32_t>& list = island_list(is.set_index);
  const int32_t local = is.local_index;
  const int32_t last = (int32_t)list.size() - 1;
  const int32_t moved_id = list[(size_t)last];
  list[(size_t)local] = moved_id;
  islands[(size_t)moved_id].local_index = local;
  list.pop_back();
  Island& gone = islands[(size_t)island_id];
  gone = Island{};
  island_pool.free(island_id);
}

// world_runtime.rs:1634-1644
void World::create_island_for_body(int32_t set_index, int32_t body_id) {
  const int32_t id = create_island(set_index);
  islands[(size_t)id].bodies.push_back(body_id);
  bodies[(size_t)body_id].island_id = id;
  bodies[(size_t)body_id].island_index = 0;
}

// island.rs:32-106: the island with more bodies survives (ties keep a); the other's bodies,
// contacts and joints follow in their order, and its removal count carries over.
int32_t World::merge_islands(int32_t a, int32_t b) {
  if (a == b) return a;
  if (a == kNullIndex) return b;
  if (b == kNullIndex) return a;
  const int32_t big = islands[(size_t)a].bodies.size() >= islands[(size_t)b].bodies.size() ? a : b;
  const int32_t small = big == a ? b : a;
  const Island gone = islands[(size_t)small];
  for (int32_t body_id : gone.bodies) {
    Body& body = bodies[(size_t)body_id];
    body.island_id = big;
    body.island_index = (int32_t)islands[(size_t)big].bodies.size();
    islands[(size_t)big].bodies.push_back(body_id);
  }
  for (const ContactLink& link : gone.contacts) {
    Contact& c = contacts[(size_t)link.contact_id];
    c.island_id = big;
    c.island_index = (int32_t)islands[(size_t)big].contacts.size();
    islands[(size_t)big].contacts.push_back(link);
  }
  for (const JointLink& link : gone.joints) {
    Joint& j = joints[(size_t)link.joint_id];
    j.island_id = big;
    j.island_index = (int32_t)islands[(size_t)big].joints.size();
    islands[(size_t)big].joints.push_back(link);
  }
  islands[(size_t)big].constraint_remove_count += gone.constraint_remove_count;
  destroy_island(gone.island_id);
  return big;
}

// island.rs:126-171: wake a sleeping side touched by an awake body, merge, add the contact.
void World::link_contact(int32_t contact_id) {
  const int32_t a = contacts[(size_t)contact_id].edges[0].body_id;
  const int32_t b = contacts[(size_t)contact_id].edges[1].body_id;
  if (bodies[(size_t)a].set_index == kAwakeSet && bodies[(size_t)b].set_index >= kFirstSleepingSet)
    wake_solver_set(bodies[(size_t)b].set_index);
  if (bodies[(size_t)b].set_index == kAwakeSet && bodies[(size_t)a].set_index >= kFirstSleepingSet)
    wake_solver_set(bodies[(size_t)a].set_index);
  const int32_t island = merge_islands(bodies[(size_t)a].island_id, bodies[(size_t)b].island_id);
  Contact& c = contacts[(size_t)contact_id];
  c.island_id = island;
  c.island_index = (int32_t)islands[(size_t)island].contacts.size();
  islands[(size_t)island].contacts.push_back(ContactLink{contact_id, a, b});
}

// island.rs:174-206
void World::unlink_contact(int32_t contact_id) {
  Contact& c = contacts[(size_t)contact_id];
  const int32_t island_id = c.island_id;
  const int32_t remove = c.island_index;
  std::vector<ContactLink>& list = islands[(size_t)island_id].contacts;
  const int32_t moved = swap_remove(list, (size_t)remove);
  if (moved != kNullIndex) contacts[(size_t)list[(size_t)remove].contact_id].island_index = remove;
  Contact& c2 = contacts[(size_t)contact_id];
  c2.island_id = kNullIndex;
  c2.island_index = kNullIndex;
  islands[(size_t)island_id].constraint_remove_count += 1;
}

// island.rs:226-261
void World::link_joint(int32_t joint_id) {
  const int32_t a = joints[(size_t)joint_id].edges[0].body_id;
  const int32_t b = joints[(size_t)joint_id].edges[1].body_id;
  if (bodies[(size_t)a].set_index == kAwakeSet && bodies[(size_t)b].set_index >= kFirstSleepingSet)
    wake_solver_set(bodies[(size_t)b].set_index);
  else if (bodies[(size_t)b].set_index == kAwakeSet && bodies[(size_t)a].set_index >= kFirstSleepingSet)
    wake_solver_set(bodies[(size_t)a].set_index);
  const int32_t island = merge_islands(bodies[(size_t)a].island_id, bodies[(size_t)b].island_id);
  Joint& j = joints[(size_t)joint_id];
  j.island_id = island;
  j.island_index = (int32_t)islands[(size_t)island].joints.size();
  islands[(size_t)island].joints.push_back(JointLink{joint_id, a, b});
}

// sleep_island.rs:62-82
void World::refresh_body_contact_indices(int32_t body_id) {
  const Body& body = bodies[(size_t)body_id];
  const int32_t enc = encode_body_sim_index(body);
  for (int32_t key = body.head_contact_key; key != kNullIndex;) {
    const int32_t edge = key & 1, cid = key >> 1;
    Contact& c = contacts[(size_t)cid];
    key = c.edges[edge].next_key;
    if (edge == 0) c.encoded_a = enc;
    else c.encoded_b = enc;
  }
}

// sleep_island.rs:85-300: union-find over the island's contacts and joints; with more than
// one component, new islands in component order take the bodies (in island order), then
// the contacts and joints (in link order).
void World::split_island(int32_t base_id) {
  const Island base = islands[(size_t)base_id];
  const size_t n = base.bodies.size();
  std::vector<int32_t> parents(n), ranks(n, 0), contact_counts(n, 0), joint_counts(n, 0);
  for (size_t i = 0; i < n; ++i) parents[i] = (int32_t)i;
  auto island_index_in_base = [&](int32_t body_id, bool* valid) {
    const int32_t k = bodies[(size_t)body_id].island_index;
    *valid = k >= 0 && (size_t)k < n;
    return k;
  };
  auto connect = [&](int32_t body_a, int32_t body_b, std::vector<int32_t>& counts) {
    bool va, vb;
    const int32_t ka = island_index_in_base(body_a, &va), kb = island_index_in_base(body_b, &vb);
    if (va && vb) {
      unite(parents, ranks, ka, kb, contact_counts, joint_counts);
      counts[(size_t)find_root(parents, ka)] += 1;
    } else if (va || vb) {
      counts[(size_t)find_root(parents, va ? ka : kb)] += 1;
    }
  };
  for (const ContactLink& l : base.contacts) connect(l.body_id_a, l.body_id_b, contact_counts);
  for (const JointLink& l : base.joints) connect(l.body_id_a, l.body_id_b, joint_counts);
  size_t components = 0;
  for (size_t i = 0; i < n; ++i) {
    parents[i] = find_root(parents, (int32_t)i);
    if (parents[i] == (int32_t)i) components += 1;
  }
  if (components == 1) {
    islands[(size_t)base_id].constraint_remove_count = 0;
    return;
  }
  islands[(size_t)base_id].bodies.clear();
  islands[(size_t)base_id].contacts.clear();
  islands[(size_t)base_id].joints.clear();
  std::vector<int32_t> root_map(n, kNullIndex);
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) {
    const int32_t root = parents[i];
    if (root_map[(size_t)root] == kNullIndex) root_map[(size_t)root] = (int32_t)count++;
  }
  std::vector<int32_t> ids(count);
  for (size_t k = 0; k < count; ++k) ids[k] = create_island(kAwakeSet);
  for (size_t i = 0; i < n; ++i) {
    const int32_t body_id = base.bodies[i];
    const int32_t island = ids[(size_t)root_map[(size_t)find_root(parents, (int32_t)i)]];
    Body& body = bodies[(size_t)body_id];
    body.island_id = island;
    body.island_index = (int32_t)islands[(size_t)island].bodies.size();
    islands[(size_t)island].bodies.push_back(body_id);
  }
  for (const ContactLink& l : base.contacts) {
    const int32_t ia = bodies[(size_t)l.body_id_a].island_id;
    const int32_t island = ia != kNullIndex ? ia : bodies[(size_t)l.body_id_b].island_id;
    Contact& c = contacts[(size_t)l.contact_id];
    c.island_id = island;
    c.island_index = (int32_t)islands[(size_t)island].contacts.size();
    islands[(size_t)island].contacts.push_back(l);
  }
  for (const JointLink& l : base.joints) {
    const int32_t ia = bodies[(size_t)l.body_id_a].island_id;
    const int32_t island = ia != kNullIndex ? ia : bodies[(size_t)l.body_id_b].island_id;
    Joint& j = joints[(size_t)l.joint_id];
    j.island_id = island;
    j.island_index = (int32_t)islands[(size_t)island].joints.size();
    islands[(size_t)island].joints.push_back(l);
  }
  destroy_island(base_id);
}

// sleep_island.rs:300-560: the island's bodies leave the awake set (swap removal; a moved
// body of another island gets its contacts' indices refreshed), their non-touching contacts
// with no awake partner go to the disabled set, touching contacts and joints leave the
// graph for the new sleeping set, and the island moves with them.
void World::try_sleep_island(int32_t island_id) {
  Island island = islands[(size_t)island_id];
  if (island.constraint_remove_count > 0 && island.bodies.size() > 1) return;
  const int32_t set_id = set_pool.alloc();
  if (set_id >= (int32_t)sleep_sets.size()) sleep_sets.resize((size_t)set_id + 1);
  SleepSet ss;
  ss.set_index = set_id;
  for (size_t i = 0; i < island.bodies.size(); ++i) {
    const int32_t body_id = island.bodies[i];
    const int32_t awake_index = bodies[(size_t)body_id].local_index;
    ss.body_sims.push_back(awake_sims[(size_t)awake_index]);
    const int32_t sleep_index = (int32_t)ss.body_sims.size() - 1;
    const int32_t moved = swap_remove(awake_sims, (size_t)awake_index);
    if (moved != kNullIndex) {
      const int32_t moved_id = awake_sims[(size_t)awake_index].body_id;
      bodies[(size_t)moved_id].local_index = awake_index;
      if (bodies[(size_t)moved_id].island_id != island_id) refresh_body_contact_indices(moved_id);
    }
    swap_remove(awake_states, (size_t)awake_index);
    Body& body = bodies[(size_t)body_id];
    body.set_index = set_id;
    body.local_index = sleep_index;
    for (int32_t key = body.head_contact_key; key != kNullIndex;) {
      const int32_t cid = key >> 1, edge = key & 1;
      Contact& c = contacts[(size_t)cid];
      key = c.edges[edge].next_key;
      if (c.color_index != kNullIndex) continue;  // touching: moved with the island below
      if (edge == 0) c.encoded_a = kNullIndex;
      else c.encoded_b = kNullIndex;
      if (bodies[(size_t)c.edges[edge ^ 1].body_id].set_index == kAwakeSet) continue;
      const int32_t local = c.local_index;
      c.set_index = kDisabledSet;
      c.local_index = (int32_t)disabled_contact_indices.size();
      disabled_contact_indices.push_back(cid);
      const int32_t moved_local = swap_remove(awake_contact_indices, (size_t)local);
      if (moved_local != kNullIndex) contacts[(size_t)awake_contact_indices[(size_t)local]].local_index = local;
    }
  }
  const int32_t overflow = kGraphColorCount - 1;
  for (const ContactLink& link : island.contacts) {
    const int32_t cid = link.contact_id;
    Contact& c = contacts[(size_t)cid];
    const int32_t color = c.color_index;
    if (color != overflow) {
      colors[color].body_set.clear(c.edges[0].body_id);
      colors[color].body_set.clear(c.edges[1].body_id);
    }
    ss.contact_indices.push_back(cid);
    const int32_t sleep_contact_index = (int32_t)ss.contact_indices.size() - 1;
    const int32_t local = c.local_index;
    if ((c.flags & CF::SIM_MESH_CONTACT) || color == overflow) {
      const int32_t moved = swap_remove(colors[color].contacts, (size_t)local);
      if (moved != kNullIndex) contacts[(size_t)colors[color].contacts[(size_t)local].contact_id].local_index = local;
    } else {
      const int32_t moved = swap_remove(colors[color].convex_contacts, (size_t)local);
      if (moved != kNullIndex) contacts[(size_t)colors[color].convex_contacts[(size_t)local]].local_index = local;
    }
    Contact& c2 = contacts[(size_t)cid];
    c2.set_index = set_id;
    c2.color_index = kNullIndex;
    c2.local_index = sleep_contact_index;
    if (c2.encoded_a >= 0) c2.encoded_a = kNullIndex;  // sleep_body_sim_index (world_runtime.rs:923-925)
    if (c2.encoded_b >= 0) c2.encoded_b = kNullIndex;
  }
  for (const JointLink& link : island.joints) {
    Joint& j = joints[(size_t)link.joint_id];
    const int32_t color = j.color_index, local = j.local_index;
    if (color != overflow) {
      colors[color].body_set.clear(j.edges[0].body_id);
      colors[color].body_set.clear(j.edges[1].body_id);
    }
    ss.joint_sims.push_back(colors[color].joint_sims[(size_t)local]);
    const int32_t sleep_joint_index = (int32_t)ss.joint_sims.size() - 1;
    const int32_t moved = swap_remove(colors[color].joint_sims, (size_t)local);
    if (moved != kNullIndex) joints[(size_t)colors[color].joint_sims[(size_t)local].joint_id].local_index = local;
    Joint& j2 = joints[(size_t)link.joint_id];
    j2.set_index = set_id;
    j2.color_index = kNullIndex;
    j2.local_index = sleep_joint_index;
  }
  const int32_t island_index = island.local_index;
  ss.island_ids.push_back(island_id);
  const int32_t moved_island = swap_remove(awake_islands, (size_t)island_index);
  if (moved_island != kNullIndex) islands[(size_t)awake_islands[(size_t)island_index]].local_index = island_index;
  island.set_index = set_id;
  island.local_index = 0;
  islands[(size_t)island_id] = island;
  if (split_island_id == island_id) split_island_id = kNullIndex;
  sleep_sets[(size_t)set_id] = std::move(ss);
}

// sleep_island.rs:563-633: take the finalize pass's split candidate for the next step, then
// try to sleep every awake island no body kept awake, last island first.
void World::process_sleep() {
  if (split_candidate != kNullIndex) split_island_id = split_candidate;
  const std::vector<int32_t> islands_copy = awake_islands;
  const std::vector<uint8_t> bits = awake_island_bits;
  for (size_t k = islands_copy.size(); k-- > 0;) {
    if (k < bits.size() && bits[k]) continue;
    try_sleep_island(islands_copy[k]);
  }
}

// constraint_graph.rs:256-323 (add_joint_to_graph)
void World::add_joint_to_graph(const JointSim& sim, int32_t joint_id) {
  Joint& j = joints[(size_t)joint_id];
  const int32_t a = j.edges[0].body_id, b = j.edges[1].body_id;
  const int32_t color = assign_joint_color(a, b, bodies[(size_t)a].type, bodies[(size_t)b].type);
  Joint& j2 = joints[(size_t)joint_id];
  j2.color_index = color;
  j2.local_index = (int32_t)colors[color].joint_sims.size();
  colors[color].joint_sims.push_back(sim);
}

// solver_set_ops.rs:21-160: bodies return to the awake set with zero velocity and a reset
// sleep timer, their disabled contacts return to the awake list, the touching contacts and
// joints re-enter the graph in their sleeping-set order, and the islands rejoin the awake set.
void World::wake_solver_set(int32_t set_index) {
  SleepSet ss = std::move(sleep_sets[(size_t)set_index]);
  sleep_sets[(size_t)set_index] = SleepSet{};
  for (const BodySim& sim : ss.body_sims) {
    Body& body = bodies[(size_t)sim.body_id];
    body.set_index = kAwakeSet;
    body.local_index = (int32_t)awake_sims.size();
    body.sleep_time = 0.0f;
    const int32_t enc = body.local_index;
    awake_sims.push_back(sim);
    BodyState st{};
    st.dq = kQuatIdentity;
    st.flags = body.flags;
    awake_states.push_back(st);
    for (int32_t key = bodies[(size_t)sim.body_id].head_contact_key; key != kNullIndex;) {
      const int32_t edge = key & 1, cid = key >> 1;
      Contact& c = contacts[(size_t)cid];
      key = c.edges[edge].next_key;
      if (c.set_index != kDisabledSet) {
        if (c.set_index == kAwakeSet) {
          if (edge == 0) c.encoded_a = enc;
          else c.encoded_b = enc;
        }
        continue;
      }
      const int32_t local = c.local_index;
      c.set_index = kAwakeSet;
      c.local_index = (int32_t)awake_contact_indices.size();
      awake_contact_indices.push_back(cid);
      if (edge == 0) c.encoded_a = enc;
      else c.encoded_b = enc;
      const int32_t moved = swap_remove(disabled_contact_indices, (size_t)local);
      if (moved != kNullIndex) contacts[(size_t)disabled_contact_indices[(size_t)local]].local_index = local;
    }
  }
  for (int32_t cid : ss.contact_indices) {
    add_contact_to_graph(cid);
    contacts[(size_t)cid].set_index = kAwakeSet;
  }
  for (const JointSim& sim : ss.joint_sims) {
    add_joint_to_graph(sim, sim.joint_id);
    joints[(size_t)sim.joint_id].set_index = kAwakeSet;
  }
  for (int32_t island_id : ss.island_ids) {
    islands[(size_t)island_id].set_index = kAwakeSet;
    islands[(size_t)island_id].local_index = (int32_t)awake_islands.size();
    awake_islands.push_back(island_id);
  }
  set_pool.free(set_index);
}

}  // namespace nslref
