// Batched stepper (see nsl_batch.h): per step, three launches over all worlds (pair
// update, collide pass, transitions and solve), each world in its own workgroup.
#include "nsl_batch.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>

#include "nsl_continuous.h"
#include "nsl_stages.h"

namespace nsl {

#define BATCH_CHECK(expr)                                                                                \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_));     \
  } while (0)

namespace {

// One world's pointers into the batch arrays.
struct WorldRefs {
  const BodyMeta* body;
  const Proxy* proxy;
  const Shape* shapes;
  const Aabb* fat;
  uint8_t* moved;
  ContactMeta* meta;
  Contact* contacts;
  uint8_t* change;
  int32_t* free_ids;
  int32_t* non_touching;
  int32_t* overflow_list;
  int32_t* colors;
  uint32_t* bits;
  uint32_t* pairs;
  BatchWorldState* ws;
  int32_t n_shapes, contact_cap, overflow_cap, color_cap, bits_words;
};

__device__ WorldRefs world_refs(const BatchDevice& b, const BatchWorldDesc& w, int world) {
  WorldRefs r;
  r.body = b.body_meta + w.body_off;
  r.proxy = b.proxies + w.shape_off;
  r.shapes = b.shapes + w.shape_off;
  r.fat = b.fat_aabbs + w.shape_off;
  r.moved = b.shape_moved + w.shape_off;
  r.meta = b.contact_meta + w.contact_off;
  r.contacts = b.contacts + w.contact_off;
  r.change = b.contact_change + w.contact_off;
  r.free_ids = b.free_ids + w.contact_off;
  r.non_touching = b.non_touching + w.contact_off;
  r.overflow_list = b.overflow_list + w.overflow_off;
  r.colors = b.color_lists + w.color_off;
  r.bits = b.color_bits + w.bits_off;
  r.pairs = b.pair_bits + w.pair_off;
  r.ws = b.wstate + world;
  r.n_shapes = w.n_shapes;
  r.contact_cap = w.contact_cap;
  r.overflow_cap = w.overflow_cap;
  r.color_cap = w.color_cap;
  r.bits_words = w.bits_words;
  return r;
}

// The world as a DeviceWorld over the shared per-item functions; `states` is the world's
// slice of the global states or its LDS copy.
__device__ DeviceWorld world_view(const BatchDevice& b, const BatchWorldDesc& w, BodyState* states) {
  DeviceWorld d;
  d.sims = b.sims + w.awake_off;
  d.states = states;
  d.static_sims = b.static_sims + w.static_off;
  d.shapes = b.shapes + w.shape_off;
  d.fat_aabbs = b.fat_aabbs + w.shape_off;
  d.shape_moved = b.shape_moved + w.shape_off;
  d.hulls = b.hulls + w.hull_off;
  d.contacts = b.contacts + w.contact_off;
  d.contact_change = b.contact_change + w.contact_off;
  d.constraints = b.constraints + w.contact_off;
  d.overflow = b.overflow + w.overflow_off;
  d.collide_ids = nullptr;
  d.contact_active = nullptr;
  d.solve_ids = b.solve_ids + w.contact_off;
  d.overflow_ids = b.overflow_ids + w.overflow_off;
  d.counters = b.counters;
  d.pairs = PairSet{nullptr, 0};
  d.joints = nullptr;
  d.joint_ids = nullptr;
  d.joint_pairs = PairSet{nullptr, 0};
  d.shape_aabbs = b.shape_aabbs + w.shape_off;
  d.fast_ids = b.fast_ids;
  d.fast_base = w.awake_off;
  d.n_awake = w.n_awake;
  d.change_list = nullptr;
  d.sleep_times = nullptr;
  d.sleep_sims = d.sims;  // never read: batch worlds do not sleep, so no encoding is -1
  return d;
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lanes_below(uint64_t mask, int lane) {
  return __popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}
__device__ __forceinline__ void wave_sync() { __syncthreads(); }  // the workgroup is one wavefront

__device__ __forceinline__ bool pair_bit(const WorldRefs& r, int a, int b) {
  const uint32_t k = (uint32_t)a * (uint32_t)r.n_shapes + (uint32_t)b;
  return (r.pairs[k >> 5] >> (k & 31)) & 1u;
}
__device__ __forceinline__ void set_pair_bit(const WorldRefs& r, int a, int b, bool on) {
  const uint32_t k = (uint32_t)a * (uint32_t)r.n_shapes + (uint32_t)b;
  if (on) atomicOr(&r.pairs[k >> 5], 1u << (k & 31));
  else atomicAnd(&r.pairs[k >> 5], ~(1u << (k & 31)));
}

__device__ __forceinline__ bool colour_has(const WorldRefs& r, int c, int body) {
  return (r.bits[c * r.bits_words + (body >> 5)] >> (body & 31)) & 1u;
}
__device__ __forceinline__ void colour_set(const WorldRefs& r, int c, int body, bool on) {
  uint32_t& w = r.bits[c * r.bits_words + (body >> 5)];
  if (on) w |= 1u << (body & 31);
  else w &= ~(1u << (body & 31));
}

// contact_creation.rs:219-425 for the contact of shapes a, b that takes id `id` and entry
// `local` of the non-touching list; the primary shape is hull > capsule > sphere.
__device__ void create_contact_at(const WorldRefs& r, int32_t a, int32_t b, int32_t id, int32_t local, bool enable_speculative) {
  auto rank = [](int32_t kind) { return kind == kShapeHull ? 2 : (kind == kShapeCapsule ? 1 : 0); };
  if (rank(r.shapes[a].kind) < rank(r.shapes[b].kind)) {
    const int32_t t = a;
    a = b;
    b = t;
  }
  const Shape& sa = r.shapes[a];
  const Shape& sb = r.shapes[b];
  const Proxy& pa = r.proxy[a];
  const Proxy& pb = r.proxy[b];
  const BodyMeta& ba = r.body[pa.body];
  const BodyMeta& bb = r.body[pb.body];
  ContactMeta m;
  m.shape_a = a;
  m.shape_b = b;
  m.color = kNullIndex;
  m.local = local;
  m.alive = 1;
  m.restitution = maxf(sa.restitution, sb.restitution);
  m.pair_type = sa.kind == kShapeHull ? (sb.kind == kShapeHull ? 0 : (sb.kind == kShapeSphere ? 1 : 2))
                                      : (sa.kind == kShapeCapsule ? (sb.kind == kShapeSphere ? 3 : 4) : 5);
  m.body_a = pa.body;
  m.body_b = pb.body;
  m.dynamic = (ba.type == kBodyDynamic ? 1 : 0) | (bb.type == kBodyDynamic ? 2 : 0);
  r.meta[id] = m;
  r.non_touching[local] = id;
  Contact c{};
  c.shape_a = a;
  c.shape_b = b;
  c.enc_a = ba.enc;
  c.enc_b = bb.enc;
  if (ba.type == kBodyStatic || bb.type == kBodyStatic) c.flags |= cflag::kStatic;
  if (enable_speculative && (pa.flags & kProxySpeculative) && (pb.flags & kProxySpeculative)) c.flags |= cflag::kSpeculative;
  if ((ba.flags & bflag::kEnableRecycling) && (bb.flags & bflag::kEnableRecycling) && (c.flags & cflag::kSpeculative))
    c.flags |= cflag::kRecycle;
  c.rolling_resistance = maxf(sa.rolling_resistance, sb.rolling_resistance) * maxf(sa.radius, sb.radius);
  c.cached_rot_a = quat_identity();
  c.cached_rot_b = quat_identity();
  r.contacts[id] = c;
  set_pair_bit(r, a < b ? a : b, a < b ? b : a, true);
}

// Swap-remove of entry `local` of `list` (length *n); returns the moved contact id or -1.
__device__ __forceinline__ int32_t swap_remove(int32_t* list, int32_t* n, int32_t local) {
  const int32_t last = --(*n);
  if (local != last) {
    list[local] = list[last];
    return list[local];
  }
  return kNullIndex;
}

// constraint_graph.rs:200-251
__device__ void remove_from_graph(const WorldRefs& r, int32_t body_a, int32_t body_b, int32_t color, int32_t local) {
  BatchWorldState* ws = r.ws;
  if (color != kOverflowColor) {
    colour_set(r, color, body_a, false);
    colour_set(r, color, body_b, false);
    const int32_t moved = swap_remove(r.colors + color * r.color_cap, &ws->color_count[color], local);
    if (moved != kNullIndex) r.meta[moved].local = local;
  } else {
    const int32_t moved = swap_remove(r.overflow_list, &ws->n_overflow, local);
    if (moved != kNullIndex) r.meta[moved].local = local;
  }
}

// constraint_graph.rs:113-197
__device__ void add_to_graph(const WorldRefs& r, int32_t id) {
  ContactMeta& c = r.meta[id];
  const int32_t a = c.body_a, b = c.body_b;
  const bool dyn_a = (c.dynamic & 1) != 0, dyn_b = (c.dynamic & 2) != 0;
  int32_t color = kOverflowColor;
  if (dyn_a && dyn_b) {
    for (int i = 0; i < kDynamicColorCount; ++i) {
      if (colour_has(r, i, a) || colour_has(r, i, b)) continue;
      colour_set(r, i, a, true);
      colour_set(r, i, b, true);
      color = i;
      break;
    }
  } else if (dyn_a || dyn_b) {
    const int32_t dyn = dyn_a ? a : b;
    for (int i = kOverflowColor - 1; i >= 1; --i) {
      if (!colour_has(r, i, dyn)) {
        colour_set(r, i, dyn, true);
        color = i;
        break;
      }
    }
  }
  BatchWorldState* ws = r.ws;
  c.color = color;
  if (color == kOverflowColor) {
    if (ws->n_overflow >= r.overflow_cap) {
      ws->error = kBatchOverflowFull;
      return;
    }
    c.local = ws->n_overflow;
    r.overflow_list[ws->n_overflow++] = id;
  } else {
    c.local = ws->color_count[color];
    r.colors[color * r.color_cap + ws->color_count[color]++] = id;
  }
}

// contact_creation.rs:428-552 and id_pool.rs:36-45
__device__ void destroy_contact(const WorldRefs& r, int32_t id) {
  ContactMeta& c = r.meta[id];
  BatchWorldState* ws = r.ws;
  if (c.color == kNullIndex) {
    const int32_t moved = swap_remove(r.non_touching, &ws->n_non_touching, c.local);
    if (moved != kNullIndex) r.meta[moved].local = c.local;
  } else {
    remove_from_graph(r, c.body_a, c.body_b, c.color, c.local);
  }
  c.alive = 0;
  c.color = kNullIndex;
  c.local = kNullIndex;
  // id_pool.rs:36-45 shrinks the pool only for id == next_index, which no live id equals
  r.free_ids[ws->n_free++] = id;
}

// collide_pass.rs:386-478: one contact's transition (contacts are visited in id order).
__device__ void apply_change(const WorldRefs& r, int32_t id, uint8_t ch) {
  if (ch == kChangeDisjoint) {
    destroy_contact(r, id);
  } else if (ch == kChangeBegan) {
    const int32_t old_local = r.meta[id].local;
    add_to_graph(r, id);
    const int32_t moved = swap_remove(r.non_touching, &r.ws->n_non_touching, old_local);
    if (moved != kNullIndex) r.meta[moved].local = old_local;
  } else if (ch == kChangeEnded) {
    ContactMeta& c = r.meta[id];
    const int32_t color = c.color, local = c.local;
    c.color = kNullIndex;
    c.local = r.ws->n_non_touching;
    r.non_touching[r.ws->n_non_touching++] = id;
    remove_from_graph(r, c.body_a, c.body_b, color, local);
  }
}

// Pair test of the update (contact_creation.rs:901-1042, shape_runtime.rs:285-292).
__device__ __forceinline__ bool new_pair(const WorldRefs& r, int m, const Proxy& pm, const Aabb& fm, int j, const Aabb& fj) {
  if (!aabb_overlaps(fm, fj)) return false;
  const Proxy pj = r.proxy[j];
  if (pm.body_type != kBodyDynamic && pj.body_type != kBodyDynamic) return false;
  if (pm.body == pj.body) return false;
  if ((pm.flags | pj.flags) & kProxySensor) return false;  // contact_creation.rs:187
  if (pm.group == pj.group && pm.group != 0) {
    if (pm.group < 0) return false;
  } else if (!((pm.mask & pj.category) != 0 && (pm.category & pj.mask) != 0)) {
    return false;
  }
  return !pair_bit(r, m < j ? m : j, m < j ? j : m);
}

// Ascending sort of n keys in LDS by the workgroup (bitonic over the next power of two;
// the padding holds all-ones keys).
__device__ void sort_keys(uint32_t* keys, int n) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int p = 1;
  while (p < n) p <<= 1;
  for (int i = n + tid; i < p; i += nt) keys[i] = ~0u;
  __syncthreads();
  for (int k = 2; k <= p; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < p; i += nt) {
        const int x = i ^ j;
        if (x > i) {
          const uint32_t a = keys[i], b = keys[x];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[x] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

struct Colours {
  int n;
  int begin[kOverflowColor];
  int end[kOverflowColor];
};

// One solve stage over the colours in colour order, a barrier after each colour: each lane
// runs item(slot, record) on its slots (slot = colour begin + lane + 64 k), and the record
// of its slot in the next round (the colour's next 64 slots or the next colour's first) is
// loaded before the current item runs, so that the load overlaps the item's work. A stage
// writes each record only from its own item, so a record loaded one round early holds the
// values the item would read.
template <class Item>
__device__ __forceinline__ void colour_stage(const DeviceWorld& d, const Colours& cs, int lane, Item item) {
  if (cs.n == 0) return;
  int a = 0, base = cs.begin[0];
  bool live = base + lane < cs.end[0];
  Constraint cur;
  if (live) cur = blocked_cons(d).load(base + lane);
  for (;;) {
    int na = a, nbase = base + kBatchLanes;
    if (nbase >= cs.end[a]) {
      na = a + 1;
      nbase = na < cs.n ? cs.begin[na] : 0;
    }
    const bool more = na < cs.n;
    const bool next_live = more && nbase + lane < cs.end[na];
    Constraint next;
    if (next_live) next = blocked_cons(d).load(nbase + lane);
    if (live) item(base + lane, cur);
    if (na != a) wave_sync();
    if (!more) break;
    a = na;
    base = nbase;
    live = next_live;
    cur = next;
  }
}

// Pair update of one world (contact_creation.rs:901-1042) by a workgroup of four
// wavefronts: every moved proxy against every shape (each wavefront takes every fourth
// 64-shape chunk of moved proxies; the fat AABBs and moved flags sit in LDS), the new keys
// sorted, and the contacts created in key order: the k-th key takes the k-th id the
// reference's pool hands out (the free list from its top, then fresh ids) and the k-th
// free entry of the non-touching list, so all keys are created at once.
__global__ __launch_bounds__(kBatchPairsBlock) void k_batch_pairs(BatchDevice b) {
  __shared__ uint32_t s_keys[kBatchMaxNewPairs];
  __shared__ float4 s_lo[kPairsLdsShapes], s_hi[kPairsLdsShapes];  // fat AABB; s_lo.w: 1 if the proxy moved
  __shared__ int s_count, s_type_count[kPairTypes], s_type_next[kPairTypes];
  const int tid = threadIdx.x, lane = tid % kBatchLanes, wave = tid / kBatchLanes;
  constexpr int kWaves = kBatchPairsBlock / kBatchLanes;
  if (blockIdx.x == 0 && tid == 0) b.counters[kCntFastList] = 0;  // the solve launch lists this step's fast bodies
  const int world = b.order[blockIdx.x];
  const BatchWorldDesc wd = b.worlds[world];
  const WorldRefs r = world_refs(b, wd, world);
  BatchWorldState* ws = r.ws;
  if (ws->error != kBatchOk) return;
  const int n_shapes = r.n_shapes;
  if (tid == 0) s_count = 0;
  if (tid < kPairTypes) s_type_count[tid] = 0;
  for (int j = tid; j < n_shapes && j < kPairsLdsShapes; j += kBatchPairsBlock) {
    const Aabb f = r.fat[j];
    s_lo[j] = make_float4(f.lo.x, f.lo.y, f.lo.z, r.moved[j] != 0 ? 1.0f : 0.0f);
    s_hi[j] = make_float4(f.hi.x, f.hi.y, f.hi.z, 0.0f);
  }
  __syncthreads();
  auto fat_of = [&](int j) {
    if (j < kPairsLdsShapes) {
      const float4 lo = s_lo[j], hi = s_hi[j];
      return Aabb{Vec3{lo.x, lo.y, lo.z}, Vec3{hi.x, hi.y, hi.z}};
    }
    return r.fat[j];
  };
  auto moved_of = [&](int j) { return j < kPairsLdsShapes ? s_lo[j].w != 0.0f : r.moved[j] != 0; };
  for (int base = wave * kBatchLanes; base < n_shapes; base += kWaves * kBatchLanes) {
    uint64_t mset = wave_ballot(base + lane < n_shapes && moved_of(base + lane));
    while (mset) {
      const int m = base + (__ffsll((unsigned long long)mset) - 1);
      mset &= mset - 1;
      const Proxy pm = r.proxy[m];
      const Aabb fm = fat_of(m);
      for (int j0 = 0; j0 < n_shapes; j0 += kBatchLanes) {
        const int j = j0 + lane;
        // a pair of two moved proxies is found from the lower one
        const bool hit = j < n_shapes && j != m && (j > m || !moved_of(j)) && new_pair(r, m, pm, fm, j, fat_of(j));
        const uint64_t hits = wave_ballot(hit);
        if (hits == 0) continue;
        int at = 0;
        if (lane == 0) at = atomicAdd(&s_count, __popcll(hits));
        at = __shfl(at, 0, kBatchLanes);
        const int k = at + lanes_below(hits, lane);
        if (hit && k < kBatchMaxNewPairs) {
          const int lo = m < j ? m : j, hi = m < j ? j : m;
          s_keys[k] = ((uint32_t)lo << 16) | (uint32_t)hi;
        }
      }
    }
  }
  __syncthreads();
  if (s_count > kBatchMaxNewPairs) {
    if (tid == 0) ws->error = kBatchPairsFull;
    return;
  }
  const int n_keys = s_count;
  if (n_keys > 0) {
    sort_keys(s_keys, n_keys);
    const int n_free = ws->n_free, next_id = ws->next_id, n_nt = ws->n_non_touching;
    // keys past the pool's capacity are not created and the world stops (contact_creation.rs:219-425)
    const int fresh = mini(maxi(n_keys - n_free, 0), maxi(r.contact_cap - next_id, 0));
    const int n_create = mini(n_keys, n_free) + fresh;
    for (int k = tid; k < n_create; k += kBatchPairsBlock) {
      const int32_t id = k < n_free ? r.free_ids[n_free - 1 - k] : next_id + (k - n_free);
      create_contact_at(r, (int32_t)(s_keys[k] >> 16), (int32_t)(s_keys[k] & 0xffffu), id, n_nt + k, b.enable_speculative != 0);
    }
    __syncthreads();
    if (tid == 0) {
      ws->n_free = maxi(n_free - n_keys, 0);
      ws->next_id = next_id + fresh;
      ws->n_non_touching = n_nt + n_create;
      if (n_create < n_keys) ws->error = kBatchContactsFull;
    }
  }
  for (int s = tid; s < n_shapes; s += kBatchPairsBlock) r.moved[s] = 0;
  __syncthreads();
  if (ws->error != kBatchOk) return;
  // the collide launch's work: this world's live contacts appended to their routine's list,
  // counted first so each list takes one atomic per world
  const int n_ids = ws->next_id;
  auto type_of = [&](int id) { return id < n_ids && r.meta[id].alive ? r.meta[id].pair_type : -1; };
  for (int base = wave * kBatchLanes; base < n_ids; base += kWaves * kBatchLanes) {
    const int type = type_of(base + lane);
    for (int t = 0; t < kPairTypes; ++t) {
      const int n = __popcll(wave_ballot(type == t));
      if (lane == 0 && n > 0) atomicAdd(&s_type_count[t], n);
    }
  }
  __syncthreads();
  if (tid < kPairTypes) s_type_next[tid] = s_type_count[tid] > 0 ? atomicAdd(&b.collide_counts[tid], s_type_count[tid]) : 0;
  __syncthreads();
  for (int base = wave * kBatchLanes; base < n_ids; base += kWaves * kBatchLanes) {
    const int id = base + lane;
    const int type = type_of(id);
    if (id < n_ids && type < 0) r.change[id] = kChangeNone;
    for (int t = 0; t < kPairTypes; ++t) {
      const uint64_t m = wave_ballot(type == t);
      if (m == 0) continue;
      int at = 0;
      if (lane == 0) at = atomicAdd(&s_type_next[t], __popcll(m));
      at = __shfl(at, 0, kBatchLanes);
      if (type == t) b.collide_items[(size_t)t * b.collide_type_cap + at + lanes_below(m, lane)] = make_int2(world, id);
    }
  }
}

// Collide pass (collide_pass.rs:541-742) over the lists the pair launch wrote: one lane per
// live contact, the contacts of one collide routine together so a wavefront's lanes take
// the same code path, and the contacts of every world spread over the whole device
// whatever the world sizes. The grid is as large as can be resident at once and strides
// over the lists.
__global__ __launch_bounds__(kBatchCollideBlock, 2) void k_batch_collide(BatchDevice b) {
  int count[kPairTypes], total = 0;
  for (int t = 0; t < kPairTypes; ++t) {
    count[t] = b.collide_counts[t];
    total += count[t];
  }
  int collided = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
    int t = 0, first = 0;
    while (k >= first + count[t]) first += count[t++];
    const int2 item = b.collide_items[(size_t)t * b.collide_type_cap + (k - first)];
    const int world = item.x, id = item.y;
    const BatchWorldDesc wd = b.worlds[world];
    const WorldRefs r = world_refs(b, wd, world);
    const DeviceWorld dg = world_view(b, wd, b.states + wd.awake_off);
    ++collided;
    const uint8_t ch = collide_contact(dg, id);
    if (ch == kChangeDisjoint) {
      const int a = r.meta[id].shape_a, c = r.meta[id].shape_b;
      set_pair_bit(r, a < c ? a : c, a < c ? c : a, false);
    }
    r.change[id] = ch;
  }
  for (int off = kBatchLanes / 2; off > 0; off >>= 1) collided += __shfl_xor(collided, off, kBatchLanes);
  if (threadIdx.x % kBatchLanes == 0 && collided) atomicAdd(&b.work[0], (unsigned long long)collided);
}

// Contact transitions in id order (one lane), the solve order, then prepare, the
// substeps with the world's body states in LDS, restitution, impulse storage and
// finalisation (solver_stages.rs:543-589, solver_integration.rs:587-799).
// Tape of one world's step, written after preparation (one wavefront): the prepared
// constraints, their points' feature ids and shape pair, the links to the previous step's
// constraints (same contact, points matched by feature id in order, as the impulse
// carry-over matches them), the colour ranges, and the body states and centres at the
// step's start. A step outside the differentiable model (overflow constraints,
// restitution, rolling resistance, tangent velocity, gyroscopic bodies, or more
// constraints or bodies than the tape holds) is marked with count -1.
__device__ void tape_record(const BatchDevice& b, int world, const BatchWorldDesc& wd, const DeviceWorld& dl, const WorldRefs& r,
                            int n_constraints, int n_ovf, bool any_restitution, const Colours& colours,
                            const BodyState* s_states) {
  const BatchTape& t = b.tape;
  const int lane = threadIdx.x;
  const int k = b.tape_step;
  const size_t ws = (size_t)k * t.n_worlds + world;
  const int n = mini(n_constraints, t.max_con);
  bool bad = false;
  for (int slot = lane; slot < n; slot += kBatchLanes) {
    const Constraint c = blocked_cons(dl).load(slot);
    const Contact& ct = dl.contacts[c.contact_id];
    t.cons[ws * t.max_con + slot] = c;
    for (int i = 0; i < 4; ++i) t.features[(ws * t.max_con + slot) * 4 + i] = i < c.point_count ? ct.points[i].feature_id : 0xffffffffu;
    const ContactMeta& m = r.meta[c.contact_id];
    t.pairs[ws * t.max_con + slot] = ((uint32_t)m.shape_a << 16) | (uint32_t)m.shape_b;
    bad = bad || c.rolling_resistance > 0.0f || c.tangent_velocity1 != 0.0f || c.tangent_velocity2 != 0.0f;
  }
  const int nb = mini(wd.n_awake, t.max_bodies);
  for (int i = lane; i < nb; i += kBatchLanes) {
    t.states[ws * t.max_bodies + i] = s_states[i];
    t.centers[ws * t.max_bodies + i] = dl.sims[i].center;
    bad = bad || (dl.sims[i].flags & bflag::kGyroscopic) != 0;
  }
  const bool ok = wave_ballot(bad) == 0 && n_ovf == 0 && !any_restitution && n_constraints <= t.max_con &&
                  wd.n_awake <= t.max_bodies;
  if (lane == 0) {
    int32_t* col = t.colours + ws * kTapeColourWords;
    col[0] = colours.n;
    for (int a = 0; a < colours.n; ++a) {
      col[1 + a] = colours.begin[a];
      col[1 + kOverflowColor + a] = colours.end[a];
    }
  }
  // links: the previous step's slot of the same contact (same id and shapes)
  const size_t wp = k > 0 ? (size_t)(k - 1) * t.n_worlds + world : 0;
  const int n_prev = k > 0 ? t.count[wp] : -1;
  for (int slot = lane; slot < n; slot += kBatchLanes) {
    int32_t link = -1;
    uint8_t pm[4] = {kTapeNoPoint, kTapeNoPoint, kTapeNoPoint, kTapeNoPoint};
    const Constraint& c = t.cons[ws * t.max_con + slot];
    const uint32_t pair = t.pairs[ws * t.max_con + slot];
    for (int j = 0; j < n_prev; ++j) {
      if (t.cons[wp * t.max_con + j].contact_id == c.contact_id && t.pairs[wp * t.max_con + j] == pair) {
        link = j;
        break;
      }
    }
    if (link >= 0) {
      uint32_t old[4];
      const int old_count = t.cons[wp * t.max_con + link].point_count;
      for (int i = 0; i < 4; ++i) old[i] = i < old_count ? t.features[(wp * t.max_con + link) * 4 + i] : 0xffffffffu;
      for (int i = 0; i < c.point_count; ++i) {
        const uint32_t f = t.features[(ws * t.max_con + slot) * 4 + i];
        for (int j = 0; j < old_count; ++j)
          if (old[j] == f) {
            pm[i] = (uint8_t)j;
            old[j] = 0xffffffffu;
            break;
          }
      }
    }
    t.links[ws * t.max_con + slot] = link;
    for (int i = 0; i < 4; ++i) t.pmap[(ws * t.max_con + slot) * 4 + i] = pm[i];
  }
  if (lane == 0) t.count[ws] = ok ? n_constraints : -1;
}

// States and centres after the last recorded step.
__global__ __launch_bounds__(kBatchLanes) void k_tape_final(BatchDevice b, int step) {
  const int world = blockIdx.x;
  const BatchWorldDesc wd = b.worlds[world];
  const BatchTape& t = b.tape;
  const size_t ws = (size_t)step * t.n_worlds + world;
  const int nb = mini(wd.n_awake, t.max_bodies);
  for (int i = threadIdx.x; i < nb; i += kBatchLanes) {
    t.states[ws * t.max_bodies + i] = b.states[wd.awake_off + i];
    t.centers[ws * t.max_bodies + i] = b.sims[wd.awake_off + i].center;
  }
}

__global__ __launch_bounds__(kBatchLanes) __attribute__((amdgpu_waves_per_eu(2))) void k_batch_solve(BatchDevice b, StepParams p) {
  // s_states holds the world's colour bits during the transitions, then its body states
  __shared__ BodyState s_states[kBatchMaxAwake];
  __shared__ BatchWorldState s_ws;
  __shared__ Colours colours;
  __shared__ int s_colour_list[kOverflowColor];  // graph colour of each non-empty colour range
  __shared__ int s_count, s_any;
  const int lane = threadIdx.x;
  if (blockIdx.x == 0 && lane < kPairTypes) b.collide_counts[lane] = 0;  // the next pair launch lists anew
  const int world = b.order[blockIdx.x];
  const BatchWorldDesc wd = b.worlds[world];
  const WorldRefs r = world_refs(b, wd, world);
  const DeviceWorld dg = world_view(b, wd, b.states + wd.awake_off);
  const DeviceWorld dl = world_view(b, wd, s_states);
  if (r.ws->error != kBatchOk) return;
  // ---- transitions in contact-id order (one lane walks the changed ids) on the world's
  // pool counters and colour bits copied to LDS, written back afterwards
  constexpr int kWsWords = (int)(sizeof(BatchWorldState) / sizeof(int32_t));
  static_assert(sizeof(BatchWorldState) % sizeof(int32_t) == 0, "world state moves as words");
  for (int i = lane; i < kWsWords; i += kBatchLanes)
    reinterpret_cast<int32_t*>(&s_ws)[i] = reinterpret_cast<const int32_t*>(r.ws)[i];
  const int bit_words = kOverflowColor * r.bits_words;
  const bool bits_in_lds = bit_words <= (int)(sizeof(s_states) / (sizeof(uint32_t)));
  uint32_t* s_bits = reinterpret_cast<uint32_t*>(s_states);
  if (bits_in_lds)
    for (int i = lane; i < bit_words; i += kBatchLanes) s_bits[i] = r.bits[i];
  wave_sync();
  WorldRefs rl = r;
  rl.ws = &s_ws;
  if (bits_in_lds) rl.bits = s_bits;
  const int next_id = s_ws.next_id;
  for (int base = 0; base < next_id; base += kBatchLanes) {
    const int id = base + lane;
    uint64_t changed = wave_ballot(id < next_id && r.change[id] != kChangeNone);
    if (lane == 0)
      while (changed) {
        const int k = base + (__ffsll((unsigned long long)changed) - 1);
        changed &= changed - 1;
        apply_change(rl, k, r.change[k]);
      }
    wave_sync();
  }
  if (lane == 0 && s_ws.error == kBatchOk) s_ws.steps += 1;
  wave_sync();
  for (int i = lane; i < kWsWords; i += kBatchLanes)
    reinterpret_cast<int32_t*>(r.ws)[i] = reinterpret_cast<const int32_t*>(&s_ws)[i];
  if (bits_in_lds)
    for (int i = lane; i < bit_words; i += kBatchLanes) r.bits[i] = s_bits[i];
  wave_sync();
  if (s_ws.error != kBatchOk) return;
  if (wd.n_awake == 0) return;

  // ---- solve order: colours 0..22 in graph order, then the overflow list
  if (lane == 0) {
    int n = 0, slots = 0;
    for (int ci = 0; ci < kOverflowColor; ++ci) {
      const int cnt = s_ws.color_count[ci];
      if (cnt > 0) {
        colours.begin[n] = slots;
        colours.end[n] = slots + cnt;
        s_colour_list[n] = ci;
        ++n;
      }
      slots += cnt;
    }
    colours.n = n;
    s_count = slots;
    s_any = 0;
  }
  for (int i = lane; i < wd.n_awake; i += kBatchLanes) s_states[i] = dg.states[i];
  wave_sync();
  const int n_constraints = s_count;
  const int n_ovf = s_ws.n_overflow;

  // ---- prepare (solver_stages.rs:543-589): each slot's contact id taken from its colour
  // list and the constraint prepared by the same lane
  {
    int points = 0;
    bool any_rest = false;
    int a = 0;
    for (int slot = lane; slot < n_constraints; slot += kBatchLanes) {
      while (slot >= colours.end[a]) ++a;
      const int32_t id = r.colors[s_colour_list[a] * r.color_cap + (slot - colours.begin[a])];
      dl.solve_ids[slot] = id;
      float restitution = 0.0f;
      points += prepare_contact(dl, blocked_cons(dl), slot, id, p, &restitution);
      any_rest = any_rest || restitution > 0.0f;
    }
    for (int k = lane; k < n_ovf; k += kBatchLanes) {
      const int32_t id = r.overflow_list[k];
      dl.overflow_ids[k] = id;
      any_rest = any_rest || r.meta[id].restitution > 0.0f;
      prepare_overflow_item(dl, k, p);
    }
    if (wave_ballot(any_rest) != 0 && lane == 0) s_any = 1;
    for (int off = 32; off > 0; off >>= 1) points += __shfl_xor(points, off, 64);
    if (lane == 0) {
      atomicAdd(&b.work[1], (unsigned long long)n_constraints);
      atomicAdd(&b.work[2], (unsigned long long)points);
      if (n_ovf) atomicAdd(&b.work[3], (unsigned long long)n_ovf);
    }
  }
  wave_sync();
  const bool any_restitution = s_any != 0;
  if (b.tape_step >= 0) {
    tape_record(b, world, wd, dl, r, n_constraints, n_ovf, any_restitution, colours, s_states);
    wave_sync();
  }
  for (int sub = 0; sub < p.sub_steps; ++sub) {
    for (int i = lane; i < wd.n_awake; i += kBatchLanes) integrate_velocities_item(dl, i, p);
    wave_sync();
    if (n_ovf > 0) {
      if (lane == 0) overflow_warm_all(dl, n_ovf);
      wave_sync();
    }
    colour_stage(dl, colours, lane, [&](int, const Constraint& c) { warm_start_record(dl, c); });
    if (n_ovf > 0) {
      if (lane == 0) overflow_solve_all(dl, n_ovf, p, 0);
      wave_sync();
    }
    colour_stage(dl, colours, lane, [&](int slot, const Constraint& c) { push_record(dl, slot, p, c); });
    for (int i = lane; i < wd.n_awake; i += kBatchLanes) integrate_positions_item(dl, i, p);
    wave_sync();
    if (n_ovf > 0) {
      if (lane == 0) overflow_solve_all(dl, n_ovf, p, 1);
      wave_sync();
    }
    colour_stage(dl, colours, lane, [&](int slot, const Constraint& c) { relax_record(dl, slot, p, c); });
  }
  if (any_restitution) {
    for (int it = 0; it < b.restitution_iterations; ++it) {
      if (n_ovf > 0) {
        if (lane == 0) overflow_solve_all(dl, n_ovf, p, 2);
        wave_sync();
      }
      for (int a = 0; a < colours.n; ++a) {
        for (int slot = colours.begin[a] + lane; slot < colours.end[a]; slot += kBatchLanes) {
          const int2 ix = blocked_cons(dl).indices(slot);
          restitution_item(dl, blocked_cons(dl), slot, p, ix.x, ix.y);
        }
        wave_sync();
      }
    }
  }
  for (int k = lane; k < n_ovf; k += kBatchLanes) overflow_store_item(dl, k);
  for (int slot = lane; slot < n_constraints; slot += kBatchLanes) store_item(dl, blocked_cons(dl), slot);
  for (int i = lane; i < wd.n_awake; i += kBatchLanes) finalize_item(dl, i, p);
  wave_sync();
  for (int i = lane; i < wd.n_awake; i += kBatchLanes) dg.states[i] = s_states[i];
}

// Continuous collision of the fast bodies the solve launch listed, each in its own world:
// without `bullets` the fast bodies against their world's static shapes, each followed by
// its world inverse inertia (solver_integration.rs:538-566); with `bullets` the bullets
// against every shape type of their world (solver_integration.rs:948-995). A body reads its
// world's static shapes (and, as a bullet, its other bodies) and writes only itself.
__global__ __launch_bounds__(kBatchLanes) void k_batch_continuous(BatchDevice b, StepParams p, int bullets) {
  const int n = b.counters[kCntFastList];
  int hits = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int g = b.fast_ids[k];
    const int world = b.awake_world[g];
    const BatchWorldDesc wd = b.worlds[world];
    const DeviceWorld d = world_view(b, wd, b.states + wd.awake_off);
    const int i = g - wd.awake_off;
    if (((d.sims[i].flags & bflag::kIsBullet) != 0) != (bullets != 0)) continue;
    const ContinuousArgs ca{b.candidates + wd.shape_off, wd.n_ccd_static, wd.n_ccd_kinematic, wd.n_ccd_dynamic,
                            b.proxies + wd.shape_off};
    const bool hit = solve_continuous_body(d, ca, i, p);
    if (!bullets) {
      BodySim& sim = d.sims[i];
      Mat3 rm = matrix_from_quat(sim.transform.q);
      sim.inv_inertia_world = mul_mm(mul_mm(rm, sim.inv_inertia_local), transpose(rm));
    }
    if (!hit) continue;
    hits += 1;
    // a body moved back to an impact leaves the differentiable rollout's model
    if (b.tape_step >= 0) b.tape.count[(size_t)b.tape_step * b.tape.n_worlds + world] = -1;
  }
  if (hits) atomicAdd(&b.counters[kCntToiHits], hits);
}

// Workgroups of the collide launch: as many as the device keeps resident at once.
int collide_grid(int device) {
  int per_cu = 0, cus = 0;
  BATCH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_batch_collide),
                                                           kBatchCollideBlock, 0));
  BATCH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  return std::max(1, per_cu) * std::max(1, cus);
}

}  // namespace

void launch_batch_step(const BatchDevice& b, const StepParams& p, int steps, hipStream_t s) {
  if (b.n_worlds <= 0) return;
  for (int k = 0; k < steps; ++k) {
    BatchDevice bk = b;
    bk.tape_step = b.tape_step >= 0 ? b.tape_step + k : -1;
    hipLaunchKernelGGL(k_batch_pairs, dim3(b.n_worlds), dim3(kBatchPairsBlock), 0, s, bk);
    hipLaunchKernelGGL(k_batch_collide, dim3(b.collide_blocks), dim3(kBatchCollideBlock), 0, s, bk);
    hipLaunchKernelGGL(k_batch_solve, dim3(b.n_worlds), dim3(kBatchLanes), 0, s, bk, p);
    if (b.enable_continuous) {
      const int blocks = std::max(1, std::min(b.n_worlds, 1024));
      hipLaunchKernelGGL(k_batch_continuous, dim3(blocks), dim3(kBatchLanes), 0, s, bk, p, 0);
      if (b.any_bullets) hipLaunchKernelGGL(k_batch_continuous, dim3(blocks), dim3(kBatchLanes), 0, s, bk, p, 1);
    }
  }
  if (b.tape_step >= 0) hipLaunchKernelGGL(k_tape_final, dim3(b.n_worlds), dim3(kBatchLanes), 0, s, b, b.tape_step + steps);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) throw std::runtime_error(std::string("batch step launch: ") + hipGetErrorString(err));
}

// ---------------------------------------------------------------------------------
// host side

BatchStepper::BatchStepper(int device) : device_(device) {
  BATCH_CHECK(hipSetDevice(device_));
  BATCH_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  BATCH_CHECK(hipEventCreate(&ev0_));
  BATCH_CHECK(hipEventCreate(&ev1_));
}

BatchStepper::~BatchStepper() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  drop_graph();
  free_tape();
  for (void* p : allocations_) (void)hipFree(p);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int32_t BatchStepper::add_world() {
  if (uploaded_) throw std::invalid_argument("worlds can only be added before the first step");
  scenes_.push_back(std::make_unique<HostScene>());
  scenes_.back()->isotropic_inertia_tolerance = options.isotropic_inertia_tolerance;
  return (int32_t)scenes_.size() - 1;
}

HostScene& BatchStepper::scene(int32_t w) {
  if (w < 0 || w >= (int32_t)scenes_.size()) throw std::invalid_argument("bad world index");
  if (uploaded_) throw std::invalid_argument("bodies and shapes can only be added before the first step");
  return *scenes_[w];
}

template <class T>
T* BatchStepper::alloc(size_t n, bool zero) {
  T* p = nullptr;
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  BATCH_CHECK(hipMalloc(&p, bytes));
  allocations_.push_back(p);
  allocation_bytes_.push_back(bytes);
  if (zero) BATCH_CHECK(hipMemsetAsync(p, 0, bytes, stream_));
  return p;
}

template <class T>
static T* upload_vec(hipStream_t s, T* dst, const std::vector<T>& v) {
  if (!v.empty()) {
    const hipError_t e = hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) throw std::runtime_error(std::string("batch upload: ") + hipGetErrorString(e));
  }
  return dst;
}

void BatchStepper::upload() {
  const int nw = (int)scenes_.size();
  desc_.assign(nw, BatchWorldDesc{});
  std::vector<BodySim> sims, statics;
  std::vector<BodyState> states;
  std::vector<BodyMeta> bmeta;
  std::vector<Shape> shapes;
  std::vector<Aabb> fat, aabbs;
  std::vector<int32_t> candidates, awake_world;
  int32_t any_bullets = 0;
  std::vector<uint8_t> moved;
  std::vector<Proxy> proxies;
  std::vector<BoxHull> hulls;
  int64_t contact_total = 0, overflow_total = 0, color_total = 0, bits_total = 0, pair_total = 0;
  for (int w = 0; w < nw; ++w) {
    const HostScene& sc = *scenes_[w];
    BatchWorldDesc& d = desc_[w];
    d.n_awake = (int32_t)sc.awake_host_.size();
    d.n_static = (int32_t)sc.static_host_.size();
    d.n_bodies = (int32_t)sc.bodies_.size();
    d.n_shapes = (int32_t)sc.shapes_.size();
    if (d.n_awake > kBatchMaxAwake) throw std::invalid_argument("a batched world holds at most 256 dynamic bodies");
    if (d.n_shapes > kBatchMaxShapes) throw std::invalid_argument("a batched world holds at most 4096 shapes");
    for (const Shape& s : sc.shapes_)
      if (s.kind == kShapeHull && s.hull < 0) throw std::invalid_argument("bad hull");
    d.awake_off = (int32_t)sims.size();
    d.static_off = (int32_t)statics.size();
    d.body_off = (int32_t)bmeta.size();
    d.shape_off = (int32_t)shapes.size();
    d.hull_off = (int32_t)hulls.size();
    sims.insert(sims.end(), sc.awake_host_.begin(), sc.awake_host_.end());
    awake_world.insert(awake_world.end(), sc.awake_host_.size(), w);
    for (const BodySim& s : sc.awake_host_)
      if (s.body_type == kBodyDynamic && (s.flags & bflag::kIsBullet)) any_bullets = 1;
    states.insert(states.end(), sc.states_host_.begin(), sc.states_host_.end());
    statics.insert(statics.end(), sc.static_host_.begin(), sc.static_host_.end());
    for (const HostScene::HBody& hb : sc.bodies_) bmeta.push_back(BodyMeta{hb.type, sc.encode(hb), hb.flags});
    shapes.insert(shapes.end(), sc.shapes_.begin(), sc.shapes_.end());
    for (int type = kBodyStatic; type <= kBodyDynamic; ++type) {
      int32_t n = 0;
      for (const HostScene::HShape& hs : sc.hshapes_)
        if (sc.bodies_[hs.body].type == type) {
          candidates.push_back(hs.id);
          n += 1;
        }
      (type == kBodyStatic ? d.n_ccd_static : type == kBodyKinematic ? d.n_ccd_kinematic : d.n_ccd_dynamic) = n;
    }
    for (const HostScene::HShape& hs : sc.hshapes_) {
      fat.push_back(hs.fat);
      aabbs.push_back(hs.aabb);
      moved.push_back(hs.proxy_moved ? 1 : 0);
      proxies.push_back(Proxy{hs.body, sc.bodies_[hs.body].type, hs.group,
                              (hs.speculative ? kProxySpeculative : 0) | (hs.sensor ? kProxySensor : 0), hs.category, hs.mask});
    }
    hulls.insert(hulls.end(), sc.hulls_.begin(), sc.hulls_.end());
    // capacities: a pile keeps a few contacts per shape; the overflow colour is rare
    d.contact_cap = std::max(64, 12 * d.n_shapes);
    d.overflow_cap = std::max(16, d.contact_cap / 4);
    d.color_cap = std::max(1, d.n_bodies);
    d.bits_words = (d.n_bodies + 31) / 32;
    d.contact_off = (int32_t)contact_total;
    d.overflow_off = (int32_t)overflow_total;
    d.color_off = (int32_t)color_total;
    d.bits_off = (int32_t)bits_total;
    d.pair_off = (int32_t)pair_total;
    // each world's records start on a block of kRecBlock slots (the blocked record layout)
    contact_total += (d.contact_cap + kRecBlock - 1) / kRecBlock * kRecBlock;
    overflow_total += d.overflow_cap;
    color_total += (int64_t)kOverflowColor * d.color_cap;
    bits_total += (int64_t)kOverflowColor * std::max(1, d.bits_words);
    pair_total += ((int64_t)d.n_shapes * d.n_shapes + 31) / 32;
    if (contact_total > INT32_MAX || color_total > INT32_MAX || pair_total > INT32_MAX)
      throw std::invalid_argument("batch too large for 32-bit offsets");
  }
  std::vector<int32_t> order(nw);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return desc_[a].n_awake > desc_[b].n_awake; });

  BatchDevice& b = d_;
  b.tape_step = -1;
  b.n_worlds = nw;
  b.enable_speculative = options.enable_speculative ? 1 : 0;
  b.restitution_iterations = options.restitution_iterations;
  b.enable_continuous = options.enable_continuous ? 1 : 0;
  b.any_bullets = any_bullets;
  b.sims = upload_vec(stream_, alloc<BodySim>(sims.size(), false), sims);
  b.states = upload_vec(stream_, alloc<BodyState>(states.size(), false), states);
  b.static_sims = upload_vec(stream_, alloc<BodySim>(statics.size(), false), statics);
  b.body_meta = upload_vec(stream_, alloc<BodyMeta>(bmeta.size(), false), bmeta);
  b.shapes = upload_vec(stream_, alloc<Shape>(shapes.size(), false), shapes);
  b.fat_aabbs = upload_vec(stream_, alloc<Aabb>(fat.size(), false), fat);
  b.shape_aabbs = upload_vec(stream_, alloc<Aabb>(aabbs.size(), false), aabbs);
  b.candidates = upload_vec(stream_, alloc<int32_t>(candidates.size(), false), candidates);
  b.fast_ids = alloc<int32_t>(std::max<size_t>(sims.size(), 1), false);
  b.awake_world = upload_vec(stream_, alloc<int32_t>(std::max<size_t>(awake_world.size(), 1), false), awake_world);
  b.shape_moved = upload_vec(stream_, alloc<uint8_t>(moved.size(), false), moved);
  b.proxies = upload_vec(stream_, alloc<Proxy>(proxies.size(), false), proxies);
  b.hulls = upload_vec(stream_, alloc<BoxHull>(hulls.size(), false), hulls);
  b.contacts = alloc<Contact>(contact_total, false);
  b.contact_change = alloc<uint8_t>(contact_total, true);
  b.contact_meta = alloc<ContactMeta>(contact_total, true);
  b.constraints = alloc<Constraint>(contact_total, false);
  b.solve_ids = alloc<int32_t>(contact_total, false);
  b.free_ids = alloc<int32_t>(contact_total, false);
  b.non_touching = alloc<int32_t>(contact_total, false);
  b.overflow = alloc<OverflowConstraint>(overflow_total, false);
  b.overflow_ids = alloc<int32_t>(overflow_total, false);
  b.overflow_list = alloc<int32_t>(overflow_total, false);
  b.color_lists = alloc<int32_t>(color_total, false);
  b.color_bits = alloc<uint32_t>(bits_total, true);
  b.pair_bits = alloc<uint32_t>(pair_total, true);
  b.worlds = upload_vec(stream_, alloc<BatchWorldDesc>(nw, false), desc_);
  b.wstate = alloc<BatchWorldState>(nw, true);
  b.order = upload_vec(stream_, alloc<int32_t>(nw, false), order);
  b.collide_type_cap = (int32_t)std::max<int64_t>(contact_total, 1);
  b.collide_items = alloc<int2>((size_t)kPairTypes * b.collide_type_cap, false);
  b.collide_counts = alloc<int32_t>(kPairTypes, true);
  resident_collide_blocks_ = collide_grid(device_);
  b.collide_blocks = tuning_.collide_blocks > 0 ? tuning_.collide_blocks : resident_collide_blocks_;
  b.counters = alloc<int32_t>(kCounterCount, true);
  b.work = alloc<unsigned long long>(4, true);
  BATCH_CHECK(hipStreamSynchronize(stream_));  // the host vectors are read by the copies
  total_awake_ = (int64_t)sims.size();
  total_contact_cap_ = contact_total;
  uploaded_ = true;
}

StepParams BatchStepper::step_params(float dt, int32_t sub_steps) const {
  StepParams p{};
  p.dt = dt;
  p.sub_steps = std::max(1, sub_steps);
  p.inv_dt = 1.0f / dt;
  p.h = dt / (float)p.sub_steps;
  p.inv_h = (float)p.sub_steps * p.inv_dt;
  const float contact_hz = minf(options.contact_hertz, 0.125f * p.inv_h);
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
  return p;
}

void BatchStepper::step(float dt, int32_t sub_steps, int32_t steps) {
  if (!(dt > 0.0f)) throw std::invalid_argument("the batched stepper needs a positive time step");
  if (steps <= 0) return;
  BATCH_CHECK(hipSetDevice(device_));
  if (!uploaded_) upload();
  const StepParams p = step_params(dt, sub_steps);
  if (tuning_.use_graph) {
    // capture before the timed region, so the host's capture does not count as device time
    const int chunk = std::min<int32_t>(steps, kBatchGraphSteps);
    if (steps >= chunk && (graph_exec_ == nullptr || graph_steps_ != chunk || graph_collide_blocks_ != d_.collide_blocks ||
                           std::memcmp(&graph_params_, &p, sizeof(StepParams)) != 0)) {
      drop_graph();
      hipGraph_t graph = nullptr;
      BATCH_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
      launch_batch_step(d_, p, chunk, stream_);
      BATCH_CHECK(hipStreamEndCapture(stream_, &graph));
      const hipError_t e = hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        graph_exec_ = nullptr;
        throw std::runtime_error(std::string("batch step graph: ") + hipGetErrorString(e));
      }
      graph_params_ = p;
      graph_steps_ = chunk;
      graph_collide_blocks_ = d_.collide_blocks;
    }
  }
  BATCH_CHECK(hipEventRecord(ev0_, stream_));
  launch_steps(p, steps);
  BATCH_CHECK(hipEventRecord(ev1_, stream_));
  BATCH_CHECK(hipEventSynchronize(ev1_));
  BATCH_CHECK(hipEventElapsedTime(&last_ms_, ev0_, ev1_));
  body_steps_ += total_awake_ * (int64_t)steps;
  device_newer_ = true;
}

// The steps of one call: whole chunks by the captured graph (when tuning_.use_graph and
// the graph matches this call's parameters), the rest by direct launches.
void BatchStepper::launch_steps(const StepParams& p, int32_t steps) {
  int32_t done = 0;
  if (tuning_.use_graph && graph_exec_ != nullptr && graph_collide_blocks_ == d_.collide_blocks &&
      std::memcmp(&graph_params_, &p, sizeof(StepParams)) == 0)
    for (; steps - done >= graph_steps_; done += graph_steps_) BATCH_CHECK(hipGraphLaunch(graph_exec_, stream_));
  if (done < steps) launch_batch_step(d_, p, steps - done, stream_);
}

void BatchStepper::drop_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  graph_exec_ = nullptr;
  graph_steps_ = 0;
}

void BatchStepper::set_tuning(const BatchTuning& t) {
  if (t.collide_blocks < 0) throw std::invalid_argument("collide_blocks must be >= 0");
  tuning_.collide_blocks = t.collide_blocks;
  tuning_.use_graph = t.use_graph != 0 ? 1 : 0;
  if (uploaded_) d_.collide_blocks = tuning_.collide_blocks > 0 ? tuning_.collide_blocks : resident_collide_blocks_;
}

std::vector<void*> BatchStepper::snapshot() {
  std::vector<void*> copies;
  try {
    for (size_t i = 0; i < allocations_.size(); ++i) {
      void* c = nullptr;
      BATCH_CHECK(hipMalloc(&c, allocation_bytes_[i]));
      copies.push_back(c);
      BATCH_CHECK(hipMemcpyAsync(c, allocations_[i], allocation_bytes_[i], hipMemcpyDeviceToDevice, stream_));
    }
    BATCH_CHECK(hipStreamSynchronize(stream_));
  } catch (...) {
    for (void* c : copies) (void)hipFree(c);
    throw;
  }
  return copies;
}

void BatchStepper::restore(const std::vector<void*>& copies) {
  for (size_t i = 0; i < allocations_.size(); ++i)
    BATCH_CHECK(hipMemcpyAsync(allocations_[i], copies[i], allocation_bytes_[i], hipMemcpyDeviceToDevice, stream_));
  BATCH_CHECK(hipStreamSynchronize(stream_));
  device_newer_ = true;
}

BatchTuning BatchStepper::autotune(float dt, int32_t sub_steps, int32_t steps) {
  if (!(dt > 0.0f) || steps <= 0) throw std::invalid_argument("autotune needs a positive time step and steps");
  BATCH_CHECK(hipSetDevice(device_));
  if (!uploaded_) upload();
  const BatchTuning start = tuning_;
  const float last_ms = last_ms_;
  const int64_t body_steps = body_steps_;
  const std::vector<void*> snap = snapshot();
  std::vector<uint64_t> first_hashes, hashes((size_t)world_count());
  BatchTuning best = start;
  double best_ms = 0.0, start_ms = 0.0;
  int candidates = 0;
  try {
    const int resident = resident_collide_blocks_;
    for (int use_graph = 0; use_graph <= 1; ++use_graph)
      for (int mult : {1, 2, 4, 8, 16}) {
        BatchTuning t;
        t.collide_blocks = resident * mult;
        t.use_graph = use_graph;
        set_tuning(t);
        restore(snap);
        step(dt, sub_steps, std::min<int32_t>(steps, kBatchGraphSteps));  // warm (and capture)
        restore(snap);
        step(dt, sub_steps, steps);
        const double ms = last_ms_ / steps;
        state_hashes(hashes.data());
        if (first_hashes.empty()) first_hashes = hashes;
        else if (hashes != first_hashes) throw std::runtime_error("batch autotune: a launch shape changed the results");
        ++candidates;
        if (t.collide_blocks == (start.collide_blocks > 0 ? start.collide_blocks : resident) && use_graph == start.use_graph)
          start_ms = ms;
        if (candidates == 1 || ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
  } catch (...) {
    set_tuning(start);
    restore(snap);
    for (void* c : snap) (void)hipFree(c);
    last_ms_ = last_ms;
    body_steps_ = body_steps;
    throw;
  }
  set_tuning(best);
  restore(snap);
  for (void* c : snap) (void)hipFree(c);
  last_ms_ = last_ms;
  body_steps_ = body_steps;
  tuning_.candidates = candidates;
  tuning_.step_ms = best_ms;
  tuning_.previous_ms = start_ms;
  return tuning_;
}

void BatchStepper::pull() {
  if (!device_newer_) return;
  std::vector<BodySim> sims((size_t)total_awake_);
  std::vector<BodyState> states((size_t)total_awake_);
  if (total_awake_ > 0) {
    BATCH_CHECK(hipMemcpyAsync(sims.data(), d_.sims, sims.size() * sizeof(BodySim), hipMemcpyDeviceToHost, stream_));
    BATCH_CHECK(hipMemcpyAsync(states.data(), d_.states, states.size() * sizeof(BodyState), hipMemcpyDeviceToHost, stream_));
  }
  BATCH_CHECK(hipStreamSynchronize(stream_));
  for (size_t w = 0; w < scenes_.size(); ++w) {
    HostScene& sc = *scenes_[w];
    const BatchWorldDesc& d = desc_[w];
    std::copy(sims.begin() + d.awake_off, sims.begin() + d.awake_off + d.n_awake, sc.awake_host_.begin());
    std::copy(states.begin() + d.awake_off, states.begin() + d.awake_off + d.n_awake, sc.states_host_.begin());
  }
  device_newer_ = false;
}

void BatchStepper::state_hashes(uint64_t* out) {
  pull();
  for (size_t w = 0; w < scenes_.size(); ++w) out[w] = scenes_[w]->state_hash();
}

void BatchStepper::get_bodies(int32_t w, double* pos, float* rot, float* v, float* wv) {
  if (w < 0 || w >= (int32_t)scenes_.size()) throw std::invalid_argument("bad world index");
  pull();
  scenes_[w]->get_bodies(pos, rot, v, wv);
}

void BatchStepper::work_counts(int64_t out[4]) {
  unsigned long long w[4] = {0, 0, 0, 0};
  if (uploaded_) BATCH_CHECK(hipMemcpy(w, d_.work, sizeof(w), hipMemcpyDeviceToHost));
  for (int i = 0; i < 4; ++i) out[i] = (int64_t)w[i];
}

int32_t BatchStepper::contact_slots(int32_t w) const {
  if (w < 0 || w >= (int32_t)scenes_.size()) throw std::invalid_argument("bad world index");
  return uploaded_ ? desc_[w].contact_cap : 0;
}

void BatchStepper::get_contact_colors(int32_t w, int32_t* color, int32_t* local, int32_t* alive) {
  const int32_t n = contact_slots(w);
  if (n == 0) return;
  std::vector<ContactMeta> meta((size_t)n);
  BATCH_CHECK(hipMemcpy(meta.data(), d_.contact_meta + desc_[w].contact_off, (size_t)n * sizeof(ContactMeta), hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < n; ++i) {
    color[i] = meta[i].alive ? meta[i].color : kNullIndex;
    local[i] = meta[i].alive ? meta[i].local : kNullIndex;
    alive[i] = meta[i].alive;
  }
}

void BatchStepper::get_shape_bounds(int32_t w, float* fat, uint8_t* moved) {
  if (w < 0 || w >= (int32_t)scenes_.size()) throw std::invalid_argument("bad world index");
  const HostScene& sc = *scenes_[w];
  const size_t n = sc.hshapes_.size();
  std::vector<Aabb> boxes(n);
  std::vector<uint8_t> flags(n);
  if (uploaded_ && n) {
    BATCH_CHECK(hipMemcpy(boxes.data(), d_.fat_aabbs + desc_[w].shape_off, n * sizeof(Aabb), hipMemcpyDeviceToHost));
    BATCH_CHECK(hipMemcpy(flags.data(), d_.shape_moved + desc_[w].shape_off, n, hipMemcpyDeviceToHost));
  } else {
    for (size_t i = 0; i < n; ++i) {
      boxes[i] = sc.hshapes_[i].fat;
      flags[i] = sc.hshapes_[i].proxy_moved ? 1 : 0;
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const Aabb& b = boxes[i];
    const float v[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
    for (int k = 0; k < 6; ++k) fat[6 * i + k] = v[k];
    moved[i] = flags[i];
  }
}

void BatchStepper::get_contact_shapes(int32_t w, int32_t* shape_a, int32_t* shape_b) {
  const int32_t n = contact_slots(w);
  if (n == 0) return;
  std::vector<ContactMeta> meta((size_t)n);
  BATCH_CHECK(hipMemcpy(meta.data(), d_.contact_meta + desc_[w].contact_off, (size_t)n * sizeof(ContactMeta), hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < n; ++i) {
    shape_a[i] = meta[i].alive ? meta[i].shape_a : kNullIndex;
    shape_b[i] = meta[i].alive ? meta[i].shape_b : kNullIndex;
  }
}

int32_t BatchStepper::error_worlds(int32_t* first_error) {
  *first_error = 0;
  if (!uploaded_) return 0;
  std::vector<BatchWorldState> ws(scenes_.size());
  BATCH_CHECK(hipMemcpy(ws.data(), d_.wstate, ws.size() * sizeof(BatchWorldState), hipMemcpyDeviceToHost));
  int32_t n = 0;
  for (const BatchWorldState& s : ws)
    if (s.error != kBatchOk) {
      if (n == 0) *first_error = s.error;
      ++n;
    }
  return n;
}

}  // namespace nsl
