// Device pair finder: the new shape pairs of one step (contact_creation.rs:901-1042).
//
// The reference queries a dynamic AABB tree with each moved proxy; the resulting pair set
// does not depend on the tree: every pair of shapes whose fat AABBs overlap, with at least
// one proxy moved, at least one dynamic body, different bodies, a passing collision filter
// (shape_runtime.rs:285-292) and no existing contact. It is returned as sorted pair keys,
// which is the order the reference creates contacts in. Here that set is found on a uniform
// grid (nsl_broadphase.hip): each proxy in the cells its fat AABB spans, pairs tested within
// a cell and reported in one cell only, proxies larger than two cells per axis tested against
// every proxy.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "nsl_device.h"

namespace nsl {

// Per shape, what the pair filter needs.
struct Proxy {
  int32_t body;
  int32_t body_type;
  int32_t group;
  int32_t flags;  // kProxy*
  uint64_t category;
  uint64_t mask;
};
constexpr int32_t kProxySpeculative = 0x1;  // the shape's enable_speculative_contact
constexpr int32_t kProxySensor = 0x2;       // a sensor: no contacts, no continuous collision

class PairFinder {
 public:
  explicit PairFinder(hipStream_t stream);
  ~PairFinder();
  PairFinder(const PairFinder&) = delete;
  PairFinder& operator=(const PairFinder&) = delete;

  // New pair keys in ascending order. Clears the moved flags and the moved counter. Few
  // moved proxies are tested directly against every proxy, more through the grid; the keys
  // do not depend on the path. `moved_hint`: the number of moved proxies if the caller
  // knows it (it picks the path without a round trip; a wrong hint costs time, not
  // results), else -1. Synchronises the stream once to three times, and once more when
  // there are new pairs.
  void find(const DeviceWorld& d, const Proxy* proxies, int n_shapes, std::vector<uint64_t>& keys,
            int64_t moved_hint = -1);

  // Candidate pairs tested by the last find (for profiling).
  int64_t last_tests() const { return host_[1]; }

 private:
  template <class T> void ensure(T*& ptr, size_t& cap, size_t n);

  hipStream_t stream_;
  int32_t* counts_ = nullptr;      // grid entries per proxy
  int32_t* offsets_ = nullptr;     // exclusive sum of counts_
  int32_t* large_ = nullptr;       // proxies tested against every proxy (or the moved ones, on the direct path)
  void* grid_ = nullptr;           // cell size and the extent statistics it comes from
  uint32_t* ekeys_in_ = nullptr;   // grid entries: cell hash
  uint32_t* ekeys_out_ = nullptr;
  uint32_t* evals_in_ = nullptr;   // grid entries: (proxy << 3) | corner offset
  uint32_t* evals_out_ = nullptr;
  size_t cap_counts_ = 0, cap_offsets_ = 0, cap_large_ = 0, cap_ekeys_in_ = 0, cap_ekeys_out_ = 0, cap_evals_in_ = 0,
         cap_evals_out_ = 0;
  uint64_t* pairs_ = nullptr;        // found keys, unsorted
  uint64_t* pairs_sorted_ = nullptr;
  size_t cap_pairs_ = 0, cap_pairs_sorted_ = 0;
  int64_t* found_ = nullptr;         // [0] pairs found, [1] candidate tests
  uint8_t* temp_ = nullptr;
  size_t cap_temp_ = 0;
  int64_t* host_ = nullptr;          // pinned: found_ copy
  int32_t* host_entries_ = nullptr;  // pinned: the moved count, then the grid's entry count
  uint64_t* host_keys_ = nullptr;    // pinned
  size_t cap_host_keys_ = 0;
};

}  // namespace nsl
