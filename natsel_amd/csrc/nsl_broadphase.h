// Device pair finder: the new shape pairs of one step (contact_creation.rs:901-1042).
//
// The reference queries a dynamic AABB tree with each moved proxy; the resulting pair set
// does not depend on the tree: every pair of shapes whose fat AABBs overlap, with at least
// one proxy moved, at least one dynamic body, different bodies, a passing collision filter
// (shape_runtime.rs:285-292) and no existing contact. Here that set is found with a sort
// along x and a load-balanced sweep over the x-overlapping pairs that hold a moved proxy,
// and returned as sorted pair keys, which is the order the reference creates contacts in.
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

  // New pair keys in ascending order. Clears the moved flags and the moved counter.
  // Synchronises the stream (once when there is no new pair, twice otherwise).
  void find(const DeviceWorld& d, const Proxy* proxies, int n_shapes, std::vector<uint64_t>& keys);

  // Candidate pairs tested by the last find (for profiling).
  int64_t last_tests() const { return host_[1]; }

 private:
  template <class T> void ensure(T*& ptr, size_t& cap, size_t n);

  hipStream_t stream_;
  uint32_t* keys_in_ = nullptr;
  uint32_t* keys_out_ = nullptr;
  int32_t* ids_in_ = nullptr;
  int32_t* ids_out_ = nullptr;
  int64_t* counts_ = nullptr;
  int64_t* offsets_ = nullptr;
  int32_t* mv_ = nullptr;         // per sorted position: moved (0/1)
  int32_t* mprefix_ = nullptr;    // exclusive sum of mv_
  int32_t* moved_pos_ = nullptr;  // the moved positions in order
  size_t cap_mv_ = 0, cap_mprefix_ = 0, cap_moved_pos_ = 0;
  size_t cap_keys_in_ = 0, cap_keys_out_ = 0, cap_ids_in_ = 0, cap_ids_out_ = 0, cap_counts_ = 0,
         cap_offsets_ = 0;
  uint64_t* pairs_ = nullptr;        // found keys, unsorted
  uint64_t* pairs_sorted_ = nullptr;
  size_t cap_pairs_ = 0, cap_pairs_sorted_ = 0;
  int64_t* found_ = nullptr;         // [0] pairs found, [1] candidate tests
  uint8_t* temp_ = nullptr;
  size_t cap_temp_ = 0;
  int64_t* host_ = nullptr;          // pinned: found_ copy
  uint64_t* host_keys_ = nullptr;    // pinned
  size_t cap_host_keys_ = 0;
};

}  // namespace nsl
