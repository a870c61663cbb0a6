// Device pair finder (see nsl_broadphase.h).
#include "nsl_broadphase.h"

#include <hipcub/hipcub.hpp>

#include <stdexcept>
#include <string>

namespace nsl {

#define BP_CHECK(expr)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_));     \
  } while (0)

namespace {

constexpr int kBlock = 256;
constexpr int kTestBlocks = 1024;

// Monotone map from float to uint32 (-0 and +0 map together, so key order is float order).
__device__ __forceinline__ uint32_t order_key(float f) {
  uint32_t u = __float_as_uint(f);
  if (u == 0x80000000u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_bp_keys(int n, const Aabb* fat, uint32_t* keys, int32_t* ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = order_key(fat[i].lo.x);
  ids[i] = i;
}

// mv[p]: 1 when the shape at sorted position p is marked moved; mv[n] = 0 closes the scan.
__global__ void k_bp_moved(int n, const int32_t* ids, const uint8_t* moved, int32_t* mv) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) mv[p] = moved[ids[p]] ? 1 : 0;
  else if (p == n) mv[n] = 0;
}

// The later shapes that overlap sorted position p along x are positions p+1 .. end-1, end
// being the first whose lower x exceeds p's upper x. A pair needs a moved proxy: a moved p
// takes all of them (counts[p] = end - p - 1); any other p takes only the moved ones, the
// entries mprefix[p+1] .. mprefix[end]-1 of moved_pos (mprefix: exclusive count of moved
// positions). Every x-overlapping pair with a moved proxy is counted exactly once, at its
// lower position. counts[n] = 0 closes the scan.
__global__ void k_bp_count(int n, const int32_t* ids, const uint32_t* keys, const Aabb* fat, const int32_t* mprefix,
                           int32_t* moved_pos, int64_t* counts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > n) return;
  if (p == n) {
    counts[n] = 0;
    return;
  }
  const bool is_moved = mprefix[p + 1] != mprefix[p];
  if (is_moved) moved_pos[mprefix[p]] = p;
  const uint32_t hi = order_key(fat[ids[p]].hi.x);
  int lo = p + 1, up = n;  // first q in [p+1, n) with keys[q] > hi
  while (lo < up) {
    const int mid = lo + ((up - lo) >> 1);
    if (keys[mid] <= hi) lo = mid + 1;
    else up = mid;
  }
  counts[p] = is_moved ? (int64_t)(lo - p - 1) : (int64_t)(mprefix[lo] - mprefix[p + 1]);
}

// One candidate (p, q) per index t < offsets[n]; p is the last position with offsets[p] <= t,
// q its (t - offsets[p])-th later overlapping position (any one when p moved, else a moved
// one).
__global__ void k_bp_test(int n, const int64_t* offsets, const int32_t* ids, const int32_t* mprefix, const int32_t* moved_pos,
                          const Aabb* fat, const Proxy* proxies, PairSet set, PairSet blocked, uint64_t* out, int64_t cap,
                          int64_t* found) {
  const int64_t total = offsets[n];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t tests = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
      const int mid = lo + ((hi - lo) >> 1);
      if (offsets[mid] <= t) lo = mid;
      else hi = mid;
    }
    const int p = lo;
    const int j = (int)(t - offsets[p]);
    const int q = mprefix[p + 1] != mprefix[p] ? p + 1 + j : moved_pos[mprefix[p + 1] + j];
    const int32_t a = ids[p], b = ids[q];
    ++tests;
    const Proxy pa = proxies[a], pb = proxies[b];
    if (pa.body_type != kBodyDynamic && pb.body_type != kBodyDynamic) continue;
    if (pa.body == pb.body) continue;
    if ((pa.flags | pb.flags) & kProxySensor) continue;  // contact_creation.rs:187
    if (!aabb_overlaps(fat[a], fat[b])) continue;
    if (pa.group == pb.group && pa.group != 0) {  // shape_runtime.rs:285-292
      if (pa.group < 0) continue;
    } else if (!((pa.mask & pb.category) != 0 && (pa.category & pb.mask) != 0)) {
      continue;
    }
    // contact_creation.rs:110-143: a joint without collide_connected between the bodies
    if (blocked.slots != nullptr && pairset_contains(blocked, pair_key(pa.body, pb.body))) continue;
    const uint64_t key = pair_key(a, b);
    if (pairset_contains(set, key)) continue;
    const int64_t slot = (int64_t)atomicAdd((unsigned long long*)&found[0], 1ull);
    if (slot < cap) out[slot] = key;
  }
  if (tests) atomicAdd((unsigned long long*)&found[1], (unsigned long long)tests);
}

}  // namespace

PairFinder::PairFinder(hipStream_t stream) : stream_(stream) {
  BP_CHECK(hipMalloc(&found_, 2 * sizeof(int64_t)));
  BP_CHECK(hipHostMalloc(&host_, 2 * sizeof(int64_t), hipHostMallocDefault));
  host_[0] = host_[1] = 0;
}

PairFinder::~PairFinder() {
  void* dev[] = {keys_in_, keys_out_, ids_in_, ids_out_, counts_, offsets_, pairs_, pairs_sorted_, found_, temp_, mv_,
                 mprefix_, moved_pos_};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (host_) (void)hipHostFree(host_);
  if (host_keys_) (void)hipHostFree(host_keys_);
}

template <class T>
void PairFinder::ensure(T*& ptr, size_t& cap, size_t n) {
  if (ptr && n <= cap) return;
  const size_t new_cap = std::max<size_t>(n + n / 2, 256);
  if (ptr) {
    BP_CHECK(hipStreamSynchronize(stream_));
    BP_CHECK(hipFree(ptr));
  }
  BP_CHECK(hipMalloc(&ptr, new_cap * sizeof(T)));
  cap = new_cap;
}

void PairFinder::find(const DeviceWorld& d, const Proxy* proxies, int n, std::vector<uint64_t>& keys) {
  keys.clear();
  host_[0] = host_[1] = 0;
  if (n < 2) {
    if (n > 0) BP_CHECK(hipMemsetAsync(d.shape_moved, 0, n, stream_));
    BP_CHECK(hipMemsetAsync(d.counters + 6, 0, sizeof(int32_t), stream_));
    return;
  }
  ensure(keys_in_, cap_keys_in_, n);
  ensure(keys_out_, cap_keys_out_, n);
  ensure(ids_in_, cap_ids_in_, n);
  ensure(ids_out_, cap_ids_out_, n);
  ensure(counts_, cap_counts_, n + 1);
  ensure(offsets_, cap_offsets_, n + 1);
  ensure(mv_, cap_mv_, n + 1);
  ensure(mprefix_, cap_mprefix_, n + 1);
  ensure(moved_pos_, cap_moved_pos_, n);
  if (!pairs_) ensure(pairs_, cap_pairs_, 4096);

  size_t sort_bytes = 0, scan_bytes = 0, mscan_bytes = 0;
  BP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_in_, keys_out_, ids_in_, ids_out_, n, 0, 32, stream_));
  BP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts_, offsets_, n + 1, stream_));
  BP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, mscan_bytes, mv_, mprefix_, n + 1, stream_));
  ensure(temp_, cap_temp_, std::max(std::max(sort_bytes, scan_bytes), mscan_bytes));

  const int nb = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_bp_keys, dim3(nb), dim3(kBlock), 0, stream_, n, d.fat_aabbs, keys_in_, ids_in_);
  sort_bytes = cap_temp_;
  BP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp_, sort_bytes, keys_in_, keys_out_, ids_in_, ids_out_, n, 0, 32, stream_));
  hipLaunchKernelGGL(k_bp_moved, dim3((n + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n, ids_out_, d.shape_moved,
                     mv_);
  mscan_bytes = cap_temp_;
  BP_CHECK(hipcub::DeviceScan::ExclusiveSum(temp_, mscan_bytes, mv_, mprefix_, n + 1, stream_));
  hipLaunchKernelGGL(k_bp_count, dim3((n + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n, ids_out_, keys_out_,
                     d.fat_aabbs, mprefix_, moved_pos_, counts_);
  scan_bytes = cap_temp_;
  BP_CHECK(hipcub::DeviceScan::ExclusiveSum(temp_, scan_bytes, counts_, offsets_, n + 1, stream_));

  for (int attempt = 0; attempt < 2; ++attempt) {
    BP_CHECK(hipMemsetAsync(found_, 0, 2 * sizeof(int64_t), stream_));
    hipLaunchKernelGGL(k_bp_test, dim3(kTestBlocks), dim3(kBlock), 0, stream_, n, offsets_, ids_out_, mprefix_, moved_pos_,
                       d.fat_aabbs, proxies, d.pairs, d.joint_pairs, pairs_, (int64_t)cap_pairs_, found_);
    BP_CHECK(hipMemcpyAsync(host_, found_, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    BP_CHECK(hipStreamSynchronize(stream_));
    if (host_[0] <= (int64_t)cap_pairs_) break;
    ensure(pairs_, cap_pairs_, (size_t)host_[0]);  // too many for the buffer: grow and test again
  }
  const int64_t found = host_[0];
  BP_CHECK(hipMemsetAsync(d.shape_moved, 0, n, stream_));
  BP_CHECK(hipMemsetAsync(d.counters + 6, 0, sizeof(int32_t), stream_));
  if (found == 0) return;
  if (found > (int64_t)cap_pairs_ || found > (int64_t)INT32_MAX) throw std::runtime_error("pair finder: too many new pairs");

  const int m = (int)found;
  ensure(pairs_sorted_, cap_pairs_sorted_, m);
  size_t key_bytes = 0;
  // Full 64-bit range: a begin_bit of 20 (the keys' zero low bits) gave unsorted output for
  // more than 1024 keys with this hipCUB.
  BP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, key_bytes, pairs_, pairs_sorted_, m, 0, 64, stream_));
  ensure(temp_, cap_temp_, key_bytes);
  key_bytes = cap_temp_;
  BP_CHECK(hipcub::DeviceRadixSort::SortKeys(temp_, key_bytes, pairs_, pairs_sorted_, m, 0, 64, stream_));
  if (!host_keys_ || (size_t)m > cap_host_keys_) {
    if (host_keys_) BP_CHECK(hipHostFree(host_keys_));
    cap_host_keys_ = std::max<size_t>((size_t)m * 2, 4096);
    BP_CHECK(hipHostMalloc(&host_keys_, cap_host_keys_ * sizeof(uint64_t), hipHostMallocDefault));
  }
  BP_CHECK(hipMemcpyAsync(host_keys_, pairs_sorted_, (size_t)m * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
  BP_CHECK(hipStreamSynchronize(stream_));
  keys.assign(host_keys_, host_keys_ + m);
  for (int i = 1; i < m; ++i)
    if (keys[i] <= keys[i - 1]) throw std::runtime_error("pair finder: keys out of order at " + std::to_string(i));
}

}  // namespace nsl
