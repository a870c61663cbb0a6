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
constexpr int kTestBlocks = 1024;  // grid-stride blocks of the large-proxy and direct tests
// At most this many moved-times-all tests: the moved proxies are tested directly instead of
// through the grid (the grid's build costs more than the tests it saves).
constexpr int64_t kDirectTests = int64_t(1) << 24;

// The grid: a cell size c (device memory: c, 1/c, and the extent statistics it is chosen
// from). A proxy whose fat AABB spans at most two cells per axis is entered in each of them
// (at most 8); a larger one (a wall, the ground, an unusually large body) is tested against
// every proxy instead. A pair of grid proxies is tested in every cell both are entered in
// and reported in one: the cell of the lower corner of their boxes' intersection, which both
// boxes contain. Every overlapping pair is thus tested at least once; the reported set is
// the same whatever the cell size.
struct GridStats {
  float sum;     // extents of the proxies of non-static bodies
  float count;
  float max_ext;
  float cell;    // chosen cell size
  float inv_cell;
  int32_t n_large;
};

__device__ __forceinline__ float extent_of(const Aabb& b) {
  return fmaxf(fmaxf(b.hi.x - b.lo.x, b.hi.y - b.lo.y), b.hi.z - b.lo.z);
}
__device__ __forceinline__ bool valid_box(const Aabb& b) {
  return b.lo.x <= b.hi.x && b.lo.y <= b.hi.y && b.lo.z <= b.hi.z && isfinite(b.lo.x) && isfinite(b.lo.y) &&
         isfinite(b.lo.z) && isfinite(b.hi.x) && isfinite(b.hi.y) && isfinite(b.hi.z);
}
__device__ __forceinline__ int cell_of(float x, float inv_cell) { return (int)floorf(x * inv_cell); }
__device__ __forceinline__ uint32_t cell_hash(int x, int y, int z) {
  return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}

// Extent statistics of the proxies of moving bodies (atomics in any order: the cell size
// only changes the work, never the pairs).
__global__ void k_grid_extent(int n, const Aabb* fat, const Proxy* proxies, GridStats* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.0f, c = 0.0f;
  if (i < n && proxies[i].body_type != kBodyStatic && valid_box(fat[i])) {
    e = extent_of(fat[i]);
    c = 1.0f;
  }
  float m = e;
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    c += __shfl_xor(c, off, 64);
    m = fmaxf(m, __shfl_xor(m, off, 64));
  }
  // the workgroup's wavefronts combined in LDS: one set of atomics per workgroup (one per
  // wavefront serialised thousands of atomics on three words)
  __shared__ float s_part[3][kBlock / 64];
  const int wave = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_part[0][wave] = e;
    s_part[1][wave] = c;
    s_part[2][wave] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / 64; ++w) {
      e += s_part[0][w];
      c += s_part[1][w];
      m = fmaxf(m, s_part[2][w]);
    }
    if (c > 0.0f) {
      atomicAdd(&g->sum, e);
      atomicAdd(&g->count, c);
      atomicMax(reinterpret_cast<unsigned int*>(&g->max_ext), __float_as_uint(m));  // non-negative: bit order is value order
    }
  }
}

// Cell size: the largest moving proxy, at most four times the mean (larger ones go to the
// brute-force list), and never below a centimetre.
__global__ void k_grid_cell(GridStats* g) {
  const float mean = g->count > 0.0f ? g->sum / g->count : 1.0f;
  float c = fminf(g->max_ext, 4.0f * mean) * 1.001f;
  if (!(c > 0.01f)) c = fmaxf(0.01f, 4.0f * mean);
  g->cell = c;
  g->inv_cell = 1.0f / c;
  g->n_large = 0;
}

// Cells per proxy (counts[n] = 0 closes the scan); proxies spanning more than two cells on
// an axis go to the large list.
__global__ void k_grid_count(int n, const Aabb* fat, GridStats* g, int32_t* counts, int32_t* large) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    counts[n] = 0;
    return;
  }
  const Aabb b = fat[i];
  int cells = 0;
  if (valid_box(b)) {
    const float ic = g->inv_cell;
    const int nx = cell_of(b.hi.x, ic) - cell_of(b.lo.x, ic) + 1;
    const int ny = cell_of(b.hi.y, ic) - cell_of(b.lo.y, ic) + 1;
    const int nz = cell_of(b.hi.z, ic) - cell_of(b.lo.z, ic) + 1;
    if (nx <= 2 && ny <= 2 && nz <= 2) {
      cells = nx * ny * nz;
    } else {
      large[atomicAdd(&g->n_large, 1)] = i;
    }
  }
  counts[i] = cells;
}

// One entry per (proxy, cell): key the cell's hash, value (proxy << 3) | corner offset.
__global__ void k_grid_emit(int n, const Aabb* fat, const GridStats* g, const int32_t* offsets, uint32_t* keys,
                            uint32_t* vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = offsets[i];
  if (offsets[i + 1] == o) return;
  const Aabb b = fat[i];
  const float ic = g->inv_cell;
  const int x0 = cell_of(b.lo.x, ic), y0 = cell_of(b.lo.y, ic), z0 = cell_of(b.lo.z, ic);
  const int nx = cell_of(b.hi.x, ic) - x0 + 1, ny = cell_of(b.hi.y, ic) - y0 + 1, nz = cell_of(b.hi.z, ic) - z0 + 1;
  for (int dz = 0; dz < nz; ++dz)
    for (int dy = 0; dy < ny; ++dy)
      for (int dx = 0; dx < nx; ++dx) {
        keys[o] = cell_hash(x0 + dx, y0 + dy, z0 + dz);
        vals[o] = ((uint32_t)i << 3) | (uint32_t)(dx | (dy << 1) | (dz << 2));
        ++o;
      }
}

// The pair filter of contact_creation.rs:170-219 for two proxies whose boxes overlap and one
// of which moved; appends the key of a new pair.
__device__ __forceinline__ void offer_pair(int32_t a, int32_t b, const Proxy* proxies, PairSet set, PairSet blocked,
                                           uint64_t* out, int64_t cap, int64_t* found) {
  const Proxy pa = proxies[a], pb = proxies[b];
  if (pa.body_type != kBodyDynamic && pb.body_type != kBodyDynamic) return;
  if (pa.body == pb.body) return;
  if ((pa.flags | pb.flags) & kProxySensor) return;  // contact_creation.rs:187
  if (pa.group == pb.group && pa.group != 0) {        // shape_runtime.rs:285-292
    if (pa.group < 0) return;
  } else if (!((pa.mask & pb.category) != 0 && (pa.category & pb.mask) != 0)) {
    return;
  }
  // contact_creation.rs:110-143: a joint without collide_connected between the bodies
  if (blocked.slots != nullptr && pairset_contains(blocked, pair_key(pa.body, pb.body))) return;
  const uint64_t key = pair_key(a, b);
  if (pairset_contains(set, key)) return;
  const int64_t slot = (int64_t)atomicAdd((unsigned long long*)&found[0], 1ull);
  if (slot < cap) out[slot] = key;
}

// Each sorted entry against the later entries of its run (same hash): pairs in the same cell,
// one of them moved, boxes overlapping, reported in their home cell.
__global__ void k_grid_test(int m, const uint32_t* keys, const uint32_t* vals, const Aabb* fat, const uint8_t* moved,
                            const GridStats* g, const Proxy* proxies, PairSet set, PairSet blocked, uint64_t* out, int64_t cap,
                            int64_t* found) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const uint32_t key = keys[s];
  const float ic = g->inv_cell;
  const uint32_t vs = vals[s];
  const int32_t a = (int32_t)(vs >> 3);
  const Aabb ba = fat[a];
  const bool ma = moved[a] != 0;
  const int cx = cell_of(ba.lo.x, ic) + (int)(vs & 1), cy = cell_of(ba.lo.y, ic) + (int)((vs >> 1) & 1),
            cz = cell_of(ba.lo.z, ic) + (int)((vs >> 2) & 1);
  int64_t tests = 0;
  for (int t = s + 1; t < m && keys[t] == key; ++t) {
    const uint32_t vt = vals[t];
    const int32_t b = (int32_t)(vt >> 3);
    if (!ma && moved[b] == 0) continue;
    const Aabb bb = fat[b];
    // the same cell (the hash may join cells)
    if (cell_of(bb.lo.x, ic) + (int)(vt & 1) != cx || cell_of(bb.lo.y, ic) + (int)((vt >> 1) & 1) != cy ||
        cell_of(bb.lo.z, ic) + (int)((vt >> 2) & 1) != cz)
      continue;
    ++tests;
    if (!aabb_overlaps(ba, bb)) continue;
    // home cell: the lower corner of the intersection
    if (cell_of(fmaxf(ba.lo.x, bb.lo.x), ic) != cx || cell_of(fmaxf(ba.lo.y, bb.lo.y), ic) != cy ||
        cell_of(fmaxf(ba.lo.z, bb.lo.z), ic) != cz)
      continue;
    offer_pair(a, b, proxies, set, blocked, out, cap, found);
  }
  if (tests) atomicAdd((unsigned long long*)&found[1], (unsigned long long)tests);
}

// Few moved proxies: the moved ones listed (any order) and each tested against every proxy
// (a pair of moved ones once, from the lower id).
__global__ void k_list_moved(int n, const uint8_t* moved, int32_t* list, int32_t* count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && moved[i] != 0) list[atomicAdd(count, 1)] = i;
}
__global__ void k_moved_all(int n, const int32_t* list, const int32_t* count, const Aabb* fat, const uint8_t* moved,
                            const Proxy* proxies, PairSet set, PairSet blocked, uint64_t* out, int64_t cap, int64_t* found) {
  const int64_t total = (int64_t)*count * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t tests = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int32_t a = list[t / n], b = (int32_t)(t % n);
    if (a == b || (moved[b] != 0 && b < a)) continue;
    const Aabb ba = fat[a], bb = fat[b];
    if (!valid_box(ba) || !valid_box(bb)) continue;
    ++tests;
    if (!aabb_overlaps(ba, bb)) continue;
    offer_pair(a, b, proxies, set, blocked, out, cap, found);
  }
  if (tests) atomicAdd((unsigned long long*)&found[1], (unsigned long long)tests);
}

// Each large proxy against every other proxy (two large ones once, from the lower id).
__global__ void k_grid_large(int n, const int32_t* large, const Aabb* fat, const uint8_t* moved, const GridStats* g,
                             const Proxy* proxies, PairSet set, PairSet blocked, uint64_t* out, int64_t cap, int64_t* found) {
  const int64_t total = (int64_t)g->n_large * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float ic = g->inv_cell;
  int64_t tests = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int32_t a = large[t / n], b = (int32_t)(t % n);
    if (a == b || (moved[a] == 0 && moved[b] == 0)) continue;
    const Aabb bb = fat[b];
    if (!valid_box(bb)) continue;
    const bool b_large = cell_of(bb.hi.x, ic) - cell_of(bb.lo.x, ic) > 1 || cell_of(bb.hi.y, ic) - cell_of(bb.lo.y, ic) > 1 ||
                         cell_of(bb.hi.z, ic) - cell_of(bb.lo.z, ic) > 1;
    if (b_large && b < a) continue;
    ++tests;
    if (!aabb_overlaps(fat[a], bb)) continue;
    offer_pair(a, b, proxies, set, blocked, out, cap, found);
  }
  if (tests) atomicAdd((unsigned long long*)&found[1], (unsigned long long)tests);
}

}  // namespace

PairFinder::PairFinder(hipStream_t stream) : stream_(stream) {
  BP_CHECK(hipMalloc(&found_, 2 * sizeof(int64_t)));
  BP_CHECK(hipHostMalloc(&host_, 2 * sizeof(int64_t), hipHostMallocDefault));
  BP_CHECK(hipHostMalloc(&host_entries_, sizeof(int32_t), hipHostMallocDefault));
  host_[0] = host_[1] = 0;
}

PairFinder::~PairFinder() {
  void* dev[] = {ekeys_in_, ekeys_out_, evals_in_, evals_out_, counts_, offsets_, large_, grid_, pairs_, pairs_sorted_, found_,
                 temp_};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (host_) (void)hipHostFree(host_);
  if (host_entries_) (void)hipHostFree(host_entries_);
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

void PairFinder::find(const DeviceWorld& d, const Proxy* proxies, int n, std::vector<uint64_t>& keys, int64_t moved_hint) {
  keys.clear();
  host_[0] = host_[1] = 0;
  if (n < 2) {
    if (n > 0) BP_CHECK(hipMemsetAsync(d.shape_moved, 0, n, stream_));
    BP_CHECK(hipMemsetAsync(d.counters + 6, 0, sizeof(int32_t), stream_));
    return;
  }
  ensure(counts_, cap_counts_, n + 1);
  ensure(offsets_, cap_offsets_, n + 1);
  ensure(large_, cap_large_, n);
  if (!grid_) BP_CHECK(hipMalloc(&grid_, sizeof(GridStats)));
  if (!pairs_) ensure(pairs_, cap_pairs_, 4096);
  const int nb = (n + kBlock - 1) / kBlock, nb1 = (n + 1 + kBlock - 1) / kBlock;
  // the moved proxies, listed and counted
  BP_CHECK(hipMemsetAsync(counts_ + n, 0, sizeof(int32_t), stream_));
  hipLaunchKernelGGL(k_list_moved, dim3(nb), dim3(kBlock), 0, stream_, n, d.shape_moved, large_, counts_ + n);
  // The path only changes the work: with a count of moved proxies from the caller that
  // clearly fits one path, the count's round trip to the host is skipped (the direct
  // kernel reads the listed count itself; the grid does not need it).
  const bool direct_by_hint = moved_hint >= 0 && moved_hint * (int64_t)n <= kDirectTests / 2;
  const bool grid_by_hint = moved_hint >= 0 && moved_hint * (int64_t)n > 2 * kDirectTests;
  int64_t n_moved = moved_hint;
  if (!direct_by_hint && !grid_by_hint) {
    BP_CHECK(hipMemcpyAsync(host_entries_, counts_ + n, sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    BP_CHECK(hipStreamSynchronize(stream_));
    n_moved = *host_entries_;
  }
  if (direct_by_hint || (!grid_by_hint && n_moved * (int64_t)n <= kDirectTests)) {
    // few moved proxies: each against every proxy, no grid
    for (int attempt = 0; attempt < 2; ++attempt) {
      BP_CHECK(hipMemsetAsync(found_, 0, 2 * sizeof(int64_t), stream_));
      hipLaunchKernelGGL(k_moved_all, dim3(kTestBlocks), dim3(kBlock), 0, stream_, n, large_, counts_ + n, d.fat_aabbs,
                         d.shape_moved, proxies, d.pairs, d.joint_pairs, pairs_, (int64_t)cap_pairs_, found_);
      BP_CHECK(hipMemcpyAsync(host_, found_, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
      BP_CHECK(hipStreamSynchronize(stream_));
      if (host_[0] <= (int64_t)cap_pairs_) break;
      ensure(pairs_, cap_pairs_, (size_t)host_[0]);
    }
  } else {
    // cell size from the extents of the moving proxies
    BP_CHECK(hipMemsetAsync(grid_, 0, sizeof(GridStats), stream_));
    hipLaunchKernelGGL(k_grid_extent, dim3(nb), dim3(kBlock), 0, stream_, n, d.fat_aabbs, proxies, (GridStats*)grid_);
    hipLaunchKernelGGL(k_grid_cell, dim3(1), dim3(1), 0, stream_, (GridStats*)grid_);
    // entries per proxy, their offsets and their count (the sort needs it on the host)
    hipLaunchKernelGGL(k_grid_count, dim3(nb1), dim3(kBlock), 0, stream_, n, d.fat_aabbs, (GridStats*)grid_, counts_, large_);
    size_t scan_bytes = 0;
    BP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts_, offsets_, n + 1, stream_));
    ensure(temp_, cap_temp_, scan_bytes);
    scan_bytes = cap_temp_;
    BP_CHECK(hipcub::DeviceScan::ExclusiveSum(temp_, scan_bytes, counts_, offsets_, n + 1, stream_));
    BP_CHECK(hipMemcpyAsync(host_entries_, offsets_ + n, sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    BP_CHECK(hipStreamSynchronize(stream_));
    const int m = *host_entries_;
    if (m > 0) {
      ensure(ekeys_in_, cap_ekeys_in_, m);
      ensure(ekeys_out_, cap_ekeys_out_, m);
      ensure(evals_in_, cap_evals_in_, m);
      ensure(evals_out_, cap_evals_out_, m);
      hipLaunchKernelGGL(k_grid_emit, dim3(nb), dim3(kBlock), 0, stream_, n, d.fat_aabbs, (const GridStats*)grid_, offsets_,
                         ekeys_in_, evals_in_);
      size_t sort_bytes = 0;
      BP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, ekeys_in_, ekeys_out_, evals_in_, evals_out_, m, 0, 32,
                                                  stream_));
      ensure(temp_, cap_temp_, sort_bytes);
      sort_bytes = cap_temp_;
      BP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp_, sort_bytes, ekeys_in_, ekeys_out_, evals_in_, evals_out_, m, 0, 32,
                                                  stream_));
    }

    for (int attempt = 0; attempt < 2; ++attempt) {
      BP_CHECK(hipMemsetAsync(found_, 0, 2 * sizeof(int64_t), stream_));
      if (m > 0)
        hipLaunchKernelGGL(k_grid_test, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, m, ekeys_out_, evals_out_,
                           d.fat_aabbs, d.shape_moved, (const GridStats*)grid_, proxies, d.pairs, d.joint_pairs, pairs_,
                           (int64_t)cap_pairs_, found_);
      hipLaunchKernelGGL(k_grid_large, dim3(kTestBlocks), dim3(kBlock), 0, stream_, n, large_, d.fat_aabbs, d.shape_moved,
                         (const GridStats*)grid_, proxies, d.pairs, d.joint_pairs, pairs_, (int64_t)cap_pairs_, found_);
      BP_CHECK(hipMemcpyAsync(host_, found_, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
      BP_CHECK(hipStreamSynchronize(stream_));
      if (host_[0] <= (int64_t)cap_pairs_) break;
      ensure(pairs_, cap_pairs_, (size_t)host_[0]);  // too many for the buffer: grow and test again
    }
  }
  const int64_t found = host_[0];
  BP_CHECK(hipMemsetAsync(d.shape_moved, 0, n, stream_));
  BP_CHECK(hipMemsetAsync(d.counters + 6, 0, sizeof(int32_t), stream_));
  if (found == 0) return;
  if (found > (int64_t)cap_pairs_ || found > (int64_t)INT32_MAX) throw std::runtime_error("pair finder: too many new pairs");

  const int nf = (int)found;
  ensure(pairs_sorted_, cap_pairs_sorted_, nf);
  size_t key_bytes = 0;
  // Full 64-bit range: a begin_bit of 20 (the keys' zero low bits) gave unsorted output for
  // more than 1024 keys with this hipCUB.
  BP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, key_bytes, pairs_, pairs_sorted_, nf, 0, 64, stream_));
  ensure(temp_, cap_temp_, key_bytes);
  key_bytes = cap_temp_;
  BP_CHECK(hipcub::DeviceRadixSort::SortKeys(temp_, key_bytes, pairs_, pairs_sorted_, nf, 0, 64, stream_));
  if (!host_keys_ || (size_t)nf > cap_host_keys_) {
    if (host_keys_) BP_CHECK(hipHostFree(host_keys_));
    cap_host_keys_ = std::max<size_t>((size_t)nf * 2, 4096);
    BP_CHECK(hipHostMalloc(&host_keys_, cap_host_keys_ * sizeof(uint64_t), hipHostMallocDefault));
  }
  BP_CHECK(hipMemcpyAsync(host_keys_, pairs_sorted_, (size_t)nf * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
  BP_CHECK(hipStreamSynchronize(stream_));
  keys.assign(host_keys_, host_keys_ + nf);
  for (int i = 1; i < nf; ++i)
    if (keys[i] <= keys[i - 1]) throw std::runtime_error("pair finder: keys out of order at " + std::to_string(i));
}

}  // namespace nsl
