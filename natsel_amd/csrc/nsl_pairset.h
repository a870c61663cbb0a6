// Device hash set of the pair keys that have a contact (the reference keeps this set on
// the host: contact_creation.rs:901-1042 checks it before creating a contact, and
// destroy_contact removes the key). Open addressing with linear probing over a
// power-of-two table of 64-bit keys; removed keys leave a tombstone until the host
// rebuilds the table.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nsl {

constexpr uint64_t kPairEmpty = ~0ull;
constexpr uint64_t kPairTomb = ~0ull - 1;

// hash_set.rs:25-40: (smaller shape << 42) | (larger shape << 20) | child (0 here)
__host__ __device__ __forceinline__ uint64_t pair_key(int32_t s1, int32_t s2) {
  const uint64_t mask = (1ull << 22) - 1;
  const uint64_t a = mask & (uint64_t)(int64_t)s1, b = mask & (uint64_t)(int64_t)s2;
  return s1 < s2 ? (a << 42) | (b << 20) : (b << 42) | (a << 20);
}

struct PairSet {
  uint64_t* slots;
  uint32_t mask;  // capacity - 1
};

__device__ __forceinline__ uint32_t pair_slot(uint64_t key, uint32_t mask) {
  uint64_t h = key ^ (key >> 33);  // 64-bit finaliser (MurmurHash3 fmix64)
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return (uint32_t)h & mask;
}

// Inserting a key that is present leaves the set unchanged. Concurrent inserts of distinct
// keys are safe; one key is never inserted by two threads at once.
__device__ inline void pairset_insert(PairSet s, uint64_t key) {
  for (;;) {
    uint32_t i = pair_slot(key, s.mask), target = ~0u;
    uint64_t expect = kPairEmpty;
    for (uint32_t probe = 0; probe <= s.mask; ++probe, i = (i + 1) & s.mask) {
      const uint64_t cur = s.slots[i];
      if (cur == key) return;
      if (cur == kPairTomb && target == ~0u) {
        target = i;
        expect = kPairTomb;
      }
      if (cur == kPairEmpty) {
        if (target == ~0u) target = i;
        break;
      }
    }
    if (target == ~0u) return;  // full: the host keeps the load factor at or below 1/2
    if (atomicCAS((unsigned long long*)&s.slots[target], (unsigned long long)expect, (unsigned long long)key) == expect) return;
  }
}

__device__ inline void pairset_erase(PairSet s, uint64_t key) {
  uint32_t i = pair_slot(key, s.mask);
  for (uint32_t probe = 0; probe <= s.mask; ++probe, i = (i + 1) & s.mask) {
    const uint64_t cur = s.slots[i];
    if (cur == key) {
      s.slots[i] = kPairTomb;
      return;
    }
    if (cur == kPairEmpty) return;
  }
}

__device__ inline bool pairset_contains(PairSet s, uint64_t key) {
  uint32_t i = pair_slot(key, s.mask);
  for (uint32_t probe = 0; probe <= s.mask; ++probe, i = (i + 1) & s.mask) {
    const uint64_t cur = s.slots[i];
    if (cur == key) return true;
    if (cur == kPairEmpty) return false;
  }
  return false;
}

}  // namespace nsl
