// Counting pass of a counting sort on the device: for items i < m with keys
// key_of(i), rank[i] = position of item i among the items of its key, and
// count[key] accumulates the key's size (count zeroed beforehand).
//
// A workgroup takes a chunk of kChunkItems consecutive items and counts them
// per key in an LDS hash table (linear probing with compare-and-swap; at most
// kChunkItems keys in kChunkSlots = 2 kChunkItems slots, so probing always
// ends), each item taking its rank within the chunk from an LDS atomic; then
// one global atomic per distinct key of the chunk reserves the chunk's range
// of that key. Per-item global atomics serialise on hot keys (clustered points
// put ~1e5 points into one bin; bin-sorted points put a wave's points into a
// few rows); a chunk of such items touches few keys.
#pragma once

#include "common.hpp"

namespace nnf {

constexpr int kChunkThreads = 256;
constexpr int kChunkPerThread = 8;
constexpr int kChunkItems = kChunkThreads * kChunkPerThread;
constexpr int kChunkSlots = 2 * kChunkItems;
constexpr int kChunkHashBits = 12;
static_assert((1 << kChunkHashBits) == kChunkSlots);

// Call from every thread of a kChunkThreads-thread workgroup, grid-striding
// over chunks; `key_of(i)` returns a uint32_t other than 0xFFFFFFFF.
template <typename KeyOf>
__device__ inline void chunk_rank(std::size_t m, uint32_t* __restrict__ count,
                                  uint32_t* __restrict__ rank, KeyOf key_of) {
  constexpr uint32_t kEmpty = 0xFFFFFFFFu;
  __shared__ uint32_t key[kChunkSlots];
  __shared__ uint32_t cnt[kChunkSlots];
  for (std::size_t c0 = static_cast<std::size_t>(blockIdx.x) * kChunkItems; c0 < m;
       c0 += static_cast<std::size_t>(gridDim.x) * kChunkItems) {
    for (int s = threadIdx.x; s < kChunkSlots; s += kChunkThreads) {
      key[s] = kEmpty;
      cnt[s] = 0;
    }
    __syncthreads();
    int slot[kChunkPerThread];
    uint32_t local[kChunkPerThread];
#pragma unroll
    for (int k = 0; k < kChunkPerThread; ++k) {
      const std::size_t i = c0 + static_cast<std::size_t>(k * kChunkThreads) + threadIdx.x;
      slot[k] = -1;
      local[k] = 0;
      if (i < m) {
        const uint32_t kk = key_of(i);
        int s = static_cast<int>((kk * 2654435761u) >> (32 - kChunkHashBits));
        for (int probe = 0; probe < kChunkSlots; ++probe) {
          const uint32_t prev = atomicCAS(&key[s], kEmpty, kk);
          if (prev == kEmpty || prev == kk) break;
          s = (s + 1) & (kChunkSlots - 1);
        }
        slot[k] = s;
        local[k] = atomicAdd(&cnt[s], 1u);
      }
    }
    __syncthreads();
    // the chunk's range of each of its keys; the slot then holds its base
    for (int s = threadIdx.x; s < kChunkSlots; s += kChunkThreads) {
      const uint32_t n = cnt[s];
      if (n != 0) cnt[s] = atomicAdd(&count[key[s]], n);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kChunkPerThread; ++k) {
      const std::size_t i = c0 + static_cast<std::size_t>(k * kChunkThreads) + threadIdx.x;
      if (slot[k] >= 0) rank[i] = cnt[slot[k]] + local[k];
    }
    __syncthreads();  // the table is cleared for the next chunk
  }
}

// Workgroups for m items (at most 8192; they stride over the chunks).
inline unsigned chunk_rank_blocks(std::size_t m) {
  std::size_t g = (m + kChunkItems - 1) / kChunkItems;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return static_cast<unsigned>(g);
}

}  // namespace nnf
