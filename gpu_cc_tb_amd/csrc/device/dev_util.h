// Device helpers for gfx950 (CDNA4, wave64).
//
// Memory model.  On gfx950 an agent-scope acquire/release fence writes back or
// invalidates the whole L2 of the issuing XCD (buffer_wbl2 sc1 / buffer_inv
// sc1), which is far too slow to pay per row access.  Every location that two
// transactions can share (CC words, table cells, per-row counters, MVCC
// version nodes, allocator counters) is therefore accessed ONLY through
// agent-scope atomic operations: relaxed atomic loads/stores compile to
// global_load/store ... sc1, which are coherent across the 8 XCDs, and RMW
// atomics execute at the device coherence point.  Ordering between two such
// accesses to different locations is enforced by waiting for completion of the
// earlier ones (s_waitcnt vmcnt(0)) plus a workgroup-scope fence that stops the
// compiler from moving memory operations across the point.  This is the same
// "order-only fence over cache-bypassing atomics" discipline ROCm's rocPRIM
// uses for its cross-workgroup look-back state on gfx942/gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../gcctb_common.h"

namespace gcctb {
namespace dev {

// ---- coherent accesses ----------------------------------------------------
// Loads and stores of shared data (cells, row words, chain heads, version
// nodes, turn counters) are system-scope relaxed atomics (global_load/store
// sc0 sc1): a store is acknowledged once performed at memory and a load reads
// from there, not from a line one XCD's L2 cached earlier.  Read-modify-write
// atomics are performed at the memory-side coherence point at agent scope.
// With that, waiting for an access to complete (s_waitcnt) orders it before
// the next one for every XCD, so publication needs no L2-wide write-back or
// invalidate (an agent-scope release/acquire fence costs both and ran the
// low-contention workload 2.6x slower).
__device__ __forceinline__ uint64_t ld(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// Compare-and-swap; on failure `expected` receives the current value.
__device__ __forceinline__ bool cas(uint64_t* p, uint64_t& expected, uint64_t desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas32(uint32_t* p, uint32_t& expected, uint32_t desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t fetch_add(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t fetch_add32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t fetch_sub(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t fetch_or(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- ordering between different locations ----------------------------------
// s_waitcnt simm16 on gfx9: vmcnt[3:0]+[15:14], expcnt[6:4], lgkmcnt[11:8].
constexpr int WAIT_VM0 = 0 | (0x7 << 4) | (0xf << 8);
// release side: data before the flag that publishes it.  Every earlier access
// of this lane has completed (stores performed at memory, see above) before
// any later one is issued; the workgroup fence keeps the compiler from
// moving accesses across it.
__device__ __forceinline__ void order_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(WAIT_VM0);
}
// acquire side: flag before the data it guards
__device__ __forceinline__ void order_acquire() {
  __builtin_amdgcn_s_waitcnt(WAIT_VM0);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---- wave / group intrinsics ----------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t l = lane_id();
  return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}
template <int G>
__device__ __forceinline__ uint64_t group_mask() {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group size");
  if constexpr (G == 64) {
    return ~0ull;
  } else {
    const uint32_t base = lane_id() & ~(uint32_t)(G - 1);
    return ((1ull << G) - 1) << base;
  }
}
// Group-wide predicates.  Every lane of the group must execute the call with
// the same control flow (the group's phase is uniform); lanes of other groups
// may be inactive.
template <int G>
__device__ __forceinline__ bool group_any(bool p) {
  return (ballot(p) & group_mask<G>()) != 0;
}
template <int G>
__device__ __forceinline__ bool group_all(bool p) {
  return (ballot(!p) & group_mask<G>()) == 0;
}
__device__ __forceinline__ uint32_t shfl32(uint32_t v, int src) {
  return (uint32_t)__shfl((int)v, src, 64);
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = shfl32((uint32_t)v, src);
  const uint32_t hi = shfl32((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}
// Broadcast from the group's first lane.
template <int G>
__device__ __forceinline__ uint64_t group_bcast64(uint64_t v) {
  return shfl64(v, (int)(lane_id() & ~(uint32_t)(G - 1)));
}
template <int G>
__device__ __forceinline__ uint64_t group_max64(uint64_t v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    const uint64_t o = shfl_xor64(v, m);
    v = o > v ? o : v;
  }
  return v;
}
template <int G>
__device__ __forceinline__ uint64_t group_min64(uint64_t v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    const uint64_t o = shfl_xor64(v, m);
    v = o < v ? o : v;
  }
  return v;
}
template <int G>
__device__ __forceinline__ uint64_t group_sum64(uint64_t v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) v += shfl_xor64(v, m);
  return v;
}

// Wave-aggregated counter allocation: every lane with `want` receives a unique
// value from *ctr; the wave issues a single atomic.  All active lanes of the
// wave must reach the call.
__device__ __forceinline__ uint64_t wave_fetch_inc(uint64_t* ctr, bool want) {
  const uint64_t m = ballot(want);
  if (m == 0) return 0;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint64_t base = 0;
  if ((int)lane_id() == leader) base = fetch_add(ctr, (uint64_t)__popcll(m));
  base = shfl64(base, leader);
  return base + (uint64_t)__popcll(m & lanemask_lt());
}
// Same with a per-lane amount (wave prefix sum + one atomic).
__device__ __forceinline__ uint64_t wave_fetch_add(uint64_t* ctr, uint32_t amount) {
  // inclusive scan over the active lanes in lane order
  const uint64_t active = ballot(true);
  uint32_t incl = amount;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)incl, d, 64);
    const int src = (int)lane_id() - d;
    if (src >= 0 && ((active >> src) & 1ull)) incl += o;
  }
  // inactive lanes in between would break the chain above; the kernels call
  // this with the whole wave active.
  const uint32_t total = shfl32(incl, 63 - __clzll((long long)active));
  const uint64_t has = ballot(amount != 0);
  if (has == 0) return 0;
  const int leader = __ffsll((unsigned long long)active) - 1;
  uint64_t base = 0;
  if ((int)lane_id() == leader) base = fetch_add(ctr, total);
  base = shfl64(base, leader);
  return base + (incl - amount);
}

__device__ __forceinline__ uint64_t now_ticks() { return wall_clock64(); }

// Find (include/index.cuh:74-91) over the hash index: the key's bucket
// holds it, or the lookup moves to the next bucket while every slot is taken.
// Find returns the index length for an absent key; on the device a miss is
// IDX_MISS instead, so that it can never be mistaken for a row when the index
// is shorter than the table (gcctb_probe maps it back to the length).
__device__ __forceinline__ uint32_t index_find(const IndexView& ix, uint32_t key) {
  if (ix.dense) return idx_dense_find(key, ix.length);
  if (key == IDX_EMPTY) return IDX_MISS;  // reserved, never a key
  uint32_t b = idx_bucket(key, ix.nbuckets);
  for (uint32_t step = 0; step < ix.nbuckets; ++step) {
    const uint32_t* bk = ix.buckets + (size_t)b * (2 * IDX_SLOTS);
    const uint4* q = reinterpret_cast<const uint4*>(bk);
    const uint4 k0 = q[0], k1 = q[1];
    const uint32_t ks[IDX_SLOTS] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
    int hit = -1;
    bool free_slot = false;
#pragma unroll
    for (int i = 0; i < IDX_SLOTS; ++i) {
      hit = ks[i] == key ? i : hit;
      free_slot |= ks[i] == IDX_EMPTY;
    }
    if (hit >= 0) return bk[IDX_SLOTS + hit];
    if (free_slot) return IDX_MISS;
    b = b + 1 == ix.nbuckets ? 0u : b + 1;
  }
  return IDX_MISS;
}

}  // namespace dev
}  // namespace gcctb
