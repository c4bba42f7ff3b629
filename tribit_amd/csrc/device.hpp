// device.hpp -- wave64 device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace tb {

// SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014); the generators'
// counter-based stream: key = mix64(seed ^ mix64(stream)),
// word i = mix64(key + i * 0xD1342543DE82EF95).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed ^ mix64(stream));
}
__host__ __device__ __forceinline__ uint64_t rand_word(uint64_t key, uint64_t i) {
  return mix64(key + i * 0xD1342543DE82EF95ull);
}

constexpr uint64_t kStreamRmat = 0x524D4154ull;  // "RMAT"
constexpr uint64_t kStreamEr = 0x45524E59ull;    // "ERNY"

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

// Orders this wave's LDS writes before its later LDS reads (all lanes).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Run-length aggregation inside a wave of sorted values: for a lane that
// heads a run of equal `key` among the wave's first `nvalid` lanes, returns
// the run length; other lanes get 0. Lanes >= nvalid must be invalid.
__device__ __forceinline__ uint32_t wave_run_length(uint32_t key, bool valid) {
  const uint32_t lane = lane_id();
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = valid && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(head);
  const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
  if (!head) return 0;
  const uint64_t after = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
  uint32_t next = after ? (uint32_t)(__ffsll((long long)after) - 1) : 64u;
  if (next > nvalid) next = nvalid;
  return next - lane;
}

// Segmented sum inside a wave of lanes sorted by `key` (equal keys
// contiguous, valid lanes first): returns true on the last valid lane of each
// run of equal keys, with *sum = the run's total of `val` up to that lane.
__device__ __forceinline__ bool wave_segment_tail(uint32_t key, bool valid, uint64_t val, uint64_t* sum) {
  const uint32_t lane = lane_id();
  uint64_t s = valid ? val : 0;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(s, o);
    const uint32_t kv = __shfl_up(key, o);
    if (lane >= (uint32_t)o && kv == key) s += t;
  }
  const uint32_t nk = __shfl_down(key, 1);
  const uint64_t vmask = __ballot(valid);
  const uint64_t above = (lane == 63) ? 0ull : (vmask >> (lane + 1));
  *sum = s;
  return valid && (!(above & 1ull) || nk != key);
}

}  // namespace tb
