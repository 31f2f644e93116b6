// Bucketed hash-table build (the AMHJ build phase).
//
// Reference: QueryProcessor::build_hash_multiway (MHJ-GPU/QueryProcessor.cuh:76-110) builds,
// for each build relation, B = floorPowerOf2(N) buckets on the first column with bucket of key
// k = k & (B-1), by a multi-pass MSB-first radix partition with atomics
// (MHJ-GPU/Indexing/radix_partitioning.cuh:17-48, :423-497). The row order inside a bucket is
// arbitrary there, and payload columns are reached through idx_in_origin gathers.
//
// Here the same bucket function is kept, but the table is built by a stable LSD radix sort of
// one 64-bit composite key per row:
//     hi32 = rotr32(key, log_b)   -> bucket bits on top, the remaining key bits below
//     lo32 = val ^ 0x80000000     -> signed value order (only sorted on when requested)
// so rows end bucket-major and key-minor (each key's rows form one contiguous run) and the
// second column is stored inline next to its key (no idx_in_origin gather on the probe).
#include "amj_internal.h"

namespace amj {

__global__ __launch_bounds__(kBlock) void pack_composite_kernel(const int32_t* __restrict__ key,
                                                                const int32_t* __restrict__ val,
                                                                uint64_t n, uint32_t log_b,
                                                                uint64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint32_t hi = rotr32((uint32_t)key[i], log_b);
    const uint32_t lo = val ? ((uint32_t)val[i] ^ 0x80000000u) : 0u;
    out[i] = ((uint64_t)hi << 32) | lo;
  }
}

__global__ __launch_bounds__(kBlock) void unpack_composite_kernel(
    const uint64_t* __restrict__ comp, uint64_t n, uint32_t log_b, int32_t* __restrict__ keys,
    int32_t* __restrict__ vals) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t c = comp[i];
    keys[i] = (int32_t)rotl32((uint32_t)(c >> 32), log_b);
    if (vals) vals[i] = (int32_t)((uint32_t)c ^ 0x80000000u);
  }
}

__device__ __forceinline__ uint64_t bucket_of(uint64_t comp, uint32_t log_b) {
  return log_b == 0 ? 0ull : (uint64_t)((uint32_t)(comp >> 32) >> (32u - log_b));
}

// ptrs[b] = first row whose bucket is >= b (ptrs[B] = n): one lower-bound search per bucket,
// so the work per thread stays logarithmic however skewed the buckets are.
__global__ __launch_bounds__(kBlock) void bucket_ptrs_kernel(const uint64_t* __restrict__ comp,
                                                             uint64_t n, uint32_t log_b,
                                                             uint64_t buckets,
                                                             uint32_t* __restrict__ ptrs) {
  for (uint64_t b = (uint64_t)blockIdx.x * kBlock + threadIdx.x; b <= buckets;
       b += (uint64_t)gridDim.x * kBlock) {
    uint64_t lo = 0, hi = n;  // first row with bucket >= b lies in [lo, hi]
    while (lo < hi) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (bucket_of(comp[mid], log_b) < b) lo = mid + 1;
      else hi = mid;
    }
    ptrs[b] = (uint32_t)lo;
  }
}

// flag := 1 if some key lies outside [0, B).
__global__ __launch_bounds__(kBlock) void key_range_kernel(const int32_t* __restrict__ key,
                                                           uint64_t n, uint32_t log_b,
                                                           uint32_t* flag) {
  bool out = false;
  const uint32_t lim = 1u << log_b;  // log_b <= 31
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    out |= (uint32_t)key[i] >= lim;
  if (__ballot(out) != 0ull && lane_id() == 0) atomicOr(flag, 1u);
}

static unsigned grid_for(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return (unsigned)g;
}

hipError_t launch_pack_composite(const int32_t* key, const int32_t* val, uint64_t n,
                                 uint32_t log_b, uint64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(pack_composite_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, key, val, n,
                     log_b, out);
  return hipGetLastError();
}
hipError_t launch_unpack_composite(const uint64_t* comp, uint64_t n, uint32_t log_b,
                                   int32_t* keys, int32_t* vals, hipStream_t s) {
  hipLaunchKernelGGL(unpack_composite_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, comp, n,
                     log_b, keys, vals);
  return hipGetLastError();
}
hipError_t launch_bucket_ptrs(const uint64_t* comp, uint64_t n, uint32_t log_b,
                              uint64_t buckets, uint32_t* ptrs, hipStream_t s) {
  if (n == 0) return hipMemsetAsync(ptrs, 0, (buckets + 1) * sizeof(uint32_t), s);
  hipLaunchKernelGGL(bucket_ptrs_kernel, dim3(grid_for(buckets + 1)), dim3(kBlock), 0, s, comp,
                     n, log_b, buckets, ptrs);
  return hipGetLastError();
}
hipError_t launch_key_range(const int32_t* key, uint64_t n, uint32_t log_b, uint32_t* flag,
                            hipStream_t s) {
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(uint32_t), s);
  if (e != hipSuccess || n == 0) return e;
  hipLaunchKernelGGL(key_range_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, key, n, log_b,
                     flag);
  return hipGetLastError();
}

}  // namespace amj
