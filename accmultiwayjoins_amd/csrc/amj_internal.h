// Internal declarations shared by the HIP kernels and the host runtime of libamj.
//
// Everything here is gfx950 (CDNA4, wave64) specific: a wave is 64 lanes, ballots are
// 64-bit, and the kernels rely on wave-uniform control flow for the frontier DFS.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/amj.h"

namespace amj {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxLevels = AMJ_MAX_BUILD_TABLES;  // build tables per query
constexpr int kMaxAttrs = AMJ_MAX_ATTRS;          // attributes in a result tuple

// ---------------------------------------------------------------------------------------
// Error plumbing: every host entry point returns an amj_status; the context keeps the
// message of the last failure.
// ---------------------------------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

void set_error(amj_ctx_t* ctx, int code, const std::string& msg);

#define AMJ_HIP_TRY(ctx, expr)                                                          \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) {                                                             \
      ::amj::set_error((ctx), AMJ_ERR_HIP,                                              \
                       std::string(#expr) + ": " + hipGetErrorString(_e));              \
      return AMJ_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define AMJ_TRY(expr)                 \
  do {                                \
    int _rc = (expr);                 \
    if (_rc != AMJ_OK) return _rc;    \
  } while (0)

// ---------------------------------------------------------------------------------------
// Device-side views.
// ---------------------------------------------------------------------------------------

// Bucketed hash table over one build relation. Rows are stored in bucket-major,
// key-minor order (and value-minor inside a key run when `vals_sorted`). The bucket of
// key k is (k & (B-1)) with B = 2^log_b, as in the reference's bucket hashing.
struct DevTable {
  const uint32_t* ptrs;   // [B+1] first row of each bucket
  const int32_t* keys;    // [n] hash-column values, bucket-major
  const int32_t* vals;    // [n] second-column values in the same order (nullptr: 1 column)
  const uint64_t* pw;     // [n+1] exclusive prefix of subtree weights (frequent-item buffer)
  uint64_t n;
  uint32_t log_b;
  uint32_t identity;      // every key lies in [0, B): bucket b holds only key b
  uint32_t vals_sorted;   // values ascending (signed) inside each key run
  uint32_t pad;
};

// One 16-byte record per row of table t (or probe row) for the frequent-item path: the value the
// row binds and the run [lo, lo+len) of table t+1 that the value leads to. Packing them keeps the
// data a child needs in one cache line per run of consecutive rows.
struct RunRef {
  int32_t val;
  uint32_t lo;
  uint32_t len;
  uint32_t pad;
};

// How a build table's second column relates to the tuple built so far.
enum ValMode : int32_t { VAL_NONE = 0, VAL_BIND = 1, VAL_CHECK = 2 };

// A fully resolved multi-way probe plan (one probe relation + n_levels build tables).
// Attribute ids are renumbered 0..n_attrs-1; the output column of attribute a is a.
struct Plan {
  int32_t n_levels;                   // build tables, >= 1
  int32_t n_attrs;                    // attributes in a result tuple
  int32_t probe_ncols;
  int32_t probe_on_host;              // probe columns are host memory: streamed in slices
  uint64_t probe_n;
  const int32_t* probe_cols[AMJ_MAX_PROBE_COLS];
  int32_t probe_attr[AMJ_MAX_PROBE_COLS];
  DevTable tab[kMaxLevels];
  int32_t key_attr[kMaxLevels];       // bound attribute probed into table t
  int32_t val_attr[kMaxLevels];       // attribute of the second column (-1: none)
  int32_t val_mode[kMaxLevels];       // ValMode
  int32_t pad1;
  // Output-window writer (frequent-item path) only:
  const RunRef* next[kMaxLevels];     // next[t][r]: binding of row r of table t and its run
                                      // one level down (table t+1, or the suffix table)
  const RunRef* probe_next;           // run of each probe row in table 0 (or the suffix table)
  const int32_t* sfx_base;            // suffix table: column k at sfx_base + k * sfx_stride
  uint64_t sfx_stride;
  int32_t sfx_cols;                   // columns of the suffix table (0..kMaxAttrs)
  int32_t pad2;
};

// ---------------------------------------------------------------------------------------
// Small device helpers (wave64).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// The dispatcher hands workgroup b of a launch to XCD b % 8. xcd_block(b, n) is a bijection of
// [0, n) that gives every XCD one contiguous range of logical blocks, so the waves resident on
// one XCD work on neighbouring data and share its L2 and address-translation caches.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t q = n / kXcds, r = n % kXcds, x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
__device__ __forceinline__ T uniform(T v) {  // value known to be equal on all lanes
  if constexpr (sizeof(T) == 8) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v));
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(((uint64_t)v) >> 32));
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)__builtin_amdgcn_readfirstlane((uint32_t)v);
  }
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t y = __shfl_up(x, (unsigned)d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t y = __shfl_up(x, (unsigned)d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014); used for the order-independent
// result checksum and by the deterministic generators.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Hash of one result tuple: chained over the attribute values in attribute order.
// The checksum of a result is the wrapping sum of tuple hashes, so it does not depend on
// the order in which tuples are produced.
__host__ __device__ __forceinline__ uint64_t tuple_hash_step(uint64_t h, int32_t v) {
  return mix64(h ^ (uint64_t)(uint32_t)v);
}

__host__ __device__ __forceinline__ uint32_t rotr32(uint32_t x, uint32_t r) {
  return r == 0 ? x : ((x >> r) | (x << (32u - r)));
}
__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, uint32_t r) {
  return r == 0 ? x : ((x << r) | (x >> (32u - r)));
}

// ---------------------------------------------------------------------------------------
// Primitives (prims.hip).
// ---------------------------------------------------------------------------------------
// Exclusive scan, in place allowed. `temp` must hold scan_temp_elems(n) elements.
size_t scan_temp_elems(uint64_t n);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp,
                              hipStream_t s);
hipError_t exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* temp,
                              hipStream_t s);
// Sum of n u64 values into *out (device pointer).
hipError_t reduce_sum_u64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* temp,
                          hipStream_t s);

// Stable LSD radix sort of 64-bit keys on the 8-bit digits selected by `digit_mask`
// (bit d set: sort on bits [8d, 8d+8)). Digits whose value is the same for every key are
// skipped automatically. Result ends in `keys` (alt is scratch of n elements).
struct RadixSortTemp {
  uint32_t* hist = nullptr;     // [256 * tiles]
  uint32_t* scan_tmp = nullptr;
  uint64_t* diff = nullptr;     // [2]
  size_t bytes = 0;
};
size_t radix_sort_temp_bytes(uint64_t n);
hipError_t radix_sort_u64(uint64_t* keys, uint64_t* alt, uint64_t n, uint32_t digit_mask,
                          void* temp, hipStream_t s);

// ---------------------------------------------------------------------------------------
// Hash build (hashbuild.hip).
// ---------------------------------------------------------------------------------------
struct HashTableDev {
  DevTable view;
  uint32_t* ptrs = nullptr;
  int32_t* keys = nullptr;
  int32_t* vals = nullptr;
  uint64_t* pw = nullptr;
  uint64_t buckets = 0;
};

// Launch helpers used by the host runtime.
hipError_t launch_pack_composite(const int32_t* key, const int32_t* val, uint64_t n,
                                 uint32_t log_b, uint64_t* out, hipStream_t s);
hipError_t launch_unpack_composite(const uint64_t* comp, uint64_t n, uint32_t log_b,
                                   int32_t* keys, int32_t* vals, hipStream_t s);
hipError_t launch_bucket_ptrs(const uint64_t* comp, uint64_t n, uint32_t log_b,
                              uint64_t buckets, uint32_t* ptrs, hipStream_t s);
hipError_t launch_key_range(const int32_t* key, uint64_t n, uint32_t log_b, uint32_t* flag,
                            hipStream_t s);

// ---------------------------------------------------------------------------------------
// Probe engine (probe.hip).
// ---------------------------------------------------------------------------------------
struct ProbeOut {
  int32_t* cols[kMaxAttrs];  // output columns (nullptr: do not materialise)
  uint64_t capacity;         // rows available in every output column
  // Direct output reserved `chunk` rows at a time (0: one reservation per staging buffer);
  // tails[c] receives the unused tail of chunk c when it is a wave's last one.
  uint32_t chunk;
  uint64_t* tails;
};

// Rows per reservation of a chunked direct-output pass by `waves` waves expecting `expected`
// results: 0 (one reservation per buffer) when the output is small, else the largest power of
// two in [256, 4096] whose worst-case unused tails (one per wave) stay within 1/16 of it.
inline uint32_t direct_chunk_rows(uint64_t expected, uint64_t waves) {
  if (expected < waves * 256ull * 16) return 0;
  uint32_t chunk = 256;
  while (chunk < 4096 && waves * (uint64_t)chunk * 2 * 16 <= expected) chunk *= 2;
  return chunk;
}

// Fill-up after a chunked direct-output pass of n results whose chunks end at `reserved`: the
// rows at [n, reserved) outside the chunks' unused tails move, in order, to the tail slots below
// n, so that the first n rows are the results. tail[c] is chunk c's unused tail and tpre its
// exclusive prefix sum, over the reserved / chunk chunks.
hipError_t launch_chunk_fill(const ProbeOut& out, int n_attrs, const uint64_t* tail,
                             const uint64_t* tpre, uint64_t n_chunks, uint32_t chunk, uint64_t n,
                             uint64_t reserved, hipStream_t s);

// Count pass: per-unit result counts (unit = 64 consecutive probe rows).
hipError_t launch_probe_count(const Plan& plan, uint64_t* unit_counts, hipStream_t s);
// Write pass: units write at unit_base[u] (exclusive scan of counts); per-block checksum
// partials go to block_sums. hash=false skips the checksum when materialising.
// With unit_base == nullptr (no output) it is the only pass: per-block result counts go to
// block_counts.
hipError_t launch_probe_write(const Plan& plan, const uint64_t* unit_base, const ProbeOut& out,
                              bool hash, uint64_t* block_sums, uint64_t* n_blocks_out,
                              hipStream_t s, uint64_t* block_counts = nullptr);
uint64_t probe_units(const Plan& plan);
uint64_t probe_write_blocks(const Plan& plan);

// Thread-per-tuple baseline (tpt.hip): count pass (write = false) stores per-lane counts in
// counts[0 .. tpt_blocks * kBlock); the write pass writes lane g's results from base[g] on.
// With hash, per-block checksum partials go to block_sums[0 .. tpt_blocks).
uint64_t tpt_blocks(const Plan& plan);
struct OaTables;
hipError_t launch_tpt(const Plan& plan, bool write, bool hash, uint64_t* counts,
                      const uint64_t* base, const ProbeOut& out, uint64_t* block_sums,
                      hipStream_t s, const OaTables* oa = nullptr);

// Frequent-item (subtree-count) buffer for linear chains.
hipError_t launch_fib_weights(const Plan& plan, int level, uint64_t* w, RunRef* next,
                              hipStream_t s);
hipError_t launch_fib_probe_weights(const Plan& plan, uint64_t* w, RunRef* next, hipStream_t s);
// Output-window writer: every wave materialises outputs [w*win, (w+1)*win) exactly; win is a
// multiple of kWindowAlign rows (the writer stages and stores aligned blocks of that many rows).
constexpr uint64_t kWindowAlign = 256;
uint64_t window_blocks(uint64_t total, uint64_t window);
bool window_plan_supported(const Plan& plan);
hipError_t launch_records_to_suffix(RunRef* rec, uint64_t n, const uint64_t* pw, hipStream_t s);
hipError_t launch_window_write(const Plan& plan, const uint64_t* probe_off, uint64_t* first_row,
                               uint64_t total, uint64_t window, const ProbeOut& out, bool hash,
                               uint64_t* block_sums, uint32_t* overflow, uint64_t* n_blocks_out,
                               hipStream_t s);

}  // namespace amj
