// Device building blocks shared by the AMHJ probe kernels (unit engine and output-window
// writer): attribute access, run lookup, wave64 scans and parent search, tuple hashing.
#pragma once

#include <type_traits>

#include "amj_internal.h"

namespace amj {

// Compile-time unrolled loop over I in [0, N): every index is a constant in the IR from the
// start, so SROA keeps per-attribute arrays in VGPRs.
template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, N>(f);
  }
}

// Attributes bound by a query of L build tables: the probe binds at most two and every table
// at most one more. Per-lane binding arrays of the frontier engines are sized by it, so a
// shallow query does not pay for the deepest one's registers.
constexpr int na_of(int L) { return L + 2 < kMaxAttrs ? L + 2 : kMaxAttrs; }

// Attribute access with a wave-uniform runtime attribute index (lowered to v_cndmask chains).
template <int I = 0, int N>
__device__ __forceinline__ int32_t get_attr(const int32_t (&b)[N], int a) {
  if constexpr (I == N - 1) return b[I];
  else return a == I ? b[I] : get_attr<I + 1>(b, a);
}
template <int N>
__device__ __forceinline__ void set_attr(int32_t (&b)[N], int a, int32_t v) {
  sfor<0, N>([&](auto i) { b[i] = (a == (int)i) ? v : b[i]; });
}

// ------------------------------------------------------------------------------------------
// Run lookup: rows of table `tb` whose key equals `key` (and whose value equals `cval` when
// `check`), returned as the contiguous run [lo, lo+len).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lower_bound_u32(const int32_t* a, uint32_t lo, uint32_t hi,
                                                    uint32_t k) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if ((uint32_t)a[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t upper_bound_u32(const int32_t* a, uint32_t lo, uint32_t hi,
                                                    uint32_t k) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if ((uint32_t)a[mid] <= k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t lower_bound_i32(const int32_t* a, uint32_t lo, uint32_t hi,
                                                    int32_t k) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t upper_bound_i32(const int32_t* a, uint32_t lo, uint32_t hi,
                                                    int32_t k) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void lookup_run(const DevTable& tb, int32_t key, bool check,
                                           int32_t cval, bool valid, uint32_t& lo,
                                           uint32_t& len) {
  lo = 0;
  len = 0;
  if (!valid) return;
  const uint32_t k = (uint32_t)key;
  const uint32_t mask = (tb.log_b == 0) ? 0u : ((1u << tb.log_b) - 1u);
  uint32_t s, e;
  if (tb.identity) {
    if (k > mask) return;  // bucket b holds only key b
    s = tb.ptrs[k];
    e = tb.ptrs[k + 1];
  } else {
    const uint32_t b = k & mask;
    s = tb.ptrs[b];
    e = tb.ptrs[b + 1];
    if (s == e) return;
    const uint32_t first = (uint32_t)tb.keys[s];
    const uint32_t last = (uint32_t)tb.keys[e - 1];
    if (k < first || k > last) return;
    if (first != k) s = lower_bound_u32(tb.keys, s + 1, e, k);
    if (last != k) e = upper_bound_u32(tb.keys, s, e - 1, k);
  }
  if (check && s < e) {
    // the run of cval inside the value-sorted key run; a second search only when cval occurs
    // more than once (duplicate pairs), so a membership test costs one search plus one load
    const uint32_t s2 = lower_bound_i32(tb.vals, s, e, cval);
    uint32_t e2 = s2;
    if (s2 < e && tb.vals[s2] == cval) {
      e2 = s2 + 1;
      if (e2 < e && tb.vals[e2] == cval) e2 = upper_bound_i32(tb.vals, e2, e, cval);
    }
    s = s2;
    e = e2;
  }
  lo = s;
  len = e - s;
}

// ------------------------------------------------------------------------------------------
// Wave64 scans with DPP row shifts and row broadcasts (VALU latency, no LDS round trips).
// ------------------------------------------------------------------------------------------
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_or0(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, BANK_MASK, false);
}

// inclusive prefix sum over the 64 lanes
__device__ __forceinline__ uint32_t wave_scan_add_dpp(uint32_t x) {
  x += dpp_or0<0x111>(x);        // row_shr:1
  x += dpp_or0<0x112>(x);        // row_shr:2
  x += dpp_or0<0x114>(x);        // row_shr:4
  x += dpp_or0<0x118>(x);        // row_shr:8
  x += dpp_or0<0x142, 0xa>(x);   // row_bcast:15 into rows 1 and 3
  x += dpp_or0<0x143, 0xc>(x);   // row_bcast:31 into rows 2 and 3
  return x;
}

// inclusive prefix maximum over the 64 lanes (values >= 0)
__device__ __forceinline__ uint32_t wave_scan_max_dpp(uint32_t x) {
  x = max(x, dpp_or0<0x111>(x));
  x = max(x, dpp_or0<0x112>(x));
  x = max(x, dpp_or0<0x114>(x));
  x = max(x, dpp_or0<0x118>(x));
  x = max(x, dpp_or0<0x142, 0xa>(x));
  x = max(x, dpp_or0<0x143, 0xc>(x));
  return x;
}

__device__ __forceinline__ uint32_t lane63(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Last lane p whose exclusive prefix is <= g (lane 0 always qualifies); 6 shuffle steps.
__device__ __forceinline__ int find_parent(uint64_t excl, uint64_t g, uint64_t T) {
  int p = 0;
  if ((T >> 32) == 0) {
    const uint32_t e32 = (uint32_t)excl, g32 = (uint32_t)g;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
      const uint32_t v = __shfl(e32, p + step, kWave);
      if (v <= g32) p += step;
    }
  } else {
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
      const uint64_t v = __shfl(excl, p + step, kWave);
      if (v <= g) p += step;
    }
  }
  return p;
}

// Parent of every result slot of the window [c, c+64) of a level whose lane j owns the
// flattened children [excl_j, excl_j + len_j): lane j marks the first window slot it covers in
// this wave's 64-entry LDS row, and a prefix maximum fills the slots in between. Requires the
// level total to fit in 32 bits.
__device__ __forceinline__ uint32_t parent_of_slots(uint32_t excl, uint32_t len, uint32_t c) {
  __shared__ uint32_t marks[kWavesPerBlock][kWave];
  const int lane = lane_id();
  uint32_t* mark = marks[threadIdx.x / kWave];
  mark[lane] = 0u;
  __asm__ __volatile__("" ::: "memory");
  const uint32_t end = excl + len;
  if (len != 0u && excl < c + (uint32_t)kWave && end > c)
    mark[excl > c ? excl - c : 0u] = (uint32_t)lane;
  __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
  const uint32_t p = mark[lane];
  __asm__ __volatile__("" ::: "memory");
  return wave_scan_max_dpp(p);
}

// Tuple hash over ceil(n_attrs/2) SplitMix64 rounds; attribute pairs are packed in one word
// (n_attrs <= N: the same value whatever the array size).
template <int N>
__device__ __forceinline__ uint64_t hash_tuple(const int32_t (&v)[N], int n_attrs) {
  uint64_t h = (uint64_t)n_attrs;
  sfor<0, (N + 1) / 2>([&](auto j) {
    constexpr int i = 2 * (int)j;
    if (i < n_attrs) {
      uint64_t hi = 0;
      if constexpr (i + 1 < N) hi = (i + 1 < n_attrs) ? (uint64_t)(uint32_t)v[i + 1] : 0ull;
      h = mix64(h ^ ((uint64_t)(uint32_t)v[i] | (hi << 32)));
    }
  });
  return h;
}

// Sort key of a unit for a heaviest-first launch order: (~bits(float(weight)) << 32) | unit.
// The bit pattern of a non-negative float orders like its value, so an ascending sort of the
// keys lists units by descending weight (ties in unit order).
__device__ __forceinline__ uint64_t weight_key(uint64_t weight, uint64_t unit) {
  const uint32_t b = __float_as_uint((float)weight);
  return ((uint64_t)(~b) << 32) | (uint32_t)unit;
}

// Unit processed by wave slot k: order[k] holds it in its low 32 bits (nullptr: unit k).
__device__ __forceinline__ uint64_t unit_of(const uint64_t* order, uint64_t k) {
  return order ? (uint64_t)(uint32_t)order[k] : k;
}

__device__ __forceinline__ void block_sum_store(uint64_t v, uint64_t* dst) {
  __shared__ uint64_t part[kWavesPerBlock];
  v = wave_sum_u64(v);
  if (lane_id() == 0) part[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) s += part[i];
    *dst = s;
  }
}

// Block sums of two per-lane values with one barrier (dst_b may be nullptr).
__device__ __forceinline__ void block_sum_store2(uint64_t a, uint64_t b, uint64_t* dst_a,
                                                 uint64_t* dst_b) {
  __shared__ uint64_t part2[2][kWavesPerBlock];
  a = wave_sum_u64(a);
  b = wave_sum_u64(b);
  if (lane_id() == 0) {
    part2[0][threadIdx.x / kWave] = a;
    part2[1][threadIdx.x / kWave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t sa = 0, sb = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) {
      sa += part2[0][i];
      sb += part2[1][i];
    }
    *dst_a = sa;
    if (dst_b) *dst_b = sb;
  }
}

#define AMJ_DISPATCH_L(P, MACRO)          \
  switch ((P).n_levels) {                 \
    case 1: MACRO(1); break;              \
    case 2: MACRO(2); break;              \
    case 3: MACRO(3); break;              \
    case 4: MACRO(4); break;              \
    case 5: MACRO(5); break;              \
    case 6: MACRO(6); break;              \
    case 7: MACRO(7); break;              \
    default: return hipErrorInvalidValue; \
  }

// Kernel variants of the result writers.
enum { MODE_COUNT = 0, MODE_WRITE = 1, MODE_WRITE_HASH = 2, MODE_HASH = 3 };
template <int MODE> constexpr bool mode_mat() { return MODE == MODE_WRITE || MODE == MODE_WRITE_HASH; }
template <int MODE> constexpr bool mode_hash() { return MODE == MODE_HASH || MODE == MODE_WRITE_HASH; }
inline int write_mode(const ProbeOut& out, bool hash) {
  if (out.cols[0] == nullptr) return MODE_HASH;
  return hash ? MODE_WRITE_HASH : MODE_WRITE;
}

inline unsigned grid_rows(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return g ? (unsigned)g : 1u;
}

}  // namespace amj
