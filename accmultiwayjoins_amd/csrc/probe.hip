// The AMHJ multi-way probe for gfx950: a wave64 "frontier DFS".
//
// Reference: AMHJ_count / AMHJ_write (MHJ-GPU/Joins/AMHJ.cuh:20-280) let one 32-lane warp walk
// one probe tuple depth-first through the bucket chains of all build tables, compacting
// matches 32 at a time (probe(), MHJ-GPU/common_kernels.cuh:27-67). Most lanes idle when buckets
// are short, and outputs are written one tuple per atomic.
//
// Here a wave owns a *frontier* of 64 partial tuples per level (one per lane):
//   * level 0 holds 64 consecutive probe rows (a "unit");
//   * every lane looks up its key in table t and gets the contiguous run [lo, lo+len) of
//     matching rows (the build layout makes each key's rows contiguous, and a compared second
//     column is found by binary search inside the run);
//   * a wave-wide exclusive scan of len gives the flattened child sequence of the level; the
//     next 64 children are created at once: lane j finds its parent with a 6-step shuffle
//     search over the scan, copies the parent's bindings with shuffles and binds the new value;
//   * the last level never materialises a frontier: its children are the result tuples and are
//     emitted 64 at a time, so each column store is one 256-byte coalesced write.
// Every emitted child is a real match, so no lane is ever idle while a level has work. The
// DFS order over probe rows is deterministic, hence so is the output order.
#include <type_traits>

#include "amj_internal.h"

namespace amj {

// ------------------------------------------------------------------------------------------
// Attribute access with a wave-uniform runtime attribute index (lowered to v_cndmask chains,
// keeping the bindings in VGPRs instead of scratch).
// ------------------------------------------------------------------------------------------
// Compile-time unrolled loop over I in [0, N): every index is a constant in the IR from the
// start, so SROA keeps per-attribute arrays in VGPRs.
template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, N>(f);
  }
}

template <int I = 0>
__device__ __forceinline__ int32_t get_attr(const int32_t (&b)[kMaxAttrs], int a) {
  if constexpr (I == kMaxAttrs - 1) return b[I];
  else return a == I ? b[I] : get_attr<I + 1>(b, a);
}
__device__ __forceinline__ void set_attr(int32_t (&b)[kMaxAttrs], int a, int32_t v) {
  sfor<0, kMaxAttrs>([&](auto i) { b[i] = (a == (int)i) ? v : b[i]; });
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
    const uint32_t s2 = lower_bound_i32(tb.vals, s, e, cval);
    e = upper_bound_i32(tb.vals, s2, e, cval);
    s = s2;
  }
  lo = s;
  len = e - s;
}

// Last lane p whose exclusive prefix is <= g (lane 0 always qualifies); 6 shuffle steps.
__device__ __forceinline__ int find_parent(uint64_t excl, uint64_t g) {
  int p = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const uint64_t v = __shfl(excl, p + step, kWave);
    if (v <= g) p += step;
  }
  return p;
}

// Tuple hash over ceil(n_attrs/2) SplitMix64 rounds; attribute pairs are packed in one word.
__device__ __forceinline__ uint64_t hash_tuple(const int32_t (&v)[kMaxAttrs], int n_attrs) {
  uint64_t h = (uint64_t)n_attrs;
  sfor<0, kMaxAttrs / 2>([&](auto j) {
    constexpr int i = 2 * (int)j;
    if (i < n_attrs) {
      const uint64_t hi = (i + 1 < n_attrs) ? (uint64_t)(uint32_t)v[i + 1] : 0ull;
      h = mix64(h ^ ((uint64_t)(uint32_t)v[i] | (hi << 32)));
    }
  });
  return h;
}

// ------------------------------------------------------------------------------------------
// Frontier state of one level (per lane) and its wave-uniform totals.
// ------------------------------------------------------------------------------------------
struct Lv {
  int32_t b[kMaxAttrs];  // bindings of this lane's partial tuple
  uint32_t lo;           // run of matches in this level's table
  uint32_t len;
  uint64_t excl;         // exclusive prefix of len over the wave
  uint64_t T;            // total children of the level (uniform)
  uint64_t eoff;         // absolute output offset of the entry (output-window writer only)
  uint64_t pwlo;         // weighted prefix at lo (output-window writer only)
};

struct Acc {
  uint64_t count;     // results counted (uniform)
  uint64_t out_base;  // absolute output row of the next result (uniform)
  uint64_t csum;      // per-lane checksum partial
  uint64_t g_end;     // output-window end (absolute)
  bool done;          // output window complete (uniform)
};

__device__ __forceinline__ void finish_level(Lv& l) {
  const uint64_t incl = wave_incl_scan_u64((uint64_t)l.len);
  l.excl = incl - l.len;
  l.T = uniform(__shfl(incl, kWave - 1, kWave));
}

__device__ __forceinline__ bool lookup_level(const Plan& P, int t, Lv& l, bool valid) {
  const int32_t key = get_attr(l.b, P.key_attr[t]);
  const bool check = P.val_mode[t] == VAL_CHECK;
  const int32_t cval = check ? get_attr(l.b, P.val_attr[t]) : 0;
  lookup_run(P.tab[t], key, check, cval, valid, l.lo, l.len);
  return valid;
}

__device__ __forceinline__ void make_level0(const Plan& P, uint64_t row0, Lv& l) {
  const uint64_t r = row0 + (uint64_t)lane_id();
  const bool valid = r < P.probe_n;
  sfor<0, kMaxAttrs>([&](auto a) { l.b[a] = 0; });
  for (int c = 0; c < P.probe_ncols; ++c)
    set_attr(l.b, P.probe_attr[c], valid ? P.probe_cols[c][r] : 0);
  lookup_level(P, 0, l, valid);
  finish_level(l);
}

// Children [c, c+64) of level t become the level-(t+1) frontier.
template <int t>
__device__ __forceinline__ void make_child(const Plan& P, const Lv& par, uint64_t c, Lv& ch) {
  const uint64_t g = c + (uint64_t)lane_id();
  const bool valid = g < par.T;
  const uint64_t gg = valid ? g : par.T - 1;
  const int p = find_parent(par.excl, gg);
  const uint32_t plo = __shfl(par.lo, p, kWave);
  const uint64_t pex = __shfl(par.excl, p, kWave);
  const uint32_t row = plo + (uint32_t)(gg - pex);
  sfor<0, kMaxAttrs>([&](auto a) { ch.b[a] = (a < P.n_attrs) ? __shfl(par.b[a], p, kWave) : 0; });
  if (P.val_mode[t] == VAL_BIND) set_attr(ch.b, P.val_attr[t], P.tab[t].vals[row]);
  lookup_level(P, t + 1, ch, valid);
  finish_level(ch);
}

// Emit the results [e0, e1) of the last level. Result e goes to absolute row out_row0 + e.
template <bool MATERIALIZE>
__device__ __forceinline__ void emit(const Plan& P, int t, const Lv& l, uint64_t e0, uint64_t e1,
                                     uint64_t out_row0, const ProbeOut& out, Acc& acc) {
  const bool bind = P.val_mode[t] == VAL_BIND;
  const int vattr = P.val_attr[t];
  const int32_t* vals = P.tab[t].vals;
  for (uint64_t c = e0; c < e1; c += kWave) {
    const uint64_t g = c + (uint64_t)lane_id();
    const bool valid = g < e1;
    const uint64_t gg = valid ? g : e1 - 1;
    const int p = find_parent(l.excl, gg);
    const uint32_t plo = __shfl(l.lo, p, kWave);
    const uint64_t pex = __shfl(l.excl, p, kWave);
    const uint32_t row = plo + (uint32_t)(gg - pex);
    int32_t v[kMaxAttrs];
    sfor<0, kMaxAttrs>([&](auto a) { v[a] = (a < P.n_attrs) ? __shfl(l.b[a], p, kWave) : 0; });
    if (bind) set_attr(v, vattr, vals[row]);
    if (valid) {
      acc.csum += hash_tuple(v, P.n_attrs);
      if (MATERIALIZE) {
        const uint64_t pos = out_row0 + g;
        if (pos < out.capacity) {
          sfor<0, kMaxAttrs>([&](auto a) {
            if (a < P.n_attrs) out.cols[a][pos] = v[a];
          });
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// Unit engine: count pass and write pass over one unit of 64 probe rows.
// ------------------------------------------------------------------------------------------
enum { MODE_COUNT = 0, MODE_WRITE = 1, MODE_HASH = 2 };

template <int L, int t, int MODE>
__device__ __forceinline__ void descend(const Plan& P, const Lv& cur, const ProbeOut& out,
                                        Acc& acc) {
  if constexpr (t == L - 1) {
    if constexpr (MODE == MODE_COUNT) {
      acc.count += cur.T;
    } else {
      emit<MODE == MODE_WRITE>(P, t, cur, 0, cur.T, acc.out_base, out, acc);
      acc.out_base += cur.T;
      acc.count += cur.T;
    }
  } else {
    for (uint64_t c = 0; c < cur.T; c += kWave) {
      Lv ch;
      make_child<t>(P, cur, c, ch);
      if (ch.T) descend<L, t + 1, MODE>(P, ch, out, acc);
    }
  }
}

template <int L>
__global__ __launch_bounds__(kBlock) void probe_count_kernel(const Plan P,
                                                             uint64_t* __restrict__ counts,
                                                             uint64_t units) {
  const uint64_t unit = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (unit >= units) return;  // whole wave leaves; this kernel has no barriers
  Lv l0;
  make_level0(P, unit * kWave, l0);
  Acc acc{};
  if (l0.T) descend<L, 0, MODE_COUNT>(P, l0, ProbeOut{}, acc);
  if (lane_id() == 0) counts[unit] = acc.count;
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

template <int L, int MODE>
__global__ __launch_bounds__(kBlock) void probe_write_kernel(const Plan P,
                                                             const uint64_t* __restrict__ base,
                                                             const ProbeOut out,
                                                             uint64_t* __restrict__ block_sums,
                                                             uint64_t units) {
  const uint64_t unit = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  Acc acc{};
  if (unit < units) {
    Lv l0;
    make_level0(P, unit * kWave, l0);
    acc.out_base = uniform(base[unit]);
    if (l0.T) descend<L, 0, MODE>(P, l0, out, acc);
  }
  block_sum_store(acc.csum, &block_sums[blockIdx.x]);
}

uint64_t probe_units(const Plan& P) { return (P.probe_n + kWave - 1) / kWave; }
uint64_t probe_write_blocks(const Plan& P) {
  const uint64_t u = probe_units(P);
  return u == 0 ? 1 : (u + kWavesPerBlock - 1) / kWavesPerBlock;
}

#define AMJ_DISPATCH_L(P, MACRO)          \
  switch ((P).n_levels) {                 \
    case 1: MACRO(1); break;              \
    case 2: MACRO(2); break;              \
    case 3: MACRO(3); break;              \
    case 4: MACRO(4); break;              \
    case 5: MACRO(5); break;              \
    default: return hipErrorInvalidValue; \
  }

hipError_t launch_probe_count(const Plan& P, uint64_t* counts, hipStream_t s) {
  const uint64_t units = probe_units(P);
  if (units == 0) return hipSuccess;
  const unsigned grid = (unsigned)((units + kWavesPerBlock - 1) / kWavesPerBlock);
#define LAUNCH(LV) \
  hipLaunchKernelGGL(probe_count_kernel<LV>, dim3(grid), dim3(kBlock), 0, s, P, counts, units)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
  return hipGetLastError();
}

hipError_t launch_probe_write(const Plan& P, const uint64_t* base, const ProbeOut& out,
                              uint64_t* block_sums, uint64_t* n_blocks_out, hipStream_t s) {
  const uint64_t units = probe_units(P);
  const unsigned grid = (unsigned)probe_write_blocks(P);
  *n_blocks_out = grid;
  const bool mat = out.cols[0] != nullptr;
#define LAUNCH(LV)                                                                          \
  if (mat)                                                                                  \
    hipLaunchKernelGGL((probe_write_kernel<LV, MODE_WRITE>), dim3(grid), dim3(kBlock), 0, s, \
                       P, base, out, block_sums, units);                                    \
  else                                                                                      \
    hipLaunchKernelGGL((probe_write_kernel<LV, MODE_HASH>), dim3(grid), dim3(kBlock), 0, s,  \
                       P, base, out, block_sums, units)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Frequent-item buffer for linear chains: the number of results below every build row.
//
// Reference: AMHJ_WS_FIB (MHJ-GPU/Joins/AMHJ_WS_FIB.cuh:29-916) with SubfixBuffer
// (MHJ-GPU/SubfixBuffer.cuh:21-98) buffers the top-256 frequent join values per column and
// reuses the count (or the written suffix) of the first wave that finished each of them.
// With 288 GB of HBM every join value can be buffered: W_t[r] = results below row r of table
// t, computed bottom-up once per query (one lookup per build row), and PW_t its exclusive
// prefix. The subtree size of any run [lo, hi) is then PW_t[hi] - PW_t[lo].
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t pw_at(const Plan& P, int t, uint32_t x) {
  return (t == P.n_levels - 1) ? (uint64_t)x : P.tab[t].pw[x];
}

__global__ __launch_bounds__(kBlock) void fib_weights_kernel(const Plan P, int t,
                                                             uint64_t* __restrict__ w) {
  const DevTable& tb = P.tab[t];
  const uint64_t n = tb.n;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r <= n;
       r += (uint64_t)gridDim.x * kBlock) {
    uint64_t x = 0;
    if (r < n) {
      uint32_t lo, len;
      lookup_run(P.tab[t + 1], tb.vals[r], false, 0, true, lo, len);
      x = pw_at(P, t + 1, lo + len) - pw_at(P, t + 1, lo);
    }
    w[r] = x;
  }
}

__global__ __launch_bounds__(kBlock) void fib_probe_weights_kernel(const Plan P,
                                                                   uint64_t* __restrict__ w) {
  const uint64_t n = P.probe_n;
  int kcol = 0;
  for (int c = 0; c < P.probe_ncols; ++c)
    if (P.probe_attr[c] == P.key_attr[0]) kcol = c;
  const int32_t* key = P.probe_cols[kcol];
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r <= n;
       r += (uint64_t)gridDim.x * kBlock) {
    uint64_t x = 0;
    if (r < n) {
      uint32_t lo, len;
      lookup_run(P.tab[0], key[r], false, 0, true, lo, len);
      x = pw_at(P, 0, lo + len) - pw_at(P, 0, lo);
    }
    w[r] = x;
  }
}

static unsigned grid_rows(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return g ? (unsigned)g : 1u;
}

hipError_t launch_fib_weights(const Plan& P, int t, uint64_t* w, hipStream_t s) {
  hipLaunchKernelGGL(fib_weights_kernel, dim3(grid_rows(P.tab[t].n + 1)), dim3(kBlock), 0, s,
                     P, t, w);
  return hipGetLastError();
}
hipError_t launch_fib_probe_weights(const Plan& P, uint64_t* w, hipStream_t s) {
  hipLaunchKernelGGL(fib_probe_weights_kernel, dim3(grid_rows(P.probe_n + 1)), dim3(kBlock), 0,
                     s, P, w);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Output-window writer: with exact subtree sizes every result has a known absolute row, so the
// output is cut into equal windows and each wave materialises one window. Skewed probe rows
// (a Zipf head key with millions of results) are spread over many waves without any queue.
// ------------------------------------------------------------------------------------------
template <int t>
__device__ __forceinline__ void make_child_w(const Plan& P, const Lv& par, uint64_t c, Lv& ch) {
  const uint64_t g = c + (uint64_t)lane_id();
  const bool valid = g < par.T;
  const uint64_t gg = valid ? g : par.T - 1;
  const int p = find_parent(par.excl, gg);
  const uint32_t plo = __shfl(par.lo, p, kWave);
  const uint64_t pex = __shfl(par.excl, p, kWave);
  const uint64_t pbase = __shfl(par.eoff - par.pwlo, p, kWave);
  const uint32_t row = plo + (uint32_t)(gg - pex);
  sfor<0, kMaxAttrs>([&](auto a) { ch.b[a] = (a < P.n_attrs) ? __shfl(par.b[a], p, kWave) : 0; });
  if (P.val_mode[t] == VAL_BIND) set_attr(ch.b, P.val_attr[t], P.tab[t].vals[row]);
  ch.eoff = pbase + pw_at(P, t, row);
  lookup_level(P, t + 1, ch, valid);
  ch.pwlo = pw_at(P, t + 1, ch.lo);
  finish_level(ch);
}

// First child of lane 0's run (in table t) whose subtree contains absolute output g0.
__device__ __forceinline__ uint64_t window_start(const Plan& P, int t, const Lv& l, uint64_t g0) {
  const uint64_t e0 = uniform(__shfl(l.eoff, 0, kWave));
  if (g0 <= e0) return 0;
  const uint64_t target = g0 - e0;
  const uint32_t lo = uniform(__shfl(l.lo, 0, kWave));
  const uint32_t len = uniform(__shfl(l.len, 0, kWave));
  if (t == P.n_levels - 1) return target;  // results of the last level are one per row
  const uint64_t base = P.tab[t].pw[lo];
  // last r in [lo, lo+len) with pw[r] - base <= target
  uint32_t a = lo, b = lo + len;  // answer in [a, b)
  while (b - a > 1) {
    const uint32_t mid = a + ((b - a) >> 1);
    if (P.tab[t].pw[mid] - base <= target) a = mid;
    else b = mid;
  }
  return (uint64_t)(a - lo);
}

template <int L, int t, bool MAT>
__device__ __forceinline__ void descend_w(const Plan& P, const Lv& cur, bool first, uint64_t g0,
                                          const ProbeOut& out, Acc& acc) {
  const uint64_t c0 = first ? window_start(P, t, cur, g0) : 0;
  if constexpr (t == L - 1) {
    const uint64_t e0 = uniform(__shfl(cur.eoff, 0, kWave));
    uint64_t e1 = cur.T;
    if (e0 + e1 >= acc.g_end) {
      e1 = acc.g_end - e0;
      acc.done = true;
    }
    if (c0 < e1) {
      emit<MAT>(P, t, cur, c0, e1, e0, out, acc);
      acc.count += e1 - c0;
    }
  } else {
    for (uint64_t c = c0; c < cur.T && !acc.done; c += kWave) {
      Lv ch;
      make_child_w<t>(P, cur, c, ch);
      if (ch.T) descend_w<L, t + 1, MAT>(P, ch, first && c == c0, g0, out, acc);
    }
  }
}

__device__ __forceinline__ uint64_t upper_bound_u64(const uint64_t* a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <int L, bool MAT>
__global__ __launch_bounds__(kBlock) void window_write_kernel(
    const Plan P, const uint64_t* __restrict__ probe_off, uint64_t total, uint64_t window,
    uint64_t n_windows, const ProbeOut out, uint64_t* __restrict__ block_sums,
    uint64_t* __restrict__ win_counts) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  Acc acc{};
  if (w < n_windows) {
    const uint64_t g0 = w * window;
    acc.g_end = (g0 + window < total) ? g0 + window : total;
    // probe row containing g0: last p with probe_off[p] <= g0
    uint64_t p0 = uniform(upper_bound_u64(probe_off, P.probe_n + 1, g0) - 1);
    bool first = true;
    while (!acc.done && p0 < P.probe_n) {
      Lv l0;
      make_level0(P, p0, l0);
      const uint64_t r = p0 + (uint64_t)lane_id();
      l0.eoff = probe_off[r < P.probe_n ? r : P.probe_n];
      l0.pwlo = pw_at(P, 0, l0.lo);
      if (l0.T) descend_w<L, 0, MAT>(P, l0, first, g0, P.n_levels ? out : out, acc);
      first = false;
      p0 += kWave;
    }
    if (lane_id() == 0) win_counts[w] = acc.count;
  }
  block_sum_store(acc.csum, &block_sums[blockIdx.x]);
}

hipError_t launch_window_write(const Plan& P, const uint64_t* probe_off, uint64_t total,
                               uint64_t window, const ProbeOut& out, uint64_t* block_sums,
                               uint64_t* n_blocks_out, hipStream_t s) {
  const uint64_t n_windows = (total + window - 1) / window;
  const uint64_t blocks = n_windows == 0 ? 1 : (n_windows + kWavesPerBlock - 1) / kWavesPerBlock;
  *n_blocks_out = blocks;
  uint64_t* win_counts = block_sums + blocks;  // caller sizes block_sums for both
  const bool mat = out.cols[0] != nullptr;
#define LAUNCH(LV)                                                                            \
  if (mat)                                                                                    \
    hipLaunchKernelGGL((window_write_kernel<LV, true>), dim3((unsigned)blocks), dim3(kBlock), \
                       0, s, P, probe_off, total, window, n_windows, out, block_sums,         \
                       win_counts);                                                           \
  else                                                                                        \
    hipLaunchKernelGGL((window_write_kernel<LV, false>), dim3((unsigned)blocks),              \
                       dim3(kBlock), 0, s, P, probe_off, total, window, n_windows, out,       \
                       block_sums, win_counts)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace amj
