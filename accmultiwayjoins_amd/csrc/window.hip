// Output-window writer for linear chain joins (the frequent-item path).
//
// With the subtree sizes of the frequent-item pass every result has a known absolute row, so
// the output is cut into equal windows and each wave materialises one window exactly: skewed
// probe rows (a Zipf head key with millions of results) are spread over many waves, and no
// count pass, queue or atomic is needed. The traversal is the frontier DFS of probe.hip, but
// every child is created from the precomputed run pointers (next_t[row] = {lo, len, PW[lo]}),
// so creating 64 children costs one round of independent loads (value, weight prefix, run
// pointer) instead of a dependent value -> bucket -> key chain.
#include "probe_common.h"

namespace amj {

// One level of the frontier. Entry j of a level-t chunk owns the children rows
// [lo_j, lo_j + len_j) of table t; its subtree starts at absolute output eoff_j.
struct LvW {
  int32_t b[kMaxAttrs];  // bindings of the entry
  uint32_t len;          // children of the entry
  uint32_t base;         // lo - excl (mod 2^32): row of flattened child g is base + g
  uint32_t excl;         // exclusive prefix of len over the wave
  uint32_t T;            // level total (uniform; always below 2^32 here)
  uint64_t pb;           // eoff - PW_t[lo]: output offset of child row r is pb + PW_t[r]
};

struct AccW {
  uint64_t csum;   // per-lane checksum partial
  uint64_t g_end;  // window end (absolute)
  bool done;       // window complete (uniform)
  bool overflow;   // a level total reached 2^32 (uniform)
};

__device__ __forceinline__ uint64_t pwt(const Plan& P, int t, uint32_t r) {
  return (t == P.n_levels - 1) ? (uint64_t)r : P.tab[t].pw[r];
}

// exclusive scan of len; returns false when the 64 runs might not sum below 2^32
__device__ __forceinline__ bool finish_w(LvW& l) {
  if (__ballot(l.len >= (1u << 26)) != 0ull) return false;
  const uint32_t incl = wave_scan_add_dpp(l.len);
  l.excl = incl - l.len;
  l.T = lane63(incl);
  return true;
}

// Level-0 chunk: probe rows [p0, p0+64).
__device__ __forceinline__ bool make_level0_w(const Plan& P, const uint64_t* probe_off,
                                              uint64_t p0, LvW& l) {
  const uint64_t r = p0 + (uint64_t)lane_id();
  const bool valid = r < P.probe_n;
  sfor<0, kMaxAttrs>([&](auto a) { l.b[a] = 0; });
  for (int c = 0; c < P.probe_ncols; ++c)
    set_attr(l.b, P.probe_attr[c], valid ? P.probe_cols[c][r] : 0);
  RunRef rr{0, 0, 0};
  uint64_t eoff = 0;
  if (valid) {
    rr = P.probe_next[r];
    eoff = probe_off[r];
  }
  l.len = rr.len;
  l.pb = eoff - rr.pwlo;
  const bool ok = finish_w(l);
  l.base = rr.lo - l.excl;
  return ok;
}

// Children [c, c+64) of level t become the level-(t+1) chunk.
template <int t>
__device__ __forceinline__ bool make_child_w(const Plan& P, const LvW& par, uint32_t c,
                                             LvW& ch) {
  const uint32_t g = c + (uint32_t)lane_id();
  const bool valid = g < par.T;
  const uint32_t p = parent_of_slots(par.excl, par.len, c);
  const uint32_t row = (uint32_t)__shfl(par.base, (int)p, kWave) + g;
  const uint64_t pbp = __shfl(par.pb, (int)p, kWave);
  sfor<0, kMaxAttrs>([&](auto a) {
    ch.b[a] = (a < P.n_attrs) ? __shfl(par.b[a], (int)p, kWave) : 0;
  });
  RunRef rr{0, 0, 0};
  uint64_t eoff = 0;
  if (valid) {
    if (P.val_mode[t] == VAL_BIND) set_attr(ch.b, P.val_attr[t], P.tab[t].vals[row]);
    rr = P.next[t][row];
    eoff = pbp + P.tab[t].pw[row];
  }
  ch.len = rr.len;
  ch.pb = eoff - rr.pwlo;
  const bool ok = finish_w(ch);
  ch.base = rr.lo - ch.excl;
  return ok;
}

// First child (flattened index) of lane 0's entry whose subtree holds absolute output g0.
__device__ __forceinline__ uint32_t window_start(const Plan& P, int t, const LvW& l,
                                                 uint64_t g0) {
  const uint32_t lo = uniform(l.base);  // excl of lane 0 is 0
  const uint64_t pb = uniform(l.pb);
  const uint64_t e0 = pb + pwt(P, t, lo);
  if (g0 <= e0) return 0;
  if (t == P.n_levels - 1) return (uint32_t)(g0 - e0);  // one result per row of the last table
  const uint32_t len = uniform(l.len);
  const uint64_t* pw = P.tab[t].pw;
  const uint64_t target = g0 - pb;  // last r in [lo, lo+len) with pw[r] <= target
  uint32_t a = lo, b = lo + len;
  while (b - a > 1) {
    const uint32_t mid = a + ((b - a) >> 1);
    if (pw[mid] <= target) a = mid;
    else b = mid;
  }
  return a - lo;
}

// Results [c0, e1) of a last-level chunk; result g goes to absolute row e0 + g. The parent
// search and value load of the next 64 results are issued before the current 64 are stored.
template <bool MAT, bool HASH>
__device__ __forceinline__ void emit_w(const Plan& P, int t,
                                       const LvW& l, uint32_t c0, uint32_t e1, uint64_t e0,
                                       const ProbeOut& out, AccW& acc) {
  const bool bind = P.val_mode[t] == VAL_BIND;
  const int vattr = P.val_attr[t];
  const int32_t* vals = P.tab[t].vals;
  const int n_attrs = P.n_attrs;
  auto locate = [&](uint32_t c, uint32_t& p, int32_t& val) {
    p = parent_of_slots(l.excl, l.len, c);
    const uint32_t g = c + (uint32_t)lane_id();
    const uint32_t row = (uint32_t)__shfl(l.base, (int)p, kWave) + g;
    val = (bind && g < e1) ? vals[row] : 0;
  };
  uint32_t p;
  int32_t val;
  locate(c0, p, val);
  for (uint32_t c = c0; c < e1;) {
    const uint32_t cn = c + (uint32_t)kWave;
    uint32_t pn = 0;
    int32_t valn = 0;
    if (cn < e1) locate(cn, pn, valn);
    int32_t v[kMaxAttrs];
    sfor<0, kMaxAttrs>([&](auto a) { v[a] = (a < n_attrs) ? __shfl(l.b[a], (int)p, kWave) : 0; });
    if (bind) set_attr(v, vattr, val);
    const uint32_t g = c + (uint32_t)lane_id();
    if (g < e1) {
      if (HASH) acc.csum += hash_tuple(v, n_attrs);
      if (MAT) {
        const uint64_t pos = e0 + g;
        sfor<0, kMaxAttrs>([&](auto a) {
          if (a < n_attrs) out.cols[a][pos] = v[a];
        });
      }
    }
    c = cn;
    p = pn;
    val = valn;
  }
}

template <int L, int t, bool MAT, bool HASH>
__device__ __forceinline__ void descend_w(const Plan& P, const LvW& cur,
                                          bool first, uint64_t g0, const ProbeOut& out,
                                          AccW& acc) {
  const uint32_t c0 = first ? window_start(P, t, cur, g0) : 0u;
  if constexpr (t == L - 1) {
    const uint64_t e0 = uniform(cur.pb) + uniform(cur.base);  // eoff of lane 0
    uint32_t e1 = cur.T;
    if (e0 + e1 >= acc.g_end) {
      e1 = (uint32_t)(acc.g_end - e0);
      acc.done = true;
    }
    if (c0 < e1) emit_w<MAT, HASH>(P, t, cur, c0, e1, e0, out, acc);
  } else {
    for (uint32_t c = c0; c < cur.T && !acc.done && !acc.overflow; c += (uint32_t)kWave) {
      LvW ch;
      if (!make_child_w<t>(P, cur, c, ch)) {
        acc.overflow = true;
        break;
      }
      if (ch.T) descend_w<L, t + 1, MAT, HASH>(P, ch, first && c == c0, g0, out, acc);
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

// Probe row holding the first result of every window (one independent search per thread).
__global__ __launch_bounds__(kBlock) void window_first_row_kernel(
    const uint64_t* __restrict__ probe_off, uint64_t n0, uint64_t window, uint64_t n_windows,
    uint64_t* __restrict__ first_row) {
  for (uint64_t w = (uint64_t)blockIdx.x * kBlock + threadIdx.x; w < n_windows;
       w += (uint64_t)gridDim.x * kBlock)
    first_row[w] = upper_bound_u64(probe_off, n0 + 1, w * window) - 1;
}

template <int L, bool MAT, bool HASH>
__global__ __launch_bounds__(kBlock) void window_write_kernel(
    const Plan P, const uint64_t* __restrict__ probe_off, const uint64_t* __restrict__ first_row,
    uint64_t total, uint64_t window, uint64_t n_windows, const ProbeOut out,
    uint64_t* __restrict__ block_sums, uint32_t* __restrict__ overflow) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  AccW acc{0, 0, false, false};
  if (w < n_windows) {
    const uint64_t g0 = w * window;
    acc.g_end = (g0 + window < total) ? g0 + window : total;
    uint64_t p0 = uniform(first_row[w]);
    bool first = true;
    while (!acc.done && !acc.overflow && p0 < P.probe_n) {
      LvW l0;
      if (!make_level0_w(P, probe_off, p0, l0)) {
        acc.overflow = true;
        break;
      }
      if (l0.T) descend_w<L, 0, MAT, HASH>(P, l0, first, g0, out, acc);
      first = false;
      p0 += kWave;
    }
    if (acc.overflow && lane_id() == 0) atomicOr(overflow, 1u);
  }
  block_sum_store(acc.csum, &block_sums[blockIdx.x]);
}

uint64_t window_blocks(uint64_t total, uint64_t window) {
  const uint64_t n_windows = (total + window - 1) / window;
  return n_windows == 0 ? 1 : (n_windows + kWavesPerBlock - 1) / kWavesPerBlock;
}

hipError_t launch_window_write(const Plan& P, const uint64_t* probe_off, uint64_t* first_row,
                               uint64_t total, uint64_t window, const ProbeOut& out, bool hash,
                               uint64_t* block_sums, uint32_t* overflow, uint64_t* n_blocks_out,
                               hipStream_t s) {
  const uint64_t n_windows = (total + window - 1) / window;
  const uint64_t blocks = window_blocks(total, window);
  *n_blocks_out = blocks;
  hipLaunchKernelGGL(window_first_row_kernel, dim3(grid_rows(n_windows)), dim3(kBlock), 0, s,
                     probe_off, P.probe_n, window, n_windows, first_row);
  const int mode = write_mode(out, hash);
#define LAUNCH_M(LV, M, H)                                                                    \
  hipLaunchKernelGGL((window_write_kernel<LV, M, H>), dim3((unsigned)blocks), dim3(kBlock), 0, \
                     s, P, probe_off, first_row, total, window, n_windows, out, block_sums,    \
                     overflow)
#define LAUNCH(LV)                                            \
  if (mode == MODE_WRITE) LAUNCH_M(LV, true, false);          \
  else if (mode == MODE_WRITE_HASH) LAUNCH_M(LV, true, true); \
  else LAUNCH_M(LV, false, true)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
#undef LAUNCH_M
  return hipGetLastError();
}

}  // namespace amj
