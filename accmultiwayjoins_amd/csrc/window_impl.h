// Output-window writer for linear chain joins (the frequent-item path).
//
// With the subtree sizes of the frequent-item pass every result has a known absolute row, so
// the output is cut into equal windows and each wave materialises one window exactly: skewed
// probe rows (a Zipf head key with millions of results) are spread over many waves, and no
// count pass, queue or atomic is needed. The traversal is the frontier DFS of probe.hip, but
// every child is created from a precomputed 16-byte record (the value the row binds and the run
// it leads to in the next table), so creating 64 children costs one round of independent loads.
//
// The last hop reads a *suffix table*: S_k holds, for every row of build table k in table
// order, all tuples (val_k, val_{k+1}, ...) below it, i.e. the frequent-item buffer
// (SubfixBuffer, MHJ-GPU/SubfixBuffer.cuh:21-98) materialised for every join value. With S_k
// resident, the results of an entry of the last frontier level are one contiguous S_k run, so
// the result stream is produced from coalesced reads. The suffix tables are built with this same
// kernel (table t as the probe, S_{t+1} as the suffix).
//
// A linear chain has a fixed attribute layout after first-occurrence numbering: the probe
// columns are attributes 0..PC-1, build table t binds attribute PC+t, and the NV suffix
// columns follow. The kernels are specialised on (L, PC, NV) so every binding lives in a
// statically indexed register and every column address is base + a * stride with a constant a.
#pragma once

#include "probe_common.h"

namespace amj {

// One level of the frontier. Entry j of a level-t chunk owns the children rows
// [lo_j, lo_j + len_j) of table t (of the suffix table at the last level).
template <int NB>
struct ChainLv {
  int32_t b[NB];   // bindings (attributes 0..NB-1)
  uint32_t len;    // children of the entry
  uint32_t base;   // lo - excl (mod 2^32): row of flattened child g is base + g
  uint32_t excl;   // exclusive prefix of len over the wave
  uint32_t T;      // level total (uniform; always below 2^32 here)
};

struct ChainOut {
  int32_t* base;     // column a starts at base + a * stride
  uint64_t stride;
};

typedef __attribute__((address_space(3))) int32_t lds_i32;

// Results are staged per wave in an LDS ring of two kBlk-row blocks (column a at
// stage[a * kStageRows ..]) and leave in whole blocks: per column, kBlk / 64 stores of 256
// aligned bytes issued back to back, so each column receives 1 KB at a time. Measured against
// 64-row blocks: 21.9-23.2 ms against 23.1-23.8 ms for the 2^32-result chain on one box,
// 24.5-24.7 against 26.5 ms on another (fewer, longer runs of writes per column; a ring of
// kBlk + 64 rows, which leaves room for one more wave per SIMD, was slower). Windows start at
// multiples of kBlk rows, so a block never straddles two waves.
constexpr int kBlk = 256;
constexpr int kLgBlk = 8;
constexpr int kStageRows = 2 * kBlk;
static_assert((1 << kLgBlk) == kBlk && kBlk % kWave == 0, "blocks are whole batches of 64 rows");
static_assert(kBlk == kWindowAlign, "windows are whole blocks");

// Full 64-result batches the last-level emitter locates before it stores any of them.
constexpr int kEmitGroup = 8;

struct AccW {
  uint64_t csum;      // per-lane checksum partial
  uint64_t next_out;  // absolute row of the next result (uniform)
  uint64_t g_end;     // window end (absolute)
  lds_i32* stage;     // this wave's staging ring
  bool done;          // window complete (uniform)
  bool overflow;      // a level total reached 2^32 (uniform)
};

// Write rows [b*kBlk, b*kBlk + n) of the staging ring to the output columns.
template <int NA>
__device__ __forceinline__ void flush_block(const ChainOut& out, lds_i32* stage, uint64_t b,
                                            uint32_t n) {
  __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
  const uint32_t lane = (uint32_t)lane_id();
  sfor<0, NA>([&](auto a) {
    sfor<0, kBlk / kWave>([&](auto k) {
      const uint32_t r = (uint32_t)k * kWave + lane;
      const uint64_t row = b * kBlk + r;
      const uint32_t slot = (uint32_t)row & (kStageRows - 1);
      if (r < n) out.base[(uint64_t)a * out.stride + row] = stage[a * kStageRows + slot];
    });
  });
  __asm__ __volatile__("" ::: "memory");
}

// Write the complete block b of the staging ring (no lane predicate, so the compiler knows the
// wave issues exactly NA * kBlk / 64 stores).
template <int NA>
__device__ __forceinline__ void flush_full_block(const ChainOut& out, lds_i32* stage,
                                                 uint64_t b) {
  __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
  const uint32_t lane = (uint32_t)lane_id();
  sfor<0, NA>([&](auto a) {
    sfor<0, kBlk / kWave>([&](auto k) {
      const uint64_t row = b * kBlk + (uint32_t)k * kWave + lane;
      const uint32_t slot = (uint32_t)row & (kStageRows - 1);
      out.base[(uint64_t)a * out.stride + row] = stage[a * kStageRows + slot];
    });
  });
  __asm__ __volatile__("" ::: "memory");
}

// exclusive scan of len; returns false when the 64 runs might not sum below 2^32
template <int NB>
__device__ __forceinline__ bool finish_w(ChainLv<NB>& l) {
  if (__ballot(l.len >= (1u << 26)) != 0ull) return false;
  const uint32_t incl = wave_scan_add_dpp(l.len);
  l.excl = incl - l.len;
  l.T = lane63(incl);
  return true;
}

template <int NA>
__device__ __forceinline__ uint64_t hash_chain(const int32_t (&v)[NA]) {
  int32_t w[kMaxAttrs];
  sfor<0, kMaxAttrs>([&](auto a) { w[a] = (a < NA) ? v[(a < NA) ? (int)a : 0] : 0; });
  return hash_tuple(w, NA);
}

// L frontier levels: level 0 = probe rows, level t = rows of table t-1; the children of level
// L-1 are suffix rows. Records: probe_next (run in table 0, or in the suffix when L == 1) and
// next[t] for t <= L-2 (run in table t+1, or in the suffix for t == L-2).
template <int L, int PC, int NV>
struct Chain {
  static constexpr int NB = PC + L - 1;  // bindings at the last level
  static constexpr int NA = NB + NV;     // result attributes

  // Level-0 chunk: probe rows [p0, p0+64).
  static __device__ __forceinline__ bool level0(const Plan& P, uint64_t p0, ChainLv<PC>& l) {
    const uint64_t r = p0 + (uint64_t)lane_id();
    const bool valid = r < P.probe_n;
    RunRef rr{0, 0, 0, 0};
    sfor<0, PC>([&](auto c) { l.b[c] = 0; });
    if (valid) {
      sfor<0, PC>([&](auto c) { l.b[c] = __builtin_nontemporal_load(&P.probe_cols[c][r]); });
      rr = P.probe_next[r];
    }
    l.len = rr.len;
    const bool ok = finish_w(l);
    l.base = rr.lo - l.excl;
    return ok;
  }

  // Children [c, c+64) of level t (rows of table t) become the level-(t+1) chunk: one 16-byte
  // record per child holds its new binding and its own run one level down.
  template <int t>
  static __device__ __forceinline__ bool child(const Plan& P, const ChainLv<PC + t>& par,
                                               uint32_t c, ChainLv<PC + t + 1>& ch) {
    const uint32_t g = c + (uint32_t)lane_id();
    const bool valid = g < par.T;
    const int p = (int)parent_of_slots(par.excl, par.len, c);
    const uint32_t row = (uint32_t)__shfl(par.base, p, kWave) + g;
    sfor<0, PC + t>([&](auto a) { ch.b[a] = __shfl(par.b[a], p, kWave); });
    RunRef rr{0, 0, 0, 0};
    if (valid) rr = P.next[t][row];
    ch.b[PC + t] = rr.val;
    ch.len = rr.len;
    const bool ok = finish_w(ch);
    ch.base = rr.lo - ch.excl;
    return ok;
  }

  // Descent for the window's first chunk at level t: lane 0's entry starts at absolute output
  // e_first and holds g0; returns the flattened index of the child holding g0 and advances
  // e_first to that child's start.
  template <int t, int NBt>
  static __device__ __forceinline__ uint32_t start(const Plan& P, const ChainLv<NBt>& l,
                                                   uint64_t g0, uint64_t& e_first) {
    const uint32_t lo = uniform(l.base);  // excl of lane 0 is 0
    if constexpr (t == L - 1) {
      return g0 <= e_first ? 0u : (uint32_t)(g0 - e_first);  // one result per suffix row
    } else {
      if (g0 <= e_first) return 0u;
      const uint64_t* pw = P.tab[t].pw;
      const uint32_t len = uniform(l.len);
      const uint64_t base = pw[lo];
      const uint64_t target = base + (g0 - e_first);  // last r in [lo, lo+len) with pw[r] <= it
      uint32_t a = lo, b = lo + len;
      while (b - a > 1) {
        const uint32_t mid = a + ((b - a) >> 1);
        if (pw[mid] <= target) a = mid;
        else b = mid;
      }
      e_first += pw[a] - base;
      return a - lo;
    }
  }

  // Results [c0, e1) of a last-level chunk; result g goes to absolute row e0 + g. Two batches
  // of 64 are located (parent search + suffix loads) before either is stored, so the second
  // batch's loads are in flight while the first batch is shuffled and written. While at least
  // kEmitGroup full batches remain, they go as one group with no per-lane predicates: every
  // lane loads, every batch completes exactly one 64-row block, so the compiler can wait for
  // each batch's loads alone (a fixed count of memory operations issued after them) and the
  // loads of the later batches of the group stay in flight while the first ones are stored.
  template <bool MAT, bool HASH>
  static __device__ __forceinline__ void emit(const Plan& P, const ChainLv<NB>& l, uint32_t c0,
                                              uint32_t e1, uint64_t e0, const ChainOut& out,
                                              AccW& acc) {
    const int32_t* sfx = P.sfx_base;
    const uint64_t sst = P.sfx_stride;
    auto locate = [&](uint32_t c, int& p, int32_t (&sv)[NV > 0 ? NV : 1]) {
      p = (int)parent_of_slots(l.excl, l.len, c);
      const uint32_t g = c + (uint32_t)lane_id();
      sfor<0, (NV > 0 ? NV : 1)>([&](auto k) { sv[k] = 0; });
      if constexpr (NV > 0) {
        const uint32_t row = (uint32_t)__shfl(l.base, p, kWave) + g;
        if (g < e1) sfor<0, NV>([&](auto k) { sv[k] = sfx[(uint64_t)k * sst + row]; });
      }
    };
    auto store = [&](uint32_t c, int p, const int32_t (&sv)[NV > 0 ? NV : 1]) {
      int32_t v[NA];
      sfor<0, NB>([&](auto a) { v[a] = __shfl(l.b[a], p, kWave); });
      if constexpr (NV > 0) sfor<0, NV>([&](auto k) { v[NB + k] = sv[k]; });
      const uint32_t g = c + (uint32_t)lane_id();
      if (g < e1) {
        if (HASH) acc.csum += hash_chain<NA>(v);
        if (MAT) {
          const uint32_t slot = (uint32_t)(e0 + g) & (kStageRows - 1);
          sfor<0, NA>([&](auto a) { acc.stage[a * kStageRows + slot] = v[a]; });
        }
      }
      if (MAT) {  // a block completed by this batch leaves for HBM
        const uint32_t m = (e1 - c) < (uint32_t)kWave ? (e1 - c) : (uint32_t)kWave;
        const uint64_t rb = e0 + c, re = rb + m;
        if ((rb >> kLgBlk) != (re >> kLgBlk))
          flush_block<NA>(out, acc.stage, rb >> kLgBlk, (uint32_t)kBlk);
      }
    };
    constexpr uint32_t GW = (uint32_t)(kEmitGroup * kWave);  // results per group
    using GroupVals = int32_t[kEmitGroup][NV > 0 ? NV : 1];
    auto locate_group = [&](uint32_t cg, int (&pg)[kEmitGroup], GroupVals& sg) {
      sfor<0, kEmitGroup>([&](auto i) {
        pg[i] = (int)parent_of_slots(l.excl, l.len, cg + (uint32_t)i * kWave);
        if constexpr (NV > 0) {
          const uint32_t row =
              (uint32_t)__shfl(l.base, pg[i], kWave) + cg + (uint32_t)i * kWave + lane_id();
          sfor<0, NV>([&](auto k) { sg[i][k] = sfx[(uint64_t)k * sst + row]; });
        }
      });
    };
    auto store_group = [&](uint32_t cg, const int (&pg)[kEmitGroup], const GroupVals& sg) {
      sfor<0, kEmitGroup>([&](auto i) {
        const uint32_t ci = cg + (uint32_t)i * kWave;
        int32_t v[NA];
        sfor<0, NB>([&](auto a) { v[a] = __shfl(l.b[a], pg[i], kWave); });
        if constexpr (NV > 0) sfor<0, NV>([&](auto k) { v[NB + k] = sg[i][k]; });
        if (HASH) acc.csum += hash_chain<NA>(v);
        if (MAT) {
          const uint32_t slot = (uint32_t)(e0 + ci + lane_id()) & (kStageRows - 1);
          sfor<0, NA>([&](auto a) { acc.stage[a * kStageRows + slot] = v[a]; });
          // the batch's rows [r, r + 64) complete r's block when r lies in its last 64 rows
          if ((((e0 + ci) >> 6) & (kBlk / kWave - 1)) == kBlk / kWave - 1)
            flush_full_block<NA>(out, acc.stage, (e0 + ci) >> kLgBlk);
        }
      });
    };
    uint32_t c = c0;
    const uint32_t n_g = (e1 - c0) / GW;
    for (uint32_t k = 0; k < n_g; ++k, c += GW) {
      int pg[kEmitGroup];
      GroupVals sg;
      locate_group(c, pg, sg);
      store_group(c, pg, sg);
    }
    for (; c < e1; c += 2u * (uint32_t)kWave) {
      const uint32_t cb = c + (uint32_t)kWave;
      int pa, pb = 0;
      int32_t sa[NV > 0 ? NV : 1], sb[NV > 0 ? NV : 1];
      locate(c, pa, sa);
      if (cb < e1) locate(cb, pb, sb);
      store(c, pa, sa);
      if (cb < e1) store(cb, pb, sb);
    }
  }

  template <int t, bool MAT, bool HASH>
  static __device__ __forceinline__ void descend(const Plan& P, const ChainLv<PC + t>& cur,
                                                 bool first, uint64_t e_first, uint64_t g0,
                                                 const ChainOut& out, AccW& acc) {
    const uint32_t c0 = first ? start<t>(P, cur, g0, e_first) : 0u;
    if constexpr (t == L - 1) {
      const uint64_t left = acc.g_end - acc.next_out;
      uint32_t e1 = cur.T;
      if ((uint64_t)(e1 - c0) >= left) {
        e1 = c0 + (uint32_t)left;
        acc.done = true;
      }
      if (c0 < e1) {
        emit<MAT, HASH>(P, cur, c0, e1, acc.next_out - c0, out, acc);
        acc.next_out += e1 - c0;
      }
    } else {
      for (uint32_t c = c0; c < cur.T && !acc.done && !acc.overflow; c += (uint32_t)kWave) {
        ChainLv<PC + t + 1> ch;
        if (!child<t>(P, cur, c, ch)) {
          acc.overflow = true;
          break;
        }
        if (ch.T) descend<t + 1, MAT, HASH>(P, ch, first && c == c0, e_first, g0, out, acc);
      }
    }
  }
};

template <int L, int PC, int NV, bool MAT, bool HASH>
__global__ __launch_bounds__(kBlock) void window_write_kernel(
    const Plan P, const uint64_t* __restrict__ probe_off, const uint64_t* __restrict__ first_row,
    uint64_t total, uint64_t window, uint64_t n_windows, const ChainOut out,
    uint64_t* __restrict__ block_sums, uint32_t* __restrict__ overflow) {
  using C = Chain<L, PC, NV>;
  constexpr int NA = Chain<L, PC, NV>::NA;
  __shared__ int32_t stage_mem[MAT ? kWavesPerBlock * NA * kStageRows : 1];
  const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);  // neighbouring windows per XCD
  // windows w, w + n_slots, ...: the grid is capped (a launch holds fewer than 2^32 lanes)
  const uint64_t n_slots = (uint64_t)gridDim.x * kWavesPerBlock;
  AccW acc{0, 0, 0, nullptr, false, false};
  acc.stage = (lds_i32*)&stage_mem[MAT ? (threadIdx.x / kWave) * NA * kStageRows : 0];
  for (uint64_t w = blk * kWavesPerBlock + threadIdx.x / kWave; w < n_windows && !acc.overflow;
       w += n_slots) {
    const uint64_t g0 = w * window;
    acc.next_out = g0;
    acc.g_end = (g0 + window < total) ? g0 + window : total;
    acc.done = false;
    uint64_t p0 = uniform(first_row[w]);
    const uint64_t e_first = uniform(probe_off[p0]);
    bool first = true;
    while (!acc.done && !acc.overflow && p0 < P.probe_n) {
      ChainLv<PC> l0;
      if (!C::level0(P, p0, l0)) {
        acc.overflow = true;
        break;
      }
      if (l0.T) C::template descend<0, MAT, HASH>(P, l0, first, e_first, g0, out, acc);
      first = false;
      p0 += kWave;
    }
    if (MAT && !acc.overflow && (acc.next_out & (kBlk - 1)) != 0)  // partial last block
      flush_block<NA>(out, acc.stage, acc.next_out >> kLgBlk,
                      (uint32_t)(acc.next_out & (kBlk - 1)));
  }
  if (acc.overflow && lane_id() == 0) atomicOr(overflow, 1u);
  block_sum_store(acc.csum, &block_sums[blockIdx.x]);
}

struct WindowArgs {
  const uint64_t* probe_off;
  const uint64_t* first_row;
  uint64_t total, window, n_windows;
  ChainOut out;
  uint64_t* block_sums;
  uint32_t* overflow;
};

template <int L, int PC, int NV>
void launch_chain(int mode, dim3 grid, hipStream_t s, const Plan& P, const WindowArgs& a) {
  if constexpr (PC + (L - 1) + NV <= kMaxAttrs) {
#define AMJ_WW(M, H)                                                                          \
  hipLaunchKernelGGL((window_write_kernel<L, PC, NV, M, H>), grid, dim3(kBlock), 0, s, P,    \
                     a.probe_off, a.first_row, a.total, a.window, a.n_windows, a.out,         \
                     a.block_sums, a.overflow)
    if (mode == MODE_WRITE) AMJ_WW(true, false);
    else if (mode == MODE_WRITE_HASH) AMJ_WW(true, true);
    else AMJ_WW(false, true);
#undef AMJ_WW
  }
}

// Dispatch over the suffix width for one (L, PC).
template <int L, int PC>
void launch_chain_nv(int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                     const WindowArgs& a) {
  switch (nv) {
    case 0: launch_chain<L, PC, 0>(mode, grid, s, P, a); break;
    case 1: launch_chain<L, PC, 1>(mode, grid, s, P, a); break;
    case 2: launch_chain<L, PC, 2>(mode, grid, s, P, a); break;
    case 3: launch_chain<L, PC, 3>(mode, grid, s, P, a); break;
    case 4: launch_chain<L, PC, 4>(mode, grid, s, P, a); break;
    case 5: launch_chain<L, PC, 5>(mode, grid, s, P, a); break;
    case 6: launch_chain<L, PC, 6>(mode, grid, s, P, a); break;
    case 7: launch_chain<L, PC, 7>(mode, grid, s, P, a); break;
    default: break;
  }
}

// One translation unit per L (window_l*.hip) instantiates these.
void launch_window_l1(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);
void launch_window_l2(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);
void launch_window_l3(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);
void launch_window_l4(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);
void launch_window_l5(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);
void launch_window_l6(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);
void launch_window_l7(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P,
                      const WindowArgs& a);

#define AMJ_DEFINE_WINDOW_L(LV)                                                                \
  void launch_window_l##LV(int pc, int nv, int mode, dim3 grid, hipStream_t s, const Plan& P, \
                           const WindowArgs& a) {                                             \
    if (pc == 1) launch_chain_nv<LV, 1>(nv, mode, grid, s, P, a);                              \
    else launch_chain_nv<LV, 2>(nv, mode, grid, s, P, a);                                      \
  }

}  // namespace amj
