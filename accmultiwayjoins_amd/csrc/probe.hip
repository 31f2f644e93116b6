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
//     next 64 children are created at once: each lane finds its parent, copies the parent's
//     bindings with shuffles and binds the new value;
//   * the last level never materialises a frontier: its children are the result tuples and are
//     emitted 64 at a time, so each column store is one 256-byte coalesced write.
// Every emitted child is a real match, so no lane is ever idle while a level has work. The
// DFS order over probe rows is deterministic, hence so is the output order.
//
// This file holds the unit engine (count pass + write pass, used for every query shape) and the
// frequent-item precomputation of the output-window writer (window.hip).
#include "probe_engine.h"

namespace amj {

// ------------------------------------------------------------------------------------------
// Unit engine: count pass and write pass over one unit of 64 probe rows.
// ------------------------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(kBlock) void probe_count_kernel(const Plan P,
                                                             uint64_t* __restrict__ counts,
                                                             uint64_t units) {
  const uint64_t unit = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (unit >= units) return;  // whole wave leaves; this kernel has no barriers
  LvL<L> l0;
  make_level0(P, unit * kWave, l0);
  Acc acc{};
  if (l0.T) descend<L, 0, MODE_COUNT>(P, l0, ProbeOut{}, acc);
  if (lane_id() == 0) counts[unit] = acc.count;
}

// Write pass at the scanned bases; or, with base == nullptr (nothing materialised), the only
// pass: it hashes every result and leaves per-block result counts in block_counts.
template <int L, int MODE>
__global__ __launch_bounds__(kBlock) void probe_write_kernel(const Plan P,
                                                             const uint64_t* __restrict__ base,
                                                             const ProbeOut out,
                                                             uint64_t* __restrict__ block_sums,
                                                             uint64_t* __restrict__ block_counts,
                                                             uint64_t units) {
  const uint64_t unit = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  Acc acc{};
  if (unit < units) {
    LvL<L> l0;
    make_level0(P, unit * kWave, l0);
    acc.out_base = base ? uniform(base[unit]) : 0;
    if (l0.T) descend<L, 0, MODE>(P, l0, out, acc);
  }
  // acc.count is uniform: lane 0 carries the wave's
  block_sum_store2(acc.csum, lane_id() == 0 ? acc.count : 0, &block_sums[blockIdx.x],
                   block_counts ? &block_counts[blockIdx.x] : nullptr);
}

uint64_t probe_units(const Plan& P) { return (P.probe_n + kWave - 1) / kWave; }
uint64_t probe_write_blocks(const Plan& P) {
  const uint64_t u = probe_units(P);
  return u == 0 ? 1 : (u + kWavesPerBlock - 1) / kWavesPerBlock;
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
                              bool hash, uint64_t* block_sums, uint64_t* n_blocks_out,
                              hipStream_t s, uint64_t* block_counts) {
  const uint64_t units = probe_units(P);
  const unsigned grid = (unsigned)probe_write_blocks(P);
  *n_blocks_out = grid;
  const int mode = write_mode(out, hash);
#define LAUNCH_M(LV, M)                                                                 \
  hipLaunchKernelGGL((probe_write_kernel<LV, M>), dim3(grid), dim3(kBlock), 0, s, P, base, \
                     out, block_sums, block_counts, units)
#define LAUNCH(LV)                                                 \
  if (mode == MODE_WRITE) LAUNCH_M(LV, MODE_WRITE);                \
  else if (mode == MODE_WRITE_HASH) LAUNCH_M(LV, MODE_WRITE_HASH); \
  else LAUNCH_M(LV, MODE_HASH)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
#undef LAUNCH_M
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
// prefix. The subtree size of any run [lo, hi) is then PW_t[hi] - PW_t[lo]. The same pass
// records next_t[r] = {lo, len, PW_{t+1}[lo]}, the run in table t+1 that row r leads to, so the
// output-window writer never repeats a hash lookup.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t pw_at(const Plan& P, int t, uint32_t x) {
  return (t == P.n_levels - 1) ? (uint64_t)x : P.tab[t].pw[x];
}

__device__ __forceinline__ uint64_t run_of(const Plan& P, int t, int32_t key, RunRef* out) {
  uint32_t lo, len;
  lookup_run(P.tab[t], key, false, 0, true, lo, len);
  out->val = key;
  out->lo = lo;
  out->len = len;
  out->pad = 0;
  return pw_at(P, t, lo + len) - pw_at(P, t, lo);
}

// rows of table t: w[r] = results below row r (w[n] = 0), next[r] = run in table t+1
__global__ __launch_bounds__(kBlock) void fib_weights_kernel(const Plan P, int t,
                                                             uint64_t* __restrict__ w,
                                                             RunRef* __restrict__ next) {
  const DevTable& tb = P.tab[t];
  const uint64_t n = tb.n;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r <= n;
       r += (uint64_t)gridDim.x * kBlock) {
    if (r < n) {
      RunRef rr;
      w[r] = run_of(P, t + 1, tb.vals[r], &rr);
      next[r] = rr;
    } else {
      w[r] = 0;
    }
  }
}

__global__ __launch_bounds__(kBlock) void fib_probe_weights_kernel(const Plan P,
                                                                   uint64_t* __restrict__ w,
                                                                   RunRef* __restrict__ next) {
  const uint64_t n = P.probe_n;
  int kcol = 0;
  for (int c = 0; c < P.probe_ncols; ++c)
    if (P.probe_attr[c] == P.key_attr[0]) kcol = c;
  const int32_t* key = P.probe_cols[kcol];
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r <= n;
       r += (uint64_t)gridDim.x * kBlock) {
    if (r < n) {
      RunRef rr;
      w[r] = run_of(P, 0, key[r], &rr);
      next[r] = rr;
    } else {
      w[r] = 0;
    }
  }
}

hipError_t launch_fib_weights(const Plan& P, int t, uint64_t* w, RunRef* next, hipStream_t s) {
  hipLaunchKernelGGL(fib_weights_kernel, dim3(grid_rows(P.tab[t].n + 1)), dim3(kBlock), 0, s,
                     P, t, w, next);
  return hipGetLastError();
}
hipError_t launch_fib_probe_weights(const Plan& P, uint64_t* w, RunRef* next, hipStream_t s) {
  hipLaunchKernelGGL(fib_probe_weights_kernel, dim3(grid_rows(P.probe_n + 1)), dim3(kBlock), 0,
                     s, P, w, next);
  return hipGetLastError();
}

}  // namespace amj
