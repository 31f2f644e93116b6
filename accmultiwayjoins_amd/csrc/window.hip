// Host-side dispatch of the output-window writer (kernels in window_impl.h, instantiated per
// frontier depth in window_l1.hip .. window_l5.hip).
#include "window_impl.h"

namespace amj {

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

// Rewrites records {val, lo, len} that point at rows of table t+1 so they point at the rows of
// its suffix table instead: row r of table t+1 owns suffix rows [PW[r], PW[r+1]).
__global__ __launch_bounds__(kBlock) void records_to_suffix_kernel(RunRef* __restrict__ rec,
                                                                   uint64_t n,
                                                                   const uint64_t* __restrict__ pw) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    RunRef r = rec[i];
    const uint64_t a = pw[r.lo], b = pw[r.lo + r.len];
    r.lo = (uint32_t)a;
    r.len = (uint32_t)(b - a);
    rec[i] = r;
  }
}

hipError_t launch_records_to_suffix(RunRef* rec, uint64_t n, const uint64_t* pw, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(records_to_suffix_kernel, dim3(grid_rows(n)), dim3(kBlock), 0, s, rec, n,
                     pw);
  return hipGetLastError();
}

// Workgroups of one writer launch; each wave loops over windows when the windows outnumber
// the grid's waves.
constexpr uint64_t kWindowMaxBlocks = 65536;
uint64_t window_blocks(uint64_t total, uint64_t window) {
  const uint64_t n_windows = (total + window - 1) / window;
  const uint64_t b = n_windows == 0 ? 1 : (n_windows + kWavesPerBlock - 1) / kWavesPerBlock;
  return b < kWindowMaxBlocks ? b : kWindowMaxBlocks;
}

// The chain layout the writer is specialised for: probe attributes 0..PC-1, table t binds
// attribute PC+t, table t+1 is keyed on the attribute table t binds.
bool window_plan_supported(const Plan& P) {
  const int L = P.n_levels, PC = P.probe_ncols;
  if (L < 1 || L > kMaxLevels || PC < 1 || PC > 2) return false;
  for (int c = 0; c < PC; ++c)
    if (P.probe_attr[c] != c) return false;
  for (int t = 0; t < L; ++t) {
    const bool last = t == L - 1;
    if (P.val_mode[t] == VAL_BIND) {
      if (P.val_attr[t] != PC + t) return false;
    } else if (!(last && P.val_mode[t] == VAL_NONE)) {
      return false;
    }
    if (t > 0 && P.key_attr[t] != P.val_attr[t - 1]) return false;
  }
  const int na = PC + (L - 1) + (P.val_mode[L - 1] == VAL_BIND ? 1 : 0);
  return na == P.n_attrs && na <= kMaxAttrs;
}

// P: n_levels frontier levels (tables 0..L-2 with pw + next records, suffix for the last hop),
// probe records in P.probe_next.
hipError_t launch_window_write(const Plan& P, const uint64_t* probe_off, uint64_t* first_row,
                               uint64_t total, uint64_t window, const ProbeOut& po, bool hash,
                               uint64_t* block_sums, uint32_t* overflow, uint64_t* n_blocks_out,
                               hipStream_t s) {
  const int L = P.n_levels, PC = P.probe_ncols, NV = P.sfx_cols;
  if (L < 1 || L > kMaxLevels || PC < 1 || PC > 2 || NV < 0 || PC + (L - 1) + NV > kMaxAttrs)
    return hipErrorInvalidValue;
  const uint64_t n_windows = (total + window - 1) / window;
  const uint64_t blocks = window_blocks(total, window);
  *n_blocks_out = blocks;
  hipLaunchKernelGGL(window_first_row_kernel, dim3(grid_rows(n_windows)), dim3(kBlock), 0, s,
                     probe_off, P.probe_n, window, n_windows, first_row);
  WindowArgs a;
  a.probe_off = probe_off;
  a.first_row = first_row;
  a.total = total;
  a.window = window;
  a.n_windows = n_windows;
  a.out = ChainOut{po.cols[0], 0};
  if (po.cols[0] && P.n_attrs > 1) a.out.stride = (uint64_t)(po.cols[1] - po.cols[0]);
  a.block_sums = block_sums;
  a.overflow = overflow;
  const int mode = write_mode(po, hash);
  const dim3 grid((unsigned)blocks);
  switch (L) {
    case 1: launch_window_l1(PC, NV, mode, grid, s, P, a); break;
    case 2: launch_window_l2(PC, NV, mode, grid, s, P, a); break;
    case 3: launch_window_l3(PC, NV, mode, grid, s, P, a); break;
    case 4: launch_window_l4(PC, NV, mode, grid, s, P, a); break;
    case 5: launch_window_l5(PC, NV, mode, grid, s, P, a); break;
    case 6: launch_window_l6(PC, NV, mode, grid, s, P, a); break;
    case 7: launch_window_l7(PC, NV, mode, grid, s, P, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace amj
