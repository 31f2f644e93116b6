// ALFTJ: leapfrog triejoin over sorted CSR tries, depth-first, wave64 frontier.
//
// Reference: ALFTJ_count / ALFTJ_write (LFTJ-GPU/ALFTJ.cuh:213-1043): a 32-lane warp claims 32
// level-0 values of the smallest relation (:280) and walks the attributes in order; at every
// attribute warp_intersection_binary (:114-167) binary-searches each element of the shortest
// list in the other lists, compacting matches into shared stacks (two relations per level, so
// triangles only; see the aliasing note at :118). Results are counted, then written in a second
// pass at group-aggregated positions.
//
// Here the attribute loop is the frontier DFS of the AMHJ probe (probe.hip): a wave holds 64
// partial bindings per attribute level; every lane keeps, for each relation whose first
// attribute is bound, the range of its children; the candidates for the next attribute are the
// entries of each lane's *shortest* participating list, flattened over the wave and enumerated
// 64 at a time, so lanes stay busy whatever the degree skew. Each candidate is checked against
// every other participating list (binary search in a child range, or a table/binary lookup in a
// level-0 list, which also opens that relation's child range). Failed candidates simply get no
// children, so no compaction is needed between levels. Any query over two-attribute relations
// with up to 4 attributes and 6 relations runs here (triangle, 4-clique, multi-way).
#include <type_traits>

#include "lftj_dfs.h"

namespace amj {

// The wave's LDS stage of the cooperative last level (kernels of a compiled shape only).
template <uint64_t SH>
__device__ __forceinline__ lds_i32* coop_stage() {
  if constexpr (SH != 0) {
    __shared__ int32_t coop_mem[kWavesPerBlock * stage_ints<SH>()];
    return (lds_i32*)&coop_mem[(threadIdx.x / kWave) * stage_ints<SH>()];
  } else {
    return nullptr;
  }
}

uint64_t lftj_static_shape(const LPlan& P) {
  if (!P.off32) return 0;
  uint64_t c = 0;
  for (int r = 0; r < P.n_rels; ++r) {
    c |= 1ull << (12 * P.x[r] + r);
    c |= 1ull << (12 * P.y[r] + 6 + r);
  }
  if (P.n_attrs == 3 && c == kShapeTri) return c;
  if (P.n_attrs == 4 && c == kShapeFour) return c;
  return 0;
}

// Unit weights for the launch order: the squared candidate count at attribute 1, summed over
// the unit's roots (what a root's second level costs when every candidate opens a list of its
// own), as the key (~bits(float(weight)) << 32) | unit (the bits of a non-negative float order
// like its value), so an ascending sort puts the heaviest units first and the dispatcher starts
// them before the light ones.
template <int A, uint64_t SH>
__global__ __launch_bounds__(kBlock) void lftj_unit_weight_kernel(const LPlan P,
                                                                  uint64_t* __restrict__ keys,
                                                                  uint64_t units) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k >= units) return;  // wave-uniform; no barriers in this kernel
  LLv<A> l0;
  level0<A, SH>(P, k, l0);
  const uint64_t w = wave_sum_u64((uint64_t)l0.dlen * (uint64_t)l0.dlen);
  if (lane_id() == 0) keys[k] = weight_key(w, k);
}

// Count pass: counts[unit] for every unit, taken in `order` (nullptr: number order).
template <int A, uint64_t SH>
__global__ __launch_bounds__(kBlock) void lftj_count_kernel(const LPlan P,
                                                            uint64_t* __restrict__ counts,
                                                            uint64_t units,
                                                            const uint64_t* __restrict__ order) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k >= units) return;  // no barriers in this kernel
  const uint64_t unit = unit_of(order, k);
  LLv<A> l0;
  level0<A, SH>(P, unit, l0);
  CountSink sink;
  if (l0.T) descend<A, 1, SH>(P, l0, sink, coop_stage<SH>());
  if (lane_id() == 0) counts[unit] = sink.count;
}

// Output-size sample over the roots k * stride + offset, k < n, in two launches. A launch lasts
// as long as its slowest wave, and a wave that counts a heavy root alone costs that root's
// whole subtree (6.3 ms for the R-MAT 22 triangle sample of 65,536 roots, 4 per wave), so each
// sampled root's candidates for attribute 1 are cut into windows of kLftjSampleWindow, counted
// by one wave each (the sum is the same: every sampled root is counted whole). First: windows[k] = the windows of sampled root k (and windows[n] = 0, so
// an exclusive scan leaves the total there).
template <int A, uint64_t SH>
__global__ __launch_bounds__(kBlock) void lftj_sample_width_kernel(const LPlan P,
                                                                   uint64_t* __restrict__ windows,
                                                                   uint64_t n, uint64_t stride,
                                                                   uint64_t offset) {
  const uint64_t k0 = ((uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave) * kWave;
  if (k0 > n) return;  // wave-uniform; no barriers in this kernel
  const uint64_t k = k0 + (uint64_t)lane_id();
  const uint64_t r = k * stride + offset;
  LLv<A> l0;
  level0_at<A, SH>(P, r, k < n && r < P.roots, l0);
  if (k < n) windows[k] = (l0.dlen + kLftjSampleWindow - 1) / kLftjSampleWindow;
  else if (k == n) windows[n] = 0;
}

// Then wave w counts window j of sampled root k, where woff (the exclusive scan of windows,
// n + 1 entries) has woff[k] <= w < woff[k + 1] and j = w - woff[k].
template <int A, uint64_t SH>
__global__ __launch_bounds__(kBlock) void lftj_sample_window_kernel(
    const LPlan P, const uint64_t* __restrict__ woff, uint64_t n, uint64_t nw, uint64_t stride,
    uint64_t offset, uint64_t* __restrict__ counts) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (w >= nw) return;  // no barriers in this kernel
  // the last k < n with woff[k] <= w (woff[0] = 0 <= w); it has a window past w - woff[k]
  uint64_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (woff[mid] <= w) lo = mid;
    else hi = mid;
  }
  const uint64_t j = w - woff[lo];
  const uint64_t r = lo * stride + offset;
  LLv<A> l0;
  level0_at<A, SH>(P, r, lane_id() == 0 && r < P.roots, l0);
  const uint64_t c0 = j * kLftjSampleWindow, c1 = c0 + kLftjSampleWindow;
  restrict_range(l0, (uint32_t)c0, c1 < 0xffffffffull ? (uint32_t)c1 : 0xffffffffu);
  CountSink sink;
  if (l0.T) descend<A, 1, SH>(P, l0, sink, coop_stage<SH>());
  if (lane_id() == 0) counts[w] = sink.count;
}

template <int A, uint64_t SH, int MODE>
__global__ __launch_bounds__(kBlock) void lftj_write_kernel(const LPlan P,
                                                            const uint64_t* __restrict__ base,
                                                            const LOut cols,
                                                            uint64_t* __restrict__ block_sums,
                                                            uint64_t* __restrict__ block_counts,
                                                            uint64_t units,
                                                            const uint64_t* __restrict__ order) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  WriteSink<MODE> sink{cols, 0, 0};
  uint64_t start = 0;
  if (k < units) {
    const uint64_t unit = unit_of(order, k);
    LLv<A> l0;
    level0<A, SH>(P, unit, l0);
    start = base ? uniform(base[unit]) : 0;
    sink.out_base = start;
    if (l0.T) descend<A, 1, SH>(P, l0, sink, coop_stage<SH>());
  }
  // out_base is uniform: lane 0 carries the wave's count
  block_sum_store2(sink.csum, lane_id() == 0 ? sink.out_base - start : 0,
                   &block_sums[blockIdx.x], block_counts ? &block_counts[blockIdx.x] : nullptr);
}

// Direct output, rows reserved `chunk` at a time (see DirectSink); every wave records the unused
// tail of its last chunk in tails[] and per-block result counts go to block_counts.
template <int A, uint64_t SH, bool HASH>
__global__ __launch_bounds__(kBlock) void lftj_direct_kernel(const LPlan P, const LOut cols,
                                                             uint64_t capacity, uint32_t chunk,
                                                             unsigned long long* __restrict__ rows,
                                                             uint64_t* __restrict__ tails,
                                                             uint64_t* __restrict__ block_sums,
                                                             uint64_t* __restrict__ block_counts,
                                                             uint64_t units,
                                                             const uint64_t* __restrict__ order) {
  __shared__ int32_t stage_mem[kWavesPerBlock * A * kLftjDirectRows];
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  DirectSink<A, HASH> sink{cols, capacity, rows,
                           (lds_i32*)&stage_mem[(threadIdx.x / kWave) * A * kLftjDirectRows], 0u,
                           0};
  sink.chunk = chunk;
  if (k < units) {
    LLv<A> l0;
    level0<A, SH>(P, unit_of(order, k), l0);
    if (l0.T) descend<A, 1, SH>(P, l0, sink, coop_stage<SH>());
    sink.flush();
    sink.tail_of(tails);
  }
  block_sum_store2(sink.csum, lane_id() == 0 ? sink.count : 0, &block_sums[blockIdx.x],
                   &block_counts[blockIdx.x]);
}

uint64_t lftj_units(const LPlan& P) {
  const uint64_t rpu = P.rpu ? P.rpu : (uint64_t)kWave;
  return (P.roots + rpu - 1) / rpu;
}

template <int V>
using IntC = std::integral_constant<int, V>;
template <uint64_t V>
using ShapeC = std::integral_constant<uint64_t, V>;

// f(IntC<A>, ShapeC<SH>) for the plan's attribute count and compiled shape (SH = 0: the masks
// are read at run time).
template <typename F>
static hipError_t dispatch(const LPlan& P, F&& f) {
  const uint64_t sh = lftj_static_shape(P);
  switch (P.n_attrs) {
    case 2: f(IntC<2>{}, ShapeC<0>{}); break;
    case 3:
      if (sh == kShapeTri) f(IntC<3>{}, ShapeC<kShapeTri>{});
      else f(IntC<3>{}, ShapeC<0>{});
      break;
    case 4:
      if (sh == kShapeFour) f(IntC<4>{}, ShapeC<kShapeFour>{});
      else f(IntC<4>{}, ShapeC<0>{});
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static dim3 grid_of(uint64_t waves) {
  return dim3((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock));
}

hipError_t launch_lftj_sample_widths(const LPlan& P, uint64_t* windows, uint64_t n,
                                     uint64_t stride, uint64_t offset, hipStream_t s) {
  const uint64_t waves = n / kWave + 1;  // n + 1 entries, 64 per wave
  return dispatch(P, [&](auto a, auto sh) {
    hipLaunchKernelGGL((lftj_sample_width_kernel<decltype(a)::value, decltype(sh)::value>),
                       grid_of(waves), dim3(kBlock), 0, s, P, windows, n, stride, offset);
  });
}

hipError_t launch_lftj_sample_windows(const LPlan& P, const uint64_t* woff, uint64_t n,
                                      uint64_t nw, uint64_t stride, uint64_t offset,
                                      uint64_t* counts, hipStream_t s) {
  if (nw == 0) return hipSuccess;
  return dispatch(P, [&](auto a, auto sh) {
    hipLaunchKernelGGL((lftj_sample_window_kernel<decltype(a)::value, decltype(sh)::value>),
                       grid_of(nw), dim3(kBlock), 0, s, P, woff, n, nw, stride, offset, counts);
  });
}

hipError_t launch_lftj_count(const LPlan& P, uint64_t* counts, const uint64_t* order,
                             hipStream_t s) {
  const uint64_t units = lftj_units(P);
  if (units == 0) return hipSuccess;
  return dispatch(P, [&](auto a, auto sh) {
    hipLaunchKernelGGL((lftj_count_kernel<decltype(a)::value, decltype(sh)::value>),
                       grid_of(units), dim3(kBlock), 0, s, P, counts, units, order);
  });
}

hipError_t launch_lftj_unit_weights(const LPlan& P, uint64_t* keys, hipStream_t s) {
  const uint64_t units = lftj_units(P);
  if (units == 0) return hipSuccess;
  return dispatch(P, [&](auto a, auto sh) {
    hipLaunchKernelGGL((lftj_unit_weight_kernel<decltype(a)::value, decltype(sh)::value>),
                       grid_of(units), dim3(kBlock), 0, s, P, keys, units);
  });
}

hipError_t launch_lftj_direct(const LPlan& P, const LOut& cols, uint64_t capacity,
                              uint32_t chunk, bool hash, unsigned long long* rows,
                              uint64_t* tails, uint64_t* block_sums, uint64_t* block_counts,
                              uint64_t* n_blocks, const uint64_t* order, hipStream_t s) {
  const uint64_t units = lftj_units(P);
  const uint64_t blocks = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  *n_blocks = blocks;
  if (chunk == 0 || chunk % kLftjDirectRows != 0) return hipErrorInvalidValue;
  return dispatch(P, [&](auto a, auto sh) {
    constexpr int AV = decltype(a)::value;
    constexpr uint64_t SV = decltype(sh)::value;
    if (hash)
      hipLaunchKernelGGL((lftj_direct_kernel<AV, SV, true>), dim3((unsigned)blocks), dim3(kBlock),
                         0, s, P, cols, capacity, chunk, rows, tails, block_sums, block_counts,
                         units, order);
    else
      hipLaunchKernelGGL((lftj_direct_kernel<AV, SV, false>), dim3((unsigned)blocks),
                         dim3(kBlock), 0, s, P, cols, capacity, chunk, rows, tails, block_sums,
                         block_counts, units, order);
  });
}

hipError_t launch_lftj_write(const LPlan& P, const uint64_t* base, const LOut& cols,
                             bool hash, uint64_t* block_sums, uint64_t* n_blocks,
                             const uint64_t* order, hipStream_t s, uint64_t* block_counts) {
  const uint64_t units = lftj_units(P);
  const uint64_t blocks = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  *n_blocks = blocks;
  const int mode = cols.c[0] ? (hash ? MODE_WRITE_HASH : MODE_WRITE) : MODE_HASH;
  return dispatch(P, [&](auto a, auto sh) {
    constexpr int AV = decltype(a)::value;
    constexpr uint64_t SV = decltype(sh)::value;
    const dim3 grid((unsigned)blocks);
    if (mode == MODE_WRITE)
      hipLaunchKernelGGL((lftj_write_kernel<AV, SV, MODE_WRITE>), grid, dim3(kBlock), 0, s, P,
                         base, cols, block_sums, block_counts, units, order);
    else if (mode == MODE_WRITE_HASH)
      hipLaunchKernelGGL((lftj_write_kernel<AV, SV, MODE_WRITE_HASH>), grid, dim3(kBlock), 0, s,
                         P, base, cols, block_sums, block_counts, units, order);
    else
      hipLaunchKernelGGL((lftj_write_kernel<AV, SV, MODE_HASH>), grid, dim3(kBlock), 0, s, P,
                         base, cols, block_sums, block_counts, units, order);
  });
}

}  // namespace amj
