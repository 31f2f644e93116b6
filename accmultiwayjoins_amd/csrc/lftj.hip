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
// every other participating list (binary search in a child range, or a dense/binary lookup in a
// level-0 list, which also opens that relation's child range). Failed candidates simply get no
// children, so no compaction is needed between levels. Any query over two-attribute relations
// with up to 4 attributes and 6 relations runs here (triangle, 4-clique, multi-way).
#include "lftj_dfs.h"

namespace amj {

// Unit weights for the launch order: the squared candidate count at attribute 1, summed over
// the unit's roots (what a root's second level costs when every candidate opens a list of its
// own), as the key (~bits(float(weight)) << 32) | unit (the bits of a non-negative float order
// like its value), so an ascending sort puts the heaviest units first and the dispatcher starts
// them before the light ones.
template <int A>
__global__ __launch_bounds__(kBlock) void lftj_unit_weight_kernel(const LPlan P,
                                                                  uint64_t* __restrict__ keys,
                                                                  uint64_t units) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k >= units) return;  // wave-uniform; no barriers in this kernel
  LLv<A> l0;
  level0(P, k, l0);
  const uint64_t w = wave_sum_u64((uint64_t)l0.dlen * (uint64_t)l0.dlen);
  if (lane_id() == 0) keys[k] = weight_key(w, k);
}

// Count pass: counts[unit] for every unit, taken in `order` (nullptr: number order).
template <int A>
__global__ __launch_bounds__(kBlock) void lftj_count_kernel(const LPlan P,
                                                            uint64_t* __restrict__ counts,
                                                            uint64_t units,
                                                            const uint64_t* __restrict__ order) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k >= units) return;  // no barriers in this kernel
  const uint64_t unit = unit_of(order, k);
  LLv<A> l0;
  level0(P, unit, l0);
  CountSink sink;
  if (l0.T) descend<A, 1>(P, l0, sink);
  if (lane_id() == 0) counts[unit] = sink.count;
}

// Output-size sample: wave k counts the results of the roots (k * per_wave + i) * stride +
// offset, i < per_wave. A few roots per wave instead of a unit of 64: the launch lasts as long
// as its slowest wave, and on a skewed graph a unit of 64 roots costs several times its
// heaviest root.
template <int A>
__global__ __launch_bounds__(kBlock) void lftj_sample_kernel(const LPlan P,
                                                             uint64_t* __restrict__ counts,
                                                             uint64_t waves, uint32_t per_wave,
                                                             uint64_t stride, uint64_t offset) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k >= waves) return;  // no barriers in this kernel
  const uint32_t lane = (uint32_t)lane_id();
  const uint64_t r = (k * per_wave + lane) * stride + offset;
  LLv<A> l0;
  level0_at(P, r, lane < per_wave && r < P.roots, l0);
  CountSink sink;
  if (l0.T) descend<A, 1>(P, l0, sink);
  if (lane == 0) counts[k] = sink.count;
}

template <int A, int MODE>
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
    level0(P, unit, l0);
    start = base ? uniform(base[unit]) : 0;
    sink.out_base = start;
    if (l0.T) descend<A, 1>(P, l0, sink);
  }
  // out_base is uniform: lane 0 carries the wave's count
  block_sum_store2(sink.csum, lane_id() == 0 ? sink.out_base - start : 0,
                   &block_sums[blockIdx.x], block_counts ? &block_counts[blockIdx.x] : nullptr);
}

template <int A, bool HASH>
__global__ __launch_bounds__(kBlock) void lftj_direct_kernel(const LPlan P, const LOut cols,
                                                             uint64_t capacity,
                                                             unsigned long long* __restrict__ rows,
                                                             uint64_t* __restrict__ block_sums,
                                                             uint64_t units,
                                                             const uint64_t* __restrict__ order) {
  __shared__ int32_t stage_mem[kWavesPerBlock * A * kLftjDirectRows];
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  DirectSink<HASH> sink{cols, capacity, rows,
                        (lds_i32*)&stage_mem[(threadIdx.x / kWave) * A * kLftjDirectRows], 0u,
                        A, 0};
  if (k < units) {
    LLv<A> l0;
    level0(P, unit_of(order, k), l0);
    if (l0.T) descend<A, 1>(P, l0, sink);
    sink.flush();
  }
  block_sum_store(sink.csum, &block_sums[blockIdx.x]);
}

uint64_t lftj_units(const LPlan& P) { return (P.roots + kWave - 1) / kWave; }

static hipError_t launch_count(const LPlan& P, uint64_t* counts, hipStream_t s, uint64_t units,
                               const uint64_t* order) {
  if (units == 0) return hipSuccess;
  const dim3 grid((unsigned)((units + kWavesPerBlock - 1) / kWavesPerBlock));
#define LC(AV)                                                                                \
  hipLaunchKernelGGL(lftj_count_kernel<AV>, grid, dim3(kBlock), 0, s, P, counts, units, order)
  switch (P.n_attrs) {
    case 2: LC(2); break;
    case 3: LC(3); break;
    case 4: LC(4); break;
    default: return hipErrorInvalidValue;
  }
#undef LC
  return hipGetLastError();
}

hipError_t launch_lftj_sample(const LPlan& P, uint64_t* counts, uint64_t waves,
                              uint32_t per_wave, uint64_t stride, uint64_t offset,
                              hipStream_t s) {
  if (waves == 0) return hipSuccess;
  const dim3 grid((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock));
#define LS(AV)                                                                                 \
  hipLaunchKernelGGL(lftj_sample_kernel<AV>, grid, dim3(kBlock), 0, s, P, counts, waves, per_wave, \
                     stride, offset)
  switch (P.n_attrs) {
    case 2: LS(2); break;
    case 3: LS(3); break;
    case 4: LS(4); break;
    default: return hipErrorInvalidValue;
  }
#undef LS
  return hipGetLastError();
}

hipError_t launch_lftj_count(const LPlan& P, uint64_t* counts, const uint64_t* order,
                             hipStream_t s) {
  return launch_count(P, counts, s, lftj_units(P), order);
}

hipError_t launch_lftj_unit_weights(const LPlan& P, uint64_t* keys, hipStream_t s) {
  const uint64_t units = lftj_units(P);
  if (units == 0) return hipSuccess;
  const dim3 grid((unsigned)((units + kWavesPerBlock - 1) / kWavesPerBlock));
  switch (P.n_attrs) {
    case 2: hipLaunchKernelGGL(lftj_unit_weight_kernel<2>, grid, dim3(kBlock), 0, s, P, keys, units); break;
    case 3: hipLaunchKernelGGL(lftj_unit_weight_kernel<3>, grid, dim3(kBlock), 0, s, P, keys, units); break;
    case 4: hipLaunchKernelGGL(lftj_unit_weight_kernel<4>, grid, dim3(kBlock), 0, s, P, keys, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lftj_direct(const LPlan& P, const LOut& cols, uint64_t capacity, bool hash,
                              unsigned long long* rows, uint64_t* block_sums,
                              uint64_t* n_blocks, const uint64_t* order, hipStream_t s) {
  const uint64_t units = lftj_units(P);
  const uint64_t blocks = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  *n_blocks = blocks;
  const dim3 grid((unsigned)blocks);
#define LD(AV, H)                                                                             \
  hipLaunchKernelGGL((lftj_direct_kernel<AV, H>), grid, dim3(kBlock), 0, s, P, cols, capacity, \
                     rows, block_sums, units, order)
  switch (P.n_attrs) {
    case 2: if (hash) LD(2, true); else LD(2, false); break;
    case 3: if (hash) LD(3, true); else LD(3, false); break;
    case 4: if (hash) LD(4, true); else LD(4, false); break;
    default: return hipErrorInvalidValue;
  }
#undef LD
  return hipGetLastError();
}

hipError_t launch_lftj_write(const LPlan& P, const uint64_t* base, const LOut& cols,
                             bool hash, uint64_t* block_sums, uint64_t* n_blocks,
                             const uint64_t* order, hipStream_t s, uint64_t* block_counts) {
  const uint64_t units = lftj_units(P);
  const uint64_t blocks = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  *n_blocks = blocks;
  const dim3 grid((unsigned)blocks);
  const int mode = cols.c[0] ? (hash ? MODE_WRITE_HASH : MODE_WRITE) : MODE_HASH;
#define LW(AV, M) \
  hipLaunchKernelGGL((lftj_write_kernel<AV, M>), grid, dim3(kBlock), 0, s, P, base, cols, block_sums, \
                     block_counts, units, order)
#define LWA(AV)                                              \
  if (mode == MODE_WRITE) LW(AV, MODE_WRITE);                \
  else if (mode == MODE_WRITE_HASH) LW(AV, MODE_WRITE_HASH); \
  else LW(AV, MODE_HASH)
  switch (P.n_attrs) {
    case 2: LWA(2); break;
    case 3: LWA(3); break;
    case 4: LWA(4); break;
    default: return hipErrorInvalidValue;
  }
#undef LWA
#undef LW
  return hipGetLastError();
}

}  // namespace amj
