// Many AMHJ queries in one launch (batched count pass, one scan, batched write pass).
//
// The reference runs one query per process and one kernel per pass (exec_AMHJ.cu); it has no
// batching. A small query (a 10 K-row probe is 157 units of 64 rows, 40 workgroups) leaves most
// of the 256 CUs idle and pays several host round trips. Here the units of every query of a
// batch are numbered consecutively; one grid covers all of them, each wave finds its query by a
// binary search over the queries' first units (wave-uniform, scalar loads of a read-only
// descriptor array) and runs the frontier DFS of probe_engine.h with that query's plan. One
// exclusive scan over all unit counts gives every unit its output row inside its own query
// (minus the query's first scanned value), and per-query checksums are summed with one atomic
// per wave (a sum mod 2^64 does not depend on the order).
#include "probe_engine.h"

namespace amj {

struct BatchQ {
  Plan P;
  ProbeOut out;     // capacity 0: count and checksum only
  uint64_t unit0;   // first unit of this query in the batch
  uint64_t units;
  uint64_t stride;  // > 1: a sample, wave k takes the rows (k * 64 + lane) * stride + stride / 2
};

__device__ __forceinline__ uint32_t query_of(const BatchQ* __restrict__ qs, uint32_t nq,
                                             uint64_t g) {
  uint32_t lo = 0, hi = nq;  // last q with unit0 <= g
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (qs[mid].unit0 <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int L>
__global__ __launch_bounds__(kBlock) void batch_count_kernel(const BatchQ* __restrict__ qs,
                                                             uint32_t nq, uint64_t unit_begin,
                                                             uint64_t total,
                                                             uint64_t* __restrict__ counts) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (w >= total) return;  // no barriers in this kernel
  const uint64_t g = unit_begin + w;
  const uint32_t q = uniform(query_of(qs, nq, g));
  const BatchQ& Q = qs[q];
  LvL<L> l0;
  if (Q.stride > 1) {
    // a sample: wave i of the query takes the rows (i * 64 + lane) * stride + stride / 2, so
    // one wave never holds a run of neighbouring rows (edges sorted by source share one hub)
    const uint64_t r = ((g - Q.unit0) * kWave + (uint64_t)lane_id()) * Q.stride + Q.stride / 2;
    make_level0_at(Q.P, r, r < Q.P.probe_n, l0);
  } else {
    make_level0(Q.P, (g - Q.unit0) * kWave, l0);
  }
  Acc acc{};
  if (l0.T) descend<L, 0, MODE_COUNT>(Q.P, l0, ProbeOut{}, acc);
  if (lane_id() == 0) counts[g] = acc.count;
}

template <int L>
__global__ __launch_bounds__(kBlock) void batch_write_kernel(const BatchQ* __restrict__ qs,
                                                             uint32_t nq, uint64_t unit_begin,
                                                             uint64_t total,
                                                             const uint64_t* __restrict__ base,
                                                             unsigned long long* __restrict__ csums) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (w >= total) return;  // no barriers in this kernel
  const uint64_t g = unit_begin + w;
  const uint32_t q = uniform(query_of(qs, nq, g));
  const BatchQ& Q = qs[q];
  LvL<L> l0;
  make_level0(Q.P, (g - Q.unit0) * kWave, l0);
  Acc acc{};
  acc.out_base = uniform(base[g] - base[Q.unit0]);
  if (l0.T) descend<L, 0, MODE_WRITE_HASH>(Q.P, l0, Q.out, acc);
  const uint64_t cs = wave_sum_u64(acc.csum);
  if (lane_id() == 0 && cs)
    __hip_atomic_fetch_add(&csums[q], (unsigned long long)cs, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Direct output (the count pass dropped): every wave gathers its results through the staging
// buffer of probe_engine.h (DirectStage), one reservation per full buffer on its query's counter.
template <int L, int t, bool MAT>
__device__ __forceinline__ void descend_direct(const Plan& P, const LvL<L>& cur, const ProbeOut& out,
                                               Acc& acc, DirectStage& st,
                                               unsigned long long* rows) {
  if constexpr (t == L - 1) {
    direct_emit<MAT, true>(P, t, cur, out, acc, st, rows);
    acc.count += cur.T;
  } else {
    for (uint64_t c = 0; c < cur.T; c += kWave) {
      LvL<L> ch;
      make_child<t>(P, cur, c, cur.T, ch);
      if (ch.T) descend_direct<L, t + 1, MAT>(P, ch, out, acc, st, rows);
    }
  }
}

// Weights of the launch's units for a heaviest-first order (squared first-table run lengths
// of each unit's probe rows), as sort keys over the launch's unit offsets w.
__global__ __launch_bounds__(kBlock) void batch_weight_kernel(const BatchQ* __restrict__ qs,
                                                              uint32_t nq, uint64_t unit_begin,
                                                              uint64_t total,
                                                              uint64_t* __restrict__ keys) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (w >= total) return;  // no barriers in this kernel
  const uint64_t g = unit_begin + w;
  const uint32_t q = uniform(query_of(qs, nq, g));
  const BatchQ& Q = qs[q];
  LvL<1> l0;  // level 0 binds the probe's attributes (ids 0 and 1) only
  make_level0(Q.P, (g - Q.unit0) * kWave, l0);
  const uint64_t wt = wave_sum_u64((uint64_t)l0.len * (uint64_t)l0.len);
  if (lane_id() == 0) keys[w] = weight_key(wt, w);
}

template <int L, bool MAT>
__global__ __launch_bounds__(kBlock) void batch_direct_kernel(
    const BatchQ* __restrict__ qs, uint32_t nq, uint64_t unit_begin, uint64_t total,
    unsigned long long* __restrict__ rows, unsigned long long* __restrict__ counts,
    unsigned long long* __restrict__ csums, const uint64_t* __restrict__ order) {
  const uint64_t k = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k >= total) return;  // no barriers in this kernel
  const uint64_t g = unit_begin + unit_of(order, k);
  const uint32_t q = uniform(query_of(qs, nq, g));
  const BatchQ& Q = qs[q];
  __shared__ int32_t stage_mem[MAT ? direct_stage_words<L>() : 1];
  DirectStage st{(lds_i32*)&stage_mem[MAT ? (threadIdx.x / kWave) * na_of(L) * kDirectRows : 0],
                 0u};
  LvL<L> l0;
  make_level0(Q.P, (g - Q.unit0) * kWave, l0);
  Acc acc{};
  if (l0.T) {  // MAT: every query of this launch materialises (one instantiation per kind)
    descend_direct<L, 0, MAT>(Q.P, l0, Q.out, acc, st, &rows[q]);
    if (MAT) direct_flush<na_of(L)>(Q.out, Q.P.n_attrs, st, &rows[q]);
  }
  const uint64_t cs = wave_sum_u64(acc.csum);
  if (lane_id() == 0) {
    if (acc.count)
      __hip_atomic_fetch_add(&counts[q], (unsigned long long)acc.count, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    if (cs)
      __hip_atomic_fetch_add(&csums[q], (unsigned long long)cs, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Per-query totals: counts[q] = base[unit0 + units] - base[unit0].
__global__ void batch_totals_kernel(const BatchQ* __restrict__ qs, uint32_t nq,
                                    const uint64_t* __restrict__ base,
                                    uint64_t* __restrict__ totals) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nq) totals[q] = base[qs[q].unit0 + qs[q].units] - base[qs[q].unit0];
}

static unsigned grid_units(uint64_t total) {
  return (unsigned)((total + kWavesPerBlock - 1) / kWavesPerBlock);
}

hipError_t launch_batch_count(int L, const void* qs, uint32_t nq, uint64_t unit_begin,
                              uint64_t total, uint64_t* counts, hipStream_t s) {
  if (total == 0) return hipSuccess;
  const BatchQ* q = (const BatchQ*)qs;
  switch (L) {
#define C(LV)                                                                              \
  case LV:                                                                                 \
    hipLaunchKernelGGL(batch_count_kernel<LV>, dim3(grid_units(total)), dim3(kBlock), 0, s, \
                       q, nq, unit_begin, total, counts);                                  \
    break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_write(int L, const void* qs, uint32_t nq, uint64_t unit_begin,
                              uint64_t total, const uint64_t* base, unsigned long long* csums,
                              hipStream_t s) {
  if (total == 0) return hipSuccess;
  const BatchQ* q = (const BatchQ*)qs;
  switch (L) {
#define C(LV)                                                                              \
  case LV:                                                                                 \
    hipLaunchKernelGGL(batch_write_kernel<LV>, dim3(grid_units(total)), dim3(kBlock), 0, s, \
                       q, nq, unit_begin, total, base, csums);                             \
    break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_weights(const void* qs, uint32_t nq, uint64_t unit_begin,
                                uint64_t total, uint64_t* keys, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_weight_kernel, dim3(grid_units(total)), dim3(kBlock), 0, s,
                     (const BatchQ*)qs, nq, unit_begin, total, keys);
  return hipGetLastError();
}

hipError_t launch_batch_direct(int L, bool mat, const void* qs, uint32_t nq,
                               uint64_t unit_begin, uint64_t total, unsigned long long* rows,
                               unsigned long long* counts, unsigned long long* csums,
                               const uint64_t* order, hipStream_t s) {
  if (total == 0) return hipSuccess;
  const BatchQ* q = (const BatchQ*)qs;
  switch (L) {
#define C(LV)                                                                                  \
  case LV:                                                                                     \
    if (mat)                                                                                   \
      hipLaunchKernelGGL((batch_direct_kernel<LV, true>), dim3(grid_units(total)), dim3(kBlock), \
                         0, s, q, nq, unit_begin, total, rows, counts, csums, order);          \
    else                                                                                       \
      hipLaunchKernelGGL((batch_direct_kernel<LV, false>), dim3(grid_units(total)),           \
                         dim3(kBlock), 0, s, q, nq, unit_begin, total, rows, counts, csums,    \
                         order);                                                               \
    break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_totals(const void* qs, uint32_t nq, const uint64_t* base,
                               uint64_t* totals, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_totals_kernel, dim3((nq + 255) / 256), dim3(256), 0, s,
                     (const BatchQ*)qs, nq, base, totals);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Captured batches: output sizes and places decided on the device, so the whole chain (size
// sample -> sizing and placement -> direct output) runs without a host round trip and can be
// recorded into a HIP graph.
// ---------------------------------------------------------------------------------------
struct PrepInfo {
  uint64_t units;     // units of the query
  uint64_t sunits;    // units sampled (0: no sample: count-only, or a size given)
  uint64_t user_cap;  // output rows given by the caller (0: from the sample)
  uint32_t n_attrs;
  uint32_t mat;       // materialised
};

// Per query k (in launch order): its output size (exact when every unit was sampled, else the
// sample scaled up by 1.25 plus 64 Ki rows), its place in the slab (column-major, 256-byte
// aligned, in launch order), and whether it fits; the direct-output descriptor gets its
// columns and capacity (none when it does not fit: it is counted, not written). One thread:
// a batch holds at most a few thousand queries.
__global__ void batch_place_kernel(const PrepInfo* __restrict__ info, uint32_t nq,
                                   const uint64_t* __restrict__ sampled, BatchQ* __restrict__ dq,
                                   int32_t* __restrict__ slab, uint64_t slab_words,
                                   uint64_t* __restrict__ cap_out, uint64_t* __restrict__ off_out,
                                   uint32_t* __restrict__ admitted) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t off = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    const PrepInfo& I = info[k];
    uint64_t cap = 0;
    if (I.mat) {
      if (I.user_cap) cap = I.user_cap;
      else if (I.sunits == I.units) cap = sampled[k];
      else cap = (uint64_t)((double)sampled[k] * (double)I.units / (double)I.sunits * 1.25) + 65536;
    }
    const uint64_t words = (cap * I.n_attrs + 63) / 64 * 64;
    const bool fits = !I.mat || off + words <= slab_words;
    for (int a = 0; a < kMaxAttrs; ++a)
      dq[k].out.cols[a] = (fits && cap && a < (int)I.n_attrs) ? slab + off + (uint64_t)a * cap
                                                              : nullptr;
    dq[k].out.capacity = fits ? cap : 0;
    cap_out[k] = cap;
    off_out[k] = fits ? off : 0;
    admitted[k] = fits ? 1u : 0u;
    if (fits) off += words;
  }
}

// Clears counters inside a captured chain (a kernel, not a memset node).
__global__ void fill_u64_kernel(uint64_t* __restrict__ p, uint64_t n, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t launch_batch_place(const void* info, uint32_t nq, const uint64_t* sampled, void* dq,
                              int32_t* slab, uint64_t slab_words, uint64_t* cap_out,
                              uint64_t* off_out, uint32_t* admitted, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_place_kernel, dim3(1), dim3(64), 0, s, (const PrepInfo*)info, nq,
                     sampled, (BatchQ*)dq, slab, slab_words, cap_out, off_out, admitted);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

size_t batch_prep_info_bytes() { return sizeof(PrepInfo); }

void batch_fill_prep_info(void* dst, uint64_t units, uint64_t sunits, uint64_t user_cap,
                          uint32_t n_attrs, bool mat) {
  PrepInfo* I = (PrepInfo*)dst;
  I->units = units;
  I->sunits = sunits;
  I->user_cap = user_cap;
  I->n_attrs = n_attrs;
  I->mat = mat ? 1u : 0u;
}

size_t batch_desc_bytes() { return sizeof(BatchQ); }

void batch_fill_desc(void* dst, const Plan& P, const ProbeOut& out, uint64_t unit0,
                     uint64_t units, uint64_t stride) {
  BatchQ* q = (BatchQ*)dst;
  q->P = P;
  q->out = out;
  q->unit0 = unit0;
  q->units = units;
  q->stride = stride;
}

}  // namespace amj
