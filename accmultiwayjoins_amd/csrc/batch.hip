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
  uint64_t stride;  // unit k of the query covers probe rows from k * stride * 64 (sampling)
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
  Lv l0;
  make_level0(Q.P, (g - Q.unit0) * Q.stride * kWave, l0);
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
  Lv l0;
  make_level0(Q.P, (g - Q.unit0) * kWave, l0);
  Acc acc{};
  acc.out_base = uniform(base[g] - base[Q.unit0]);
  if (l0.T) descend<L, 0, MODE_WRITE_HASH>(Q.P, l0, Q.out, acc);
  const uint64_t cs = wave_sum_u64(acc.csum);
  if (lane_id() == 0 && cs)
    __hip_atomic_fetch_add(&csums[q], (unsigned long long)cs, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Direct output (the count pass dropped). A wave gathers its results in an LDS buffer of
// kDirectRows rows and reserves room in its query's output with one atomic per full buffer, so
// the output stays dense and the query's counter sees one update per kDirectRows results
// (tiny last-level frontiers, as in FOUR, would otherwise make one update per few results).
// A query's output was sized from a sample: rows past its capacity are not written, and the
// reserved total tells the host that the query must run again.
constexpr uint32_t kDirectRows = 256;
typedef __attribute__((address_space(3))) int32_t lds_i32;

struct DirectStage {
  lds_i32* buf;   // this wave's rows: attribute a of row i at buf[a * kDirectRows + i]
  uint32_t fill;  // rows held (uniform)
};

__device__ __forceinline__ void direct_flush(const ProbeOut& out, int na, DirectStage& st,
                                             unsigned long long* rows) {
  if (st.fill == 0) return;
  unsigned long long r0 = 0;
  if (lane_id() == 0)
    r0 = __hip_atomic_fetch_add(rows, (unsigned long long)st.fill, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t base = uniform((uint64_t)r0);
  __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
  for (uint32_t i = (uint32_t)lane_id(); i < st.fill; i += kWave) {
    const uint64_t pos = base + i;
    if (pos < out.capacity)
      sfor<0, kMaxAttrs>([&](auto a) {
        if (a < na) out.cols[a][pos] = st.buf[a * kDirectRows + i];
      });
  }
  __asm__ __volatile__("" ::: "memory");
  st.fill = 0;
}

// Results [0, T) of a last level into the wave's buffer (and the checksum).
template <bool MAT>
__device__ __forceinline__ void direct_emit(const Plan& P, int t, const Lv& l,
                                            const ProbeOut& out, Acc& acc, DirectStage& st,
                                            unsigned long long* rows) {
  const bool bind = P.val_mode[t] == VAL_BIND;
  const int vattr = P.val_attr[t];
  const int32_t* vals = P.tab[t].vals;
  const uint64_t e1 = l.T;
  for (uint64_t c = 0; c < e1; c += kWave) {
    const uint64_t g = c + (uint64_t)lane_id();
    const bool valid = g < e1;
    const uint64_t gg = valid ? g : e1 - 1;
    const int p = parent_lane(l, c, gg);
    const uint32_t plo = __shfl(l.lo, p, kWave);
    const uint32_t pex = (uint32_t)__shfl((uint32_t)l.excl, p, kWave);
    const uint32_t row = plo + ((uint32_t)gg - pex);
    int32_t v[kMaxAttrs];
    sfor<0, kMaxAttrs>([&](auto a) { v[a] = (a < P.n_attrs) ? __shfl(l.b[a], p, kWave) : 0; });
    if (bind) set_attr(v, vattr, vals[row]);
    if (valid) acc.csum += hash_tuple(v, P.n_attrs);
    if (MAT) {
      const uint32_t m = e1 - c < (uint64_t)kWave ? (uint32_t)(e1 - c) : (uint32_t)kWave;
      if (st.fill + m > kDirectRows) direct_flush(out, P.n_attrs, st, rows);
      __asm__ __volatile__("" ::: "memory");
      if (valid)
        sfor<0, kMaxAttrs>([&](auto a) {
          if (a < P.n_attrs) st.buf[a * kDirectRows + st.fill + lane_id()] = v[a];
        });
      __asm__ __volatile__("" ::: "memory");
      st.fill += m;
    }
  }
}

template <int L, int t, bool MAT>
__device__ __forceinline__ void descend_direct(const Plan& P, const Lv& cur, const ProbeOut& out,
                                               Acc& acc, DirectStage& st,
                                               unsigned long long* rows) {
  if constexpr (t == L - 1) {
    direct_emit<MAT>(P, t, cur, out, acc, st, rows);
    acc.count += cur.T;
  } else {
    for (uint64_t c = 0; c < cur.T; c += kWave) {
      Lv ch;
      make_child<t>(P, cur, c, cur.T, ch);
      if (ch.T) descend_direct<L, t + 1, MAT>(P, ch, out, acc, st, rows);
    }
  }
}

template <int L, bool MAT>
__global__ __launch_bounds__(kBlock) void batch_direct_kernel(
    const BatchQ* __restrict__ qs, uint32_t nq, uint64_t unit_begin, uint64_t total,
    unsigned long long* __restrict__ rows, unsigned long long* __restrict__ counts,
    unsigned long long* __restrict__ csums) {
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (w >= total) return;  // no barriers in this kernel
  const uint64_t g = unit_begin + w;
  const uint32_t q = uniform(query_of(qs, nq, g));
  const BatchQ& Q = qs[q];
  __shared__ int32_t stage_mem[MAT ? kWavesPerBlock * kMaxAttrs * kDirectRows : 1];
  DirectStage st{(lds_i32*)&stage_mem[MAT ? (threadIdx.x / kWave) * kMaxAttrs * kDirectRows : 0],
                 0u};
  Lv l0;
  make_level0(Q.P, (g - Q.unit0) * kWave, l0);
  Acc acc{};
  if (l0.T) {  // MAT: every query of this launch materialises (one instantiation per kind)
    descend_direct<L, 0, MAT>(Q.P, l0, Q.out, acc, st, &rows[q]);
    if (MAT) direct_flush(Q.out, Q.P.n_attrs, st, &rows[q]);
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
    C(1) C(2) C(3) C(4) C(5)
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
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_direct(int L, bool mat, const void* qs, uint32_t nq,
                               uint64_t unit_begin, uint64_t total, unsigned long long* rows,
                               unsigned long long* counts, unsigned long long* csums,
                               hipStream_t s) {
  if (total == 0) return hipSuccess;
  const BatchQ* q = (const BatchQ*)qs;
  switch (L) {
#define C(LV)                                                                                  \
  case LV:                                                                                     \
    if (mat)                                                                                   \
      hipLaunchKernelGGL((batch_direct_kernel<LV, true>), dim3(grid_units(total)), dim3(kBlock), \
                         0, s, q, nq, unit_begin, total, rows, counts, csums);                 \
    else                                                                                       \
      hipLaunchKernelGGL((batch_direct_kernel<LV, false>), dim3(grid_units(total)),           \
                         dim3(kBlock), 0, s, q, nq, unit_begin, total, rows, counts, csums);   \
    break;
    C(1) C(2) C(3) C(4) C(5)
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
