// Thread-per-tuple multi-way hash join: the baseline the wave engines are measured against
// (MHJ_count / MHJ_write, MHJ-GPU/Joins/MHJ.cuh:17-321). Every lane takes probe rows
// grid-stride and walks the nested runs of the build tables on its own, depth first, with no
// cooperation inside the wave: lanes of one wave diverge as their runs differ in length, a
// heavy key serialises on one lane, and the result columns are written one scattered 4-byte
// store per lane. It runs the same plans and produces the same tuples (in another row order)
// and the same checksum as the wave engines; it exists as the ablation of their design.
#include "probe_common.h"

namespace amj {

struct TptSink {
  uint64_t count;    // results seen by this lane
  uint64_t pos;      // next output row (write pass)
  uint64_t csum;     // per-lane checksum partial
};

template <bool WRITE, bool HASH>
__device__ __forceinline__ void tpt_emit(const Plan& P, const int32_t (&v)[kMaxAttrs],
                                         const ProbeOut& out, TptSink& k) {
  if (HASH) k.csum += hash_tuple(v, P.n_attrs);
  if (WRITE) {
    sfor<0, kMaxAttrs>([&](auto a) {
      if ((int)a < P.n_attrs) out.cols[a][k.pos] = v[a];
    });
    ++k.pos;
  }
  ++k.count;
}

// Table t: the run of the bound key (and of the compared value), each row binding its value.
template <int t, int L, bool WRITE, bool HASH>
__device__ __forceinline__ void tpt_descend(const Plan& P, int32_t (&v)[kMaxAttrs],
                                            const ProbeOut& out, TptSink& k) {
  const DevTable& tb = P.tab[t];
  const int mode = P.val_mode[t];
  const int va = P.val_attr[t];
  uint32_t lo, len;
  lookup_run(tb, get_attr(v, P.key_attr[t]), mode == VAL_CHECK,
             mode == VAL_CHECK ? get_attr(v, va) : 0, true, lo, len);
  for (uint32_t r = lo; r < lo + len; ++r) {
    if (mode == VAL_BIND) set_attr(v, va, tb.vals[r]);
    if constexpr (t == L - 1) tpt_emit<WRITE, HASH>(P, v, out, k);
    else tpt_descend<t + 1, L, WRITE, HASH>(P, v, out, k);
  }
}

// One lane per probe row, rows gtid, gtid + lanes, ... . Count pass: counts[gtid] = results of
// the lane's rows. Write pass: the lane writes its results from row base[gtid] on.
template <int L, bool WRITE, bool HASH>
__global__ __launch_bounds__(kBlock) void tpt_kernel(const Plan P, uint64_t* __restrict__ counts,
                                                     const uint64_t* __restrict__ base,
                                                     const ProbeOut out,
                                                     uint64_t* __restrict__ block_sums) {
  const uint64_t gtid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t lanes = (uint64_t)gridDim.x * kBlock;
  TptSink k{0, WRITE ? base[gtid] : 0, 0};
  for (uint64_t r = gtid; r < P.probe_n; r += lanes) {
    int32_t v[kMaxAttrs];
    sfor<0, kMaxAttrs>([&](auto a) { v[a] = 0; });
    for (int c = 0; c < P.probe_ncols; ++c) set_attr(v, P.probe_attr[c], P.probe_cols[c][r]);
    tpt_descend<0, L, WRITE, HASH>(P, v, out, k);
  }
  if (!WRITE) counts[gtid] = k.count;
  if (HASH) block_sum_store(k.csum, &block_sums[blockIdx.x]);
}

uint64_t tpt_blocks(const Plan& P) {
  const uint64_t need = (P.probe_n + kBlock - 1) / kBlock;
  const uint64_t cap = 2048;  // 8 workgroups of 256 lanes per CU
  return need == 0 ? 1 : (need < cap ? need : cap);
}

template <int L>
static void launch_tpt_l(bool write, bool hash, uint64_t blocks, hipStream_t s, const Plan& P,
                         uint64_t* counts, const uint64_t* base, const ProbeOut& out,
                         uint64_t* block_sums) {
  const dim3 g((unsigned)blocks), b(kBlock);
  if (write && hash)
    hipLaunchKernelGGL((tpt_kernel<L, true, true>), g, b, 0, s, P, counts, base, out, block_sums);
  else if (write)
    hipLaunchKernelGGL((tpt_kernel<L, true, false>), g, b, 0, s, P, counts, base, out,
                       block_sums);
  else if (hash)
    hipLaunchKernelGGL((tpt_kernel<L, false, true>), g, b, 0, s, P, counts, base, out,
                       block_sums);
  else
    hipLaunchKernelGGL((tpt_kernel<L, false, false>), g, b, 0, s, P, counts, base, out,
                       block_sums);
}

hipError_t launch_tpt(const Plan& P, bool write, bool hash, uint64_t* counts,
                      const uint64_t* base, const ProbeOut& out, uint64_t* block_sums,
                      hipStream_t s) {
  const uint64_t blocks = tpt_blocks(P);
  switch (P.n_levels) {
    case 1: launch_tpt_l<1>(write, hash, blocks, s, P, counts, base, out, block_sums); break;
    case 2: launch_tpt_l<2>(write, hash, blocks, s, P, counts, base, out, block_sums); break;
    case 3: launch_tpt_l<3>(write, hash, blocks, s, P, counts, base, out, block_sums); break;
    case 4: launch_tpt_l<4>(write, hash, blocks, s, P, counts, base, out, block_sums); break;
    case 5: launch_tpt_l<5>(write, hash, blocks, s, P, counts, base, out, block_sums); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace amj
