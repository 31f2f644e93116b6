// Thread-per-tuple multi-way hash join: the baseline the wave engines are measured against
// (MHJ_count / MHJ_write, MHJ-GPU/Joins/MHJ.cuh:17-321). Every lane takes probe rows
// grid-stride and walks the nested runs of the build tables on its own, depth first, with no
// cooperation inside the wave: lanes of one wave diverge as their runs differ in length, a
// heavy key serialises on one lane, and the result columns are written one scattered 4-byte
// store per lane. It runs the same plans and produces the same tuples (in another row order)
// and the same checksum as the wave engines; it exists as the ablation of their design.
//
// With open addressing (opts.engine = AMJ_ENGINE_OPEN_ADDR, the reference's AMHJ_open_addr.cuh
// over Indexing/linear.cuh) every build table is re-indexed before the probe as a linear-probing
// table of ceilPow2(2n) 64-bit slots (load factor at most 1/2, OA_HT_RATIO), slot = key << 32 |
// row, filled with compare-and-swap from the MurmurHash3 finalizer of the key; a lookup walks
// the cluster from the key's slot to the first empty one, and a key's rows are spread along it
// rather than contiguous. The traversal is the same lane-per-probe-row walk, so the pair
// (thread per tuple, open addressing) is the ablation of (wave frontier, bucketed runs).
#include "probe_common.h"
#include "tpt.h"

namespace amj {

// MurmurHash3's 32-bit finalizer (A. Appleby, public domain), as murmur_hash in
// MHJ-GPU/common_kernels.cuh:12-19.
__device__ __forceinline__ uint32_t fmix32(uint32_t k) {
  k ^= k >> 16;
  k *= 0x85ebca6bu;
  k ^= k >> 13;
  k *= 0xc2b2ae35u;
  k ^= k >> 16;
  return k;
}

constexpr uint64_t kOaEmpty = ~0ull;  // a filled slot holds a row index below 2^32 - 1

__global__ __launch_bounds__(kBlock) void oa_insert_kernel(const int32_t* __restrict__ keys,
                                                           uint64_t n, uint64_t* __restrict__ slots,
                                                           uint32_t mask) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * kBlock) {
    const uint32_t key = (uint32_t)keys[r];
    const uint64_t slot = ((uint64_t)key << 32) | r;
    uint32_t h = fmix32(key) & mask;
    // at most half the slots are filled, so an empty one is always found
    while (atomicCAS((unsigned long long*)&slots[h], (unsigned long long)kOaEmpty,
                     (unsigned long long)slot) != (unsigned long long)kOaEmpty)
      h = (h + 1) & mask;
  }
}

hipError_t launch_oa_insert(const int32_t* keys, uint64_t n, uint64_t* slots, uint32_t mask,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(oa_insert_kernel, dim3((unsigned)g), dim3(kBlock), 0, s, keys, n, slots,
                     mask);
  return hipGetLastError();
}

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
template <int t, int L, bool WRITE, bool HASH, bool OA>
__device__ __forceinline__ void tpt_descend(const Plan& P, const OaTables& oa,
                                            int32_t (&v)[kMaxAttrs], const ProbeOut& out,
                                            TptSink& k) {
  const DevTable& tb = P.tab[t];
  const int mode = P.val_mode[t];
  const int va = P.val_attr[t];
  const int32_t key = get_attr(v, P.key_attr[t]);
  if constexpr (OA) {
    // the cluster from the key's slot to the first empty slot holds every row of the key
    const uint64_t* slots = oa.slots[t];
    const uint32_t mask = oa.mask[t];
    const int32_t cval = mode == VAL_CHECK ? get_attr(v, va) : 0;
    for (uint32_t h = fmix32((uint32_t)key) & mask;; h = (h + 1) & mask) {
      const uint64_t sl = slots[h];
      if (sl == kOaEmpty) break;
      if ((int32_t)(uint32_t)(sl >> 32) != key) continue;
      const uint32_t r = (uint32_t)sl;
      if (mode == VAL_CHECK && tb.vals[r] != cval) continue;
      if (mode == VAL_BIND) set_attr(v, va, tb.vals[r]);
      if constexpr (t == L - 1) tpt_emit<WRITE, HASH>(P, v, out, k);
      else tpt_descend<t + 1, L, WRITE, HASH, OA>(P, oa, v, out, k);
    }
  } else {
    uint32_t lo, len;
    lookup_run(tb, key, mode == VAL_CHECK, mode == VAL_CHECK ? get_attr(v, va) : 0, true, lo,
               len);
    for (uint32_t r = lo; r < lo + len; ++r) {
      if (mode == VAL_BIND) set_attr(v, va, tb.vals[r]);
      if constexpr (t == L - 1) tpt_emit<WRITE, HASH>(P, v, out, k);
      else tpt_descend<t + 1, L, WRITE, HASH, OA>(P, oa, v, out, k);
    }
  }
}

// One lane per probe row, rows gtid, gtid + lanes, ... . Count pass: counts[gtid] = results of
// the lane's rows. Write pass: the lane writes its results from row base[gtid] on.
template <int L, bool WRITE, bool HASH, bool OA>
__global__ __launch_bounds__(kBlock) void tpt_kernel(const Plan P, const OaTables oa,
                                                     uint64_t* __restrict__ counts,
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
    tpt_descend<0, L, WRITE, HASH, OA>(P, oa, v, out, k);
  }
  if (!WRITE) counts[gtid] = k.count;
  if (HASH) block_sum_store(k.csum, &block_sums[blockIdx.x]);
}

uint64_t tpt_blocks(const Plan& P) {
  const uint64_t need = (P.probe_n + kBlock - 1) / kBlock;
  const uint64_t cap = 2048;  // 8 workgroups of 256 lanes per CU
  return need == 0 ? 1 : (need < cap ? need : cap);
}

template <int L, bool OA>
static void launch_tpt_l(bool write, bool hash, uint64_t blocks, hipStream_t s, const Plan& P,
                         const OaTables& oa, uint64_t* counts, const uint64_t* base,
                         const ProbeOut& out, uint64_t* block_sums) {
  const dim3 g((unsigned)blocks), b(kBlock);
  if (write && hash)
    hipLaunchKernelGGL((tpt_kernel<L, true, true, OA>), g, b, 0, s, P, oa, counts, base, out,
                       block_sums);
  else if (write)
    hipLaunchKernelGGL((tpt_kernel<L, true, false, OA>), g, b, 0, s, P, oa, counts, base, out,
                       block_sums);
  else if (hash)
    hipLaunchKernelGGL((tpt_kernel<L, false, true, OA>), g, b, 0, s, P, oa, counts, base, out,
                       block_sums);
  else
    hipLaunchKernelGGL((tpt_kernel<L, false, false, OA>), g, b, 0, s, P, oa, counts, base, out,
                       block_sums);
}

hipError_t launch_tpt(const Plan& P, bool write, bool hash, uint64_t* counts,
                      const uint64_t* base, const ProbeOut& out, uint64_t* block_sums,
                      hipStream_t s, const OaTables* oa) {
  const uint64_t blocks = tpt_blocks(P);
  const OaTables none{};
#define LT(LV)                                                                                 \
  if (oa) launch_tpt_l<LV, true>(write, hash, blocks, s, P, *oa, counts, base, out, block_sums); \
  else launch_tpt_l<LV, false>(write, hash, blocks, s, P, none, counts, base, out, block_sums)
  switch (P.n_levels) {
    case 1: LT(1); break;
    case 2: LT(2); break;
    case 3: LT(3); break;
    case 4: LT(4); break;
    case 5: LT(5); break;
    case 6: LT(6); break;
    case 7: LT(7); break;
    default: return hipErrorInvalidValue;
  }
#undef LT
  return hipGetLastError();
}

}  // namespace amj
