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
#include "lftj_common.h"

namespace amj {

// Level 0: 64 consecutive level-0 values of driver0.
template <int A>
__device__ __forceinline__ void level0(const LPlan& P, uint64_t unit, LLv<A>& l) {
  const uint64_t r = unit * kWave + (uint64_t)lane_id();
  const bool valid = r < P.roots;
  sfor<0, A>([&](auto a) { l.b[a] = 0; });
  sfor<0, kR>([&](auto q) { l.lo[q] = 0u; l.hi[q] = 0u; });
  int32_t v = 0;
  const LRel& d = P.rel[P.driver0];
  if (valid) v = d.d0[r];
  bool ok = valid;
  if (valid) ok = check_and_open(P, 0, v, -1, true, l);
  l.b[0] = v;
  choose_driver(P, 1, ok, l);
  finish(l);
}

template <int A, int i, int MODE>
__device__ __forceinline__ void descend(const LPlan& P, const LLv<A>& par,
                                        const LOut& cols, LAcc& acc) {
  for (uint32_t c = 0; c < par.T; c += (uint32_t)kWave) {
    const uint32_t g = c + (uint32_t)lane_id();
    const bool valid = g < par.T;
    const int p = (int)parent_of_slots(par.excl, par.dlen, c);
    const uint32_t k = g - (uint32_t)__shfl(par.excl, p, kWave);
    const int32_t* src = (const int32_t*)__shfl((uint64_t)par.dptr, p, kWave);
    const int drel = __shfl(par.drel, p, kWave);
    LLv<A> ch;
    sfor<0, A>([&](auto a) { ch.b[a] = __shfl(par.b[a], p, kWave); });
    sfor<0, kR>([&](auto q) {
      ch.lo[q] = __shfl(par.lo[q], p, kWave);
      ch.hi[q] = __shfl(par.hi[q], p, kWave);
    });
    const int32_t v = valid ? src[k] : 0;
    const bool drel_root = P.child_mask[i] == 0u;
    bool ok = valid && check_and_open(P, i, v, drel, drel_root, ch);
    ch.b[i] = v;
    if constexpr (i == A - 1) {
      const uint64_t m = __ballot(ok);
      const uint32_t n = (uint32_t)__popcll(m);
      if (MODE != MODE_COUNT && ok) {
        if (mode_hash<MODE>()) {
          int32_t w[kMaxAttrs];
          sfor<0, kMaxAttrs>([&](auto a) { w[a] = (a < A) ? ch.b[(a < A) ? (int)a : 0] : 0; });
          acc.csum += hash_tuple(w, A);
        }
        if (mode_mat<MODE>()) {
          const uint64_t pos = acc.out_base + mbcnt(m);
          sfor<0, A>([&](auto a) { cols.c[a][pos] = ch.b[a]; });
        }
      }
      acc.count += n;
      acc.out_base += n;
    } else {
      choose_driver(P, i + 1, ok, ch);
      finish(ch);
      if (ch.T) descend<A, i + 1, MODE>(P, ch, cols, acc);
    }
  }
}

template <int A>
__global__ __launch_bounds__(kBlock) void lftj_count_kernel(const LPlan P,
                                                            uint64_t* __restrict__ counts,
                                                            uint64_t units) {
  const uint64_t unit = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (unit >= units) return;  // no barriers in this kernel
  LLv<A> l0;
  level0(P, unit, l0);
  LAcc acc{0, 0, 0};
  if (l0.T) descend<A, 1, MODE_COUNT>(P, l0, LOut{}, acc);
  if (lane_id() == 0) counts[unit] = acc.count;
}

template <int A, int MODE>
__global__ __launch_bounds__(kBlock) void lftj_write_kernel(const LPlan P,
                                                            const uint64_t* __restrict__ base,
                                                            const LOut cols,
                                                            uint64_t* __restrict__ block_sums,
                                                            uint64_t units) {
  const uint64_t unit = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  LAcc acc{0, 0, 0};
  if (unit < units) {
    LLv<A> l0;
    level0(P, unit, l0);
    acc.out_base = uniform(base[unit]);
    if (l0.T) descend<A, 1, MODE>(P, l0, cols, acc);
  }
  block_sum_store(acc.csum, &block_sums[blockIdx.x]);
}

uint64_t lftj_units(const LPlan& P) { return (P.roots + kWave - 1) / kWave; }

hipError_t launch_lftj_count(const LPlan& P, uint64_t* counts, hipStream_t s) {
  const uint64_t units = lftj_units(P);
  if (units == 0) return hipSuccess;
  const dim3 grid((unsigned)((units + kWavesPerBlock - 1) / kWavesPerBlock));
  switch (P.n_attrs) {
    case 2: hipLaunchKernelGGL(lftj_count_kernel<2>, grid, dim3(kBlock), 0, s, P, counts, units); break;
    case 3: hipLaunchKernelGGL(lftj_count_kernel<3>, grid, dim3(kBlock), 0, s, P, counts, units); break;
    case 4: hipLaunchKernelGGL(lftj_count_kernel<4>, grid, dim3(kBlock), 0, s, P, counts, units); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lftj_write(const LPlan& P, const uint64_t* base, const LOut& cols,
                             bool hash, uint64_t* block_sums, uint64_t* n_blocks, hipStream_t s) {
  const uint64_t units = lftj_units(P);
  const uint64_t blocks = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  *n_blocks = blocks;
  const dim3 grid((unsigned)blocks);
  const int mode = cols.c[0] ? (hash ? MODE_WRITE_HASH : MODE_WRITE) : MODE_HASH;
#define LW(AV, M) \
  hipLaunchKernelGGL((lftj_write_kernel<AV, M>), grid, dim3(kBlock), 0, s, P, base, cols, block_sums, units)
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
