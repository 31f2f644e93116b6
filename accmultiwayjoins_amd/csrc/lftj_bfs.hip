// Breadth-first leapfrog triejoin (the reference's LFTJ_BFS).
//
// Reference: LFTJ_BFS::evaluate (LFTJ-GPU/LFTJ_BFS.cuh:457-500) runs the triangle query level by
// level with library primitives: np_join_np intersects the level-0 lists (moderngpu
// sorted_search + cub select, :143-186), p_join_np expands children with a load-balancing search
// and binary searches them (:196-332), p_join_p expands both sides and counts matching pairs
// (:340-454, with a host OpenMP load-balance search). It returns an accumulator that nothing
// updates, so the count is never reported (:496-499).
//
// Here every level but the last is materialised explicitly: level 0 keeps the surviving roots
// of the smallest level-0 list; at level i every frontier entry picks its shortest candidate
// list, a scan of those lengths flattens all candidates, and one thread per candidate finds its
// entry by binary search over the scan, checks the candidate against the other lists and writes
// the extended entry, compacted by a scan of the survivor flags (candidates in batches sized to
// the memory budget). The last attribute is counted from the materialised frontier 64 entries
// per wave, by the DFS engine's last level (lftj_dfs.h; for a triangle or a 4-clique its
// cooperative form, the other lists staged in LDS): one thread per last-level candidate had to
// find its entry by a binary search over the whole frontier first. The count and checksum are
// returned; any query the DFS engine runs also runs here, and the two must agree.
#include "lftj_dfs.h"
#include "runtime.h"

namespace amj {

template <int A>
struct BfsEntry {
  int32_t b[A];
  uint32_t lo[kR];
  uint32_t hi[kR];
};

template <int A>
__device__ __forceinline__ void to_lv(const BfsEntry<A>& e, LLv<A>& l) {
  sfor<0, A>([&](auto a) { l.b[a] = e.b[a]; });
  sfor<0, kR>([&](auto r) {
    l.lo[r] = e.lo[r];
    l.hi[r] = e.hi[r];
  });
}
template <int A>
__device__ __forceinline__ void from_lv(const LLv<A>& l, BfsEntry<A>& e) {
  sfor<0, A>([&](auto a) { e.b[a] = l.b[a]; });
  sfor<0, kR>([&](auto r) {
    e.lo[r] = l.lo[r];
    e.hi[r] = l.hi[r];
  });
}

// Level 0: every level-0 value of driver0 that passes the other level-0 lists.
template <int A>
__global__ __launch_bounds__(kBlock) void bfs_roots_kernel(const LPlan P, BfsEntry<A>* out,
                                                           uint32_t* __restrict__ keep) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < P.roots;
       r += (uint64_t)gridDim.x * kBlock) {
    LLv<A> l;
    sfor<0, A>([&](auto a) { l.b[a] = 0; });
    sfor<0, kR>([&](auto q) { l.lo[q] = 0u; l.hi[q] = 0u; });
    const int32_t v = P.rel[P.driver0].d0[r];
    const bool ok = check_and_open(P, 0, v, -1, true, l);
    l.b[0] = v;
    BfsEntry<A> e;
    from_lv(l, e);
    out[r] = e;
    keep[r] = ok ? 1u : 0u;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bfs_compact_kernel(const T* __restrict__ in,
                                                             const uint32_t* __restrict__ keep,
                                                             const uint64_t* __restrict__ pos,
                                                             uint64_t n, T* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    if (keep[i]) out[pos[i]] = in[i];
}

__global__ __launch_bounds__(kBlock) void bfs_u32_to_u64_kernel(const uint32_t* __restrict__ a,
                                                                uint64_t n,
                                                                uint64_t* __restrict__ b) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * kBlock)
    b[i] = i < n ? a[i] : 0ull;
}

// Candidate count of every entry for attribute i.
template <int A>
__global__ __launch_bounds__(kBlock) void bfs_sizes_kernel(const LPlan P, int i,
                                                           const BfsEntry<A>* __restrict__ f,
                                                           uint64_t n, uint64_t* __restrict__ len) {
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e <= n;
       e += (uint64_t)gridDim.x * kBlock) {
    if (e == n) {
      len[e] = 0;
      continue;
    }
    LLv<A> l;
    to_lv(f[e], l);
    choose_driver(P, i, true, l);
    len[e] = l.dlen;
  }
}

// Candidates [g0, g1) of attribute i (not the last): entry by binary search over the scanned
// lengths; the extended entry and its survivor flag are written at g - g0.
template <int A>
__global__ __launch_bounds__(kBlock) void bfs_expand_kernel(
    const LPlan P, int i, const BfsEntry<A>* __restrict__ f, uint64_t n,
    const uint64_t* __restrict__ off, uint64_t g0, uint64_t g1, BfsEntry<A>* __restrict__ next,
    uint32_t* __restrict__ keep) {
  for (uint64_t g = g0 + (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < g1;
       g += (uint64_t)gridDim.x * kBlock) {
    uint64_t lo = 0, hi = n;  // last e with off[e] <= g
    while (hi - lo > 1) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (off[mid] <= g) lo = mid;
      else hi = mid;
    }
    LLv<A> l;
    to_lv(f[lo], l);
    choose_driver(P, i, true, l);
    const int32_t v = l.dptr[g - off[lo]];
    const bool ok = check_and_open(P, i, v, l.drel, P.child_mask[i] == 0u, l);
    sfor<0, A>([&](auto a) {  // l.b[i] = v without a dynamically indexed (scratch) array
      if ((int)a == i) l.b[a] = v;
    });
    BfsEntry<A> e;
    from_lv(l, e);
    next[g - g0] = e;
    keep[g - g0] = ok ? 1u : 0u;
  }
}

// The last attribute over a materialised frontier of n entries, 64 per wave: count and
// checksum (per block: block_sums[2b], block_sums[2b + 1]).
template <int A, uint64_t SH>
__global__ __launch_bounds__(kBlock) void bfs_last_kernel(const LPlan P,
                                                          const BfsEntry<A>* __restrict__ f,
                                                          uint64_t n,
                                                          uint64_t* __restrict__ block_sums) {
  __shared__ int32_t coop_mem[SH != 0 ? kWavesPerBlock * stage_ints<SH>() : 1];
  lds_i32* stage =
      SH != 0 ? (lds_i32*)&coop_mem[(threadIdx.x / kWave) * stage_ints<SH>()] : nullptr;
  const uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  CountHashSink<true> sink;
  if (w * kWave < n) {  // wave-uniform
    const uint64_t e = w * kWave + (uint64_t)lane_id();
    const bool valid = e < n;
    LLv<A> l;
    if (valid) {
      to_lv(f[e], l);
    } else {
      sfor<0, A>([&](auto a) { l.b[a] = 0; });
      sfor<0, kR>([&](auto q) { l.lo[q] = 0u; l.hi[q] = 0u; });
    }
    choose_driver<A, SH>(P, A - 1, valid, l);
    finish(l);
    if (l.T) descend<A, A - 1, SH>(P, l, sink, stage);
  }
  block_sum_store(lane_id() == 0 ? sink.count : 0, &block_sums[2 * blockIdx.x]);
  block_sum_store(sink.csum, &block_sums[2 * blockIdx.x + 1]);
}

template <int A>
static hipError_t launch_bfs_last(const LPlan& P, const BfsEntry<A>* f, uint64_t n,
                                  uint64_t* block_sums, unsigned blocks, hipStream_t s) {
  const uint64_t sh = lftj_static_shape(P);
  if constexpr (A == 3) {
    if (sh == kShapeTri) {
      hipLaunchKernelGGL((bfs_last_kernel<3, kShapeTri>), dim3(blocks), dim3(kBlock), 0, s, P,
                         f, n, block_sums);
      return hipGetLastError();
    }
  }
  if constexpr (A == 4) {
    if (sh == kShapeFour) {
      hipLaunchKernelGGL((bfs_last_kernel<4, kShapeFour>), dim3(blocks), dim3(kBlock), 0, s, P,
                         f, n, block_sums);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((bfs_last_kernel<A, 0>), dim3(blocks), dim3(kBlock), 0, s, P, f, n,
                     block_sums);
  return hipGetLastError();
}

static unsigned grid_n(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 16384) g = 16384;
  return g ? (unsigned)g : 1u;
}

// Device buffer that grows by doubling (keeps its contents).
struct GrowBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
static int grow(amj_ctx_t* ctx, GrowBuf& b, size_t need, size_t keep_bytes) {
  if (b.bytes >= need) return AMJ_OK;
  size_t want = std::max(need, b.bytes * 2);
  void* q = nullptr;
  AMJ_TRY(dmalloc(ctx, &q, want));
  if (b.p && keep_bytes)
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(q, b.p, keep_bytes, hipMemcpyDeviceToDevice, ctx->stream));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  dfree(ctx, b.p, b.bytes);
  b.p = q;
  b.bytes = want;
  return AMJ_OK;
}

template <int A>
static int bfs_run(amj_ctx_t* ctx, const LPlan& P, uint64_t budget, uint64_t* count,
                   uint64_t* checksum) {
  using E = BfsEntry<A>;
  hipStream_t s = ctx->stream;
  GrowBuf cur, nxt, tmp, keep, pos, len;
  uint64_t n = P.roots;
  int rc = AMJ_OK;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    for (GrowBuf* b : {&cur, &nxt, &tmp, &keep, &pos, &len}) dfree(ctx, b->p, b->bytes);
  };
  void* stmp = nullptr;
  uint64_t total = 0, csum = 0;
  do {
    // level 0
    if ((rc = grow(ctx, tmp, std::max<uint64_t>(n, 1) * sizeof(E), 0)) != AMJ_OK) break;
    if ((rc = grow(ctx, keep, (n + 1) * 4, 0)) != AMJ_OK) break;
    if ((rc = grow(ctx, pos, (n + 1) * 8, 0)) != AMJ_OK) break;
    if ((rc = scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(n + 1) * 8, &stmp)) != AMJ_OK) break;
    hipLaunchKernelGGL(bfs_roots_kernel<A>, dim3(grid_n(n)), dim3(kBlock), 0, s, P, (E*)tmp.p,
                       (uint32_t*)keep.p);
    hipLaunchKernelGGL(bfs_u32_to_u64_kernel, dim3(grid_n(n + 1)), dim3(kBlock), 0, s,
                       (const uint32_t*)keep.p, n, (uint64_t*)pos.p);
    AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)pos.p, (uint64_t*)pos.p, n + 1,
                                        (uint64_t*)stmp, s));
    uint64_t m = 0;
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&m, (uint64_t*)pos.p + n, 8, hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
    if ((rc = grow(ctx, cur, std::max<uint64_t>(m, 1) * sizeof(E), 0)) != AMJ_OK) break;
    hipLaunchKernelGGL(bfs_compact_kernel<E>, dim3(grid_n(n)), dim3(kBlock), 0, s,
                       (const E*)tmp.p, (const uint32_t*)keep.p, (const uint64_t*)pos.p, n,
                       (E*)cur.p);
    n = m;
    for (int i = 1; i < A && n > 0; ++i) {
      const bool last = i == A - 1;
      if (last) {  // counted from the frontier, 64 entries per wave
        const uint64_t waves = (n + kWave - 1) / kWave;
        const uint64_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
        void* bs = nullptr;
        if ((rc = scratch(ctx, SLOT_BLOCK_SUMS, blocks * 16, &bs)) != AMJ_OK) break;
        AMJ_HIP_TRY(ctx, launch_bfs_last<A>(P, (const E*)cur.p, n, (uint64_t*)bs,
                                            (unsigned)blocks, s));
        std::vector<uint64_t> h((size_t)blocks * 2);
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(h.data(), bs, h.size() * 8, hipMemcpyDeviceToHost, s));
        AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
        for (uint64_t b = 0; b < blocks; ++b) {
          total += h[2 * b];
          csum += h[2 * b + 1];
        }
        break;
      }
      if ((rc = grow(ctx, len, (n + 1) * 8, 0)) != AMJ_OK) break;
      if ((rc = scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(n + 1) * 8, &stmp)) != AMJ_OK) break;
      hipLaunchKernelGGL(bfs_sizes_kernel<A>, dim3(grid_n(n + 1)), dim3(kBlock), 0, s, P, i,
                         (const E*)cur.p, n, (uint64_t*)len.p);
      AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)len.p, (uint64_t*)len.p, n + 1,
                                          (uint64_t*)stmp, s));
      uint64_t C = 0;
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(&C, (uint64_t*)len.p + n, 8, hipMemcpyDeviceToHost, s));
      AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
      // candidates per batch: bounded by the budget (entry + flag + position per candidate)
      const uint64_t per = sizeof(E) + 4 + 8;
      const uint64_t batch =
          std::min<uint64_t>(std::max<uint64_t>(budget / 4 / per, 1ull << 20), 1ull << 27);
      uint64_t next_n = 0;
      for (uint64_t g0 = 0; g0 < C; g0 += batch) {
        const uint64_t g1 = std::min<uint64_t>(C, g0 + batch);
        const uint64_t nb = g1 - g0;
        const unsigned grid = grid_n(nb);
        if ((rc = grow(ctx, tmp, nb * sizeof(E), 0)) != AMJ_OK) break;
        if ((rc = grow(ctx, keep, (nb + 1) * 4, 0)) != AMJ_OK) break;
        if ((rc = grow(ctx, pos, (nb + 1) * 8, 0)) != AMJ_OK) break;
        if ((rc = scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(nb + 1) * 8, &stmp)) != AMJ_OK)
          break;
        hipLaunchKernelGGL((bfs_expand_kernel<A>), dim3(grid), dim3(kBlock), 0, s, P, i,
                           (const E*)cur.p, n, (const uint64_t*)len.p, g0, g1, (E*)tmp.p,
                           (uint32_t*)keep.p);
        hipLaunchKernelGGL(bfs_u32_to_u64_kernel, dim3(grid_n(nb + 1)), dim3(kBlock), 0, s,
                           (const uint32_t*)keep.p, nb, (uint64_t*)pos.p);
        AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)pos.p, (uint64_t*)pos.p, nb + 1,
                                            (uint64_t*)stmp, s));
        uint64_t m2 = 0;
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(&m2, (uint64_t*)pos.p + nb, 8, hipMemcpyDeviceToHost, s));
        AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
        if ((rc = grow(ctx, nxt, std::max<uint64_t>(next_n + m2, 1) * sizeof(E),
                       next_n * sizeof(E))) != AMJ_OK)
          break;
        hipLaunchKernelGGL(bfs_compact_kernel<E>, dim3(grid), dim3(kBlock), 0, s,
                           (const E*)tmp.p, (const uint32_t*)keep.p, (const uint64_t*)pos.p, nb,
                           (E*)nxt.p + next_n);
        next_n += m2;
      }
      if (rc != AMJ_OK) break;
      std::swap(cur, nxt);
      n = next_n;
    }
  } while (false);
  if (rc == AMJ_OK) {
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
    *count = total;
    *checksum = csum;
  }
  cleanup();
  return rc;
}

int run_lftj_bfs(amj_ctx_t* ctx, const LPlan& P, uint64_t budget, uint64_t* count,
                 uint64_t* checksum) {
  switch (P.n_attrs) {
    case 2: return bfs_run<2>(ctx, P, budget, count, checksum);
    case 3: return bfs_run<3>(ctx, P, budget, count, checksum);
    case 4: return bfs_run<4>(ctx, P, budget, count, checksum);
    default:
      set_error(ctx, AMJ_ERR_ARG, "LFTJ queries need 2..4 attributes");
      return AMJ_ERR_ARG;
  }
}

}  // namespace amj
