// Pairwise joins with materialised intermediates (opts.engine = AMJ_ENGINE_PW): the comparison
// point of the multi-way design.
//
// Reference: PW() (MHJ-GPU/HashJoinEPFL/hashJoinEPFL.cuh:174-270) joins the chain one build
// table at a time (right-deep), each step a partitioned hash join (hashJoinEPFL: radix-partition
// both sides so that a build partition fits shared memory, then join partition pairs there) that
// returns pairs of matching row indexes; the intermediate keeps one row-index array per joined
// relation (late materialisation), the next join column is gathered through them, and the
// result columns are gathered at the end.
//
// Here, per build table t:
//   * the intermediate's join column K is gathered from the relation that bound the attribute;
//   * partitions of the build side are ranges of 2^shift consecutive buckets of its bucketed
//     table (rows ordered by key & (B-1), so a partition is a contiguous row range, about 1024
//     rows); the intermediate is partitioned by the same bits with one stable radix sort of
//     (partition, index) pairs;
//   * one workgroup per partition stages the partition's bucket pointers and keys in LDS and
//     looks up each of its intermediate tuples there (the key's run inside its bucket), counts
//     in one pass and writes (intermediate row, build row) pairs at scanned positions in a
//     second;
//   * the row-index arrays of the relations joined so far are gathered through the pairs.
// Finally every attribute is gathered through its relation's index array and hashed (and
// written when materialising). Chains whose tables each bind one new attribute (or none) run
// here; compared columns (TRI/FOUR) do not.
#include <algorithm>
#include <cstring>

#include "probe_common.h"
#include "runtime.h"

namespace amj {
namespace {

constexpr uint32_t kPwPartRows = 1024;   // build rows per partition on average
constexpr uint32_t kPwLdsRows = 4096;    // partitions up to this many rows are staged in LDS
constexpr uint32_t kPwMaxShift = 10;     // at most 1024 buckets per partition

// dst[i] = src[idx[i]] (idx == nullptr: src[i]).
__global__ __launch_bounds__(kBlock) void pw_gather_kernel(const int32_t* __restrict__ src,
                                                           const uint32_t* __restrict__ idx,
                                                           uint64_t m, int32_t* __restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * kBlock)
    dst[i] = src[idx ? idx[i] : i];
}

// idx_out[k] = idx_in[sel[k]] (idx_in == nullptr: sel[k]).
__global__ __launch_bounds__(kBlock) void pw_compose_kernel(const uint32_t* __restrict__ idx_in,
                                                            const uint32_t* __restrict__ sel,
                                                            uint64_t m,
                                                            uint32_t* __restrict__ idx_out) {
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * kBlock)
    idx_out[k] = idx_in ? idx_in[sel[k]] : sel[k];
}

// (partition of K[i]) in the top np_bits, i below.
__global__ __launch_bounds__(kBlock) void pw_part_kernel(const int32_t* __restrict__ K, uint64_t m,
                                                         uint32_t bmask, uint32_t shift,
                                                         uint32_t np_bits,
                                                         uint64_t* __restrict__ comp) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t p = (((uint32_t)K[i]) & bmask) >> shift;
    comp[i] = np_bits ? ((p << (64 - np_bits)) | i) : i;
  }
}

// Partition p of the build table: buckets [p << shift, (p + 1) << shift). Each intermediate
// tuple of the partition (sorted positions [sptr[p], sptr[p+1])) finds its key's run in its
// bucket; count pass: counts[j]; write pass: pairs at base[j].
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void pw_join_kernel(
    const DevTable T, uint32_t shift, const uint64_t* __restrict__ scomp,
    const uint32_t* __restrict__ sptr, const int32_t* __restrict__ K,
    uint64_t* __restrict__ counts, const uint64_t* __restrict__ base,
    uint32_t* __restrict__ out_i, uint32_t* __restrict__ out_r) {
  __shared__ uint32_t lptr[(1u << kPwMaxShift) + 1];
  __shared__ int32_t lkey[kPwLdsRows];
  const uint32_t p = blockIdx.x;
  const uint64_t B = 1ull << T.log_b;
  const uint64_t b0 = (uint64_t)p << shift;
  const uint64_t b1 = std::min<uint64_t>(b0 + (1ull << shift), B);
  const uint32_t nb = (uint32_t)(b1 - b0);
  const uint32_t r0 = T.ptrs[b0], r1 = T.ptrs[b1];
  const bool staged = r1 - r0 <= kPwLdsRows;
  for (uint32_t x = threadIdx.x; x <= nb; x += kBlock) lptr[x] = T.ptrs[b0 + x];
  if (staged)
    for (uint32_t x = threadIdx.x; x < r1 - r0; x += kBlock) lkey[x] = T.keys[r0 + x];
  __syncthreads();
  const uint32_t bmask = (uint32_t)(B - 1);
  for (uint32_t j = sptr[p] + threadIdx.x; j < sptr[p + 1]; j += kBlock) {
    const uint32_t i = (uint32_t)scomp[j];
    const int32_t key = K[i];
    const uint32_t b = (((uint32_t)key) & bmask) - (uint32_t)b0;
    uint32_t s = lptr[b], e = lptr[b + 1];
    // the key's rows inside its bucket (keys ascending as unsigned values there)
    if (staged) {
      const uint32_t* lk = (const uint32_t*)lkey;
      uint32_t lo = s - r0, hi = e - r0;
      while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (lk[mid] < (uint32_t)key) lo = mid + 1;
        else hi = mid;
      }
      uint32_t lo2 = lo, hi2 = e - r0;
      while (lo2 < hi2) {
        const uint32_t mid = lo2 + ((hi2 - lo2) >> 1);
        if (lk[mid] <= (uint32_t)key) lo2 = mid + 1;
        else hi2 = mid;
      }
      s = lo + r0;
      e = lo2 + r0;
    } else {
      const uint32_t lo = lower_bound_u32(T.keys, s, e, (uint32_t)key);
      e = upper_bound_u32(T.keys, lo, e, (uint32_t)key);
      s = lo;
    }
    if (!WRITE) {
      counts[j] = e - s;
    } else {
      uint64_t pos = base[j];
      for (uint32_t r = s; r < e; ++r, ++pos) {
        out_i[pos] = i;
        out_r[pos] = r;
      }
    }
  }
}

struct PwCols {
  const int32_t* src[kMaxAttrs];   // column holding attribute a
  const uint32_t* idx[kMaxAttrs];  // its row of intermediate tuple k (nullptr: k itself)
  int32_t* out[kMaxAttrs];         // output column (nullptr: not written)
};

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void pw_emit_kernel(const PwCols C, int n_attrs, uint64_t m,
                                                         uint64_t* __restrict__ block_sums) {
  uint64_t csum = 0;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * kBlock) {
    int32_t v[kMaxAttrs];
    sfor<0, kMaxAttrs>([&](auto a) {
      v[a] = (int)a < n_attrs ? C.src[a][C.idx[a] ? C.idx[a][k] : k] : 0;
      if (WRITE && (int)a < n_attrs) C.out[a][k] = v[a];
    });
    csum += hash_tuple(v, n_attrs);
  }
  block_sum_store(csum, &block_sums[blockIdx.x]);
}

unsigned pw_grid(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  return g ? (unsigned)g : 1u;
}

// Device buffers of one query, freed on every exit.
struct PwBufs {
  amj_ctx_t* ctx;
  std::vector<std::pair<void*, size_t>> held;
  explicit PwBufs(amj_ctx_t* c) : ctx(c) {}
  ~PwBufs() {
    if (!held.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (auto& h : held) dfree(ctx, h.first, h.second);
  }
  template <typename T>
  int get(uint64_t n, T** p) {
    void* q = nullptr;
    const size_t bytes = std::max<uint64_t>(n, 1) * sizeof(T);
    AMJ_TRY(dmalloc(ctx, &q, bytes));
    held.emplace_back(q, bytes);
    *p = (T*)q;
    return AMJ_OK;
  }
  void drop(void* p) {
    for (size_t x = 0; x < held.size(); ++x)
      if (held[x].first == p) {
        (void)hipStreamSynchronize(ctx->stream);
        dfree(ctx, held[x].first, held[x].second);
        held.erase(held.begin() + x);
        return;
      }
  }
};

}  // namespace

int run_pw(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const int L = P.n_levels;
  for (int t = 0; t < L; ++t)
    if (P.val_mode[t] == VAL_CHECK) {
      set_error(ctx, AMJ_ERR_ARG, "engine AMJ_ENGINE_PW joins chains whose tables bind a new "
                                  "attribute (no compared columns: TRI/FOUR are not pairwise)");
      return AMJ_ERR_ARG;
    }
  if (P.probe_n >= 0xffffffffull) {
    set_error(ctx, AMJ_ERR_ARG, "engine AMJ_ENGINE_PW: the probe relation exceeds 2^32 rows");
    return AMJ_ERR_ARG;
  }
  // binder of every attribute: relation (-1: probe, t: table t's second column) and column
  int bind_rel[kMaxAttrs], bind_col[kMaxAttrs];
  for (int a = 0; a < kMaxAttrs; ++a) bind_rel[a] = -2;
  for (int c = 0; c < P.probe_ncols; ++c) {
    bind_rel[P.probe_attr[c]] = -1;
    bind_col[P.probe_attr[c]] = c;
  }
  for (int t = 0; t < L; ++t)
    if (P.val_mode[t] == VAL_BIND) bind_rel[P.val_attr[t]] = t;
  auto column_of = [&](int a) -> const int32_t* {
    return bind_rel[a] == -1 ? P.probe_cols[bind_col[a]] : P.tab[bind_rel[a]].vals;
  };
  PwBufs B(ctx);
  // idx[r + 1]: row of relation r (-1: probe) of every intermediate tuple (nullptr: identity)
  std::vector<uint32_t*> idx(L + 1, nullptr);
  uint64_t m = P.probe_n;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  for (int t = 0; t < L && m > 0; ++t) {
    const DevTable& T = P.tab[t];
    // 1. the join column of the intermediate
    int32_t* K = nullptr;
    AMJ_TRY(B.get(m, &K));
    const int ka = P.key_attr[t];
    hipLaunchKernelGGL(pw_gather_kernel, dim3(pw_grid(m)), dim3(kBlock), 0, s, column_of(ka),
                       idx[bind_rel[ka] + 1], m, K);
    AMJ_HIP_TRY(ctx, hipGetLastError());
    // 2. partitions: 2^shift buckets each, about kPwPartRows build rows
    uint32_t shift = 0;  // rows of a partition with shift + 1: n * 2^(shift + 1) / B
    while (shift < kPwMaxShift && shift < T.log_b &&
           ((T.n << (shift + 1)) >> T.log_b) <= kPwPartRows)
      ++shift;
    const uint32_t np_bits = T.log_b - shift;
    const uint64_t parts = 1ull << np_bits;
    uint64_t *comp = nullptr, *alt = nullptr;
    uint32_t* sptr = nullptr;
    uint8_t* stmp = nullptr;
    AMJ_TRY(B.get(m, &comp));
    AMJ_TRY(B.get(m, &alt));
    AMJ_TRY(B.get(parts + 1, &sptr));
    AMJ_TRY(B.get(radix_sort_temp_bytes(m), &stmp));
    hipLaunchKernelGGL(pw_part_kernel, dim3(pw_grid(m)), dim3(kBlock), 0, s, K, m,
                       (uint32_t)((1ull << T.log_b) - 1), shift, np_bits, comp);
    AMJ_HIP_TRY(ctx, hipGetLastError());
    if (np_bits) {
      const uint32_t first_digit = (64 - np_bits) / 8;
      const uint32_t mask = 0xffu & ~((1u << first_digit) - 1u);
      AMJ_HIP_TRY(ctx, radix_sort_u64(comp, alt, m, mask, stmp, s));
    }
    AMJ_HIP_TRY(ctx, launch_bucket_ptrs(comp, m, np_bits, parts, sptr, s));
    B.drop(alt);
    // 3. count, scan, write the pairs
    uint64_t *cnt = nullptr, *tmp = nullptr;
    AMJ_TRY(B.get(m + 1, &cnt));
    AMJ_TRY(B.get(scan_temp_elems(m + 1), &tmp));
    AMJ_HIP_TRY(ctx, hipMemsetAsync(cnt + m, 0, 8, s));
    hipLaunchKernelGGL(pw_join_kernel<false>, dim3((unsigned)parts), dim3(kBlock), 0, s, T,
                       shift, (const uint64_t*)comp, (const uint32_t*)sptr, (const int32_t*)K,
                       cnt, (const uint64_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr);
    AMJ_HIP_TRY(ctx, hipGetLastError());
    AMJ_HIP_TRY(ctx, exclusive_scan_u64(cnt, cnt, m + 1, tmp, s));
    uint64_t m2 = 0;
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&m2, cnt + m, 8, hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (m2 >= 0xffffffffull) {
      set_error(ctx, AMJ_ERR_ARG, "engine AMJ_ENGINE_PW: an intermediate exceeds 2^32 rows");
      return AMJ_ERR_ARG;
    }
    uint32_t *pi = nullptr, *pr = nullptr;
    AMJ_TRY(B.get(m2, &pi));
    AMJ_TRY(B.get(m2, &pr));
    if (m2)
      hipLaunchKernelGGL(pw_join_kernel<true>, dim3((unsigned)parts), dim3(kBlock), 0, s, T,
                         shift, (const uint64_t*)comp, (const uint32_t*)sptr,
                         (const int32_t*)K, (uint64_t*)nullptr, (const uint64_t*)cnt, pi, pr);
    AMJ_HIP_TRY(ctx, hipGetLastError());
    // 4. the index arrays of the relations joined so far, through the pairs
    for (int r = -1; r < t; ++r) {
      uint32_t* nx = nullptr;
      AMJ_TRY(B.get(m2, &nx));
      if (m2)
        hipLaunchKernelGGL(pw_compose_kernel, dim3(pw_grid(m2)), dim3(kBlock), 0, s,
                           (const uint32_t*)idx[r + 1], (const uint32_t*)pi, m2, nx);
      if (idx[r + 1]) B.drop(idx[r + 1]);
      idx[r + 1] = nx;
    }
    idx[t + 1] = pr;
    for (void* x : {(void*)K, (void*)comp, (void*)sptr, (void*)stmp, (void*)cnt, (void*)tmp,
                    (void*)pi})
      B.drop(x);
    m = m2;
  }
  if (L > 0 && m > 0 && !idx[L]) m = 0;  // an empty intermediate stopped the loop
  // 5. result: every attribute gathered through its relation's index array, hashed, written
  res->count = m;
  const bool mat = o->materialize != 0;
  if (mat && m > 0) {
    if (m * (uint64_t)P.n_attrs * 4 > budget_left(ctx, o)) {
      set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(m) +
                                        " tuples does not fit the HBM budget (use materialize=0)");
      return AMJ_ERR_NOMEM;
    }
    AMJ_TRY(alloc_output(ctx, m, P.n_attrs, res));
  }
  PwCols C{};
  for (int a = 0; a < P.n_attrs; ++a) {
    C.src[a] = column_of(a);
    C.idx[a] = idx[bind_rel[a] + 1];
    C.out[a] = mat && m > 0 ? res->cols[a] : nullptr;
  }
  const unsigned g = pw_grid(m);
  uint64_t* bsum = nullptr;
  AMJ_TRY(B.get(g, &bsum));
  if (m > 0) {
    if (mat) hipLaunchKernelGGL(pw_emit_kernel<true>, dim3(g), dim3(kBlock), 0, s, C, P.n_attrs, m,
                                bsum);
    else hipLaunchKernelGGL(pw_emit_kernel<false>, dim3(g), dim3(kBlock), 0, s, C, P.n_attrs, m,
                            bsum);
    AMJ_HIP_TRY(ctx, hipGetLastError());
    AMJ_TRY(finish_checksum(ctx, bsum, g, &res->checksum));
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = res->probe_ms;
  res->probe_kernel_launches = (uint64_t)L * 2 + 1;
  res->rows_written = mat ? m : 0;
  res->path = AMJ_PATH_PW;
  res->passes = (uint32_t)L;
  return AMJ_OK;
}

}  // namespace amj
