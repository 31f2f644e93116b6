// Sorted CSR trie build for the leapfrog triejoin (ALFTJ).
//
// Reference: constructSortedTrie<D,C,2> (LFTJ-GPU/multi_partitioning.cuh:129-293) sorts the
// (col0, col1) pairs with two cub radix sorts, marks run boundaries (extractMasks, :104-121) and
// compacts the distinct level-0 values and their child offsets with segmented selects; the last
// level is the sorted col1. It keeps duplicate pairs (a duplicate edge is a duplicate child).
//
// Here: one stable LSD radix sort of the 64-bit composite ((col0 ^ 2^31) << 32 | (col1 ^ 2^31))
// orders the pairs by (col0, col1) in signed order; duplicates are dropped (the triejoin
// intersects sets); a scan of the level-0 head flags gives the distinct values and their
// offsets. When the level-0 values are small non-negative integers (vertex ids) a table
// value -> child range is added so that "is v a level-0 value, and what are its children" is one
// 8-byte load.
#include "amj_internal.h"
#include "lftj.h"
#include "runtime.h"

namespace amj {

__global__ __launch_bounds__(kBlock) void trie_pack_kernel(const int32_t* __restrict__ c0,
                                                           const int32_t* __restrict__ c1,
                                                           uint64_t n, uint64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    out[i] = ((uint64_t)((uint32_t)c0[i] ^ 0x80000000u) << 32) |
             (uint64_t)((uint32_t)c1[i] ^ 0x80000000u);
}

// keep[i] = first occurrence of a pair (input sorted)
__global__ __launch_bounds__(kBlock) void trie_unique_flags_kernel(const uint64_t* __restrict__ k,
                                                                   uint64_t n,
                                                                   uint32_t* __restrict__ f) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    f[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

// compact the distinct pairs: d1[pos] = col1; head[pos] = 1 when col0 starts a new run
__global__ __launch_bounds__(kBlock) void trie_compact_kernel(const uint64_t* __restrict__ k,
                                                              uint64_t n,
                                                              const uint32_t* __restrict__ pos,
                                                              const uint32_t* __restrict__ keep,
                                                              int32_t* __restrict__ d1,
                                                              uint32_t* __restrict__ head) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    if (!keep[i]) continue;
    const uint32_t p = pos[i];
    d1[p] = (int32_t)((uint32_t)k[i] ^ 0x80000000u);
    head[p] = (i == 0 || (k[i] >> 32) != (k[i - 1] >> 32)) ? 1u : 0u;
  }
}

// level 0: d0[j] and off[j] for the j-th run head (hpos = exclusive scan of head)
__global__ __launch_bounds__(kBlock) void trie_level0_kernel(const uint64_t* __restrict__ k,
                                                             uint64_t n,
                                                             const uint32_t* __restrict__ pos,
                                                             const uint32_t* __restrict__ keep,
                                                             const uint32_t* __restrict__ head,
                                                             const uint32_t* __restrict__ hpos,
                                                             int32_t* __restrict__ d0,
                                                             uint32_t* __restrict__ off) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    if (!keep[i]) continue;
    const uint32_t p = pos[i];
    if (!head[p]) continue;
    const uint32_t j = hpos[p];
    d0[j] = (int32_t)((uint32_t)(k[i] >> 32) ^ 0x80000000u);
    off[j] = p;
  }
}

// vrange[v] = child range of level-0 value v (the table is zeroed first: absent values keep the
// empty range [0, 0)); every level-0 value has at least one child.
__global__ __launch_bounds__(kBlock) void trie_vrange_kernel(const int32_t* __restrict__ d0,
                                                             const uint32_t* __restrict__ off,
                                                             uint32_t n0,
                                                             uint2* __restrict__ vrange) {
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n0;
       j += (uint64_t)gridDim.x * kBlock)
    vrange[(uint32_t)d0[j]] = make_uint2(off[j], off[j + 1]);
}

// longest child run (the largest candidate list the triejoin frontier can take from this trie)
__global__ __launch_bounds__(kBlock) void trie_max_run_kernel(const uint32_t* __restrict__ off,
                                                              uint32_t n0,
                                                              uint32_t* __restrict__ out) {
  __shared__ uint32_t red[kBlock];
  uint32_t m = 0;
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n0;
       j += (uint64_t)gridDim.x * kBlock)
    m = max(m, off[j + 1] - off[j]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

static unsigned grid_for(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return g ? (unsigned)g : 1u;
}

// Longest child run, optional value -> child range table, timing and the device view. Expects d0,
// off, d1, n0 and n1 set; ev[0] recorded at the start of the build.
static int finish_trie(amj_ctx_t* ctx, TrieDev* out) {
  hipStream_t s = ctx->stream;
  const uint32_t n0 = out->n0;
  uint32_t max_run = 0;
  if (n0) {
    void* mr = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_FLAG, 256, &mr));
    AMJ_HIP_TRY(ctx, hipMemsetAsync(mr, 0, 4, s));
    hipLaunchKernelGGL(trie_max_run_kernel, dim3(std::min(grid_for(n0), 1024u)), dim3(kBlock), 0,
                       s, (const uint32_t*)out->off, n0, (uint32_t*)mr);
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&max_run, mr, 4, hipMemcpyDeviceToHost, s));
    // value -> child range table when level-0 values are small non-negative integers
    int32_t lo_v = 0, hi_v = 0;
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&lo_v, out->d0, 4, hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&hi_v, out->d0 + n0 - 1, 4, hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
    const uint64_t span = (uint64_t)(uint32_t)hi_v + 1;
    if (lo_v >= 0 && span <= std::max<uint64_t>(4ull * n0, 1ull << 20)) {
      AMJ_TRY(dmalloc(ctx, (void**)&out->vrange, span * 8));
      out->vrange_n = (uint32_t)span;
      AMJ_HIP_TRY(ctx, hipMemsetAsync(out->vrange, 0, span * 8, s));
      hipLaunchKernelGGL(trie_vrange_kernel, dim3(grid_for(n0)), dim3(kBlock), 0, s,
                         (const int32_t*)out->d0, (const uint32_t*)out->off, n0, out->vrange);
    }
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
  out->build_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
  out->max_run = max_run;
  out->view = LRel{out->d0, out->off, out->d1, out->vrange, n0, out->vrange_n};
  return AMJ_OK;
}

int build_trie(amj_ctx_t* ctx, const int32_t* c0, const int32_t* c1, uint64_t n,
               TrieDev* out) {
  *out = TrieDev{};
  if (n >= (1ull << 32)) {
    set_error(ctx, AMJ_ERR_ARG, "trie build: relations must have fewer than 2^32 rows");
    return AMJ_ERR_ARG;
  }
  hipStream_t s = ctx->stream;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
  if (n == 0) {
    AMJ_TRY(dmalloc(ctx, (void**)&out->off, 4));
    AMJ_HIP_TRY(ctx, hipMemsetAsync(out->off, 0, 4, s));
    return finish_trie(ctx, out);
  }
  void *keys = nullptr, *alt = nullptr, *tmp = nullptr, *keep = nullptr, *pos = nullptr,
       *head = nullptr, *hpos = nullptr, *stmp = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_SORT_KEYS, n * 8, &keys));
  AMJ_TRY(scratch(ctx, SLOT_SORT_ALT, n * 8, &alt));
  AMJ_TRY(scratch(ctx, SLOT_SORT_TEMP, radix_sort_temp_bytes(n), &tmp));
  AMJ_TRY(scratch(ctx, SLOT_TRIE0, (n + 1) * 4, &keep));
  AMJ_TRY(scratch(ctx, SLOT_TRIE1, (n + 1) * 4, &pos));
  AMJ_TRY(scratch(ctx, SLOT_TRIE2, (n + 1) * 4, &head));
  AMJ_TRY(scratch(ctx, SLOT_TRIE3, (n + 1) * 4, &hpos));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(n + 1) * 8, &stmp));
  hipLaunchKernelGGL(trie_pack_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, c0, c1, n,
                     (uint64_t*)keys);
  AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)keys, (uint64_t*)alt, n, 0xffu, tmp, s));
  hipLaunchKernelGGL(trie_unique_flags_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     (const uint64_t*)keys, n, (uint32_t*)keep);
  AMJ_HIP_TRY(ctx, hipMemsetAsync((uint32_t*)keep + n, 0, 4, s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u32((uint32_t*)keep, (uint32_t*)pos, n + 1, (uint32_t*)stmp, s));
  uint32_t m = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&m, (uint32_t*)pos + n, 4, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  AMJ_TRY(dmalloc(ctx, (void**)&out->d1, (uint64_t)m * 4));
  out->n1 = m;
  hipLaunchKernelGGL(trie_compact_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     (const uint64_t*)keys, n, (const uint32_t*)pos, (const uint32_t*)keep,
                     out->d1, (uint32_t*)head);
  AMJ_HIP_TRY(ctx, hipMemsetAsync((uint32_t*)head + m, 0, 4, s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u32((uint32_t*)head, (uint32_t*)hpos, (uint64_t)m + 1,
                                      (uint32_t*)stmp, s));
  uint32_t n0 = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&n0, (uint32_t*)hpos + m, 4, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  AMJ_TRY(dmalloc(ctx, (void**)&out->d0, (uint64_t)n0 * 4));
  out->n0 = n0;
  AMJ_TRY(dmalloc(ctx, (void**)&out->off, ((uint64_t)n0 + 1) * 4));
  hipLaunchKernelGGL(trie_level0_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     (const uint64_t*)keys, n, (const uint32_t*)pos, (const uint32_t*)keep,
                     (const uint32_t*)head, (const uint32_t*)hpos, out->d0, out->off);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(out->off + n0, &m, 4, hipMemcpyHostToDevice, s));
  return finish_trie(ctx, out);
}

// Host CSR (e.g. a deserialised .tr file) -> device trie. Level 0 must be strictly increasing
// and every child run sorted (duplicate children are kept, as in the reference's tries).
int trie_from_host(amj_ctx_t* ctx, uint64_t n0, const int32_t* d0, const uint64_t* off,
                   uint64_t n1, const int32_t* d1, TrieDev* out) {
  *out = TrieDev{};
  if (n0 >= (1ull << 32) || n1 >= (1ull << 32)) {
    set_error(ctx, AMJ_ERR_ARG, "trie levels must hold fewer than 2^32 values");
    return AMJ_ERR_ARG;
  }
  if (off[0] != 0 || off[n0] != n1) {
    set_error(ctx, AMJ_ERR_ARG, "trie offsets must start at 0 and end at the child count");
    return AMJ_ERR_ARG;
  }
  std::vector<uint32_t> off32(n0 + 1);
  for (uint64_t j = 0; j < n0; ++j) {
    if (off[j + 1] < off[j] || (j > 0 && d0[j] <= d0[j - 1])) {
      set_error(ctx, AMJ_ERR_ARG, "trie level 0 must be strictly increasing with monotone offsets");
      return AMJ_ERR_ARG;
    }
    for (uint64_t k = off[j] + 1; k < off[j + 1]; ++k)
      if (d1[k] < d1[k - 1]) {
        set_error(ctx, AMJ_ERR_ARG, "trie child lists must be sorted");
        return AMJ_ERR_ARG;
      }
    off32[j] = (uint32_t)off[j];
  }
  off32[n0] = (uint32_t)n1;
  hipStream_t s = ctx->stream;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
  AMJ_TRY(dmalloc(ctx, (void**)&out->d0, n0 * 4));
  out->n0 = (uint32_t)n0;
  AMJ_TRY(dmalloc(ctx, (void**)&out->off, (n0 + 1) * 4));
  AMJ_TRY(dmalloc(ctx, (void**)&out->d1, n1 * 4));
  out->n1 = (uint32_t)n1;
  if (n0) AMJ_HIP_TRY(ctx, hipMemcpyAsync(out->d0, d0, n0 * 4, hipMemcpyHostToDevice, s));
  if (n1) AMJ_HIP_TRY(ctx, hipMemcpyAsync(out->d1, d1, n1 * 4, hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(out->off, off32.data(), (n0 + 1) * 4, hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  return finish_trie(ctx, out);
}

void free_trie(amj_ctx_t* ctx, TrieDev* t) {
  if (!t) return;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  dfree(ctx, t->d0, (uint64_t)t->n0 * 4);
  dfree(ctx, t->off, ((uint64_t)t->n0 + 1) * 4);
  dfree(ctx, t->d1, (uint64_t)t->n1 * 4);
  dfree(ctx, t->vrange, (uint64_t)t->vrange_n * 8);
  *t = TrieDev{};
}

}  // namespace amj
