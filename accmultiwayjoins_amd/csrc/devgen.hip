// Synthetic columns generated on the device, and device buffers for callers.
//
// No counterpart in the reference (its writers are host or CUDA programs that write files, and
// are unseeded or time-seeded: MHJ-GPU/uniform_writer_.cu:114, skewed_writer_.cu:55). These are
// counter-based like the host generators (gen.cpp): value i of stream s under seed x is a pure
// function of (x, s, i) built from SplitMix64, so a column is the same whatever the launch shape.
//   * uniform over [lo, hi): the host generator's formula (amj_gen_uniform), bit for bit;
//   * Zipf(s) over k ranks by rejection-inversion (W. Hormann and G. Derflinger, "Rejection-
//     inversion to generate variates from monotone discrete distributions", ACM TOMACS 6(3),
//     1996): O(1) work per value and no table, so columns of 10^7..10^8 values over as many keys
//     take milliseconds; rank r maps to key ((r - 1) * a + b) mod k with gcd(a, k) = 1, a seeded
//     bijection that scatters the frequent ranks over the key range. The values are not those of
//     the host's table-driven Zipf generator (same law, other draws).
#include <cmath>
#include <utility>
#include <vector>

#include "runtime.h"

namespace amj {
namespace {

__device__ __forceinline__ uint64_t gen_draw(uint64_t base, uint64_t i) {
  return mix64(base + i * 0x9e3779b97f4a7c15ull);
}
__device__ __forceinline__ uint64_t gen_below(uint64_t r, uint64_t range) {
  return __umul64hi(r, range);
}
__device__ __forceinline__ double gen_unit(uint64_t r) {
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(kBlock) void gen_uniform_kernel(int32_t* __restrict__ out, uint64_t n,
                                                             int64_t lo, uint64_t range,
                                                             uint64_t base) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    out[i] = (int32_t)(lo + (int64_t)gen_below(gen_draw(base, i), range));
}

// Rejection-inversion for P(r) proportional to r^-s, r = 1..k (s >= 0). H is the integral of
// h(x) = x^-s from 1 to x and Hinv its inverse, written through log1p/expm1 forms that stay
// exact as s -> 1.
struct ZipfRI {
  double s;
  double h_int_x1;  // H(1.5) - 1
  double h_int_k;   // H(k + 0.5)
  double s_const;   // 2 - Hinv(H(2.5) - h(2)): acceptance without the second test
  uint64_t k;
};

__host__ __device__ inline double zr_helper1(double x) {  // log1p(x) / x
  return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
__host__ __device__ inline double zr_helper2(double x) {  // expm1(x) / x
  return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
__host__ __device__ inline double zr_h(double s, double x) { return exp(-s * log(x)); }
__host__ __device__ inline double zr_H(double s, double x) {
  const double lx = log(x);
  return zr_helper2((1.0 - s) * lx) * lx;
}
__host__ __device__ inline double zr_Hinv(double s, double x) {
  double t = x * (1.0 - s);
  if (t < -1.0) t = -1.0;  // rounding at the lower end of the support
  return exp(zr_helper1(t) * x);
}

__global__ __launch_bounds__(kBlock) void gen_zipf_kernel(int32_t* __restrict__ out, uint64_t n,
                                                          ZipfRI z, uint64_t a, uint64_t b,
                                                          uint64_t base) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    uint64_t rank = 1;
    for (uint32_t attempt = 0; attempt < 64; ++attempt) {  // accepts with probability > 0.5
      const double u01 = gen_unit(gen_draw(base + attempt * 0x632be59bd9b4e019ull, i));
      const double u = z.h_int_k + u01 * (z.h_int_x1 - z.h_int_k);
      const double x = zr_Hinv(z.s, u);
      double kk = floor(x + 0.5);
      if (kk < 1.0) kk = 1.0;
      if (kk > (double)z.k) kk = (double)z.k;
      rank = (uint64_t)kk;
      if (kk - x <= z.s_const || u >= zr_H(z.s, kk + 0.5) - zr_h(z.s, kk)) break;
    }
    out[i] = (int32_t)(((rank - 1) % z.k * a + b) % z.k);
  }
}

// R-MAT edge e: `scale` quadrant choices with probabilities (a, b, c, 1 - a - b - c), both
// endpoints relabelled by a seeded permutation; the undirected edge as min << 32 | max, a
// self-loop as ~0 (dropped after the sort).
__global__ __launch_bounds__(kBlock) void rmat_edge_kernel(uint64_t* __restrict__ e, uint64_t m,
                                                           int scale, double a, double ab,
                                                           double abc, uint64_t base,
                                                           const uint32_t* __restrict__ relabel) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * kBlock) {
    uint64_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      const double r = gen_unit(gen_draw(base, i * (uint64_t)scale + (uint64_t)l));
      const uint64_t bu = r >= ab ? 1 : 0;
      const uint64_t bv = (r >= a && r < ab) || r >= abc ? 1 : 0;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    const uint64_t x = relabel[u], y = relabel[v];
    e[i] = x == y ? ~0ull : ((x < y ? x : y) << 32) | (x < y ? y : x);
  }
}

// flag[i] = 1 for the first copy of each edge (the input is sorted); flag[n] = 0.
__global__ __launch_bounds__(kBlock) void first_copy_kernel(const uint64_t* __restrict__ e,
                                                            uint64_t n, uint32_t* __restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * kBlock)
    flag[i] = i < n && e[i] != ~0ull && (i == 0 || e[i] != e[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void compact_kernel(const uint64_t* __restrict__ e, uint64_t n,
                                                         const uint32_t* __restrict__ flag,
                                                         const uint32_t* __restrict__ pos,
                                                         uint64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    if (flag[i]) out[pos[i]] = e[i];
}

__global__ __launch_bounds__(kBlock) void degree_kernel(const uint64_t* __restrict__ e, uint64_t k,
                                                        uint32_t* __restrict__ deg) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * kBlock) {
    atomicAdd(&deg[(uint32_t)(e[i] >> 32)], 1u);
    atomicAdd(&deg[(uint32_t)e[i]], 1u);
  }
}

// u -> v iff (deg u, u) < (deg v, v)
__global__ __launch_bounds__(kBlock) void orient_kernel(uint64_t* __restrict__ e, uint64_t k,
                                                        const uint32_t* __restrict__ deg) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint32_t u = (uint32_t)(e[i] >> 32), v = (uint32_t)e[i];
    const bool fwd = deg[u] < deg[v] || (deg[u] == deg[v] && u < v);
    e[i] = fwd ? (((uint64_t)u << 32) | v) : (((uint64_t)v << 32) | u);
  }
}

__global__ __launch_bounds__(kBlock) void split_kernel(const uint64_t* __restrict__ e, uint64_t k,
                                                       int32_t* __restrict__ src,
                                                       int32_t* __restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * kBlock) {
    src[i] = (int32_t)(e[i] >> 32);
    dst[i] = (int32_t)(uint32_t)e[i];
  }
}

uint64_t gen_stream_base(uint64_t seed, uint64_t stream) {  // as the host generators
  auto mix = [](uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  return mix(mix(seed) ^ (stream * 0xd1b54a32d192ed03ull));
}

unsigned gen_grid(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  return g ? (unsigned)g : 1u;
}

// Seeded permutation of [0, k): Fisher-Yates driven by the counter stream (as the host
// generator's relabelling, so both give the same graph).
void host_permutation(uint32_t* perm, uint64_t k, uint64_t seed, uint64_t stream) {
  auto mix = [](uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  for (uint64_t i = 0; i < k; ++i) perm[i] = (uint32_t)i;
  const uint64_t base = gen_stream_base(seed, stream);
  for (uint64_t i = k; i > 1; --i) {
    const uint64_t r = mix(base + i * 0x9e3779b97f4a7c15ull);
    const uint64_t j = (uint64_t)(((unsigned __int128)r * i) >> 64);
    std::swap(perm[i - 1], perm[j]);
  }
}

// Device temporaries of one call, freed on every exit.
struct GenBufs {
  amj_ctx_t* ctx;
  std::vector<std::pair<void*, size_t>> held;
  explicit GenBufs(amj_ctx_t* c) : ctx(c) {}
  ~GenBufs() {
    if (!held.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (auto& h : held) dfree(ctx, h.first, h.second);
  }
  int get(size_t bytes, void** p) {
    AMJ_TRY(dmalloc(ctx, p, bytes ? bytes : 8));
    held.emplace_back(*p, bytes ? bytes : 8);
    return AMJ_OK;
  }
};

uint64_t gcd_u64(uint64_t x, uint64_t y) {
  while (y) {
    const uint64_t t = x % y;
    x = y;
    y = t;
  }
  return x;
}

}  // namespace
}  // namespace amj

using namespace amj;

extern "C" {

int amj_device_alloc(amj_ctx_t* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return AMJ_ERR_ARG;
  *out = nullptr;
  return dmalloc(ctx, out, bytes ? bytes : 1);
}

void amj_device_free(amj_ctx_t* ctx, void* p, size_t bytes) {
  if (ctx && p) dfree(ctx, p, bytes ? bytes : 1);
}

int amj_memcpy(amj_ctx_t* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || ((!dst || !src) && bytes)) return AMJ_ERR_ARG;
  if (bytes == 0) return AMJ_OK;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return AMJ_OK;
}

int amj_gen_device_uniform(amj_ctx_t* ctx, int32_t* out, uint64_t n, int64_t lo, int64_t hi,
                           uint64_t seed, uint64_t stream) {
  if (!ctx || (!out && n)) return AMJ_ERR_ARG;
  if (!(lo < hi && lo >= INT32_MIN && hi - 1 <= INT32_MAX)) {
    set_error(ctx, AMJ_ERR_ARG, "uniform column: need lo < hi within the int32 range");
    return AMJ_ERR_ARG;
  }
  if (n == 0) return AMJ_OK;
  hipLaunchKernelGGL(gen_uniform_kernel, dim3(gen_grid(n)), dim3(kBlock), 0, ctx->stream, out, n,
                     lo, (uint64_t)(hi - lo), gen_stream_base(seed, stream));
  AMJ_HIP_TRY(ctx, hipGetLastError());
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return AMJ_OK;
}

int amj_gen_device_zipf(amj_ctx_t* ctx, int32_t* out, uint64_t n, uint64_t k, double s,
                        uint64_t seed, uint64_t stream) {
  if (!ctx || (!out && n)) return AMJ_ERR_ARG;
  if (k == 0 || k > (1ull << 31) || !(s >= 0.0) || s > 16.0) {
    set_error(ctx, AMJ_ERR_ARG, "Zipf column: need 1 <= k <= 2^31 keys and 0 <= s <= 16");
    return AMJ_ERR_ARG;
  }
  if (n == 0) return AMJ_OK;
  ZipfRI z;
  z.s = s;
  z.k = k;
  z.h_int_x1 = zr_H(s, 1.5) - 1.0;
  z.h_int_k = zr_H(s, (double)k + 0.5);
  z.s_const = 2.0 - zr_Hinv(s, zr_H(s, 2.5) - zr_h(s, 2.0));
  // seeded bijection of [0, k): multiplier coprime to k, offset below k
  const uint64_t r = gen_stream_base(seed, stream ^ 0x5a5a5a5aull);
  uint64_t a = k > 2 ? 1 + (r % (k - 1)) : 1;
  while (gcd_u64(a, k) != 1) a = a + 1 < k ? a + 1 : 1;
  const uint64_t b = (r >> 32) % k;
  hipLaunchKernelGGL(gen_zipf_kernel, dim3(gen_grid(n)), dim3(kBlock), 0, ctx->stream, out, n, z,
                     a, b, gen_stream_base(seed, stream));
  AMJ_HIP_TRY(ctx, hipGetLastError());
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return AMJ_OK;
}

int amj_gen_device_rmat(amj_ctx_t* ctx, int scale, uint32_t edge_factor, double a, double b,
                        double c, uint64_t seed, int32_t* out_src, int32_t* out_dst,
                        uint64_t* edges) {
  if (!ctx || !out_src || !out_dst || !edges) return AMJ_ERR_ARG;
  if (scale < 1 || scale > 30 || edge_factor == 0 || edge_factor > 64 || !(a >= 0 && b >= 0 &&
      c >= 0 && a + b + c <= 1.0)) {
    set_error(ctx, AMJ_ERR_ARG, "R-MAT: need 1 <= scale <= 30, 1 <= edge factor <= 64 and "
                                "probabilities a, b, c >= 0 with a + b + c <= 1");
    return AMJ_ERR_ARG;
  }
  const uint64_t nv = 1ull << scale, m = (uint64_t)edge_factor << scale;
  if (m >= (1ull << 32)) {
    set_error(ctx, AMJ_ERR_ARG, "R-MAT: at most 2^32 - 1 generated edges");
    return AMJ_ERR_ARG;
  }
  hipStream_t s = ctx->stream;
  std::vector<uint32_t> relabel(nv);
  host_permutation(relabel.data(), nv, seed, 0x72656c62ull);
  GenBufs B(ctx);
  void *e = nullptr, *alt = nullptr, *flag = nullptr, *pos = nullptr, *deg = nullptr,
       *perm = nullptr, *stmp = nullptr, *ctmp = nullptr;
  AMJ_TRY(B.get(m * 8, &e));
  AMJ_TRY(B.get(m * 8, &alt));
  AMJ_TRY(B.get((m + 1) * 4, &flag));
  AMJ_TRY(B.get((m + 1) * 4, &pos));
  AMJ_TRY(B.get(nv * 4, &deg));
  AMJ_TRY(B.get(nv * 4, &perm));
  AMJ_TRY(B.get(radix_sort_temp_bytes(m), &stmp));
  AMJ_TRY(B.get(scan_temp_elems(m + 1) * 4, &ctmp));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(perm, relabel.data(), nv * 4, hipMemcpyHostToDevice, s));
  const unsigned g = gen_grid(m);
  // 1. edges, undirected; 2. sort, drop self-loops and repeats
  hipLaunchKernelGGL(rmat_edge_kernel, dim3(g), dim3(kBlock), 0, s, (uint64_t*)e, m, scale, a,
                     a + b, a + b + c, gen_stream_base(seed, 0x726d6174ull),
                     (const uint32_t*)perm);
  AMJ_HIP_TRY(ctx, hipGetLastError());
  AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)e, (uint64_t*)alt, m, 0xffu, stmp, s));
  hipLaunchKernelGGL(first_copy_kernel, dim3(gen_grid(m + 1)), dim3(kBlock), 0, s,
                     (const uint64_t*)e, m, (uint32_t*)flag);
  AMJ_HIP_TRY(ctx, hipGetLastError());
  AMJ_HIP_TRY(ctx, exclusive_scan_u32((const uint32_t*)flag, (uint32_t*)pos, m + 1,
                                      (uint32_t*)ctmp, s));
  hipLaunchKernelGGL(compact_kernel, dim3(g), dim3(kBlock), 0, s, (const uint64_t*)e, m,
                     (const uint32_t*)flag, (const uint32_t*)pos, (uint64_t*)alt);
  AMJ_HIP_TRY(ctx, hipGetLastError());
  uint32_t k32 = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&k32, (uint32_t*)pos + m, 4, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  const uint64_t k = k32;
  // 3. degrees of the simple graph, orientation, sort by (src, dst)
  AMJ_HIP_TRY(ctx, hipMemsetAsync(deg, 0, nv * 4, s));
  if (k) {
    hipLaunchKernelGGL(degree_kernel, dim3(gen_grid(k)), dim3(kBlock), 0, s,
                       (const uint64_t*)alt, k, (uint32_t*)deg);
    hipLaunchKernelGGL(orient_kernel, dim3(gen_grid(k)), dim3(kBlock), 0, s, (uint64_t*)alt, k,
                       (const uint32_t*)deg);
    AMJ_HIP_TRY(ctx, hipGetLastError());
    AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)alt, (uint64_t*)e, k, 0xffu, stmp, s));
    hipLaunchKernelGGL(split_kernel, dim3(gen_grid(k)), dim3(kBlock), 0, s,
                       (const uint64_t*)alt, k, out_src, out_dst);
    AMJ_HIP_TRY(ctx, hipGetLastError());
  }
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  *edges = k;
  return AMJ_OK;
}

}  // extern "C"
