// Deterministic synthetic workload generators (host, OpenMP).
//
// The reference's generators are unseeded or time-seeded (MHJ-GPU/uniform_writer_.cu:114,
// MHJ-GPU/skewed_writer_.cu:55, MHJ-GPU/data_generator.cuh:113), so their data cannot be
// reproduced. These are counter-based instead: value i of stream s under seed x is a pure
// function of (x, s, i) built from SplitMix64, so the output does not depend on the number
// of threads or on the order in which rows are produced.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/amj_gen.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
inline uint64_t stream_base(uint64_t seed, uint64_t stream) {
  return mix64(mix64(seed) ^ (stream * 0xd1b54a32d192ed03ull));
}
inline uint64_t draw(uint64_t base, uint64_t i) { return mix64(base + i * 0x9e3779b97f4a7c15ull); }
inline uint64_t below(uint64_t r, uint64_t range) {
  return (uint64_t)(((unsigned __int128)r * range) >> 64);
}
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Seeded permutation of [0, k) (Fisher-Yates driven by the counter stream).
void permutation(uint32_t* perm, uint64_t k, uint64_t seed, uint64_t stream) {
  for (uint64_t i = 0; i < k; ++i) perm[i] = (uint32_t)i;
  const uint64_t base = stream_base(seed, stream);
  for (uint64_t i = k; i > 1; --i) {
    const uint64_t j = below(draw(base, i), i);
    std::swap(perm[i - 1], perm[j]);
  }
}

// LSD radix sort of u64 keys with per-thread 16-bit digit histograms.
void radix_sort_host(uint64_t* a, uint64_t n, int key_bits) {
  std::vector<uint64_t> tmp(n);
  uint64_t* src = a;
  uint64_t* dst = tmp.data();
  const int T = omp_get_max_threads();
  std::vector<uint64_t> hist((size_t)T * 65536);
  for (int shift = 0; shift < key_bits; shift += 16) {
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      const uint64_t lo = n * t / T, hi = n * (t + 1) / T;
      uint64_t* h = &hist[(size_t)t * 65536];
      for (uint64_t i = lo; i < hi; ++i) ++h[(src[i] >> shift) & 0xffff];
#pragma omp barrier
#pragma omp single
      {
        uint64_t sum = 0;
        for (int d = 0; d < 65536; ++d)
          for (int tt = 0; tt < T; ++tt) {
            const uint64_t c = hist[(size_t)tt * 65536 + d];
            hist[(size_t)tt * 65536 + d] = sum;
            sum += c;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) dst[h[(src[i] >> shift) & 0xffff]++] = src[i];
    }
    std::swap(src, dst);
  }
  if (src != a) std::memcpy(a, src, n * sizeof(uint64_t));
}

uint64_t unique_sorted(uint64_t* a, uint64_t n) {
  if (n == 0) return 0;
  uint64_t w = 1;
  for (uint64_t i = 1; i < n; ++i)
    if (a[i] != a[w - 1]) a[w++] = a[i];
  return w;
}

}  // namespace

extern "C" {

void amj_gen_uniform(int32_t* out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed,
                     uint64_t stream) {
  const uint64_t base = stream_base(seed, stream);
  const uint64_t range = (uint64_t)(hi - lo);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    out[i] = (int32_t)(lo + (int64_t)below(draw(base, (uint64_t)i), range));
}

int amj_gen_zipf(int32_t* out, uint64_t n, uint64_t k, double s, uint64_t seed,
                 uint64_t stream) {
  if (k == 0 || k > (1ull << 31)) return -1;
  std::vector<uint32_t> perm(k);
  permutation(perm.data(), k, seed, stream ^ 0x5a5a5a5aull);
  const uint64_t base = stream_base(seed, stream);
  if (s == 0.0) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
      out[i] = (int32_t)perm[below(draw(base, (uint64_t)i), k)];
    return 0;
  }
  // inverse-CDF sampling over ranks 1..k with P(rank r) proportional to r^-s
  std::vector<double> cdf(k);
  double acc = 0.0;
  for (uint64_t r = 0; r < k; ++r) {
    acc += std::pow((double)(r + 1), -s);
    cdf[r] = acc;
  }
  for (uint64_t r = 0; r < k; ++r) cdf[r] /= acc;
  cdf[k - 1] = 1.0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const double u = unit(draw(base, (uint64_t)i));
    const uint64_t rank = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    out[i] = (int32_t)perm[rank < k ? rank : k - 1];
  }
  return 0;
}

void amj_gen_rmat(int32_t* src, int32_t* dst, int scale, uint64_t m, double a, double b,
                  double c, uint64_t seed) {
  const uint64_t base = stream_base(seed, 0x726d6174ull);
  const uint64_t nv = 1ull << scale;
  std::vector<uint32_t> relabel(nv);
  permutation(relabel.data(), nv, seed, 0x72656c62ull);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m; ++e) {
    uint64_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      const double r = unit(draw(base, (uint64_t)e * (uint64_t)scale + (uint64_t)l));
      const uint64_t bu = r >= ab ? 1 : 0;
      const uint64_t bv = (r >= a && r < ab) || r >= abc ? 1 : 0;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    src[e] = (int32_t)relabel[u];
    dst[e] = (int32_t)relabel[v];
  }
}

uint64_t amj_graph_orient(const int32_t* src, const int32_t* dst, uint64_t m, uint64_t nv,
                          int32_t* out_src, int32_t* out_dst) {
  // 1. symmetrise + drop self-loops + dedupe: undirected edge {u<v} as (u<<32|v)
  std::vector<uint64_t> e(m);
  uint64_t k = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    const uint32_t u = (uint32_t)src[i], v = (uint32_t)dst[i];
    e[i] = u == v ? ~0ull : (((uint64_t)std::min(u, v) << 32) | std::max(u, v));
  }
  int bits = 64;
  radix_sort_host(e.data(), m, bits);
  k = unique_sorted(e.data(), m);
  if (k && e[k - 1] == ~0ull) --k;
  // 2. degrees in the simple undirected graph
  std::vector<uint32_t> deg(nv, 0);
  for (uint64_t i = 0; i < k; ++i) {
    ++deg[e[i] >> 32];
    ++deg[(uint32_t)e[i]];
  }
  // 3. orient u -> v iff (deg u, u) < (deg v, v), then sort by (src, dst)
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)k; ++i) {
    const uint32_t u = (uint32_t)(e[i] >> 32), v = (uint32_t)e[i];
    const bool fwd = deg[u] < deg[v] || (deg[u] == deg[v] && u < v);
    e[i] = fwd ? (((uint64_t)u << 32) | v) : (((uint64_t)v << 32) | u);
  }
  radix_sort_host(e.data(), k, 64);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)k; ++i) {
    out_src[i] = (int32_t)(e[i] >> 32);
    out_dst[i] = (int32_t)(uint32_t)e[i];
  }
  return k;
}

}  // extern "C"
