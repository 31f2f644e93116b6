// Deterministic synthetic workload generators (host, OpenMP).
//
// The reference's generators are unseeded or time-seeded (MHJ-GPU/uniform_writer_.cu:114,
// MHJ-GPU/skewed_writer_.cu:55, MHJ-GPU/data_generator.cuh:113), so their data cannot be
// reproduced. These are counter-based instead: value i of stream s under seed x is a pure
// function of (x, s, i) built from SplitMix64, so the output does not depend on the number
// of threads or on the order in which rows are produced.
#include <omp.h>

#include <algorithm>
#include <climits>
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


// Rows [0, n) in a seeded random order: row i is keyed by (32 random bits, i) and the keys are
// sorted, so the order is a pure function of (seed, stream). n < 2^32.
std::vector<uint32_t> shuffled_rows(uint64_t n, uint64_t seed, uint64_t stream) {
  std::vector<uint64_t> k(n);
  const uint64_t base = stream_base(seed, stream);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    k[i] = (draw(base, (uint64_t)i) & 0xffffffff00000000ull) | (uint64_t)i;
  radix_sort_host(k.data(), n, 64);
  std::vector<uint32_t> r(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) r[i] = (uint32_t)k[i];
  return r;
}

int bits_for(uint64_t x) {  // bits needed for values < x
  int b = 0;
  while (b < 64 && (x - 1) >> b) ++b;
  return b;
}

// n distinct values of [0, p), sorted, chosen uniformly under (seed, stream). Dense spaces
// (p <= 4n) take the first n of a seeded shuffle of [0, p); sparse ones draw the missing
// number of values in rounds until n distinct remain. Returns false if n > p or the rounds
// do not finish.
bool distinct_sample(uint64_t n, uint64_t p, uint64_t seed, uint64_t stream,
                     std::vector<uint64_t>& out) {
  out.clear();
  if (n > p) return false;
  if (n == 0) return true;
  if (p <= 4 * n) {
    if (p > 0xffffffffull) return false;
    const std::vector<uint32_t> r = shuffled_rows(p, seed, stream);
    out.assign(r.begin(), r.begin() + n);
    radix_sort_host(out.data(), n, bits_for(p));
    return true;
  }
  out.reserve(n + n / 2);
  for (uint64_t round = 0; out.size() < n; ++round) {
    if (round == 64) return false;
    const uint64_t have = out.size(), d = n - have;
    const uint64_t base = stream_base(seed, stream ^ (round * 0x632be59bd9b4e019ull));
    out.resize(have + d);
#pragma omp parallel for schedule(static) if (d > 65536)
    for (int64_t i = 0; i < (int64_t)d; ++i) out[have + i] = below(draw(base, (uint64_t)i), p);
    radix_sort_host(out.data(), out.size(), bits_for(p));
    out.resize(unique_sorted(out.data(), out.size()));
  }
  return true;
}

// The same for one small group, single-threaded (many groups run in parallel).
bool distinct_sample_serial(uint64_t n, uint64_t p, uint64_t seed, uint64_t stream,
                            std::vector<uint64_t>& out) {
  out.clear();
  if (n > p) return false;
  if (n == 0) return true;
  if (p <= 4 * n) {
    std::vector<uint64_t> k(p);
    const uint64_t base = stream_base(seed, stream);
    for (uint64_t i = 0; i < p; ++i) k[i] = (draw(base, i) & 0xffffffff00000000ull) | i;
    std::sort(k.begin(), k.end());
    out.resize(n);
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)k[i];
    std::sort(out.begin(), out.end());
    return true;
  }
  for (uint64_t round = 0; out.size() < n; ++round) {
    if (round == 64) return false;
    const uint64_t have = out.size(), d = n - have;
    const uint64_t base = stream_base(seed, stream ^ (round * 0x632be59bd9b4e019ull));
    for (uint64_t i = 0; i < d; ++i) out.push_back(below(draw(base, i), p));
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }
  return true;
}

// Zipf(s) ranks 1..k mapped to values: value of rank r is perm[r-1] (shuffle) or r-1.
int zipf_values(int32_t* out, uint64_t n, uint64_t k, double s, uint64_t seed, uint64_t stream,
                bool shuffle) {
  if (k == 0 || k > (1ull << 31)) return -1;
  std::vector<uint32_t> perm(k);
  if (shuffle) permutation(perm.data(), k, seed, stream ^ 0x5a5a5a5aull);
  else for (uint64_t i = 0; i < k; ++i) perm[i] = (uint32_t)i;
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

bool valid_range(int64_t lo, int64_t hi) {
  return lo < hi && lo >= INT32_MIN && hi - 1 <= INT32_MAX;
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
  return zipf_values(out, n, k, s, seed, stream, true);
}

int amj_gen_zipf_mapped(int32_t* out, uint64_t n, uint64_t k, double s, uint64_t seed,
                        uint64_t stream, int shuffle) {
  return zipf_values(out, n, k, s, seed, stream, shuffle != 0);
}

int amj_gen_unique_column(int32_t* out, uint64_t n, uint64_t range, uint64_t seed,
                          uint64_t stream) {
  if (range == 0 || range > (1ull << 31) || n > 0xffffffffull) return -1;
  const std::vector<uint32_t> r = shuffled_rows(n, seed, stream);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = (int32_t)(r[i] % range);
  return 0;
}

int amj_gen_unique_pairs(int32_t* key0, int32_t* key1, uint64_t n, int64_t lo0, int64_t hi0,
                         int64_t lo1, int64_t hi1, uint64_t seed, int shuffle) {
  if (!valid_range(lo0, hi0) || !valid_range(lo1, hi1) || n > 0xffffffffull) return -1;
  const uint64_t r0 = (uint64_t)(hi0 - lo0), r1 = (uint64_t)(hi1 - lo1);
  const unsigned __int128 p128 = (unsigned __int128)r0 * r1;
  if ((unsigned __int128)n >= p128) return -1;  // fewer pairs than requested (or exactly as many)
  std::vector<uint64_t> idx;
  if (!distinct_sample(n, (uint64_t)p128, seed, 0x70616972ull, idx)) return -2;
  std::vector<uint32_t> order;
  if (shuffle) order = shuffled_rows(n, seed, 0x73687566ull);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const uint64_t x = idx[shuffle ? order[i] : (uint64_t)i];
    key0[i] = (int32_t)(lo0 + (int64_t)(x / r1));
    key1[i] = (int32_t)(lo1 + (int64_t)(x % r1));
  }
  return 0;
}

int amj_gen_zipf_pairs(int32_t* key0, int32_t* key1, uint64_t n, int64_t lo0, int64_t hi0,
                       int64_t lo1, int64_t hi1, uint64_t seed, int skew_col, double z,
                       int shuffle) {
  if (!valid_range(lo0, hi0) || !valid_range(lo1, hi1) || n > 0xffffffffull ||
      (skew_col != 0 && skew_col != 1) || z < 0)
    return -1;
  const int64_t slo = skew_col == 0 ? lo0 : lo1, shi = skew_col == 0 ? hi0 : hi1;
  const int64_t ulo = skew_col == 0 ? lo1 : lo0, uhi = skew_col == 0 ? hi1 : hi0;
  const uint64_t rs = (uint64_t)(shi - slo), ru = (uint64_t)(uhi - ulo);
  if ((unsigned __int128)n >= (unsigned __int128)rs * ru) return -1;
  // 1. the skewed column, sorted: groups of equal values
  std::vector<int32_t> sk(n);
  if (zipf_values(sk.data(), n, rs, z, seed, 0x7a697066ull, true) != 0) return -1;
  {
    std::vector<uint64_t> k(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) k[i] = (uint32_t)sk[i];
    radix_sort_host(k.data(), n, bits_for(rs));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) sk[i] = (int32_t)k[i];
  }
  std::vector<uint64_t> starts;
  for (uint64_t i = 0; i < n; ++i)
    if (i == 0 || sk[i] != sk[i - 1]) starts.push_back(i);
  starts.push_back(n);
  const int64_t n_groups = (int64_t)starts.size() - 1;
  // 2. every group of c equal skewed values gets c distinct uniform values, sorted
  std::vector<int32_t> un(n);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (int64_t g = 0; g < n_groups; ++g) {
    const uint64_t b = starts[g], c = starts[g + 1] - b;
    std::vector<uint64_t> v;
    if (!distinct_sample_serial(c, ru, seed, 0x756e6966ull ^ ((uint64_t)(uint32_t)sk[b] << 20), v)) {
      bad = 1;
      continue;
    }
    for (uint64_t j = 0; j < c; ++j) un[b + j] = (int32_t)(ulo + (int64_t)v[j]);
  }
  if (bad) return -2;
  std::vector<uint32_t> order;
  if (shuffle) order = shuffled_rows(n, seed, 0x73687566ull);
  int32_t* scol = skew_col == 0 ? key0 : key1;
  int32_t* ucol = skew_col == 0 ? key1 : key0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const uint64_t r = shuffle ? order[i] : (uint64_t)i;
    scol[i] = (int32_t)(slo + sk[r]);
    ucol[i] = un[r];
  }
  return 0;
}

int amj_gen_seq_pairs(int32_t* key0, int32_t* key1, uint64_t n, int64_t lo0, int64_t hi0) {
  if (!valid_range(lo0, hi0)) return -1;
  const uint64_t per = n / (uint64_t)(hi0 - lo0);  // values of key1 per key0 value
  if (per == 0) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    key0[i] = (int32_t)(lo0 + (int64_t)((uint64_t)i / per));
    key1[i] = (int32_t)((uint64_t)i % per);
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
