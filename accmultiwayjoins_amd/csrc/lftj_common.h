// Per-candidate logic of the leapfrog triejoin, shared by the DFS (lftj.hip) and BFS
// (lftj_bfs.hip) engines: list membership, level-0 lookups, choice of the candidate list.
#pragma once

#include "lftj.h"
#include "probe_common.h"

namespace amj {

constexpr int kR = kMaxLftjRels;

template <int A>
struct LLv {
  int32_t b[A];        // bindings of attributes 0..level-1
  uint32_t lo[kR];     // child range [lo, hi) in d1 of relation r (x_r < level <= y_r)
  uint32_t hi[kR];
  const int32_t* dptr;  // candidate list of this entry for the next attribute
  uint32_t dlen;        // its length (0: no candidates, e.g. a failed entry)
  int32_t drel;         // relation providing the candidates
  uint32_t excl;        // exclusive prefix of dlen over the wave
  uint32_t T;           // level total (uniform)
};

struct LAcc {
  uint64_t count;     // results (uniform)
  uint64_t out_base;  // absolute row of the next result (uniform)
  uint64_t csum;      // per-lane checksum partial
};

// Query shape fixed at compile time, or 0: read from the plan at run time. For attribute i,
// bits [12i, 12i + 6) hold the relations rooted at i and bits [12i + 6, 12i + 12) the relations
// whose child attribute is i. A kernel of a fixed shape drops every test and range of the
// relations that do not take part at a level, and addresses the trie arrays with 32-bit element
// offsets (the host runs it only when every array holds fewer than 2^30 entries).
template <uint64_t SH>
__device__ __forceinline__ uint32_t rmask(const LPlan& P, int i) {
  if constexpr (SH != 0) return (uint32_t)(SH >> (12 * i)) & 63u;
  else return P.root_mask[i];
}
template <uint64_t SH>
__device__ __forceinline__ uint32_t cmask(const LPlan& P, int i) {
  if constexpr (SH != 0) return (uint32_t)(SH >> (12 * i + 6)) & 63u;
  else return P.child_mask[i];
}

// base[idx] from global memory (device or pinned host buffers; never LDS), so the load is a
// global load rather than a flat one, which would also wait for the wave's LDS operations.
// With OFF32 the byte offset is a 32-bit value added to the base.
template <bool OFF32, typename T>
__device__ __forceinline__ T ld_at(const T* base, uint32_t idx) {
  typedef __attribute__((address_space(1))) const T gT;
  const gT* g = (const gT*)base;
  if constexpr (OFF32)
    return *(const gT*)((__attribute__((address_space(1))) const char*)g +
                        (uint32_t)(idx * (uint32_t)sizeof(T)));
  else
    return g[idx];
}

// Child range [lo, hi) of value v in rel (false: v is not a level-0 value). With the value-
// indexed range table one 8-byte load answers; otherwise a binary search of d0, then off.
template <bool OFF32>
__device__ __forceinline__ bool root_range(const LRel& rel, int32_t v, uint32_t& lo,
                                           uint32_t& hi) {
  if (rel.vrange) {
    const bool in = (uint32_t)v < rel.vrange_n;
    // x (the range start) is the low word of the 8-byte entry
    const uint64_t r = in ? ld_at<OFF32>((const uint64_t*)rel.vrange, (uint32_t)v) : 0ull;
    lo = (uint32_t)r;
    hi = (uint32_t)(r >> 32);
    return hi > lo;
  }
  uint32_t b = 0, n = rel.n0;
  while (n > 1) {
    const uint32_t half = n >> 1;
    b = ld_at<OFF32>(rel.d0, b + half) <= v ? b + half : b;
    n -= half;
  }
  const bool f = rel.n0 != 0 && ld_at<OFF32>(rel.d0, b) == v;
  lo = f ? ld_at<OFF32>(rel.off, b) : 0u;
  hi = f ? ld_at<OFF32>(rel.off, b + 1) : 0u;
  return f;
}

// Is v in the ascending run d1[lo, end)? Halving without branches: each step keeps the half
// whose first element is <= v, so the loop runs ceil(log2(end - lo)) times, then one compare.
template <bool OFF32>
__device__ __forceinline__ bool child_has(const int32_t* d1, uint32_t lo, uint32_t end,
                                          int32_t v) {
  uint32_t n = end - lo;
  while (n > 1) {
    const uint32_t half = n >> 1;
    lo = ld_at<OFF32>(d1, lo + half) <= v ? lo + half : lo;
    n -= half;
  }
  return n != 0 && ld_at<OFF32>(d1, lo) == v;
}

template <int A>
__device__ __forceinline__ void finish(LLv<A>& l) {
  const uint32_t incl = wave_scan_add_dpp(l.dlen);
  l.excl = incl - l.dlen;
  l.T = lane63(incl);
}

// Candidate list for attribute i of an entry: the shortest child range among the relations
// whose second attribute is i; a level-0 list when no relation reaches i from a bound attribute.
template <int A, uint64_t SH = 0>
__device__ __forceinline__ void choose_driver(const LPlan& P, int i, bool ok, LLv<A>& l) {
  uint32_t best = 0xffffffffu;
  int br = -1;
  const int32_t* bp = nullptr;
  const uint32_t cm = cmask<SH>(P, i);
  sfor<0, kR>([&](auto r) {
    if ((cm >> r) & 1u) {
      const uint32_t len = l.hi[r] - l.lo[r];
      if (len < best) {
        best = len;
        br = (int)r;
        bp = P.rel[r].d1 + l.lo[r];
      }
    }
  });
  if (cm == 0u) {
    const uint32_t rm = rmask<SH>(P, i);
    sfor<0, kR>([&](auto r) {
      if (((rm >> r) & 1u) && P.rel[r].n0 < best) {
        best = P.rel[r].n0;
        br = (int)r;
        bp = P.rel[r].d0;
      }
    });
  }
  l.dptr = bp;
  l.drel = br;
  l.dlen = (ok && br >= 0) ? best : 0u;
}

// Check candidate v for attribute i against every participating list except the driver's and
// open the child ranges of the relations rooted at i.
template <int A, uint64_t SH = 0>
__device__ __forceinline__ bool check_and_open(const LPlan& P, int i, int32_t v, int drel,
                                               bool drel_is_root, LLv<A>& l) {
  constexpr bool kOff32 = SH != 0;
  bool ok = true;
  const uint32_t cm = cmask<SH>(P, i), rm = rmask<SH>(P, i);
  sfor<0, kR>([&](auto r) {
    if (((cm >> r) & 1u) && !(drel == (int)r && !drel_is_root))
      ok = ok && child_has<kOff32>(P.rel[r].d1, l.lo[r], l.hi[r], v);
  });
  sfor<0, kR>([&](auto r) {
    if ((rm >> r) & 1u) {
      uint32_t lo = 0, hi = 0;
      const bool f = ok && root_range<kOff32>(P.rel[r], v, lo, hi);
      ok = ok && f;
      l.lo[r] = f ? lo : 0u;
      l.hi[r] = f ? hi : 0u;
    }
  });
  return ok;
}

constexpr int bit_count(uint32_t m) {
  int c = 0;
  for (int r = 0; r < 32; ++r) c += (int)((m >> r) & 1u);
  return c;
}
// Index of the k-th set bit of m (-1: none).
constexpr int nth_bit(uint32_t m, int k) {
  for (int r = 0; r < 32; ++r)
    if ((m >> r) & 1u) {
      if (k == 0) return r;
      --k;
    }
  return -1;
}

// check_and_open for a compiled shape at attribute I. The candidate came from the child list of
// relation drel (one of I's child relations; I = 0 has none). Every other child list is searched
// once: search k takes the k-th child relation left after removing drel, chosen with selects,
// so lanes whose candidates come from different lists run the same loop.
template <int A, uint64_t SH, int I>
__device__ __forceinline__ bool check_and_open_fixed(const LPlan& P, int32_t v, int drel,
                                                     LLv<A>& l) {
  constexpr uint32_t cm = (uint32_t)(SH >> (12 * I + 6)) & 63u;
  constexpr uint32_t rm = (uint32_t)(SH >> (12 * I)) & 63u;
  constexpr int nc = bit_count(cm);
  bool ok = true;
  if constexpr (nc > 1) {
    int di = 0;  // position of the driver among the child relations
    sfor<1, nc>([&](auto j) {
      if (drel == nth_bit(cm, (int)j)) di = (int)j;
    });
    sfor<0, nc - 1>([&](auto k) {
      constexpr int ra = nth_bit(cm, (int)k), rb = nth_bit(cm, (int)k + 1);
      const bool next = (int)k >= di;
      const int32_t* d1 = next ? P.rel[rb].d1 : P.rel[ra].d1;
      const uint32_t lo = next ? l.lo[rb] : l.lo[ra];
      const uint32_t hi = next ? l.hi[rb] : l.hi[ra];
      ok = ok && child_has<false>(d1, lo, hi, v);
    });
  }
  sfor<0, kR>([&](auto r) {
    if constexpr (((rm >> r) & 1u) != 0u) {
      uint32_t lo = 0, hi = 0;
      const bool f = ok && root_range<true>(P.rel[r], v, lo, hi);
      ok = ok && f;
      l.lo[r] = f ? lo : 0u;
      l.hi[r] = f ? hi : 0u;
    }
  });
  return ok;
}

}  // namespace amj
