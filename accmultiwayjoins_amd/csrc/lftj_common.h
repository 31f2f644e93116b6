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

// Is v a level-0 value of rel; its row when it is.
__device__ __forceinline__ bool root_row(const LRel& rel, int32_t v, uint32_t& row) {
  if (rel.dense) {
    if ((uint32_t)v >= rel.dense_n) return false;
    row = rel.dense[(uint32_t)v];
    return row != 0xffffffffu;
  }
  uint32_t lo = 0, hi = rel.n0;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (rel.d0[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  row = lo;
  return lo < rel.n0 && rel.d0[lo] == v;
}

__device__ __forceinline__ bool child_has(const int32_t* d1, uint32_t lo, uint32_t end,
                                          int32_t v) {
  uint32_t hi = end;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (d1[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo < end && d1[lo] == v;
}

template <int A>
__device__ __forceinline__ void finish(LLv<A>& l) {
  const uint32_t incl = wave_scan_add_dpp(l.dlen);
  l.excl = incl - l.dlen;
  l.T = lane63(incl);
}

// Candidate list for attribute i of an entry: the shortest child range among the relations
// whose second attribute is i; a level-0 list when no relation reaches i from a bound attribute.
template <int A>
__device__ __forceinline__ void choose_driver(const LPlan& P, int i, bool ok, LLv<A>& l) {
  uint32_t best = 0xffffffffu;
  int br = -1;
  const int32_t* bp = nullptr;
  const uint32_t cm = P.child_mask[i];
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
    const uint32_t rm = P.root_mask[i];
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
template <int A>
__device__ __forceinline__ bool check_and_open(const LPlan& P, int i, int32_t v, int drel,
                                               bool drel_is_root, LLv<A>& l) {
  bool ok = true;
  const uint32_t cm = P.child_mask[i], rm = P.root_mask[i];
  sfor<0, kR>([&](auto r) {
    if (((cm >> r) & 1u) && !(drel == (int)r && !drel_is_root))
      ok = ok && child_has(P.rel[r].d1, l.lo[r], l.hi[r], v);
  });
  sfor<0, kR>([&](auto r) {
    if ((rm >> r) & 1u) {
      uint32_t row = 0;
      const bool f = ok && root_row(P.rel[r], v, row);
      ok = ok && f;
      l.lo[r] = f ? P.rel[r].off[row] : 0u;
      l.hi[r] = f ? P.rel[r].off[row + 1] : 0u;
    }
  });
  return ok;
}

}  // namespace amj
