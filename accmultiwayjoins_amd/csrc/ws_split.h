// Range arithmetic of work sharing, shared by the persistent AMHJ and ALFTJ engines and by the
// CPU test of the protocol (plain constexpr C++: host and device code compile the same lines).
//
// A wave walking the flattened children [c, c1) of a frontier level with more than split + 64
// left keeps [c, c + split) and shares the rest as k pieces of `split` children (the last one
// shorter). When the queue grants only got <= k slots, the wave publishes the LAST got pieces and
// keeps everything before them, so the kept range and the published pieces tile [c, c1) exactly.
// Reference: the split of long ranges into BUC_THRESHOLD pieces pushed to a TaskBook
// (LFTJ-GPU/ALFTJ.cuh:464-499; MHJ-GPU/Joins/AMHJ_WS.cuh:168-187).
#pragma once

#include <stdint.h>

namespace amj {

// Does a wave at child c of [c, c1) share work?
constexpr bool ws_should_split(uint64_t c, uint64_t c1, uint64_t split) {
  return split != 0 && c1 - c > split + 64;
}

// Pieces the rest [c + split, c1) is cut into (call only when ws_should_split holds).
constexpr uint64_t ws_pieces(uint64_t c, uint64_t c1, uint64_t split) {
  return (c1 - (c + split) + split - 1) / split;
}

// Piece j (0 <= j < got) of the got published ones: [*p0, *p1).
constexpr void ws_piece(uint64_t c, uint64_t c1, uint64_t split, uint64_t k, uint64_t got,
                        uint64_t j, uint64_t* p0, uint64_t* p1) {
  *p0 = c + split + (k - got + j) * split;
  *p1 = *p0 + split < c1 ? *p0 + split : c1;
}

// End of the range the wave keeps after publishing got of the k pieces (all of it when none).
constexpr uint64_t ws_keep_end(uint64_t c, uint64_t c1, uint64_t split, uint64_t k,
                               uint64_t got) {
  return got ? c + split + (k - got) * split : c1;
}

// Slots granted by one fetch-add on the queue tail: tail was the old value, cap the capacity.
constexpr uint64_t ws_granted(uint64_t tail, uint64_t want, uint64_t cap) {
  return tail >= cap ? 0 : (want < cap - tail ? want : cap - tail);
}

}  // namespace amj
