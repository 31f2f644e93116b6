// Lane-per-probe-row engines (tpt.hip): the thread-per-tuple baseline over the bucketed tables,
// and its open-addressing variant over per-query linear-probing tables.
#pragma once

#include "amj_internal.h"

namespace amj {

// Linear-probing tables of one query, per level: slot = key << 32 | row (row into the level's
// bucketed table, whose vals give the second column), ~0 when empty; mask = slots - 1.
struct OaTables {
  const uint64_t* slots[kMaxLevels];
  uint32_t mask[kMaxLevels];
};

// Fill a table of mask + 1 slots (all ~0 beforehand) with rows [0, n) of `keys`.
hipError_t launch_oa_insert(const int32_t* keys, uint64_t n, uint64_t* slots, uint32_t mask,
                            hipStream_t s);

}  // namespace amj
