// Leapfrog triejoin (ALFTJ) declarations: CSR tries and the query plan over them.
#pragma once

#include "amj_internal.h"

namespace amj {

constexpr int kMaxLftjAttrs = 4;  // triangle (3) and 4-clique (4)
constexpr int kMaxLftjRels = 6;   // the 4-clique query joins six relations

// Two-level sorted trie of a relation R(x, y) (IndexedTrie<D,C> with 2 attributes,
// LFTJ-GPU/IndexedTrie.cuh:21-54): d0 = distinct x ascending, off[j] .. off[j+1] the children of
// d0[j] in d1 (ascending, distinct).
struct LRel {
  const int32_t* d0;
  const uint32_t* off;  // [n0 + 1]
  const int32_t* d1;
  const uint2* vrange;  // optional: vrange[v] = child range [x, y) of value v (empty: absent)
  uint32_t n0;
  uint32_t vrange_n;
};

struct TrieDev {
  LRel view{};
  int32_t* d0 = nullptr;
  uint32_t* off = nullptr;
  int32_t* d1 = nullptr;
  uint2* vrange = nullptr;
  uint32_t n0 = 0, n1 = 0, vrange_n = 0;
  uint32_t max_run = 0;  // longest child list
  double build_ms = 0.0;
};

// Query: attribute order 0..n_attrs-1 (the reference's attr_order, exec_LFTJ.cu:230-231);
// relation r joins attributes x[r] < y[r].
struct LPlan {
  int32_t n_attrs;
  int32_t n_rels;
  LRel rel[kMaxLftjRels];
  int32_t x[kMaxLftjRels];
  int32_t y[kMaxLftjRels];
  uint32_t root_mask[kMaxLftjAttrs];   // relations whose first attribute is i
  uint32_t child_mask[kMaxLftjAttrs];  // relations whose second attribute is i
  int32_t driver0;                     // root relation whose level 0 enumerates attribute 0
  int32_t off32;                       // every trie array is addressable with 32-bit byte offsets
  uint64_t roots;                      // level-0 size of driver0
  uint32_t rpu;                        // roots per unit of the DFS (1..64; 0: 64)
  uint32_t pad;
};

// Query shapes with kernels of their own: relations listed by (x, y) ascending (the order the
// plan puts them in). The code has, for attribute i, bit 12i + r set when relation r is rooted
// at i and bit 12i + 6 + r when i is its child attribute.
constexpr uint64_t lftj_shape_code(const int (*xy)[2], int n) {
  uint64_t c = 0;
  for (int r = 0; r < n; ++r) {
    c |= 1ull << (12 * xy[r][0] + r);
    c |= 1ull << (12 * xy[r][1] + 6 + r);
  }
  return c;
}
constexpr int kTriXY[3][2] = {{0, 1}, {0, 2}, {1, 2}};
constexpr int kFourXY[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
constexpr uint64_t kShapeTri = lftj_shape_code(kTriXY, 3);    // triangle
constexpr uint64_t kShapeFour = lftj_shape_code(kFourXY, 6);  // 4-clique
// The plan's compiled shape (kShapeTri, kShapeFour), or 0: kernels read the masks at run time.
uint64_t lftj_static_shape(const LPlan& P);

int build_trie(amj_ctx_t* ctx, const int32_t* c0, const int32_t* c1, uint64_t n, TrieDev* out);
int trie_from_host(amj_ctx_t* ctx, uint64_t n0, const int32_t* d0, const uint64_t* off,
                   uint64_t n1, const int32_t* d1, TrieDev* out);
void free_trie(amj_ctx_t* ctx, TrieDev* t);

// DFS engine (lftj.hip): per-unit counts of P.rpu roots, then the write pass at scanned bases.
// `order` (nullptr: units in number order) is a launch order from launch_lftj_unit_weights
// sorted ascending: wave slot k runs the unit in the low 32 bits of order[k].
uint64_t lftj_units(const LPlan& P);
// keys[unit] = (~weight << 32) | unit for every unit (see lftj_unit_weight_kernel).
hipError_t launch_lftj_unit_weights(const LPlan& P, uint64_t* keys, hipStream_t s);
hipError_t launch_lftj_count(const LPlan& P, uint64_t* unit_counts, const uint64_t* order,
                             hipStream_t s);
struct LOut {
  int32_t* c[kMaxLftjAttrs];  // result columns (c[0] == nullptr: checksum only)
};
// With unit_base == nullptr (no output) it is the only pass: per-block result counts go to
// block_counts.
hipError_t launch_lftj_write(const LPlan& P, const uint64_t* unit_base, const LOut& cols,
                             bool hash, uint64_t* block_sums, uint64_t* n_blocks,
                             const uint64_t* order, hipStream_t s,
                             uint64_t* block_counts = nullptr);
// Output-size sample over the roots k * stride + offset, k < n, in windows of
// kLftjSampleWindow candidates for attribute 1: windows[k] = the windows of root k for k < n
// and windows[n] = 0 (n + 1 entries); then, with woff their exclusive scan (woff[n] = nw),
// counts[w] = the results of window w (nw entries).
constexpr uint32_t kLftjSampleWindow = 64;
hipError_t launch_lftj_sample_widths(const LPlan& P, uint64_t* windows, uint64_t n,
                                     uint64_t stride, uint64_t offset, hipStream_t s);
hipError_t launch_lftj_sample_windows(const LPlan& P, const uint64_t* woff, uint64_t n,
                                      uint64_t nw, uint64_t stride, uint64_t offset,
                                      uint64_t* counts, hipStream_t s);
// Direct output: one pass, rows reserved `chunk` at a time (a multiple of 256) with atomics
// into *rows (zeroed by the caller); rows at or past `capacity` are dropped, so
// *rows > capacity means the pass must run again. tails[capacity / chunk] (zeroed by the
// caller) receives the unused tail of every wave's last chunk, block_counts the per-block result
// counts; when *rows <= capacity, launch_chunk_fill (with the exclusive scan of the tails of the
// *rows / chunk reserved chunks) makes the first n = sum(block_counts) rows the results.
hipError_t launch_lftj_direct(const LPlan& P, const LOut& cols, uint64_t capacity,
                              uint32_t chunk, bool hash, unsigned long long* rows,
                              uint64_t* tails, uint64_t* block_sums, uint64_t* block_counts,
                              uint64_t* n_blocks, const uint64_t* order, hipStream_t s);

// Work sharing (lftj_ws.hip): persistent waves over the units (taken in `order`, nullptr: number
// order) that split long candidate runs into tasks. `args` is a WsArgs (ws_queue.h) whose task
// buffer holds cap records of lftj_ws_task_bytes(). mode: MODE_COUNT / MODE_HASH count (and
// hash) into the control block; MODE_WRITE / MODE_WRITE_HASH write with direct output, rows
// reserved `chunk` at a time (0: one reservation per 256-row buffer) with the unused tail of
// every wave's last chunk in tails[] (zeroed by the caller), as launch_lftj_direct; the control
// block's count is then the result count and out_rows the reserved rows.
uint64_t lftj_ws_task_bytes();
int lftj_ws_blocks_per_cu(const LPlan& P, int mode);
hipError_t launch_lftj_ws(const LPlan& P, const LOut& cols, int mode, const void* args,
                          const uint64_t* order, unsigned grid, hipStream_t s,
                          uint32_t chunk = 0, uint64_t* tails = nullptr);

// Level-0 arrays of a relation readable from the host.
struct HostView {
  const int32_t* d0;
  const uint32_t* off;
  uint32_t n0;
};
// Out-of-core evaluation (lftj_ooc.hip): per combination of level-0 value ranges, count and
// checksum summed; parts_hint ranges per attribute (0: from the budget).
int run_lftj_ooc(amj_ctx_t* ctx, const LPlan& P, const HostView* hv, const amj_opts_t* o,
                 uint64_t parts_hint, amj_result_t* res, uint64_t* combos_out);

// BFS engine (lftj_bfs.hip): count and checksum, level by level within mem_budget bytes.
int run_lftj_bfs(amj_ctx_t* ctx, const LPlan& P, uint64_t mem_budget, uint64_t* count,
                 uint64_t* checksum);

}  // namespace amj

// A built trie as handed out by the C ABI: the device trie plus the attribute ids of its
// two levels (IndexedTrie::attr_list, LFTJ-GPU/IndexedTrie.cuh:27).
struct amj_trie {
  amj::TrieDev dev;               // device arrays; after alftj_trie_offload the view points to
  uint32_t attr_ids[2] = {0, 1};  // the pinned host copies below and the device ones are freed
  bool host = false;
  int32_t* h_d0 = nullptr;
  uint32_t* h_off = nullptr;
  int32_t* h_d1 = nullptr;
};
