// Host-side runtime objects of libamj (context, relations, hash tables, device buffers).
#pragma once

#include <hip/hip_runtime.h>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "amj_internal.h"

struct amj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // copy stream for out-of-core streaming
  std::string last_error;
  hipDeviceProp_t props{};
  uint64_t allocated = 0;      // bytes currently allocated by this context
  uint64_t peak = 0;
  hipEvent_t ev[8]{};

  // Reusable scratch buffers keyed by purpose.
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  std::map<int, Buf> scratch;
  // Parked result buffers, reused by later results that fit them (pool_put / alloc_output).
  std::vector<Buf> out_pool;
  uint64_t pool_bytes = 0;
  // Result arena (arena_reserve): one device allocation that every result is carved from,
  // first fit over free ranges kept in address order and merged on release.
  struct Arena {
    char* base = nullptr;
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free;  // offset -> bytes
    uint64_t free_bytes = 0;
  } arena;
  // Autotuned options per query shape (table and probe handles, rows, materialise).
  struct Tuned {
    amj_opts_t o;
    double ms;
  };
  std::map<std::vector<uint64_t>, Tuned> tuned;
};

struct amj_relation {
  uint32_t num_cols = 0;
  uint64_t rows = 0;
  int32_t* cols[AMJ_MAX_PROBE_COLS] = {nullptr, nullptr};
  uint32_t attr_ids[AMJ_MAX_PROBE_COLS] = {0, 0};
  bool owned = false;
  bool host = false;  // cols are host memory (out-of-core probe relation)
};

struct amj_hash_table {
  amj::HashTableDev dev;
  const amj_relation* rel = nullptr;
  uint32_t key_col = 0;
  bool vals_sorted = false;
  double build_ms = 0.0;
};

namespace amj {

enum ScratchSlot {
  SLOT_SORT_KEYS = 1,
  SLOT_SORT_ALT,
  SLOT_SORT_TEMP,
  SLOT_UNIT_COUNTS,
  SLOT_SCAN_TEMP,
  SLOT_BLOCK_SUMS,
  SLOT_SUM,
  SLOT_FLAG,
  SLOT_PROBE_OFF,
  SLOT_PW0,  // SLOT_PW0 + t for level t
  SLOT_PW_END = SLOT_PW0 + kMaxLevels,
  SLOT_NEXT0,  // SLOT_NEXT0 + t: run pointers of table t into table t+1
  SLOT_NEXT_END = SLOT_NEXT0 + kMaxLevels,
  SLOT_PROBE_NEXT,
  SLOT_SFX0,  // SLOT_SFX0 + t: suffix table of build table t
  SLOT_SFX_END = SLOT_SFX0 + kMaxLevels,
  SLOT_FIB_W,
  SLOT_OOC_A,
  SLOT_OOC_B,
  SLOT_MISC0,
  SLOT_MISC1,
  SLOT_MISC2,
  SLOT_MISC3,
  SLOT_TRIE0,
  SLOT_TRIE1,
  SLOT_TRIE2,
  SLOT_TRIE3,
  SLOT_PROBE_SORTED,  // frequent-item path: probe columns grouped by join key
  SLOT_DYN_CTL,    // persistent probe: control block, task records, task flags
  SLOT_DYN_TASKS,
  SLOT_DYN_READY,
  SLOT_TUNE_SAMPLE,  // the autotuner's probe samples (the engines it runs use the others)
  SLOT_LFTJ_SAMPLE,  // the triejoin's output-size sample: window offsets of the sampled roots
};

int dmalloc(amj_ctx_t* ctx, void** p, size_t bytes);
void dfree(amj_ctx_t* ctx, void* p, size_t bytes);
// Park a result buffer for reuse (at most 256 buffers and a third of the device, oldest
// out first).
void pool_put(amj_ctx_t* ctx, void* p, size_t bytes);
void pool_clear(amj_ctx_t* ctx);
// Reserve `bytes` of device memory as the context's result arena (results come only from it
// afterwards; AMJ_ERR_NOMEM when it is full). Once per context.
int arena_reserve(amj_ctx_t* ctx, uint64_t bytes);
// Free an arena that holds no result (AMJ_ERR_STATE otherwise).
int arena_drop(amj_ctx_t* ctx);
void arena_release(amj_ctx_t* ctx, void* p, uint64_t bytes);
// Largest result the arena can hold now (0 without an arena).
uint64_t arena_largest(const amj_ctx_t* ctx);
int scratch(amj_ctx_t* ctx, int slot, size_t bytes, void** out);

// A table as used by one query: the same physical table may be used several times with
// different attribute ids (the TRI and FOUR queries alias one edge table).
struct TableUse {
  const amj_hash_table* ht;
  uint32_t key_attr_id;
  int64_t val_attr_id;  // -1: single-column table
};

struct ProbeUse {
  const amj_relation* rel;
  uint32_t attr_ids[AMJ_MAX_PROBE_COLS];
};

// Resolve attribute numbering and per-level modes; reorders compare-only tables to the
// earliest level at which both of their attributes are bound.
int make_plan(amj_ctx_t* ctx, const ProbeUse& probe, const std::vector<TableUse>& uses,
              Plan* plan, std::vector<uint32_t>* out_attr_ids, std::vector<int>* order);

int build_hash_table(amj_ctx_t* ctx, const amj_relation* rel, uint32_t key_col, int sort_vals,
                     amj_hash_table** out);

int execute_plan(amj_ctx_t* ctx, Plan& plan, const std::vector<uint32_t>& attr_ids,
                 const amj_opts_t* opts, amj_result_t* res);
// Plan of amj_execute: probe relation and hash tables as given.
int make_execute_plan(amj_ctx_t* ctx, const amj_relation* probe, amj_hash_table* const* hts,
                      uint32_t n_hts, Plan* plan, std::vector<uint32_t>* ids);

uint32_t floor_log2_u64(uint64_t v);
double elapsed_ms(hipEvent_t a, hipEvent_t b);

// Helpers shared by the execution drivers (runtime.cpp, fib.cpp).
bool fib_eligible(const Plan& P);
int alloc_output(amj_ctx_t* ctx, uint64_t rows, int n_attrs, amj_result_t* res);
int finish_checksum(amj_ctx_t* ctx, const uint64_t* block_sums, uint64_t nb, uint64_t* out);
uint64_t budget_left(amj_ctx_t* ctx, const amj_opts_t* o);
// Frequent-item path: suffix tables + output-window writer. AMJ_ERR_STATE asks the caller to
// use the count/write path instead (a frontier level would exceed 2^32 children).
int run_fib(amj_ctx_t* ctx, const Plan& plan, const amj_opts_t* o, amj_result_t* res);
// Plan of exec-AMHJ TRI / FOUR over one edge relation and its table (built with sort_vals=1).
int make_pattern_plan(amj_ctx_t* ctx, const amj_relation* edges, const amj_hash_table* ht,
                      int query, Plan* plan, std::vector<uint32_t>* ids);
// Pairwise joins with materialised intermediates (opts.engine = AMJ_ENGINE_PW).
int run_pw(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res);
// Persistent probe with work sharing (o->ws) and/or direct result output (o->dro).
int run_dynamic(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res);
// Out-of-core probe: the probe relation streamed from host memory in slices (o->ooc, or a
// host-resident probe relation).
int run_ooc(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res);

}  // namespace amj
