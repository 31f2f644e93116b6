/*
 * amj.h -- C ABI of libamj, the MI355X-native multi-way join library (AMHJ + ALFTJ).
 *
 * The reference exposes its operators as templated C++ classes compiled into each CLI.
 * Each entry point below replaces one of them with plain pointers, sizes and status codes:
 *
 *   amj_relation_create      Relation<D,C>::init + read_rel_cols_mmap
 *                            (MHJ-GPU/Relation.cuh:9-48, common-utils/file_io.h:90-121)
 *   amj_hash_table_build     QueryProcessor::build_hash_multiway / build_hash_single
 *                            (MHJ-GPU/QueryProcessor.cuh:76-110, :112-169) and
 *                            RadixPartitioner::splitKI (MHJ-GPU/Indexing/radix_partitioning.cuh:576-579)
 *   amj_execute              OOC<D,C,single_join_key,AlgoType>::execute
 *                            (MHJ-GPU/ooc.cuh:77-165), dispatching like exec_AMHJ.cu:163-198
 *   amj_chain_join           exec-AMHJ NORMAL (MHJ-GPU/exec_AMHJ.cu:70-89, :101-116, :163-203)
 *   amj_pattern_join         exec-AMHJ TRI / FOUR (MHJ-GPU/exec_AMHJ.cu:32-69, :118-145)
 *   amj_trie_build           constructSortedTrie<D,C,2> (LFTJ-GPU/multi_partitioning.cuh:129-293)
 *   alftj_run                OOCWrapper::execute_with_{DFS,BFS} -> LFTJ_Base::evaluate
 *                            (LFTJ-GPU/ooc_wrapper.cuh:404-427, LFTJ-GPU/LFTJ_base.h:20)
 *
 * Conventions: every function returns AMJ_OK (0) or a negative amj_status instead of the
 * reference's exit(); amj_last_error() returns the message. The library owns all device
 * memory it allocates; callers own their inputs. A context owns one HIP stream and must not
 * be used from two host threads at once.
 */
#ifndef AMJ_H_
#define AMJ_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMJ_MAX_BUILD_TABLES 7   /* MHJ-GPU/conf.h:16-18 MAX_NUM_BUILDTABLES, at the value the
                                    reference sets for TPC-H Q8 */
#define AMJ_MAX_ATTRS 8          /* MHJ-GPU/conf.h:16-20 MAX_NUM_RES_ATTRS, likewise (Q8) */
/* Row limits: a build table holds fewer than 2^32 rows (bucket pointers and row indexes are
 * 32-bit), and so does a trie level; probe relations and results are counted in 64 bits. */
#define AMJ_MAX_PROBE_COLS 2     /* MHJ-GPU/conf.h:19 MAX_NUM_ATTRS_IN_BUILD_TABLE */

typedef enum {
  AMJ_OK = 0,
  AMJ_ERR_ARG = -1,        /* invalid argument / unsupported query shape */
  AMJ_ERR_HIP = -2,        /* HIP runtime failure */
  AMJ_ERR_NOMEM = -3,      /* device memory or HBM budget exhausted */
  AMJ_ERR_NODEV = -4,      /* no usable GPU */
  AMJ_ERR_STATE = -5       /* object used in the wrong state */
} amj_status;

typedef enum { AMJ_QUERY_NORMAL = 0, AMJ_QUERY_TRI = 1, AMJ_QUERY_FOUR = 2 } amj_query;
typedef enum { ALFTJ_BFS = 0, ALFTJ_DFS = 1 } alftj_algo;   /* exec_LFTJ.cu:256-260 */

typedef struct amj_ctx amj_ctx_t;
typedef struct amj_relation amj_relation_t;
typedef struct amj_hash_table amj_hash_table_t;
typedef struct amj_trie amj_trie_t;
typedef struct amj_batch amj_batch_t;

/* A relation in the reference's column layout (Relation<D,C>, MHJ-GPU/Relation.cuh:9-48):
 * data[attr][row] int32, plus the attribute id of each column (attr_list). */
typedef struct {
  uint32_t num_cols;
  uint64_t rows;
  const int32_t* const* cols;   /* num_cols column pointers */
  const uint32_t* attr_ids;     /* num_cols attribute ids (may be NULL for TRI/FOUR edges) */
  int on_device;                /* 0: host memory, copied to the device at creation;
                                   1: device pointers that stay valid;
                                   AMJ_REL_HOST: host memory that stays there (pinned memory
                                   from amj_host_alloc streams fastest): a probe relation
                                   used out of core, read in slices */
} amj_rel_t;
#define AMJ_REL_HOST 2

/* Query options: the exec-AMHJ switches (exec_AMHJ.cu:216-245) plus this library's knobs. */
typedef struct {
  int ooc;                      /* stream the probe relation in slices (ooc.cuh:166-306) */
  int ws;                       /* work sharing between waves: the persistent AMHJ engine
                                   (long runs of a frontier level shared as tasks), and ALFTJ
                                   DFS (long candidate runs shared as tasks, ALFTJ.cuh:464-499);
                                   BFS and out-of-core ALFTJ do not use it */
  int dro;                      /* direct result output (single pass, no count pass): AMHJ,
                                   ALFTJ DFS and batches; row order then depends on timing */
  int fib;                      /* frequent-item buffer (subtree counts of join values) */
  int materialize;              /* 1: write result columns; 0: count + checksum only */
  int checksum;                 /* 1: compute the result checksum (always on when
                                   materialize = 0) */
  uint64_t hbm_budget_bytes;    /* 0: use what the device has free */
  uint64_t ooc_slice_rows;      /* 0: derive from the budget */
  uint64_t out_chunk_rows;      /* fib: rows per output window, rounded up to a multiple of 256 (0: 4,096) */
  uint64_t dro_capacity_rows;   /* dro: initial output allocation (0: sampled estimate) */
  uint32_t ws_split_threshold;  /* ws: split runs longer than this (0: default) */
  int fib_suffix;               /* fib: 0 size-based suffix tables, 1 none, 2 deepest that fits */
  int probe_order;              /* fib: 0 group probe rows by join key when the probe has at
                                   least 2^20 rows, 1 keep the input order (the unit engine's
                                   output order), 2 always group */
  uint32_t engine;              /* AMJ_ENGINE_WAVE (0): the wave engines above;
                                   AMJ_ENGINE_TPT: thread-per-tuple baseline (MHJ_count /
                                   MHJ_write, MHJ-GPU/Joins/MHJ.cuh:17-321), for comparison */
  uint32_t ws_stall_ms;         /* ws (AMHJ and ALFTJ persistent engines): a launch in which no
                                   wave reports progress for this long stops with
                                   AMJ_ERR_STATE instead of hanging (0: 20,000 ms); working
                                   waves report progress as they go, so a slow pass is not
                                   mistaken for a stuck one */
  uint32_t ws_claim_units;      /* ws: most units (64 probe rows or roots each) a wave claims at
                                   once (0: 64); claims shrink with the remaining units: half
                                   of them per wave and at least 4 units in launches of 64 or
                                   more units per wave, else a quarter and at least 1 */
  uint32_t reserved[2];
} amj_opts_t;
#define AMJ_ENGINE_WAVE 0
#define AMJ_ENGINE_TPT 1
#define AMJ_ENGINE_OPEN_ADDR 2   /* the thread-per-tuple walk over per-query open-addressing
                                    (linear probing) tables: AMHJ_open_addr.cuh with
                                    Indexing/linear.cuh, for comparison; same results */
#define AMJ_ENGINE_PW 3          /* pairwise joins with materialised intermediates, each a
                                    partitioned hash join probed from LDS: PW() of
                                    HashJoinEPFL/hashJoinEPFL.cuh:174, for comparison; chains
                                    that bind one new attribute per table */

/* Result of one join. Columns are device memory owned by the library until
 * amj_result_free(); column a holds attribute a (renumbered in first-occurrence order). */
typedef struct {
  uint64_t count;               /* number of result tuples */
  uint64_t checksum;            /* sum over tuples of the chained SplitMix64 hash, mod 2^64 */
  uint32_t num_cols;
  int32_t* cols[AMJ_MAX_ATTRS];
  uint64_t rows_written;        /* rows materialised (== count when materialize=1) */
  uint32_t attr_ids[AMJ_MAX_ATTRS];  /* original attribute id of each output column */
  double build_ms;              /* hash build / trie build time (device, HIP events) */
  double probe_ms;              /* probe + output time (device, HIP events) */
  double probe_kernel_ms;       /* time of the dominant probe kernel only */
  uint64_t probe_kernel_launches;
  uint32_t path;                /* which engine ran (AMJ_PATH_*) */
  uint32_t reserved[3];         /* library bookkeeping of the result columns */
  uint32_t detail;              /* fib: suffix-table depth; ws: tasks shared between waves */
  uint32_t passes;              /* probe traversals run (dro: 1, or 2 after an overflow) */
} amj_result_t;

#define AMJ_PATH_PLAIN 1        /* count pass + scan + write pass (materialize = 0: one
                                   pass that counts and hashes) */
#define AMJ_PATH_FIB 2          /* subtree-count buffer + output-window writer */
#define AMJ_PATH_WS 3
#define AMJ_PATH_DRO 4
#define AMJ_PATH_LFTJ_DFS 5
#define AMJ_PATH_LFTJ_BFS 6
#define AMJ_PATH_WS_DRO 7
#define AMJ_PATH_BATCH 8        /* many queries per launch (amj_batch_run) */
#define AMJ_PATH_OOC 9          /* probe relation streamed in slices (opts.ooc / AMJ_REL_HOST) */
#define AMJ_PATH_LFTJ_OOC 10    /* triejoin per combination of level-0 value ranges */
#define AMJ_PATH_TPT 11         /* thread-per-tuple baseline (opts.engine = AMJ_ENGINE_TPT) */
#define AMJ_PATH_OPEN_ADDR 12   /* open-addressing tables (opts.engine = AMJ_ENGINE_OPEN_ADDR) */
#define AMJ_PATH_PW 13          /* pairwise partitioned joins (opts.engine = AMJ_ENGINE_PW) */

/* ---- context ------------------------------------------------------------------------ */
int amj_ctx_create(int device, amj_ctx_t** out);
void amj_ctx_destroy(amj_ctx_t* ctx);
const char* amj_last_error(const amj_ctx_t* ctx);
int amj_sync(amj_ctx_t* ctx);
/* Pinned host memory (hipHostMalloc) for relations kept on the host (AMJ_REL_HOST): the
 * out-of-core probe streams it to the device at PCIe rate, overlapped with the probe. */
int amj_host_alloc(size_t bytes, void** out);
/* Device bytes held by the context now and at its peak (the HBM budget counts the former). */
int amj_memory_info(amj_ctx_t* ctx, uint64_t* allocated, uint64_t* peak);
void amj_host_free(void* p);
int amj_device_info(amj_ctx_t* ctx, char* name, size_t name_len, uint64_t* hbm_bytes,
                    int* cu_count);
void amj_default_opts(amj_opts_t* opts);
const char* amj_version(void);

/* Device memory for caller-owned columns (relations created with on_device = 1 from them):
 * counted in amj_memory_info; free with the same size. */
int amj_device_alloc(amj_ctx_t* ctx, size_t bytes, void** out);
void amj_device_free(amj_ctx_t* ctx, void* p, size_t bytes);
/* Copy bytes between any host and device memory, on the context's stream (synchronous). */
int amj_memcpy(amj_ctx_t* ctx, void* dst, const void* src, size_t bytes);
/* Synthetic int32 columns written on the device (no counterpart in the reference, whose
 * generators write files and are not seeded reproducibly). Value i is a pure function of
 * (seed, stream, i). uniform: over [lo, hi), the same values as the host amj_gen_uniform.
 * zipf: Zipf(s) over k keys (0 <= s <= 16, k <= 2^31) by rejection-inversion sampling, rank r
 * mapped to key ((r - 1) * a + b) mod k by a seeded bijection. Synchronous on the context's
 * stream. */
int amj_gen_device_uniform(amj_ctx_t* ctx, int32_t* out, uint64_t n, int64_t lo, int64_t hi,
                           uint64_t seed, uint64_t stream);
int amj_gen_device_zipf(amj_ctx_t* ctx, int32_t* out, uint64_t n, uint64_t k, double s,
                        uint64_t seed, uint64_t stream);
/* R-MAT graph on the device: edge_factor << scale edges with quadrant probabilities (a, b, c),
 * vertices relabelled by a seeded permutation, then symmetrised, self-loops and duplicates
 * dropped, oriented u -> v iff (deg u, u) < (deg v, v) and sorted by (src, dst): the same edges
 * as the host generator (amj_gen_rmat + amj_graph_orient in amj_gen.h). out_src and out_dst
 * are device memory of edge_factor << scale int32 each; *edges gets the edges kept. */
int amj_gen_device_rmat(amj_ctx_t* ctx, int scale, uint32_t edge_factor, double a, double b,
                        double c, uint64_t seed, int32_t* out_src, int32_t* out_dst,
                        uint64_t* edges);

/* ---- relations and indexes ---------------------------------------------------------- */
int amj_relation_create(amj_ctx_t* ctx, const amj_rel_t* rel, amj_relation_t** out);
void amj_relation_destroy(amj_ctx_t* ctx, amj_relation_t* rel);
/* Build the bucketed hash table of `rel` on column key_col, with B = floorPowerOf2(rows)
 * buckets (QueryProcessor.cuh:86). sort_vals=1 also orders the other column inside each
 * key run (needed when that column is compared rather than bound). */
int amj_hash_table_build(amj_ctx_t* ctx, const amj_relation_t* rel, uint32_t key_col,
                         int sort_vals, amj_hash_table_t** out);
void amj_hash_table_destroy(amj_ctx_t* ctx, amj_hash_table_t* ht);
/* Size and build statistics of a hash table (HashTable::get_ht_size, Relation.cuh:235-247). */
int amj_hash_table_info(amj_ctx_t* ctx, const amj_hash_table_t* ht, uint64_t* rows,
                        uint64_t* buckets, int* identity, double* build_ms);

/* ---- AMHJ ----------------------------------------------------------------------------- */
/* OOC::execute: rels[0] is the probe relation, ht[i] is built on rels[i+1] (its attribute
 * list is `attr_ids` of that relation, hash column first). Output attributes follow the
 * reference's first-occurrence rule (exec_AMHJ.cu:76-88). */
int amj_execute(amj_ctx_t* ctx, const amj_relation_t* probe, amj_hash_table_t* const* hts,
                uint32_t n_hts, const amj_opts_t* opts, amj_result_t* res);
/* exec-AMHJ NORMAL in one call: relations in chain order, rels[0] = probe. */
int amj_chain_join(amj_ctx_t* ctx, const amj_rel_t* rels, uint32_t n, const amj_opts_t* opts,
                   amj_result_t* res);
/* exec-AMHJ TRI / FOUR over one two-column edge relation (exec_AMHJ.cu:32-69). */
int amj_pattern_join(amj_ctx_t* ctx, const amj_rel_t* edges, int query, const amj_opts_t* opts,
                     amj_result_t* res);
/* TRI / FOUR over a resident edge relation and its table (key column 0, sort_vals=1), the
 * resident counterpart of amj_pattern_join (build_hash_single, QueryProcessor.cuh:112-169). */
int amj_execute_pattern(amj_ctx_t* ctx, const amj_relation_t* edges, const amj_hash_table_t* ht,
                        int query, const amj_opts_t* opts, amj_result_t* res);
/* amj_execute's count + checksum (materialize = 0) with no device allocation. No counterpart
 * in the reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Runs the plain engine's
 * one-pass count-and-hash launch with its temporaries in a fixed 8.4 MB device array (held on
 * every GPU the library runs on until the process ends); the probe relation must be resident
 * and hold at most 2^27 rows (AMJ_ERR_ARG beyond). opts may be NULL; otherwise materialize,
 * ooc, ws, dro and fib must be 0 and the engine AMJ_ENGINE_WAVE. Calls that overlap on one GPU
 * share that array and race. res gets count and checksum (no columns). */
int amj_execute_noalloc(amj_ctx_t* ctx, const amj_relation_t* probe,
                        amj_hash_table_t* const* hts, uint32_t n_hts, const amj_opts_t* opts,
                        amj_result_t* res);
int amj_result_free(amj_ctx_t* ctx, amj_result_t* res);
/* Copy rows [row0, row0+n) of result column `col` to host memory. */
int amj_result_copy_col(amj_ctx_t* ctx, const amj_result_t* res, uint32_t col, uint64_t row0,
                        uint64_t n, int32_t* host_out);

/* ---- autotuner ---------------------------------------------------------------------------- */
/* No counterpart in the reference (its schedule is fixed at compile time, MHJ-GPU/conf.h).
 * Times candidate engine options with the caller's materialise setting on two evenly strided
 * samples of the probe relation (at most 2 M and 4 M rows; the relation itself up to 4 M rows):
 * the unit engine, the persistent engine with small and large tasks and claims, direct output,
 * and for frequent-item chains the output-window size (1 Ki / 4 Ki / 16 Ki rows) with and
 * without suffix tables. Each candidate's device time is extrapolated from the two samples to
 * the whole relation (affine in the probe rows: fixed costs such as suffix tables count once).
 * The fastest is written into *opts (materialize and checksum kept) and remembered in the
 * context for the same query shape; tuned_ms is its extrapolated device time. */
int amj_autotune(amj_ctx_t* ctx, const amj_relation_t* probe, amj_hash_table_t* const* hts,
                 uint32_t n_hts, amj_opts_t* opts, double* tuned_ms);
void amj_autotune_forget(amj_ctx_t* ctx);

/* ---- batches: many AMHJ queries per launch --------------------------------------------- */
/* No counterpart in the reference (one query per process, exec_AMHJ.cu). Queries over
 * resident relations and hash tables are queued, then run together: one count launch and one
 * write launch per group of equal depth, one scan, two host round trips per batch. Each result
 * equals the one amj_execute gives for the same query (count, checksum, tuples in the same
 * order). Queries added with opts->dro = 1 skip the count pass instead (the direct result
 * output of AMHJ_DRO, Joins/AMHJ_DRO.cuh:18-707, applied to the whole batch): each output is
 * sized from a strided sample of the query's probe rows (or from opts->dro_capacity_rows),
 * one launch per depth reserves rows with one atomic per 256 staged rows, taking its units
 * heaviest first, and a query whose results overrun its size is run
 * again exactly; counts and checksums are the same, the row order is not. dro queries are
 * admitted in the order they were added while their outputs fit the budget: the rest is not
 * run and amj_batch_result returns AMJ_ERR_NOMEM for it. Results are taken with
 * amj_batch_result and released with amj_result_free. */
int amj_batch_create(amj_ctx_t* ctx, amj_batch_t** out);
void amj_batch_destroy(amj_batch_t* batch);
/* Queue a NORMAL query (the arguments of amj_execute); *id numbers it within the batch.
 * opts->materialize, opts->dro and opts->dro_capacity_rows are used (the batch engine is the
 * unit engine; ws, fib and ooc do not apply; opts->engine must be AMJ_ENGINE_WAVE). */
int amj_batch_add(amj_batch_t* batch, const amj_relation_t* probe, amj_hash_table_t* const* hts,
                  uint32_t n_hts, const amj_opts_t* opts, uint32_t* id);
/* Queue a TRI / FOUR query over an edge relation and its table (key column 0, sort_vals=1). */
int amj_batch_add_pattern(amj_batch_t* batch, const amj_relation_t* edges,
                          const amj_hash_table_t* ht, int query, const amj_opts_t* opts,
                          uint32_t* id);
/* Bytes the batch's materialised results may take (default: what the device has free);
 * amj_batch_run returns AMJ_ERR_NOMEM, with nothing run, when they would take more. */
int amj_batch_set_budget(amj_batch_t* batch, uint64_t bytes);
int amj_batch_run(amj_batch_t* batch);
/* Capture (no counterpart in the reference): record the batch's whole direct-output chain into
 * a HIP graph. Every query must have been added with opts->dro = 1. Output sizes and places are
 * decided on the device: a strided size sample of every materialised query, one scan, a
 * placement kernel that sizes each output (exact when every unit is sampled, else the sample
 * scaled up by 1.25 + 64 Ki rows) and places it in one result slab of slab_bytes (0: the batch
 * budget, else 1 GiB) owned by the batch, then one direct-output launch per depth and kind:
 * no host round trip inside the chain. The relations and tables must outlive the batch.
 * amj_batch_replay launches the graph and reads the per-query totals once; a materialised
 * query that did not fit the slab or overran its size runs again exactly (outside the graph).
 * Results (amj_batch_result) of an admitted query are views into the slab, valid until the
 * next replay or amj_batch_destroy (amj_result_free on them releases nothing); amj_batch_run
 * does not apply to a captured batch. Counts and checksums equal amj_batch_run's. */
int amj_batch_capture(amj_batch_t* batch, uint64_t slab_bytes);
int amj_batch_replay(amj_batch_t* batch);
int amj_batch_result(amj_batch_t* batch, uint32_t id, amj_result_t* res);
/* Queries, device time of the run, and the bytes its materialised results take (also set when
 * amj_batch_run returned AMJ_ERR_NOMEM). */
int amj_batch_info(const amj_batch_t* batch, uint32_t* queries, double* run_ms,
                   uint64_t* result_bytes);

/* ---- serving ----------------------------------------------------------------------------- */
/* No counterpart in the reference. A server owns a context and a worker thread: relations are
 * registered once (resident, hashed on column 0 with value-sorted runs), requests from any host
 * thread are served with continuous batching (each cycle runs, as one batch, every request that
 * arrived while the previous batch was on the GPU, up to max_batch, as dro batch queries: one
 * direct-output launch per depth, plus one sample launch for queries never served before —
 * a repeated query is sized from its last exact count), and results are admitted against
 * hbm_budget_bytes (bytes of results held by clients until amj_server_release; 0 = half the
 * device; a result counts with its whole output buffer, which direct output sizes from an
 * estimate). plain = 1 serves one query per launch instead (amj_execute's default engine, the
 * comparison path); plain = 2 one query per launch with direct output (opts.dro = 1), which
 * separates the gain of batching from that of single-pass output. The result budget is
 * reserved as one arena at creation; amj_server_set_budget reserves it again at another size
 * before the first request (e.g. once the resident relations are registered). */
typedef struct amj_server amj_server_t;
int amj_server_create(int device, uint64_t hbm_budget_bytes, uint32_t max_batch, int plain,
                      amj_server_t** out);
void amj_server_destroy(amj_server_t* server);
/* Message of the last failed call on this server (valid until the next call). */
const char* amj_server_last_error(amj_server_t* server);
int amj_server_add_relation(amj_server_t* server, const amj_rel_t* rel, uint32_t* rel_id);
/* NORMAL chain over registered relations (rel_ids[0] probes, the others are hashed on their
 * first column), or TRI / FOUR over a registered edge relation. Returns at once. */
int amj_server_submit(amj_server_t* server, const uint32_t* rel_ids, uint32_t n, int materialize,
                      uint64_t* ticket);
int amj_server_submit_pattern(amj_server_t* server, uint32_t edge_rel, int query, int materialize,
                              uint64_t* ticket);
/* Block until the request is served; latency_ms = submission to completion. */
int amj_server_wait(amj_server_t* server, uint64_t ticket, amj_result_t* res, double* latency_ms);
/* The same, with queue_ms = submission to the start of the launch or batch that served it. */
int amj_server_wait_ex(amj_server_t* server, uint64_t ticket, amj_result_t* res,
                       double* latency_ms, double* queue_ms);
/* Batched mode: NORMAL requests whose probe relation has at least min_probe_rows rows run
 * alone, with the engine options amj_autotune picks once per class of request (depth,
 * materialise, probe size within a factor of 4); the others go into the batch. 0: off. */
int amj_server_set_autotune(amj_server_t* server, uint64_t min_probe_rows);
/* Requests run alone with tuned options, and the classes tuned so far. */
int amj_server_tuning(amj_server_t* server, uint64_t* tuned_alone, uint64_t* classes);
/* Batched mode (no counterpart in the reference): TRI / FOUR requests over an edge relation of
 * at least min_edges rows run alone through the leapfrog triejoin (exec-LFTJ's SELF aliasing of
 * one trie, built on the first such request and kept resident with the relation): triangles by
 * the DFS unit engine (direct output when materialised), 4-cliques counted with work sharing.
 * Count and checksum equal the hash-join pattern engine's. 0 (default): off. */
int amj_server_set_triejoin(amj_server_t* server, uint64_t min_edges);
/* Requests served through the triejoin so far. */
int amj_server_triejoin_runs(amj_server_t* server, uint64_t* runs);
/* Set the result budget before the first request (AMJ_ERR_STATE afterwards); the arena is
 * reserved again at that size then. */
int amj_server_set_budget(amj_server_t* server, uint64_t hbm_budget_bytes);
/* Device bytes held for registered relations and their tables, the result budget, and the
 * device memory free now (any pointer may be NULL). */
int amj_server_memory(amj_server_t* server, uint64_t* resident_bytes, uint64_t* budget_bytes,
                      uint64_t* device_free_bytes);
int amj_server_release(amj_server_t* server, amj_result_t* res);
int amj_server_copy_col(amj_server_t* server, const amj_result_t* res, uint32_t col,
                        uint64_t row0, uint64_t n, int32_t* host_out);
int amj_server_stats(amj_server_t* server, uint64_t* batches, uint64_t* served, uint64_t* splits,
                     double* busy_ms, uint64_t* bytes_held);

/* ---- ALFTJ (leapfrog triejoin) ---------------------------------------------------------- */
/* Sorted two-level CSR trie of a two-column relation (constructSortedTrie<D,C,2>,
 * LFTJ-GPU/multi_partitioning.cuh:129-293, IndexedTrie, LFTJ-GPU/IndexedTrie.cuh:21-54):
 * level 0 = distinct col0 values, level 1 = their sorted distinct col1 children. Duplicate
 * pairs are dropped (the triejoin intersects sets). attr_ids of `rel` name the two levels. */
int alftj_trie_build(amj_ctx_t* ctx, const amj_rel_t* rel, amj_trie_t** out);
/* The same from a relation already resident on the device (amj_relation_create). */
int alftj_trie_build_resident(amj_ctx_t* ctx, const amj_relation_t* rel, amj_trie_t** out);
void alftj_trie_destroy(amj_ctx_t* ctx, amj_trie_t* trie);
/* Level sizes, the longest child list and the build time of a trie. */
int alftj_trie_info(amj_ctx_t* ctx, const amj_trie_t* trie, uint64_t* level0, uint64_t* children,
                    uint64_t* max_children, double* build_ms);
/* Copy a level to host memory: level 0 values (n0) and offsets (n0 + 1), or level 1 values. */
int alftj_trie_copy_level(amj_ctx_t* ctx, const amj_trie_t* trie, int level, int32_t* values,
                          uint32_t* offsets);
/* LFTJ_Base::evaluate via OOCWrapper::execute_with_{DFS,BFS} (LFTJ-GPU/LFTJ_base.h:20-22,
 * LFTJ-GPU/ooc_wrapper.cuh:404-427): join the tries on attributes attr_order[0..n_attrs-1]
 * (2..4 attributes, 1..6 tries; each trie's level-0 attribute precedes its level-1 attribute
 * in attr_order). Output column i holds attribute attr_order[i]. algo = ALFTJ_DFS
 * materialises (opts->materialize) and checksums; ALFTJ_BFS returns count and checksum. */
int alftj_run(amj_ctx_t* ctx, amj_trie_t* const* tries, uint32_t n, const uint32_t* attr_order,
              uint32_t n_attrs, int algo, const amj_opts_t* opts, amj_result_t* res);
/* As alftj_run, with the attributes of trie i given as trie_attrs[2i], trie_attrs[2i+1] instead
 * of the trie's own: one trie can then appear several times under different attributes, as
 * SELF mode aliases Trie[0] (exec_LFTJ.cu:198-222). */
int alftj_run_aliased(amj_ctx_t* ctx, amj_trie_t* const* tries, uint32_t n,
                      const uint32_t* trie_attrs, const uint32_t* attr_order, uint32_t n_attrs,
                      int algo, const amj_opts_t* opts, amj_result_t* res);
/* exec-LFTJ SELF mode (exec_LFTJ.cu:198-231): one trie over `edges`, aliased as R(0,1),
 * S(1,2), T(0,2) for TRI; FOUR adds (2,3), (0,3), (1,3) (the 4-clique of exec_AMHJ.cu:59). */
int alftj_pattern(amj_ctx_t* ctx, const amj_rel_t* edges, int query, int algo,
                  const amj_opts_t* opts, amj_result_t* res);
/* A trie from host CSR arrays (IndexedTrie data/offsets of a 2-level trie,
 * IndexedTrie.cuh:21-54): d0[n0] strictly increasing, off[n0+1] (off[0] = 0, off[n0] = n1),
 * d1[n1] sorted within each run. Duplicate children are kept. attr_ids may be NULL ({0,1}). */
int alftj_trie_from_csr(amj_ctx_t* ctx, uint64_t n0, const int32_t* d0, const uint64_t* off,
                        uint64_t n1, const int32_t* d1, const uint32_t* attr_ids,
                        amj_trie_t** out);
/* .tr files (IndexedTrie::serialization / init_with_deserialization, IndexedTrie.cuh:122-197):
 * u32 num_attrs, u64 data_len[], u32 attr_list[], int32 levels, u64 offsets. Two levels only. */
int alftj_trie_save(amj_ctx_t* ctx, const amj_trie_t* trie, const char* path);
/* Move a trie to pinned host memory (the device copy is freed). Queries over such tries run
 * out of core (OOCWrapper, LFTJ-GPU/ooc_wrapper.cuh:14-427): per combination of level-0 value
 * ranges, the slices streamed on a side stream; opts.ooc = 1 does the same for device tries,
 * and opts.ooc_slice_rows sets the level-0 values per range (0: from the HBM budget). Out of
 * core, alftj_run returns count and checksum (no materialised tuples). */
int alftj_trie_offload(amj_ctx_t* ctx, amj_trie_t* trie);
int alftj_trie_load(amj_ctx_t* ctx, const char* path, amj_trie_t** out);

/* ---- files and result fingerprint (host side, no context needed) ----------------------- */
/* .db relation files (read_rel_cols_mmap, common-utils/file_io.h:90-121, names as in
 * QueryProcessor.cuh:41-73): <rows>_<attr0>_<attr1>.db holds column 0 then column 1, raw
 * little-endian int32. parse_name fills rows, num_cols and up to max_cols attribute ids. */
int amj_db_parse_name(const char* path, uint64_t* rows, uint32_t* num_cols, uint32_t* attr_ids,
                      uint32_t max_cols);
int amj_db_read(const char* path, uint64_t rows, uint32_t num_cols, int32_t* const* cols);
/* Write a host relation as <dir>/<rows>_<attr0>_<attr1>.db; the path goes to path_out. */
int amj_db_write(const char* dir, const amj_rel_t* rel, char* path_out, size_t path_len);
/* Message of the last failed call that took no context (file functions). */
const char* amj_file_error(void);
/* res_stat (common-utils/res_stat.h:32-133): MD5 of the text "[h0, h1, ...]\n" of the histogram
 * of (v & 0xFFFFF) over all values of all columns, up to the largest such value. hex_out gets
 * 32 hex digits and a NUL. amj_result_md5 histograms a materialised result on the device;
 * amj_res_stat_md5 takes host columns. */
int amj_result_md5(amj_ctx_t* ctx, const amj_result_t* res, char* hex_out);
int amj_res_stat_md5(const int32_t* const* cols, uint32_t num_cols, uint64_t rows,
                     char* hex_out);

#ifdef __cplusplus
}
#endif
#endif /* AMJ_H_ */
