// Serving loop (amj_server_*): relations and their hash tables stay resident on the device;
// requests (NORMAL chains, TRI, FOUR over resident relations) are queued by any host thread and
// served by one worker thread with continuous batching: every cycle takes all requests that
// arrived while the previous batch was on the GPU (up to max_batch), runs them as one batch
// (amj_batch_run, every query with direct output: no count pass; a query seen before is
// sized from its last exact count, a new one from a sample), and admits their results against
// an HBM budget (results held by clients count until they are released). A batch whose results would overrun the budget is split in
// halves; a single request that does not fit fails with AMJ_ERR_NOMEM. The plain mode runs the
// same requests one query per launch (the engines behind amj_execute), for comparison.
//
// Locking: `mu` guards the request queue, the finished results and the counters; `gpu` guards
// the device context and is held by the worker while a batch runs, so submissions, waits and
// releases never wait for the GPU. Released results are returned to the context's pool by the
// worker at the start of its next cycle.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "runtime.h"

using namespace amj;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

struct amj_server {
  amj_ctx_t* ctx = nullptr;
  uint64_t budget = 0;  // bytes of results that may be held at once
  uint32_t max_batch = 256;
  int plain = 0;        // 1: one query per launch; 2: one query per launch with direct output
  bool arena_ready = false;  // the result arena is reserved (under gpu)
  // Batched mode: NORMAL requests whose probe relation has at least tune_rows rows run alone
  // with the options amj_autotune picks for their class (depth, materialise, probe size class:
  // one tuning per class, under gpu); 0 = every request goes into the batch.
  uint64_t tune_rows = 0;
  std::map<std::vector<uint64_t>, amj_opts_t> tuned_opts;
  uint64_t tuned_alone = 0;  // under mu
  // Batched mode: TRI / FOUR requests over an edge relation of at least trie_edges rows run
  // alone through the leapfrog triejoin, over a trie built once per relation and kept resident
  // (under gpu); 0 = every pattern request goes into the batch.
  uint64_t trie_edges = 0;
  uint64_t trie_alone = 0;  // under mu

  std::mutex mu, gpu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;

  struct Resident {
    amj_relation_t* rel = nullptr;
    amj_hash_table_t* ht = nullptr;
    amj_trie_t* trie = nullptr;  // built on the first triejoin request over the relation
  };
  std::vector<Resident> rels;          // under gpu
  std::atomic<uint32_t> n_rels{0};

  struct Req {
    uint64_t ticket;
    int query;                   // AMJ_QUERY_*
    std::vector<uint32_t> ids;   // NORMAL: chain, ids[0] probes; TRI/FOUR: the edge relation
    int materialize;
    double t_submit;
    double t_start = 0;  // when the launch (plain) or batch that serves it began
  };
  struct Done {
    int status = AMJ_OK;
    std::string err;
    amj_result_t res{};
    double t_submit = 0, t_start = 0, t_done = 0;
    uint64_t bytes = 0;
  };
  struct Outcome {
    uint64_t ticket;
    Done d;
  };
  // Exact result size of every (query, relation ids) served so far (under gpu): relations are
  // immutable while registered, so a repeated query's output is sized exactly and its batch
  // needs no sample launch for it.
  std::map<std::vector<uint32_t>, uint64_t> result_rows;
  static std::vector<uint32_t> shape(const Req& q) {
    std::vector<uint32_t> k{(uint32_t)q.query};
    k.insert(k.end(), q.ids.begin(), q.ids.end());
    return k;
  }
  std::deque<Req> pending;             // under mu
  std::map<uint64_t, Done> done;       // under mu
  std::vector<amj_result_t> released;  // under mu: returned to the pool by the worker
  uint64_t next_ticket = 1;
  uint64_t in_use = 0;                 // bytes of results handed out and not yet released
  uint64_t batches = 0, served = 0, splits = 0;
  double busy_ms = 0.0;
  std::string last_err;

  // Device bytes a result holds: its whole output buffer (a direct-output buffer is sized
  // from an estimate and is larger than the rows it holds).
  static uint64_t result_bytes(const amj_result_t& r) {
    if ((r.reserved[2] == 1 || r.reserved[2] == 2) && r.cols[0])
      return (uint64_t)r.reserved[0] | ((uint64_t)r.reserved[1] << 32);
    return r.rows_written * (uint64_t)r.num_cols * 4;
  }

  // --- worker side (gpu held) -------------------------------------------------------------
  struct Cycle {
    std::vector<Outcome> out;
    std::vector<Req> deferred;  // fit the budget, not the room left now: run after releases
    uint64_t deferred_need = 0; // bytes the first deferred request needs
    uint64_t room;
    uint64_t batches = 0, splits = 0, tuned = 0, tries = 0;
  };

  static Outcome failed(const Req& q, int rc, const std::string& msg) {
    Outcome o;
    o.ticket = q.ticket;
    o.d.status = rc;
    o.d.err = msg;
    o.d.t_submit = q.t_submit;
    o.d.t_done = now_ms();
    o.d.t_start = q.t_start ? q.t_start : o.d.t_done;
    return o;
  }

  void finished(Cycle& c, const Req& q, const amj_result_t& res) {
    Outcome o;
    o.ticket = q.ticket;
    o.d.res = res;
    o.d.t_submit = q.t_submit;
    o.d.t_start = q.t_start;
    o.d.t_done = now_ms();
    o.d.bytes = result_bytes(res);
    c.room -= std::min(c.room, o.d.bytes);
    c.out.push_back(o);
  }

  void run_group(Cycle& c, std::vector<Req> reqs) {
    if (reqs.empty()) return;
    if (!c.deferred.empty()) {  // nothing overtakes a deferred request: results finish in order
      c.deferred.insert(c.deferred.end(), reqs.begin(), reqs.end());
      return;
    }
    if (plain) {
      for (auto& q : reqs) {
        if (!c.deferred.empty()) c.deferred.push_back(q);  // keep the order behind a deferral
        else run_one(c, q);
      }
      return;
    }
    // In submission order: runs of batchable requests go as one batch each, large chains alone
    // with their tuned engine, large graph patterns alone through the triejoin. A request
    // never starts before an earlier one: were a later result admitted first, an earlier
    // request could wait for room that only the later result's release frees, while a client
    // waits for the earlier result before it releases the later one.
    std::vector<Req> run;
    auto flush = [&] {
      if (!run.empty()) run_batch_group(c, run);
      run.clear();
    };
    for (auto& q : reqs) {
      const uint64_t rows = rels[q.ids[0]].rel->rows;
      const bool tuned = tune_rows && q.query == AMJ_QUERY_NORMAL && rows >= tune_rows;
      const bool trie = trie_edges && q.query != AMJ_QUERY_NORMAL && rows >= trie_edges;
      if (tuned || trie) {
        flush();
        if (!c.deferred.empty()) c.deferred.push_back(q);
        else if (tuned) run_one(c, q, true);
        else run_trie(c, q);
      } else if (!c.deferred.empty()) {
        c.deferred.push_back(q);
      } else {
        run.push_back(q);
      }
    }
    flush();
  }

  // Batchable requests (in order) as one batch, split in halves while their results do not fit.
  void run_batch_group(Cycle& c, std::vector<Req> reqs) {
    if (reqs.empty()) return;
    if (!c.deferred.empty()) {
      c.deferred.insert(c.deferred.end(), reqs.begin(), reqs.end());
      return;
    }
    const double t0 = now_ms();
    for (auto& q : reqs) q.t_start = t0;
    uint64_t need = 0;
    const int rc = run_batch(c, reqs, &need);
    if (rc == AMJ_ERR_NOMEM && reqs.size() > 1) {
      ++c.splits;
      std::vector<Req> a(reqs.begin(), reqs.begin() + reqs.size() / 2);
      std::vector<Req> b(reqs.begin() + reqs.size() / 2, reqs.end());
      run_batch_group(c, a);
      run_batch_group(c, b);
    } else if (rc == AMJ_ERR_NOMEM && (need > budget || c.room >= budget)) {
      // larger than the budget, or no room even with nothing held: waiting cannot help
      c.out.push_back(failed(reqs[0], rc, "the result (" + std::to_string(need) +
                                              " bytes) exceeds the server's HBM budget"));
    } else if (rc == AMJ_ERR_NOMEM) {
      if (c.deferred.empty()) c.deferred_need = need;
      c.deferred.push_back(reqs[0]);
    }
  }

  void run_one(Cycle& c, Req q, bool tuned = false) {
    q.t_start = now_ms();
    amj_opts_t o;
    amj_default_opts(&o);
    o.materialize = q.materialize;
    o.dro = plain == 2 && q.materialize ? 1 : 0;  // direct output where something is written
    amj_result_t res;
    int rc = AMJ_OK;
    if (q.query == AMJ_QUERY_NORMAL) {
      std::vector<amj_hash_table_t*> hts;
      for (size_t i = 1; i < q.ids.size(); ++i) hts.push_back(rels[q.ids[i]].ht);
      if (tuned) {
        uint64_t cls = 0;
        while ((1ull << (2 * cls)) < rels[q.ids[0]].rel->rows && cls < 31) ++cls;
        const std::vector<uint64_t> key{(uint64_t)hts.size(), (uint64_t)q.materialize, cls};
        auto it = tuned_opts.find(key);
        if (it == tuned_opts.end()) {
          amj_opts_t t = o;
          rc = amj_autotune(ctx, rels[q.ids[0]].rel, hts.data(), (uint32_t)hts.size(), &t,
                            nullptr);
          if (rc == AMJ_OK) it = tuned_opts.emplace(key, t).first;
        }
        if (rc == AMJ_OK) o = it->second;
        o.materialize = q.materialize;
        o.checksum = 1;
        ++c.tuned;
      }
      o.hbm_budget_bytes = ctx->allocated + c.room;
      if (rc == AMJ_OK)
        rc = amj_execute(ctx, rels[q.ids[0]].rel, hts.data(), (uint32_t)hts.size(), &o, &res);
    } else {
      o.hbm_budget_bytes = ctx->allocated + c.room;
      Plan plan;
      std::vector<uint32_t> ids;
      rc = make_pattern_plan(ctx, rels[q.ids[0]].rel, rels[q.ids[0]].ht, q.query, &plan, &ids);
      if (rc == AMJ_OK) rc = execute_plan(ctx, plan, ids, &o, &res);
    }
    if (rc == AMJ_ERR_NOMEM && c.room < budget) {
      // the result may fit once held results are released: wait until nothing is held
      c.deferred_need = budget;
      c.deferred.push_back(q);
      return;
    }
    ++c.batches;
    if (rc != AMJ_OK) c.out.push_back(failed(q, rc, ctx->last_error));
    else finished(c, q, res);
  }

  // A TRI / FOUR request alone through the leapfrog triejoin (alftj_run_aliased over the edge
  // relation's trie, aliased as in exec-LFTJ's SELF mode): a triangle with the unit engine
  // (direct output when materialised), a 4-clique counted with work sharing. Count and checksum
  // are those of the hash-join pattern engine (the same result set).
  void run_trie(Cycle& c, Req q) {
    q.t_start = now_ms();
    Resident& r = rels[q.ids[0]];
    int rc = AMJ_OK;
    if (!r.trie) rc = alftj_trie_build_resident(ctx, r.rel, &r.trie);
    amj_result_t res;
    if (rc == AMJ_OK) {
      amj_opts_t o;
      amj_default_opts(&o);
      o.materialize = q.materialize;
      o.checksum = 1;
      o.dro = q.materialize ? 1 : 0;
      o.ws = q.query == AMJ_QUERY_FOUR && !q.materialize ? 1 : 0;
      o.hbm_budget_bytes = ctx->allocated + c.room;
      static const uint32_t tri[6] = {0, 1, 1, 2, 0, 2};
      static const uint32_t four[12] = {0, 1, 1, 2, 2, 3, 0, 3, 0, 2, 1, 3};
      static const uint32_t order[4] = {0, 1, 2, 3};
      const bool is_tri = q.query == AMJ_QUERY_TRI;
      amj_trie_t* ts[6] = {r.trie, r.trie, r.trie, r.trie, r.trie, r.trie};
      rc = alftj_run_aliased(ctx, ts, is_tri ? 3 : 6, is_tri ? tri : four, order, is_tri ? 3 : 4,
                             ALFTJ_DFS, &o, &res);
    }
    if (rc == AMJ_ERR_NOMEM && c.room < budget) {
      c.deferred_need = budget;
      c.deferred.push_back(q);
      return;
    }
    ++c.batches;
    ++c.tries;
    if (rc != AMJ_OK) c.out.push_back(failed(q, rc, ctx->last_error));
    else finished(c, q, res);
  }

  // AMJ_ERR_NOMEM (with reqs reduced to the requests still to run, *need the bytes their
  // results take): the results overrun the room.
  int run_batch(Cycle& c, std::vector<Req>& reqs, uint64_t* need) {
    amj_opts_t o;
    amj_default_opts(&o);
    o.dro = 1;  // one pass per batch: outputs sized from a sample, overruns rerun exactly
    amj_batch_t* b = nullptr;
    amj_batch_create(ctx, &b);
    amj_batch_set_budget(b, c.room);
    std::vector<Req> planned;
    std::vector<uint32_t> pid;
    for (const Req& q : reqs) {
      o.materialize = q.materialize;
      const auto known = result_rows.find(shape(q));
      o.dro_capacity_rows = known != result_rows.end() ? std::max<uint64_t>(known->second, 1) : 0;
      uint32_t id = 0;
      int rc;
      if (q.query == AMJ_QUERY_NORMAL) {
        std::vector<amj_hash_table_t*> hts;
        for (size_t k = 1; k < q.ids.size(); ++k) hts.push_back(rels[q.ids[k]].ht);
        rc = amj_batch_add(b, rels[q.ids[0]].rel, hts.data(), (uint32_t)hts.size(), &o, &id);
      } else {
        rc = amj_batch_add_pattern(b, rels[q.ids[0]].rel, rels[q.ids[0]].ht, q.query, &o, &id);
      }
      if (rc != AMJ_OK) {
        c.out.push_back(failed(q, rc, ctx->last_error));  // a malformed request fails alone
      } else {
        planned.push_back(q);
        pid.push_back(id);
      }
    }
    const int rc = amj_batch_run(b);
    if (rc == AMJ_ERR_NOMEM) {
      amj_batch_info(b, nullptr, nullptr, need);
      amj_batch_destroy(b);
      reqs.swap(planned);
      return AMJ_ERR_NOMEM;
    }
    ++c.batches;
    for (size_t i = 0; i < planned.size(); ++i) {
      if (rc != AMJ_OK) {
        c.out.push_back(failed(planned[i], rc, ctx->last_error));
        continue;
      }
      amj_result_t res;
      const int r = c.deferred.empty() ? amj_batch_result(b, pid[i], &res) : AMJ_ERR_NOMEM;
      if (r == AMJ_ERR_NOMEM) {  // not admitted: runs after results are released
        if (c.deferred.empty()) c.deferred_need = 0;
        c.deferred.push_back(planned[i]);
      } else if (r != AMJ_OK) {
        c.out.push_back(failed(planned[i], r, ctx->last_error));
      } else {
        result_rows[shape(planned[i])] = res.count;
        finished(c, planned[i], res);
      }
    }
    amj_batch_destroy(b);
    return AMJ_OK;
  }

  void loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    uint64_t gate = 0;         // room the request at the head of the queue waits for
    bool wait_release = false;  // the head was deferred: retry once something is released
    for (;;) {
      cv_work.wait(lk, [&] {
        const uint64_t room = budget > in_use ? budget - in_use : 0;
        return stop || !released.empty() || (!pending.empty() && !wait_release && room >= gate);
      });
      if (stop && pending.empty()) return;
      std::vector<amj_result_t> rel;
      rel.swap(released);
      std::vector<Req> reqs;
      while (!pending.empty() && reqs.size() < max_batch) {
        reqs.push_back(pending.front());
        pending.pop_front();
      }
      Cycle c;
      c.room = budget > in_use ? budget - in_use : 0;
      lk.unlock();
      const double t0 = now_ms();
      {
        std::lock_guard<std::mutex> g(gpu);
        for (auto& r : rel) amj_result_free(ctx, &r);
        if (!arena_ready && !reqs.empty()) {  // a reservation that failed at creation
          const int rc = arena_reserve(ctx, budget);
          if (rc == AMJ_OK) {
            arena_ready = true;
          } else {
            for (auto& q : reqs)
              c.out.push_back(failed(q, rc, "cannot reserve the result budget: " +
                                                ctx->last_error));
            reqs.clear();
          }
        }
        run_group(c, reqs);
      }
      const double t1 = now_ms();
      lk.lock();
      if (!reqs.empty()) busy_ms += t1 - t0;
      batches += c.batches;
      splits += c.splits;
      tuned_alone += c.tuned;
      trie_alone += c.tries;
      for (auto& o : c.out) {
        if (o.d.status == AMJ_OK) {
          in_use += o.d.bytes;
          ++served;
        }
        done[o.ticket] = o.d;
      }
      // deferred requests go back to the head of the queue, in order
      for (auto it = c.deferred.rbegin(); it != c.deferred.rend(); ++it) pending.push_front(*it);
      gate = c.deferred.empty() ? 0 : c.deferred_need;
      wait_release = !c.deferred.empty();
      if (stop && !c.deferred.empty()) {  // nobody will release any more: fail them
        for (auto& q : c.deferred) done[q.ticket] = failed(q, AMJ_ERR_NOMEM, "server stopped").d;
        pending.clear();
      }
      cv_done.notify_all();
    }
  }
};

extern "C" {

int amj_server_create(int device, uint64_t hbm_budget_bytes, uint32_t max_batch, int plain,
                      amj_server_t** out) {
  if (!out) return AMJ_ERR_ARG;
  *out = nullptr;
  amj_ctx_t* ctx = nullptr;
  AMJ_TRY(amj_ctx_create(device, &ctx));
  auto* s = new amj_server();
  s->ctx = ctx;
  s->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)ctx->props.totalGlobalMem / 2;
  if (plain < 0 || plain > 2) {
    delete s;
    amj_ctx_destroy(ctx);
    return AMJ_ERR_ARG;
  }
  // Every result is carved from one reservation of the budget: no device allocation (and no
  // free, whose memory this device clears before handing it out again, which costs seconds
  // for tens of GB) per request. Reserved here, outside any request's latency.
  const int rc = arena_reserve(ctx, s->budget);
  if (rc != AMJ_OK) {
    delete s;
    amj_ctx_destroy(ctx);
    return rc;
  }
  s->arena_ready = true;
  s->max_batch = max_batch ? max_batch : 256;
  s->plain = plain;
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return AMJ_OK;
}

void amj_server_destroy(amj_server_t* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& kv : s->done)
    if (kv.second.status == AMJ_OK) amj_result_free(s->ctx, &kv.second.res);
  for (auto& r : s->released) amj_result_free(s->ctx, &r);
  for (auto& r : s->rels) {
    if (r.trie) alftj_trie_destroy(s->ctx, r.trie);
    amj_hash_table_destroy(s->ctx, r.ht);
    amj_relation_destroy(s->ctx, r.rel);
  }
  amj_ctx_destroy(s->ctx);
  delete s;
}

const char* amj_server_last_error(amj_server_t* s) {
  if (!s) return "no server";
  std::lock_guard<std::mutex> lk(s->mu);
  return s->last_err.c_str();
}

int amj_server_add_relation(amj_server_t* s, const amj_rel_t* rel, uint32_t* rel_id) {
  if (!s || !rel) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> g(s->gpu);
  amj_server::Resident r;
  int rc = amj_relation_create(s->ctx, rel, &r.rel);
  // hash on column 0 with value-sorted runs: serves chains and TRI/FOUR alike
  if (rc == AMJ_OK) rc = amj_hash_table_build(s->ctx, r.rel, 0, 1, &r.ht);
  if (rc != AMJ_OK) {
    amj_relation_destroy(s->ctx, r.rel);
    std::lock_guard<std::mutex> lk(s->mu);
    s->last_err = s->ctx->last_error;
    return rc;
  }
  s->rels.push_back(r);
  if (rel_id) *rel_id = (uint32_t)(s->rels.size() - 1);
  s->n_rels.store((uint32_t)s->rels.size());
  return AMJ_OK;
}

static int submit(amj_server_t* s, int query, const uint32_t* ids, uint32_t n, int materialize,
                  uint64_t* ticket) {
  std::lock_guard<std::mutex> lk(s->mu);
  for (uint32_t i = 0; i < n; ++i)
    if (ids[i] >= s->n_rels.load()) {
      s->last_err = "unknown relation id";
      return AMJ_ERR_ARG;
    }
  amj_server::Req q;
  q.ticket = s->next_ticket++;
  q.query = query;
  q.ids.assign(ids, ids + n);
  q.materialize = materialize;
  q.t_submit = now_ms();
  s->pending.push_back(q);
  if (ticket) *ticket = q.ticket;
  s->cv_work.notify_one();
  return AMJ_OK;
}

int amj_server_submit(amj_server_t* s, const uint32_t* rel_ids, uint32_t n, int materialize,
                      uint64_t* ticket) {
  if (!s || !rel_ids || n < 2) return AMJ_ERR_ARG;
  return submit(s, AMJ_QUERY_NORMAL, rel_ids, n, materialize, ticket);
}

int amj_server_submit_pattern(amj_server_t* s, uint32_t edge_rel, int query, int materialize,
                              uint64_t* ticket) {
  if (!s || (query != AMJ_QUERY_TRI && query != AMJ_QUERY_FOUR)) return AMJ_ERR_ARG;
  return submit(s, query, &edge_rel, 1, materialize, ticket);
}

int amj_server_set_autotune(amj_server_t* s, uint64_t min_probe_rows) {
  if (!s) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> g(s->gpu);
  s->tune_rows = min_probe_rows;
  return AMJ_OK;
}

int amj_server_set_triejoin(amj_server_t* s, uint64_t min_edges) {
  if (!s) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> g(s->gpu);
  s->trie_edges = min_edges;
  return AMJ_OK;
}

int amj_server_triejoin_runs(amj_server_t* s, uint64_t* runs) {
  if (!s || !runs) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *runs = s->trie_alone;
  return AMJ_OK;
}

int amj_server_tuning(amj_server_t* s, uint64_t* tuned_alone, uint64_t* classes) {
  if (!s) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> g(s->gpu);
  std::lock_guard<std::mutex> lk(s->mu);
  if (tuned_alone) *tuned_alone = s->tuned_alone;
  if (classes) *classes = s->tuned_opts.size();
  return AMJ_OK;
}

int amj_server_set_budget(amj_server_t* s, uint64_t hbm_budget_bytes) {
  if (!s || hbm_budget_bytes == 0) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> g(s->gpu);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->next_ticket != 1) {
    s->last_err = "the budget can only be set before the first request";
    return AMJ_ERR_STATE;
  }
  // the arena is reserved again at the new size now, not inside a request
  int rc = arena_drop(s->ctx);
  if (rc == AMJ_OK) {
    s->arena_ready = false;
    rc = arena_reserve(s->ctx, hbm_budget_bytes);
    s->arena_ready = rc == AMJ_OK;
  }
  if (rc != AMJ_OK) {
    s->last_err = s->ctx->last_error;
    return rc;
  }
  s->budget = hbm_budget_bytes;
  return AMJ_OK;
}

int amj_server_memory(amj_server_t* s, uint64_t* resident_bytes, uint64_t* budget_bytes,
                      uint64_t* device_free_bytes) {
  if (!s) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> g(s->gpu);
  const uint64_t arena = s->arena_ready ? s->ctx->arena.size : 0;
  if (resident_bytes)
    *resident_bytes = s->ctx->allocated > arena ? s->ctx->allocated - arena : 0;
  if (budget_bytes) *budget_bytes = s->budget;
  if (device_free_bytes) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) f = 0;
    *device_free_bytes = f;
  }
  return AMJ_OK;
}

int amj_server_wait(amj_server_t* s, uint64_t ticket, amj_result_t* res, double* latency_ms) {
  return amj_server_wait_ex(s, ticket, res, latency_ms, nullptr);
}

int amj_server_wait_ex(amj_server_t* s, uint64_t ticket, amj_result_t* res, double* latency_ms,
                       double* queue_ms) {
  if (!s || !res) return AMJ_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (ticket == 0 || ticket >= s->next_ticket) return AMJ_ERR_ARG;
  s->cv_done.wait(lk, [&] { return s->done.count(ticket) != 0; });
  auto it = s->done.find(ticket);
  const int rc = it->second.status;
  if (latency_ms) *latency_ms = it->second.t_done - it->second.t_submit;
  if (queue_ms) *queue_ms = it->second.t_start - it->second.t_submit;
  if (rc == AMJ_OK) *res = it->second.res;
  else s->last_err = it->second.err;
  s->done.erase(it);  // the result's bytes stay counted until amj_server_release
  return rc;
}

int amj_server_release(amj_server_t* s, amj_result_t* res) {
  if (!s || !res) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t b = amj_server::result_bytes(*res);
  s->in_use -= std::min(s->in_use, b);
  s->released.push_back(*res);
  std::memset(res->cols, 0, sizeof(res->cols));
  res->rows_written = 0;
  s->cv_work.notify_one();
  return AMJ_OK;
}

int amj_server_copy_col(amj_server_t* s, const amj_result_t* res, uint32_t col, uint64_t row0,
                        uint64_t n, int32_t* host_out) {
  if (!s || !res || !host_out || col >= res->num_cols || !res->cols[col] ||
      row0 + n > res->rows_written)
    return AMJ_ERR_ARG;
  if (n == 0) return AMJ_OK;
  // a blocking copy on the null stream: needs nothing of the worker's context
  return hipMemcpy(host_out, res->cols[col] + row0, n * 4, hipMemcpyDeviceToHost) == hipSuccess
             ? AMJ_OK
             : AMJ_ERR_HIP;
}

int amj_server_stats(amj_server_t* s, uint64_t* batches, uint64_t* served, uint64_t* splits,
                     double* busy_ms, uint64_t* bytes_held) {
  if (!s) return AMJ_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (batches) *batches = s->batches;
  if (served) *served = s->served;
  if (splits) *splits = s->splits;
  if (busy_ms) *busy_ms = s->busy_ms;
  if (bytes_held) *bytes_held = s->in_use;
  return AMJ_OK;
}

}  // extern "C"
