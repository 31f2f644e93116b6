// exec-AMHJ: the reference's AMHJ command line on libamj.
//
//   exec-AMHJ <NORMAL|TRI|FOUR> <dataset-path> <ooc 0/1> <ws 0/1> <dro 0/1> <fib 0/1> [--count-only]
//
// Reference: main / test_general (MHJ-GPU/exec_AMHJ.cu:10-245). NORMAL reads every .db file of
// the directory and asks for its position in the chain on stdin ("Enter the table order: ",
// QueryProcessor.cuh:62-63); position 0 is the probe relation and every other relation is hashed
// on its first column. When stdin has no answer, the order is inferred: the probe is the relation
// whose first attribute no other relation has, and every next relation is one whose first
// (hash) attribute is already bound. TRI and FOUR take one two-column edge file. The result is
// materialised and fingerprinted as the reference does (res_stat, "Md5sum of result histogram");
// --count-only skips the materialisation and reports count and checksum.
#include <algorithm>
#include <iostream>
#include <set>

#include "cli_common.h"

using namespace cli;

static std::vector<int> infer_chain_order(const std::vector<HostRel>& rels) {
  const int n = (int)rels.size();
  std::vector<int> order;
  std::vector<bool> used(n, false);
  int probe = -1;
  for (int i = 0; i < n && probe < 0; ++i) {
    bool shared = false;
    for (int j = 0; j < n; ++j)
      if (j != i && std::count(rels[j].attrs.begin(), rels[j].attrs.end(), rels[i].attrs[0]))
        shared = true;
    if (!shared) probe = i;
  }
  if (probe < 0) probe = 0;
  order.push_back(probe);
  used[probe] = true;
  std::set<uint32_t> bound(rels[probe].attrs.begin(), rels[probe].attrs.end());
  while ((int)order.size() < n) {
    int next = -1;
    for (int i = 0; i < n && next < 0; ++i)
      if (!used[i] && bound.count(rels[i].attrs[0])) next = i;
    if (next < 0) return {};
    used[next] = true;
    order.push_back(next);
    bound.insert(rels[next].attrs.begin(), rels[next].attrs.end());
  }
  return order;
}

int main(int argc, char** argv) {
  if (argc < 7)
    die("usage: exec-AMHJ <NORMAL|TRI|FOUR> <dataset-path> <ooc> <ws> <dro> <fib> [--count-only]");
  const std::string qt = argv[1], path = argv[2];
  amj_opts_t o;
  amj_default_opts(&o);
  o.ooc = std::atoi(argv[3]);
  o.ws = std::atoi(argv[4]);
  o.dro = std::atoi(argv[5]);
  o.fib = std::atoi(argv[6]);
  const bool count_only = has_flag(argc, argv, 7, "--count-only");
  o.materialize = count_only ? 0 : 1;
  o.checksum = 1;
  int query = -1;
  if (qt == "NORMAL") query = AMJ_QUERY_NORMAL;
  else if (qt == "TRI") query = AMJ_QUERY_TRI;
  else if (qt == "FOUR") query = AMJ_QUERY_FOUR;
  else die("Wrong query type");
  log_info("Query type: %s, ooc=%d, ws=%d, dro=%d, fib=%d", qt.c_str(), o.ooc, o.ws, o.dro, o.fib);

  Timer wall;
  std::vector<HostRel> rels;
  if (query == AMJ_QUERY_NORMAL) {
    std::vector<HostRel> files;
    for (const auto& f : list_files(path))
      if (ends_with(f, ".db")) files.push_back(load_db(path + "/" + f));
    if (files.empty()) die("The directory is empty");
    const int n = (int)files.size();
    std::vector<int> pos(n, -1);
    bool answered = true;
    for (int i = 0; i < n; ++i) {
      std::printf("Enter the table order: ");
      std::fflush(stdout);
      int idx = -1;
      if (!(std::cin >> idx)) {
        answered = false;
        break;
      }
      if (idx < 0 || idx >= n) die("table order %d out of range", idx);
      pos[i] = idx;
    }
    std::printf("\n");
    std::vector<int> order(n, -1);  // order[k] = file at chain position k
    if (answered) {
      for (int i = 0; i < n; ++i) {
        if (order[pos[i]] >= 0) die("table order %d given twice", pos[i]);
        order[pos[i]] = i;
      }
    } else {
      order = infer_chain_order(files);
      if (order.empty()) die("cannot infer a chain order; give it on stdin");
      log_info("No table order on stdin: inferred from the attribute chain");
    }
    for (int k = 0; k < n; ++k) {
      rels.push_back(std::move(files[order[k]]));
      log_info("Table %d: %s", k, rels.back().path.c_str());
    }
  } else {
    rels.push_back(load_db(path));
    if (rels[0].cols.size() != 2) die("TRI/FOUR need a two-column edge file");
  }

  amj_ctx_t* ctx = open_ctx();
  std::vector<amj_rel_t> views;
  for (auto& r : rels) views.push_back(r.view());
  // out of core: the probe relation stays in host memory and is streamed in slices
  if (o.ooc && query == AMJ_QUERY_NORMAL) views[0].on_device = AMJ_REL_HOST;
  amj_result_t res;
  const int rc = query == AMJ_QUERY_NORMAL
                     ? amj_chain_join(ctx, views.data(), (uint32_t)views.size(), &o, &res)
                     : amj_pattern_join(ctx, views.data(), query, &o, &res);
  if (rc != AMJ_OK) die("join failed (%d): %s", rc, amj_last_error(ctx));
  static const char* paths[] = {"?",        "plain",    "fib",    "ws",       "dro",
                                "lftj-dfs", "lftj-bfs", "ws-dro", "batch",    "ooc",
                                "lftj-ooc", "tpt",      "open-addr", "pw"};
  const uint32_t npaths = sizeof(paths) / sizeof(paths[0]);
  log_info("Engine: %s", res.path < npaths ? paths[res.path] : "?");
  log_info("Total number of outputs: %llu", (unsigned long long)res.count);
  log_info("---------------------------------");
  log_info("Result checking...");
  if (!count_only) {
    char hex[33];
    if (amj_result_md5(ctx, &res, hex) != AMJ_OK) die("fingerprint failed: %s", amj_last_error(ctx));
    log_info("Md5sum of result histogram: %s", hex);
  }
  log_info("checksum=0x%016llx", (unsigned long long)res.checksum);
  log_info("---------------------------------");
  log_info("Time and Mem stat...");
  log_info("Build kernel time: %.2f ms", res.build_ms);
  log_info("Probe kernel time: %.2f ms", res.probe_ms);
  log_info("Total kernel time: %.0f ms", res.build_ms + res.probe_ms);
  log_info("Total CPU time: %.2f ms", wall.ms());
  amj_result_free(ctx, &res);
  amj_ctx_destroy(ctx);
  return 0;
}
