// exec-LFTJ: the reference's leapfrog triejoin command line on libamj.
//
//   exec-LFTJ <dataset-path> <0=BFS|1=DFS> <ooc 0/1> <ws 0/1> [--count-only]
//
// Reference: main / query_processing (LFTJ-GPU/exec_LFTJ.cu:62-267). The directory holds .db
// relation files and/or .tr trie files (a .tr file is used instead of the .db of the same
// name). One dataset: SELF mode, the triangle query over one trie aliased as R(0,1), S(1,2),
// T(0,2) (:198-222). Three or more: MULTI_WAY, one trie per dataset whose levels are its
// attributes in increasing id order (:145-156), attribute order 0, 1, ... (:230-231). Two
// datasets are rejected as in the reference. The count is logged as "Output count"
// (ALFTJ.cuh:1248); DFS also materialises the result unless --count-only is given.
#include <algorithm>
#include <map>

#include "cli_common.h"

using namespace cli;

int main(int argc, char** argv) {
  if (argc < 5) die("usage: exec-LFTJ <dataset-path> <0=BFS|1=DFS> <ooc> <ws> [--count-only]");
  const std::string dir = argv[1];
  const int algo = std::atoi(argv[2]);
  if (algo != ALFTJ_BFS && algo != ALFTJ_DFS) die("Unsupported algorithm type");
  amj_opts_t o;
  amj_default_opts(&o);
  o.ooc = std::atoi(argv[3]);
  o.ws = std::atoi(argv[4]);
  const bool count_only = has_flag(argc, argv, 5, "--count-only");
  o.materialize = (algo == ALFTJ_DFS && !count_only) ? 1 : 0;
  log_info("LFTJ %s, ooc=%d, ws=%d", algo == ALFTJ_DFS ? "DFS" : "BFS", o.ooc, o.ws);

  // datasets: name without suffix -> has a .tr file
  std::vector<std::string> names;
  std::map<std::string, bool> is_trie;
  for (const auto& f : list_files(dir)) {
    if (!ends_with(f, ".db") && !ends_with(f, ".tr")) continue;
    const std::string base = f.substr(0, f.size() - 3);
    if (!is_trie.count(base)) names.push_back(base);
    is_trie[base] = is_trie[base] || ends_with(f, ".tr");
  }
  if (names.empty() || names.size() == 2) die("Wrong number of datasets: %zu", names.size());
  const bool self = names.size() == 1;

  Timer wall;
  amj_ctx_t* ctx = open_ctx();
  std::vector<amj_trie_t*> tries;
  std::vector<uint32_t> trie_attrs;
  std::map<uint32_t, bool> attrs;
  double build_ms = 0.0;
  for (const auto& nm : names) {
    uint64_t rows = 0;
    uint32_t nc = 0, ids[8];
    if (amj_db_parse_name((nm + ".db").c_str(), &rows, &nc, ids, 8) != AMJ_OK)
      die("%s", amj_file_error());
    if (nc != 2) die("%s: tries of %u levels are not supported (two only)", nm.c_str(), nc);
    amj_trie_t* t = nullptr;
    if (is_trie[nm]) {
      log_info("Read from Trie file: %s/%s.tr", dir.c_str(), nm.c_str());
      if (alftj_trie_load(ctx, (dir + "/" + nm + ".tr").c_str(), &t) != AMJ_OK)
        die("%s", amj_last_error(ctx));
    } else {
      HostRel r = load_db(dir + "/" + nm + ".db");
      if (r.attrs[0] > r.attrs[1]) {  // levels follow increasing attribute ids
        std::swap(r.attrs[0], r.attrs[1]);
        std::swap(r.cols[0], r.cols[1]);
      }
      amj_rel_t v = r.view();
      if (alftj_trie_build(ctx, &v, &t) != AMJ_OK) die("%s", amj_last_error(ctx));
    }
    uint64_t n0 = 0, n1 = 0, mx = 0;
    double ms = 0.0;
    alftj_trie_info(ctx, t, &n0, &n1, &mx, &ms);
    build_ms += ms;
    log_info("Trie %s: %llu level-0 values, %llu children, longest list %llu", nm.c_str(),
             (unsigned long long)n0, (unsigned long long)n1, (unsigned long long)mx);
    tries.push_back(t);
    const uint32_t a = std::min(ids[0], ids[1]), b = std::max(ids[0], ids[1]);
    trie_attrs.push_back(a);
    trie_attrs.push_back(b);
    attrs[a] = attrs[b] = true;
  }
  std::vector<uint32_t> order;
  std::vector<amj_trie_t*> use;
  if (self) {  // triangle: Trie[0] as (0,1), (1,2), (0,2)
    use = {tries[0], tries[0], tries[0]};
    trie_attrs = {0, 1, 1, 2, 0, 2};
    order = {0, 1, 2};
  } else {
    use = tries;
    for (uint32_t i = 0; i < attrs.size(); ++i) order.push_back(i);
  }
  log_info("Read file & Trie construction kernel time: %.2f ms", build_ms);
  log_info("Query: #datasets=%zu, #join tables=%zu, #join attrs=%zu", names.size(), use.size(),
           order.size());
  amj_result_t res;
  const int rc = alftj_run_aliased(ctx, use.data(), (uint32_t)use.size(), trie_attrs.data(),
                                   order.data(), (uint32_t)order.size(), algo, &o, &res);
  if (rc != AMJ_OK) die("LFTJ failed (%d): %s", rc, amj_last_error(ctx));
  log_info("Output count: %llu.", (unsigned long long)res.count);
  log_info("checksum=0x%016llx", (unsigned long long)res.checksum);
  log_info("LFTJ Total kernel time: %.2f ms", res.probe_ms);
  log_info("Total wall time: %.2f s", wall.ms() / 1000.0);
  amj_result_free(ctx, &res);
  for (auto* t : tries) alftj_trie_destroy(ctx, t);
  amj_ctx_destroy(ctx);
  return 0;
}
