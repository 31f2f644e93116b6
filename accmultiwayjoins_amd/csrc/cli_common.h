// Shared pieces of the exec-AMHJ / exec-LFTJ command-line drivers: logging in the reference's
// "INFO ... message" style, directory listing in readdir order, and .db loading.
#pragma once

#include <dirent.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "amj.h"

namespace cli {

inline void log_info(const char* fmt, ...) {
  std::printf("INFO  ");
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::printf("\n");
  std::fflush(stdout);
}

[[noreturn]] inline void die(const char* fmt, ...) {
  std::fprintf(stderr, "ERROR ");
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
  std::fprintf(stderr, "\n");
  std::exit(1);
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Directory entries in readdir order (list_files, common-utils/helper.h:112-127).
inline std::vector<std::string> list_files(const std::string& dir) {
  std::vector<std::string> out;
  DIR* d = opendir(dir.c_str());
  if (!d) die("Cannot open the directory: %s", dir.c_str());
  while (dirent* e = readdir(d)) {
    const std::string n = e->d_name;
    if (n != "." && n != "..") out.push_back(n);
  }
  closedir(d);
  return out;
}

inline bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

// A host relation read from a .db file.
struct HostRel {
  std::string path;
  uint64_t rows = 0;
  std::vector<uint32_t> attrs;
  std::vector<std::vector<int32_t>> cols;
  std::vector<const int32_t*> ptrs;

  amj_rel_t view() {
    ptrs.clear();
    for (auto& c : cols) ptrs.push_back(c.data());
    amj_rel_t r{};
    r.num_cols = (uint32_t)cols.size();
    r.rows = rows;
    r.cols = ptrs.data();
    r.attr_ids = attrs.data();
    r.on_device = 0;
    return r;
  }
};

inline HostRel load_db(const std::string& path) {
  HostRel r;
  r.path = path;
  uint32_t nc = 0;
  uint32_t ids[8];
  if (amj_db_parse_name(path.c_str(), &r.rows, &nc, ids, 8) != AMJ_OK) die("%s", amj_file_error());
  if (nc > 8) die("too many columns in %s", path.c_str());
  r.attrs.assign(ids, ids + nc);
  r.cols.assign(nc, std::vector<int32_t>(r.rows));
  std::vector<int32_t*> p;
  for (auto& c : r.cols) p.push_back(c.data());
  Timer t;
  if (amj_db_read(path.c_str(), r.rows, nc, p.data()) != AMJ_OK) die("%s", amj_file_error());
  log_info("Read file: %s, size: %.1f MB, time: %.2f s.", path.c_str(),
           (double)r.rows * nc * 4 / 1048576.0, t.ms() / 1000.0);
  return r;
}

inline amj_ctx_t* open_ctx() {
  amj_ctx_t* ctx = nullptr;
  const int rc = amj_ctx_create(0, &ctx);
  if (rc != AMJ_OK) die("no usable GPU (amj_ctx_create returned %d)", rc);
  char name[256];
  uint64_t hbm = 0;
  int cus = 0;
  amj_device_info(ctx, name, sizeof(name), &hbm, &cus);
  log_info("Device: %s, %d CUs, %.1f GB", name, cus, (double)hbm / 1e9);
  return ctx;
}

inline bool has_flag(int argc, char** argv, int from, const char* flag) {
  for (int i = from; i < argc; ++i)
    if (std::string(argv[i]) == flag) return true;
  return false;
}

}  // namespace cli
