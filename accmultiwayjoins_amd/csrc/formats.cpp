// On-disk formats of the reference: .db relation files and .tr trie files.
//
// .db (read_rel_cols_mmap, common-utils/file_io.h:90-121; names parsed by
// QueryProcessor::load_multi_rels, MHJ-GPU/QueryProcessor.cuh:41-73): the file name is
// <rows>_<attr0>_<attr1>...db and the content is column 0, then column 1, ..., each `rows`
// little-endian int32 values.
//
// .tr (IndexedTrie::serialization / init_with_deserialization,
// LFTJ-GPU/IndexedTrie.cuh:122-197): u32 num_attrs, u64 data_len[num_attrs],
// u32 attr_list[num_attrs], then the int32 values of every level, then the u64 offsets
// (data_len[l] + 1 of them) of every level but the last.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lftj.h"
#include "runtime.h"

namespace amj {

namespace {

thread_local std::string g_file_error;  // message of the last failed context-free call

int file_error(amj_ctx_t* ctx, int code, const std::string& msg) {
  if (ctx) set_error(ctx, code, msg);
  else g_file_error = msg;
  return code;
}

std::string base_name(const char* path) {
  std::string p(path);
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? p : p.substr(s + 1);
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool read_all(FILE* f, void* p, size_t bytes) {
  return bytes == 0 || std::fread(p, 1, bytes, f) == bytes;
}
bool write_all(FILE* f, const void* p, size_t bytes) {
  return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes;
}

}  // namespace

const char* last_file_error() { return g_file_error.c_str(); }

}  // namespace amj

using namespace amj;

extern "C" {

const char* amj_file_error(void) { return last_file_error(); }

int amj_db_parse_name(const char* path, uint64_t* rows, uint32_t* num_cols, uint32_t* attr_ids,
                      uint32_t max_cols) {
  if (!path || !rows || !num_cols) return AMJ_ERR_ARG;
  std::string name = base_name(path);
  if (name.size() < 4 || name.compare(name.size() - 3, 3, ".db") != 0)
    return file_error(nullptr, AMJ_ERR_ARG, "not a .db file name: " + name);
  name.resize(name.size() - 3);
  std::vector<std::string> parts;
  size_t start = 0;
  for (;;) {
    const size_t u = name.find('_', start);
    parts.push_back(name.substr(start, u == std::string::npos ? std::string::npos : u - start));
    if (u == std::string::npos) break;
    start = u + 1;
  }
  auto parse_u64 = [](const std::string& s, uint64_t* v) {
    if (s.empty() || s.size() > 19) return false;
    uint64_t x = 0;
    for (char c : s) {
      if (c < '0' || c > '9') return false;
      x = x * 10 + (uint64_t)(c - '0');
    }
    *v = x;
    return true;
  };
  uint64_t n = 0;
  if (parts.size() < 2 || !parse_u64(parts[0], &n))
    return file_error(nullptr, AMJ_ERR_ARG, "expected <rows>_<attr0>[_<attr1>...].db: " + name);
  const uint32_t nc = (uint32_t)parts.size() - 1;
  for (uint32_t c = 0; c < nc; ++c) {
    uint64_t a = 0;
    if (!parse_u64(parts[c + 1], &a) || a > 0xffffffffull)
      return file_error(nullptr, AMJ_ERR_ARG, "bad attribute id in " + name);
    if (attr_ids && c < max_cols) attr_ids[c] = (uint32_t)a;
  }
  *rows = n;
  *num_cols = nc;
  return AMJ_OK;
}

int amj_db_read(const char* path, uint64_t rows, uint32_t num_cols, int32_t* const* cols) {
  if (!path || (!cols && num_cols)) return AMJ_ERR_ARG;
  File f;
  f.f = std::fopen(path, "rb");
  if (!f.f) return file_error(nullptr, AMJ_ERR_ARG, std::string("cannot open ") + path + ": " +
                                                         std::strerror(errno));
  for (uint32_t c = 0; c < num_cols; ++c)
    if (!read_all(f.f, cols[c], rows * 4))
      return file_error(nullptr, AMJ_ERR_ARG, std::string("short read in ") + path);
  return AMJ_OK;
}

int amj_db_write(const char* dir, const amj_rel_t* rel, char* path_out, size_t path_len) {
  if (!dir || !rel || !rel->cols || rel->num_cols == 0 || rel->on_device) return AMJ_ERR_ARG;
  std::string name = std::string(dir) + "/" + std::to_string(rel->rows);
  for (uint32_t c = 0; c < rel->num_cols; ++c)
    name += "_" + std::to_string(rel->attr_ids ? rel->attr_ids[c] : c);
  name += ".db";
  File f;
  f.f = std::fopen(name.c_str(), "wb");
  if (!f.f) return file_error(nullptr, AMJ_ERR_ARG, "cannot create " + name + ": " +
                                                         std::strerror(errno));
  for (uint32_t c = 0; c < rel->num_cols; ++c)
    if (!write_all(f.f, rel->cols[c], rel->rows * 4))
      return file_error(nullptr, AMJ_ERR_ARG, "short write to " + name);
  if (path_out && path_len) std::snprintf(path_out, path_len, "%s", name.c_str());
  return AMJ_OK;
}

int alftj_trie_save(amj_ctx_t* ctx, const amj_trie_t* t, const char* path) {
  if (!ctx || !t || !path) return AMJ_ERR_ARG;
  const uint64_t n0 = t->dev.n0, n1 = t->dev.n1;
  std::vector<int32_t> d0(n0), d1(n1);
  std::vector<uint32_t> off32(n0 + 1);
  hipStream_t s = ctx->stream;
  if (n0) AMJ_HIP_TRY(ctx, hipMemcpyAsync(d0.data(), t->dev.view.d0, n0 * 4, hipMemcpyDefault, s));
  if (n1) AMJ_HIP_TRY(ctx, hipMemcpyAsync(d1.data(), t->dev.view.d1, n1 * 4, hipMemcpyDefault, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(off32.data(), t->dev.view.off, (n0 + 1) * 4, hipMemcpyDefault, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  std::vector<uint64_t> off(off32.begin(), off32.end());
  File f;
  f.f = std::fopen(path, "wb");
  if (!f.f) return file_error(ctx, AMJ_ERR_ARG, std::string("cannot create ") + path);
  const uint32_t na = 2;
  const uint64_t len[2] = {n0, n1};
  const uint32_t attrs[2] = {t->attr_ids[0], t->attr_ids[1]};
  if (!write_all(f.f, &na, 4) || !write_all(f.f, len, 16) || !write_all(f.f, attrs, 8) ||
      !write_all(f.f, d0.data(), n0 * 4) || !write_all(f.f, d1.data(), n1 * 4) ||
      !write_all(f.f, off.data(), (n0 + 1) * 8))
    return file_error(ctx, AMJ_ERR_ARG, std::string("short write to ") + path);
  return AMJ_OK;
}

int alftj_trie_from_csr(amj_ctx_t* ctx, uint64_t n0, const int32_t* d0, const uint64_t* off,
                        uint64_t n1, const int32_t* d1, const uint32_t* attr_ids,
                        amj_trie_t** out) {
  if (!ctx || !out || (n0 && (!d0 || !off)) || (n1 && !d1)) return AMJ_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<amj_trie> t(new amj_trie());
  if (attr_ids) {
    t->attr_ids[0] = attr_ids[0];
    t->attr_ids[1] = attr_ids[1];
  }
  AMJ_TRY(trie_from_host(ctx, n0, d0, off, n1, d1, &t->dev));
  *out = t.release();
  return AMJ_OK;
}

int alftj_trie_load(amj_ctx_t* ctx, const char* path, amj_trie_t** out) {
  if (!ctx || !path || !out) return AMJ_ERR_ARG;
  *out = nullptr;
  File f;
  f.f = std::fopen(path, "rb");
  if (!f.f) return file_error(ctx, AMJ_ERR_ARG, std::string("cannot open ") + path);
  uint32_t na = 0;
  if (!read_all(f.f, &na, 4)) return file_error(ctx, AMJ_ERR_ARG, "truncated .tr header");
  if (na != 2)
    return file_error(ctx, AMJ_ERR_ARG,
                      "only two-level tries are supported (file has " + std::to_string(na) + ")");
  uint64_t len[2];
  uint32_t attrs[2];
  if (!read_all(f.f, len, 16) || !read_all(f.f, attrs, 8))
    return file_error(ctx, AMJ_ERR_ARG, "truncated .tr header");
  if (len[0] >= (1ull << 32) || len[1] >= (1ull << 32))
    return file_error(ctx, AMJ_ERR_ARG, "trie levels must hold fewer than 2^32 values");
  std::vector<int32_t> d0(len[0]), d1(len[1]);
  std::vector<uint64_t> off(len[0] + 1);
  if (!read_all(f.f, d0.data(), len[0] * 4) || !read_all(f.f, d1.data(), len[1] * 4) ||
      !read_all(f.f, off.data(), (len[0] + 1) * 8))
    return file_error(ctx, AMJ_ERR_ARG, std::string("truncated .tr file ") + path);
  return alftj_trie_from_csr(ctx, len[0], d0.data(), off.data(), len[1], d1.data(), attrs, out);
}

}  // extern "C"
