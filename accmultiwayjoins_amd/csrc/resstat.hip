// Result fingerprint of the reference: res_stat (common-utils/res_stat.h:32-133) histograms the
// low 20 bits of every value of every result column, prints the histogram up to the largest
// value seen as "[h0, h1, ...]\n" and logs the MD5 of that text. The histogram runs here on the
// device (one atomic per value into 2^20 u64 bins); the text and the digest are built on the host.
#include <cstdio>
#include <cstring>

#include "md5.h"
#include "runtime.h"

namespace amj {

constexpr uint32_t kValBits = 20;  // MAX_VAL_BITS, res_stat.h:16
constexpr uint32_t kBins = 1u << kValBits;

__global__ __launch_bounds__(kBlock) void res_hist_kernel(const int32_t* __restrict__ col,
                                                          uint64_t n,
                                                          unsigned long long* __restrict__ hist) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    atomicAdd(&hist[(uint32_t)col[i] & (kBins - 1)], 1ull);
}

// "[h0, h1, ..., hmax]\n" (the prettyprint format of a std::vector) and its MD5
static std::string digest_of_histogram(const uint64_t* hist) {
  uint32_t last = 0;
  for (uint32_t v = 0; v < kBins; ++v)
    if (hist[v]) last = v;
  Md5 md5;
  std::string chunk;
  chunk.reserve(1 << 16);
  chunk.push_back('[');
  char num[32];
  for (uint32_t v = 0; v <= last; ++v) {
    const int k = std::snprintf(num, sizeof(num), v ? ", %llu" : "%llu", (unsigned long long)hist[v]);
    chunk.append(num, (size_t)k);
    if (chunk.size() > 60000) {
      md5.update(chunk.data(), chunk.size());
      chunk.clear();
    }
  }
  chunk.append("]\n");
  md5.update(chunk.data(), chunk.size());
  return md5.hexdigest();
}

int result_md5(amj_ctx_t* ctx, const int32_t* const* dcols, uint32_t ncols, uint64_t rows,
               std::string* hex) {
  hipStream_t s = ctx->stream;
  void* h = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_MISC3, (size_t)kBins * 8, &h));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(h, 0, (size_t)kBins * 8, s));
  uint64_t g = (rows + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  for (uint32_t c = 0; c < ncols && rows > 0; ++c)
    hipLaunchKernelGGL(res_hist_kernel, dim3((unsigned)g), dim3(kBlock), 0, s, dcols[c], rows,
                       (unsigned long long*)h);
  AMJ_HIP_TRY(ctx, hipGetLastError());
  std::vector<uint64_t> host(kBins);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(host.data(), h, (size_t)kBins * 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  *hex = digest_of_histogram(host.data());
  return AMJ_OK;
}

}  // namespace amj

using namespace amj;

extern "C" {

int amj_result_md5(amj_ctx_t* ctx, const amj_result_t* res, char* hex_out) {
  if (!ctx || !res || !hex_out) return AMJ_ERR_ARG;
  if (res->count && res->rows_written != res->count) {
    set_error(ctx, AMJ_ERR_STATE, "the fingerprint needs a materialised result");
    return AMJ_ERR_STATE;
  }
  std::string hex;
  AMJ_TRY(result_md5(ctx, res->cols, res->num_cols, res->rows_written, &hex));
  std::memcpy(hex_out, hex.c_str(), 33);
  return AMJ_OK;
}

int amj_res_stat_md5(const int32_t* const* cols, uint32_t num_cols, uint64_t rows,
                     char* hex_out) {
  if ((!cols && num_cols) || !hex_out) return AMJ_ERR_ARG;
  std::vector<uint64_t> hist(kBins, 0);
  for (uint32_t c = 0; c < num_cols; ++c)
    for (uint64_t i = 0; i < rows; ++i) ++hist[(uint32_t)cols[c][i] & (kBins - 1)];
  const std::string hex = digest_of_histogram(hist.data());
  std::memcpy(hex_out, hex.c_str(), 33);
  return AMJ_OK;
}

}  // extern "C"
