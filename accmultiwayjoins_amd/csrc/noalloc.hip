// amj_execute_noalloc: the count + checksum of amj_execute without any device allocation.
//
// It runs the plain engine's one-pass count-and-hash launch (probe_write_kernel in MODE_HASH,
// the count + checksum path of amj_execute) and the two reductions of its per-block partials,
// with every temporary in one fixed device array instead of the context's scratch buffers:
// per-block checksum partials, per-block counts, the reduction's tile partials and the two
// totals. The array is file-scope device memory, so each GPU the library runs on holds it until
// the process ends; its address on a device comes from hipGetSymbolAddress, taken once per
// device with that device current (in host code the array's name is not a device address).
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU race. Nothing in
// the array has to start at zero: every block writes its partials and the reductions write
// their totals.
#include <cstring>
#include <mutex>

#include "runtime.h"

namespace {

constexpr uint64_t kNoAllocMaxProbeRows = 1ull << 27;                          // 134,217,728
constexpr uint64_t kNoAllocMaxBlocks = ((kNoAllocMaxProbeRows + 63) / 64 + 3) / 4;  // 2^19
constexpr uint64_t kNoAllocTempWords = kNoAllocMaxBlocks / 2048 + 2;            // reduction
// layout in 64-bit words: block checksums | block counts | reduction temp | totals
constexpr uint64_t kOffSums = 0;
constexpr uint64_t kOffCounts = kOffSums + kNoAllocMaxBlocks;
constexpr uint64_t kOffTemp = kOffCounts + kNoAllocMaxBlocks;
constexpr uint64_t kOffTotals = (kOffTemp + kNoAllocTempWords + 31) / 32 * 32;
constexpr uint64_t kNoAllocWords = kOffTotals + 32;  // 8.4 MB

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t* g_ptr[kMaxDevices] = {nullptr};

}  // namespace

static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

// Device address of this device's copy of `scratch`, looked up once per device.
static int noalloc_scratch(amj_ctx_t* ctx, uint64_t** out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) {
    amj::set_error(ctx, AMJ_ERR_ARG, "amj_execute_noalloc: device id out of range");
    return AMJ_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      amj::set_error(ctx, AMJ_ERR_HIP,
                     std::string("amj_execute_noalloc: hipGetSymbolAddress failed: ") +
                         hipGetErrorString(e));
      return AMJ_ERR_HIP;
    }
    g_ptr[dev] = (uint64_t*)p;
  }
  *out = g_ptr[dev];
  return AMJ_OK;
}

int amj_execute_noalloc(amj_ctx_t* ctx, const amj_relation_t* probe,
                        amj_hash_table_t* const* hts, uint32_t n_hts, const amj_opts_t* opts,
                        amj_result_t* res) {
  if (!ctx || !probe || !hts || !res) return AMJ_ERR_ARG;
  if (opts && (opts->materialize || opts->ooc || opts->ws || opts->dro || opts->fib ||
               opts->engine != AMJ_ENGINE_WAVE)) {
    amj::set_error(ctx, AMJ_ERR_ARG,
                   "amj_execute_noalloc computes count + checksum only: materialize, ooc, ws, "
                   "dro and fib must be 0 and the engine AMJ_ENGINE_WAVE");
    return AMJ_ERR_ARG;
  }
  amj::Plan P;
  std::vector<uint32_t> ids;
  AMJ_TRY(amj::make_execute_plan(ctx, probe, hts, n_hts, &P, &ids));
  if (P.probe_on_host) {
    amj::set_error(ctx, AMJ_ERR_ARG, "amj_execute_noalloc needs a device-resident probe");
    return AMJ_ERR_ARG;
  }
  if (P.probe_n > kNoAllocMaxProbeRows) {
    amj::set_error(ctx, AMJ_ERR_ARG,
                   "amj_execute_noalloc: probe relation of " + std::to_string(P.probe_n) +
                       " rows exceeds its limit of " + std::to_string(kNoAllocMaxProbeRows) +
                       " rows");
    return AMJ_ERR_ARG;
  }
  const uint64_t nb = amj::probe_write_blocks(P);
  if (nb > kNoAllocMaxBlocks) {  // implied by the row limit; checked before any launch
    amj::set_error(ctx, AMJ_ERR_STATE, "amj_execute_noalloc: block partials exceed the array");
    return AMJ_ERR_STATE;
  }
  uint64_t* base = nullptr;
  AMJ_TRY(noalloc_scratch(ctx, &base));
  std::memset(res, 0, sizeof(*res));
  res->num_cols = (uint32_t)P.n_attrs;
  for (int a = 0; a < P.n_attrs; ++a) res->attr_ids[a] = ids[a];
  hipStream_t s = ctx->stream;
  uint64_t* sums = base + kOffSums;
  uint64_t* counts = base + kOffCounts;
  uint64_t* temp = base + kOffTemp;
  uint64_t* totals = base + kOffTotals;
  uint64_t nb_used = 0;
  uint64_t host[2] = {0, 0};
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  AMJ_HIP_TRY(ctx, amj::launch_probe_write(P, nullptr, amj::ProbeOut{}, true, sums, &nb_used, s,
                                           counts));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  AMJ_HIP_TRY(ctx, amj::reduce_sum_u64(sums, nb_used, totals + 0, temp, s));
  AMJ_HIP_TRY(ctx, amj::reduce_sum_u64(counts, nb_used, totals + 1, temp, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(host, totals, sizeof(host), hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  res->checksum = host[0];
  res->count = host[1];
  res->probe_ms = amj::elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = amj::elapsed_ms(ctx->ev[2], ctx->ev[4]);
  res->probe_kernel_launches = 1;
  res->rows_written = 0;
  res->path = AMJ_PATH_PLAIN;
  res->passes = 1;
  return AMJ_OK;
}
