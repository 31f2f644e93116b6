// Device-wide primitives written for gfx950: exclusive scan (reduce-then-scan), u64
// reduction and a stable LSD radix sort of 64-bit keys with wave64 ballot multisplit.
//
// These replace the reference's uses of cub::DeviceScan::ExclusiveSum, DeviceReduce::Sum and
// DeviceRadixSort::SortPairs (common-utils/cuda/primitives.cuh:258-510); the algorithms
// are the textbook ones (Merrill & Garland's reduce-then-scan, LSD radix sort), written
// from scratch for 64-lane waves.
#include "amj_internal.h"

namespace amj {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048 elements per workgroup

size_t scan_temp_elems(uint64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T y = __shfl_up(x, (unsigned)d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

// Block-wide exclusive scan of one value per thread; returns the block total too.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* total, T* lds_wave) {
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  T incl = wave_incl_scan<T>(v);
  if (lane == kWave - 1) lds_wave[w] = incl;
  __syncthreads();
  T wave_off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; ++i) {
    T s = lds_wave[i];
    if (i < w) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + incl - v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scan_reduce_kernel(const T* __restrict__ in,
                                                             uint64_t n, T* __restrict__ part) {
  __shared__ T lds[kWavesPerBlock];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  T sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) sum += in[idx];
  }
  T total;
  (void)block_excl_scan<T>(sum, &total, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the per-tile partials, in place.
template <typename T>
__global__ __launch_bounds__(kBlock) void scan_partials_kernel(T* part, uint64_t m) {
  __shared__ T lds[kWavesPerBlock];
  T carry = 0;
  for (uint64_t base = 0; base < m; base += kScanTile) {
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + i;
      v[i] = idx < m ? part[idx] : (T)0;
      sum += v[i];
    }
    T total;
    T off = block_excl_scan<T>(sum, &total, lds) + carry;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + i;
      if (idx < m) part[idx] = off;
      off += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) part[m] = carry;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scan_down_kernel(const T* __restrict__ in, T* out,
                                                           uint64_t n,
                                                           const T* __restrict__ part) {
  __shared__ T lds[kWavesPerBlock];
  __shared__ T tile[kScanTile];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  // striped coalesced load into LDS, blocked scan out of LDS
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    tile[i * kBlock + threadIdx.x] = idx < n ? in[idx] : (T)0;
  }
  __syncthreads();
  T v[kScanItems];
  T sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = tile[threadIdx.x * kScanItems + i];
    sum += v[i];
  }
  T total;
  T off = block_excl_scan<T>(sum, &total, lds) + part[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    tile[threadIdx.x * kScanItems + i] = off;
    off += v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) out[idx] = tile[i * kBlock + threadIdx.x];
  }
}

template <typename T>
static hipError_t exclusive_scan_t(const T* in, T* out, uint64_t n, T* temp, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_reduce_kernel<T>, dim3((unsigned)tiles), dim3(kBlock), 0, s, in, n,
                     temp);
  hipLaunchKernelGGL(scan_partials_kernel<T>, dim3(1), dim3(kBlock), 0, s, temp, tiles);
  hipLaunchKernelGGL(scan_down_kernel<T>, dim3((unsigned)tiles), dim3(kBlock), 0, s, in, out,
                     n, temp);
  return hipGetLastError();
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp,
                              hipStream_t s) {
  return exclusive_scan_t<uint32_t>(in, out, n, temp, s);
}
hipError_t exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* temp,
                              hipStream_t s) {
  return exclusive_scan_t<uint64_t>(in, out, n, temp, s);
}

hipError_t reduce_sum_u64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* temp,
                          hipStream_t s) {
  if (n == 0) return hipMemsetAsync(out, 0, sizeof(uint64_t), s);
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_reduce_kernel<uint64_t>, dim3((unsigned)tiles), dim3(kBlock), 0, s,
                     in, n, temp);
  hipLaunchKernelGGL(scan_partials_kernel<uint64_t>, dim3(1), dim3(kBlock), 0, s, temp, tiles);
  return hipMemcpyAsync(out, temp + tiles, sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
}

// ---------------------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------------------
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;  // 4096 keys per workgroup
constexpr int kRadix = 256;

static inline uint64_t sort_tiles(uint64_t n) { return (n + kSortTile - 1) / kSortTile; }

size_t radix_sort_temp_bytes(uint64_t n) {
  const uint64_t tiles = sort_tiles(n);
  const uint64_t hist = (uint64_t)kRadix * tiles;
  return 256 + sizeof(uint32_t) * (hist + 1) + sizeof(uint32_t) * scan_temp_elems(hist) + 256;
}

// diff[0] = OR over all keys of (key ^ key[0]): the bits in which the keys differ;
// diff[1] = 1 when some key is greater than the next (the keys are not in order).
__global__ __launch_bounds__(kBlock) void key_diff_kernel(const uint64_t* __restrict__ keys,
                                                          uint64_t n, uint64_t* diff) {
  const uint64_t k0 = keys[0];
  uint64_t acc = 0;
  uint32_t down = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t k = keys[i];
    acc |= k ^ k0;
    if (i + 1 < n && k > keys[i + 1]) down = 1;
  }
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) acc |= __shfl_xor(acc, d, kWave);
  const bool any_down = __ballot(down != 0) != 0ull;
  __shared__ uint64_t part[kWavesPerBlock];
  __shared__ uint32_t part_down[kWavesPerBlock];
  if (lane_id() == 0) {
    part[threadIdx.x / kWave] = acc;
    part_down[threadIdx.x / kWave] = any_down ? 1u : 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t v = 0;
    uint32_t dn = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) {
      v |= part[i];
      dn |= part_down[i];
    }
    if (v) atomicOr((unsigned long long*)diff, (unsigned long long)v);
    if (dn) atomicOr((unsigned long long*)diff + 1, 1ull);
  }
}

__global__ __launch_bounds__(kBlock) void radix_hist_kernel(const uint64_t* __restrict__ keys,
                                                            uint64_t n, int shift,
                                                            uint32_t* __restrict__ hist,
                                                            uint64_t tiles) {
  __shared__ uint32_t cnt[kRadix];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortItems; ++i) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[(uint32_t)(keys[idx] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of one 4096-key tile. Wave w owns the tile's keys [w*1024, (w+1)*1024), 64
// per round, so (wave, round, lane) is the input order. Each wave ranks its keys with a ballot
// multisplit and a wave-private running count per digit (no block barrier between rounds);
// the keys are then placed in digit order in LDS and leave in that order, so consecutive
// lanes write consecutive slots of a bucket's run instead of 8-byte writes all over the
// output.
__global__ __launch_bounds__(kBlock) void radix_scatter_kernel(
    const uint64_t* __restrict__ keys, uint64_t* __restrict__ out, uint64_t n, int shift,
    const uint32_t* __restrict__ hist_scanned, uint64_t tiles) {
  constexpr int kPerWave = kSortTile / kWavesPerBlock;
  __shared__ uint64_t tile[kSortTile];
  __shared__ uint32_t wrun[kWavesPerBlock][kRadix];  // per-wave digit counts, then offsets
  __shared__ uint32_t loff[kRadix];                  // first slot of each digit in the tile
  __shared__ uint32_t goff[kRadix];                  // first output row of each digit
  __shared__ uint32_t scan_lds[kWavesPerBlock];
  const int tid = threadIdx.x;
  const int w = tid / kWave;
  const int lane = lane_id();
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; ++i) wrun[i][tid] = 0;
  goff[tid] = hist_scanned[(uint64_t)tid * tiles + blockIdx.x];
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * kSortTile;
  const uint64_t base = t0 + (uint64_t)w * kPerWave;
  uint64_t key[kSortItems];
  uint32_t rk[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const uint64_t idx = base + (uint64_t)r * kWave + lane;
    key[r] = idx < n ? keys[idx] : 0ull;
  }
  uint32_t* run = wrun[w];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = base + (uint64_t)r * kWave + lane < n;
    const uint32_t d = (uint32_t)(key[r] >> shift) & 0xffu;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t rank = mbcnt(m);
    __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
    const uint32_t before = run[d];
    rk[r] = before + rank;
    __asm__ __volatile__("" ::: "memory");
    if (valid && rank == 0) run[d] = before + (uint32_t)__popcll(m);
    __asm__ __volatile__("" ::: "memory");
  }
  __syncthreads();
  {  // digit tid: tile offset, then each wave's first slot for it
    uint32_t c[kWavesPerBlock], tot = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) {
      c[i] = wrun[i][tid];
      tot += c[i];
    }
    uint32_t all;
    uint32_t p = block_excl_scan<uint32_t>(tot, &all, scan_lds);
    loff[tid] = p;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) {
      wrun[i][tid] = p;
      p += c[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (base + (uint64_t)r * kWave + lane < n) {
      const uint32_t d = (uint32_t)(key[r] >> shift) & 0xffu;
      tile[run[d] + rk[r]] = key[r];
    }
  }
  __syncthreads();
  const uint32_t tile_n = (uint32_t)(n - t0 < (uint64_t)kSortTile ? n - t0 : (uint64_t)kSortTile);
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t j = (uint32_t)i * kBlock + tid;
    if (j < tile_n) {
      const uint64_t k = tile[j];
      const uint32_t d = (uint32_t)(k >> shift) & 0xffu;
      out[(uint64_t)goff[d] + (j - loff[d])] = k;
    }
  }
}

hipError_t radix_sort_u64(uint64_t* keys, uint64_t* alt, uint64_t n, uint32_t digit_mask,
                          void* temp, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  if (n >= (1ull << 32)) return hipErrorInvalidValue;
  const uint64_t tiles = sort_tiles(n);
  const uint64_t hist_n = (uint64_t)kRadix * tiles;
  char* t = (char*)temp;
  uint64_t* diff = (uint64_t*)t;
  uint32_t* hist = (uint32_t*)(t + 256);
  uint32_t* scan_tmp = hist + hist_n + 1;

  hipError_t e = hipMemsetAsync(diff, 0, 2 * sizeof(uint64_t), s);
  if (e != hipSuccess) return e;
  const unsigned grid = (unsigned)std::min<uint64_t>((n + kBlock - 1) / kBlock, 1024);
  hipLaunchKernelGGL(key_diff_kernel, dim3(grid), dim3(kBlock), 0, s, keys, n, diff);
  uint64_t hd[2] = {0, 0};
  e = hipMemcpyAsync(hd, diff, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const uint64_t hdiff = hd[0];
  // Keys already in order whose differing bits all lie in sorted digits: the stable sort is the
  // identity (an edge list stored as sorted CSR, for instance, skips all its passes).
  uint64_t sorted_bits = 0;
  for (int d = 0; d < 8; ++d)
    if ((digit_mask >> d) & 1u) sorted_bits |= 0xffull << (8 * d);
  if (hd[1] == 0 && (hdiff & ~sorted_bits) == 0) return hipSuccess;

  uint64_t* src = keys;
  uint64_t* dst = alt;
  for (int d = 0; d < 8; ++d) {
    if (!((digit_mask >> d) & 1u)) continue;
    if (((hdiff >> (8 * d)) & 0xffull) == 0) continue;  // digit constant: pass is identity
    const int shift = 8 * d;
    hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)tiles), dim3(kBlock), 0, s, src, n,
                       shift, hist, tiles);
    e = exclusive_scan_u32(hist, hist, hist_n, scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)tiles), dim3(kBlock), 0, s, src,
                       dst, n, shift, hist, tiles);
    std::swap(src, dst);
  }
  if (src != keys)
    e = hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// Fill-up after a chunked direct pass of n results whose chunks end at `reserved`: the rows at
// [n, reserved) outside the chunks' unused tails move, in order, to the tail slots below n, so
// the n results are dense. tail[c] is chunk c's unused tail (its last tail[c] rows) and tpre
// its exclusive prefix sum, for the reserved / chunk chunks. Reads and writes never overlap
// (sources lie at or past n, destinations below it).
__global__ __launch_bounds__(kBlock) void chunk_fill_kernel(const ProbeOut out, int n_attrs,
                                                            const uint64_t* __restrict__ tail,
                                                            const uint64_t* __restrict__ tpre,
                                                            uint64_t n_chunks, uint32_t chunk,
                                                            uint64_t n, uint64_t reserved) {
  // tail slots below n: every tail of the chunks before n's chunk, and the part of that
  // chunk's tail below n
  const uint64_t cn = n / chunk;
  uint64_t below = cn < n_chunks ? tpre[cn] : tpre[n_chunks - 1] + tail[n_chunks - 1];
  if (cn < n_chunks) {
    const uint64_t keep = chunk - tail[cn], o = n % chunk;
    below += o > keep ? o - keep : 0;
  }
  for (uint64_t p = n + (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < reserved;
       p += (uint64_t)gridDim.x * kBlock) {
    const uint64_t c = p / chunk, o = p % chunk;
    if (o >= chunk - tail[c]) continue;  // p is in a tail
    // results in [n, p): positions minus the tail slots among them
    const uint64_t r = (p - n) - (tpre[c] - below);
    // the r-th tail slot: chunk c2 = last chunk with tpre[c2] <= r
    uint64_t lo = 0, hi = cn < n_chunks ? cn + 1 : n_chunks;
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (tpre[mid] <= r) lo = mid;
      else hi = mid;
    }
    const uint64_t dst = lo * chunk + (chunk - tail[lo]) + (r - tpre[lo]);
    for (int a = 0; a < n_attrs; ++a) out.cols[a][dst] = out.cols[a][p];
  }
}

hipError_t launch_chunk_fill(const ProbeOut& out, int n_attrs, const uint64_t* tail,
                             const uint64_t* tpre, uint64_t n_chunks, uint32_t chunk, uint64_t n,
                             uint64_t reserved, hipStream_t s) {
  if (reserved <= n || n_chunks == 0 || chunk == 0) return hipSuccess;
  uint64_t g = (reserved - n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(chunk_fill_kernel, dim3((unsigned)g), dim3(kBlock), 0, s, out, n_attrs,
                     tail, tpre, n_chunks, chunk, n, reserved);
  return hipGetLastError();
}

}  // namespace amj
