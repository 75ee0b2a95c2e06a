// Device-wide exclusive scan and stable LSB radix sort for gfx950 (wave64).
// Element counts may live in device memory so a whole substep runs without a
// host round trip; grids are sized from host-side capacities and surplus
// workgroups find nothing to do.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zsmpm_device.hpp"
#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096
constexpr int kScanTopThreads = 1024;
constexpr int kScanTopItems = 16;

__device__ __forceinline__ uint32_t load_n(const uint32_t* n_dev, uint32_t n_cap) {
  if (!n_dev) return n_cap;
  const uint32_t n = *n_dev;
  return n < n_cap ? n : n_cap;
}

// inclusive scan across the wave
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive scan across the workgroup of NT threads; returns the exclusive
// prefix of `v` and the workgroup total in `total`
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total, uint32_t* lds) {
  constexpr int kWaves = NT / kWave;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == kWave - 1) lds[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < kWaves ? lds[lane] : 0u;
    const uint32_t wi = wave_incl_scan(w);
    if (lane < kWaves) lds[kWaves + lane] = wi - w;
    if (lane == kWaves - 1) lds[2 * kWaves] = wi;
  }
  __syncthreads();
  const uint32_t res = lds[kWaves + wave] + inc - v;
  total = lds[2 * kWaves];
  __syncthreads();
  return res;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                   const uint32_t* n_dev, uint32_t n_cap,
                                                                   uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t lds[2 * (kScanThreads / kWave) + 1];
  const uint32_t n = load_n(n_dev, n_cap);
  const uint32_t base = blockIdx.x * (uint32_t)kScanTile;
  uint32_t s = 0;
  if (base < n) {
    const uint32_t start = base + threadIdx.x * kScanItems;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t idx = start + i;
      if (idx < n) s += in[idx];
    }
  }
  uint32_t total;
  block_excl_scan<kScanThreads>(s, total, lds);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// single workgroup: exclusive scan of up to kScanTopThreads*kScanTopItems tile sums
__global__ __launch_bounds__(kScanTopThreads) void scan_top_kernel(uint32_t* __restrict__ tile_sums,
                                                                   uint32_t ntiles,
                                                                   uint32_t* __restrict__ total_out) {
  __shared__ uint32_t lds[2 * (kScanTopThreads / kWave) + 1];
  uint32_t vals[kScanTopItems];
  const uint32_t start = threadIdx.x * kScanTopItems;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanTopItems; ++i) {
    const uint32_t idx = start + i;
    vals[i] = idx < ntiles ? tile_sums[idx] : 0u;
    s += vals[i];
  }
  uint32_t total;
  uint32_t run = block_excl_scan<kScanTopThreads>(s, total, lds);
#pragma unroll
  for (int i = 0; i < kScanTopItems; ++i) {
    const uint32_t idx = start + i;
    if (idx < ntiles) tile_sums[idx] = run;
    run += vals[i];
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(const uint32_t* __restrict__ in,
                                                                 uint32_t* __restrict__ out,
                                                                 const uint32_t* n_dev, uint32_t n_cap,
                                                                 const uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t lds[2 * (kScanThreads / kWave) + 1];
  const uint32_t n = load_n(n_dev, n_cap);
  const uint32_t base = blockIdx.x * (uint32_t)kScanTile;
  if (base >= n) return;
  const uint32_t start = base + threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t idx = start + i;
    vals[i] = idx < n ? in[idx] : 0u;
    s += vals[i];
  }
  uint32_t total;
  uint32_t run = block_excl_scan<kScanThreads>(s, total, lds) + tile_sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t idx = start + i;
    if (idx < n) out[idx] = run;
    run += vals[i];
  }
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;  // 2048
constexpr int kRsWaves = kRsThreads / kWave;

__global__ __launch_bounds__(kRsThreads) void rs_hist_kernel(const uint32_t* __restrict__ keys,
                                                             const uint32_t* n_dev, uint32_t n_cap,
                                                             int shift, uint32_t* __restrict__ hist,
                                                             uint32_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = load_n(n_dev, n_cap);
  const uint32_t base = blockIdx.x * (uint32_t)kRsTile;
  if (base < n) {
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
      const uint32_t idx = base + r * kRsThreads + threadIdx.x;
      if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: within a tile, elements keep their (round, thread) order.
__global__ __launch_bounds__(kRsThreads) void rs_scatter_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint32_t* n_dev,
    uint32_t n_cap, int shift, const uint32_t* __restrict__ hist_scanned, uint32_t ntiles,
    uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[kRsWaves][256];
  __shared__ uint32_t gbase[256];
  const uint32_t n = load_n(n_dev, n_cap);
  const uint32_t base = blockIdx.x * (uint32_t)kRsTile;
  if (base >= n) return;
  run[threadIdx.x] = 0;
#pragma unroll
  for (int w = 0; w < kRsWaves; ++w) wc[w][threadIdx.x] = 0;
  gbase[threadIdx.x] = hist_scanned[(size_t)threadIdx.x * ntiles + blockIdx.x];
  __syncthreads();
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
  for (int r = 0; r < kRsItems; ++r) {
    const uint32_t idx = base + r * kRsThreads + threadIdx.x;
    const bool valid = idx < n;
    const uint32_t key = valid ? keys[idx] : 0u;
    const uint32_t val = valid ? vals[idx] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t b = ballot(on);
      peers &= on ? b : ~b;
    }
    const uint32_t below = (uint32_t)__popcll(peers & lt);
    const bool leader = valid && below == 0;
    if (leader) wc[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + below;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
      const uint32_t dst = gbase[d] + pos;
      okeys[dst] = key;
      ovals[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kRsWaves; ++w) {
      add += wc[w][threadIdx.x];
      wc[w][threadIdx.x] = 0;
    }
    run[threadIdx.x] += add;
    __syncthreads();
  }
  (void)lane;
}

}  // namespace

uint32_t scan_tiles_for(uint32_t n_cap) { return (n_cap + kScanTile - 1) / kScanTile; }
uint32_t radix_tiles_for(uint32_t n_cap) { return (n_cap + kRsTile - 1) / kRsTile; }

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_cap,
                              uint32_t* tile_sums, uint32_t* total_out, hipStream_t s) {
  if (n_cap == 0) return hipSuccess;
  const uint32_t ntiles = scan_tiles_for(n_cap);
  if (ntiles > (uint32_t)(kScanTopThreads * kScanTopItems)) return hipErrorInvalidValue;
  scan_reduce_kernel<<<ntiles, kScanThreads, 0, s>>>(in, n_dev, n_cap, tile_sums);
  scan_top_kernel<<<1, kScanTopThreads, 0, s>>>(tile_sums, ntiles, total_out);
  scan_down_kernel<<<ntiles, kScanThreads, 0, s>>>(in, out, n_dev, n_cap, tile_sums);
  return hipGetLastError();
}

size_t radix_sort_scratch_words(uint32_t n_cap) {
  const uint32_t nt = radix_tiles_for(n_cap);
  const size_t hist = (size_t)256 * nt;
  return hist * 2 + scan_tiles_for((uint32_t)hist) + 16;
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                            const uint32_t* n_dev, uint32_t n_cap, int bits, uint32_t* scratch,
                            hipStream_t s, bool* result_in_alt) {
  *result_in_alt = false;
  if (n_cap == 0 || bits <= 0) return hipSuccess;
  const uint32_t nt = radix_tiles_for(n_cap);
  const size_t hsz = (size_t)256 * nt;
  if (hsz > 0xffffffffull) return hipErrorInvalidValue;
  uint32_t* hist = scratch;
  uint32_t* hist_scan = scratch + hsz;
  uint32_t* tsums = scratch + 2 * hsz;
  uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
  for (int shift = 0; shift < bits; shift += 8) {
    rs_hist_kernel<<<nt, kRsThreads, 0, s>>>(ka, n_dev, n_cap, shift, hist, nt);
    hipError_t e = exclusive_scan_u32(hist, hist_scan, nullptr, (uint32_t)hsz, tsums, nullptr, s);
    if (e != hipSuccess) return e;
    rs_scatter_kernel<<<nt, kRsThreads, 0, s>>>(ka, va, n_dev, n_cap, shift, hist_scan, nt, kb, vb);
    uint32_t* t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
    *result_in_alt = !*result_in_alt;
  }
  return hipGetLastError();
}

}  // namespace zsmpm
