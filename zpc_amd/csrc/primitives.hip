// Device-wide exclusive scan and stable LSB radix sort for gfx950 (wave64).
//
// Element counts may live in device memory, so a whole substep runs without a
// host round trip. Work follows the device-side count, not the host-side
// capacity: a fixed grid of at most kGroups workgroups is launched and
// workgroup g takes the contiguous range [g per, (g+1) per) of the n elements,
// per = ceil(n / grid). A scan is reduce -> top -> down; a radix pass is
// histogram -> one-workgroup scan of the digit counts -> stable scatter.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zsmpm_device.hpp"
#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kGroups = 128;       // grid of the range-per-workgroup kernels
constexpr int kThreads = 256;
constexpr int kItems = 16;         // scan: items per thread per step (blocked)
constexpr int kStep = kThreads * kItems;
constexpr int kSmallThreads = 1024;  // one-workgroup scan of up to kSmallThreads * kSmallItems values
constexpr int kSmallItems = 32;

__device__ __forceinline__ uint32_t load_n(const uint32_t* n_dev, uint32_t n_cap) {
  if (!n_dev) return n_cap;
  const uint32_t n = *n_dev;
  return n < n_cap ? n : n_cap;
}

// range [lo, hi) of workgroup g out of G over n elements
__device__ __forceinline__ void group_range(uint32_t n, uint32_t G, uint32_t g, uint32_t& lo, uint32_t& hi) {
  const uint32_t per = (n + G - 1) / G;
  const uint64_t l = (uint64_t)g * per;
  lo = l < n ? (uint32_t)l : n;
  hi = (uint64_t)lo + per < n ? lo + per : n;
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

template <int NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* lds) {
  uint32_t total;
  block_excl_scan<NT>(v, total, lds);
  return total;
}

// kPopc: scan the population counts of the input words instead of the words
template <bool kPopc>
__device__ __forceinline__ uint32_t scan_in(uint32_t v) {
  return kPopc ? (uint32_t)__popc(v) : v;
}

template <bool kPopc>
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                               const uint32_t* n_dev, uint32_t n_cap,
                                                               uint32_t* __restrict__ group_sums) {
  __shared__ uint32_t lds[2 * (kThreads / kWave) + 1];
  const uint32_t n = load_n(n_dev, n_cap);
  uint32_t lo, hi;
  group_range(n, gridDim.x, blockIdx.x, lo, hi);
  uint32_t s = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kThreads) s += scan_in<kPopc>(in[i]);
  const uint32_t total = block_sum<kThreads>(s, lds);
  if (threadIdx.x == 0) group_sums[blockIdx.x] = total;
}

// one workgroup: exclusive scan of up to kSmallThreads * kN values in place
// (thread t owns values [t kN, (t+1) kN), read and written as uint4 when kN
// is a multiple of 4; vals must then be 16-byte aligned); total to
// *total_out when given
template <int kN>
__global__ __launch_bounds__(kSmallThreads) void scan_small_kernel(uint32_t* __restrict__ vals, uint32_t count,
                                                                   uint32_t* __restrict__ total_out) {
  __shared__ uint32_t lds[2 * (kSmallThreads / kWave) + 1];
  const uint32_t start = threadIdx.x * kN;
  uint32_t v[kN];
  if (kN % 4 == 0 && start + kN <= count) {
#pragma unroll
    for (int i = 0; i < kN; i += 4) {
      const uint4 q = *reinterpret_cast<const uint4*>(vals + start + i);
      v[i] = q.x;
      v[i + 1] = q.y;
      v[i + 2] = q.z;
      v[i + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kN; ++i) v[i] = start + i < count ? vals[start + i] : 0u;
  }
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kN; ++i) s += v[i];
  uint32_t total;
  uint32_t run = block_excl_scan<kSmallThreads>(s, total, lds);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    const uint32_t x = v[i];
    v[i] = run;
    run += x;
  }
  if (kN % 4 == 0 && start + kN <= count) {
#pragma unroll
    for (int i = 0; i < kN; i += 4)
      *reinterpret_cast<uint4*>(vals + start + i) = make_uint4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < kN; ++i)
      if (start + i < count) vals[start + i] = v[i];
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

template <bool kPopc>
__global__ __launch_bounds__(kThreads) void scan_down_kernel(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, const uint32_t* n_dev,
                                                             uint32_t n_cap, const uint32_t* __restrict__ group_base) {
  __shared__ uint32_t lds[2 * (kThreads / kWave) + 1];
  const uint32_t n = load_n(n_dev, n_cap);
  uint32_t lo, hi;
  group_range(n, gridDim.x, blockIdx.x, lo, hi);
  uint32_t carry = group_base[blockIdx.x];
  for (uint32_t base = lo; base < hi; base += kStep) {  // uniform across the workgroup
    const uint32_t start = base + threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint32_t idx = start + i;
      v[i] = idx < hi ? scan_in<kPopc>(in[idx]) : 0u;
      s += v[i];
    }
    uint32_t total;
    uint32_t run = block_excl_scan<kThreads>(s, total, lds) + carry;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint32_t idx = start + i;
      if (idx < hi) out[idx] = run;
      run += v[i];
    }
    carry += total;
  }
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsWaves = kThreads / kWave;

// digit histogram of workgroup g's range, written digit-major: hist[d G + g]
template <int kBits>
__global__ __launch_bounds__(kThreads) void rs_hist_kernel(const uint32_t* __restrict__ keys,
                                                           const uint32_t* n_dev, uint32_t n_cap, int shift,
                                                           uint32_t* __restrict__ hist) {
  constexpr uint32_t kBins = 1u << kBits;
  __shared__ uint32_t h[kBins];
  for (uint32_t d = threadIdx.x; d < kBins; d += kThreads) h[d] = 0;
  __syncthreads();
  const uint32_t n = load_n(n_dev, n_cap);
  uint32_t lo, hi;
  group_range(n, gridDim.x, blockIdx.x, lo, hi);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kThreads) atomicAdd(&h[(keys[i] >> shift) & (kBins - 1u)], 1u);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kBins; d += kThreads) hist[(size_t)d * gridDim.x + blockIdx.x] = h[d];
}

// Stable scatter of workgroup g's range: rounds of kThreads elements in
// order; within a round, equal digits keep thread order (ballot matching).
template <int kBits>
__global__ __launch_bounds__(kThreads) void rs_scatter_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint32_t* n_dev, uint32_t n_cap,
    int shift, const uint32_t* __restrict__ hist_scanned, uint32_t* __restrict__ okeys,
    uint32_t* __restrict__ ovals) {
  constexpr uint32_t kBins = 1u << kBits;
  __shared__ uint32_t run[kBins];
  __shared__ uint32_t wc[kRsWaves][kBins];
  const uint32_t n = load_n(n_dev, n_cap);
  uint32_t lo, hi;
  group_range(n, gridDim.x, blockIdx.x, lo, hi);
  for (uint32_t d = threadIdx.x; d < kBins; d += kThreads) {
    run[d] = hist_scanned[(size_t)d * gridDim.x + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kRsWaves; ++w) wc[w][d] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
  for (uint32_t base = lo; base < hi; base += kThreads) {  // uniform across the workgroup
    const uint32_t idx = base + threadIdx.x;
    const bool valid = idx < hi;
    const uint32_t key = valid ? keys[idx] : 0u;
    const uint32_t val = valid ? vals[idx] : 0u;
    const uint32_t d = (key >> shift) & (kBins - 1u);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kBits; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t b = ballot(on);
      peers &= on ? b : ~b;
    }
    const uint32_t below = (uint32_t)__popcll(peers & lt);
    if (valid && below == 0) wc[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + below;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
      okeys[pos] = key;
      ovals[pos] = val;
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < kBins; e += kThreads) {
      uint32_t add = 0;
#pragma unroll
      for (int w = 0; w < kRsWaves; ++w) {
        add += wc[w][e];
        wc[w][e] = 0;
      }
      run[e] += add;
    }
    __syncthreads();
  }
}

// grid for n_cap elements: one workgroup per kStep elements, at most kGroups
uint32_t groups_for_cap(uint32_t n_cap) {
  const uint32_t g = (n_cap + kStep - 1) / kStep;
  return g < 1 ? 1 : (g > (uint32_t)kGroups ? (uint32_t)kGroups : g);
}

}  // namespace

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_cap,
                              uint32_t* scratch, uint32_t* total_out, hipStream_t s) {
  if (n_cap == 0) return hipSuccess;
  const uint32_t G = groups_for_cap(n_cap);
  scan_reduce_kernel<false><<<G, kThreads, 0, s>>>(in, n_dev, n_cap, scratch);
  scan_small_kernel<1><<<1, kSmallThreads, 0, s>>>(scratch, G, total_out);
  scan_down_kernel<false><<<G, kThreads, 0, s>>>(in, out, n_dev, n_cap, scratch);
  return hipGetLastError();
}

hipError_t exclusive_scan_popc_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                                   hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t G = groups_for_cap(n);
  scan_reduce_kernel<true><<<G, kThreads, 0, s>>>(in, nullptr, n, scratch);
  scan_small_kernel<1><<<1, kSmallThreads, 0, s>>>(scratch, G, nullptr);
  scan_down_kernel<true><<<G, kThreads, 0, s>>>(in, out, nullptr, n, scratch);
  return hipGetLastError();
}

size_t scan_scratch_words() { return kGroups; }

size_t radix_sort_scratch_words(uint32_t n_cap) {
  (void)n_cap;
  return (size_t)256 * kGroups;
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                            const uint32_t* n_dev, uint32_t n_cap, int bits, uint32_t* scratch,
                            hipStream_t s, bool* result_in_alt) {
  *result_in_alt = false;
  if (n_cap == 0 || bits <= 0) return hipSuccess;
  static_assert(256 * kGroups <= kSmallThreads * kSmallItems, "digit histogram exceeds the one-workgroup scan");
  static_assert(512 * (kGroups / 2) <= kSmallThreads * kSmallItems, "digit histogram exceeds the one-workgroup scan");
  static_assert(kGroups <= kSmallThreads, "group sums exceed the one-workgroup scan");
  // 17- and 18-bit keys take two 9-bit passes instead of three 8-bit ones, on
  // half the workgroups so that the 512-digit histogram fits the scan
  const bool nine = bits > 16 && bits <= 18;
  const uint32_t G8 = groups_for_cap(n_cap);
  const uint32_t G = nine ? (G8 > (uint32_t)(kGroups / 2) ? (uint32_t)(kGroups / 2) : G8) : G8;
  const int step = nine ? 9 : 8;
  uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
  for (int shift = 0; shift < bits; shift += step) {
    if (nine) {
      rs_hist_kernel<9><<<G, kThreads, 0, s>>>(ka, n_dev, n_cap, shift, scratch);
      scan_small_kernel<kSmallItems><<<1, kSmallThreads, 0, s>>>(scratch, 512 * G, nullptr);
      rs_scatter_kernel<9><<<G, kThreads, 0, s>>>(ka, va, n_dev, n_cap, shift, scratch, kb, vb);
    } else {
      rs_hist_kernel<8><<<G, kThreads, 0, s>>>(ka, n_dev, n_cap, shift, scratch);
      scan_small_kernel<kSmallItems><<<1, kSmallThreads, 0, s>>>(scratch, 256 * G, nullptr);
      rs_scatter_kernel<8><<<G, kThreads, 0, s>>>(ka, va, n_dev, n_cap, shift, scratch, kb, vb);
    }
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
