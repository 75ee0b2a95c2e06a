// The implicit step's operators on gfx950: zpc's ImplicitMPMSystem mass term
// and preconditioner, the force hand-over of G2P2GTransfer (its gather half
// is a bucket kernel beside G2P in transfer.hip, its force scatter P2G's; the
// projection lives beside the colliders there too) and the vector work of
// the preconditioned conjugate gradient. Vectors hold 3 floats per grid node, node = block * 64 +
// cell, for the blocks the table holds (the count is read on the device).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kIT = 256;

__device__ __forceinline__ int64_t node_count(const TableView& t, int32_t grid_cap) {
  const int32_t nb = *t.cnt;
  return (int64_t)(nb < grid_cap ? nb : grid_cap) * kBlockCells;
}

__device__ __forceinline__ float node_mass(const float* grid, int64_t node) {
  return grid[((node >> 6) * kChannels) * kBlockCells + (node & 63)];
}

// grid channels 1..3 (the folded P2G force) -> r, 3 floats per node
__global__ __launch_bounds__(kIT) void add_force_kernel(const float* __restrict__ grid, TableView t, int32_t grid_cap,
                                                        float* r) {
  const int64_t n = node_count(t, grid_cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t b = i >> 6;
    const int64_t cell = i & 63;
#pragma unroll
    for (int d = 0; d < 3; ++d) r[3 * i + d] += grid[(b * kChannels + 1 + d) * kBlockCells + cell];
  }
}

__global__ __launch_bounds__(kIT) void fill_kernel(float* a, TableView t, int32_t grid_cap, float value) {
  const int64_t n = 3 * node_count(t, grid_cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) a[i] = value;
}

// ForceDtSqrPlusMass (simulation/mpm/ImplicitMPM.hpp:17-33): out_i = (out_i
// dt dt + m) in_i where the node has mass
__global__ __launch_bounds__(kIT) void force_dt2_mass_kernel(const float* __restrict__ grid, TableView t,
                                                             int32_t grid_cap, float dt, const float* in, float* out) {
#pragma clang fp contract(off)
  const int64_t n = 3 * node_count(t, grid_cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const float mass = node_mass(grid, i / 3);
    if (mass > 0.f) out[i] = (out[i] * dt * dt + mass) * in[i];
  }
}

// DivPernodeMass (:129-155): out_i = in_i / m where the node has mass
__global__ __launch_bounds__(kIT) void div_mass_kernel(const float* __restrict__ grid, TableView t, int32_t grid_cap,
                                                       const float* in, float* out) {
  const int64_t n = 3 * node_count(t, grid_cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const float mass = node_mass(grid, i / 3);
    if (mass > 0.f) out[i] = in[i] / mass;
  }
}

// c = a_coef * a + b_coef * b (LinearCombineOp, math/linear/LinearOperators.hpp:32-39)
__global__ __launch_bounds__(kIT) void lincomb_kernel(TableView t, int32_t grid_cap, float ca, const float* a,
                                                      float cb, const float* b, float* c) {
#pragma clang fp contract(off)
  const int64_t n = 3 * node_count(t, grid_cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) c[i] = ca * a[i] + cb * b[i];
}

// dot product, summed in double (per thread, across the wave, one atomic
// per workgroup)
__global__ __launch_bounds__(kIT) void dot_kernel(TableView t, int32_t grid_cap, const float* a, const float* b,
                                                  double* out) {
  const int64_t n = 3 * node_count(t, grid_cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    s += (double)a[i] * (double)b[i];
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
  __shared__ double ws[kIT / kWave];
  if (lane_id() == 0) ws[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double g = 0.0;
    for (int w = 0; w < kIT / kWave; ++w) g += ws[w];
    atomicAdd(out, g);
  }
}

__global__ void zero_doubles_kernel(double* a, int n) {
  if (threadIdx.x < n) a[threadIdx.x] = 0.0;
}

int groups_for_nodes() { return 2048; }

}  // namespace

hipError_t implicit_add_force(const float* grid, int32_t grid_cap, const TableView& t, float* r, hipStream_t s) {
  add_force_kernel<<<groups_for_nodes(), kIT, 0, s>>>(grid, t, grid_cap, r);
  return hipGetLastError();
}

hipError_t implicit_force_dt2_mass(const float* grid, int32_t grid_cap, const TableView& t, float dt, const float* in,
                                   float* out, hipStream_t s) {
  force_dt2_mass_kernel<<<groups_for_nodes(), kIT, 0, s>>>(grid, t, grid_cap, dt, in, out);
  return hipGetLastError();
}

hipError_t implicit_precondition(const float* grid, int32_t grid_cap, const TableView& t, const float* in, float* out,
                                 hipStream_t s) {
  div_mass_kernel<<<groups_for_nodes(), kIT, 0, s>>>(grid, t, grid_cap, in, out);
  return hipGetLastError();
}

hipError_t implicit_lincomb(const TableView& t, int32_t grid_cap, float ca, const float* a, float cb, const float* b,
                            float* c, hipStream_t s) {
  lincomb_kernel<<<groups_for_nodes(), kIT, 0, s>>>(t, grid_cap, ca, a, cb, b, c);
  return hipGetLastError();
}

hipError_t implicit_fill(const TableView& t, int32_t grid_cap, float* a, float value, hipStream_t s) {
  fill_kernel<<<groups_for_nodes(), kIT, 0, s>>>(a, t, grid_cap, value);
  return hipGetLastError();
}

hipError_t implicit_dot(const TableView& t, int32_t grid_cap, const float* a, const float* b, double* out,
                        hipStream_t s) {
  zero_doubles_kernel<<<1, 64, 0, s>>>(out, 1);
  dot_kernel<<<groups_for_nodes(), kIT, 0, s>>>(t, grid_cap, a, b, out);
  return hipGetLastError();
}

}  // namespace zsmpm
