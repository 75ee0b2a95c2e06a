// The implicit step's operators on gfx950: the gather half of G2P2GTransfer
// (the stress at F = (I + dt C(v)) F_old for a grid velocity vector v; its
// force scatter is P2G's), zpc's ImplicitMPMSystem mass term and
// preconditioner (the projection lives beside the colliders in transfer.hip)
// and the vector work of its preconditioned conjugate gradient. Vectors hold 3 floats per grid node, node = block * 64 +
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

// The gather half of G2P2GTransfer::operator() (simulation/transfer/G2P2G.hpp:49-110),
// one thread per particle: C from the node velocities v (the 27 stencil nodes
// lie in at most 8 blocks, each looked up once), F = (I + dt C) F_old (not
// stored), and the model's vol P F^T (logJp read, not written back) into
// contrib. The scatter half is P2G's force scatter, driven by the C ABI layer.
// Rounded as written, as the oracle is.
__global__ __launch_bounds__(kIT) void g2p2g_gather_kernel(ParticlesView ps, TableView t, Model m, float dt, float dx,
                                                           const float* __restrict__ v, float* contrib, uint32_t* err) {
#pragma clang fp contract(off)
  const float dx_inv = 1.f / dx;
  const float D_inv = 4.f * dx_inv * dx_inv;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < ps.n; p += stride) {
    Arena ar;
    make_arena(ps.x[3 * p], ps.x[3 * p + 1], ps.x[3 * p + 2], dx, ar);
    int32_t blk[2][2][2];
    const int32_t b0 = ar.base[0] >> 2, b1 = ar.base[1] >> 2, b2 = ar.base[2] >> 2;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int qi = q >> 2, qj = (q >> 1) & 1, qk = q & 1;
      const bool need = ((ar.base[0] & 3) + 2 >= 4 || qi == 0) && ((ar.base[1] & 3) + 2 >= 4 || qj == 0) &&
                        ((ar.base[2] & 3) + 2 >= 4 || qk == 0);
      int32_t bn = -1;
      if (need) {
        bn = table_find(t, b0 + qi, b1 + qj, b2 + qk);
        if (bn < 0) atomicOr(err, kErrMissingBlock);
      }
      blk[qi][qj][qk] = bn;
    }
    float C[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const int32_t c0 = ar.base[0] + i, c1 = ar.base[1] + j, c2 = ar.base[2] + k;
          const int32_t bn = blk[(c0 >> 2) - b0][(c1 >> 2) - b1][(c2 >> 2) - b2];
          if (bn < 0) continue;
          const int cell = ((c0 & 3) << 4) | ((c1 & 3) << 2) | (c2 & 3);
          const float xixp[3] = {(float)i * dx - ar.lp[0], (float)j * dx - ar.lp[1], (float)k * dx - ar.lp[2]};
          float W = 1.f;
          W *= ar.w[0][i];
          W *= ar.w[1][j];
          W *= ar.w[2][k];
          const float* vi = v + 3 * ((int64_t)bn * kBlockCells + cell);
#pragma unroll
          for (int d = 0; d < 9; ++d) C[d] += W * vi[d % 3] * xixp[d / 3] * D_inv;
        }
    float tmp[9], F[9], PF[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) tmp[d] = C[d] * dt + ((d & 0x3) ? 0.f : 1.f);
    const float* Fo = ps.F + 9 * p;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        F[i + 3 * j] = tmp[i] * Fo[3 * j] + tmp[i + 3] * Fo[3 * j + 1] + tmp[i + 6] * Fo[3 * j + 2];
    bool f64 = false;
    if (m.kind == 0) {
      stress_fixedcorotated(m, F, PF, f64);
    } else {
      float lj = ps.logJp[p];  // read, not written back (a local copy in zpc)
      stress_sand(m, F, lj, PF, nullptr, f64);
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) contrib[9 * p + d] = PF[d];
  }
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

hipError_t g2p2g_gather(const ParticlesView& ps, const TableView& t, const Model& m, float dt, float dx,
                        const float* v, float* contrib, uint32_t* err, hipStream_t s) {
  if (ps.n == 0) return hipSuccess;
  const int64_t g = (ps.n + kIT - 1) / kIT;
  g2p2g_gather_kernel<<<(int)(g > 65536 ? 65536 : g), kIT, 0, s>>>(ps, t, m, dt, dx, v, contrib, err);
  return hipGetLastError();
}

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
