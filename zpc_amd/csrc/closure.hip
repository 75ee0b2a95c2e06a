// Learned stress closure for the Drucker-Prager sand model: the MLP on BF16
// MFMA (inference, training gradient, Adam), the harvest of training records
// from the plain SVD return mapping, and the closure stress pass with its SVD
// fallback queue.
//
// With A = C - I (C = F^T F), n = |A|_F and U = A / n, the network maps
// invariants of U, n, logJp and the material (9 inputs) to 7 outputs: the
// Hencky strain E = log(C)/2 and the projected strain Lp of the return
// mapping as isotropic polynomials E / n = psi0 I + psi1 U + psi2 U^2,
// Lp / n = rho0 I + rho1 U + rho2 U^2, and the plastic multiplier. The
// stress pass starts one Newton step for X = log C from X0 = 2 E_pred; X0 is
// a polynomial in C, so log C = X0 + log(I + D), D = exp(-X0) C - I, holds
// exactly and the step's error is the series truncation, bounded a
// posteriori. The return mapping of physics/ConstitutiveModel_Vol_dP.hpp:
// 241-325 acts on principal values only and is applied to E in tensor form;
// vol P F_proj^T = vol F S F^T with S = tau C^-1, tau = 2 mu Lp + lambda
// tr(Lp) I. Particles whose bound exceeds the tolerance (or |D| >= 1/2,
// |X| > 1, det F <= 0) are queued and take the plain SVD path.
//
// Network layout on MFMA: every layer computes W^T (as the A operand) times
// the activations (as the B operand, one 16-particle tile per column block),
// so a layer's accumulator (rows = units, lane = particle) is the next
// layer's B operand without data movement. v_mfma_f32_16x16x16_bf16 takes 4
// consecutive K per lane group, v_mfma_f32_16x16x32_bf16 8: the 32-wide K
// steps use the hidden-unit order hid(ks, g, e) (four rows of two 16-unit
// tiles per lane group), which the weight operands follow.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zsmpm_device.hpp"
#include "zsmpm_internal.hpp"

namespace zsmpm {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));

constexpr int kIn = kMlpIn, kH = 64, kOut = kMlpOut, kRec = kMlpRec;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut, kNP = kOffB3 + kOut;
static_assert(kNP == kMlpParams, "parameter layout");
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

// bfloat16 bits of a float, round to nearest, ties to even
// float -> bfloat16, round to nearest even: gfx950's v_cvt_pk_bf16_f32 rounds
// two values in one instruction (the integer form, (u + 0x7fff + lsb) >> 16,
// gives the same bits for finite values in about six operations per value)
typedef __bf16 zs_bf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  const zs_bf2 v = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ uint32_t bf16_bits(float f) { return pack2(f, 0.f) & 0xffffu; }
__device__ __forceinline__ v4s pack4(const float (&v)[4]) {
  v2u u;
  u.x = pack2(v[0], v[1]);
  u.y = pack2(v[2], v[3]);
  return __builtin_bit_cast(v4s, u);
}
__device__ __forceinline__ v4s pack4_bits(const uint32_t (&b)[4]) {
  v2u u;
  u.x = b[0] | (b[1] << 16);
  u.y = b[2] | (b[3] << 16);
  return __builtin_bit_cast(v4s, u);
}
__device__ __forceinline__ v8bf pack8(const float (&v)[8]) {
  v4u u;
  u.x = pack2(v[0], v[1]);
  u.y = pack2(v[2], v[3]);
  u.z = pack2(v[4], v[5]);
  u.w = pack2(v[6], v[7]);
  return __builtin_bit_cast(v8bf, u);
}
__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma32(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hidden unit of slot e (0..7) of lane group g in the 32-wide K step ks
__device__ __forceinline__ int hid(int ks, int g, int e) { return e < 4 ? 32 * ks + 4 * g + e : 32 * ks + 16 + 4 * g + (e - 4); }

// tanh form of GELU; tanh(u) = 1 - 2 / (e^{2u} + 1) with the hardware
// exponential and reciprocal (1 ulp each; e^{2u} = inf gives tanh = 1)
// tanh-form GELU, 0.5 x (1 + tanh u) with u = C (x + A x^3), evaluated as
// x s with s = sigmoid(2 u) = 1 / (1 + 2^(x (k1 + k2 x^2))): one exp2 and one
// rcp on the hardware units and five other operations. Its derivative is
// s + 2 C x s (1 - s) (1 + 3 A x^2). Large |x| saturate to 0 and x.
constexpr float kGeluK1 = (float)(-2.0 * 0.7978845608028654 * 1.4426950408889634);
constexpr float kGeluK2 = (float)(-2.0 * 0.7978845608028654 * 0.044715 * 1.4426950408889634);
__device__ __forceinline__ float gelu_s(float x) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * fmaf(kGeluK2, x * x, kGeluK1)));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float sg = gelu_s(x);
  return sg + 2.f * kGeluC * x * sg * (1.f - sg) * fmaf(3.f * kGeluA, x * x, 1.f);
}

// The forward weights of one lane, loaded once per wave from the float
// master copy and rounded to bfloat16.
// The biases stay in a per-wave LDS block (kNetBias floats: b1 at 0, b2 at
// 64, b3 at 128 padded with zeros to 16), read as one 16-byte vector per tile
// row: in registers they took 36 VGPRs, which with the weights held the
// network kernels to two waves per SIMD. The sums are the same.
constexpr int kNetBias = 144;

// W2 stays in one LDS block per workgroup, in the MFMA operand layout
// (kNetW2 vectors: [m-tile mt][K step ks][lane]), read per tile as 16-byte
// vectors; in registers it took 32 VGPRs. Same operands, same sums.
constexpr int kNetW2 = 4 * 2 * 64;

struct NetRegs {
  v4s a1[4];      // W1^T, m-tile mt: out 16 mt + (l & 15), in 4 g + e
  v8bf a3[2];     // W3^T: out (l & 15) < 7, in hid(ks, g, e)
  const float* bias;  // the wave's LDS bias block, 16-byte aligned
  const v8bf* w2;     // the workgroup's LDS block of W2^T: out 16 mt + (l & 15), in hid(ks, g, e)
};

// wb: the calling wave's LDS block of kNetBias floats (16-byte aligned);
// w2: the workgroup's LDS block of kNetW2 vectors, written by its first wave.
// Every thread of the workgroup must call it (it ends with a barrier).
__device__ __forceinline__ void load_net(const float* __restrict__ P, NetRegs& r, float* wb, v8bf* w2) {
  const int l = lane_id(), i = l & 15, g = l >> 4;
  wb[l] = P[kOffB1 + l];
  wb[kH + l] = P[kOffB2 + l];
  if (l < 16) wb[2 * kH + l] = l < kOut ? P[kOffB3 + l] : 0.f;
  r.bias = wb;
  r.w2 = w2;
  if (threadIdx.x < kWave) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        float w[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = P[kOffW2 + hid(ks, g, e) * kH + 16 * mt + i];
        w2[(2 * mt + ks) * kWave + l] = pack8(w);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      v[e] = k < kIn ? P[kOffW1 + k * kH + 16 * mt + i] : 0.f;
    }
    r.a1[mt] = pack4(v);
  }
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    float w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = i < kOut ? P[kOffW3 + hid(ks, g, e) * kOut + i] : 0.f;
    r.a3[ks] = pack8(w);
  }
}

// bias of accumulator rows 4 g + e (e = 0..3) of row block `base`; the empty
// asm makes the index opaque so that the compiler reloads the vector per
// tile instead of hoisting all biases back into registers
__device__ __forceinline__ float4 net_bias(const NetRegs& r, int base) {
  int q = (base >> 2) + (lane_id() >> 4);
  asm volatile("" : "+v"(q));
  return reinterpret_cast<const float4*>(r.bias)[q];
}

// Intermediates the backward pass needs.
struct Acts {
  v4f z1[4], z2[4];        // pre-activations, rows 16 mt + 4 g + e
  uint32_t a1[4][4], a2[4][4];  // bfloat16 bits of the activations, same rows
};

// One 16-particle tile: x holds features 4 g .. 4 g + 3 of particle (l & 15);
// returns outputs 4 g + e of that particle. Every lane must be active.
template <bool kKeep>
__device__ __forceinline__ v4f net_tile(const NetRegs& r, v4s x, Acts* acts) {
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
  float a1[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    v4f d = mfma16(r.a1[mt], x, zero);
    const float4 bb = net_bias(r, 16 * mt);
    const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      d[e] += bv[e];
      a1[mt][e] = gelu(d[e]);
      if constexpr (kKeep) acts->a1[mt][e] = bf16_bits(a1[mt][e]);
    }
    if constexpr (kKeep) acts->z1[mt] = d;
  }
  v8bf b2[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const float v[8] = {a1[2 * ks][0],     a1[2 * ks][1],     a1[2 * ks][2],     a1[2 * ks][3],
                        a1[2 * ks + 1][0], a1[2 * ks + 1][1], a1[2 * ks + 1][2], a1[2 * ks + 1][3]};
    b2[ks] = pack8(v);
  }
  float a2[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    int q = 2 * mt * kWave + lane_id();
    asm volatile("" : "+v"(q));  // reload per tile: keep W2 out of registers
    v4f d = mfma32(r.w2[q], b2[0], zero);
    d = mfma32(r.w2[q + kWave], b2[1], d);
    const float4 bb = net_bias(r, kH + 16 * mt);
    const float bv2[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      d[e] += bv2[e];
      a2[mt][e] = gelu(d[e]);
      if constexpr (kKeep) acts->a2[mt][e] = bf16_bits(a2[mt][e]);
    }
    if constexpr (kKeep) acts->z2[mt] = d;
  }
  v8bf b3[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const float v[8] = {a2[2 * ks][0],     a2[2 * ks][1],     a2[2 * ks][2],     a2[2 * ks][3],
                        a2[2 * ks + 1][0], a2[2 * ks + 1][1], a2[2 * ks + 1][2], a2[2 * ks + 1][3]};
    b3[ks] = pack8(v);
  }
  v4f y = mfma32(r.a3[0], b3[0], zero);
  y = mfma32(r.a3[1], b3[1], y);
  {
    const float4 bb = net_bias(r, 2 * kH);
    const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) y[e] += bv[e];
  }
  return y;
}

// ---------------------------------------------------------------------------
// inference on feature rows [n][9] -> outputs [n][7]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_infer_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                        int64_t n, float* __restrict__ Y) {
  __shared__ float4 nbias[256 / kWave][kNetBias / 4];
  __shared__ v8bf nw2[kNetW2];
  NetRegs r;
  load_net(P, r, reinterpret_cast<float*>(nbias[threadIdx.x / kWave]), nw2);
  const int l = lane_id(), j = l & 15, g = l >> 4;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t w0 = uniform((int64_t)(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
  for (int64_t t = w0; t * 16 < n; t += nw) {
    const int64_t s = t * 16 + j;
    float f[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      f[e] = (s < n && k < kIn) ? X[s * kIn + k] : 0.f;
    }
    const v4f y = net_tile<false>(r, pack4(f), nullptr);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = 4 * g + e;
      if (s < n && o < kOut) Y[s * kOut + o] = y[e];
    }
  }
}

// ---------------------------------------------------------------------------
// training gradient: per-workgroup partial sums of the parameter gradient of
// the mean per-sample loss (see mlp_reduce_kernel)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                      const float* __restrict__ R, int64_t n,
                                                      float* __restrict__ partial, float* __restrict__ loss_part) {
  __shared__ uint16_t sx[16][16], sa1[16][kH], sa2[16][kH], sdz1[16][kH], sdz2[16][kH], sdy[16][16];
  __shared__ float sy[16][8], sdyf[16][8];
  __shared__ float4 nbias[kNetBias / 4];
  __shared__ v8bf nw2[kNetW2];
  NetRegs r;
  load_net(P, r, reinterpret_cast<float*>(nbias), nw2);
  const int l = lane_id(), j = l & 15, g = l >> 4;
  // backward operands: W3 as A (rows hidden 16 mt + (l & 15), K = outputs
  // 4 g + e) and W2 as A (rows in-unit 16 mt + (l & 15), K = out-units 16 ks + 4 g + e)
  v4s a3b[4], a2b[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = 4 * g + e;
      v[e] = o < kOut ? P[kOffW3 + (16 * mt + j) * kOut + o] : 0.f;
    }
    a3b[mt] = pack4(v);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = P[kOffW2 + (16 * mt + j) * kH + 16 * ks + 4 * g + e];
      a2b[mt][ks] = pack4(v);
    }
  }
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
  v4f gw1[4], gw2[4][4], gw3[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gw1[a] = zero;
    gw3[a] = zero;
#pragma unroll
    for (int b = 0; b < 4; ++b) gw2[a][b] = zero;
  }
  float gb1[4][4] = {}, gb2[4][4] = {}, gb3[4] = {};
  float lsum = 0.f;
  const float inv_n = 1.f / (float)n;

  for (int64_t t = blockIdx.x; t * 16 < n; t += gridDim.x) {
    const int64_t s = t * 16 + j;
    const bool valid = s < n;
    float f[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * g + e;
      f[e] = (valid && k < kIn) ? X[s * kIn + k] : 0.f;
      sx[j][k] = (uint16_t)bf16_bits(f[e]);
    }
    Acts act;
    const v4f y = net_tile<true>(r, pack4(f), &act);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sa1[j][16 * mt + 4 * g + e] = (uint16_t)act.a1[mt][e];
        sa2[j][16 * mt + 4 * g + e] = (uint16_t)act.a2[mt][e];
      }
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (4 * g + e < 8) sy[j][4 * g + e] = y[e];
    wave_lds_barrier();
    // per-sample loss in eigen-space (lanes 0..15, one sample each)
    if (l < 16) {
      float dy[kOut] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (valid) {
        float rec[kRec];
#pragma unroll
        for (int k = 0; k < kRec; ++k) rec[k] = R[s * kRec + k];
        float yy[kOut];
#pragma unroll
        for (int o = 0; o < kOut; ++o) yy[o] = sy[l][o];
        float loss = 0.f;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const float mm = rec[q];
          const float re = yy[0] + yy[1] * mm + yy[2] * mm * mm - rec[3 + q];
          const float rl = yy[3] + yy[4] * mm + yy[5] * mm * mm - rec[6 + q];
          loss += re * re + rl * rl;
          dy[0] += 2.f * re;
          dy[1] += 2.f * re * mm;
          dy[2] += 2.f * re * mm * mm;
          dy[3] += 2.f * rl;
          dy[4] += 2.f * rl * mm;
          dy[5] += 2.f * rl * mm * mm;
        }
        const float rd = yy[6] - rec[9];
        loss += rd * rd;
        dy[6] = 2.f * rd;
        lsum += loss;
      }
#pragma unroll
      for (int o = 0; o < kOut; ++o) sdyf[l][o] = dy[o] * inv_n;
      sdyf[l][7] = 0.f;
    }
    wave_lds_barrier();
    // dY in accumulator layout (rows = outputs 4 g + e)
    float dyv[4];
    uint32_t dyb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = 4 * g + e;
      dyv[e] = o < kOut ? sdyf[j][o] : 0.f;
      gb3[e] += dyv[e];
      dyb[e] = bf16_bits(dyv[e]);
      sdy[j][o] = (uint16_t)dyb[e];
    }
    const v4s dyB = pack4_bits(dyb);
    // dZ2 = (W3 dY) * gelu'(z2)
    v4s dz2B[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const v4f da = mfma16(a3b[mt], dyB, zero);
      uint32_t b[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float dz = da[e] * gelu_grad(act.z2[mt][e]);
        gb2[mt][e] += dz;
        b[e] = bf16_bits(dz);
        sdz2[j][16 * mt + 4 * g + e] = (uint16_t)b[e];
      }
      dz2B[mt] = pack4_bits(b);
    }
    // dZ1 = (W2 dZ2) * gelu'(z1)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      v4f da = zero;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) da = mfma16(a2b[mt][ks], dz2B[ks], da);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float dz = da[e] * gelu_grad(act.z1[mt][e]);
        gb1[mt][e] += dz;
        sdz1[j][16 * mt + 4 * g + e] = (uint16_t)bf16_bits(dz);
      }
    }
    wave_lds_barrier();
    // weight gradients, K = the 16 samples of the tile (4 per lane group)
    {
      uint32_t b[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) b[e] = sdy[4 * g + e][j];
      const v4s aY = pack4_bits(b);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = sa2[4 * g + e][16 * nt + j];
        gw3[nt] = mfma16(aY, pack4_bits(b), gw3[nt]);
      }
      v4s bA1[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = sa1[4 * g + e][16 * nt + j];
        bA1[nt] = pack4_bits(b);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) b[e] = sx[4 * g + e][j];
      const v4s bX = pack4_bits(b);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = sdz2[4 * g + e][16 * mt + j];
        const v4s aZ2 = pack4_bits(b);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) gw2[mt][nt] = mfma16(aZ2, bA1[nt], gw2[mt][nt]);
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = sdz1[4 * g + e][16 * mt + j];
        gw1[mt] = mfma16(pack4_bits(b), bX, gw1[mt]);
      }
    }
    wave_lds_barrier();
  }
  // bias sums over the 16 samples (lanes) of each lane group
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        gb1[mt][e] += __shfl_xor(gb1[mt][e], o, kWave);
        gb2[mt][e] += __shfl_xor(gb2[mt][e], o, kWave);
      }
      gb3[e] += __shfl_xor(gb3[e], o, kWave);
    }
    lsum += __shfl_xor(lsum, o, kWave);
  }
  float* out = partial + (int64_t)blockIdx.x * kNP;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int u = 16 * mt + 4 * g + e;  // accumulator row
      if (j < kIn) out[kOffW1 + j * kH + u] = gw1[mt][e];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) out[kOffW2 + (16 * nt + j) * kH + u] = gw2[mt][nt][e];
      if (j == 0) {
        out[kOffB1 + u] = gb1[mt][e];
        out[kOffB2 + u] = gb2[mt][e];
      }
    }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int o = 4 * g + e;
    if (o < kOut) {
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) out[kOffW3 + (16 * nt + j) * kOut + o] = gw3[nt][e];
      if (j == 0) out[kOffB3 + o] = gb3[e];
    }
  }
  if (l == 0) loss_part[blockIdx.x] = lsum * inv_n;
}

// Sum of the partials; optional gradient output; optional Adam update
// (Kingma & Ba 2015) with bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t.
// A workgroup takes kRedParams parameters; its kRedParts slices of threads
// sum every kRedParts-th partial (four loads in flight per thread), and the
// slices are combined through LDS. Workgroup 0 also sums the loss partials.
constexpr int kRedParams = 32;
constexpr int kRedParts = 256 / kRedParams;

__global__ __launch_bounds__(256) void mlp_reduce_kernel(float* __restrict__ P, float* __restrict__ m,
                                                         float* __restrict__ v, const float* __restrict__ partial,
                                                         const float* __restrict__ loss_part, int G, float lr,
                                                         float bc1, float bc2, float* __restrict__ grad_out,
                                                         float* __restrict__ loss_out) {
  __shared__ float red[kRedParts][kRedParams];
  __shared__ float lred[256 / kWave];
  const int pl = (int)threadIdx.x % kRedParams, part = (int)threadIdx.x / kRedParams;
  const int i = (int)blockIdx.x * kRedParams + pl;
  const bool loss_here = blockIdx.x == 0 && loss_out != nullptr;
  if (loss_here) {
    float ls = 0.f;
    for (int b = (int)threadIdx.x; b < G; b += (int)blockDim.x) ls += loss_part[b];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) ls += __shfl_xor(ls, o, kWave);
    if (lane_id() == 0) lred[threadIdx.x / kWave] = ls;
  }
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
  if (i < kNP) {
    int b = part;
    for (; b + 3 * kRedParts < G; b += 4 * kRedParts) {
      a0 += partial[(int64_t)b * kNP + i];
      a1 += partial[(int64_t)(b + kRedParts) * kNP + i];
      a2 += partial[(int64_t)(b + 2 * kRedParts) * kNP + i];
      a3 += partial[(int64_t)(b + 3 * kRedParts) * kNP + i];
    }
    for (; b < G; b += kRedParts) a0 += partial[(int64_t)b * kNP + i];
  }
  red[part][pl] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (loss_here && threadIdx.x == 0) {
    float ls = 0.f;
#pragma unroll
    for (int w = 0; w < 256 / kWave; ++w) ls += lred[w];
    *loss_out = ls;
  }
  if (part != 0 || i >= kNP) return;
  float gsum = 0.f;
#pragma unroll
  for (int q = 0; q < kRedParts; ++q) gsum += red[q][pl];
  if (grad_out) grad_out[i] = gsum;
  if (P && m && v) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.f - b1) * gsum;
    const float vi = b2 * v[i] + (1.f - b2) * gsum * gsum;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

// ---------------------------------------------------------------------------
// symmetric 3x3 helpers: s[0..5] = (00, 11, 22, 01, 02, 12)
// ---------------------------------------------------------------------------
struct Sym {
  float a[6];
};
__device__ __forceinline__ Sym sym_mul(const Sym& p, const Sym& q) {
  // upper triangle of p q (symmetric when p and q commute)
  const float* x = p.a;
  const float* y = q.a;
  Sym r;
  r.a[0] = x[0] * y[0] + x[3] * y[3] + x[4] * y[4];
  r.a[1] = x[3] * y[3] + x[1] * y[1] + x[5] * y[5];
  r.a[2] = x[4] * y[4] + x[5] * y[5] + x[2] * y[2];
  r.a[3] = x[0] * y[3] + x[3] * y[1] + x[4] * y[5];
  r.a[4] = x[0] * y[4] + x[3] * y[5] + x[4] * y[2];
  r.a[5] = x[3] * y[4] + x[1] * y[5] + x[5] * y[2];
  return r;
}
__device__ __forceinline__ float sym_norm2(const Sym& p) {
  const float* x = p.a;
  return x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + 2.f * (x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
}
__device__ __forceinline__ float sym_det(const Sym& p) {
  const float* x = p.a;
  return x[0] * (x[1] * x[2] - x[5] * x[5]) - x[3] * (x[3] * x[2] - x[5] * x[4]) + x[4] * (x[3] * x[5] - x[1] * x[4]);
}

// C - I = G + G^T + G^T G, G = F - I (F column-major)
__device__ __forceinline__ Sym c_minus_i(const float* F) {
  float G[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) G[i][k] = F[i + 3 * k] - (i == k ? 1.f : 0.f);
  auto c = [&](int i, int k) { return G[i][k] + G[k][i] + G[0][i] * G[0][k] + G[1][i] * G[1][k] + G[2][i] * G[2][k]; };
  Sym A;
  A.a[0] = c(0, 0);
  A.a[1] = c(1, 1);
  A.a[2] = c(2, 2);
  A.a[3] = c(0, 1);
  A.a[4] = c(0, 2);
  A.a[5] = c(1, 2);
  return A;
}

__device__ __forceinline__ float friction(const Model& m) {
  return m.yield_surface * (3.f * m.lambda + 2.f * m.mu) / (2.f * m.mu);
}

// the 9 network inputs; returns n = |C - I|_F (at least 1e-30)
__device__ __forceinline__ float closure_features(const Sym& A, float lj, const Model& m, float (&f)[kIn]) {
  const float nrm = fmaxf(sqrtf(sym_norm2(A)), 1e-30f);
  const float inv = 1.f / nrm;
  Sym U;
#pragma unroll
  for (int k = 0; k < 6; ++k) U.a[k] = A.a[k] * inv;
  const Sym U2 = sym_mul(U, U);
  f[0] = U.a[0] + U.a[1] + U.a[2];
  f[1] = U2.a[0] * U.a[0] + U2.a[1] * U.a[1] + U2.a[2] * U.a[2] +
         2.f * (U2.a[3] * U.a[3] + U2.a[4] * U.a[4] + U2.a[5] * U.a[5]);
  f[2] = sym_det(U);
  f[3] = logf(nrm) * 0.125f + 1.f;
  f[4] = fminf(nrm, 1.f);
  f[5] = fminf(fmaxf(lj * inv, -8.f), 8.f) * 0.125f;
  f[6] = friction(m);
  f[7] = fminf(fmaxf(m.cohesion * inv, -8.f), 8.f) * 0.125f;
  f[8] = m.lambda / (m.lambda + 2.f * m.mu);
  return nrm;
}

// The return mapping on principal Hencky strains (the P2G sand branch):
// projected strains, logJp, plastic multiplier.
__device__ __forceinline__ void project_principal(const float (&e)[3], float lj, const Model& m, float (&lp)[3],
                                                  float& dg) {
  const float eps[3] = {e[0] - m.cohesion, e[1] - m.cohesion, e[2] - m.cohesion};
  const float tr = eps[0] + eps[1] + eps[2] + lj;
  const float t3 = tr / 3.f;
  const float eh[3] = {eps[0] - t3, eps[1] - t3, eps[2] - t3};
  const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
  dg = ehn + friction(m) * tr;
  if (tr >= 0.f) {
    lp[0] = lp[1] = lp[2] = m.cohesion;
  } else if (dg > 0.f) {
    const float r = dg / ehn;
#pragma unroll
    for (int q = 0; q < 3; ++q) lp[q] = e[q] - r * eh[q];
  } else {
#pragma unroll
    for (int q = 0; q < 3; ++q) lp[q] = e[q];
  }
}

// ---------------------------------------------------------------------------
// harvest: features and eigen-space target records from the plain SVD path
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void harvest_kernel(Model m, const float* __restrict__ F,
                                                      const float* __restrict__ logJp, int64_t n,
                                                      float* __restrict__ X, float* __restrict__ Rout) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float Fp[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Fp[k] = F[9 * p + k];
  const float lj = logJp ? logJp[p] : 0.f;
  const Sym A = c_minus_i(Fp);
  float f[kIn];
  const float nrm = closure_features(A, lj, m, f);
#pragma unroll
  for (int k = 0; k < kIn; ++k) X[p * kIn + k] = f[k];
  float U[3][3], S[3], V[3][3];
  bool f64 = false;
  svd_colmajor(Fp, U, S, V, m.fp64_gap, f64);
  const float inv = 1.f / nrm;
  float e[3], lp[3], dg;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float s = fabsf(S[q]);
    Rout[p * kRec + q] = (s - 1.f) * (s + 1.f) * inv;  // eigenvalue of U
    e[q] = logf(fmaxf(s, 1e-4f));
  }
  project_principal(e, lj, m, lp, dg);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    Rout[p * kRec + 3 + q] = e[q] * inv;
    Rout[p * kRec + 6 + q] = lp[q] * inv;
  }
  Rout[p * kRec + 9] = fminf(fmaxf(dg * inv, -8.f), 8.f) * 0.125f;
}

// ---------------------------------------------------------------------------
// the closure stress pass
// ---------------------------------------------------------------------------
constexpr float kResidualFloor = 1e-6f;  // float32 rounding of the step, relative

// vol P F_proj^T from psi (network or zero start); returns false where the
// particle must take the SVD path (logJp then stays untouched).
__device__ __forceinline__ bool closure_stress(const Model& m, const float* F, const Sym& A, float nrm,
                                               const float (&psi)[3], float tol, float& lj, float* PF) {
  const Sym A2 = sym_mul(A, A);
  const float c0 = 2.f * nrm * psi[0], c1 = 2.f * psi[1], c2 = 2.f * psi[2] / nrm;
  Sym X0;
#pragma unroll
  for (int k = 0; k < 6; ++k) X0.a[k] = c1 * A.a[k] + c2 * A2.a[k] + (k < 3 ? c0 : 0.f);
  // Z = exp(-X0) - I: degree-6 Taylor of exp(Y) - I, Y = -X0 / 4, squared twice
  Sym Y;
#pragma unroll
  for (int k = 0; k < 6; ++k) Y.a[k] = -0.25f * X0.a[k];
  Sym z;
#pragma unroll
  for (int k = 0; k < 6; ++k) z.a[k] = Y.a[k] * (1.f / 6.f) + (k < 3 ? 1.f : 0.f);
  const float inv_k[4] = {1.f / 5.f, 1.f / 4.f, 1.f / 3.f, 1.f / 2.f};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const Sym t = sym_mul(Y, z);
#pragma unroll
    for (int k = 0; k < 6; ++k) z.a[k] = t.a[k] * inv_k[q] + (k < 3 ? 1.f : 0.f);
  }
  z = sym_mul(Y, z);
#pragma unroll
  for (int sq = 0; sq < 2; ++sq) {
    const Sym t = sym_mul(z, z);
#pragma unroll
    for (int k = 0; k < 6; ++k) z.a[k] = 2.f * z.a[k] + t.a[k];
  }
  // D = Z + A + Z A, X = X0 + D - D^2 / 2 + D^3 / 3
  Sym D;
  {
    const Sym t = sym_mul(z, A);
#pragma unroll
    for (int k = 0; k < 6; ++k) D.a[k] = z.a[k] + A.a[k] + t.a[k];
  }
  const float d = sqrtf(sym_norm2(D));
  const Sym D2 = sym_mul(D, D);
  const Sym D3 = sym_mul(D2, D);
  Sym X;
#pragma unroll
  for (int k = 0; k < 6; ++k) X.a[k] = X0.a[k] + D.a[k] - 0.5f * D2.a[k] + D3.a[k] * (1.f / 3.f);
  const float xn = sqrtf(sym_norm2(X));
  const float y = 0.25f * sqrtf(sym_norm2(X0));
  const float y2 = y * y, y7 = y2 * y2 * y2 * y;
  const float ser = d * d * d * d / (4.f * fmaxf(1.f - d, 1e-30f));
  const float ex = 4.f * __expf(4.f * y) * y7 * (1.f / 5040.f) * (1.f + nrm);
  const float res = (ser + ex) / fmaxf(xn, 1e-30f) + kResidualFloor;
  const float detF = F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
                     F[6] * (F[1] * F[5] - F[4] * F[2]);
  if (!(res <= tol) || !(d < 0.5f) || !(xn <= 1.f) || !(detF > 0.f)) return false;
  // return mapping on E = X / 2 in tensor form
  Sym E;
#pragma unroll
  for (int k = 0; k < 6; ++k) E.a[k] = 0.5f * X.a[k];
  const float tre = E.a[0] + E.a[1] + E.a[2] - 3.f * m.cohesion;
  const float tr = tre + lj;
  const float t3 = tr / 3.f;
  Sym eh;  // eps - tr/3 I, eps = E - c I
#pragma unroll
  for (int k = 0; k < 6; ++k) eh.a[k] = E.a[k] - (k < 3 ? m.cohesion + t3 : 0.f);
  const float ehn = sqrtf(sym_norm2(eh));
  const float dg = ehn + friction(m) * tr;
  Sym Lp;
  if (tr >= 0.f) {
#pragma unroll
    for (int k = 0; k < 6; ++k) Lp.a[k] = k < 3 ? m.cohesion : 0.f;
    if (m.volume_correction) lj = m.beta * tre + lj;
  } else {
    lj = 0.f;
    const float r = dg > 0.f ? dg / ehn : 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) Lp.a[k] = E.a[k] - r * eh.a[k];
  }
  // tau = 2 mu Lp + lambda tr(Lp) I; S = tau C^-1; vol F S F^T
  const float trl = Lp.a[0] + Lp.a[1] + Lp.a[2];
  Sym tau;
#pragma unroll
  for (int k = 0; k < 6; ++k) tau.a[k] = 2.f * m.mu * Lp.a[k] + (k < 3 ? m.lambda * trl : 0.f);
  Sym Cm;
#pragma unroll
  for (int k = 0; k < 6; ++k) Cm.a[k] = A.a[k] + (k < 3 ? 1.f : 0.f);
  const float* c = Cm.a;
  const float detC = sym_det(Cm);
  const float idet = 1.f / detC;
  Sym Ci;  // adjugate / det
  Ci.a[0] = (c[1] * c[2] - c[5] * c[5]) * idet;
  Ci.a[1] = (c[0] * c[2] - c[4] * c[4]) * idet;
  Ci.a[2] = (c[0] * c[1] - c[3] * c[3]) * idet;
  Ci.a[3] = (c[4] * c[5] - c[3] * c[2]) * idet;
  Ci.a[4] = (c[3] * c[5] - c[4] * c[1]) * idet;
  Ci.a[5] = (c[3] * c[4] - c[0] * c[5]) * idet;
  const Sym S = sym_mul(tau, Ci);
  const float Sf[3][3] = {{S.a[0], S.a[3], S.a[4]}, {S.a[3], S.a[1], S.a[5]}, {S.a[4], S.a[5], S.a[2]}};
  float T[3][3];  // F S
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) T[i][k] = F[i] * Sf[0][k] + F[i + 3] * Sf[1][k] + F[i + 6] * Sf[2][k];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      PF[i + 3 * k] = m.volume * (T[i][0] * F[k] + T[i][1] * F[k + 3] + T[i][2] * F[k + 6]);
  return true;
}

constexpr int kCThreads = 256;
constexpr int kCWaves = kCThreads / kWave;

// Whether one step from X0 = 0 certifies the particle a priori: the step
// gives log1p(A) to degree 3, with error below |A|^4 / (4 (1 - |A|)) and
// |log C| >= |A| (1 - |A|/2).
__device__ __forceinline__ bool zero_start_certifies(float nrm, float tol) {
  if (!(nrm < 0.5f)) return false;
  const float b = nrm * nrm * nrm / (4.f * (1.f - nrm) * (1.f - 0.5f * nrm)) + kResidualFloor;
  return b <= 0.5f * tol;
}

// One wave per 64 consecutive particles. kNet: start from the network's
// prediction (else from X0 = 0); a wave whose particles the zero start
// certifies a priori skips the network unless `always`. Rejected particles
// are appended to the fallback queue (queue[0] counts them) and keep their
// logJp.
template <bool kNet>
__global__ __launch_bounds__(kCThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void closure_kernel(Model m, ParticlesView ps, float* __restrict__ contrib,
                                                            const float* __restrict__ P, float tol, int always,
                                                            int32_t* __restrict__ fb, uint32_t* queue,
                                                            uint32_t* __restrict__ wq, uint32_t* wq_n) {
  __shared__ float4 stage4[kCWaves][stage_floats(9) / 4];
  __shared__ float sy[kCWaves][kWave][4];
  const int wave = uniform((int32_t)(threadIdx.x / kWave)), lane = lane_id();
  float* stage = reinterpret_cast<float*>(stage4[wave]);
  const int64_t total_waves = (ps.n + kWave - 1) / kWave;
  // kNet with wave flags: only the flagged waves, which the zero-start pass
  // could not certify (their count goes to wq_n); without kNet the flags
  // receive that verdict for every wave. Flags instead of a queue: a queue
  // needs one same-address atomic per wave, which serialises when most waves
  // need the network (262K per pass at 16M particles).
  const bool from_queue = kNet && wq != nullptr;
  const int64_t nw = total_waves;
  uint32_t nproc = 0;
  __shared__ float4 nbias[kNet ? kCWaves : 1][kNet ? kNetBias / 4 : 1];
  __shared__ v8bf nw2[kNet ? kNetW2 : 1];
  NetRegs r;
  if constexpr (kNet) load_net(P, r, reinterpret_cast<float*>(nbias[wave]), nw2);
  const int64_t nwaves = (int64_t)gridDim.x * kCWaves;
  for (int64_t i = (int64_t)blockIdx.x * kCWaves + wave; i < nw; i += nwaves) {
    const int64_t w = i;
    if (from_queue) {
      if (uniform((int32_t)wq[w]) == 0) continue;
      ++nproc;
    }
    const int64_t p0 = w * kWave;
    const int32_t cw = (int32_t)min((int64_t)kWave, ps.n - p0);
    const int64_t p = p0 + lane;
    float F[9], PF[9];
    float lj = lane < cw ? ps.logJp[p] : 0.f;
    wave_load_rows<9>(ps.F, p0, cw, stage, F);
    const Sym A = c_minus_i(F);
    float f[kIn];
    const float nrm = closure_features(A, lj, m, f);
    float psi[3] = {0.f, 0.f, 0.f};
    bool use_net = kNet;
    if constexpr (kNet) {
      use_net = from_queue || always || ballot(lane < cw && !zero_start_certifies(nrm, tol)) != 0;
    } else if (wq != nullptr) {
      const bool net_wave = ballot(lane < cw && !zero_start_certifies(nrm, tol)) != 0;
      if (lane == 0) wq[w] = net_wave ? 1u : 0u;
      if (net_wave) continue;  // the network pass computes this wave
    }
    if (use_net) {
      // features of particle 16 t + (l & 15) in the B layout: through LDS
      float* sf = reinterpret_cast<float*>(stage4[wave]);  // 64 x 9 floats fit the row stage
#pragma unroll
      for (int k = 0; k < kIn; ++k) sf[lane * kIn + k] = f[k];
      wave_lds_barrier();
      const int j = lane & 15, g = lane >> 4;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        float x[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = 4 * g + e;
          x[e] = k < kIn ? sf[(16 * t + j) * kIn + k] : 0.f;
        }
        v4f y = {0.f, 0.f, 0.f, 0.f};
        if constexpr (kNet) y = net_tile<false>(r, pack4(x), nullptr);
        if (g == 0) {
#pragma unroll
          for (int e = 0; e < 3; ++e) sy[wave][16 * t + j][e] = y[e];
        }
      }
      wave_lds_barrier();
#pragma unroll
      for (int e = 0; e < 3; ++e) psi[e] = sy[wave][lane][e];
      wave_lds_barrier();  // the stage is reused by the row store below
    }
    bool ok = false;
    if (lane < cw) {
      ok = closure_stress(m, F, A, nrm, psi, tol, lj, PF);
      if (ok) {
        ps.logJp[p] = lj;
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) PF[k] = 0.f;
      }
    }
    // the wave's rejected particles go to the fallback queue with one atomic
    const uint64_t rej = ballot(lane < cw && !ok);
    if (rej != 0) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(queue, (uint32_t)__popcll(rej));
      base = (uint32_t)__shfl((int)base, 0, kWave);
      if ((rej >> lane) & 1ull) fb[base + (uint32_t)__popcll(rej & ((1ull << lane) - 1ull))] = (int32_t)p;
    }
    wave_store_rows<9>(contrib, p0, cw, stage, PF);
  }
  if (from_queue && lane == 0 && nproc) atomicAdd(wq_n, nproc);
}

__global__ void reset_queue_kernel(uint32_t* queue, uint32_t* queue2) {
  *queue = 0u;
  if (queue2) *queue2 = 0u;
}

// the plain SVD return mapping for the queued particles; also adds the
// pass's network waves (the queue length netq, or net_all) to the count
__global__ __launch_bounds__(256) void fallback_kernel(Model m, ParticlesView ps, float* __restrict__ contrib,
                                                       const int32_t* __restrict__ fb, const uint32_t* queue,
                                                       uint32_t* counters, const uint32_t* netq, uint32_t net_all) {
  const uint32_t nq = min(*queue, (uint32_t)ps.n);
  uint32_t fp64_used = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    const int64_t p = fb[i];
    float F[9], PF[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = ps.F[9 * p + k];
    float lj = ps.logJp[p];
    bool f64 = false;
    stress_sand(m, F, lj, PF, nullptr, f64);
    ps.logJp[p] = lj;
#pragma unroll
    for (int k = 0; k < 9; ++k) contrib[9 * p + k] = PF[k];
    fp64_used += f64 ? 1u : 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(&counters[kCtrFallback], nq);
    const uint32_t nw = netq ? *netq : net_all;
    if (nw) atomicAdd(&counters[kCtrNetWaves], nw);
  }
  if (fp64_used) atomicAdd(&counters[kCtrFp64], fp64_used);
}

int groups(int64_t threads, int per) { return (int)((threads + per - 1) / per); }

}  // namespace

hipError_t mlp_infer(const float* P, const float* X, int64_t n, float* Y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t tiles = (n + 15) / 16;
  const int64_t want = (tiles + 3) / 4;
  const int g = (int)(want < 4096 ? want : 4096);
  mlp_infer_kernel<<<g, 256, 0, s>>>(P, X, n, Y);
  return hipGetLastError();
}

hipError_t mlp_grad(const float* P, const float* X, const float* R, int64_t n, float* partial, float* loss_part,
                    int G, hipStream_t s) {
  if (n <= 0 || G <= 0) return hipErrorInvalidValue;
  mlp_grad_kernel<<<G, 64, 0, s>>>(P, X, R, n, partial, loss_part);
  return hipGetLastError();
}

hipError_t mlp_reduce(float* P, float* m, float* v, const float* partial, const float* loss_part, int G, float lr,
                      float bc1, float bc2, float* grad_out, float* loss_out, hipStream_t s) {
  mlp_reduce_kernel<<<groups(kNP, kRedParams), 256, 0, s>>>(P, m, v, partial, loss_part, G, lr, bc1, bc2,
                                                            grad_out, loss_out);
  return hipGetLastError();
}

hipError_t mlp_harvest(const Model& m, const float* F, const float* logJp, int64_t n, float* X, float* R,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  harvest_kernel<<<groups(n, 256), 256, 0, s>>>(m, F, logJp, n, X, R);
  return hipGetLastError();
}

hipError_t stress_closure(const Model& m, const ClosureView& cl, const ParticlesView& ps, float* contrib,
                          Workspace& ws, hipStream_t s) {
  if (ps.n <= 0) return hipSuccess;
  uint32_t* queue = ws.counters + kCtrFbQueue;
  uint32_t* wq_n = ws.counters + kCtrNetQueue;
  // With a network and without always_net: the zero-start pass takes every
  // wave (80 VGPRs, six waves per SIMD) and flags the waves it cannot
  // certify; the network pass then takes only those. Each wave gets exactly
  // what the single network pass would give it.
  const bool two_pass = cl.params && !cl.always_net;
  reset_queue_kernel<<<1, 1, 0, s>>>(queue, two_pass ? wq_n : nullptr);
  const int64_t waves = (ps.n + kWave - 1) / kWave;
  const int want = groups(waves * kWave, kCThreads);
  const int g_zero = want < 65536 ? want : 65536, g_net = want < 2048 ? want : 2048;
  if (two_pass) {
    closure_kernel<false><<<g_zero, kCThreads, 0, s>>>(m, ps, contrib, nullptr, cl.tol, 0, ws.fb, queue, ws.wq, wq_n);
    // 1024 workgroups of four waves fill every SIMD at the network's four
    // waves per SIMD; they read every wave's flag and take the flagged ones
    const int g_queue = g_net < 1024 ? g_net : 1024;
    closure_kernel<true><<<g_queue, kCThreads, 0, s>>>(m, ps, contrib, cl.params, cl.tol, 0, ws.fb, queue, ws.wq,
                                                       wq_n);
  } else if (cl.params) {
    closure_kernel<true><<<g_net, kCThreads, 0, s>>>(m, ps, contrib, cl.params, cl.tol, cl.always_net, ws.fb, queue,
                                                     nullptr, nullptr);
  } else {
    closure_kernel<false><<<g_zero, kCThreads, 0, s>>>(m, ps, contrib, nullptr, cl.tol, 0, ws.fb, queue, nullptr,
                                                       nullptr);
  }
  const uint32_t net_all = (cl.params && !two_pass) ? (uint32_t)waves : 0u;
  fallback_kernel<<<persistent_groups(), 256, 0, s>>>(m, ps, contrib, ws.fb, queue, ws.counters,
                                                      two_pass ? wq_n : nullptr, net_all);
  return hipGetLastError();
}

hipError_t closure_launch_limits() {
  const void* kernels[] = {reinterpret_cast<const void*>(&mlp_infer_kernel),
                           reinterpret_cast<const void*>(&mlp_grad_kernel),
                           reinterpret_cast<const void*>(&closure_kernel<true>),
                           reinterpret_cast<const void*>(&closure_kernel<false>),
                           reinterpret_cast<const void*>(&fallback_kernel)};
  const int threads[] = {256, 64, kCThreads, kCThreads, 256};
  for (int k = 0; k < 5; ++k) {
    hipFuncAttributes attr;
    const hipError_t e = hipFuncGetAttributes(&attr, kernels[k]);
    if (e != hipSuccess) return e;
    if (attr.maxThreadsPerBlock < threads[k]) return hipErrorInvalidConfiguration;
  }
  return hipSuccess;
}

}  // namespace zsmpm
