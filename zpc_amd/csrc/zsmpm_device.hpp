// Device-side building blocks of the gfx950 MLS-MPM substep: block keys and
// the zpc table hash, 64-lane wavefront helpers, the 3x3 SVD and the two
// constitutive models. Everything here is __device__ code for CDNA4 (wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rowspan.hpp"

namespace zsmpm {

constexpr int kWave = 64;
constexpr int kChannels = 7;        // m, v(3), rhs(3): Grids{{m,1},{v,3},{rhs,3}}
constexpr int kBlockCells = 64;     // 4^3 cells per block
constexpr int32_t kKeySentinel = 0x7fffffff;
constexpr int32_t kProbeStep = 127;
constexpr uint64_t kEmpty64 = ~0ull;
constexpr int32_t kPackBias = 1 << 20;  // packed keys hold components in [-2^20, 2^20)

// error bits raised by kernels, read back by the host at synchronisation points
enum : uint32_t {
  kErrTableFull = 1u << 0,    // a probe walked the whole table
  kErrKeyRange = 1u << 1,     // a block key does not fit the 21-bit packing
  kErrArena = 1u << 2,        // a stencil fell outside its block's 8^3 arena
  kErrCapacity = 1u << 3,     // more blocks / runs than the buffers hold
  kErrMissingBlock = 1u << 4  // a neighbour block was not active (table not enlarged)
};

// ---------------------------------------------------------------------------
// zpc's table hash (container/HashTable.hpp:496-500 + math/Hash.hpp:17-27):
// 64-bit hash_combine seeded by the sign-extended first component, truncated
// to i32, reduced with (h % T + T) % T (:359).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t zpc_hash(int32_t k0, int32_t k1, int32_t k2) {
  uint64_t r = (uint64_t)(int64_t)k0;
  r ^= (uint64_t)(int64_t)k1 + 0x9e3779b97f4a7c15ull + (r << 12) + (r >> 4);
  r ^= (uint64_t)(int64_t)k2 + 0x9e3779b97f4a7c15ull + (r << 12) + (r >> 4);
  return (int32_t)(uint32_t)r;
}
__device__ __forceinline__ int32_t zpc_home(int32_t k0, int32_t k1, int32_t k2, int32_t T) {
  const int32_t h = zpc_hash(k0, k1, k2);
  return (h % T + T) % T;
}

// Block key of a position (SparsityOp.hpp:72-77): c = floor(x*dxinv + disp) + offset,
// key = floor(c / blockLen). Rounded exactly as written (no contraction), so the
// integer result matches the host path bit for bit.
__device__ __forceinline__ void block_key(float x0, float x1, float x2, float dxinv, int blockLen,
                                          int offset, float disp, int32_t& k0, int32_t& k1,
                                          int32_t& k2) {
#pragma clang fp contract(off)
  const float xs[3] = {x0, x1, x2};
  int32_t out[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float t = xs[d] * dxinv;
    const int32_t c = (int32_t)floorf(t + disp) + offset;
    if ((blockLen & (blockLen - 1)) == 0) {
      out[d] = c >> (__ffs(blockLen) - 1);  // floor division by a power of two
    } else {
      const int32_t b = c + (c < 0 ? -blockLen + 1 : 0);
      out[d] = b / blockLen;
    }
  }
  k0 = out[0];
  k1 = out[1];
  k2 = out[2];
}

__device__ __forceinline__ bool key_packable(int32_t k0, int32_t k1, int32_t k2) {
  return k0 >= -kPackBias && k0 < kPackBias && k1 >= -kPackBias && k1 < kPackBias &&
         k2 >= -kPackBias && k2 < kPackBias;
}
__device__ __forceinline__ uint64_t pack_key(int32_t k0, int32_t k1, int32_t k2) {
  return ((uint64_t)(uint32_t)(k0 + kPackBias) << 42) | ((uint64_t)(uint32_t)(k1 + kPackBias) << 21) |
         (uint64_t)(uint32_t)(k2 + kPackBias);
}
__device__ __forceinline__ void unpack_key(uint64_t p, int32_t& k0, int32_t& k1, int32_t& k2) {
  k0 = (int32_t)((p >> 42) & 0x1fffff) - kPackBias;
  k1 = (int32_t)((p >> 21) & 0x1fffff) - kPackBias;
  k2 = (int32_t)(p & 0x1fffff) - kPackBias;
}
// murmur3 finaliser: slot hash of the scratch (dedup) table
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z ^= z >> 33;
  z *= 0xff51afd7ed558ccdull;
  z ^= z >> 33;
  z *= 0xc4ceb9fe1a85ec53ull;
  z ^= z >> 33;
  return z;
}

// ---------------------------------------------------------------------------
// wavefront helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) {
  return __shfl(v, src, kWave);
}

// Value of the first active lane as a wave-uniform (scalar) value. Called
// with the whole wave active, that is lane 0.
__device__ __forceinline__ int32_t uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Wave-uniform, called with the whole wave active: do the valid lanes hold
// one contiguous run p0 + lane (lane 0 valid)? p0 receives lane 0's index.
__device__ __forceinline__ bool wave_contiguous(int64_t p, bool valid, int64_t& p0) {
  p0 = uniform(p);
  const uint64_t vm = ballot(valid);
  return (vm & 1ull) != 0 && ballot(valid && p == p0 + lane_id()) == vm;
}

// Orders a wave's own LDS writes before its later LDS reads by other lanes
// (the LDS serves one wave's instructions in issue order; this keeps the
// compiler from moving accesses across the point).
__device__ __forceinline__ void wave_lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// Particle rows are streamed: every kernel reads a row once and writes it
// once, and the next reader is a later kernel. Their 16-byte vector loads and
// stores are non-temporal, so the streams do not evict what is re-used from
// L2 (staged grid velocities, table and bucket arrays). On the 16M sand
// column this took the substep from 2.46 to 2.36 ms (G2P 0.86 -> 0.82 ms).
typedef float zs_f4 __attribute__((ext_vector_type(4)));

// Split-phase load of rows q0 .. q0+cnt-1 (cnt <= 64) of W floats: issue()
// puts the wave's global vector loads in flight (no LDS, no wait), finish()
// transposes them through the wave's LDS stage so lane l < cnt receives row
// q0 + l. Several RowLoads issued back to back cost one memory round trip.
// Both calls are made by the whole wave with wave-uniform q0, cnt.
template <int W>
struct RowLoad {
  static constexpr int kVec = (W * kWave / 4 + kWave - 1) / kWave;  // float4 per lane
  RowSpan r;
  float4 v[kVec];
  float head, tail;
  __device__ __forceinline__ void issue(const float* __restrict__ base, int64_t q0, int32_t cnt) {
    const int lane = lane_id();
    r = row_span((uint64_t)(uintptr_t)base, q0, cnt, W);
    const int64_t hn = r.a4 - r.a;
    const int64_t t0 = r.a4 + 4 * r.nvec;
    head = lane < hn ? base[r.a + lane] : 0.f;
    tail = lane < r.b - t0 ? base[t0 + lane] : 0.f;
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
      const int64_t j = lane + (int64_t)k * kWave;
      if (j < r.nvec) {
        const zs_f4 t = __builtin_nontemporal_load(reinterpret_cast<const zs_f4*>(base + r.a4 + 4 * j));
        v[k] = make_float4(t.x, t.y, t.z, t.w);
      } else {
        v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }
  __device__ __forceinline__ void finish(float* stage, int32_t cnt, float (&out)[W]) {
    const int lane = lane_id();
    const int64_t hn = r.a4 - r.a;
    const int64_t t0 = r.a4 + 4 * r.nvec;
    if (lane < hn) stage[r.shift + lane] = head;
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
      const int64_t j = lane + (int64_t)k * kWave;
      if (j < r.nvec) *reinterpret_cast<float4*>(stage + r.shift + hn + 4 * j) = v[k];
    }
    if (lane < r.b - t0) stage[r.shift + (t0 - r.a) + lane] = tail;
    wave_lds_barrier();
    if (lane < cnt) {
#pragma unroll
      for (int k = 0; k < W; ++k) out[k] = stage[r.shift + lane * W + k];
    }
    wave_lds_barrier();
  }
};

// One-shot form: issue and finish.
template <int W>
__device__ __forceinline__ void wave_load_rows(const float* __restrict__ base, int64_t q0, int32_t cnt,
                                               float* stage, float (&out)[W]) {
  RowLoad<W> ld;
  ld.issue(base, q0, cnt);
  ld.finish(stage, cnt, out);
}

// The store counterpart: lane l < cnt provides row q0 + l.
template <int W>
__device__ __forceinline__ void wave_store_rows(float* __restrict__ base, int64_t q0, int32_t cnt, float* stage,
                                                const float (&in)[W]) {
  const int lane = lane_id();
  const RowSpan r = row_span((uint64_t)(uintptr_t)base, q0, cnt, W);
  if (lane < cnt) {
#pragma unroll
    for (int k = 0; k < W; ++k) stage[r.shift + lane * W + k] = in[k];
  }
  wave_lds_barrier();
  const int64_t head = r.a4 - r.a;
  if (lane < head) base[r.a + lane] = stage[r.shift + lane];
  for (int64_t j = lane; j < r.nvec; j += kWave)
    __builtin_nontemporal_store(*reinterpret_cast<const zs_f4*>(stage + r.shift + head + 4 * j),
                                reinterpret_cast<zs_f4*>(base + r.a4 + 4 * j));
  const int64_t t0 = r.a4 + 4 * r.nvec;
  if (lane < r.b - t0) base[t0 + lane] = stage[r.shift + (t0 - r.a) + lane];
  wave_lds_barrier();
}

// ---------------------------------------------------------------------------
// 3x3 SVD: McAdams et al. 2011 ("Computing the SVD of 3x3 matrices with
// minimal branching and elementary floating point operations"), the method
// zpc uses (math/matrix/SVD.hpp). Four sweeps of approximate Jacobi
// conjugations on A^T A accumulate V as a quaternion; B = A V is column-sorted
// by norm (V kept a rotation by sign flips) and Givens QR gives U and the
// singular values (sigma_3 carries the sign of det A). Matrices are row-major
// a[r][c] here.
// ---------------------------------------------------------------------------
struct Svd3 {
  static constexpr float kFourGammaSq = 5.8284273147583007813f;  // 4 (3 + 2 sqrt 2) in float
  static constexpr float kSinPi8 = 0.3826834261417388916f;       // sin(pi/8) rounded to float
  static constexpr float kCosPi8 = 0.9238795042037963867f;       // cos(pi/8) rounded to float
  static constexpr float kTiny = 1.e-20f;
  static constexpr float kSmall = 1.e-12f;

  // hardware reciprocal square root (v_rsq_f32, about 1 ulp): zpc's SVD calls
  // zs::rsqrt (math/matrix/SVD.hpp:113), which is ::rsqrtf on its CUDA backend
  // and 1/sqrtf on the host (ZpcMathUtils.hpp:805-846)
  __device__ static __forceinline__ float rsq(float v) { return __builtin_amdgcn_rsqf(v); }

  // approximate Jacobi conjugation for pivot (P, Q), third index R (quaternion axis)
  template <int P, int Q, int R>
  __device__ static __forceinline__ void conjugate(float (&s)[3][3], float& qs, float (&qv)[3]) {
    float sh = s[Q][P] * 0.5f;
    float t5 = s[P][P] - s[Q][Q];
    const bool big = sh * sh >= kTiny;
    sh = big ? sh : 0.0f;
    float ch = big ? t5 : 1.0f;
    float t1 = sh * sh;
    float t2 = ch * ch;
    const float w = rsq(t1 + t2);
    sh = w * sh;
    ch = w * ch;
    const bool pi8 = t2 <= kFourGammaSq * t1;
    sh = pi8 ? kSinPi8 : sh;
    ch = pi8 ? kCosPi8 : ch;
    t1 = sh * sh;
    t2 = ch * ch;
    const float c = t2 - t1;
    float sn = ch * sh;
    sn = sn + sn;
    const float sc = t1 + t2;
    s[R][R] = s[R][R] * sc * sc;
    float rp = s[R][P] * sc, rq = s[R][Q] * sc;
    const float a1 = sn * rp, a2 = sn * rq;
    rp = c * rp + a2;
    rq = c * rq - a1;
    const float ss = sn * sn;
    const float cc = c * c;
    const float pp = s[P][P], qq = s[Q][Q], qp = s[Q][P];
    const float cs = c * sn;
    const float two_qp_cs = (qp + qp) * cs;
    s[P][P] = pp * cc + qq * ss + two_qp_cs;
    s[Q][Q] = qq * cc + pp * ss - two_qp_cs;
    s[Q][P] = qp * (cc - ss) - t5 * cs;
    s[P][Q] = s[Q][P];
    s[R][P] = rp;
    s[P][R] = rp;
    s[R][Q] = rq;
    s[Q][R] = rq;
    constexpr int A = R, B = (R + 1) % 3, Cc = (R + 2) % 3;
    const float u1 = sh * qv[B], u2 = sh * qv[Cc], u3 = sh * qv[A], shs = sh * qs;
    qs = ch * qs - u3;
    const float na = ch * qv[A] + shs;
    const float nb = ch * qv[B] + u2;
    const float nc = ch * qv[Cc] - u1;
    qv[A] = na;
    qv[B] = nb;
    qv[Cc] = nc;
  }

  template <int I, int J, int NEG>
  __device__ static __forceinline__ void sort_step(float (&b)[3][3], float (&v)[3][3], float (&n)[3]) {
    const bool sw = n[I] < n[J];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float bi = b[r][I], bj = b[r][J], vi = v[r][I], vj = v[r][J];
      b[r][I] = sw ? bj : bi;
      b[r][J] = sw ? bi : bj;
      v[r][I] = sw ? vj : vi;
      v[r][J] = sw ? vi : vj;
    }
    const float ni = n[I], nj = n[J];
    n[I] = sw ? nj : ni;
    n[J] = sw ? ni : nj;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      b[r][NEG] = sw ? -b[r][NEG] : b[r][NEG];
      v[r][NEG] = sw ? -v[r][NEG] : v[r][NEG];
    }
  }

  // one Newton step on rsqrt: r (1.5 - 0.5 x r^2), written as the method does
  __device__ static __forceinline__ float newton_rsq(float x) {
    const float r = rsq(x);
    const float h = r * 0.5f;
    return (r + h) - x * (r * (r * h));
  }

  template <int P, int Q>
  __device__ static __forceinline__ void givens_qr(float (&b)[3][3], float (&u)[3][3]) {
    const float a_qp = b[Q][P], a_pp = b[P][P];
    float sh = (a_qp * a_qp >= kSmall) ? a_qp : 0.0f;
    float ch = fmaxf(fmaxf(-a_pp, a_pp), kSmall);
    const bool nonneg = a_pp >= 0.f;
    const float rr = ch * ch + sh * sh;
    ch = ch + newton_rsq(rr) * rr;
    const float tch = nonneg ? ch : sh;
    const float tsh = nonneg ? sh : ch;
    const float w = newton_rsq(tch * tch + tsh * tsh);
    ch = tch * w;
    sh = tsh * w;
    const float c = ch * ch - sh * sh;
    float sn = sh * ch;
    sn = sn + sn;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float bp = b[P][j], bq = b[Q][j];
      b[P][j] = c * bp + sn * bq;
      b[Q][j] = c * bq - sn * bp;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float up = u[i][P], uq = u[i][Q];
      u[i][P] = c * up + sn * uq;
      u[i][Q] = c * uq - sn * up;
    }
  }

  __device__ static __forceinline__ void svd(const float (&a)[3][3], float (&U)[3][3], float (&S)[3],
                                             float (&V)[3][3]) {
    float s[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const float t = a[0][i] * a[0][j] + a[1][i] * a[1][j] + a[2][i] * a[2][j];
        s[i][j] = t;
        s[j][i] = t;
      }
    float qs = 1.f, qv[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int sweep = 0; sweep < 4; ++sweep) {
      conjugate<0, 1, 2>(s, qs, qv);
      conjugate<1, 2, 0>(s, qs, qv);
      conjugate<2, 0, 1>(s, qs, qv);
    }
    const float qn = newton_rsq(qs * qs + qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2]);
    qs *= qn;
    qv[0] *= qn;
    qv[1] *= qn;
    qv[2] *= qn;
    const float xx = qv[0] * qv[0], yy = qv[1] * qv[1], zz = qv[2] * qv[2], ww = qs * qs;
    float v[3][3];
    v[0][0] = ww + xx - yy - zz;
    v[1][1] = ww - xx + yy - zz;
    v[2][2] = ww - xx - yy + zz;
    const float tx = qv[0] + qv[0], ty = qv[1] + qv[1], tz = qv[2] + qv[2];
    const float sx = qs * tx, sy = qs * ty, sz = qs * tz;
    const float xy = qv[1] * tx, yz = qv[2] * ty, xz = qv[0] * tz;
    v[0][1] = xy - sz;
    v[1][0] = xy + sz;
    v[1][2] = yz - sx;
    v[2][1] = yz + sx;
    v[2][0] = xz - sy;
    v[0][2] = xz + sy;
    float b[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) b[r][c] = a[r][0] * v[0][c] + a[r][1] * v[1][c] + a[r][2] * v[2][c];
    float n[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = b[0][c] * b[0][c] + b[1][c] * b[1][c] + b[2][c] * b[2][c];
    sort_step<0, 1, 1>(b, v, n);
    sort_step<0, 2, 0>(b, v, n);
    sort_step<1, 2, 2>(b, v, n);
    float u[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    givens_qr<0, 1>(b, u);
    givens_qr<0, 2>(b, u);
    givens_qr<1, 2>(b, u);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        U[i][j] = u[i][j];
        V[i][j] = v[i][j];
      }
    S[0] = b[0][0];
    S[1] = b[1][1];
    S[2] = b[2][2];
  }
};

// FP64 SVD for nearly repeated singular values: cyclic Jacobi with exact
// rotations on A^T A until the off-diagonal mass is below 1e-34 of the
// diagonal, then the same column sort and Givens QR conventions.
__device__ inline void svd3_fp64(const float (&af)[3][3], float (&U)[3][3], float (&S)[3],
                                 float (&V)[3][3]) {
  double a[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) a[i][j] = (double)af[i][j];
  double s[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s[i][j] = a[0][i] * a[0][j] + a[1][i] * a[1][j] + a[2][i] * a[2][j];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = s[0][1] * s[0][1] + s[0][2] * s[0][2] + s[1][2] * s[1][2];
    const double dia = s[0][0] * s[0][0] + s[1][1] * s[1][1] + s[2][2] * s[2][2];
    if (off <= 1e-34 * dia) break;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int p = (k == 2) ? 0 : k, q = (k == 0) ? 1 : 2;
      const double spq = s[p][q];
      if (spq == 0.0) continue;
      const double theta = (s[q][q] - s[p][p]) / (2.0 * spq);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double srp = s[r][p], srq = s[r][q];
        s[r][p] = c * srp - sn * srq;
        s[r][q] = sn * srp + c * srq;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double spr = s[p][r], sqr = s[q][r];
        s[p][r] = c * spr - sn * sqr;
        s[q][r] = sn * spr + c * sqr;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double vp = v[r][p], vq = v[r][q];
        v[r][p] = c * vp - sn * vq;
        v[r][q] = sn * vp + c * vq;
      }
    }
  }
  double b[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) b[r][c] = a[r][0] * v[0][c] + a[r][1] * v[1][c] + a[r][2] * v[2][c];
  double n[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) n[c] = b[0][c] * b[0][c] + b[1][c] * b[1][c] + b[2][c] * b[2][c];
  auto sort_step = [&](int i, int j, int neg) {
    if (n[i] < n[j]) {
      for (int r = 0; r < 3; ++r) {
        double t = b[r][i]; b[r][i] = b[r][j]; b[r][j] = t;
        t = v[r][i]; v[r][i] = v[r][j]; v[r][j] = t;
      }
      double t = n[i]; n[i] = n[j]; n[j] = t;
      for (int r = 0; r < 3; ++r) { b[r][neg] = -b[r][neg]; v[r][neg] = -v[r][neg]; }
    }
  };
  sort_step(0, 1, 1);
  sort_step(0, 2, 0);
  sort_step(1, 2, 2);
  double u[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  auto givens = [&](int p, int q) {
    const double x = b[p][p], y = b[q][p];
    const double r = sqrt(x * x + y * y);
    double c = 1.0, sn = 0.0;  // c = x / r, s = y / r leaves the pivot at r >= 0
    if (r != 0.0) {
      c = x / r;
      sn = y / r;
    }
    for (int j = 0; j < 3; ++j) {
      const double bp = b[p][j], bq = b[q][j];
      b[p][j] = c * bp + sn * bq;
      b[q][j] = -sn * bp + c * bq;
    }
    for (int i = 0; i < 3; ++i) {
      const double up = u[i][p], uq = u[i][q];
      u[i][p] = c * up + sn * uq;
      u[i][q] = -sn * up + c * uq;
    }
  };
  givens(0, 1);
  givens(0, 2);
  givens(1, 2);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      U[i][j] = (float)u[i][j];
      V[i][j] = (float)v[i][j];
    }
  S[0] = (float)b[0][0];
  S[1] = (float)b[1][1];
  S[2] = (float)b[2][2];
}

// SVD of a column-major F (F[i + 3j] = F_ij) with the FP64 path for nearly
// repeated singular values: when min_i |sigma_i - sigma_{i+1}| < gap * |sigma_1|
// (gap > 0) the FP32 result is replaced by the FP64 one.
__device__ __forceinline__ void svd_colmajor(const float* F, float (&U)[3][3], float (&S)[3],
                                             float (&V)[3][3], float fp64_gap, bool& used_fp64) {
  float a[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) a[i][j] = F[i + 3 * j];
  Svd3::svd(a, U, S, V);
  used_fp64 = false;
  if (fp64_gap > 0.f) {
    const float g = fminf(fabsf(S[0] - S[1]), fabsf(S[1] - S[2]));
    if (g < fp64_gap * fabsf(S[0])) {
      svd3_fp64(a, U, S, V);
      used_fp64 = true;
    }
  }
}

// ---------------------------------------------------------------------------
// constitutive models (physics/ConstitutiveModel_Vol_dP.hpp:10-46, 241-325)
// ---------------------------------------------------------------------------
struct Model {
  int32_t kind;  // 0 fixed-corotated, 1 Drucker-Prager sand
  float volume, mu, lambda, cohesion, beta, yield_surface;
  int32_t volume_correction;
  float fp64_gap;
};

// P F^T volume for P = U diag(Ph) V^T (the tail both models share).
__device__ __forceinline__ void pft_volume(const float (&U)[3][3], const float (&Ph)[3],
                                           const float (&V)[3][3], const float* F, float vol,
                                           float* PF) {
  float P[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = Ph[0] * U[i][0] * V[j][0] + Ph[1] * U[i][1] * V[j][1] + Ph[2] * U[i][2] * V[j][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      PF[i + 3 * j] = (P[i][0] * F[j] + P[i][1] * F[j + 3] + P[i][2] * F[j + 6]) * vol;
}

// Fixed-corotated: Ph_i = 2 mu (s_i - 1) + lambda (J - 1) prod_{j!=i} s_j.
__device__ __forceinline__ void stress_fixedcorotated(const Model& m, const float* F, float* PF,
                                                      bool& fp64) {
  float U[3][3], S[3], V[3][3];
  svd_colmajor(F, U, S, V, m.fp64_gap, fp64);
  const float J = S[0] * S[1] * S[2];
  const float smu = 2.f * m.mu;
  const float slam = m.lambda * (J - 1.f);
  const float Ph[3] = {smu * (S[0] - 1.f) + slam * (S[1] * S[2]), smu * (S[1] - 1.f) + slam * (S[0] * S[2]),
                       smu * (S[2] - 1.f) + slam * (S[0] * S[1])};
  pft_volume(U, Ph, V, F, m.volume, PF);
}

// Drucker-Prager with volume correction; F is projected locally (the stored F
// is left alone), logJp is updated. Returns the case: 1 inside the cone,
// 2 cone tip, 3 projected to the cone surface, 0 otherwise.
__device__ __forceinline__ int stress_sand(const Model& m, const float* F, float& logJp, float* PF,
                                           float* Fproj, bool& fp64) {
  float U[3][3], S[3], V[3][3];
  svd_colmajor(F, U, S, V, m.fp64_gap, fp64);
  const float smu = 2.f * m.mu;
  float eps[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float a = fmaxf(fabsf(S[i]), 1e-4f);
    eps[i] = logf(a) - m.cohesion;
  }
  const float sum_eps = eps[0] + eps[1] + eps[2];
  const float tr = sum_eps + logJp;
  const float third = tr / 3.f;
  const float eh[3] = {eps[0] - third, eps[1] - third, eps[2] - third};
  const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
  float ns[3] = {0.f, 0.f, 0.f};
  int which = 0;
  if (tr >= 0.f) {
    const float e = expf(m.cohesion);
    ns[0] = ns[1] = ns[2] = e;
    if (m.volume_correction) logJp = m.beta * sum_eps + logJp;
    which = 2;
  } else if (m.mu != 0.f) {
    logJp = 0.f;
    const float dg = ehn + (3.f * m.lambda + smu) / smu * tr * m.yield_surface;
    if (dg <= 0.f) {
#pragma unroll
      for (int i = 0; i < 3; ++i) ns[i] = expf(eps[i] + m.cohesion);
      which = 1;
    } else {
      const float r = dg / ehn;
#pragma unroll
      for (int i = 0; i < 3; ++i) ns[i] = expf(eps[i] - r * eh[i] + m.cohesion);
      which = 3;
    }
  }
  if (which != 0 && !Fproj) {
    // With the projected F = U diag(ns) V^T and P = U diag(Ph) V^T,
    // P F^T = U diag(Ph) V^T V diag(ns) U^T = U diag(tau) U^T (V^T V = I to
    // rounding), tau_i = Ph_i ns_i = 2 mu ln ns_i + lambda sum_j ln ns_j, the
    // Kirchhoff stress: no division, and ln ns_i is the exponent the
    // projection chose (no exp / log round trip). Symmetric: 6 entries. It
    // agrees with the composed form to 1e-6 of the largest entry.
    float H[3];
    if (which == 2) {
      H[0] = H[1] = H[2] = m.cohesion;
    } else if (which == 1) {
#pragma unroll
      for (int i = 0; i < 3; ++i) H[i] = eps[i] + m.cohesion;
    } else {
      const float r = (ehn + (3.f * m.lambda + smu) / smu * tr * m.yield_surface) / ehn;
#pragma unroll
      for (int i = 0; i < 3; ++i) H[i] = eps[i] - r * eh[i] + m.cohesion;
    }
    const float trl = H[0] + H[1] + H[2];
    const float tau[3] = {(smu * H[0] + m.lambda * trl) * m.volume, (smu * H[1] + m.lambda * trl) * m.volume,
                          (smu * H[2] + m.lambda * trl) * m.volume};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const float e = U[i][0] * tau[0] * U[j][0] + U[i][1] * tau[1] * U[j][1] + U[i][2] * tau[2] * U[j][2];
        PF[i + 3 * j] = e;
        PF[j + 3 * i] = e;
      }
    return which;
  }
  // projected F = U diag(ns) V^T (column-major)
  float Fp[9];
  if (which != 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Fp[i + 3 * j] = U[i][0] * ns[0] * V[j][0] + U[i][1] * ns[1] * V[j][1] + U[i][2] * ns[2] * V[j][2];
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) Fp[i] = F[i];
  }
  const float ls[3] = {logf(ns[0]), logf(ns[1]), logf(ns[2])};
  const float trl = ls[0] + ls[1] + ls[2];
  const float Ph[3] = {(smu * ls[0] + m.lambda * trl) / ns[0], (smu * ls[1] + m.lambda * trl) / ns[1],
                       (smu * ls[2] + m.lambda * trl) / ns[2]};
  pft_volume(U, Ph, V, Fp, m.volume, PF);
  if (Fproj)
#pragma unroll
    for (int i = 0; i < 9; ++i) Fproj[i] = Fp[i];
  return which;
}

// ---------------------------------------------------------------------------
// quadratic B-spline arena (simulation/Utils.hpp:51-70): X = x / dx,
// base = floor(X - 0.5), local = X - base, weights per axis, local *= dx.
// ---------------------------------------------------------------------------
struct Arena {
  int32_t base[3];
  float lp[3];   // local position * dx
  float w[3][3]; // w[axis][node]
};

__device__ __forceinline__ void make_arena(float x0, float x1, float x2, float dx, Arena& ar) {
#pragma clang fp contract(off)
  const float xs[3] = {x0, x1, x2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float X = xs[d] / dx;
    const int32_t b = (int32_t)floorf(X - 0.5f);
    const float l = X - (float)b;
    const float d0 = l - (float)(int32_t)floorf(l - 0.5f);
    ar.base[d] = b;
    ar.w[d][0] = 0.5f * (1.5f - d0) * (1.5f - d0);
    const float d1 = d0 - 1.0f;
    ar.w[d][1] = 0.75f - d1 * d1;
    const float zz = 0.5f + d1;
    ar.w[d][2] = 0.5f * zz * zz;
    ar.lp[d] = l * dx;
  }
}

}  // namespace zsmpm
