// zsmpm oracle: TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of littlemine/zpc's explicit MLS-MPM substep on the sparse
// 4^3-block grid, with sequential (seq_exec) semantics. It is the checker the
// HIP path is compared against, and the OpenMP build of it (oracle_omp.cpp) is
// the host-core baseline. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product never links it.
//
// Every function cites the reference file:line it restates (paths are relative
// to zpc's include/zensim/). The file is compiled with -ffp-contract=off so the
// float expressions round exactly as written, in the order the reference writes
// them.
//
// Pinning (what checks each part, beyond retyping the formula):
//   do_hash, next_2pow, lame, svd3_mcadams, stress_*, arena weights/diff/coord,
//   mat3 products: bit-exact against oracle/_ref (the reference's own headers
//   compiled from /root/reference) and against golden vectors under
//   tests/golden/ generated from it.
//   Table insert/query, activation, enlarge, bucket sort: brute-force
//   set/first-occurrence/std::stable_sort restatements in the tests, plus the
//   table invariants (bijection, activeKeys[idx] == key, probe reachability).
//   P2G / grid update / boundary / G2P: the P2G force against a finite-
//   difference gradient of the discrete elastic energy (sign, dt, D^-1 and
//   the stress transpose), APIC angular-momentum transfer (C's orientation),
//   G2P against affine grid fields with a non-symmetric F (x, v, C and the
//   F-update order), free fall (gravity axis, sign and scale, maxVel),
//   slip/separate/sticky plane projections, conservation of mass and
//   momentum, a single-particle known answer.
//   svd3_jacobi64 (the FP64 twin's SVD): reconstruction / orthogonality /
//   ordering invariants and agreement with numpy's LAPACK SVD.
#pragma once
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace zso {

using i32 = int32_t;
using u32 = uint32_t;
using i64 = int64_t;
using u64 = uint64_t;

// ----------------------------------------------------------------------------
// hashing
// ----------------------------------------------------------------------------

/// container/HashTable.hpp:496-500 (do_hash) composed with math/Hash.hpp:17-27
/// (hash_combine for an 8-byte seed). The seed starts as key[0] sign-extended
/// to 64 bits; each later component is sign-extended before the add; the
/// result is truncated to i32.
inline i32 do_hash(i32 k0, i32 k1, i32 k2) {
  u64 ret = (u64)(i64)k0;
  ret ^= ((u64)(i64)k1 + 0x9e3779b97f4a7c15ull + (ret << 12) + (ret >> 4));
  ret ^= ((u64)(i64)k2 + 0x9e3779b97f4a7c15ull + (ret << 12) + (ret >> 4));
  return (i32)(u32)ret;
}

/// container/HashTable.hpp:359 / :381: (h % T + T) % T with C++ remainder.
inline i32 home_slot(i32 k0, i32 k1, i32 k2, i32 T) {
  i32 h = do_hash(k0, k1, k2);
  return (h % T + T) % T;
}

/// math/bit/Bits.h:166-184 (bit_length, bit_count, next_2pow).
inline i64 next_2pow(i64 n) {
  if (n <= 0) return 1;
  i64 m = n - 1, len = 0;
  while (m > 0) { m >>= 1; ++len; }
  return (i64)1 << len;
}

/// container/HashTable.hpp:88-91 (reserve_ratio_v = 16).
inline i64 evaluate_table_size(i64 entryCnt) {
  if (entryCnt == 0) return 0;
  return next_2pow(entryCnt) * 16;
}

constexpr i32 kKeySentinel = INT32_MAX;  // deduce_numeric_max<i32>, HashTable.hpp:67
constexpr i32 kProbeStep = 127;          // HashTable.hpp:363

/// HashTable<i32,3,int> in zpc's layout (HashTable.hpp:40-58, 200-205):
/// keys[T] (vec<i32,3>), indices[T], status[T], _cnt, _activeKeys[T].
struct Table {
  i32 T = 0;
  i32 cnt = 0;
  std::vector<i32> keys, indices, status, active;
  /// set when a query probe would land on slot == T, where the reference's
  /// wrap test (HashTable.hpp:453-454, '>' instead of '>=') reads keys[T].
  bool query_wrap_quirk_hit = false;

  void resize(i32 tableSize) {
    T = tableSize;
    keys.assign((size_t)3 * T, kKeySentinel);
    indices.assign(T, -1);
    status.assign(T, -1);
    active.assign((size_t)3 * T, 0);
    cnt = 0;
  }
  /// CleanSparsity, simulation/sparsity/SparsityOp.hpp:41-55.
  void clear() {
    std::fill(keys.begin(), keys.end(), kKeySentinel);
    std::fill(indices.begin(), indices.end(), -1);
    std::fill(status.begin(), status.end(), -1);
    cnt = 0;
  }
  bool slot_is(i32 s, const i32 k[3]) const {
    return keys[3 * s] == k[0] && keys[3 * s + 1] == k[1] && keys[3 * s + 2] == k[2];
  }
  bool slot_empty(i32 s) const {
    return keys[3 * s] == kKeySentinel && keys[3 * s + 1] == kKeySentinel &&
           keys[3 * s + 2] == kKeySentinel;
  }
  /// HashTableView::insert (host), container/HashTable.hpp:380-398, with the
  /// lock-protected compare/store of atomicKeyCAS (:546-564) taken serially.
  i32 insert(const i32 k[3]) {
    i32 s = home_slot(k[0], k[1], k[2], T);
    while (!(slot_empty(s) || slot_is(s, k))) s = (s + kProbeStep) % T;
    if (slot_empty(s)) {
      keys[3 * s] = k[0];
      keys[3 * s + 1] = k[1];
      keys[3 * s + 2] = k[2];
      i32 localno = cnt++;
      indices[s] = localno;
      active[3 * localno] = k[0];
      active[3 * localno + 1] = k[1];
      active[3 * localno + 2] = k[2];
      return localno;
    }
    return -1;
  }
  /// HashTableView::query, container/HashTable.hpp:447-456. The probe wraps
  /// with a true modulo (the same sequence insert uses); a probe that the
  /// reference would have sent to keys[T] is recorded in query_wrap_quirk_hit.
  i32 query(const i32 k[3]) {
    i32 s = home_slot(k[0], k[1], k[2], T);
    while (true) {
      if (slot_is(s, k)) return indices[s];
      if (indices[s] == -1) return -1;
      i64 nxt = (i64)s + kProbeStep;
      if (nxt == T) query_wrap_quirk_hit = true;
      s = (i32)(nxt % T);
    }
  }
};

// ----------------------------------------------------------------------------
// sparsity (block activation) and the particle->block bucket sort
// ----------------------------------------------------------------------------

/// ComputeSparsity / SpatiallyCount / SpatiallyDistribute key, SparsityOp.hpp:72-77
/// (also :138-143, :177-182): c = floor(x*dxinv + disp) + offset, then a
/// floor division by blockLen written as a biased truncating division.
template <typename T>
inline void block_key(const T* x, T dxinv, int blockLen, int offset, T disp, i32 out[3]) {
  for (int d = 0; d < 3; ++d) {
    i32 c = (i32)std::floor(x[d] * dxinv + disp) + offset;
    i32 b = c + (c < 0 ? -blockLen + 1 : 0);
    out[d] = b / blockLen;
  }
}

/// ComputeSparsity over range(N) in particle order, SparsityOp.hpp:57-86
/// (dxinv = (T)1.0 / dx at :65).
template <typename T>
void compute_sparsity(Table& t, const T* x, i64 n, T dx, int blockLen, int offset, T disp) {
  T dxinv = (T)1.0 / dx;
  i32 k[3];
  for (i64 p = 0; p < n; ++p) {
    block_key(x + 3 * p, dxinv, blockLen, offset, disp, k);
    t.insert(k);
  }
}

/// EnlargeSparsity over range(n0) (n0 = table size before the call),
/// SparsityOp.hpp:88-112: for each active key insert key + {lo..hi-1}^3 in
/// (dx, dy, dz) loop order.
inline void enlarge_sparsity(Table& t, const i32 lo[3], const i32 hi[3]) {
  const i32 n0 = t.cnt;
  for (i32 i = 0; i < n0; ++i) {
    const i32 b0 = t.active[3 * i], b1 = t.active[3 * i + 1], b2 = t.active[3 * i + 2];
    for (i32 dx = lo[0]; dx < hi[0]; ++dx)
      for (i32 dy = lo[1]; dy < hi[1]; ++dy)
        for (i32 dz = lo[2]; dz < hi[2]; ++dz) {
          i32 k[3] = {b0 + dx, b1 + dy, b2 + dz};
          t.insert(k);
        }
  }
}

/// index_buckets_for_particles (simulation/particle/Query.tpp:30-56) with the
/// activation table already built: SpatiallyCount (SparsityOp.hpp:114-151),
/// exclusive_scan (omp/execution/ExecutionPolicy.hpp:341-419) and
/// SpatiallyDistribute (:153-195), in seq order so each bucket lists its
/// particles by increasing index. counts/offsets hold table.cnt + 1 entries.
template <typename T>
void bucket_sort(Table& t, const T* x, i64 n, T dx, int blockLen, int offset, T disp,
                 i32* counts, i32* offsets, i32* indices) {
  const i32 nb = t.cnt + 1;
  T dxinv = (T)1.0 / dx;
  std::vector<i32> bucket(n);
  for (i32 b = 0; b < nb; ++b) counts[b] = 0;
  i32 k[3];
  for (i64 p = 0; p < n; ++p) {
    block_key(x + 3 * p, dxinv, blockLen, offset, disp, k);
    bucket[p] = t.query(k);
    counts[bucket[p]]++;
  }
  i32 run = 0;
  for (i32 b = 0; b < nb; ++b) { offsets[b] = run; run += counts[b]; }
  std::vector<i32> tmp(nb, 0);
  for (i64 p = 0; p < n; ++p) {
    i32 b = bucket[p];
    indices[offsets[b] + tmp[b]++] = (i32)p;
  }
}

// ----------------------------------------------------------------------------
// small matrix helpers (column-major vec<T,9>, math/matrix/MatrixUtils.h)
// ----------------------------------------------------------------------------

/// matmul_mat_diag_matT_3D, MatrixUtils.h:26-47: out = mat1 * diag * mat2^T.
template <typename T>
inline void matmul_mat_diag_matT(T* out, const T* m1, const T* dg, const T* m2) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      out[i + 3 * j] = m1[i] * dg[0] * m2[j] + m1[i + 3] * dg[1] * m2[j + 3] +
                       m1[i + 6] * dg[2] * m2[j + 6];
}

/// matrixMatrixMultiplication3d, MatrixUtils.h:136-146: c = a * b.
template <typename T>
inline void matmul3(const T* a, const T* b, T* c) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      c[i + 3 * j] = a[i] * b[3 * j] + a[i + 3] * b[3 * j + 1] + a[i + 6] * b[3 * j + 2];
}

// ----------------------------------------------------------------------------
// 3x3 SVD
// ----------------------------------------------------------------------------

/// McAdams, Selle, Tamstorf, Teran, Sifakis, "Computing the Singular Value
/// Decomposition of 3x3 matrices with minimal branching and elementary
/// floating point operations" (UW-Madison TR 1690, 2011), as zpc runs it on
/// the host: math/matrix/SVD.hpp:15-1026 with rsqrt = 1/sqrtf
/// (ZpcMathUtils.hpp:842-846). The reference spells each step out with bit
/// masks; this restatement uses exact selects, which produce the same bits.
///
/// A is row-major a[r][c] (the argument order of svd_3d); U, V come back
/// row-major too, S holds the three singular values (S[2] may be negative).
struct McAdams {
  // Constants of the method: 4*gamma^2 with gamma = 3 + 2*sqrt(2) rounded to
  // float, and sin/cos(pi/8) (SVD.hpp:27-32 store the last two as bit
  // patterns 1053028117 and 1064076127).
  static constexpr float four_gamma_sq = 5.8284273147583007813f;
  static float sin_pi8() { u32 b = 1053028117u; float f; std::memcpy(&f, &b, 4); return f; }
  static float cos_pi8() { u32 b = 1064076127u; float f; std::memcpy(&f, &b, 4); return f; }
  static constexpr float tiny = 1.e-20f;
  static constexpr float small = 1.e-12f;

  static float rsqrt(float v) { return 1.f / std::sqrt(v); }

  /// One approximate Jacobi conjugation on the symmetric S for pivot (p, q)
  /// with r the third index; quaternion (qs, qv[3]) accumulates V. The axis
  /// of the quaternion rotation is r. (SVD.hpp:96-185 for (1,2), :187-271 for
  /// (2,3), :273-365 for (3,1).)
  static void conjugate(float s[3][3], int p, int q, int r, float& qs, float qv[3]) {
    float sh = s[q][p] * 0.5f;
    float t5 = s[p][p] - s[q][q];
    float t2 = sh * sh;
    bool big = t2 >= tiny;
    sh = big ? sh : 0.0f;
    float ch = big ? t5 : 1.0f;
    float t1 = sh * sh;
    t2 = ch * ch;
    float t3 = t1 + t2;
    float t4 = rsqrt(t3);
    sh = t4 * sh;
    ch = t4 * ch;
    t1 = four_gamma_sq * t1;
    bool use_pi8 = t2 <= t1;
    sh = use_pi8 ? sin_pi8() : sh;
    ch = use_pi8 ? cos_pi8() : ch;
    t1 = sh * sh;
    t2 = ch * ch;
    float c = t2 - t1;
    float sn = ch * sh;
    sn = sn + sn;
    // Givens conjugation; the rotation is scaled by (sh^2+ch^2) twice on s_rr.
    t3 = t1 + t2;
    s[r][r] = s[r][r] * t3;
    s[r][p] = s[r][p] * t3;
    s[r][q] = s[r][q] * t3;
    s[r][r] = s[r][r] * t3;
    t1 = sn * s[r][p];
    t2 = sn * s[r][q];
    s[r][p] = c * s[r][p];
    s[r][q] = c * s[r][q];
    s[r][p] = t2 + s[r][p];
    s[r][q] = s[r][q] - t1;
    t2 = sn * sn;
    t1 = s[q][q] * t2;
    t3 = s[p][p] * t2;
    t4 = c * c;
    s[p][p] = s[p][p] * t4;
    s[q][q] = s[q][q] * t4;
    s[p][p] = s[p][p] + t1;
    s[q][q] = s[q][q] + t3;
    t4 = t4 - t2;
    t2 = s[q][p] + s[q][p];
    s[q][p] = s[q][p] * t4;
    t4 = c * sn;
    t2 = t2 * t4;
    t5 = t5 * t4;
    s[p][p] = s[p][p] + t2;
    s[q][p] = s[q][p] - t5;
    s[q][q] = s[q][q] - t2;
    // mirror the symmetric entries the next pivots read
    s[p][r] = s[r][p];
    s[q][r] = s[r][q];
    s[p][q] = s[q][p];
    // quaternion q <- q * (ch + sh e_r); (a,b,c) = (r, r+1, r+2) cyclic
    const int a = r, b = (r + 1) % 3, cc = (r + 2) % 3;
    float u1 = sh * qv[b];
    float u2 = sh * qv[cc];
    float u3 = sh * qv[a];
    float shs = sh * qs;
    qs = ch * qs;
    qv[0] = ch * qv[0];
    qv[1] = ch * qv[1];
    qv[2] = ch * qv[2];
    qv[a] = qv[a] + shs;
    qs = qs - u3;
    qv[b] = qv[b] + u2;
    qv[cc] = qv[cc] - u1;
  }

  /// One QR Givens step zeroing b[q][p] against b[p][p] and accumulating U
  /// (SVD.hpp:675-788, :790-903, :905-1001).
  static void givens_qr(float b[3][3], float u[3][3], int p, int q) {
    float a_qp = b[q][p], a_pp = b[p][p];
    float sh = a_qp * a_qp;
    sh = (sh >= small) ? a_qp : 0.0f;
    float ch = 0.f - a_pp;
    ch = (a_pp > ch) ? a_pp : ch;  // math::max(x, y) = y > x ? y : x
    ch = (small > ch) ? small : ch;
    bool nonneg = a_pp >= 0.f;
    float t1 = ch * ch;
    float t2 = sh * sh;
    t2 = t1 + t2;
    t1 = rsqrt(t2);
    float t4 = t1 * 0.5f;
    float t3 = t1 * t4;
    t3 = t1 * t3;
    t3 = t2 * t3;
    t1 = t1 + t4;
    t1 = t1 - t3;
    t1 = t1 * t2;
    ch = ch + t1;
    if (!nonneg) { float tmp = ch; ch = sh; sh = tmp; }
    t1 = ch * ch;
    t2 = sh * sh;
    t2 = t1 + t2;
    t1 = rsqrt(t2);
    t4 = t1 * 0.5f;
    t3 = t1 * t4;
    t3 = t1 * t3;
    t3 = t2 * t3;
    t1 = t1 + t4;
    t1 = t1 - t3;
    ch = ch * t1;
    sh = sh * t1;
    float c = ch * ch;
    float sn = sh * sh;
    c = c - sn;
    sn = sh * ch;
    sn = sn + sn;
    for (int j = 0; j < 3; ++j) {
      float x1 = sn * b[p][j];
      float x2 = sn * b[q][j];
      b[p][j] = c * b[p][j];
      b[q][j] = c * b[q][j];
      b[p][j] = b[p][j] + x2;
      b[q][j] = b[q][j] - x1;
    }
    for (int i = 0; i < 3; ++i) {
      float x1 = sn * u[i][p];
      float x2 = sn * u[i][q];
      u[i][p] = c * u[i][p];
      u[i][q] = c * u[i][q];
      u[i][p] = u[i][p] + x2;
      u[i][q] = u[i][q] - x1;
    }
  }

  /// Swap columns i and j of B and V (and their norms) when n_i < n_j, then
  /// negate column `neg` so V stays a rotation (SVD.hpp:489-664).
  static void sort_step(float b[3][3], float v[3][3], float n[3], int i, int j, int neg) {
    if (n[i] < n[j]) {
      for (int r = 0; r < 3; ++r) {
        std::swap(b[r][i], b[r][j]);
        std::swap(v[r][i], v[r][j]);
      }
      std::swap(n[i], n[j]);
      for (int r = 0; r < 3; ++r) {
        b[r][neg] = b[r][neg] * -1.f;
        v[r][neg] = v[r][neg] * -1.f;
      }
    } else {
      for (int r = 0; r < 3; ++r) {
        b[r][neg] = b[r][neg] * 1.f;
        v[r][neg] = v[r][neg] * 1.f;
      }
    }
  }

  static void svd(const float a_in[3][3], float U[3][3], float S[3], float V[3][3]) {
    float a[3][3];
    std::memcpy(a, a_in, sizeof(a));
    // normal equations S = A^T A (SVD.hpp:53-87)
    float s[3][3];
    for (int j = 0; j < 3; ++j)
      for (int i = j; i < 3; ++i) {
        float acc = a[0][i] * a[0][j];
        float t = a[1][i] * a[1][j];
        acc = t + acc;
        t = a[2][i] * a[2][j];
        acc = t + acc;
        s[i][j] = acc;
        s[j][i] = acc;
      }
    float qs = 1.f, qv[3] = {0.f, 0.f, 0.f};
    for (int sweep = 0; sweep < 4; ++sweep) {  // SVD.hpp:95
      conjugate(s, 0, 1, 2, qs, qv);
      conjugate(s, 1, 2, 0, qs, qv);
      conjugate(s, 2, 0, 1, qs, qv);
    }
    // normalize the quaternion with one Newton step on rsqrt (SVD.hpp:368-390)
    float t2 = qs * qs;
    float t1 = qv[0] * qv[0];
    t2 = t1 + t2;
    t1 = qv[1] * qv[1];
    t2 = t1 + t2;
    t1 = qv[2] * qv[2];
    t2 = t1 + t2;
    t1 = rsqrt(t2);
    float t4 = t1 * 0.5f;
    float t3 = t1 * t4;
    t3 = t1 * t3;
    t3 = t2 * t3;
    t1 = t1 + t4;
    t1 = t1 - t3;
    qs = qs * t1;
    qv[0] = qv[0] * t1;
    qv[1] = qv[1] * t1;
    qv[2] = qv[2] * t1;
    // quaternion -> V (SVD.hpp:392-430)
    float x2 = qv[0] * qv[0], y2 = qv[1] * qv[1], z2 = qv[2] * qv[2];
    float v[3][3];
    v[0][0] = qs * qs;
    v[1][1] = v[0][0] - x2;
    v[2][2] = v[1][1] - y2;
    v[2][2] = v[2][2] + z2;
    v[1][1] = v[1][1] + y2;
    v[1][1] = v[1][1] - z2;
    v[0][0] = v[0][0] + x2;
    v[0][0] = v[0][0] - y2;
    v[0][0] = v[0][0] - z2;
    float tx = qv[0] + qv[0], ty = qv[1] + qv[1], tz = qv[2] + qv[2];
    v[2][1] = qs * tx;
    v[0][2] = qs * ty;
    v[1][0] = qs * tz;
    tx = qv[1] * tx;
    ty = qv[2] * ty;
    tz = qv[0] * tz;
    v[0][1] = tx - v[1][0];
    v[1][2] = ty - v[2][1];
    v[2][0] = tz - v[0][2];
    v[1][0] = tx + v[1][0];
    v[2][1] = ty + v[2][1];
    v[0][2] = tz + v[0][2];
    // B = A V, row by row (SVD.hpp:432-486)
    float b[3][3];
    for (int r = 0; r < 3; ++r) {
      float a1 = a[r][0], a2 = a[r][1], a3 = a[r][2];
      float c0 = v[0][0] * a1, c1 = v[0][1] * a1, c2 = v[0][2] * a1;
      c0 = c0 + v[1][0] * a2;
      c0 = c0 + v[2][0] * a3;
      c1 = c1 + v[1][1] * a2;
      c1 = c1 + v[2][1] * a3;
      c2 = c2 + v[1][2] * a2;
      c2 = c2 + v[2][2] * a3;
      b[r][0] = c0;
      b[r][1] = c1;
      b[r][2] = c2;
    }
    // column norms and sorting (SVD.hpp:488-664)
    float n[3];
    for (int c = 0; c < 3; ++c) {
      float acc = b[0][c] * b[0][c];
      float t = b[1][c] * b[1][c];
      acc = acc + t;
      t = b[2][c] * b[2][c];
      acc = acc + t;
      n[c] = acc;
    }
    sort_step(b, v, n, 0, 1, 1);
    sort_step(b, v, n, 0, 2, 0);
    sort_step(b, v, n, 1, 2, 2);
    // QR by Givens rotations (SVD.hpp:666-1001)
    float u[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    givens_qr(b, u, 0, 1);
    givens_qr(b, u, 0, 2);
    givens_qr(b, u, 1, 2);
    std::memcpy(U, u, sizeof(u));
    std::memcpy(V, v, sizeof(v));
    S[0] = b[0][0];
    S[1] = b[1][1];
    S[2] = b[2][2];
  }
};

/// FP64 SVD for the double-precision twin: cyclic Jacobi on A^T A with exact
/// rotations to convergence, then the same column sort (V kept a rotation)
/// and Givens QR as McAdams, so the conventions match (sigma sorted, U and V
/// rotations, sigma_3 carries the sign of det A). Own implementation; pinned
/// by reconstruction/orthogonality invariants and against LAPACK via numpy.
inline void svd3_jacobi64(const double a[3][3], double U[3][3], double S[3], double V[3][3]) {
  double s[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      s[i][j] = a[0][i] * a[0][j] + a[1][i] * a[1][j] + a[2][i] * a[2][j];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = s[0][1] * s[0][1] + s[0][2] * s[0][2] + s[1][2] * s[1][2];
    double dia = s[0][0] * s[0][0] + s[1][1] * s[1][1] + s[2][2] * s[2][2];
    if (off <= 1e-34 * dia || off == 0.0) break;
    const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (auto& pq : pairs) {
      int p = pq[0], q = pq[1];
      if (s[p][q] == 0.0) continue;
      double theta = (s[q][q] - s[p][p]) / (2.0 * s[p][q]);
      double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
      for (int k = 0; k < 3; ++k) {  // S <- J^T S J
        double skp = s[k][p], skq = s[k][q];
        s[k][p] = c * skp - sn * skq;
        s[k][q] = sn * skp + c * skq;
      }
      for (int k = 0; k < 3; ++k) {
        double spk = s[p][k], sqk = s[q][k];
        s[p][k] = c * spk - sn * sqk;
        s[q][k] = sn * spk + c * sqk;
      }
      for (int k = 0; k < 3; ++k) {  // V <- V J
        double vkp = v[k][p], vkq = v[k][q];
        v[k][p] = c * vkp - sn * vkq;
        v[k][q] = sn * vkp + c * vkq;
      }
    }
  }
  double b[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) b[r][c] = a[r][0] * v[0][c] + a[r][1] * v[1][c] + a[r][2] * v[2][c];
  double n[3];
  for (int c = 0; c < 3; ++c) n[c] = b[0][c] * b[0][c] + b[1][c] * b[1][c] + b[2][c] * b[2][c];
  auto sort_step = [&](int i, int j, int neg) {
    if (n[i] < n[j]) {
      for (int r = 0; r < 3; ++r) { std::swap(b[r][i], b[r][j]); std::swap(v[r][i], v[r][j]); }
      std::swap(n[i], n[j]);
      for (int r = 0; r < 3; ++r) { b[r][neg] = -b[r][neg]; v[r][neg] = -v[r][neg]; }
    }
  };
  sort_step(0, 1, 1);
  sort_step(0, 2, 0);
  sort_step(1, 2, 2);
  double u[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  auto givens = [&](int p, int q) {
    double x = b[p][p], y = b[q][p];
    double r = std::hypot(x, y);
    double c, sn;
    // c = x / r, s = y / r leaves the pivot at r >= 0
    if (r == 0.0) { c = 1.0; sn = 0.0; } else { c = x / r; sn = y / r; }
    for (int j = 0; j < 3; ++j) {
      double bp = b[p][j], bq = b[q][j];
      b[p][j] = c * bp + sn * bq;
      b[q][j] = -sn * bp + c * bq;
    }
    for (int i = 0; i < 3; ++i) {
      double up = u[i][p], uq = u[i][q];
      u[i][p] = c * up + sn * uq;
      u[i][q] = -sn * up + c * uq;
    }
  };
  givens(0, 1);
  givens(0, 2);
  givens(1, 2);
  std::memcpy(U, u, sizeof(u));
  std::memcpy(V, v, sizeof(v));
  S[0] = b[0][0];
  S[1] = b[1][1];
  S[2] = b[2][2];
}

/// svd_3d called as the stress functions call it
/// (physics/ConstitutiveModel_Vol_dP.hpp:14-16): F column-major in, U and V
/// column-major out.
inline void svd_colmajor(const float* F, float* Uc, float* S, float* Vc) {
  float a[3][3], u[3][3], v[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = F[i + 3 * j];
  McAdams::svd(a, u, S, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { Uc[i + 3 * j] = u[i][j]; Vc[i + 3 * j] = v[i][j]; }
}
inline void svd_colmajor(const double* F, double* Uc, double* S, double* Vc) {
  double a[3][3], u[3][3], v[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = F[i + 3 * j];
  svd3_jacobi64(a, u, S, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { Uc[i + 3 * j] = u[i][j]; Vc[i + 3 * j] = v[i][j]; }
}

// ----------------------------------------------------------------------------
// constitutive models
// ----------------------------------------------------------------------------

/// lame_parameters, physics/ConstitutiveModel.hpp:34-38. For float, mu is
/// formed in double (0.5 is a double literal) from the float (1 + nu).
inline void lame(float E, float nu, float& mu, float& lam) {
  float onep = 1 + nu;
  mu = (float)(0.5 * (double)E / (double)onep);
  lam = E * nu / ((1 + nu) * (1 - 2 * nu));
}
inline void lame(double E, double nu, double& mu, double& lam) {
  mu = 0.5 * E / (1 + nu);
  lam = E * nu / ((1 + nu) * (1 - 2 * nu));
}

/// PF^T * volume, the tail shared by both stress functions
/// (ConstitutiveModel_Vol_dP.hpp:35-45, :315-324).
template <typename T>
inline void pft_volume(const T* P, const T* F, T volume, T* PF) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      PF[i + 3 * j] = (P[i] * F[j] + P[i + 3] * F[j + 3] + P[i + 6] * F[j + 6]) * volume;
}

/// compute_stress_fixedcorotated, physics/ConstitutiveModel_Vol_dP.hpp:10-46.
template <typename T>
void stress_fixedcorotated(T volume, T mu, T lambda, const T* F, T* PF) {
  T U[9], S[3], V[9];
  svd_colmajor(F, U, S, V);
  T J = S[0] * S[1] * S[2];
  T scaled_mu = (T)2.f * mu;
  T scaled_lambda = lambda * (J - (T)1.f);
  T Ph[3];
  Ph[0] = scaled_mu * (S[0] - (T)1.f) + scaled_lambda * (S[1] * S[2]);
  Ph[1] = scaled_mu * (S[1] - (T)1.f) + scaled_lambda * (S[0] * S[2]);
  Ph[2] = scaled_mu * (S[2] - (T)1.f) + scaled_lambda * (S[0] * S[1]);
  T P[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      P[i + 3 * j] = Ph[0] * U[i] * V[j] + Ph[1] * U[i + 3] * V[j + 3] + Ph[2] * U[i + 6] * V[j + 6];
  pft_volume(P, F, volume, PF);
}

inline float olog(float v) { return std::log(v); }
inline double olog(double v) { return std::log(v); }
inline float oexp(float v) { return std::exp(v); }
inline double oexp(double v) { return std::exp(v); }
/// zs::math::q_rsqrt(float), math/MathUtils.h:258-269: Lomont's rsqrt seed
/// (magic 0x5f375a86, "Fast Inverse Square Root", 2003) and two Newton steps.
inline float zpc_q_rsqrt(float number) {
  float x2 = number * 0.5f, y = number;
  u32 i;
  std::memcpy(&i, &y, 4);
  i = 0x5f375a86u - (i >> 1);
  std::memcpy(&y, &i, 4);
  y = y * (1.5f - (x2 * y * y));
  y = y * (1.5f - (x2 * y * y));
  return y;
}
/// zs::math::sqrt(float), math/MathUtils.h:286-323: the software square
/// root zpc's host path uses for the sand model's deviatoric norm
/// (ConstitutiveModel_Vol_dP.hpp:272-274); the seed above, two Newton steps
/// on rsqrt and one residual correction, with 2^26 / 2^-13 scaling. Without
/// contraction it is not always correctly rounded, which is the behaviour
/// being restated.
inline float zpc_sqrt(float arg) {
  if (arg < 0.0f) {
    u32 b = 0x7fffffffu;
    float q;
    std::memcpy(&q, &b, 4);
    return q;
  }
  const float inf = std::numeric_limits<float>::infinity();
  if (arg == 0.0f || !((arg < 0 ? -arg : arg) < inf)) return arg + arg;
  arg = arg * 0x1.0p+26f;
  float rsq = zpc_q_rsqrt(arg);
  rsq = ((-0.5f * arg * rsq) * rsq + 0.5f) * rsq + rsq;
  rsq = ((-0.5f * arg * rsq) * rsq + 0.5f) * rsq + rsq;
  float sqt = rsq * arg;
  float err = sqt * -sqt + arg;
  sqt = (0.5f * rsq * err + sqt);
  return sqt * 0x1.0p-13f;
}
inline float osqrt(float v) { return zpc_sqrt(v); }
inline double osqrt(double v) { return std::sqrt(v); }

/// compute_stress_sand (Drucker-Prager with volume correction),
/// physics/ConstitutiveModel_Vol_dP.hpp:241-325. F is projected in a local
/// copy only; logJp is updated in place. When Fproj is non-null the locally
/// projected F is stored there (used by the learned-closure data harvest).
/// Returns the case taken: 1 inside the cone, 2 cone tip, 3 projected onto
/// the cone surface, 0 when mu == 0 and the trace is negative.
template <typename T>
int stress_sand(T volume, T mu, T lambda, T cohesion, T beta, T yieldSurface, bool volCorrection,
                T& logJp, const T* F_in, T* PF, T* Fproj = nullptr) {
  T F[9];
  for (int i = 0; i < 9; ++i) F[i] = F_in[i];
  T U[9], S[3], V[9];
  svd_colmajor(F, U, S, V);
  T scaled_mu = T(2) * mu;
  T epsilon[3], New_S[3] = {0, 0, 0};
  T New_F[9];
  int which = 0;
  for (int i = 0; i < 3; ++i) {
    T abs_S = S[i] > 0 ? S[i] : -S[i];
    abs_S = abs_S > 1e-4 ? abs_S : (T)1e-4;
    epsilon[i] = olog(abs_S) - cohesion;
  }
  T sum_epsilon = epsilon[0] + epsilon[1] + epsilon[2];
  T trace_epsilon = sum_epsilon + logJp;
  T epsilon_hat[3];
  for (int i = 0; i < 3; ++i) epsilon_hat[i] = epsilon[i] - (trace_epsilon / (T)3);
  T epsilon_hat_norm = osqrt(epsilon_hat[0] * epsilon_hat[0] + epsilon_hat[1] * epsilon_hat[1] +
                             epsilon_hat[2] * epsilon_hat[2]);
  if (trace_epsilon >= (T)0) {
    New_S[0] = New_S[1] = New_S[2] = oexp(cohesion);
    matmul_mat_diag_matT(New_F, U, New_S, V);
    for (int i = 0; i < 9; ++i) F[i] = New_F[i];
    if (volCorrection) logJp = beta * sum_epsilon + logJp;
    which = 2;
  } else if (mu != 0) {
    logJp = 0;
    T delta_gamma =
        epsilon_hat_norm + ((T)3 * lambda + scaled_mu) / scaled_mu * trace_epsilon * yieldSurface;
    T H[3];
    if (delta_gamma <= 0) {
      for (int i = 0; i < 3; ++i) H[i] = epsilon[i] + cohesion;
      which = 1;
    } else {
      for (int i = 0; i < 3; ++i)
        H[i] = epsilon[i] - (delta_gamma / epsilon_hat_norm) * epsilon_hat[i] + cohesion;
      which = 3;
    }
    for (int i = 0; i < 3; ++i) New_S[i] = oexp(H[i]);
    matmul_mat_diag_matT(New_F, U, New_S, V);
    for (int i = 0; i < 9; ++i) F[i] = New_F[i];
  }
  T New_S_log[3] = {olog(New_S[0]), olog(New_S[1]), olog(New_S[2])};
  T P_hat[3];
  T trace_log_S = New_S_log[0] + New_S_log[1] + New_S_log[2];
  for (int i = 0; i < 3; ++i) P_hat[i] = (scaled_mu * New_S_log[i] + lambda * trace_log_S) / New_S[i];
  T P[9];
  matmul_mat_diag_matT(P, U, P_hat, V);
  pft_volume(P, F, volume, PF);
  if (Fproj)
    for (int i = 0; i < 9; ++i) Fproj[i] = F[i];
  return which;
}

// ----------------------------------------------------------------------------
// particles, grid, material, colliders
// ----------------------------------------------------------------------------

enum MaterialKind : int { kFixedCorotated = 0, kDruckerPrager = 1 };

/// FixedCorotatedConfig / DruckerPragerConfig, physics/ConstitutiveModel.hpp:736-754.
template <typename T>
struct Material {
  int kind = kFixedCorotated;
  T volume = 1, E = 5e4, nu = 0.4;
  T cohesion = 0, beta = 1, yieldSurface = (T)0.816496580927726f * 2.f * 0.5f / (3.f - 0.5f);
  bool volumeCorrection = true;
};

/// Particles<f32,3> attributes (simulation/init/Scene.cpp:210-221), one
/// array per attribute: x vec3, v vec3, C vec9 and F vec9 column-major, m,
/// logJp. Non-owning.
template <typename T>
struct ParticlesView {
  i64 n = 0;
  T* x = nullptr;
  T* v = nullptr;
  T* C = nullptr;
  T* F = nullptr;
  T* m = nullptr;
  T* logJp = nullptr;
};

/// Grids<f32,3,4> with channels {m, v(3), rhs(3)} (simulation/mpm/Simulator.cpp:117)
/// in TileVector layout (container/TileVector.hpp:108,397): element (ch, cell)
/// of block b at (b*7 + ch)*64 + cell, cell = (x<<4)|(y<<2)|z
/// (geometry/Structure.hpp:849-860).
constexpr int kChannels = 7;
inline i64 gidx(i64 b, int ch, int cell) { return (b * kChannels + ch) * 64 + cell; }

enum ColliderType : int { kSticky = 0, kSlip = 1, kSeparate = 2 };

/// Collider<AnalyticLevelSet<Plane>> with an identity transform
/// (geometry/Collider.h:80-107, geometry/AnalyticLevelSet.h:11-50).
template <typename T>
struct PlaneCollider {
  T origin[3], normal[3];
  int type = kSticky;
};

enum Geometry : int { kPlane = 0, kCuboid = 1, kSphere = 2, kCylinder = 3 };

/// Collider<AnalyticLevelSet<geometry>> with its rigid motion: the material
/// point X of the level set sits at x = R s X + b (geometry/Collider.h:10-190;
/// the level sets of geometry/AnalyticLevelSet.h:11-225). p0/p1: plane origin
/// and normal; cuboid min and max; sphere center and p1[0] = radius; cylinder
/// bottom, p1[0] = radius, p1[1] = length, along `axis`. R column-major.
template <typename T>
struct GeneralCollider {
  int type = kSticky, geometry = kPlane, axis = 1;
  T p0[3] = {0, 0, 0}, p1[3] = {0, 0, 0};
  T R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  T omega[3] = {0, 0, 0};
  T s = 1, dsdt = 0;
  T b[3] = {0, 0, 0}, dbdt[3] = {0, 0, 0};
};

/// Signed distances of the analytic level sets in material space
/// (AnalyticLevelSet.h: Plane :26-29, Cuboid :81-88, Sphere :142-145,
/// Cylinder :180-207), evaluated in T in the order written there.
template <typename T>
T level_set_sdf(const GeneralCollider<T>& c, const T X[3]) {
  switch (c.geometry) {
    case kPlane:
      return c.p1[0] * (X[0] - c.p0[0]) + c.p1[1] * (X[1] - c.p0[1]) + c.p1[2] * (X[2] - c.p0[2]);
    case kCuboid: {
      T point[3], mx = 0;
      for (int d = 0; d < 3; ++d) {
        const T center = (c.p0[d] + c.p1[d]) / (T)2;
        point[d] = std::fabs(X[d] - center) - (c.p1[d] - c.p0[d]) / (T)2;
        mx = d == 0 ? point[d] : std::max(mx, point[d]);
      }
      for (int d = 0; d < 3; ++d)
        if (point[d] < 0) point[d] = 0;
      const T len = std::sqrt(point[0] * point[0] + point[1] * point[1] + point[2] * point[2]);
      return (mx < 0 ? mx : (T)0) + len;
    }
    case kSphere: {
      const T o[3] = {X[0] - c.p0[0], X[1] - c.p0[1], X[2] - c.p0[2]};
      return std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) - c.p1[0];
    }
    default: {  // cylinder
      const int ax = c.axis;
      T dr[2];
      for (int k = 0, i = 0; k < 3; ++k)
        if (k != ax) dr[i++] = X[k] - c.p0[k];
      const T disR = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
      const T r = c.p1[0], len = c.p1[1];
      const bool outside = disR > r;
      if (X[ax] < c.p0[ax]) {
        const T disL = c.p0[ax] - X[ax];
        return outside ? std::sqrt((disR - r) * (disR - r) + disL * disL) : disL;
      }
      if (X[ax] > c.p0[ax] + len) {
        const T disL = X[ax] - (c.p0[ax] + len);
        return outside ? std::sqrt((disR - r) * (disR - r) + disL * disL) : disL;
      }
      if (outside) return disR - r;
      const T disL = std::min(c.p0[ax] + len - X[ax], X[ax] - c.p0[ax]);
      return -std::min(disL, r - disR);
    }
  }
}

/// Normals in material space: the plane's normal as given, the sphere's
/// outward direction (zero within 1e-7 squared of the center), central
/// differences of step 1e-6, normalised, for the cuboid and the cylinder
/// (AnalyticLevelSet.h :30-33, :89-103, :146-150, :208-222).
template <typename T>
void level_set_normal(const GeneralCollider<T>& c, const T X[3], T n[3]) {
  if (c.geometry == kPlane) {
    for (int d = 0; d < 3; ++d) n[d] = c.p1[d];
    return;
  }
  if (c.geometry == kSphere) {
    const T o[3] = {X[0] - c.p0[0], X[1] - c.p0[1], X[2] - c.p0[2]};
    const T l2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    if (l2 < (T)1e-7) {
      n[0] = n[1] = n[2] = 0;
      return;
    }
    const T l = std::sqrt(l2);
    for (int d = 0; d < 3; ++d) n[d] = o[d] / l;
    return;
  }
  const T eps = (T)1e-6;
  T diff[3];
  for (int i = 0; i < 3; ++i) {
    T v1[3] = {X[0], X[1], X[2]}, v2[3] = {X[0], X[1], X[2]};
    v1[i] = X[i] + eps;
    v2[i] = X[i] - eps;
    diff[i] = (level_set_sdf(c, v1) - level_set_sdf(c, v2)) / (eps + eps);
  }
  const T l = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
  for (int d = 0; d < 3; ++d) n[d] = diff[d] / l;
}

/// Collider::resolveCollision (geometry/Collider.h:81-107) with erosion 0 and
/// the analytic level sets' zero material velocity: where phi(X) < 0 the node
/// takes the object velocity omega x (x - b) + (s'/s)(x - b) + b' (sticky), or
/// loses the normal part of its velocity relative to it (slip; separate only
/// when approaching).
template <typename T>
void resolve_general(const GeneralCollider<T>& c, const T x[3], T v[3]) {
  const T xmb[3] = {x[0] - c.b[0], x[1] - c.b[1], x[2] - c.b[2]};
  const T inv_s = (T)1 / c.s;
  T X[3];
  for (int i = 0; i < 3; ++i)
    X[i] = (c.R[0 + 3 * i] * xmb[0] + c.R[1 + 3 * i] * xmb[1] + c.R[2 + 3 * i] * xmb[2]) * inv_s;
  if (!(level_set_sdf(c, X) < (T)0)) return;
  const T rate = c.dsdt * inv_s;
  const T wx[3] = {c.omega[1] * xmb[2] - c.omega[2] * xmb[1], c.omega[2] * xmb[0] - c.omega[0] * xmb[2],
                   c.omega[0] * xmb[1] - c.omega[1] * xmb[0]};
  T vo[3];
  for (int d = 0; d < 3; ++d) vo[d] = wx[d] + rate * xmb[d] + c.dbdt[d];
  if (c.type == kSticky) {
    for (int d = 0; d < 3; ++d) v[d] = vo[d];
    return;
  }
  for (int d = 0; d < 3; ++d) v[d] -= vo[d];
  T nm[3], n[3];
  level_set_normal(c, X, nm);
  for (int r = 0; r < 3; ++r) n[r] = c.R[r] * nm[0] + c.R[r + 3] * nm[1] + c.R[r + 6] * nm[2];
  const T proj = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  if ((c.type == kSeparate && proj < 0) || c.type == kSlip)
    for (int d = 0; d < 3; ++d) v[d] -= proj * n[d];
  for (int d = 0; d < 3; ++d) v[d] += vo[d];
}

// ----------------------------------------------------------------------------
// transfer operators
// ----------------------------------------------------------------------------

/// LocalArena::init for the collocated quadratic kernel, simulation/Utils.hpp:51-70,
/// base_node (math/curve/InterpolationKernel.hpp:48-57) and
/// quadratic_bspline_weights (:95-130). X = pos / dx by division.
template <typename T>
struct Arena {
  i32 corner[3];
  T localPos[3];  // already scaled by dx
  T w[3][3];      // w[axis][node]
  T dx;
  void init(T dx_, const T* pos) {
    dx = dx_;
    T X[3];
    for (int d = 0; d < 3; ++d) X[d] = pos[d] / dx;
    for (int d = 0; d < 3; ++d) corner[d] = (i32)std::floor(X[d] - (T)0.5 * 1);
    for (int d = 0; d < 3; ++d) localPos[d] = X[d] - (T)corner[d];
    for (int d = 0; d < 3; ++d) {
      T d0 = localPos[d] - (T)(i32)std::floor(localPos[d] - (T)0.5 * 1);
      w[d][0] = (T)0.5 * ((T)1.5 - d0) * ((T)1.5 - d0);
      T d1 = d0 - (T)1.0;
      w[d][1] = (T)0.75 - d1 * d1;
      T zz = (T)0.5 + d1;
      w[d][2] = (T)0.5 * zz * zz;
    }
    for (int d = 0; d < 3; ++d) localPos[d] *= dx;
  }
  /// weight_impl (Utils.hpp:99-105): ret = 1; ret *= w(0,i); ret *= w(1,j); ret *= w(2,k).
  T weight(int i, int j, int k) const {
    T r = 1;
    r *= w[0][i];
    r *= w[1][j];
    r *= w[2][k];
    return r;
  }
  /// diff (Utils.hpp:163-168): offset * dx - localPos.
  void diff(int i, int j, int k, T* out) const {
    out[0] = (T)i * dx - localPos[0];
    out[1] = (T)j * dx - localPos[1];
    out[2] = (T)k * dx - localPos[2];
  }
};

/// unpack_coord_in_grid with a table, simulation/Utils.hpp:20-28, and
/// coord_to_cellid (geometry/Structure.hpp:849-860).
inline void unpack_coord(Table& t, const i32 coord[3], i32& block, int& cellid) {
  i32 loc[3], key[3];
  for (int d = 0; d < 3; ++d) {
    loc[d] = coord[d] & 3;
    key[d] = (coord[d] - loc[d]) / 4;
  }
  block = t.query(key);
  cellid = (loc[0] << 4) | (loc[1] << 2) | loc[2];
}

/// Per-particle stress contribution, P2G.hpp:83-104 (before the scatter).
template <typename T>
inline void p2g_stress(const Material<T>& mat, T mu, T lambda, const T* F, T& logJp, T* contrib,
                       T dt, T D_inv) {
  if (mat.kind == kFixedCorotated) {
    stress_fixedcorotated(mat.volume, mu, lambda, F, contrib);
  } else {
    stress_sand(mat.volume, mu, lambda, mat.cohesion, mat.beta, mat.yieldSurface,
                mat.volumeCorrection, logJp, F, contrib);
  }
  for (int i = 0; i < 9; ++i) contrib[i] = contrib[i] * -dt * D_inv;
}

/// P2GTransfer::operator() over range(N) in particle order,
/// simulation/transfer/P2G.hpp:45-127 (APIC, quadratic). Writes back only logJp.
template <typename T>
void p2g(const Material<T>& mat, T dt, T dx, ParticlesView<T> ps, Table& t, T* grid) {
  const T dx_inv = (T)1.0 / dx;
  const T D_inv = (T)4.f * dx_inv * dx_inv;
  T mu, lambda;
  lame(mat.E, mat.nu, mu, lambda);
  for (i64 p = 0; p < ps.n; ++p) {
    const T* pos = ps.x + 3 * p;
    const T* vel = ps.v + 3 * p;
    const T mass = ps.m[p];
    const T* C = ps.C + 9 * p;
    T contrib[9];
    T logJp = ps.logJp ? ps.logJp[p] : (T)0;
    p2g_stress(mat, mu, lambda, ps.F + 9 * p, logJp, contrib, dt, D_inv);
    if (mat.kind == kDruckerPrager) ps.logJp[p] = logJp;
    Arena<T> ar;
    ar.init(dx, pos);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          i32 coord[3] = {ar.corner[0] + i, ar.corner[1] + j, ar.corner[2] + k};
          i32 b;
          int cell;
          unpack_coord(t, coord, b, cell);
          T xixp[3];
          ar.diff(i, j, k, xixp);
          T W = ar.weight(i, j, k);
          grid[gidx(b, 0, cell)] += mass * W;
          for (int d = 0; d < 3; ++d) {
            grid[gidx(b, 1 + d, cell)] +=
                W * mass * (vel[d] + (C[d] * xixp[0] + C[3 + d] * xixp[1] + C[6 + d] * xixp[2]));
            grid[gidx(b, 4 + d, cell)] +=
                (contrib[d] * xixp[0] + contrib[3 + d] * xixp[1] + contrib[6 + d] * xixp[2]) * W;
          }
        }
  }
}

/// CleanGridBlocks over Collapse{nBlocks, 64}, simulation/grid/GridOp.hpp:54-69.
template <typename T>
void grid_clear(T* grid, i64 nblocks) {
  std::fill(grid, grid + nblocks * kChannels * 64, (T)0);
}

/// The explicit driver's fold of the P2G force channels into momentum
/// (rhs channels 4..6 are written by P2G.hpp:120-123 but never read by
/// ComputeGridBlockVelocity, GridOp.hpp:90-104): mv += rhs.
template <typename T>
void fold_rhs(T* grid, i64 nblocks) {
  for (i64 b = 0; b < nblocks; ++b)
    for (int d = 0; d < 3; ++d)
      for (int c = 0; c < 64; ++c) grid[gidx(b, 1 + d, c)] = grid[gidx(b, 1 + d, c)] + grid[gidx(b, 4 + d, c)];
}

/// ComputeGridBlockVelocity over Collapse{nBlocks, 64}, GridOp.hpp:71-110:
/// if m != 0, v = mv * (1/m) + extf * dt and maxVel = max(maxVel, |v|^2).
template <typename T>
void grid_velocity(T* grid, i64 nblocks, T dt, const T extf[3], T& maxVel) {
  for (i64 b = 0; b < nblocks; ++b)
    for (int c = 0; c < 64; ++c) {
      T mass = grid[gidx(b, 0, c)];
      if (mass != (T)0) {
        mass = (T)1 / mass;
        T v[3];
        for (int d = 0; d < 3; ++d) v[d] = grid[gidx(b, 1 + d, c)] * mass + extf[d] * dt;
        for (int d = 0; d < 3; ++d) grid[gidx(b, 1 + d, c)] = v[d];
        T r = 0;
        r += v[0] * v[0];
        r += v[1] * v[1];
        r += v[2] * v[2];
        if (maxVel < r) maxVel = r;
      }
    }
}

/// ApplyBoundaryConditionOnGridBlocks, GridOp.hpp:112-164, once per collider
/// in list order; Collider::resolveCollision (geometry/Collider.h:80-107)
/// for a static plane: v_object = 0, phi = n . (x - origin).
template <typename T>
void grid_boundary(T* grid, const Table& t, i64 nblocks, T dx, const PlaneCollider<T>* cols,
                   int ncols) {
  for (int ci = 0; ci < ncols; ++ci) {
    const PlaneCollider<T>& col = cols[ci];
    for (i64 b = 0; b < nblocks; ++b)
      for (int c = 0; c < 64; ++c) {
        if (!(grid[gidx(b, 0, c)] > 0)) continue;
        const i32 cc[3] = {(c >> 4) & 3, (c >> 2) & 3, c & 3};
        T pos[3];
        for (int d = 0; d < 3; ++d)
          pos[d] = ((T)t.active[3 * b + d] * (T)4 + (T)cc[d]) * dx;
        T phi = col.normal[0] * (pos[0] - col.origin[0]) + col.normal[1] * (pos[1] - col.origin[1]) +
                col.normal[2] * (pos[2] - col.origin[2]);
        if (phi < (T)0) {
          T v[3] = {grid[gidx(b, 1, c)], grid[gidx(b, 2, c)], grid[gidx(b, 3, c)]};
          if (col.type == kSticky) {
            v[0] = v[1] = v[2] = 0;
          } else {
            T proj = col.normal[0] * v[0] + col.normal[1] * v[1] + col.normal[2] * v[2];
            if ((col.type == kSeparate && proj < 0) || col.type == kSlip)
              for (int d = 0; d < 3; ++d) v[d] -= proj * col.normal[d];
          }
          for (int d = 0; d < 3; ++d) grid[gidx(b, 1 + d, c)] = v[d];
        }
      }
  }
}

/// ApplyBoundaryConditionOnGridBlocks (GridOp.hpp:112-164) for general
/// colliders, once per collider in list order, on nodes with m > 0.
template <typename T>
void grid_boundary_general(T* grid, const Table& t, i64 nblocks, T dx, const GeneralCollider<T>* cols, int ncols) {
  for (int ci = 0; ci < ncols; ++ci)
    for (i64 b = 0; b < nblocks; ++b)
      for (int c = 0; c < 64; ++c) {
        if (!(grid[gidx(b, 0, c)] > 0)) continue;
        const i32 cc[3] = {(c >> 4) & 3, (c >> 2) & 3, c & 3};
        T pos[3], v[3];
        for (int d = 0; d < 3; ++d) {
          pos[d] = ((T)t.active[3 * b + d] * (T)4 + (T)cc[d]) * dx;
          v[d] = grid[gidx(b, 1 + d, c)];
        }
        resolve_general(cols[ci], pos, v);
        for (int d = 0; d < 3; ++d) grid[gidx(b, 1 + d, c)] = v[d];
      }
}

/// GridAngularMomentum over Collapse{nBlocks, 64}, simulation/grid/GridOp.hpp:216-259:
/// for every node with m != 0, x = (blockkey * 4 + coord(cell)) * dx and
/// x cross mv are formed in T; sum[0..2] += (double)(x cross mv),
/// sum[3..5] += (double)mv (atomic_add into the caller's six doubles).
template <typename T>
void grid_angular_momentum(const T* grid, const Table& t, i64 nblocks, T dx, int mChn, int mvChn,
                           double sum[6]) {
  for (i64 b = 0; b < nblocks; ++b)
    for (int c = 0; c < 64; ++c) {
      const T mass = grid[gidx(b, mChn, c)];
      if (mass == (T)0) continue;
      const i32 cc[3] = {(c >> 4) & 3, (c >> 2) & 3, c & 3};
      T x[3], mv[3];
      for (int d = 0; d < 3; ++d) {
        x[d] = ((T)t.active[3 * b + d] * (T)4 + (T)cc[d]) * dx;
        mv[d] = grid[gidx(b, mvChn + d, c)];
      }
      const T res[3] = {x[1] * mv[2] - x[2] * mv[1], x[2] * mv[0] - x[0] * mv[2], x[0] * mv[1] - x[1] * mv[0]};
      for (int d = 0; d < 3; ++d) sum[d] += (double)res[d];
      for (int d = 0; d < 3; ++d) sum[3 + d] += (double)mv[d];
    }
}

/// G2PTransfer::operator() over range(N), simulation/transfer/G2P.hpp:43-84.
template <typename T>
void g2p(T dt, T dx, ParticlesView<T> ps, Table& t, const T* grid) {
  const T dx_inv = (T)1 / dx;
  const T D_inv = (T)4.f * dx_inv * dx_inv;
  for (i64 p = 0; p < ps.n; ++p) {
    T* pos = ps.x + 3 * p;
    T vel[3] = {0, 0, 0};
    T C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    Arena<T> ar;
    ar.init(dx, pos);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          i32 coord[3] = {ar.corner[0] + i, ar.corner[1] + j, ar.corner[2] + k};
          i32 b;
          int cell;
          unpack_coord(t, coord, b, cell);
          T xixp[3];
          ar.diff(i, j, k, xixp);
          T W = ar.weight(i, j, k);
          T vi[3] = {grid[gidx(b, 1, cell)], grid[gidx(b, 2, cell)], grid[gidx(b, 3, cell)]};
          for (int d = 0; d < 3; ++d) vel[d] += vi[d] * W;
          for (int d = 0; d < 9; ++d) C[d] += W * vi[d % 3] * xixp[d / 3] * D_inv;
        }
    for (int d = 0; d < 3; ++d) pos[d] += vel[d] * dt;
    T tmp[9], F[9];
    const T* oldF = ps.F + 9 * p;
    for (int d = 0; d < 9; ++d) tmp[d] = C[d] * dt + ((d & 0x3) ? (T)0.f : (T)1.f);
    matmul3(tmp, oldF, F);
    for (int d = 0; d < 9; ++d) ps.F[9 * p + d] = F[d];
    for (int d = 0; d < 3; ++d) ps.v[3 * p + d] = vel[d];
    for (int d = 0; d < 9; ++d) ps.C[9 * p + d] = C[d];
  }
}

// ----------------------------------------------------------------------------
// P2C2G / G2C2P: particle -> cell -> node transfers over cell-level buckets
// ----------------------------------------------------------------------------

/// The IndexBuckets of index_buckets_for_particles(pol, particles, dx, 0.5)
/// (simulation/particle/Query.tpp:9-56): a table of cells built with
/// blockLen 1, offset 0 and displacement 0.5 (the cell of x is
/// floor(x / dx + 0.5), the collocated bucketCoord of
/// container/IndexBuckets.hpp:96-104) and each cell's particles by
/// increasing index.
struct CellBuckets {
  Table t;
  std::vector<i32> counts, offsets, indices;
  template <typename T>
  void build(const T* x, i64 n, T dx) {
    t.resize((i32)evaluate_table_size(n > 0 ? n : 1));
    compute_sparsity(t, x, n, dx, 1, 0, (T)0.5);
    counts.assign((size_t)t.cnt + 1, 0);
    offsets.assign((size_t)t.cnt + 1, 0);
    indices.assign((size_t)n, 0);
    bucket_sort(t, x, n, dx, 1, 0, (T)0.5, counts.data(), offsets.data(), indices.data());
  }
};

/// The per-axis factor of P2C2G / G2C2P (P2C2G.hpp:86-90, G2C2P.hpp:235-238):
/// Xrel = x - lower_trunc(x dx_inv + 0.5) dx, the offset from the nearest
/// node, and Dinv = 2 / (dx^2 - 2 Xrel^2).
template <typename T>
inline void cell_dinv(const T* pos, T dx, T dx_inv, T Dinv[3]) {
  for (int d = 0; d < 3; ++d) {
    const T xr = pos[d] - (T)(i32)std::floor(pos[d] * dx_inv + (T)0.5) * dx;
    Dinv[d] = (T)2 / (dx * dx - (T)2 * xr * xr);
  }
}

/// P2C2GTransfer::operator()(blockid, cellid) over Collapse{table.cnt, 64}
/// (simulation/transfer/P2C2G.hpp:53-181), fixed-corotated and Drucker-Prager
/// models. Stage 1 gathers, for the cell at coord (centre (coord + 0.5) dx),
/// the particles of the 27 buckets coord - 1 .. coord + 1 that lie within dx
/// on every axis, with C scaled per column by Dinv, the model's vol P F^T
/// scaled by Dinv (-dt) plus m C, and the tent weight prod(1 - |xc - xp| / dx)
/// into m_c, mv_c, QDinv_c and QDinvXp_c; stage 2 adds (m_c, mv_c + QDinv_c
/// x_i - QDinvXp_c) / 8 to the 2x2x2 nodes x_i = (coord + {0,1}^3) dx of the
/// cell that have a block. Writes channels 0..3 of the grid.
///
/// logJp: the reference's functor reads a particle's logJp, runs the return
/// mapping and writes it back from every cell that gathers the particle (up
/// to eight, concurrently on the GPU), so the stored value depends on the
/// interleaving. This restatement takes the interleaving in which every
/// gather reads the logJp the transfer started with: the stress is the same
/// from every cell and the value written is the once-updated one.
template <typename T>
void p2c2g(const Material<T>& mat, T dt, T dx, ParticlesView<T> ps, Table& t, CellBuckets& ib, T* grid) {
  const T dx_inv = (T)1.0 / dx;
  T mu, lambda;
  lame(mat.E, mat.nu, mu, lambda);
  std::vector<T> lj0;
  if (ps.logJp) lj0.assign(ps.logJp, ps.logJp + ps.n);
  for (i32 b = 0; b < t.cnt; ++b)
    for (int cellid = 0; cellid < 64; ++cellid) {
      T m_c = 0, mv_c[3] = {0, 0, 0}, QDinv_c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, QDinvXp_c[3] = {0, 0, 0};
      const i32 coord[3] = {t.active[3 * b] * 4 + (cellid >> 4), t.active[3 * b + 1] * 4 + ((cellid >> 2) & 3),
                            t.active[3 * b + 2] * 4 + (cellid & 3)};
      T posc[3];
      for (int d = 0; d < 3; ++d) posc[d] = ((T)coord[d] + (T)0.5) * dx;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            const i32 bc[3] = {coord[0] - 1 + i, coord[1] - 1 + j, coord[2] - 1 + k};
            const i32 bno = ib.t.query(bc);
            if (bno < 0) continue;
            for (i32 st = ib.offsets[bno], ed = ib.offsets[bno + 1]; st != ed; ++st) {
              const i64 p = ib.indices[st];
              const T* posp = ps.x + 3 * p;
              bool in = true;
              for (int d = 0; d < 3; ++d)
                if (std::abs(posp[d] - posc[d]) > dx) in = false;
              if (!in) continue;
              T Dinv[3];
              cell_dinv(posp, dx, dx_inv, Dinv);
              const T* vel = ps.v + 3 * p;
              const T mass = ps.m[p];
              T C[9], contrib[9];
              for (int d = 0; d < 9; ++d) C[d] = ps.C[9 * p + d] * Dinv[d / 3];
              if (mat.kind == kFixedCorotated) {
                stress_fixedcorotated(mat.volume, mu, lambda, ps.F + 9 * p, contrib);
              } else {
                T logJp = lj0[p];
                stress_sand(mat.volume, mu, lambda, mat.cohesion, mat.beta, mat.yieldSurface,
                            mat.volumeCorrection, logJp, ps.F + 9 * p, contrib);
                ps.logJp[p] = logJp;
              }
              for (int d = 0; d < 9; ++d) contrib[d] = contrib[d] * (Dinv[d / 3] * -dt);
              for (int d = 0; d < 9; ++d) contrib[d] = contrib[d] + C[d] * mass;
              T Wpc = 1;
              for (int d = 0; d < 3; ++d) Wpc *= (T)1 - std::abs((posc[d] - posp[d]) * dx_inv);
              m_c += mass * Wpc;
              for (int d = 0; d < 3; ++d) {
                mv_c[d] += mass * vel[d] * Wpc;
                QDinvXp_c[d] += (contrib[d] * posp[0] + contrib[3 + d] * posp[1] + contrib[6 + d] * posp[2]) * Wpc;
              }
              for (int d = 0; d < 9; ++d) QDinv_c[d] += contrib[d] * Wpc;
            }
          }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int k = 0; k < 2; ++k) {
            const i32 ci[3] = {coord[0] + i, coord[1] + j, coord[2] + k};
            const T posi[3] = {(T)ci[0] * dx, (T)ci[1] * dx, (T)ci[2] * dx};
            i32 nb;
            int cell;
            unpack_coord(t, ci, nb, cell);
            if (nb < 0) continue;
            const T Wci = (T)(1. / 8);
            grid[gidx(nb, 0, cell)] += m_c * Wci;
            for (int d = 0; d < 3; ++d)
              grid[gidx(nb, 1 + d, cell)] +=
                  (mv_c[d] + ((QDinv_c[d] * posi[0] + QDinv_c[3 + d] * posi[1] + QDinv_c[6 + d] * posi[2]) -
                              QDinvXp_c[d])) *
                  Wci;
          }
    }
}

/// PreG2C2PTransfer (G2C2P.hpp:197-212: v = 0, B = 0), G2C2PTransfer over
/// Collapse{table.cnt, 64} (:58-151: the cell's v_c and v_c x^T from its
/// 2x2x2 nodes with weight 1/8, then for each particle of the 27 buckets
/// around it within dx on every axis, v_p += W v_c and
/// B_p += W (v_c x^T - v_c x_p^T), W the tent weight, zero beyond one cell)
/// and PostG2C2PTransfer (:214-271: C = B scaled per column by Dinv at the
/// old position, F = (I + dt C) F, x += v dt; B stays in the C attribute).
/// The grid's channels 1..3 hold the node velocities.
template <typename T>
void g2c2p(T dt, T dx, ParticlesView<T> ps, Table& t, CellBuckets& ib, const T* grid) {
  const T dx_inv = (T)1 / dx;
  for (i64 p = 0; p < ps.n; ++p) {
    for (int d = 0; d < 3; ++d) ps.v[3 * p + d] = 0;
    for (int d = 0; d < 9; ++d) ps.C[9 * p + d] = 0;
  }
  for (i32 b = 0; b < t.cnt; ++b)
    for (int cellid = 0; cellid < 64; ++cellid) {
      const i32 coord[3] = {t.active[3 * b] * 4 + (cellid >> 4), t.active[3 * b + 1] * 4 + ((cellid >> 2) & 3),
                            t.active[3 * b + 2] * 4 + (cellid & 3)};
      T v_c[3] = {0, 0, 0}, vx_c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int k = 0; k < 2; ++k) {
            const i32 ci[3] = {coord[0] + i, coord[1] + j, coord[2] + k};
            const T posi[3] = {(T)ci[0] * dx, (T)ci[1] * dx, (T)ci[2] * dx};
            i32 nb;
            int cell;
            unpack_coord(t, ci, nb, cell);
            if (nb < 0) continue;
            const T W = (T)(1. / 8);
            const T vi[3] = {grid[gidx(nb, 1, cell)], grid[gidx(nb, 2, cell)], grid[gidx(nb, 3, cell)]};
            for (int d = 0; d < 3; ++d) v_c[d] += vi[d] * W;
            for (int d = 0; d < 9; ++d) vx_c[d] += W * vi[d % 3] * posi[d / 3];
          }
      T posc[3];
      for (int d = 0; d < 3; ++d) posc[d] = ((T)coord[d] + (T)0.5) * dx;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            const i32 bc[3] = {coord[0] - 1 + i, coord[1] - 1 + j, coord[2] - 1 + k};
            const i32 bno = ib.t.query(bc);
            if (bno < 0) continue;
            for (i32 st = ib.offsets[bno], ed = ib.offsets[bno + 1]; st != ed; ++st) {
              const i64 p = ib.indices[st];
              const T* posp = ps.x + 3 * p;
              bool in = true;
              for (int d = 0; d < 3; ++d)
                if (std::abs(posp[d] - posc[d]) > dx) in = false;
              if (!in) continue;
              T W = 1;
              for (int d = 0; d < 3; ++d) {
                const T xabs = std::abs((posc[d] - posp[d]) * dx_inv);
                W *= xabs <= (T)1 ? (T)1 - xabs : (T)0;
              }
              for (int d = 0; d < 3; ++d) ps.v[3 * p + d] += v_c[d] * W;
              for (int d = 0; d < 9; ++d) ps.C[9 * p + d] += W * (vx_c[d] - v_c[d % 3] * posp[d / 3]);
            }
          }
    }
  for (i64 p = 0; p < ps.n; ++p) {
    T* pos = ps.x + 3 * p;
    T Dinv[3];
    cell_dinv(pos, dx, dx_inv, Dinv);
    T tmp[9], F[9];
    for (int d = 0; d < 9; ++d) tmp[d] = ps.C[9 * p + d] * Dinv[d / 3] * dt + ((d & 0x3) ? (T)0.f : (T)1.f);
    matmul3(tmp, ps.F + 9 * p, F);
    for (int d = 0; d < 9; ++d) ps.F[9 * p + d] = F[d];
    for (int d = 0; d < 3; ++d) pos[d] += ps.v[3 * p + d] * dt;
  }
}

// ----------------------------------------------------------------------------
// one explicit substep, stage order of the reconstructed driver
// ----------------------------------------------------------------------------

/// Inputs of one substep. The activation convention (blockLen 4, offset -2,
/// displacement 0.5) and the enlargement {0,0,0}..{2,2,2} are the ones that
/// make every 3x3x3 quadratic stencil fall inside active blocks
/// (SparsityOp.hpp:62-79, :88-112).
template <typename T>
struct StepParams {
  T dt = (T)1e-4, dx = (T)1 / 128;
  T extf[3] = {0, (T)-9.8, 0};
  Material<T> mat;
  const PlaneCollider<T>* cols = nullptr;
  int ncols = 0;
  const GeneralCollider<T>* gcols = nullptr;  // applied after the planes, in order
  int ngcols = 0;
  i32 tableSize = 0;
};

/// Outputs of one substep beyond the updated particles.
template <typename T>
struct StepOutputs {
  Table table;
  std::vector<i32> counts, offsets, indices;
  std::vector<T> grid;  // after P2G, before the fold/update (for parity on m, mv)
  std::vector<T> gridv; // after the update and boundary (node velocities)
  T maxVel = 0;
};

// ----------------------------------------------------------------------------
// implicit step: G2P2G and the ImplicitMPMSystem operators, the PCG solve
// ----------------------------------------------------------------------------

/// G2P2GTransfer::operator() over range(N), simulation/transfer/G2P2G.hpp:49-146
/// (APIC, quadratic): C from the node vector v (3 floats per node, node =
/// block * 64 + cell), F = (I + dt C) F_old (not stored), contrib = the
/// model's vol P F^T times D_inv (logJp read, not written back), and
/// r[3 node + d] += W (contrib[d] xixp0 + contrib[3+d] xixp1 + contrib[6+d] xixp2).
template <typename T>
void g2p2g(const Material<T>& mat, T dt, T dx, ParticlesView<T> ps, Table& t, const T* v, T* r) {
  const T dx_inv = (T)1 / dx;
  const T D_inv = (T)4.f * dx_inv * dx_inv;
  T mu, lambda;
  lame(mat.E, mat.nu, mu, lambda);
  for (i64 p = 0; p < ps.n; ++p) {
    Arena<T> ar;
    ar.init(dx, ps.x + 3 * p);
    T C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          i32 coord[3] = {ar.corner[0] + i, ar.corner[1] + j, ar.corner[2] + k};
          i32 b;
          int cell;
          unpack_coord(t, coord, b, cell);
          T xixp[3];
          ar.diff(i, j, k, xixp);
          const T W = ar.weight(i, j, k);
          const T* vi = v + 3 * ((i64)b * 64 + cell);
          for (int d = 0; d < 9; ++d) C[d] += W * vi[d % 3] * xixp[d / 3] * D_inv;
        }
    T tmp[9], F[9], contrib[9];
    for (int d = 0; d < 9; ++d) tmp[d] = C[d] * dt + ((d & 0x3) ? (T)0.f : (T)1.f);
    matmul3(tmp, ps.F + 9 * p, F);
    if (mat.kind == kFixedCorotated) {
      stress_fixedcorotated(mat.volume, mu, lambda, F, contrib);
    } else {
      T logJp = ps.logJp[p];
      stress_sand(mat.volume, mu, lambda, mat.cohesion, mat.beta, mat.yieldSurface, mat.volumeCorrection, logJp,
                  F, contrib);
    }
    for (int d = 0; d < 9; ++d) contrib[d] = contrib[d] * D_inv;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          i32 coord[3] = {ar.corner[0] + i, ar.corner[1] + j, ar.corner[2] + k};
          i32 b;
          int cell;
          unpack_coord(t, coord, b, cell);
          T xixp[3];
          ar.diff(i, j, k, xixp);
          const T W = ar.weight(i, j, k);
          T* ri = r + 3 * ((i64)b * 64 + cell);
          for (int d = 0; d < 3; ++d)
            ri[d] += W * (contrib[d] * xixp[0] + contrib[3 + d] * xixp[1] + contrib[6 + d] * xixp[2]);
        }
  }
}

/// ImplicitMPMSystem::multiply (simulation/mpm/ImplicitMPM.hpp:35-63): out = 0,
/// out += G2P2G(in), then out_i = (out_i dt dt + m) in_i on nodes with m > 0
/// (ForceDtSqrPlusMass, :17-33); other nodes keep the force.
template <typename T>
void implicit_multiply(const Material<T>& mat, T dt, T dx, ParticlesView<T> ps, Table& t, const T* grid,
                       i64 nblocks, const T* in, T* out) {
  const i64 n = nblocks * 64 * 3;
  for (i64 i = 0; i < n; ++i) out[i] = 0;
  g2p2g(mat, dt, dx, ps, t, in, out);
  for (i64 i = 0; i < n; ++i) {
    const i64 node = i / 3;
    const T mass = grid[gidx(node / 64, 0, (int)(node % 64))];
    if (mass > 0) out[i] = (out[i] * dt * dt + mass) * in[i];
  }
}

/// ImplicitMPMSystem::project (:103-127) with its Projector (:65-101), once per
/// boundary in list order (planes, then general colliders): nodes with m > 0
/// resolve the collision on their vector, other nodes are cleared. Without
/// boundaries nothing changes.
template <typename T>
void implicit_project(const T* grid, const Table& t, i64 nblocks, T dx, const PlaneCollider<T>* cols, int ncols,
                      const GeneralCollider<T>* gcols, int ngcols, T* dof) {
  for (int ci = 0; ci < ncols + ngcols; ++ci)
    for (i64 b = 0; b < nblocks; ++b)
      for (int c = 0; c < 64; ++c) {
        T* v = dof + 3 * (b * 64 + c);
        if (!(grid[gidx(b, 0, c)] > 0)) {
          v[0] = v[1] = v[2] = 0;
          continue;
        }
        const i32 cc[3] = {(c >> 4) & 3, (c >> 2) & 3, c & 3};
        T pos[3];
        for (int d = 0; d < 3; ++d) pos[d] = ((T)t.active[3 * b + d] * (T)4 + (T)cc[d]) * dx;
        if (ci < ncols) {
          const PlaneCollider<T>& col = cols[ci];
          const T phi = col.normal[0] * (pos[0] - col.origin[0]) + col.normal[1] * (pos[1] - col.origin[1]) +
                        col.normal[2] * (pos[2] - col.origin[2]);
          if (phi < (T)0) {
            if (col.type == kSticky) {
              v[0] = v[1] = v[2] = 0;
            } else {
              const T proj = col.normal[0] * v[0] + col.normal[1] * v[1] + col.normal[2] * v[2];
              if ((col.type == kSeparate && proj < 0) || col.type == kSlip)
                for (int d = 0; d < 3; ++d) v[d] -= proj * col.normal[d];
            }
          }
        } else {
          resolve_general(gcols[ci - ncols], pos, v);
        }
      }
}

/// ImplicitMPMSystem::precondition (:129-155, DivPernodeMass): out_i = in_i / m
/// on nodes with m > 0; other entries of out are left alone.
template <typename T>
void implicit_precondition(const T* grid, i64 nblocks, const T* in, T* out) {
  const i64 n = nblocks * 64 * 3;
  for (i64 i = 0; i < n; ++i) {
    const i64 node = i / 3;
    const T mass = grid[gidx(node / 64, 0, (int)(node % 64))];
    if (mass > 0) out[i] = in[i] / mass;
  }
}

/// ConjugateGradient::solve (math/linear/ConjugateGradient.hpp:70-175) on the
/// system above: r = project(b - A x), q = M r, p = q, zTrk = r.q,
/// tolerance min(relTol sqrt(zTrk), tol); per iteration temp = project(A p),
/// alpha = zTrk / temp.p, x += alpha p, r -= alpha temp, q = M r,
/// beta = (q.r) / zTrk, p = q + beta p. Dot products are summed in T in index
/// order. Returns the iterations taken.
template <typename T>
int implicit_solve(const StepParams<T>& sp, ParticlesView<T> ps, Table& t, const T* grid, i64 nblocks, T* x,
                   const T* b, int maxIters, T tol, T relTol, T* residual) {
  const i64 n = nblocks * 64 * 3;
  std::vector<T> r(n), p(n), q(n, 0), temp(n);
  auto dot = [n](const T* a, const T* c) {
    T s = 0;
    for (i64 i = 0; i < n; ++i) s += a[i] * c[i];
    return s;
  };
  auto A = [&](const T* in, T* out) { implicit_multiply(sp.mat, sp.dt, sp.dx, ps, t, grid, nblocks, in, out); };
  auto project = [&](T* v) { implicit_project(grid, t, nblocks, sp.dx, sp.cols, sp.ncols, sp.gcols, sp.ngcols, v); };
  A(x, temp.data());
  for (i64 i = 0; i < n; ++i) r[i] = b[i] - temp[i];
  project(r.data());
  implicit_precondition(grid, nblocks, r.data(), q.data());
  p = q;
  T zTrk = dot(r.data(), q.data());
  T norm = std::sqrt(zTrk);
  const T localTol = std::min(relTol * norm, tol);
  int iter = 0;
  for (; iter != maxIters; ++iter) {
    if (norm <= localTol) break;
    A(p.data(), temp.data());
    project(temp.data());
    const T alpha = zTrk / dot(temp.data(), p.data());
    for (i64 i = 0; i < n; ++i) x[i] = alpha * p[i] + x[i];
    for (i64 i = 0; i < n; ++i) r[i] = -alpha * temp[i] + r[i];
    implicit_precondition(grid, nblocks, r.data(), q.data());
    const T zLast = zTrk;
    zTrk = dot(q.data(), r.data());
    const T beta = zTrk / zLast;
    for (i64 i = 0; i < n; ++i) p[i] = beta * p[i] + q[i];
    norm = std::sqrt(zTrk);
  }
  if (residual) *residual = norm;
  return iter;
}

/// Vector<T>::reorder(pol, map, false_c) applied to every attribute
/// (container/Vector.hpp:617-646, the gather form ordered[i] = vs[map[i]]),
/// with an optional particle-id array carried along.
template <typename T>
void reorder_particles(ParticlesView<T> ps, const i32* map, i32* ids) {
  const i64 n = ps.n;
  auto gather = [&](T* a, int w) {
    if (!a) return;
    std::vector<T> tmp((size_t)n * w);
    for (i64 i = 0; i < n; ++i)
      for (int k = 0; k < w; ++k) tmp[(size_t)i * w + k] = a[(i64)map[i] * w + k];
    std::copy(tmp.begin(), tmp.end(), a);
  };
  gather(ps.x, 3);
  gather(ps.v, 3);
  gather(ps.C, 9);
  gather(ps.F, 9);
  gather(ps.m, 1);
  gather(ps.logJp, 1);
  if (ids) {
    std::vector<i32> tmp(n);
    for (i64 i = 0; i < n; ++i) tmp[i] = ids[map[i]];
    std::copy(tmp.begin(), tmp.end(), ids);
  }
}

/// One substep: clear table, activate, enlarge, bucket-sort, clear grid, P2G,
/// fold, grid velocity, boundary, G2P (the stage list of the reconstructed
/// explicit driver). With `reorder`, the driver then reorders every particle
/// attribute by the substep's bucket sort (the Vector::reorder gather), so the
/// particles are stored block by block for the next substep.
template <typename T>
void substep(const StepParams<T>& sp, ParticlesView<T> ps, StepOutputs<T>& out,
             bool keep_grids = true, bool reorder = false, i32* ids = nullptr) {
  Table& t = out.table;
  if (t.T != sp.tableSize) t.resize(sp.tableSize);
  t.clear();
  compute_sparsity(t, ps.x, ps.n, sp.dx, 4, -2, (T)0.5f);
  const i32 lo[3] = {0, 0, 0}, hi[3] = {2, 2, 2};
  enlarge_sparsity(t, lo, hi);
  const i64 nb = t.cnt;
  out.counts.assign(nb + 1, 0);
  out.offsets.assign(nb + 1, 0);
  out.indices.assign(ps.n, 0);
  bucket_sort(t, ps.x, ps.n, sp.dx, 4, -2, (T)0.5f, out.counts.data(), out.offsets.data(),
              out.indices.data());
  std::vector<T> grid(nb * kChannels * 64);
  grid_clear(grid.data(), nb);
  p2g(sp.mat, sp.dt, sp.dx, ps, t, grid.data());
  if (keep_grids) out.grid = grid;
  fold_rhs(grid.data(), nb);
  T maxVel = 0;
  grid_velocity(grid.data(), nb, sp.dt, sp.extf, maxVel);
  grid_boundary(grid.data(), t, nb, sp.dx, sp.cols, sp.ncols);
  grid_boundary_general(grid.data(), t, nb, sp.dx, sp.gcols, sp.ngcols);
  out.maxVel = maxVel;
  g2p(sp.dt, sp.dx, ps, t, grid.data());
  if (keep_grids) out.gridv = std::move(grid);
  if (reorder) reorder_particles(ps, out.indices.data(), ids);
}

}  // namespace zso
