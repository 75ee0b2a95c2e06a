// Device views and per-thread helpers of zpc's bucketed hash table
// (bht<i32, 3, int, 16>, container/Bht.hpp) and the SparseGrid<3, f32, 8> on
// it (geometry/SparseGrid.hpp): universal hashing, bucket queries, trilinear
// sampling with the background value, the signed-distance normal and the
// level-set boundary of the grid update.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zsmpm {

constexpr int kBhtBucket = 16;                 // bucket_size of bht<i32, 3, int, 16>
constexpr int kBhtThreshold = kBhtBucket - 2;  // a bucket takes keys while it holds at most 14
constexpr int32_t kBhtSentinelKey = 0x3f3f3f3f;
constexpr int32_t kBhtFailure = INT32_MIN;
constexpr uint32_t kBhtPrime = 4294967291u;

struct BhtView {
  int32_t T = 0;   // slots, a multiple of 16
  int32_t nb = 0;  // buckets
  int4* keys = nullptr;        // [T] storage keys (x, y, z, pad)
  int32_t* indices = nullptr;  // [T]
  int32_t* status = nullptr;   // [T]
  int32_t* cnt = nullptr;      // [1]
  int32_t* active = nullptr;   // [3T]
  int32_t* success = nullptr;  // [1]
  uint32_t h[6] = {};          // (hashx, hashy) of hf0, hf1, hf2
};

// universal_hash_base on vec<i32,3> (py_interop/HashUtils.hpp:22-41): every
// component through ((hx ^ k) + hy) mod p, combined with the u32
// hash_combine of math/Hash.hpp:18-26
__device__ __forceinline__ uint32_t bht_hash(uint32_t hx, uint32_t hy, int32_t k0, int32_t k1, int32_t k2) {
  uint32_t r = ((hx ^ (uint32_t)k0) + hy) % kBhtPrime;
  const uint32_t s1 = ((hx ^ (uint32_t)k1) + hy) % kBhtPrime;
  r ^= s1 + 0x9e3779b9u + (r << 6) + (r >> 2);
  const uint32_t s2 = ((hx ^ (uint32_t)k2) + hy) % kBhtPrime;
  r ^= s2 + 0x9e3779b9u + (r << 6) + (r >> 2);
  return r;
}

__device__ __forceinline__ int32_t bht_bucket(const BhtView& t, int level, int32_t k0, int32_t k1, int32_t k2) {
  return (int32_t)(bht_hash(t.h[2 * level], t.h[2 * level + 1], k0, k1, k2) % (uint32_t)t.nb);
}

// BHTView::query (Bht.hpp:551-580) by one thread: the key's index, or -1
__device__ __forceinline__ int32_t bht_query(const BhtView& t, int32_t k0, int32_t k1, int32_t k2) {
  if (t.nb == 0) return -1;
  for (int l = 0; l < 3; ++l) {
    const int32_t off = bht_bucket(t, l, k0, k1, k2) * kBhtBucket;
    int found = -1;
#pragma unroll
    for (int j = 0; j < kBhtBucket; ++j) {
      const int4 s = t.keys[off + j];
      if (found < 0 && s.x == k0 && s.y == k1 && s.z == k2) found = j;
    }
    if (found >= 0) return t.indices[off + found];
  }
  return -1;
}

// SparseGrid<3, f32, 8>: block origins in the bht, cell values in a
// TileVector<f32, 512> (block b, channel c, cell k at (b nch + c) 512 + k,
// k = (x 8 + y) 8 + z), the index-to-world transform in the row-vector
// convention (world = [X, 1] M) and its inverse, the background value.
struct SpgView {
  BhtView t;
  const float* data = nullptr;
  int32_t nch = 0;
  float i2w[16];
  float w2i[16];
  float background = 0.f;
};

__device__ __forceinline__ void spg_apply(const float* M, float x0, float x1, float x2, float out[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int j = 0; j < 3; ++j) out[j] = x0 * M[j] + x1 * M[4 + j] + x2 * M[8 + j] + M[12 + j];
}

// valueOr at an integer coordinate (SparseGridView::valueOr, :340-352)
__device__ __forceinline__ float spg_value_or(const SpgView& g, int chn, int32_t c0, int32_t c1, int32_t c2) {
  const int32_t b = bht_query(g.t, c0 & ~7, c1 & ~7, c2 & ~7);
  if (b < 0) return g.background;
  const int32_t k = ((c0 & 7) * 8 + (c1 & 7)) * 8 + (c2 & 7);
  return g.data[((int64_t)b * g.nch + chn) * 512 + k];
}

// iSample, kernel_e::linear (GridArena + xlerp): corner = floor(X), the 2^3
// cell values (background where a block is inactive), lerp a + (b - a) t with
// x outermost. The eight values share at most eight block queries; a corner
// inside one block (the common case) needs one.
__device__ __forceinline__ void spg_isample_n(const SpgView& g, int chn, int width, const float X[3], float* out) {
#pragma clang fp contract(off)
  int32_t c[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = (int32_t)floorf(X[d]);
    t[d] = X[d] - (float)c[d];
  }
  int32_t blk[2][2][2];
  const bool one = ((c[0] & 7) != 7) && ((c[1] & 7) != 7) && ((c[2] & 7) != 7);
  if (one) {
    const int32_t b = bht_query(g.t, c[0] & ~7, c[1] & ~7, c[2] & ~7);
#pragma unroll
    for (int i = 0; i < 8; ++i) blk[i >> 2][(i >> 1) & 1][i & 1] = b;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int32_t a0 = c[0] + (i >> 2), a1 = c[1] + ((i >> 1) & 1), a2 = c[2] + (i & 1);
      blk[i >> 2][(i >> 1) & 1][i & 1] = bht_query(g.t, a0 & ~7, a1 & ~7, a2 & ~7);
    }
  }
  for (int w = 0; w < width; ++w) {
    float a[2][2][2];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int i0 = i >> 2, i1 = (i >> 1) & 1, i2 = i & 1;
      const int32_t b = blk[i0][i1][i2];
      const int32_t a0 = c[0] + i0, a1 = c[1] + i1, a2 = c[2] + i2;
      const int32_t k = ((a0 & 7) * 8 + (a1 & 7)) * 8 + (a2 & 7);
      a[i0][i1][i2] = b < 0 ? g.background : g.data[((int64_t)b * g.nch + chn + w) * 512 + k];
    }
    float yz[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float p = a[i][0][0] + (a[i][0][1] - a[i][0][0]) * t[2];
      const float q = a[i][1][0] + (a[i][1][1] - a[i][1][0]) * t[2];
      yz[i] = p + (q - p) * t[1];
    }
    out[w] = yz[0] + (yz[1] - yz[0]) * t[0];
  }
}

__device__ __forceinline__ float spg_wsample(const SpgView& g, int chn, float x0, float x1, float x2) {
#pragma clang fp contract(off)
  float X[3], v;
  spg_apply(g.w2i, x0, x1, x2, X);
  spg_isample_n(g, chn, 1, X, &v);
  return v;
}

// The level-set boundary (LevelSetBoundary<SparseLevelSet> in
// ApplyBoundaryConditionOnGridBlocks, GridOp.hpp:127-134, with
// Collider::resolveCollision, Collider.h:77-107; the boundary's collider keeps
// the identity motion): inside where the sampled signed distance is negative;
// the object velocity is the sampled "v" channels (zero without them); sticky
// takes it, slip removes the normal part of the relative velocity, separate
// only when approaching. The normal is do_getNormal (SparseGrid.hpp:524-538):
// central differences with eps = voxelSize(0) / 4, normalized.
struct LevelSetBoundary {
  int32_t on = 0;
  int32_t sdf = 0;
  int32_t vchn = -1;
  int32_t type = 0;
  float eps = 0.f;
  SpgView g;
};

__device__ __forceinline__ void levelset_resolve(const LevelSetBoundary& ls, const float x[3], float v[3]) {
#pragma clang fp contract(off)
  if (!(spg_wsample(ls.g, ls.sdf, x[0], x[1], x[2]) < 0.f)) return;
  float vo[3] = {0.f, 0.f, 0.f};
  if (ls.vchn >= 0) {
    float X[3];
    spg_apply(ls.g.w2i, x[0], x[1], x[2], X);
    spg_isample_n(ls.g, ls.vchn, 3, X, vo);
  }
  if (ls.type == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = vo[d];
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] -= vo[d];
  float n[3];
  const float e = ls.eps;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float p[3] = {x[0], x[1], x[2]}, q[3] = {x[0], x[1], x[2]};
    p[i] = x[i] + e;
    q[i] = x[i] - e;
    n[i] = (spg_wsample(ls.g, ls.sdf, p[0], p[1], p[2]) - spg_wsample(ls.g, ls.sdf, q[0], q[1], q[2])) / (e + e);
  }
  const float len = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) n[i] = n[i] / len;
  const float proj = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  if ((ls.type == 2 && proj < 0.f) || ls.type == 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] -= proj * n[d];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] += vo[d];
}

}  // namespace zsmpm
