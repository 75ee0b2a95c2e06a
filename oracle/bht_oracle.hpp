// Test infrastructure, not product code: the oracle's restatement of zpc's
// bucketed hash table (container/Bht.hpp), the SparseGrid built on it
// (geometry/SparseGrid.hpp) and the level-set boundary that samples it
// (geometry/Collider.h:10-246, simulation/grid/GridOp.hpp:112-164). Only the
// tests and the bench's host baseline use it, as the checker of the HIP path.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace zso {

// MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998), the engine of
// std::mt19937 that bht's constructor draws its hash parameters from
// (container/Bht.hpp:160-166): seeding by the published recurrence
// x_i = 1812433253 (x_{i-1} xor (x_{i-1} >> 30)) + i, the twist, the tempering.
struct Mt19937 {
  uint32_t mt[624];
  int idx = 624;
  explicit Mt19937(uint32_t seed) {
    mt[0] = seed;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  }
  uint32_t operator()() {
    if (idx >= 624) {
      for (int i = 0; i < 624; ++i) {
        const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
        mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      idx = 0;
    }
    uint32_t y = mt[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
};

constexpr uint32_t kPrimeDivisor = 4294967291u;  // universal_hash, container/Bcht.hpp:36
constexpr int kBucket = 16;                       // bht<i32, 3, int, 16>
constexpr int kThreshold = kBucket - 2;           // Bht.hpp:35
constexpr int32_t kBhtKeySentinel = 0x3f3f3f3f;   // keys byte-filled with 0x3f (Bht.hpp:103,127-131)
constexpr int32_t kBhtFailure = INT32_MIN;        // failure_token_v (Bht.hpp:136)

// universal_hash(std::mt19937&) x 3 (Bht.hpp:160-166; Bcht.hpp:39-43):
// hashx = rng() % p (at least 1), hashy = rng() % p, for hf0, hf1, hf2
inline void bht_hash_params(uint32_t seed, uint32_t out[6]) {
  Mt19937 rng(seed);
  for (int h = 0; h < 3; ++h) {
    uint32_t x = rng() % kPrimeDivisor;
    if (x < 1) x = 1;
    const uint32_t y = rng() % kPrimeDivisor;
    out[2 * h] = x;
    out[2 * h + 1] = y;
  }
}

// universal_hash_base on vec<i32,3> (py_interop/HashUtils.hpp:22-41): each
// component through ((hx ^ k) + hy) % p in u32, combined by the u32
// hash_combine of math/Hash.hpp:18-26
inline uint32_t bht_hash(uint32_t hx, uint32_t hy, const int32_t k[3]) {
  auto sub = [&](int32_t v) -> uint32_t { return ((hx ^ (uint32_t)v) + hy) % kPrimeDivisor; };
  uint32_t ret = sub(k[0]);
  for (int d = 1; d < 3; ++d) ret ^= (sub(k[d]) + 0x9e3779b9u + (ret << 6) + (ret >> 2));
  return ret;
}

// evaluateTableSize (Bht.hpp:153-157): 2 next_pow2(n), rounded up to a
// multiple of the bucket size -- by a whole extra bucket when it already is one
inline int64_t bht_table_size(int64_t n) {
  if (n <= 0) return 0;
  int64_t p = 1;
  while (p < n) p <<= 1;
  const int64_t m = 2 * p;
  return m + (kBucket - m % kBucket);
}

struct Bht {
  int32_t T = 0;
  std::vector<int32_t> keys;     // [4 T]: storage_key_type of vec<i32,3>, 16 bytes (x, y, z, pad)
  std::vector<int32_t> indices;  // [T]
  std::vector<int32_t> status;   // [T]
  int32_t cnt = 0;
  std::vector<int32_t> active;   // [3 T]
  int32_t success = 1;
  uint32_t h[6] = {};

  void init(int32_t table_size, const uint32_t hp[6]) {
    T = table_size;
    keys.assign((size_t)4 * T, kBhtKeySentinel);
    indices.assign((size_t)T, -1);
    status.assign((size_t)T, -1);
    active.assign((size_t)3 * T, 0);
    cnt = 0;
    success = 1;
    for (int i = 0; i < 6; ++i) h[i] = hp[i];
  }
  // bht::reset(clearCnt = true) (Bht.hpp:316-320): keys to the sentinel,
  // status to -1, indices kept
  void clear() {
    for (auto& k : keys) k = kBhtKeySentinel;
    for (auto& s : status) s = -1;
    cnt = 0;
    success = 1;
  }
  int64_t nbuckets() const { return T / kBucket; }
  int64_t bucket_offset(int level, const int32_t k[3]) const {
    return (int64_t)(bht_hash(h[2 * level], h[2 * level + 1], k) % (uint32_t)nbuckets()) * kBucket;
  }
  bool key_at(int64_t slot, const int32_t k[3]) const {
    return keys[4 * slot] == k[0] && keys[4 * slot + 1] == k[1] && keys[4 * slot + 2] == k[2];
  }
  bool empty_at(int64_t slot) const {
    return keys[4 * slot] == kBhtKeySentinel && keys[4 * slot + 1] == kBhtKeySentinel &&
           keys[4 * slot + 2] == kBhtKeySentinel;
  }

  // BHTView::insert, the host path (Bht.hpp:496-548): scan the bucket of hf0
  // for the key or the first empty slot; insert there while the bucket holds
  // at most `threshold` keys, else move on to hf1, then hf2. Returns the new
  // index, -1 if the key is present, the failure token when all three buckets
  // are full or the index comes within 20 of the table size.
  int32_t insert(const int32_t k[3]) {
    if (nbuckets() == 0) return kBhtFailure;
    int iter = 0, load = 0;
    int64_t off = bucket_offset(0, k);
    while (iter < 3) {
      for (; load != kBucket; ++load) {
        if (key_at(off + load, k)) {
          load = -1;
          break;
        }
        if (empty_at(off + load)) break;
      }
      if (load < 0) return -1;
      if (load <= kThreshold) {
        const int64_t s = off + load;
        keys[4 * s] = k[0];
        keys[4 * s + 1] = k[1];
        keys[4 * s + 2] = k[2];
        int32_t localno = cnt++;
        indices[s] = localno;
        active[3 * (size_t)localno] = k[0];
        active[3 * (size_t)localno + 1] = k[1];
        active[3 * (size_t)localno + 2] = k[2];
        if (localno >= T - 20) {
          success = 0;
          localno = kBhtFailure;
        }
        return localno;
      }
      ++iter;
      load = 0;
      if (iter < 3) off = bucket_offset(iter, k);
    }
    success = 0;
    return kBhtFailure;
  }

  // BHTView::query (Bht.hpp:551-580): the index, or -1
  int32_t query(const int32_t k[3]) const {
    if (nbuckets() == 0) return -1;
    for (int iter = 0; iter < 3; ++iter) {
      const int64_t off = bucket_offset(iter, k);
      for (int loc = 0; loc < kBucket; ++loc)
        if (key_at(off + loc, k)) return indices[off + loc];
    }
    return -1;
  }
};

// SparseGrid<3, f32, 8> (geometry/SparseGrid.hpp:16-48): a bht of block
// origins (multiples of 8), a TileVector<f32, 512> of the blocks' cells
// (block b, channel c, cell k at (b nch + c) 512 + k, k = (x 8 + y) 8 + z,
// SparseGridView::local_coord_to_offset :277-283), the index-to-world
// transform (row-vector convention: world = [X, 1] M, the translation in row
// 3; SparseGridView::indexToWorld :255-262) and the background value.
struct SparseGrid {
  Bht table;
  int32_t nch = 1;
  std::vector<float> data;
  float i2w[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float w2i[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float background = 0.f;

  static void apply(const float M[16], const float x[3], float out[3]) {
    for (int j = 0; j < 3; ++j) out[j] = x[0] * M[j] + x[1] * M[4 + j] + x[2] * M[8 + j] + M[12 + j];
  }
  // SparseGridView::insert / query's block key (:439-455): X = floor(w2i x + 1/2), minus X & 7
  void block_key(const float x[3], int32_t k[3]) const {
    float X[3];
    apply(w2i, x, X);
    for (int d = 0; d < 3; ++d) {
      const int32_t c = (int32_t)std::floor(X[d] + 0.5f);
      k[d] = c - (c & 7);
    }
  }
  // valueOr at an integer coordinate (:340-352 via decomposeCoord :302-308)
  float value_or(int chn, const int32_t c[3], float def) const {
    int32_t k[3], cell[3];
    for (int d = 0; d < 3; ++d) {
      cell[d] = c[d] & 7;
      k[d] = c[d] - cell[d];
    }
    const int32_t b = table.query(k);
    if (b == -1) return def;
    const int64_t cno = (cell[0] * 8 + cell[1]) * 8 + cell[2];
    return data[((size_t)b * nch + chn) * 512 + cno];
  }
  // iSample, kernel_e::linear (:458-470; GridArena, math/curve/InterpolationKernel.hpp:274-470):
  // corner = floor(X), the 2^3 values around it (background where the block
  // is inactive), xlerp with lerp(a, b, t) = a + (b - a) t, x outermost
  float isample(int chn, const float X[3]) const {
    int32_t corner[3];
    float t[3];
    for (int d = 0; d < 3; ++d) {
      corner[d] = (int32_t)std::floor(X[d]);
      t[d] = X[d] - (float)corner[d];
    }
    float a[2][2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) {
          const int32_t c[3] = {corner[0] + i, corner[1] + j, corner[2] + k};
          a[i][j][k] = value_or(chn, c, background);
        }
    auto lerp = [](float t, float p, float q) { return p + (q - p) * t; };
    float yz[2];
    for (int i = 0; i < 2; ++i)
      yz[i] = lerp(t[1], lerp(t[2], a[i][0][0], a[i][0][1]), lerp(t[2], a[i][1][0], a[i][1][1]));
    return lerp(t[0], yz[0], yz[1]);
  }
  float wsample(int chn, const float x[3]) const {
    float X[3];
    apply(w2i, x, X);
    return isample(chn, X);
  }
  // voxelSize(0) (:389-397): the length of row 0 of the transform's 3 x 3 part
  float voxel_size0() const {
    float s = 0.f;
    for (int d = 0; d < 3; ++d) s += i2w[d] * i2w[d];
    return std::sqrt(s);
  }
  // do_getNormal (:524-538): central differences of the signed distance with
  // eps = voxelSize(0) / 4, normalized
  void normal(int sdf, const float x[3], float n[3]) const {
    const float eps = voxel_size0() / 4.f;
    for (int i = 0; i < 3; ++i) {
      float p[3] = {x[0], x[1], x[2]}, q[3] = {x[0], x[1], x[2]};
      p[i] = x[i] + eps;
      q[i] = x[i] - eps;
      n[i] = (wsample(sdf, p) - wsample(sdf, q)) / (eps + eps);
    }
    const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; ++i) n[i] = n[i] / len;
  }
};

// ApplyBoundaryConditionOnGridBlocks with a LevelSetBoundary<SparseLevelSet>
// (GridOp.hpp:127-134): the boundary's collider keeps the identity motion, so
// Collider::resolveCollision (Collider.h:77-107) reduces to: where the signed
// distance at the node is negative, the object velocity is the grid's
// material velocity ("v" channels, zero without them); sticky takes it, slip
// removes the normal part of the velocity relative to it, separate only when
// approaching. Returns whether the node was inside.
inline bool levelset_resolve(const SparseGrid& g, int sdf, int vchn, int type, const float x[3], float v[3]) {
  if (!(g.wsample(sdf, x) < 0.f)) return false;
  float vo[3] = {0.f, 0.f, 0.f};
  if (vchn >= 0)
    for (int d = 0; d < 3; ++d) vo[d] = g.wsample(vchn + d, x);
  if (type == 0) {
    for (int d = 0; d < 3; ++d) v[d] = vo[d];
    return true;
  }
  for (int d = 0; d < 3; ++d) v[d] -= vo[d];
  float n[3];
  g.normal(sdf, x, n);
  const float proj = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  if ((type == 2 && proj < 0.f) || type == 1)
    for (int d = 0; d < 3; ++d) v[d] -= proj * n[d];
  for (int d = 0; d < 3; ++d) v[d] += vo[d];
  return true;
}

}  // namespace zso
