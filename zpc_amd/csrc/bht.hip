// zpc's bucketed hash table (bht<i32, 3, int, 16>, container/Bht.hpp) and the
// SparseGrid<3, f32, 8> stored in it (geometry/SparseGrid.hpp) on gfx950.
//
// Insertion is deterministic: its result is the sequential insertion of the
// keys in input order (the host path of BHTView::insert, Bht.hpp:496-548).
// Repeats and keys already present are found by a deduplication hash; the new
// keys are ranked by first occurrence; then each key's bucket is the first of
// its three hash buckets holding at most 14 keys of lower rank (with the keys
// already there). That condition only looks at lower ranks, so iterating it
// from "every key in its first bucket" reaches the sequential answer: one
// round decides every key whose buckets stay below 15 keys, and each further
// round settles the next link of an overflow chain. A round is a stable sort
// of (bucket, rank) and a binary search per candidate bucket. Slots are the
// rank order within each bucket, after the keys it already held; indices
// continue the table's count in rank order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zsmpm.h"
#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kBT = 256;
constexpr uint64_t kEmpty64 = ~0ull;
constexpr uint32_t kNone = 0xffffffffu;    // an input key out of range
constexpr uint32_t kRepeat = 0xfffffffeu;  // an input equal to the one before it
constexpr int kMaxRounds = 64;

int groups_n(int64_t n, int per = kBT) {
  const int64_t g = (n + per - 1) / per;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  // splitmix64's finaliser (Steele, Lea & Flood, OOPSLA 2014)
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

__device__ __forceinline__ bool key_in_range(int32_t k) { return k >= -(1 << 20) && k < (1 << 20); }

__device__ __forceinline__ uint64_t pack_key(int32_t k0, int32_t k1, int32_t k2) {
  return ((uint64_t)(uint32_t)(k0 + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(k1 + (1 << 20)) << 21) |
         (uint64_t)(uint32_t)(k2 + (1 << 20));
}

__global__ __launch_bounds__(kBT) void bht_clear_kernel(BhtView t) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < t.T; i += stride) {
    t.keys[i] = make_int4(kBhtSentinelKey, kBhtSentinelKey, kBhtSentinelKey, kBhtSentinelKey);
    t.status[i] = -1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *t.cnt = 0;
    *t.success = 1;
  }
}

__global__ __launch_bounds__(kBT) void dedup_clear_kernel(uint64_t* skey, uint32_t* smin, int64_t S) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S; i += stride) {
    skey[i] = kEmpty64;
    smin[i] = kNone;
  }
}

// every input key into the deduplication hash (minimum input index per key)
__global__ __launch_bounds__(kBT) void dedup_insert_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                           uint64_t* skey, uint32_t* smin, int64_t S,
                                                           uint32_t* slot_of, uint32_t* err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // whole waves iterate together, so the lanes can compare keys with their
  // neighbour before touching the hash
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const bool live = i < n;
    int32_t k0 = 0, k1 = 0, k2 = 0;
    if (live) {
      k0 = keys[3 * i];
      k1 = keys[3 * i + 1];
      k2 = keys[3 * i + 2];
    }
    const bool valid = live && key_in_range(k0) && key_in_range(k1) && key_in_range(k2);
    const uint64_t p = valid ? pack_key(k0, k1, k2) : kEmpty64;
    // a lane whose left neighbour holds the same key is not its first
    // occurrence: only the first lane of each run goes to the hash (runs are
    // common: neighbouring particles share a block)
    const uint64_t left = __shfl_up(p, 1, kWave);
    const bool repeat = lane_id() > 0 && valid && left == p;
    if (!live) continue;
    if (!valid) {
      atomicOr(err, kErrKeyRange);
      slot_of[i] = kNone;
      continue;
    }
    if (repeat) {
      slot_of[i] = kRepeat;
      continue;
    }
    uint64_t h = mix64(p) & (uint64_t)(S - 1);
    for (;;) {
      const uint64_t cur = skey[h];
      if (cur == p) break;
      if (cur == kEmpty64) {
        const uint64_t prev = atomicCAS((unsigned long long*)&skey[h], (unsigned long long)kEmpty64,
                                        (unsigned long long)p);
        if (prev == kEmpty64 || prev == p) break;
      }
      h = (h + 1) & (uint64_t)(S - 1);
    }
    atomicMin(&smin[h], (uint32_t)i);
    slot_of[i] = (uint32_t)h;
  }
}

// 1 for the first occurrence of a key the table does not hold yet
__global__ __launch_bounds__(kBT) void first_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                    const uint32_t* slot_of, const uint32_t* smin, BhtView t,
                                                    uint32_t* flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t s = slot_of[i];
    uint32_t f = 0;
    if (s < kRepeat && smin[s] == (uint32_t)i)
      f = bht_query(t, keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]) < 0 ? 1u : 0u;
    flag[i] = f;
  }
}

__global__ __launch_bounds__(kBT) void compact_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                      const uint32_t* flag, const uint32_t* rank, int32_t* nk,
                                                      uint32_t* level) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!flag[i]) continue;
    const uint32_t r = rank[i];
    nk[3 * (int64_t)r] = keys[3 * i];
    nk[3 * (int64_t)r + 1] = keys[3 * i + 1];
    nk[3 * (int64_t)r + 2] = keys[3 * i + 2];
    level[r] = 0;
  }
}

// keys each bucket already holds (slots fill from 0 without gaps)
__global__ __launch_bounds__(kBT) void loads_kernel(BhtView t, uint32_t* L0, uint32_t* c0) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < t.nb; b += stride) {
    uint32_t c = 0;
    for (int j = 0; j < kBhtBucket; ++j) {
      const int4 s = t.keys[b * kBhtBucket + j];
      c += (s.x == kBhtSentinelKey && s.y == kBhtSentinelKey && s.z == kBhtSentinelKey) ? 0u : 1u;
    }
    L0[b] = c;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *c0 = (uint32_t)*t.cnt;
}

__global__ __launch_bounds__(kBT) void assign_kernel(const int32_t* __restrict__ nk, const uint32_t* level,
                                                     uint32_t m, BhtView t, uint32_t* bkey, uint32_t* bval) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m; r += stride) {
    const uint32_t l = level[r];
    bkey[r] = l < 3 ? (uint32_t)bht_bucket(t, (int)l, nk[3 * r], nk[3 * r + 1], nk[3 * r + 2]) : (uint32_t)t.nb;
    bval[r] = (uint32_t)r;
  }
}

__global__ __launch_bounds__(kBT) void seg_clear_kernel(uint32_t* seg_start, uint32_t* seg_end, int64_t nb1,
                                                        uint32_t* changed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb1; b += stride) {
    seg_start[b] = kNone;
    seg_end[b] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *changed = 0;
}

__global__ __launch_bounds__(kBT) void seg_kernel(const uint32_t* sk, uint32_t m, uint32_t* seg_start,
                                                  uint32_t* seg_end) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < m; p += stride) {
    const uint32_t b = sk[p];
    if (p == 0 || sk[p - 1] != b) seg_start[b] = (uint32_t)p;
    if (p == (int64_t)m - 1 || sk[p + 1] != b) seg_end[b] = (uint32_t)p + 1;
  }
}

// number of keys of rank < r in bucket b under the current assignment
__device__ __forceinline__ uint32_t lower_ranks(const uint32_t* sv, const uint32_t* seg_start,
                                                const uint32_t* seg_end, uint32_t b, uint32_t r) {
  const uint32_t s = seg_start[b];
  if (s == kNone) return 0;
  uint32_t lo = s, hi = seg_end[b];
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sv[mid] < r)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - s;
}

__global__ __launch_bounds__(kBT) void relevel_kernel(const int32_t* __restrict__ nk, uint32_t* level, uint32_t m,
                                                      BhtView t, const uint32_t* L0, const uint32_t* seg_start,
                                                      const uint32_t* seg_end, const uint32_t* sv,
                                                      uint32_t* changed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m; r += stride) {
    uint32_t nl = 3;
    for (int l = 0; l < 3; ++l) {
      const uint32_t b = (uint32_t)bht_bucket(t, l, nk[3 * r], nk[3 * r + 1], nk[3 * r + 2]);
      if (L0[b] + lower_ranks(sv, seg_start, seg_end, b, (uint32_t)r) <= (uint32_t)kBhtThreshold) {
        nl = (uint32_t)l;
        break;
      }
    }
    if (nl != level[r]) {
      level[r] = nl;
      atomicOr(changed, 1u);
    }
  }
}

// every new key to its slot: the rank order within its bucket after the keys
// the bucket held; indices continue the count in rank order
__global__ __launch_bounds__(kBT) void place_kernel(const int32_t* __restrict__ nk, uint32_t m, const uint32_t* sk,
                                                    const uint32_t* sv, const uint32_t* seg_start,
                                                    const uint32_t* L0, const uint32_t* c0, BhtView t,
                                                    uint32_t* err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < m; p += stride) {
    const uint32_t b = sk[p];
    if (b >= (uint32_t)t.nb) {
      atomicOr(err, kErrTableFull);
      *t.success = 0;
      continue;
    }
    const uint32_t r = sv[p];
    const int64_t slot = (int64_t)b * kBhtBucket + L0[b] + (p - seg_start[b]);
    const int32_t idx = (int32_t)(*c0 + r);
    const int32_t k0 = nk[3 * (int64_t)r], k1 = nk[3 * (int64_t)r + 1], k2 = nk[3 * (int64_t)r + 2];
    t.keys[slot] = make_int4(k0, k1, k2, kBhtSentinelKey);
    t.indices[slot] = idx;
    if (idx < t.T) {
      t.active[3 * (int64_t)idx] = k0;
      t.active[3 * (int64_t)idx + 1] = k1;
      t.active[3 * (int64_t)idx + 2] = k2;
    }
    if (idx >= t.T - 20) {
      atomicOr(err, kErrTableFull);
      *t.success = 0;
    }
  }
}

// the sequential insertion itself, one key after another (the host path of
// BHTView::insert): the reference mode of zsmpm_bht_insert__rocm
__global__ void serial_place_kernel(const int32_t* __restrict__ nk, uint32_t m, uint32_t* level, BhtView t,
                                    uint32_t* err) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (uint32_t r = 0; r < m; ++r) {
    const int32_t k0 = nk[3 * (int64_t)r], k1 = nk[3 * (int64_t)r + 1], k2 = nk[3 * (int64_t)r + 2];
    uint32_t lv = 3;
    for (int l = 0; l < 3; ++l) {
      const int64_t off = (int64_t)bht_bucket(t, l, k0, k1, k2) * kBhtBucket;
      int load = 0;
      while (load < kBhtBucket) {
        const int4 s = t.keys[off + load];
        if (s.x == kBhtSentinelKey && s.y == kBhtSentinelKey && s.z == kBhtSentinelKey) break;
        ++load;
      }
      if (load <= kBhtThreshold) {
        const int32_t idx = (*t.cnt)++;
        t.keys[off + load] = make_int4(k0, k1, k2, kBhtSentinelKey);
        t.indices[off + load] = idx;
        if (idx < t.T) {
          t.active[3 * (int64_t)idx] = k0;
          t.active[3 * (int64_t)idx + 1] = k1;
          t.active[3 * (int64_t)idx + 2] = k2;
        }
        if (idx >= t.T - 20) {
          atomicOr(err, kErrTableFull);
          *t.success = 0;
        }
        lv = (uint32_t)l;
        break;
      }
    }
    if (lv == 3) {
      atomicOr(err, kErrTableFull);
      *t.success = 0;
    }
    level[r] = lv;
  }
}

// per input: its new index, -1 for a key already present or repeated, the
// failure token where it could not be placed (or lies within 20 of the size)
__global__ __launch_bounds__(kBT) void out_kernel(int64_t n, const uint32_t* slot_of, const uint32_t* flag,
                                                  const uint32_t* rank, const uint32_t* level, const uint32_t* c0,
                                                  int32_t T, int32_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    int32_t o = -1;
    if (slot_of[i] == kNone) {
      o = kBhtFailure;
    } else if (flag[i]) {
      const uint32_t r = rank[i];
      const int64_t idx = (int64_t)*c0 + r;
      o = (level[r] == 3 || idx >= T - 20) ? kBhtFailure : (int32_t)idx;
    }
    out[i] = o;
  }
}

__global__ void finish_kernel(BhtView t, const uint32_t* c0, const uint32_t* m) { *t.cnt = (int32_t)(*c0 + *m); }

// BHTView::query for many keys: 16 lanes per key compare one bucket's 16
// slots at once (a 256-byte read), as zpc's tile_query does with a
// cooperative-groups tile (Bht.hpp:582-620); four keys per wave.
__global__ __launch_bounds__(kBT) void query_kernel(BhtView t, const int32_t* __restrict__ keys, int64_t n,
                                                    int32_t* out) {
  const int64_t total = n * kBhtBucket;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t q = g >> 4;
    const int j = (int)(g & 15);
    const int32_t k0 = keys[3 * q], k1 = keys[3 * q + 1], k2 = keys[3 * q + 2];
    const int sh = (int)(threadIdx.x & 63) & ~15;  // the group's first lane in the wave
    int32_t res = -1;
    for (int l = 0; l < 3; ++l) {
      const int64_t off = (int64_t)bht_bucket(t, l, k0, k1, k2) * kBhtBucket;
      const int4 s = t.keys[off + j];
      const uint64_t bal = __ballot(s.x == k0 && s.y == k1 && s.z == k2);
      const uint32_t mine = (uint32_t)(bal >> sh) & 0xffffu;
      if (mine) {
        res = t.indices[off + __ffs(mine) - 1];
        break;
      }
    }
    if (j == 0) out[q] = res;
  }
}

// SparseGridView::insert's block key of each world position (:439-446):
// X = floor(w2i x + 1/2), minus X & 7
__global__ __launch_bounds__(kBT) void spg_keys_kernel(SpgView g, const float* __restrict__ x, int64_t n,
                                                       int32_t* keys) {
#pragma clang fp contract(off)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    float X[3];
    spg_apply(g.w2i, x[3 * i], x[3 * i + 1], x[3 * i + 2], X);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int32_t c = (int32_t)floorf(X[d] + 0.5f);
      keys[3 * i + d] = c - (c & 7);
    }
  }
}

__global__ __launch_bounds__(kBT) void spg_sample_kernel(SpgView g, int chn, int width, const float* __restrict__ x,
                                                         int64_t n, float* out) {
#pragma clang fp contract(off)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    float X[3], v[3];
    spg_apply(g.w2i, x[3 * i], x[3 * i + 1], x[3 * i + 2], X);
    spg_isample_n(g, chn, width, X, v);
    for (int w = 0; w < width; ++w) out[width * i + w] = v[w];
  }
}

int64_t pow2_at_least(int64_t v) {
  int64_t p = 1024;
  while (p < v) p <<= 1;
  return p;
}

// scratch layout of one insertion of n keys into a table of nb buckets
struct Carve {
  uint64_t* skey;
  uint32_t *smin, *slot_of, *flag, *rank, *total, *c0, *changed, *level, *bkey, *bval, *bkey2, *bval2, *L0, *seg_start,
      *seg_end, *scan, *radix;
  int32_t* nk;
  int32_t* pkeys;
  int64_t S;
};

size_t carve(void* base, int64_t n, int64_t nb, Carve* c) {
  const int64_t S = pow2_at_least(2 * (n > 0 ? n : 1));
  const int64_t nn = n > 0 ? n : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    off = (off + 255) & ~(size_t)255;
    void* p = base ? (char*)base + off : nullptr;
    off += bytes;
    return p;
  };
  Carve k;
  k.S = S;
  k.skey = (uint64_t*)take(8 * (size_t)S);
  k.smin = (uint32_t*)take(4 * (size_t)S);
  k.slot_of = (uint32_t*)take(4 * (size_t)nn);
  k.flag = (uint32_t*)take(4 * (size_t)nn);
  k.rank = (uint32_t*)take(4 * (size_t)nn);
  k.total = (uint32_t*)take(4);
  k.c0 = (uint32_t*)take(4);
  k.changed = (uint32_t*)take(4);
  k.level = (uint32_t*)take(4 * (size_t)nn);
  k.bkey = (uint32_t*)take(4 * (size_t)nn);
  k.bval = (uint32_t*)take(4 * (size_t)nn);
  k.bkey2 = (uint32_t*)take(4 * (size_t)nn);
  k.bval2 = (uint32_t*)take(4 * (size_t)nn);
  k.L0 = (uint32_t*)take(4 * (size_t)(nb + 1));
  k.seg_start = (uint32_t*)take(4 * (size_t)(nb + 1));
  k.seg_end = (uint32_t*)take(4 * (size_t)(nb + 1));
  k.scan = (uint32_t*)take(4 * scan_scratch_words());
  k.radix = (uint32_t*)take(4 * radix_sort_scratch_words((uint32_t)nn));
  k.nk = (int32_t*)take(12 * (size_t)nn);
  k.pkeys = (int32_t*)take(12 * (size_t)nn);
  if (c) *c = k;
  return off + 256;
}

int32_t hip_code(hipError_t e) { return e == hipSuccess ? ZSMPM_OK : ZSMPM_DEVICE_ERROR; }

}  // namespace

size_t bht_scratch_bytes(int64_t n, int32_t nb) { return carve(nullptr, n, nb, nullptr); }

hipError_t bht_clear(const BhtView& t, hipStream_t s) {
  bht_clear_kernel<<<groups_n(t.T), kBT, 0, s>>>(t);
  return hipGetLastError();
}

int32_t bht_insert(const BhtView& t, const int32_t* keys, int64_t n, int32_t* out, bool serial, void* scratch,
                   uint32_t* err, hipStream_t s, int32_t* rounds) {
  if (rounds) *rounds = 0;
  if (n == 0) return ZSMPM_OK;
  if (n > (int64_t)0x7fffffff) return ZSMPM_CAPACITY;
  Carve c;
  carve(scratch, n, t.nb, &c);
  hipError_t e = hipSuccess;
  dedup_clear_kernel<<<groups_n(c.S), kBT, 0, s>>>(c.skey, c.smin, c.S);
  dedup_insert_kernel<<<groups_n(n), kBT, 0, s>>>(keys, n, c.skey, c.smin, c.S, c.slot_of, err);
  first_kernel<<<groups_n(n), kBT, 0, s>>>(keys, n, c.slot_of, c.smin, t, c.flag);
  e = exclusive_scan_u32(c.flag, c.rank, nullptr, (uint32_t)n, c.scan, c.total, s);
  if (e != hipSuccess) return ZSMPM_DEVICE_ERROR;
  compact_kernel<<<groups_n(n), kBT, 0, s>>>(keys, n, c.flag, c.rank, c.nk, c.level);
  loads_kernel<<<groups_n(t.nb), kBT, 0, s>>>(t, c.L0, c.c0);
  uint32_t m = 0;
  if ((e = hipMemcpyAsync(&m, c.total, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return ZSMPM_DEVICE_ERROR;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return ZSMPM_DEVICE_ERROR;
  bool done = m == 0;
  if (!done && serial) {
    serial_place_kernel<<<1, 1, 0, s>>>(c.nk, m, c.level, t, err);
    if (out) out_kernel<<<groups_n(n), kBT, 0, s>>>(n, c.slot_of, c.flag, c.rank, c.level, c.c0, t.T, out);
    return hip_code(hipGetLastError());
  }
  const int bits = [&] {
    int b = 1;
    while (((int64_t)1 << b) <= t.nb) ++b;
    return b;
  }();
  uint32_t *sk = nullptr, *sv = nullptr;
  for (int round = 0; !done && round < kMaxRounds; ++round) {
    assign_kernel<<<groups_n(m), kBT, 0, s>>>(c.nk, c.level, m, t, c.bkey, c.bval);
    bool alt = false;
    e = radix_sort_pairs(c.bkey, c.bval, c.bkey2, c.bval2, nullptr, m, bits, c.radix, s, &alt);
    if (e != hipSuccess) return ZSMPM_DEVICE_ERROR;
    sk = alt ? c.bkey2 : c.bkey;
    sv = alt ? c.bval2 : c.bval;
    seg_clear_kernel<<<groups_n(t.nb + 1), kBT, 0, s>>>(c.seg_start, c.seg_end, (int64_t)t.nb + 1, c.changed);
    seg_kernel<<<groups_n(m), kBT, 0, s>>>(sk, m, c.seg_start, c.seg_end);
    relevel_kernel<<<groups_n(m), kBT, 0, s>>>(c.nk, c.level, m, t, c.L0, c.seg_start, c.seg_end, sv, c.changed);
    uint32_t changed = 0;
    if (hipMemcpyAsync(&changed, c.changed, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
    if (hipStreamSynchronize(s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
    if (rounds) *rounds = round + 1;
    done = changed == 0;
  }
  if (!done) {
    // an overflow chain longer than the rounds: the sequential pass decides
    serial_place_kernel<<<1, 1, 0, s>>>(c.nk, m, c.level, t, err);
    if (out) out_kernel<<<groups_n(n), kBT, 0, s>>>(n, c.slot_of, c.flag, c.rank, c.level, c.c0, t.T, out);
    return hip_code(hipGetLastError());
  }
  if (m > 0) place_kernel<<<groups_n(m), kBT, 0, s>>>(c.nk, m, sk, sv, c.seg_start, c.L0, c.c0, t, err);
  if (out) out_kernel<<<groups_n(n), kBT, 0, s>>>(n, c.slot_of, c.flag, c.rank, c.level, c.c0, t.T, out);
  if (m > 0) finish_kernel<<<1, 1, 0, s>>>(t, c.c0, c.total);
  return hip_code(hipGetLastError());
}

hipError_t bht_query_many(const BhtView& t, const int32_t* keys, int64_t n, int32_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  query_kernel<<<groups_n(n * kBhtBucket), kBT, 0, s>>>(t, keys, n, out);
  return hipGetLastError();
}

int32_t spg_activate(const SpgView& g, const float* x, int64_t n, void* scratch, uint32_t* err, hipStream_t s) {
  if (n == 0) return ZSMPM_OK;
  Carve c;
  carve(scratch, n, g.t.nb, &c);
  spg_keys_kernel<<<groups_n(n), kBT, 0, s>>>(g, x, n, c.pkeys);
  if (hipGetLastError() != hipSuccess) return ZSMPM_DEVICE_ERROR;
  return bht_insert(g.t, c.pkeys, n, nullptr, false, scratch, err, s, nullptr);
}

hipError_t spg_sample(const SpgView& g, int chn, int width, const float* x, int64_t n, float* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  spg_sample_kernel<<<groups_n(n), kBT, 0, s>>>(g, chn, width, x, n, out);
  return hipGetLastError();
}

}  // namespace zsmpm
