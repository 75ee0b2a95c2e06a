// Block activation, neighbour enlargement and the particle -> block bucket
// sort on zpc's HashTable<i32,3,int> layout, for gfx950.
//
// zpc inserts keys with per-slot spin locks and an atomic block counter
// (container/HashTable.hpp:356-398), so under its parallel policies the block
// numbers and slot layout depend on timing. Here they are deterministic and
// equal to zpc's sequential semantics:
//   * numbering: a key gets the next block number in order of its first
//     occurrence (particle index for activation, (block, dx, dy, dz) loop
//     order for enlargement) -- found with a dedup hash holding the minimum
//     order per key; each new key's order sets one bit of a bitmap, and an
//     exclusive popcount scan of the bitmap ranks the new keys;
//   * slots: keys are placed by priority linear probing on the reference
//     probe sequence (home = (h % T + T) % T, step 127) with priority = block
//     number; the result equals inserting serially in block-number order
//     (Shun & Blelloch, "Phase-Concurrent Hash Tables for Determinism",
//     SPAA 2014), i.e. the layout zpc's seq_exec path produces.
// The bucket sort is stable by particle index (zpc's seq semantics of
// SpatiallyCount/SpatiallyDistribute): particles are cut into runs of equal
// bucket, runs are radix-sorted by bucket (stable) and every particle lands at
// its run's destination plus its offset in the run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zsmpm_device.hpp"
#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kThreads = 256;
constexpr int kTileItems = 16;
constexpr int kTile = kThreads * kTileItems;  // 4096 particles per bucket-sort tile

inline int groups_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 65535 * 8) g = 65535 * 8;
  return (int)g;
}

// --------------------------------------------------------------- clearing
__global__ void clear_table_kernel(TableView t, uint64_t* skey, uint32_t* sorder, uint32_t* dmask, uint32_t ts,
                                   uint32_t* counters, uint32_t* zero, int64_t nzero) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)((ts + 31u) >> 5); e += stride) dmask[e] = 0u;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nzero; e += stride) zero[e] = 0u;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < t.T; e += stride) {
    t.keys[3 * e] = kKeySentinel;
    t.keys[3 * e + 1] = kKeySentinel;
    t.keys[3 * e + 2] = kKeySentinel;
    t.indices[e] = -1;
    t.status[e] = -1;
  }
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ts; e += stride) {
    skey[e] = kEmpty64;
    sorder[e] = 0xffffffffu;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *t.cnt = 0;
    counters[kCtrList] = 0;
  }
}

__global__ void clear_scratch_kernel(uint64_t* skey, uint32_t* sorder, uint32_t* dmask, uint32_t ts,
                                     const int32_t* cnt, uint32_t* counters, int32_t* slot_block,
                                     uint32_t* bitmap, int64_t bm_n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)((ts + 31u) >> 5); e += stride) dmask[e] = 0u;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < bm_n; e += stride) bitmap[e] = 0u;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ts; e += stride) {
    skey[e] = kEmpty64;
    sorder[e] = 0xffffffffu;
    if (slot_block) slot_block[e] = -1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counters[kCtrList] = 0;
    counters[kCtrCnt0] = (uint32_t)*cnt;
  }
}

// ----------------------------------------------------------- dedup inserts
// The dedup scratch of one insertion pass: packed keys and minimum orders per
// slot, and one bit per slot the pass claimed, so the pass's keys are found
// by reading ts/32 words instead of all ts slots. (A list of claimed slots
// behind one atomic counter was slower: tens of thousands of same-address
// atomics serialize.)
struct Dedup {
  uint64_t* skey;
  uint32_t* sorder;
  uint32_t ts;
  uint32_t* err;
  uint32_t* dmask;  // [ts / 32]
};

__device__ __forceinline__ uint32_t dedup_home(uint64_t pk, uint32_t ts) { return (uint32_t)mix64(pk) & (ts - 1u); }

// Finishes an insert whose probe at slot s returned prev: probes on until the
// slot is free or holds pk (ts probes in all), then lowers the slot's order;
// the insert that claims a free slot marks it. Returns the slot holding pk, or
// ~0u when the scratch table is full.
__device__ __forceinline__ uint32_t dedup_finish(uint64_t pk, uint32_t order, uint32_t s, uint64_t prev,
                                                 const Dedup& d) {
  for (uint32_t probe = 1; prev != kEmpty64 && prev != pk; ++probe) {
    if (probe >= d.ts) {
      atomicOr(d.err, kErrTableFull);
      return ~0u;
    }
    s = (s + 1u) & (d.ts - 1u);
    prev = atomicCAS((unsigned long long*)&d.skey[s], (unsigned long long)kEmpty64, (unsigned long long)pk);
  }
  atomicMin(&d.sorder[s], order);
  if (prev == kEmpty64) atomicOr(&d.dmask[s >> 5], 1u << (s & 31u));
  return s;
}

__device__ __forceinline__ uint32_t dedup_insert(uint64_t pk, uint32_t order, const Dedup& d) {
  const uint32_t s = dedup_home(pk, d.ts);
  const uint64_t prev =
      atomicCAS((unsigned long long*)&d.skey[s], (unsigned long long)kEmpty64, (unsigned long long)pk);
  return dedup_finish(pk, order, s, prev, d);
}

constexpr int kActItems = 4;  // 64-particle groups per wave, their loads and first probes issued together

// ComputeSparsity candidates: one per particle, order = particle index. Each
// distinct key of a 64-particle group is inserted once, by its lowest lane
// (the smallest particle index holding it), so hot keys see one atomic per
// group. A wave takes kActItems groups and issues their x loads, then their
// first probes, together: the kernel is bound by those round trips, not by
// bytes. The dedup hash keeps the minimum order per key, so the result does
// not depend on the order of the inserts. With pslot, every particle's
// scratch slot (the identity of its key) is recorded for the bucket sort.
// kMulti (a batch, ss): particle p takes its scene's 1/dx, and its key is
// shifted by the scene's offset (a key outside the scene's own range
// [-512, 511) along axis 0 is a key-range error)
template <bool kMulti>
__global__ __launch_bounds__(kThreads) void activate_kernel(const float* __restrict__ x, int64_t n,
                                                            float dxinv, int blockLen, int offset,
                                                            float disp, Dedup dd, uint32_t* pslot,
                                                            SceneSet ss) {
  uint32_t* err = dd.err;
  const int64_t span = (int64_t)kThreads * kActItems;
  const int64_t stride = (int64_t)gridDim.x * span;
  const int lane = lane_id();
  __shared__ int64_t sfirst[kMulti ? kMaxScenes + 1 : 1];
  if constexpr (kMulti) {
    for (int i = threadIdx.x; i <= ss.n; i += kThreads) sfirst[i] = ss.first[i];
    __syncthreads();
  }
  for (int64_t base = (int64_t)blockIdx.x * span; base < n; base += stride) {
    float cx[kActItems], cy[kActItems], cz[kActItems];
#pragma unroll
    for (int r = 0; r < kActItems; ++r) {
      const int64_t p = base + (int64_t)r * kThreads + threadIdx.x;
      cx[r] = cy[r] = cz[r] = 0.f;
      if (p < n) {
        cx[r] = x[3 * p];
        cy[r] = x[3 * p + 1];
        cz[r] = x[3 * p + 2];
      }
    }
    uint64_t pk[kActItems];
    int leader[kActItems];
    bool head[kActItems];
#pragma unroll
    for (int r = 0; r < kActItems; ++r) {
      const int64_t p = base + (int64_t)r * kThreads + threadIdx.x;
      pk[r] = kEmpty64;
      if (p < n) {
        int32_t k0, k1, k2;
        bool in_range = true;
        if constexpr (kMulti) {
          const int32_t sc = scene_of_particle(sfirst, ss.n, p);
          block_key(cx[r], cy[r], cz[r], ss.p[sc].dx_inv, blockLen, offset, disp, k0, k1, k2);
          // below 511, so the enlargement's key + 1 stays in the scene's range
          in_range = k0 >= -kSceneKeyHalf && k0 < kSceneKeyHalf - 1;
          k0 += sc * kSceneKeyStride;
        } else {
          block_key(cx[r], cy[r], cz[r], dxinv, blockLen, offset, disp, k0, k1, k2);
        }
        if (in_range && key_packable(k0, k1, k2)) {
          pk[r] = pack_key(k0, k1, k2);
        } else {
          atomicOr(err, kErrKeyRange);
        }
      }
      // peel off one distinct key per iteration: the lowest pending lane leads
      bool pending = pk[r] != kEmpty64;
      bool h = false;
      int my_leader = lane;
      for (uint64_t m = ballot(pending); m != 0; m = ballot(pending)) {
        const int ld = __ffsll((long long)m) - 1;
        const uint64_t lk = shfl(pk[r], ld);
        const bool same = pending && pk[r] == lk;
        my_leader = same ? ld : my_leader;
        h = h || lane == ld;
        pending = pending && !same;
      }
      head[r] = h;
      leader[r] = my_leader;
    }
    uint32_t slot[kActItems];
    uint64_t prev[kActItems];
#pragma unroll
    for (int r = 0; r < kActItems; ++r) {
      slot[r] = ~0u;
      prev[r] = kEmpty64;
      if (head[r]) {
        slot[r] = dedup_home(pk[r], dd.ts);
        prev[r] = atomicCAS((unsigned long long*)&dd.skey[slot[r]], (unsigned long long)kEmpty64,
                            (unsigned long long)pk[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < kActItems; ++r) {
      const int64_t p = base + (int64_t)r * kThreads + threadIdx.x;
      if (head[r]) slot[r] = dedup_finish(pk[r], (uint32_t)p, slot[r], prev[r], dd);
    }
    if (pslot) {
#pragma unroll
      for (int r = 0; r < kActItems; ++r) {
        const int64_t p = base + (int64_t)r * kThreads + threadIdx.x;
        const uint32_t ls = (uint32_t)__shfl((int)slot[r], leader[r], kWave);
        if (p < n) pslot[p] = pk[r] != kEmpty64 ? ls : ~0u;
      }
    }
  }
}

// EnlargeSparsity candidates: key_i + (dx, dy, dz) for i < n0 in loop order.
__global__ __launch_bounds__(kThreads) void enlarge_kernel(const int32_t* __restrict__ active,
                                                           const uint32_t* counters, int32_t lo0, int32_t lo1,
                                                           int32_t lo2, int32_t e0, int32_t e1, int32_t e2,
                                                           Dedup dd) {
  uint32_t* err = dd.err;
  const int64_t n0 = counters[kCtrCnt0];
  const int64_t nd = (int64_t)e0 * e1 * e2;
  const int64_t total = n0 * nd;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < total; c += stride) {
    const int64_t i = c / nd;
    const int32_t j = (int32_t)(c - i * nd);
    const int32_t dz = j % e2, dy = (j / e2) % e1, dx = j / (e1 * e2);
    const int32_t k0 = active[3 * i] + lo0 + dx, k1 = active[3 * i + 1] + lo1 + dy,
                  k2 = active[3 * i + 2] + lo2 + dz;
    if (!key_packable(k0, k1, k2)) {
      atomicOr(err, kErrKeyRange);
      continue;
    }
    (void)dedup_insert(pack_key(k0, k1, k2), (uint32_t)c, dd);
  }
}

constexpr int kCollectGroups = 1024;
constexpr int kCollectBuf = 2048;

__device__ __forceinline__ void list_put(uint32_t slot, uint32_t key, uint32_t val, uint32_t* lkey,
                                         uint32_t* lval, uint32_t list_cap, uint32_t* err) {
  if (slot < list_cap) {
    lkey[slot] = key;
    lval[slot] = val;
  } else {
    atomicOr(err, kErrCapacity);
  }
}

// Gather the dedup entries whose key is not in the table yet: (order, slot),
// and mark each one's order in the first-occurrence bitmap. Only the slots
// the pass claimed (set bits of dmask) are visited, not all ts. Each
// workgroup takes one contiguous range of dmask words, collects its entries in
// LDS and reserves list space with a single global atomic (entries beyond the
// LDS buffer go straight to the list). The list order is irrelevant: ranks
// come from the bitmap.
__global__ __launch_bounds__(kThreads) void collect_new_kernel(const uint64_t* __restrict__ skey,
                                                               const uint32_t* __restrict__ sorder,
                                                               const uint32_t* __restrict__ dmask,
                                                               uint32_t ts, TableView t, uint32_t* counters,
                                                               uint32_t* lkey, uint32_t* lval,
                                                               uint32_t list_cap, uint32_t* err,
                                                               uint32_t* bitmap, uint32_t bm_words) {
  __shared__ uint32_t bkey[kCollectBuf], bval[kCollectBuf];
  __shared__ uint32_t bcount, gbase;
  const bool table_empty = counters[kCtrCnt0] == 0;
  const int64_t nw = (int64_t)((ts + 31u) >> 5);
  const int64_t per = (nw + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = min(nw, lo + per);
  if (threadIdx.x == 0) bcount = 0;
  __syncthreads();
  const int lane = lane_id();
  for (int64_t base = lo; base < hi; base += blockDim.x) {  // uniform across the workgroup
    const int64_t j = base + threadIdx.x;
    uint32_t bits = j < hi ? dmask[j] : 0u;  // slots beyond ts are never marked
    while (ballot(bits != 0u) != 0) {  // one claimed slot per lane per step (wave-uniform loop)
      bool take = false;
      uint32_t s = 0;
      if (bits != 0u) {
        s = (uint32_t)(j * 32) + (uint32_t)(__ffs(bits) - 1);
        bits &= bits - 1u;
        const uint64_t pk = skey[s];
        if (pk != kEmpty64) {
          take = true;
          if (!table_empty) {
            int32_t k0, k1, k2;
            unpack_key(pk, k0, k1, k2);
            take = table_find(t, k0, k1, k2) < 0;
          }
        }
      }
      const uint64_t m = ballot(take);
      if (m == 0) continue;
      const int leader = __ffsll((long long)m) - 1;
      uint32_t wbase = 0;
      if (lane == leader) wbase = atomicAdd(&bcount, (uint32_t)__popcll(m));
      wbase = shfl(wbase, leader);
      if (take) {
        // mark the key's order: its rank among the new keys is the number of
        // marked orders below it (orders of distinct keys are distinct)
        const uint32_t o = sorder[s];
        if ((o >> 5) < bm_words) atomicOr(&bitmap[o >> 5], 1u << (o & 31u));
        const uint32_t slot = wbase + (uint32_t)__popcll(m & lanemask_lt());
        if (slot < (uint32_t)kCollectBuf) {
          bkey[slot] = sorder[s];
          bval[slot] = (uint32_t)s;
        } else {
          list_put(atomicAdd(&counters[kCtrList], 1u), sorder[s], (uint32_t)s, lkey, lval, list_cap, err);
        }
      }
    }
  }
  __syncthreads();
  const uint32_t nloc = min(bcount, (uint32_t)kCollectBuf);
  if (threadIdx.x == 0) gbase = nloc ? atomicAdd(&counters[kCtrList], nloc) : 0u;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nloc; i += blockDim.x)
    list_put(gbase + i, bkey[i], bval[i], lkey, lval, list_cap, err);
}

// Number the new keys cnt0 + rank, record _activeKeys and place them by
// priority linear probing on indices[] (lower block number wins a slot).
__global__ __launch_bounds__(kThreads) void assign_kernel(const uint32_t* __restrict__ lorder,
                                                          const uint32_t* __restrict__ lslot,
                                                          const uint32_t* __restrict__ bitmap,
                                                          const uint32_t* __restrict__ bm_prefix, uint32_t bm_words,
                                                          const uint64_t* __restrict__ skey,
                                                          TableView t, const uint32_t* counters,
                                                          uint32_t list_cap, uint32_t* err, int32_t* slot_block) {
  const uint32_t nl = min(counters[kCtrList], list_cap);
  const uint32_t cnt0 = counters[kCtrCnt0];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nl; i += stride) {
    const uint32_t sl = lslot[i];
    const uint32_t o = lorder[i];
    if ((o >> 5) >= bm_words) {
      atomicOr(err, kErrCapacity);
      continue;
    }
    // first-occurrence rank: new keys with a smaller order
    const uint32_t below = bm_prefix[o >> 5] + (uint32_t)__popc(bitmap[o >> 5] & ((1u << (o & 31u)) - 1u));
    const uint64_t pk = skey[sl];
    int32_t k0, k1, k2;
    unpack_key(pk, k0, k1, k2);
    const uint32_t rank = cnt0 + below;
    if (rank >= (uint32_t)t.T) {
      atomicOr(err, kErrTableFull);
      continue;
    }
    if (slot_block) slot_block[sl] = (int32_t)rank;  // the block number of the key in scratch slot sl
    t.active[3 * (int64_t)rank] = k0;
    t.active[3 * (int64_t)rank + 1] = k1;
    t.active[3 * (int64_t)rank + 2] = k2;
    // carried element: starts as this key; a displaced lower-priority entry
    // continues from the slot after the one it lost
    uint32_t cur = rank;
    int32_t s = zpc_home(k0, k1, k2, t.T);
    int64_t steps = 0;
    const int64_t limit = 2ll * t.T + 64;
    while (true) {
      const uint32_t old = atomicMin((uint32_t*)&t.indices[s], cur);
      if (old == 0xffffffffu) break;
      if (old > cur) cur = old;
      s = probe_next(s, t.T);
      if (++steps > limit) {
        atomicOr(err, kErrTableFull);
        break;
      }
    }
  }
}

// Write keys[] of every newly placed entry: block r probes from its home
// until it finds the slot holding r.
__global__ __launch_bounds__(kThreads) void fill_keys_kernel(TableView t, const uint32_t* counters,
                                                             uint32_t list_cap, uint32_t* err) {
  const uint32_t nl = min(counters[kCtrList], list_cap);
  const uint32_t cnt0 = counters[kCtrCnt0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // nothing in this kernel reads the count
    const int64_t c = (int64_t)cnt0 + nl;
    *t.cnt = (int32_t)(c < t.T ? c : t.T);
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nl; i += stride) {
    const int32_t r = (int32_t)(cnt0 + (uint32_t)i);
    if (r >= t.T) continue;
    const int32_t k0 = t.active[3 * (int64_t)r], k1 = t.active[3 * (int64_t)r + 1],
                  k2 = t.active[3 * (int64_t)r + 2];
    int32_t s = zpc_home(k0, k1, k2, t.T);
    int32_t probe = 0;
    for (; probe < t.T; ++probe) {
      if (t.indices[s] == r) break;
      s = probe_next(s, t.T);
    }
    if (probe == t.T) {
      atomicOr(err, kErrTableFull);
      continue;
    }
    t.keys[3 * (int64_t)s] = k0;
    t.keys[3 * (int64_t)s + 1] = k1;
    t.keys[3 * (int64_t)s + 2] = k2;
  }
}

__global__ void query_kernel(TableView t, const int32_t* __restrict__ keys, int64_t n, int32_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = table_find(t, keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]);
}

// ------------------------------------------------------------ bucket sort
__global__ void zero_u32_kernel(uint32_t* p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

constexpr uint32_t kHeadBit = 0x80000000u;  // bkt[p] bit 31: p starts a run

// Bucket of every particle (the block number its key maps to), per-bucket
// counts, run heads and the number of run heads per tile. A run head is a
// particle whose key differs from the previous particle's; bkt[p] carries the
// head flag in bit 31 so every later pass sees the same runs. Lanes query the
// table only at run heads (and lane 0), the rest take the head's bucket.
// Particles of a tile are visited in order: round r, thread t -> tile*kTile +
// r*kThreads + t.
__global__ __launch_bounds__(kThreads) void bucket_kernel(const float* __restrict__ x, int64_t n, float dxinv,
                                                          int blockLen, int offset, float disp, TableView t,
                                                          uint32_t* __restrict__ bkt, uint32_t* counts,
                                                          int32_t cap_blocks, uint32_t* __restrict__ tile_heads,
                                                          uint32_t* err) {
  __shared__ uint32_t wsum[kThreads / kWave];
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int lane = lane_id();
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t heads = 0;
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      const bool valid = p < n;
      int32_t k0 = 0, k1 = 0, k2 = 0;
      if (valid) block_key(x[3 * p], x[3 * p + 1], x[3 * p + 2], dxinv, blockLen, offset, disp, k0, k1, k2);
      int32_t q0 = __shfl_up(k0, 1, kWave), q1 = __shfl_up(k1, 1, kWave), q2 = __shfl_up(k2, 1, kWave);
      if (lane == 0 && valid && p > 0)
        block_key(x[3 * (p - 1)], x[3 * (p - 1) + 1], x[3 * (p - 1) + 2], dxinv, blockLen, offset, disp, q0,
                  q1, q2);
      const bool head = valid && (p == 0 || q0 != k0 || q1 != k1 || q2 != k2);
      const bool lookup = valid && (head || lane == 0);
      int32_t b = -1;
      if (lookup) {
        b = table_find(t, k0, k1, k2);
        if (b < 0) atomicOr(err, kErrMissingBlock);
        if (b >= cap_blocks) {
          atomicOr(err, kErrCapacity);
          b = -1;
        }
      }
      const uint64_t lm = ballot(lookup);
      const uint64_t le = lanemask_lt() | (1ull << lane);
      const int src = 63 - __clzll(lm & le);
      b = shfl(b, src < 0 ? 0 : src);
      if (valid) bkt[p] = (uint32_t)(b < 0 ? 0 : b) | (head ? kHeadBit : 0u);
      heads += head ? 1u : 0u;
      // per-bucket counts: one atomic per (lookup segment) in the wave
      const uint64_t vm = ballot(valid);
      if (lookup && b >= 0) {
        const uint64_t above = lm & ~le;
        const int next = above ? (__ffsll((long long)above) - 1) : 64;
        const uint64_t seg = (next >= 64 ? ~0ull : ((1ull << next) - 1ull)) & ~((1ull << lane) - 1ull);
        atomicAdd(&counts[b], (uint32_t)__popcll(seg & vm));
      }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) heads += __shfl_xor(heads, o, kWave);
    if (lane == 0) wsum[threadIdx.x / kWave] = heads;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int i = 0; i < kThreads / kWave; ++i) tot += wsum[i];
      tile_heads[tile] = tot;
    }
    __syncthreads();
  }
}

// The same outputs as bucket_kernel from the scratch slots recorded by the
// activation that built this table (activate_kernel with pslot): a particle's
// slot identifies its key, and slot_block maps it to the block number, so
// neither x nor the table is read again.
__global__ __launch_bounds__(kThreads) void bucket_slots_kernel(const uint32_t* __restrict__ pslot, int64_t n,
                                                                const int32_t* __restrict__ slot_block, uint32_t ts,
                                                                uint32_t* __restrict__ bkt, uint32_t* counts,
                                                                int32_t cap_blocks, uint32_t* __restrict__ tile_heads,
                                                                uint32_t* err) {
  __shared__ uint32_t wsum[kThreads / kWave];
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int lane = lane_id();
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t heads = 0;
    // all rounds' slots (and, for lane 0, the slot before the wave) load
    // together, then all run heads' block lookups: two memory round trips per
    // tile instead of two per round
    uint32_t sl[kTileItems], pv[kTileItems];
    int32_t bl[kTileItems];
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      sl[r] = p < n ? pslot[p] : ~0u;
      pv[r] = (lane == 0 && p < n && p > 0) ? pslot[p - 1] : ~0u;
    }
    uint32_t headm = 0;
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      const bool valid = p < n;
      uint32_t prev = __shfl_up(sl[r], 1, kWave);
      if (lane == 0) prev = pv[r];
      const bool head = valid && (p == 0 || sl[r] != prev);
      headm |= head ? (1u << r) : 0u;
      const bool lookup = valid && (head || lane == 0);
      bl[r] = lookup ? (sl[r] < ts ? slot_block[sl[r]] : -1) : -1;
    }
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      const bool valid = p < n;
      const bool head = (headm >> r) & 1u;
      const bool lookup = valid && (head || lane == 0);
      int32_t b = -1;
      if (lookup) {
        b = bl[r];
        if (b < 0) atomicOr(err, kErrMissingBlock);
        if (b >= cap_blocks) {
          atomicOr(err, kErrCapacity);
          b = -1;
        }
      }
      const uint64_t lm = ballot(lookup);
      const uint64_t le = lanemask_lt() | (1ull << lane);
      const int src = 63 - __clzll(lm & le);
      b = shfl(b, src < 0 ? 0 : src);
      if (valid) bkt[p] = (uint32_t)(b < 0 ? 0 : b) | (head ? kHeadBit : 0u);
      heads += head ? 1u : 0u;
      const uint64_t vm = ballot(valid);
      if (lookup && b >= 0) {
        const uint64_t above = lm & ~le;
        const int next = above ? (__ffsll((long long)above) - 1) : 64;
        const uint64_t seg = (next >= 64 ? ~0ull : ((1ull << next) - 1ull)) & ~((1ull << lane) - 1ull);
        atomicAdd(&counts[b], (uint32_t)__popcll(seg & vm));
      }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) heads += __shfl_xor(heads, o, kWave);
    if (lane == 0) wsum[threadIdx.x / kWave] = heads;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int i = 0; i < kThreads / kWave; ++i) tot += wsum[i];
      tile_heads[tile] = tot;
    }
    __syncthreads();
  }
}

// Run ids of one round of a tile (round r, thread t -> particle tile*kTile +
// r*kThreads + t): `carry` holds the heads before the round; returns the id of
// the run containing this thread's particle (the number of heads at or before
// it, minus one) and adds the round's heads to carry. Workgroup-uniform call.
__device__ __forceinline__ uint32_t round_run_id(bool head, uint32_t& carry, uint32_t* lds) {
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const uint64_t hm = ballot(head);
  if (lane == 0) lds[wave] = (uint32_t)__popcll(hm);
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    const uint32_t c = lds[w];
    before += w < wave ? c : 0u;
    total += c;
  }
  __syncthreads();
  const uint64_t le = lanemask_lt() | (1ull << lane);
  const uint32_t rid = carry + before + (uint32_t)__popcll(hm & le) - 1u;
  carry += total;
  return rid;
}

__global__ __launch_bounds__(kThreads) void run_emit_kernel(const uint32_t* __restrict__ bkt, int64_t n,
                                                            const uint32_t* __restrict__ tile_base,
                                                            uint32_t* run_start, uint32_t* rkey, uint32_t* rval) {
  __shared__ uint32_t lds[kThreads / kWave];
  const int64_t ntiles = (n + kTile - 1) / kTile;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t carry = tile_base[tile];
    uint32_t v[kTileItems];  // all rounds load before the first barrier
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      v[r] = p < n ? bkt[p] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      const bool head = p < n && (v[r] & kHeadBit);
      const uint32_t rid = round_run_id(head, carry, lds);
      if (head && rid < (uint32_t)n) {
        run_start[rid] = (uint32_t)p;
        rkey[rid] = v[r] & ~kHeadBit;
        rval[rid] = rid;
      }
    }
  }
}

// Lengths of the runs in sorted order.
__global__ void run_len_kernel(const uint32_t* __restrict__ sorted_ids, const uint32_t* __restrict__ run_start,
                               const uint32_t* counters, int64_t n, uint32_t* run_len) {
  const uint32_t R = min(counters[kCtrRuns], (uint32_t)n);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < R; i += stride) {
    const uint32_t id = sorted_ids[i];
    if (id >= R) {  // cannot happen when the run passes agree; never index with it
      run_len[i] = 0;
      continue;
    }
    const uint32_t end = (id + 1 < R) ? run_start[id + 1] : (uint32_t)n;
    const uint32_t st = run_start[id];
    run_len[i] = (end >= st && end <= (uint32_t)n) ? end - st : 0u;
  }
}

__global__ void run_dest_kernel(const uint32_t* __restrict__ sorted_ids, const uint32_t* __restrict__ dest_sorted,
                                const uint32_t* counters, int64_t n, uint32_t* run_dest) {
  const uint32_t R = min(counters[kCtrRuns], (uint32_t)n);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < R; i += stride) {
    const uint32_t id = sorted_ids[i];
    if (id < R) run_dest[id] = dest_sorted[i];
  }
}

__global__ __launch_bounds__(kThreads) void particle_scatter_kernel(const uint32_t* __restrict__ bkt, int64_t n,
                                                                    const uint32_t* __restrict__ tile_base,
                                                                    const uint32_t* __restrict__ run_start,
                                                                    const uint32_t* __restrict__ run_dest,
                                                                    const uint32_t* counters,
                                                                    int32_t* __restrict__ indices) {
  __shared__ uint32_t lds[kThreads / kWave];
  const uint32_t R = min(counters[kCtrRuns], (uint32_t)n);
  const int64_t ntiles = (n + kTile - 1) / kTile;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t carry = tile_base[tile];
    // three phases, each with its loads in flight together: the tile's
    // bucket words, the run ids (LDS scans only), the runs' destinations
    uint32_t w[kTileItems], rid[kTileItems];
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      w[r] = p < n ? bkt[p] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      rid[r] = round_run_id(p < n && (w[r] & kHeadBit), carry, lds);
    }
    uint32_t dest[kTileItems], start[kTileItems];
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      const bool ok = p < n && rid[r] < R;
      dest[r] = ok ? run_dest[rid[r]] : 0u;
      start[r] = ok ? run_start[rid[r]] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
      const int64_t p = tile * kTile + (int64_t)r * kThreads + threadIdx.x;
      if (p < n && rid[r] < R) {
        const uint64_t dst = (uint64_t)dest[r] + ((uint32_t)p - start[r]);
        if (dst < (uint64_t)n) indices[dst] = (int32_t)p;
      }
    }
  }
}

// Neighbour blocks of every active block: nbr[8 b + d] = index of
// activeKeys[b] + (d>>2&1, d>>1&1, d&1), -1 where absent.
__global__ void neighbours_kernel(TableView t, int32_t* __restrict__ nbr, uint32_t cap) {
  const int64_t total = 8ll * (int64_t)min((uint32_t)*t.cnt, cap);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += stride) {
    const int64_t b = i >> 3;
    const int d = (int)(i & 7);
    nbr[i] = table_find(t, t.active[3 * b] + ((d >> 2) & 1), t.active[3 * b + 1] + ((d >> 1) & 1),
                        t.active[3 * b + 2] + (d & 1));
  }
}

__global__ void set_counter_kernel(uint32_t* counters, int slot, const uint32_t* src) {
  counters[slot] = *src;
}

int bits_for(uint64_t maxval) {
  int b = 0;
  while (b < 32 && (maxval >> b) != 0) ++b;
  return b;
}

Dedup dedup_of(const Workspace& ws) { return Dedup{ws.skey, ws.sorder, ws.ts_cap, ws.err, ws.dmask}; }

// words of the first-occurrence bitmap for orders below order_bound (orders
// beyond the bitmap are reported by assign)
uint64_t bitmap_words(const Workspace& ws, uint64_t order_bound) {
  const uint64_t words = order_bound / 32 + 1;
  return words > ws.bm_words ? ws.bm_words : words;
}

// dedup entries -> ranked, placed and keyed table entries (shared tail of
// activation and enlargement); the bitmap was cleared with the dedup scratch
hipError_t finish_insertion(const TableView& t, Workspace& ws, uint64_t order_bound, hipStream_t s,
                            int32_t* slot_block = nullptr) {
  const uint64_t words = bitmap_words(ws, order_bound);
  collect_new_kernel<<<kCollectGroups, kThreads, 0, s>>>(ws.skey, ws.sorder, ws.dmask, ws.ts_cap, t, ws.counters,
                                                         ws.la_key,
                                                         ws.la_val, ws.list_cap, ws.err, ws.bitmap, (uint32_t)words);
  hipError_t e = exclusive_scan_popc_u32(ws.bitmap, ws.bm_prefix, (uint32_t)words, ws.scan_tmp, s);
  if (e != hipSuccess) return e;
  const int gl = groups_for(ws.list_cap, kThreads);
  assign_kernel<<<gl, kThreads, 0, s>>>(ws.la_key, ws.la_val, ws.bitmap, ws.bm_prefix, (uint32_t)words, ws.skey, t,
                                        ws.counters, ws.list_cap, ws.err, slot_block);
  fill_keys_kernel<<<gl, kThreads, 0, s>>>(t, ws.counters, ws.list_cap, ws.err);
  return hipGetLastError();
}

}  // namespace

hipError_t table_clear(const TableView& t, Workspace& ws, hipStream_t s, uint32_t* zero, int64_t nzero) {
  const int64_t m = (int64_t)t.T > (int64_t)ws.ts_cap ? t.T : ws.ts_cap;
  clear_table_kernel<<<groups_for(m, kThreads), kThreads, 0, s>>>(t, ws.skey, ws.sorder, ws.dmask, ws.ts_cap,
                                                                   ws.counters, zero, zero ? nzero : 0);
  return hipGetLastError();
}

hipError_t table_activate(const TableView& t, Workspace& ws, const float* x, int64_t n, float dx,
                          int blockLen, int offset, float disp, hipStream_t s, bool record_slots,
                          const SceneSet* ss) {
  if (blockLen <= 0) return hipErrorInvalidValue;
  int32_t* slot_block = record_slots ? ws.slot_block : nullptr;
  const uint64_t order_bound = (uint64_t)(n > 0 ? n : 1);
  clear_scratch_kernel<<<groups_for(ws.ts_cap, kThreads), kThreads, 0, s>>>(
      ws.skey, ws.sorder, ws.dmask, ws.ts_cap, t.cnt, ws.counters, slot_block, ws.bitmap,
      (int64_t)bitmap_words(ws, order_bound));
  const float dxinv = (float)1.0 / dx;
  if (n > 0) {
    const int g = groups_for(n, kThreads * kActItems);
    uint32_t* ps = record_slots ? ws.pslot : nullptr;
    if (ss && ss->n > 0)
      activate_kernel<true><<<g, kThreads, 0, s>>>(x, n, dxinv, blockLen, offset, disp, dedup_of(ws), ps, *ss);
    else
      activate_kernel<false><<<g, kThreads, 0, s>>>(x, n, dxinv, blockLen, offset, disp, dedup_of(ws), ps,
                                                    SceneSet{});
  }
  return finish_insertion(t, ws, order_bound, s, slot_block);
}

hipError_t table_enlarge(const TableView& t, Workspace& ws, const int32_t lo[3], const int32_t hi[3],
                         int32_t n0_cap, hipStream_t s) {
  const int32_t e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
  if (e0 <= 0 || e1 <= 0 || e2 <= 0) return hipSuccess;
  const int64_t total_cap = (int64_t)n0_cap * e0 * e1 * e2;
  const uint64_t order_bound = (uint64_t)(total_cap > 0 ? total_cap : 1);
  clear_scratch_kernel<<<groups_for(ws.ts_cap, kThreads), kThreads, 0, s>>>(
      ws.skey, ws.sorder, ws.dmask, ws.ts_cap, t.cnt, ws.counters, nullptr, ws.bitmap,
      (int64_t)bitmap_words(ws, order_bound));
  enlarge_kernel<<<groups_for(total_cap, kThreads), kThreads, 0, s>>>(t.active, ws.counters, lo[0], lo[1], lo[2],
                                                                       e0, e1, e2, dedup_of(ws));
  return finish_insertion(t, ws, order_bound, s);
}

hipError_t block_neighbours(const TableView& t, Workspace& ws, hipStream_t s) {
  const int g = groups_for(8ll * ws.ts_cap, kThreads);
  neighbours_kernel<<<g < 1024 ? g : 1024, kThreads, 0, s>>>(t, ws.nbr, ws.ts_cap);
  return hipGetLastError();
}

hipError_t table_query(const TableView& t, const int32_t* keys, int64_t n, int32_t* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  query_kernel<<<groups_for(n, kThreads), kThreads, 0, s>>>(t, keys, n, out);
  return hipGetLastError();
}

namespace {

// after the bucket pass (bkt with run heads, counts, tile_heads): stable
// placement of the particles through the runs
hipError_t place_runs(Workspace& ws, int64_t n, const SortView& out, int64_t ntiles, int gt, hipStream_t s) {
  const int64_t nb1 = (int64_t)out.cap_blocks + 1;
  uint32_t* counts = (uint32_t*)out.counts;
  hipError_t e = exclusive_scan_u32(ws.tile_heads, ws.tile_base, nullptr, (uint32_t)ntiles, ws.scan_tmp,
                                    ws.counters + kCtrRuns, s);
  if (e != hipSuccess) return e;
  run_emit_kernel<<<gt, kThreads, 0, s>>>(ws.bkt, n, ws.tile_base, ws.run_start, ws.rk_a, ws.rv_a);
  bool alt = false;
  e = radix_sort_pairs(ws.rk_a, ws.rv_a, ws.rk_b, ws.rv_b, ws.counters + kCtrRuns, (uint32_t)n,
                       bits_for((uint64_t)out.cap_blocks - 1), ws.rs_scratch, s, &alt);
  if (e != hipSuccess) return e;
  const uint32_t* sorted_ids = alt ? ws.rv_b : ws.rv_a;
  uint32_t* spare = alt ? ws.rk_a : ws.rk_b;  // free buffer for the sorted dest scan
  const int gr = groups_for(n, kThreads) < 1024 ? groups_for(n, kThreads) : 1024;  // grid-stride over R <= n runs
  run_len_kernel<<<gr, kThreads, 0, s>>>(sorted_ids, ws.run_start, ws.counters, n, ws.run_len);
  e = exclusive_scan_u32(ws.run_len, spare, ws.counters + kCtrRuns, (uint32_t)n, ws.scan_tmp, nullptr, s);
  if (e != hipSuccess) return e;
  run_dest_kernel<<<gr, kThreads, 0, s>>>(sorted_ids, spare, ws.counters, n, ws.run_dest);
  e = exclusive_scan_u32(counts, (uint32_t*)out.offsets, nullptr, (uint32_t)nb1, ws.scan_tmp, nullptr, s);
  if (e != hipSuccess) return e;
  particle_scatter_kernel<<<gt, kThreads, 0, s>>>(ws.bkt, n, ws.tile_base, ws.run_start, ws.run_dest,
                                                  ws.counters, out.indices);
  return hipGetLastError();
}

}  // namespace

hipError_t bucket_sort(const TableView& t, Workspace& ws, const float* x, int64_t n, float dx, int blockLen,
                       int offset, float disp, const SortView& out, hipStream_t s) {
  if (n > (int64_t)ws.n_cap) return hipErrorInvalidValue;
  const int64_t nb1 = (int64_t)out.cap_blocks + 1;
  uint32_t* counts = (uint32_t*)out.counts;
  zero_u32_kernel<<<groups_for(nb1, kThreads), kThreads, 0, s>>>(counts, nb1);
  if (n <= 0) {
    return exclusive_scan_u32(counts, (uint32_t*)out.offsets, nullptr, (uint32_t)nb1, ws.scan_tmp, nullptr, s);
  }
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int gt = (int)(ntiles < 65536 ? ntiles : 65536);
  const float dxinv = (float)1.0 / dx;
  bucket_kernel<<<gt, kThreads, 0, s>>>(x, n, dxinv, blockLen, offset, disp, t, ws.bkt, counts, out.cap_blocks,
                                        ws.tile_heads, ws.err);
  return place_runs(ws, n, out, ntiles, gt, s);
}

hipError_t bucket_sort_slots(Workspace& ws, int64_t n, const SortView& out, hipStream_t s, bool counts_zeroed) {
  if (n > (int64_t)ws.n_cap) return hipErrorInvalidValue;
  const int64_t nb1 = (int64_t)out.cap_blocks + 1;
  uint32_t* counts = (uint32_t*)out.counts;
  if (!counts_zeroed) zero_u32_kernel<<<groups_for(nb1, kThreads), kThreads, 0, s>>>(counts, nb1);
  if (n <= 0) {
    return exclusive_scan_u32(counts, (uint32_t*)out.offsets, nullptr, (uint32_t)nb1, ws.scan_tmp, nullptr, s);
  }
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int gt = (int)(ntiles < 65536 ? ntiles : 65536);
  bucket_slots_kernel<<<gt, kThreads, 0, s>>>(ws.pslot, n, ws.slot_block, ws.ts_cap, ws.bkt, counts,
                                              out.cap_blocks, ws.tile_heads, ws.err);
  return place_runs(ws, n, out, ntiles, gt, s);
}

}  // namespace zsmpm
