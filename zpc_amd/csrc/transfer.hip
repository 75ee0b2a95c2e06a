// P2G, grid update and G2P for the gfx950 MLS-MPM substep.
//
// Work is organised per particle bucket: the 4^3 block a particle's key maps
// to (SparsityOp.hpp:72-77 with blockLen 4, offset -2, displacement 0.5). A
// particle with key cell c has its quadratic stencil base at c+1 (c or c+2
// only at rounding ties), so the stencils of a bucket whose key block starts
// at cell 4b lie in nodes [4b, 4b+8)^3, all inside the 2x2x2 blocks
// b + {0,1}^3 that EnlargeSparsity{0..2} activates.
//
// P2G. On gfx950 LDS float atomics (ds_add_f32) retire about one lane every
// three clocks, roughly 30x slower than a plain LDS read-modify-write, so the
// scatter avoids them:
//   A  every lane takes one particle (full occupancy): stress (3x3 SVD and
//      the constitutive model), then an affine record -- mass, B and dx*A with
//      node value(o) = W(o) (B + dx A o) -- kept in LDS, and a base-cell id
//      (one of the 64 main cells [4b+1, 4b+5)^3, or "overflow");
//   B  a counting sort of the chunk by base cell (LDS integer atomics);
//   C  lane l of a wave sums its share of main cell l's particles in
//      registers, one x-slab of the 27-node stencil at a time, and adds the
//      node sums to the wave's own 6^3 node arena with plain float4
//      read-modify-writes (the 64 lanes address 64 different nodes);
//      overflow particles scatter with global atomics;
//   then the four wave arenas are summed and flushed with one global float
//   atomic per touched node and channel (zpc's CUDA P2G issues 189 global
//   atomics per particle, cuda/simulation/transfer/P2G.hpp:98-115).
// The substep's P2G adds the force term to the momentum channels (the
// explicit driver's v = (mv + rhs)/m convention) and the grid carries
// {m, mv + rhs}; the stage-wise P2G keeps the rhs channels 4..6 as zpc lays
// them out, from two launches: momentum only, then the force only.
//
// G2P stages the 8^3 node velocities of a bucket in LDS (float4 per node) and
// gathers from there.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "zsmpm_device.hpp"
#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kMainCells = 64;    // base cells [4b+1, 4b+5)^3
constexpr int kArena6 = 216;      // nodes [4b+1, 4b+7)^3
constexpr int kArena8 = 512;      // nodes [4b, 4b+8)^3
// G2P velocity stage: node (r0, r1, r2) at r0*73 + r1*9 + r2, so a float4
// node's LDS bank group is (r0 + r1 + r2) mod 8 rather than r2 alone
constexpr int kVelSY = 9, kVelSX = 73, kVelNodes = 7 * kVelSX + 7 * kVelSY + 8;
__device__ __forceinline__ int vel_at(int r0, int r1, int r2) { return r0 * kVelSX + r1 * kVelSY + r2; }
// one workgroup per bucket. Two waves: a bucket's 8 staged velocity blocks
// serve ~3 rounds of 64 particles per wave, and eight such workgroups per CU
// (the register limit, 4 waves per SIMD) keep more buckets' loads in flight
// than four workgroups of four waves (0.82 -> 0.79 ms at 16M particles)
constexpr int kG2PThreads = 128;
constexpr int kP2GThreads = 256;  // one workgroup per bucket
constexpr int kP2GWaves = kP2GThreads / kWave;
// Launch tiles the autotuner chooses between. P2G tile 0: four waves and
// chunks of Rec::kChunk particles (a whole bucket of up to 576, three
// workgroups per CU); tile 1: two waves and chunks of 256 (six workgroups per
// CU, for sparser buckets). G2P: 128 or 256 threads per bucket.
constexpr int p2g_threads(int tile) { return tile == 0 ? kP2GThreads : 128; }
constexpr int p2g_chunk(int tile, int fold_chunk) { return tile == 0 ? fold_chunk : 256; }

inline int groups_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 65535 * 8) g = 65535 * 8;
  return (int)g;
}

__device__ __forceinline__ int64_t grid_at(int32_t b, int ch, int cell) {
  return ((int64_t)b * kChannels + ch) * kBlockCells + cell;
}

// node r in [0, 8)^3 relative to 4b -> (neighbour block of b + {0,1}^3, cell)
__device__ __forceinline__ void node_home(int r0, int r1, int r2, int& nb, int& cell) {
  nb = ((r0 >> 2) << 2) | ((r1 >> 2) << 1) | (r2 >> 2);
  cell = ((r0 & 3) << 4) | ((r1 & 3) << 2) | (r2 & 3);
}

__global__ void grid_clear_kernel(float* grid, const int32_t* nblocks, int32_t cap, uint32_t* counters) {
  if (counters && blockIdx.x == 0 && threadIdx.x == 0) {
    counters[kCtrQueue] = 0u;
    counters[kCtrMaxVel] = 0u;
  }
  const int32_t nb = min(*nblocks, cap);
  const int64_t total = (int64_t)nb * kChannels * kBlockCells / 4;
  float4* g4 = reinterpret_cast<float4*>(grid);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += stride)
    g4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// active block count, clipped to what the grid and bucket arrays hold
__device__ __forceinline__ int32_t bounded_blocks(const TableView& t, const SortView& sv, int32_t grid_cap,
                                                  uint32_t* err) {
  const int32_t nb = *t.cnt;
  const int32_t cap = min(grid_cap, sv.cap_blocks);
  if (nb > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrCapacity);
    return cap;
  }
  return nb;
}

// P2G hands buckets out in order from a queue (one atomic per bucket), so a
// workgroup that drew small buckets takes more: bucket sizes range from 0
// (enlarged-only blocks) to several hundred particles. Every workgroup leaves
// once the queue passes the bucket count.
__global__ void reset_queue_kernel(uint32_t* counters) { counters[kCtrQueue] = 0u; }

// neighbour blocks key + (d>>2&1, d>>1&1, d&1) of bucket b, from the
// per-substep neighbour table
__device__ __forceinline__ void load_neighbours(const int32_t* __restrict__ nbr_tab, int32_t b, int32_t* nbr,
                                                uint32_t* err) {
  if (threadIdx.x < 8) {
    const int32_t r = nbr_tab[8 * (int64_t)b + threadIdx.x];
    nbr[threadIdx.x] = r;
    if (r < 0) atomicOr(err, kErrMissingBlock);
  }
}

// Count, offset, block key and scene of bucket b (0 for b >= nb or a bucket
// whose range is not within the particles), written to m[0..6] by one
// thread. In a batch of nscenes > 0 scenes the key is the scene's own
// (offset removed).
__device__ __forceinline__ void bucket_meta(int32_t b, int32_t nb, const SortView& sv, const TableView& t,
                                            int64_t n, int32_t* m, int32_t nscenes = 0) {
  int32_t cnt = 0, off = 0, k0 = 0, k1 = 0, k2 = 0, sc = 0;
  if (b < nb) {
    cnt = sv.counts[b];
    off = sv.offsets[b];
    if (cnt <= 0 || off < 0 || (int64_t)off + cnt > n) cnt = 0;
    k0 = t.active[3 * (int64_t)b];
    k1 = t.active[3 * (int64_t)b + 1];
    k2 = t.active[3 * (int64_t)b + 2];
    if (nscenes > 0) {
      sc = scene_of_key(k0, nscenes);
      k0 -= sc * kSceneKeyStride;
    }
  }
  m[0] = b;
  m[1] = cnt;
  m[2] = off;
  m[3] = k0;
  m[4] = k1;
  m[5] = k2;
  m[6] = sc;
}

// Main base cell (i, j, k) in [0,4)^3 <-> lane of the scatter rounds. Lane
// bits are (k0, k1, i0, j0, i1, j1): the 8 consecutive lanes one LDS b128
// access serves together cover 4 values of k and 2 of i, 36 nodes apart in
// the 6^3 arena (36 = 4 mod 8), so they fall in 8 distinct bank groups.
__device__ __forceinline__ int cell_lane(int i, int j, int k) {
  return k | ((i & 1) << 2) | ((j & 1) << 3) | ((i >> 1) << 4) | ((j >> 1) << 5);
}
__device__ __forceinline__ void lane_cell(int l, int& i, int& j, int& k) {
  k = l & 3;
  i = ((l >> 2) & 1) | (((l >> 4) & 1) << 1);
  j = ((l >> 3) & 1) | (((l >> 5) & 1) << 1);
}

// quadratic B-spline weights from the local position l in [0.5, 1.5)
// (quadratic_bspline_weights, math/curve/InterpolationKernel.hpp:95-130)
__device__ __forceinline__ void bspline_w(float l, float (&w)[3]) {
#pragma clang fp contract(off)
  const float d0 = l - (float)(int32_t)floorf(l - 0.5f);
  w[0] = 0.5f * (1.5f - d0) * (1.5f - d0);
  const float d1 = d0 - 1.0f;
  w[1] = 0.75f - d1 * d1;
  const float zz = 0.5f + d1;
  w[2] = 0.5f * zz * zz;
}

// Node weights of the particle -> cell -> node kernel (P2C2GTransfer,
// simulation/transfer/P2C2G.hpp:143-147 and the 1/8 of :164; G2C2PTransfer,
// G2C2P.hpp:81-84, :112-119): the tent weights 1 - t and t to the two cell
// centres within reach (t = l - 0.5 from the local position l in [0.5, 1.5)),
// each spread evenly over its cell's corners, per axis (1 - t) / 2, 1 / 2 and
// t / 2 on the nodes base .. base + 2.
__device__ __forceinline__ void cell_w(float l, float (&w)[3]) {
#pragma clang fp contract(off)
  const float t = l - 0.5f;
  w[0] = (1.f - t) * 0.5f;
  w[1] = 0.5f;
  w[2] = t * 0.5f;
}

template <bool kCell>
__device__ __forceinline__ void stencil_w(float l, float (&w)[3]) {
  if constexpr (kCell)
    cell_w(l, w);
  else
    bspline_w(l, w);
}

// Per-axis factor of the cell kernel (P2C2G.hpp:86-90, G2C2P.hpp:235-238):
// Dinv = 2 / (dx^2 - 2 Xrel^2), Xrel = x - floor(x dx_inv + 0.5) dx the offset
// from the nearest node; 1 / Dinv is the kernel's second moment on that axis.
__device__ __forceinline__ void cell_dinv(const float (&x)[3], float dx, float dx_inv, float (&Dinv)[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float xr = x[d] - (float)(int32_t)floorf(x[d] * dx_inv + 0.5f) * dx;
    Dinv[d] = 2.f / (dx * dx - 2.f * xr * xr);
  }
}

// Record layout (float4 units): q0 = (m, B0, B1, B2), q1 = (l0, l1, l2, A8),
// q2 = A0..A3, q3 = A4..A7 with A column-major (A[d + 3e]) already times dx.
struct Rec {
  static constexpr int kQuads = 4;
  // particles staged per pass. 576 (nine waves) holds a whole bucket at 8
  // particles per cell (up to 562 in the 16M sand column; 39% of its buckets
  // exceed 512 and took a nearly empty second pass, 0.80 -> 0.765 ms) and
  // keeps the workgroup at 53 KB of LDS, three per CU
  static constexpr int kChunk = 576;
};

// Per-particle constitutive part of P2G (P2G.hpp:57-100): vol P F^T from the
// stored F (fixed-corotated or the Drucker-Prager return mapping, which
// writes logJp back). Particles in array order, 64 consecutive per wave, so
// rows always move as 16-byte vectors.
// kKind: 0 fixed-corotated, 1 Drucker-Prager, 2 each particle its scene's
// model (a batch, ss)
template <int kKind>
__global__ __launch_bounds__(kThreads) void stress_kernel(Model m, ParticlesView ps, float* __restrict__ contrib,
                                                          uint32_t* counters, SceneSet ss) {
  __shared__ float4 stage4[kWaves][stage_floats(9) / 4];
  __shared__ int64_t sfirst[kKind == 2 ? kMaxScenes + 1 : 1];
  const int wave = uniform((int32_t)(threadIdx.x / kWave)), lane = lane_id();
  float* stage = reinterpret_cast<float*>(stage4[wave]);
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  uint32_t fp64_used = 0;
  if constexpr (kKind == 2) {
    for (int i = threadIdx.x; i <= ss.n; i += kThreads) sfirst[i] = ss.first[i];
    __syncthreads();
  }
  // The next group's F rows and logJp are in flight while this group's SVD
  // runs (stress pass 0.70 -> 0.67 ms beside the sparsity stages at 16M
  // particles). Two groups ahead made the pass faster still (0.61 ms) but
  // took enough of the machine to slow the bucket sort beside it by more.
  RowLoad<9> ld;
  int64_t w = (int64_t)blockIdx.x * kWaves + wave;
  float lj_next = 0.f;
  if (w * kWave < ps.n) {
    const int64_t q0 = w * kWave;
    const int32_t qw = (int32_t)min((int64_t)kWave, ps.n - q0);
    ld.issue(ps.F, q0, qw);
    if (kKind == 1 && lane < qw) lj_next = ps.logJp[q0 + lane];
  }
  for (; w * kWave < ps.n; w += nwaves) {
    const int64_t p0 = w * kWave;
    const int32_t cw = (int32_t)min((int64_t)kWave, ps.n - p0);
    const int64_t p = p0 + lane;
    float F[9], PF[9];
    float lj = lj_next;
    ld.finish(stage, cw, F);
    const int64_t wn = w + nwaves;
    if (wn * kWave < ps.n) {
      const int64_t q0 = wn * kWave;
      const int32_t qw = (int32_t)min((int64_t)kWave, ps.n - q0);
      ld.issue(ps.F, q0, qw);
      if (kKind == 1 && lane < qw) lj_next = ps.logJp[q0 + lane];
    }
    if (lane < cw) {
      bool f64 = false;
      if (kKind == 0) {
        stress_fixedcorotated(m, F, PF, f64);
      } else if (kKind == 1) {
        stress_sand(m, F, lj, PF, nullptr, f64);
        ps.logJp[p] = lj;
      } else {
        const Model ms = ss.p[scene_of_particle(sfirst, ss.n, p)].model;
        if (ms.kind == 0) {
          stress_fixedcorotated(ms, F, PF, f64);
        } else {
          lj = ps.logJp[p];
          stress_sand(ms, F, lj, PF, nullptr, f64);
          ps.logJp[p] = lj;
        }
      }
      fp64_used += f64 ? 1u : 0u;
    }
    wave_store_rows<9>(contrib, p0, cw, stage, PF);
  }
  if (fp64_used) atomicAdd(&counters[kCtrFp64], fp64_used);
}

// kMode 0: mass and momentum with the force folded in (channels 0..3);
// 1: mass and momentum only (the stress is not read); 2: the force only, as
// if every particle were massless (v, C and m are not read), into channels
// fch .. fch + 2. kMulti (a batch, ss): each bucket's dx and dt are its
// scene's. kTile: the launch tile (p2g_threads, p2g_chunk).
// kCell (mode 0, one scene): P2C2GTransfer's particle -> cell -> node kernel
// (cell_w) in place of the quadratic B-spline, with C and the stress scaled
// per column by the cell kernel's Dinv.
template <int kMode, bool kMulti, bool kCell = false, int kTile = 0>
__global__ __launch_bounds__(p2g_threads(kTile)) void p2g_kernel(ParticlesView ps, const float* __restrict__ contrib,
                                                       TableView t, SortView sv, const int32_t* __restrict__ nbr_tab,
                                                       float* __restrict__ grid, int32_t grid_cap, float dt,
                                                       float dx, uint32_t* err, uint32_t* queue, SceneSet ss,
                                                       int32_t fch) {
  static_assert(kMode >= 0 && kMode <= 2 && (!kCell || kMode == 0), "P2G modes");
  constexpr bool kMom = kMode != 2;     // mass, v and C are read
  constexpr bool kStress = kMode != 1;  // the stress is read
  constexpr int kQ = Rec::kQuads;
  constexpr int kT = p2g_threads(kTile);
  constexpr int kWv = kT / kWave;
  constexpr int kChunk = p2g_chunk(kTile, Rec::kChunk);
  static_assert(kT % kWave == 0 && kChunk % kWave == 0,
                "whole waves; a chunk is a whole number of waves");
  __shared__ float4 arena[kWv][kArena6];
  // records in component-major order: component k of record i at
  // rec[k * kChunk + i], so consecutive records are consecutive float4 (no
  // LDS bank conflicts when lanes write consecutive records)
  __shared__ float4 rec[kChunk * kQ];
  __shared__ uint16_t list[kChunk];
  __shared__ uint8_t cellof[kChunk];
  __shared__ int32_t ccount[kMainCells + 1], cstart[kMainCells + 1], cfill[kMainCells + 1];
  __shared__ int32_t nbr[8];

  const int32_t nb = bounded_blocks(t, sv, grid_cap, err);
  const float dx_inv = (float)1.0 / dx;
  const float D_inv = 4.f * dx_inv * dx_inv;
  const float scale = -dt * D_inv;
  const int wave = uniform((int32_t)(threadIdx.x / kWave)), lane = lane_id();
  // Buckets come from the queue one ahead: thread 0 pops the next bucket
  // after the last chunk's counting sort and loads its count, offset and key
  // while phase C runs (bmeta, published by the barrier that ends the chunk);
  // every thread then issues the loads of its particle indices in the next
  // bucket's first chunk before the flush, so they are in flight across the
  // flush and the next bucket's set-up instead of starting phase A.
  constexpr int kPre = (kChunk + kT - 1) / kT;  // chunk slots per thread
  __shared__ int32_t bmeta[7];  // bucket, count, offset, key, scene
  const int32_t nscenes = kMulti ? ss.n : 0;
  if (threadIdx.x == 0)
    bucket_meta((int32_t)min(atomicAdd(queue, 1u), 0x7fffffffu), nb, sv, t, ps.n, bmeta, nscenes);
  for (int c = threadIdx.x; c <= kMainCells; c += kT) {
    ccount[c] = 0;
    cfill[c] = 0;
  }
  __syncthreads();
  int32_t pre[kPre];
  auto prefetch = [&]() {
    const int32_t lim = min(bmeta[1], kChunk), po = bmeta[2];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
      const int32_t i = threadIdx.x + k * kT;
      pre[k] = i < lim ? sv.indices[po + i] : -1;
    }
  };
  prefetch();

  for (;;) {
    const int32_t b = bmeta[0], cnt = bmeta[1], off = bmeta[2];
    if (b >= nb) break;
    const int32_t o0 = bmeta[3] * 4, o1 = bmeta[4] * 4, o2 = bmeta[5] * 4;
    if (cnt == 0) {  // empty or out-of-range bucket: take the next one now
      __syncthreads();
      if (threadIdx.x == 0)
        bucket_meta((int32_t)min(atomicAdd(queue, 1u), 0x7fffffffu), nb, sv, t, ps.n, bmeta, nscenes);
      __syncthreads();
      prefetch();
      continue;
    }
    // the bucket's grid spacing and force scale (its scene's, in a batch)
    float bdx = dx, bscale = scale;
    if constexpr (kMulti) {
      const SceneParams& sp = ss.p[bmeta[6]];
      bdx = sp.dx;
      const float bdx_inv = (float)1.0 / bdx;
      bscale = -sp.dt * (4.f * bdx_inv * bdx_inv);
    }
    load_neighbours(nbr_tab, b, nbr, err);
    uint32_t nxt = 0;  // thread 0: the next bucket, popped during the last chunk

    for (int32_t c0 = 0; c0 < cnt; c0 += kChunk) {
      const int32_t nch = min(kChunk, cnt - c0);
      if (c0 > 0) {
        // later chunks clear the cell counters here; a bucket's first chunk
        // finds them cleared by the previous bucket's flush (or the prologue)
        for (int c = threadIdx.x; c <= kMainCells; c += kT) {  // 64 main cells + overflow
          ccount[c] = 0;
          cfill[c] = 0;
        }
        __syncthreads();
      }
      // ---- A: stress + affine record per particle. Each wave takes 64
      // consecutive chunk slots (i = threadIdx.x + k * kT, as phase B
      // expects); a contiguous run of particles is read with 16-byte vectors,
      // staged in the wave's own (not yet written) record slots.
#pragma unroll
      for (int kk = 0; kk < kPre; ++kk) {
        const int32_t i0 = wave * kWave + kk * kT;
        if (i0 >= nch) break;
        const int32_t i = i0 + lane;
        const int32_t cw = min(kWave, nch - i0);
        const bool valid = lane < cw;
        const int64_t p = valid ? (int64_t)(c0 == 0 ? pre[kk] : sv.indices[off + c0 + i]) : -1;
        const bool bad = valid && (p < 0 || p >= ps.n);
        if (bad) {
          atomicOr(err, kErrCapacity);
          cellof[i] = 255;
          list[i] = 0;
        }
        const bool live = valid && !bad;
        int64_t p0;
        // the wave's arena is free during the first chunk's phase A (it is
        // cleared right after): it stages the vector row loads
        const bool contig = c0 == 0 && wave_contiguous(p, live, p0) && ballot(bad) == 0;
        float* stage = reinterpret_cast<float*>(arena[wave]);
        float xv[3] = {0.f, 0.f, 0.f}, vv[3] = {0.f, 0.f, 0.f};
        float PF[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float C[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const float mass = (live && kMom) ? ps.m[p] : 0.f;
        if (contig) {
          RowLoad<3> lx, lv;
          RowLoad<9> lp, lc;
          // the wave issues its row loads ahead of other waves' scatter
          // arithmetic (0.763 -> 0.757 ms at 16M particles)
          __builtin_amdgcn_s_setprio(2);
          lx.issue(ps.x, p0, cw);
          if (kMom) lv.issue(ps.v, p0, cw);
          if (kStress) lp.issue(contrib, p0, cw);
          if (kMom) lc.issue(ps.C, p0, cw);
          __builtin_amdgcn_s_setprio(0);
          lx.finish(stage, cw, xv);
          if (kMom) lv.finish(stage, cw, vv);
          if (kStress) lp.finish(stage, cw, PF);
          if (kMom) lc.finish(stage, cw, C);
        } else if (live) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            xv[d] = ps.x[3 * p + d];
            if (kMom) vv[d] = ps.v[3 * p + d];
          }
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            if (kStress) PF[k] = contrib[9 * p + k];
            if (kMom) C[k] = ps.C[9 * p + k];
          }
        }
        if (!live) continue;
        const float x0 = xv[0], x1 = xv[1], x2 = xv[2];
        const float v0 = vv[0], v1 = vv[1], v2 = vv[2];
        // base node and local position in grid units (LocalArena::init, Utils.hpp:51-70)
        float l[3];
        int32_t base[3];
        {
#pragma clang fp contract(off)
          const float xs[3] = {x0, x1, x2};
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const float X = xs[d] / bdx;
            base[d] = (int32_t)floorf(X - 0.5f);
            l[d] = X - (float)base[d];
          }
        }
        const int r0 = base[0] - o0, r1 = base[1] - o1, r2 = base[2] - o2;  // in [0, 6)
        int cell;
        if (r0 >= 1 && r0 <= 4 && r1 >= 1 && r1 <= 4 && r2 >= 1 && r2 <= 4) {
          cell = cell_lane(r0 - 1, r1 - 1, r2 - 1);
        } else if (r0 >= 0 && r0 <= 5 && r1 >= 0 && r1 <= 5 && r2 >= 0 && r2 <= 5) {
          cell = kMainCells;  // overflow: scattered with global atomics
        } else {
          atomicOr(err, kErrArena);
          cell = 255;
        }
        // affine record: m v + (m C + contrib') (o - l) dx
        float A[9], Bv[3];
        const float mv[3] = {mass * v0, mass * v1, mass * v2};
        if constexpr (kCell) {
          // P2C2G.hpp:86-96, 138-140: C and vol P F^T scaled per column by
          // Dinv, the stress by -dt, summed with m C
#pragma clang fp contract(off)
          float Dinv[3];
          cell_dinv(xv, bdx, dx_inv, Dinv);
#pragma unroll
          for (int k = 0; k < 9; ++k) A[k] = (PF[k] * (Dinv[k / 3] * -dt) + (C[k] * Dinv[k / 3]) * mass) * bdx;
        } else {
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            if (kMode == 0) A[k] = (mass * C[k] + PF[k] * bscale) * bdx;
            if (kMode == 1) A[k] = mass * C[k] * bdx;
            if (kMode == 2) A[k] = PF[k] * bscale * bdx;
          }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) Bv[d] = mv[d] - (A[d] * l[0] + A[d + 3] * l[1] + A[d + 6] * l[2]);
        rec[0 * kChunk + i] = make_float4(mass, Bv[0], Bv[1], Bv[2]);
        rec[1 * kChunk + i] = make_float4(l[0], l[1], l[2], A[8]);
        rec[2 * kChunk + i] = make_float4(A[0], A[1], A[2], A[3]);
        rec[3 * kChunk + i] = make_float4(A[4], A[5], A[6], A[7]);
        cellof[i] = (uint8_t)cell;
        if (cell <= kMainCells) atomicAdd(&ccount[cell], 1);
        list[i] = (uint16_t)((r0 << 6) | (r1 << 3) | r2);  // base offsets; the list is rebuilt in B
      }
      if (c0 == 0) {
        // the first chunk's loads are done with the wave's arena as stage:
        // clear it for the bucket's accumulation
        wave_lds_barrier();
        for (int k = lane; k < kArena6; k += kWave) arena[wave][k] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      __syncthreads();
      // ---- B: counting sort of the chunk by base cell
      if (threadIdx.x < kWave) {
        const int32_t c = (threadIdx.x < kMainCells) ? ccount[threadIdx.x] : 0;
        int32_t inc = c;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          const int32_t u = __shfl_up(inc, o, kWave);
          if (lane >= o) inc += u;
        }
        cstart[threadIdx.x] = inc - c;
        if (threadIdx.x == kWave - 1) cstart[kMainCells] = inc;  // overflow particles follow the main cells
      }
      // each thread keeps the base offsets and cells of its own particles
      // across the rebuild of list[]
      uint16_t keep[kPre];
      uint8_t keepc[kPre];
#pragma unroll
      for (int k = 0; k < kPre; ++k) {
        const int32_t i = threadIdx.x + k * kT;
        keep[k] = i < nch ? list[i] : (uint16_t)0;
        keepc[k] = i < nch ? cellof[i] : (uint8_t)255;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kPre; ++k) {
        const int32_t i = threadIdx.x + k * kT;
        if (i < nch && keepc[k] <= kMainCells) {
          const int32_t slot = atomicAdd(&cfill[keepc[k]], 1);
          list[cstart[keepc[k]] + slot] = (uint16_t)i;
        }
      }
      __syncthreads();
      const bool last = c0 + kChunk >= cnt;
      if (last && threadIdx.x == 0) nxt = atomicAdd(queue, 1u);
      // ---- C: rounds, lane = main cell
      {
        const int32_t lcount = ccount[lane];
        const int32_t lstart = cstart[lane];
        int lr0, lr1, lr2;  // base - (4b+1) of the lane's main cell
        lane_cell(lane, lr0, lr1, lr2);
        float4* ar = arena[wave];
        // All particles of one base cell touch the same 27 nodes: the lane
        // sums its share of its cell's particles (every kWv-th, starting
        // at the wave index) in registers and adds the node sums to the
        // wave's arena once per cell, instead of 27 LDS read-modify-writes
        // per particle. One pass per x-slab of the stencil keeps the
        // accumulators at 9 nodes x 4 channels.
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          float acc[9][4];
#pragma unroll
          for (int k = 0; k < 9; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.f;
          for (int32_t r = wave; r < lcount; r += kWv) {
            const int32_t i = list[lstart + r];
            const float4 q0 = rec[i], q1 = rec[kChunk + i], q2 = rec[2 * kChunk + i], q3 = rec[3 * kChunk + i];
            float w0[3], w1[3], w2[3];
            stencil_w<kCell>(q1.x, w0);
            stencil_w<kCell>(q1.y, w1);
            stencil_w<kCell>(q1.z, w2);
            const float fa = (float)a;
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
              const float w01 = w0[a] * w1[bb];
              const float fb = (float)bb;
              const float s0 = q0.y + q2.x * fa + q2.w * fb;
              const float s1 = q0.z + q2.y * fa + q3.x * fb;
              const float s2 = q0.w + q2.z * fa + q3.y * fb;
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const float W = w01 * w2[c];
                const float fc = (float)c;
                float* e = acc[bb * 3 + c];
                if (kMom) e[0] += q0.x * W;
                e[1] += W * (s0 + q3.z * fc);
                e[2] += W * (s1 + q3.w * fc);
                e[3] += W * (s2 + q1.w * fc);
              }
            }
          }
          if (wave < lcount) {
#pragma unroll
            for (int bb = 0; bb < 3; ++bb)
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const int n = (lr0 + a) * 36 + (lr1 + bb) * 6 + (lr2 + c);
                const float* e = acc[bb * 3 + c];
                // Lanes of one step touch distinct nodes, but a node a lane
                // updates in this step may be one another lane updated in an
                // earlier step. The LDS unit serves a wave's accesses in issue
                // order; this barrier keeps the compiler from hoisting this
                // step's read above earlier steps' writes.
                asm volatile("" ::: "memory");
                float4 v = ar[n];
                v.x += e[0];
                v.y += e[1];
                v.z += e[2];
                v.w += e[3];
                ar[n] = v;
              }
          }
        }
      }
      // ---- overflow particles: global atomics straight into the grid
#pragma unroll
      for (int k = 0; k < kPre; ++k) {
        const int32_t i = threadIdx.x + k * kT;
        if (i >= nch || keepc[k] != kMainCells) continue;
        const int r0 = (keep[k] >> 6) & 7, r1 = (keep[k] >> 3) & 7, r2 = keep[k] & 7;
        const float4 q0 = rec[i], q1 = rec[kChunk + i], q2 = rec[2 * kChunk + i], q3 = rec[3 * kChunk + i];
        const float Am[9] = {q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w, q1.w};
        float w0[3], w1[3], w2[3];
        stencil_w<kCell>(q1.x, w0);
        stencil_w<kCell>(q1.y, w1);
        stencil_w<kCell>(q1.z, w2);
        for (int a = 0; a < 3; ++a)
          for (int bb = 0; bb < 3; ++bb)
            for (int c = 0; c < 3; ++c) {
              const float W = w0[a] * w1[bb] * w2[c];
              int nbi, cell;
              node_home(r0 + a, r1 + bb, r2 + c, nbi, cell);
              const int32_t gb = nbr[nbi];
              if (gb < 0 || gb >= grid_cap) continue;
              float* gp = grid + grid_at(gb, 0, cell);
              const float fa = (float)a, fb = (float)bb, fc = (float)c;
              float* gv = kMom ? gp + kBlockCells : gp + fch * kBlockCells;  // momentum or force channels
              if (kMom) unsafeAtomicAdd(gp, q0.x * W);
              unsafeAtomicAdd(gv, W * (q0.y + Am[0] * fa + Am[3] * fb + Am[6] * fc));
              unsafeAtomicAdd(gv + kBlockCells, W * (q0.z + Am[1] * fa + Am[4] * fb + Am[7] * fc));
              unsafeAtomicAdd(gv + 2 * kBlockCells, W * (q0.w + Am[2] * fa + Am[5] * fb + Am[8] * fc));
            }
      }
      if (last && threadIdx.x == 0) bucket_meta((int32_t)min(nxt, 0x7fffffffu), nb, sv, t, ps.n, bmeta, nscenes);
      __syncthreads();
    }
    prefetch();  // the next bucket's first-chunk indices (bmeta is the next bucket's now)
    // the next bucket's first chunk finds its cell counters cleared (they are
    // not read from here on; the barrier after the flush orders the clear)
    for (int c = threadIdx.x; c <= kMainCells; c += kT) {
      ccount[c] = 0;
      cfill[c] = 0;
    }
    // ---- flush: sum the wave arenas, one global atomic per touched node/channel
    for (int n = threadIdx.x; n < kArena6; n += kT) {
      float4 s = arena[0][n];
#pragma unroll
      for (int w = 1; w < kWv; ++w) {
        const float4 u = arena[w][n];
        s.x += u.x;
        s.y += u.y;
        s.z += u.z;
        s.w += u.w;
      }
      // every node a particle touched, massless particles included (zpc adds
      // all 27 nodes, P2G.hpp:114-123); untouched nodes hold exact zeros
      const bool touched = s.x != 0.f || s.y != 0.f || s.z != 0.f || s.w != 0.f;
      if (touched) {
        int nbi, cell;
        node_home(n / 36 + 1, (n / 6) % 6 + 1, n % 6 + 1, nbi, cell);
        const int32_t gb = nbr[nbi];
        if (gb >= 0 && gb < grid_cap) {
          float* gp = grid + grid_at(gb, 0, cell);
          float* gv = kMom ? gp + kBlockCells : gp + fch * kBlockCells;  // momentum or force channels
          if (kMom) unsafeAtomicAdd(gp, s.x);
          unsafeAtomicAdd(gv, s.y);
          unsafeAtomicAdd(gv + kBlockCells, s.z);
          unsafeAtomicAdd(gv + 2 * kBlockCells, s.w);
        }
      }
    }
    __syncthreads();
  }
}

// v = (mv [+ rhs]) / m + extf dt where m != 0 (GridOp.hpp:71-110), max |v|^2,
// then the colliders in order on nodes with m > 0 (GridOp.hpp:112-164,
// Collider.h:80-107). fold adds the rhs channels first (the explicit
// driver's convention when P2G kept them separate).
// Signed distance of a general collider's level set at the material-space
// point X (AnalyticLevelSet.h: Plane :26-29, Cuboid :81-88, Sphere :142-145,
// Cylinder :180-207), rounded as written.
__device__ __forceinline__ float level_set_sdf(const GeneralCollider& c, const float (&X)[3]) {
#pragma clang fp contract(off)
  if (c.geometry == 0) {
    return c.p1[0] * (X[0] - c.p0[0]) + c.p1[1] * (X[1] - c.p0[1]) + c.p1[2] * (X[2] - c.p0[2]);
  }
  if (c.geometry == 1) {
    float point[3], mx = 0.f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float center = (c.p0[d] + c.p1[d]) / 2.f;
      point[d] = fabsf(X[d] - center) - (c.p1[d] - c.p0[d]) / 2.f;
      mx = d == 0 ? point[d] : fmaxf(mx, point[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (point[d] < 0.f) point[d] = 0.f;
    const float len = sqrtf(point[0] * point[0] + point[1] * point[1] + point[2] * point[2]);
    return (mx < 0.f ? mx : 0.f) + len;
  }
  if (c.geometry == 2) {
    const float o0 = X[0] - c.p0[0], o1 = X[1] - c.p0[1], o2 = X[2] - c.p0[2];
    return sqrtf(o0 * o0 + o1 * o1 + o2 * o2) - c.p1[0];
  }
  const int ax = c.axis;
  const float dr0 = X[ax == 0 ? 1 : 0] - c.p0[ax == 0 ? 1 : 0];
  const float dr1 = X[ax == 2 ? 1 : 2] - c.p0[ax == 2 ? 1 : 2];
  const float disR = sqrtf(dr0 * dr0 + dr1 * dr1);
  const float r = c.p1[0], len = c.p1[1];
  const bool outside = disR > r;
  if (X[ax] < c.p0[ax]) {
    const float disL = c.p0[ax] - X[ax];
    return outside ? sqrtf((disR - r) * (disR - r) + disL * disL) : disL;
  }
  if (X[ax] > c.p0[ax] + len) {
    const float disL = X[ax] - (c.p0[ax] + len);
    return outside ? sqrtf((disR - r) * (disR - r) + disL * disL) : disL;
  }
  if (outside) return disR - r;
  const float disL = fminf(c.p0[ax] + len - X[ax], X[ax] - c.p0[ax]);
  return -fminf(disL, r - disR);
}

// Material-space normal (AnalyticLevelSet.h :30-33, :89-103, :146-150,
// :208-222): the plane's normal as given, the sphere's outward direction
// (zero within 1e-7 squared of the center), normalised central differences
// of step 1e-6 for the cuboid and the cylinder.
__device__ __forceinline__ void level_set_normal(const GeneralCollider& c, const float (&X)[3], float (&n)[3]) {
#pragma clang fp contract(off)
  if (c.geometry == 0) {
    n[0] = c.p1[0];
    n[1] = c.p1[1];
    n[2] = c.p1[2];
    return;
  }
  if (c.geometry == 2) {
    const float o[3] = {X[0] - c.p0[0], X[1] - c.p0[1], X[2] - c.p0[2]};
    const float l2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    if (l2 < 1e-7f) {
      n[0] = n[1] = n[2] = 0.f;
      return;
    }
    const float l = sqrtf(l2);
    n[0] = o[0] / l;
    n[1] = o[1] / l;
    n[2] = o[2] / l;
    return;
  }
  const float eps = 1e-6f;
  float diff[3];
  for (int i = 0; i < 3; ++i) {
    float v1[3] = {X[0], X[1], X[2]}, v2[3] = {X[0], X[1], X[2]};
    v1[i] = X[i] + eps;
    v2[i] = X[i] - eps;
    diff[i] = (level_set_sdf(c, v1) - level_set_sdf(c, v2)) / (eps + eps);
  }
  const float l = sqrtf(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
  n[0] = diff[0] / l;
  n[1] = diff[1] / l;
  n[2] = diff[2] / l;
}

// Collider::resolveCollision (Collider.h:81-107) with erosion 0 and the
// analytic level sets' zero material velocity.
__device__ __forceinline__ void resolve_general(const GeneralCollider& c, const float (&x)[3], float (&v)[3]) {
#pragma clang fp contract(off)
  const float xmb[3] = {x[0] - c.b[0], x[1] - c.b[1], x[2] - c.b[2]};
  const float inv_s = 1.f / c.s;
  float X[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    X[i] = (c.R[0 + 3 * i] * xmb[0] + c.R[1 + 3 * i] * xmb[1] + c.R[2 + 3 * i] * xmb[2]) * inv_s;
  if (!(level_set_sdf(c, X) < 0.f)) return;
  const float rate = c.dsdt * inv_s;
  const float wx[3] = {c.omega[1] * xmb[2] - c.omega[2] * xmb[1], c.omega[2] * xmb[0] - c.omega[0] * xmb[2],
                       c.omega[0] * xmb[1] - c.omega[1] * xmb[0]};
  float vo[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) vo[d] = wx[d] + rate * xmb[d] + c.dbdt[d];
  if (c.type == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = vo[d];
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] -= vo[d];
  float nm[3], n[3];
  level_set_normal(c, X, nm);
#pragma unroll
  for (int r = 0; r < 3; ++r) n[r] = c.R[r] * nm[0] + c.R[r + 3] * nm[1] + c.R[r + 6] * nm[2];
  const float proj = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  if ((c.type == 2 && proj < 0.f) || c.type == 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] -= proj * n[d];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] += vo[d];
}

// Node velocity and colliders of one node (mass != 0) of the block with key
// k; cols is the kernel argument or a scene's colliders in a batch.
// kLocal: k0_local is the key's first component with a batch's scene offset
// removed; otherwise the table's own key is read.
template <bool kLocal = false, typename Cols>
__device__ __forceinline__ void node_velocity(float* gp, float mass, int fold, float e0, float e1, float e2,
                                              const Cols& cols, const GeneralColliders& gen, const TableView& t,
                                              int32_t b, int32_t k0_local, int cell, float dx, float& vmax) {
  // rounded as written (no FMA contraction): a node on a collider plane has
  // phi = 0 in exact arithmetic, and the sign of the rounded phi must be the
  // one the host path computes
#pragma clang fp contract(off)
  const float inv = 1.f / mass;
  float mv0 = gp[1 * kBlockCells], mv1 = gp[2 * kBlockCells], mv2 = gp[3 * kBlockCells];
  if (fold) {
    mv0 = mv0 + gp[4 * kBlockCells];
    mv1 = mv1 + gp[5 * kBlockCells];
    mv2 = mv2 + gp[6 * kBlockCells];
  }
  float v[3] = {mv0 * inv + e0, mv1 * inv + e1, mv2 * inv + e2};
  vmax = fmaxf(vmax, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (mass > 0.f && (cols.n > 0 || gen.n > 0 || gen.ls.on)) {
    const int32_t k0 = kLocal ? k0_local : t.active[3 * (int64_t)b];
    const int32_t k1 = t.active[3 * (int64_t)b + 1], k2 = t.active[3 * (int64_t)b + 2];
    const float pos[3] = {((float)k0 * 4.f + (float)((cell >> 4) & 3)) * dx,
                          ((float)k1 * 4.f + (float)((cell >> 2) & 3)) * dx,
                          ((float)k2 * 4.f + (float)(cell & 3)) * dx};
    for (int c = 0; c < cols.n; ++c) {
      const PlaneCollider& pc = cols.c[c];
      const float phi = pc.normal[0] * (pos[0] - pc.origin[0]) + pc.normal[1] * (pos[1] - pc.origin[1]) +
                        pc.normal[2] * (pos[2] - pc.origin[2]);
      if (phi < 0.f) {
        if (pc.type == 0) {
          v[0] = v[1] = v[2] = 0.f;
        } else {
          const float proj = pc.normal[0] * v[0] + pc.normal[1] * v[1] + pc.normal[2] * v[2];
          if ((pc.type == 2 && proj < 0.f) || pc.type == 1) {
            v[0] -= proj * pc.normal[0];
            v[1] -= proj * pc.normal[1];
            v[2] -= proj * pc.normal[2];
          }
        }
      }
    }
    for (int c = 0; c < gen.n; ++c) resolve_general(gen.c[c], pos, v);
    if (gen.ls.on) levelset_resolve(gen.ls, pos, v);
  }
  gp[1 * kBlockCells] = v[0];
  gp[2 * kBlockCells] = v[1];
  gp[3 * kBlockCells] = v[2];
}

template <bool kMulti>
__global__ __launch_bounds__(kThreads) void grid_update_kernel(float* __restrict__ grid, int32_t grid_cap,
                                                               TableView t, float dt, float g0, float g1,
                                                               float g2, Colliders cols, float dx, int fold,
                                                               uint32_t* counters, uint32_t* err, SceneSet ss,
                                                               GeneralColliders gen) {
  int32_t nb = *t.cnt;
  if (nb > grid_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrCapacity);
    nb = grid_cap;
  }
  const int64_t nodes = (int64_t)nb * kBlockCells;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float vmax = 0.f;
  const float e0 = g0 * dt, e1 = g1 * dt, e2 = g2 * dt;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nodes; base += stride) {
    const int64_t i = base + threadIdx.x;
    if (i < nodes) {
      const int32_t b = (int32_t)(i >> 6);
      const int cell = (int)(i & 63);
      float* gp = grid + grid_at(b, 0, cell);
      const float mass = gp[0];
      if (mass != 0.f) {
        if constexpr (kMulti) {
          // the node's scene (from its block key): its dt, gravity, dx and colliders
          int32_t k0 = t.active[3 * (int64_t)b];
          const int32_t sc = scene_of_key(k0, ss.n);
          const SceneParams& sp = ss.p[sc];
          k0 -= sc * kSceneKeyStride;
          node_velocity<true>(gp, mass, fold, sp.g[0] * sp.dt, sp.g[1] * sp.dt, sp.g[2] * sp.dt, sp.cols, gen, t, b,
                              k0, cell, sp.dx, vmax);
        } else {
          node_velocity(gp, mass, fold, e0, e1, e2, cols, gen, t, b, 0, cell, dx, vmax);
        }
      } else if (fold) {
        // the fold covers every node, massless ones included (their
        // momentum channels then hold mv + rhs, as after a separate fold)
#pragma unroll
        for (int d = 1; d <= 3; ++d) gp[d * kBlockCells] = gp[d * kBlockCells] + gp[(d + 3) * kBlockCells];
      }
    }
  }
  // max |v|^2: wave, then workgroup, then one atomic per workgroup -- and only
  // when it raises the running maximum (same-address atomics serialise)
  __shared__ float wmax[kWaves];
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, kWave));
  if (lane_id() == 0) wmax[threadIdx.x / kWave] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    float g = wmax[0];
    for (int w = 1; w < kWaves; ++w) g = fmaxf(g, wmax[w]);
    const uint32_t bits = __float_as_uint(g);
    if (g > 0.f && bits > __atomic_load_n(&counters[kCtrMaxVel], __ATOMIC_RELAXED))
      atomicMax(&counters[kCtrMaxVel], bits);
  }
}

// GridAngularMomentum (GridOp.hpp:216-259): for every node with m != 0,
// x = (blockkey * 4 + coord(cell)) * dx; sum[0..2] += (double)(x cross mv)
// formed in float, sum[3..5] += (double)mv. Per-thread double partials, a
// wave and workgroup reduction, six double atomics per workgroup.
__global__ __launch_bounds__(kThreads) void angular_momentum_kernel(const float* __restrict__ grid, int32_t grid_cap,
                                                                    TableView t, float dx, int mchn, int mvchn,
                                                                    double* sum, uint32_t* err) {
#pragma clang fp contract(off)
  int32_t nb = *t.cnt;
  if (nb > grid_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrCapacity);
    nb = grid_cap;
  }
  const int64_t nodes = (int64_t)nb * kBlockCells;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes; i += stride) {
    const int32_t b = (int32_t)(i >> 6);
    const int cell = (int)(i & 63);
    const float mass = grid[grid_at(b, mchn, cell)];
    if (mass == 0.f) continue;
    const float x0 = ((float)t.active[3 * (int64_t)b] * 4.f + (float)((cell >> 4) & 3)) * dx;
    const float x1 = ((float)t.active[3 * (int64_t)b + 1] * 4.f + (float)((cell >> 2) & 3)) * dx;
    const float x2 = ((float)t.active[3 * (int64_t)b + 2] * 4.f + (float)(cell & 3)) * dx;
    const float m0 = grid[grid_at(b, mvchn, cell)], m1 = grid[grid_at(b, mvchn + 1, cell)],
                m2 = grid[grid_at(b, mvchn + 2, cell)];
    acc[0] += (double)(x1 * m2 - x2 * m1);
    acc[1] += (double)(x2 * m0 - x0 * m2);
    acc[2] += (double)(x0 * m1 - x1 * m0);
    acc[3] += (double)m0;
    acc[4] += (double)m1;
    acc[5] += (double)m2;
  }
  __shared__ double part[kWaves][6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, kWave);
  }
  const int wave = threadIdx.x / kWave;
  if (lane_id() == 0)
    for (int k = 0; k < 6; ++k) part[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w) s += part[w][threadIdx.x];
    if (s != 0.0) atomicAdd(&sum[threadIdx.x], s);
  }
}

// kNext: -1 plain G2P; 0 / 1 (reorder only): also the next substep's
// fixed-corotated / Drucker-Prager stress from the updated F (NextStress).
// kCell (in place, one scene): G2C2PTransfer with its pre and post passes
// (G2C2P.hpp:58-151, 197-271): the cell kernel's weights (cell_w), B =
// sum W v_i (x_i - x_p)^T stored as C, and F updated with B scaled per column
// by Dinv at the old position.
template <bool kReorder, int kNext, bool kMulti = false, bool kCell = false, int kT = kG2PThreads>
__global__ __launch_bounds__(kT) void g2p_kernel(ParticlesView ps, ParticlesView out, TableView t,
                                                          SortView sv, const int32_t* __restrict__ nbr_tab,
                                                          const float* __restrict__ grid, int32_t grid_cap,
                                                          float dt_arg, float dx_arg, NextStress ns, uint32_t* err,
                                                          uint32_t* counters, SceneSet ss) {
  static_assert(kNext < 0 || kReorder, "the fused stress writes in output slot order");
  uint32_t fp64_used = 0;
  __shared__ float4 vel[kVelNodes];
  __shared__ int32_t nbr[8];
  __shared__ float4 stage4[kT / kWave][stage_floats(9) / 4];
  const int32_t nb = bounded_blocks(t, sv, grid_cap, err);
  const float dx_inv_arg = (float)1 / dx_arg;
  const float D_inv_arg = 4.f * dx_inv_arg * dx_inv_arg;
  const int wave = uniform((int32_t)(threadIdx.x / kWave)), lane = lane_id();
  float* stage = reinterpret_cast<float*>(stage4[wave]);
  for (int32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const int32_t cnt = sv.counts[b];
    const int32_t off = sv.offsets[b];
    if (cnt <= 0 || off < 0 || (int64_t)off + cnt > ps.n) continue;
    load_neighbours(nbr_tab, b, nbr, err);
    int32_t k0 = t.active[3 * (int64_t)b];
    // the bucket's grid spacing and time step (its scene's, in a batch)
    float dx = dx_arg, dt = dt_arg, D_inv = D_inv_arg;
    if constexpr (kMulti) {
      const int32_t sc = scene_of_key(k0, ss.n);
      k0 -= sc * kSceneKeyStride;
      dx = ss.p[sc].dx;
      dt = ss.p[sc].dt;
      const float dx_inv = (float)1 / dx;
      D_inv = 4.f * dx_inv * dx_inv;
    }
    const int32_t o0 = k0 * 4, o1 = t.active[3 * (int64_t)b + 1] * 4, o2 = t.active[3 * (int64_t)b + 2] * 4;
    // the wave's first 64 particle indices load while the velocities are staged
    int32_t first = -1;
    if (wave * kWave + lane < cnt) first = sv.indices[off + wave * kWave + lane];
    __syncthreads();
    // stage node velocities: slot s -> (neighbour block s>>6, cell s&63), stored at vel_at(r0, r1, r2)
    for (int s = threadIdx.x; s < kArena8; s += kT) {
      const int nbi = s >> 6, cell = s & 63;
      const int r0 = ((nbi >> 2) & 1) * 4 + ((cell >> 4) & 3);
      const int r1 = ((nbi >> 1) & 1) * 4 + ((cell >> 2) & 3);
      const int r2 = (nbi & 1) * 4 + (cell & 3);
      const int32_t gb = nbr[nbi];
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gb >= 0 && gb < grid_cap) {
        const float* gp = grid + grid_at(gb, 1, cell);
        v = make_float4(gp[0], gp[kBlockCells], gp[2 * kBlockCells], 0.f);
      }
      vel[vel_at(r0, r1, r2)] = v;
    }
    __syncthreads();
    // each wave takes 64 consecutive bucket slots at a time (wave-uniform loop)
    for (int32_t i0 = wave * kWave; i0 < cnt; i0 += kT) {
      const int32_t cw = min(kWave, cnt - i0);
      const bool valid = lane < cw;
      const int64_t p = valid ? (int64_t)(i0 == wave * kWave ? first : sv.indices[off + i0 + lane]) : -1;
      const bool bad = valid && (p < 0 || p >= ps.n);
      if (bad) atomicOr(err, kErrCapacity);
      const bool live = valid && !bad;
      int64_t p0;
      const bool contig = wave_contiguous(p, live, p0) && ballot(bad) == 0;
      float x[3] = {0.f, 0.f, 0.f}, F[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      // the attributes the reorder only moves, loaded with the rest
      float mcopy = 0.f, ljcopy = 0.f;
      int32_t idcopy = 0;
      if (kReorder && live) {
        mcopy = ps.m[p];
        if (ps.logJp) ljcopy = ns.lj_src ? ns.lj_src[p] : ps.logJp[p];
        idcopy = ps.id ? ps.id[p] : (int32_t)p;
      }
      if (contig) {
        RowLoad<3> lx;
        RowLoad<9> lf;
        lx.issue(ps.x, p0, cw);
        lf.issue(ps.F, p0, cw);
        lx.finish(stage, cw, x);
        lf.finish(stage, cw, F);
      } else if (live) {
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = ps.x[3 * p + d];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = ps.F[9 * p + k];
      }
      Arena ar;
      make_arena(x[0], x[1], x[2], dx, ar);
      if constexpr (kCell) {
#pragma clang fp contract(off)
#pragma unroll
        for (int d = 0; d < 3; ++d) cell_w(x[d] / dx - (float)ar.base[d], ar.w[d]);
      }
      const int lb0 = ar.base[0] - o0, lb1 = ar.base[1] - o1, lb2 = ar.base[2] - o2;
      const bool outside = live && (lb0 < 0 || lb0 > 5 || lb1 < 0 || lb1 > 5 || lb2 < 0 || lb2 > 5);
      if (outside) atomicOr(err, kErrArena);
      const bool ok = live && !outside;
      float v[3] = {0.f, 0.f, 0.f};
      float C[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float xn[3], Fn[9];
      if (ok) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float xa = (float)a * dx - ar.lp[0];
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) {
            const float xb = (float)bb * dx - ar.lp[1];
            const float wab = ar.w[0][a] * ar.w[1][bb];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float xc = (float)c * dx - ar.lp[2];
              const float W = wab * ar.w[2][c];
              const float4 vi = vel[vel_at(lb0 + a, lb1 + bb, lb2 + c)];
              v[0] += vi.x * W;
              v[1] += vi.y * W;
              v[2] += vi.z * W;
              // the cell kernel keeps B (no D_inv): C = B Dinv is formed below
              const float w0 = kCell ? W * vi.x : W * vi.x * D_inv;
              const float w1 = kCell ? W * vi.y : W * vi.y * D_inv;
              const float w2 = kCell ? W * vi.z : W * vi.z * D_inv;
              C[0] += w0 * xa;
              C[1] += w1 * xa;
              C[2] += w2 * xa;
              C[3] += w0 * xb;
              C[4] += w1 * xb;
              C[5] += w2 * xb;
              C[6] += w0 * xc;
              C[7] += w1 * xc;
              C[8] += w2 * xc;
            }
          }
        }
        // F <- (I + dt C) F, column-major (G2P.hpp:75-78); the cell kernel's
        // C is B scaled per column by Dinv (G2C2P.hpp:231-253)
        float Cs[9];
        if constexpr (kCell) {
#pragma clang fp contract(off)
          float Dinv[3];
          cell_dinv(x, dx, (float)1 / dx, Dinv);
#pragma unroll
          for (int k = 0; k < 9; ++k) Cs[k] = C[k] * Dinv[k / 3];
        } else {
#pragma unroll
          for (int k = 0; k < 9; ++k) Cs[k] = C[k];
        }
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float a0 = Cs[r] * dt + (r == 0 ? 1.f : 0.f);
            const float a1 = Cs[r + 3] * dt + (r == 1 ? 1.f : 0.f);
            const float a2 = Cs[r + 6] * dt + (r == 2 ? 1.f : 0.f);
            Fn[r + 3 * j] = a0 * F[3 * j] + a1 * F[3 * j + 1] + a2 * F[3 * j + 2];
          }
#pragma unroll
        for (int d = 0; d < 3; ++d) xn[d] = x[d] + v[d] * dt;
      } else {
        // a particle outside its bucket's arena keeps its state (the error is
        // reported at the next sync)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          xn[d] = x[d];
          v[d] = outside ? ps.v[3 * p + d] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          Fn[k] = F[k];
          C[k] = outside ? ps.C[9 * p + k] : 0.f;
        }
      }
      if (kReorder) {
        // destination slots off + i0 + lane are contiguous by construction
        const int64_t q = (int64_t)off + i0 + lane;
        if (live) {
          out.m[q] = mcopy;
          if (ps.logJp) out.logJp[q] = ljcopy;
          if (out.id) out.id[q] = idcopy;
        }
        wave_store_rows<3>(out.x, (int64_t)off + i0, cw, stage, xn);
        wave_store_rows<3>(out.v, (int64_t)off + i0, cw, stage, v);
        wave_store_rows<9>(out.C, (int64_t)off + i0, cw, stage, C);
        wave_store_rows<9>(out.F, (int64_t)off + i0, cw, stage, Fn);
        if constexpr (kNext >= 0) {
          // the next substep's P2G stress from F_new and the logJp this
          // substep leaves (P2G.hpp:57-100 on the stored F)
          float PF[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
          if (live) {
            bool f64 = false;
            if (kNext == 0) {
              stress_fixedcorotated(ns.m, Fn, PF, f64);
            } else {
              float lj = ljcopy;
              stress_sand(ns.m, Fn, lj, PF, nullptr, f64);
              ns.lj_next[q] = lj;
            }
            fp64_used += f64 ? 1u : 0u;
          }
          wave_store_rows<9>(ns.contrib, (int64_t)off + i0, cw, stage, PF);
        }
      } else if (contig && ballot(outside) == 0) {
        wave_store_rows<3>(ps.x, p0, cw, stage, xn);
        wave_store_rows<3>(ps.v, p0, cw, stage, v);
        wave_store_rows<9>(ps.C, p0, cw, stage, C);
        wave_store_rows<9>(ps.F, p0, cw, stage, Fn);
      } else if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          ps.x[3 * p + d] = xn[d];
          ps.v[3 * p + d] = v[d];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          ps.C[9 * p + k] = C[k];
          ps.F[9 * p + k] = Fn[k];
        }
      }
    }
    __syncthreads();
  }
  if (kNext >= 0 && fp64_used) atomicAdd(&counters[kCtrFp64], fp64_used);
}

// The gather half of G2PTransfer's implicit twin, G2P2GTransfer::operator()
// (simulation/transfer/G2P2G.hpp:49-84), per bucket as G2P: the bucket's
// 8^3 node velocities from the vector v (3 floats per node, node = block * 64
// + cell) are staged in LDS, and each particle takes C from its 27 staged
// nodes and writes F = (I + dt C) F_old to Fout[9 p]. The model's stress of
// that F is the ordinary stress pass (full occupancy; the SVD inside this
// kernel held it to three waves per SIMD), the force scatter P2G's.
__global__ __launch_bounds__(kG2PThreads) void g2p2g_gather_kernel(ParticlesView ps, TableView t, SortView sv,
                                                                   const int32_t* __restrict__ nbr_tab,
                                                                   const float* __restrict__ v, int32_t grid_cap,
                                                                   float dt, float dx, float* __restrict__ Fout,
                                                                   uint32_t* err) {
  __shared__ float4 vel[kVelNodes];
  __shared__ int32_t nbr[8];
  __shared__ float4 stage4[kG2PThreads / kWave][stage_floats(9) / 4];
  const int32_t nb = bounded_blocks(t, sv, grid_cap, err);
  const float dx_inv = (float)1 / dx;
  const float D_inv = 4.f * dx_inv * dx_inv;
  const int wave = uniform((int32_t)(threadIdx.x / kWave)), lane = lane_id();
  float* stage = reinterpret_cast<float*>(stage4[wave]);
  for (int32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const int32_t cnt = sv.counts[b];
    const int32_t off = sv.offsets[b];
    if (cnt <= 0 || off < 0 || (int64_t)off + cnt > ps.n) continue;
    if (threadIdx.x < 8) {
      const int32_t r = nbr_tab[8 * (int64_t)b + threadIdx.x];
      nbr[threadIdx.x] = r;
      if (r < 0) atomicOr(err, kErrMissingBlock);
    }
    const int32_t o0 = t.active[3 * (int64_t)b] * 4, o1 = t.active[3 * (int64_t)b + 1] * 4,
                  o2 = t.active[3 * (int64_t)b + 2] * 4;
    int32_t first = -1;
    if (wave * kWave + lane < cnt) first = sv.indices[off + wave * kWave + lane];
    __syncthreads();
    for (int s = threadIdx.x; s < kArena8; s += kG2PThreads) {
      const int nbi = s >> 6, cell = s & 63;
      const int r0 = ((nbi >> 2) & 1) * 4 + ((cell >> 4) & 3);
      const int r1 = ((nbi >> 1) & 1) * 4 + ((cell >> 2) & 3);
      const int r2 = (nbi & 1) * 4 + (cell & 3);
      const int32_t gb = nbr[nbi];
      float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gb >= 0 && gb < grid_cap) {
        const float* vp = v + 3 * ((int64_t)gb * kBlockCells + cell);
        w = make_float4(vp[0], vp[1], vp[2], 0.f);
      }
      vel[vel_at(r0, r1, r2)] = w;
    }
    __syncthreads();
    for (int32_t i0 = wave * kWave; i0 < cnt; i0 += kG2PThreads) {
      const int32_t cw = min(kWave, cnt - i0);
      const bool valid = lane < cw;
      const int64_t p = valid ? (int64_t)(i0 == wave * kWave ? first : sv.indices[off + i0 + lane]) : -1;
      const bool bad = valid && (p < 0 || p >= ps.n);
      if (bad) atomicOr(err, kErrCapacity);
      const bool live = valid && !bad;
      int64_t p0;
      const bool contig = wave_contiguous(p, live, p0) && ballot(bad) == 0;
      float x[3] = {0.f, 0.f, 0.f}, F[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (contig) {
        RowLoad<3> lx;
        RowLoad<9> lf;
        lx.issue(ps.x, p0, cw);
        lf.issue(ps.F, p0, cw);
        lx.finish(stage, cw, x);
        lf.finish(stage, cw, F);
      } else if (live) {
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = ps.x[3 * p + d];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = ps.F[9 * p + k];
      }
      Arena ar;
      make_arena(x[0], x[1], x[2], dx, ar);
      const int lb0 = ar.base[0] - o0, lb1 = ar.base[1] - o1, lb2 = ar.base[2] - o2;
      const bool outside = live && (lb0 < 0 || lb0 > 5 || lb1 < 0 || lb1 > 5 || lb2 < 0 || lb2 > 5);
      if (outside) atomicOr(err, kErrArena);
      const bool ok = live && !outside;
      float Fn[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) Fn[k] = F[k];
      if (ok) {
        float C[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float xa = (float)a * dx - ar.lp[0];
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) {
            const float xb = (float)bb * dx - ar.lp[1];
            const float wab = ar.w[0][a] * ar.w[1][bb];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float xc = (float)c * dx - ar.lp[2];
              const float W = wab * ar.w[2][c];
              const float4 vi = vel[vel_at(lb0 + a, lb1 + bb, lb2 + c)];
              const float w0 = W * vi.x * D_inv, w1 = W * vi.y * D_inv, w2 = W * vi.z * D_inv;
              C[0] += w0 * xa;
              C[1] += w1 * xa;
              C[2] += w2 * xa;
              C[3] += w0 * xb;
              C[4] += w1 * xb;
              C[5] += w2 * xb;
              C[6] += w0 * xc;
              C[7] += w1 * xc;
              C[8] += w2 * xc;
            }
          }
        }
        // F = (I + dt C) F_old, column-major (G2P2G.hpp:81-84)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float a0 = C[r] * dt + (r == 0 ? 1.f : 0.f);
            const float a1 = C[r + 3] * dt + (r == 1 ? 1.f : 0.f);
            const float a2 = C[r + 6] * dt + (r == 2 ? 1.f : 0.f);
            Fn[r + 3 * j] = a0 * F[3 * j] + a1 * F[3 * j + 1] + a2 * F[3 * j + 2];
          }
      }
      if (contig && ballot(outside) == 0) {
        wave_store_rows<9>(Fout, p0, cw, stage, Fn);
      } else if (live) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Fout[9 * p + k] = Fn[k];
      }
    }
    __syncthreads();
  }
}

// Vector::reorder gather (container/Vector.hpp:617-646) of every attribute
__global__ __launch_bounds__(kThreads) void reorder_kernel(ParticlesView in, ParticlesView out,
                                                           const int32_t* __restrict__ map, uint32_t* err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < out.n; i += stride) {
    const int64_t p = map[i];
    if (p < 0 || p >= in.n) {
      atomicOr(err, kErrCapacity);
      continue;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      out.x[3 * i + d] = in.x[3 * p + d];
      out.v[3 * i + d] = in.v[3 * p + d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      out.C[9 * i + k] = in.C[9 * p + k];
      out.F[9 * i + k] = in.F[9 * p + k];
    }
    out.m[i] = in.m[p];
    if (in.logJp) out.logJp[i] = in.logJp[p];
    if (out.id) out.id[i] = in.id ? in.id[p] : (int32_t)p;
  }
}

// ImplicitMPMSystem::Projector (simulation/mpm/ImplicitMPM.hpp:65-101) for
// every boundary at once: a node with mass resolves its vector against the
// plane colliders, the general colliders and the level-set boundary in that
// order; a node without mass is cleared.
__global__ __launch_bounds__(kThreads) void implicit_project_kernel(const float* __restrict__ grid, int32_t grid_cap,
                                                                    TableView t, float dx, Colliders cols,
                                                                    GeneralColliders gen, float* dof) {
#pragma clang fp contract(off)
  int32_t nb = *t.cnt;
  if (nb > grid_cap) nb = grid_cap;
  const int64_t nodes = (int64_t)nb * kBlockCells;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes; i += stride) {
    const int32_t b = (int32_t)(i >> 6);
    const int cell = (int)(i & 63);
    float* dv = dof + 3 * i;
    if (!(grid[grid_at(b, 0, cell)] > 0.f)) {
      dv[0] = dv[1] = dv[2] = 0.f;
      continue;
    }
    const int32_t k0 = t.active[3 * (int64_t)b], k1 = t.active[3 * (int64_t)b + 1], k2 = t.active[3 * (int64_t)b + 2];
    const float pos[3] = {((float)k0 * 4.f + (float)((cell >> 4) & 3)) * dx,
                          ((float)k1 * 4.f + (float)((cell >> 2) & 3)) * dx,
                          ((float)k2 * 4.f + (float)(cell & 3)) * dx};
    float v[3] = {dv[0], dv[1], dv[2]};
    for (int c = 0; c < cols.n; ++c) {
      const PlaneCollider& pc = cols.c[c];
      const float phi = pc.normal[0] * (pos[0] - pc.origin[0]) + pc.normal[1] * (pos[1] - pc.origin[1]) +
                        pc.normal[2] * (pos[2] - pc.origin[2]);
      if (phi < 0.f) {
        if (pc.type == 0) {
          v[0] = v[1] = v[2] = 0.f;
        } else {
          const float proj = pc.normal[0] * v[0] + pc.normal[1] * v[1] + pc.normal[2] * v[2];
          if ((pc.type == 2 && proj < 0.f) || pc.type == 1) {
            v[0] -= proj * pc.normal[0];
            v[1] -= proj * pc.normal[1];
            v[2] -= proj * pc.normal[2];
          }
        }
      }
    }
    for (int c = 0; c < gen.n; ++c) resolve_general(gen.c[c], pos, v);
    if (gen.ls.on) levelset_resolve(gen.ls, pos, v);
    dv[0] = v[0];
    dv[1] = v[1];
    dv[2] = v[2];
  }
}

}  // namespace

hipError_t implicit_project(const float* grid, int32_t grid_cap, const TableView& t, float dx, const Colliders& cols,
                            const GeneralColliders& gen, float* dof, hipStream_t s) {
  if (cols.n == 0 && gen.n == 0 && !gen.ls.on) return hipSuccess;  // no boundary: nothing changes
  implicit_project_kernel<<<persistent_groups(), kThreads, 0, s>>>(grid, grid_cap, t, dx, cols, gen, dof);
  return hipGetLastError();
}

// Dynamic LDS the stress pass reserves (and does not use) when it runs beside
// the sparsity stages: with its 9,280 bytes of static LDS, three workgroups
// fit in a CU's 160 KiB and a fourth does not. Uncapped, its workgroups (120
// VGPRs, four waves per SIMD) fill the register files and the sparsity
// kernels wait for CUs; capped, both finish earlier (measured on the 16M
// sand column: 2.52 -> 2.46 ms per substep; a cap of two is slower again).
constexpr size_t kStressLdsReserve = 40960;

hipError_t check_launch_limits(int device) {
  int max_lds = 0;
  hipError_t e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  if (e != hipSuccess) return e;
  const void* kernels[] = {
      reinterpret_cast<const void*>(&stress_kernel<0>), reinterpret_cast<const void*>(&stress_kernel<1>),
      reinterpret_cast<const void*>(&stress_kernel<2>), reinterpret_cast<const void*>(&p2g_kernel<0, false>),
      reinterpret_cast<const void*>(&p2g_kernel<1, false>), reinterpret_cast<const void*>(&p2g_kernel<0, true>),
      reinterpret_cast<const void*>(&p2g_kernel<1, true>),
      reinterpret_cast<const void*>(&g2p_kernel<false, -1>), reinterpret_cast<const void*>(&g2p_kernel<true, -1>),
      reinterpret_cast<const void*>(&g2p_kernel<true, 0>), reinterpret_cast<const void*>(&g2p_kernel<true, 1>),
      reinterpret_cast<const void*>(&g2p_kernel<true, -1, true>),
      reinterpret_cast<const void*>(&g2p_kernel<false, -1, true>),
      reinterpret_cast<const void*>(&p2g_kernel<0, false, true>),
      reinterpret_cast<const void*>(&g2p_kernel<false, -1, false, true>),
      reinterpret_cast<const void*>(&p2g_kernel<0, false, false, 1>),
      reinterpret_cast<const void*>(&p2g_kernel<0, true, false, 1>),
      reinterpret_cast<const void*>(&g2p_kernel<true, -1, false, false, 256>),
      reinterpret_cast<const void*>(&g2p_kernel<true, -1, true, false, 256>),
      reinterpret_cast<const void*>(&g2p2g_gather_kernel),
      reinterpret_cast<const void*>(&p2g_kernel<2, false>),
      reinterpret_cast<const void*>(&p2g_kernel<2, true>),
      reinterpret_cast<const void*>(&g2p_kernel<true, -1, false, true>)};
  for (int i = 0; i < (int)(sizeof(kernels) / sizeof(kernels[0])); ++i) {
    hipFuncAttributes attr;
    e = hipFuncGetAttributes(&attr, kernels[i]);
    if (e != hipSuccess) return e;
    const size_t extra = i < 3 ? kStressLdsReserve : 0;  // the stress kernels' reserve
    if (attr.sharedSizeBytes + extra > (size_t)max_lds) return hipErrorInvalidConfiguration;
  }
  return closure_launch_limits();
}

hipError_t grid_clear(float* grid, const int32_t* nblocks_dev, int32_t cap_blocks, hipStream_t s,
                      uint32_t* counters) {
  const int64_t total = (int64_t)cap_blocks * kChannels * kBlockCells / 4;
  grid_clear_kernel<<<groups_for(total, kThreads), kThreads, 0, s>>>(grid, nblocks_dev, cap_blocks, counters);
  return hipGetLastError();
}

hipError_t stress(const Model& m, const ParticlesView& ps, Workspace& ws, hipStream_t s, const ClosureView* closure,
                  float* contrib, bool beside_sparsity, const SceneSet* ss) {
  if (ps.n <= 0) return hipSuccess;
  float* out = contrib ? contrib : ws.contrib;
  const int64_t waves = (ps.n + kWave - 1) / kWave;
  const int g = groups_for(waves * kWave, kThreads);
  const size_t lds = beside_sparsity ? kStressLdsReserve : 0;
  if (ss && ss->n > 0) {
    if (closure) return hipErrorNotSupported;
    stress_kernel<2><<<g, kThreads, lds, s>>>(m, ps, out, ws.counters, *ss);
    return hipGetLastError();
  }
  if (closure && m.kind == 1 && m.mu != 0.f) return stress_closure(m, *closure, ps, out, ws, s);
  if (m.kind == 0)
    stress_kernel<0><<<g, kThreads, lds, s>>>(m, ps, out, ws.counters, SceneSet{});
  else
    stress_kernel<1><<<g, kThreads, lds, s>>>(m, ps, out, ws.counters, SceneSet{});
  return hipGetLastError();
}

hipError_t p2g_scatter(const ParticlesView& ps, const TableView& t, const SortView& sv, float* grid,
                       int32_t grid_cap, float dt, float dx, bool fold, Workspace& ws, hipStream_t s,
                       bool reset_queue, const SceneSet* ss) {
  const int g = ws.grid_p2g > 0 ? ws.grid_p2g : persistent_groups();
  if (reset_queue) reset_queue_kernel<<<1, 1, 0, s>>>(ws.counters);
  uint32_t* q = ws.counters + kCtrQueue;
  const bool multi = ss && ss->n > 0;
  const SceneSet sset = multi ? *ss : SceneSet{};
  const bool tile1 = ws.p2g_tile == 1;  // the folded launches take the tuned tile
  if (fold && ws.cell && !multi)  // the substep's P2C2G (one tile)
    p2g_kernel<0, false, true><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx,
                                                         ws.err, q, sset, 1);
  else if (fold && multi && tile1)
    p2g_kernel<0, true, false, 1><<<g, p2g_threads(1), 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx,
                                                               ws.err, q, sset, 1);
  else if (fold && multi)
    p2g_kernel<0, true><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err, q,
                                                  sset, 1);
  else if (fold && tile1)
    p2g_kernel<0, false, false, 1><<<g, p2g_threads(1), 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt,
                                                                dx, ws.err, q, sset, 1);
  else if (fold)
    p2g_kernel<0, false><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err, q,
                                                   sset, 1);
  else if (multi) {
    // the force kept in channels 4..6: momentum, then the force
    p2g_kernel<1, true><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err, q,
                                                  sset, 1);
    reset_queue_kernel<<<1, 1, 0, s>>>(ws.counters);
    p2g_kernel<2, true><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err, q,
                                                  sset, 4);
  } else {
    p2g_kernel<1, false><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err, q,
                                                   sset, 1);
    reset_queue_kernel<<<1, 1, 0, s>>>(ws.counters);
    p2g_kernel<2, false><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err, q,
                                                   sset, 4);
  }
  return hipGetLastError();
}

hipError_t p2c2g_scatter(const ParticlesView& ps, const TableView& t, const SortView& sv, float* grid,
                         int32_t grid_cap, float dt, float dx, Workspace& ws, hipStream_t s) {
  const int g = ws.grid_p2g > 0 ? ws.grid_p2g : persistent_groups();
  reset_queue_kernel<<<1, 1, 0, s>>>(ws.counters);
  p2g_kernel<0, false, true><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx,
                                                       ws.err, ws.counters + kCtrQueue, SceneSet{}, 1);
  return hipGetLastError();
}

hipError_t p2g_force_scatter(const ParticlesView& ps, const TableView& t, const SortView& sv, float* grid,
                             int32_t grid_cap, float dt, float dx, Workspace& ws, hipStream_t s) {
  const int g = ws.grid_p2g > 0 ? ws.grid_p2g : persistent_groups();
  reset_queue_kernel<<<1, 1, 0, s>>>(ws.counters);
  p2g_kernel<2, false><<<g, kP2GThreads, 0, s>>>(ps, ws.contrib, t, sv, ws.nbr, grid, grid_cap, dt, dx, ws.err,
                                                 ws.counters + kCtrQueue, SceneSet{}, 1);
  return hipGetLastError();
}

hipError_t g2p2g_gather(const ParticlesView& ps, const TableView& t, const SortView& sv, const float* v,
                        int32_t grid_cap, float dt, float dx, float* Fout, Workspace& ws, hipStream_t s) {
  if (ps.n == 0) return hipSuccess;
  const int g = ws.grid_g2p > 0 ? ws.grid_g2p : persistent_groups();
  g2p2g_gather_kernel<<<g, kG2PThreads, 0, s>>>(ps, t, sv, ws.nbr, v, grid_cap, dt, dx, Fout, ws.err);
  return hipGetLastError();
}

hipError_t g2c2p(const ParticlesView& ps, const TableView& t, const SortView& sv, const float* grid, int32_t grid_cap,
                 float dt, float dx, Workspace& ws, hipStream_t s) {
  const int g = ws.grid_g2p > 0 ? ws.grid_g2p : persistent_groups();
  g2p_kernel<false, -1, false, true><<<g, kG2PThreads, 0, s>>>(ps, ps, t, sv, ws.nbr, grid, grid_cap, dt, dx,
                                                               NextStress{}, ws.err, ws.counters, SceneSet{});
  return hipGetLastError();
}

hipError_t grid_update(float* grid, int32_t grid_cap, const TableView& t, float dt, const float extf[3],
                       const Colliders& cols, float dx, int fold_rhs, Workspace& ws, hipStream_t s,
                       const SceneSet* ss, const GeneralColliders* gen) {
  const GeneralColliders g = gen ? *gen : GeneralColliders{};
  if (ss && ss->n > 0)
    grid_update_kernel<true><<<persistent_groups(), kThreads, 0, s>>>(grid, grid_cap, t, dt, extf[0], extf[1],
                                                                     extf[2], cols, dx, fold_rhs, ws.counters,
                                                                     ws.err, *ss, g);
  else
    grid_update_kernel<false><<<persistent_groups(), kThreads, 0, s>>>(grid, grid_cap, t, dt, extf[0], extf[1],
                                                                      extf[2], cols, dx, fold_rhs, ws.counters,
                                                                      ws.err, SceneSet{}, g);
  return hipGetLastError();
}

hipError_t grid_angular_momentum(const float* grid, int32_t grid_cap, const TableView& t, float dx, int mchn,
                                 int mvchn, double* sum, Workspace& ws, hipStream_t s) {
  angular_momentum_kernel<<<1024, kThreads, 0, s>>>(grid, grid_cap, t, dx, mchn, mvchn, sum, ws.err);
  return hipGetLastError();
}

hipError_t g2p(const ParticlesView& ps, const ParticlesView* out, const TableView& t, const SortView& sv,
               const float* grid, int32_t grid_cap, float dt, float dx, Workspace& ws, hipStream_t s,
               const NextStress* next, const SceneSet* ss) {
  const int g = ws.grid_g2p > 0 ? ws.grid_g2p : persistent_groups();
  NextStress ns{};
  if (next) ns = *next;
  const SceneSet sset = ss ? *ss : SceneSet{};
  if (ws.cell) {  // the substep's G2C2P (no batch, no fused stress; one tile)
    if ((next && next->contrib) || sset.n > 0) return hipErrorInvalidValue;
    if (out)
      g2p_kernel<true, -1, false, true><<<g, kG2PThreads, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx,
                                                                  ns, ws.err, ws.counters, sset);
    else
      g2p_kernel<false, -1, false, true><<<g, kG2PThreads, 0, s>>>(ps, ps, t, sv, ws.nbr, grid, grid_cap, dt, dx,
                                                                   ns, ws.err, ws.counters, sset);
  } else if (next && next->contrib) {
    if (!out || sset.n > 0) return hipErrorInvalidValue;
    if (ns.m.kind == 0)
      g2p_kernel<true, 0><<<g, kG2PThreads, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns, ws.err,
                                                    ws.counters, sset);
    else
      g2p_kernel<true, 1><<<g, kG2PThreads, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns, ws.err,
                                                    ws.counters, sset);
  } else if (out && sset.n > 0 && ws.g2p_threads == 256) {
    g2p_kernel<true, -1, true, false, 256><<<g, 256, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns,
                                                             ws.err, ws.counters, sset);
  } else if (out && sset.n > 0) {
    g2p_kernel<true, -1, true><<<g, kG2PThreads, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns, ws.err,
                                                         ws.counters, sset);
  } else if (out && ws.g2p_threads == 256) {
    g2p_kernel<true, -1, false, false, 256><<<g, 256, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns,
                                                              ws.err, ws.counters, sset);
  } else if (out) {
    g2p_kernel<true, -1><<<g, kG2PThreads, 0, s>>>(ps, *out, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns, ws.err,
                                                   ws.counters, sset);
  } else if (sset.n > 0) {
    g2p_kernel<false, -1, true><<<g, kG2PThreads, 0, s>>>(ps, ps, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns, ws.err,
                                                          ws.counters, sset);
  } else {
    g2p_kernel<false, -1><<<g, kG2PThreads, 0, s>>>(ps, ps, t, sv, ws.nbr, grid, grid_cap, dt, dx, ns, ws.err,
                                                    ws.counters, sset);
  }
  return hipGetLastError();
}

hipError_t reorder_particles(const ParticlesView& in, const ParticlesView& out, const int32_t* map,
                             Workspace& ws, hipStream_t s) {
  if (out.n <= 0) return hipSuccess;
  reorder_kernel<<<groups_for(out.n, kThreads), kThreads, 0, s>>>(in, out, map, ws.err);
  return hipGetLastError();
}

}  // namespace zsmpm
