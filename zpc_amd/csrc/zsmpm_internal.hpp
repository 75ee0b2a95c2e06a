// Host-side declarations shared by the zsmpm HIP translation units: non-owning
// device views in zpc's container layouts, the scratch workspace a policy
// owns, and the stage launchers the C ABI strings together.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/zsmpm.h"
#include "zsmpm_device.hpp"
#include "bht_device.hpp"

namespace zsmpm {

/// HashTable<i32,3,int> arrays (container/HashTable.hpp:40-58, 200-205).
struct TableView {
  int32_t T = 0;
  int32_t* keys = nullptr;     // [3T]
  int32_t* indices = nullptr;  // [T]
  int32_t* status = nullptr;   // [T]
  int32_t* cnt = nullptr;      // [1]
  int32_t* active = nullptr;   // [3T] _activeKeys
};

/// Particles<f32,3> attributes (one array per attribute, C and F column-major).
struct ParticlesView {
  int64_t n = 0;
  float* x = nullptr;
  float* v = nullptr;
  float* C = nullptr;
  float* F = nullptr;
  float* m = nullptr;
  float* logJp = nullptr;
  int32_t* id = nullptr;  // optional particle ids, carried by reordering
};

/// Bucket-sort outputs (IndexBuckets, container/IndexBuckets.hpp:11-66).
struct SortView {
  int32_t* counts = nullptr;   // [B + 1]
  int32_t* offsets = nullptr;  // [B + 1]
  int32_t* indices = nullptr;  // [N]
  int32_t cap_blocks = 0;      // capacity of counts/offsets minus one
};

struct PlaneCollider {
  int32_t type;  // 0 sticky, 1 slip, 2 separate
  float origin[3];
  float normal[3];
};
constexpr int kMaxColliders = 8;
struct Colliders {
  int32_t n = 0;
  PlaneCollider c[kMaxColliders];
};

// Collider<AnalyticLevelSet<geometry>> with its rigid motion (zsmpm_collider).
struct GeneralCollider {
  int32_t type, geometry, axis;
  float p0[3], p1[3];
  float R[9];
  float omega[3];
  float s, dsdt;
  float b[3], dbdt[3];
};
struct GeneralColliders {
  int32_t n = 0;
  GeneralCollider c[kMaxColliders];
  LevelSetBoundary ls;  // a SparseGrid signed-distance boundary, applied after the analytic ones
};

// A batch of independent scenes in one table, one grid and one set of
// particle arrays. Scene s owns the particles [first[s], first[s + 1]); its
// block keys are shifted by s * kSceneKeyStride along axis 0, so no block is
// shared between scenes and a block's scene follows from its key. Each scene
// has its own model, dx, dt, gravity and colliders. The sorted driver keeps
// every scene's particles in its own range: blocks are numbered in order of
// first occurrence, so all of scene s's buckets follow scene s - 1's.
constexpr int32_t kSceneKeyShift = 10;
constexpr int32_t kSceneKeyStride = 1 << kSceneKeyShift;
// a scene's own particle keys lie in [-512, 511) along axis 0, its enlarged
// keys in [-512, 512)
constexpr int32_t kSceneKeyHalf = kSceneKeyStride / 2;
constexpr int32_t kMaxScenes = 256;

struct SceneParams {
  Model model;
  float dx, dx_inv, dt;
  float g[3];
  Colliders cols;
};

struct SceneSet {
  int32_t n = 0;                   // 0: one scene, parameters from the launch arguments
  const SceneParams* p = nullptr;  // [n] device
  const int64_t* first = nullptr;  // [n + 1] device: first[0] = 0, first[n] = particle count
};

// scene of a block from the first component of its key, clamped to the
// batch's n scenes
__device__ __forceinline__ int32_t scene_of_key(int32_t k0, int32_t n) {
  return min(max((k0 + kSceneKeyHalf) >> kSceneKeyShift, 0), n - 1);
}

// scene of particle p: the last s with first[s] <= p
__device__ __forceinline__ int32_t scene_of_particle(const int64_t* first, int32_t n, int64_t p) {
  int32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (first[mid] <= p) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// counters[] slots in the workspace
enum : int {
  kCtrList = 0,     // unique new keys found by a dedup pass
  kCtrCnt0 = 1,     // table size before the current insertion pass
  kCtrRuns = 2,     // particle runs in the bucket sort
  kCtrFp64 = 3,     // particles that took the FP64 SVD path
  kCtrFallback = 4, // particles that took the SVD fallback of the learned closure
  kCtrMaxVel = 5,   // max |v|^2 as float bits
  kCtrQueue = 6,    // next bucket of the running per-bucket launch
  kCtrFbQueue = 7,  // length of the closure's SVD fallback queue (per pass)
  kCtrNetQueue = 8, // waves the closure's network pass took (per pass)
  kCtrNetWaves = 9, // waves that evaluated the closure's network (all passes)
  kCtrBhtRounds = 10, // placement rounds of the last bht insertion (host-side, filled by the counters call)
  kCtrCount = 16
};

// learned closure network: 9 inputs, 64 x 64 hidden (GELU), 7 outputs;
// training records of 10 floats (eigen-space targets)
constexpr int kMlpIn = 9, kMlpOut = 7, kMlpRec = 10;
constexpr int kMlpParams = 9 * 64 + 64 + 64 * 64 + 64 + 64 * 7 + 7;  // 5255
constexpr int kMlpGroups = 512;  // workgroups (and gradient partials) of one training pass

// The closure a stress pass uses for Drucker-Prager particles: the network
// weights, or null for Newton from X0 = 0, and the residual tolerance.
struct ClosureView {
  const float* params = nullptr;
  float tol = 1e-4f;
  int32_t always_net = 0;  // evaluate the network even where X0 = 0 certifies a priori
};

/// Scratch device memory owned by a policy (the get_temporary_memory_source
/// counterpart). Sized by reserve(); kernels read live counts from `counters`.
struct Workspace {
  uint32_t n_cap = 0;       // particles
  uint32_t ts_cap = 0;      // dedup table slots (power of two)
  uint64_t* skey = nullptr;     // [ts_cap] packed keys, ~0 = empty
  uint32_t* sorder = nullptr;   // [ts_cap] first-occurrence order (min)
  uint32_t* la_key = nullptr;   // [list_cap] sort buffers for unique keys
  uint32_t* la_val = nullptr;
  uint32_t* lb_key = nullptr;
  uint32_t* lb_val = nullptr;
  uint32_t list_cap = 0;
  uint32_t* bkt = nullptr;      // [n_cap] bucket of each particle
  uint32_t* run_start = nullptr;  // [n_cap]
  uint32_t* rk_a = nullptr;     // [n_cap] run sort keys/vals (ping-pong)
  uint32_t* rv_a = nullptr;
  uint32_t* rk_b = nullptr;
  uint32_t* rv_b = nullptr;
  uint32_t* run_len = nullptr;  // [n_cap]
  uint32_t* run_dest = nullptr; // [n_cap]
  uint32_t* tile_heads = nullptr;  // [n_cap / 4096 + 1]
  uint32_t* tile_base = nullptr;
  uint32_t* rs_scratch = nullptr;
  size_t rs_words = 0;
  uint32_t* scan_tmp = nullptr;  // tile sums for scans
  size_t scan_words = 0;
  uint32_t* counters = nullptr;  // [kCtrCount]
  uint32_t* err = nullptr;       // [1] error bits
  float* contrib = nullptr;      // [9 n_cap] vol P F^T of each particle (column-major)
  int32_t* nbr = nullptr;        // [8 ts_cap] block index of key_b + (d>>2&1, d>>1&1, d&1), or -1
  int32_t* fb = nullptr;         // [n_cap] particles queued for the closure's SVD fallback
  uint32_t* wq = nullptr;        // [n_cap / 64 + 2] per wave: 1 where the closure's network must run
  int32_t grid_p2g = 0;          // P2G / G2P workgroups (0: persistent_groups()), set by the policy
  int32_t grid_g2p = 0;
  int32_t p2g_tile = 0;     // folded P2G launch tile: 0 four waves / 576-particle chunks, 1 two waves / 256
  int32_t g2p_threads = 0;  // sorted G2P threads per bucket: 256, else 128
  bool cell = false;        // the substep's transfers are P2C2G / G2C2P (set by the substep driver only)
  float* lj_next[2] = {nullptr, nullptr};  // [n_cap] each: logJp caches of the fused multi-substep driver
  uint32_t* bitmap = nullptr;    // [bm_words] one bit per insertion order: the orders of new keys
  uint32_t* bm_prefix = nullptr; // [bm_words] exclusive prefix of the bit counts
  uint32_t bm_words = 0;
  uint32_t* pslot = nullptr;     // [n_cap] scratch slot of each particle's key (recording activation)
  int32_t* slot_block = nullptr; // [ts_cap] block number of the key in each scratch slot
  uint32_t* dmask = nullptr;     // [ts_cap / 32 + 1] one bit per dedup slot claimed by the current pass
};

// The next substep's constitutive part computed inside G2P from the updated
// F (the sorted multi-substep driver): vol P F^T and the return-mapped logJp
// of every particle, written in output slot order.
struct NextStress {
  Model m;
  const float* lj_src = nullptr;  // logJp carried to the output (the previous G2P's cache), or null: ps.logJp
  float* contrib = nullptr;       // [9 n] next substep's vol P F^T; null: not computed (last substep)
  float* lj_next = nullptr;       // [n] next substep's logJp after the return mapping
};

// Next slot of zpc's probe sequence, (s + 127) % T (HashTable.hpp:362), for
// any table size T >= 1 and s in [0, T).
__device__ __forceinline__ int32_t probe_next(int32_t s, int32_t T) {
  return (int32_t)(((uint32_t)s + (uint32_t)kProbeStep) % (uint32_t)T);
}

// HashTableView::query (HashTable.hpp:447-456) with the modular wrap; at most
// T probes.
__device__ __forceinline__ int32_t table_find(const TableView& t, int32_t k0, int32_t k1, int32_t k2) {
  int32_t s = zpc_home(k0, k1, k2, t.T);
  for (int32_t probe = 0; probe < t.T; ++probe) {
    const int32_t* kp = t.keys + 3 * (int64_t)s;
    if (kp[0] == k0 && kp[1] == k1 && kp[2] == k2) return t.indices[s];
    if (t.indices[s] == -1) return -1;
    s = probe_next(s, t.T);
  }
  return -1;
}

// ---------------------------------------------------------------- primitives
size_t scan_scratch_words();
size_t radix_sort_scratch_words(uint32_t n_cap);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_cap,
                              uint32_t* tile_sums, uint32_t* total_out, hipStream_t s);
// exclusive scan of the population counts of n words
hipError_t exclusive_scan_popc_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                                   hipStream_t s);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                            const uint32_t* n_dev, uint32_t n_cap, int bits, uint32_t* scratch,
                            hipStream_t s, bool* result_in_alt);

// ---------------------------------------------------------------- sparsity
// zero[0, nzero) is cleared by the same launch (the substep's bucket counts)
hipError_t table_clear(const TableView& t, Workspace& ws, hipStream_t s, uint32_t* zero = nullptr,
                       int64_t nzero = 0);
// record_slots: also fill ws.pslot / ws.slot_block for bucket_sort_slots
// (the table must be empty: every key is then new)
// ss (a batch): each particle's key uses its scene's dx and is shifted by the
// scene's key offset
hipError_t table_activate(const TableView& t, Workspace& ws, const float* x, int64_t n, float dx,
                          int blockLen, int offset, float disp, hipStream_t s, bool record_slots = false,
                          const SceneSet* ss = nullptr);
hipError_t table_enlarge(const TableView& t, Workspace& ws, const int32_t lo[3], const int32_t hi[3],
                         int32_t n0_cap, hipStream_t s);
hipError_t table_query(const TableView& t, const int32_t* keys, int64_t n, int32_t* out, hipStream_t s);
hipError_t bucket_sort(const TableView& t, Workspace& ws, const float* x, int64_t n, float dx,
                       int blockLen, int offset, float disp, const SortView& out, hipStream_t s);
// bucket_sort of the particles a recording activation just inserted into an
// emptied table (blockLen 4, offset -2, displacement 0.5 as that activation)
// counts_zeroed: out.counts[0, cap_blocks] is already zero (table_clear did it)
hipError_t bucket_sort_slots(Workspace& ws, int64_t n, const SortView& out, hipStream_t s,
                             bool counts_zeroed = false);
// ws.nbr for every active block (the 2x2x2 blocks a bucket's stencils cover)
hipError_t block_neighbours(const TableView& t, Workspace& ws, hipStream_t s);

// ---------------------------------------------------------------- transfer
// every transfer kernel's static LDS fits the device's per-workgroup limit
hipError_t check_launch_limits(int device);
// with counters: also resets the P2G bucket queue and the max |v|^2 counter
// (then p2g_scatter is called with reset_queue = false)
hipError_t grid_clear(float* grid, const int32_t* nblocks_dev, int32_t cap_blocks, hipStream_t s,
                      uint32_t* counters = nullptr);
// P2G = stress (per particle, writes ws.contrib and logJp) + scatter (per bucket).
// beside_sparsity: the pass runs on a second stream next to the sparsity
// stages, and leaves room on every CU for them (at most three of its
// workgroups per CU)
// ss (a batch): each particle uses its scene's model (m is not read; no closure)
hipError_t stress(const Model& m, const ParticlesView& ps, Workspace& ws, hipStream_t s,
                  const ClosureView* closure = nullptr, float* contrib = nullptr, bool beside_sparsity = false,
                  const SceneSet* ss = nullptr);

// ---------------------------------------------------------------- closure
// stress pass of the learned closure into contrib (SVD fallback queued)
hipError_t stress_closure(const Model& m, const ClosureView& cl, const ParticlesView& ps, float* contrib,
                          Workspace& ws, hipStream_t s);
hipError_t mlp_infer(const float* P, const float* X, int64_t n, float* Y, hipStream_t s);
hipError_t mlp_grad(const float* P, const float* X, const float* R, int64_t n, float* partial, float* loss_part,
                    int G, hipStream_t s);
hipError_t mlp_reduce(float* P, float* m, float* v, const float* partial, const float* loss_part, int G, float lr,
                      float bc1, float bc2, float* grad_out, float* loss_out, hipStream_t s);
hipError_t mlp_harvest(const Model& m, const float* F, const float* logJp, int64_t n, float* X, float* R,
                       hipStream_t s);
hipError_t closure_launch_limits();
// ss (a batch, here and below): dx, dt, gravity and colliders of each
// bucket's or node's scene, read from its block key
hipError_t p2g_scatter(const ParticlesView& ps, const TableView& t, const SortView& sv, float* grid,
                       int32_t grid_cap, float dt, float dx, bool fold, Workspace& ws, hipStream_t s,
                       bool reset_queue = true, const SceneSet* ss = nullptr);
// P2C2GTransfer as a scatter of each particle's cell-kernel stencil (folded
// into m and mv, channels 0..3) with the stress already in ws.contrib
hipError_t p2c2g_scatter(const ParticlesView& ps, const TableView& t, const SortView& sv, float* grid,
                         int32_t grid_cap, float dt, float dx, Workspace& ws, hipStream_t s);
// G2C2PTransfer with its pre and post passes, in place: v, B (in C), F, x
hipError_t g2c2p(const ParticlesView& ps, const TableView& t, const SortView& sv, const float* grid, int32_t grid_cap,
                 float dt, float dx, Workspace& ws, hipStream_t s);
// gen: general colliders applied after the plane colliders (nullptr: none)
hipError_t grid_update(float* grid, int32_t grid_cap, const TableView& t, float dt, const float extf[3],
                       const Colliders& cols, float dx, int fold_rhs, Workspace& ws, hipStream_t s,
                       const SceneSet* ss = nullptr, const GeneralColliders* gen = nullptr);
// GridAngularMomentum: sum[0..2] += x cross mv, sum[3..5] += mv over nodes with mass
hipError_t grid_angular_momentum(const float* grid, int32_t grid_cap, const TableView& t, float dx, int mchn,
                                 int mvchn, double* sum, Workspace& ws, hipStream_t s);
// out == nullptr updates ps in place; otherwise the updated particles (and
// m, logJp, id) are written to out in bucket order (slot = sorted position)
hipError_t g2p(const ParticlesView& ps, const ParticlesView* out, const TableView& t, const SortView& sv,
               const float* grid, int32_t grid_cap, float dt, float dx, Workspace& ws, hipStream_t s,
               const NextStress* next = nullptr, const SceneSet* ss = nullptr);
// out.attr[i] = in.attr[map[i]] for every attribute; out.id[i] = in.id ? in.id[map[i]] : map[i]
hipError_t reorder_particles(const ParticlesView& in, const ParticlesView& out, const int32_t* map,
                             Workspace& ws, hipStream_t s);

// grid-stride launch width for kernels that loop over a device-resident count
inline int persistent_groups() { return 4096; }

// ---- implicit step (implicit.hip; the projection in transfer.hip) ----
// Vectors hold 3 floats per node (node = block * 64 + cell) for the blocks the
// table holds (read on the device), at most grid_cap.
// P2G's folded scatter of the stress in ws.contrib alone (as if every
// particle were massless; v, C and m are not read): G2P2G's force half
hipError_t p2g_force_scatter(const ParticlesView& ps, const TableView& t, const SortView& sv, float* grid,
                             int32_t grid_cap, float dt, float dx, Workspace& ws, hipStream_t s);
// G2P2G's gather half, per bucket: Fout[9 p] = (I + dt C(v)) F_old
hipError_t g2p2g_gather(const ParticlesView& ps, const TableView& t, const SortView& sv, const float* v,
                        int32_t grid_cap, float dt, float dx, float* Fout, Workspace& ws, hipStream_t s);
// r += grid channels 1..3 (a folded P2G force) node by node
hipError_t implicit_add_force(const float* grid, int32_t grid_cap, const TableView& t, float* r, hipStream_t s);
// out_i = (out_i dt^2 + m) in_i where the node has mass
hipError_t implicit_force_dt2_mass(const float* grid, int32_t grid_cap, const TableView& t, float dt, const float* in,
                                   float* out, hipStream_t s);
hipError_t implicit_project(const float* grid, int32_t grid_cap, const TableView& t, float dx, const Colliders& cols,
                            const GeneralColliders& gen, float* dof, hipStream_t s);
hipError_t implicit_precondition(const float* grid, int32_t grid_cap, const TableView& t, const float* in, float* out,
                                 hipStream_t s);
hipError_t implicit_lincomb(const TableView& t, int32_t grid_cap, float ca, const float* a, float cb, const float* b,
                            float* c, hipStream_t s);
hipError_t implicit_dot(const TableView& t, int32_t grid_cap, const float* a, const float* b, double* out,
                        hipStream_t s);
hipError_t implicit_fill(const TableView& t, int32_t grid_cap, float* a, float value, hipStream_t s);

// ---- bht and SparseGrid (bht.hip) ----
size_t bht_scratch_bytes(int64_t n, int32_t nb);
hipError_t bht_clear(const BhtView& t, hipStream_t s);
// Inserts n keys with the sequential semantics (serial: one thread inserts
// them one after another); out[i] (may be null) as BHTView::insert returns
// it. Synchronises `s` (once, and once per placement round). rounds: the
// placement rounds taken. Returns a zsmpm result code.
int32_t bht_insert(const BhtView& t, const int32_t* keys, int64_t n, int32_t* out, bool serial, void* scratch,
                   uint32_t* err, hipStream_t s, int32_t* rounds);
hipError_t bht_query_many(const BhtView& t, const int32_t* keys, int64_t n, int32_t* out, hipStream_t s);
int32_t spg_activate(const SpgView& g, const float* x, int64_t n, void* scratch, uint32_t* err, hipStream_t s);
hipError_t spg_sample(const SpgView& g, int chn, int width, const float* x, int64_t n, float* out, hipStream_t s);

// ---- particle layout and IO adapters (io.hip) ----
/// One attribute of a TileVector<f32, L> particle set: its first channel, its
/// width (1, 3 or 9) and the matching per-attribute array; transpose: the 9
/// channels hold a vec<f32,3,3> row-major while the array row is column-major.
struct TileAttr {
  int32_t chn = 0;
  int32_t width = 0;
  int32_t transpose = 0;
  float* soa = nullptr;
};
constexpr int kTileAttrs = 7;
struct TileAttrs {
  int64_t n = 0;
  int32_t lane_shift = 0;  // log2 of the lane width L
  int32_t nch = 0;         // channels per element
  int32_t count = 0;
  TileAttr attr[kTileAttrs];
};
// to_tiles: tiles <- arrays; else arrays <- tiles
hipError_t tiles_convert(const TileAttrs& a, float* tiles, bool to_tiles, hipStream_t s);
// gather: dst tile t = src tile map[t]; scatter: dst tile map[t] = src tile t
hipError_t tiles_reorder(const float* src, float* dst, const int32_t* map, int64_t ntiles, int64_t tile_floats,
                         bool scatter, uint32_t* err, hipStream_t s);
// Host-side file formats; each synchronises `s` and returns a zsmpm result code.
int32_t checkpoint_save(const char* path, const ParticlesView& p, const TableView* t, double time, int64_t step,
                        hipStream_t s);
int32_t checkpoint_info(const char* path, zsmpm_checkpoint_info* out);
int32_t checkpoint_load(const char* path, const ParticlesView& p, const TableView* t, zsmpm_checkpoint_info* out,
                        hipStream_t s);
int32_t write_bgeo(const char* path, const float* pos, int64_t n, const float* attr, int32_t width,
                   const char* name, hipStream_t s);

}  // namespace zsmpm
