// C ABI of the zsmpm library (declared in include/zsmpm.h): policies with
// their scratch workspace, descriptor validation and the stage entry points.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <new>
#include <random>
#include <vector>

#include "../../include/zsmpm.h"
#include "zsmpm_internal.hpp"

using namespace zsmpm;

struct zsmpm_policy_s {
  int devid = 0;
  hipStream_t stream = nullptr;
  // second stream of the substep: the stress pass runs on it while the
  // sparsity stages run on `stream` (created on first use, default priority)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // captured pair of substeps (ZSMPM_SUBSTEPS_GRAPH): the capture stream, the
  // executable graph and the bytes of the arguments it was captured with
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;
  uint64_t ws_epoch = 0;  // bumped whenever the scratch is reallocated
  // learned closure of the stress pass (zsmpm_policy_closure__rocm)
  bool closure_on = false;
  ClosureView closure{};
  float* mlp_partial = nullptr;  // [kMlpGroups * kMlpParams] gradient partials
  float* mlp_loss = nullptr;     // [kMlpGroups]
  // batched scenes (zsmpm_substeps_batched__rocm): the scene table on the
  // device and the host bytes it was uploaded from
  SceneParams* scene_p = nullptr;   // [kMaxScenes]
  int64_t* scene_first = nullptr;   // [kMaxScenes + 1]
  std::vector<unsigned char> scene_bytes;
  // general colliders of every grid update (zsmpm_policy_colliders__rocm)
  GeneralColliders gcols{};
  // zsmpm_table_activate_noalloc__rocm ran on this policy: its device-side
  // errors are in the device's fixed scratch array
  bool noalloc_used = false;
  // scratch grid and zero masses of G2P2G's force scatter (grown on demand)
  void* g2p2g_buf = nullptr;
  size_t g2p2g_bytes = 0;
  // vectors of the implicit solve (grown on demand)
  void* cg_buf = nullptr;
  size_t cg_bytes = 0;
  // scratch of bht insertions (grown on demand) and the rounds the last took
  void* bht_buf = nullptr;
  size_t bht_bytes = 0;
  int32_t bht_rounds = 0;
  // P2G / G2P workgroup counts picked by ZSMPM_SUBSTEPS_AUTOTUNE (0: default)
  int32_t tuned_p2g = 0, tuned_g2p = 0;
  int32_t tuned_p2g_tile = 0, tuned_g2p_threads = 0;  // launch tiles (0: defaults)
  int32_t transfer = 0;  // the substep drivers' transfers: ZSMPM_TRANSFER_APIC or ZSMPM_TRANSFER_CELL
  Workspace ws{};
  std::vector<void*> allocs;
  int64_t reserved_n = 0;
  int32_t reserved_T = 0;
  // internal buckets for calls that pass none
  int32_t* own_counts = nullptr;
  int32_t* own_offsets = nullptr;
  int32_t* own_indices = nullptr;
  int32_t own_cap_blocks = 0;
  // stage timing (pol.profile(true)): event pairs recorded on the stream,
  // read back and accumulated at the next synchronisation
  bool profile = false;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> ev_free;
  std::vector<Pending> pending;
  double stage_ms[ZSMPM_STAGE_COUNT] = {};
  uint64_t stage_calls[ZSMPM_STAGE_COUNT] = {};
};

namespace {

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

template <typename D>
int32_t check_header(const D* d) {
  if (!d) return ZSMPM_INVALID_ARGUMENT;
  if (d->hdr.major != ZSMPM_ABI_MAJOR) return ZSMPM_INCOMPATIBLE_MAJOR;
  if (d->hdr.size < sizeof(D)) return ZSMPM_INSUFFICIENT_SIZE;
  return ZSMPM_OK;
}

int32_t from_hip(hipError_t e) {
  if (e == hipSuccess) return ZSMPM_OK;
  if (e == hipErrorInvalidValue) return ZSMPM_INVALID_ARGUMENT;
  if (e == hipErrorOutOfMemory) return ZSMPM_CAPACITY;
  return ZSMPM_DEVICE_ERROR;
}

// Several bits can be raised by one cause (a bucket beyond the grid capacity
// also leaves its particles outside the arena they land in), so causes are
// reported before their consequences.
int32_t from_err_bits(uint32_t b) {
  if (b & kErrTableFull) return ZSMPM_TABLE_FULL;
  if (b & kErrKeyRange) return ZSMPM_KEY_RANGE;
  if (b & kErrCapacity) return ZSMPM_CAPACITY;
  if (b & kErrMissingBlock) return ZSMPM_MISSING_BLOCK;
  if (b & kErrArena) return ZSMPM_STENCIL_OUTSIDE;
  return ZSMPM_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

template <typename T>
hipError_t alloc(zsmpm_policy_s* p, T** out, size_t count) {
  void* ptr = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&ptr, count * sizeof(T));
  if (e != hipSuccess) return e;
  p->allocs.push_back(ptr);
  *out = static_cast<T*>(ptr);
  return hipSuccess;
}

hipEvent_t take_event(zsmpm_policy_s* p) {
  if (!p->ev_free.empty()) {
    hipEvent_t e = p->ev_free.back();
    p->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one stage with events on the policy's stream when profiling is on
// (the StreamExecutionTimer of cuda/Cuda.h:161-178, kept per stage).
struct StageScope {
  zsmpm_policy_s* p;
  int stage;
  hipStream_t s;
  hipEvent_t a = nullptr;
  StageScope(zsmpm_policy_s* pol, int st) : StageScope(pol, st, pol->stream) {}
  StageScope(zsmpm_policy_s* pol, int st, hipStream_t stream) : p(pol), stage(st), s(stream) {
    if (!p->profile) return;
    a = take_event(p);
    if (a && hipEventRecord(a, s) != hipSuccess) {
      p->ev_free.push_back(a);
      a = nullptr;
    }
  }
  ~StageScope() {
    if (!a) return;
    hipEvent_t b = take_event(p);
    if (!b || hipEventRecord(b, s) != hipSuccess) {
      p->ev_free.push_back(a);
      if (b) p->ev_free.push_back(b);
      return;
    }
    p->pending.push_back({stage, a, b});
  }
};

// after the stream is idle: fold the recorded event pairs into the totals
void harvest_events(zsmpm_policy_s* p) {
  for (auto& q : p->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) {
      p->stage_ms[q.stage] += ms;
      p->stage_calls[q.stage] += 1;
    }
    p->ev_free.push_back(q.a);
    p->ev_free.push_back(q.b);
  }
  p->pending.clear();
}

void destroy_events(zsmpm_policy_s* p) {
  harvest_events(p);
  for (hipEvent_t e : p->ev_free) (void)hipEventDestroy(e);
  p->ev_free.clear();
  if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
  if (p->join_ev) (void)hipEventDestroy(p->join_ev);
  if (p->side) (void)hipStreamDestroy(p->side);
  p->fork_ev = p->join_ev = nullptr;
  p->side = nullptr;
}

hipError_t ensure_side(zsmpm_policy_s* p) {
  if (p->side) return hipSuccess;
  hipError_t e = hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    if (p->join_ev) (void)hipEventDestroy(p->join_ev);
    p->fork_ev = p->join_ev = nullptr;
    p->side = nullptr;
  }
  return e;
}

// Work enqueued on the side stream after fork() is ordered after everything
// already on the policy's stream; the policy's stream waits for it when the
// object goes out of scope (every return path), so a sync of the policy's
// stream still covers all of a call's work.
struct SideFork {
  zsmpm_policy_s* p;
  bool open = false;
  explicit SideFork(zsmpm_policy_s* pol) : p(pol) {}
  hipError_t fork() {
    hipError_t e = hipEventRecord(p->fork_ev, p->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(p->side, p->fork_ev, 0);
    open = e == hipSuccess;
    return e;
  }
  hipError_t join() {
    if (!open) return hipSuccess;
    open = false;
    hipError_t e = hipEventRecord(p->join_ev, p->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(p->stream, p->join_ev, 0);
    if (e != hipSuccess) (void)hipStreamSynchronize(p->side);
    return e;
  }
  ~SideFork() { (void)join(); }
};

void destroy_graph(zsmpm_policy_s* p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  p->gexec = nullptr;
  p->gkey.clear();
}

void free_all(zsmpm_policy_s* p) {
  destroy_graph(p);  // the graph holds scratch addresses
  ++p->ws_epoch;
  for (void* a : p->allocs) (void)hipFree(a);
  p->allocs.clear();
  p->ws = Workspace{};
  p->mlp_partial = p->mlp_loss = nullptr;
  p->own_counts = p->own_offsets = p->own_indices = nullptr;
  p->own_cap_blocks = 0;
  p->reserved_n = 0;
  p->reserved_T = 0;
}

int32_t reserve(zsmpm_policy_s* p, int64_t n, int32_t T) {
  if (n < 0 || T < 0) return ZSMPM_INVALID_ARGUMENT;
  if (T > 0 && !is_pow2(T)) return ZSMPM_INVALID_ARGUMENT;
  if (n <= p->reserved_n && T <= p->reserved_T && !p->allocs.empty()) return ZSMPM_OK;
  if (n > (int64_t)0x7fffffff) return ZSMPM_CAPACITY;
  const int64_t nn = n > p->reserved_n ? n : p->reserved_n;
  const int32_t TT = T > p->reserved_T ? T : p->reserved_T;
  DeviceGuard g(p->devid);
  (void)hipStreamSynchronize(p->stream);
  free_all(p);
  Workspace& w = p->ws;
  w.n_cap = (uint32_t)(nn > 0 ? nn : 1);
  w.ts_cap = (uint32_t)(TT > 0 ? TT : 1024);
  w.list_cap = w.ts_cap;
  hipError_t e = hipSuccess;
#define ZA(ptr, cnt) \
  if (e == hipSuccess) e = alloc(p, &(ptr), (size_t)(cnt))
  ZA(w.skey, w.ts_cap);
  ZA(w.sorder, w.ts_cap);
  ZA(w.la_key, w.list_cap);
  ZA(w.la_val, w.list_cap);
  ZA(w.lb_key, w.list_cap);
  ZA(w.lb_val, w.list_cap);
  ZA(w.bkt, w.n_cap);
  ZA(w.run_start, w.n_cap);
  ZA(w.rk_a, w.n_cap);
  ZA(w.rv_a, w.n_cap);
  ZA(w.rk_b, w.n_cap);
  ZA(w.rv_b, w.n_cap);
  ZA(w.run_len, w.n_cap);
  ZA(w.run_dest, w.n_cap);
  ZA(w.tile_heads, w.n_cap / 4096 + 2);
  ZA(w.tile_base, w.n_cap / 4096 + 2);
  const uint32_t big = w.n_cap > w.ts_cap ? w.n_cap : w.ts_cap;
  w.rs_words = radix_sort_scratch_words(big);
  ZA(w.rs_scratch, w.rs_words);
  w.scan_words = 16384 + 64;
  ZA(w.scan_tmp, w.scan_words);
  ZA(w.counters, kCtrCount);
  ZA(w.err, 4);
  ZA(w.contrib, (size_t)9 * w.n_cap);
  ZA(w.nbr, (size_t)8 * w.ts_cap);
  ZA(w.fb, w.n_cap);
  ZA(w.wq, w.n_cap / 64 + 2);
  ZA(w.lj_next[0], w.n_cap);
  ZA(w.lj_next[1], w.n_cap);
  {
    // insertion orders: particle indices (< n) or enlargement candidates (< 8 T)
    const uint64_t orders = (uint64_t)w.n_cap > 8ull * w.ts_cap ? (uint64_t)w.n_cap : 8ull * w.ts_cap;
    w.bm_words = (uint32_t)(orders / 32 + 2);
  }
  ZA(w.bitmap, w.bm_words);
  ZA(w.bm_prefix, w.bm_words);
  ZA(w.pslot, w.n_cap);
  ZA(w.slot_block, w.ts_cap);
  ZA(w.dmask, w.ts_cap / 32 + 1);
  p->own_cap_blocks = (int32_t)w.ts_cap;
  ZA(p->own_counts, (size_t)p->own_cap_blocks + 1);
  ZA(p->own_offsets, (size_t)p->own_cap_blocks + 1);
  ZA(p->own_indices, w.n_cap);
#undef ZA
  if (e == hipSuccess) e = hipMemsetAsync(w.counters, 0, sizeof(uint32_t) * kCtrCount, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(w.err, 0, sizeof(uint32_t) * 4, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) {
    free_all(p);
    return from_hip(e);
  }
  p->reserved_n = nn;
  p->reserved_T = TT;
  return ZSMPM_OK;
}

int32_t view_table(const zsmpm_table* d, TableView& t) {
  int32_t r = check_header(d);
  if (r) return r;
  if (!is_pow2(d->table_size) || !d->keys || !d->indices || !d->status || !d->cnt || !d->active_keys)
    return ZSMPM_INVALID_ARGUMENT;
  t.T = d->table_size;
  t.keys = d->keys;
  t.indices = d->indices;
  t.status = d->status;
  t.cnt = d->cnt;
  t.active = d->active_keys;
  return ZSMPM_OK;
}

int32_t view_particles(const zsmpm_particles* d, ParticlesView& ps, bool need_state) {
  int32_t r = check_header(d);
  if (r) return r;
  if (d->n < 0 || !d->x) return ZSMPM_INVALID_ARGUMENT;
  if (need_state && (!d->v || !d->C || !d->F || !d->m)) return ZSMPM_INVALID_ARGUMENT;
  ps.n = d->n;
  ps.x = d->x;
  ps.v = d->v;
  ps.C = d->C;
  ps.F = d->F;
  ps.m = d->m;
  ps.logJp = d->logJp;
  ps.id = d->id;
  return ZSMPM_OK;
}

bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  if (!a || !b || na == 0 || nb == 0) return false;
  const char* pa = static_cast<const char*>(a);
  const char* pb = static_cast<const char*>(b);
  return pa < pb + nb && pb < pa + na;
}

// a full particle set for reordered output, disjoint from `in`
int32_t check_out_particles(const ParticlesView& in, const ParticlesView& out) {
  if (out.n != in.n || !out.v || !out.C || !out.F || !out.m) return ZSMPM_INVALID_ARGUMENT;
  if (in.logJp && !out.logJp) return ZSMPM_INVALID_ARGUMENT;
  const size_t n = (size_t)in.n;
  const void* ia[] = {in.x, in.v, in.C, in.F, in.m, in.logJp, in.id};
  const size_t is[] = {12 * n, 12 * n, 36 * n, 36 * n, 4 * n, 4 * n, 4 * n};
  const void* oa[] = {out.x, out.v, out.C, out.F, out.m, out.logJp, out.id};
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 7; ++j)
      if (overlaps(ia[i], is[i], oa[j], is[j])) return ZSMPM_INVALID_ARGUMENT;
  return ZSMPM_OK;
}

int32_t view_grid(const zsmpm_grid* d) {
  int32_t r = check_header(d);
  if (r) return r;
  if (!d->data || d->capacity_blocks <= 0 || !(d->dx > 0.f)) return ZSMPM_INVALID_ARGUMENT;
  return ZSMPM_OK;
}

int32_t view_buckets(const zsmpm_buckets* d, SortView& sv) {
  int32_t r = check_header(d);
  if (r) return r;
  if (!d->counts || !d->offsets || !d->indices || d->capacity_blocks <= 0) return ZSMPM_INVALID_ARGUMENT;
  sv.counts = d->counts;
  sv.offsets = d->offsets;
  sv.indices = d->indices;
  sv.cap_blocks = d->capacity_blocks;
  return ZSMPM_OK;
}

int32_t make_model(const zsmpm_model* d, Model& m) {
  int32_t r = check_header(d);
  if (r) return r;
  if (d->kind != ZSMPM_FIXED_COROTATED && d->kind != ZSMPM_DRUCKER_PRAGER) return ZSMPM_UNSUPPORTED_OPERATION;
  m.kind = d->kind;
  m.volume = d->volume;
  // lame_parameters (physics/ConstitutiveModel.hpp:34-38); mu is formed in double
  const float onep = 1 + d->nu;
  m.mu = (float)(0.5 * (double)d->E / (double)onep);
  m.lambda = d->E * d->nu / ((1 + d->nu) * (1 - 2 * d->nu));
  m.cohesion = d->cohesion;
  m.beta = d->beta;
  m.yield_surface = d->yield_surface;
  m.volume_correction = d->volume_correction;
  m.fp64_gap = d->svd_fp64_gap;
  return ZSMPM_OK;
}

int32_t make_colliders(const zsmpm_plane_collider* c, int32_t n, Colliders& cols) {
  if (n < 0 || n > kMaxColliders || (n > 0 && !c)) return ZSMPM_INVALID_ARGUMENT;
  cols.n = n;
  for (int i = 0; i < n; ++i) {
    if (c[i].type < 0 || c[i].type > 2) return ZSMPM_INVALID_ARGUMENT;
    cols.c[i].type = c[i].type;
    for (int d = 0; d < 3; ++d) {
      cols.c[i].origin[d] = c[i].origin[d];
      cols.c[i].normal[d] = c[i].normal[d];
    }
  }
  return ZSMPM_OK;
}

int32_t ensure(zsmpm_policy_s* p, int64_t n, int32_t T) {
  if (n <= p->reserved_n && T <= p->reserved_T && !p->allocs.empty()) return ZSMPM_OK;
  return reserve(p, n > p->reserved_n ? n : p->reserved_n, T > p->reserved_T ? T : p->reserved_T);
}

SortView own_buckets(zsmpm_policy_s* p) {
  SortView sv;
  sv.counts = p->own_counts;
  sv.offsets = p->own_offsets;
  sv.indices = p->own_indices;
  sv.cap_blocks = p->own_cap_blocks;
  return sv;
}

__global__ void set_maxvel_kernel(uint32_t* counters) { counters[kCtrMaxVel] = 0u; }
// atomic_max(maxVel, m) (execution/Atomics.hpp:308-350) with float
// semantics for any caller-initialised value: m = max |v|^2 >= 0 replaces a
// smaller value, a negative one (sign bit set) and a NaN.
__global__ void copy_maxvel_kernel(const uint32_t* counters, float* out) {
  const uint32_t mb = counters[kCtrMaxVel];
  const float m = __uint_as_float(mb);
  uint32_t* o = reinterpret_cast<uint32_t*>(out);
  uint32_t cur = __atomic_load_n(o, __ATOMIC_RELAXED);
  while (!(__uint_as_float(cur) >= m)) {
    const uint32_t prev = atomicCAS(o, cur, mb);
    if (prev == cur) break;
    cur = prev;
  }
}

#define ZS_TRY(expr)              \
  do {                            \
    int32_t _r = (expr);          \
    if (_r != ZSMPM_OK) return _r; \
  } while (0)
#define ZS_HIP(expr) ZS_TRY(from_hip(expr))

// The first-occurrence bitmap holds one bit per insertion order; an
// enlargement by nd offsets numbers its candidates up to T * nd, which can
// exceed what reserve() sized for (particle counts and 8 T). Grows both
// bitmap arrays (the old ones are released with the rest of the scratch).
int32_t ensure_bitmap(zsmpm_policy_s* p, uint64_t orders) {
  const uint64_t words = orders / 32 + 2;
  if (words <= p->ws.bm_words) return ZSMPM_OK;
  if (words > 0xffffffffull) return ZSMPM_INVALID_ARGUMENT;
  ZS_HIP(hipStreamSynchronize(p->stream));
  uint32_t *bm = nullptr, *pre = nullptr;
  ZS_HIP(alloc(p, &bm, (size_t)words));
  ZS_HIP(alloc(p, &pre, (size_t)words));
  destroy_graph(p);  // a captured graph holds the old addresses
  ++p->ws_epoch;
  p->ws.bitmap = bm;
  p->ws.bm_prefix = pre;
  p->ws.bm_words = (uint32_t)words;
  return ZSMPM_OK;
}

int32_t run_buckets(zsmpm_policy_s* p, const TableView& t, const ParticlesView& ps, float dx, SortView& sv) {
  return from_hip(bucket_sort(t, p->ws, ps.x, ps.n, dx, 4, -2, 0.5f, sv, p->stream));
}

// Multi-substep chaining of the sorted driver: the stress of substep k+1 is
// computed by G2P of substep k from the updated F.
struct Chain {
  bool cached_stress = false;     // ws.contrib already holds this substep's stress
  const float* lj_src = nullptr;  // this substep's logJp (the previous G2P's cache)
  bool compute_next = false;      // G2P computes the next substep's stress
  float* lj_next = nullptr;       // where G2P puts the next substep's logJp
};

// The explicit driver's substep (stage order of the reconstructed zpc
// driver); with `particles_out`, G2P also performs the Vector::reorder gather.
// With ss (a batch, validated by its entry point) every scene's model, dt,
// gravity and colliders come from the device scene table, and model, dt,
// extf and colliders are not read.
int32_t run_substep(zsmpm_policy_s* pol, const zsmpm_model* model, const zsmpm_particles* particles,
                    const zsmpm_particles* particles_out, const zsmpm_table* table, const zsmpm_buckets* buckets,
                    const zsmpm_grid* grid, float dt, const float extf[3], const zsmpm_plane_collider* colliders,
                    int32_t ncolliders, const Chain* chain = nullptr, const SceneSet* ss = nullptr) {
  const bool multi = ss && ss->n > 0;
  if (!pol || (!multi && !extf)) return ZSMPM_INVALID_ARGUMENT;
  Model m{};
  ParticlesView ps, po;
  TableView t;
  Colliders cols;
  SortView sv;
  ZS_TRY(view_particles(particles, ps, true));
  if (particles_out) {
    ZS_TRY(view_particles(particles_out, po, true));
    ZS_TRY(check_out_particles(ps, po));
  }
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  if (!multi) {
    ZS_TRY(make_model(model, m));
    ZS_TRY(make_colliders(colliders, ncolliders, cols));
    if (m.kind == ZSMPM_DRUCKER_PRAGER && !ps.logJp) return ZSMPM_INVALID_ARGUMENT;
  } else if (pol->closure_on || (chain && (chain->cached_stress || chain->compute_next))) {
    return ZSMPM_UNSUPPORTED_OPERATION;
  }
  const bool cell = pol->transfer == ZSMPM_TRANSFER_CELL;
  if (cell && (multi || (chain && chain->compute_next))) return ZSMPM_UNSUPPORTED_OPERATION;
  const float no_extf[3] = {0.f, 0.f, 0.f};
  const float* g3 = multi ? no_extf : extf;
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, ps.n, t.T));
  if (buckets) {
    ZS_TRY(view_buckets(buckets, sv));
  } else {
    sv = own_buckets(pol);
  }
  const float dx = grid->dx;
  hipStream_t s = pol->stream;
  Workspace& ws = pol->ws;
  ws.grid_p2g = pol->tuned_p2g;
  ws.grid_g2p = pol->tuned_g2p;
  ws.p2g_tile = pol->tuned_p2g_tile;
  ws.g2p_threads = pol->tuned_g2p_threads;
  // the transfer scheme holds for this substep's launches only (the
  // stage-wise entry points keep their own transfers)
  struct CellScope {
    Workspace& w;
    CellScope(Workspace& w_, bool on) : w(w_) { w.cell = on; }
    ~CellScope() { w.cell = false; }
  } cell_scope(ws, cell);
  const int32_t lo[3] = {0, 0, 0}, hi[3] = {2, 2, 2};
  // The stress pass reads F and logJp and writes ws.contrib; the sparsity
  // stages read x only. It runs on the side stream beside them and P2G waits
  // for both.
  SideFork side(pol);
  if (!chain || !chain->cached_stress) {
    ZS_HIP(ensure_side(pol));
    ZS_HIP(side.fork());
    StageScope sc(pol, ZSMPM_STAGE_STRESS, pol->side);
    ZS_HIP(zsmpm::stress(m, ps, ws, pol->side, pol->closure_on ? &pol->closure : nullptr, nullptr,
                         /*beside_sparsity=*/true, ss));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_TABLE_CLEAR);
    // the same launch clears the bucket counts the bucket pass adds to
    ZS_HIP(table_clear(t, ws, s, reinterpret_cast<uint32_t*>(sv.counts), (int64_t)sv.cap_blocks + 1));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_ACTIVATE);
    ZS_HIP(table_activate(t, ws, ps.x, ps.n, dx, 4, -2, 0.5f, s, /*record_slots=*/true, ss));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_ENLARGE);
    ZS_HIP(table_enlarge(t, ws, lo, hi, t.T / 8 > 0 ? t.T / 8 : 1, s));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_BUCKET_SORT);
    // the activation above recorded every particle's key slot in the table it
    // just built from empty: the bucket pass reads those instead of x
    ZS_HIP(bucket_sort_slots(ws, ps.n, sv, s, /*counts_zeroed=*/true));
    ZS_HIP(block_neighbours(t, ws, s));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_GRID_CLEAR);
    // also resets the P2G bucket queue and the max |v|^2 counter
    ZS_HIP(zsmpm::grid_clear(grid->data, t.cnt, grid->capacity_blocks, s, ws.counters));
  }
  ZS_HIP(side.join());
  {
    // P2G folds the force into the momentum channels (the driver's
    // v = (mv + rhs)/m convention), so the update must not add channels 4..6
    StageScope sc(pol, ZSMPM_STAGE_P2G);
    ZS_HIP(zsmpm::p2g_scatter(ps, t, sv, grid->data, grid->capacity_blocks, dt, dx, true, ws, s,
                              /*reset_queue=*/false, ss));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_GRID_UPDATE);
    ZS_HIP(zsmpm::grid_update(grid->data, grid->capacity_blocks, t, dt, g3, cols, dx, 0, ws, s, ss, &pol->gcols));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_G2P);
    NextStress ns{};
    ns.m = m;
    if (chain) {
      ns.lj_src = chain->lj_src;
      if (chain->compute_next) {
        ns.contrib = ws.contrib;
        ns.lj_next = chain->lj_next;
      }
    }
    ZS_HIP(zsmpm::g2p(ps, particles_out ? &po : nullptr, t, sv, grid->data, grid->capacity_blocks, dt, dx, ws, s,
                      &ns, ss));
  }
  return ZSMPM_OK;
}

template <typename D>
void key_put(std::vector<unsigned char>& k, const D* d, size_t bytes = sizeof(D)) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(d);
  if (!d) {
    k.push_back(0);
    return;
  }
  k.push_back(1);
  k.insert(k.end(), b, b + bytes);
}

// Captures the substeps a -> b and b -> a into one executable graph on the
// policy's capture stream, unless the graph kept on the policy was captured
// with the same arguments and scratch. Profiling is off while capturing.
int32_t capture_pair(zsmpm_policy_s* pol, const zsmpm_model* model, const zsmpm_particles* a,
                     const zsmpm_particles* b, const zsmpm_table* table, const zsmpm_buckets* buckets,
                     const zsmpm_grid* grid, float dt, const float extf[3], const zsmpm_plane_collider* colliders,
                     int32_t ncolliders, const SceneSet* ss = nullptr) {
  const bool multi = ss && ss->n > 0;
  if ((!multi && (!model || !extf)) || !table || !grid || ncolliders < 0 || ncolliders > kMaxColliders)
    return ZSMPM_INVALID_ARGUMENT;
  std::vector<unsigned char> key;
  key_put(key, &multi);
  if (multi) key.insert(key.end(), pol->scene_bytes.begin(), pol->scene_bytes.end());
  key_put(key, model);
  key_put(key, a);
  key_put(key, b);
  key_put(key, table);
  key_put(key, buckets);
  key_put(key, grid);
  key_put(key, &dt);
  key_put(key, extf, 3 * sizeof(float));
  key_put(key, &ncolliders);
  if (ncolliders > 0 && colliders) key_put(key, colliders, (size_t)ncolliders * sizeof(zsmpm_plane_collider));
  key_put(key, &pol->ws_epoch);
  key_put(key, &pol->closure_on);
  key_put(key, &pol->closure);
  key_put(key, &pol->tuned_p2g);
  key_put(key, &pol->tuned_g2p);
  key_put(key, &pol->tuned_p2g_tile);
  key_put(key, &pol->tuned_g2p_threads);
  key_put(key, &pol->transfer);
  key_put(key, &pol->gcols);
  if (pol->gexec && key == pol->gkey) return ZSMPM_OK;
  destroy_graph(pol);
  ZS_HIP(ensure_side(pol));
  if (!pol->cap) ZS_HIP(hipStreamCreateWithFlags(&pol->cap, hipStreamNonBlocking));
  const hipStream_t saved = pol->stream;
  const bool prof = pol->profile;
  pol->stream = pol->cap;
  pol->profile = false;
  int32_t r = from_hip(hipStreamBeginCapture(pol->cap, hipStreamCaptureModeThreadLocal));
  const bool begun = r == ZSMPM_OK;
  if (r == ZSMPM_OK)
    r = run_substep(pol, model, a, b, table, buckets, grid, dt, extf, colliders, ncolliders, nullptr, ss);
  if (r == ZSMPM_OK)
    r = run_substep(pol, model, b, a, table, buckets, grid, dt, extf, colliders, ncolliders, nullptr, ss);
  hipGraph_t g = nullptr;
  if (begun) {
    const hipError_t e = hipStreamEndCapture(pol->cap, &g);
    if (r == ZSMPM_OK) r = from_hip(e);
  }
  pol->stream = saved;
  pol->profile = prof;
  if (r == ZSMPM_OK && g) r = from_hip(hipGraphInstantiate(&pol->gexec, g, nullptr, nullptr, 0));
  if (g) (void)hipGraphDestroy(g);
  if (r != ZSMPM_OK) {
    pol->gexec = nullptr;
    return r;
  }
  pol->gkey.swap(key);
  return ZSMPM_OK;
}

// ------------------------------------------------------ allocation-free
// zsmpm_table_activate_noalloc__rocm runs the activation on one fixed array
// of the library's own device memory, `scratch` (each GPU holds its own copy
// until the process ends), instead of policy scratch that hipMalloc provides.
// Sub-buffers at 256-byte aligned offsets: the dedup hash (keys, orders), the
// two new-key lists, the claimed-slot mask, the first-occurrence bitmap and
// its prefix, the popcount scan's tile sums, the counters and the error word.
constexpr uint32_t kNoAllocSlots = 1u << 21;      // largest table (and dedup hash)
constexpr int64_t kNoAllocParticles = 1ll << 24;  // most particles per call
constexpr size_t na_align(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr uint32_t kNaBmWords = (uint32_t)(kNoAllocParticles / 32 + 2);
constexpr uint32_t kNaScanWords = 1024;
constexpr size_t kNaSkey = 0;
constexpr size_t kNaSorder = na_align(kNaSkey + 8ull * kNoAllocSlots);
constexpr size_t kNaLkey = na_align(kNaSorder + 4ull * kNoAllocSlots);
constexpr size_t kNaLval = na_align(kNaLkey + 4ull * kNoAllocSlots);
constexpr size_t kNaDmask = na_align(kNaLval + 4ull * kNoAllocSlots);
constexpr size_t kNaBitmap = na_align(kNaDmask + 4ull * (kNoAllocSlots / 32 + 1));
constexpr size_t kNaPrefix = na_align(kNaBitmap + 4ull * kNaBmWords);
constexpr size_t kNaScan = na_align(kNaPrefix + 4ull * kNaBmWords);
constexpr size_t kNaCounters = na_align(kNaScan + 4ull * kNaScanWords);
constexpr size_t kNaErr = na_align(kNaCounters + 4ull * kCtrCount);
constexpr size_t kNoAllocBytes = na_align(kNaErr + 16);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The device address of `scratch` on each device, looked up once per device
// (with that device current) by the first policy created on it. The host-side
// name `scratch` is not a device address; only these pointers are used.
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};
bool g_na_tried[kNaMaxDevices] = {};

void noalloc_setup(int devid) {
  if (devid < 0 || devid >= kNaMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_tried[devid]) return;
  g_na_tried[devid] = true;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_na_ptr[devid] = ptr;
}

unsigned char* noalloc_base(int devid) {
  if (devid < 0 || devid >= kNaMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_na_mu);
  return static_cast<unsigned char*>(g_na_ptr[devid]);
}

// ---------------------------------------------------------------- batches

// Validates a batch's scenes against the combined particle arrays and
// uploads the scene table to the policy (only when its bytes changed; the
// stream is drained first, since queued or captured substeps read it).
int32_t upload_scenes(zsmpm_policy_s* pol, const zsmpm_batch_scene* sc, int32_t ns, const ParticlesView& pa,
                      const ParticlesView& pb, SceneSet& out) {
  if (ns <= 0 || ns > kMaxScenes) return ZSMPM_INVALID_ARGUMENT;
  std::vector<SceneParams> params((size_t)ns);
  std::vector<int64_t> first((size_t)ns + 1);
  int64_t next = 0;
  for (int32_t i = 0; i < ns; ++i) {
    ZS_TRY(check_header(&sc[i]));
    if (sc[i].first != next || sc[i].count < 0) return ZSMPM_INVALID_ARGUMENT;
    if (!(sc[i].dx > 0.f) || !isfinite(sc[i].dx) || !isfinite(sc[i].dt)) return ZSMPM_INVALID_ARGUMENT;
    for (int d = 0; d < 3; ++d)
      if (!isfinite(sc[i].extf[d])) return ZSMPM_INVALID_ARGUMENT;
    SceneParams& p = params[(size_t)i];
    memset(&p, 0, sizeof(p));
    ZS_TRY(make_model(&sc[i].model, p.model));
    if (p.model.kind == ZSMPM_DRUCKER_PRAGER && sc[i].count > 0 && (!pa.logJp || !pb.logJp))
      return ZSMPM_INVALID_ARGUMENT;
    ZS_TRY(make_colliders(sc[i].colliders, sc[i].ncolliders, p.cols));
    p.dx = sc[i].dx;
    p.dx_inv = (float)1.0 / sc[i].dx;  // as the single-scene activation forms it
    p.dt = sc[i].dt;
    for (int d = 0; d < 3; ++d) p.g[d] = sc[i].extf[d];
    first[(size_t)i] = next;
    next += sc[i].count;
  }
  if (next != pa.n) return ZSMPM_INVALID_ARGUMENT;
  first[(size_t)ns] = next;
  std::vector<unsigned char> bytes(sizeof(int32_t) + params.size() * sizeof(SceneParams) +
                                   first.size() * sizeof(int64_t));
  memcpy(bytes.data(), &ns, sizeof(int32_t));
  memcpy(bytes.data() + sizeof(int32_t), params.data(), params.size() * sizeof(SceneParams));
  memcpy(bytes.data() + sizeof(int32_t) + params.size() * sizeof(SceneParams), first.data(),
         first.size() * sizeof(int64_t));
  if (!pol->scene_p) {
    void* a = nullptr;
    void* b = nullptr;
    ZS_HIP(hipMalloc(&a, sizeof(SceneParams) * kMaxScenes));
    const hipError_t e = hipMalloc(&b, sizeof(int64_t) * (kMaxScenes + 1));
    if (e != hipSuccess) {
      (void)hipFree(a);
      return from_hip(e);
    }
    pol->scene_p = static_cast<SceneParams*>(a);
    pol->scene_first = static_cast<int64_t*>(b);
    pol->scene_bytes.clear();
  }
  if (bytes != pol->scene_bytes) {
    ZS_HIP(hipStreamSynchronize(pol->stream));
    ZS_HIP(hipMemcpy(pol->scene_p, params.data(), params.size() * sizeof(SceneParams), hipMemcpyHostToDevice));
    ZS_HIP(hipMemcpy(pol->scene_first, first.data(), first.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    pol->scene_bytes.swap(bytes);
  }
  out.n = ns;
  out.p = pol->scene_p;
  out.first = pol->scene_first;
  return ZSMPM_OK;
}

// nsteps sorted substeps a -> b -> a ... with the multi-substep flags (the
// shared body of zsmpm_substeps_sorted__rocm and the batched entry point)
int32_t substeps_driver(zsmpm_policy_s* pol, const zsmpm_model* model, const zsmpm_particles* a,
                        const zsmpm_particles* b, const zsmpm_table* table, const zsmpm_buckets* buckets,
                        const zsmpm_grid* grid, float dt, const float extf[3], const zsmpm_plane_collider* colliders,
                        int32_t ncolliders, int32_t nsteps, int32_t flags, const SceneSet* ss) {
  const bool fuse = (flags & ZSMPM_SUBSTEPS_FUSE_STRESS) != 0;
  const bool graph = (flags & ZSMPM_SUBSTEPS_GRAPH) != 0;
  // Autotuning: the first 8 substeps (real ones, in pairs a -> b -> a) run
  // with four P2G and four G2P launch configurations (tile and workgroup
  // count), timed per kernel with the stage events; each kernel keeps its
  // fastest. P2G: four waves with 576-particle chunks at 1,024 and 2,048
  // workgroups, two waves with 256-particle chunks at 2,048 and 4,096; G2P:
  // 128 and 256 threads per bucket at 32,768 and 65,536 workgroups.
  int32_t done = 0;
  if ((flags & ZSMPM_SUBSTEPS_AUTOTUNE) && nsteps >= 8 && !fuse) {
    DeviceGuard g(pol->devid);
    const int32_t cand_p2g[4] = {1024, 2048, 2048, 4096}, cand_g2p[4] = {32768, 65536, 32768, 65536};
    const int32_t cand_p2g_tile[4] = {0, 0, 1, 1}, cand_g2p_threads[4] = {128, 128, 256, 256};
    const bool prof = pol->profile;
    ZS_HIP(hipStreamSynchronize(pol->stream));
    harvest_events(pol);
    pol->profile = true;
    double best_p = 0.0, best_g = 0.0;
    int32_t pick_p = 0, pick_g = 0, pick_pt = 0, pick_gt = 0;
    int32_t r = ZSMPM_OK;
    for (int c = 0; c < 4 && r == ZSMPM_OK; ++c) {
      pol->tuned_p2g = cand_p2g[c];
      pol->tuned_g2p = cand_g2p[c];
      pol->tuned_p2g_tile = cand_p2g_tile[c];
      pol->tuned_g2p_threads = cand_g2p_threads[c];
      const double p0 = pol->stage_ms[ZSMPM_STAGE_P2G], g0 = pol->stage_ms[ZSMPM_STAGE_G2P];
      for (int k = 0; k < 2 && r == ZSMPM_OK; ++k, ++done)
        r = run_substep(pol, model, (done & 1) ? b : a, (done & 1) ? a : b, table, buckets, grid, dt, extf,
                        colliders, ncolliders, nullptr, ss);
      if (r == ZSMPM_OK) r = from_hip(hipStreamSynchronize(pol->stream));
      harvest_events(pol);
      const double dp = pol->stage_ms[ZSMPM_STAGE_P2G] - p0, dg = pol->stage_ms[ZSMPM_STAGE_G2P] - g0;
      if (c == 0 || dp < best_p) {
        best_p = dp;
        pick_p = cand_p2g[c];
        pick_pt = cand_p2g_tile[c];
      }
      if (c == 0 || dg < best_g) {
        best_g = dg;
        pick_g = cand_g2p[c];
        pick_gt = cand_g2p_threads[c];
      }
    }
    pol->profile = prof;
    if (r != ZSMPM_OK) return r;
    pol->tuned_p2g = pick_p;
    pol->tuned_g2p = pick_g;
    pol->tuned_p2g_tile = pick_pt;
    pol->tuned_g2p_threads = pick_gt;
  }
  nsteps -= done;  // an even count: the particles are back in a
  if (graph) {
    DeviceGuard g(pol->devid);
    if (nsteps >= 2) {
      ZS_TRY(capture_pair(pol, model, a, b, table, buckets, grid, dt, extf, colliders, ncolliders, ss));
      for (int32_t k = 0; k + 1 < nsteps; k += 2) ZS_HIP(hipGraphLaunch(pol->gexec, pol->stream));
    }
    if (nsteps & 1)
      ZS_TRY(run_substep(pol, model, a, b, table, buckets, grid, dt, extf, colliders, ncolliders, nullptr, ss));
    return ZSMPM_OK;
  }
  for (int32_t k = 0; k < nsteps; ++k) {
    Chain c;
    if (fuse) {
      c.cached_stress = k > 0;
      c.lj_src = k > 0 ? pol->ws.lj_next[(k - 1) & 1] : nullptr;
      c.compute_next = k + 1 < nsteps;
      c.lj_next = pol->ws.lj_next[k & 1];
    }
    const zsmpm_particles* in = (k & 1) ? b : a;
    const zsmpm_particles* out = (k & 1) ? a : b;
    ZS_TRY(run_substep(pol, model, in, out, table, buckets, grid, dt, extf, colliders, ncolliders, &c, ss));
  }
  return ZSMPM_OK;
}

void destroy_policy(zsmpm_policy_s* p) {
  (void)hipStreamSynchronize(p->stream);
  destroy_events(p);
  free_all(p);
  if (p->scene_p) (void)hipFree(p->scene_p);
  if (p->scene_first) (void)hipFree(p->scene_first);
  if (p->bht_buf) (void)hipFree(p->bht_buf);
  if (p->cg_buf) (void)hipFree(p->cg_buf);
  if (p->g2p2g_buf) (void)hipFree(p->g2p2g_buf);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  delete p;
}

}  // namespace

extern "C" {

int32_t zsmpm_abi_version(zsmpm_abi_header* out) {
  if (!out) return ZSMPM_INVALID_ARGUMENT;
  out->size = sizeof(zsmpm_abi_header);
  out->major = ZSMPM_ABI_MAJOR;
  out->minor = ZSMPM_ABI_MINOR;
  return ZSMPM_OK;
}

int32_t zsmpm_policy__rocm(int32_t devid, void* stream, zsmpm_policy* out) {
  if (!out) return ZSMPM_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || devid < 0 || devid >= ndev) return ZSMPM_INVALID_ARGUMENT;
  {
    DeviceGuard g(devid);
    if (check_launch_limits(devid) != hipSuccess) return ZSMPM_UNSUPPORTED_OPERATION;
    noalloc_setup(devid);  // the device's fixed scratch address, once per device
  }
  zsmpm_policy_s* p = new (std::nothrow) zsmpm_policy_s();
  if (!p) return ZSMPM_ERROR;
  p->devid = devid;
  p->stream = static_cast<hipStream_t>(stream);
  *out = p;
  return ZSMPM_OK;
}

int32_t zsmpm_del_policy__rocm(zsmpm_policy pol) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  {
    DeviceGuard g(pol->devid);
    destroy_policy(pol);
  }
  return ZSMPM_OK;
}

int32_t zsmpm_policy_reserve__rocm(zsmpm_policy pol, int64_t max_particles, int32_t table_size) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  return reserve(pol, max_particles, table_size);
}

int32_t zsmpm_policy_sync(zsmpm_policy pol) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  ZS_HIP(hipStreamSynchronize(pol->stream));
  harvest_events(pol);
  uint32_t bits = 0;
  if (pol->ws.err) {
    ZS_HIP(hipMemcpy(&bits, pol->ws.err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (bits) ZS_HIP(hipMemset(pol->ws.err, 0, sizeof(uint32_t)));
  }
  unsigned char* na = pol->noalloc_used ? noalloc_base(pol->devid) : nullptr;
  if (na) {
    // the allocation-free activation's error word (shared by every policy on
    // this device that called it)
    uint32_t nb = 0;
    ZS_HIP(hipMemcpy(&nb, na + kNaErr, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (nb) ZS_HIP(hipMemset(na + kNaErr, 0, sizeof(uint32_t)));
    bits |= nb;
  }
  return from_err_bits(bits);
}

int32_t zsmpm_policy_counters(zsmpm_policy pol, uint32_t* out) {
  if (!pol || !out) return ZSMPM_INVALID_ARGUMENT;
  memset(out, 0, sizeof(uint32_t) * kCtrCount);
  if (!pol->ws.counters) return ZSMPM_OK;
  DeviceGuard g(pol->devid);
  ZS_HIP(hipStreamSynchronize(pol->stream));
  ZS_HIP(hipMemcpy(out, pol->ws.counters, sizeof(uint32_t) * kCtrCount, hipMemcpyDeviceToHost));
  out[kCtrBhtRounds] = (uint32_t)pol->bht_rounds;
  return ZSMPM_OK;
}

int32_t zsmpm_policy_profile(zsmpm_policy pol, int32_t enable) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  pol->profile = enable != 0;
  return ZSMPM_OK;
}

int32_t zsmpm_policy_stage_times(zsmpm_policy pol, double* ms, uint64_t* calls, int32_t n, int32_t reset) {
  if (!pol || n < 0 || (n > 0 && (!ms || !calls))) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  ZS_HIP(hipStreamSynchronize(pol->stream));
  harvest_events(pol);
  for (int i = 0; i < n; ++i) {
    ms[i] = i < ZSMPM_STAGE_COUNT ? pol->stage_ms[i] : 0.0;
    calls[i] = i < ZSMPM_STAGE_COUNT ? pol->stage_calls[i] : 0;
  }
  if (reset) {
    for (int i = 0; i < ZSMPM_STAGE_COUNT; ++i) {
      pol->stage_ms[i] = 0.0;
      pol->stage_calls[i] = 0;
    }
  }
  return ZSMPM_OK;
}

int32_t zsmpm_table_clear__rocm(zsmpm_policy pol, const zsmpm_table* table) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, pol->reserved_n, t.T));
  StageScope sc(pol, ZSMPM_STAGE_TABLE_CLEAR);
  return from_hip(table_clear(t, pol->ws, pol->stream));
}

int32_t zsmpm_table_activate__rocm(zsmpm_policy pol, const zsmpm_table* table, const float* x, int64_t n,
                                   float dx, int32_t block_len, int32_t offset, float displacement) {
  if (!pol || n < 0 || (n > 0 && !x) || block_len <= 0 || !(dx > 0.f)) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, n, t.T));
  StageScope sc(pol, ZSMPM_STAGE_ACTIVATE);
  return from_hip(table_activate(t, pol->ws, x, n, dx, block_len, offset, displacement, pol->stream));
}

int32_t zsmpm_table_activate_noalloc__rocm(zsmpm_policy pol, const zsmpm_table* table, const float* x, int64_t n,
                                           float dx, int32_t block_len, int32_t offset, float displacement) {
  if (!pol || n < 0 || (n > 0 && !x) || block_len <= 0 || !(dx > 0.f)) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  if (n > kNoAllocParticles || (uint32_t)t.T > kNoAllocSlots) return ZSMPM_INSUFFICIENT_SIZE;
  unsigned char* base = noalloc_base(pol->devid);
  if (!base) return ZSMPM_DEVICE_ERROR;
  // views into the device's fixed scratch array; the launches below clear
  // what they read before it is written (dedup hash, mask, bitmap, counters)
  Workspace w{};
  w.n_cap = (uint32_t)(n > 0 ? n : 1);
  w.ts_cap = (uint32_t)t.T > 1024u ? (uint32_t)t.T : 1024u;
  w.list_cap = w.ts_cap;
  w.skey = reinterpret_cast<uint64_t*>(base + kNaSkey);
  w.sorder = reinterpret_cast<uint32_t*>(base + kNaSorder);
  w.la_key = reinterpret_cast<uint32_t*>(base + kNaLkey);
  w.la_val = reinterpret_cast<uint32_t*>(base + kNaLval);
  w.dmask = reinterpret_cast<uint32_t*>(base + kNaDmask);
  w.bitmap = reinterpret_cast<uint32_t*>(base + kNaBitmap);
  w.bm_prefix = reinterpret_cast<uint32_t*>(base + kNaPrefix);
  w.bm_words = kNaBmWords;
  w.scan_tmp = reinterpret_cast<uint32_t*>(base + kNaScan);
  w.scan_words = kNaScanWords;
  w.counters = reinterpret_cast<uint32_t*>(base + kNaCounters);
  w.err = reinterpret_cast<uint32_t*>(base + kNaErr);
  DeviceGuard g(pol->devid);
  pol->noalloc_used = true;
  return from_hip(table_activate(t, w, x, n, dx, block_len, offset, displacement, pol->stream));
}

int32_t zsmpm_table_enlarge__rocm(zsmpm_policy pol, const zsmpm_table* table, const int32_t lo[3],
                                  const int32_t hi[3]) {
  if (!pol || !lo || !hi) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  // candidate c = block * nd + offset is its insertion order, a 32-bit value
  // below T * nd, one bit each in the first-occurrence bitmap
  int64_t nd = 1;
  for (int d = 0; d < 3; ++d) {
    if (hi[d] <= lo[d]) return ZSMPM_OK;  // empty offset range: nothing to insert (EnlargeSparsity loops)
    if (hi[d] - lo[d] > 1024) return ZSMPM_INVALID_ARGUMENT;
    nd *= hi[d] - lo[d];
  }
  if ((int64_t)t.T * nd >= (int64_t)1 << 32) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, pol->reserved_n, t.T));
  ZS_TRY(ensure_bitmap(pol, (uint64_t)t.T * (uint64_t)nd));
  StageScope sc(pol, ZSMPM_STAGE_ENLARGE);
  return from_hip(table_enlarge(t, pol->ws, lo, hi, t.T, pol->stream));
}

int32_t zsmpm_grid_angular_momentum__rocm(zsmpm_policy pol, const zsmpm_grid* grid, const zsmpm_table* table,
                                          int32_t m_channel, int32_t mv_channel, double* sum) {
  if (!pol || !sum || m_channel < 0 || m_channel >= 7 || mv_channel < 0 || mv_channel + 3 > 7)
    return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, pol->reserved_n, t.T));
  return from_hip(grid_angular_momentum(grid->data, grid->capacity_blocks, t, grid->dx, m_channel, mv_channel, sum,
                                        pol->ws, pol->stream));
}

int32_t zsmpm_table_query__rocm(zsmpm_policy pol, const zsmpm_table* table, const int32_t* keys, int64_t n,
                                int32_t* out) {
  if (!pol || n < 0 || (n > 0 && (!keys || !out))) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  DeviceGuard g(pol->devid);
  return from_hip(table_query(t, keys, n, out, pol->stream));
}

int32_t zsmpm_bucket_sort__rocm(zsmpm_policy pol, const zsmpm_table* table, const float* x, int64_t n, float dx,
                                int32_t block_len, int32_t offset, float displacement,
                                const zsmpm_buckets* buckets) {
  if (!pol || n < 0 || (n > 0 && !x) || block_len <= 0 || !(dx > 0.f)) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  SortView sv;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_buckets(buckets, sv));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, n, t.T));
  StageScope sc(pol, ZSMPM_STAGE_BUCKET_SORT);
  return from_hip(bucket_sort(t, pol->ws, x, n, dx, block_len, offset, displacement, sv, pol->stream));
}

int32_t zsmpm_grid_clear__rocm(zsmpm_policy pol, const zsmpm_grid* grid, const zsmpm_table* table) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_grid(grid));
  ZS_TRY(view_table(table, t));
  DeviceGuard g(pol->devid);
  StageScope sc(pol, ZSMPM_STAGE_GRID_CLEAR);
  return from_hip(zsmpm::grid_clear(grid->data, t.cnt, grid->capacity_blocks, pol->stream));
}

int32_t zsmpm_p2g__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                        const zsmpm_table* table, const zsmpm_buckets* buckets, const zsmpm_grid* grid,
                        float dt) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ParticlesView ps;
  TableView t;
  ZS_TRY(make_model(model, m));
  ZS_TRY(view_particles(particles, ps, true));
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  if (m.kind == ZSMPM_DRUCKER_PRAGER && !ps.logJp) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, ps.n, t.T));
  SortView sv;
  if (buckets) {
    ZS_TRY(view_buckets(buckets, sv));
  } else {
    sv = own_buckets(pol);
    StageScope sc(pol, ZSMPM_STAGE_BUCKET_SORT);
    ZS_TRY(run_buckets(pol, t, ps, grid->dx, sv));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_STRESS);
    ZS_HIP(zsmpm::stress(m, ps, pol->ws, pol->stream, pol->closure_on ? &pol->closure : nullptr));
  }
  StageScope sc(pol, ZSMPM_STAGE_P2G);
  ZS_HIP(block_neighbours(t, pol->ws, pol->stream));
  return from_hip(
      zsmpm::p2g_scatter(ps, t, sv, grid->data, grid->capacity_blocks, dt, grid->dx, false, pol->ws, pol->stream));
}

int32_t zsmpm_grid_update__rocm(zsmpm_policy pol, const zsmpm_grid* grid, const zsmpm_table* table, float dt,
                                const float extf[3], const zsmpm_plane_collider* colliders, int32_t ncolliders,
                                int32_t fold_rhs, float* max_vel_sqr) {
  if (!pol || !extf) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  Colliders cols;
  ZS_TRY(view_grid(grid));
  ZS_TRY(view_table(table, t));
  ZS_TRY(make_colliders(colliders, ncolliders, cols));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, pol->reserved_n, t.T));
  StageScope sc(pol, ZSMPM_STAGE_GRID_UPDATE);
  set_maxvel_kernel<<<1, 1, 0, pol->stream>>>(pol->ws.counters);
  ZS_HIP(zsmpm::grid_update(grid->data, grid->capacity_blocks, t, dt, extf, cols, grid->dx, fold_rhs, pol->ws,
                            pol->stream, nullptr, &pol->gcols));
  if (max_vel_sqr) copy_maxvel_kernel<<<1, 1, 0, pol->stream>>>(pol->ws.counters, max_vel_sqr);
  return from_hip(hipGetLastError());
}

int32_t zsmpm_g2p__rocm(zsmpm_policy pol, const zsmpm_particles* particles, const zsmpm_table* table,
                        const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  ParticlesView ps;
  TableView t;
  ZS_TRY(view_particles(particles, ps, true));
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, ps.n, t.T));
  SortView sv;
  if (buckets) {
    ZS_TRY(view_buckets(buckets, sv));
  } else {
    sv = own_buckets(pol);
    StageScope sc(pol, ZSMPM_STAGE_BUCKET_SORT);
    ZS_TRY(run_buckets(pol, t, ps, grid->dx, sv));
  }
  StageScope sc(pol, ZSMPM_STAGE_G2P);
  ZS_HIP(block_neighbours(t, pol->ws, pol->stream));
  return from_hip(
      zsmpm::g2p(ps, nullptr, t, sv, grid->data, grid->capacity_blocks, dt, grid->dx, pol->ws, pol->stream));
}

int32_t zsmpm_p2c2g__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                          const zsmpm_table* table, const zsmpm_buckets* buckets, const zsmpm_grid* grid,
                          float dt) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ParticlesView ps;
  TableView t;
  ZS_TRY(make_model(model, m));
  ZS_TRY(view_particles(particles, ps, true));
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  if (m.kind == ZSMPM_DRUCKER_PRAGER && !ps.logJp) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, ps.n, t.T));
  SortView sv;
  if (buckets) {
    ZS_TRY(view_buckets(buckets, sv));
  } else {
    sv = own_buckets(pol);
    StageScope sc(pol, ZSMPM_STAGE_BUCKET_SORT);
    ZS_TRY(run_buckets(pol, t, ps, grid->dx, sv));
  }
  {
    StageScope sc(pol, ZSMPM_STAGE_STRESS);
    ZS_HIP(zsmpm::stress(m, ps, pol->ws, pol->stream));
  }
  StageScope sc(pol, ZSMPM_STAGE_P2G);
  ZS_HIP(block_neighbours(t, pol->ws, pol->stream));
  return from_hip(
      zsmpm::p2c2g_scatter(ps, t, sv, grid->data, grid->capacity_blocks, dt, grid->dx, pol->ws, pol->stream));
}

int32_t zsmpm_g2c2p__rocm(zsmpm_policy pol, const zsmpm_particles* particles, const zsmpm_table* table,
                          const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  ParticlesView ps;
  TableView t;
  ZS_TRY(view_particles(particles, ps, true));
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, ps.n, t.T));
  SortView sv;
  if (buckets) {
    ZS_TRY(view_buckets(buckets, sv));
  } else {
    sv = own_buckets(pol);
    StageScope sc(pol, ZSMPM_STAGE_BUCKET_SORT);
    ZS_TRY(run_buckets(pol, t, ps, grid->dx, sv));
  }
  StageScope sc(pol, ZSMPM_STAGE_G2P);
  ZS_HIP(block_neighbours(t, pol->ws, pol->stream));
  return from_hip(zsmpm::g2c2p(ps, t, sv, grid->data, grid->capacity_blocks, dt, grid->dx, pol->ws, pol->stream));
}

int32_t zsmpm_substep__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                            const zsmpm_table* table, const zsmpm_buckets* buckets, const zsmpm_grid* grid,
                            float dt, const float extf[3], const zsmpm_plane_collider* colliders,
                            int32_t ncolliders) {
  return run_substep(pol, model, particles, nullptr, table, buckets, grid, dt, extf, colliders, ncolliders);
}

int32_t zsmpm_substep_sorted__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* in,
                                   const zsmpm_particles* out, const zsmpm_table* table,
                                   const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt,
                                   const float extf[3], const zsmpm_plane_collider* colliders,
                                   int32_t ncolliders) {
  if (!out) return ZSMPM_INVALID_ARGUMENT;
  return run_substep(pol, model, in, out, table, buckets, grid, dt, extf, colliders, ncolliders);
}

int32_t zsmpm_substeps_sorted__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* a,
                                    const zsmpm_particles* b, const zsmpm_table* table,
                                    const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt,
                                    const float extf[3], const zsmpm_plane_collider* colliders, int32_t ncolliders,
                                    int32_t nsteps, int32_t flags, int32_t* final_in_b) {
  if (!pol || !a || !b || nsteps < 0 ||
      (flags & ~(ZSMPM_SUBSTEPS_FUSE_STRESS | ZSMPM_SUBSTEPS_GRAPH | ZSMPM_SUBSTEPS_AUTOTUNE)))
    return ZSMPM_INVALID_ARGUMENT;
  const bool fuse = (flags & ZSMPM_SUBSTEPS_FUSE_STRESS) != 0;
  const bool graph = (flags & ZSMPM_SUBSTEPS_GRAPH) != 0;
  if (fuse && graph) return ZSMPM_INVALID_ARGUMENT;
  if (fuse && pol->closure_on) return ZSMPM_UNSUPPORTED_OPERATION;  // G2P's fused stress is the SVD path
  if (final_in_b) *final_in_b = nsteps & 1;
  if (nsteps == 0) return ZSMPM_OK;
  {
    // validate once; run_substep repeats the cheap checks per substep
    ParticlesView pa, pb;
    ZS_TRY(view_particles(a, pa, true));
    ZS_TRY(view_particles(b, pb, true));
    ZS_TRY(check_out_particles(pa, pb));
    DeviceGuard g(pol->devid);
    ZS_TRY(ensure(pol, pa.n, table ? table->table_size : 0));
  }
  return substeps_driver(pol, model, a, b, table, buckets, grid, dt, extf, colliders, ncolliders, nsteps, flags,
                         nullptr);
}

int32_t zsmpm_reorder__rocm(zsmpm_policy pol, const zsmpm_buckets* buckets, const zsmpm_particles* in,
                            const zsmpm_particles* out) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  SortView sv;
  ParticlesView pi, po;
  ZS_TRY(view_buckets(buckets, sv));
  ZS_TRY(view_particles(in, pi, true));
  ZS_TRY(view_particles(out, po, true));
  ZS_TRY(check_out_particles(pi, po));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, pi.n, pol->reserved_T));
  return from_hip(reorder_particles(pi, po, sv.indices, pol->ws, pol->stream));
}

// ---- implicit step ----

// G2P2GTransfer in two passes: the gather half writes each particle's
// vol P F^T (at F = (I + dt C(v)) F_old) into the workspace's stress array;
// P2G's force scatter -- one atomic per node and channel after binning in LDS,
// in place of zpc's 81 per particle -- with zero masses and dt = -1 (its force
// scale -dt D becomes G2P2G's D) accumulates the force into a scratch grid,
// which is added to r node by node.
// sorted: the policy's bucket sort and neighbour table already hold this
// table and these particles (the solver sorts once for all its products)
static int32_t run_g2p2g(zsmpm_policy_s* pol, const Model& m, const ParticlesView& ps, const TableView& t,
                         const zsmpm_grid* grid, float dt, const float* v, float* r, bool sorted = false) {
  hipStream_t s = pol->stream;
  ZS_TRY(ensure(pol, ps.n, t.T));
  const size_t gfloats = (size_t)grid->capacity_blocks * kChannels * kBlockCells;
  // scratch grid, the updated F and a logJp copy the return mapping may
  // write (G2P2G reads logJp and leaves it)
  const size_t galign = (gfloats + 63) & ~(size_t)63;
  const size_t nalign = ((size_t)ps.n + 63) & ~(size_t)63;
  const size_t need = 4 * (galign + 10 * nalign) + 512;
  if (need > pol->g2p2g_bytes) {
    ZS_HIP(hipStreamSynchronize(s));
    if (pol->g2p2g_buf) (void)hipFree(pol->g2p2g_buf);
    pol->g2p2g_buf = nullptr;
    pol->g2p2g_bytes = 0;
    ZS_HIP(hipMalloc(&pol->g2p2g_buf, need));
    pol->g2p2g_bytes = need;
  }
  float* tg = (float*)pol->g2p2g_buf;
  float* Fn = tg + galign;
  float* lj = Fn + 9 * nalign;
  SortView sv = own_buckets(pol);
  if (!sorted) {
    ZS_TRY(run_buckets(pol, t, ps, grid->dx, sv));
    ZS_HIP(block_neighbours(t, pol->ws, s));
  }
  ZS_HIP(g2p2g_gather(ps, t, sv, v, grid->capacity_blocks, dt, grid->dx, Fn, pol->ws, s));
  // the model's vol P F^T of the updated F into the workspace's stress array
  ParticlesView pf = ps;
  pf.F = Fn;
  if (m.kind == 1) {
    if (ps.n > 0)
      ZS_HIP(hipMemcpyAsync(lj, ps.logJp, sizeof(float) * (size_t)ps.n, hipMemcpyDeviceToDevice, s));
    pf.logJp = lj;
  }
  ZS_HIP(zsmpm::stress(m, pf, pol->ws, s));
  ZS_HIP(grid_clear(tg, t.cnt, grid->capacity_blocks, s, nullptr));
  ZS_HIP(p2g_force_scatter(ps, t, sv, tg, grid->capacity_blocks, -1.f, grid->dx, pol->ws, s));
  ZS_HIP(implicit_add_force(tg, grid->capacity_blocks, t, r, s));
  return ZSMPM_OK;
}

// ImplicitMPMSystem::multiply: out = 0, out += G2P2G(in), the mass term
static int32_t run_implicit_multiply(zsmpm_policy_s* pol, const Model& m, const ParticlesView& ps, const TableView& t,
                                     const zsmpm_grid* grid, float dt, const float* in, float* out,
                                     bool sorted = false) {
  ZS_HIP(implicit_fill(t, grid->capacity_blocks, out, 0.f, pol->stream));
  ZS_TRY(run_g2p2g(pol, m, ps, t, grid, dt, in, out, sorted));
  return from_hip(implicit_force_dt2_mass(grid->data, grid->capacity_blocks, t, dt, in, out, pol->stream));
}

int32_t zsmpm_g2p2g__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                          const zsmpm_table* table, const zsmpm_grid* grid, float dt, const float* v, float* r) {
  if (!pol || !v || !r) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ParticlesView ps;
  TableView t;
  ZS_TRY(make_model(model, m));
  ZS_TRY(view_particles(particles, ps, true));
  if (m.kind == 1 && !ps.logJp) return ZSMPM_INVALID_ARGUMENT;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  DeviceGuard g(pol->devid);
  return run_g2p2g(pol, m, ps, t, grid, dt, v, r);
}

int32_t zsmpm_implicit_multiply__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                                      const zsmpm_table* table, const zsmpm_grid* grid, float dt, const float* in,
                                      float* out) {
  if (!pol || !in || !out || in == out) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ParticlesView ps;
  TableView t;
  ZS_TRY(make_model(model, m));
  ZS_TRY(view_particles(particles, ps, true));
  if (m.kind == 1 && !ps.logJp) return ZSMPM_INVALID_ARGUMENT;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  DeviceGuard g(pol->devid);
  return run_implicit_multiply(pol, m, ps, t, grid, dt, in, out);
}

int32_t zsmpm_implicit_project__rocm(zsmpm_policy pol, const zsmpm_table* table, const zsmpm_grid* grid,
                                     const zsmpm_plane_collider* colliders, int32_t ncolliders, float* dof) {
  if (!pol || !dof) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  Colliders cols;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  ZS_TRY(make_colliders(colliders, ncolliders, cols));
  DeviceGuard g(pol->devid);
  return from_hip(implicit_project(grid->data, grid->capacity_blocks, t, grid->dx, cols, pol->gcols, dof, pol->stream));
}

int32_t zsmpm_implicit_precondition__rocm(zsmpm_policy pol, const zsmpm_table* table, const zsmpm_grid* grid,
                                          const float* in, float* out) {
  if (!pol || !in || !out) return ZSMPM_INVALID_ARGUMENT;
  TableView t;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  DeviceGuard g(pol->devid);
  return from_hip(implicit_precondition(grid->data, grid->capacity_blocks, t, in, out, pol->stream));
}

int32_t zsmpm_implicit_solve__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                                   const zsmpm_table* table, const zsmpm_grid* grid, float dt,
                                   const zsmpm_plane_collider* colliders, int32_t ncolliders, float* x, const float* b,
                                   int32_t max_iters, float tol, float rel_tol, int32_t* iters, float* residual) {
  if (!pol || !x || !b || max_iters < 0 || !(tol >= 0.f) || !(rel_tol >= 0.f)) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ParticlesView ps;
  TableView t;
  Colliders cols;
  ZS_TRY(make_model(model, m));
  ZS_TRY(view_particles(particles, ps, true));
  if (m.kind == 1 && !ps.logJp) return ZSMPM_INVALID_ARGUMENT;
  ZS_TRY(view_table(table, t));
  ZS_TRY(view_grid(grid));
  ZS_TRY(make_colliders(colliders, ncolliders, cols));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, 0, 0));
  hipStream_t s = pol->stream;
  // the solver's four vectors and a dot-product cell, sized by the grid
  const size_t nv = (size_t)grid->capacity_blocks * kBlockCells * 3;
  const size_t need = 4 * nv * sizeof(float) + 256;
  if (need > pol->cg_bytes) {
    ZS_HIP(hipStreamSynchronize(s));
    if (pol->cg_buf) (void)hipFree(pol->cg_buf);
    pol->cg_buf = nullptr;
    pol->cg_bytes = 0;
    ZS_HIP(hipMalloc(&pol->cg_buf, need));
    pol->cg_bytes = need;
  }
  float* r = (float*)pol->cg_buf;
  float* p = r + nv;
  float* q = p + nv;
  float* tmp = q + nv;
  double* dd = (double*)(tmp + nv);
  const float* gd = grid->data;
  const int32_t cap = grid->capacity_blocks;
  auto dot = [&](const float* a, const float* c, float* outv) -> int32_t {
    ZS_HIP(implicit_dot(t, cap, a, c, dd, s));
    double h = 0.0;
    ZS_HIP(hipMemcpyAsync(&h, dd, sizeof h, hipMemcpyDeviceToHost, s));
    ZS_HIP(hipStreamSynchronize(s));
    *outv = (float)h;
    return ZSMPM_OK;
  };
  // the particles and the table stay put during the solve: one bucket sort
  // and neighbour table serve every product
  ZS_TRY(ensure(pol, ps.n, t.T));
  SortView sv_solve = own_buckets(pol);
  ZS_TRY(run_buckets(pol, t, ps, grid->dx, sv_solve));
  ZS_HIP(block_neighbours(t, pol->ws, s));
  auto A = [&](const float* in, float* outv) -> int32_t {
    return run_implicit_multiply(pol, m, ps, t, grid, dt, in, outv, true);
  };
  auto project = [&](float* v) -> int32_t {
    return from_hip(implicit_project(gd, cap, t, grid->dx, cols, pol->gcols, v, s));
  };
  // ConjugateGradient::solve (math/linear/ConjugateGradient.hpp:70-175)
  ZS_TRY(A(x, tmp));
  ZS_HIP(implicit_lincomb(t, cap, 1.f, b, -1.f, tmp, r, s));  // r = b - A x
  ZS_TRY(project(r));
  ZS_HIP(implicit_fill(t, cap, q, 0.f, s));                   // q = 0 where the preconditioner leaves it
  ZS_HIP(implicit_precondition(gd, cap, t, r, q, s));
  ZS_HIP(implicit_lincomb(t, cap, 1.f, q, 0.f, q, p, s));     // p = q
  float zTrk = 0.f;
  ZS_TRY(dot(r, q, &zTrk));
  float norm = sqrtf(zTrk);
  const float localTol = fminf(rel_tol * norm, tol);
  int32_t it = 0;
  for (; it != max_iters; ++it) {
    if (norm <= localTol) break;
    ZS_TRY(A(p, tmp));
    ZS_TRY(project(tmp));
    float tp = 0.f;
    ZS_TRY(dot(tmp, p, &tp));
    const float alpha = zTrk / tp;
    ZS_HIP(implicit_lincomb(t, cap, alpha, p, 1.f, x, x, s));
    ZS_HIP(implicit_lincomb(t, cap, -alpha, tmp, 1.f, r, r, s));
    ZS_HIP(implicit_precondition(gd, cap, t, r, q, s));
    const float zLast = zTrk;
    ZS_TRY(dot(q, r, &zTrk));
    const float beta = zTrk / zLast;
    ZS_HIP(implicit_lincomb(t, cap, beta, p, 1.f, q, p, s));
    norm = sqrtf(zTrk);
  }
  if (iters) *iters = it;
  if (residual) *residual = norm;
  return ZSMPM_OK;
}

static int32_t view_bht(const zsmpm_bht* d, BhtView& t) {
  ZS_TRY(check_header(d));
  if (d->table_size < kBhtBucket || d->table_size % kBhtBucket != 0 || !d->keys || !d->indices || !d->status ||
      !d->cnt || !d->active_keys || !d->success || ((uintptr_t)d->keys % 16) != 0)
    return ZSMPM_INVALID_ARGUMENT;
  t.T = d->table_size;
  t.nb = d->table_size / kBhtBucket;
  t.keys = reinterpret_cast<int4*>(d->keys);
  t.indices = d->indices;
  t.status = d->status;
  t.cnt = d->cnt;
  t.active = d->active_keys;
  t.success = d->success;
  for (int i = 0; i < 6; ++i) t.h[i] = d->hash[i];
  return ZSMPM_OK;
}

static int32_t view_spg(const zsmpm_sparse_grid* d, SpgView& g) {
  ZS_TRY(check_header(d));
  ZS_TRY(view_bht(&d->table, g.t));
  if (d->num_channels < 1 || d->capacity_blocks < 0 || (d->capacity_blocks > 0 && !d->data))
    return ZSMPM_INVALID_ARGUMENT;
  g.data = d->data;
  g.nch = d->num_channels;
  memcpy(g.i2w, d->i2w, sizeof g.i2w);
  memcpy(g.w2i, d->w2i, sizeof g.w2i);
  g.background = d->background;
  return ZSMPM_OK;
}

// the policy's bht scratch, grown (after draining the stream) when too small
static int32_t bht_scratch(zsmpm_policy_s* pol, int64_t n, int32_t nb, void** out) {
  const size_t need = bht_scratch_bytes(n, nb);
  if (need > pol->bht_bytes) {
    ZS_HIP(hipStreamSynchronize(pol->stream));
    if (pol->bht_buf) (void)hipFree(pol->bht_buf);
    pol->bht_buf = nullptr;
    pol->bht_bytes = 0;
    ZS_HIP(hipMalloc(&pol->bht_buf, need));
    pol->bht_bytes = need;
  }
  *out = pol->bht_buf;
  return ZSMPM_OK;
}

int64_t zsmpm_bht_table_size(int64_t n) {
  if (n <= 0) return 0;
  int64_t p = 1;
  while (p < n) p <<= 1;
  const int64_t m = 2 * p;
  return m + (kBhtBucket - m % kBhtBucket);
}

void zsmpm_bht_hash_params(uint32_t seed, uint32_t out[6]) {
  if (!out) return;
  std::mt19937 rng(seed);
  for (int h = 0; h < 3; ++h) {
    uint32_t x = (uint32_t)rng() % kBhtPrime;
    if (x < 1) x = 1;
    out[2 * h] = x;
    out[2 * h + 1] = (uint32_t)rng() % kBhtPrime;
  }
}

int32_t zsmpm_bht_clear__rocm(zsmpm_policy pol, const zsmpm_bht* table) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  BhtView t;
  ZS_TRY(view_bht(table, t));
  DeviceGuard g(pol->devid);
  return from_hip(bht_clear(t, pol->stream));
}

int32_t zsmpm_bht_insert__rocm(zsmpm_policy pol, const zsmpm_bht* table, const int32_t* keys, int64_t n,
                               int32_t* out_index, int32_t flags) {
  if (!pol || n < 0 || (n > 0 && !keys) || (flags & ~ZSMPM_BHT_SERIAL)) return ZSMPM_INVALID_ARGUMENT;
  if (n > (int64_t)0x7fffffff) return ZSMPM_CAPACITY;
  BhtView t;
  ZS_TRY(view_bht(table, t));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, 0, 0));
  void* sc = nullptr;
  ZS_TRY(bht_scratch(pol, n, t.nb, &sc));
  return bht_insert(t, keys, n, out_index, (flags & ZSMPM_BHT_SERIAL) != 0, sc, pol->ws.err, pol->stream,
                    &pol->bht_rounds);
}

int32_t zsmpm_bht_query__rocm(zsmpm_policy pol, const zsmpm_bht* table, const int32_t* keys, int64_t n,
                              int32_t* out) {
  if (!pol || n < 0 || (n > 0 && (!keys || !out))) return ZSMPM_INVALID_ARGUMENT;
  BhtView t;
  ZS_TRY(view_bht(table, t));
  DeviceGuard g(pol->devid);
  return from_hip(bht_query_many(t, keys, n, out, pol->stream));
}

int32_t zsmpm_spg_activate__rocm(zsmpm_policy pol, const zsmpm_sparse_grid* grid, const float* x, int64_t n) {
  if (!pol || n < 0 || (n > 0 && !x)) return ZSMPM_INVALID_ARGUMENT;
  if (n > (int64_t)0x7fffffff) return ZSMPM_CAPACITY;
  SpgView gv;
  ZS_TRY(view_spg(grid, gv));
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, 0, 0));
  void* sc = nullptr;
  ZS_TRY(bht_scratch(pol, n, gv.t.nb, &sc));
  return spg_activate(gv, x, n, sc, pol->ws.err, pol->stream);
}

int32_t zsmpm_spg_sample__rocm(zsmpm_policy pol, const zsmpm_sparse_grid* grid, int32_t channel, int32_t width,
                               const float* x, int64_t n, float* out) {
  if (!pol || n < 0 || (n > 0 && (!x || !out)) || width < 1 || width > 3) return ZSMPM_INVALID_ARGUMENT;
  SpgView gv;
  ZS_TRY(view_spg(grid, gv));
  if (channel < 0 || channel + width > gv.nch) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  return from_hip(spg_sample(gv, channel, width, x, n, out, pol->stream));
}

int32_t zsmpm_policy_levelset_boundary__rocm(zsmpm_policy pol, const zsmpm_sparse_grid* grid, int32_t sdf_channel,
                                             int32_t v_channel, int32_t type) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  if (!grid) {
    pol->gcols.ls = LevelSetBoundary{};
    return ZSMPM_OK;
  }
  LevelSetBoundary ls;
  ZS_TRY(view_spg(grid, ls.g));
  if (sdf_channel < 0 || sdf_channel >= ls.g.nch || v_channel < -1 || (v_channel >= 0 && v_channel + 3 > ls.g.nch) ||
      type < ZSMPM_STICKY || type > ZSMPM_SEPARATE)
    return ZSMPM_INVALID_ARGUMENT;
  ls.on = 1;
  ls.sdf = sdf_channel;
  ls.vchn = v_channel;
  ls.type = type;
  // voxelSize(0): the length of row 0 of the transform's 3 x 3 part (SparseGrid.hpp:389-397)
  const float v0 = sqrtf(grid->i2w[0] * grid->i2w[0] + grid->i2w[1] * grid->i2w[1] + grid->i2w[2] * grid->i2w[2]);
  ls.eps = v0 / 4.f;
  pol->gcols.ls = ls;
  return ZSMPM_OK;
}

// The attributes a tile set names, checked against the arrays of p.
static int32_t make_tile_attrs(const zsmpm_tile_particles* t, const zsmpm_particles* p, TileAttrs& a) {
  ZS_TRY(check_header(t));
  ZS_TRY(check_header(p));
  if (t->n < 0 || t->n != p->n || !is_pow2(t->lane_width) || t->lane_width > 1024 || t->num_channels < 1 ||
      (t->n > 0 && !t->data) || (t->matrix_order != ZSMPM_COLUMN_MAJOR && t->matrix_order != ZSMPM_ROW_MAJOR))
    return ZSMPM_INVALID_ARGUMENT;
  const int32_t tr = t->matrix_order == ZSMPM_ROW_MAJOR ? 1 : 0;
  const TileAttr list[kTileAttrs] = {{t->x, 3, 0, p->x},     {t->v, 3, 0, p->v},
                                     {t->C, 9, tr, p->C},    {t->F, 9, tr, p->F},
                                     {t->m, 1, 0, p->m},     {t->logJp, 1, 0, p->logJp},
                                     {t->id, 1, 0, (float*)p->id}};
  std::vector<char> used((size_t)t->num_channels, 0);
  a = TileAttrs{};
  a.n = t->n;
  a.nch = t->num_channels;
  while ((1 << a.lane_shift) < t->lane_width) ++a.lane_shift;
  for (const TileAttr& x : list) {
    if (x.chn == -1) continue;
    if (x.chn < 0 || x.chn > t->num_channels - x.width || (t->n > 0 && !x.soa)) return ZSMPM_INVALID_ARGUMENT;
    for (int32_t c = x.chn; c < x.chn + x.width; ++c) {
      if (used[c]) return ZSMPM_INVALID_ARGUMENT;
      used[c] = 1;
    }
    a.attr[a.count++] = x;
  }
  return ZSMPM_OK;
}

int32_t zsmpm_particles_from_tiles__rocm(zsmpm_policy pol, const zsmpm_tile_particles* src,
                                         const zsmpm_particles* dst) {
  if (!pol || !src || !dst) return ZSMPM_INVALID_ARGUMENT;
  TileAttrs a;
  ZS_TRY(make_tile_attrs(src, dst, a));
  DeviceGuard g(pol->devid);
  return from_hip(tiles_convert(a, src->data, false, pol->stream));
}

int32_t zsmpm_particles_to_tiles__rocm(zsmpm_policy pol, const zsmpm_particles* src,
                                       const zsmpm_tile_particles* dst) {
  if (!pol || !src || !dst) return ZSMPM_INVALID_ARGUMENT;
  TileAttrs a;
  ZS_TRY(make_tile_attrs(dst, src, a));
  DeviceGuard g(pol->devid);
  return from_hip(tiles_convert(a, dst->data, true, pol->stream));
}

int32_t zsmpm_tiles_reorder__rocm(zsmpm_policy pol, const float* src, float* dst, const int32_t* map,
                                  int64_t num_tiles, int32_t lane_width, int32_t num_channels, int32_t scatter) {
  if (!pol || num_tiles < 0 || !is_pow2(lane_width) || lane_width > 1024 || num_channels < 1 ||
      (num_tiles > 0 && (!src || !dst || !map)))
    return ZSMPM_INVALID_ARGUMENT;
  const int64_t tf = (int64_t)lane_width * num_channels;
  if (num_tiles > 0 && overlaps(src, (size_t)(num_tiles * tf) * 4, dst, (size_t)(num_tiles * tf) * 4))
    return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, 0, 0));
  return from_hip(tiles_reorder(src, dst, map, num_tiles, tf, scatter != 0, pol->ws.err, pol->stream));
}

int32_t zsmpm_checkpoint_save__rocm(zsmpm_policy pol, const char* path, const zsmpm_particles* particles,
                                    const zsmpm_table* table, double time, int64_t step) {
  if (!pol || !path) return ZSMPM_INVALID_ARGUMENT;
  ParticlesView ps;
  ZS_TRY(view_particles(particles, ps, true));
  TableView t;
  if (table) ZS_TRY(view_table(table, t));
  DeviceGuard g(pol->devid);
  return checkpoint_save(path, ps, table ? &t : nullptr, time, step, pol->stream);
}

int32_t zsmpm_checkpoint_info__rocm(const char* path, zsmpm_checkpoint_info* out) {
  if (!path || !out) return ZSMPM_INVALID_ARGUMENT;
  return checkpoint_info(path, out);
}

int32_t zsmpm_checkpoint_load__rocm(zsmpm_policy pol, const char* path, const zsmpm_particles* particles,
                                    const zsmpm_table* table, zsmpm_checkpoint_info* out) {
  if (!pol || !path) return ZSMPM_INVALID_ARGUMENT;
  ParticlesView ps;
  ZS_TRY(view_particles(particles, ps, true));
  TableView t;
  if (table) ZS_TRY(view_table(table, t));
  DeviceGuard g(pol->devid);
  return checkpoint_load(path, ps, table ? &t : nullptr, out, pol->stream);
}

int32_t zsmpm_write_partio__rocm(zsmpm_policy pol, const char* path, const float* pos, int64_t n,
                                 const float* attr, int32_t attr_width, const char* attr_name) {
  if (!pol || !path || n < 0 || (n > 0 && !pos)) return ZSMPM_INVALID_ARGUMENT;
  if (attr && (!attr_name || !attr_name[0] || strlen(attr_name) > 0xffff || (attr_width != 1 && attr_width != 3)))
    return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  return write_bgeo(path, pos, n, attr, attr_width, attr_name, pol->stream);
}

int32_t zsmpm_substeps_batched__rocm(zsmpm_policy pol, const zsmpm_batch_scene* scenes, int32_t nscenes,
                                     const zsmpm_particles* a, const zsmpm_particles* b, const zsmpm_table* table,
                                     const zsmpm_buckets* buckets, const zsmpm_grid* grid, int32_t nsteps,
                                     int32_t flags, int32_t* final_in_b) {
  if (!pol || !scenes || !a || !b || !table || !grid || nsteps < 0 ||
      (flags & ~(ZSMPM_SUBSTEPS_GRAPH | ZSMPM_SUBSTEPS_AUTOTUNE)))
    return ZSMPM_INVALID_ARGUMENT;
  if (pol->closure_on) return ZSMPM_UNSUPPORTED_OPERATION;  // the learned closure serves single scenes
  ParticlesView pa, pb;
  ZS_TRY(view_particles(a, pa, true));
  ZS_TRY(view_particles(b, pb, true));
  ZS_TRY(check_out_particles(pa, pb));
  if (final_in_b) *final_in_b = nsteps & 1;
  DeviceGuard g(pol->devid);
  SceneSet ss;
  ZS_TRY(upload_scenes(pol, scenes, nscenes, pa, pb, ss));
  if (nsteps == 0) return ZSMPM_OK;
  ZS_TRY(ensure(pol, pa.n, table->table_size));
  return substeps_driver(pol, nullptr, a, b, table, buckets, grid, 0.f, nullptr, nullptr, 0, nsteps, flags, &ss);
}

int32_t zsmpm_policy_colliders__rocm(zsmpm_policy pol, const zsmpm_collider* colliders, int32_t n) {
  if (!pol || n < 0 || n > kMaxColliders || (n > 0 && !colliders)) return ZSMPM_INVALID_ARGUMENT;
  GeneralColliders g{};
  for (int32_t i = 0; i < n; ++i) {
    const zsmpm_collider& c = colliders[i];
    if (c.type < ZSMPM_STICKY || c.type > ZSMPM_SEPARATE || c.geometry < ZSMPM_GEOMETRY_PLANE ||
        c.geometry > ZSMPM_GEOMETRY_CYLINDER || c.axis < 0 || c.axis > 2 || !(c.s > 0.f) || !isfinite(c.s) ||
        !isfinite(c.dsdt))
      return ZSMPM_INVALID_ARGUMENT;
    const float* vals[] = {c.p0, c.p1, c.omega, c.b, c.dbdt};
    for (const float* v : vals)
      for (int d = 0; d < 3; ++d)
        if (!isfinite(v[d])) return ZSMPM_INVALID_ARGUMENT;
    for (int k = 0; k < 9; ++k)
      if (!isfinite(c.R[k])) return ZSMPM_INVALID_ARGUMENT;
    GeneralCollider& o = g.c[i];
    o.type = c.type;
    o.geometry = c.geometry;
    o.axis = c.axis;
    memcpy(o.p0, c.p0, sizeof o.p0);
    memcpy(o.p1, c.p1, sizeof o.p1);
    memcpy(o.R, c.R, sizeof o.R);
    memcpy(o.omega, c.omega, sizeof o.omega);
    o.s = c.s;
    o.dsdt = c.dsdt;
    memcpy(o.b, c.b, sizeof o.b);
    memcpy(o.dbdt, c.dbdt, sizeof o.dbdt);
  }
  g.n = n;
  g.ls = pol->gcols.ls;  // the level-set boundary stays
  pol->gcols = g;
  return ZSMPM_OK;
}

// P2G takes buckets from a work queue and G2P strides over them, so any
// positive workgroup count is correct; the bound keeps a grid launchable.
constexpr int32_t kMaxLaunchGroups = 1 << 20;

int32_t zsmpm_policy_launch_widths__rocm(zsmpm_policy pol, int32_t* p2g_groups, int32_t* g2p_groups) {
  if (!pol || !p2g_groups || !g2p_groups) return ZSMPM_INVALID_ARGUMENT;
  *p2g_groups = pol->tuned_p2g > 0 ? pol->tuned_p2g : persistent_groups();
  *g2p_groups = pol->tuned_g2p > 0 ? pol->tuned_g2p : persistent_groups();
  return ZSMPM_OK;
}

int32_t zsmpm_policy_set_launch_widths__rocm(zsmpm_policy pol, int32_t p2g_groups, int32_t g2p_groups) {
  if (!pol || p2g_groups < 0 || g2p_groups < 0 || p2g_groups > kMaxLaunchGroups || g2p_groups > kMaxLaunchGroups)
    return ZSMPM_INVALID_ARGUMENT;
  pol->tuned_p2g = p2g_groups;
  pol->tuned_g2p = g2p_groups;
  return ZSMPM_OK;
}

int32_t zsmpm_policy_transfer_scheme__rocm(zsmpm_policy pol, int32_t scheme) {
  if (!pol || (scheme != ZSMPM_TRANSFER_APIC && scheme != ZSMPM_TRANSFER_CELL)) return ZSMPM_INVALID_ARGUMENT;
  pol->transfer = scheme;
  return ZSMPM_OK;
}

int32_t zsmpm_policy_launch_tiles__rocm(zsmpm_policy pol, int32_t* p2g_chunk, int32_t* g2p_threads) {
  if (!pol || !p2g_chunk || !g2p_threads) return ZSMPM_INVALID_ARGUMENT;
  *p2g_chunk = pol->tuned_p2g_tile == 1 ? 256 : 576;
  *g2p_threads = pol->tuned_g2p_threads == 256 ? 256 : 128;
  return ZSMPM_OK;
}

int32_t zsmpm_policy_set_launch_tiles__rocm(zsmpm_policy pol, int32_t p2g_chunk, int32_t g2p_threads) {
  if (!pol || (p2g_chunk != 0 && p2g_chunk != 576 && p2g_chunk != 256) ||
      (g2p_threads != 0 && g2p_threads != 128 && g2p_threads != 256))
    return ZSMPM_INVALID_ARGUMENT;
  pol->tuned_p2g_tile = p2g_chunk == 256 ? 1 : 0;
  pol->tuned_g2p_threads = g2p_threads == 256 ? 256 : 0;
  return ZSMPM_OK;
}

int32_t zsmpm_policy_closure__rocm(zsmpm_policy pol, const zsmpm_closure* closure) {
  if (!pol) return ZSMPM_INVALID_ARGUMENT;
  if (!closure) {
    pol->closure_on = false;
    pol->closure = ClosureView{};
    return ZSMPM_OK;
  }
  ZS_TRY(check_header(closure));
  if (!(closure->tol > 0.f)) return ZSMPM_INVALID_ARGUMENT;
  if (closure->flags & ~ZSMPM_CLOSURE_ALWAYS_NETWORK) return ZSMPM_INVALID_ARGUMENT;
  pol->closure.params = closure->params;
  pol->closure.tol = closure->tol;
  pol->closure.always_net = (closure->flags & ZSMPM_CLOSURE_ALWAYS_NETWORK) ? 1 : 0;
  pol->closure_on = true;
  return ZSMPM_OK;
}

int32_t zsmpm_stress__rocm(zsmpm_policy pol, const zsmpm_model* model, const float* F, float* logJp, int64_t n,
                           float* PF) {
  if (!pol || n < 0 || (n > 0 && (!F || !PF))) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ZS_TRY(make_model(model, m));
  if (m.kind == ZSMPM_DRUCKER_PRAGER && n > 0 && !logJp) return ZSMPM_INVALID_ARGUMENT;
  if (n == 0) return ZSMPM_OK;
  DeviceGuard g(pol->devid);
  ZS_TRY(ensure(pol, n, pol->reserved_T));
  ParticlesView ps;
  ps.n = n;
  ps.F = const_cast<float*>(F);
  ps.logJp = logJp;
  StageScope sc(pol, ZSMPM_STAGE_STRESS);
  return from_hip(zsmpm::stress(m, ps, pol->ws, pol->stream, pol->closure_on ? &pol->closure : nullptr, PF));
}

int32_t zsmpm_mlp_harvest__rocm(zsmpm_policy pol, const zsmpm_model* model, const float* F, const float* logJp,
                                int64_t n, float* features, float* records) {
  if (!pol || n < 0 || (n > 0 && (!F || !features || !records))) return ZSMPM_INVALID_ARGUMENT;
  Model m;
  ZS_TRY(make_model(model, m));
  if (m.kind != ZSMPM_DRUCKER_PRAGER || m.mu == 0.f) return ZSMPM_UNSUPPORTED_OPERATION;
  DeviceGuard g(pol->devid);
  return from_hip(mlp_harvest(m, F, logJp, n, features, records, pol->stream));
}

int32_t zsmpm_mlp_infer__rocm(zsmpm_policy pol, const float* params, const float* features, int64_t n,
                              float* outputs) {
  if (!pol || n < 0 || (n > 0 && (!params || !features || !outputs))) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  return from_hip(mlp_infer(params, features, n, outputs, pol->stream));
}

namespace {
int32_t mlp_pass(zsmpm_policy_s* pol, float* params, float* adam_m, float* adam_v, const float* features,
                 const float* records, int64_t n, float lr, int32_t step, float* grad, float* loss) {
  if (!pol || n <= 0 || !params || !features || !records) return ZSMPM_INVALID_ARGUMENT;
  DeviceGuard g(pol->devid);
  if (!pol->mlp_partial) {
    ZS_HIP(alloc(pol, &pol->mlp_partial, (size_t)kMlpGroups * kMlpParams));
    ZS_HIP(alloc(pol, &pol->mlp_loss, (size_t)kMlpGroups));
  }
  const int64_t tiles = (n + 15) / 16;
  const int G = (int)(tiles < kMlpGroups ? tiles : kMlpGroups);
  ZS_HIP(mlp_grad(params, features, records, n, pol->mlp_partial, pol->mlp_loss, G, pol->stream));
  float bc1 = 1.f, bc2 = 1.f;
  if (adam_m) {
    bc1 = (float)(1.0 - pow(0.9, (double)step));
    bc2 = (float)(1.0 - pow(0.999, (double)step));
  }
  return from_hip(mlp_reduce(adam_m ? params : nullptr, adam_m, adam_v, pol->mlp_partial, pol->mlp_loss, G, lr, bc1,
                             bc2, grad, loss, pol->stream));
}
}  // namespace

int32_t zsmpm_mlp_grad__rocm(zsmpm_policy pol, const float* params, const float* features, const float* records,
                             int64_t n, float* grad, float* loss) {
  if (!grad) return ZSMPM_INVALID_ARGUMENT;
  return mlp_pass(pol, const_cast<float*>(params), nullptr, nullptr, features, records, n, 0.f, 0, grad, loss);
}

int32_t zsmpm_mlp_train_step__rocm(zsmpm_policy pol, float* params, float* adam_m, float* adam_v,
                                   const float* features, const float* records, int64_t n, float lr, int32_t step,
                                   float* loss) {
  if (!adam_m || !adam_v || step < 1 || !(lr > 0.f)) return ZSMPM_INVALID_ARGUMENT;
  return mlp_pass(pol, params, adam_m, adam_v, features, records, n, lr, step, nullptr, loss);
}

}  // extern "C"
