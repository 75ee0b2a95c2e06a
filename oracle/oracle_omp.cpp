// zsmpm oracle, OpenMP build: TEST INFRASTRUCTURE ONLY. It is the host-core
// baseline timed beside the HIP path (the cpu_baseline leg of bench.py).
//
// It runs the same substep as zsmpm_oracle.hpp with zpc's OpenMP execution
// policy choices (omp/execution/ExecutionPolicy.hpp:33-135): every stage is a
// `#pragma omp parallel for` over particles, table slots or grid blocks;
// inserts take the per-slot spin lock of HashTableView::atomicKeyCAS (host,
// container/HashTable.hpp:546-564) and an atomic counter for the block
// number (:385); P2G scatters with the CAS-loop float atomic_add
// (execution/Atomics.hpp:72-85); the velocity maximum uses the CAS-loop
// atomic_max (:336-350); exclusive_scan is the two-pass per-thread scan
// (omp/execution/ExecutionPolicy.hpp:341-419). Block numbers and in-bucket
// order are therefore nondeterministic, as in the reference.
#include <omp.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <thread>
#include <vector>

#include "zsmpm_oracle.hpp"

using namespace zso;

extern "C" {
typedef struct {
  double dt, dx;
  double extf[3];
  int32_t mat_kind;
  double volume, E, nu, cohesion, beta, yield_surface;
  int32_t volume_correction;
  int32_t table_size;
  int32_t ncols;
  double col_origin[8][3];
  double col_normal[8][3];
  int32_t col_type[8];
} zso_params;
}

namespace {

inline float atomic_add_f32(float* dest, float val) {
  uint32_t* p = reinterpret_cast<uint32_t*>(dest);
  uint32_t oldv = __atomic_load_n(p, __ATOMIC_RELAXED);
  for (;;) {
    float of;
    std::memcpy(&of, &oldv, 4);
    float nf = of + val;
    uint32_t nv;
    std::memcpy(&nv, &nf, 4);
    if (__atomic_compare_exchange_n(p, &oldv, nv, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST))
      return of;
  }
}

inline void atomic_max_f32(float* dest, float val) {
  uint32_t* p = reinterpret_cast<uint32_t*>(dest);
  uint32_t oldv = __atomic_load_n(p, __ATOMIC_RELAXED);
  for (;;) {
    float of;
    std::memcpy(&of, &oldv, 4);
    if (!(of < val)) return;
    uint32_t nv;
    std::memcpy(&nv, &val, 4);
    if (__atomic_compare_exchange_n(p, &oldv, nv, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) return;
  }
}

struct ConcurrentTable {
  int32_t T = 0;
  std::vector<int32_t> keys, indices, status, active;
  int32_t cnt = 0;

  void resize(int32_t t) {
    T = t;
    keys.assign((size_t)3 * T, kKeySentinel);
    indices.assign(T, -1);
    status.assign(T, -1);
    active.assign((size_t)3 * T, 0);
  }
  /// atomicKeyCAS (host), HashTable.hpp:546-564: spin on the slot lock,
  /// read the stored key, store ours if the slot is empty, unlock.
  void key_cas(int32_t s, const int32_t k[3], int32_t out[3]) {
    for (;;) {
      int32_t expect = -1;
      if (__atomic_compare_exchange_n(&status[s], &expect, 0, false, __ATOMIC_SEQ_CST,
                                      __ATOMIC_SEQ_CST)) {
        volatile int32_t* kp = &keys[3 * s];
        out[0] = kp[0];
        out[1] = kp[1];
        out[2] = kp[2];
        if (out[0] == kKeySentinel && out[1] == kKeySentinel && out[2] == kKeySentinel) {
          kp[0] = k[0];
          kp[1] = k[1];
          kp[2] = k[2];
        }
        __atomic_exchange_n(&status[s], -1, __ATOMIC_SEQ_CST);
        return;
      }
    }
  }
  /// HashTableView::insert (host), HashTable.hpp:380-398.
  int32_t insert(const int32_t k[3]) {
    int32_t s = home_slot(k[0], k[1], k[2], T);
    int32_t stored[3];
    key_cas(s, k, stored);
    auto is_sent = [](const int32_t* v) {
      return v[0] == kKeySentinel && v[1] == kKeySentinel && v[2] == kKeySentinel;
    };
    auto is_key = [&](const int32_t* v) { return v[0] == k[0] && v[1] == k[1] && v[2] == k[2]; };
    while (!(is_sent(stored) || is_key(stored))) {
      s = (s + kProbeStep) % T;
      key_cas(s, k, stored);
    }
    if (is_sent(stored)) {
      int32_t localno = __atomic_fetch_add(&cnt, 1, __ATOMIC_SEQ_CST);
      indices[s] = localno;
      active[3 * localno] = k[0];
      active[3 * localno + 1] = k[1];
      active[3 * localno + 2] = k[2];
      return localno;
    }
    return -1;
  }
  /// HashTableView::query, HashTable.hpp:447-456 (modular wrap).
  int32_t query(const int32_t k[3]) const {
    int32_t s = home_slot(k[0], k[1], k[2], T);
    for (;;) {
      const int32_t* kp = &keys[3 * s];
      if (kp[0] == k[0] && kp[1] == k[1] && kp[2] == k[2]) return indices[s];
      if (indices[s] == -1) return -1;
      s = (int32_t)(((int64_t)s + kProbeStep) % T);
    }
  }
};

/// exclusive_scan, omp/execution/ExecutionPolicy.hpp:341-419: per-thread
/// partial sums, a serial scan of those, then a per-thread local scan.
void omp_exclusive_scan(const int32_t* in, int32_t* out, int64_t n, int nthreads) {
  std::vector<int64_t> part(nthreads + 1, 0);
#pragma omp parallel num_threads(nthreads)
  {
    int tid = omp_get_thread_num(), nt = omp_get_num_threads();
    int64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += in[i];
    part[tid + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int t = 1; t <= nt; ++t) part[t] += part[t - 1];
    int64_t run = part[tid];
    for (int64_t i = lo; i < hi; ++i) {
      int32_t v = in[i];
      out[i] = (int32_t)run;
      run += v;
    }
  }
}

struct OmpSim {
  StepParams<float> sp;
  std::vector<PlaneCollider<float>> cols;
  ConcurrentTable table;
  std::vector<int32_t> counts, tmp, offsets, indices;
  std::vector<float> grid;
  int nthreads = 1;

  void substep(ParticlesView<float> ps) {
    const int64_t n = ps.n;
    ConcurrentTable& t = table;
    const int nt = nthreads;
    if (t.T != sp.tableSize) t.resize(sp.tableSize);
    // CleanSparsity over range(tableSize)
#pragma omp parallel for num_threads(nt)
    for (int64_t e = 0; e < t.T; ++e) {
      t.keys[3 * e] = t.keys[3 * e + 1] = t.keys[3 * e + 2] = kKeySentinel;
      t.indices[e] = -1;
      t.status[e] = -1;
    }
    t.cnt = 0;
    // ComputeSparsity over range(N)
    const float dxinv = (float)1.0 / sp.dx;
#pragma omp parallel for num_threads(nt)
    for (int64_t p = 0; p < n; ++p) {
      int32_t k[3];
      block_key(ps.x + 3 * p, dxinv, 4, -2, 0.5f, k);
      t.insert(k);
    }
    // EnlargeSparsity over range(table.size())
    const int32_t n0 = t.cnt;
#pragma omp parallel for num_threads(nt)
    for (int32_t i = 0; i < n0; ++i) {
      const int32_t b0 = t.active[3 * i], b1 = t.active[3 * i + 1], b2 = t.active[3 * i + 2];
      for (int32_t dx = 0; dx < 2; ++dx)
        for (int32_t dy = 0; dy < 2; ++dy)
          for (int32_t dz = 0; dz < 2; ++dz) {
            int32_t k[3] = {b0 + dx, b1 + dy, b2 + dz};
            t.insert(k);
          }
    }
    const int64_t nb = t.cnt;
    // index buckets: SpatiallyCount, exclusive_scan, SpatiallyDistribute
    counts.assign(nb + 1, 0);
    tmp.assign(nb + 1, 0);
    offsets.assign(nb + 1, 0);
    indices.assign(n, 0);
#pragma omp parallel for num_threads(nt)
    for (int64_t p = 0; p < n; ++p) {
      int32_t k[3];
      block_key(ps.x + 3 * p, dxinv, 4, -2, 0.5f, k);
      __atomic_fetch_add(&counts[t.query(k)], 1, __ATOMIC_SEQ_CST);
    }
    omp_exclusive_scan(counts.data(), offsets.data(), nb + 1, nt);
#pragma omp parallel for num_threads(nt)
    for (int64_t p = 0; p < n; ++p) {
      int32_t k[3];
      block_key(ps.x + 3 * p, dxinv, 4, -2, 0.5f, k);
      int32_t b = t.query(k);
      int32_t dst = __atomic_fetch_add(&tmp[b], 1, __ATOMIC_SEQ_CST);
      indices[offsets[b] + dst] = (int32_t)p;
    }
    // CleanGridBlocks over Collapse{nb, 64}
    grid.resize((size_t)nb * kChannels * 64);
#pragma omp parallel for num_threads(nt)
    for (int64_t b = 0; b < nb; ++b)
      for (int e = 0; e < kChannels * 64; ++e) grid[b * kChannels * 64 + e] = 0.f;
    // P2GTransfer over range(N) with atomic scatter
    const float dx = sp.dx, dt = sp.dt;
    const float dx_inv = (float)1.0 / dx;
    const float D_inv = 4.f * dx_inv * dx_inv;
    float mu, lambda;
    lame(sp.mat.E, sp.mat.nu, mu, lambda);
#pragma omp parallel for num_threads(nt)
    for (int64_t p = 0; p < n; ++p) {
      const float* pos = ps.x + 3 * p;
      const float* vel = ps.v + 3 * p;
      const float mass = ps.m[p];
      const float* C = ps.C + 9 * p;
      float contrib[9];
      float logJp = ps.logJp ? ps.logJp[p] : 0.f;
      p2g_stress(sp.mat, mu, lambda, ps.F + 9 * p, logJp, contrib, dt, D_inv);
      if (sp.mat.kind == kDruckerPrager) ps.logJp[p] = logJp;
      Arena<float> ar;
      ar.init(dx, pos);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            int32_t loc[3], key[3];
            const int32_t coord[3] = {ar.corner[0] + i, ar.corner[1] + j, ar.corner[2] + k};
            for (int d = 0; d < 3; ++d) {
              loc[d] = coord[d] & 3;
              key[d] = (coord[d] - loc[d]) / 4;
            }
            const int64_t b = t.query(key);
            const int cell = (loc[0] << 4) | (loc[1] << 2) | loc[2];
            float xixp[3];
            ar.diff(i, j, k, xixp);
            const float W = ar.weight(i, j, k);
            atomic_add_f32(&grid[gidx(b, 0, cell)], mass * W);
            for (int d = 0; d < 3; ++d) {
              atomic_add_f32(&grid[gidx(b, 1 + d, cell)],
                             W * mass * (vel[d] + (C[d] * xixp[0] + C[3 + d] * xixp[1] + C[6 + d] * xixp[2])));
              atomic_add_f32(&grid[gidx(b, 4 + d, cell)],
                             (contrib[d] * xixp[0] + contrib[3 + d] * xixp[1] + contrib[6 + d] * xixp[2]) * W);
            }
          }
    }
    // fold rhs, ComputeGridBlockVelocity, ApplyBoundaryConditionOnGridBlocks
    float maxVel = 0.f;
#pragma omp parallel for num_threads(nt)
    for (int64_t b = 0; b < nb; ++b)
      for (int c = 0; c < 64; ++c) {
        for (int d = 0; d < 3; ++d) grid[gidx(b, 1 + d, c)] = grid[gidx(b, 1 + d, c)] + grid[gidx(b, 4 + d, c)];
        float mass = grid[gidx(b, 0, c)];
        if (mass != 0.f) {
          mass = 1.f / mass;
          float v[3];
          for (int d = 0; d < 3; ++d) v[d] = grid[gidx(b, 1 + d, c)] * mass + sp.extf[d] * dt;
          for (int d = 0; d < 3; ++d) grid[gidx(b, 1 + d, c)] = v[d];
          float r = 0.f;
          r += v[0] * v[0];
          r += v[1] * v[1];
          r += v[2] * v[2];
          atomic_max_f32(&maxVel, r);
        }
      }
    for (int ci = 0; ci < sp.ncols; ++ci) {
      const PlaneCollider<float>& col = sp.cols[ci];
#pragma omp parallel for num_threads(nt)
      for (int64_t b = 0; b < nb; ++b)
        for (int c = 0; c < 64; ++c) {
          if (!(grid[gidx(b, 0, c)] > 0)) continue;
          const int32_t cc[3] = {(c >> 4) & 3, (c >> 2) & 3, c & 3};
          float pos[3];
          for (int d = 0; d < 3; ++d) pos[d] = ((float)t.active[3 * b + d] * 4.f + (float)cc[d]) * dx;
          float phi = col.normal[0] * (pos[0] - col.origin[0]) + col.normal[1] * (pos[1] - col.origin[1]) +
                      col.normal[2] * (pos[2] - col.origin[2]);
          if (phi < 0.f) {
            float v[3] = {grid[gidx(b, 1, c)], grid[gidx(b, 2, c)], grid[gidx(b, 3, c)]};
            if (col.type == kSticky) {
              v[0] = v[1] = v[2] = 0.f;
            } else {
              float proj = col.normal[0] * v[0] + col.normal[1] * v[1] + col.normal[2] * v[2];
              if ((col.type == kSeparate && proj < 0) || col.type == kSlip)
                for (int d = 0; d < 3; ++d) v[d] -= proj * col.normal[d];
            }
            for (int d = 0; d < 3; ++d) grid[gidx(b, 1 + d, c)] = v[d];
          }
        }
    }
    // G2PTransfer over range(N)
#pragma omp parallel for num_threads(nt)
    for (int64_t p = 0; p < n; ++p) {
      float* pos = ps.x + 3 * p;
      float vel[3] = {0, 0, 0};
      float C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      Arena<float> ar;
      ar.init(dx, pos);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            int32_t loc[3], key[3];
            const int32_t coord[3] = {ar.corner[0] + i, ar.corner[1] + j, ar.corner[2] + k};
            for (int d = 0; d < 3; ++d) {
              loc[d] = coord[d] & 3;
              key[d] = (coord[d] - loc[d]) / 4;
            }
            const int64_t b = t.query(key);
            const int cell = (loc[0] << 4) | (loc[1] << 2) | loc[2];
            float xixp[3];
            ar.diff(i, j, k, xixp);
            const float W = ar.weight(i, j, k);
            const float vi[3] = {grid[gidx(b, 1, cell)], grid[gidx(b, 2, cell)], grid[gidx(b, 3, cell)]};
            for (int d = 0; d < 3; ++d) vel[d] += vi[d] * W;
            for (int d = 0; d < 9; ++d) C[d] += W * vi[d % 3] * xixp[d / 3] * D_inv;
          }
      for (int d = 0; d < 3; ++d) pos[d] += vel[d] * dt;
      float tmpm[9], F[9];
      for (int d = 0; d < 9; ++d) tmpm[d] = C[d] * dt + ((d & 0x3) ? 0.f : 1.f);
      matmul3(tmpm, ps.F + 9 * p, F);
      for (int d = 0; d < 9; ++d) ps.F[9 * p + d] = F[d];
      for (int d = 0; d < 3; ++d) ps.v[3 * p + d] = vel[d];
      for (int d = 0; d < 9; ++d) ps.C[9 * p + d] = C[d];
    }
  }
};

}  // namespace

extern "C" {

/// zpc's default OpenMP thread count, hardware_concurrency - 1
/// (omp/execution/ExecutionPolicy.hpp:1245-1252).
int32_t zso_omp_default_threads() {
  unsigned c = std::thread::hardware_concurrency();
  return c > 1 ? (int32_t)(c - 1) : 1;
}

/// Run `warmup` untimed and `steps` timed substeps in place with `nthreads`
/// OpenMP threads; returns the wall seconds of the timed substeps and the
/// final active block count in *nblocks.
double zso_omp_run_f32(const zso_params* prm, int32_t warmup, int32_t steps, int32_t nthreads,
                       int64_t n, float* x, float* v, float* C, float* F, float* m, float* logJp,
                       int64_t* nblocks) {
  OmpSim sim;
  sim.nthreads = nthreads > 0 ? nthreads : zso_omp_default_threads();
  sim.sp.dt = (float)prm->dt;
  sim.sp.dx = (float)prm->dx;
  for (int d = 0; d < 3; ++d) sim.sp.extf[d] = (float)prm->extf[d];
  sim.sp.mat.kind = prm->mat_kind;
  sim.sp.mat.volume = (float)prm->volume;
  sim.sp.mat.E = (float)prm->E;
  sim.sp.mat.nu = (float)prm->nu;
  sim.sp.mat.cohesion = (float)prm->cohesion;
  sim.sp.mat.beta = (float)prm->beta;
  sim.sp.mat.yieldSurface = (float)prm->yield_surface;
  sim.sp.mat.volumeCorrection = prm->volume_correction != 0;
  sim.sp.tableSize = prm->table_size;
  sim.cols.resize(prm->ncols);
  for (int c = 0; c < prm->ncols; ++c) {
    for (int d = 0; d < 3; ++d) {
      sim.cols[c].origin[d] = (float)prm->col_origin[c][d];
      sim.cols[c].normal[d] = (float)prm->col_normal[c][d];
    }
    sim.cols[c].type = prm->col_type[c];
  }
  sim.sp.cols = sim.cols.data();
  sim.sp.ncols = prm->ncols;
  ParticlesView<float> ps;
  ps.n = n;
  ps.x = x;
  ps.v = v;
  ps.C = C;
  ps.F = F;
  ps.m = m;
  ps.logJp = logJp;
  for (int s = 0; s < warmup; ++s) sim.substep(ps);
  auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; ++s) sim.substep(ps);
  auto t1 = std::chrono::steady_clock::now();
  if (nblocks) *nblocks = sim.table.cnt;
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
