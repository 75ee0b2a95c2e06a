// zsmpm oracle C API: TEST INFRASTRUCTURE ONLY (ctypes entry points over
// zsmpm_oracle.hpp for tests/, smoke() and the cpu_baseline leg of bench.py).
// Sequential semantics; see zsmpm_oracle.hpp for the reference citations.
#include <cstdint>
#include <cstring>
#include <memory>

#include "zsmpm_oracle.hpp"
#include "bht_oracle.hpp"

using namespace zso;

extern "C" {

/// Scene/step parameters shared by the f32 and f64 entry points. The f32
/// entry points round each field to float once, the same rounding the
/// product's C ABI applies to its float arguments.
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
  int32_t ngcols;  /* general colliders, applied after the planes */
  struct {
    int32_t type, geometry, axis;
    double p0[3], p1[3], R[9], omega[3], s, dsdt, b[3], dbdt[3];
  } gcols[8];
} zso_params;
}

namespace {

template <typename T>
struct ColStore {
  std::vector<PlaneCollider<T>> planes;
  std::vector<GeneralCollider<T>> general;
};

template <typename T>
StepParams<T> to_params(const zso_params* p, ColStore<T>& store) {
  std::vector<PlaneCollider<T>>& cols = store.planes;
  StepParams<T> sp;
  sp.dt = (T)p->dt;
  sp.dx = (T)p->dx;
  for (int d = 0; d < 3; ++d) sp.extf[d] = (T)p->extf[d];
  sp.mat.kind = p->mat_kind;
  sp.mat.volume = (T)p->volume;
  sp.mat.E = (T)p->E;
  sp.mat.nu = (T)p->nu;
  sp.mat.cohesion = (T)p->cohesion;
  sp.mat.beta = (T)p->beta;
  sp.mat.yieldSurface = (T)p->yield_surface;
  sp.mat.volumeCorrection = p->volume_correction != 0;
  sp.tableSize = p->table_size;
  cols.resize(p->ncols);
  for (int c = 0; c < p->ncols; ++c) {
    for (int d = 0; d < 3; ++d) {
      cols[c].origin[d] = (T)p->col_origin[c][d];
      cols[c].normal[d] = (T)p->col_normal[c][d];
    }
    cols[c].type = p->col_type[c];
  }
  sp.cols = cols.data();
  sp.ncols = p->ncols;
  std::vector<GeneralCollider<T>>& gc = store.general;
  gc.resize(p->ngcols);
  for (int c = 0; c < p->ngcols; ++c) {
    const auto& q = p->gcols[c];
    gc[c].type = q.type;
    gc[c].geometry = q.geometry;
    gc[c].axis = q.axis;
    for (int d = 0; d < 3; ++d) {
      gc[c].p0[d] = (T)q.p0[d];
      gc[c].p1[d] = (T)q.p1[d];
      gc[c].omega[d] = (T)q.omega[d];
      gc[c].b[d] = (T)q.b[d];
      gc[c].dbdt[d] = (T)q.dbdt[d];
    }
    for (int k = 0; k < 9; ++k) gc[c].R[k] = (T)q.R[k];
    gc[c].s = (T)q.s;
    gc[c].dsdt = (T)q.dsdt;
  }
  sp.gcols = gc.data();
  sp.ngcols = p->ngcols;
  return sp;
}

template <typename T>
struct Ctx {
  StepOutputs<T> out;
  ColStore<T> cols;
};

template <typename T>
ParticlesView<T> view(int64_t n, T* x, T* v, T* C, T* F, T* m, T* logJp) {
  ParticlesView<T> ps;
  ps.n = n;
  ps.x = x;
  ps.v = v;
  ps.C = C;
  ps.F = F;
  ps.m = m;
  ps.logJp = logJp;
  return ps;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- math pieces
int32_t zso_hash(int32_t k0, int32_t k1, int32_t k2) { return do_hash(k0, k1, k2); }
int32_t zso_home_slot(int32_t k0, int32_t k1, int32_t k2, int32_t T) { return home_slot(k0, k1, k2, T); }
int64_t zso_evaluate_table_size(int64_t n) { return evaluate_table_size(n); }

void zso_svd_f32(int64_t n, const float* F, float* U, float* S, float* V) {
  for (int64_t i = 0; i < n; ++i) svd_colmajor(F + 9 * i, U + 9 * i, S + 3 * i, V + 9 * i);
}
void zso_svd_f64(int64_t n, const double* F, double* U, double* S, double* V) {
  for (int64_t i = 0; i < n; ++i) svd_colmajor(F + 9 * i, U + 9 * i, S + 3 * i, V + 9 * i);
}
void zso_lame_f32(float E, float nu, float* mu, float* lam) { lame(E, nu, *mu, *lam); }
void zso_lame_f64(double E, double nu, double* mu, double* lam) { lame(E, nu, *mu, *lam); }

void zso_stress_fc_f32(int64_t n, float vol, float mu, float lam, const float* F, float* PF) {
  for (int64_t i = 0; i < n; ++i) stress_fixedcorotated(vol, mu, lam, F + 9 * i, PF + 9 * i);
}
void zso_stress_fc_f64(int64_t n, double vol, double mu, double lam, const double* F, double* PF) {
  for (int64_t i = 0; i < n; ++i) stress_fixedcorotated(vol, mu, lam, F + 9 * i, PF + 9 * i);
}
void zso_stress_sand_f32(int64_t n, float vol, float mu, float lam, float coh, float beta, float ys,
                         int32_t volcorr, float* logJp, const float* F, float* PF, float* Fproj,
                         int32_t* which) {
  for (int64_t i = 0; i < n; ++i) {
    int w = stress_sand(vol, mu, lam, coh, beta, ys, volcorr != 0, logJp[i], F + 9 * i, PF + 9 * i,
                        Fproj ? Fproj + 9 * i : nullptr);
    if (which) which[i] = w;
  }
}
void zso_stress_sand_f64(int64_t n, double vol, double mu, double lam, double coh, double beta,
                         double ys, int32_t volcorr, double* logJp, const double* F, double* PF,
                         double* Fproj, int32_t* which) {
  for (int64_t i = 0; i < n; ++i) {
    int w = stress_sand(vol, mu, lam, coh, beta, ys, volcorr != 0, logJp[i], F + 9 * i, PF + 9 * i,
                        Fproj ? Fproj + 9 * i : nullptr);
    if (which) which[i] = w;
  }
}

/// LocalArena for n positions: corner (3), weights (27 in ndrange order),
/// diffs (27 x 3).
void zso_arena_f32(int64_t n, float dx, const float* x, int32_t* corner, float* w, float* diff) {
  for (int64_t p = 0; p < n; ++p) {
    Arena<float> ar;
    ar.init(dx, x + 3 * p);
    for (int d = 0; d < 3; ++d) corner[3 * p + d] = ar.corner[d];
    int e = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k, ++e) {
          w[27 * p + e] = ar.weight(i, j, k);
          ar.diff(i, j, k, diff + 81 * p + 3 * e);
        }
  }
}

void zso_block_keys_f32(int64_t n, const float* x, float dx, int32_t blockLen, int32_t offset,
                        float disp, int32_t* keys) {
  float dxinv = (float)1.0 / dx;
  for (int64_t p = 0; p < n; ++p) block_key(x + 3 * p, dxinv, blockLen, offset, disp, keys + 3 * p);
}

// ---------------------------------------------------------------- contexts
void* zso_ctx_new_f32() { return new Ctx<float>(); }
void* zso_ctx_new_f64() { return new Ctx<double>(); }
/// The general collider `idx` of prm (in T): phi of material-space points
/// X [n, 3], and resolve_general on world-space nodes x [n, 3] with
/// velocities v [n, 3] (updated in place).
void zso_level_set_sdf_f64(const zso_params* prm, int32_t idx, int64_t n, const double* X, double* phi) {
  ColStore<double> st;
  auto sp = to_params<double>(prm, st);
  for (int64_t i = 0; i < n; ++i) phi[i] = level_set_sdf(sp.gcols[idx], X + 3 * i);
}
void zso_level_set_normal_f64(const zso_params* prm, int32_t idx, int64_t n, const double* X, double* nrm) {
  ColStore<double> st;
  auto sp = to_params<double>(prm, st);
  for (int64_t i = 0; i < n; ++i) level_set_normal(sp.gcols[idx], X + 3 * i, nrm + 3 * i);
}
void zso_resolve_general_f64(const zso_params* prm, int32_t idx, int64_t n, const double* x, double* v) {
  ColStore<double> st;
  auto sp = to_params<double>(prm, st);
  for (int64_t i = 0; i < n; ++i) resolve_general(sp.gcols[idx], x + 3 * i, v + 3 * i);
}
void zso_resolve_general_f32(const zso_params* prm, int32_t idx, int64_t n, const float* x, float* v) {
  ColStore<float> st;
  auto sp = to_params<float>(prm, st);
  for (int64_t i = 0; i < n; ++i) resolve_general(sp.gcols[idx], x + 3 * i, v + 3 * i);
}

void zso_ctx_free_f32(void* c) { delete static_cast<Ctx<float>*>(c); }
void zso_ctx_free_f64(void* c) { delete static_cast<Ctx<double>*>(c); }

/// sizes: [T, cnt, n_indices, n_grid_floats, quirk_hit]
void zso_ctx_sizes_f32(void* c, int64_t* s) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  s[0] = ctx->out.table.T;
  s[1] = ctx->out.table.cnt;
  s[2] = (int64_t)ctx->out.indices.size();
  s[3] = (int64_t)ctx->out.grid.size();
  s[4] = ctx->out.table.query_wrap_quirk_hit ? 1 : 0;
}
void zso_ctx_sizes_f64(void* c, int64_t* s) {
  auto* ctx = static_cast<Ctx<double>*>(c);
  s[0] = ctx->out.table.T;
  s[1] = ctx->out.table.cnt;
  s[2] = (int64_t)ctx->out.indices.size();
  s[3] = (int64_t)ctx->out.grid.size();
  s[4] = ctx->out.table.query_wrap_quirk_hit ? 1 : 0;
}

/// Copy out the table (keys 3T, indices T, status T, activeKeys 3*cnt).
void zso_ctx_table_f32(void* c, int32_t* keys, int32_t* indices, int32_t* status, int32_t* active) {
  auto& t = static_cast<Ctx<float>*>(c)->out.table;
  std::memcpy(keys, t.keys.data(), sizeof(int32_t) * 3 * (size_t)t.T);
  std::memcpy(indices, t.indices.data(), sizeof(int32_t) * (size_t)t.T);
  std::memcpy(status, t.status.data(), sizeof(int32_t) * (size_t)t.T);
  std::memcpy(active, t.active.data(), sizeof(int32_t) * 3 * (size_t)t.cnt);
}
/// Copy out the bucket sort (counts, offsets: cnt + 1; indices: n).
void zso_ctx_sort_f32(void* c, int32_t* counts, int32_t* offsets, int32_t* indices) {
  auto& o = static_cast<Ctx<float>*>(c)->out;
  std::memcpy(counts, o.counts.data(), sizeof(int32_t) * o.counts.size());
  std::memcpy(offsets, o.offsets.data(), sizeof(int32_t) * o.offsets.size());
  std::memcpy(indices, o.indices.data(), sizeof(int32_t) * o.indices.size());
}
/// Copy out the grids: after P2G (m, mv, rhs) and after update + boundary.
float zso_ctx_grids_f32(void* c, float* grid_p2g, float* grid_v) {
  auto& o = static_cast<Ctx<float>*>(c)->out;
  if (grid_p2g) std::memcpy(grid_p2g, o.grid.data(), sizeof(float) * o.grid.size());
  if (grid_v) std::memcpy(grid_v, o.gridv.data(), sizeof(float) * o.gridv.size());
  return o.maxVel;
}
double zso_ctx_grids_f64(void* c, double* grid_p2g, double* grid_v) {
  auto& o = static_cast<Ctx<double>*>(c)->out;
  if (grid_p2g) std::memcpy(grid_p2g, o.grid.data(), sizeof(double) * o.grid.size());
  if (grid_v) std::memcpy(grid_v, o.gridv.data(), sizeof(double) * o.gridv.size());
  return o.maxVel;
}

/// Replace the grids (after P2G, and after update + boundary) with caller
/// data of table.cnt * 7 * 64 floats: synthetic node values for the known-
/// answer checks of the grid update and G2P.
void zso_ctx_set_grid_f32(void* c, const float* grid_p2g, const float* grid_v) {
  auto& o = static_cast<Ctx<float>*>(c)->out;
  const size_t nf = (size_t)o.table.cnt * kChannels * 64;
  if (grid_p2g) o.grid.assign(grid_p2g, grid_p2g + nf);
  if (grid_v) o.gridv.assign(grid_v, grid_v + nf);
}

/// GridAngularMomentum of the context's grid after P2G (which = 0) or after
/// the update (which = 1), accumulated into sum[6].
void zso_ctx_angular_momentum_f32(void* c, int32_t which, float dx, int32_t mchn, int32_t mvchn, double* sum) {
  auto& o = static_cast<Ctx<float>*>(c)->out;
  const auto& g = which ? o.gridv : o.grid;
  grid_angular_momentum(g.data(), o.table, (i64)(g.size() / (kChannels * 64)), dx, mchn, mvchn, sum);
}

// ---------------------------------------------------------------- stages
/// CleanSparsity + ComputeSparsity into the context's table of size T.
void zso_activate_f32(void* c, int32_t T, int64_t n, const float* x, float dx, int32_t blockLen,
                      int32_t offset, float disp) {
  auto& t = static_cast<Ctx<float>*>(c)->out.table;
  if (t.T != T) t.resize(T);
  t.clear();
  compute_sparsity(t, x, n, dx, blockLen, offset, disp);
}
void zso_enlarge_f32(void* c, const int32_t* lo, const int32_t* hi) {
  enlarge_sparsity(static_cast<Ctx<float>*>(c)->out.table, lo, hi);
}
void zso_bucket_sort_f32(void* c, int64_t n, const float* x, float dx, int32_t blockLen,
                         int32_t offset, float disp) {
  auto& o = static_cast<Ctx<float>*>(c)->out;
  o.counts.assign(o.table.cnt + 1, 0);
  o.offsets.assign(o.table.cnt + 1, 0);
  o.indices.assign(n, 0);
  bucket_sort(o.table, x, n, dx, blockLen, offset, disp, o.counts.data(), o.offsets.data(),
              o.indices.data());
}
/// Clean the grid of table.cnt blocks and run P2G; the grid stays in the context.
void zso_p2g_f32(void* c, const zso_params* prm, int64_t n, float* x, float* v, float* C, float* F,
                 float* m, float* logJp) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  auto& o = ctx->out;
  o.grid.assign((size_t)o.table.cnt * kChannels * 64, 0.f);
  p2g(sp.mat, sp.dt, sp.dx, view<float>(n, x, v, C, F, m, logJp), o.table, o.grid.data());
}
/// FP64 P2G of double particle arrays on the f32 context's table (so the
/// grid layout matches); grid_out holds table.cnt * 7 * 64 doubles.
void zso_p2g_f64_on(void* c32, const zso_params* prm, int64_t n, double* x, double* v, double* C, double* F,
                    double* m, double* logJp, double* grid_out) {
  auto* ctx = static_cast<Ctx<float>*>(c32);
  ColStore<double> cols;
  auto sp = to_params<double>(prm, cols);
  const int64_t nb = ctx->out.table.cnt;
  for (int64_t i = 0; i < nb * kChannels * 64; ++i) grid_out[i] = 0.0;
  p2g(sp.mat, sp.dt, sp.dx, view<double>(n, x, v, C, F, m, logJp), ctx->out.table, grid_out);
}

/// Fold, grid velocity and boundary on a copy of the P2G grid; returns maxVel.
float zso_grid_update_f32(void* c, const zso_params* prm) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  auto& o = ctx->out;
  o.gridv = o.grid;
  const int64_t nb = o.table.cnt;
  fold_rhs(o.gridv.data(), nb);
  float mv = 0;
  grid_velocity(o.gridv.data(), nb, sp.dt, sp.extf, mv);
  grid_boundary(o.gridv.data(), o.table, nb, sp.dx, sp.cols, sp.ncols);
  grid_boundary_general(o.gridv.data(), o.table, nb, sp.dx, sp.gcols, sp.ngcols);
  o.maxVel = mv;
  return mv;
}
void zso_g2p_f32(void* c, const zso_params* prm, int64_t n, float* x, float* v, float* C, float* F) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  g2p(sp.dt, sp.dx, view<float>(n, x, v, C, F, nullptr, nullptr), ctx->out.table, ctx->out.gridv.data());
}

/// P2C2G (zsmpm_oracle.hpp p2c2g) into a cleaned grid on the context's table,
/// kept as the context's P2G grid; the cell buckets are built from x.
void zso_p2c2g_f32(void* c, const zso_params* prm, int64_t n, float* x, float* v, float* C, float* F, float* m,
                   float* logJp) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  auto& o = ctx->out;
  o.grid.assign((size_t)o.table.cnt * kChannels * 64, 0.f);
  CellBuckets ib;
  ib.build(x, n, sp.dx);
  p2c2g(sp.mat, sp.dt, sp.dx, view<float>(n, x, v, C, F, m, logJp), o.table, ib, o.grid.data());
}
/// FP64 P2C2G of double particles on the f32 context's table; grid_out holds
/// table.cnt * 7 * 64 doubles.
void zso_p2c2g_f64_on(void* c32, const zso_params* prm, int64_t n, double* x, double* v, double* C, double* F,
                      double* m, double* logJp, double* grid_out) {
  auto* ctx = static_cast<Ctx<float>*>(c32);
  ColStore<double> cols;
  auto sp = to_params<double>(prm, cols);
  const int64_t nb = ctx->out.table.cnt;
  for (int64_t i = 0; i < nb * kChannels * 64; ++i) grid_out[i] = 0.0;
  CellBuckets ib;
  ib.build(x, n, sp.dx);
  p2c2g(sp.mat, sp.dt, sp.dx, view<double>(n, x, v, C, F, m, logJp), ctx->out.table, ib, grid_out);
}
/// G2C2P on the context's updated grid (zso_grid_update_f32 or set_grids).
void zso_g2c2p_f32(void* c, const zso_params* prm, int64_t n, float* x, float* v, float* C, float* F) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  CellBuckets ib;
  ib.build(x, n, sp.dx);
  g2c2p(sp.dt, sp.dx, view<float>(n, x, v, C, F, nullptr, nullptr), ctx->out.table, ib, ctx->out.gridv.data());
}
/// FP64 G2C2P of double particles on the f32 context's table with a double
/// node-velocity grid (table.cnt * 7 * 64).
void zso_g2c2p_f64_on(void* c32, const zso_params* prm, int64_t n, double* x, double* v, double* C, double* F,
                      const double* gridv) {
  auto* ctx = static_cast<Ctx<float>*>(c32);
  ColStore<double> cols;
  auto sp = to_params<double>(prm, cols);
  CellBuckets ib;
  ib.build(x, n, sp.dx);
  g2c2p(sp.dt, sp.dx, view<double>(n, x, v, C, F, nullptr, nullptr), ctx->out.table, ib, gridv);
}

/// Run `steps` substeps in place; the context keeps the last substep's
/// table, sort and grids.
void zso_run_f32(void* c, const zso_params* prm, int32_t steps, int64_t n, float* x, float* v,
                 float* C, float* F, float* m, float* logJp) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  for (int s = 0; s < steps; ++s)
    substep(sp, view<float>(n, x, v, C, F, m, logJp), ctx->out, s == steps - 1);
}
void zso_run_f64(void* c, const zso_params* prm, int32_t steps, int64_t n, double* x, double* v,
                 double* C, double* F, double* m, double* logJp) {
  auto* ctx = static_cast<Ctx<double>*>(c);
  auto sp = to_params<double>(prm, ctx->cols);
  for (int s = 0; s < steps; ++s)
    substep(sp, view<double>(n, x, v, C, F, m, logJp), ctx->out, s == steps - 1);
}

/// As zso_run_*, reordering the particles by each substep's bucket sort
/// after G2P; ids (may be NULL) is permuted along.
void zso_run_sorted_f32(void* c, const zso_params* prm, int32_t steps, int64_t n, float* x, float* v,
                        float* C, float* F, float* m, float* logJp, int32_t* ids) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  for (int s = 0; s < steps; ++s)
    substep(sp, view<float>(n, x, v, C, F, m, logJp), ctx->out, s == steps - 1, true, ids);
}
void zso_run_sorted_f64(void* c, const zso_params* prm, int32_t steps, int64_t n, double* x, double* v,
                        double* C, double* F, double* m, double* logJp, int32_t* ids) {
  auto* ctx = static_cast<Ctx<double>*>(c);
  auto sp = to_params<double>(prm, ctx->cols);
  for (int s = 0; s < steps; ++s)
    substep(sp, view<double>(n, x, v, C, F, m, logJp), ctx->out, s == steps - 1, true, ids);
}

/// The Vector::reorder gather on every attribute with map[n].
void zso_reorder_f32(int64_t n, const int32_t* map, float* x, float* v, float* C, float* F, float* m,
                     float* logJp, int32_t* ids) {
  reorder_particles(view<float>(n, x, v, C, F, m, logJp), map, ids);
}

// ---- implicit step (zsmpm_oracle.hpp) on the context's table and P2G grid ----

void zso_g2p2g_f32(void* c, const zso_params* prm, int64_t n, float* x, float* F, float* logJp, const float* v,
                   float* r) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  g2p2g(sp.mat, sp.dt, sp.dx, view<float>(n, x, nullptr, nullptr, F, nullptr, logJp), ctx->out.table, v, r);
}
void zso_implicit_multiply_f32(void* c, const zso_params* prm, int64_t n, float* x, float* F, float* logJp,
                               const float* in, float* out) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  implicit_multiply(sp.mat, sp.dt, sp.dx, view<float>(n, x, nullptr, nullptr, F, nullptr, logJp), ctx->out.table,
                    ctx->out.grid.data(), (i64)ctx->out.table.cnt, in, out);
}
void zso_implicit_project_f32(void* c, const zso_params* prm, float* dof) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  implicit_project(ctx->out.grid.data(), ctx->out.table, (i64)ctx->out.table.cnt, sp.dx, sp.cols, sp.ncols,
                   sp.gcols, sp.ngcols, dof);
}
void zso_implicit_precondition_f32(void* c, const float* in, float* out) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  implicit_precondition(ctx->out.grid.data(), (i64)ctx->out.table.cnt, in, out);
}
int32_t zso_implicit_solve_f32(void* c, const zso_params* prm, int64_t n, float* x, float* F, float* logJp,
                               float* xdof, const float* b, int32_t max_iters, float tol, float rel_tol,
                               float* residual) {
  auto* ctx = static_cast<Ctx<float>*>(c);
  auto sp = to_params<float>(prm, ctx->cols);
  return implicit_solve(sp, view<float>(n, x, nullptr, nullptr, F, nullptr, logJp), ctx->out.table,
                        ctx->out.grid.data(), (i64)ctx->out.table.cnt, xdof, b, max_iters, tol, rel_tol, residual);
}

// ---- bht, SparseGrid, level-set boundary (bht_oracle.hpp) ----

/// n raw draws of MT19937 seeded with `seed`.
void zso_mt19937(uint32_t seed, int64_t n, uint32_t* out) {
  Mt19937 r(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = r();
}
void zso_bht_hash_params(uint32_t seed, uint32_t* out6) { bht_hash_params(seed, out6); }
void zso_bht_hash(uint32_t hx, uint32_t hy, int64_t n, const int32_t* keys, uint32_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = bht_hash(hx, hy, keys + 3 * i);
}
int64_t zso_bht_table_size(int64_t n) { return bht_table_size(n); }

namespace {
// a Bht over caller arrays (copied in; copied back by bht_store)
Bht bht_load(int32_t T, const uint32_t* h6, const int32_t* keys, const int32_t* indices, const int32_t* status,
             int32_t cnt, const int32_t* active) {
  Bht b;
  b.init(T, h6);
  if (keys) std::memcpy(b.keys.data(), keys, sizeof(int32_t) * 4 * (size_t)T);
  if (indices) std::memcpy(b.indices.data(), indices, sizeof(int32_t) * (size_t)T);
  if (status) std::memcpy(b.status.data(), status, sizeof(int32_t) * (size_t)T);
  if (active) std::memcpy(b.active.data(), active, sizeof(int32_t) * 3 * (size_t)T);
  b.cnt = cnt;
  return b;
}
}  // namespace

/// Sequential inserts of n keys into the table held by the arrays (in place):
/// out[i] = BHTView::insert's return for key i.
void zso_bht_insert(int32_t T, const uint32_t* h6, int32_t* keys, int32_t* indices, int32_t* status,
                    int32_t* cnt, int32_t* active, int32_t* success, int64_t n, const int32_t* in,
                    int32_t* out) {
  Bht b = bht_load(T, h6, keys, indices, status, *cnt, active);
  b.success = *success;
  for (int64_t i = 0; i < n; ++i) out[i] = b.insert(in + 3 * i);
  std::memcpy(keys, b.keys.data(), sizeof(int32_t) * 4 * (size_t)T);
  std::memcpy(indices, b.indices.data(), sizeof(int32_t) * (size_t)T);
  std::memcpy(status, b.status.data(), sizeof(int32_t) * (size_t)T);
  std::memcpy(active, b.active.data(), sizeof(int32_t) * 3 * (size_t)T);
  *cnt = b.cnt;
  *success = b.success;
}

void zso_bht_query(int32_t T, const uint32_t* h6, const int32_t* keys, const int32_t* indices, int64_t n,
                   const int32_t* q, int32_t* out) {
  Bht b = bht_load(T, h6, keys, indices, nullptr, 0, nullptr);
  for (int64_t i = 0; i < n; ++i) out[i] = b.query(q + 3 * i);
}

namespace {
SparseGrid spg_load(int32_t T, const uint32_t* h6, const int32_t* keys, const int32_t* indices, int32_t nblocks,
                    int32_t nch, const float* data, const float* i2w, const float* w2i, float background) {
  SparseGrid g;
  g.table = bht_load(T, h6, keys, indices, nullptr, 0, nullptr);
  g.nch = nch;
  g.data.assign(data, data + (size_t)nblocks * nch * 512);
  std::memcpy(g.i2w, i2w, sizeof g.i2w);
  std::memcpy(g.w2i, w2i, sizeof g.w2i);
  g.background = background;
  return g;
}
}  // namespace

/// SparseGridView::insert's block key of each world position.
void zso_spg_block_keys(const float* w2i, int64_t n, const float* x, int32_t* out) {
  SparseGrid g;
  std::memcpy(g.w2i, w2i, sizeof g.w2i);
  for (int64_t i = 0; i < n; ++i) g.block_key(x + 3 * i, out + 3 * i);
}

/// wSample of channel chn at n world positions.
void zso_spg_wsample(int32_t T, const uint32_t* h6, const int32_t* keys, const int32_t* indices, int32_t nblocks,
                     int32_t nch, const float* data, const float* i2w, const float* w2i, float background,
                     int32_t chn, int64_t n, const float* x, float* out) {
  SparseGrid g = spg_load(T, h6, keys, indices, nblocks, nch, data, i2w, w2i, background);
  for (int64_t i = 0; i < n; ++i) out[i] = g.wsample(chn, x + 3 * i);
}

/// do_getNormal at n world positions.
void zso_spg_normal(int32_t T, const uint32_t* h6, const int32_t* keys, const int32_t* indices, int32_t nblocks,
                    int32_t nch, const float* data, const float* i2w, const float* w2i, float background,
                    int32_t sdf, int64_t n, const float* x, float* out) {
  SparseGrid g = spg_load(T, h6, keys, indices, nblocks, nch, data, i2w, w2i, background);
  for (int64_t i = 0; i < n; ++i) g.normal(sdf, x + 3 * i, out + 3 * i);
}

/// The level-set boundary on n nodes (positions x, velocities v in place);
/// inside[i] = 1 where the node was inside.
void zso_levelset_resolve(int32_t T, const uint32_t* h6, const int32_t* keys, const int32_t* indices,
                          int32_t nblocks, int32_t nch, const float* data, const float* i2w, const float* w2i,
                          float background, int32_t sdf, int32_t vchn, int32_t type, int64_t n, const float* x,
                          float* v, int32_t* inside) {
  SparseGrid g = spg_load(T, h6, keys, indices, nblocks, nch, data, i2w, w2i, background);
  for (int64_t i = 0; i < n; ++i) inside[i] = levelset_resolve(g, sdf, vchn, type, x + 3 * i, v + 3 * i) ? 1 : 0;
}

}  // extern "C"
