// TEST INFRASTRUCTURE ONLY. Thin extern "C" wrappers over littlemine/zpc's own
// math headers, compiled from /root/reference/include by oracle/Makefile into
// oracle/_ref/libzpc_refmath.so. It is the checker for the oracle's
// restatement of the pieces of the hot path that compile on their own:
// svd_3d, the two stress functions, lame_parameters, the quadratic
// LocalArena, hash_combine, next_2pow, the 3x3 products and bht's universal
// hash (py_interop/HashUtils.hpp). The shim adds no arithmetic of its own;
// each wrapper only moves arrays in and out. zref_mt19937 draws from the C++
// standard library's std::mt19937, the engine bht seeds its hashes with.
#include <cstddef>
#include <cstdint>
#include <random>

#include "zensim/math/Hash.hpp"
#include "zensim/math/bit/Bits.h"
#include "zensim/math/matrix/MatrixUtils.h"
#include "zensim/physics/ConstitutiveModel.hpp"
#include "zensim/physics/ConstitutiveModel_Vol_dP.hpp"
#include "zensim/physics/SoundSpeedCfl.hpp"
#include "zensim/simulation/Utils.hpp"
#include "zensim/math/Vec.h"
#include "zensim/py_interop/HashUtils.hpp"

extern "C" {

/// math/matrix/SVD.hpp:15-26 argument order (F column-major in, U, V column-major out)
/// exactly as physics/ConstitutiveModel_Vol_dP.hpp:14-16 calls it.
void zref_svd(int64_t n, const float* F, float* U, float* S, float* V) {
  for (int64_t p = 0; p < n; ++p) {
    const float* f = F + 9 * p;
    float* u = U + 9 * p;
    float* s = S + 3 * p;
    float* v = V + 9 * p;
    zs::math::svd_3d(f[0], f[3], f[6], f[1], f[4], f[7], f[2], f[5], f[8], u[0], u[3], u[6], u[1],
                     u[4], u[7], u[2], u[5], u[8], s[0], s[1], s[2], v[0], v[3], v[6], v[1], v[4],
                     v[7], v[2], v[5], v[8]);
  }
}

void zref_lame(float E, float nu, float* mu, float* lam) {
  auto [m, l] = zs::lame_parameters(E, nu);
  *mu = m;
  *lam = l;
}

void zref_stress_fc(int64_t n, float vol, float mu, float lam, const float* F, float* PF) {
  for (int64_t p = 0; p < n; ++p) {
    zs::vec<float, 9> f, pf;
    for (int i = 0; i < 9; ++i) f[i] = F[9 * p + i];
    zs::compute_stress_fixedcorotated(vol, mu, lam, f, pf);
    for (int i = 0; i < 9; ++i) PF[9 * p + i] = pf[i];
  }
}

void zref_stress_sand(int64_t n, float vol, float mu, float lam, float coh, float beta, float ys,
                      int32_t volcorr, float* logJp, const float* F, float* PF, float* Fproj) {
  for (int64_t p = 0; p < n; ++p) {
    zs::vec<float, 9> f, pf;
    for (int i = 0; i < 9; ++i) f[i] = F[9 * p + i];
    float lj = logJp[p];
    zs::compute_stress_sand(vol, mu, lam, coh, beta, ys, volcorr != 0, lj, f, pf);
    logJp[p] = lj;
    for (int i = 0; i < 9; ++i) PF[9 * p + i] = pf[i];
    if (Fproj)
      for (int i = 0; i < 9; ++i) Fproj[9 * p + i] = f[i];
  }
}

/// simulation/Utils.hpp make_local_arena (collocated, quadratic): corner, the 27
/// weights and diffs in ndrange order.
void zref_arena(int64_t n, float dx, const float* x, int32_t* corner, float* w, float* diff) {
  for (int64_t p = 0; p < n; ++p) {
    zs::vec<float, 3> pos{x[3 * p], x[3 * p + 1], x[3 * p + 2]};
    auto arena = zs::make_local_arena(dx, pos);
    int e = 0;
    for (auto loc : arena.range()) {
      auto c = arena.coord(loc);
      auto d = arena.diff(loc);
      if (e == 0)
        for (int k = 0; k < 3; ++k) corner[3 * p + k] = c[k];
      w[27 * p + e] = arena.weight(loc);
      for (int k = 0; k < 3; ++k) diff[81 * p + 3 * e + k] = d[k];
      ++e;
    }
  }
}

/// hash_combine composed as HashTableView::do_hash does (HashTable.hpp:496-500).
int32_t zref_hash(int32_t k0, int32_t k1, int32_t k2) {
  size_t ret = k0;
  zs::hash_combine(ret, k1);
  zs::hash_combine(ret, k2);
  return static_cast<int32_t>(ret);
}

int64_t zref_next_2pow(int64_t n) { return zs::next_2pow(n); }

void zref_matmul3(const float* a, const float* b, float* c) {
  zs::matrixMatrixMultiplication3d(a, b, c);
}
void zref_mat_diag_matT(const float* m1, const float* d, const float* m2, float* out) {
  zs::matmul_mat_diag_matT_3D(out, m1, d, m2);
}

float zref_cfl_dt(float E, float nu, float rho, float dx, float cfl) {
  return zs::evaluate_timestep_linear_elasticity(E, nu, rho, dx, cfl);
}

/// universal_hash_base<vec<i32,3>>{hx, hy} of n keys (py_interop/HashUtils.hpp:22-41)
void zref_universal_hash(uint32_t hx, uint32_t hy, int64_t n, const int32_t* keys, uint32_t* out) {
  const zs::universal_hash_base<zs::vec<int, 3>> h{hx, hy};
  for (int64_t i = 0; i < n; ++i) out[i] = h(zs::vec<int, 3>{keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]});
}

/// n draws of std::mt19937 seeded with `seed`
void zref_mt19937(uint32_t seed, int64_t n, uint32_t* out) {
  std::mt19937 r(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = (uint32_t)r();
}
}
