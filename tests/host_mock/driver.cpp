// Drives the C ABI's host logic against the mock runtime: policy life cycle,
// graph capture of the sorted driver, closure selection, batched scenes (scene
// table validation and upload, capture, autotuning) and teardown. Exit status 0
// and every call returning ZSMPM_OK (or the documented code) is the pass.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "zsmpm.h"
template <typename D> D mk() { D d{}; d.hdr = {sizeof(D), ZSMPM_ABI_MAJOR, ZSMPM_ABI_MINOR}; return d; }
struct Sc {
  zsmpm_model model; zsmpm_particles a, b; zsmpm_table t; zsmpm_buckets bk; zsmpm_grid g; zsmpm_plane_collider cols[6];
};
int main(int argc, char** argv) {
  zsmpm_policy pol;
  printf("policy %d\n", zsmpm_policy__rocm(0, nullptr, &pol));
  const int ns = 6;
  std::vector<Sc> sc(ns);
  for (int i = 0; i < ns; ++i) {
    Sc& s = sc[i];
    s.model = mk<zsmpm_model>(); s.model.kind = i & 1; s.model.volume = 1e-9f; s.model.E = 1e4f; s.model.nu = 0.3f; s.model.beta = 1; s.model.yield_surface = 0.3f; s.model.volume_correction = 1;
    int64_t n = 2000 + 3000 * i;
    for (zsmpm_particles* p : {&s.a, &s.b}) {
      *p = mk<zsmpm_particles>(); p->n = n;
      p->x = (float*)calloc(3 * n, 4); p->v = (float*)calloc(3 * n, 4); p->C = (float*)calloc(9 * n, 4); p->F = (float*)calloc(9 * n, 4);
      p->m = (float*)calloc(n, 4); p->logJp = (float*)calloc(n, 4); p->id = (int32_t*)calloc(n, 4);
    }
    s.t = mk<zsmpm_table>(); s.t.table_size = 16384 << (i & 1);
    s.t.keys = (int32_t*)calloc(3 * s.t.table_size, 4); s.t.indices = (int32_t*)calloc(s.t.table_size, 4); s.t.status = (int32_t*)calloc(s.t.table_size, 4);
    s.t.cnt = (int32_t*)calloc(1, 4); s.t.active_keys = (int32_t*)calloc(3 * s.t.table_size, 4);
    s.bk = mk<zsmpm_buckets>(); s.bk.capacity_blocks = s.t.table_size / 16;
    s.bk.counts = (int32_t*)calloc(s.bk.capacity_blocks + 1, 4); s.bk.offsets = (int32_t*)calloc(s.bk.capacity_blocks + 1, 4); s.bk.indices = (int32_t*)calloc(n, 4);
    s.g = mk<zsmpm_grid>(); s.g.dx = 1.f / 128; s.g.capacity_blocks = s.t.table_size / 16; s.g.data = (float*)calloc((size_t)s.g.capacity_blocks * 7 * 64, 4);
    for (int k = 0; k < 6; ++k) { s.cols[k] = {}; s.cols[k].normal[k / 2] = 1; }
  }
  int fails = 0;
  auto expect = [&](const char* what, int r, int want) {
    printf("%s -> %d\n", what, r);
    if (r != want) { printf("  expected %d\n", want); ++fails; }
  };
  int32_t inb = 0;
  // the sorted driver: eager, fused, graph (capture + cached replays), odd count
  Sc& s0 = sc[0];
  float g3[3] = {0.f, -9.8f, 0.f};
  expect("substeps eager", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 4, 0, &inb), 0);
  expect("substeps graph", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 6, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  expect("substeps graph again", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 5, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  expect("odd count ends in b", inb, 1);
  expect("autotune", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 10, ZSMPM_SUBSTEPS_AUTOTUNE | ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  int32_t wp = 0, wg = 0;
  expect("launch widths", zsmpm_policy_launch_widths__rocm(pol, &wp, &wg), 0);
  expect("autotuned widths are candidates", (wp == 512 || wp == 768 || wp == 1024 || wp == 2048) && (wg == 8192 || wg == 16384 || wg == 32768 || wg == 65536), 1);
  expect("set widths", zsmpm_policy_set_launch_widths__rocm(pol, 300, 5000), 0);
  expect("widths read back", zsmpm_policy_launch_widths__rocm(pol, &wp, &wg) == 0 && wp == 300 && wg == 5000, 1);
  expect("negative width rejected", zsmpm_policy_set_launch_widths__rocm(pol, -1, 5000), ZSMPM_INVALID_ARGUMENT);
  expect("huge width rejected", zsmpm_policy_set_launch_widths__rocm(pol, 1, (1 << 20) + 1), ZSMPM_INVALID_ARGUMENT);
  expect("substeps with set widths", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 4, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  expect("graph + fuse rejected", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 2, ZSMPM_SUBSTEPS_GRAPH | ZSMPM_SUBSTEPS_FUSE_STRESS, &inb), ZSMPM_INVALID_ARGUMENT);
  // closure selection
  zsmpm_closure cl = mk<zsmpm_closure>();
  cl.params = nullptr; cl.tol = 1e-4f;
  expect("closure on", zsmpm_policy_closure__rocm(pol, &cl), 0);
  expect("closure + fuse unsupported", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 2, ZSMPM_SUBSTEPS_FUSE_STRESS, &inb), ZSMPM_UNSUPPORTED_OPERATION);
  cl.tol = 0.f;
  expect("closure tol 0 rejected", zsmpm_policy_closure__rocm(pol, &cl), ZSMPM_INVALID_ARGUMENT);
  expect("closure off", zsmpm_policy_closure__rocm(pol, nullptr), 0);
  // batched scenes: every scene's particles in one set of arrays
  int64_t ntot = 0;
  for (int i = 0; i < ns; ++i) ntot += sc[i].a.n;
  zsmpm_particles ba[2];
  for (zsmpm_particles& p : ba) {
    p = mk<zsmpm_particles>(); p.n = ntot;
    p.x = (float*)calloc(3 * ntot, 4); p.v = (float*)calloc(3 * ntot, 4); p.C = (float*)calloc(9 * ntot, 4); p.F = (float*)calloc(9 * ntot, 4);
    p.m = (float*)calloc(ntot, 4); p.logJp = (float*)calloc(ntot, 4); p.id = (int32_t*)calloc(ntot, 4);
  }
  Sc& big = sc[ns - 1];  // its table, buckets and grid serve the batch (buckets sized for it)
  std::vector<zsmpm_batch_scene> d(ns);
  int64_t first = 0;
  for (int i = 0; i < ns; ++i) {
    zsmpm_batch_scene& z = d[i]; z = mk<zsmpm_batch_scene>();
    z.first = first; z.count = sc[i].a.n; first += z.count;
    z.model = sc[i].model; z.dx = (i & 1) ? 1.f / 256 : 1.f / 128; z.dt = 1e-4f; z.extf[1] = -9.8f;
    z.ncolliders = 6;
    for (int k = 0; k < 6; ++k) z.colliders[k] = sc[i].cols[k];
  }
  zsmpm_buckets bbk = mk<zsmpm_buckets>(); bbk.capacity_blocks = big.bk.capacity_blocks;
  bbk.counts = big.bk.counts; bbk.offsets = big.bk.offsets; bbk.indices = (int32_t*)calloc(ntot, 4);
  expect("batched eager", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 3, 0, &inb), 0);
  expect("batched odd ends in b", inb, 1);
  expect("batched graph + autotune", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 12, ZSMPM_SUBSTEPS_GRAPH | ZSMPM_SUBSTEPS_AUTOTUNE, &inb), 0);
  expect("batched graph replayed", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 4, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  d[1].dt = 2e-4f;  // a changed scene table is uploaded again and recaptured
  expect("batched new scene table", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 2, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  expect("batched fuse rejected", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 2, ZSMPM_SUBSTEPS_FUSE_STRESS, &inb), ZSMPM_INVALID_ARGUMENT);
  d[2].first += 1;
  expect("batched gap rejected", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 1, 0, &inb), ZSMPM_INVALID_ARGUMENT);
  d[2].first -= 1;
  expect("batched 257 scenes rejected", zsmpm_substeps_batched__rocm(pol, d.data(), 257, &ba[0], &ba[1], &big.t, &bbk, &big.g, 1, 0, &inb), ZSMPM_INVALID_ARGUMENT);
  d[3].ncolliders = 9;
  expect("batched 9 colliders rejected", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 1, 0, &inb), ZSMPM_INVALID_ARGUMENT);
  d[3].ncolliders = 6;
  cl.tol = 1e-4f;
  expect("closure on again", zsmpm_policy_closure__rocm(pol, &cl), 0);
  expect("batched closure unsupported", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 1, 0, &inb), ZSMPM_UNSUPPORTED_OPERATION);
  expect("closure off again", zsmpm_policy_closure__rocm(pol, nullptr), 0);
  expect("batched after closure off", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 2, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  // allocation-free activation: limits, then calls on the fixed scratch
  expect("noalloc activate", zsmpm_table_activate_noalloc__rocm(pol, &big.t, ba[0].x, ntot, 1.f / 128, 4, -2, 0.5f), 0);
  expect("noalloc activate smaller", zsmpm_table_activate_noalloc__rocm(pol, &sc[0].t, sc[0].a.x, sc[0].a.n, 1.f / 128, 4, -2, 0.5f), 0);
  expect("noalloc too many particles", zsmpm_table_activate_noalloc__rocm(pol, &big.t, ba[0].x, (1ll << 24) + 1, 1.f / 128, 4, -2, 0.5f), ZSMPM_INSUFFICIENT_SIZE);
  {
    zsmpm_table huge = big.t;
    huge.table_size = 1 << 22;  // descriptor only: rejected before any access
    expect("noalloc table too large", zsmpm_table_activate_noalloc__rocm(pol, &huge, ba[0].x, ntot, 1.f / 128, 4, -2, 0.5f), ZSMPM_INSUFFICIENT_SIZE);
  }
  expect("sync", zsmpm_policy_sync(pol), 0);
  expect("del", zsmpm_del_policy__rocm(pol), 0);
  for (zsmpm_particles& p : ba) { free(p.x); free(p.v); free(p.C); free(p.F); free(p.m); free(p.logJp); free(p.id); }
  free(bbk.indices);
  for (Sc& s : sc) {
    for (zsmpm_particles* p : {&s.a, &s.b}) { free(p->x); free(p->v); free(p->C); free(p->F); free(p->m); free(p->logJp); free(p->id); }
    free(s.t.keys); free(s.t.indices); free(s.t.status); free(s.t.cnt); free(s.t.active_keys);
    free(s.bk.counts); free(s.bk.offsets); free(s.bk.indices); free(s.g.data);
  }
  printf("%s\n", fails ? "FAILED" : "PASSED");
  return fails ? 1 : 0;
}
