// Drives the C ABI's host logic against the mock runtime: policy life cycle,
// graph capture of the sorted driver, closure selection, batched scenes (lane
// assignment, capture, autotuning bookkeeping) and teardown. Exit status 0
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
  std::vector<zsmpm_scene> d(ns);
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
    zsmpm_scene& z = d[i]; z = mk<zsmpm_scene>();
    z.model = &s.model; z.a = &s.a; z.b = &s.b; z.table = &s.t; z.buckets = &s.bk; z.grid = &s.g; z.dt = 1e-4f; z.extf[1] = -9.8f; z.colliders = s.cols; z.ncolliders = 6;
  }
  int fails = 0;
  auto expect = [&](const char* what, int r, int want) {
    printf("%s -> %d\n", what, r);
    if (r != want) { printf("  expected %d\n", want); ++fails; }
  };
  int32_t used = 0, inb = 0; float tune[4];
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
  expect("graph + fuse rejected", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 2, ZSMPM_SUBSTEPS_GRAPH | ZSMPM_SUBSTEPS_FUSE_STRESS, &inb), ZSMPM_INVALID_ARGUMENT);
  // closure selection
  zsmpm_closure cl = mk<zsmpm_closure>();
  cl.params = nullptr; cl.tol = 1e-4f;
  expect("closure on", zsmpm_policy_closure__rocm(pol, &cl), 0);
  expect("closure + fuse unsupported", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 2, ZSMPM_SUBSTEPS_FUSE_STRESS, &inb), ZSMPM_UNSUPPORTED_OPERATION);
  cl.tol = 0.f;
  expect("closure tol 0 rejected", zsmpm_policy_closure__rocm(pol, &cl), ZSMPM_INVALID_ARGUMENT);
  expect("closure off", zsmpm_policy_closure__rocm(pol, nullptr), 0);
  // batched scenes
  const int want_batch = argc > 1 ? atoi(argv[1]) : 0;
  for (int lanes : {1, 3}) {
    char what[64]; snprintf(what, sizeof what, "batched lanes %d", lanes);
    expect(what, zsmpm_substeps_batched__rocm(pol, d.data(), ns, 4, lanes, &used, tune, &inb), want_batch);
    if (want_batch == 0) expect("lanes used", used, lanes);
  }
  expect("batched autotune", zsmpm_substeps_batched__rocm(pol, d.data(), 5, 25, 0, &used, tune, &inb), want_batch);
  if (want_batch == 0) {
    expect("autotune timed 1, 2, 4 lanes", tune[0] > 0 && tune[1] > 0 && tune[2] > 0 && tune[3] < 0, 1);
    expect("odd batch ends in b", inb, 1);
  }
  expect("batched 9 lanes rejected", zsmpm_substeps_batched__rocm(pol, d.data(), ns, 2, 9, &used, tune, &inb), ZSMPM_INVALID_ARGUMENT);
  expect("sync", zsmpm_policy_sync(pol), 0);
  expect("del", zsmpm_del_policy__rocm(pol), 0);
  for (Sc& s : sc) {
    for (zsmpm_particles* p : {&s.a, &s.b}) { free(p->x); free(p->v); free(p->C); free(p->F); free(p->m); free(p->logJp); free(p->id); }
    free(s.t.keys); free(s.t.indices); free(s.t.status); free(s.t.cnt); free(s.t.active_keys);
    free(s.bk.counts); free(s.bk.offsets); free(s.bk.indices); free(s.g.data);
  }
  printf("%s\n", fails ? "FAILED" : "PASSED");
  return fails ? 1 : 0;
}
