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
  // the particle -> cell -> grid transfers: argument checks, own bucket sort
  expect("p2c2g", zsmpm_p2c2g__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.bk, &s0.g, 1e-4f), 0);
  expect("p2c2g own buckets", zsmpm_p2c2g__rocm(pol, &s0.model, &s0.a, &s0.t, nullptr, &s0.g, 1e-4f), 0);
  expect("p2c2g null grid", zsmpm_p2c2g__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.bk, nullptr, 1e-4f), ZSMPM_INVALID_ARGUMENT);
  expect("g2c2p", zsmpm_g2c2p__rocm(pol, &s0.a, &s0.t, &s0.bk, &s0.g, 1e-4f), 0);
  expect("g2c2p null particles", zsmpm_g2c2p__rocm(pol, nullptr, &s0.t, &s0.bk, &s0.g, 1e-4f), ZSMPM_INVALID_ARGUMENT);
  {
    zsmpm_model dp = s0.model; dp.kind = ZSMPM_DRUCKER_PRAGER;
    zsmpm_particles nolj = s0.a; nolj.logJp = nullptr;
    expect("p2c2g sand without logJp", zsmpm_p2c2g__rocm(pol, &dp, &nolj, &s0.t, &s0.bk, &s0.g, 1e-4f), ZSMPM_INVALID_ARGUMENT);
  }
  expect("launch widths", zsmpm_policy_launch_widths__rocm(pol, &wp, &wg), 0);
  expect("autotuned widths are candidates", (wp == 1024 || wp == 2048 || wp == 4096) && (wg == 32768 || wg == 65536), 1);
  int32_t tc = 0, tt = 0;
  expect("launch tiles", zsmpm_policy_launch_tiles__rocm(pol, &tc, &tt), 0);
  expect("autotuned tiles are candidates", (tc == 576 || tc == 256) && (tt == 128 || tt == 256), 1);
  expect("set tiles", zsmpm_policy_set_launch_tiles__rocm(pol, 256, 256), 0);
  expect("tiles read back", zsmpm_policy_launch_tiles__rocm(pol, &tc, &tt) == 0 && tc == 256 && tt == 256, 1);
  expect("substeps with set tiles", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 4, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  expect("odd chunk rejected", zsmpm_policy_set_launch_tiles__rocm(pol, 300, 128), ZSMPM_INVALID_ARGUMENT);
  expect("odd threads rejected", zsmpm_policy_set_launch_tiles__rocm(pol, 576, 64), ZSMPM_INVALID_ARGUMENT);
  expect("bad transfer scheme", zsmpm_policy_transfer_scheme__rocm(pol, 2), ZSMPM_INVALID_ARGUMENT);
  expect("cell transfer", zsmpm_policy_transfer_scheme__rocm(pol, ZSMPM_TRANSFER_CELL), 0);
  expect("cell substeps, graph", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 3, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
  expect("cell + fuse unsupported", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 2, ZSMPM_SUBSTEPS_FUSE_STRESS, &inb), ZSMPM_UNSUPPORTED_OPERATION);
  expect("apic transfer", zsmpm_policy_transfer_scheme__rocm(pol, ZSMPM_TRANSFER_APIC), 0);
  expect("default tiles", zsmpm_policy_set_launch_tiles__rocm(pol, 0, 0) == 0 && zsmpm_policy_launch_tiles__rocm(pol, &tc, &tt) == 0 && tc == 576 && tt == 128, 1);
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
  expect("cell transfer for a batch", zsmpm_policy_transfer_scheme__rocm(pol, ZSMPM_TRANSFER_CELL), 0);
  expect("batched cell unsupported", zsmpm_substeps_batched__rocm(pol, d.data(), ns, &ba[0], &ba[1], &big.t, &bbk, &big.g, 1, 0, &inb), ZSMPM_UNSUPPORTED_OPERATION);
  expect("apic transfer back", zsmpm_policy_transfer_scheme__rocm(pol, ZSMPM_TRANSFER_APIC), 0);
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
  // layout adapters: argument validation (no kernel runs under the mock)
  {
    Sc& s0 = sc[0];
    const int64_t n = s0.a.n, L = 32, nch = 27;
    std::vector<float> tiles((size_t)((n + L - 1) / L * nch * L));
    zsmpm_tile_particles tp = mk<zsmpm_tile_particles>();
    tp.n = n; tp.lane_width = (int32_t)L; tp.num_channels = (int32_t)nch; tp.data = tiles.data();
    tp.x = 0; tp.v = 3; tp.C = 6; tp.F = 15; tp.m = 24; tp.logJp = 25; tp.id = 26; tp.matrix_order = ZSMPM_ROW_MAJOR;
    expect("from tiles", zsmpm_particles_from_tiles__rocm(pol, &tp, &s0.b), 0);
    expect("to tiles", zsmpm_particles_to_tiles__rocm(pol, &s0.a, &tp), 0);
    zsmpm_tile_particles bad = tp; bad.F = 14;
    expect("overlapping channels rejected", zsmpm_particles_from_tiles__rocm(pol, &bad, &s0.b), ZSMPM_INVALID_ARGUMENT);
    bad = tp; bad.m = 27;
    expect("channel past the end rejected", zsmpm_particles_from_tiles__rocm(pol, &bad, &s0.b), ZSMPM_INVALID_ARGUMENT);
    bad = tp; bad.lane_width = 24;
    expect("lane width not a power of two rejected", zsmpm_particles_from_tiles__rocm(pol, &bad, &s0.b), ZSMPM_INVALID_ARGUMENT);
    bad = tp; bad.n = n - 1;
    expect("count mismatch rejected", zsmpm_particles_from_tiles__rocm(pol, &bad, &s0.b), ZSMPM_INVALID_ARGUMENT);
    bad = tp; bad.x = -1; bad.v = -1;
    expect("absent attributes skipped", zsmpm_particles_from_tiles__rocm(pol, &bad, &s0.b), 0);
    const int64_t nt = (n + L - 1) / L;
    std::vector<float> tiles2(tiles.size());
    std::vector<int32_t> map((size_t)nt);
    for (int64_t t = 0; t < nt; ++t) map[t] = (int32_t)(nt - 1 - t);
    expect("tile reorder", zsmpm_tiles_reorder__rocm(pol, tiles.data(), tiles2.data(), map.data(), nt, 32, 27, 0), 0);
    expect("tile reorder in place rejected", zsmpm_tiles_reorder__rocm(pol, tiles.data(), tiles.data(), map.data(), nt, 32, 27, 1), ZSMPM_INVALID_ARGUMENT);
  }
  // implicit step: argument validation, the solver's vectors
  {
    Sc& s0 = sc[0];
    const size_t nv = (size_t)s0.g.capacity_blocks * 64 * 3;
    std::vector<float> xv(nv), bv(nv), ov(nv);
    expect("g2p2g", zsmpm_g2p2g__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.g, 1e-4f, xv.data(), ov.data()), 0);
    expect("implicit multiply", zsmpm_implicit_multiply__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.g, 1e-4f, xv.data(), ov.data()), 0);
    expect("implicit multiply in place rejected", zsmpm_implicit_multiply__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.g, 1e-4f, xv.data(), xv.data()), ZSMPM_INVALID_ARGUMENT);
    expect("implicit project", zsmpm_implicit_project__rocm(pol, &s0.t, &s0.g, s0.cols, 6, ov.data()), 0);
    expect("implicit precondition", zsmpm_implicit_precondition__rocm(pol, &s0.t, &s0.g, ov.data(), xv.data()), 0);
    int32_t its = -1;
    float res = -1.f;
    expect("implicit solve", zsmpm_implicit_solve__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.g, 1e-4f, s0.cols, 6, xv.data(), bv.data(), 10, 1e-6f, 0.5f, &its, &res), 0);
    expect("implicit solve reports", its >= 0 && res >= 0.f, 1);
    expect("implicit solve bad tol", zsmpm_implicit_solve__rocm(pol, &s0.model, &s0.a, &s0.t, &s0.g, 1e-4f, s0.cols, 6, xv.data(), bv.data(), 10, -1.f, 0.5f, &its, &res), ZSMPM_INVALID_ARGUMENT);
  }
  // bht and SparseGrid: argument validation and the policy's scratch
  {
    expect("bht table size", zsmpm_bht_table_size(1000) == 2064 && zsmpm_bht_table_size(0) == 0, 1);
    uint32_t hp[6];
    zsmpm_bht_hash_params(ZSMPM_BHT_SEED, hp);
    expect("bht hash params", hp[0] == 1872583848u && hp[1] == 794921487u, 1);
    const int32_t T = 2064;
    zsmpm_bht bt = mk<zsmpm_bht>();
    bt.table_size = T;
    std::vector<int32_t> bk(4 * T + 4), bi(T), bs(T), bc(1), ba(3 * T), bsu(1);
    int32_t* kal = bk.data();
    while ((uintptr_t)kal % 16) ++kal;
    bt.keys = kal; bt.indices = bi.data(); bt.status = bs.data(); bt.cnt = bc.data(); bt.active_keys = ba.data(); bt.success = bsu.data();
    for (int i = 0; i < 6; ++i) bt.hash[i] = hp[i];
    std::vector<int32_t> keys(3 * 500, 1), outi(500);
    expect("bht clear", zsmpm_bht_clear__rocm(pol, &bt), 0);
    expect("bht insert", zsmpm_bht_insert__rocm(pol, &bt, keys.data(), 500, outi.data(), 0), 0);
    expect("bht insert serial", zsmpm_bht_insert__rocm(pol, &bt, keys.data(), 500, nullptr, ZSMPM_BHT_SERIAL), 0);
    expect("bht insert bad flags", zsmpm_bht_insert__rocm(pol, &bt, keys.data(), 500, nullptr, 4), ZSMPM_INVALID_ARGUMENT);
    expect("bht insert grows the scratch", zsmpm_bht_insert__rocm(pol, &bt, keys.data(), 400, nullptr, 0), 0);
    std::vector<int32_t> big_keys(3 * 100000, 0);
    expect("bht insert larger", zsmpm_bht_insert__rocm(pol, &bt, big_keys.data(), 100000, nullptr, 0), 0);
    expect("bht query", zsmpm_bht_query__rocm(pol, &bt, keys.data(), 500, outi.data()), 0);
    zsmpm_bht bad = bt; bad.table_size = 2060;
    expect("bht size not a multiple of 16", zsmpm_bht_query__rocm(pol, &bad, keys.data(), 500, outi.data()), ZSMPM_INVALID_ARGUMENT);
    bad = bt; bad.keys = kal + 1;
    expect("bht keys misaligned", zsmpm_bht_clear__rocm(pol, &bad), ZSMPM_INVALID_ARGUMENT);
    zsmpm_sparse_grid sg = mk<zsmpm_sparse_grid>();
    sg.table = bt; sg.num_channels = 4; sg.capacity_blocks = 8;
    std::vector<float> sd(8 * 4 * 512), xs(3 * 300), so(3 * 300);
    sg.data = sd.data();
    for (int i = 0; i < 16; ++i) sg.i2w[i] = sg.w2i[i] = (i % 5 == 0) ? 1.f : 0.f;
    expect("spg activate", zsmpm_spg_activate__rocm(pol, &sg, xs.data(), 300), 0);
    expect("spg sample", zsmpm_spg_sample__rocm(pol, &sg, 1, 3, xs.data(), 300, so.data()), 0);
    expect("spg sample past the channels", zsmpm_spg_sample__rocm(pol, &sg, 2, 3, xs.data(), 300, so.data()), ZSMPM_INVALID_ARGUMENT);
    expect("levelset boundary", zsmpm_policy_levelset_boundary__rocm(pol, &sg, 0, 1, ZSMPM_SLIP), 0);
    expect("levelset boundary bad channel", zsmpm_policy_levelset_boundary__rocm(pol, &sg, 0, 2, ZSMPM_SLIP), ZSMPM_INVALID_ARGUMENT);
    Sc& s0 = sc[0];
    float g3[3] = {0.f, -9.8f, 0.f};
    expect("substeps with the levelset boundary", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 4, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
    expect("general colliders keep the boundary", zsmpm_policy_colliders__rocm(pol, nullptr, 0), 0);
    expect("levelset boundary off", zsmpm_policy_levelset_boundary__rocm(pol, nullptr, 0, -1, 0), 0);
    expect("substeps after", zsmpm_substeps_sorted__rocm(pol, &s0.model, &s0.a, &s0.b, &s0.t, &s0.bk, &s0.g, 1e-4f, g3, s0.cols, 6, 4, ZSMPM_SUBSTEPS_GRAPH, &inb), 0);
    uint32_t ctr[16];
    expect("counters", zsmpm_policy_counters(pol, ctr), 0);
    expect("bht rounds counter", ctr[10] == 1u, 1);
  }
  // checkpoint round trip through a real file (mock device memory is host memory)
  {
    Sc& s0 = sc[0];
    const int64_t n = s0.a.n;
    for (int64_t i = 0; i < 3 * n; ++i) { s0.a.x[i] = 0.25f * i; s0.a.v[i] = -1.f * i; }
    for (int64_t i = 0; i < 9 * n; ++i) { s0.a.C[i] = 0.5f * i; s0.a.F[i] = 1.f + i; }
    for (int64_t i = 0; i < n; ++i) { s0.a.m[i] = 2.f * i; s0.a.logJp[i] = -0.5f * i; s0.a.id[i] = (int32_t)(n - i); }
    for (int32_t i = 0; i < 3 * s0.t.table_size; ++i) { s0.t.keys[i] = i - 7; s0.t.active_keys[i] = 3 * i; }
    for (int32_t i = 0; i < s0.t.table_size; ++i) { s0.t.indices[i] = i ^ 5; s0.t.status[i] = -1; }
    *s0.t.cnt = 1234;
    const char* path = "zsmpm_ck_probe.bin";
    expect("checkpoint save", zsmpm_checkpoint_save__rocm(pol, path, &s0.a, &s0.t, 1.5, 42), 0);
    zsmpm_checkpoint_info info{};
    expect("checkpoint info", zsmpm_checkpoint_info__rocm(path, &info), 0);
    expect("info fields", info.n == n && info.table_size == s0.t.table_size && info.fields == 127 && info.time == 1.5 && info.step == 42, 1);
    zsmpm_table t2 = s0.t;
    std::vector<int32_t> k2(3 * s0.t.table_size), i2(s0.t.table_size), st2(s0.t.table_size), c2(1), a2(3 * s0.t.table_size);
    t2.keys = k2.data(); t2.indices = i2.data(); t2.status = st2.data(); t2.cnt = c2.data(); t2.active_keys = a2.data();
    zsmpm_checkpoint_info got{};
    expect("checkpoint load", zsmpm_checkpoint_load__rocm(pol, path, &s0.b, &t2, &got), 0);
    bool same = got.step == 42 && c2[0] == 1234;
    for (int64_t i = 0; i < 3 * n; ++i) same = same && s0.b.x[i] == s0.a.x[i] && s0.b.v[i] == s0.a.v[i];
    for (int64_t i = 0; i < 9 * n; ++i) same = same && s0.b.C[i] == s0.a.C[i] && s0.b.F[i] == s0.a.F[i];
    for (int64_t i = 0; i < n; ++i) same = same && s0.b.m[i] == s0.a.m[i] && s0.b.logJp[i] == s0.a.logJp[i] && s0.b.id[i] == s0.a.id[i];
    for (int32_t i = 0; i < 3 * s0.t.table_size; ++i) same = same && k2[i] == s0.t.keys[i] && a2[i] == s0.t.active_keys[i];
    for (int32_t i = 0; i < s0.t.table_size; ++i) same = same && i2[i] == s0.t.indices[i] && st2[i] == s0.t.status[i];
    expect("checkpoint contents", same, 1);
    zsmpm_particles fewer = s0.b; fewer.n = n - 1;
    expect("load into another count rejected", zsmpm_checkpoint_load__rocm(pol, path, &fewer, &t2, nullptr), ZSMPM_INVALID_ARGUMENT);
    zsmpm_particles nolj = s0.b; nolj.logJp = nullptr;
    expect("load without logJp rejected", zsmpm_checkpoint_load__rocm(pol, path, &nolj, &t2, nullptr), ZSMPM_INVALID_ARGUMENT);
    expect("load without the table rejected", zsmpm_checkpoint_load__rocm(pol, path, &s0.b, nullptr, nullptr), ZSMPM_INVALID_ARGUMENT);
    if (FILE* f = fopen(path, "r+b")) {  // flip one byte of the body
      fseek(f, 48 + 100, SEEK_SET);
      int c = fgetc(f);
      fseek(f, 48 + 100, SEEK_SET);
      fputc(c ^ 0x40, f);
      fclose(f);
    }
    expect("corrupted checkpoint rejected", zsmpm_checkpoint_load__rocm(pol, path, &s0.b, &t2, nullptr), ZSMPM_IO_ERROR);
    remove(path);
    expect("missing file", zsmpm_checkpoint_info__rocm(path, &info), ZSMPM_IO_ERROR);
    // bgeo: header + 16 bytes per point (+ the attribute) + 2
    const char* geo = "zsmpm_probe.bgeo";
    expect("bgeo positions", zsmpm_write_partio__rocm(pol, geo, s0.a.x, n, nullptr, 0, nullptr), 0);
    FILE* f = fopen(geo, "rb");
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    fclose(f);
    expect("bgeo size", sz == 41 + 16 * n + 2, 1);
    expect("bgeo stress", zsmpm_write_partio__rocm(pol, geo, s0.a.x, n, s0.a.m, 1, "stress"), 0);
    expect("bgeo bad width", zsmpm_write_partio__rocm(pol, geo, s0.a.x, n, s0.a.m, 2, "stress"), ZSMPM_INVALID_ARGUMENT);
    remove(geo);
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
