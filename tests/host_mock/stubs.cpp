// The kernels' translation units replaced by no-ops that check the stream they
// are given: the ABI layer's host logic (abi.hip) is what is under test.
#include <hip/hip_runtime.h>
#include "zsmpm_internal.hpp"
extern "C" hipError_t hipLaunchKernel(const void*, dim3, dim3, void**, size_t, hipStream_t s);
namespace zsmpm {
static hipError_t touch(hipStream_t s) { return hipLaunchKernel(nullptr, dim3(1), dim3(1), nullptr, 0, s); }
// a batch's scene table reaches the kernels whole: device arrays set, the
// particle ranges contiguous from 0 (the mock's device memory is host memory)
static hipError_t scenes_ok(const SceneSet* ss, hipStream_t s) {
  if (ss && ss->n > 0) {
    if (!ss->p || !ss->first || ss->n > kMaxScenes || ss->first[0] != 0) return hipErrorInvalidValue;
    for (int i = 0; i < ss->n; ++i)
      if (ss->first[i + 1] < ss->first[i] || !(ss->p[i].dx > 0.f)) return hipErrorInvalidValue;
  }
  return touch(s);
}
hipError_t table_clear(const TableView&, Workspace&, hipStream_t s, uint32_t*, int64_t) { return touch(s); }
hipError_t table_activate(const TableView&, Workspace&, const float*, int64_t, float, int, int, float, hipStream_t s, bool, const SceneSet* ss) { return scenes_ok(ss, s); }
hipError_t table_enlarge(const TableView&, Workspace&, const int32_t*, const int32_t*, int32_t, hipStream_t s) { return touch(s); }
hipError_t table_query(const TableView&, const int32_t*, int64_t, int32_t*, hipStream_t s) { return touch(s); }
hipError_t bucket_sort(const TableView&, Workspace&, const float*, int64_t, float, int, int, float, const SortView&, hipStream_t s) { return touch(s); }
hipError_t bucket_sort_slots(Workspace&, int64_t, const SortView&, hipStream_t s, bool) { return touch(s); }
hipError_t block_neighbours(const TableView&, Workspace&, hipStream_t s) { return touch(s); }
hipError_t check_launch_limits(int) { return hipSuccess; }
hipError_t grid_clear(float*, const int32_t*, int32_t, hipStream_t s, uint32_t*) { return touch(s); }
hipError_t stress(const Model&, const ParticlesView&, Workspace&, hipStream_t s, const ClosureView* cl, float*, bool, const SceneSet* ss) { return ss && ss->n > 0 && cl ? hipErrorNotSupported : scenes_ok(ss, s); }
hipError_t stress_closure(const Model&, const ClosureView&, const ParticlesView&, float*, Workspace&, hipStream_t s) { return touch(s); }
hipError_t mlp_infer(const float*, const float*, int64_t, float*, hipStream_t s) { return touch(s); }
hipError_t mlp_grad(const float*, const float*, const float*, int64_t, float*, float*, int, hipStream_t s) { return touch(s); }
hipError_t mlp_reduce(float*, float*, float*, const float*, const float*, int, float, float, float, float*, float*, hipStream_t s) { return touch(s); }
hipError_t mlp_harvest(const Model&, const float*, const float*, int64_t, float*, float*, hipStream_t s) { return touch(s); }
hipError_t closure_launch_limits() { return hipSuccess; }
hipError_t p2g_scatter(const ParticlesView&, const TableView&, const SortView&, float*, int32_t, float, float, bool, Workspace&, hipStream_t s, bool, const SceneSet* ss) { return scenes_ok(ss, s); }
hipError_t grid_update(float*, int32_t, const TableView&, float, const float*, const Colliders&, float, int, Workspace&, hipStream_t s, const SceneSet* ss, const GeneralColliders* gen) { return gen && gen->n > kMaxColliders ? hipErrorInvalidValue : scenes_ok(ss, s); }
hipError_t g2p(const ParticlesView&, const ParticlesView*, const TableView&, const SortView&, const float*, int32_t, float, float, Workspace&, hipStream_t s, const NextStress*, const SceneSet* ss) { return scenes_ok(ss, s); }
hipError_t grid_angular_momentum(const float*, int32_t, const TableView&, float, int, int, double*, Workspace&, hipStream_t s) { return touch(s); }
hipError_t p2c2g_scatter(const ParticlesView&, const TableView&, const SortView&, float*, int32_t, float, float, Workspace&, hipStream_t s) { return touch(s); }
hipError_t p2g_force_scatter(const ParticlesView&, const TableView&, const SortView&, float*, int32_t, float, float, Workspace&, hipStream_t s) { return touch(s); }
hipError_t g2c2p(const ParticlesView&, const TableView&, const SortView&, const float*, int32_t, float, float, Workspace&, hipStream_t s) { return touch(s); }
hipError_t reorder_particles(const ParticlesView&, const ParticlesView&, const int32_t*, Workspace&, hipStream_t s) { return touch(s); }
hipError_t g2p2g_gather(const ParticlesView&, const TableView&, const SortView&, const float*, int32_t, float, float, float*, Workspace&, hipStream_t s) { return touch(s); }
hipError_t implicit_add_force(const float*, int32_t, const TableView&, float*, hipStream_t s) { return touch(s); }
hipError_t implicit_force_dt2_mass(const float*, int32_t, const TableView&, float, const float* in, float* out, hipStream_t s) { return in == out ? hipErrorInvalidValue : touch(s); }
hipError_t implicit_project(const float*, int32_t, const TableView&, float, const Colliders&, const GeneralColliders&, float*, hipStream_t s) { return touch(s); }
hipError_t implicit_precondition(const float*, int32_t, const TableView&, const float*, float*, hipStream_t s) { return touch(s); }
hipError_t implicit_lincomb(const TableView&, int32_t, float, const float*, float, const float*, float*, hipStream_t s) { return touch(s); }
hipError_t implicit_dot(const TableView&, int32_t, const float*, const float*, double*, hipStream_t s) { return touch(s); }
hipError_t implicit_fill(const TableView&, int32_t, float*, float, hipStream_t s) { return touch(s); }
size_t bht_scratch_bytes(int64_t n, int32_t nb) { return 1024 + 64 * (size_t)(n > 0 ? n : 1) + 16 * (size_t)nb; }
hipError_t bht_clear(const BhtView&, hipStream_t s) { return touch(s); }
int32_t bht_insert(const BhtView& t, const int32_t*, int64_t, int32_t*, bool, void* scratch, uint32_t*, hipStream_t s, int32_t* rounds) {
  if (!scratch || t.nb <= 0 || t.T != 16 * t.nb) return -5;
  if (rounds) *rounds = 1;
  return touch(s) == hipSuccess ? 0 : -6;
}
hipError_t bht_query_many(const BhtView&, const int32_t*, int64_t, int32_t*, hipStream_t s) { return touch(s); }
int32_t spg_activate(const SpgView& g, const float*, int64_t, void* scratch, uint32_t*, hipStream_t s) {
  return !scratch || g.nch < 1 ? -5 : (touch(s) == hipSuccess ? 0 : -6);
}
hipError_t spg_sample(const SpgView&, int, int, const float*, int64_t, float*, hipStream_t s) { return touch(s); }
}
namespace zsmpm {
size_t radix_sort_scratch_words(uint32_t) { return 1024; }
size_t scan_scratch_words() { return 1024; }
}
