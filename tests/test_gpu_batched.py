"""Batched scenes on the GPU (zsmpm_substeps_batched__rocm): independent scenes
of mixed material (fixed-corotated, Drucker-Prager) and grid spacing (1/128,
1/256) in one table, one bucket sort and one grid, every stage one launch over
all scenes. Each scene is held to the oracle run of that scene alone.

- One sorted substep of six scenes: per scene, the active block keys (scene
  offset removed, in block-number order) equal that scene's own oracle table;
  the sort permutation inside the scene's particle range equals the oracle's
  exactly; x, v, C, F, m and logJp are within 1e-5 (FP64-anchored where
  float32 summation order decides).
- 26 substeps with the launch-width autotuning (8 substeps) and graph
  replays of the rest, then one more substep as an odd call: x, v and F within
  1e-3 of the per-scene oracle runs, ids still a permutation of each scene's
  own range.
- Argument errors: scene ranges that are not contiguous, counts that do not
  add up to the particle count, more than 256 scenes, a selected learned
  closure.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu

TOL_1STEP = 1e-5
TOL_MANY = 1e-3


def rel_err(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-30))


def check_field(name, gpu, ref32, ref64, tol):
    e32 = rel_err(gpu, ref32)
    if e32 < tol:
        return
    assert ref64 is not None, f"{name}: {e32:.3e} vs FP32 oracle"
    e_gpu = rel_err(gpu, ref64)
    e_ref = rel_err(ref32, ref64)
    assert e_gpu <= max(tol, 2.0 * e_ref), (
        f"{name}: gpu-vs-fp32 {e32:.3e}, gpu-vs-fp64 {e_gpu:.3e}, fp32-vs-fp64 {e_ref:.3e}")


def _rocm():
    from zpc_amd import rocm

    return rocm


def _oracle(sc, dtype=np.float32):
    return O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp).astype(dtype)


@pytest.fixture(scope="module")
def six():
    scenes = S.batch_scenes(count=6, lo=2000, hi=20000)
    # every scene moves and deforms from the first substep on
    return [S.perturb(sc, seed=40 + i, f_scale=0.02, c_scale=2.0, v_scale=0.3) for i, sc in enumerate(scenes)]


def test_batch_scene_mix(six):
    kinds = {sc.material.kind for sc in six}
    dxs = {sc.dx for sc in six}
    print("scenes:", [(sc.name, sc.n) for sc in six], flush=True)
    assert kinds == {S.FIXED_COROTATED, S.DRUCKER_PRAGER} and len(dxs) == 2


def test_batched_one_substep_exact_and_fields(six):
    rocm = _rocm()
    b = rocm.Batch(six)
    b.substeps(1)
    b.sync()
    act = b.table.host()["active"]
    scene_of = (act[:, 0] + 512) >> 10
    ids = b.ids()
    for i, sc in enumerate(six):
        ctx = O.Context()
        ps = _oracle(sc)
        ctx.run(O.make_params(sc), ps, 1, reorder=True)
        ref = ctx.table()
        mine = act[scene_of == i].copy()
        mine[:, 0] -= i * rocm.SCENE_KEY_STRIDE
        assert np.array_equal(mine, ref["active"]), f"{sc.name}: block keys or their order differ"
        lo, hi = int(b.first[i]), int(b.first[i + 1])
        assert np.array_equal(ids[lo:hi] - lo, ps.ids), f"{sc.name}: sort permutation differs"
        got = b.scene_state(i)
        ps64 = _oracle(sc, np.float64)
        O.Context(np.float64).run(O.make_params(sc), ps64, 1, reorder=True)
        names = ["x", "v", "C", "F", "m"] + (["logJp"] if sc.logJp is not None else [])
        for name in names:
            check_field(f"{sc.name} {name}", got[name], ps.by_id(name), ps64.by_id(name), TOL_1STEP)


def test_batched_autotune_graph_and_odd_call(six):
    rocm = _rocm()
    b = rocm.Batch(six)
    b.substeps(26, graph=True, autotune=True)
    b.substeps(1, graph=True)
    b.sync()
    steps = 27
    p2g, g2p = b.pol.launch_widths()
    assert p2g in (512, 768, 1024, 2048) and g2p in (8192, 16384, 32768, 65536)
    ids = b.ids()
    for i, sc in enumerate(six):
        lo, hi = int(b.first[i]), int(b.first[i + 1])
        assert np.array_equal(np.sort(ids[lo:hi]), np.arange(lo, hi)), f"{sc.name}: ids left the scene's range"
        got = b.scene_state(i)
        ps = _oracle(sc)
        O.Context().run(O.make_params(sc), ps, steps)
        for name in ("x", "v", "F"):
            e = rel_err(got[name], ps.by_id(name))
            assert e < TOL_MANY, f"{sc.name} {name} after {steps} substeps: {e:.3e}"
        disp = ps.x - sc.x.ravel()
        de = np.max(np.abs(got["x"].ravel() - ps.x)) / np.max(np.abs(disp))
        assert de < TOL_MANY, f"{sc.name} displacement error {de:.3e}"


def test_batched_argument_errors(six):
    rocm = _rocm()
    b = rocm.Batch(six[:3])
    L, d = b.L, b._descs
    bad = (rocm.BatchSceneDesc * 3)()
    C.memmove(bad, d, C.sizeof(bad))

    def call(descs, n, nsteps=1):
        in_b = C.c_int32(0)
        return L.zsmpm_substeps_batched__rocm(b.pol.h, descs, n, C.byref(b.parts.desc), C.byref(b.spare.desc),
                                              C.byref(b.table.desc), C.byref(b.buckets.desc), C.byref(b.grid.desc),
                                              nsteps, 0, C.byref(in_b))

    bad[1].first += 1  # a gap between scene 0 and scene 1
    assert call(bad, 3) == -5
    C.memmove(bad, d, C.sizeof(bad))
    bad[2].count -= 1  # counts no longer add up to the particle count
    assert call(bad, 3) == -5
    assert call(d, 0) == -5
    assert call(d, 257) == -5
    b.pol.set_closure(None)
    assert call(d, 3) == -4  # no learned closure for batches
    b.pol.set_closure(enable=False)
    assert call(d, 3) == 0
    b.sync()
