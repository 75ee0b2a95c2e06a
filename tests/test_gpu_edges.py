"""Edge cases of the gfx950 path through the C ABI: ragged particle counts
(one particle, partial waves and tiles), empty inputs and out-of-range keys.

Integer work must match the oracle exactly; fields within the tolerance of
the parity tests (relative error against the FP32 oracle, FP64-anchored where
float32 summation order decides).
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

from test_gpu_parity import TOL_1STEP, _assert_table_equal, _oracle_particles, _rocm, check_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sand():
    return S.perturb(S.sand_scene(dims=(16, 32, 16)), seed=31, f_scale=0.02)


def _subset(sc, idx, name):
    idx = np.asarray(idx)
    return dataclasses.replace(
        sc, name=name, x=np.ascontiguousarray(sc.x[idx]), v=np.ascontiguousarray(sc.v[idx]),
        C=np.ascontiguousarray(sc.C[idx]), F=np.ascontiguousarray(sc.F[idx]), m=np.ascontiguousarray(sc.m[idx]),
        logJp=None if sc.logJp is None else np.ascontiguousarray(sc.logJp[idx]))


@pytest.mark.parametrize("k", [1, 63, 65, 4097])
def test_ragged_sorted_substeps(sand, k):
    """k particles spread over the column (a seeded subset, in index order):
    one particle, a partial wave, one past a wave and one past a bucket-sort
    tile. Two sorted substeps: table, sort and ids exact, fields within
    tolerance."""
    rng = np.random.default_rng(1000 + k)
    sc = _subset(sand, np.sort(rng.choice(sand.n, size=k, replace=False)), f"sand_{k}")
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    sim.substeps(2)
    sim.sync()
    prm = O.make_params(sc)
    ps = _oracle_particles(sc)
    ctx = O.Context()
    ctx.run(prm, ps, 2, reorder=True)
    _assert_table_equal(sim.table.host(), ctx.table())
    counts, offsets, indices = ctx.sort()
    gc, go, gi = sim.buckets.host(ctx.sizes()["cnt"], sc.n)
    assert np.array_equal(gc, counts) and np.array_equal(go, offsets) and np.array_equal(gi, indices)
    got = sim.parts.host()
    assert np.array_equal(got["id"], ps.ids)
    assert np.array_equal(got["m"], ps.m)
    ps64 = _oracle_particles(sc).astype(np.float64)
    O.Context(np.float64).run(prm, ps64, 2, reorder=True)
    for name, ref in (("x", ps.x), ("v", ps.v), ("C", ps.C), ("F", ps.F), ("logJp", ps.logJp)):
        check_field(f"{name} (n={k})", got[name], ref, getattr(ps64, name), TOL_1STEP)


def test_single_particle_at_rest_known_answer():
    """One particle at rest with F = I, C = 0 and no gravity: P2G puts m W(o)
    on its 27 nodes (weights of the quadratic B-spline, summing to m) and no
    momentum; G2P brings back v = 0, C = 0 and leaves x and F unchanged."""
    sc = S.cube_scene(cells=4)
    sc = _subset(sc, [int(sc.n // 2)], "one")
    sc = dataclasses.replace(sc, gravity=(0.0, 0.0, 0.0))
    sc.v[:] = 0.0
    sc.C[:] = 0.0
    sc.F[:] = np.eye(3, dtype=np.float32).ravel()
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    g = sim.grid.data[:sim.nblocks()].cpu().numpy().astype(np.float64)  # [blocks, 7, 64]
    mass = g[:, 0, :]
    assert np.count_nonzero(mass) == 27
    assert abs(float(mass.sum()) - float(sc.m[0])) <= 1e-6 * float(sc.m[0])
    # the 27 masses are m w0(i) w1(j) w2(k), the quadratic B-spline weights
    X = sc.x[0].astype(np.float64) / sc.dx
    base = np.floor(X - 0.5)
    d = X - base
    w = np.stack([0.5 * (1.5 - d) ** 2, 0.75 - (d - 1.0) ** 2, 0.5 * (d - 0.5) ** 2])  # [3 nodes, 3 axes]
    want = np.sort((sc.m[0] * np.einsum("i,j,k->ijk", w[:, 0], w[:, 1], w[:, 2])).ravel())
    assert np.allclose(np.sort(mass[mass != 0]), want, rtol=1e-5, atol=0)
    assert np.all(np.abs(g[:, 1:4, :]) <= 1e-7 * float(sc.m[0]))  # no momentum
    sim.grid_update()
    sim.g2p()
    sim.sync()
    got = sim.parts.host()
    # the stress of F = I vanishes up to rounding, so v and C do too
    assert np.allclose(got["v"], 0.0, atol=1e-6)
    assert np.allclose(got["C"], 0.0, atol=1e-4)
    assert np.allclose(got["x"], sc.x, rtol=0, atol=1e-7)
    assert np.allclose(got["F"].ravel(), np.eye(3, dtype=np.float32).ravel(), atol=1e-6)


def _with_n(desc, n):
    d = type(desc).from_buffer_copy(desc)
    d.n = n
    return d


def test_empty_inputs(sand):
    """n = 0 with valid buffers: activation leaves the table empty, the bucket
    sort writes all-zero counts and offsets, and a sorted substep runs and
    activates nothing."""
    rocm = _rocm()
    sim = rocm.Simulator(sand, reorder=True)
    L = rocm.lib()
    sim.clear_table()
    sim.activate(n=0)
    sim.enlarge()
    sim.sync()
    assert sim.table.host()["cnt"] == 0
    sim.buckets.counts.fill_(7)
    sim.buckets.offsets.fill_(7)
    rocm.check(L.zsmpm_bucket_sort__rocm(sim.pol.h, C.byref(sim.table.desc), sim.parts.x.data_ptr(), 0,
                                         sim.grid.dx, 4, -2, 0.5, C.byref(sim.buckets.desc)), "bucket_sort")
    sim.sync()
    assert int(sim.buckets.counts.abs().sum()) == 0
    assert int(sim.buckets.offsets.abs().sum()) == 0
    a = _with_n(sim.parts.desc, 0)
    b = _with_n(sim.spare.desc, 0)
    x0 = sim.parts.x.clone()
    rocm.check(L.zsmpm_substep_sorted__rocm(sim.pol.h, C.byref(sim.model), C.byref(a), C.byref(b),
                                            C.byref(sim.table.desc), C.byref(sim.buckets.desc),
                                            C.byref(sim.grid.desc), float(sand.dt), C.byref(sim.extf), sim.cols,
                                            sim.ncols), "substep_sorted")
    sim.sync()
    assert sim.table.host()["cnt"] == 0
    assert bool((sim.parts.x == x0).all())


def test_empty_transfers_and_implicit(sand):
    """n = 0 on an empty table: P2C2G, G2C2P and G2P2G succeed, touch no
    particle and add nothing."""
    rocm = _rocm()
    sim = rocm.Simulator(sand)
    L = rocm.lib()
    sim.clear_table()
    sim.activate(n=0)
    sim.enlarge()
    sim.sync()
    assert sim.table.host()["cnt"] == 0
    e = _with_n(sim.parts.desc, 0)
    x0 = sim.parts.x.clone()
    sim.grid.data.fill_(0.0)
    rocm.check(L.zsmpm_p2c2g__rocm(sim.pol.h, C.byref(sim.model), C.byref(e), C.byref(sim.table.desc), None,
                                   C.byref(sim.grid.desc), float(sand.dt)), "p2c2g")
    rocm.check(L.zsmpm_g2c2p__rocm(sim.pol.h, C.byref(e), C.byref(sim.table.desc), None, C.byref(sim.grid.desc),
                                   float(sand.dt)), "g2c2p")
    v = sim.dof_vector(1.0)
    r = sim.dof_vector(0.0)
    rocm.check(L.zsmpm_g2p2g__rocm(sim.pol.h, C.byref(sim.model), C.byref(e), C.byref(sim.table.desc),
                                   C.byref(sim.grid.desc), float(sand.dt), v.data_ptr(), r.data_ptr()), "g2p2g")
    sim.sync()
    assert bool((sim.parts.x == x0).all())
    assert float(sim.grid.data.abs().sum()) == 0.0
    assert float(r.abs().sum()) == 0.0


def test_key_range_reported(sand):
    """A particle whose block key leaves [-2^20, 2^20) is reported as
    ZSMPM_KEY_RANGE by the activation and by the substep."""
    sc = _subset(sand, np.arange(256), "far")
    sc.x[100] = np.array([1.0e6, 0.5, 0.5], np.float32)
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    with pytest.raises(rocm.ZsmpmError) as ei:
        sim.sync()
    assert ei.value.code == -8
    sim2 = rocm.Simulator(sc, reorder=True)
    with pytest.raises(rocm.ZsmpmError) as ei:
        sim2.substep()
        sim2.sync()
    assert ei.value.code == -8
