"""Parity of the gfx950 HIP path (through the C ABI) with the oracle.

Integer work (hash table slots, block numbers, bucket sort) must match exactly;
floating-point fields must match within the stated relative tolerance, where
the error of a field is max|gpu - oracle| / max|oracle| over the field.
"""
import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu

TOL_1STEP = 1e-5
TOL_100STEP = 1e-3


def rel_err(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = max(float(np.max(np.abs(b))), 1e-30) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / den) if b.size else 0.0


def check_field(name, gpu, ref32, ref64, tol):
    """gpu must match the FP32 oracle within tol, or -- where float32 summation
    order decides (forces that cancel across ~200 contributions per node) -- be
    at least as close to exact arithmetic (the FP64 twin) as the FP32 oracle is,
    within a factor 2."""
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


def _oracle_particles(sc):
    return O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)


def _assert_table_equal(got, ref):
    assert got["cnt"] == ref["cnt"]
    assert np.array_equal(got["indices"], ref["indices"]), "slot -> block index differs"
    assert np.array_equal(got["keys"], ref["keys"]), "slot keys differ"
    assert np.array_equal(got["status"], ref["status"])
    assert np.array_equal(got["active"], ref["active"]), "_activeKeys differ"


@pytest.fixture(scope="module")
def cube():
    return S.perturb(S.cube_scene(cells=20), seed=11)


@pytest.fixture(scope="module")
def sand():
    return S.perturb(S.sand_scene(dims=(16, 32, 16)), seed=12, f_scale=0.02)


def test_abi_version():
    rocm = _rocm()
    h = rocm.Header()
    assert rocm.lib().zsmpm_abi_version(h) == 0
    assert (h.major, h.minor) == (1, 0)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_activation_enlarge_exact(scene, request):
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    ctx = O.Context()
    sim.clear_table()
    sim.activate()
    sim.sync()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    _assert_table_equal(sim.table.host(), ctx.table())
    sim.enlarge()
    sim.sync()
    ctx.enlarge()
    _assert_table_equal(sim.table.host(), ctx.table())


def test_enlarge_27_offsets_exact():
    """An enlargement by all 27 neighbour offsets (lo -1, hi 2) with more active
    blocks than 8 T / 27 (343 blocks in a table of 1,024 slots), so that the
    candidates' insertion orders reach past 8 T while the table still holds
    the result (729 blocks)."""
    sc = S.perturb(S.cube_scene(cells=24), seed=11)
    sc.table_size = 1024
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    ctx = O.Context()
    sim.clear_table()
    sim.activate()
    sim.sync()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    n0 = ctx.sizes()["cnt"]
    print(f"27-offset enlargement: {n0} active blocks, T = {sc.table_size}", flush=True)
    assert n0 * 27 > 8 * sc.table_size
    _assert_table_equal(sim.table.host(), ctx.table())
    sim.enlarge(lo=(-1, -1, -1), hi=(2, 2, 2))
    sim.sync()
    ctx.enlarge((-1, -1, -1), (2, 2, 2))
    assert ctx.sizes()["cnt"] == 729
    _assert_table_equal(sim.table.host(), ctx.table())


def test_activation_shuffled_order_exact(cube):
    """First-occurrence numbering must follow particle order, not memory luck."""
    rng = np.random.default_rng(3)
    perm = rng.permutation(cube.n)
    x = np.ascontiguousarray(cube.x[perm])
    import torch

    rocm = _rocm()
    sim = rocm.Simulator(cube)
    xd = torch.as_tensor(x).cuda()
    sim.clear_table()
    sim.activate(x=xd, n=cube.n)
    sim.enlarge()
    sim.sync()
    ctx = O.Context()
    ctx.activate(cube.table_size, x, cube.dx)
    ctx.enlarge()
    _assert_table_equal(sim.table.host(), ctx.table())


def test_query_matches_table(cube):
    import torch

    rocm = _rocm()
    sim = rocm.Simulator(cube)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.sync()
    t = sim.table.host()
    keys = np.concatenate([t["active"], t["active"] + np.array([[1000, 0, 0]], np.int32)]).astype(np.int32)
    kd = torch.as_tensor(keys).cuda().contiguous()
    out = torch.zeros(keys.shape[0], dtype=torch.int32, device="cuda")
    rocm.check(rocm.lib().zsmpm_table_query__rocm(sim.pol.h, sim.table.desc, kd.data_ptr(), keys.shape[0],
                                                  out.data_ptr()), "query")
    sim.sync()
    o = out.cpu().numpy()
    cnt = t["cnt"]
    assert np.array_equal(o[:cnt], np.arange(cnt))
    assert np.all(o[cnt:] == -1)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_bucket_sort_exact(scene, request):
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.sync()
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    counts, offsets, indices = ctx.sort()
    nb = ctx.sizes()["cnt"]
    gc, go, gi = sim.buckets.host(nb, sc.n)
    assert np.array_equal(gc, counts)
    assert np.array_equal(go, offsets)
    assert np.array_equal(gi, indices)


def test_bucket_sort_shuffled_exact(cube):
    """Shuffled particles: nearly every particle starts its own run."""
    import torch

    rng = np.random.default_rng(17)
    perm = rng.permutation(cube.n)
    x = np.ascontiguousarray(cube.x[perm])
    rocm = _rocm()
    sim = rocm.Simulator(cube)
    sim.parts.x.copy_(torch.as_tensor(x).cuda())
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.sync()
    ctx = O.Context()
    ctx.activate(cube.table_size, x, cube.dx)
    ctx.enlarge()
    ctx.bucket_sort(x, cube.dx)
    counts, offsets, indices = ctx.sort()
    gc, go, gi = sim.buckets.host(ctx.sizes()["cnt"], cube.n)
    assert sim.pol.counters()[2] > cube.n // 2  # runs
    assert np.array_equal(gc, counts)
    assert np.array_equal(go, offsets)
    assert np.array_equal(gi, indices)


@pytest.mark.parametrize("cap_log2", [17, 18])
def test_bucket_sort_wide_buckets_exact(cube, cap_log2):
    """Bucket capacities of 2^17 and 2^18 blocks: the runs are radix-sorted
    in two 9-bit passes (three 8-bit passes below 2^16 keys and above 2^18).
    Shuffled particles make nearly every particle its own run."""
    import torch

    rng = np.random.default_rng(23 + cap_log2)
    perm = rng.permutation(cube.n)
    x = np.ascontiguousarray(cube.x[perm])
    rocm = _rocm()
    sim = rocm.Simulator(cube)
    sim.buckets = rocm.DeviceBuckets(1 << cap_log2, cube.n)
    sim.parts.x.copy_(torch.as_tensor(x).cuda())
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.sync()
    ctx = O.Context()
    ctx.activate(cube.table_size, x, cube.dx)
    ctx.enlarge()
    ctx.bucket_sort(x, cube.dx)
    counts, offsets, indices = ctx.sort()
    gc, go, gi = sim.buckets.host(ctx.sizes()["cnt"], cube.n)
    assert sim.pol.counters()[2] > cube.n // 2  # runs
    assert np.array_equal(gc, counts)
    assert np.array_equal(go, offsets)
    assert np.array_equal(gi, indices)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_stages_parity(scene, request):
    """P2G grid (m, mv, rhs), grid update (velocities) and G2P state, stage by stage."""
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    nb = sim.nblocks()
    g_p2g = sim.grid.data[:nb].cpu().numpy()
    lj_gpu = None if sim.parts.logJp is None else sim.parts.logJp.cpu().numpy()
    sim.grid_update()
    sim.sync()
    g_v = sim.grid.data[:nb].cpu().numpy()
    sim.g2p()
    sim.sync()
    got = sim.parts.host()

    ps = _oracle_particles(sc)
    prm = O.make_params(sc)
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    ps_p2g64 = _oracle_particles(sc).astype(np.float64)
    ref_p2g64 = ctx.p2g_f64(prm, ps_p2g64)
    ctx.p2g(prm, ps)
    ref_p2g, _, _ = ctx.grids()
    if lj_gpu is not None:
        check_field("logJp", lj_gpu, ps.logJp, ps_p2g64.logJp, TOL_1STEP)
    ctx.grid_update(prm)
    _, ref_v, _ = ctx.grids()
    ctx.g2p(prm, ps)
    assert ref_p2g.shape == g_p2g.shape
    for ch, name in ((0, "m"), (slice(1, 4), "mv"), (slice(4, 7), "rhs")):
        check_field(f"P2G {name}", g_p2g[:, ch], ref_p2g[:, ch], ref_p2g64[:, ch], TOL_1STEP)
    ps64 = _oracle_particles(sc).astype(np.float64)
    ctx64 = O.Context(np.float64)
    ctx64.run(prm, ps64, 1)
    _, ref_v64, _ = ctx64.grids()
    if ref_v64.shape == ref_v.shape:
        check_field("grid velocity", g_v[:, 1:4], ref_v[:, 1:4], ref_v64[:, 1:4], TOL_1STEP)
    else:
        assert rel_err(g_v[:, 1:4], ref_v[:, 1:4]) < TOL_1STEP
    for name, ref in (("x", ps.x), ("v", ps.v), ("C", ps.C), ("F", ps.F)):
        check_field(f"G2P {name}", got[name], ref, getattr(ps64, name), TOL_1STEP)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_substep_parity(scene, request):
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    sim.substep()
    sim.sync()
    ps = _oracle_particles(sc)
    ctx = O.Context()
    ctx.run(O.make_params(sc), ps, 1)
    _assert_table_equal(sim.table.host(), ctx.table())
    counts, offsets, indices = ctx.sort()
    gc, go, gi = sim.buckets.host(ctx.sizes()["cnt"], sc.n)
    assert np.array_equal(gc, counts) and np.array_equal(go, offsets) and np.array_equal(gi, indices)
    got = sim.parts.host()
    ps64 = _oracle_particles(sc).astype(np.float64)
    O.Context(np.float64).run(O.make_params(sc), ps64, 1)
    for name, ref in (("x", ps.x), ("v", ps.v), ("C", ps.C), ("F", ps.F)):
        check_field(name, got[name], ref, getattr(ps64, name), TOL_1STEP)
    if ps.logJp is not None:
        check_field("logJp", got["logJp"], ps.logJp, ps64.logJp, TOL_1STEP)


@pytest.mark.parametrize("scene,steps", [("cube", 100), ("sand", 100)])
def test_multistep_drift(scene, steps, request):
    base = request.getfixturevalue(scene)
    sc = S.cube_scene(cells=20) if scene == "cube" else S.sand_scene(dims=(16, 32, 16))
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    for _ in range(steps):
        sim.substep()
    sim.sync()
    ps = _oracle_particles(sc)
    ctx = O.Context()
    ctx.run(O.make_params(sc), ps, steps)
    got = sim.parts.host()
    for name, ref in (("x", ps.x), ("v", ps.v), ("F", ps.F)):
        e = rel_err(got[name], ref)
        assert e < TOL_100STEP, f"{name} after {steps}: {e:.3e}"
    # displacement error relative to the motion itself
    disp_ref = ps.x - sc.x.ravel()
    disp_err = np.max(np.abs(got["x"].ravel() - ps.x)) / max(np.max(np.abs(disp_ref)), 1e-30)
    assert disp_err < TOL_100STEP, f"displacement error {disp_err:.3e}"
    del base


def test_reorder_gather_exact(sand):
    """zsmpm_reorder__rocm is Vector::reorder's gather on every attribute."""
    rocm = _rocm()
    sim = rocm.Simulator(sand, reorder=True)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.reorder_particles()
    sim.sync()
    _, _, idx = sim.buckets.host(sim.nblocks(), sand.n)
    ps = _oracle_particles(sand)
    ps.reorder(idx)
    got = sim.parts.host()
    for k in ("x", "v", "C", "F", "m", "logJp"):
        assert np.array_equal(got[k].ravel(), getattr(ps, k)), k
    assert np.array_equal(got["id"], idx)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_sorted_substep_parity(scene, request):
    """The substep with the particle reorder fused into G2P: same table and
    sort as the oracle, particles in the oracle's bucket order (ids exact),
    fields within tolerance."""
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    sim.substep()
    sim.sync()
    prm = O.make_params(sc)
    ps = _oracle_particles(sc)
    ctx = O.Context()
    ctx.run(prm, ps, 1, reorder=True)
    _assert_table_equal(sim.table.host(), ctx.table())
    counts, offsets, indices = ctx.sort()
    gc, go, gi = sim.buckets.host(ctx.sizes()["cnt"], sc.n)
    assert np.array_equal(gc, counts) and np.array_equal(go, offsets) and np.array_equal(gi, indices)
    got = sim.parts.host()
    assert np.array_equal(got["id"], ps.ids)
    assert np.array_equal(got["m"], ps.m)
    ps64 = _oracle_particles(sc).astype(np.float64)
    O.Context(np.float64).run(prm, ps64, 1, reorder=True)
    assert np.array_equal(ps64.ids, ps.ids)
    for name, ref in (("x", ps.x), ("v", ps.v), ("C", ps.C), ("F", ps.F)):
        check_field(name, got[name], ref, getattr(ps64, name), TOL_1STEP)
    if ps.logJp is not None:
        check_field("logJp", got["logJp"], ps.logJp, ps64.logJp, TOL_1STEP)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_sorted_multistep_drift(scene):
    """100 reordered substeps against the plain (never reordered) oracle run,
    compared particle by particle through the carried ids."""
    steps = 100
    sc = S.cube_scene(cells=20) if scene == "cube" else S.sand_scene(dims=(16, 32, 16))
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    for _ in range(steps):
        sim.substep()
    sim.sync()
    got = sim.parts.host_by_id()
    ps = _oracle_particles(sc)
    O.Context().run(O.make_params(sc), ps, steps)
    for name in ("x", "v", "F"):
        e = rel_err(got[name], getattr(ps, name))
        assert e < TOL_100STEP, f"{name} after {steps}: {e:.3e}"
    disp_ref = ps.x - sc.x.ravel()
    disp_err = np.max(np.abs(got["x"].ravel() - ps.x)) / max(np.max(np.abs(disp_ref)), 1e-30)
    assert disp_err < TOL_100STEP, f"displacement error {disp_err:.3e}"
    assert np.array_equal(got["m"], ps.m)


@pytest.mark.parametrize("scene", ["cube", "sand"])
@pytest.mark.parametrize("steps,tol", [(2, TOL_1STEP), (5, 1e-4)])
def test_fused_substeps_match_single_substeps(scene, steps, tol, request):
    """zsmpm_substeps_sorted__rocm (the next substep's stress computed inside
    G2P) and the single sorted substeps both stay on the oracle, logJp
    included. Two GPU runs differ by float atomic order in P2G, so each is
    held to the oracle (FP64-anchored where float32 order decides), not to
    the other bit for bit."""
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    fused = rocm.Simulator(sc, reorder=True)
    fused.substeps(steps, fuse_stress=True)
    single = rocm.Simulator(sc, reorder=True)
    for _ in range(steps):
        single.substep()
    fused.sync()
    single.sync()
    a = fused.parts.host_by_id()
    b = single.parts.host_by_id()
    ps = _oracle_particles(sc)
    O.Context().run(O.make_params(sc), ps, steps, reorder=True)
    ps64 = _oracle_particles(sc).astype(np.float64)
    O.Context(np.float64).run(O.make_params(sc), ps64, steps, reorder=True)
    for name in ("x", "v", "C", "F", "m", "logJp"):
        if a[name] is None:
            continue
        check_field(f"{name} (fused vs oracle)", a[name], ps.by_id(name), ps64.by_id(name), tol)
        check_field(f"{name} (single vs oracle)", b[name], ps.by_id(name), ps64.by_id(name), tol)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_graph_substeps_replays(scene, request):
    """ZSMPM_SUBSTEPS_GRAPH: the captured pair of substeps is replayed three
    times with eager work (another simulator's substeps, host read-backs)
    between the replays; the state after every replay, and after an odd call
    (one replay plus one ordinary substep), stays on the oracle."""
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    other = rocm.Simulator(sc, reorder=True)
    ps = _oracle_particles(sc)
    ps64 = _oracle_particles(sc).astype(np.float64)
    params = O.make_params(sc)
    done = 0
    for calls, tol in ((2, TOL_1STEP), (2, 1e-4), (2, 1e-4), (3, 1e-4)):
        sim.substeps(calls, graph=True)
        other.substep()
        other.sync()
        sim.sync()
        got = sim.parts.host_by_id()
        O.Context().run(params, ps, calls, reorder=True)
        O.Context(np.float64).run(params, ps64, calls, reorder=True)
        done += calls
        for name in ("x", "v", "C", "F", "m", "logJp"):
            if got[name] is None:
                continue
            check_field(f"{name} after {done} substeps", got[name], ps.by_id(name), ps64.by_id(name), tol)


@pytest.mark.parametrize("scene", ["cube", "sand"])
def test_autotuned_substeps(scene, request):
    """ZSMPM_SUBSTEPS_AUTOTUNE: the 8 tuning substeps (four P2G / G2P
    workgroup counts) and the graph-replayed rest are all real substeps;
    the state after 10 stays on the oracle and the counts kept are
    candidates."""
    sc = request.getfixturevalue(scene)
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    sim.substeps(10, graph=True, autotune=True)
    sim.sync()
    p2g, g2p = sim.pol.launch_widths()
    assert p2g in (512, 768, 1024, 2048) and g2p in (8192, 16384, 32768, 65536)
    got = sim.parts.host_by_id()
    ps = _oracle_particles(sc)
    O.Context().run(O.make_params(sc), ps, 10, reorder=True)
    ps64 = _oracle_particles(sc).astype(np.float64)
    O.Context(np.float64).run(O.make_params(sc), ps64, 10, reorder=True)
    for name in ("x", "v", "F"):
        check_field(f"{name} after 10 autotuned substeps", got[name], ps.by_id(name), ps64.by_id(name), 1e-4)


def test_set_launch_widths_substeps(sand):
    """Fixed P2G / G2P workgroup counts (zsmpm_policy_set_launch_widths__rocm),
    odd ones included: P2G drains its work queue and G2P strides over the
    buckets whatever the count, so the substeps stay on the oracle."""
    rocm = _rocm()
    sim = rocm.Simulator(sand, reorder=True)
    sim.pol.set_launch_widths(37, 1000)
    assert sim.pol.launch_widths() == (37, 1000)
    sim.substeps(4, graph=True)
    sim.sync()
    got = sim.parts.host_by_id()
    ps = _oracle_particles(sand)
    O.Context().run(O.make_params(sand), ps, 4, reorder=True)
    ps64 = _oracle_particles(sand).astype(np.float64)
    O.Context(np.float64).run(O.make_params(sand), ps64, 4, reorder=True)
    for name in ("x", "v", "F"):
        check_field(f"{name} after 4 substeps at 37 / 1000 workgroups", got[name], ps.by_id(name),
                    ps64.by_id(name), 1e-4)
    with pytest.raises(rocm.ZsmpmError):
        sim.pol.set_launch_widths(-1, 1000)


@pytest.mark.parametrize("tiles", [(256, 256), (576, 256), (256, 128)])
def test_launch_tiles_substeps(sand, tiles):
    """Each P2G chunk / G2P workgroup shape the autotuner chooses between
    (zsmpm_policy_set_launch_tiles__rocm), with graph replays: the substeps
    stay on the oracle."""
    rocm = _rocm()
    sim = rocm.Simulator(sand, reorder=True)
    sim.pol.set_launch_tiles(*tiles)
    assert sim.pol.launch_tiles() == tiles
    sim.substeps(4, graph=True)
    sim.sync()
    got = sim.parts.host_by_id()
    ps = _oracle_particles(sand)
    O.Context().run(O.make_params(sand), ps, 4, reorder=True)
    ps64 = _oracle_particles(sand).astype(np.float64)
    O.Context(np.float64).run(O.make_params(sand), ps64, 4, reorder=True)
    for name in ("x", "v", "F"):
        check_field(f"{name} after 4 substeps at tiles {tiles}", got[name], ps.by_id(name), ps64.by_id(name), 1e-4)
    with pytest.raises(rocm.ZsmpmError):
        sim.pol.set_launch_tiles(300, 128)


def _first_occurrence(packed):
    """Distinct values of `packed` in order of first occurrence."""
    u, first = np.unique(packed, return_index=True)
    return u[np.argsort(first, kind="stable")]


def _pack(k):
    k = k.astype(np.int64) + (1 << 20)
    return (k[:, 0] << 42) | (k[:, 1] << 21) | k[:, 2]


def test_full_size_properties():
    """The 16,777,216-particle sand column (the bench workload) after two
    sorted substeps, stage by stage: block numbering after activation and
    enlargement equals a numpy restatement of first-occurrence order; the
    bucket sort equals a stable argsort of the block numbers; P2G conserves
    mass and momentum (the stress terms sum to zero over a particle's
    stencil); a sorted substep keeps the particle multiset (ids a permutation,
    total mass)."""
    sc = S.sand_scene()
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    sim.substeps(2)
    sim.sync()
    x = sim.parts.x.cpu().numpy()
    sim.clear_table()
    sim.activate()
    sim.sync()
    t = sim.table.host()
    keys = O.block_keys(x, sc.dx)
    pk = _pack(keys)
    order = _first_occurrence(pk)
    assert t["cnt"] == order.size
    assert np.array_equal(_pack(t["active"]), order)
    sim.enlarge()
    sim.sync()
    t = sim.table.host()
    offs = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], np.int64)
    akeys = t["active"][: order.size].astype(np.int64)
    cand = _pack((akeys[:, None, :] + offs[None]).reshape(-1, 3))
    new = _first_occurrence(cand[~np.isin(cand, order)])
    want = np.concatenate([order, new])
    assert t["cnt"] == want.size
    assert np.array_equal(_pack(t["active"]), want)
    sim.bucket_sort()
    sim.sync()
    nb = t["cnt"]
    srt = np.argsort(want, kind="stable")
    block = srt[np.searchsorted(want[srt], pk)]
    counts, offsets, indices = sim.buckets.host(nb, sc.n)
    bc = np.bincount(block, minlength=nb)
    assert np.array_equal(counts[:nb], bc)
    assert np.array_equal(offsets[:nb], np.concatenate([[0], np.cumsum(bc)[:-1]]))
    assert np.array_equal(indices, np.argsort(block, kind="stable"))
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    g = sim.grid.data[:nb].cpu().numpy().astype(np.float64)
    m = sim.parts.m.cpu().numpy().astype(np.float64)
    v = sim.parts.v.cpu().numpy().astype(np.float64)
    assert abs(g[:, 0].sum() - m.sum()) <= 1e-5 * m.sum()
    mom_p = (m[:, None] * v).sum(0)
    mom_g = g[:, 1:4].sum((0, 2)) + g[:, 4:7].sum((0, 2))
    scale = (m[:, None] * np.abs(v)).sum()
    assert np.all(np.abs(mom_g - mom_p) <= 1e-4 * scale), (mom_g, mom_p, scale)
    # one sorted substep: ids stay a permutation, the mass is moved, not changed
    sim.substeps(1)
    sim.sync()
    ids = sim.parts.id.cpu().numpy()
    assert np.array_equal(np.sort(ids), np.arange(sc.n, dtype=ids.dtype))
    assert abs(float(sim.parts.m.double().sum().cpu()) - m.sum()) <= 1e-9 * m.sum()


def test_fused_substeps_drift_100(sand):
    """100 fused substeps of the sand column against the plain oracle, per id:
    x, v and F within 1e-3. logJp is not bounded here: it follows the
    return-mapping branch, which roundoff alone can flip for sand at rest, so
    two oracle runs that differ only in summation order already disagree by
    ~0.3 (relative to its max) after 100 substeps; it is checked at 1 and 5
    substeps instead."""
    sc = S.sand_scene(dims=(16, 32, 16))
    rocm = _rocm()
    sim = rocm.Simulator(sc, reorder=True)
    sim.substeps(60, fuse_stress=True)
    sim.substeps(40)
    sim.sync()
    got = sim.parts.host_by_id()
    ps = _oracle_particles(sc)
    O.Context().run(O.make_params(sc), ps, 100)
    for name in ("x", "v", "F"):
        e = rel_err(got[name], getattr(ps, name))
        assert e < TOL_100STEP, f"{name} after 100: {e:.3e}"
    del sand


@pytest.mark.parametrize("T", [32, 64, 128, 256])
def test_small_tables_exact(T):
    """Probe wrap for tables smaller than the probe step (127)."""
    rocm = _rocm()
    sc = S.tiny_scene(n_cells=(2, 3, 2), ppc=2)
    sc.table_size = T
    sim = rocm.Simulator(sc, grid_capacity=64)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.sync()
    ctx = O.Context()
    ctx.activate(T, sc.x, sc.dx)
    ctx.enlarge()
    ref = ctx.table()
    assert ref["cnt"] <= T
    _assert_table_equal(sim.table.host(), ref)


def test_table_full_reported():
    rocm = _rocm()
    sc = S.cube_scene(cells=12)
    sc.table_size = 16  # far too small for the ~64 blocks this scene activates
    sim = rocm.Simulator(sc, grid_capacity=64)
    sim.clear_table()
    sim.activate()
    with pytest.raises(rocm.ZsmpmError) as ei:
        sim.sync()
    assert ei.value.code == -7


def test_grid_capacity_reported():
    rocm = _rocm()
    sc = S.cube_scene(cells=12)
    sim = rocm.Simulator(sc, grid_capacity=4)
    with pytest.raises(rocm.ZsmpmError) as ei:
        sim.substep()
        sim.sync()
    assert ei.value.code in (-11,)
