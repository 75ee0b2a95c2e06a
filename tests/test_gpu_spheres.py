"""Two colliding elastic spheres (fixed-corotated, E = 5e4, nu = 0.4,
rho = 1e3, 8 particles per cell, +-2 m/s along x, dx = 1/256) on the GPU.

- Shrunk (r = 0.03, centres 0.08 apart, 30K particles, contact after about
  50 substeps): one sorted substep against the oracle (table, bucket sort and
  ids exact; fields within 1e-5, FP64-anchored where float32 summation order
  decides) and 200 substeps through the collision within 1e-3.
- Full size (r = 0.15 at x = 0.3 and 0.7, 3.79M particles): the first
  substep's table and bucket sort exact against the oracle; then 400
  substeps, through first contact at about 220, checked through size-
  independent properties: every particle id present once, the momentum
  balance P(t) = P(0) + M g t, the spheres slowed by the contact.
"""
import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu


def _rocm():
    from zpc_amd import rocm

    return rocm


def rel_err(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-30))


def check_field(name, gpu, ref32, ref64, tol):
    e32 = rel_err(gpu, ref32)
    if e32 < tol:
        return
    e_gpu = rel_err(gpu, ref64)
    e_ref = rel_err(ref32, ref64)
    assert e_gpu <= max(tol, 2.0 * e_ref), f"{name}: gpu-vs-fp32 {e32:.3e}, gpu-vs-fp64 {e_gpu:.3e}, fp32-vs-fp64 {e_ref:.3e}"


@pytest.fixture(scope="module")
def small():
    return S.spheres_scene(r=0.03, centers=(0.46, 0.54))


def _oracle(sc, dtype=np.float32):
    return O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp).astype(dtype)


def test_spheres_one_substep(small):
    rocm = _rocm()
    print(f"spheres r=0.03: {small.n} particles", flush=True)
    sim = rocm.Simulator(small, reorder=True)
    sim.substep()
    sim.sync()
    ctx = O.Context()
    ps = _oracle(small)
    ctx.run(O.make_params(small), ps, 1, reorder=True)
    t = sim.table.host()
    r = ctx.table()
    assert t["cnt"] == r["cnt"]
    for k in ("keys", "indices", "status", "active"):
        assert np.array_equal(t[k], r[k]), k
    counts, offsets, indices = ctx.sort()
    gc, go, gi = sim.buckets.host(r["cnt"], small.n)
    assert np.array_equal(gc, counts) and np.array_equal(go, offsets) and np.array_equal(gi, indices)
    got = sim.parts.host()
    assert np.array_equal(got["id"], ps.ids)
    ps64 = _oracle(small, np.float64)
    O.Context(np.float64).run(O.make_params(small), ps64, 1, reorder=True)
    for name in ("x", "v", "C", "F"):
        check_field(name, got[name], getattr(ps, name), getattr(ps64, name), 1e-5)


def test_spheres_200_substeps_through_contact(small):
    rocm = _rocm()
    steps = 200
    sim = rocm.Simulator(small, reorder=True)
    sim.substeps(steps, graph=True)
    sim.sync()
    got = sim.parts.host_by_id()
    ps = _oracle(small)
    O.Context().run(O.make_params(small), ps, steps)
    n1 = int(np.sum(small.x[:, 0] < 0.5))
    vx1 = float(np.mean(ps.v.reshape(-1, 3)[:n1, 0]))
    print(f"200 substeps: oracle mean vx of sphere 1 = {vx1:.3f} m/s (2.0 before contact)", flush=True)
    assert vx1 < 1.9, "the spheres must have collided within the run"
    for name in ("x", "v", "F"):
        e = rel_err(got[name], getattr(ps, name))
        assert e < 1e-3, f"{name} after {steps} substeps: {e:.3e}"
    disp = ps.x - small.x.ravel()
    de = np.max(np.abs(got["x"].ravel() - ps.x)) / np.max(np.abs(disp))
    assert de < 1e-3, f"displacement error {de:.3e}"


def test_spheres_full_size_properties():
    rocm = _rocm()
    import torch

    sc = S.spheres_scene()
    n = sc.n
    print(f"spheres r=0.15: {n} particles, table {sc.table_size}", flush=True)
    # first substep's sparsity against the oracle (integer work, exact)
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    r = ctx.table()
    counts, offsets, indices = ctx.sort()
    sim = rocm.Simulator(sc, reorder=True)
    sim.substep()
    sim.sync()
    t = sim.table.host()
    assert t["cnt"] == r["cnt"]
    for k in ("keys", "indices", "active"):
        assert np.array_equal(t[k], r[k]), k
    gc, go, gi = sim.buckets.host(r["cnt"], n)
    assert np.array_equal(gc, counts) and np.array_equal(go, offsets) and np.array_equal(gi, indices)
    print(f"first substep exact: {r['cnt']} blocks", flush=True)
    m = torch.as_tensor(sc.m).cuda().double()
    P0 = (torch.as_tensor(sc.v).cuda().double() * m[:, None]).sum(0).cpu().numpy()
    M = float(m.sum().item())
    steps = 400
    sim.substeps(steps - 1, graph=True)
    sim.sync()
    ids = sim.parts.id.cpu().numpy()
    assert np.array_equal(np.sort(ids), np.arange(n)), "every particle id present once"
    mm = sim.parts.m.double()
    P = (sim.parts.v.reshape(-1, 3).double() * mm[:, None]).sum(0).cpu().numpy()
    P_exp = P0 + M * np.array(sc.gravity) * sc.dt * steps
    scale = M * 2.0
    err = np.max(np.abs(P - P_exp)) / scale
    print(f"{steps} substeps: momentum balance error {err:.2e} of M*2 m/s", flush=True)
    assert err < 1e-4, (P, P_exp)
    v = sim.parts.v.reshape(-1, 3).cpu().numpy()
    first = ids < int(np.sum(sc.x[:, 0] < 0.5))
    vx1 = float(np.mean(v[first, 0]))
    print(f"mean vx of sphere 1 after {steps} substeps: {vx1:.3f} m/s", flush=True)
    assert vx1 < 1.95, "contact slowed the spheres"
    assert np.all(np.isfinite(v))
