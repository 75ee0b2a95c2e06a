"""GPU parity for the grid update's colliders and max |v|^2, the grid angular
momentum, and the FP64 SVD path, through the C ABI against the oracle.

- zsmpm_grid_update__rocm on a synthetic grid (random masses with massless
  nodes, momenta and forces) with a sticky, slip or separate plane, oblique
  or through a node layer: node velocities within 1e-6 of the oracle's
  (GridOp.hpp:71-164, Collider.h:80-107) and max |v|^2 equal to the
  oracle's maxVel, merged into a caller value that starts negative.
- max |v|^2 of a whole substep against the oracle substep's maxVel.
- zsmpm_grid_angular_momentum__rocm (GridOp.hpp:216-259) on the GPU's own
  P2G grid: equal to the oracle's sum over the same grid (the terms are
  formed identically; only the double summation order differs), and equal to
  the APIC angular momentum of the particles.
- svd_fp64_gap: with a gap that sends every particle to the FP64 SVD the
  stress equals the FP64 oracle's (its own Jacobi SVD) to float rounding on
  the golden inputs (repeated, rank-deficient, negative-determinant and
  random F); at gap 1e-6 each particle's stress is bit-equal to either the
  FP32 or the FP64 path, and the count of FP64 ones is what counter 3 says.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu


def _rocm():
    from zpc_amd import rocm

    return rocm


def _torch():
    import torch

    return torch


def _seeded_grid(table_cnt, seed):
    rng = np.random.default_rng(seed)
    g = np.zeros((table_cnt, 7, 64), np.float32)
    g[:, 0, :] = rng.uniform(0.5, 2.0, size=(table_cnt, 64))
    g[:, 0, ::7] = 0.0
    g[:, 1:4, :] = rng.normal(size=(table_cnt, 3, 64))
    g[:, 4:7, :] = 0.1 * rng.normal(size=(table_cnt, 3, 64))
    return g


@pytest.mark.parametrize("kind", [S.STICKY, S.SLIP, S.SEPARATE])
@pytest.mark.parametrize("plane", ["oblique", "node_layer"])
def test_grid_update_colliders_and_max_vel(kind, plane):
    rocm, torch = _rocm(), _torch()
    if plane == "oblique":
        nrm = np.array([1.0, 2.0, -0.5])
        origin = (0.5, 0.5, 0.5)
    else:
        nrm = np.array([0.0, 1.0, 0.0])
        origin = (0.3, 16 / 32, 0.7)
    nrm = (nrm / np.linalg.norm(nrm)).astype(np.float32)
    sc = S.tiny_scene(n_cells=(6, 6, 6), lo=(13, 13, 13), ppc=2)
    sc.colliders = [S.Collider(origin, tuple(float(c) for c in nrm), kind)]
    sc.dt = 1e-3
    print(f"collider kind {kind} plane {plane}: {sc.n} particles", flush=True)
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.sync()
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    B = ctx.sizes()["cnt"]
    assert sim.nblocks() == B
    g = _seeded_grid(B, 31)
    sim.grid.data[:B].copy_(torch.as_tensor(g).cuda())
    mv = torch.full((1,), -1.0, dtype=torch.float32, device="cuda")  # a negative start loses to any |v|^2
    sim.grid_update(fold=1, max_vel=mv)
    sim.sync()
    got = sim.grid.data[:B].cpu().numpy()
    prm = O.make_params(sc)
    ctx.set_grids(grid_p2g=g)
    ref_max = ctx.grid_update(prm)
    _, ref, _ = ctx.grids()
    scale = float(np.max(np.abs(ref[:, 1:4])))
    err = float(np.max(np.abs(got[:, 1:4] - ref[:, 1:4]))) / scale
    assert err < 1e-6, f"node velocities: {err:.3e}"
    assert np.array_equal(got[:, 0], ref[:, 0]), "mass channel untouched"
    gm = float(mv.item())
    assert abs(gm - ref_max) <= 1e-6 * ref_max, (gm, ref_max)
    # a larger caller value survives the merge
    mv.fill_(1e30)
    sim.grid.data[:B].copy_(torch.as_tensor(g).cuda())
    sim.grid_update(fold=1, max_vel=mv)
    sim.sync()
    assert float(mv.item()) == np.float32(1e30)


def test_substep_max_vel_matches_oracle():
    rocm = _rocm()
    sc = S.perturb(S.cube_scene(cells=16), seed=41)
    sim = rocm.Simulator(sc)
    sim.substep()
    got = sim.max_vel_sqr()
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx = O.Context()
    ctx.run(O.make_params(sc), ps, 1)
    _, _, ref = ctx.grids()
    assert abs(got - ref) <= 1e-5 * ref, (got, ref)
    dt = sim.adaptive_dt(0.5)
    assert 0.0 < dt <= sim.evaluated_dt(0.5)
    assert dt <= 0.5 * sc.dx / np.sqrt(ref) * (1 + 1e-6)


def test_grid_angular_momentum_matches_oracle_and_apic():
    rocm = _rocm()
    rng = np.random.default_rng(5)
    sc = S.perturb(S.cube_scene(cells=12), seed=43, c_scale=20.0)
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    B = sim.nblocks()
    g = sim.grid.data[:B].cpu().numpy()
    got = sim.angular_momentum()
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.set_grids(grid_p2g=g)
    ref = ctx.angular_momentum(sc.dx)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), (got, ref)
    # APIC: grid angular momentum after P2G equals the particles'
    x, v, m = sc.x.astype(np.float64), sc.v.astype(np.float64), sc.m.astype(np.float64)
    Cd = sc.C.astype(np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)
    spin = np.stack([Cd[:, 2, 1] - Cd[:, 1, 2], Cd[:, 0, 2] - Cd[:, 2, 0], Cd[:, 1, 0] - Cd[:, 0, 1]], 1)
    exp_L = np.sum(m[:, None] * (np.cross(x, v) + sc.dx ** 2 / 4.0 * spin), axis=0)
    exp_P = np.sum(m[:, None] * v, axis=0)
    assert np.max(np.abs(got[3:] - exp_P)) < 1e-5 * np.max(np.abs(exp_P))
    assert np.max(np.abs(got[:3] - exp_L)) < 1e-5 * np.max(np.abs(exp_L))
    del rng


def _golden_F():
    d = np.load(__file__.replace("test_gpu_pins.py", "golden/refmath.npz"), allow_pickle=False)
    rng = np.random.default_rng(77)
    rand = (np.eye(3).reshape(9) + 0.2 * rng.normal(size=(2000, 9))).astype(np.float32)
    return np.concatenate([d["svd_F"], rand]).astype(np.float32)


def _stress(pol, mat, gap, F, logJp=None):
    rocm, torch = _rocm(), _torch()
    model = rocm.model_desc(mat, svd_fp64_gap=gap)
    Fd = torch.as_tensor(F).cuda().contiguous()
    lj = None if logJp is None else torch.as_tensor(logJp).cuda().contiguous()
    PF = torch.empty_like(Fd)
    before = int(pol.counters()[3])
    rocm.check(pol.L.zsmpm_stress__rocm(pol.h, C.byref(model), Fd.data_ptr(), None if lj is None else lj.data_ptr(),
                                        F.shape[0], PF.data_ptr()), "stress")
    pol.sync()
    used = int(pol.counters()[3]) - before
    return PF.cpu().numpy(), (None if lj is None else lj.cpu().numpy()), used


@pytest.mark.parametrize("kind", ["fixed_corotated", "drucker_prager"])
def test_fp64_svd_path(kind):
    rocm = _rocm()
    pol = rocm.Policy()
    pol.reserve(4096, 1024)
    F = _golden_F()
    n = F.shape[0]
    print(f"FP64 SVD path, {kind}: {n} deformation gradients", flush=True)
    if kind == "fixed_corotated":
        mat = S.Material(kind=S.FIXED_COROTATED, E=5e4, nu=0.4, rho=1e3, volume=1e-6)
        lj = None
    else:
        mat = S.Material.drucker_prager(3.5e5, 0.3, 1e3, volume=1e-6)
        lj = (0.02 * np.random.default_rng(3).normal(size=n)).astype(np.float32)
    mu, lam = O.lame(mat.E, mat.nu)
    L = O.lib()
    F64 = F.astype(np.float64)
    ref = np.zeros_like(F64)
    if kind == "fixed_corotated":
        L.zso_stress_fc_f64(n, float(mat.volume), float(mu), float(lam), F64, ref)
    else:
        ref, lj_ref, _, _ = O.stress_sand64(mat.volume, mu, lam, 0.0, 1.0, mat.yield_surface, 1, lj, F64)
    all64, lj_all, used_all = _stress(pol, mat, 1e30, F, lj)
    assert used_all == n
    # per row, relative to |ref| plus the modulus scale 2 mu vol: a rotation's
    # stress is zero, and there the float32 output carries the rounding of
    # terms of size 2 mu vol (sigma - 1) that cancel
    scale = np.max(np.abs(ref), axis=1) + 2.0 * mu * mat.volume
    err = np.max(np.abs(all64 - ref), axis=1) / scale
    # the constitutive tail after the SVD stays float32 (only the SVD is FP64):
    # for nearly singular F the sand model divides by the clamped singular
    # value (>= 1e-4) and multiplies it back, so its rounding grows with the
    # condition number sigma_max / max(sigma_min, 1e-4)
    sv = np.linalg.svd(F64.reshape(-1, 3, 3), compute_uv=False)
    kappa = sv[:, 0] / np.maximum(sv[:, 2], 1e-4)
    tol = 5e-6 * np.maximum(1.0, kappa)
    bad = err >= tol
    assert not np.any(bad), (f"FP64 path vs FP64 oracle: {int(bad.sum())} rows, worst {err.max():.3e} at "
                             f"{int(err.argmax())} (kappa {kappa[int(err.argmax())]:.3e})")
    if lj is not None:
        assert np.max(np.abs(lj_all - lj_ref)) < 1e-6
    plain, _, used0 = _stress(pol, mat, 0.0, F, lj)
    assert used0 == 0
    gated, _, used = _stress(pol, mat, 1e-6, F, lj)
    as64 = np.all(gated == all64, axis=1)
    as32 = np.all(gated == plain, axis=1)
    assert np.all(as64 | as32), "every particle takes exactly one of the two paths"
    only64 = as64 & ~as32
    assert only64.sum() <= used <= as64.sum(), (only64.sum(), used, as64.sum())
    assert used > 0, "the repeated-singular-value inputs reach the FP64 path at gap 1e-6"
