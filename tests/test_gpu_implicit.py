"""The implicit step's operators on the GPU (implicit.hip and the projection in
transfer.hip) through the C ABI against the oracle (zsmpm_oracle.hpp):
G2P2G forces for random and affine grid velocity fields, the system's
multiply / precondition / project, and the preconditioned conjugate gradient
(iterations and solution), for the fixed-corotated cube and the sand column.
Float atomics and double dot products make the GPU's sums differ from the
oracle's sequential ones in the last bits, and a node's force is a sum of
~200 contributions that partly cancel (the P2G force channels show the same
2.5e-4 spread between the float32 and float64 oracles): forces are held to
2e-4 of their largest entry, the system's products to 1e-4, solutions to 1e-3.
"""
import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu


def _rocm():
    from zpc_amd import rocm

    return rocm


def _scene(kind):
    if kind == "sand":
        return S.perturb(S.sand_scene(dims=(8, 16, 8)), seed=2, f_scale=0.02, c_scale=1.0, v_scale=0.2)
    return S.perturb(S.cube_scene(cells=12), seed=2, f_scale=0.02, c_scale=1.0, v_scale=0.2)


def _both(sc):
    rocm = _rocm()
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    prm = O.make_params(sc)
    ctx.p2g(prm, ps)  # sand's P2G writes logJp back, on the GPU as here
    grid, _, _ = ctx.grids()
    assert sim.nblocks() == grid.shape[0]
    return sim, ctx, prm, grid, ps


def _copy(ps):
    return O.HostParticles(ps.x, ps.v, ps.C, ps.F, ps.m, ps.logJp)


def _dev(a):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).cuda()


def _vec(sim, host, n):
    v = sim.dof_vector()
    v[:n].copy_(_dev(host))
    return v


def _close(got, want, tol, what):
    scale = max(float(np.max(np.abs(want))), 1e-30)
    err = float(np.max(np.abs(got.astype(np.float64) - want)))
    print(f"{what}: {err / scale:.3e}", flush=True)
    assert err / scale < tol, f"{what}: {err / scale:.3e}"


@pytest.mark.parametrize("kind", ["cube", "sand"])
def test_g2p2g_matches_oracle(kind):
    sc = _scene(kind)
    sim, ctx, prm, grid, ps1 = _both(sc)
    n = grid.shape[0] * 64 * 3
    rng = np.random.default_rng(3)
    for field in ("random", "zero"):
        v = rng.normal(size=n).astype(np.float32) * 0.3 if field == "random" else np.zeros(n, np.float32)
        want = ctx.g2p2g(prm, _copy(ps1), v)
        r = sim.dof_vector()
        sim.g2p2g(_vec(sim, v, n), r)
        sim.sync()
        _close(r[:n].cpu().numpy(), want, 2e-4, f"{kind} g2p2g {field}")


@pytest.mark.parametrize("kind", ["cube", "sand"])
def test_multiply_precondition_project_match_oracle(kind):
    sc = _scene(kind)
    sim, ctx, prm, grid, ps1 = _both(sc)
    n = grid.shape[0] * 64 * 3
    rng = np.random.default_rng(4)
    vin = rng.normal(size=n).astype(np.float32)
    want = ctx.implicit_multiply(prm, _copy(ps1), vin)
    out = sim.dof_vector()
    din = _vec(sim, vin, n)
    sim.implicit_multiply(din, out)
    sim.sync()
    _close(out[:n].cpu().numpy(), want, 1e-4, f"{kind} multiply")
    pre = sim.dof_vector(7.0)
    sim.implicit_precondition(din, pre)
    sim.sync()
    # in / m with the GPU's own P2G masses (they differ from the oracle's in the
    # last bits, which matters at the stencil fringe where masses are tiny)
    mg = np.repeat(sim.grid.data[:grid.shape[0], 0, :].cpu().numpy().reshape(-1), 3)
    wpre = np.where(mg > 0, vin / np.where(mg > 0, mg, 1.0).astype(np.float32), np.float32(7.0))
    assert np.allclose(pre[:n].cpu().numpy(), wpre, rtol=1e-6, atol=0), f"{kind} precondition"
    assert np.array_equal(mg > 0, np.repeat(grid[:, 0, :].reshape(-1), 3) > 0), f"{kind} nodes with mass"
    proj = _vec(sim, vin, n)
    sim.implicit_project(proj)
    sim.sync()
    assert np.array_equal(proj[:n].cpu().numpy(), ctx.implicit_project(prm, vin)), f"{kind} project"


@pytest.mark.parametrize("kind,dt_scale", [("cube", 1), ("sand", 1), ("cube", 100), ("sand", 30)])
def test_conjugate_gradient_matches_oracle(kind, dt_scale):
    """dt_scale > 1: the force term f dt^2 weighs against the mass and the
    solve takes two or three iterations to its tolerance. (At larger steps the
    system, (f(v) dt^2 + m) v per entry, is far from symmetric positive
    definite and the iteration stagnates or diverges, in the oracle as here.)"""
    sc = _scene(kind)
    sc.dt *= dt_scale
    sim, ctx, prm, grid, ps1 = _both(sc)
    n = grid.shape[0] * 64 * 3
    m = np.repeat(grid[:, 0, :].reshape(-1), 3)
    has = m > 0
    v = np.where(has, grid[:, 1:4, :].transpose(0, 2, 1).reshape(-1) / np.where(has, m, 1.0), 0.0)
    b = (m * v).astype(np.float32)
    for max_iters in (1, 3, 8):
        xw, itw, resw = ctx.implicit_solve(prm, _copy(ps1), np.zeros(n, np.float32), b, max_iters=max_iters)
        x = sim.dof_vector()
        it, res = sim.implicit_solve(x, _vec(sim, b, n), max_iters=max_iters)
        sim.sync()
        print(f"{kind}, at most {max_iters}: gpu {it} iterations, residual {res:.4e}; "
              f"oracle {itw}, {resw:.4e}", flush=True)
        assert it == itw
        _close(x[:n].cpu().numpy(), xw, 1e-3, f"{kind} x after {it} iterations")
        # the residual is a difference of near-equal vectors: held relative to
        # the starting one, sqrt(b . M b)
        r0 = float(np.sqrt(np.sum(np.where(has, b.astype(np.float64) ** 2 / np.where(has, m, 1.0), 0.0))))
        assert abs(res - resw) <= 1e-4 * r0
