"""The oracle's implicit step (G2P2GTransfer, ImplicitMPMSystem, the
conjugate gradient; zsmpm_oracle.hpp) pinned against what fixes it:

- G2P2G at a zero velocity field is the P2G force (already pinned by the
  energy gradient) divided by -dt: the same stress, weights and arena;
- G2P2G of an affine field v = A x + b is the P2G force at F = (I + dt A) F_old
  (C recovers A exactly for the quadratic kernel);
- multiply: (f dt^2 + m) v on nodes with mass, the force elsewhere;
  precondition: v / m on nodes with mass, the rest untouched; project: sticky
  walls zero, massless nodes cleared, nothing without boundaries;
- the solve: without stiffness the system is diag(m) and one preconditioned
  iteration gives b / m; with stiffness the true residual of the returned x is
  below the tolerance the solver stopped at (scaled), and iterations stop at
  the relative tolerance.
"""
import numpy as np

from oracle import oracle as O
from zpc_amd import scenes as S


def _setup(sc):
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    prm = O.make_params(sc)
    ctx.p2g(prm, ps)
    grid, _, _ = ctx.grids()
    return ctx, prm, grid


def _node_positions(ctx, dx):
    act = ctx.table()["active"].astype(np.int64)
    cell = np.arange(64)
    off = np.stack([(cell >> 4) & 3, (cell >> 2) & 3, cell & 3], -1)
    return ((act[:, None, :] * 4 + off[None]).astype(np.float32) * np.float32(dx)).reshape(-1, 3)


def _scene(kind):
    if kind == "sand":
        return S.perturb(S.sand_scene(dims=(8, 8, 8)), seed=2, f_scale=0.02, c_scale=1.0, v_scale=0.2)
    return S.perturb(S.cube_scene(cells=8), seed=2, f_scale=0.02, c_scale=1.0, v_scale=0.2)


def test_g2p2g_at_rest_is_p2g_force_over_minus_dt():
    for kind in ("cube", "sand"):
        sc = _scene(kind)
        ctx, prm, grid = _setup(sc)
        ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
        r = ctx.g2p2g(prm, ps, np.zeros(grid.shape[0] * 64 * 3, np.float32))
        f = (grid[:, 4:7, :].transpose(0, 2, 1).reshape(-1) / -np.float32(sc.dt)).astype(np.float64)
        scale = np.max(np.abs(f))
        assert scale > 0
        assert np.max(np.abs(r - f)) / scale < 1e-5, kind


def test_g2p2g_affine_field_is_force_at_updated_F():
    sc = _scene("cube")
    ctx, prm, grid = _setup(sc)
    A = np.array([[0.3, -1.1, 0.4], [0.9, 0.2, -0.5], [-0.7, 0.6, 0.1]], np.float64)
    b = np.array([0.2, -0.3, 0.05])
    xs = _node_positions(ctx, sc.dx).astype(np.float64)
    v = (xs @ A.T + b).astype(np.float32).reshape(-1)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    r = ctx.g2p2g(prm, ps, v)
    # the P2G force with F' = (I + dt A) F (column-major rows: F[:, 3c + r] = F(r, c))
    Fm = sc.F.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)
    F2 = (np.eye(3) + sc.dt * A)[None] @ Fm
    sc2 = S.perturb(_scene("cube"), seed=0, f_scale=0.0, c_scale=0.0, v_scale=0.0)
    sc2.x, sc2.v, sc2.C, sc2.m = sc.x, sc.v, sc.C, sc.m
    sc2.F = F2.transpose(0, 2, 1).reshape(-1, 9).astype(np.float32)
    _, _, grid2 = _setup(sc2)
    f = grid2[:, 4:7, :].transpose(0, 2, 1).reshape(-1).astype(np.float64) / -sc.dt
    assert np.max(np.abs(r - f)) / np.max(np.abs(f)) < 1e-4
    # and a different field gives a different force (the test sees C)
    r0 = ctx.g2p2g(prm, ps, np.zeros_like(v))
    assert np.max(np.abs(r - r0)) / np.max(np.abs(r0)) > 1e-3


def test_multiply_precondition_project_closed_forms():
    sc = _scene("sand")
    ctx, prm, grid = _setup(sc)
    n = grid.shape[0] * 64 * 3
    rng = np.random.default_rng(0)
    vin = rng.normal(size=n).astype(np.float32)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    f = ctx.g2p2g(prm, ps, vin)
    out = ctx.implicit_multiply(prm, ps, vin)
    m = np.repeat(grid[:, 0, :].reshape(-1), 3)
    has = m > 0
    want = np.where(has, (f * np.float32(sc.dt) * np.float32(sc.dt) + m) * vin, f)
    assert np.allclose(out, want, rtol=1e-6, atol=1e-12)
    pre = ctx.implicit_precondition(vin, np.full(n, 7.0, np.float32))
    assert np.allclose(pre[has], vin[has] / m[has], rtol=1e-6) and np.all(pre[~has] == 7.0)
    # project: the scene's sticky walls zero their nodes, massless nodes are cleared
    proj = ctx.implicit_project(prm, vin)
    assert np.all(proj[~has] == 0.0)
    xs = np.repeat(_node_positions(ctx, sc.dx), 1, axis=0)
    inside_wall = np.zeros(len(xs), bool)
    for c in sc.colliders:
        inside_wall |= (xs - np.asarray(c.origin, np.float32)) @ np.asarray(c.normal, np.float32) < 0
    wall = np.repeat(inside_wall, 3) & has
    assert np.all(proj[wall] == 0.0)
    keep = has & ~np.repeat(inside_wall, 3)
    assert np.array_equal(proj[keep], vin[keep])
    sc_free = _scene("sand")
    sc_free.colliders = []
    prm_free = O.make_params(sc_free)
    assert np.array_equal(ctx.implicit_project(prm_free, vin), vin)


def test_solve_without_stiffness_is_the_mass_inverse():
    sc = _scene("cube")
    sc.material.E = 0.0
    sc.colliders = []
    ctx, prm, grid = _setup(sc)
    n = grid.shape[0] * 64 * 3
    rng = np.random.default_rng(1)
    b = rng.normal(size=n).astype(np.float32)
    x0 = rng.normal(size=n).astype(np.float32) * 0.1
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    x, it, res = ctx.implicit_solve(prm, ps, x0, b)
    m = np.repeat(grid[:, 0, :].reshape(-1), 3)
    has = m > 0
    # one iteration solves it up to rounding; the absolute tolerance (1e-6 on
    # sqrt(r.M r), M = 1/m) may ask for one or two more
    assert 1 <= it <= 3
    assert np.allclose(x[has], b[has] / m[has], rtol=1e-4)
    assert np.array_equal(x[~has], x0[~has])


def test_solve_reduces_the_true_residual():
    sc = _scene("cube")
    ctx, prm, grid = _setup(sc)
    n = grid.shape[0] * 64 * 3
    m = np.repeat(grid[:, 0, :].reshape(-1), 3)
    has = m > 0
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    # b = m v for the grid velocity the explicit update would give
    v = np.where(has, grid[:, 1:4, :].transpose(0, 2, 1).reshape(-1) / np.where(has, m, 1.0), 0.0).astype(np.float32)
    b = (m * v).astype(np.float32)
    x0 = np.zeros(n, np.float32)

    def true_res(x):
        r = ctx.implicit_project(prm, b - ctx.implicit_multiply(prm, ps, x))
        return float(np.sqrt(np.sum(np.where(has, r * r / np.where(has, m, 1.0), 0.0))))

    r0 = true_res(x0)
    x, it, res = ctx.implicit_solve(prm, ps, x0, b, max_iters=50, tol=1e-6, rel_tol=0.5)
    assert it >= 1 and res <= 0.5 * r0 * 1.0001
    assert true_res(x) < 0.6 * r0
