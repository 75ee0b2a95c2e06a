"""CPU checks that pin the oracle's transfer arithmetic (P2G, grid update,
boundary, G2P) against what the mathematics fixes, so that a plausible
mistake -- a dropped term, a wrong sign, scale or axis, a transposed operand
-- fails at least one of them:

- the P2G force (rhs channels) is -dt times the gradient of the discrete
  elastic energy sum_p V psi(F_p(x_hat)) with respect to the node positions,
  by central finite differences in float64 (sign, dt, D^-1 = 4/dx^2, the
  volume and P F^T versus P^T F);
- P2G transfers APIC angular momentum exactly: grid sum x_i x (m v)_i equals
  sum_p m_p (x_p x v_p + dx^2/4 eps:C_p^T) (C's orientation in P2G);
- G2P of an affine grid field v(x) = A x + b returns v = A x_p + b, C = A,
  x + dt v and F <- (I + dt A) F for a non-symmetric F (the update order);
- free fall: v0 + g dt on the y axis, C = 0, x + dt v, maxVel = |v|^2;
- plane colliders: sticky zeroes, slip removes the normal component,
  separate removes it only when it points into the collider; nodes outside
  or on the plane, and massless nodes, are untouched.

References: simulation/transfer/P2G.hpp:45-127, G2P.hpp:43-84,
simulation/grid/GridOp.hpp:71-164,216-259, geometry/Collider.h:80-107.
"""
import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S


def _params(dx, dt, mat, gravity=(0.0, 0.0, 0.0), colliders=(), table_size=1 << 12):
    sc = S.Scene("pin", dx, dt, mat, np.zeros((1, 3), np.float32), np.zeros((1, 3), np.float32),
                 np.zeros((1, 9), np.float32), np.zeros((1, 9), np.float32), np.zeros(1, np.float32), None,
                 gravity=gravity, colliders=list(colliders), table_size=table_size)
    return O.make_params(sc)


def _cm(a):
    """column-major rows of 9 -> (n, 3, 3) matrices"""
    return np.asarray(a, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)


def _to_cm(M):
    return np.ascontiguousarray(np.asarray(M).transpose(0, 2, 1)).reshape(-1, 9)


def _weights(x, dx):
    """Quadratic B-spline stencil in float64: base node, 27 weights and the
    27 offsets x_i - x_p (simulation/Utils.hpp:51-70 restated from the
    spline's definition)."""
    X = np.asarray(x, np.float64) / dx
    base = np.floor(X - 0.5).astype(np.int64)
    d0 = X - base
    w1d = np.stack([0.5 * (1.5 - d0) ** 2, 0.75 - (d0 - 1.0) ** 2, 0.5 * (d0 - 0.5) ** 2], axis=-1)  # (n, 3, 3)
    nodes, W, diff = [], [], []
    for i in range(3):
        for j in range(3):
            for k in range(3):
                off = np.array([i, j, k])
                nodes.append(base + off)
                W.append(w1d[:, 0, i] * w1d[:, 1, j] * w1d[:, 2, k])
                diff.append((base + off) * dx - np.asarray(x, np.float64))
    return np.stack(nodes, 1), np.stack(W, 1), np.stack(diff, 1)


def _node_positions(table, dx):
    """(nblocks, 64, 3) world positions of every node of the active blocks."""
    cells = np.arange(64)
    cc = np.stack([(cells >> 4) & 3, (cells >> 2) & 3, cells & 3], 1)
    keys = table["active"].astype(np.float64)
    return (keys[:, None, :] * 4 + cc[None]) * dx


def _node_index(table):
    """{(i, j, k) node coordinate: (block, cell)}"""
    out = {}
    for b, key in enumerate(table["active"]):
        for c in range(64):
            co = (int(key[0]) * 4 + ((c >> 4) & 3), int(key[1]) * 4 + ((c >> 2) & 3), int(key[2]) * 4 + (c & 3))
            out[co] = (b, c)
    return out


# ------------------------------------------------------------------ energies (float64, from the definitions)

def _psi_fc(F, mu, lam):
    """fixed-corotated energy mu |F - R|^2 + lam/2 (J - 1)^2, R the polar rotation"""
    U, s, Vt = np.linalg.svd(F)
    R = U @ Vt
    J = np.linalg.det(F)
    return mu * np.sum((F - R) ** 2, axis=(1, 2)) + 0.5 * lam * (J - 1.0) ** 2


def _psi_hencky(F, mu, lam):
    """Hencky energy mu |log S|^2 + lam/2 tr(log S)^2 (the elastic part of the
    Drucker-Prager model, cohesion 0)"""
    s = np.linalg.svd(F, compute_uv=False)
    e = np.log(s)
    return mu * np.sum(e * e, axis=1) + 0.5 * lam * np.sum(e, axis=1) ** 2


def _fd_force(x, F, dx, vol, psi):
    """dE/du_i for every node i of the particles' stencils, E(u) = sum_p vol
    psi((I + sum_i u_i g_ip^T) F_p), g_ip = (4/dx^2) w_ip (x_i - x_p) (the MLS
    gradient), by central differences."""
    nodes, W, diff = _weights(x, dx)
    g = (4.0 / dx ** 2) * W[..., None] * diff  # (n, 27, 3)
    Fm = _cm(F)
    n = Fm.shape[0]
    out = {}
    h = 1e-6 * dx
    for p in range(n):
        for e in range(27):
            key = tuple(int(v) for v in nodes[p, e])
            if key in out:
                continue
            # every particle whose stencil holds this node
            grad = np.zeros(3)
            for q in range(n):
                hit = np.where(np.all(nodes[q] == nodes[p, e], axis=1))[0]
                if hit.size == 0:
                    continue
                gq = g[q, hit[0]]
                for a in range(3):
                    ea = np.zeros(3)
                    ea[a] = 1.0
                    Fp = (np.eye(3) + h * np.outer(ea, gq)) @ Fm[q]
                    Fn = (np.eye(3) - h * np.outer(ea, gq)) @ Fm[q]
                    grad[a] += vol * (psi(Fp[None])[0] - psi(Fn[None])[0]) / (2 * h)
            out[key] = grad
    return out


def _run_p2g(x, F, dx, dt, mat, v=None, Cm=None, m=None, logJp=None):
    n = x.shape[0]
    v = np.zeros((n, 3), np.float32) if v is None else v
    Cm = np.zeros((n, 9), np.float32) if Cm is None else Cm
    m = np.ones(n, np.float32) if m is None else m
    ps = O.HostParticles(x, v, Cm, F, m, logJp)
    prm = _params(dx, dt, mat)
    ctx = O.Context()
    ctx.activate(prm.table_size, x, dx)
    ctx.enlarge()
    ctx.p2g(prm, ps)
    g, _, _ = ctx.grids()
    return ctx, g, ps


@pytest.mark.parametrize("model", ["fixed_corotated", "drucker_prager_elastic"])
def test_p2g_force_is_energy_gradient(model):
    rng = np.random.default_rng(101)
    dx = 1.0 / 64
    dt = 1e-3
    n = 5
    x = (0.5 + 3 * dx * rng.uniform(-1, 1, size=(n, 3))).astype(np.float32)
    if model == "fixed_corotated":
        mat = S.Material(kind=S.FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=2e-6)
        Fm = np.eye(3) + 0.15 * rng.normal(size=(n, 3, 3))
        logJp = None
    else:
        # compression inside the cone: the return mapping leaves F elastic
        mat = S.Material.drucker_prager(3.5e5, 0.3, 1e3, volume=2e-6)
        Q1 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
        Q2 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
        Q1 *= np.sign(np.linalg.det(Q1))[:, None, None]
        Q2 *= np.sign(np.linalg.det(Q2))[:, None, None]
        sig = rng.uniform(0.96, 0.985, size=(n, 3))
        Fm = Q1 @ (sig[:, :, None] * np.eye(3)) @ np.transpose(Q2, (0, 2, 1))
        logJp = np.zeros(n, np.float32)
    F = _to_cm(Fm).astype(np.float32)
    mu, lam = O.lame(mat.E, mat.nu)
    if model != "fixed_corotated":
        _, _, _, which = O.stress_sand(mat.volume, mu, lam, 0.0, 1.0, mat.yield_surface, 1, logJp, F)
        assert np.all(which == 1), which  # every particle in the elastic case
    ctx, g, _ = _run_p2g(x, F, dx, dt, mat, logJp=logJp)
    table = ctx.table()
    idx = _node_index(table)
    psi = (lambda A: _psi_fc(A, mu, lam)) if model == "fixed_corotated" else (lambda A: _psi_hencky(A, mu, lam))
    grads = _fd_force(x, F, dx, mat.volume, psi)
    exp = np.array([-dt * grads[k] for k in grads])
    got = np.array([[g[idx[k][0], 4 + a, idx[k][1]] for a in range(3)] for k in grads], np.float64)
    scale = np.max(np.abs(exp))
    assert scale > 0
    err = np.max(np.abs(got - exp)) / scale
    assert err < 2e-4, f"rhs vs -dt dE/dx: {err:.3e}"
    # nodes outside every stencil carry no force
    touched = np.zeros(g.shape[0] * 64, bool)
    for k in grads:
        touched[idx[k][0] * 64 + idx[k][1]] = True
    rest = g[:, 4:7, :].transpose(0, 2, 1).reshape(-1, 3)[~touched]
    assert np.all(rest == 0.0)


def test_p2g_force_fails_on_planted_mistakes():
    """The energy-gradient check distinguishes the correct force from the
    mistakes it is meant to catch: a flipped sign, a missing dt or D^-1 and
    P^T F in place of P F^T all miss it by far more than the tolerance."""
    rng = np.random.default_rng(7)
    dx = 1.0 / 64
    dt = 1e-3
    n = 3
    x = (0.5 + 2 * dx * rng.uniform(-1, 1, size=(n, 3))).astype(np.float32)
    mat = S.Material(kind=S.FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=2e-6)
    Fm = np.eye(3) + 0.2 * rng.normal(size=(n, 3, 3))
    mu, lam = O.lame(mat.E, mat.nu)
    grads = _fd_force(x, _to_cm(Fm), dx, mat.volume, lambda A: _psi_fc(A, mu, lam))
    exp = np.array([-dt * grads[k] for k in grads])
    # P from the oracle's pinned stress (vol P F^T) in two forms
    PF = _cm(O.stress_fc(mat.volume, mu, lam, _to_cm(Fm).astype(np.float32)))  # vol P F^T
    P = PF @ np.linalg.inv(np.transpose(Fm, (0, 2, 1)))                       # vol P
    nodes, W, diff = _weights(x, dx)

    def force(tau, scale):
        out = {}
        for p in range(n):
            for e in range(27):
                key = tuple(int(v) for v in nodes[p, e])
                out[key] = out.get(key, 0.0) + scale * W[p, e] * (tau[p] @ diff[p, e])
        return np.array([out[k] for k in grads])

    D = 4.0 / dx ** 2
    right = force(PF, -dt * D)
    assert np.max(np.abs(right - exp)) / np.max(np.abs(exp)) < 2e-4
    for wrong in (force(PF, dt * D), force(PF, -D), force(PF, -dt), force(np.transpose(P, (0, 2, 1)) @ Fm, -dt * D)):
        assert np.max(np.abs(wrong - exp)) / np.max(np.abs(exp)) > 1e-2


def test_p2g_transfers_apic_angular_momentum():
    """Grid sum x_i x (m v)_i after P2G equals sum_p m_p (x_p x v_p + dx^2/4
    eps:C_p^T) exactly in arithmetic (Jiang et al. 2015, APIC); the oracle's
    GridAngularMomentum (GridOp.hpp:216-259) reads it off the grid. With
    C_p^T in place of C_p the C term changes sign."""
    rng = np.random.default_rng(5)
    dx = 1.0 / 128
    n = 200
    x = (0.5 + 6 * dx * rng.uniform(-1, 1, size=(n, 3))).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    Cm = (20.0 * rng.normal(size=(n, 9))).astype(np.float32)
    m = rng.uniform(0.5, 2.0, size=n).astype(np.float32)
    F = np.tile(np.eye(3).reshape(1, 9), (n, 1)).astype(np.float32)
    mat = S.Material(kind=S.FIXED_COROTATED, E=0.0, nu=0.3, rho=1e3, volume=1e-6)
    ctx, g, _ = _run_p2g(x, F, dx, 1e-4, mat, v=v, Cm=Cm, m=m)
    L = ctx.angular_momentum(dx)
    xd, vd, md = x.astype(np.float64), v.astype(np.float64), m.astype(np.float64)
    Cd = _cm(Cm)  # C[d][e] = C[d + 3e]
    spin = np.stack([Cd[:, 2, 1] - Cd[:, 1, 2], Cd[:, 0, 2] - Cd[:, 2, 0], Cd[:, 1, 0] - Cd[:, 0, 1]], 1)
    exp_L = np.sum(md[:, None] * (np.cross(xd, vd) + dx * dx / 4.0 * spin), axis=0)
    exp_P = np.sum(md[:, None] * vd, axis=0)
    assert np.max(np.abs(L[3:] - exp_P)) < 1e-5 * np.max(np.abs(exp_P))
    assert np.max(np.abs(L[:3] - exp_L)) < 1e-5 * np.max(np.abs(exp_L))
    # the spin term is not negligible here, so a transposed C would show
    assert np.max(np.abs(np.sum(md[:, None] * dx * dx / 4.0 * spin, 0))) > 1e-3 * np.max(np.abs(exp_L))
    # mass and the whole P2G grid: only stencil nodes of mass
    assert abs(float(np.sum(g[:, 0, :], dtype=np.float64)) - float(np.sum(md))) < 1e-5 * float(np.sum(md))


def test_g2p_affine_field_known_answer():
    """G2P of v(x) = A (x - xc) + b with a non-symmetric A and F: quadratic
    B-splines reproduce affine fields, so v_p = A (x_p - xc) + b and C_p = A;
    x <- x + dt v and F <- (I + dt A) F (G2P.hpp:65-78). F (I + dt A), x + v
    dt^2 or C^T would be off by orders of magnitude more than the
    tolerance."""
    rng = np.random.default_rng(9)
    dx = 1.0 / 64
    dt = 1e-2
    n = 64
    xc = np.array([0.5, 0.45, 0.55])
    x = (xc + 5 * dx * rng.uniform(-1, 1, size=(n, 3))).astype(np.float32)
    A = 8.0 * rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    Fm = np.eye(3) + 0.3 * rng.normal(size=(n, 3, 3))
    F = _to_cm(Fm).astype(np.float32)
    v0 = np.zeros((n, 3), np.float32)
    C0 = np.zeros((n, 9), np.float32)
    mat = S.Material(kind=S.FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=1e-6)
    prm = _params(dx, dt, mat)
    ctx = O.Context()
    ctx.activate(prm.table_size, x, dx)
    ctx.enlarge()
    table = ctx.table()
    pos = _node_positions(table, dx)  # (B, 64, 3)
    vel = (pos - xc) @ A.T + b
    gv = np.zeros((table["cnt"], 7, 64), np.float32)
    gv[:, 0, :] = 1.0
    gv[:, 1:4, :] = vel.transpose(0, 2, 1)
    ctx.set_grids(grid_v=gv)
    ps = O.HostParticles(x, v0, C0, F, np.ones(n, np.float32), None)
    ctx.g2p(prm, ps)
    xd = x.astype(np.float64)
    v_exp = (xd - xc) @ A.T + b
    v_got = ps.v.reshape(n, 3)
    assert np.max(np.abs(v_got - v_exp)) < 2e-5 * np.max(np.abs(v_exp))
    C_got = _cm(ps.C)
    assert np.max(np.abs(C_got - A[None])) < 2e-4 * np.max(np.abs(A))
    x_exp = xd + dt * v_exp
    assert np.max(np.abs(ps.x.reshape(n, 3) - x_exp)) < 1e-6
    F_exp = (np.eye(3) + dt * A)[None] @ _cm(F)
    F_got = _cm(ps.F)
    assert np.max(np.abs(F_got - F_exp)) < 2e-5 * np.max(np.abs(F_exp))
    # the wrong update order differs by far more than that
    F_wrong = _cm(F) @ (np.eye(3) + dt * A)[None]
    assert np.max(np.abs(F_wrong - F_exp)) > 1e-2 * np.max(np.abs(F_exp))


def test_free_fall_known_answer():
    """Unstressed particles (F = I) with v0 under gravity (0, -9.8, 0): one
    substep gives v = v0 + g dt on every particle, C = 0, x + dt v, F = I and
    maxVel = |v0 + g dt|^2 (GridOp.hpp:71-110)."""
    rng = np.random.default_rng(3)
    dx = 1.0 / 64
    dt = 2e-3
    n = 40
    x = (0.5 + 4 * dx * rng.uniform(-1, 1, size=(n, 3))).astype(np.float32)
    v0 = np.array([0.3, -0.2, 0.1])
    v = np.tile(v0, (n, 1)).astype(np.float32)
    F = np.tile(np.eye(3).reshape(1, 9), (n, 1)).astype(np.float32)
    mat = S.Material(kind=S.FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=dx ** 3 / 8)
    sc = S.Scene("fall", dx, dt, mat, x, v, np.zeros((n, 9), np.float32), F, np.full(n, 0.5, np.float32), None,
                 gravity=(0.0, -9.8, 0.0), colliders=[], table_size=1 << 12)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, None)
    ctx = O.Context()
    ctx.run(O.make_params(sc), ps, 1)
    v1 = v0 + np.array([0.0, -9.8, 0.0]) * dt
    assert np.max(np.abs(ps.v.reshape(n, 3) - v1)) < 1e-6 * np.linalg.norm(v1)
    assert np.max(np.abs(ps.C)) < 1e-3 * np.linalg.norm(v1) / dx
    assert np.max(np.abs(ps.x.reshape(n, 3) - (x.astype(np.float64) + dt * v1))) < 1e-6
    assert np.max(np.abs(ps.F.reshape(n, 9) - np.eye(3).reshape(9))) < 1e-5
    _, _, maxvel = ctx.grids()
    assert abs(maxvel - float(v1 @ v1)) < 1e-6 * float(v1 @ v1)


def test_max_vel_is_max_over_massive_nodes():
    """maxVel is the largest |v|^2 among the updated nodes that carry mass (no
    colliders, so the boundary leaves them as the update wrote them)."""
    sc = S.perturb(S.tiny_scene(n_cells=(4, 3, 5)), seed=21)
    sc.colliders = []
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx = O.Context()
    ctx.run(O.make_params(sc), ps, 1)
    g, gv, maxvel = ctx.grids()
    has = g[:, 0, :] != 0
    v2 = np.sum(gv[:, 1:4, :].astype(np.float64) ** 2, axis=1)
    assert abs(maxvel - float(np.max(v2[has]))) <= 1e-6 * maxvel


def _collider_case(kind, normal, origin, seed=31):
    rng = np.random.default_rng(seed)
    dx = 1.0 / 32
    n = 60
    x = (0.5 + 3 * dx * rng.uniform(-1, 1, size=(n, 3))).astype(np.float32)
    mat = S.Material(kind=S.FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=1e-6)
    nrm = np.asarray(normal, np.float64)
    nrm = (nrm / np.linalg.norm(nrm)).astype(np.float32)
    col = S.Collider(tuple(float(o) for o in origin), tuple(float(c) for c in nrm), kind)
    prm = _params(dx, 1e-3, mat, gravity=(0.0, -9.8, 0.0), colliders=[col])
    ctx = O.Context()
    ctx.activate(prm.table_size, x, dx)
    ctx.enlarge()
    table = ctx.table()
    B = table["cnt"]
    g = np.zeros((B, 7, 64), np.float32)
    g[:, 0, :] = rng.uniform(0.5, 2.0, size=(B, 64))
    g[:, 0, ::7] = 0.0  # some massless nodes
    g[:, 1:4, :] = rng.normal(size=(B, 3, 64))
    g[:, 4:7, :] = 0.1 * rng.normal(size=(B, 3, 64))
    ctx.set_grids(grid_p2g=g)
    ctx.grid_update(prm)
    _, gv, _ = ctx.grids()
    pos = _node_positions(table, dx).reshape(-1, 3)
    phi = (pos - np.asarray(origin, np.float64)) @ nrm.astype(np.float64)
    m = g[:, 0, :].reshape(-1)
    mv = (g[:, 1:4, :] + g[:, 4:7, :]).astype(np.float64).transpose(0, 2, 1).reshape(-1, 3)
    vin = mv / np.where(m[:, None] != 0, m[:, None], 1.0) + np.array([0.0, -9.8, 0.0]) * 1e-3
    vout = gv[:, 1:4, :].transpose(0, 2, 1).reshape(-1, 3).astype(np.float64)
    return phi, m, mv, vin, vout, nrm.astype(np.float64)


@pytest.mark.parametrize("kind", [S.STICKY, S.SLIP, S.SEPARATE])
def test_plane_collider_projection(kind):
    # oblique plane through the middle; and an axis plane through a node layer
    for normal, origin in (((1.0, 2.0, -0.5), (0.5, 0.5, 0.5)), ((0.0, 1.0, 0.0), (0.3, 16 / 32, 0.7))):
        phi, m, mv, vin, vout, nrm = _collider_case(kind, normal, origin)
        massless = m == 0
        assert np.array_equal(vout[massless], mv[massless].astype(np.float32)), "massless nodes untouched"
        live = ~massless
        inside = live & (phi < -1e-7)
        outside = live & (phi > 1e-7)
        assert inside.sum() > 50 and outside.sum() > 50
        tol = 2e-6 * (1.0 + np.abs(vin))
        assert np.all(np.abs(vout[outside] - vin[outside]) <= tol[outside]), "nodes outside keep v"
        vn_in = vin @ nrm
        vn_out = vout @ nrm
        vt_in = vin - vn_in[:, None] * nrm
        vt_out = vout - vn_out[:, None] * nrm
        if kind == S.STICKY:
            assert np.all(vout[inside] == 0.0)
        elif kind == S.SLIP:
            assert np.all(np.abs(vn_out[inside]) < 1e-5)
            assert np.all(np.abs(vt_out[inside] - vt_in[inside]) < 1e-5)
        else:
            into = inside & (vn_in < 0)
            away = inside & (vn_in > 0)
            assert into.sum() > 10 and away.sum() > 10
            assert np.all(np.abs(vn_out[into]) < 1e-5)
            assert np.all(np.abs(vt_out[into] - vt_in[into]) < 1e-5)
            assert np.all(np.abs(vout[away] - vin[away]) <= tol[away])
        if normal == (0.0, 1.0, 0.0):
            on = live & (np.abs(phi) < 1e-9)
            assert on.sum() > 0
            assert np.all(np.abs(vout[on] - vin[on]) <= tol[on]), "a node on the plane is not inside"


def test_cfl_dt_golden():
    """evaluate_timestep_linear_elasticity (physics/SoundSpeedCfl.hpp:6-15),
    bit-exact against zpc's header on the configs' materials and others."""
    d = np.load(__file__.replace("test_oracle_transfer.py", "golden/refmath.npz"), allow_pickle=False)
    for row, dt in zip(d["cfl_in"], d["cfl_dt"]):
        mat = S.Material(E=float(row[0]), nu=float(row[1]), rho=float(row[2]))
        assert np.float32(S.cfl_dt(mat, float(row[3]), float(row[4]))) == dt
