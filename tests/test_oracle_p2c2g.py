"""The oracle's particle -> cell -> grid transfers (P2C2GTransfer,
G2C2PTransfer with its pre and post passes; zsmpm_oracle.hpp p2c2g / g2c2p)
pinned against what fixes them, not against a retyped formula:

- the oracle gathers per cell (tent weights to cell centres, then 1/8 to the
  cell's corners, absolute positions); the checks scatter per particle with
  node weights taken as the convolution of the two, and with D, the kernel's
  second-moment matrix sum_i W (x_i - x_p)(x_i - x_p)^T, computed from those
  weights numerically (the MLS-MPM form: momentum m v + (m B - dt vol P F^T)
  D^-1 (x_i - x_p), and back B = sum_i W v_i (x_i - x_p)^T, C = B D^-1);
- an affine velocity field v = A x + b survives P2C2G, the grid velocity and
  G2C2P exactly: v_p = A x_p + b, F = (I + dt A) F, B = A D;
- total mass and momentum are conserved by the transfer (zeroth and first
  moments of the weights);
- Drucker-Prager's logJp is the once-updated value (one return mapping from
  the value the transfer started with).
"""

import numpy as np

from oracle import oracle as O
from zpc_amd import scenes as S

def _scene(kind, cells=8):
    if kind == "sand":
        return S.perturb(S.sand_scene(dims=(cells, cells, cells)), seed=5, f_scale=0.02, c_scale=1.0, v_scale=0.2)
    return S.perturb(S.cube_scene(cells=cells), seed=5, f_scale=0.02, c_scale=1.0, v_scale=0.2)

def _ctx(sc):
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    return ctx

def _cols(a):
    """column-major vec9 rows -> [n, 3, 3] matrices M[r, c] = a[r + 3c]"""
    return np.asarray(a, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)

def _stencil(x, dx):
    """Per-particle node weights of the cell kernel, from its definition: the
    tent weight 1 - |x_c - x_p| / dx to the cell centres within one cell, then
    1/8 to each of the cell's 8 corner nodes. Returns (node coords [n, 27, 3],
    W [n, 27], x_i - x_p [n, 27, 3])."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    X = x / dx
    n = x.shape[0]
    lo = np.floor(X - 0.5).astype(np.int64)  # the lowest cell within reach
    acc = {}
    for cx in range(3):
        for cy in range(3):
            for cz in range(3):
                cell = lo + np.array([cx, cy, cz])
                tent = np.prod(np.clip(1.0 - np.abs(cell + 0.5 - X), 0.0, None), axis=1)
                for nx in range(2):
                    for ny in range(2):
                        for nz in range(2):
                            o = (cx + nx, cy + ny, cz + nz)
                            if o[0] > 2 or o[1] > 2 or o[2] > 2:
                                assert np.all(tent == 0.0)  # the third cell is out of reach
                                continue
                            acc[o] = acc.get(o, 0.0) + tent / 8.0
    offs = sorted(acc)
    assert len(offs) == 27
    W = np.stack([acc[o] for o in offs], 1)
    nodes = lo[:, None, :] + np.array(offs)[None]
    rel = nodes * dx - x[:, None, :]
    assert np.allclose(W.sum(1), 1.0)
    return nodes, W, rel

def _rows(ctx, nodes):
    """grid (block row, cell) of node coordinates; every node must be active"""
    act = ctx.table()["active"].astype(np.int64)
    enc = lambda k: ((k[..., 0] + 4096) * 8192 + (k[..., 1] + 4096)) * 8192 + (k[..., 2] + 4096)
    keys = enc(act)
    order = np.argsort(keys)
    blk = nodes >> 2
    e = enc(blk)
    pos = np.searchsorted(keys[order], e)
    pos = np.minimum(pos, len(keys) - 1)
    assert np.all(keys[order][pos] == e), "a stencil node outside the active blocks"
    row = order[pos]
    cell = ((nodes[..., 0] & 3) << 4) | ((nodes[..., 1] & 3) << 2) | (nodes[..., 2] & 3)
    return row, cell

def _stress64(sc, F):
    mu, lam = O.lame(sc.material.E, sc.material.nu)
    F = np.ascontiguousarray(F, np.float64).reshape(-1, 9)
    PF = np.zeros_like(F)
    O.lib().zso_stress_fc_f64(F.shape[0], sc.material.volume, mu, lam, F, PF)
    return PF

def _moment_inv(W, rel):
    D = np.einsum("pi,pia,pib->pab", W, rel, rel)
    return np.linalg.inv(D), D

def test_p2c2g_matches_per_particle_mls_scatter():
    sc = _scene("cube")
    ctx = _ctx(sc)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp).astype(np.float64)
    prm = O.make_params(sc)
    g = ctx.p2c2g_f64(prm, ps)
    nodes, W, rel = _stencil(sc.x, sc.dx)
    row, cell = _rows(ctx, nodes)
    Dinv, _ = _moment_inv(W, rel)
    m = sc.m.astype(np.float64)
    B = _cols(sc.C)
    PF = _cols(_stress64(sc, sc.F))
    Q = (m[:, None, None] * B - sc.dt * PF) @ Dinv
    mom = m[:, None, None] * sc.v.astype(np.float64).reshape(-1, 1, 3) + np.einsum("pab,pib->pia", Q, rel)
    want = np.zeros((g.shape[0], 7, 64))
    np.add.at(want, (row, 0, cell), W * m[:, None])
    for d in range(3):
        np.add.at(want, (row, 1 + d, cell), W * mom[..., d])
    assert np.allclose(g[:, 0], want[:, 0], rtol=1e-10, atol=1e-14 * m.max())
    scale = np.abs(want[:, 1:4]).max()
    err = np.abs(g[:, 1:4] - want[:, 1:4]).max() / scale
    assert err < 1e-9, err
    # conservation: total mass, total momentum (first moments vanish)
    assert np.isclose(g[:, 0].sum(), m.sum(), rtol=1e-12)
    pm = (m[:, None] * sc.v.astype(np.float64)).sum(0)
    assert np.allclose(g[:, 1:4].sum(axis=(0, 2)), pm, rtol=1e-9, atol=1e-9 * np.abs(mom).max())
    # the float32 restatement agrees up to rounding (the absolute positions of
    # the per-cell form cost ~|x| / dx ulps in the affine terms)
    ps32 = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx.p2c2g(prm, ps32)
    g32, _, _ = ctx.grids()
    assert np.abs(g32[:, 1:4] - g[:, 1:4]).max() / scale < 2e-4
    assert np.abs(g32[:, 0] - g[:, 0]).max() / g[:, 0].max() < 1e-5

def test_affine_field_round_trip():
    sc = _scene("cube")
    sc.material.E = 0.0  # no stress: the grid velocity is the transported field
    ctx = _ctx(sc)
    prm = O.make_params(sc)
    A = np.array([[0.3, -1.1, 0.4], [0.9, 0.2, -0.5], [-0.7, 0.6, 0.1]])
    b = np.array([0.2, -0.3, 0.05])
    x = sc.x.astype(np.float64).reshape(-1, 3)
    nodes, W, rel = _stencil(sc.x, sc.dx)
    _, D = _moment_inv(W, rel)
    Bm = A[None] @ D  # B = A D, so that C = B D^-1 = A
    ps = O.HostParticles(sc.x, x @ A.T + b, Bm.transpose(0, 2, 1).reshape(-1, 9), sc.F, sc.m, sc.logJp,
                         dtype=np.float64)
    g = ctx.p2c2g_f64(prm, ps)
    gv = np.zeros_like(g)
    has = g[:, 0] > 0
    for d in range(3):
        gv[:, 1 + d] = np.where(has, g[:, 1 + d] / np.where(has, g[:, 0], 1.0), 0.0)
    # every node with mass carries the field
    act = ctx.table()["active"].astype(np.int64)
    cellc = np.stack([(np.arange(64) >> 4) & 3, (np.arange(64) >> 2) & 3, np.arange(64) & 3], -1)
    xi = (act[:, None, :] * 4 + cellc[None]) * sc.dx
    fv = xi @ A.T + b
    for d in range(3):
        assert np.allclose(gv[:, 1 + d][has], fv[..., d][has], rtol=0, atol=1e-11)
    F0 = _cols(sc.F)
    ctx.g2c2p_f64(prm, ps, gv)
    assert np.allclose(ps.v.reshape(-1, 3), x @ A.T + b, rtol=0, atol=1e-11)
    assert np.allclose(_cols(ps.C), Bm, rtol=0, atol=1e-13)  # B stays as stored
    F1 = (np.eye(3) + sc.dt * A)[None] @ F0
    assert np.allclose(_cols(ps.F), F1, rtol=0, atol=1e-12)
    assert np.allclose(ps.x.reshape(-1, 3), x + (x @ A.T + b) * sc.dt, rtol=0, atol=1e-13)

def test_g2c2p_matches_per_particle_apic_gather():
    sc = _scene("cube")
    ctx = _ctx(sc)
    prm = O.make_params(sc)
    rng = np.random.default_rng(7)
    nb = ctx.sizes()["cnt"]
    gv = np.zeros((nb, 7, 64))
    gv[:, 1:4] = rng.normal(size=(nb, 3, 64))
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp, dtype=np.float64)
    ctx.g2c2p_f64(prm, ps, gv)
    nodes, W, rel = _stencil(sc.x, sc.dx)
    row, cell = _rows(ctx, nodes)
    Dinv, _ = _moment_inv(W, rel)
    vi = np.stack([gv[row, 1 + d, cell] for d in range(3)], -1)  # [n, 27, 3]
    v = np.einsum("pi,pid->pd", W, vi)
    B = np.einsum("pi,pia,pib->pab", W, vi, rel)
    x0 = sc.x.astype(np.float64).reshape(-1, 3)
    assert np.allclose(ps.v.reshape(-1, 3), v, rtol=0, atol=1e-12)
    assert np.allclose(_cols(ps.C), B, rtol=0, atol=1e-14)
    Fw = (np.eye(3)[None] + sc.dt * (B @ Dinv)) @ _cols(sc.F)
    assert np.allclose(_cols(ps.F), Fw, rtol=0, atol=1e-12)
    assert np.allclose(ps.x.reshape(-1, 3), x0 + v * sc.dt, rtol=0, atol=1e-14)
    # and the float32 restatement up to rounding
    ps32 = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx.set_grids(grid_v=gv.astype(np.float32))
    ctx.g2c2p(prm, ps32)
    assert np.abs(ps32.v.reshape(-1, 3) - v).max() / np.abs(v).max() < 1e-5
    assert np.abs(_cols(ps32.C) - B).max() / np.abs(B).max() < 1e-4

def test_sand_logjp_updated_once():
    sc = _scene("sand")
    ctx = _ctx(sc)
    prm = O.make_params(sc)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp, dtype=np.float64)
    lj0 = ps.logJp.copy()
    ctx.p2c2g_f64(prm, ps)
    mat = sc.material
    mu, lam = O.lame(mat.E, mat.nu)
    _, lj1, _, which = O.stress_sand64(mat.volume, mu, lam, mat.cohesion, mat.beta, mat.yield_surface,
                                       mat.volume_correction, lj0, sc.F.astype(np.float64))
    # the cone-tip case (2) accumulates, so a second return mapping would show
    assert np.any(which == 2) and np.any(which != 2)
    assert np.array_equal(ps.logJp, lj1)
