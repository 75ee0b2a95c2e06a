"""CPU check of the node-value decomposition the GPU P2G uses.

The HIP P2G evaluates each stencil node as W(o) * (B + A o) with
A = dx * (m C [+ contrib']) and B = m v - A l (l = local position in grid
units, o in {0,1,2}^3), instead of zpc's W m (v + C xixp) and
W contrib' xixp with xixp = (o - l) dx (simulation/transfer/P2G.hpp:104-123).
This restates both forms in numpy (float64) and checks that they agree and
that the sum over a tiny scene equals the oracle's P2G grid.
"""
import numpy as np

from oracle import oracle as O
from zpc_amd import scenes as S


def _per_particle_terms(sc, ps_after, logJp_in):
    dx = np.float64(sc.dx)
    x = sc.x.astype(np.float64)
    X = x / dx
    base = np.floor(X - 0.5)
    l = X - base
    mu, lam = O.lame(np.float32(sc.material.E), np.float32(sc.material.nu))
    if sc.material.kind == S.FIXED_COROTATED:
        PF = O.stress_fc(np.float32(sc.material.volume), mu, lam, sc.F).astype(np.float64)
    else:
        PF, _, _, _ = O.stress_sand(np.float32(sc.material.volume), mu, lam, 0.0, 1.0,
                                    np.float32(sc.material.yield_surface), 1, logJp_in, sc.F)
        PF = PF.astype(np.float64)
    dinv = 4.0 / (dx * dx)
    contrib = PF * (-np.float64(sc.dt) * dinv)
    return x, base.astype(np.int64), l, contrib


def _weights(l):
    d0 = l - np.floor(l - 0.5)
    w0 = 0.5 * (1.5 - d0) ** 2
    d1 = d0 - 1.0
    w1 = 0.75 - d1 * d1
    w2 = 0.5 * (0.5 + d1) ** 2
    return np.stack([w0, w1, w2], axis=-1)  # (n, 3 axes, 3 nodes)


def test_affine_form_matches_reference_form():
    sc = S.perturb(S.tiny_scene(), seed=4)
    x, base, l, contrib = _per_particle_terms(sc, None, sc.logJp)
    dx = float(sc.dx)
    m = sc.m.astype(np.float64)
    v = sc.v.astype(np.float64)
    Cm = sc.C.astype(np.float64).reshape(-1, 3, 3, order="F")  # C[d, e] = C[d + 3e]
    cf = contrib.reshape(-1, 3, 3, order="F")
    w = _weights(l)
    A_mv = dx * m[:, None, None] * Cm
    A_f = dx * cf
    B_mv = m[:, None] * v - np.einsum("nde,ne->nd", A_mv, l)
    B_f = -np.einsum("nde,ne->nd", A_f, l)
    A_fold = A_mv + A_f
    B_fold = B_mv + B_f
    for a in range(3):
        for b in range(3):
            for c in range(3):
                o = np.array([a, b, c], np.float64)
                W = w[:, 0, a] * w[:, 1, b] * w[:, 2, c]
                xixp = (o[None] - l) * dx
                ref_mv = (W * m)[:, None] * (v + np.einsum("nde,ne->nd", Cm, xixp))
                ref_f = W[:, None] * np.einsum("nde,ne->nd", cf, xixp)
                got_mv = W[:, None] * (B_mv + A_mv @ o)
                got_f = W[:, None] * (B_f + A_f @ o)
                got_fold = W[:, None] * (B_fold + A_fold @ o)
                scale = np.abs(ref_mv).max() + np.abs(ref_f).max()
                assert np.abs(got_mv - ref_mv).max() <= 1e-12 * scale
                assert np.abs(got_f - ref_f).max() <= 1e-12 * scale
                assert np.abs(got_fold - (ref_mv + ref_f)).max() <= 1e-12 * scale


def test_affine_form_sums_to_oracle_grid():
    """Dense-grid sum of the affine form equals the oracle's sparse P2G grid."""
    sc = S.perturb(S.tiny_scene(material=S.Material.drucker_prager(3.5e5, 0.3, 1e3, volume=(1 / 32) ** 3 / 2)),
                   seed=9)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    prm = O.make_params(sc)
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.p2g(prm, ps)
    g, _, _ = ctx.grids()
    t = ctx.table()
    x, base, l, contrib = _per_particle_terms(sc, None, sc.logJp)
    dx = float(sc.dx)
    m = sc.m.astype(np.float64)
    v = sc.v.astype(np.float64)
    Cm = sc.C.astype(np.float64).reshape(-1, 3, 3, order="F")
    A = dx * (m[:, None, None] * Cm + contrib.reshape(-1, 3, 3, order="F"))
    B = m[:, None] * v - np.einsum("nde,ne->nd", A, l)
    w = _weights(l)
    dense = {}
    for p in range(sc.n):
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    o = np.array([a, b, c], np.float64)
                    W = w[p, 0, a] * w[p, 1, b] * w[p, 2, c]
                    node = tuple(base[p] + np.array([a, b, c]))
                    acc = dense.setdefault(node, np.zeros(4))
                    acc[0] += W * m[p]
                    acc[1:] += W * (B[p] + A[p] @ o)
    blk = {tuple(k): i for i, k in enumerate(t["active"])}
    for node, acc in dense.items():
        bk = tuple(np.floor_divide(np.array(node), 4))
        cell = ((node[0] & 3) << 4) | ((node[1] & 3) << 2) | (node[2] & 3)
        gv = g[blk[bk], :, cell].astype(np.float64)
        assert abs(gv[0] - acc[0]) <= 1e-5 * abs(acc[0]) + 1e-12
        np.testing.assert_allclose(gv[1:4] + gv[4:7], acc[1:], rtol=1e-4, atol=1e-6 * np.abs(acc[1:]).max())
