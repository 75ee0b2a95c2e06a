"""CPU tests pinning the oracle (the checker the HIP path is compared to).

Every oracle function is checked against something other than its own
formula: the reference's own math headers (golden vectors under tests/golden/
and, when built, oracle/_ref), closed forms, invariants, brute force on tiny
inputs, and conservation laws.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "refmath.npz")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN, allow_pickle=False)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else a.dtype),
                                                 b.view(np.uint32 if b.dtype == np.float32 else b.dtype))


# ------------------------------------------------------------------ golden vectors (zpc's headers)

def test_svd_golden(gold):
    U, S_, V = O.svd(gold["svd_F"])
    assert bits_equal(U, gold["svd_U"]) and bits_equal(S_, gold["svd_S"]) and bits_equal(V, gold["svd_V"])


def test_lame_golden(gold):
    for (E, nu), (mu, lam) in zip(gold["E_nu"], gold["lame"]):
        m2, l2 = O.lame(float(E), float(nu))
        assert np.float32(m2) == mu and np.float32(l2) == lam


def test_stress_fc_golden(gold):
    mu, lam = gold["lame"][2]
    PF = O.stress_fc(float(gold["fc_vol"]), float(mu), float(lam), gold["svd_F"])
    assert bits_equal(PF, gold["fc_PF"])


def test_stress_sand_golden(gold):
    mu, lam = gold["lame"][2]
    PF, lj, Fp, which = O.stress_sand(float(gold["fc_vol"]), float(mu), float(lam), 0.0, 1.0,
                                      float(gold["sand_ys"]), 1, gold["sand_logJp_in"], gold["sand_F"])
    assert bits_equal(PF, gold["sand_PF"])
    assert bits_equal(lj, gold["sand_logJp_out"])
    assert bits_equal(Fp, gold["sand_Fproj"])
    # the fixture exercises every return-mapping case
    assert set(np.unique(which)) >= {1, 2, 3}


def test_arena_golden(gold):
    corner, w, diff = O.arena(float(gold["arena_dx"]), gold["arena_x"])
    assert np.array_equal(corner, gold["arena_corner"])
    assert bits_equal(w, gold["arena_w"]) and bits_equal(diff, gold["arena_diff"])


def test_hash_golden(gold):
    L = O.lib()
    got = np.array([L.zso_hash(int(a), int(b), int(c)) for a, b, c in gold["hash_keys"]], np.int32)
    assert np.array_equal(got, gold["hash_vals"])


def test_next_2pow_golden(gold):
    got = [S.next_2pow(int(v)) for v in gold["n2p_in"]]
    assert got == list(gold["n2p_out"])
    L = O.lib()
    for v, e in zip(gold["n2p_in"], gold["n2p_out"]):
        assert L.zso_evaluate_table_size(int(v)) == 16 * int(e)


# ------------------------------------------------------------------ against the built reference headers

# The check loads oracle/_ref only when it runs (never at collection), so a
# GPU-only test run maps nothing built from the reference.
@pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="oracle/_ref not built (reference tree absent)")
def test_oracle_matches_reference_headers_random():
    R = O.ref()
    rng = np.random.default_rng(1)
    n = 20000
    F = np.concatenate([rng.normal(size=(n // 2, 9)),
                        np.eye(3).reshape(9) + 0.02 * rng.normal(size=(n // 2, 9))]).astype(np.float32)
    U, S_, V = O.svd(F)
    Ur, Sr, Vr = np.zeros_like(F), np.zeros((n, 3), np.float32), np.zeros_like(F)
    R.zref_svd(n, F, Ur, Sr, Vr)
    assert bits_equal(U, Ur) and bits_equal(S_, Sr) and bits_equal(V, Vr)
    mu, lam = O.lame(3.5e5, 0.3)
    lj = (0.05 * rng.normal(size=n)).astype(np.float32)
    PF, lj2, Fp, _ = O.stress_sand(1e-6, mu, lam, 0.0, 1.0, 0.3266, 1, lj, F)
    PFr, lj3, Fpr = np.zeros_like(F), lj.copy(), np.zeros_like(F)
    R.zref_stress_sand(n, 1e-6, mu, lam, 0.0, 1.0, 0.3266, 1, lj3, F, PFr, Fpr.ctypes.data_as(C.c_void_p))
    assert bits_equal(PF, PFr) and bits_equal(lj2, lj3) and bits_equal(Fp, Fpr)
    x = rng.uniform(-1, 2, size=(n, 3)).astype(np.float32)
    c1, w1, d1 = O.arena(1 / 256, x)
    c2, w2, d2 = np.zeros_like(c1), np.zeros_like(w1), np.zeros_like(d1)
    R.zref_arena(n, 1 / 256, x, c2, w2, d2)
    assert np.array_equal(c1, c2) and bits_equal(w1, w2) and bits_equal(d1, d2)


# ------------------------------------------------------------------ SVD invariants and the FP64 twin

def _cm(a):  # column-major flat -> matrix
    return a.reshape(-1, 3, 3).transpose(0, 2, 1)


def test_svd_invariants():
    rng = np.random.default_rng(2)
    F = (np.eye(3).reshape(9) + 0.3 * rng.normal(size=(5000, 9))).astype(np.float32)
    U, S_, V = O.svd(F)
    Um, Vm, Fm = _cm(U).astype(np.float64), _cm(V).astype(np.float64), _cm(F).astype(np.float64)
    rec = Um @ (S_[:, :, None] * np.transpose(Vm, (0, 2, 1)))
    # zpc's fixed 4-sweep approximate Jacobi (SVD.hpp:95) is not fully converged
    # on a small fraction of inputs -- nearly repeated singular values in
    # particular -- so the bound is on the distribution, not the worst case
    e = np.abs(rec - Fm).max(axis=(1, 2)) / np.abs(Fm).max(axis=(1, 2))
    assert np.median(e) < 1e-6 and np.percentile(e, 99) < 2e-5 and e.max() < 1e-2
    eye = np.eye(3)
    assert np.max(np.abs(Um @ np.transpose(Um, (0, 2, 1)) - eye)) < 1e-5
    assert np.max(np.abs(Vm @ np.transpose(Vm, (0, 2, 1)) - eye)) < 1e-5
    assert np.all(np.linalg.det(Um) > 0) and np.all(np.linalg.det(Vm) > 0)
    assert np.all(S_[:, 0] >= S_[:, 1] - 1e-6) and np.all(S_[:, 1] >= np.abs(S_[:, 2]) - 1e-6)


def test_svd_fp64_matches_lapack():
    rng = np.random.default_rng(3)
    F = np.concatenate([rng.normal(size=(2000, 9)), np.eye(3).reshape(1, 9).repeat(5, 0),
                        np.diag([2.0, 2.0, 0.5]).reshape(1, 9)])
    U, S_, V = O.svd64(F)
    sv = np.linalg.svd(_cm(F), compute_uv=False)
    assert np.max(np.abs(np.abs(S_) - sv)) < 1e-12 * max(1.0, sv.max())
    Um, Vm = _cm(U), _cm(V)
    rec = Um @ (S_[:, :, None] * np.transpose(Vm, (0, 2, 1)))
    assert np.max(np.abs(rec - _cm(F))) < 1e-12
    assert np.all(np.linalg.det(Um) > 0) and np.all(np.linalg.det(Vm) > 0)
    # sigma_1, sigma_2 >= 0 and sigma_3 carries the sign of det F
    assert np.all(S_[:, :2] >= 0)
    det = np.linalg.det(_cm(F))
    nz = np.abs(S_[:, 2]) > 1e-12
    assert np.all(np.sign(S_[nz, 2]) == np.sign(det[nz]))


def test_fixed_corotated_rest_and_rotation_invariance():
    mu, lam = O.lame(1e4, 0.3)
    eye = np.eye(3, dtype=np.float32).reshape(1, 9)
    assert np.all(O.stress_fc(1.0, mu, lam, eye) == 0)
    rng = np.random.default_rng(4)
    F = (np.eye(3).reshape(9) + 0.1 * rng.normal(size=(200, 9))).astype(np.float32)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.linalg.det(q))
    QF = np.transpose(q @ _cm(F).astype(np.float64), (0, 2, 1)).reshape(-1, 9).astype(np.float32)
    PF = _cm(O.stress_fc(1.0, mu, lam, F)).astype(np.float64)
    PF_q = _cm(O.stress_fc(1.0, mu, lam, QF)).astype(np.float64)
    # P F^T transforms as Q (P F^T) Q^T under F -> Q F
    assert np.max(np.abs(q @ PF @ q.T - PF_q)) < 2e-3 * max(np.abs(PF).max(), 1.0)


def test_fixed_corotated_energy_gradient():
    """P = dPsi/dF for Psi = mu |F - R|^2 + lam/2 (J-1)^2, by central differences in float64."""
    mu, lam = 3.0, 5.0
    rng = np.random.default_rng(5)

    def psi(Fm):
        u, s, vt = np.linalg.svd(Fm)
        if np.linalg.det(u @ vt) < 0:
            s = s.copy()
            s[-1] = -s[-1]
        return mu * np.sum((s - 1) ** 2) + 0.5 * lam * (np.prod(s) - 1) ** 2

    for _ in range(5):
        Fm = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        Fflat = Fm.T.reshape(1, 9)
        PF = _cm(O.lib() and np.asarray(_stress_fc64(Fflat, mu, lam)))[0]
        P = PF @ np.linalg.inv(Fm.T)
        num = np.zeros((3, 3))
        h = 1e-6
        for i in range(3):
            for j in range(3):
                d = np.zeros((3, 3))
                d[i, j] = h
                num[i, j] = (psi(Fm + d) - psi(Fm - d)) / (2 * h)
        assert np.max(np.abs(P - num)) < 1e-5 * max(1.0, np.abs(num).max())


def _stress_fc64(F, mu, lam):
    F = np.ascontiguousarray(F, np.float64)
    PF = np.zeros_like(F)
    O.lib().zso_stress_fc_f64(F.shape[0], 1.0, mu, lam, F, PF)
    return PF


def test_sand_return_map_cases():
    mu, lam = O.lame(3.5e5, 0.3)
    ys = S.Material.drucker_prager(3.5e5, 0.3, 1e3).yield_surface
    # pure expansion with logJp = 0 -> cone tip: projected F = I, logJp += sum log s
    F = (np.eye(3) * 1.02).reshape(1, 9).astype(np.float32)
    PF, lj, Fp, which = O.stress_sand(1.0, mu, lam, 0.0, 1.0, ys, 1, np.zeros(1, np.float32), F)
    assert which[0] == 2
    assert np.allclose(Fp, np.eye(3).reshape(1, 9), atol=1e-6)
    assert abs(lj[0] - 3 * np.log(np.float32(1.02))) < 1e-5
    assert np.max(np.abs(PF)) < 1e-3  # stress-free at the tip
    # uniform compression stays inside the cone: F unchanged, logJp reset to 0
    F = (np.eye(3) * 0.99).reshape(1, 9).astype(np.float32)
    PF, lj, Fp, which = O.stress_sand(1.0, mu, lam, 0.0, 1.0, ys, 1, np.full(1, 0.01, np.float32), F)
    assert which[0] == 1 and lj[0] == 0 and np.allclose(Fp, F, atol=1e-6)
    # strong shear with compression projects onto the cone surface: the projected
    # deviatoric Hencky strain equals the cone radius
    F = np.array([[0.98, 0.2, 0], [0, 0.98, 0], [0, 0, 0.98]], np.float32).T.reshape(1, 9)
    PF, lj, Fp, which = O.stress_sand(1.0, mu, lam, 0.0, 1.0, ys, 1, np.zeros(1, np.float32), F)
    assert which[0] == 3
    s = np.linalg.svd(_cm(Fp)[0], compute_uv=False)
    e = np.log(s)
    eh = e - e.mean()
    tr = np.log(np.linalg.svd(_cm(F)[0], compute_uv=False)).sum()
    radius = -(3 * lam + 2 * mu) / (2 * mu) * tr * ys
    assert abs(np.linalg.norm(eh) - radius) < 1e-4 * max(radius, 1e-6)


# ------------------------------------------------------------------ table, activation, bucket sort

def _py_hash(k0, k1, k2):
    k0, k1, k2 = int(k0), int(k1), int(k2)
    M = (1 << 64) - 1
    r = k0 & M
    r ^= ((k1 & M) + 0x9E3779B97F4A7C15 + ((r << 12) & M) + (r >> 4)) & M
    r ^= ((k2 & M) + 0x9E3779B97F4A7C15 + ((r << 12) & M) + (r >> 4)) & M
    v = r & 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def _c_mod(a, b):
    return a - b * int(a / b) if a >= 0 else -((-a) - b * int((-a) / b))


def _py_insert_serial(keys, T):
    """Serial insertion in order with zpc's probe sequence: returns (slot -> index, active)."""
    idx = [-1] * T
    slot_key = [None] * T
    active = []
    for k in keys:
        h = _py_hash(*k)
        s = (_c_mod(h, T) + T) % T
        while slot_key[s] is not None and slot_key[s] != k:
            s = (s + 127) % T
        if slot_key[s] is None:
            slot_key[s] = k
            idx[s] = len(active)
            active.append(k)
    return idx, slot_key, active


def test_hash_closed_form_matches_oracle():
    rng = np.random.default_rng(6)
    L = O.lib()
    for k in rng.integers(-2 ** 31, 2 ** 31 - 1, size=(2000, 3)):
        k = tuple(int(v) for v in k)
        assert L.zso_hash(*k) == _py_hash(*k)
        for T in (16, 1024, 1 << 20):
            h = _py_hash(*k)
            assert L.zso_home_slot(*k, T) == (_c_mod(h, T) + T) % T


@pytest.mark.parametrize("T", [64, 256, 4096])
def test_activation_and_enlarge_match_serial_insertion(T):
    sc = S.tiny_scene(n_cells=(3, 4, 3), ppc=3)
    ctx = O.Context()
    ctx.activate(T, sc.x, sc.dx)
    keys = [tuple(int(v) for v in k) for k in O.block_keys(sc.x, sc.dx)]
    idx, slot_key, active = _py_insert_serial(keys, T)
    t = ctx.table()
    assert t["cnt"] == len(active)
    assert np.array_equal(t["indices"], np.array(idx, np.int32))
    assert [tuple(a) for a in t["active"]] == active
    # first-occurrence numbering
    firsts = []
    for k in keys:
        if k not in firsts:
            firsts.append(k)
    assert active == firsts
    # enlargement: key + {0,1}^3 in (dx, dy, dz) order, block by block
    ctx.enlarge()
    cand = [(a[0] + i, a[1] + j, a[2] + k) for a in active for i in (0, 1) for j in (0, 1) for k in (0, 1)]
    idx2, _, active2 = _py_insert_serial(keys + cand, T)
    t2 = ctx.table()
    assert [tuple(a) for a in t2["active"]] == active2
    assert np.array_equal(t2["indices"], np.array(idx2, np.int32))
    # every active key is found by a probe walk from its home slot
    for i, k in enumerate(active2):
        s = (_c_mod(_py_hash(*k), T) + T) % T
        while tuple(t2["keys"][s]) != k:
            assert t2["indices"][s] != -1
            s = (s + 127) % T
        assert t2["indices"][s] == i


def test_bucket_sort_is_stable_counting_sort():
    sc = S.perturb(S.cube_scene(cells=6), seed=2)
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    counts, offsets, indices = ctx.sort()
    t = ctx.table()
    lut = {tuple(k): i for i, k in enumerate(t["active"])}
    b = np.array([lut[tuple(k)] for k in O.block_keys(sc.x, sc.dx)])
    assert np.array_equal(indices, np.argsort(b, kind="stable"))
    assert np.array_equal(counts[:-1], np.bincount(b, minlength=t["cnt"]))
    assert counts[-1] == 0 and offsets[-1] == sc.n
    assert np.array_equal(offsets[:-1], np.concatenate([[0], np.cumsum(counts[:-2])]))


def test_block_key_closed_form():
    rng = np.random.default_rng(8)
    x = rng.uniform(-0.3, 1.3, size=(5000, 3)).astype(np.float32)
    dx = np.float32(1 / 128)
    k = O.block_keys(x, float(dx))
    c = np.floor(x * np.float32(1.0 / dx) + np.float32(0.5)).astype(np.int64) - 2
    assert np.array_equal(k, np.floor_divide(c, 4))


# ------------------------------------------------------------------ transfer operators

def _p2g_grid(sc):
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.p2g(O.make_params(sc), ps)
    g, _, _ = ctx.grids()
    return ctx, g, ps


def test_p2g_conserves_mass_and_momentum():
    sc = S.perturb(S.cube_scene(cells=8), seed=3)
    ctx, g, ps = _p2g_grid(sc)
    m = sc.m.astype(np.float64)
    assert abs(g[:, 0].sum(dtype=np.float64) - m.sum()) < 1e-5 * m.sum()
    p = (m[:, None] * sc.v).sum(0)
    # APIC: sum_i W (x_i - x_p) = 0, so grid momentum equals particle momentum
    assert np.max(np.abs(g[:, 1:4].sum(axis=(0, 2), dtype=np.float64) - p)) < 1e-5 * np.abs(m[:, None] * sc.v).sum()
    # internal forces sum to zero (sum_i W (x_i - x_p) = 0 again)
    f = g[:, 4:7].sum(axis=(0, 2), dtype=np.float64)
    assert np.max(np.abs(f)) < 1e-4 * np.abs(g[:, 4:7]).sum()


def test_p2g_single_particle_known_answer():
    dx = 1.0 / 64
    x = np.array([[0.5 + 0.3 * dx, 0.5 + 0.7 * dx, 0.5 + 0.5 * dx]], np.float32)
    mat = S.Material(E=0.0, nu=0.3, volume=dx ** 3)
    sc = S.Scene("one", dx, 1e-4, mat, x, np.array([[1.0, -2.0, 0.5]], np.float32), np.zeros((1, 9), np.float32),
                 np.eye(3, dtype=np.float32).reshape(1, 9), np.array([2.0], np.float32), None,
                 table_size=1024)
    ctx, g, _ = _p2g_grid(sc)
    t = ctx.table()
    X = x[0].astype(np.float64) / dx
    base = np.floor(X - 0.5)
    d = X - base
    w = np.stack([0.5 * (1.5 - d) ** 2, 0.75 - (d - 1) ** 2, 0.5 * (d - 0.5) ** 2])  # (node, axis)
    lut = {tuple(k): i for i, k in enumerate(t["active"])}
    for a in range(3):
        for b in range(3):
            for c in range(3):
                node = (base + [a, b, c]).astype(int)
                bk = tuple(np.floor_divide(node, 4))
                cell = ((node[0] & 3) << 4) | ((node[1] & 3) << 2) | (node[2] & 3)
                W = w[a, 0] * w[b, 1] * w[c, 2]
                assert abs(g[lut[bk], 0, cell] - 2.0 * W) < 1e-6
                assert np.allclose(g[lut[bk], 1:4, cell], 2.0 * W * np.array([1.0, -2.0, 0.5]), atol=1e-6)
    assert abs(g[:, 0].sum() - 2.0) < 1e-6


def test_affine_velocity_round_trip():
    """P2G then G2P with no force and no gravity reproduces v = A x + b and C = A."""
    rng = np.random.default_rng(9)
    sc = S.cube_scene(cells=6)
    Aff = 0.5 * rng.normal(size=(3, 3))
    bvec = rng.normal(size=3)
    sc.v = (sc.x.astype(np.float64) @ Aff.T + bvec).astype(np.float32)
    sc.C = np.tile(Aff.T.reshape(1, 9), (sc.n, 1)).astype(np.float32)  # column-major C[d + 3e] = A[d, e]
    sc.material = S.Material(E=0.0, nu=0.3, volume=sc.material.volume)
    sc.gravity = (0.0, 0.0, 0.0)
    sc.colliders = []
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    x0 = sc.x.copy()
    ctx = O.Context()
    ctx.run(O.make_params(sc), ps, 1)
    v = ps.v.reshape(-1, 3)
    vexp = x0.astype(np.float64) @ Aff.T + bvec
    assert np.max(np.abs(v - vexp)) < 2e-5 * np.abs(vexp).max()
    Cg = ps.C.reshape(-1, 9)
    assert np.max(np.abs(Cg - Aff.T.reshape(1, 9))) < 5e-3 * np.abs(Aff).max()


def test_sticky_walls_zero_boundary_velocity():
    sc = S.cube_scene(cells=6)
    sc.x = sc.x - np.float32(0.23)  # push the cube against the low walls
    sc.v[:] = np.array([-1.0, -1.0, -1.0], np.float32)
    ctx = O.Context()
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx.run(O.make_params(sc), ps, 1)
    _, gv, _ = ctx.grids()
    t = ctx.table()
    pos = (t["active"][:, None, :] * 4 + np.stack([(np.arange(64) >> 4) & 3, (np.arange(64) >> 2) & 3,
                                                    np.arange(64) & 3], 1)[None]) * sc.dx
    inside_wall = np.any(pos < 3 * sc.dx, axis=2)
    has_mass = gv[:, 0, :] > 0
    sel = inside_wall & has_mass
    assert sel.sum() > 0
    assert np.all(gv[:, 1:4, :].transpose(0, 2, 1)[sel] == 0)


def test_omp_baseline_matches_seq_oracle():
    sc = S.perturb(S.cube_scene(cells=10), seed=21)
    a = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    b = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    O.Context().run(O.make_params(sc), a, 2)
    secs, nt, nb = O.omp_run(O.make_params(sc), b, 0, 2, threads=4)
    assert secs > 0 and nt == 4 and nb > 0
    for f in ("x", "v", "C", "F"):
        ra, rb = getattr(a, f), getattr(b, f)
        assert np.max(np.abs(ra - rb)) <= 1e-5 * max(np.abs(ra).max(), 1e-30), f


def test_fp64_twin_close_to_fp32():
    sc = S.perturb(S.sand_scene(dims=(8, 16, 8)), seed=5, f_scale=0.02)
    a = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    b = a.astype(np.float64)
    O.Context().run(O.make_params(sc), a, 3)
    O.Context(np.float64).run(O.make_params(sc), b, 3)
    for f in ("x", "v", "F"):
        ra, rb = getattr(a, f), getattr(b, f)
        assert np.max(np.abs(ra - rb)) <= 1e-4 * np.abs(rb).max(), f


def test_reorder_is_vector_gather():
    """Vector::reorder(gather): ordered[i] = vs[map[i]] for every attribute."""
    sc = S.perturb(S.sand_scene(dims=(4, 6, 4)), seed=8)
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    order = np.random.default_rng(4).permutation(ps.n).astype(np.int32)
    before = {k: getattr(ps, k).reshape(ps.n, -1).copy() for k in ("x", "v", "C", "F", "m", "logJp")}
    ps.reorder(order)
    for k, a in before.items():
        assert np.array_equal(getattr(ps, k).reshape(ps.n, -1), a[order]), k
    assert np.array_equal(ps.ids, order)
    ps.reorder(np.argsort(order).astype(np.int32))
    assert np.array_equal(ps.ids, np.arange(ps.n))


def test_sorted_substep_is_substep_then_bucket_gather():
    """One sorted substep = the plain substep followed by the gather of every
    attribute by that substep's bucket sort (bit for bit); later substeps only
    change the summation order, so they stay within float rounding."""
    sc = S.perturb(S.cube_scene(cells=10), seed=31)
    a = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    b = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    prm = O.make_params(sc)
    ca, cb = O.Context(), O.Context()
    ca.run(prm, a, 1)
    cb.run(prm, b, 1, reorder=True)
    _, _, idx = ca.sort()
    assert np.array_equal(b.ids, idx)
    for f in ("x", "v", "C", "F", "m"):
        assert np.array_equal(b.by_id(f), getattr(a, f)), f
    ca.run(prm, a, 4)
    cb.run(prm, b, 4, reorder=True)
    for f in ("x", "v", "F"):
        ra = getattr(a, f)
        assert np.max(np.abs(b.by_id(f) - ra)) <= 1e-5 * np.abs(ra).max(), f


# ------------------------------------------------------------------ the C ABI library

def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "zsmpm.h")).read()
    return sorted(set(re.findall(r"^\s*int32_t\s+(zsmpm_\w+)\s*\(", txt, re.M)))


def test_hip_library_exports_every_declared_symbol():
    from zpc_amd import rocm

    if not os.path.exists(rocm.LIB_PATH):
        pytest.skip("libzsmpm_rocm.so not built")
    L = C.CDLL(rocm.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 16
    for s in syms:
        assert hasattr(L, s), s
    # the ctypes binding declares every symbol too
    for s in syms:
        assert s in open(os.path.join(ROOT, "zpc_amd", "rocm.py")).read(), s


def test_abi_version_without_gpu():
    from zpc_amd import rocm

    if not os.path.exists(rocm.LIB_PATH):
        pytest.skip("libzsmpm_rocm.so not built")
    h = rocm.Header()
    assert rocm.lib().zsmpm_abi_version(h) == 0
    assert (h.major, h.minor, h.size) == (1, 0, 8)
