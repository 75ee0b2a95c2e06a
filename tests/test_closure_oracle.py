"""Pins of the learned-closure oracle (oracle/mlp_oracle.py) on the CPU.

- the backward pass against central finite differences of the loss (float64,
  no bfloat16 rounding), so a dropped term or a transposed operand fails;
- bfloat16 rounding against known answers (ties to even);
- the closure (series Newton steps + tensor return mapping) against the SVD
  return mapping of the oracle's restatement of
  physics/ConstitutiveModel_Vol_dP.hpp:241-325 in float64, including cone
  tip, yielding and elastic particles;
- the a-posteriori residual bound against the true error (log C from eigh);
- the eigen-space targets and the projection's cases and logJp against the
  same SVD return mapping.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle as O


def _sand():
    E, nu = 3.5e5, 0.3
    mu, lam = O.lame(E, nu)
    ys = 0.816496580927726 * 2.0 * np.sin(np.pi / 6) / (3.0 - np.sin(np.pi / 6))
    return dict(vol=1e-9, mu=float(mu), lam=float(lam), ys=float(ys), coh=0.0, beta=1.0, volcorr=True)


def _random_F(n, scale, seed):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, 3, 3)) * scale[:, None, None]
    Fm = np.eye(3) + G
    return np.ascontiguousarray(Fm.transpose(0, 2, 1)).reshape(n, 9)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 3.0e38, 1e-40], np.float32)
    got = M.bf16(x)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, np.float32(3.0e38), 0.0])
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: the even one is 1 + 2^-6
    assert got[0] == want[0] and got[1] == want[1] and got[2] == want[2] and got[3] == want[3]
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8
    assert abs(got[6]) < 1e-39


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64)
    p += rng.normal(size=p.shape) * 0.05  # nonzero biases
    X = rng.normal(size=(16, M.NIN))
    R = rng.normal(size=(16, M.NREC))
    R[:, 0:3] = rng.uniform(-1, 1, size=(16, 3))

    def loss(q):
        return M.loss_and_grad(q.astype(np.float32), X, R, emulate=False)[0]

    p32 = p.astype(np.float32)
    _, g = M.loss_and_grad(p32, X, R, emulate=False)
    idx = rng.choice(M.NPARAMS, 200, replace=False)
    # every tensor gets checked at its first and last entry
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    idx = np.unique(np.concatenate([idx, offs[:-1], offs[1:] - 1]))
    h = 1e-3
    for i in idx:
        a = p32.astype(np.float64).copy()
        b = a.copy()
        a[i] += h
        b[i] -= h
        # the loss sees float32 parameters: divide by the step float32 holds
        fd = (loss(a) - loss(b)) / (np.float32(a[i]) - np.float32(b[i]))
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_matches_closed_form():
    p = np.ones(4)
    m = np.zeros(4)
    v = np.zeros(4)
    g = np.array([1.0, -2.0, 0.5, 0.0])
    M.adam(p, m, v, g, 1, 0.1)
    # after one step the bias-corrected moments are g and g^2: the update is lr * sign(g)
    assert np.allclose(p, 1.0 - 0.1 * np.sign(g) * (np.abs(g) / (np.abs(g) + 1e-8)))


@pytest.mark.parametrize("scale", [1e-4, 1e-2, 0.1])
def test_closure_matches_svd_return_mapping(scale):
    """Zero start, Newton to convergence: the tensor return mapping equals
    the SVD one on every accepted particle (all three cases present)."""
    prm = _sand()
    n = 512
    F = _random_F(n, np.full(n, scale), seed=int(scale * 1e5) + 5)
    rng = np.random.default_rng(9)
    lj = rng.normal(size=n) * 3 * scale
    PF, lj_new, res, ok = M.closure(F, lj, prm["vol"], prm["mu"], prm["lam"], prm["ys"], prm["coh"],
                                    prm["beta"], prm["volcorr"], y=None, iters=8)
    assert ok.mean() > 0.95, f"accepted {ok.mean():.3f}"
    PF64, lj64, _, which = O.stress_sand64(prm["vol"], prm["mu"], prm["lam"], prm["coh"], prm["beta"],
                                           prm["ys"], prm["volcorr"], lj, F)
    assert set(np.unique(which[ok])) >= ({1, 2, 3} if scale >= 1e-2 else {1, 2})
    den = np.max(np.abs(PF64[ok]))
    # the converged Newton iterate carries the exp truncation, below 1e-7 for |X| <= 1
    assert np.max(np.abs(PF[ok] - PF64[ok])) / den < 1e-7
    assert np.max(np.abs(lj_new[ok] - lj64[ok])) <= 1e-10 + 1e-8 * np.max(np.abs(lj64))


def test_closure_rejects_inverted_and_large():
    prm = _sand()
    F = _random_F(3, np.array([1e-3, 1e-3, 1e-3]), seed=2)
    Fm = F.reshape(3, 3, 3).transpose(0, 2, 1).copy()
    Fm[0, :, 0] *= -1.0  # det < 0
    Fm[1] *= 2.0  # |log C| > 1
    F = np.ascontiguousarray(Fm.transpose(0, 2, 1)).reshape(3, 9)
    _, _, _, ok = M.closure(F, np.zeros(3), prm["vol"], prm["mu"], prm["lam"], prm["ys"], prm["coh"],
                            prm["beta"], prm["volcorr"], iters=3)
    assert list(ok) == [False, False, True]


def _exact_coefficients(F):
    """psi with psi0 + psi1 m_i + psi2 m_i^2 = e_i / n (Vandermonde, float64)."""
    R = M.targets(F, None, 1.0, 1.0, 0.3, 0.0)
    V = np.stack([np.ones_like(R[:, 0:3]), R[:, 0:3], R[:, 0:3] ** 2], 2)
    return np.linalg.solve(V, R[:, 3:6, None])[:, :, 0]


@pytest.mark.parametrize("noise", [0.0, 1e-3, 3e-2])
def test_residual_bound_bounds_the_error(noise):
    """One step from a coaxial guess: the a-posteriori bound is at least the
    true error of X against log C from eigh; exact coefficients pass at 1e-8."""
    n = 400
    F = _random_F(n, np.geomspace(1e-3, 0.12, n), seed=7)
    psi = _exact_coefficients(F)
    rng = np.random.default_rng(8)
    psi = psi + noise * rng.normal(size=psi.shape) * np.abs(psi).max(1, keepdims=True)
    y = np.concatenate([psi, np.zeros((n, 4))], 1)
    A = M.c_minus_i(F)
    X0 = M.initial_log(y, A)
    X, d = M.newton(X0, A, 1)
    err = np.sqrt(np.sum((X - 2 * M.hencky(F)) ** 2, (1, 2)))
    xn = np.sqrt(np.sum(X * X, (1, 2)))
    res = M.residual_bound(d, np.sqrt(np.sum(X0 * X0, (1, 2))), np.sqrt(np.sum(A * A, (1, 2))), xn)
    assert np.all(err / xn <= res + 1e-12)
    if noise == 0.0:
        assert res.max() < 1e-6


def test_targets_and_projection_match_oracle_cases():
    prm = _sand()
    n = 256
    F = _random_F(n, np.geomspace(1e-5, 0.2, n), seed=4)
    rng = np.random.default_rng(5)
    lj = rng.normal(size=n) * 1e-2
    R = M.targets(F, lj, prm["mu"], prm["lam"], prm["ys"], prm["coh"])
    f, nrm = M.features(F, lj, prm["mu"], prm["lam"], prm["ys"], prm["coh"])
    # the eigenvalues m_i are those of U: trace and tr U^3 agree with the features
    assert np.allclose(R[:, 0:3].sum(1), f[:, 0])
    assert np.allclose((R[:, 0:3] ** 3).sum(1), f[:, 1])
    assert np.allclose((R[:, 0:3] ** 2).sum(1), 1.0)
    E = M.hencky(F)
    assert np.allclose(np.sort(np.linalg.eigvalsh(E), 1), np.sort(R[:, 3:6] * nrm[:, None], 1),
                       rtol=0, atol=1e-13)
    PF64, lj64, _, which = O.stress_sand64(prm["vol"], prm["mu"], prm["lam"], prm["coh"], prm["beta"], prm["ys"],
                                           prm["volcorr"], lj, F)
    _, tre, dg, case = M.project(R[:, 3:6] * nrm[:, None], lj, prm["mu"], prm["lam"], prm["ys"], prm["coh"])
    assert np.array_equal(case, which)
    assert np.allclose(np.where(case == 2, lj + tre, 0.0), lj64, rtol=0, atol=1e-12)
