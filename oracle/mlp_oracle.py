"""Learned Drucker-Prager stress closure: CPU restatement. TEST INFRASTRUCTURE ONLY.

The closure replaces the SVD of the plain return mapping
(physics/ConstitutiveModel_Vol_dP.hpp:241-325, the sand branch of
simulation/transfer/P2G.hpp:84-100) for a particle. With A = C - I,
C = F^T F (invariant under F -> Q F), n = |A|_F and U = A / n:

1. features: invariants of U (tr U, tr U^3, det U), the scale n, logJp,
   cohesion and the material (9 inputs);
2. an MLP 9 -> 64 -> 64 -> 7 (GELU, bfloat16 operands, float accumulation)
   predicts the Hencky strain E = log(C)/2 and the projected strain Lp of the
   return mapping as isotropic polynomials (Cayley-Hamilton):
   E / n = psi0 I + psi1 U + psi2 U^2, Lp / n = rho0 I + rho1 U + rho2 U^2,
   and the plastic multiplier dg (7 outputs). Both tensors are coaxial with C
   by construction;
3. one Newton step X = X0 + log1p(D), D = exp(-X0) C - I, from X0 = 2 E_pred.
   X0 commutes with C, so log C = X0 + log(I + D) exactly and the step's only
   errors are the degree-3 log1p and the exp truncations (degree-6 Taylor of
   exp(-X0/4), squared twice), bounded by |D|^4 / (4 (1 - |D|)) and
   4 e^(4y) y^7 / 7! (1 + |A|), y = |X0| / 4;
4. the return mapping in tensor form on E = X / 2 (it acts on principal values
   only, so its projection is the same polynomial map), the Kirchhoff stress
   tau = 2 mu Lp + lambda tr(Lp) I in the frame of C, S = tau C^-1 and
   vol P F_proj^T = vol F S F^T (equal to the reference's U diag(P_hat sigma)
   U^T since all of it is coaxial with C);
5. fallback to the SVD path where the bound relative to |X| exceeds the
   tolerance, |D| >= 1/2, |X| > 1 or det F <= 0.

The network is trained on (features, eigen-space targets) pairs: for the
eigenvalues m_i of U, the loss is sum_i (psi0 + psi1 m_i + psi2 m_i^2 - e_i/n)^2
+ the same for Lp + (dg_pred - dg)^2, the Frobenius error of the coaxial
tensors. Only ``tests/`` uses this module. Everything here is float64 numpy;
the device rounds to bfloat16 exactly where ``bf16`` is applied below.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 9, 64, 7
NREC = 10  # target record: m_1..3, e_1..3 / n, lp_1..3 / n, dg feature
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NOUT)), ("b3", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 5255
SQ2 = np.sqrt(2.0)
GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def unpack(params):
    """Views of the flat parameter vector: W1[in][out], b1, W2, b2, W3, b3."""
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = params[o:o + k].reshape(shape)
        o += k
    return out


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, float32."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    v = unpack(p)
    for name, shape in SHAPES:
        if name.startswith("W"):
            a = 1.0 / np.sqrt(shape[0])
            v[name][...] = rng.uniform(-a, a, shape).astype(np.float32)
    return p


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + GELU_A * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def forward(params, X, emulate=True):
    """Network outputs [n, 7] and the intermediates of the backward pass.
    emulate: round operands to bfloat16 where the device does."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    w = unpack(np.asarray(params, np.float32))
    x = r(X)
    W1, W2, W3 = r(w["W1"]), r(w["W2"]), r(w["W3"])
    z1 = x @ W1 + w["b1"]
    a1 = r(gelu(z1))
    z2 = a1 @ W2 + w["b2"]
    a2 = r(gelu(z2))
    y = a2 @ W3 + w["b3"]
    return y, (x, z1, a1, z2, a2, W1, W2, W3)


def loss_terms(y, R):
    """Per-sample loss and its gradient with respect to the outputs."""
    y = np.asarray(y, np.float64)
    R = np.asarray(R, np.float64)
    m = R[:, 0:3]
    pe = y[:, 0:1] + y[:, 1:2] * m + y[:, 2:3] * m * m
    pl = y[:, 3:4] + y[:, 4:5] * m + y[:, 5:6] * m * m
    re = pe - R[:, 3:6]
    rl = pl - R[:, 6:9]
    rd = y[:, 6] - R[:, 9]
    loss = np.sum(re * re, 1) + np.sum(rl * rl, 1) + rd * rd
    g = np.zeros_like(y)
    g[:, 0], g[:, 1], g[:, 2] = 2 * re.sum(1), 2 * (re * m).sum(1), 2 * (re * m * m).sum(1)
    g[:, 3], g[:, 4], g[:, 5] = 2 * rl.sum(1), 2 * (rl * m).sum(1), 2 * (rl * m * m).sum(1)
    g[:, 6] = 2 * rd
    return loss, g


def loss_and_grad(params, X, R, emulate=True):
    """Mean per-sample loss over the batch and its gradient in the flat
    parameter layout."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    n = X.shape[0]
    y, (x, z1, a1, z2, a2, W1, W2, W3) = forward(params, X, emulate)
    ls, gy = loss_terms(y, R)
    loss = float(ls.sum() / n)
    dy = gy / n
    dyr = r(dy)
    g = {"W3": a2.T @ dyr, "b3": dy.sum(0)}
    dz2 = (dyr @ W3.T) * gelu_grad(z2)
    dz2r = r(dz2)
    g["W2"], g["b2"] = a1.T @ dz2r, dz2.sum(0)
    dz1 = (dz2r @ W2.T) * gelu_grad(z1)
    dz1r = r(dz1)
    g["W1"], g["b1"] = x.T @ dz1r, dz1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return loss, flat


def adam(params, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in place; step counts from 1."""
    m[...] = b1 * m + (1 - b1) * grad
    v[...] = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    params[...] = params - lr * mh / (np.sqrt(vh) + eps)


# ---------------------------------------------------------------------------
# features, targets and the closure
# ---------------------------------------------------------------------------

def _mat(F):
    """Column-major rows of 9 (F[i + 3 j] = F_ij) -> [n, 3, 3] float64."""
    return np.asarray(F, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)


def _flat(M):
    return np.ascontiguousarray(M.transpose(0, 2, 1)).reshape(-1, 9)


def c_minus_i(F):
    """C - I = G + G^T + G^T G with G = F - I (no cancellation near F = I)."""
    G = _mat(F) - np.eye(3)
    return G + G.transpose(0, 2, 1) + G.transpose(0, 2, 1) @ G


def friction(mu, lam, ys):
    """The return mapping's k in dg = |eps_hat| + k tr(eps) (P2G sand branch)."""
    return ys * (3.0 * lam + 2.0 * mu) / (2.0 * mu)


NRM_MIN = 1e-30


def _scale(A):
    return np.maximum(np.sqrt(np.sum(A * A, (1, 2))), NRM_MIN)


def features(F, logJp, mu, lam, ys, coh):
    """[n, 9]: tr U, tr U^3, det U, log(n)/8 + 1, min(n, 1),
    clamp(logJp/n, -8, 8)/8, k, clamp(c/n, -8, 8)/8, lambda/(lambda + 2 mu)."""
    A = c_minus_i(F)
    nrm = _scale(A)
    U = A / nrm[:, None, None]
    U2 = U @ U
    lj = np.asarray(logJp, np.float64) if logJp is not None else np.zeros(A.shape[0])
    f = np.zeros((A.shape[0], NIN))
    f[:, 0] = np.trace(U, axis1=1, axis2=2)
    f[:, 1] = np.sum(U2 * U, (1, 2))
    f[:, 2] = np.linalg.det(U)
    f[:, 3] = np.log(nrm) / 8.0 + 1.0
    f[:, 4] = np.minimum(nrm, 1.0)
    f[:, 5] = np.clip(lj / nrm, -8.0, 8.0) / 8.0
    f[:, 6] = friction(mu, lam, ys)
    f[:, 7] = np.clip(coh / nrm, -8.0, 8.0) / 8.0
    f[:, 8] = lam / (lam + 2.0 * mu)
    return f, nrm


def project(e, logJp, mu, lam, ys, coh):
    """The sand return mapping on principal Hencky strains e [n, 3]:
    projected strains, updated logJp (volume correction on), plastic
    multiplier and case (1 inside, 2 tip, 3 projected)."""
    eps = e - coh
    tre = eps.sum(1)
    tr = tre + logJp
    eh = eps - (tr / 3.0)[:, None]
    ehn = np.sqrt(np.sum(eh * eh, 1))
    dg = ehn + friction(mu, lam, ys) * tr
    tip = tr >= 0
    yl = ~tip & (dg > 0)
    r = np.where(yl, dg / np.where(ehn > 0, ehn, 1.0), 0.0)
    lp = np.where(yl[:, None], e - r[:, None] * eh, e)
    lp = np.where(tip[:, None], coh, lp)
    case = np.where(tip, 2, np.where(yl, 3, 1))
    return lp, tre, dg, case


def targets(F, logJp, mu, lam, ys, coh):
    """[n, 10] target records: eigenvalues m_i of U, e_i / n, lp_i / n and
    clamp(dg / n, -8, 8) / 8 (float64, eigh)."""
    A = c_minus_i(F)
    nrm = _scale(A)
    m = np.linalg.eigvalsh(A / nrm[:, None, None])
    e = 0.5 * np.log1p(nrm[:, None] * m)
    lj = np.asarray(logJp, np.float64) if logJp is not None else np.zeros(A.shape[0])
    lp, _, dg, _ = project(e, lj, mu, lam, ys, coh)
    R = np.zeros((A.shape[0], NREC))
    R[:, 0:3] = m
    R[:, 3:6] = e / nrm[:, None]
    R[:, 6:9] = lp / nrm[:, None]
    R[:, 9] = np.clip(dg / nrm, -8.0, 8.0) / 8.0
    return R


def hencky(F):
    """E = log(C) / 2 from the eigen-decomposition of C (float64)."""
    C = np.eye(3) + c_minus_i(F)
    w, Q = np.linalg.eigh(C)
    return Q @ (0.5 * np.log(np.maximum(w, 1e-300))[:, :, None] * Q.transpose(0, 2, 1))


def expm1_neg(X):
    """exp(-X) - I: degree-6 Taylor of exp(-X/4) - I, then twice
    (I + z)^2 - I = 2 z + z^2."""
    Y = -0.25 * X
    I = np.eye(3)
    z = I + Y / 6.0
    for k in (5.0, 4.0, 3.0, 2.0):
        z = I + (Y @ z) / k
    z = Y @ z
    z = 2.0 * z + z @ z
    return 2.0 * z + z @ z


def newton(X, A, iters=1):
    """Newton steps X += log1p(D), D = sym(exp(-X) (I + A)) - I, log1p to
    degree 3; returns X and |D| of the last step."""
    d = np.zeros(X.shape[0])
    for _ in range(iters):
        Z = expm1_neg(X)
        D = Z + A + Z @ A
        D = 0.5 * (D + D.transpose(0, 2, 1))
        d = np.sqrt(np.sum(D * D, (1, 2)))
        D2 = D @ D
        X = X + D - 0.5 * D2 + (D2 @ D) / 3.0
    return X, d


def residual_bound(d, x0n, an, xn, floor=1e-12):
    """Bound of |X - log C| after the step, relative to |X|; floor: the
    relative rounding level of the arithmetic (float64 here)."""
    y = 0.25 * x0n
    ser = d ** 4 / (4.0 * np.maximum(1.0 - d, 1e-30))
    ex = 4.0 * np.exp(4.0 * y) * y ** 7 / 5040.0 * (1.0 + an)
    return (ser + ex) / np.maximum(xn, NRM_MIN) + floor


def initial_log(y, A):
    """X0 = 2 n (psi0 I + psi1 U + psi2 U^2) = 2 (n psi0 I + psi1 A + psi2 A^2 / n)."""
    nrm = _scale(A)
    y = np.asarray(y, np.float64)
    I = np.eye(3)
    return 2.0 * (nrm[:, None, None] * y[:, 0, None, None] * I + y[:, 1, None, None] * A
                  + (y[:, 2] / nrm)[:, None, None] * (A @ A))


def closure(F, logJp, vol, mu, lam, ys, coh, beta, volcorr, y=None, iters=1, tol=1e-4):
    """vol P F_proj^T (rows of 9, column-major), the updated logJp, the
    residual bound and the accept mask of the closure. y: network outputs
    [n, 7] (None: start from X0 = 0)."""
    A = c_minus_i(F)
    n = A.shape[0]
    X0 = np.zeros((n, 3, 3)) if y is None else initial_log(y, A)
    X, d = newton(X0, A, iters)
    xn = np.sqrt(np.sum(X * X, (1, 2)))
    an = np.sqrt(np.sum(A * A, (1, 2)))
    res = residual_bound(d, np.sqrt(np.sum(X0 * X0, (1, 2))), an, xn)
    detF = np.linalg.det(_mat(F))
    ok = (res <= tol) & (d < 0.5) & (xn <= 1.0) & (detF > 0)
    E = 0.5 * X
    I = np.eye(3)
    lj = np.asarray(logJp, np.float64).copy()
    eps = E - coh * I
    tre = np.trace(eps, axis1=1, axis2=2)
    tr = tre + lj
    eh = eps - (tr / 3.0)[:, None, None] * I
    ehn = np.sqrt(np.sum(eh * eh, (1, 2)))
    dg = ehn + friction(mu, lam, ys) * tr
    tip = tr >= 0
    yl = ~tip & (dg > 0)
    r = np.where(yl, dg / np.where(ehn > 0, ehn, 1.0), 0.0)
    Lp = np.where(yl[:, None, None], E - r[:, None, None] * eh, E)
    Lp = np.where(tip[:, None, None], coh * I, Lp)
    new_lj = np.where(tip, (beta * tre + lj) if volcorr else lj, 0.0)
    trl = np.trace(Lp, axis1=1, axis2=2)
    tau = 2.0 * mu * Lp + lam * trl[:, None, None] * I
    S = tau @ np.linalg.inv(I + A)
    S = 0.5 * (S + S.transpose(0, 2, 1))
    Fm = _mat(F)
    PF = vol * (Fm @ S @ Fm.transpose(0, 2, 1))
    return _flat(PF), new_lj, res, ok
