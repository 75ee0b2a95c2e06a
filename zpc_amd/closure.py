"""The learned Drucker-Prager stress closure: data, training and array entry points.

The network and the closure are described in include/zsmpm.h. Everything
here runs on the GPU through the C ABI (zsmpm_mlp_harvest__rocm,
zsmpm_mlp_train_step__rocm, zsmpm_mlp_infer__rocm, zsmpm_mlp_grad__rocm,
zsmpm_stress__rocm); torch only holds the device buffers.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import rocm

NIN, NOUT, NREC, NPARAMS = rocm.MLP_INPUTS, rocm.MLP_OUTPUTS, rocm.MLP_RECORD, rocm.MLP_PARAMS
_SHAPES = [(NIN, 64), (64,), (64, 64), (64,), (64, NOUT), (NOUT,)]


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and zero biases, in the
    flat layout W1[9][64], b1, W2[64][64], b2, W3[64][7], b3 (float32)."""
    rng = np.random.default_rng(seed)
    out = []
    for shape in _SHAPES:
        if len(shape) == 2:
            a = 1.0 / np.sqrt(shape[0])
            out.append(rng.uniform(-a, a, shape).astype(np.float32).ravel())
        else:
            out.append(np.zeros(shape, np.float32))
    p = np.concatenate(out)
    assert p.size == NPARAMS
    return p


def synthetic_states(n: int, seed: int, lo: float = 1e-5, hi: float = 0.25):
    """Deformation gradients F = I + G (column-major rows of 9) with Gaussian G
    whose scale is log-uniform in [lo, hi], and logJp ~ N(0, (2 scale)^2)."""
    rng = np.random.default_rng(seed)
    scale = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    G = rng.normal(size=(n, 3, 3)) * scale[:, None, None]
    Fm = np.eye(3) + G
    F = np.ascontiguousarray(Fm.transpose(0, 2, 1)).reshape(n, 9).astype(np.float32)
    lj = (rng.normal(size=n) * 2.0 * scale).astype(np.float32)
    return F, lj


def _pol_model(pol: rocm.Policy, mat, model=None):
    return model if model is not None else rocm.model_desc(mat)


def harvest(pol: rocm.Policy, model: rocm.ModelDesc, F, logJp):
    """(features [n, 9], records [n, 10]) device tensors from the plain path."""
    torch = rocm._torch()
    n = F.shape[0]
    X = torch.empty((n, NIN), dtype=torch.float32, device=F.device)
    R = torch.empty((n, NREC), dtype=torch.float32, device=F.device)
    rocm.check(pol.L.zsmpm_mlp_harvest__rocm(pol.h, C.byref(model), F.data_ptr(),
                                             None if logJp is None else logJp.data_ptr(), n, X.data_ptr(),
                                             R.data_ptr()), "mlp_harvest")
    return X, R


def infer(pol: rocm.Policy, params, X):
    torch = rocm._torch()
    Y = torch.empty((X.shape[0], NOUT), dtype=torch.float32, device=X.device)
    rocm.check(pol.L.zsmpm_mlp_infer__rocm(pol.h, params.data_ptr(), X.data_ptr(), X.shape[0], Y.data_ptr()),
               "mlp_infer")
    return Y


def grad(pol: rocm.Policy, params, X, R):
    """(gradient [NPARAMS], loss [1]) device tensors of the mean loss."""
    torch = rocm._torch()
    g = torch.empty(NPARAMS, dtype=torch.float32, device=X.device)
    loss = torch.empty(1, dtype=torch.float32, device=X.device)
    rocm.check(pol.L.zsmpm_mlp_grad__rocm(pol.h, params.data_ptr(), X.data_ptr(), R.data_ptr(), X.shape[0],
                                          g.data_ptr(), loss.data_ptr()), "mlp_grad")
    return g, loss


def stress(pol: rocm.Policy, model: rocm.ModelDesc, F, logJp):
    """vol P F^T rows [n, 9] with the policy's closure; logJp is updated in place."""
    torch = rocm._torch()
    PF = torch.empty_like(F)
    rocm.check(pol.L.zsmpm_stress__rocm(pol.h, C.byref(model), F.data_ptr(),
                                        None if logJp is None else logJp.data_ptr(), F.shape[0], PF.data_ptr()),
               "stress")
    return PF


@dataclass
class Trained:
    params: object            # device float32 [NPARAMS]: the checkpoint with the lowest eval_loss
    losses: List[float] = field(default_factory=list)  # batch loss every `log_every` steps
    seconds: float = 0.0      # GPU training time (harvest and checkpoint evaluations excluded)
    samples: int = 0
    eval_loss: float = float("nan")  # mean loss of `params` on the evaluation sample
    best_step: int = 0        # Adam steps behind `params`
    eval_samples: int = 0
    restores: int = 0         # times a diverging checkpoint was rolled back (rate halved each time)


def eval_loss(pol: rocm.Policy, params, X, R) -> float:
    """Mean loss of params over the samples (X, R)."""
    _, loss = grad(pol, params, X, R)
    return float(loss.item())


def train(pol: rocm.Policy, model: rocm.ModelDesc, F, logJp, steps: int = 2000, batch: int = 8192,
          lr: float = 2e-3, seed: int = 0, log_every: int = 100, params=None, check_every: int = 250,
          eval_max: int = 1 << 16) -> Trained:
    """Adam on minibatches drawn from the states (F, logJp) (device tensors):
    the records come from zsmpm_mlp_harvest__rocm, every step is one
    zsmpm_mlp_train_step__rocm. The learning rate decays by cosine to lr/20.
    Every check_every steps (and at the end) the mean loss over a fixed
    strided sample of at most eval_max states is evaluated, outside the timed
    segments, and the parameters with the lowest one are returned: one bad
    stretch of minibatches cannot leave a worse network behind. When a
    checkpoint's loss exceeds 1.5 times the best, the parameters and Adam
    moments go back to the best checkpoint and the learning rate is halved.
    The Adam step count goes back with them: the bias corrections of later
    steps must match the moments they correct."""
    torch = rocm._torch()
    dev = F.device
    X, R = harvest(pol, model, F, logJp)
    n = X.shape[0]
    P = params.clone() if params is not None else torch.as_tensor(init_params(seed)).to(dev)
    m = torch.zeros_like(P)
    v = torch.zeros_like(P)
    loss = torch.zeros(steps, dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    ev = torch.arange(0, n, max(n // eval_max, 1), device=dev)[:eval_max]
    Xe, Re = X.index_select(0, ev).contiguous(), R.index_select(0, ev).contiguous()
    best, best_P, best_step = eval_loss(pol, P, Xe, Re), P.clone(), 0
    best_m, best_v = m.clone(), v.clone()
    t_adam, best_t = 0, 0  # Adam steps behind (m, v) and behind the best checkpoint's moments
    restores = 0
    scale = 1.0
    secs = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        idx = torch.randint(0, n, (batch,), device=dev, generator=gen)
        xb = X.index_select(0, idx)
        rb = R.index_select(0, idx)
        a = 0.5 * (1.0 + np.cos(np.pi * k / max(steps - 1, 1)))
        rate = lr * scale * (0.05 + 0.95 * a)
        t_adam += 1
        rocm.check(pol.L.zsmpm_mlp_train_step__rocm(pol.h, P.data_ptr(), m.data_ptr(), v.data_ptr(),
                                                    xb.data_ptr(), rb.data_ptr(), batch, float(rate), t_adam,
                                                    loss[k:k + 1].data_ptr()), "mlp_train_step")
        if (k + 1) % check_every == 0 or k + 1 == steps:
            torch.cuda.synchronize()
            secs += time.perf_counter() - t0
            e = eval_loss(pol, P, Xe, Re)
            if e < best:
                best, best_P, best_step = e, P.clone(), k + 1
                best_m, best_v, best_t = m.clone(), v.clone(), t_adam
            elif not (e <= 1.5 * best):  # diverging (or NaN): back to the best checkpoint, half the rate
                P.copy_(best_P)
                m.copy_(best_m)
                v.copy_(best_v)
                t_adam = best_t
                scale *= 0.5
                restores += 1
            torch.cuda.synchronize()
            t0 = time.perf_counter()
    lh = loss.cpu().numpy()
    return Trained(best_P, [float(lh[i]) for i in range(0, steps, log_every)] + [float(lh[-1])], secs, n,
                   eval_loss=best, best_step=best_step, eval_samples=int(ev.numel()), restores=restores)


def training_states(sc, sim_steps: int = 0, synthetic: int = 1 << 18, seed: int = 0, device="cuda"):
    """States to train on: the scene's particles (after sim_steps plain
    substeps) and `synthetic` random states covering large strains."""
    torch = rocm._torch()
    Fs, ljs = [], []
    if sim_steps > 0:
        sim = rocm.Simulator(sc, reorder=True)
        sim.substeps(sim_steps)
        sim.sync()
        Fs.append(sim.parts.F.reshape(-1, 9).clone())
        ljs.append(sim.parts.logJp.clone() if sim.parts.logJp is not None
                   else torch.zeros(sc.n, dtype=torch.float32, device=device))
        del sim
    else:
        Fs.append(torch.as_tensor(np.ascontiguousarray(sc.F.reshape(-1, 9), dtype=np.float32)).to(device))
        lj0 = sc.logJp if sc.logJp is not None else np.zeros(sc.n, np.float32)
        ljs.append(torch.as_tensor(np.ascontiguousarray(lj0, dtype=np.float32)).to(device))
    if synthetic > 0:
        F, lj = synthetic_states(synthetic, seed)
        Fs.append(torch.as_tensor(F).to(device))
        ljs.append(torch.as_tensor(lj).to(device))
    return torch.cat(Fs).contiguous(), torch.cat(ljs).contiguous()
