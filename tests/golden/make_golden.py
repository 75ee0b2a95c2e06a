"""Write tests/golden/refmath.npz from zpc's own math headers.

Runs in the build container only: it loads oracle/_ref/libzpc_refmath.so,
which oracle/Makefile compiles from /root/reference/include (svd_3d, the two
stress functions, lame_parameters, the quadratic LocalArena, hash_combine,
next_2pow, the CFL time step). The inputs are seeded and cover the cases that matter for the
path: random, near-identity, rotation, rank-deficient, repeated singular
values, negative determinant; Drucker-Prager inputs in all three return-map
cases; hash keys that are negative, large and adjacent. Only numbers are
written; the fixture holds inputs and expected outputs.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def _matrices(rng):
    out = []
    out.append(rng.normal(size=(64, 9)))
    out.append((np.eye(3).reshape(9) + 1e-3 * rng.normal(size=(64, 9))))
    # rotations
    q = rng.normal(size=(32, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y),
                  2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x),
                  2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], axis=1)
    out.append(R)
    # rank deficient
    a = rng.normal(size=(32, 3, 1)) @ rng.normal(size=(32, 1, 3))
    out.append(a.reshape(32, 9))
    b = rng.normal(size=(32, 3, 2)) @ rng.normal(size=(32, 2, 3))
    out.append(b.reshape(32, 9))
    # repeated singular values and negative determinant
    for s in ([1.0, 1.0, 1.0], [2.0, 2.0, 0.5], [1.5, 0.7, 0.7], [1.0, 1.0, -1.0], [0.9, 0.8, -0.3]):
        U = np.linalg.qr(rng.normal(size=(16, 3, 3)))[0]
        V = np.linalg.qr(rng.normal(size=(16, 3, 3)))[0]
        M = U @ np.diag(s) @ np.transpose(V, (0, 2, 1))
        out.append(np.transpose(M, (0, 2, 1)).reshape(16, 9))  # column-major storage
    out.append(np.eye(3).reshape(1, 9))
    out.append(np.diag([3.0, -2.0, 0.5]).reshape(1, 9))
    return np.concatenate(out).astype(np.float32)


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libzpc_refmath.so is not built (needs the reference tree)")
    rng = np.random.default_rng(20261202)
    F = _matrices(rng)
    n = F.shape[0]
    U = np.zeros_like(F)
    S = np.zeros((n, 3), np.float32)
    V = np.zeros_like(F)
    R.zref_svd(n, F, U, S, V)

    E_nu = np.array([[1e4, 0.3], [5e4, 0.4], [3.5e5, 0.3], [1e6, 0.45]], np.float32)
    lame = np.zeros_like(E_nu)
    for i, (E, nu) in enumerate(E_nu):
        mu, lam = C.c_float(), C.c_float()
        R.zref_lame(float(E), float(nu), C.byref(mu), C.byref(lam))
        lame[i] = (mu.value, lam.value)

    mu, lam = lame[2]
    vol = np.float32((1 / 512) ** 3 / 8)
    PF_fc = np.zeros_like(F)
    R.zref_stress_fc(n, float(vol), float(mu), float(lam), F, PF_fc)

    # Drucker-Prager: near-identity deformations with logJp spanning all cases
    Fs = (np.eye(3).reshape(9) + 0.05 * rng.normal(size=(96, 9))).astype(np.float32)
    Fs[:32] = (np.eye(3).reshape(9) * 1.03 + 0.005 * rng.normal(size=(32, 9))).astype(np.float32)  # expansion
    lj_in = np.concatenate([np.zeros(32), 0.1 * rng.normal(size=64)]).astype(np.float32)
    lj_out = lj_in.copy()
    PF_sand = np.zeros_like(Fs)
    Fp_sand = np.zeros_like(Fs)
    ys = np.float32(0.816496580927726 * 2 * 0.5 / (3 - 0.5))
    R.zref_stress_sand(Fs.shape[0], float(vol), float(mu), float(lam), 0.0, 1.0, float(ys), 1, lj_out, Fs,
                       PF_sand, Fp_sand.ctypes.data_as(C.c_void_p))

    dx = np.float32(1 / 512)
    x = rng.uniform(0.0, 1.0, size=(128, 3)).astype(np.float32)
    x[:8] = (np.floor(x[:8] / dx) + 0.5) * dx  # exactly on a half-cell
    corner = np.zeros((128, 3), np.int32)
    w = np.zeros((128, 27), np.float32)
    diff = np.zeros((128, 27, 3), np.float32)
    R.zref_arena(128, float(dx), x, corner, w, diff)

    keys = np.concatenate([
        rng.integers(-2 ** 31, 2 ** 31 - 1, size=(64, 3)),
        rng.integers(-300, 300, size=(64, 3)),
        np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, -1, -1], [2 ** 31 - 1] * 3, [-2 ** 31] * 3]),
    ]).astype(np.int32)
    hashes = np.array([R.zref_hash(int(a), int(b), int(c)) for a, b, c in keys], np.int32)
    n2p_in = np.array([1, 2, 3, 5, 17, 1000, 36465, 65536, 65537], np.int64)
    n2p = np.array([R.zref_next_2pow(int(v)) for v in n2p_in], np.int64)

    # evaluate_timestep_linear_elasticity (physics/SoundSpeedCfl.hpp:6-15):
    # the configs' materials and dx values, and a spread of others
    cfl_in = np.array([[1e4, 0.3, 1e3, 1 / 128, 0.5], [5e4, 0.4, 1e3, 1 / 256, 0.5],
                       [3.5e5, 0.3, 1e3, 1 / 512, 0.5], [1e6, 0.45, 2e3, 1 / 64, 0.3],
                       [2e3, 0.1, 5e2, 1 / 32, 0.9], [1e8, 0.2, 8e3, 1 / 1024, 0.6]], np.float32)
    cfl_dt = np.array([R.zref_cfl_dt(*[float(v) for v in row]) for row in cfl_in], np.float32)

    np.savez_compressed(os.path.join(HERE, "refmath.npz"), svd_F=F, svd_U=U, svd_S=S, svd_V=V, E_nu=E_nu,
                        lame=lame, fc_vol=vol, fc_PF=PF_fc, sand_F=Fs, sand_logJp_in=lj_in, sand_logJp_out=lj_out,
                        sand_PF=PF_sand, sand_Fproj=Fp_sand, sand_ys=ys, arena_dx=dx, arena_x=x,
                        arena_corner=corner, arena_w=w, arena_diff=diff, hash_keys=keys, hash_vals=hashes,
                        n2p_in=n2p_in, n2p_out=n2p, cfl_in=cfl_in, cfl_dt=cfl_dt)
    print("wrote", os.path.join(HERE, "refmath.npz"))


if __name__ == "__main__":
    main()
