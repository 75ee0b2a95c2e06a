"""ctypes bindings for the zsmpm oracle: TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of littlemine/zpc's MLS-MPM substep
(``zsmpm_oracle.hpp``); ``_ref/libzpc_refmath.so`` wraps zpc's own math
headers when the reference tree was present at build time. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
this module, always as the checker or the baseline, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libzsmpm_oracle.so")
OMP_PATH = os.path.join(HERE, "lib", "libzsmpm_oracle_omp.so")
REF_PATH = os.path.join(HERE, "_ref", "libzpc_refmath.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class GeneralColliderParams(C.Structure):
    """One general collider of ``zso_params`` (oracle_capi.cpp)."""

    _fields_ = [("type", C.c_int32), ("geometry", C.c_int32), ("axis", C.c_int32),
                ("p0", C.c_double * 3), ("p1", C.c_double * 3), ("R", C.c_double * 9), ("omega", C.c_double * 3),
                ("s", C.c_double), ("dsdt", C.c_double), ("b", C.c_double * 3), ("dbdt", C.c_double * 3)]


def general_collider_params(c) -> GeneralColliderParams:
    """GeneralColliderParams of a scenes.GeneralCollider."""
    g = GeneralColliderParams()
    g.type, g.geometry, g.axis = int(c.kind), int(c.geometry), int(c.axis)
    for d in range(3):
        g.p0[d], g.p1[d] = float(c.p0[d]), float(c.p1[d])
        g.omega[d], g.b[d], g.dbdt[d] = float(c.omega[d]), float(c.b[d]), float(c.dbdt[d])
    for k in range(9):
        g.R[k] = float(c.R[k])
    g.s, g.dsdt = float(c.s), float(c.dsdt)
    return g


class Params(C.Structure):
    """Mirror of ``zso_params`` (oracle_capi.cpp)."""

    _fields_ = [
        ("dt", C.c_double), ("dx", C.c_double), ("extf", C.c_double * 3),
        ("mat_kind", C.c_int32), ("volume", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("cohesion", C.c_double), ("beta", C.c_double), ("yield_surface", C.c_double),
        ("volume_correction", C.c_int32), ("table_size", C.c_int32), ("ncols", C.c_int32),
        ("col_origin", (C.c_double * 3) * 8), ("col_normal", (C.c_double * 3) * 8),
        ("col_type", C.c_int32 * 8),
        ("ngcols", C.c_int32), ("gcols", GeneralColliderParams * 8),
    ]


def build(ref: bool = True) -> None:
    """Compile the oracle (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all" if ref else "lib/libzsmpm_oracle.so"],
                   check=True)


_lib = None
_omp = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.zso_hash.argtypes = [C.c_int32] * 3
        L.zso_hash.restype = C.c_int32
        L.zso_home_slot.argtypes = [C.c_int32] * 4
        L.zso_home_slot.restype = C.c_int32
        L.zso_evaluate_table_size.argtypes = [C.c_int64]
        L.zso_evaluate_table_size.restype = C.c_int64
        L.zso_svd_f32.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p]
        L.zso_svd_f64.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p]
        L.zso_lame_f32.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.zso_lame_f64.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.zso_stress_fc_f32.argtypes = [C.c_int64, C.c_float, C.c_float, C.c_float, _f32p, _f32p]
        L.zso_stress_fc_f64.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, _f64p, _f64p]
        L.zso_stress_sand_f32.argtypes = [C.c_int64] + [C.c_float] * 6 + [C.c_int32, _f32p, _f32p,
                                                                           _f32p, C.c_void_p, C.c_void_p]
        L.zso_stress_sand_f64.argtypes = [C.c_int64] + [C.c_double] * 6 + [C.c_int32, _f64p, _f64p,
                                                                            _f64p, C.c_void_p, C.c_void_p]
        L.zso_arena_f32.argtypes = [C.c_int64, C.c_float, _f32p, _i32p, _f32p, _f32p]
        L.zso_block_keys_f32.argtypes = [C.c_int64, _f32p, C.c_float, C.c_int32, C.c_int32, C.c_float, _i32p]
        for s in ("f32", "f64"):
            getattr(L, f"zso_ctx_new_{s}").restype = C.c_void_p
            getattr(L, f"zso_ctx_free_{s}").argtypes = [C.c_void_p]
            getattr(L, f"zso_ctx_sizes_{s}").argtypes = [C.c_void_p, _i64p]
        L.zso_ctx_table_f32.argtypes = [C.c_void_p, _i32p, _i32p, _i32p, _i32p]
        L.zso_ctx_sort_f32.argtypes = [C.c_void_p, _i32p, _i32p, _i32p]
        L.zso_ctx_grids_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.zso_ctx_grids_f32.restype = C.c_float
        L.zso_ctx_grids_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.zso_ctx_grids_f64.restype = C.c_double
        L.zso_ctx_set_grid_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.zso_ctx_angular_momentum_f32.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.c_int32,
                                                   _f64p]
        L.zso_activate_f32.argtypes = [C.c_void_p, C.c_int32, C.c_int64, _f32p, C.c_float, C.c_int32,
                                       C.c_int32, C.c_float]
        L.zso_enlarge_f32.argtypes = [C.c_void_p, _i32p, _i32p]
        L.zso_bucket_sort_f32.argtypes = [C.c_void_p, C.c_int64, _f32p, C.c_float, C.c_int32, C.c_int32,
                                          C.c_float]
        L.zso_p2g_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, _f32p, _f32p,
                                  _f32p, C.c_void_p]
        L.zso_p2g_f64_on.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f64p, _f64p, _f64p, _f64p,
                                     _f64p, C.c_void_p, _f64p]
        L.zso_p2c2g_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p,
                                    C.c_void_p]
        L.zso_p2c2g_f64_on.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f64p, _f64p, _f64p, _f64p,
                                       _f64p, C.c_void_p, _f64p]
        L.zso_g2c2p_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, _f32p, _f32p]
        L.zso_g2c2p_f64_on.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f64p, _f64p, _f64p, _f64p,
                                       _f64p]
        L.zso_grid_update_f32.argtypes = [C.c_void_p, C.POINTER(Params)]
        L.zso_grid_update_f32.restype = C.c_float
        L.zso_g2p_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, _f32p, _f32p]
        L.zso_run_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_int64, _f32p, _f32p,
                                  _f32p, _f32p, _f32p, C.c_void_p]
        L.zso_run_f64.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_int64, _f64p, _f64p,
                                  _f64p, _f64p, _f64p, C.c_void_p]
        L.zso_run_sorted_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_int64, _f32p, _f32p,
                                         _f32p, _f32p, _f32p, C.c_void_p, _i32p]
        L.zso_run_sorted_f64.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_int64, _f64p, _f64p,
                                         _f64p, _f64p, _f64p, C.c_void_p, _i32p]
        L.zso_reorder_f32.argtypes = [C.c_int64, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_void_p, _i32p]
        for name in ("zso_level_set_sdf_f64", "zso_resolve_general_f64"):
            getattr(L, name).argtypes = [C.POINTER(Params), C.c_int32, C.c_int64, _f64p, _f64p]
        L.zso_level_set_normal_f64.argtypes = [C.POINTER(Params), C.c_int32, C.c_int64, _f64p, _f64p]
        L.zso_resolve_general_f32.argtypes = [C.POINTER(Params), C.c_int32, C.c_int64, _f32p, _f32p]
        L.zso_g2p2g_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, C.c_void_p, _f32p,
                                    _f32p]
        L.zso_implicit_multiply_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, C.c_void_p,
                                                _f32p, _f32p]
        L.zso_implicit_project_f32.argtypes = [C.c_void_p, C.POINTER(Params), _f32p]
        L.zso_implicit_precondition_f32.argtypes = [C.c_void_p, _f32p, _f32p]
        L.zso_implicit_solve_f32.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int64, _f32p, _f32p, C.c_void_p,
                                             _f32p, _f32p, C.c_int32, C.c_float, C.c_float, C.POINTER(C.c_float)]
        L.zso_implicit_solve_f32.restype = C.c_int32
        _u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        L.zso_mt19937.argtypes = [C.c_uint32, C.c_int64, _u32p]
        L.zso_bht_hash_params.argtypes = [C.c_uint32, _u32p]
        L.zso_bht_hash.argtypes = [C.c_uint32, C.c_uint32, C.c_int64, _i32p, _u32p]
        L.zso_bht_table_size.argtypes = [C.c_int64]
        L.zso_bht_table_size.restype = C.c_int64
        L.zso_bht_insert.argtypes = [C.c_int32, _u32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, C.c_int64, _i32p,
                                     _i32p]
        L.zso_bht_query.argtypes = [C.c_int32, _u32p, _i32p, _i32p, C.c_int64, _i32p, _i32p]
        L.zso_spg_block_keys.argtypes = [_f32p, C.c_int64, _f32p, _i32p]
        spg = [C.c_int32, _u32p, _i32p, _i32p, C.c_int32, C.c_int32, _f32p, _f32p, _f32p, C.c_float]
        L.zso_spg_wsample.argtypes = spg + [C.c_int32, C.c_int64, _f32p, _f32p]
        L.zso_spg_normal.argtypes = spg + [C.c_int32, C.c_int64, _f32p, _f32p]
        L.zso_levelset_resolve.argtypes = spg + [C.c_int32, C.c_int32, C.c_int32, C.c_int64, _f32p, _f32p, _i32p]
        _lib = L
    return _lib


def omp():
    global _omp
    if _omp is None:
        if not os.path.exists(OMP_PATH):
            build(ref=False)
            subprocess.run(["make", "-s", "-C", HERE, "lib/libzsmpm_oracle_omp.so"], check=True)
        L = C.CDLL(OMP_PATH)
        L.zso_omp_default_threads.restype = C.c_int32
        L.zso_omp_run_f32.argtypes = [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                                      _f32p, _f32p, _f32p, _f32p, _f32p, C.c_void_p, C.POINTER(C.c_int64)]
        L.zso_omp_run_f32.restype = C.c_double
        _omp = L
    return _omp


def ref():
    """zpc's own math headers behind the shim, or None when not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        L = C.CDLL(REF_PATH)
        L.zref_svd.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p]
        L.zref_lame.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.zref_stress_fc.argtypes = [C.c_int64, C.c_float, C.c_float, C.c_float, _f32p, _f32p]
        L.zref_stress_sand.argtypes = [C.c_int64] + [C.c_float] * 6 + [C.c_int32, _f32p, _f32p, _f32p,
                                                                       C.c_void_p]
        L.zref_arena.argtypes = [C.c_int64, C.c_float, _f32p, _i32p, _f32p, _f32p]
        L.zref_hash.argtypes = [C.c_int32] * 3
        L.zref_hash.restype = C.c_int32
        L.zref_next_2pow.argtypes = [C.c_int64]
        L.zref_next_2pow.restype = C.c_int64
        L.zref_matmul3.argtypes = [_f32p, _f32p, _f32p]
        L.zref_mat_diag_matT.argtypes = [_f32p, _f32p, _f32p, _f32p]
        L.zref_cfl_dt.argtypes = [C.c_float] * 5
        L.zref_cfl_dt.restype = C.c_float
        _u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        if hasattr(L, "zref_universal_hash"):
            L.zref_universal_hash.argtypes = [C.c_uint32, C.c_uint32, C.c_int64, _i32p, _u32p]
            L.zref_mt19937.argtypes = [C.c_uint32, C.c_int64, _u32p]
        _ref = L
    return _ref


# ----------------------------------------------------------------------------- helpers

def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_params(scene, dtype=np.float32) -> Params:
    """Build the C params block from a scene description (zpc_amd.scenes.Scene)."""
    p = Params()
    p.dt = float(scene.dt)
    p.dx = float(scene.dx)
    for d in range(3):
        p.extf[d] = float(scene.gravity[d])
    m = scene.material
    p.mat_kind = int(m.kind)
    p.volume = float(m.volume)
    p.E = float(m.E)
    p.nu = float(m.nu)
    p.cohesion = float(m.cohesion)
    p.beta = float(m.beta)
    p.yield_surface = float(m.yield_surface)
    p.volume_correction = int(m.volume_correction)
    p.table_size = int(scene.table_size)
    p.ncols = len(scene.colliders)
    for i, c in enumerate(scene.colliders):
        for d in range(3):
            p.col_origin[i][d] = float(c.origin[d])
            p.col_normal[i][d] = float(c.normal[d])
        p.col_type[i] = int(c.kind)
    gen = list(getattr(scene, "general_colliders", None) or [])
    p.ngcols = len(gen)
    for i, c in enumerate(gen):
        p.gcols[i] = general_collider_params(c)
    return p


def svd(F: np.ndarray):
    F = np.ascontiguousarray(F, dtype=np.float32).reshape(-1, 9)
    n = F.shape[0]
    U = np.zeros((n, 9), np.float32)
    S = np.zeros((n, 3), np.float32)
    V = np.zeros((n, 9), np.float32)
    lib().zso_svd_f32(n, F, U, S, V)
    return U, S, V


def svd64(F: np.ndarray):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 9)
    n = F.shape[0]
    U = np.zeros((n, 9))
    S = np.zeros((n, 3))
    V = np.zeros((n, 9))
    lib().zso_svd_f64(n, F, U, S, V)
    return U, S, V


def lame(E, nu):
    mu, lam = C.c_float(), C.c_float()
    lib().zso_lame_f32(E, nu, C.byref(mu), C.byref(lam))
    return mu.value, lam.value


def stress_fc(vol, mu, lam, F):
    F = np.ascontiguousarray(F, dtype=np.float32).reshape(-1, 9)
    PF = np.zeros_like(F)
    lib().zso_stress_fc_f32(F.shape[0], vol, mu, lam, F, PF)
    return PF


def _one_collider(c) -> Params:
    p = Params()
    p.ngcols = 1
    p.gcols[0] = general_collider_params(c)
    return p


def level_set_sdf(c, X):
    """phi of the general collider c (scenes.GeneralCollider) at material-space points X [n, 3], float64."""
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 3)
    out = np.zeros(X.shape[0])
    lib().zso_level_set_sdf_f64(C.byref(_one_collider(c)), 0, X.shape[0], X, out)
    return out


def level_set_normal(c, X):
    """Material-space normals of the general collider c at X [n, 3], float64."""
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 3)
    out = np.zeros_like(X)
    lib().zso_level_set_normal_f64(C.byref(_one_collider(c)), 0, X.shape[0], X, out)
    return out


def resolve_general(c, x, v, dtype=np.float64):
    """Collider::resolveCollision of c on world-space nodes x [n, 3] with velocities v [n, 3]; returns the new v."""
    x = np.ascontiguousarray(x, dtype).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype).reshape(-1, 3).copy()
    fn = lib().zso_resolve_general_f64 if dtype == np.float64 else lib().zso_resolve_general_f32
    fn(C.byref(_one_collider(c)), 0, x.shape[0], x, v)
    return v


def stress_sand(vol, mu, lam, coh, beta, ys, volcorr, logJp, F):
    F = np.ascontiguousarray(F, dtype=np.float32).reshape(-1, 9)
    lj = np.ascontiguousarray(logJp, dtype=np.float32).copy()
    PF = np.zeros_like(F)
    Fp = np.zeros_like(F)
    which = np.zeros(F.shape[0], np.int32)
    lib().zso_stress_sand_f32(F.shape[0], vol, mu, lam, coh, beta, ys, int(volcorr), lj, F, PF, _ptr(Fp),
                              _ptr(which))
    return PF, lj, Fp, which


def stress_sand64(vol, mu, lam, coh, beta, ys, volcorr, logJp, F):
    """The float64 twin of stress_sand (Jacobi SVD)."""
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 9)
    lj = np.ascontiguousarray(logJp, dtype=np.float64).copy()
    PF = np.zeros_like(F)
    Fp = np.zeros_like(F)
    which = np.zeros(F.shape[0], np.int32)
    lib().zso_stress_sand_f64(F.shape[0], vol, mu, lam, coh, beta, ys, int(volcorr), lj, F, PF, _ptr(Fp),
                              _ptr(which))
    return PF, lj, Fp, which


def arena(dx, x):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    corner = np.zeros((n, 3), np.int32)
    w = np.zeros((n, 27), np.float32)
    diff = np.zeros((n, 27, 3), np.float32)
    lib().zso_arena_f32(n, dx, x, corner, w, diff)
    return corner, w, diff


def block_keys(x, dx, block_len=4, offset=-2, disp=0.5):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    k = np.zeros((x.shape[0], 3), np.int32)
    lib().zso_block_keys_f32(x.shape[0], x, dx, block_len, offset, disp, k)
    return k


class Context:
    """One oracle substep context (f32 or f64); keeps the last substep's outputs."""

    def __init__(self, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.sfx = "f32" if self.dtype == np.float32 else "f64"
        self.L = lib()
        self.h = getattr(self.L, f"zso_ctx_new_{self.sfx}")()

    def __del__(self):
        try:
            getattr(self.L, f"zso_ctx_free_{self.sfx}")(self.h)
        except Exception:
            pass

    def sizes(self):
        s = np.zeros(5, np.int64)
        getattr(self.L, f"zso_ctx_sizes_{self.sfx}")(self.h, s)
        return {"T": int(s[0]), "cnt": int(s[1]), "n": int(s[2]), "grid": int(s[3]),
                "quirk": bool(s[4])}

    # stage-level
    def activate(self, T, x, dx, block_len=4, offset=-2, disp=0.5):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        self.L.zso_activate_f32(self.h, T, x.shape[0], x, dx, block_len, offset, disp)

    def enlarge(self, lo=(0, 0, 0), hi=(2, 2, 2)):
        self.L.zso_enlarge_f32(self.h, np.array(lo, np.int32), np.array(hi, np.int32))

    def bucket_sort(self, x, dx, block_len=4, offset=-2, disp=0.5):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        self.L.zso_bucket_sort_f32(self.h, x.shape[0], x, dx, block_len, offset, disp)

    def table(self):
        s = self.sizes()
        T, cnt = s["T"], s["cnt"]
        keys = np.zeros((T, 3), np.int32)
        idx = np.zeros(T, np.int32)
        st = np.zeros(T, np.int32)
        act = np.zeros((max(cnt, 1), 3), np.int32)
        self.L.zso_ctx_table_f32(self.h, keys, idx, st, act)
        return {"T": T, "cnt": cnt, "keys": keys, "indices": idx, "status": st,
                "active": act[:cnt], "quirk": s["quirk"]}

    def sort(self):
        s = self.sizes()
        cnt, n = s["cnt"], s["n"]
        counts = np.zeros(cnt + 1, np.int32)
        offsets = np.zeros(cnt + 1, np.int32)
        indices = np.zeros(max(n, 1), np.int32)
        self.L.zso_ctx_sort_f32(self.h, counts, offsets, indices)
        return counts, offsets, indices[:n]

    def grids(self):
        s = self.sizes()
        g = np.zeros(s["grid"], self.dtype)
        gv = np.zeros(s["grid"], self.dtype)
        mv = getattr(self.L, f"zso_ctx_grids_{self.sfx}")(self.h, _ptr(g), _ptr(gv))
        return g.reshape(-1, 7, 64), gv.reshape(-1, 7, 64), float(mv)

    def set_grids(self, grid_p2g=None, grid_v=None):
        """Replace the P2G grid and/or the updated grid (table.cnt x 7 x 64)."""
        n = self.sizes()["cnt"] * 7 * 64
        a = None if grid_p2g is None else np.ascontiguousarray(grid_p2g, np.float32).reshape(-1)
        b = None if grid_v is None else np.ascontiguousarray(grid_v, np.float32).reshape(-1)
        assert a is None or a.size == n
        assert b is None or b.size == n
        self.L.zso_ctx_set_grid_f32(self.h, _ptr(a), _ptr(b))

    def angular_momentum(self, dx, which=0, m_channel=0, mv_channel=1):
        """GridAngularMomentum (GridOp.hpp:216-259) of the P2G grid (which=0)
        or the updated grid (which=1): [x cross mv (3), mv (3)] in float64."""
        out = np.zeros(6)
        self.L.zso_ctx_angular_momentum_f32(self.h, which, dx, m_channel, mv_channel, out)
        return out

    def p2g(self, params, ps):
        self.L.zso_p2g_f32(self.h, C.byref(params), ps.n, ps.x, ps.v, ps.C, ps.F, ps.m, _ptr(ps.logJp))

    def p2g_f64(self, params, ps64):
        """FP64 P2G of double particles on this (f32) context's table."""
        n = ps64.n
        g = np.zeros(self.sizes()["cnt"] * 7 * 64)
        self.L.zso_p2g_f64_on(self.h, C.byref(params), n, ps64.x, ps64.v, ps64.C, ps64.F, ps64.m,
                              _ptr(ps64.logJp), g)
        return g.reshape(-1, 7, 64)

    # particle -> cell -> grid transfers (P2C2G / G2C2P) on this context's table
    def p2c2g(self, params, ps):
        """P2C2G into a cleaned grid, kept as this context's P2G grid."""
        self.L.zso_p2c2g_f32(self.h, C.byref(params), ps.n, ps.x, ps.v, ps.C, ps.F, ps.m, _ptr(ps.logJp))

    def p2c2g_f64(self, params, ps64):
        """FP64 P2C2G of double particles on this (f32) context's table."""
        g = np.zeros(self.sizes()["cnt"] * 7 * 64)
        self.L.zso_p2c2g_f64_on(self.h, C.byref(params), ps64.n, ps64.x, ps64.v, ps64.C, ps64.F, ps64.m,
                                _ptr(ps64.logJp), g)
        return g.reshape(-1, 7, 64)

    def g2c2p(self, params, ps):
        """G2C2P (pre, per-cell gather, post) on this context's updated grid."""
        self.L.zso_g2c2p_f32(self.h, C.byref(params), ps.n, ps.x, ps.v, ps.C, ps.F)

    def g2c2p_f64(self, params, ps64, gridv):
        """FP64 G2C2P of double particles with a double node-velocity grid."""
        g = np.ascontiguousarray(gridv, np.float64).reshape(-1)
        assert g.size == self.sizes()["cnt"] * 7 * 64
        self.L.zso_g2c2p_f64_on(self.h, C.byref(params), ps64.n, ps64.x, ps64.v, ps64.C, ps64.F, g)

    def grid_update(self, params):
        return self.L.zso_grid_update_f32(self.h, C.byref(params))

    # implicit step on this context's table and P2G grid (vectors: 3 floats per node)
    def g2p2g(self, params, ps, v):
        r = np.zeros(self.sizes()["cnt"] * 64 * 3, np.float32)
        self.L.zso_g2p2g_f32(self.h, C.byref(params), ps.n, ps.x, ps.F, _ptr(ps.logJp),
                             np.ascontiguousarray(v, np.float32), r)
        return r

    def implicit_multiply(self, params, ps, vin):
        out = np.zeros(self.sizes()["cnt"] * 64 * 3, np.float32)
        self.L.zso_implicit_multiply_f32(self.h, C.byref(params), ps.n, ps.x, ps.F, _ptr(ps.logJp),
                                         np.ascontiguousarray(vin, np.float32), out)
        return out

    def implicit_project(self, params, dof):
        d = np.ascontiguousarray(dof, np.float32).copy()
        self.L.zso_implicit_project_f32(self.h, C.byref(params), d)
        return d

    def implicit_precondition(self, vin, vout):
        o = np.ascontiguousarray(vout, np.float32).copy()
        self.L.zso_implicit_precondition_f32(self.h, np.ascontiguousarray(vin, np.float32), o)
        return o

    def implicit_solve(self, params, ps, x0, b, max_iters=1000, tol=1e-6, rel_tol=0.5):
        x = np.ascontiguousarray(x0, np.float32).copy()
        res = C.c_float(0)
        it = self.L.zso_implicit_solve_f32(self.h, C.byref(params), ps.n, ps.x, ps.F, _ptr(ps.logJp), x,
                                           np.ascontiguousarray(b, np.float32), max_iters, tol, rel_tol,
                                           C.byref(res))
        return x, int(it), float(res.value)

    def g2p(self, params, ps):
        self.L.zso_g2p_f32(self.h, C.byref(params), ps.n, ps.x, ps.v, ps.C, ps.F)

    def run(self, params, ps, steps, reorder=False):
        """Run `steps` substeps in place. With reorder, the particles (and
        ps.ids) are gathered into bucket order after each substep's G2P."""
        if reorder:
            fn = getattr(self.L, f"zso_run_sorted_{self.sfx}")
            fn(self.h, C.byref(params), steps, ps.n, ps.x, ps.v, ps.C, ps.F, ps.m, _ptr(ps.logJp), ps.ids)
            return
        fn = getattr(self.L, f"zso_run_{self.sfx}")
        fn(self.h, C.byref(params), steps, ps.n, ps.x, ps.v, ps.C, ps.F, ps.m, _ptr(ps.logJp))


class HostParticles:
    """Host particle state in zpc's attribute layout (x, v, C, F, m, logJp)."""

    def __init__(self, x, v, Cm, F, m, logJp, dtype=np.float32):
        dt = np.dtype(dtype)
        self.n = int(np.asarray(x).reshape(-1, 3).shape[0])
        self.x = np.ascontiguousarray(x, dt).reshape(-1).copy()
        self.v = np.ascontiguousarray(v, dt).reshape(-1).copy()
        self.C = np.ascontiguousarray(Cm, dt).reshape(-1).copy()
        self.F = np.ascontiguousarray(F, dt).reshape(-1).copy()
        self.m = np.ascontiguousarray(m, dt).reshape(-1).copy()
        self.logJp = None if logJp is None else np.ascontiguousarray(logJp, dt).reshape(-1).copy()
        self.ids = np.arange(self.n, dtype=np.int32)

    def astype(self, dtype):
        out = HostParticles(self.x, self.v, self.C, self.F, self.m, self.logJp, dtype)
        out.ids = self.ids.copy()
        return out

    def reorder(self, order):
        """Vector::reorder gather of every attribute (and ids) by `order`."""
        order = np.ascontiguousarray(order, np.int32)
        assert self.x.dtype == np.float32 and order.shape == (self.n,)
        lib().zso_reorder_f32(self.n, order, self.x, self.v, self.C, self.F, self.m, _ptr(self.logJp), self.ids)

    def by_id(self, name):
        """Attribute `name` in particle-id order."""
        a = getattr(self, name)
        w = a.size // self.n
        out = np.empty_like(a).reshape(self.n, w)
        out[self.ids] = a.reshape(self.n, w)
        return out.reshape(-1)


def omp_run(params, ps: HostParticles, warmup: int, steps: int, threads: int = 0):
    """Time the OpenMP baseline; returns (seconds, threads_used, nblocks)."""
    L = omp()
    nt = threads if threads > 0 else L.zso_omp_default_threads()
    nb = C.c_int64(0)
    secs = L.zso_omp_run_f32(C.byref(params), warmup, steps, nt, ps.n, ps.x, ps.v, ps.C, ps.F, ps.m,
                             _ptr(ps.logJp), C.byref(nb))
    return float(secs), int(nt), int(nb.value)


# ----------------------------------------------------------------------------- bht / SparseGrid

BHT_SEED = 2  # std::mt19937 rng(2) in bht's constructor (container/Bht.hpp:160-166)
BHT_SENTINEL = 0x3F3F3F3F
BHT_FAILURE = -(2 ** 31)


def mt19937(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().zso_mt19937(seed, n, out)
    return out


def bht_hash_params(seed: int = BHT_SEED) -> np.ndarray:
    out = np.zeros(6, np.uint32)
    lib().zso_bht_hash_params(seed, out)
    return out


def bht_hash(hx: int, hy: int, keys) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.int32).reshape(-1, 3)
    out = np.zeros(len(k), np.uint32)
    lib().zso_bht_hash(hx, hy, len(k), k, out)
    return out


def bht_table_size(n: int) -> int:
    return int(lib().zso_bht_table_size(n))


class HostBht:
    """bht<i32, 3, int, 16> on the host with zpc's sequential insert."""

    def __init__(self, table_size: int, hash_params=None):
        T = int(table_size)
        self.T = T
        self.h = np.ascontiguousarray(bht_hash_params() if hash_params is None else hash_params, np.uint32)
        self.keys = np.full(4 * T, BHT_SENTINEL, np.int32)
        self.indices = np.full(T, -1, np.int32)
        self.status = np.full(T, -1, np.int32)
        self.cnt = np.zeros(1, np.int32)
        self.active = np.zeros(3 * T, np.int32)
        self.success = np.ones(1, np.int32)

    def insert(self, keys) -> np.ndarray:
        k = np.ascontiguousarray(keys, np.int32).reshape(-1)
        out = np.zeros(len(k) // 3, np.int32)
        lib().zso_bht_insert(self.T, self.h, self.keys, self.indices, self.status, self.cnt, self.active,
                             self.success, len(out), k, out)
        return out

    def query(self, keys) -> np.ndarray:
        k = np.ascontiguousarray(keys, np.int32).reshape(-1)
        out = np.zeros(len(k) // 3, np.int32)
        lib().zso_bht_query(self.T, self.h, self.keys, self.indices, len(out), k, out)
        return out


class HostSparseGrid:
    """SparseGrid<3, f32, 8>: a HostBht of block origins, (blocks, nch, 512)
    cell values, the index-to-world transform (4 x 4, row-vector convention)
    and the background value."""

    def __init__(self, table: HostBht, data, i2w, background: float = 0.0):
        self.table = table
        self.data = np.ascontiguousarray(data, np.float32)
        self.nblocks, self.nch = self.data.shape[0], self.data.shape[1]
        self.i2w = np.ascontiguousarray(i2w, np.float32).reshape(16)
        self.w2i = np.linalg.inv(np.asarray(i2w, np.float64).reshape(4, 4)).astype(np.float32).reshape(16)
        self.background = float(background)

    def _args(self):
        t = self.table
        return (t.T, t.h, t.keys, t.indices, self.nblocks, self.nch, self.data.reshape(-1), self.i2w, self.w2i,
                self.background)

    def wsample(self, chn: int, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        out = np.zeros(len(x) // 3, np.float32)
        lib().zso_spg_wsample(*self._args(), chn, len(out), x, out)
        return out

    def normal(self, sdf: int, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        out = np.zeros(len(x), np.float32)
        lib().zso_spg_normal(*self._args(), sdf, len(x) // 3, x, out)
        return out.reshape(-1, 3)

    def resolve(self, sdf: int, vchn: int, kind: int, x, v):
        """The level-set boundary on nodes at x with velocities v: (v', inside)."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        v = np.ascontiguousarray(v, np.float32).reshape(-1).copy()
        inside = np.zeros(len(x) // 3, np.int32)
        lib().zso_levelset_resolve(*self._args(), sdf, vchn, kind, len(inside), x, v, inside)
        return v.reshape(-1, 3), inside.astype(bool)


def spg_block_keys(w2i, x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    out = np.zeros(len(x), np.int32)
    lib().zso_spg_block_keys(np.ascontiguousarray(w2i, np.float32).reshape(16), len(x) // 3, x, out)
    return out.reshape(-1, 3)
