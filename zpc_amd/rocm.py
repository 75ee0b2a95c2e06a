"""ctypes bindings of the zsmpm C ABI (include/zsmpm.h) and device containers.

Device buffers are torch tensors on the HIP device (torch is only the
allocator and stream provider here); every computation runs in the hand-written
gfx950 kernels of libzsmpm_rocm.so. The containers keep zpc's layouts:
HashTable<i32,3,int>, Grids<f32,3,4>{m,v,rhs}, Particles<f32,3> attributes and
IndexBuckets. Loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzsmpm_rocm.so")

ABI_MAJOR = 1
ABI_MINOR = 0

RESULTS = {
    0: "OK", -1: "ERROR", -2: "INCOMPATIBLE_MAJOR", -3: "INSUFFICIENT_SIZE", -4: "UNSUPPORTED_OPERATION",
    -5: "INVALID_ARGUMENT", -6: "DEVICE_ERROR", -7: "TABLE_FULL", -8: "KEY_RANGE", -9: "STENCIL_OUTSIDE",
    -10: "MISSING_BLOCK", -11: "CAPACITY", -12: "IO_ERROR",
}


class ZsmpmError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {RESULTS.get(code, code)} ({code})")
        self.code = code


class Header(C.Structure):
    _fields_ = [("size", C.c_uint32), ("major", C.c_uint16), ("minor", C.c_uint16)]


def _hdr(struct_type):
    return Header(C.sizeof(struct_type), ABI_MAJOR, ABI_MINOR)


class TableDesc(C.Structure):
    _fields_ = [("hdr", Header), ("table_size", C.c_int32), ("keys", C.c_void_p), ("indices", C.c_void_p),
                ("status", C.c_void_p), ("cnt", C.c_void_p), ("active_keys", C.c_void_p)]


class ParticlesDesc(C.Structure):
    _fields_ = [("hdr", Header), ("n", C.c_int64), ("x", C.c_void_p), ("v", C.c_void_p), ("C", C.c_void_p),
                ("F", C.c_void_p), ("m", C.c_void_p), ("logJp", C.c_void_p), ("id", C.c_void_p)]


class GridDesc(C.Structure):
    _fields_ = [("hdr", Header), ("dx", C.c_float), ("capacity_blocks", C.c_int32), ("data", C.c_void_p)]


class BucketsDesc(C.Structure):
    _fields_ = [("hdr", Header), ("capacity_blocks", C.c_int32), ("counts", C.c_void_p),
                ("offsets", C.c_void_p), ("indices", C.c_void_p)]


class ModelDesc(C.Structure):
    _fields_ = [("hdr", Header), ("kind", C.c_int32), ("volume", C.c_float), ("E", C.c_float), ("nu", C.c_float),
                ("cohesion", C.c_float), ("beta", C.c_float), ("yield_surface", C.c_float),
                ("volume_correction", C.c_int32), ("svd_fp64_gap", C.c_float)]


class ClosureDesc(C.Structure):
    _fields_ = [("hdr", Header), ("params", C.c_void_p), ("tol", C.c_float), ("flags", C.c_int32)]


# learned closure network sizes (include/zsmpm.h)
MLP_INPUTS, MLP_HIDDEN, MLP_OUTPUTS, MLP_RECORD, MLP_PARAMS = 9, 64, 7, 10, 5255

MAX_COLLIDERS = 8
MAX_SCENES = 256
SCENE_KEY_STRIDE = 1024  # a batch shifts scene s's block keys by s * 1024 along axis 0


class PlaneCollider(C.Structure):
    _fields_ = [("type", C.c_int32), ("origin", C.c_float * 3), ("normal", C.c_float * 3)]


class ColliderDesc(C.Structure):
    """zsmpm_collider: an analytic level-set collider with its rigid motion."""

    _fields_ = [("type", C.c_int32), ("geometry", C.c_int32), ("axis", C.c_int32), ("p0", C.c_float * 3),
                ("p1", C.c_float * 3), ("R", C.c_float * 9), ("omega", C.c_float * 3), ("s", C.c_float),
                ("dsdt", C.c_float), ("b", C.c_float * 3), ("dbdt", C.c_float * 3)]


def collider_desc(c) -> ColliderDesc:
    """ColliderDesc of a scenes.GeneralCollider."""
    d = ColliderDesc()
    d.type, d.geometry, d.axis = int(c.kind), int(c.geometry), int(c.axis)
    for k in range(3):
        d.p0[k], d.p1[k] = float(c.p0[k]), float(c.p1[k])
        d.omega[k], d.b[k], d.dbdt[k] = float(c.omega[k]), float(c.b[k]), float(c.dbdt[k])
    for k in range(9):
        d.R[k] = float(c.R[k])
    d.s, d.dsdt = float(c.s), float(c.dsdt)
    return d


class BatchSceneDesc(C.Structure):
    _fields_ = [("hdr", Header), ("first", C.c_int64), ("count", C.c_int64), ("model", ModelDesc),
                ("dx", C.c_float), ("dt", C.c_float), ("extf", C.c_float * 3), ("ncolliders", C.c_int32),
                ("colliders", PlaneCollider * MAX_COLLIDERS)]


COLUMN_MAJOR, ROW_MAJOR = 0, 1


class BhtDesc(C.Structure):
    """zsmpm_bht: bht<i32, 3, int, 16>."""

    _fields_ = [("hdr", Header), ("table_size", C.c_int32), ("keys", C.c_void_p), ("indices", C.c_void_p),
                ("status", C.c_void_p), ("cnt", C.c_void_p), ("active_keys", C.c_void_p), ("success", C.c_void_p),
                ("hash", C.c_uint32 * 6)]


class SparseGridDesc(C.Structure):
    """zsmpm_sparse_grid: SparseGrid<3, f32, 8>."""

    _fields_ = [("hdr", Header), ("table", BhtDesc), ("data", C.c_void_p), ("num_channels", C.c_int32),
                ("capacity_blocks", C.c_int32), ("i2w", C.c_float * 16), ("w2i", C.c_float * 16),
                ("background", C.c_float)]


TRANSFER_APIC, TRANSFER_CELL = 0, 1
BHT_SEED = 2
BHT_SERIAL = 1


class TileParticlesDesc(C.Structure):
    """zsmpm_tile_particles: particles in a TileVector<f32, L>."""

    _fields_ = [("hdr", Header), ("n", C.c_int64), ("lane_width", C.c_int32), ("num_channels", C.c_int32),
                ("data", C.c_void_p), ("x", C.c_int32), ("v", C.c_int32), ("C", C.c_int32), ("F", C.c_int32),
                ("m", C.c_int32), ("logJp", C.c_int32), ("id", C.c_int32), ("matrix_order", C.c_int32)]


class CheckpointInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("table_size", C.c_int32), ("fields", C.c_int32), ("time", C.c_double),
                ("step", C.c_int64)]


_lib = None

# zsmpm_stage order
STAGES = ("table_clear", "activate", "enlarge", "bucket_sort", "grid_clear", "p2g", "grid_update", "g2p", "stress")


def lib():
    """Load libzsmpm_rocm.so (built by zpc_amd.build) and declare its ABI."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m zpc_amd.build` "
                           "(or __graft_entry__.build()) before using the HIP path")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    P = C.POINTER
    sig = {
        "zsmpm_abi_version": [P(Header)],
        "zsmpm_policy__rocm": [i32, vp, P(vp)],
        "zsmpm_del_policy__rocm": [vp],
        "zsmpm_policy_reserve__rocm": [vp, i64, i32],
        "zsmpm_policy_sync": [vp],
        "zsmpm_policy_counters": [vp, P(C.c_uint32)],
        "zsmpm_policy_profile": [vp, i32],
        "zsmpm_policy_stage_times": [vp, P(C.c_double), P(C.c_uint64), i32, i32],
        "zsmpm_table_clear__rocm": [vp, P(TableDesc)],
        "zsmpm_table_activate__rocm": [vp, P(TableDesc), vp, i64, f32, i32, i32, f32],
        "zsmpm_table_activate_noalloc__rocm": [vp, P(TableDesc), vp, i64, f32, i32, i32, f32],
        "zsmpm_table_enlarge__rocm": [vp, P(TableDesc), P(C.c_int32 * 3), P(C.c_int32 * 3)],
        "zsmpm_table_query__rocm": [vp, P(TableDesc), vp, i64, vp],
        "zsmpm_bucket_sort__rocm": [vp, P(TableDesc), vp, i64, f32, i32, i32, f32, P(BucketsDesc)],
        "zsmpm_grid_clear__rocm": [vp, P(GridDesc), P(TableDesc)],
        "zsmpm_p2g__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(TableDesc), P(BucketsDesc), P(GridDesc), f32],
        "zsmpm_grid_update__rocm": [vp, P(GridDesc), P(TableDesc), f32, P(C.c_float * 3), P(PlaneCollider), i32,
                                    i32, vp],
        "zsmpm_grid_angular_momentum__rocm": [vp, P(GridDesc), P(TableDesc), i32, i32, vp],
        "zsmpm_g2p__rocm": [vp, P(ParticlesDesc), P(TableDesc), P(BucketsDesc), P(GridDesc), f32],
        "zsmpm_p2c2g__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(TableDesc), P(BucketsDesc), P(GridDesc), f32],
        "zsmpm_g2c2p__rocm": [vp, P(ParticlesDesc), P(TableDesc), P(BucketsDesc), P(GridDesc), f32],
        "zsmpm_substep__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(TableDesc), P(BucketsDesc), P(GridDesc), f32,
                                P(C.c_float * 3), P(PlaneCollider), i32],
        "zsmpm_substep_sorted__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(ParticlesDesc), P(TableDesc),
                                       P(BucketsDesc), P(GridDesc), f32, P(C.c_float * 3), P(PlaneCollider), i32],
        "zsmpm_reorder__rocm": [vp, P(BucketsDesc), P(ParticlesDesc), P(ParticlesDesc)],
        "zsmpm_substeps_sorted__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(ParticlesDesc), P(TableDesc),
                                        P(BucketsDesc), P(GridDesc), f32, P(C.c_float * 3), P(PlaneCollider), i32,
                                        i32, i32, P(C.c_int32)],
        "zsmpm_policy_closure__rocm": [vp, P(ClosureDesc)],
        "zsmpm_policy_launch_widths__rocm": [vp, P(C.c_int32), P(C.c_int32)],
        "zsmpm_policy_launch_tiles__rocm": [vp, P(C.c_int32), P(C.c_int32)],
        "zsmpm_policy_transfer_scheme__rocm": [vp, i32],
        "zsmpm_policy_set_launch_tiles__rocm": [vp, i32, i32],
        "zsmpm_policy_set_launch_widths__rocm": [vp, C.c_int32, C.c_int32],
        "zsmpm_policy_colliders__rocm": [vp, P(ColliderDesc), i32],
        "zsmpm_substeps_batched__rocm": [vp, P(BatchSceneDesc), i32, P(ParticlesDesc), P(ParticlesDesc),
                                         P(TableDesc), P(BucketsDesc), P(GridDesc), i32, i32, P(C.c_int32)],
        "zsmpm_g2p2g__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(TableDesc), P(GridDesc), f32, vp, vp],
        "zsmpm_implicit_multiply__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(TableDesc), P(GridDesc), f32, vp,
                                          vp],
        "zsmpm_implicit_project__rocm": [vp, P(TableDesc), P(GridDesc), P(PlaneCollider), i32, vp],
        "zsmpm_implicit_precondition__rocm": [vp, P(TableDesc), P(GridDesc), vp, vp],
        "zsmpm_implicit_solve__rocm": [vp, P(ModelDesc), P(ParticlesDesc), P(TableDesc), P(GridDesc), f32,
                                       P(PlaneCollider), i32, vp, vp, i32, f32, f32, P(C.c_int32), P(C.c_float)],
        "zsmpm_bht_clear__rocm": [vp, P(BhtDesc)],
        "zsmpm_bht_insert__rocm": [vp, P(BhtDesc), vp, i64, vp, i32],
        "zsmpm_bht_query__rocm": [vp, P(BhtDesc), vp, i64, vp],
        "zsmpm_spg_activate__rocm": [vp, P(SparseGridDesc), vp, i64],
        "zsmpm_spg_sample__rocm": [vp, P(SparseGridDesc), i32, i32, vp, i64, vp],
        "zsmpm_policy_levelset_boundary__rocm": [vp, P(SparseGridDesc), i32, i32, i32],
        "zsmpm_particles_from_tiles__rocm": [vp, P(TileParticlesDesc), P(ParticlesDesc)],
        "zsmpm_particles_to_tiles__rocm": [vp, P(ParticlesDesc), P(TileParticlesDesc)],
        "zsmpm_tiles_reorder__rocm": [vp, vp, vp, vp, i64, i32, i32, i32],
        "zsmpm_checkpoint_save__rocm": [vp, C.c_char_p, P(ParticlesDesc), P(TableDesc), C.c_double, i64],
        "zsmpm_checkpoint_info__rocm": [C.c_char_p, P(CheckpointInfo)],
        "zsmpm_checkpoint_load__rocm": [vp, C.c_char_p, P(ParticlesDesc), P(TableDesc), P(CheckpointInfo)],
        "zsmpm_write_partio__rocm": [vp, C.c_char_p, vp, i64, vp, i32, C.c_char_p],
        "zsmpm_stress__rocm": [vp, P(ModelDesc), vp, vp, i64, vp],
        "zsmpm_mlp_harvest__rocm": [vp, P(ModelDesc), vp, vp, i64, vp, vp],
        "zsmpm_mlp_infer__rocm": [vp, vp, vp, i64, vp],
        "zsmpm_mlp_grad__rocm": [vp, vp, vp, vp, i64, vp, vp],
        "zsmpm_mlp_train_step__rocm": [vp, vp, vp, vp, vp, vp, i64, f32, i32, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    L.zsmpm_bht_table_size.argtypes = [i64]
    L.zsmpm_bht_table_size.restype = i64
    L.zsmpm_bht_hash_params.argtypes = [C.c_uint32, P(C.c_uint32 * 6)]
    L.zsmpm_bht_hash_params.restype = None
    _lib = L
    return L


def check(code: int, what: str) -> None:
    if code != 0:
        raise ZsmpmError(code, what)


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the zsmpm HIP path needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


class Policy:
    """zsmpm_policy: device, stream and scratch (CudaExecutionPolicy's counterpart)."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        self.L = lib()
        self.device = device
        self.stream = stream if stream is not None else torch.cuda.current_stream(device)
        h = C.c_void_p()
        check(self.L.zsmpm_policy__rocm(device, C.c_void_p(self.stream.cuda_stream), C.byref(h)), "policy")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.zsmpm_del_policy__rocm(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, n: int, table_size: int):
        check(self.L.zsmpm_policy_reserve__rocm(self.h, n, table_size), "reserve")

    def sync(self):
        check(self.L.zsmpm_policy_sync(self.h), "device")

    def counters(self) -> np.ndarray:
        out = (C.c_uint32 * 16)()
        check(self.L.zsmpm_policy_counters(self.h, out), "counters")
        return np.frombuffer(out, dtype=np.uint32).copy()

    def profile(self, enable: bool = True):
        """Time every stage with HIP events on the policy's stream."""
        check(self.L.zsmpm_policy_profile(self.h, 1 if enable else 0), "profile")

    def set_closure(self, params=None, tol: float = 1e-4, enable: bool = True, always_network: bool = False):
        """Stress passes of Drucker-Prager models use the learned closure:
        params (a device float32 tensor of MLP_PARAMS) or None for Newton from
        X0 = 0; enable=False restores the plain SVD return mapping;
        always_network evaluates the network even in waves that one step from
        X0 = 0 certifies a priori. The policy reads params at every pass: keep
        the tensor alive."""
        if not enable:
            check(self.L.zsmpm_policy_closure__rocm(self.h, None), "closure")
            self._closure_params = None
            return
        d = ClosureDesc()
        d.hdr = _hdr(ClosureDesc)
        d.params = None if params is None else params.data_ptr()
        d.tol = float(tol)
        d.flags = 1 if always_network else 0
        check(self.L.zsmpm_policy_closure__rocm(self.h, C.byref(d)), "closure")
        self._closure_params = params

    def set_colliders(self, colliders):
        """General colliders (scenes.GeneralCollider) of every later grid update
        on this policy, after the call's plane colliders; [] removes them."""
        cols = list(colliders or [])
        arr = (ColliderDesc * max(len(cols), 1))(*[collider_desc(c) for c in cols])
        check(self.L.zsmpm_policy_colliders__rocm(self.h, arr if cols else None, len(cols)), "colliders")

    def launch_widths(self):
        """(P2G, G2P) workgroup counts the policy launches with."""
        p, g = C.c_int32(0), C.c_int32(0)
        check(self.L.zsmpm_policy_launch_widths__rocm(self.h, C.byref(p), C.byref(g)), "launch_widths")
        return p.value, g.value

    def set_launch_widths(self, p2g: int, g2p: int):
        """Fixes the P2G and G2P workgroup counts of later launches (0: default)."""
        check(self.L.zsmpm_policy_set_launch_widths__rocm(self.h, int(p2g), int(g2p)), "set_launch_widths")

    def set_transfer_scheme(self, scheme: int):
        """Transfers of later substeps: TRANSFER_APIC (P2G / G2P) or TRANSFER_CELL (P2C2G / G2C2P)."""
        check(self.L.zsmpm_policy_transfer_scheme__rocm(self.h, int(scheme)), "transfer_scheme")

    def launch_tiles(self):
        """(P2G chunk, G2P threads per bucket) of the sorted substep's launches."""
        a, b = C.c_int32(0), C.c_int32(0)
        check(self.L.zsmpm_policy_launch_tiles__rocm(self.h, C.byref(a), C.byref(b)), "launch_tiles")
        return a.value, b.value

    def set_launch_tiles(self, p2g_chunk: int, g2p_threads: int):
        """Fixes the P2G chunk (576 or 256) and G2P threads (128 or 256); 0: default."""
        check(self.L.zsmpm_policy_set_launch_tiles__rocm(self.h, int(p2g_chunk), int(g2p_threads)),
              "set_launch_tiles")

    # ---- layout and IO adapters ----
    def particles_from_tiles(self, tiles: "DeviceTiles", parts: "DeviceParticles", matrix_order: int = ROW_MAJOR):
        """parts' arrays <- the attributes the tile set holds (one HBM pass)."""
        check(self.L.zsmpm_particles_from_tiles__rocm(self.h, C.byref(tiles.desc(parts, matrix_order)),
                                                      C.byref(parts.desc)), "particles_from_tiles")

    def particles_to_tiles(self, parts: "DeviceParticles", tiles: "DeviceTiles", matrix_order: int = ROW_MAJOR):
        """The tile set's channels of parts' attributes <- parts."""
        check(self.L.zsmpm_particles_to_tiles__rocm(self.h, C.byref(parts.desc),
                                                    C.byref(tiles.desc(parts, matrix_order))), "particles_to_tiles")

    def tiles_reorder(self, src: "DeviceTiles", dst: "DeviceTiles", tile_map, scatter: bool = False):
        """TileVector::reorderTiles: gather dst tile t = src tile map[t]
        (scatter: dst tile map[t] = src tile t); map a device int32 tensor."""
        if (src.L, src.nch, src.n) != (dst.L, dst.nch, dst.n):
            raise ValueError("tile sets of different shapes")
        check(self.L.zsmpm_tiles_reorder__rocm(self.h, src.data.data_ptr(), dst.data.data_ptr(), tile_map.data_ptr(),
                                               src.ntiles, src.L, src.nch, 1 if scatter else 0), "tiles_reorder")

    def checkpoint_save(self, path: str, parts: "DeviceParticles", table: Optional["DeviceTable"] = None,
                        time: float = 0.0, step: int = 0):
        """Writes parts (and the table) to one checkpoint file; synchronises."""
        check(self.L.zsmpm_checkpoint_save__rocm(self.h, os.fsencode(path), C.byref(parts.desc),
                                                 None if table is None else C.byref(table.desc), float(time),
                                                 int(step)), "checkpoint_save")

    def checkpoint_load(self, path: str, parts: "DeviceParticles", table: Optional["DeviceTable"] = None):
        """Fills parts (and the table) from a checkpoint; returns its info."""
        info = CheckpointInfo()
        check(self.L.zsmpm_checkpoint_load__rocm(self.h, os.fsencode(path), C.byref(parts.desc),
                                                 None if table is None else C.byref(table.desc), C.byref(info)),
              "checkpoint_load")
        return {"n": info.n, "table_size": info.table_size, "fields": info.fields, "time": info.time,
                "step": info.step}

    def write_partio(self, path: str, pos, attr=None, attr_name: Optional[str] = None):
        """write_partio (positions) / _with_stress (attr of n floats, named
        "stress") / _with_grid (attr of n x 3, named "force"): a bgeo file.
        pos and attr are device float32 tensors."""
        n = int(pos.shape[0])
        width = 0 if attr is None else (1 if attr.dim() == 1 else int(attr.shape[1]))
        check(self.L.zsmpm_write_partio__rocm(self.h, os.fsencode(path), pos.data_ptr(), n,
                                              None if attr is None else attr.data_ptr(), width,
                                              None if attr_name is None else attr_name.encode()), "write_partio")

    def set_levelset_boundary(self, grid: Optional["DeviceSparseGrid"], sdf_channel: int = 0, v_channel: int = -1,
                              kind: int = 0):
        """LevelSetBoundary<SparseLevelSet> of every later grid update (None
        removes it); the grid's buffers are read by every update."""
        if grid is None:
            check(self.L.zsmpm_policy_levelset_boundary__rocm(self.h, None, 0, -1, 0), "levelset_boundary")
            self._ls_grid = None
            return
        check(self.L.zsmpm_policy_levelset_boundary__rocm(self.h, C.byref(grid.desc()), int(sdf_channel),
                                                          int(v_channel), int(kind)), "levelset_boundary")
        self._ls_grid = grid

    def stage_times(self, reset: bool = False) -> dict:
        """{stage: (total ms, calls)} accumulated since the last reset."""
        n = len(STAGES)
        ms = (C.c_double * n)()
        calls = (C.c_uint64 * n)()
        check(self.L.zsmpm_policy_stage_times(self.h, ms, calls, n, 1 if reset else 0), "stage_times")
        return {name: (float(ms[i]), int(calls[i])) for i, name in enumerate(STAGES)}


class DeviceTable:
    """HashTable<i32,3,int> buffers (container/HashTable.hpp:40-58, 200-205)."""

    def __init__(self, table_size: int, device="cuda"):
        torch = _torch()
        T = int(table_size)
        self.T = T
        self.keys = torch.full((T, 3), 2 ** 31 - 1, dtype=torch.int32, device=device)
        self.indices = torch.full((T,), -1, dtype=torch.int32, device=device)
        self.status = torch.full((T,), -1, dtype=torch.int32, device=device)
        self.cnt = torch.zeros(1, dtype=torch.int32, device=device)
        self.active = torch.zeros((T, 3), dtype=torch.int32, device=device)
        self.desc = TableDesc(_hdr(TableDesc), T, _ptr(self.keys), _ptr(self.indices), _ptr(self.status),
                              _ptr(self.cnt), _ptr(self.active))

    def host(self):
        cnt = int(self.cnt.item())
        return {"T": self.T, "cnt": cnt, "keys": self.keys.cpu().numpy(), "indices": self.indices.cpu().numpy(),
                "status": self.status.cpu().numpy(), "active": self.active[:cnt].cpu().numpy()}


class DeviceParticles:
    """Particles<f32,3> attributes on the device."""

    def __init__(self, x, v, Cm, F, m, logJp=None, device="cuda", ids=None):
        torch = _torch()

        def dev(a, shape, dtype=np.float32):
            return torch.as_tensor(np.ascontiguousarray(a, dtype)).reshape(shape).to(device).contiguous()

        n = int(np.asarray(x).reshape(-1, 3).shape[0])
        self.n = n
        self.x = dev(x, (n, 3))
        self.v = dev(v, (n, 3))
        self.C = dev(Cm, (n, 9))
        self.F = dev(F, (n, 9))
        self.m = dev(m, (n,))
        self.logJp = None if logJp is None else dev(logJp, (n,))
        self.id = None if ids is None else dev(ids, (n,), np.int32)
        self._describe()

    def _describe(self):
        self.desc = ParticlesDesc(_hdr(ParticlesDesc), self.n, _ptr(self.x), _ptr(self.v), _ptr(self.C),
                                  _ptr(self.F), _ptr(self.m), _ptr(self.logJp), _ptr(self.id))

    @staticmethod
    def from_scene(sc: S.Scene, device="cuda", ids=False):
        return DeviceParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp, device,
                               np.arange(sc.n, dtype=np.int32) if ids else None)

    def empty_like(self):
        """Uninitialised buffers of the same shape (the reorder target)."""
        torch = _torch()
        out = DeviceParticles.__new__(DeviceParticles)
        out.n = self.n
        for k in ("x", "v", "C", "F", "m", "logJp", "id"):
            a = getattr(self, k)
            setattr(out, k, None if a is None else torch.empty_like(a))
        out._describe()
        return out

    def host(self):
        out = {k: getattr(self, k).cpu().numpy() for k in ("x", "v", "C", "F", "m")}
        out["logJp"] = None if self.logJp is None else self.logJp.cpu().numpy()
        out["id"] = None if self.id is None else self.id.cpu().numpy()
        return out

    def host_by_id(self):
        """host() with every attribute scattered back to particle-id order."""
        h = self.host()
        ids = h.pop("id")
        if ids is None:
            return h
        out = {}
        for k, a in h.items():
            if a is None:
                out[k] = None
                continue
            b = np.empty_like(a)
            b[ids] = a
            out[k] = b
        return out


def checkpoint_info(path: str) -> dict:
    """The header of a checkpoint file (sizes its load)."""
    info = CheckpointInfo()
    check(lib().zsmpm_checkpoint_info__rocm(os.fsencode(path), C.byref(info)), "checkpoint_info")
    return {"n": info.n, "table_size": info.table_size, "fields": info.fields, "time": info.time,
            "step": info.step}


class DeviceTiles:
    """A TileVector<f32, L> of n elements: element i, channel c at
    (i // L * nch + c) * L + i % L (container/TileVector.hpp:108,397). tags:
    [(name, width)] in channel order, as zpc's PropertyTag lists; attributes
    named x, v, C, F, m, logJp and id map to the particle arrays."""

    NAMES = ("x", "v", "C", "F", "m", "logJp", "id")

    def __init__(self, tags, n: int, lane_width: int = 32, device="cuda"):
        torch = _torch()
        self.tags = [(str(a), int(w)) for a, w in tags]
        self.offsets, off = {}, 0
        for a, w in self.tags:
            self.offsets[a] = off
            off += w
        self.nch, self.L, self.n = off, int(lane_width), int(n)
        self.ntiles = (self.n + self.L - 1) // self.L
        self.data = torch.zeros(max(self.ntiles, 1) * self.nch * self.L, dtype=torch.float32, device=device)

    def desc(self, parts=None, matrix_order: int = ROW_MAJOR) -> TileParticlesDesc:
        d = TileParticlesDesc()
        d.hdr = _hdr(TileParticlesDesc)
        d.n, d.lane_width, d.num_channels, d.data = self.n, self.L, self.nch, self.data.data_ptr()
        for a in self.NAMES:
            named = a in self.offsets and (parts is None or getattr(parts, a) is not None)
            setattr(d, a, self.offsets[a] if named else -1)
        d.matrix_order = int(matrix_order)
        return d

    def host(self) -> np.ndarray:
        """(ntiles, nch, L) view of the buffer on the host."""
        return self.data.cpu().numpy().reshape(max(self.ntiles, 1), self.nch, self.L)


def bht_table_size(n: int) -> int:
    return int(lib().zsmpm_bht_table_size(int(n)))


def bht_hash_params(seed: int = BHT_SEED):
    out = (C.c_uint32 * 6)()
    lib().zsmpm_bht_hash_params(seed, C.byref(out))
    return [int(v) for v in out]


class DeviceBht:
    """bht<i32, 3, int, 16> buffers on the device, cleared."""

    def __init__(self, table_size: int, pol: Optional[Policy] = None, hash_params=None, device="cuda"):
        torch = _torch()
        T = int(table_size)
        if T < 16 or T % 16:
            raise ValueError("bht table size: a positive multiple of 16 (bht_table_size)")
        self.T = T
        self.keys = torch.full((T, 4), 0x3F3F3F3F, dtype=torch.int32, device=device)
        self.indices = torch.full((T,), -1, dtype=torch.int32, device=device)
        self.status = torch.full((T,), -1, dtype=torch.int32, device=device)
        self.cnt = torch.zeros(1, dtype=torch.int32, device=device)
        self.active = torch.zeros((T, 3), dtype=torch.int32, device=device)
        self.success = torch.ones(1, dtype=torch.int32, device=device)
        self.hash = list(bht_hash_params() if hash_params is None else hash_params)
        self.pol = pol if pol is not None else Policy()
        self.desc = BhtDesc(_hdr(BhtDesc), T, _ptr(self.keys), _ptr(self.indices), _ptr(self.status),
                            _ptr(self.cnt), _ptr(self.active), _ptr(self.success), (C.c_uint32 * 6)(*self.hash))

    def clear(self):
        check(self.pol.L.zsmpm_bht_clear__rocm(self.pol.h, C.byref(self.desc)), "bht_clear")

    def insert(self, keys, serial: bool = False, out=None):
        """Inserts device int32 keys (n, 3); returns the (n,) insert results."""
        torch = _torch()
        n = int(keys.shape[0])
        if out is None:
            out = torch.empty(max(n, 1), dtype=torch.int32, device=keys.device)
        check(self.pol.L.zsmpm_bht_insert__rocm(self.pol.h, C.byref(self.desc), keys.data_ptr(), n, out.data_ptr(),
                                                BHT_SERIAL if serial else 0), "bht_insert")
        return out[:n]

    def query(self, keys, out=None):
        torch = _torch()
        n = int(keys.shape[0])
        if out is None:
            out = torch.empty(max(n, 1), dtype=torch.int32, device=keys.device)
        check(self.pol.L.zsmpm_bht_query__rocm(self.pol.h, C.byref(self.desc), keys.data_ptr(), n, out.data_ptr()),
              "bht_query")
        return out[:n]

    def size(self) -> int:
        return int(self.cnt.item())


class DeviceSparseGrid:
    """SparseGrid<3, f32, 8>: a DeviceBht of block origins, a TileVector<f32,
    512> of (capacity, nch, 512) cell values, the index-to-world transform (4
    x 4, row-vector convention) and the background value."""

    def __init__(self, capacity_blocks: int, num_channels: int = 1, i2w=None, background: float = 0.0,
                 pol: Optional[Policy] = None, device="cuda"):
        torch = _torch()
        self.cap = int(capacity_blocks)
        self.nch = int(num_channels)
        self.table = DeviceBht(bht_table_size(self.cap), pol)
        self.pol = self.table.pol
        self.data = torch.zeros((max(self.cap, 1), self.nch, 512), dtype=torch.float32, device=device)
        M = np.eye(4) if i2w is None else np.asarray(i2w, np.float64).reshape(4, 4)
        self.i2w = M.astype(np.float32)
        self.w2i = np.linalg.inv(M).astype(np.float32)
        self.background = float(background)

    def desc(self) -> SparseGridDesc:
        d = SparseGridDesc()
        d.hdr = _hdr(SparseGridDesc)
        d.table = self.table.desc
        d.data, d.num_channels, d.capacity_blocks = self.data.data_ptr(), self.nch, self.cap
        for i in range(16):
            d.i2w[i] = float(self.i2w.reshape(16)[i])
            d.w2i[i] = float(self.w2i.reshape(16)[i])
        d.background = self.background
        return d

    def activate(self, x):
        """SparseGridView::insert of every world position of x (device, (n, 3))."""
        check(self.pol.L.zsmpm_spg_activate__rocm(self.pol.h, C.byref(self.desc()), x.data_ptr(), int(x.shape[0])),
              "spg_activate")

    def sample(self, channel: int, x, width: int = 1):
        """wSample / wPack of [channel, channel + width) at world positions x."""
        torch = _torch()
        n = int(x.shape[0])
        out = torch.empty((max(n, 1), width), dtype=torch.float32, device=x.device)
        check(self.pol.L.zsmpm_spg_sample__rocm(self.pol.h, C.byref(self.desc()), int(channel), int(width),
                                                x.data_ptr(), n, out.data_ptr()), "spg_sample")
        return out[:n, 0] if width == 1 else out[:n]

    def cell_coords(self):
        """(nblocks, 512, 3) integer coordinates of every active cell (the
        block origin plus the cell's (x, y, z), k = (x 8 + y) 8 + z)."""
        nb = self.table.size()
        org = self.table.active[:nb].cpu().numpy().astype(np.int64)
        k = np.arange(512)
        off = np.stack([k >> 6, (k >> 3) & 7, k & 7], -1)
        return org[:, None, :] + off[None, :, :]


class DeviceGrid:
    """Grids<f32,3,4>{m:1, v:3, rhs:3}: (capacity, 7, 64) floats."""

    def __init__(self, dx: float, capacity_blocks: int, device="cuda"):
        torch = _torch()
        self.dx = float(dx)
        self.cap = int(capacity_blocks)
        self.data = torch.zeros((self.cap, 7, 64), dtype=torch.float32, device=device)
        self.desc = GridDesc(_hdr(GridDesc), self.dx, self.cap, _ptr(self.data))


class DeviceBuckets:
    """IndexBuckets counts/offsets (capacity + 1) and indices (n)."""

    def __init__(self, capacity_blocks: int, n: int, device="cuda"):
        torch = _torch()
        self.cap = int(capacity_blocks)
        self.counts = torch.zeros(self.cap + 1, dtype=torch.int32, device=device)
        self.offsets = torch.zeros(self.cap + 1, dtype=torch.int32, device=device)
        self.indices = torch.zeros(max(int(n), 1), dtype=torch.int32, device=device)
        self.desc = BucketsDesc(_hdr(BucketsDesc), self.cap, _ptr(self.counts), _ptr(self.offsets),
                                _ptr(self.indices))

    def host(self, nblocks: int, n: int):
        return (self.counts[: nblocks + 1].cpu().numpy(), self.offsets[: nblocks + 1].cpu().numpy(),
                self.indices[:n].cpu().numpy())


def model_desc(mat: S.Material, svd_fp64_gap: float = 0.0) -> ModelDesc:
    return ModelDesc(_hdr(ModelDesc), int(mat.kind), float(mat.volume), float(mat.E), float(mat.nu),
                     float(mat.cohesion), float(mat.beta), float(mat.yield_surface), int(mat.volume_correction),
                     float(svd_fp64_gap))


def colliders_array(cols: Sequence[S.Collider]):
    arr = (PlaneCollider * max(len(cols), 1))()
    for i, c in enumerate(cols):
        arr[i].type = int(c.kind)
        for d in range(3):
            arr[i].origin[d] = float(c.origin[d])
            arr[i].normal[d] = float(c.normal[d])
    return arr


class Simulator:
    """Drives the C-ABI substep for one scene: the explicit driver zpc itself
    lacks, in the stage order its operators imply (table clear, activation,
    enlargement, bucket sort, grid clear, P2G, grid update, G2P). Device state
    stays resident."""

    def __init__(self, sc: S.Scene, device: int = 0, grid_capacity: Optional[int] = None,
                 svd_fp64_gap: float = 0.0, reorder: bool = False, ids: Optional[bool] = None):
        self.sc = sc
        self.pol = Policy(device)
        T = int(sc.table_size)
        self.table = DeviceTable(T)
        cap = int(grid_capacity) if grid_capacity else max(T // 16, 64)
        self.grid = DeviceGrid(sc.dx, cap)
        # reorder: particles are kept in bucket order (ping-pong buffers); ids
        # (by default with the reorder) are carried with them, so that a
        # particle can be found again
        self.reorder = bool(reorder)
        self.parts = DeviceParticles.from_scene(sc, ids=self.reorder if ids is None else bool(ids))
        self.spare = self.parts.empty_like() if self.reorder else None
        self.buckets = DeviceBuckets(cap, sc.n)
        self.model = model_desc(sc.material, svd_fp64_gap)
        self.cols = colliders_array(sc.colliders)
        self.ncols = len(sc.colliders)
        self.extf = (C.c_float * 3)(*[float(g) for g in sc.gravity])
        self.pol.reserve(sc.n, T)
        self.L = lib()
        if sc.general_colliders:
            self.pol.set_colliders(sc.general_colliders)

    def substep(self, dt: Optional[float] = None):
        dt = float(self.sc.dt if dt is None else dt)
        if self.reorder:
            check(self.L.zsmpm_substep_sorted__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                                    C.byref(self.spare.desc), C.byref(self.table.desc),
                                                    C.byref(self.buckets.desc), C.byref(self.grid.desc), dt,
                                                    C.byref(self.extf), self.cols, self.ncols), "substep_sorted")
            self.parts, self.spare = self.spare, self.parts
            return
        check(self.L.zsmpm_substep__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                         C.byref(self.table.desc), C.byref(self.buckets.desc),
                                         C.byref(self.grid.desc), dt, C.byref(self.extf), self.cols, self.ncols),
              "substep")

    def substeps(self, n: int, dt: Optional[float] = None, fuse_stress: bool = False, graph: bool = False,
                 autotune: bool = False):
        """n substeps in one call (sorted driver); fuse_stress: G2P computes the
        next substep's stress; graph: replay a captured hipGraph of two
        substeps (kept by the library while the arguments stay the same);
        autotune: the first 8 substeps pick the P2G / G2P workgroup counts."""
        if not self.reorder:
            for _ in range(n):
                self.substep(dt)
            return
        dt = float(self.sc.dt if dt is None else dt)
        in_b = C.c_int32(0)
        check(self.L.zsmpm_substeps_sorted__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                                 C.byref(self.spare.desc), C.byref(self.table.desc),
                                                 C.byref(self.buckets.desc), C.byref(self.grid.desc), dt,
                                                 C.byref(self.extf), self.cols, self.ncols, int(n),
                                                 (1 if fuse_stress else 0) | (2 if graph else 0) |
                                                 (4 if autotune else 0),
                                                 C.byref(in_b)),
              "substeps_sorted")
        if in_b.value:
            self.parts, self.spare = self.spare, self.parts

    # stage-level operators (zpc functor equivalents)
    def clear_table(self):
        check(self.L.zsmpm_table_clear__rocm(self.pol.h, C.byref(self.table.desc)), "table_clear")

    def activate(self, x=None, n=None, block_len=4, offset=-2, disp=0.5):
        x = self.parts.x if x is None else x
        n = self.parts.n if n is None else n
        check(self.L.zsmpm_table_activate__rocm(self.pol.h, C.byref(self.table.desc), x.data_ptr(), n,
                                                self.grid.dx, block_len, offset, disp), "activate")

    def enlarge(self, lo=(0, 0, 0), hi=(2, 2, 2)):
        lo_ = (C.c_int32 * 3)(*lo)
        hi_ = (C.c_int32 * 3)(*hi)
        check(self.L.zsmpm_table_enlarge__rocm(self.pol.h, C.byref(self.table.desc), C.byref(lo_), C.byref(hi_)),
              "enlarge")

    def bucket_sort(self, block_len=4, offset=-2, disp=0.5):
        check(self.L.zsmpm_bucket_sort__rocm(self.pol.h, C.byref(self.table.desc), self.parts.x.data_ptr(),
                                             self.parts.n, self.grid.dx, block_len, offset, disp,
                                             C.byref(self.buckets.desc)), "bucket_sort")

    def clear_grid(self):
        check(self.L.zsmpm_grid_clear__rocm(self.pol.h, C.byref(self.grid.desc), C.byref(self.table.desc)),
              "grid_clear")

    def p2g(self, dt=None):
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_p2g__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                     C.byref(self.table.desc), C.byref(self.buckets.desc), C.byref(self.grid.desc),
                                     dt), "p2g")

    def grid_update(self, dt=None, fold=1, max_vel=None):
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_grid_update__rocm(self.pol.h, C.byref(self.grid.desc), C.byref(self.table.desc), dt,
                                             C.byref(self.extf), self.cols, self.ncols, fold,
                                             None if max_vel is None else max_vel.data_ptr()), "grid_update")

    def angular_momentum(self, m_channel: int = 0, mv_channel: int = 1) -> np.ndarray:
        """GridAngularMomentum (GridOp.hpp:216-259) of the grid as it stands:
        [sum x cross mv (3), sum mv (3)] over nodes with mass, float64."""
        torch = _torch()
        out = torch.zeros(6, dtype=torch.float64, device=self.grid.data.device)
        check(self.L.zsmpm_grid_angular_momentum__rocm(self.pol.h, C.byref(self.grid.desc),
                                                       C.byref(self.table.desc), m_channel, mv_channel,
                                                       out.data_ptr()), "grid_angular_momentum")
        self.pol.sync()
        return out.cpu().numpy()

    def g2p(self, dt=None):
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_g2p__rocm(self.pol.h, C.byref(self.parts.desc), C.byref(self.table.desc),
                                     C.byref(self.buckets.desc), C.byref(self.grid.desc), dt), "g2p")

    def p2c2g(self, dt=None):
        """P2C2GTransfer into the (cleared) grid: m and mv through the cells."""
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_p2c2g__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                       C.byref(self.table.desc), C.byref(self.buckets.desc),
                                       C.byref(self.grid.desc), dt), "p2c2g")

    def g2c2p(self, dt=None):
        """G2C2PTransfer with its pre and post passes on the particles, in place."""
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_g2c2p__rocm(self.pol.h, C.byref(self.parts.desc), C.byref(self.table.desc),
                                       C.byref(self.buckets.desc), C.byref(self.grid.desc), dt), "g2c2p")

    def reorder_particles(self):
        """Vector::reorder of every attribute by the current bucket sort."""
        if self.spare is None:
            self.spare = self.parts.empty_like()
        if self.parts.id is not None and self.spare.id is None:
            raise RuntimeError("reorder target lacks ids")
        check(self.L.zsmpm_reorder__rocm(self.pol.h, C.byref(self.buckets.desc), C.byref(self.parts.desc),
                                         C.byref(self.spare.desc)), "reorder")
        self.parts, self.spare = self.spare, self.parts

    def sync(self):
        self.pol.sync()

    # ---- implicit step (G2P2G, ImplicitMPMSystem, the conjugate gradient) ----
    def dof_vector(self, fill: float = 0.0):
        """A grid vector: 3 floats per node of the grid's capacity."""
        torch = _torch()
        return torch.full((self.grid.cap * 64 * 3,), float(fill), dtype=torch.float32, device=self.grid.data.device)

    def g2p2g(self, v, r, dt=None):
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_g2p2g__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                       C.byref(self.table.desc), C.byref(self.grid.desc), dt, v.data_ptr(),
                                       r.data_ptr()), "g2p2g")

    def implicit_multiply(self, inp, out, dt=None):
        dt = float(self.sc.dt if dt is None else dt)
        check(self.L.zsmpm_implicit_multiply__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                                   C.byref(self.table.desc), C.byref(self.grid.desc), dt,
                                                   inp.data_ptr(), out.data_ptr()), "implicit_multiply")

    def implicit_project(self, dof):
        check(self.L.zsmpm_implicit_project__rocm(self.pol.h, C.byref(self.table.desc), C.byref(self.grid.desc),
                                                  self.cols, self.ncols, dof.data_ptr()), "implicit_project")

    def implicit_precondition(self, inp, out):
        check(self.L.zsmpm_implicit_precondition__rocm(self.pol.h, C.byref(self.table.desc), C.byref(self.grid.desc),
                                                       inp.data_ptr(), out.data_ptr()), "implicit_precondition")

    def implicit_solve(self, x, b, dt=None, max_iters: int = 1000, tol: float = 1e-6, rel_tol: float = 0.5):
        """ConjugateGradient::solve of the implicit system; x in place.
        Returns (iterations, residual)."""
        dt = float(self.sc.dt if dt is None else dt)
        it, res = C.c_int32(0), C.c_float(0)
        check(self.L.zsmpm_implicit_solve__rocm(self.pol.h, C.byref(self.model), C.byref(self.parts.desc),
                                                C.byref(self.table.desc), C.byref(self.grid.desc), dt, self.cols,
                                                self.ncols, x.data_ptr(), b.data_ptr(), int(max_iters), float(tol),
                                                float(rel_tol), C.byref(it), C.byref(res)), "implicit_solve")
        return it.value, res.value

    def save_checkpoint(self, path: str, time: float = 0.0, step: int = 0, table: bool = True):
        """The particle state (and the table) as one checkpoint file."""
        self.pol.checkpoint_save(path, self.parts, self.table if table else None, time, step)

    def load_checkpoint(self, path: str) -> dict:
        """Resumes from a checkpoint written for a scene of the same size."""
        info = checkpoint_info(path)
        return self.pol.checkpoint_load(path, self.parts, self.table if info["table_size"] else None)

    def nblocks(self) -> int:
        return int(self.table.cnt.item())

    def evaluated_dt(self, cfl: float = 0.5) -> float:
        """The builder's time step, cfl * dx / sound speed of the scene's model
        (MPMSimulator::evaluatedDt, simulation/mpm/Simulator.cpp:53-64)."""
        return S.cfl_dt(self.sc.material, self.sc.dx, cfl)

    def max_vel_sqr(self) -> float:
        """max |v|^2 over the grid nodes with mass after the last grid update
        (MPMSimulator::getMaxVel, simulation/mpm/Simulator.hpp:19-26); syncs."""
        self.pol.sync()
        bits = np.array([self.pol.counters()[5]], np.uint32)
        return float(bits.view(np.float32)[0])

    def adaptive_dt(self, cfl: float = 0.5) -> float:
        """The next substep's dt: the builder's CFL step, lowered so that the
        fastest grid node of the last update moves at most cfl * dx."""
        dt = self.evaluated_dt(cfl)
        v2 = self.max_vel_sqr()
        if v2 > 0.0:
            dt = min(dt, float(np.float32(cfl) * np.float32(self.sc.dx) / np.sqrt(np.float32(v2))))
        return dt


class Batch:
    """Independent scenes advanced together by zsmpm_substeps_batched__rocm:
    their particles concatenated into one set of arrays (scene s at
    [first[s], first[s] + n_s), ids = positions in the concatenation), one
    table, one bucket sort and one grid, every stage one launch over all
    scenes. Sorted driver (ping-pong buffers)."""

    def __init__(self, scenes: Sequence[S.Scene], device: int = 0, table_size: Optional[int] = None,
                 grid_capacity: Optional[int] = None):
        if not 0 < len(scenes) <= MAX_SCENES:
            raise ValueError(f"a batch holds 1..{MAX_SCENES} scenes")
        self.scenes = list(scenes)
        self.pol = Policy(device)
        self.L = lib()
        counts = [int(sc.n) for sc in self.scenes]
        self.first = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        n = int(self.first[-1])
        self.n = n
        cat = lambda key, shape: np.concatenate([np.asarray(getattr(sc, key), np.float32).reshape(shape)
                                                 for sc in self.scenes])
        plastic = any(sc.logJp is not None for sc in self.scenes)
        lj = None
        if plastic:
            lj = np.concatenate([np.zeros(sc.n, np.float32) if sc.logJp is None else
                                 np.asarray(sc.logJp, np.float32).reshape(-1) for sc in self.scenes])
        self.parts = DeviceParticles(cat("x", (-1, 3)), cat("v", (-1, 3)), cat("C", (-1, 9)), cat("F", (-1, 9)),
                                     cat("m", (-1,)), lj, ids=np.arange(n, dtype=np.int32))
        self.spare = self.parts.empty_like()
        # the scenes' blocks side by side: the table sized for their sum
        est = sum(max(int(sc.table_size) // 16, 1) for sc in self.scenes)
        T = int(table_size) if table_size else 16 * S.next_2pow(est)
        self.table = DeviceTable(T)
        cap = int(grid_capacity) if grid_capacity else max(T // 16, 64)
        self.grid = DeviceGrid(min(sc.dx for sc in self.scenes), cap)
        self.buckets = DeviceBuckets(cap, n)
        descs = (BatchSceneDesc * len(self.scenes))()
        for i, sc in enumerate(self.scenes):
            d = descs[i]
            d.hdr = _hdr(BatchSceneDesc)
            d.first = int(self.first[i])
            d.count = counts[i]
            d.model = model_desc(sc.material)
            d.dx = float(sc.dx)
            d.dt = float(sc.dt)
            for k in range(3):
                d.extf[k] = float(sc.gravity[k])
            if len(sc.colliders) > MAX_COLLIDERS:
                raise ValueError("at most 8 colliders per scene")
            d.ncolliders = len(sc.colliders)
            for j, c in enumerate(sc.colliders):
                d.colliders[j].type = int(c.kind)
                for k in range(3):
                    d.colliders[j].origin[k] = float(c.origin[k])
                    d.colliders[j].normal[k] = float(c.normal[k])
        self._descs = descs
        self.pol.reserve(n, T)

    def substeps(self, n: int, graph: bool = False, autotune: bool = False):
        """n sorted substeps of every scene (graph: replays of a captured
        two-substep hipGraph; autotune: the first 8 pick the P2G / G2P
        workgroup counts)."""
        in_b = C.c_int32(0)
        check(self.L.zsmpm_substeps_batched__rocm(self.pol.h, self._descs, len(self.scenes),
                                                  C.byref(self.parts.desc), C.byref(self.spare.desc),
                                                  C.byref(self.table.desc), C.byref(self.buckets.desc),
                                                  C.byref(self.grid.desc), int(n),
                                                  (2 if graph else 0) | (4 if autotune else 0), C.byref(in_b)),
              "substeps_batched")
        if in_b.value:
            self.parts, self.spare = self.spare, self.parts

    def sync(self):
        self.pol.sync()

    def nblocks(self) -> int:
        return int(self.table.cnt.item())

    def scene_state(self, i: int) -> dict:
        """Scene i's particles in its own id order (ids relative to the scene)."""
        h = self.parts.host_by_id()
        lo, hi = int(self.first[i]), int(self.first[i + 1])
        return {k: (None if a is None else a[lo:hi]) for k, a in h.items()}

    def ids(self) -> np.ndarray:
        return self.parts.id.cpu().numpy()

    @property
    def particles(self) -> int:
        return self.n
