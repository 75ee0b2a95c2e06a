"""The allocation-free activation (zsmpm_table_activate_noalloc__rocm) against
zsmpm_table_activate__rocm through the C ABI: the table arrays (keys,
indices, status, count, active keys) must be equal, bit for bit, on two calls
in a row, the second with a smaller input (the fixed scratch keeps the first
call's contents), and on a table that already holds keys. Sizes beyond the
fixed scratch are rejected; a key-range error raised on the fixed scratch is
reported by the policy's next sync.
"""
import numpy as np
import pytest

from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu

INSUFFICIENT_SIZE, KEY_RANGE = -3, -8


def _rocm():
    from zpc_amd import rocm

    return rocm


def _tables_equal(a, b, what):
    ha, hb = a.host(), b.host()
    assert ha["cnt"] == hb["cnt"], (what, ha["cnt"], hb["cnt"])
    for k in ("keys", "indices", "status", "active"):
        assert np.array_equal(ha[k], hb[k]), f"{what}: {k} differs"
    return ha["cnt"]


def _activate_both(rocm, pol_ref, pol_na, sc, T, x=None):
    import torch

    x = torch.as_tensor(np.ascontiguousarray(sc.x if x is None else x, np.float32)).cuda().contiguous()
    n = x.shape[0]
    ref, got = rocm.DeviceTable(T), rocm.DeviceTable(T)
    rocm.check(pol_ref.L.zsmpm_table_activate__rocm(pol_ref.h, rocm.C.byref(ref.desc), x.data_ptr(), n, sc.dx, 4,
                                                    -2, 0.5), "activate")
    rocm.check(pol_na.L.zsmpm_table_activate_noalloc__rocm(pol_na.h, rocm.C.byref(got.desc), x.data_ptr(), n, sc.dx,
                                                           4, -2, 0.5), "activate_noalloc")
    pol_ref.sync()
    pol_na.sync()
    return ref, got, x


def test_noalloc_activation_equals_allocating_entry_point():
    rocm = _rocm()
    pol_ref, pol_na = rocm.Policy(), rocm.Policy()
    big = S.sand_scene(dims=(64, 96, 32))  # 1.57M particles
    small = S.cube_scene(cells=12)
    print(f"first call: {big.n} particles, table {big.table_size}; second call: {small.n} particles, "
          f"table {small.table_size}", flush=True)
    ref, got, _ = _activate_both(rocm, pol_ref, pol_na, big, big.table_size)
    cnt = _tables_equal(ref, got, "first call")
    assert cnt > 0
    ref2, got2, _ = _activate_both(rocm, pol_ref, pol_na, small, small.table_size)
    assert _tables_equal(ref2, got2, "second, smaller call") > 0
    # a table that already holds keys: new keys are numbered after them
    shifted = (small.x + np.float32(3 * small.dx * 4)).astype(np.float32)
    import torch

    xs = torch.as_tensor(shifted).cuda().contiguous()
    rocm.check(pol_ref.L.zsmpm_table_activate__rocm(pol_ref.h, rocm.C.byref(ref2.desc), xs.data_ptr(), small.n,
                                                    small.dx, 4, -2, 0.5), "activate")
    rocm.check(pol_na.L.zsmpm_table_activate_noalloc__rocm(pol_na.h, rocm.C.byref(got2.desc), xs.data_ptr(), small.n,
                                                           small.dx, 4, -2, 0.5), "activate_noalloc")
    pol_ref.sync()
    pol_na.sync()
    _tables_equal(ref2, got2, "insertion into a filled table")


def test_noalloc_limits_and_errors():
    rocm = _rocm()
    import torch

    pol = rocm.Policy()
    sc = S.cube_scene(cells=8)
    x = torch.as_tensor(sc.x).cuda().contiguous()
    t = rocm.DeviceTable(sc.table_size)
    L = pol.L
    # more particles or slots than the fixed scratch holds: rejected before any launch
    assert L.zsmpm_table_activate_noalloc__rocm(pol.h, rocm.C.byref(t.desc), x.data_ptr(), (1 << 24) + 1, sc.dx, 4,
                                                -2, 0.5) == INSUFFICIENT_SIZE
    huge = rocm.TableDesc.from_buffer_copy(t.desc)
    huge.table_size = 1 << 22
    assert L.zsmpm_table_activate_noalloc__rocm(pol.h, rocm.C.byref(huge), x.data_ptr(), sc.n, sc.dx, 4, -2,
                                                0.5) == INSUFFICIENT_SIZE
    # a key outside [-2^20, 2^20): reported by the next sync, then cleared
    far = torch.full((4, 3), 0.5, dtype=torch.float32, device="cuda")
    far[0, 0] = 1e6
    assert L.zsmpm_table_activate_noalloc__rocm(pol.h, rocm.C.byref(t.desc), far.data_ptr(), 4, 1.0 / 128, 4, -2,
                                                0.5) == 0
    assert L.zsmpm_policy_sync(pol.h) == KEY_RANGE
    assert L.zsmpm_policy_sync(pol.h) == 0
    t2 = rocm.DeviceTable(sc.table_size)
    assert L.zsmpm_table_activate_noalloc__rocm(pol.h, rocm.C.byref(t2.desc), x.data_ptr(), sc.n, sc.dx, 4, -2,
                                                0.5) == 0
    pol.sync()
    assert t2.host()["cnt"] > 0
