"""Per-field error of the GPU path against the FP64 oracle on the 65,536-particle
sand column (16 x 32 x 16 cells at dx = 1/512, 8 particles per cell, the 16M
column's material, dt = 5e-5, sticky walls).

For every field the error is max|a - ref64| / scale over the field, where
ref64 is the oracle run in float64 from the same float32 initial state and
scale is max|ref64|, raised for fields that vanish at rest to their physical
scale: C to max|v_ref64| / dx (a velocity gradient), logJp (a dimensionless
log of a volume ratio) to 1, so its entry is an absolute error. The
float32 oracle's own error against ref64 is printed beside it: the rounding
floor any float32 implementation of the same algorithm reaches. Reported:
- particles x, v, C, F, logJp after 1 and after 100 sorted substeps of the
  column as built (at rest, F = I) and after 1 substep of a perturbed copy
  (deformed F, random v and C: every term of the transfer nonzero);
- the grid after one P2G of the perturbed copy: mass, momentum and the force
  channels (stage-level calls, rhs kept unfolded).
Bounds: 1e-5 after one substep and for the grid, 1e-3 after 100 substeps, or
(where float32 summation order decides) no worse than twice the float32
oracle's own error. The table is printed and, when the environment names a
report folder in ZSMPM_REPORT_DIR, written there as field_errors.md.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from zpc_amd import scenes as S

pytestmark = pytest.mark.gpu


def _err(a, ref, floor=0.0):
    a = np.asarray(a, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    return float(np.max(np.abs(a - ref)) / max(float(np.max(np.abs(ref))), floor, 1e-300))


def _rocm():
    from zpc_amd import rocm

    return rocm


def _oracle(sc, steps, dtype):
    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp).astype(dtype)
    O.Context(dtype).run(O.make_params(sc), ps, steps, reorder=True)
    return ps


ROWS = []


def _record(case, field, gpu, r32, r64, tol, floor=0.0):
    e_gpu = _err(gpu, r64, floor)
    e_32 = _err(r32, r64, floor)
    scale = max(float(np.max(np.abs(np.asarray(r64, np.float64)))), floor)
    ROWS.append((case, field, scale, e_gpu, e_32))
    print(f"{case:34s} {field:6s} scale {scale:.2e}  gpu {e_gpu:.2e}  fp32 oracle {e_32:.2e}", flush=True)
    assert e_gpu <= max(tol, 2.0 * e_32), f"{case} {field}: gpu {e_gpu:.3e}, fp32 oracle {e_32:.3e}"


@pytest.mark.parametrize("steps,perturbed", [(1, False), (100, False), (1, True)])
def test_particle_fields_against_fp64(steps, perturbed):
    rocm = _rocm()
    sc = S.sand_scene(dims=(16, 32, 16))
    if perturbed:
        sc = S.perturb(sc, seed=31, f_scale=0.02, c_scale=2.0, v_scale=0.3)
    case = f"{'perturbed' if perturbed else 'at rest'} column, {steps} substep{'s' if steps > 1 else ''}"
    sim = rocm.Simulator(sc, reorder=True)
    sim.substeps(steps)
    sim.sync()
    got = sim.parts.host_by_id()
    r32 = _oracle(sc, steps, np.float32)
    r64 = _oracle(sc, steps, np.float64)
    tol = 1e-5 if steps == 1 else 1e-3
    v64 = r64.by_id("v")
    floors = {"C": float(np.max(np.abs(v64))) / sc.dx, "logJp": 1.0}
    for name in ("x", "v", "C", "F", "logJp"):
        _record(case, name, got[name], r32.by_id(name), r64.by_id(name), tol, floors.get(name, 0.0))


def test_grid_fields_against_fp64():
    rocm = _rocm()
    sc = S.perturb(S.sand_scene(dims=(16, 32, 16)), seed=31, f_scale=0.02, c_scale=2.0, v_scale=0.3)
    sim = rocm.Simulator(sc)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    B = sim.nblocks()
    g = sim.grid.data[:B].cpu().numpy()
    prm = O.make_params(sc)
    ctx = O.Context()
    ctx.activate(sc.table_size, sc.x, sc.dx)
    ctx.enlarge()
    ctx.bucket_sort(sc.x, sc.dx)
    assert ctx.sizes()["cnt"] == B
    ps32 = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    ctx.p2g(prm, ps32)
    g32, _, _ = ctx.grids()
    ps64 = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp).astype(np.float64)
    g64 = ctx.p2g_f64(prm, ps64)
    case = "perturbed column, one P2G (grid)"
    _record(case, "m", g[:, 0], g32[:, 0], g64[:, 0], 1e-5)
    _record(case, "mv", g[:, 1:4], g32[:, 1:4], g64[:, 1:4], 1e-5)
    _record(case, "force", g[:, 4:7], g32[:, 4:7], g64[:, 4:7], 1e-5)


def test_zz_write_report():
    """Writes the table of the tests above (run in the same session)."""
    if not ROWS:
        pytest.skip("no rows recorded in this session")
    lines = ["| case | field | scale | GPU vs FP64 | FP32 oracle vs FP64 |", "|---|---|---|---|---|"]
    lines += [f"| {c} | {f} | {sc:.2e} | {e:.2e} | {r:.2e} |" for c, f, sc, e, r in ROWS]
    text = "\n".join(lines) + "\n"
    print(text, flush=True)
    out = os.environ.get("ZSMPM_REPORT_DIR", "")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "field_errors.md"), "w") as fh:
            fh.write(text)
