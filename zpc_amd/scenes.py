"""Synthetic MLS-MPM scenes (host-side, deterministic).

Each scene is a seeded particle set in zpc's attribute layout (x, v, C, F, m,
logJp; C and F column-major 3x3) plus the step parameters the substep needs.
Particles are generated cell-major (i, j, k) with a fixed number per cell, each
jittered uniformly inside its cell by a splitmix64 stream (24-bit uniforms);
initial state follows Scene::commit (simulation/init/Scene.cpp:155-226):
v = 0 (or the scene's initial velocity), F = I, C = 0, logJp = logJp0,
m = rho * volume with volume = dx^3 / ppc.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

FIXED_COROTATED = 0
DRUCKER_PRAGER = 1
STICKY, SLIP, SEPARATE = 0, 1, 2

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start+1 .. start+count of the splitmix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform24(seed: int, count: int, start: int = 0) -> np.ndarray:
    """24-bit uniforms in [0, 1) as float64 (exact)."""
    return (splitmix64(seed, count, start) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


@dataclasses.dataclass
class Material:
    """FixedCorotatedConfig / DruckerPragerConfig (physics/ConstitutiveModel.hpp:736-754)."""

    kind: int = FIXED_COROTATED
    E: float = 5e4
    nu: float = 0.4
    rho: float = 1e3
    volume: float = 1.0
    cohesion: float = 0.0
    beta: float = 1.0
    yield_surface: float = float(np.float32(0.816496580927726) * np.float32(2.0) * np.float32(0.5) /
                                 (np.float32(3.0) - np.float32(0.5)))
    volume_correction: bool = True
    logJp0: float = 0.0

    @staticmethod
    def drucker_prager(E, nu, rho, friction_deg=30.0, **kw):
        s = math.sin(math.radians(friction_deg))
        ys = math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)
        return Material(kind=DRUCKER_PRAGER, E=E, nu=nu, rho=rho, yield_surface=ys, **kw)


@dataclasses.dataclass
class Collider:
    """Static plane collider (geometry/Collider.h:80-107, AnalyticLevelSet Plane)."""

    origin: Sequence[float]
    normal: Sequence[float]
    kind: int = STICKY


PLANE, CUBOID, SPHERE, CYLINDER = 0, 1, 2, 3


@dataclasses.dataclass
class GeneralCollider:
    """Collider<AnalyticLevelSet<geometry>> with a rigid motion
    (geometry/Collider.h:10-190, geometry/AnalyticLevelSet.h): the level set's
    material point X sits at x = R s X + b, and its velocity there is
    omega x (x - b) + (dsdt / s)(x - b) + dbdt. p0 / p1: plane origin and
    normal, cuboid min and max, sphere center and p1[0] = radius, cylinder
    bottom with p1[0] = radius and p1[1] = length along `axis`. R column-major."""

    geometry: int
    p0: Sequence[float]
    p1: Sequence[float]
    kind: int = STICKY
    axis: int = 1
    R: Sequence[float] = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    omega: Sequence[float] = (0.0, 0.0, 0.0)
    s: float = 1.0
    dsdt: float = 0.0
    b: Sequence[float] = (0.0, 0.0, 0.0)
    dbdt: Sequence[float] = (0.0, 0.0, 0.0)


@dataclasses.dataclass
class Scene:
    name: str
    dx: float
    dt: float
    material: Material
    x: np.ndarray          # (N, 3) float32
    v: np.ndarray          # (N, 3) float32
    C: np.ndarray          # (N, 9) float32
    F: np.ndarray          # (N, 9) float32
    m: np.ndarray          # (N,) float32
    logJp: Optional[np.ndarray]  # (N,) float32 or None
    gravity: Sequence[float] = (0.0, -9.8, 0.0)
    colliders: List[Collider] = dataclasses.field(default_factory=list)
    table_size: int = 0
    steps: int = 100
    # applied after the plane colliders, in order (at most 8)
    general_colliders: List[GeneralCollider] = dataclasses.field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def box_walls(dx: float, cells: int = 3, kind: int = STICKY) -> List[Collider]:
    """Six planes `cells` grid cells inside the faces of the unit cube."""
    lo, hi = cells * dx, 1.0 - cells * dx
    out = []
    for d in range(3):
        o = [0.0, 0.0, 0.0]
        n = [0.0, 0.0, 0.0]
        o[d], n[d] = lo, 1.0
        out.append(Collider(tuple(o), tuple(n), kind))
        o = [0.0, 0.0, 0.0]
        n = [0.0, 0.0, 0.0]
        o[d], n[d] = hi, -1.0
        out.append(Collider(tuple(o), tuple(n), kind))
    return out


def sound_speed(E: float, nu: float, rho: float) -> np.float32:
    """evaluate_soundspeed_linear_elasticity (physics/SoundSpeedCfl.hpp:6-10) in
    float32: c = sqrt(E (1 - nu) / ((1 + nu) (1 - 2 nu) rho)), the root taken by
    math::sqrtNewtonRaphson (math/MathUtils.h:239-251: Newton from 1 until
    successive iterates differ by at most max(1e-6 n, 128 eps))."""
    f = np.float32
    E, nu, rho = f(E), f(nu), f(rho)
    n = E * (f(1) - nu) / ((f(1) + nu) * (f(1) - f(2) * nu) * rho)
    eps = f(128) * np.finfo(np.float32).eps
    if n < -eps:
        return f("nan")
    if n < eps:
        return f(0)
    xn = f(1)
    xnp1 = f(0.5) * (xn + n / xn)
    tol = max(n * f(1e-6), eps)
    while abs(xnp1 - xn) > tol:
        xn = xnp1
        xnp1 = f(0.5) * (xn + n / xn)
    return xnp1


def cfl_dt(material: "Material", dx: float, cfl: float = 0.5) -> float:
    """evaluate_timestep_linear_elasticity (physics/SoundSpeedCfl.hpp:12-15):
    cfl * dx / c in float32 -- the time step MPMSimulator's builder evaluates
    per model (simulation/mpm/Simulator.cpp:53-64)."""
    f = np.float32
    return float(f(cfl) * f(dx) / sound_speed(material.E, material.nu, material.rho))


def next_2pow(n: int) -> int:
    """math/bit/Bits.h:166-184."""
    return 1 if n <= 1 else 1 << (n - 1).bit_length()


def evaluate_table_size(n: int) -> int:
    """container/HashTable.hpp:88-91 (reserve ratio 16)."""
    return 0 if n == 0 else next_2pow(n) * 16


def jittered_cells(seed: int, lo_cell: Sequence[int], dims: Sequence[int], dx: float, ppc: int,
                   keep=None) -> np.ndarray:
    """Cell-major (i, j, k) jittered particles, `ppc` per cell, float32 positions."""
    nx, ny, nz = (int(d) for d in dims)
    ncell = nx * ny * nz
    n = ncell * ppc
    u = uniform24(seed, 3 * n).reshape(n, 3)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    cells = np.repeat(cells, ppc, axis=0) + np.asarray(lo_cell, np.float64)
    pos = (cells + u) * dx
    if keep is not None:
        pos = pos[keep(pos)]
    return pos.astype(np.float32)


def _state(x: np.ndarray, material: Material, vel=(0.0, 0.0, 0.0)):
    n = x.shape[0]
    v = np.tile(np.asarray(vel, np.float32), (n, 1))
    C = np.zeros((n, 9), np.float32)
    F = np.tile(np.eye(3, dtype=np.float32).reshape(9), (n, 1))
    m = np.full(n, np.float32(material.rho * material.volume), np.float32)
    logJp = (np.full(n, np.float32(material.logJp0), np.float32)
             if material.kind == DRUCKER_PRAGER else None)
    return v, C, F, m, logJp


def _table_size_for(x: np.ndarray, dx: float, slack: float = 2.0) -> int:
    """T = evaluate_table_size(nExpected) with nExpected = slack x the enlarged block count."""
    c = np.floor(x.astype(np.float32) * np.float32(1.0 / dx) + np.float32(0.5)).astype(np.int64) - 2
    b = np.floor_divide(c, 4) + (1 << 20)  # non-negative, 21 bits per axis

    def pack(k):
        return (k[..., 0] << 42) | (k[..., 1] << 21) | k[..., 2]

    keys = np.unique(pack(b))
    kb = np.stack([keys >> 42, (keys >> 21) & ((1 << 21) - 1), keys & ((1 << 21) - 1)], axis=1)
    offs = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], np.int64)
    enl = np.unique(pack(kb[:, None, :] + offs[None]))
    return evaluate_table_size(int(math.ceil(slack * enl.shape[0])))


def cube_scene(cells: int = 40, seed: int = 20261201 + 1, dx: float = 1.0 / 128, ppc: int = 8,
               steps: int = 100) -> Scene:
    """Config 1: jittered cube of cells^3 cells at 8 ppc, lower corner (0.25,)*3,
    fixed-corotated E=1e4 nu=0.3 rho=1e3, dt=1e-4, sticky box walls."""
    lo = int(round(0.25 / dx))
    x = jittered_cells(seed, (lo, lo, lo), (cells,) * 3, dx, ppc)
    mat = Material(kind=FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=dx ** 3 / ppc)
    v, C, F, m, logJp = _state(x, mat)
    return Scene(f"cube{cells}", dx, 1e-4, mat, x, v, C, F, m, logJp, colliders=box_walls(dx),
                 table_size=_table_size_for(x, dx), steps=steps)


def spheres_scene(seed: int = 20261201 + 2, dx: float = 1.0 / 256, r: float = 0.15, ppc: int = 8,
                  speed: float = 2.0, steps: int = 200, centers=(0.3, 0.7)) -> Scene:
    """Config 2: two elastic spheres r=0.15 at (0.3,0.5,0.5) and (0.7,0.5,0.5)
    moving at +-2 m/s along x; reference default fixed-corotated E=5e4 nu=0.4.
    Particles of the first sphere come first. `centers` moves the spheres
    along x (smaller scenes that collide within a few hundred substeps)."""
    mat = Material(kind=FIXED_COROTATED, E=5e4, nu=0.4, rho=1e3, volume=dx ** 3 / ppc)
    xs, vs = [], []
    for ci, (cx, sgn) in enumerate(((centers[0], 1.0), (centers[1], -1.0))):
        c = np.array([cx, 0.5, 0.5])
        lo = np.floor((c - r) / dx).astype(int)
        hi = np.ceil((c + r) / dx).astype(int)
        dims = hi - lo
        x = jittered_cells(seed + 1000 * ci, lo, dims, dx, ppc,
                           keep=lambda p, c=c: np.sum((p - c) ** 2, axis=1) <= r * r)
        xs.append(x)
        vs.append(np.tile(np.array([sgn * speed, 0, 0], np.float32), (x.shape[0], 1)))
    x = np.concatenate(xs)
    _, C, F, m, logJp = _state(x, mat)
    v = np.concatenate(vs).astype(np.float32)
    return Scene("spheres", dx, 1e-4, mat, x, v, C, F, m, logJp, colliders=box_walls(dx),
                 table_size=_table_size_for(x, dx), steps=steps)


def sand_scene(dims=(128, 256, 64), seed: int = 20261201 + 3, dx: float = 1.0 / 512, ppc: int = 8,
               steps: int = 100, dt: float = 5e-5) -> Scene:
    """Config 3: Drucker-Prager sand column (friction 30 deg, E=3.5e5, nu=0.3,
    rho=1e3) of dims cells at dx=1/512, 8 ppc (128x256x64 -> 16,777,216
    particles), resting on the floor, centred in x and z."""
    nx, ny, nz = dims
    n_cells = int(round(1.0 / dx))
    lo = ((n_cells - nx) // 2, 4, (n_cells - nz) // 2)
    x = jittered_cells(seed, lo, dims, dx, ppc)
    mat = Material.drucker_prager(3.5e5, 0.3, 1e3, 30.0, volume=dx ** 3 / ppc)
    v, C, F, m, logJp = _state(x, mat)
    return Scene(f"sand{nx}x{ny}x{nz}", dx, dt, mat, x, v, C, F, m, logJp, colliders=box_walls(dx),
                 table_size=_table_size_for(x, dx), steps=steps)


def batch_scenes(count: int = 64, seed: int = 20261201 + 4, lo: int = 10_000, hi: int = 500_000,
                 ppc: int = 8) -> List[Scene]:
    """Config 4: `count` independent scenes with particle counts log-uniform in
    [lo, hi] (seeded), each a jittered box or sphere at 8 ppc resting in the
    middle of the unit cube, half fixed-corotated (E = 1e4, nu = 0.3) and half
    Drucker-Prager sand (E = 3.5e5, nu = 0.3, 30 deg), dx 1/128 or 1/256,
    dt = 1e-4, sticky walls, gravity along -y."""
    out = []
    for i in range(count):
        u = uniform24(seed + 7919 * i, 4)
        n_target = math.exp(math.log(lo) + float(u[0]) * (math.log(hi) - math.log(lo)))
        sphere = u[1] >= 0.5
        sand = u[2] >= 0.5
        dx = 1.0 / 128 if u[3] < 0.5 else 1.0 / 256
        cells = n_target / ppc
        if sphere:
            r = (3.0 * cells / (4.0 * math.pi)) ** (1.0 / 3.0) * dx
            c = np.array([0.5, 0.5, 0.5])
            lo_c = np.floor((c - r) / dx).astype(int)
            hi_c = np.ceil((c + r) / dx).astype(int)
            x = jittered_cells(seed + 31 * i + 1, lo_c, hi_c - lo_c, dx, ppc,
                               keep=lambda p, c=c, r=r: np.sum((p - c) ** 2, axis=1) <= r * r)
        else:
            side = max(2, int(round(cells ** (1.0 / 3.0))))
            lo_c = [int(round(0.5 / dx)) - side // 2] * 3
            x = jittered_cells(seed + 31 * i + 1, lo_c, (side,) * 3, dx, ppc)
        if sand:
            mat = Material.drucker_prager(3.5e5, 0.3, 1e3, volume=dx ** 3 / ppc)
        else:
            mat = Material(kind=FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=dx ** 3 / ppc)
        v, C, F, m, logJp = _state(x, mat)
        out.append(Scene(f"batch{i}_{'sphere' if sphere else 'box'}_{'sand' if sand else 'fc'}", dx, 1e-4, mat,
                         x, v, C, F, m, logJp, colliders=box_walls(dx), table_size=_table_size_for(x, dx)))
    return out


def tiny_scene(n_cells=(3, 3, 3), seed: int = 7, dx: float = 1.0 / 32, ppc: int = 2,
               material: Optional[Material] = None, vel=(0.3, -0.2, 0.1), lo=(12, 13, 14)) -> Scene:
    """A few dozen particles for brute-force checks."""
    x = jittered_cells(seed, lo, n_cells, dx, ppc)
    mat = material or Material(kind=FIXED_COROTATED, E=1e4, nu=0.3, rho=1e3, volume=dx ** 3 / ppc)
    v, C, F, m, logJp = _state(x, mat, vel)
    return Scene("tiny", dx, 1e-4, mat, x, v, C, F, m, logJp, colliders=box_walls(dx),
                 table_size=_table_size_for(x, dx), steps=1)


def strained(scene: Scene, seed: int = 0, compression: float = 0.05, shear: float = 0.03,
             noise: float = 0.002, closed: bool = True) -> Scene:
    """Return a copy at rest under a smooth initial strain: every particle's F is
    (1 - compression) I plus `shear` in the x-y entry, plus per-entry noise
    uniform in [-noise, noise] (numpy's seeded PCG64, fast at 16M particles).
    Strains of this size are beyond what one Newton step from X0 = 0
    certifies a priori, so the learned closure evaluates its network. With
    `closed` the colliders are six sticky planes on the faces of the
    particles' bounding box (a lid included): the compressed body stays
    compressed. Released instead, the sand flows apart (zpc never projects F
    of sand, so |C - I| passes 1 within tens of substeps). x, v, C, m and
    logJp are as in the scene."""
    rng = np.random.default_rng(seed)
    s = dataclasses.replace(scene)
    if closed:
        lo = scene.x.min(axis=0).astype(np.float64)
        hi = scene.x.max(axis=0).astype(np.float64)
        cols = []
        for d in range(3):
            nrm = [0.0, 0.0, 0.0]
            nrm[d] = 1.0
            cols.append(Collider(tuple(float(v) if i == d else 0.0 for i, v in enumerate(lo)), tuple(nrm), STICKY))
            nrm = [0.0, 0.0, 0.0]
            nrm[d] = -1.0
            cols.append(Collider(tuple(float(v) if i == d else 0.0 for i, v in enumerate(hi)), tuple(nrm), STICKY))
        s.colliders = cols
    F = rng.uniform(-noise, noise, size=(scene.n, 9)).astype(np.float32)
    F[:, 0] += 1.0 - compression
    F[:, 4] += 1.0 - compression
    F[:, 8] += 1.0 - compression
    F[:, 3] += shear  # column-major: entry (0, 1)
    s.F = F
    return s


def perturb(scene: Scene, seed: int = 99, f_scale: float = 0.05, c_scale: float = 5.0,
            v_scale: float = 0.5) -> Scene:
    """Return a copy with a deformed F, nonzero C and random v (seeded), so a
    single substep exercises every term of the transfer."""
    n = scene.n
    u = uniform24(seed, 21 * n).reshape(n, 21) * 2.0 - 1.0
    s = dataclasses.replace(scene)
    s.F = (np.eye(3).reshape(1, 9) + f_scale * u[:, :9]).astype(np.float32)
    s.C = (c_scale * u[:, 9:18]).astype(np.float32)
    s.v = (v_scale * u[:, 18:21]).astype(np.float32)
    s.x = scene.x.copy()
    s.m = scene.m.copy()
    s.logJp = None if scene.logJp is None else (0.05 * u[:, 0]).astype(np.float32)
    return s
