"""The oracle's general colliders (Collider<AnalyticLevelSet<...>> with a rigid
motion; geometry/Collider.h:10-190, geometry/AnalyticLevelSet.h:11-225)
pinned by closed forms, not by restating the code:
- signed distances of plane, cuboid, sphere and cylinder at points whose
  distance is known (inside: minus the nearest face; outside: the distance
  to the nearest face, edge or corner);
- normals: the sphere's radial direction, the cuboid's and cylinder's face
  normals from central differences;
- a sticky moving object gives its velocity omega x (x - b) + (s'/s)(x - b)
  + b' inside and leaves outside nodes alone;
- slip removes, separate removes only when approaching, the normal part of
  the velocity relative to the moving object;
- the transform: a cuboid rotated 90 degrees about z and scaled by 2 covers
  the world points it must and no others;
- a plane general collider with identity motion equals the plane collider
  path of the substep.
"""
import numpy as np

from oracle import oracle as O
from zpc_amd import scenes as S


def _rot_z(theta):
    c, s = np.cos(theta), np.sin(theta)
    M = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    return tuple(M.T.reshape(9))  # column-major


def test_sdf_closed_forms():
    sph = S.GeneralCollider(S.SPHERE, (0.1, -0.2, 0.3), (0.5, 0, 0))
    X = np.array([[0.1, -0.2, 0.3], [0.6, -0.2, 0.3], [0.1, 0.8, 0.3], [0.1 + 0.3, -0.2 + 0.4, 0.3]])
    assert np.allclose(O.level_set_sdf(sph, X), [-0.5, 0.0, 0.5, 0.0], atol=1e-14)
    cub = S.GeneralCollider(S.CUBOID, (-1.0, -2.0, -3.0), (1.0, 2.0, 3.0))
    X = np.array([[0.0, 0.0, 0.0],      # center: 1 from the x faces
                  [0.5, 1.8, 0.0],      # nearest face y = 2 at 0.2
                  [1.5, 0.0, 0.0],      # outside the x face by 0.5
                  [2.0, 3.0, 0.0],      # outside an edge: (1, 1)
                  [1.3, 2.4, 3.5]])     # outside a corner: (0.3, 0.4, 0.5)
    want = [-1.0, -0.2, 0.5, np.sqrt(2.0), np.sqrt(0.09 + 0.16 + 0.25)]
    assert np.allclose(O.level_set_sdf(cub, X), want, atol=1e-14)
    cyl = S.GeneralCollider(S.CYLINDER, (0.0, 0.0, 0.0), (1.0, 4.0, 0), axis=1)  # along y, radius 1, length 4
    X = np.array([[0.0, 2.0, 0.0],      # axis, mid-height: nearest is the side at 1
                  [0.0, 0.25, 0.0],     # near the bottom cap: 0.25
                  [0.5, 2.0, 0.0],      # 0.5 from the side
                  [3.0, 2.0, 0.0],      # 2 outside the side
                  [0.0, -1.0, 0.0],     # 1 below the bottom cap, inside the radius
                  [2.0, 5.0, 0.0]])     # beyond the top cap and the radius: (1, 1)
    want = [-1.0, -0.25, -0.5, 2.0, 1.0, np.sqrt(2.0)]
    assert np.allclose(O.level_set_sdf(cyl, X), want, atol=1e-14)
    pl = S.GeneralCollider(S.PLANE, (0.0, 0.5, 0.0), (0.0, 1.0, 0.0))
    assert np.allclose(O.level_set_sdf(pl, np.array([[3.0, 0.2, -1.0], [0.0, 0.9, 0.0]])), [-0.3, 0.4])


def test_normals():
    sph = S.GeneralCollider(S.SPHERE, (0.0, 0.0, 0.0), (1.0, 0, 0))
    X = np.array([[0.3, 0.4, 0.0], [0.0, 0.0, 0.0]])
    n = O.level_set_normal(sph, X)
    assert np.allclose(n[0], [0.6, 0.8, 0.0]) and np.allclose(n[1], 0.0)
    cub = S.GeneralCollider(S.CUBOID, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    n = O.level_set_normal(cub, np.array([[0.95, 0.1, -0.2], [0.1, -0.97, 0.3]]))
    assert np.allclose(n, [[1, 0, 0], [0, -1, 0]], atol=1e-6)
    cyl = S.GeneralCollider(S.CYLINDER, (0.0, 0.0, 0.0), (1.0, 4.0, 0), axis=1)
    n = O.level_set_normal(cyl, np.array([[0.0, 2.0, 0.9], [0.1, 3.9, 0.0]]))
    assert np.allclose(n, [[0, 0, 1], [0, 1, 0]], atol=1e-6)


def test_sticky_moving_object_velocity():
    rng = np.random.default_rng(0)
    omega, b, dbdt, s, dsdt = np.array([0.3, -1.2, 2.0]), np.array([0.4, 0.5, 0.6]), np.array([1.0, -2.0, 0.5]), 1.5, 0.7
    c = S.GeneralCollider(S.SPHERE, (0.0, 0.0, 0.0), (0.1, 0, 0), kind=S.STICKY, R=_rot_z(0.4), omega=tuple(omega),
                          s=s, dsdt=dsdt, b=tuple(b), dbdt=tuple(dbdt))
    x = b + rng.uniform(-0.3, 0.3, size=(400, 3))
    v = rng.normal(size=(400, 3))
    got = O.resolve_general(c, x, v)
    inside = np.linalg.norm(x - b, axis=1) < 0.1 * s  # world radius = s r
    assert inside.sum() > 5 and (~inside).sum() > 5
    xmb = x - b
    want = np.cross(omega, xmb) + (dsdt / s) * xmb + dbdt
    assert np.allclose(got[inside], want[inside], atol=1e-12)
    assert np.array_equal(got[~inside], v[~inside])


def test_slip_and_separate_relative_to_the_object():
    rng = np.random.default_rng(1)
    dbdt = np.array([0.5, 0.0, -0.25])
    x = np.array([0.2, 0.3, 0.1]) + rng.uniform(-0.05, 0.05, size=(300, 3))
    v = rng.normal(size=(300, 3))
    for kind in (S.SLIP, S.SEPARATE):
        c = S.GeneralCollider(S.SPHERE, (0.2, 0.3, 0.1), (0.2, 0, 0), kind=kind, dbdt=tuple(dbdt))
        got = O.resolve_general(c, x, v)
        n = (x - np.array([0.2, 0.3, 0.1]))
        n /= np.linalg.norm(n, axis=1)[:, None]
        rel_in = np.sum((v - dbdt) * n, axis=1)
        rel_out = np.sum((got - dbdt) * n, axis=1)
        tang_in = (v - dbdt) - rel_in[:, None] * n
        tang_out = (got - dbdt) - rel_out[:, None] * n
        assert np.allclose(tang_in, tang_out, atol=1e-12), "the tangential part is kept"
        if kind == S.SLIP:
            assert np.allclose(rel_out, 0.0, atol=1e-12)
        else:
            approaching = rel_in < 0
            assert approaching.any() and (~approaching).any()
            assert np.allclose(rel_out[approaching], 0.0, atol=1e-12)
            assert np.allclose(rel_out[~approaching], rel_in[~approaching], atol=1e-12)


def test_rotated_scaled_cuboid_covers_its_image():
    # material box [-1, 1] x [-0.5, 0.5] x [-0.5, 0.5], rotated 90 deg about z,
    # scaled by 2 and moved to b: in world space |x - b| <= (1, 2, 1)
    b = np.array([0.5, 0.5, 0.5])
    c = S.GeneralCollider(S.CUBOID, (-1.0, -0.5, -0.5), (1.0, 0.5, 0.5), kind=S.STICKY, R=_rot_z(np.pi / 2), s=2.0,
                          b=tuple(b), dbdt=(7.0, 7.0, 7.0))
    rng = np.random.default_rng(2)
    x = b + rng.uniform(-3, 3, size=(2000, 3))
    got = O.resolve_general(c, x, np.zeros_like(x))
    d = np.abs(x - b)
    inside = (d[:, 0] < 1.0) & (d[:, 1] < 2.0) & (d[:, 2] < 1.0)
    assert inside.sum() > 50
    assert np.allclose(got[inside], 7.0) and np.allclose(got[~inside], 0.0)


def test_plane_general_collider_equals_plane_collider_substep():
    sc = S.perturb(S.cube_scene(cells=10), seed=3)
    sc.colliders = [S.Collider((0.5, 0.44, 0.5), (0.0, 1.0, 0.0), S.SLIP)]
    ps_a = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    O.Context().run(O.make_params(sc), ps_a, 2)
    sc2 = S.perturb(S.cube_scene(cells=10), seed=3)
    sc2.colliders = []
    sc2.general_colliders = [S.GeneralCollider(S.PLANE, (0.5, 0.44, 0.5), (0.0, 1.0, 0.0), kind=S.SLIP)]
    ps_b = O.HostParticles(sc2.x, sc2.v, sc2.C, sc2.F, sc2.m, sc2.logJp)
    O.Context().run(O.make_params(sc2), ps_b, 2)
    for name in ("x", "v", "C", "F"):
        assert np.array_equal(getattr(ps_a, name), getattr(ps_b, name)), name
