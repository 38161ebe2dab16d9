"""Position pins of the oracle's back projectors.

The BP of a one-hot detector pixel must be non-zero exactly on the voxels
whose centre projects (along the ray through it: from the source for
cone / fan, along the ray direction for par3d / parallel) within one pixel
of the hot pixel's centre.  The projection is a closed-form line / plane
intersection in world units from the raw geometry vectors, independent of
the oracle's BP coefficients (cuda/3d/cone_bp.cu:221-302,
cuda/3d/par3d_bp.cu:202-244, cuda/2d/fan_bp.cu:225-279,
cuda/2d/par_bp.cu:163-191) and of the volume normalisation
(src/GeometryUtil3D.cpp:569-593, src/GeometryUtil2D.cpp:181-209).

Unlike the adjoint-criterion pins (dominated by the data means), these fail
for a mirrored detector axis, a one-pixel shift, swapped u / v or x / y, a
flipped image row axis, or a wrong window offset; each test also plants
such a mistake in the oracle's input and checks that the pin catches it.
"""
import numpy as np

from oracle import geometry as OG, proj2d as O2, proj3d as O3


def _support_ok(bp, coords, centre, margin=0.02):
    d = [np.abs(a - c) for a, c in zip(coords, centre)]
    inside = np.logical_and.reduce([x < 1.0 - margin for x in d])
    outside = np.logical_or.reduce([x > 1.0 + margin for x in d])
    assert inside.sum() > 3
    return bool((bp[inside] != 0).all() and (bp[outside] == 0).all())


# volume with a shifted, non-centred window and unequal sizes per axis
# (x: 30 from -14, y: 34 from -19, z: 26 from -10), unit voxels
VG3 = {"GridColCount": 30, "GridRowCount": 34, "GridSliceCount": 26,
       "option": {"WindowMinX": -14.0, "WindowMaxX": 16.0, "WindowMinY": -19.0, "WindowMaxY": 15.0,
                  "WindowMinZ": -10.0, "WindowMaxZ": 16.0}}
SHAPE3 = (26, 34, 30)


def _world3():
    xs = -14 + np.arange(30) + 0.5
    ys = -19 + np.arange(34) + 0.5
    zs = -10 + np.arange(26) + 0.5
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X, Y, Z], axis=-1)


def _cone_coords(raw, W):
    s, d, u, v = raw[0:3], raw[3:6], raw[6:9], raw[9:12]
    M = np.stack([np.broadcast_to(u, W.shape), np.broadcast_to(v, W.shape), -(W - s)], axis=-1)
    sol = np.linalg.solve(M, np.broadcast_to(s - d, W.shape)[..., None])[..., 0]
    return sol[..., 0], sol[..., 1]


def _par3d_coords(raw, W):
    r, d, u, v = raw[0:3], raw[3:6], raw[6:9], raw[9:12]
    M = np.stack([np.broadcast_to(u, W.shape), np.broadcast_to(v, W.shape),
                  -np.broadcast_to(r, W.shape)], axis=-1)
    sol = np.linalg.solve(M, (W - d)[..., None])[..., 0]
    return sol[..., 0], sol[..., 1]


NU, NV, U0, V0 = 60, 40, 37, 12


def _one_hot(na, a, u0=U0, v0=V0):
    p = np.zeros((NV, na, NU))
    p[v0, a, u0] = 1.0
    return p


def test_cone_bp_positions():
    angles = np.array([0.3, 1.9, 4.1, 5.5], dtype=np.float32)
    raw = OG.gen_cone_projections(angles, NU, NV, 120.0, 60.0, 1.0, 1.0)
    vec, vs = OG.normalise_3d(raw, "cone", VG3)
    W = _world3()
    for kernel in ("default", "fdk", "2d"):
        for a in range(len(angles)):
            bp = O3.cone_bp(_one_hot(len(angles), a), vec, SHAPE3, vs, kernel=kernel)
            assert _support_ok(bp, _cone_coords(raw[a], W), (U0 + 0.5, V0 + 0.5)), (kernel, a)
    # planted mistakes: a mirrored u axis, and a one-pixel shift in v
    bad = O3.cone_bp(_one_hot(len(angles), 1, u0=NU - 1 - U0), vec, SHAPE3, vs)
    assert not _support_ok(bad, _cone_coords(raw[1], W), (U0 + 0.5, V0 + 0.5))
    bad = O3.cone_bp(_one_hot(len(angles), 2, v0=V0 + 1), vec, SHAPE3, vs)
    assert not _support_ok(bad, _cone_coords(raw[2], W), (U0 + 0.5, V0 + 0.5))


def test_cone_bp_positions_anisotropic_voxels():
    """Non-cube voxels (x 1.25, y 0.8, z 1.5): the voxel centres in world
    units follow the window, the BP support follows them."""
    vg = {"GridColCount": 24, "GridRowCount": 40, "GridSliceCount": 16,
          "option": {"WindowMinX": -15.0, "WindowMaxX": 15.0, "WindowMinY": -18.0, "WindowMaxY": 14.0,
                     "WindowMinZ": -9.0, "WindowMaxZ": 15.0}}
    xs = -15 + (np.arange(24) + 0.5) * 1.25
    ys = -18 + (np.arange(40) + 0.5) * 0.8
    zs = -9 + (np.arange(16) + 0.5) * 1.5
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    W = np.stack([X, Y, Z], axis=-1)
    angles = np.array([0.7, 2.6], dtype=np.float32)
    raw = OG.gen_cone_projections(angles, NU, NV, 110.0, 55.0, 1.0, 1.0)
    vec, vs = OG.normalise_3d(raw, "cone", vg)
    for a in range(len(angles)):
        bp = O3.cone_bp(_one_hot(len(angles), a), vec, (16, 40, 24), vs)
        assert _support_ok(bp, _cone_coords(raw[a], W), (U0 + 0.5, V0 + 0.5)), a


def test_par3d_bp_positions():
    angles = np.array([0.2, 1.3, 2.9], dtype=np.float32)
    raw = OG.gen_par3d_projections(angles, NU, NV, 1.0, 1.0)
    vec, vs = OG.normalise_3d(raw, "parallel3d", VG3)
    W = _world3()
    for a in range(len(angles)):
        bp = O3.par3d_bp(_one_hot(len(angles), a), vec, SHAPE3, vs)
        assert _support_ok(bp, _par3d_coords(raw[a], W), (U0 + 0.5, V0 + 0.5)), a
    bad = O3.par3d_bp(_one_hot(len(angles), 0, u0=NU - 1 - U0), vec, SHAPE3, vs)
    assert not _support_ok(bad, _par3d_coords(raw[0], W), (U0 + 0.5, V0 + 0.5))


# 2D: 28 rows x 36 columns, window x in [-20, 16], y in [-11, 17]; image
# row r sits at world y = 17 - (r + 0.5)
VG2 = {"GridColCount": 36, "GridRowCount": 28,
       "option": {"WindowMinX": -20.0, "WindowMaxX": 16.0, "WindowMinY": -11.0, "WindowMaxY": 17.0}}
ND2, D0 = 50, 31


def _world2():
    xs = -20 + np.arange(36) + 0.5
    ys = 17 - (np.arange(28) + 0.5)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    return np.stack([X, Y], axis=-1)


def _coords2(raw, W, fan):
    r, d, u = raw[0:2], raw[2:4], raw[4:6]
    dirn = (W - r) if fan else np.broadcast_to(r, W.shape)
    base = np.broadcast_to(r, W.shape) if fan else W
    M = np.stack([np.broadcast_to(u, W.shape), -dirn], axis=-1)
    return np.linalg.solve(M, (base - d)[..., None])[..., 0, 0]


def _hot2(na, a, d0=D0):
    s = np.zeros((na, ND2))
    s[a, d0] = 1.0
    return s


def test_fan_bp_positions():
    angles = np.array([0.4, 2.2, 3.7, 5.9], dtype=np.float32)
    raw = OG.gen_fan_projections(angles, ND2, 90.0, 45.0, 1.0)
    vec, osc = OG.normalise_2d(raw, "fan", VG2)
    W = _world2()
    for fbp in (False, True):
        for a in range(len(angles)):
            bp = O2.fan_bp(_hot2(len(angles), a), vec, (28, 36), osc, fbp_weighted=fbp)
            assert _support_ok(bp, (_coords2(raw[a], W, True),), (D0 + 0.5,)), (fbp, a)
    bad = O2.fan_bp(_hot2(len(angles), 1, d0=ND2 - 1 - D0), vec, (28, 36), osc)
    assert not _support_ok(bad, (_coords2(raw[1], W, True),), (D0 + 0.5,))
    # a flipped image row axis (rows at world y = -11 + r + 0.5) is caught
    bp = O2.fan_bp(_hot2(len(angles), 0), vec, (28, 36), osc)
    assert not _support_ok(bp[::-1], (_coords2(raw[0], W, True),), (D0 + 0.5,))


def test_par_bp_positions():
    angles = np.array([0.4, 1.2, 2.7], dtype=np.float32)
    raw = OG.gen_par_projections(angles, ND2, 1.0)
    vec, osc = OG.normalise_2d(raw, "parallel", VG2)
    W = _world2()
    for a in range(len(angles)):
        bp = O2.par_bp(_hot2(len(angles), a), vec, (28, 36), osc)
        assert _support_ok(bp, (_coords2(raw[a], W, False),), (D0 + 0.5,)), a
    bad = O2.par_bp(_hot2(len(angles), 2, d0=D0 - 1), vec, (28, 36), osc)
    assert not _support_ok(bad, (_coords2(raw[2], W, False),), (D0 + 0.5,))
