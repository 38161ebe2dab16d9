"""Pins of the FP64 oracle (CPU only).

Each test checks an oracle function against something it does not restate:
a closed form, a different formula for the same quantity (2D vs 3D, cone vs
parallel limit), the reference's own known answers, or an invariant.  A
plausible transcription slip (sign, index, transposed operand, dropped
factor) fails at least one of them.
"""
import math

import numpy as np
import pytest

from oracle import algorithms as OA, fdk as OF, geometry as OG, proj2d as O2, proj3d as O3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


# ------------------------------------------------------------- geometry ---
def test_cone_vectors_closed_form():
    """genConeProjections at angle 0 and pi/2 (src/GeometryUtil3D.cpp:42-83)."""
    v = OG.gen_cone_projections([0.0, np.pi / 2], 4, 2, 100.0, 50.0, 1.5, 2.0)
    assert np.allclose(v[0], [0, -100, 0, -3.0, 50, -2.0, 1.5, 0, 0, 0, 0, 2.0])
    # rotation by +90 degrees: (x, y) -> (-y, x)
    assert np.allclose(v[1], [100, 0, 0, -50, -3.0, -2.0, 0, 1.5, 0, 0, 0, 2.0], atol=1e-5)


def test_normalisation_translate_scale():
    vg = {"GridRowCount": 10, "GridColCount": 20, "GridSliceCount": 5,
          "option": {"WindowMinX": -5, "WindowMaxX": 15, "WindowMinY": 0, "WindowMaxY": 20,
                     "WindowMinZ": -2.5, "WindowMaxZ": 2.5}}
    v = np.array([[1.0, 2, 3, 4, 5, 6, 1, 0, 0, 0, 0, 1]])
    out, vs = OG.normalise_3d(v, "cone", vg)
    # voxel (1, 2, 1); centre (5, 10, 0): src (1-5, (2-10)/2, 3), corner (4-5, (5-10)/2, 6)
    assert vs == (1.0, 2.0, 1.0)
    assert np.allclose(out[0], [-4, -4, 3, -1, -2.5, 6, 1, 0, 0, 0, 0, 1])


# ------------------------------------------------------------ 2D FP/BP ---
def test_par_fp_angle0_column_sums():
    """python/astra/tests.py:27-48: at angle 0 a parallel FP is the column sum."""
    rng = np.random.default_rng(0)
    vol = rng.random((24, 24))
    vg = OG.vol_geom_2d(24, 24)
    vec, osc = OG.normalise_2d(OG.gen_par_projections([0.0], 24, 1.0), "par", vg)
    assert np.allclose(O2.par_fp(vol, vec, 24, osc)[0], vol.sum(axis=0), atol=1e-12)


def test_par2d_equals_par3d_single_slice():
    """2D parallel FP/BP (angle/offset trig form, row axis -y) against the 3D
    parallel projector (vector form, no row flip) on a one-slice volume."""
    rng = np.random.default_rng(1)
    n, nd = 20, 28
    angles = np.array([0.1, 0.5, 1.1, 2.0, 2.9])  # away from 45 degree ties
    vol = rng.random((n, n))
    v2, osc = OG.normalise_2d(OG.gen_par_projections(angles, nd, 0.9), "par", OG.vol_geom_2d(n, n))
    s2 = O2.par_fp(vol, v2, nd, osc)
    v3, vs = OG.normalise_3d(OG.gen_par3d_projections(angles, nd, 1, 0.9, 1.0), "par3d",
                             OG.vol_geom_3d(n, n, 1))
    vol3 = vol[::-1][None]  # 2D row r is world y = -(r - n/2 + 0.5)
    s3 = O3.par3d_fp(vol3, v3, nd, 1, vs)[0]
    assert rel(s2, s3) < 1e-6
    y = rng.random((len(angles), nd))
    b2 = O2.par_bp(y, v2, (n, n), osc)
    b3 = O3.par3d_bp(y[None], v3, (1, n, n), vs)[0][::-1]
    # the 2D BP weight is |r|/|d| = 1/w; the 3D one is 1/|u x v| = 1/(w * 1)
    assert rel(b2, b3) < 1e-6


def test_fan2d_equals_cone_single_slice():
    """2D fan FP (per-ray orientation) against the cone FP (per-angle
    orientation) on a one-slice volume, for a narrow fan whose rays all share
    the central ray's orientation."""
    rng = np.random.default_rng(2)
    n, nd = 16, 12
    angles = np.array([0.2, 1.0, 2.4, 3.6, 5.0])
    vol = rng.random((n, n))
    v2, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, 300.0, 200.0, 1.0), "fan",
                              OG.vol_geom_2d(n, n))
    for a in range(len(angles)):
        hor = {O2.fan_ray_is_horizontal(v2[a], t) for t in range(nd)}
        assert len(hor) == 1
    s2 = O2.fan_fp(vol, v2, nd, osc)
    v3, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, 1, 300.0, 200.0, 1.0, 1.0),
                             "cone", OG.vol_geom_3d(n, n, 1))
    s3 = O3.cone_fp(vol[::-1][None], v3, nd, 1, vs)[0]
    assert rel(s2, s3) < 1e-6


def test_fan_bp_matches_cone_2d_weighting():
    """The cone '2d_weighting' BP is built to approximate the fan BP on a
    single slice (cuda/3d/cone_bp.cu:263-270); with a one-row detector of
    unit height they coincide."""
    rng = np.random.default_rng(3)
    n, nd = 16, 24
    angles = np.linspace(0, 2 * np.pi, 7, endpoint=False)
    v2, osc = OG.normalise_2d(OG.gen_fan_projections(angles, nd, 100.0, 60.0, 1.0), "fan",
                              OG.vol_geom_2d(n, n))
    v3, vs = OG.normalise_3d(OG.gen_cone_projections(angles, nd, 1, 100.0, 60.0, 1.0, 1.0),
                             "cone", OG.vol_geom_3d(n, n, 1))
    y = rng.random((len(angles), nd))
    b2 = O2.fan_bp(y, v2, (n, n), osc)
    b3 = O3.cone_bp(y[None], v3, (1, n, n), vs, kernel="2d")[0][::-1]
    assert rel(b2, b3) < 1e-6


@pytest.mark.parametrize("geom", ["par", "fan"])
def test_2d_bp_is_approximate_adjoint(geom):
    """<A x, y> ~ <x, A^T y> for the unmatched BP (tests/python/functional/
    test_line2d.py:591-640 accepts 1e-1 for the GPU kernels)."""
    rng = np.random.default_rng(4)
    n, nd, na = 32, 48, 40
    vg = OG.vol_geom_2d(n, n)
    if geom == "par":
        ang = np.linspace(0, np.pi, na, endpoint=False)
        v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "par", vg)
        A = lambda x: O2.par_fp(x, v, nd, osc)
        AT = lambda y: O2.par_bp(y, v, (n, n), osc)
    else:
        ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
        v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, nd, 200.0, 100.0, 1.0), "fan", vg)
        A = lambda x: O2.fan_fp(x, v, nd, osc)
        AT = lambda y: O2.fan_bp(y, v, (n, n), osc)
    x = rng.random((n, n))
    y = rng.random((na, nd))
    lhs, rhs = np.sum(A(x) * y), np.sum(x * AT(y))
    assert abs(lhs - rhs) / abs(lhs) < 5e-2


# ------------------------------------------------------------ 3D FP/BP ---
def test_cone_fp_ones_is_chord_length():
    """For a constant volume every sample inside reads 1: the ray sum is the
    exact chord length n0 * |d| / |d_principal| (rays that stay inside)."""
    n = 24
    vec, vs = OG.normalise_3d(OG.gen_cone_projections([0.3, 2.0], 6, 6, 400.0, 200.0, 1.0, 1.0),
                              "cone", OG.vol_geom_3d(n, n, n))
    p = O3.cone_fp(np.ones((n, n, n)), vec, 6, 6, vs)
    for a in range(2):
        d0 = O3.cone_direction(vec[a], 6, 6)
        for iv in range(6):
            for iu in range(6):
                det = vec[a, 3:6] + (iu + .5) * vec[a, 6:9] + (iv + .5) * vec[a, 9:12]
                dvec = det - vec[a, 0:3]
                assert abs(p[iv, a, iu] - n * np.linalg.norm(dvec) / abs(dvec[d0])) < 1e-9


def test_cone_tends_to_par3d():
    """Cone with the source at 1e7 and the detector through the origin is
    the parallel geometry up to O(1e-6): FP, and BP including weights."""
    rng = np.random.default_rng(5)
    ang = [0.4, 1.3, 2.5]
    vc, _ = OG.normalise_3d(OG.gen_cone_projections(ang, 12, 10, 1e7, 0.0, 1.0, 1.0), "cone",
                            OG.vol_geom_3d(10, 12, 8))
    vp, _ = OG.normalise_3d(OG.gen_par3d_projections(ang, 12, 10, 1.0, 1.0), "par3d",
                            OG.vol_geom_3d(10, 12, 8))
    x = rng.random((8, 10, 12))
    assert rel(O3.cone_fp(x, vc, 12, 10), O3.par3d_fp(x, vp, 12, 10)) < 1e-5
    y = rng.random((10, 3, 12))
    assert rel(O3.cone_bp(y, vc, (8, 10, 12)), O3.par3d_bp(y, vp, (8, 10, 12))) < 1e-5


def test_par3d_fp_angle0_sums():
    rng = np.random.default_rng(6)
    x = rng.random((6, 8, 10))
    v, vs = OG.normalise_3d(OG.gen_par3d_projections([0.0], 10, 6, 1.0, 1.0), "par3d",
                            OG.vol_geom_3d(8, 10, 6))
    assert np.allclose(O3.par3d_fp(x, v, 10, 6, vs)[:, 0, :], x.sum(axis=1), atol=1e-12)


def test_cone_bp_is_approximate_adjoint():
    """3D adjoint check (tests/python/functional/test_rec_scaling.py:147-174
    accepts 1e-1)."""
    rng = np.random.default_rng(7)
    n, na, nd = 16, 24, 24
    v, vs = OG.normalise_3d(OG.gen_cone_projections(np.linspace(0, 2 * np.pi, na, False),
                                                    nd, nd, 80.0, 40.0, 1.0, 1.0),
                            "cone", OG.vol_geom_3d(n, n, n))
    x = rng.random((n, n, n))
    y = rng.random((nd, na, nd))
    lhs = np.sum(O3.cone_fp(x, v, nd, nd, vs) * y)
    rhs = np.sum(x * O3.cone_bp(y, v, (n, n, n), vs))
    assert abs(lhs - rhs) / abs(lhs) < 5e-2


def test_noncube_fp_scale_matches_physical_length():
    """SCALE_NONCUBE: for a constant volume the FP is the physical chord
    length through the box, whatever the voxel aspect ratio."""
    vg = {"GridRowCount": 8, "GridColCount": 8, "GridSliceCount": 8,
          "option": {"WindowMinX": -8, "WindowMaxX": 8, "WindowMinY": -4, "WindowMaxY": 4,
                     "WindowMinZ": -2, "WindowMaxZ": 2}}
    ang = [0.25]
    vec, vs = OG.normalise_3d(OG.gen_par3d_projections(ang, 3, 3, 0.5, 0.5), "par3d", vg)
    p = O3.par3d_fp(np.ones((8, 8, 8)), vec, 3, 3, vs)
    # ray along (-sin a, cos a, 0); principal axis y (|cos| > |sin|): the box
    # is 8 units tall in y, so the chord is 8 / cos a for rays that stay inside
    assert abs(p[1, 0, 1] - 8.0 / math.cos(0.25)) < 1e-9


# --------------------------------------------------------------- filter ---
def test_ramp_filter_real_space_kernel():
    """The inverse DFT of the Ram-Lak spectrum is the reference test's
    real-space kernel 2h: 1/2 at 0, -2/(pi j)^2 at odd j
    (tests/python/unit/test_algorithms_2d.py:67-76)."""
    nd = 24
    H = OF.ramp_filter(nd)
    n = 2 * (len(H) - 1)
    k = np.fft.irfft(H, n=n)
    assert abs(k[0] - 0.5) < 1e-12
    for j in range(1, n // 2):
        want = -2.0 / (math.pi * j) ** 2 if j % 2 else 0.0
        assert abs(k[j] - want) < 1e-12
        assert abs(k[n - j] - want) < 1e-12


def test_filter_rows_is_zero_padded_convolution():
    """FFT filtering of zero-padded rows = direct circular convolution with
    the real-space kernel, truncated to the row (cuda/2d/fft.cu:98-131)."""
    rng = np.random.default_rng(8)
    nd = 20
    H = OF.ramp_filter(nd)
    n = 2 * (len(H) - 1)
    kern = np.fft.irfft(H, n=n)
    rows = rng.random((3, nd))
    got = OF.filter_rows(rows, H)
    want = np.zeros_like(rows)
    for r in range(3):
        for i in range(nd):
            want[r, i] = sum(rows[r, j] * kern[(i - j) % n] for j in range(nd))
    assert np.allclose(got, want, atol=1e-12)


def test_fdk_preweight_centre_pixel():
    """devFDK_preweight closed form at the detector centre: W / D."""
    na, nu, nv = 10, 4, 4
    vec, _ = OG.normalise_3d(OG.gen_cone_projections(np.linspace(0, 2 * np.pi, na, False),
                                                     nu, nv, 300.0, 100.0, 1.0, 1.0),
                             "cone", OG.vol_geom_3d(8, 8, 8))
    w = OF.fdk_preweight(np.ones((nv, na, nu)), vec)
    D = 400.0
    W = D * (D / (1.0 * 300.0)) * (math.pi / 2) / na
    # pixels (1,1),(1,2),(2,1),(2,2) sit at u, v = +-0.5
    assert abs(w[1, 0, 1] - W / math.sqrt(D * D + 0.5)) < 1e-12


@pytest.mark.slow
def test_fdk_cube_scaling():
    """FDK of a unit cube recovers density 1 +- 5e-2 in its centre
    (tests/python/functional/test_rec_scaling.py:89-145)."""
    n, na, nd = 32, 90, 48
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vec, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nd, nd, 200.0, 100.0, 1.0, 1.0),
                              "cone", OG.vol_geom_3d(n, n, n))
    vol = np.zeros((n, n, n))
    vol[12:20, 12:20, 12:20] = 1.0
    proj = O3.cone_fp(vol, vec, nd, nd, vs)
    rec = OF.fdk(proj, vec, (n, n, n), vs)
    assert abs(rec[14:18, 14:18, 14:18].mean() - 1.0) < 5e-2


def test_fbp_square_scaling():
    """Parallel and fan FBP of a square recover 1 +- 5e-2 in its centre."""
    n = 64
    vol = np.zeros((n, n))
    vol[26:38, 26:38] = 1.0
    vg = OG.vol_geom_2d(n, n)
    ang = np.linspace(0, np.pi, 120, endpoint=False)
    v, osc = OG.normalise_2d(OG.gen_par_projections(ang, 96, 1.0), "par", vg)
    rec = OF.fbp_par(O2.par_fp(vol, v, 96, osc), v, (n, n), osc, 1.0)
    assert abs(rec[29:35, 29:35].mean() - 1.0) < 5e-2
    ang = np.linspace(0, 2 * np.pi, 180, endpoint=False)
    v, osc = OG.normalise_2d(OG.gen_fan_projections(ang, 128, 300.0, 200.0, 1.0), "fan", vg)
    rec = OF.fbp_fan(O2.fan_fp(vol, v, 128, osc), v, (n, n), osc, 1.0)
    assert abs(rec[29:35, 29:35].mean() - 1.0) < 5e-2


# ------------------------------------------------------------ algorithms -
def test_sirt_and_cgls_reduce_the_residual():
    rng = np.random.default_rng(9)
    n, nd = 24, 36
    ang = np.linspace(0, np.pi, 30, endpoint=False)
    v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "par", OG.vol_geom_2d(n, n))
    A = lambda x: O2.par_fp(x, v, nd, osc)
    AT = lambda y: O2.par_bp(y, v, (n, n), osc)
    x_true = rng.random((n, n))
    b = A(x_true)
    x0 = np.zeros((n, n))
    r0 = OA.residual_norm(A, b, x0)
    r_s = [OA.residual_norm(A, b, OA.sirt(A, AT, b, x0, k)) for k in (2, 8)]
    r_c = [OA.residual_norm(A, b, OA.cgls(A, AT, b, x0, k)) for k in (2, 8)]
    assert r0 > r_s[0] > r_s[1]
    assert r0 > r_c[0] > r_c[1]
    # SIRT with a non-negativity / upper bound keeps the iterate inside
    xc = OA.sirt(A, AT, b, x0, 5, min_c=0.2, max_c=0.7)
    assert xc.min() >= 0.2 and xc.max() <= 0.7


# ---------------------------------------------------------------- matched adjoints
# The oracle's exact transposes (cone_fp_T, par3d_fp_T, par_fp_T, fan_fp_T)
# are pinned in FP64 by the dot-product identity <A x, y> = <x, A^T y> on
# random data (a wrong index, weight or scale breaks it far above 1e-12),
# and on a tiny case against the explicit matrix of the forward projector.

def _dot_rel(fp, fpT, xshape, yshape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(xshape)
    y = rng.standard_normal(yshape)
    lhs = float(np.sum(fp(x) * y))
    rhs = float(np.sum(x * fpT(y)))
    return abs(lhs - rhs) / abs(lhs)


def test_cone_transpose_dot_product_fp64():
    n, na, nu, nv = 12, 9, 14, 13
    v, vs = OG.normalise_3d(OG.gen_cone_projections(np.linspace(0, 2 * np.pi, na, False),
                                                    nu, nv, 30.0, 20.0, 1.5, 1.2),
                            "cone", OG.vol_geom_3d(n, n, n))
    rel = _dot_rel(lambda x: O3.cone_fp(x, v, nu, nv, vs), lambda y: O3.cone_fp_T(y, v, (n, n, n), vs),
                   (n, n, n), (nv, na, nu), 1)
    assert rel < 1e-12


def test_cone_vec_transpose_all_directions_fp64():
    # steep rays put angles in all three direction groups
    n, nu, nv = 10, 12, 12
    ang = np.linspace(0, np.pi, 6, endpoint=False)
    vecs = []
    for a in ang:
        s = np.array([np.sin(a) * 20, -np.cos(a) * 20, 0.0])
        vecs.append(np.concatenate([s, -s / 2.0, [np.cos(a), np.sin(a), 0.0], [0, 0, 1.0]]))
    vecs.append(np.array([0, 0, 25.0, 0, 0, -10.0, 1.0, 0, 0, 0, 1.0, 0]))  # along z
    v, vs = OG.normalise_3d(OG.conevec_from_astra(np.array(vecs), nu, nv), "cone", OG.vol_geom_3d(n, n, n))
    assert {O3.cone_direction(x, nu, nv) for x in v} == {0, 1, 2}
    rel = _dot_rel(lambda x: O3.cone_fp(x, v, nu, nv, vs), lambda y: O3.cone_fp_T(y, v, (n, n, n), vs),
                   (n, n, n), (nv, len(v), nu), 2)
    assert rel < 1e-12


def test_par3d_transpose_dot_product_fp64():
    n, na, nu, nv = 10, 8, 13, 12
    v, vs = OG.normalise_3d(OG.gen_par3d_projections(np.linspace(0, np.pi, na, False), nu, nv, 1.1, 0.9),
                            "parallel3d", OG.vol_geom_3d(n, n, n))
    rel = _dot_rel(lambda x: O3.par3d_fp(x, v, nu, nv, vs), lambda y: O3.par3d_fp_T(y, v, (n, n, n), vs),
                   (n, n, n), (nv, na, nu), 3)
    assert rel < 1e-12


def test_supersampled_3d_transposes_dot_product_fp64():
    n, na, nu, nv = 9, 5, 10, 11
    ang = np.linspace(0, 2 * np.pi, na, False)
    v, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nu, nv, 30.0, 20.0, 1.5, 1.2),
                            "cone", OG.vol_geom_3d(n, n, n))
    rel = _dot_rel(lambda x: O3.cone_fp(x, v, nu, nv, vs, ss=2),
                   lambda y: O3.cone_fp_T(y, v, (n, n, n), vs, ss=2), (n, n, n), (nv, na, nu), 4)
    assert rel < 1e-12
    v3, vs3 = OG.normalise_3d(OG.gen_par3d_projections(ang, nu, nv, 1.1, 0.9), "parallel3d",
                              OG.vol_geom_3d(n, n, n))
    rel = _dot_rel(lambda x: O3.par3d_fp(x, v3, nu, nv, vs3, ss=3),
                   lambda y: O3.par3d_fp_T(y, v3, (n, n, n), vs3, ss=3), (n, n, n), (nv, na, nu), 5)
    assert rel < 1e-12


def test_2d_transposes_dot_product_fp64():
    n, na, nd = 17, 11, 25
    ang = np.linspace(0, np.pi, na, False)
    vg = OG.vol_geom_2d(n, n)
    v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.3), "parallel", vg)
    rel = _dot_rel(lambda x: O2.par_fp(x, v, nd, osc), lambda y: O2.par_fp_T(y, v, (n, n), osc),
                   (n, n), (na, nd), 4)
    assert rel < 1e-12
    v, osc = OG.normalise_2d(OG.gen_fan_projections(np.linspace(0, 2 * np.pi, na, False), nd, 40.0, 20.0, 1.4),
                             "fan", vg)
    rel = _dot_rel(lambda x: O2.fan_fp(x, v, nd, osc), lambda y: O2.fan_fp_T(y, v, (n, n), osc),
                   (n, n), (na, nd), 5)
    assert rel < 1e-12


def test_cone_transpose_equals_explicit_matrix():
    """Brute force on a tiny case: the matrix of cone_fp (one column per
    unit voxel) transposed equals cone_fp_T applied to unit projections."""
    n, na, nu, nv = 4, 3, 5, 4
    v, vs = OG.normalise_3d(OG.gen_cone_projections(np.linspace(0, 2 * np.pi, na, False),
                                                    nu, nv, 12.0, 8.0, 1.5, 1.5),
                            "cone", OG.vol_geom_3d(n, n, n))
    cols = []
    for i in range(n ** 3):
        e = np.zeros(n ** 3)
        e[i] = 1.0
        cols.append(O3.cone_fp(e.reshape(n, n, n), v, nu, nv, vs).ravel())
    A = np.stack(cols, axis=1)
    rows = []
    for j in range(nv * na * nu):
        e = np.zeros(nv * na * nu)
        e[j] = 1.0
        rows.append(O3.cone_fp_T(e.reshape(nv, na, nu), v, (n, n, n), vs).ravel())
    AT = np.stack(rows, axis=0)
    assert np.abs(A).max() > 0
    np.testing.assert_allclose(AT, A, rtol=0, atol=1e-12 * np.abs(A).max())


# ---------------------------------------------------------------- filter windows
# The windows of genFilter (src/Filters.cpp:151-472) over the nb = N/2 + 1
# bins, checked against independent library windows (scipy.signal.windows)
# where the reference's formula is a textbook window on nb (or 2 nb - 1)
# points, and against closed-form values at w = pi / the centre bin for the
# frequency-domain ones.

def _win(name, n_det=50, **kw):
    n = OF.next_pow2(2 * n_det)
    nb = n // 2 + 1
    return OF.filter_window(name, nb, n, **kw), nb, n


def test_windows_equal_library_windows():
    from scipy.signal import windows as W
    w, nb, _ = _win("gaussian")
    np.testing.assert_allclose(w, W.gaussian(nb, 0.4 * (nb - 1) / 2)[1:], atol=1e-12)
    np.testing.assert_allclose(_win("blackman")[0], W.blackman(2 * nb - 1)[1:nb], atol=1e-12)
    np.testing.assert_allclose(_win("blackman-harris")[0], W.blackmanharris(nb)[1:], atol=1e-12)
    np.testing.assert_allclose(_win("triangular")[0], W.bartlett(nb)[1:], atol=1e-12)
    np.testing.assert_allclose(_win("barlett-hann")[0], W.barthann(nb)[1:], atol=1e-12)
    np.testing.assert_allclose(_win("lanczos")[0], W.lanczos(nb)[1:], atol=1e-12)
    np.testing.assert_allclose(_win("tukey")[0], W.tukey(nb + 1, 0.5)[1:nb], atol=1e-12)
    # cosine-sum windows with Nuttall's published coefficients (Nuttall 1981)
    np.testing.assert_allclose(_win("nuttall")[0],
                               W.general_cosine(nb, [0.355768, 0.487396, 0.144232, 0.012604])[1:],
                               atol=1e-12)
    np.testing.assert_allclose(_win("blackman-nuttall")[0],
                               W.general_cosine(nb, [0.3635819, 0.4891775, 0.1365995, 0.0106411])[1:],
                               atol=1e-12)


def test_frequency_windows_closed_forms():
    # last bin: w = pi (d = 1)
    assert abs(_win("hann")[0][-1]) < 1e-15
    assert abs(_win("hamming")[0][-1] - 0.08) < 1e-15
    assert abs(_win("cosine")[0][-1]) < 1e-15
    assert abs(_win("shepp-logan")[0][-1] - 2.0 / math.pi) < 1e-15
    # d = 0.5 stretches the window: hann reaches 0 at w = pi / 2 (bin N/4)
    w, nb, n = _win("hann", d=0.5)
    assert abs(w[n // 4 - 1]) < 1e-15
    # kaiser (the reference evaluates J0): 1 at the centre bin, 1/J0(3 pi) at the end
    from scipy.special import j0
    w, nb, n = _win("kaiser")
    assert abs(w[(nb - 1) // 2 - 1] - 1.0) < 1e-12
    assert abs(w[-1] - 1.0 / j0(3 * math.pi)) < 1e-9
    # parzen: continuous at q = 1/2 (0.25 from both branches), 0 at q = 1
    w, nb, n = _win("parzen")
    assert abs(w[(nb - 1) // 2 - 1] - 0.25) < 1e-15 and abs(w[-1]) < 1e-15


def test_windowed_filter_cutoff_and_ramlak():
    H = OF.windowed_filter("ram-lak", 40)
    np.testing.assert_allclose(H, OF.ramp_filter(40), atol=1e-15)
    H = OF.windowed_filter("hamming", 40, d=0.5)
    n = OF.next_pow2(80)
    w = 2 * math.pi * np.arange(n // 2 + 1) / n
    assert np.all(H[w > math.pi * 0.5] == 0) and np.all(H[1:][w[1:] <= math.pi * 0.5] != 0)


def test_real_space_filter_is_centred_convolution():
    """A real-space kernel of width W is centred at W/2: filtering equals
    the linear convolution sum_j row[i - (j - W/2)] kernel[j]."""
    rng = np.random.default_rng(11)
    n_det, width = 37, 21
    row = rng.standard_normal(n_det)
    kern = rng.standard_normal(width)
    got = OF.filter_rows(row[None, :], OF.custom_filter(kern[None, :], n_det, True))[0]
    ref = np.zeros(n_det)
    for i in range(n_det):
        for j in range(width):
            t = i - (j - width // 2)
            if 0 <= t < n_det:
                ref[i] += row[t] * kern[j]
    np.testing.assert_allclose(got, ref, atol=1e-12)


def test_per_row_filters_index_rows_modulo():
    rng = np.random.default_rng(12)
    rows = rng.standard_normal((2, 3, 30))  # e.g. [v][angle][u]
    Hs = np.stack([OF.windowed_filter(t, 30) for t in ("hann", "cosine", "parzen")])
    got = OF.filter_rows(rows, Hs)
    for v in range(2):
        for a in range(3):
            np.testing.assert_allclose(got[v, a], OF.filter_rows(rows[v, a], Hs[a]), atol=1e-12)


# ---------------------------------------------------------------- Parker weights

def test_parker_conjugate_rays_sum_to_one():
    """Short-scan weights of conjugate rays add up to one: in the fan
    geometry ray (a, U) is the line of ray (a + pi - 2 gamma, -U) with
    gamma = atan(U / (SOD + ODD)) (checked on the oracle's own geometry),
    and the Parker weights of the two sum to 1 over the whole scan."""
    sod, odd, nd = 100.0, 50.0, 41
    D = sod + odd
    U = 10.0
    g = math.atan(U / D)
    a = 0.3
    vecs = OG.gen_fan_projections(np.array([a, a + math.pi - 2 * g]), nd, sod, odd, 1.0)

    def line(v, u):
        s = np.array(v[0:2])
        d = np.array(v[2:4]) + (u + 0.5) * np.array(v[4:6])
        n = (d - s) / np.linalg.norm(d - s)
        return n, s[0] * n[1] - s[1] * n[0]

    u = nd / 2.0 - 0.5 + U  # detector coordinate of U (pixel 1)
    n1, o1 = line(vecs[0], u)
    n2, o2 = line(vecs[1], nd - 1 - u)
    assert min(np.abs(n1 + n2).max() + abs(o1 + o2), np.abs(n1 - n2).max() + abs(o1 - o2)) < 1e-4
    fan = 0.3
    rng = np.random.default_rng(21)
    gam = rng.uniform(-0.99 * fan, 0.99 * fan, 4000)
    beta = rng.uniform(1e-6, math.pi + 2 * gam)  # conjugate stays inside the scan
    tot = OF.parker_weight(beta, gam, fan) + OF.parker_weight(beta + math.pi - 2 * gam, -gam, fan)
    np.testing.assert_allclose(tot, 1.0, atol=1e-12)


def test_parker_weights_scan_range_scale():
    """Equally spaced scan over pi + 2 fan: weights in [0, range/pi], zero
    at the first angle, and the angle-sum of the weights of each detector
    column equals the count of a half turn (n_a * pi / range) times the
    scale, up to discretisation."""
    sod, odd, nd, du = 100.0, 50.0, 64, 1.0
    fan = math.atan(du * nd / 2 / (sod + odd))
    na = 400
    ang = np.linspace(0, math.pi + 2 * fan, na, endpoint=False)
    w = OF.parker_weights(ang, nd, sod, odd, du)
    scale = (ang[-1] - ang[0]) / (na - 1) * na / math.pi
    assert np.all(w >= 0) and np.all(w <= scale + 1e-12) and np.all(w[0] == 0)
    col = w.sum(axis=0) / scale            # sum over angles of w / scale
    np.testing.assert_allclose(col, na * math.pi / (ang[1] - ang[0]) / na, rtol=2e-2)


# ---------------------------------------------------------------- supersampling
# Sub-ray / sub-voxel sampling written out explicitly in the oracle is
# checked against the average of the plain projector over shifted detector
# or volume geometries (an independent construction of the same samples).

def _shift_window_3d(n, d):
    vg = OG.vol_geom_3d(n, n, n)
    vg["option"] = {"WindowMinX": -n / 2 + d[0], "WindowMaxX": n / 2 + d[0],
                    "WindowMinY": -n / 2 + d[1], "WindowMaxY": n / 2 + d[1],
                    "WindowMinZ": -n / 2 + d[2], "WindowMaxZ": n / 2 + d[2]}
    return vg


def test_cone_fp_supersampling_equals_shifted_detectors():
    n, nu, nv, ss = 12, 10, 9, 3
    ang = np.array([0.2, 1.3, 2.0])  # away from the 45-degree direction switch
    v, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nu, nv, 40.0, 20.0, 1.5, 1.2), "cone",
                            OG.vol_geom_3d(n, n, n))
    vol = np.random.default_rng(41).random((n, n, n))
    ref = np.zeros((nv, len(ang), nu))
    for k in range(ss):
        for l in range(ss):
            w = v.copy()
            w[:, 3:6] += ((k + 0.5) / ss - 0.5) * v[:, 6:9] + ((l + 0.5) / ss - 0.5) * v[:, 9:12]
            ref += O3.cone_fp(vol, w, nu, nv, vs)
    np.testing.assert_allclose(O3.cone_fp(vol, v, nu, nv, vs, ss=ss), ref / ss ** 2, rtol=1e-12)


def test_bp_supersampling_equals_shifted_volumes():
    n, na, nu, ss = 8, 5, 12, 2
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    proj = np.random.default_rng(42).random((nu, na, nu))
    vecs = OG.gen_cone_projections(ang, nu, nu, 30.0, 15.0, 1.0, 1.0)
    v0, vs = OG.normalise_3d(vecs, "cone", OG.vol_geom_3d(n, n, n))
    pv0, _ = OG.normalise_3d(OG.gen_par3d_projections(ang, nu, nu, 1.0, 1.0), "parallel3d",
                             OG.vol_geom_3d(n, n, n))
    refc = np.zeros((n, n, n))
    refp = np.zeros((n, n, n))
    offs = [(a, b, c) for a in range(ss) for b in range(ss) for c in range(ss)]
    for a, b, c in offs:
        d = tuple((q + 0.5) / ss - 0.5 for q in (a, b, c))
        vc, _ = OG.normalise_3d(vecs, "cone", _shift_window_3d(n, d))
        refc += O3.cone_bp(proj, vc, (n, n, n), vs)
        vp, _ = OG.normalise_3d(OG.gen_par3d_projections(ang, nu, nu, 1.0, 1.0), "parallel3d",
                                _shift_window_3d(n, d))
        refp += O3.par3d_bp(proj, vp, (n, n, n), vs)
    np.testing.assert_allclose(O3.cone_bp(proj, v0, (n, n, n), vs, ss=ss), refc / len(offs), rtol=1e-12)
    np.testing.assert_allclose(O3.par3d_bp(proj, pv0, (n, n, n), vs, ss=ss), refp / len(offs),
                               rtol=1e-12, atol=1e-12)


def test_2d_supersampling_equals_shifted_geometries():
    n, na, nd, ss = 16, 7, 24, 3
    ang = np.linspace(0.1, 0.5, na)  # near-vertical rays: no orientation change
    vol = np.random.default_rng(43).random((n, n))
    vg = OG.vol_geom_2d(n, n)
    for kind, vecs in (("parallel", OG.gen_par_projections(ang, nd, 1.2)),
                       ("fan", OG.gen_fan_projections(ang, nd, 400.0, 200.0, 1.0))):
        v, osc = OG.normalise_2d(vecs, kind, vg)
        fp = O2.par_fp if kind == "parallel" else O2.fan_fp
        ref = sum(fp(vol, O2._shift_det(v, (k + 0.5) / ss - 0.5), nd, osc) for k in range(ss)) / ss
        np.testing.assert_allclose(fp(vol, v, nd, osc, rays=ss), ref, rtol=1e-12, atol=1e-12)
        y = np.random.default_rng(44).random((na, nd))
        bp = O2.par_bp if kind == "parallel" else O2.fan_bp
        refb = 0
        for a in range(ss):
            for b in range(ss):
                # sub-pixel (dx, dy) in (column, row) units sits at world
                # (dx, -dy) from the pixel centre: move the geometry by the
                # opposite amount
                dx, dy = (a + 0.5) / ss - 0.5, (b + 0.5) / ss - 0.5
                vv = np.array(v, dtype=np.float64)
                vv[:, 2:4] -= (dx, -dy)
                if kind == "fan":
                    vv[:, 0:2] -= (dx, -dy)
                refb = refb + bp(y, vv, (n, n), osc)
        np.testing.assert_allclose(bp(y, v, (n, n), osc, ss=ss), refb / ss ** 2, rtol=1e-10)


def test_sum_square_weights_equals_squared_matrix_rows():
    """sum_square_weights = row sums of the element-wise squared system
    matrix of the par3d FP (built column by column from unit volumes)."""
    n, na, nu, nv = 4, 3, 6, 5
    ang = np.array([0.3, 1.2, 2.5])
    v, vs = OG.normalise_3d(OG.gen_par3d_projections(ang, nu, nv, 0.9, 1.1), "parallel3d",
                            OG.vol_geom_3d(n, n, n))
    cols = []
    for i in range(n ** 3):
        e = np.zeros(n ** 3)
        e[i] = 1.0
        cols.append(O3.par3d_fp(e.reshape(n, n, n), v, nu, nv, vs, out_scale=1.3).ravel())
    A = np.stack(cols, axis=1)
    ref = (A * A).sum(axis=1).reshape(nv, na, nu)
    got = O3.par3d_fp_sumsq(v, nu, nv, (n, n, n), vs, out_scale=1.3)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------- EM / SART

def test_em_fixed_point_and_sart_reduces_residual():
    """With the exact transpose, EM leaves the solution of a consistent
    positive system in place (every ray through a positive pixel has ratio
    1) and decreases the Kullback-Leibler divergence from any start; SART
    sweeps reduce the residual."""
    n, na, nd = 16, 20, 24
    ang = np.linspace(0, np.pi, na, endpoint=False)
    v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.0), "parallel", OG.vol_geom_2d(n, n))
    truth = 0.5 + np.random.default_rng(61).random((n, n))
    fp = lambda x: O2.par_fp(x, v, nd, osc)
    bp = lambda y: O2.par_bp(y, v, (n, n), osc)
    bpT = lambda y: O2.par_fp_T(y, v, (n, n), osc)
    b = fp(truth)
    np.testing.assert_allclose(OA.em(fp, bpT, b, truth, 3), truth, rtol=1e-10)

    def kl(x):
        t = fp(x)
        m = b > 0
        return float(np.sum(b[m] * np.log(b[m] / t[m]) - b[m] + t[m]) + np.sum(t[~m]))
    x = np.ones((n, n))
    prev = kl(x)
    for _ in range(5):
        x = OA.em(fp, bpT, b, x, 1)
        cur = kl(x)
        assert cur < prev
        prev = cur
    line_w = OA._inv(fp(np.ones((n, n))))
    fpa = lambda x, a: O2.par_fp(x, v[a:a + 1], nd, osc)[0]
    bpa = lambda r, a: O2.sart_bp(r, v[a], (n, n), "parallel", osc)
    r0 = np.linalg.norm(b)
    x1 = OA.sart(fpa, bpa, b, np.zeros((n, n)), list(range(na)), line_w=line_w)
    x2 = OA.sart(fpa, bpa, b, x1, list(range(na)), line_w=line_w)
    r1, r2 = np.linalg.norm(b - fp(x1)), np.linalg.norm(b - fp(x2))
    assert r2 < r1 < 0.5 * r0


def test_sart_parallel_bp_is_the_unweighted_bp():
    """For parallel beams SART's BP of one row is the full BP of that row
    divided by the angle's weight |r| / |d|."""
    n, nd = 15, 21
    ang = np.array([0.7])
    v, osc = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.3), "parallel", OG.vol_geom_2d(n, n))
    row = np.random.default_rng(62).random(nd)
    rx, ry, dsx, dsy, ux, uy = v[0]
    w = math.hypot(rx, ry) / abs(ux * ry - uy * rx)
    np.testing.assert_allclose(O2.sart_bp(row, v[0], (n, n), "parallel"),
                               O2.par_bp(row[None, :], v, (n, n)) / w, rtol=1e-12)


# ---------------------------------------------------------------- cyl_cone_vec
# The cylindrical-detector cone projector (cylcone_fp / cylcone_bp) is pinned
# by (1) its flat-detector limit: as the radius grows the cylinder section
# becomes the flat cone_vec detector with the same centre and pixel vectors,
# and the FP and BP reduce to cone_fp / cone_bp (independent formulas: ray
# setup in world units vs normalised vectors, quadratic cylinder hit + asin
# vs the projective map), with anisotropic voxels and a shifted window;
# (2) the BP's detector mapping inverting the FP's pixel rays exactly at a
# strongly curved detector; (3) the reference's adjoint criterion.

def _cyl_vecs(na, sod, odd, R, du=1.1, dv=0.9, tilt=0.0):
    out = []
    for t in np.linspace(0, 2 * np.pi, na, endpoint=False):
        c, s = np.cos(t), np.sin(t)
        src = np.array([s * sod, -c * sod, tilt * sod])
        det = np.array([-s * odd, c * odd, -tilt * odd])
        out.append(np.r_[src, det, np.array([c, s, 0.0]) * du, np.array([0.0, 0.0, dv]), R])
    return np.array(out)


def test_cylcone_flat_detector_limit():
    n, na, nu, nv = 14, 5, 18, 16
    vg = OG.vol_geom_3d(n, n, n)
    vg["option"] = {"WindowMinX": -8.0, "WindowMaxX": 9.5, "WindowMinY": -7.0, "WindowMaxY": 7.0,
                    "WindowMinZ": -5.0, "WindowMaxZ": 6.2}
    vec13 = _cyl_vecs(na, 60.0, 30.0, 1e6, tilt=0.05)
    tv, vs = OG.translate_cylcone(vec13, vg)
    cv, cvs = OG.normalise_3d(OG.conevec_from_astra(vec13[:, :12], nu, nv), "cone", vg)
    assert vs == cvs
    rng = np.random.default_rng(81)
    vol = rng.random((n, n, n))
    a = O3.cylcone_fp(vol, tv, nu, nv, vs)
    b = O3.cone_fp(vol, cv, nu, nv, cvs)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5
    y = rng.random((nv, na, nu))
    a = O3.cylcone_bp(y, tv, (n, n, n), vs)
    b = O3.cone_bp(y, cv, (n, n, n), cvs)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5


def test_cylcone_bp_mapping_inverts_fp_rays():
    nu, nv = 40, 24
    vec13 = _cyl_vecs(7, 50.0, 40.0, 90.0, du=2.0, dv=1.3, tilt=0.1)   # 50 deg fan
    rng = np.random.default_rng(82)
    for vec in vec13:
        src = vec[0:3]
        pix = O3.cylcone_pixels(vec, nu, nv)
        iu = rng.integers(0, nu, 30)
        iv = rng.integers(0, nv, 30)
        tau = rng.uniform(0.2, 0.9, 30)
        P = [pix[k][iv, iu] for k in range(3)]
        W = [src[k] + tau * (P[k] - src[k]) for k in range(3)]
        fU, fV, t, ok = O3.cylcone_detector_coords(vec, nu, nv, *W)
        assert ok.all()
        assert np.abs(fU - (iu + 0.5)).max() < 1e-9 and np.abs(fV - (iv + 0.5)).max() < 1e-9


def test_cylcone_bp_adjoint_criterion():
    # the reference's BP-vs-FP adjoint criterion (<= 1e-1,
    # tests/python/functional/test_line2d.py:591-640)
    n, na, nu, nv = 12, 8, 20, 16
    vg = OG.vol_geom_3d(n, n, n)
    tv, vs = OG.translate_cylcone(_cyl_vecs(na, 40.0, 30.0, 70.0), vg)
    rng = np.random.default_rng(83)
    x = rng.random((n, n, n))
    y = rng.random((nv, na, nu))
    lhs = float(np.sum(O3.cylcone_fp(x, tv, nu, nv, vs) * y))
    rhs = float(np.sum(x * O3.cylcone_bp(y, tv, (n, n, n), vs)))
    assert abs(lhs - rhs) / abs(lhs) < 1e-1


def test_vector_forms_reproduce_angle_geometries():
    # *_vec rows hold the detector centre; the conversions to corner form
    # (DetS = centre - n/2 u [- n_v/2 v]) must reproduce the angle-based
    # generators' corner vectors exactly (a dropped half or sign fails)
    ang = np.linspace(0.1, 5.9, 11)
    c3 = OG.gen_cone_projections(ang, 13, 9, 70.0, 30.0, 1.3, 0.7)
    centre = c3.copy()
    centre[:, 3:6] += 0.5 * 13 * c3[:, 6:9] + 0.5 * 9 * c3[:, 9:12]
    assert np.allclose(OG.conevec_from_astra(centre, 13, 9), c3, rtol=0, atol=1e-12)
    for gen, conv in ((lambda: OG.gen_fan_projections(ang, 17, 60.0, 25.0, 1.1), OG.fanvec_from_astra),
                      (lambda: OG.gen_par_projections(ang, 17, 1.1), OG.parvec_from_astra)):
        v = gen()
        centre = v.copy()
        centre[:, 2:4] += 0.5 * 17 * v[:, 4:6]
        assert np.allclose(conv(centre, 17), v, rtol=0, atol=1e-5)


# ------------------------------------------------- 2D line-integral answers
@pytest.mark.parametrize("kind", ["parallel", "parallel_vec", "fanflat", "fanflat_vec"])
def test_2d_fp_matches_linear_kernel_line_integrals(kind):
    """The reference's kernel test (tests/python/functional/test_line2d.py:
    399-513): rectangles of ones in random windows and geometries against
    the exact line integral of the linearly interpolated image; the
    reference asks 2e-3 of its CPU linear kernel (2e-2 with 8-bit CUDA
    weights), the exact-interpolation oracle meets 5e-4."""
    import astra
    import line2d_known_answer as K
    rng = np.random.default_rng(123)
    for _ in range(8):
        vol, proj, data, rect = K.random_case(kind, rng)
        vg = astra.create_vol_geom(*vol)
        pg = astra.create_proj_geom(*proj)
        ok, nv, osc = K.oracle_vectors(proj, vg)
        shape = astra.functions.geom_size(pg)
        vert = K.fan_kernel_vertical(nv, shape[1]) if ok == "fan" else None
        ref = K.linear_kernel_sinogram(proj, rect, shape, vertical=vert)
        got = (O2.par_fp if ok == "par" else O2.fan_fp)(data, nv, shape[1], osc)
        assert np.abs(got - ref).max() <= 5e-4
