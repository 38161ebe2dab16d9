"""astra.experimental.projectPoint / getProjectedBBox on the host
(python/astra/experimental.pyx:229-241).

Pinned independently of the formulas: a world point on the ray of detector
pixel (i, j) projects to (i, j) exactly (flat detectors: the pixel centre
from the vector definition; cylindrical detector: the pixel centre on the
cylinder as the FP samples it), the angle-based geometries agree with
their vector forms, and the projected bounding box contains every interior
point's row (attained at a corner for flat detectors).
"""
import numpy as np
import pytest


def _astra():
    import astra
    return astra


def _cone():
    a = _astra()
    return a.create_proj_geom("cone", 1.3, 0.9, 24, 32, np.linspace(0.2, 2 * np.pi, 9, endpoint=False),
                              120.0, 70.0)


def _par3d():
    a = _astra()
    return a.create_proj_geom("parallel3d", 1.1, 0.8, 24, 32, np.linspace(0.1, np.pi, 9, endpoint=False))


def _tilted_vectors(rng, n, cone=True):
    vecs = np.zeros((n, 12))
    for k in range(n):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        u = np.cross(d, rng.normal(size=3))
        u /= np.linalg.norm(u)
        v = np.cross(d, u)
        src = d * 110.0 if cone else -d
        det = -d * 60.0 + rng.normal(size=3)
        vecs[k] = np.concatenate([src, det, u * 1.2, v * 0.85])
    return vecs


@pytest.mark.parametrize("kind", ["cone", "par3d"])
def test_angle_form_equals_vector_form(kind):
    a = _astra()
    pg = _cone() if kind == "cone" else _par3d()
    pv = a.geom_2vec(pg)
    rng = np.random.default_rng(1)
    for _ in range(40):
        x, y, z = rng.uniform(-20, 20, 3)
        i = int(rng.integers(9))
        u1, v1 = a.experimental.projectPoint(pg, x, y, z, i)
        u2, v2 = a.experimental.projectPoint(pv, x, y, z, i)
        assert abs(u1 - u2) < 2e-5 * max(1.0, abs(u2)) and abs(v1 - v2) < 2e-5 * max(1.0, abs(v2))


@pytest.mark.parametrize("cone", [True, False])
def test_point_on_pixel_ray_projects_to_pixel(cone):
    a = _astra()
    rng = np.random.default_rng(2)
    nu, nv = 32, 24
    vecs = _tilted_vectors(rng, 6, cone)
    pg = a.create_proj_geom("cone_vec" if cone else "parallel3d_vec", nv, nu, vecs)
    for k in range(6):
        s, d, u, v = vecs[k, 0:3], vecs[k, 3:6], vecs[k, 6:9], vecs[k, 9:12]
        for i, j in ((0, 0), (31, 23), (7, 15), (20, 3)):
            pix = d + (i - nu / 2 + 0.5) * u + (j - nv / 2 + 0.5) * v
            for lam in (0.3, 0.8):
                pt = s + lam * (pix - s) if cone else pix + (lam * 50 - 20) * s
                gu, gv = a.experimental.projectPoint(pg, *pt, k)
                assert abs(gu - i) < 1e-8 and abs(gv - j) < 1e-8, (k, i, j, gu, gv)


def _cyl_geom(radius=200.0):
    a = _astra()
    base = a.geom_2vec(a.create_proj_geom("cone", 1.0, 1.0, 20, 45, np.linspace(0, 2 * np.pi, 12,
                                                                                endpoint=False),
                                          100.0, 100.0))["Vectors"]
    vecs = np.hstack([base, np.full((12, 1), radius)])
    return a.create_proj_geom("cyl_cone_vec", 20, 45, vecs), vecs


def test_cylinder_point_on_pixel_ray_projects_to_pixel():
    from oracle import proj3d as O3
    a = _astra()
    pg, vecs = _cyl_geom()
    for k in range(12):
        P = O3.cylcone_pixels(vecs[k], 45, 20)
        s = vecs[k, 0:3]
        for i, j in ((0, 0), (44, 19), (22, 10), (5, 17)):
            pix = np.array([P[0][j, i], P[1][j, i], P[2][j, i]])
            for lam in (0.35, 0.7):
                gu, gv = a.experimental.projectPoint(pg, *(s + lam * (pix - s)), k)
                assert abs(gu - i) < 1e-8 and abs(gv - j) < 1e-8, (k, i, j, gu, gv)


def test_cylinder_ray_missing_detector():
    """A source outside a small cylinder: a ray that passes beside it gives
    u = +-inf (by the side it passes) and v = NaN."""
    a = _astra()
    vec = np.array([[0, -100, 0, 0, 10, 0, 1, 0, 0, 0, 0, 1, 20.0]])
    pg = a.create_proj_geom("cyl_cone_vec", 8, 16, vec)
    u1, v1 = a.experimental.projectPoint(pg, 60, 0, 0, 0)
    u2, v2 = a.experimental.projectPoint(pg, -60, 0, 0, 0)
    assert np.isinf(u1) and np.isinf(u2) and np.sign(u1) != np.sign(u2)
    assert np.isnan(v1) and np.isnan(v2)
    uc, vc = a.experimental.projectPoint(pg, 0, 0, 0, 0)
    assert abs(uc - 7.5) < 1e-12 and abs(vc - 3.5) < 1e-12


@pytest.mark.parametrize("kind", ["cone", "cone_vec", "par3d", "cyl"])
def test_projected_bbox_contains_interior(kind):
    a = _astra()
    if kind == "cone":
        pg = _cone()
    elif kind == "cone_vec":
        pg = a.geom_2vec(_cone())
    elif kind == "par3d":
        pg = _par3d()
    else:
        pg = _cyl_geom()[0]
    box = (-9.0, 7.0, -5.0, 11.0, -6.0, 4.0)
    vmin, vmax = a.experimental.getProjectedBBox(pg, *box)
    assert vmin < vmax
    rng = np.random.default_rng(3)
    na = a.geom_size(pg, 1)
    for _ in range(300):
        p = [rng.uniform(box[0], box[1]), rng.uniform(box[2], box[3]), rng.uniform(box[4], box[5])]
        _, v = a.experimental.projectPoint(pg, *p, int(rng.integers(na)))
        assert vmin - 1e-6 <= v <= vmax + 1e-6
    if kind != "cyl":
        corners = [a.experimental.projectPoint(pg, x, y, z, i)[1]
                   for x in box[0:2] for y in box[2:4] for z in box[4:6] for i in range(na)]
        assert np.isclose(min(corners), vmin) and np.isclose(max(corners), vmax)


def test_project_point_index_checked():
    a = _astra()
    with pytest.raises(a.AstraError):
        a.experimental.projectPoint(_cone(), 0, 0, 0, 9)
