"""Projection / volume geometry dictionaries -> normalised projection vectors.

Plays the role of the reference's geometry classes plus convertAstraGeometry
(src/GeometryUtil2D.cpp:40-396, src/GeometryUtil3D.cpp:42-805): the angle
based geometries are expanded to per-angle vectors, vector geometries are
converted from detector-centre to detector-corner form, and everything is
translated / scaled so that the volume is centred at 0 with unit voxels.
Storage precision follows the reference structs: 2D vectors float32, 3D
vectors float64, angles float32.
"""
import numpy as np

from .log import AstraError

f32 = np.float32

PARALLEL_2D = ("parallel", "parallel_vec")
FAN_2D = ("fanflat", "fanflat_vec")
PAR_3D = ("parallel3d", "parallel3d_vec")
CONE_3D = ("cone", "cone_vec")


def is_3d_proj(pg):
    return pg["type"] in PAR_3D + CONE_3D + ("cyl_cone_vec",)


def is_3d_vol(vg):
    return "GridSliceCount" in vg


OPTION_KEYS = ("option", "options", "Option", "Options")


def options_of(cfg):
    """The options of a config or geometry dictionary: 'option', 'options',
    'Option' and 'Options' are synonyms (python/astra/utils.pyx:121)."""
    out = {}
    for k in OPTION_KEYS:
        out.update(cfg.get(k) or {})
    return out


def config_form(geom):
    """A geometry as the reference's get_geometry returns it: the
    geometry's configuration read back, where a volume's window is under
    'options' (python/astra/utils.pyx:213-235, src/VolumeGeometry3D.cpp:
    290-305); projection geometries come back with their own keys."""
    if "GridColCount" not in geom:
        return dict(geom)
    out = {"GridColCount": int(geom["GridColCount"]), "GridRowCount": int(geom["GridRowCount"])}
    axes = [("X", geom["GridColCount"]), ("Y", geom["GridRowCount"])]
    if "GridSliceCount" in geom:
        out["GridSliceCount"] = int(geom["GridSliceCount"])
        axes.append(("Z", geom["GridSliceCount"]))
    opts = {}
    for ax, n in axes:
        lo, hi = _window(geom, ax, n)
        opts["WindowMin" + ax], opts["WindowMax" + ax] = float(lo), float(hi)
    out["options"] = opts
    return out


# ------------------------------------------------------------- volumes ---
def _window(vg, axis, n):
    o = options_of(vg)
    lo = o.get("WindowMin" + axis, -n / 2.0)
    hi = o.get("WindowMax" + axis, n / 2.0)
    return f32(lo), f32(hi)


def vol_shape(vg):
    """Array shape of a volume: (rows, cols) or (slices, rows, cols)."""
    if is_3d_vol(vg):
        return (int(vg["GridSliceCount"]), int(vg["GridRowCount"]), int(vg["GridColCount"]))
    return (int(vg["GridRowCount"]), int(vg["GridColCount"]))


def pixel_lengths(vg):
    """Voxel sizes (x, y[, z]) as float32, (max - min) / count."""
    axes = [("X", vg["GridColCount"]), ("Y", vg["GridRowCount"])]
    if is_3d_vol(vg):
        axes.append(("Z", vg["GridSliceCount"]))
    out = []
    for ax, n in axes:
        lo, hi = _window(vg, ax, n)
        out.append(f32((hi - lo) / f32(n)))
    return out


# --------------------------------------------------------- projections ---
def proj_shape(pg):
    """Array shape: (angles, dets) in 2D, (det rows, angles, det cols) in 3D."""
    t = pg["type"]
    if t in ("parallel", "fanflat"):
        return (len(np.atleast_1d(pg["ProjectionAngles"])), int(pg["DetectorCount"]))
    if t in ("parallel_vec", "fanflat_vec"):
        return (np.asarray(pg["Vectors"]).shape[0], int(pg["DetectorCount"]))
    if t in ("parallel3d", "cone"):
        return (int(pg["DetectorRowCount"]), len(np.atleast_1d(pg["ProjectionAngles"])),
                int(pg["DetectorColCount"]))
    if t in ("parallel3d_vec", "cone_vec", "cyl_cone_vec"):
        return (int(pg["DetectorRowCount"]), np.asarray(pg["Vectors"]).shape[0],
                int(pg["DetectorColCount"]))
    raise AstraError("unsupported projection geometry type '%s'" % t)


def _angles(pg):
    return np.atleast_1d(np.asarray(pg["ProjectionAngles"], dtype=f32)).astype(np.float64)


def vectors_2d(pg):
    """(kind, float32 [N, 6]) in detector-corner form, before normalisation.

    parallel: ray = R(a)(0, 1), corner = R(a)(-n w / 2, 0), u = R(a)(w, 0)
    (genParProjections, src/GeometryUtil2D.cpp:40-79); fanflat: src =
    R(a)(0, -SOD), corner = R(a)(-n w / 2, ODD), u = R(a)(w, 0) with float32
    base vectors (genFanProjections, :82-107); *_vec: corner = centre -
    n/2 u (src/FanFlatVecProjectionGeometry2D.cpp:103-125).
    """
    t = pg["type"]
    n = int(pg["DetectorCount"])
    if t in ("parallel", "fanflat"):
        a = _angles(pg)
        c, s = np.cos(a), np.sin(a)
        w = float(f32(pg["DetectorWidth"]))
        if t == "parallel":
            base = [(0.0, 1.0), (float(f32(n * w * -0.5)), 0.0), (w, 0.0)]
        else:
            wf = f32(pg["DetectorWidth"])
            base = [(0.0, float(f32(-pg["DistanceOriginSource"]))),
                    (float(f32(f32(n) * wf / f32(-2.0))), float(f32(pg["DistanceOriginDetector"]))),
                    (float(wf), 0.0)]
        out = np.empty((len(a), 6), dtype=f32)
        for k, (bx, by) in enumerate(base):
            out[:, 2 * k] = bx * c - by * s
            out[:, 2 * k + 1] = bx * s + by * c
        return ("par" if t == "parallel" else "fan"), out
    if t in ("parallel_vec", "fanflat_vec"):
        v = np.asarray(pg["Vectors"], dtype=np.float64)
        if v.ndim != 2 or v.shape[1] != 6:
            raise AstraError("Vectors should be an Nx6 matrix")
        out = np.empty(v.shape, dtype=f32)
        out[:, 0:2] = v[:, 0:2]
        out[:, 4:6] = v[:, 4:6]
        out[:, 2:4] = v[:, 2:4] - 0.5 * n * out[:, 4:6].astype(np.float64)
        return ("par" if t == "parallel_vec" else "fan"), out
    raise AstraError("projection geometry '%s' is not a 2D parallel/fan geometry" % t)


def normalise_2d(kind, vecs, vg):
    """Translate by -window centre, scale by 1/pixel (float32 storage); return
    (vectors, output_scale = pixel length) (src/GeometryUtil2D.cpp:151-179)."""
    px, py = pixel_lengths(vg)
    if abs(float(px) - float(py)) > 1e-5:
        raise AstraError("the 2D GPU projectors require square pixels")
    lox, hix = _window(vg, "X", vg["GridColCount"])
    loy, hiy = _window(vg, "Y", vg["GridRowCount"])
    dx = float(-(lox + hix) / f32(2))
    dy = float(-(loy + hiy) / f32(2))
    fac = float(f32(1.0) / px)
    out = vecs.astype(np.float64)
    if kind == "fan":
        out[:, 0] += dx
        out[:, 1] += dy
    out[:, 2] += dx
    out[:, 3] += dy
    out = out.astype(f32).astype(np.float64) * fac
    return out.astype(f32), float(px)


def vectors_3d(pg):
    """(kind, float64 [N, 12]) in detector-corner form, before normalisation
    (genConeProjections / genPar3DProjections, src/GeometryUtil3D.cpp:42-172;
    src/ConeVecProjectionGeometry3D.cpp:82-106)."""
    t = pg["type"]
    nu, nv = int(pg["DetectorColCount"]), int(pg["DetectorRowCount"])
    if t in ("cone", "parallel3d"):
        a = _angles(pg)
        c, s = np.cos(a), np.sin(a)
        du = float(f32(pg["DetectorSpacingX"]))
        dv = float(f32(pg["DetectorSpacingY"]))
        corner = (nu * du * -0.5, 0.0, nv * dv * -0.5)
        if t == "cone":
            sod = float(f32(pg["DistanceOriginSource"]))
            odd = float(f32(pg["DistanceOriginDetector"]))
            base = [(0.0, -sod, 0.0), (corner[0], odd, corner[2]), (du, 0.0, 0.0), (0.0, 0.0, dv)]
        else:
            base = [(0.0, 1.0, 0.0), corner, (du, 0.0, 0.0), (0.0, 0.0, dv)]
        out = np.empty((len(a), 12))
        for k, (bx, by, bz) in enumerate(base):
            out[:, 3 * k] = bx * c - by * s
            out[:, 3 * k + 1] = bx * s + by * c
            out[:, 3 * k + 2] = bz
        return ("cone" if t == "cone" else "par3d"), out
    if t in ("cone_vec", "parallel3d_vec"):
        v = np.array(pg["Vectors"], dtype=np.float64)
        if v.ndim != 2 or v.shape[1] != 12:
            raise AstraError("Vectors should be an Nx12 matrix")
        v[:, 3:6] = v[:, 3:6] - 0.5 * nv * v[:, 9:12] - 0.5 * nu * v[:, 6:9]
        return ("cone" if t == "cone_vec" else "par3d"), v
    if t == "cyl_cone_vec":
        # 13-tuples (src, detector centre, u, v, radius), kept as given
        # (src/CylConeVecProjectionGeometry3D.cpp:80-112)
        v = np.array(pg["Vectors"], dtype=np.float64)
        if v.ndim != 2 or v.shape[1] != 13:
            raise AstraError("Vectors should be an Nx13 matrix")
        if not (v[:, 12] > 0).all():
            raise AstraError("cyl_cone_vec: the detector radius must be positive")
        return "cylcone", v
    raise AstraError("projection geometry '%s' is not a 3D parallel/cone geometry" % t)


def normalise_3d(kind, vecs, vg):
    """Translate by -window centre, scale per axis by 1/voxel size; return
    (vectors, vol_scale) (src/GeometryUtil3D.cpp:569-593).  The cylindrical
    detector is translated only: its kernels work in world units with the
    voxel sizes (convertAstraGeometry_unscaled_internal, :595-616)."""
    d, fac = [], []
    for ax, n in (("X", vg["GridColCount"]), ("Y", vg["GridRowCount"]), ("Z", vg["GridSliceCount"])):
        lo, hi = _window(vg, ax, n)
        d.append(float(-(lo + hi) / f32(2)))
        fac.append(float(f32(1.0) / f32((hi - lo) / f32(n))))
    out = np.array(vecs, dtype=np.float64)
    if kind == "cylcone":
        out[:, 0:3] += d
        out[:, 3:6] += d
        return out, tuple(float(p) for p in pixel_lengths(vg))
    if kind == "cone":
        out[:, 0:3] += d
    out[:, 3:6] += d
    for k in range(4):
        out[:, 3 * k:3 * k + 3] *= fac
    return out, tuple(float(p) for p in pixel_lengths(vg))


# ------------------------------------------------------- geom_2vec -------
def geom_2vec(pg):
    """Convert an angle-based geometry to its vector form
    (python/astra/functions.py:167-275)."""
    t = pg["type"]
    if t in ("parallel", "fanflat"):
        kind, v = vectors_2d(pg)
        v = v.astype(np.float64)
        n = int(pg["DetectorCount"])
        v[:, 2:4] += 0.5 * n * v[:, 4:6]
        return {"type": t + "_vec", "DetectorCount": n, "Vectors": v}
    if t in ("cone", "parallel3d"):
        kind, v = vectors_3d(pg)
        nu, nv = int(pg["DetectorColCount"]), int(pg["DetectorRowCount"])
        v[:, 3:6] += 0.5 * nv * v[:, 9:12] + 0.5 * nu * v[:, 6:9]
        return {"type": ("cone_vec" if t == "cone" else "parallel3d_vec"),
                "DetectorRowCount": nv, "DetectorColCount": nu, "Vectors": v}
    if t.endswith("_vec"):
        return dict(pg)
    raise AstraError("geom_2vec: unsupported geometry type '%s'" % t)


# ------------------------------------------------ point projection -------
def _det_index(offset, spacing, n):
    """detectorOffset{X,Y}To{Col,Row}IndexFloat
    (include/astra/ProjectionGeometry3D.h:442-458): float32 arithmetic."""
    return (np.asarray(offset, dtype=f32) / f32(spacing) + f32((f32(n) - f32(1.0)) / f32(2.0))).astype(np.float64)


def _cyl_frame(vecs, flip):
    """Cylinder centre c, radial a, tangential b and axis vectors of
    cyl_cone_vec rows [N, 13]; `flip` orients a away from the source as
    getCylConeAxes does (src/GeometryUtil3D.cpp:561-583)."""
    s, d, u, v, r = vecs[:, 0:3], vecs[:, 3:6], vecs[:, 6:9], vecs[:, 9:12], vecs[:, 12:13]
    a = -np.cross(u, v) * (r / (np.linalg.norm(u, axis=1, keepdims=True) *
                                np.linalg.norm(v, axis=1, keepdims=True)))
    if flip:
        swap = np.linalg.norm(d - a - s, axis=1) > np.linalg.norm(d + a - s, axis=1)
        a = np.where(swap[:, None], -a, a)
    return d - a, a, u * (r / np.linalg.norm(u, axis=1, keepdims=True)), v


def _dot(a, b):
    return np.einsum("...k,...k->...", a, b)


def project_points_3d(pg, pts):
    """(u, v), each [projections, points]: detector index coordinates (0 =
    centre of the first column / row) of world points pts [P, 3] in every
    projection: projectPoint of the 3D geometry classes
    (src/ParallelProjectionGeometry3D.cpp:153-167,
    src/ConeProjectionGeometry3D.cpp:176-200,
    src/ParallelVecProjectionGeometry3D.cpp:201-215,
    src/ConeVecProjectionGeometry3D.cpp:199-214,
    src/CylConeVecProjectionGeometry3D.cpp:343-448).  A cylindrical
    detector that a ray misses gives u = -inf / +inf (the side the ray
    passes) and v = NaN."""
    t = pg["type"]
    p = np.atleast_2d(np.asarray(pts, dtype=np.float64))[None, :, :]  # [1, P, 3]
    nu, nv = int(pg["DetectorColCount"]), int(pg["DetectorRowCount"])
    if t in ("parallel3d", "cone"):
        a = _angles(pg)[:, None]
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        lateral = np.cos(a) * x + np.sin(a) * y
        su, sv = pg["DetectorSpacingX"], pg["DetectorSpacingY"]
        if t == "parallel3d":
            return _det_index(lateral, su, nu), _det_index(np.broadcast_to(z, lateral.shape), sv, nv)
        sod = float(f32(pg["DistanceOriginSource"]))
        odd = float(f32(pg["DistanceOriginDetector"]))
        mag = (sod + odd) / (-np.sin(a) * x + np.cos(a) * y + sod)
        return _det_index(lateral * mag, su, nu), _det_index(z * mag, sv, nv)
    _, vec = vectors_3d(pg)
    if t in ("parallel3d_vec", "cone_vec"):
        src, corner = vec[:, None, 0:3], vec[:, None, 3:6]
        u, v = vec[:, None, 6:9], vec[:, None, 9:12]
        if t == "parallel3d_vec":
            # the point moves along the ray onto corner + a u + b v
            den = _dot(src, np.cross(u, v))
            rel = p - corner
            return (_dot(np.cross(v, src), rel) / den - 0.5, _dot(np.cross(src, u), rel) / den - 0.5)
        w = corner - src
        den = _dot(np.cross(u, v), p - src)
        return ((_dot(np.cross(v, w), p) + _dot(src, np.cross(corner, v))) / den - 0.5,
                (_dot(np.cross(w, u), p) - _dot(src, np.cross(corner, u))) / den - 0.5)
    if t == "cyl_cone_vec":
        c, a, b, axis = _cyl_frame(vec, flip=False)
        s, d, u, v, r = vec[:, None, 0:3], vec[:, None, 3:6], vec[:, None, 6:9], vec[:, None, 9:12], \
            vec[:, None, 12]
        an = (a / np.linalg.norm(a, axis=1, keepdims=True))[:, None]
        bn = (b / np.linalg.norm(b, axis=1, keepdims=True))[:, None]
        cn = (axis / np.linalg.norm(axis, axis=1, keepdims=True))[:, None]
        c = c[:, None]
        ray = p - s
        ra, rb, rc = _dot(ray, an), _dot(ray, bn), _dot(ray, cn)
        sa, sb, sc = _dot(s - c, an), _dot(s - c, bn), _dot(s - c, cn)
        # |(t ra + sa, t rb + sb)| = r, the far root
        qa, qb, qc = ra * ra + rb * rb, 2.0 * (ra * sa + rb * sb), sa * sa + sb * sb - r * r
        disc = qb * qb - 4.0 * qa * qc
        hit = disc >= 0
        tt = (-qb + np.sqrt(np.where(hit, disc, 0.0))) / (2.0 * qa)
        with np.errstate(invalid="ignore"):
            theta = np.arcsin((tt * rb + sb) / r)
        dth = np.linalg.norm(u, axis=-1) / r
        uu = theta / dth + 0.5 * nu - 0.5
        vv = (tt * rc + sc) / np.linalg.norm(v, axis=-1) + 0.5 * nv - 0.5
        side = np.linalg.det(np.stack([u, d - s, v], axis=-2))
        side = side * np.linalg.det(np.stack(np.broadcast_arrays(ray, d - s, v), axis=-2))
        miss_u = np.where(side < 0, -np.inf, np.inf)
        return np.where(hit, uu, miss_u), np.where(hit, vv, np.nan)
    raise AstraError("projectPoint: unsupported geometry type '%s'" % t)


def project_point_3d(pg, x, y, z, i):
    """(u, v) of one world point in projection i (see project_points_3d)."""
    u, v = project_points_3d(pg, [[x, y, z]])
    return float(u[i, 0]), float(v[i, 0])


def _bbox_cyl(vecs, nu, nv, corners, corner_us):
    """getProjectedBBoxSingleAngle of the cylindrical detector for every
    projection (src/CylConeVecProjectionGeometry3D.cpp:201-340): u from the
    projected corners and column index 0; v from the corners projected onto
    a near and a far plane normal to the ray through column 0's centre."""
    s = vecs[:, 0:3]
    c, a, b, axis = _cyl_frame(vecs, flip=True)
    axis_n = axis / np.linalg.norm(axis, axis=1, keepdims=True)
    dth = np.linalg.norm(vecs[:, 6:9], axis=1) / vecs[:, 12]
    us = np.concatenate([np.zeros((len(vecs), 1)), corner_us], axis=1)      # [N, 9]
    theta = np.clip((us + 0.5 - 0.5 * nu) * dth[:, None], -np.pi / 2, np.pi / 2)
    pts = c[:, None] + a[:, None] * np.cos(theta)[..., None] + b[:, None] * np.sin(theta)[..., None]
    ray = pts[:, 0] - c
    ray = ray / np.linalg.norm(ray, axis=1, keepdims=True)
    depth = _dot(pts - s[:, None], ray[:, None])                           # [N, 9]
    far = depth[:, 0]
    near = depth.min(axis=1)
    rel = corners[None] - s[:, None]                                       # [N, 8, 3]
    along = _dot(rel, ray[:, None])
    lat = _dot(rel, axis_n[:, None])
    vs = np.concatenate([lat * near[:, None] / along, lat * far[:, None] / along], axis=1) + 0.5 * nv - 0.5
    return us.min(axis=1), us.max(axis=1), vs.min(axis=1), vs.max(axis=1)


def projected_bbox_3d(pg, xmin, xmax, ymin, ymax, zmin, zmax):
    """(umin, umax, vmin, vmax) over all projections of the box's corners
    (CProjectionGeometry3D::getProjectedBBox,
    src/ProjectionGeometry3D.cpp:162-236)."""
    corners = np.array([[cx, cy, cz] for cx in (xmin, xmax) for cy in (ymin, ymax) for cz in (zmin, zmax)],
                       dtype=np.float64)
    nu, nv = int(pg["DetectorColCount"]), int(pg["DetectorRowCount"])
    u, v = project_points_3d(pg, corners)
    if pg["type"] == "cyl_cone_vec":
        umin, umax, vmin, vmax = _bbox_cyl(vectors_3d(pg)[1], nu, nv, corners, u)
    else:
        umin, umax, vmin, vmax = u.min(axis=1), u.max(axis=1), v.min(axis=1), v.max(axis=1)
    return float(umin.min()), float(umax.max()), float(vmin.min()), float(vmax.max())
