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


# ------------------------------------------------------------- volumes ---
def _window(vg, axis, n):
    o = vg.get("option", {})
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
    return float(f32(f32(offset) / f32(spacing)) + f32((f32(n) - f32(1.0)) / f32(2.0)))


def _cyl_frame(vec, flip):
    """Cylinder centre, radial (a), tangential (b) and axis vectors of one
    cyl_cone_vec row; `flip` orients a away from the source as
    getCylConeAxes does (src/GeometryUtil3D.cpp:561-583)."""
    s, d, u, v, r = vec[0:3], vec[3:6], vec[6:9], vec[9:12], vec[12]
    a = -np.cross(u, v) * (r / (np.linalg.norm(u) * np.linalg.norm(v)))
    if flip and np.linalg.norm(d - a - s) > np.linalg.norm(d + a - s):
        a = -a
    return d - a, a, u * (r / np.linalg.norm(u)), v


def project_point_3d(pg, x, y, z, i):
    """(u, v) detector index coordinates (0 = centre of the first column /
    row) of world point (x, y, z) in projection i: projectPoint of the 3D
    geometry classes (src/ParallelProjectionGeometry3D.cpp:153-167,
    src/ConeProjectionGeometry3D.cpp:176-200,
    src/ParallelVecProjectionGeometry3D.cpp:201-215,
    src/ConeVecProjectionGeometry3D.cpp:199-214,
    src/CylConeVecProjectionGeometry3D.cpp:343-448).  A cylindrical
    detector that the ray misses gives u = -inf / +inf (the side it passes)
    and v = NaN."""
    t = pg["type"]
    p = np.array([x, y, z], dtype=np.float64)
    nu, nv = int(pg["DetectorColCount"]), int(pg["DetectorRowCount"])
    if t in ("parallel3d", "cone"):
        a = float(_angles(pg)[i])
        su, sv = pg["DetectorSpacingX"], pg["DetectorSpacingY"]
        lateral = np.cos(a) * x + np.sin(a) * y
        if t == "parallel3d":
            return _det_index(lateral, su, nu), _det_index(z, sv, nv)
        sod = float(f32(pg["DistanceOriginSource"]))
        odd = float(f32(pg["DistanceOriginDetector"]))
        depth = -np.sin(a) * x + np.cos(a) * y + sod
        mag = (sod + odd) / depth
        return _det_index(lateral * mag, su, nu), _det_index(z * mag, sv, nv)
    if t in ("parallel3d_vec", "cone_vec"):
        _, vec = vectors_3d(pg)
        src, corner, u, v = vec[i, 0:3], vec[i, 3:6], vec[i, 6:9], vec[i, 9:12]
        if t == "parallel3d_vec":
            # the point moves along the ray onto corner + a u + b v
            den = float(np.dot(src, np.cross(u, v)))
            rel = p - corner
            return (float(np.dot(np.cross(v, src), rel)) / den - 0.5,
                    float(np.dot(np.cross(src, u), rel)) / den - 0.5)
        w = corner - src
        den = float(np.dot(np.cross(u, v), p - src))
        return ((float(np.dot(np.cross(v, w), p)) + float(np.dot(src, np.cross(corner, v)))) / den - 0.5,
                (float(np.dot(np.cross(w, u), p)) - float(np.dot(src, np.cross(corner, u)))) / den - 0.5)
    if t == "cyl_cone_vec":
        _, vec = vectors_3d(pg)
        row = vec[i]
        s, d, u, v, r = row[0:3], row[3:6], row[6:9], row[9:12], row[12]
        c, a, b, axis = _cyl_frame(row, flip=False)
        an, bn, cn = a / np.linalg.norm(a), b / np.linalg.norm(b), axis / np.linalg.norm(axis)
        ray = p - s
        ra, rb, rc = np.dot(ray, an), np.dot(ray, bn), np.dot(ray, cn)
        sa, sb, sc = np.dot(s - c, an), np.dot(s - c, bn), np.dot(s - c, cn)
        # |(t ra + sa, t rb + sb)| = r, the far root
        qa, qb, qc = ra * ra + rb * rb, 2.0 * (ra * sa + rb * sb), sa * sa + sb * sb - r * r
        disc = qb * qb - 4.0 * qa * qc
        if disc >= 0:
            tt = (-qb + np.sqrt(disc)) / (2.0 * qa)
            theta = np.arcsin((tt * rb + sb) / r)
            dth = np.linalg.norm(u) / r
            return (float(theta / dth + 0.5 * nu - 0.5),
                    float((tt * rc + sc) / np.linalg.norm(v) + 0.5 * nv - 0.5))
        side = np.linalg.det(np.array([u, d - s, v])) * np.linalg.det(np.array([ray, d - s, v]))
        return (float("-inf") if side < 0 else float("inf")), float("nan")
    raise AstraError("projectPoint: unsupported geometry type '%s'" % t)


def _bbox_cyl_single(vec, nu, nv, corners, corner_us):
    """getProjectedBBoxSingleAngle of the cylindrical detector
    (src/CylConeVecProjectionGeometry3D.cpp:201-340): u from the projected
    corners and column index 0; v from the corners projected onto a near and
    a far plane normal to the ray through column 0's centre."""
    s = vec[0:3]
    c, a, b, axis = _cyl_frame(vec, flip=True)
    axis_n = axis / np.linalg.norm(axis)
    dth = np.linalg.norm(vec[6:9]) / vec[12]
    us = [0.0] + list(corner_us)
    ray = None
    near = far = 0.0
    for uu in us:
        theta = min(max((uu + 0.5 - 0.5 * nu) * dth, -np.pi / 2), np.pi / 2)
        pt = c + a * np.cos(theta) + b * np.sin(theta)
        if ray is None:
            ray = (pt - c) / np.linalg.norm(pt - c)
            near = far = float(np.dot(pt - s, ray))
        else:
            near = min(near, float(np.dot(pt - s, ray)))
    vmin, vmax = 1e20, -1e20
    for x in corners:
        along = float(np.dot(x - s, ray))
        lat = float(np.dot(x - s, axis_n))
        for plane in (near, far):
            vv = lat * plane / along + 0.5 * nv - 0.5
            vmin, vmax = min(vmin, vv), max(vmax, vv)
    return min(us), max(us), vmin, vmax


def projected_bbox_3d(pg, xmin, xmax, ymin, ymax, zmin, zmax):
    """(umin, umax, vmin, vmax) over all projections of the box's corners
    (CProjectionGeometry3D::getProjectedBBox,
    src/ProjectionGeometry3D.cpp:162-236)."""
    corners = [np.array([cx, cy, cz], dtype=np.float64)
               for cx in (xmin, xmax) for cy in (ymin, ymax) for cz in (zmin, zmax)]
    nu, nv = int(pg["DetectorColCount"]), int(pg["DetectorRowCount"])
    na = proj_shape(pg)[1]
    vecs = vectors_3d(pg)[1] if pg["type"] == "cyl_cone_vec" else None
    g = None
    for i in range(na):
        if vecs is not None:
            b = _bbox_cyl_single(vecs[i], nu, nv, corners,
                                 [project_point_3d(pg, *p, i)[0] for p in corners])
        else:
            uv = [project_point_3d(pg, *p, i) for p in corners]
            us, vs = [q[0] for q in uv], [q[1] for q in uv]
            b = (min(us), max(us), min(vs), max(vs))
        g = b if g is None else (min(g[0], b[0]), max(g[1], b[1]), min(g[2], b[2]), max(g[3], b[3]))
    return g
