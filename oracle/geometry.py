"""Oracle geometry: angle -> projection vectors, and volume normalisation.

Test infrastructure only (see ``oracle/__init__.py``).

Conventions (reference ``include/astra/GeometryUtil2D.h:53-116``,
``include/astra/GeometryUtil3D.h:47-123``):

* 2D vectors are stored as float32 (the reference's ``SParProjection`` /
  ``SFanProjection`` hold ``float``); rows are ``[rayX, rayY, dSX, dSY, uX, uY]``
  (parallel) or ``[srcX, srcY, dSX, dSY, uX, uY]`` (fan).
* 3D vectors are float64 (``SConeProjection`` / ``SPar3DProjection`` hold
  ``double``); rows are ``[sX,sY,sZ, dSX,dSY,dSZ, uX,uY,uZ, vX,vY,vZ]`` where
  ``s`` is the source (cone) or ray direction (par3d) and ``dS`` is the
  detector corner ("bottom left"), ``u``/``v`` one pixel edge each.
* Angles are float32, as the reference stores them (``const float *pfAngles``).
"""
import numpy as np

f32 = np.float32


def _rot2(x, y, a):
    """ROTATE0 macro: (x cos a - y sin a, x sin a + y cos a), in double."""
    a = np.float64(a)
    return x * np.cos(a) - y * np.sin(a), x * np.sin(a) + y * np.cos(a)


def gen_par_projections(angles, n_det, det_size):
    """``genParProjections`` (src/GeometryUtil2D.cpp:40-79), float32 storage."""
    angles = np.asarray(angles, dtype=f32)
    base_ray = (0.0, 1.0)
    base_ds = (float(f32(n_det * float(det_size) * -0.5)), 0.0)
    base_u = (float(f32(det_size)), 0.0)
    out = np.zeros((len(angles), 6), dtype=np.float64)
    for i, a in enumerate(angles):
        for k, (bx, by) in enumerate((base_ray, base_ds, base_u)):
            x, y = _rot2(bx, by, a)
            out[i, 2 * k] = f32(x)
            out[i, 2 * k + 1] = f32(y)
    return out


def gen_fan_projections(angles, n_det, src_origin, origin_det, det_size):
    """``genFanProjections`` (src/GeometryUtil2D.cpp:82-107), float32 storage.

    The base vectors are float32 there too (``float fSrcY0 = -fOriginSource``,
    ``fDetSX0 = iProjDets * fDetUX0 / -2.0f``).
    """
    angles = np.asarray(angles, dtype=f32)
    sx0, sy0 = f32(0.0), f32(-src_origin)
    ux0, uy0 = f32(det_size), f32(0.0)
    dsx0 = f32(f32(n_det) * ux0 / f32(-2.0))
    dsy0 = f32(origin_det)
    out = np.zeros((len(angles), 6), dtype=np.float64)
    for i, a in enumerate(angles):
        for k, (bx, by) in enumerate(((sx0, sy0), (dsx0, dsy0), (ux0, uy0))):
            x, y = _rot2(float(bx), float(by), a)
            out[i, 2 * k] = f32(x)
            out[i, 2 * k + 1] = f32(y)
    return out


def fanvec_from_astra(vectors, n_det):
    """``fanflat_vec`` rows (src, detector centre, u) -> corner form.

    Follows src/FanFlatVecProjectionGeometry2D.cpp:103-125:
    DetS = centre - 0.5 * nDet * u.
    """
    v = np.asarray(vectors, dtype=np.float64)
    out = np.zeros((v.shape[0], 6))
    out[:, 0:2] = v[:, 0:2].astype(f32)
    out[:, 4:6] = v[:, 4:6].astype(f32)
    out[:, 2:4] = (v[:, 2:4] - 0.5 * n_det * out[:, 4:6]).astype(f32)
    return out


def parvec_from_astra(vectors, n_det):
    """``parallel_vec`` rows (ray, detector centre, u) -> corner form."""
    return fanvec_from_astra(vectors, n_det)


def normalise_2d(vecs, kind, vol_geom):
    """``convertAstraGeometry_internal`` (src/GeometryUtil2D.cpp:151-179).

    Translates by minus the window centre and scales by 1/pixel, all stored
    back as float32.  Returns ``(vecs, output_scale)`` where output_scale is
    the pixel length (``fOutputScale *= pixelLengthX``).
    """
    minx, maxx, miny, maxy, nx, ny = _vol2d_window(vol_geom)
    px = f32((f32(maxx) - f32(minx)) / f32(nx))
    py = f32((f32(maxy) - f32(miny)) / f32(ny))
    if abs(float(px) - float(py)) > 1e-5:
        raise ValueError("non-square pixels are not supported by the 2D GPU path")
    dx = float(f32(-(f32(minx) + f32(maxx)) / f32(2)))
    dy = float(f32(-(f32(miny) + f32(maxy)) / f32(2)))
    fac = float(f32(f32(1.0) / px))
    out = np.array(vecs, dtype=np.float64)
    if kind == "fan":
        out[:, 0] = out[:, 0] + dx
        out[:, 1] = out[:, 1] + dy
    out[:, 2] = out[:, 2] + dx
    out[:, 3] = out[:, 3] + dy
    out = out.astype(f32).astype(np.float64)
    out = (out * fac).astype(f32).astype(np.float64)
    return out, float(px)


def _vol2d_window(vg):
    nx = vg["GridColCount"]
    ny = vg["GridRowCount"]
    o = vg.get("option", {})
    minx = o.get("WindowMinX", -nx / 2.0)
    maxx = o.get("WindowMaxX", nx / 2.0)
    miny = o.get("WindowMinY", -ny / 2.0)
    maxy = o.get("WindowMaxY", ny / 2.0)
    return minx, maxx, miny, maxy, nx, ny


def gen_cone_projections(angles, n_u, n_v, sod, odd, du, dv):
    """``genConeProjections`` (src/GeometryUtil3D.cpp:42-83), float64 storage."""
    angles = np.asarray(angles, dtype=f32)
    sod, odd, du, dv = (float(f32(x)) for x in (sod, odd, du, dv))
    base = [
        (0.0, -sod, 0.0),
        (n_u * du * -0.5, odd, n_v * dv * -0.5),
        (du, 0.0, 0.0),
        (0.0, 0.0, dv),
    ]
    out = np.zeros((len(angles), 12))
    for i, a in enumerate(angles):
        for k, (bx, by, bz) in enumerate(base):
            x, y = _rot2(bx, by, a)
            out[i, 3 * k:3 * k + 3] = (x, y, bz)
    return out


def gen_par3d_projections(angles, n_u, n_v, du, dv):
    """``genPar3DProjections`` (src/GeometryUtil3D.cpp:133-172)."""
    angles = np.asarray(angles, dtype=f32)
    du, dv = float(f32(du)), float(f32(dv))
    base = [
        (0.0, 1.0, 0.0),
        (n_u * du * -0.5, 0.0, n_v * dv * -0.5),
        (du, 0.0, 0.0),
        (0.0, 0.0, dv),
    ]
    out = np.zeros((len(angles), 12))
    for i, a in enumerate(angles):
        for k, (bx, by, bz) in enumerate(base):
            x, y = _rot2(bx, by, a)
            out[i, 3 * k:3 * k + 3] = (x, y, bz)
    return out


def conevec_from_astra(vectors, n_u, n_v):
    """``cone_vec`` rows (src, detector centre, u, v) -> corner form.

    src/ConeVecProjectionGeometry3D.cpp:82-106: DetS = c - 0.5 nV v - 0.5 nU u.
    """
    v = np.asarray(vectors, dtype=np.float64)
    out = v.copy()
    out[:, 3:6] = v[:, 3:6] - 0.5 * n_v * v[:, 9:12] - 0.5 * n_u * v[:, 6:9]
    return out


def normalise_3d(vecs, kind, vol_geom):
    """``convertAstraGeometry_internal`` (src/GeometryUtil3D.cpp:569-593).

    Translate (float32 window centre) then scale by float32 1/voxel size.
    Returns ``(vecs, vol_scale)``.
    """
    win = _vol3d_window(vol_geom)
    (minx, maxx, nx), (miny, maxy, ny), (minz, maxz, nz) = win
    d = [float(f32(-(f32(lo) + f32(hi)) / f32(2))) for lo, hi, _ in win]
    pix = [f32((f32(hi) - f32(lo)) / f32(n)) for lo, hi, n in win]
    fac = [float(f32(f32(1.0) / p)) for p in pix]
    out = np.array(vecs, dtype=np.float64)
    if kind == "cone":
        out[:, 0:3] += d
    out[:, 3:6] += d
    for k in range(4):
        out[:, 3 * k:3 * k + 3] *= fac
    return out, tuple(float(p) for p in pix)


def translate_cylcone(vecs, vol_geom):
    """``convertAstraGeometry_unscaled_internal`` (src/GeometryUtil3D.cpp:
    595-616) for ``cyl_cone_vec``: translate source and detector centre by
    the float32 window centre; no scaling (the kernels take voxel sizes).
    Rows are 13-tuples (src, detector centre, u, v, radius).  Returns
    ``(vecs, vol_scale)``."""
    win = _vol3d_window(vol_geom)
    d = [float(f32(-(f32(lo) + f32(hi)) / f32(2))) for lo, hi, _ in win]
    pix = [f32((f32(hi) - f32(lo)) / f32(n)) for lo, hi, n in win]
    out = np.array(vecs, dtype=np.float64)
    out[:, 0:3] += d
    out[:, 3:6] += d
    return out, tuple(float(p) for p in pix)


def _vol3d_window(vg):
    nx = vg["GridColCount"]
    ny = vg["GridRowCount"]
    nz = vg["GridSliceCount"]
    o = vg.get("option", {})
    return (
        (o.get("WindowMinX", -nx / 2.0), o.get("WindowMaxX", nx / 2.0), nx),
        (o.get("WindowMinY", -ny / 2.0), o.get("WindowMaxY", ny / 2.0), ny),
        (o.get("WindowMinZ", -nz / 2.0), o.get("WindowMaxZ", nz / 2.0), nz),
    )


def vol_geom_2d(n_rows, n_cols):
    return {"GridRowCount": n_rows, "GridColCount": n_cols}


def vol_geom_3d(n_rows, n_cols, n_slices):
    return {"GridRowCount": n_rows, "GridColCount": n_cols, "GridSliceCount": n_slices}
