"""Oracle 3D projectors (float64), restating the reference CUDA sampling.

Test infrastructure only (see ``oracle/__init__.py``).

Layouts: volume ``[z][y][x]``, projections ``[v][angle][u]``
(cuda/3d/cone_fp.cu:229, cuda/3d/cone_bp.cu:132).  Vectors are the
normalised rows from ``oracle.geometry.normalise_3d``; ``vol_scale`` is the
voxel size triple (``SVolScale3D``) and ``out_scale`` is
``SProjectorParams3D::fOutputScale``.
"""
import math

import numpy as np

from .proj2d import tex2d, tex2d_T

f32 = np.float32

# principal-axis coordinate permutations of DIR_X / DIR_Y / DIR_Z
# (cuda/3d/cone_fp.cu:61-98): (c0, c1, c2) = axes of (slice, dim1, dim2)
_DIRS = {0: (0, 1, 2), 1: (1, 0, 2), 2: (2, 0, 1)}


def _slice_plane(vol, d, s):
    """2D plane of slice s along axis d, indexed [dim2][dim1]."""
    if d == 0:
        return vol[:, :, s]          # [z][y]
    if d == 1:
        return vol[:, s, :]          # [z][x]
    return vol[s, :, :]              # [y][x]


def _shape_xyz(vol):
    nz, ny, nx = vol.shape
    return (nx, ny, nz)


def cone_direction(vec, n_u, n_v):
    """Per-angle principal axis of ``ConeFP_Array_internal``
    (cuda/3d/cone_fp.cu:378-387): largest |src - detector centre| component,
    compared as float32 with ties going to the lower axis."""
    s = vec[0:3]
    c = vec[3:6] + n_u * vec[6:9] * 0.5 + n_v * vec[9:12] * 0.5
    dX, dY, dZ = (abs(f32(s[k] - c[k])) for k in range(3))
    if dX >= dY and dX >= dZ:
        return 0
    if dY >= dX and dY >= dZ:
        return 1
    return 2


def par3d_direction(vec):
    """``Par3DFP_Array_internal`` (cuda/3d/par3d_fp.cu:474-488)."""
    dX, dY, dZ = (abs(f32(vec[k])) for k in range(3))
    if dX >= dY and dX >= dZ:
        return 0
    if dY >= dX and dY >= dZ:
        return 1
    return 2


def _fp_scale(d, vol_scale, out_scale, par3d=False):
    """SCALE_CUBE / SCALE_NONCUBE factors (cuda/3d/cone_fp.cu:100-112,
    342-370; cuda/3d/par3d_fp.cu:442-472).  Returns (s1, s2, outscale)."""
    vx, vy, vz = (f32(x) for x in vol_scale)
    cube = abs(float(vx) / float(vy) - 1.0) <= 1e-5 and abs(float(vx) / float(vz) - 1.0) <= 1e-5
    if cube:
        return 1.0, 1.0, float(f32(out_scale) * vx)
    if d == 0:
        s1, s2, o = vy / vx, vz / vx, vx
    elif d == 1:
        # par3d keeps the reference's fScale2 = volScale.fY / volScale.fY
        # (cuda/3d/par3d_fp.cu:459); cone uses fZ / fY (cone_fp.cu:363)
        s1, s2, o = vx / vy, (vy / vy if par3d else vz / vy), vy
    else:
        s1, s2, o = vx / vz, vy / vz, vz
    return float(s1 * s1), float(s2 * s2), float(f32(out_scale) * o)


def _ray_sum(vol, d, a1, a2, b1, b2):
    """Sum over every slice of the bilinear sample at (f1, f2)."""
    n = _shape_xyz(vol)
    c0, c1, c2 = _DIRS[d]
    n0, n1, n2 = n[c0], n[c1], n[c2]
    acc = np.zeros_like(a1)
    for s in range(n0):
        p = s - 0.5 * n0 + 0.5
        f1 = a1 * p + b1 + 0.5 * n1
        f2 = a2 * p + b2 + 0.5 * n2
        acc += tex2d(_slice_plane(vol, d, s), f1, f2)
    return acc


def _sub_offsets(ss):
    """Sub-ray / sub-voxel offsets (k + 0.5) / ss within a unit cell."""
    return (np.arange(ss) + 0.5) / ss


def cone_fp(vol, vecs, n_u, n_v, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, ss=1):
    """Cone-beam ray-driven FP: ``cone_FP_t`` (cuda/3d/cone_fp.cu:154-231).

    Ray to the centre of detector pixel (u, v) (DetS + (u+.5)U + (v+.5)V),
    written as (c1, c2) = a * c0 + b along the principal axis, sampled at
    every slice centre; the sum is scaled by sqrt(a1^2 s1 + a2^2 s2 + 1).
    ss > 1: ss x ss sub-rays through DetS + (u + (k+.5)/ss) U + (v + (l+.5)/ss) V,
    averaged (``cone_FP_SS_t``, cuda/3d/cone_fp.cu:233-323).
    """
    vol = np.asarray(vol, dtype=np.float64)
    nA = vecs.shape[0]
    proj = np.zeros((n_v, nA, n_u))
    for a in range(nA):
        vec = vecs[a]
        d = cone_direction(vec, n_u, n_v)
        c0, c1, c2 = _DIRS[d]
        src = vec[0:3]
        s1, s2, osc = _fp_scale(d, vol_scale, out_scale)
        acc = np.zeros((n_v, n_u))
        for ou in _sub_offsets(ss):
            for ov in _sub_offsets(ss):
                uu, vv = np.meshgrid(np.arange(n_u) + ou, np.arange(n_v) + ov)
                det = [vec[3 + k] + uu * vec[6 + k] + vv * vec[9 + k] for k in range(3)]
                den = src[c0] - det[c0]
                a1 = (src[c1] - det[c1]) / den
                a2 = (src[c2] - det[c2]) / den
                b1 = src[c1] - a1 * src[c0]
                b2 = src[c2] - a2 * src[c0]
                acc += _ray_sum(vol, d, a1, a2, b1, b2) * np.sqrt(a1 * a1 * s1 + a2 * a2 * s2 + 1.0)
        proj[:, a, :] = acc * osc / (ss * ss)
    return proj


def par3d_fp(vol, vecs, n_u, n_v, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, ss=1):
    """Parallel-beam 3D FP: ``par3D_FP_t`` (cuda/3d/par3d_fp.cu:152-230);
    ss > 1 as in ``par3D_FP_SS_t`` (:234-325)."""
    vol = np.asarray(vol, dtype=np.float64)
    nA = vecs.shape[0]
    proj = np.zeros((n_v, nA, n_u))
    for a in range(nA):
        vec = vecs[a]
        d = par3d_direction(vec)
        c0, c1, c2 = _DIRS[d]
        ray = vec[0:3]
        a1 = ray[c1] / ray[c0]
        a2 = ray[c2] / ray[c0]
        s1, s2, osc = _fp_scale(d, vol_scale, out_scale, par3d=True)
        acc = np.zeros((n_v, n_u))
        for ou in _sub_offsets(ss):
            for ov in _sub_offsets(ss):
                uu, vv = np.meshgrid(np.arange(n_u) + ou, np.arange(n_v) + ov)
                det = [vec[3 + k] + uu * vec[6 + k] + vv * vec[9 + k] for k in range(3)]
                b1 = det[c1] - a1 * det[c0]
                b2 = det[c2] - a2 * det[c0]
                acc += _ray_sum(vol, d, np.full_like(b1, a1), np.full_like(b2, a2), b1, b2)
        proj[:, a, :] = acc * math.sqrt(a1 * a1 * s1 + a2 * a2 * s2 + 1.0) * osc / (ss * ss)
    return proj


def _ray_scatter(vol, d, a1, a2, b1, b2, vals):
    """Transpose of ``_ray_sum``: add vals x bilinear weights of every slice
    sample into vol (in place)."""
    n = _shape_xyz(vol)
    c0, c1, c2 = _DIRS[d]
    n0, n1, n2 = n[c0], n[c1], n[c2]
    for s in range(n0):
        p = s - 0.5 * n0 + 0.5
        f1 = a1 * p + b1 + 0.5 * n1
        f2 = a2 * p + b2 + 0.5 * n2
        plane = _slice_plane(vol, d, s)
        plane += tex2d_T(plane.shape, f1, f2, vals)


def cone_fp_T(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, ss=1):
    """Exact transpose of ``cone_fp`` (the matched adjoint: the same rays,
    samples and weights as the forward projection, scattered); ss x ss
    sub-rays per pixel as in ``cone_fp``."""
    proj = np.asarray(proj, dtype=np.float64)
    n_v, nA, n_u = proj.shape
    vol = np.zeros(shape)
    for a in range(nA):
        vec = vecs[a]
        d = cone_direction(vec, n_u, n_v)
        c0, c1, c2 = _DIRS[d]
        src = vec[0:3]
        s1, s2, osc = _fp_scale(d, vol_scale, out_scale)
        for ou in _sub_offsets(ss):
            for ov in _sub_offsets(ss):
                uu, vv = np.meshgrid(np.arange(n_u) + ou, np.arange(n_v) + ov)
                det = [vec[3 + k] + uu * vec[6 + k] + vv * vec[9 + k] for k in range(3)]
                den = src[c0] - det[c0]
                a1 = (src[c1] - det[c1]) / den
                a2 = (src[c2] - det[c2]) / den
                b1 = src[c1] - a1 * src[c0]
                b2 = src[c2] - a2 * src[c0]
                vals = proj[:, a, :] * np.sqrt(a1 * a1 * s1 + a2 * a2 * s2 + 1.0) * osc / (ss * ss)
                _ray_scatter(vol, d, a1, a2, b1, b2, vals)
    return vol


def par3d_fp_T(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, ss=1):
    """Exact transpose of ``par3d_fp`` (ss x ss sub-rays per pixel)."""
    proj = np.asarray(proj, dtype=np.float64)
    n_v, nA, n_u = proj.shape
    vol = np.zeros(shape)
    for a in range(nA):
        vec = vecs[a]
        d = par3d_direction(vec)
        c0, c1, c2 = _DIRS[d]
        ray = vec[0:3]
        a1 = ray[c1] / ray[c0]
        a2 = ray[c2] / ray[c0]
        s1, s2, osc = _fp_scale(d, vol_scale, out_scale, par3d=True)
        vals = proj[:, a, :] * math.sqrt(a1 * a1 * s1 + a2 * a2 * s2 + 1.0) * osc / (ss * ss)
        for ou in _sub_offsets(ss):
            for ov in _sub_offsets(ss):
                uu, vv = np.meshgrid(np.arange(n_u) + ou, np.arange(n_v) + ov)
                det = [vec[3 + k] + uu * vec[6 + k] + vv * vec[9 + k] for k in range(3)]
                b1 = det[c1] - a1 * det[c0]
                b2 = det[c2] - a2 * det[c0]
                _ray_scatter(vol, d, np.full_like(b1, a1), np.full_like(b2, a2), b1, b2, vals)
    return vol


def _dir_weight_sq(f, n):
    """``dirWeights`` (cuda/3d/par3d_fp.cu:329-341): sum of the squared
    linear weights of texture coordinate f on an axis of n texels."""
    f = np.asarray(f, dtype=np.float64)
    t = f + 0.5 - np.floor(f + 0.5)
    return np.select([f <= -0.5, f <= 0.5, f <= n - 0.5, f <= n + 0.5],
                     [0.0, (f + 0.5) ** 2, t * t + (1 - t) ** 2, (n + 0.5 - f) ** 2], 0.0)


def par3d_fp_sumsq(vecs, n_u, n_v, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0):
    """par3d FP with the sum_square_weights kernel (``par3D_FP_SumSqW_t``,
    cuda/3d/par3d_fp.cu:344-419): per ray, the sum over slices of the squared
    bilinear weights, times the squared distance weight."""
    nz, ny, nx = shape
    n = (nx, ny, nz)
    nA = vecs.shape[0]
    proj = np.zeros((n_v, nA, n_u))
    uu, vv = np.meshgrid(np.arange(n_u) + 0.5, np.arange(n_v) + 0.5)
    for a in range(nA):
        vec = vecs[a]
        d = par3d_direction(vec)
        c0, c1, c2 = _DIRS[d]
        n0, n1, n2 = n[c0], n[c1], n[c2]
        ray = vec[0:3]
        det = [vec[3 + k] + uu * vec[6 + k] + vv * vec[9 + k] for k in range(3)]
        a1 = ray[c1] / ray[c0]
        a2 = ray[c2] / ray[c0]
        b1 = det[c1] - a1 * det[c0]
        b2 = det[c2] - a2 * det[c0]
        s1, s2, osc = _fp_scale(d, vol_scale, out_scale, par3d=True)
        acc = np.zeros_like(b1)
        for sl in range(n0):
            p = sl - 0.5 * n0 + 0.5
            acc += _dir_weight_sq(a1 * p + b1 + 0.5 * n1, n1) * _dir_weight_sq(a2 * p + b2 + 0.5 * n2, n2)
        proj[:, a, :] = acc * (math.sqrt(a1 * a1 * s1 + a2 * a2 * s2 + 1.0) * osc) ** 2
    return proj


# ---------------------------------------------------------------- BP ----

def _det3x(b, c):
    return b[1] * c[2] - b[2] * c[1]


def _det3y(b, c):
    return -(b[0] * c[2] - b[2] * c[0])


def _det3z(b, c):
    return b[0] * c[1] - b[1] * c[0]


def _det3(a, b, c):
    return a[0] * _det3x(b, c) + a[1] * _det3y(b, c) + a[2] * _det3z(b, c)


def _scaled_cross_norm(a, b, sc):
    """``scaled_cross3(a, b, sc).norm()`` (include/astra/cuda/3d/util3d.h:82-88)."""
    cx = _det3x(a, b) * sc[1] * sc[2]
    cy = _det3y(a, b) * sc[0] * sc[2]
    cz = _det3z(a, b) * sc[0] * sc[1]
    return math.sqrt(cx * cx + cy * cy + cz * cz)


def cone_bp_params(vec, kernel, vol_scale):
    """``DevConeParams`` (cuda/3d/cone_bp.cu:221-302): NumU, NumV, Den as
    (x, y, z, w) coefficient quadruples, float64."""
    s = np.asarray(vec[0:3], dtype=np.float64)
    d = np.asarray(vec[3:6], dtype=np.float64)
    u = np.asarray(vec[6:9], dtype=np.float64)
    v = np.asarray(vec[9:12], dtype=np.float64)
    smd = s - d
    if kernel == "fdk":
        scale = 1.0 / _det3(u, v, s)
    elif kernel == "2d":
        scale = _scaled_cross_norm(u, (0.0, 0.0, 1.0), vol_scale) / _det3(u, v, smd)
    else:
        scale = math.sqrt(_scaled_cross_norm(u, v, vol_scale)) / _det3(u, v, smd)
    num_u = (scale * _det3x(v, smd), scale * _det3y(v, smd), scale * _det3z(v, smd), scale * _det3(s, v, d))
    num_v = (-scale * _det3x(u, smd), -scale * _det3y(u, smd), -scale * _det3z(u, smd), -scale * _det3(s, u, d))
    den = (-scale * _det3x(u, v), -scale * _det3y(u, v), -scale * _det3z(u, v), scale * _det3(u, v, s))
    return num_u, num_v, den


def _voxel_grid(shape, off=(0.5, 0.5, 0.5)):
    """Voxel (sub-)positions: X - nx/2 + off_x etc. (off = 0.5: centres)."""
    nz, ny, nx = shape
    X = np.arange(nx) - 0.5 * nx + off[0]
    Y = np.arange(ny) - 0.5 * ny + off[1]
    Z = np.arange(nz) - 0.5 * nz + off[2]
    return np.meshgrid(Z, Y, X, indexing="ij")


def _sub_voxels(ss):
    o = _sub_offsets(ss)
    return [(a, b, c) for a in o for b in o for c in o]


def cone_bp(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, kernel="default", ss=1):
    """Cone-beam voxel-driven BP: ``dev_cone_BP<KERNEL>``
    (cuda/3d/cone_bp.cu:55-133) and its output scale (``ConeBP_Array``,
    :322-328).  kernel: 'default' (weight fr^2), 'fdk' (Den.w = 1, fr^2,
    output / voxel x), '2d' (weight fr).  ss > 1 (default kernel only, as
    ``ConeBP_Array`` dispatches, :351-370): ss^3 sub-voxels at offsets
    (k + 0.5) / ss, averaged (``dev_cone_BP_SS``, :138-217)."""
    proj = np.asarray(proj, dtype=np.float64)
    if kernel != "default":
        ss = 1
    vol = np.zeros(shape)
    for a in range(vecs.shape[0]):
        nu, nv, dn = cone_bp_params(vecs[a], kernel, vol_scale)
        for off in _sub_voxels(ss):
            fz, fy, fx = _voxel_grid(shape, off)
            U = nu[3] + fx * nu[0] + fy * nu[1] + fz * nu[2]
            V = nv[3] + fx * nv[0] + fy * nv[1] + fz * nv[2]
            D = (1.0 if kernel == "fdk" else dn[3]) + fx * dn[0] + fy * dn[1] + fz * dn[2]
            fr = 1.0 / D
            val = tex2d(proj[:, a, :], U * fr, V * fr)
            vol += (fr if kernel == "2d" else fr * fr) * val / ss ** 3
    vx, vy, vz = (f32(x) for x in vol_scale)
    if kernel == "fdk":
        osc = float(f32(out_scale) / vx)
    else:
        osc = float(f32(out_scale) * (vx * vy * vz))
    return vol * osc


def par3d_bp_params(vec, kernel, vol_scale):
    """``DevPar3DParams`` + scale (cuda/3d/par3d_bp.cu:202-244)."""
    r = np.asarray(vec[0:3], dtype=np.float64)
    d = np.asarray(vec[3:6], dtype=np.float64)
    u = np.asarray(vec[6:9], dtype=np.float64)
    v = np.asarray(vec[9:12], dtype=np.float64)
    den = _det3(r, u, v)
    num_u = (-_det3x(r, v) / den, -_det3y(r, v) / den, -_det3z(r, v) / den, -_det3(r, d, v) / den)
    num_v = (_det3x(r, u) / den, _det3y(r, u) / den, _det3z(r, u) / den, _det3(r, d, u) / den)
    if kernel == "2d":
        s = 1.0 / _scaled_cross_norm(u, (0.0, 0.0, 1.0), vol_scale)
    else:
        s = 1.0 / _scaled_cross_norm(u, v, vol_scale)
    return num_u, num_v, s


def par3d_bp(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, kernel="default", ss=1):
    """Parallel-beam 3D BP: ``dev_par3D_BP`` (cuda/3d/par3d_bp.cu:55-120),
    output scale volScale x*y*z (:263); ss > 1 as ``dev_par3D_BP_SS``
    (:123-198)."""
    proj = np.asarray(proj, dtype=np.float64)
    vol = np.zeros(shape)
    for a in range(vecs.shape[0]):
        nu, nv, s = par3d_bp_params(vecs[a], kernel, vol_scale)
        for off in _sub_voxels(ss):
            fz, fy, fx = _voxel_grid(shape, off)
            U = nu[3] + fx * nu[0] + fy * nu[1] + fz * nu[2]
            V = nv[3] + fx * nv[0] + fy * nv[1] + fz * nv[2]
            vol += tex2d(proj[:, a, :], U, V) * s / ss ** 3
    vx, vy, vz = (f32(x) for x in vol_scale)
    return vol * float(f32(out_scale) * (vx * vy * vz))


# ------------------------------------------------ cylindrical detector ------
def cyl_axes(vec):
    """``getCylConeAxes`` (src/GeometryUtil3D.cpp:542-561): radial a (from
    the cylinder axis to the detector centre, pointing away from the
    source), tangential b = u R / |u|, cylinder centre c = d - a."""
    s = np.asarray(vec[0:3], dtype=np.float64)
    d = np.asarray(vec[3:6], dtype=np.float64)
    u = np.asarray(vec[6:9], dtype=np.float64)
    v = np.asarray(vec[9:12], dtype=np.float64)
    R = float(vec[12])
    a = -np.cross(u, v) * (R / (np.linalg.norm(u) * np.linalg.norm(v)))
    if np.linalg.norm(d - a - s) > np.linalg.norm(d + a - s):
        a = -a
    return a, u * (R / np.linalg.norm(u)), d - a


def cylcone_direction(vec):
    """Principal axis of ``ConeCylFP_Array_internal`` (cuda/3d/cone_cyl.cu:
    486-500): largest |src - detector centre| component (world units,
    compared as float32, ties to the lower axis)."""
    dX, dY, dZ = (abs(f32(vec[k] - vec[3 + k])) for k in range(3))
    if dX >= dY and dX >= dZ:
        return 0
    if dY >= dX and dY >= dZ:
        return 1
    return 2


def cylcone_pixels(vec, n_u, n_v):
    """World positions [3][n_v][n_u] of the detector pixel centres:
    c + a cos(theta) + b sin(theta) + v (row - n_v/2 + 0.5), theta =
    (col - n_u/2 + 0.5) |u| / R (cuda/3d/cone_cyl.cu:243-258)."""
    a, b, c = cyl_axes(vec)
    v = np.asarray(vec[9:12], dtype=np.float64)
    dth = np.linalg.norm(vec[6:9]) / float(vec[12])
    uu, vv = np.meshgrid(np.arange(n_u), np.arange(n_v))
    th = (uu - 0.5 * n_u + 0.5) * dth
    fv = vv - 0.5 * n_v + 0.5
    return [c[k] + a[k] * np.cos(th) + b[k] * np.sin(th) + v[k] * fv for k in range(3)]


def cylcone_fp(vol, vecs, n_u, n_v, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0):
    """Cone-beam FP onto a cylindrical detector: ``cylcone_FP_t``
    (cuda/3d/cone_cyl.cu:197-298), with ideal bilinear sampling.  Vectors are
    the translated (unscaled) 13-tuples of ``translate_cylcone``; the ray
    from the source to each pixel centre is written per slice in texel
    units of the principal-axis group and weighted by its world length per
    slice, sqrt(a1^2 + a2^2 + 1) x voxel size along the axis."""
    vol = np.asarray(vol, dtype=np.float64)
    vs = [float(f32(x)) for x in vol_scale]
    nA = vecs.shape[0]
    proj = np.zeros((n_v, nA, n_u))
    for a in range(nA):
        vec = vecs[a]
        d = cylcone_direction(vec)
        c0, c1, c2 = _DIRS[d]
        src = vec[0:3]
        det = cylcone_pixels(vec, n_u, n_v)
        A1 = (src[c1] - det[c1]) / (src[c0] - det[c0])
        A2 = (src[c2] - det[c2]) / (src[c0] - det[c0])
        dist = np.sqrt(A1 * A1 + A2 * A2 + 1.0) * vs[c0]
        a1 = A1 / vs[c1]
        a2 = A2 / vs[c2]
        b1 = src[c1] / vs[c1] - a1 * src[c0]
        b2 = src[c2] / vs[c2] - a2 * src[c0]
        proj[:, a, :] = _ray_sum(vol, d, a1 * vs[c0], a2 * vs[c0], b1, b2) * dist * out_scale
    return proj


def cylcone_bp_params(vec):
    """``DevCylConeParamsBP`` (cuda/3d/cone_cyl_bp.cu:165-230), float64:
    rows (x, y, z, w) of the source-relative ray in the cylinder frame
    (radial, tangential, axial), scaled by sqrt(|u x v|) so that the ray
    density is 1 / t^2; the source in that frame; 1 / angular step; 1/|v|."""
    s = np.asarray(vec[0:3], dtype=np.float64)
    u = np.asarray(vec[6:9], dtype=np.float64)
    v = np.asarray(vec[9:12], dtype=np.float64)
    R = float(vec[12])
    a, b, c = cyl_axes(vec)
    an, bn, vn = a / np.linalg.norm(a), b / np.linalg.norm(b), v / np.linalg.norm(v)
    sc = math.sqrt(np.linalg.norm(np.cross(u, v)))
    rows = [np.r_[e * sc, -np.dot(s, e) * sc] for e in (an, bn, vn)]
    sm = [float(np.dot(s - c, e)) for e in (an, bn, vn)]
    return rows, sm, np.linalg.norm(u) / R, 1.0 / np.linalg.norm(v), R


def cylcone_detector_coords(vec, n_u, n_v, X, Y, Z):
    """Texture coordinates (fU, fV) on the cylindrical detector of the ray
    from the source through world points (X, Y, Z) (translated frame), the
    ray parameter t of the hit and the hit mask (cuda/3d/cone_cyl_bp.cu:
    104-134)."""
    (ca, cb, cc), (sa, sb, sc), dth, inv_nv, R = cylcone_bp_params(vec)
    ra = ca[3] + X * ca[0] + Y * ca[1] + Z * ca[2]
    rb = cb[3] + X * cb[0] + Y * cb[1] + Z * cb[2]
    rc = cc[3] + X * cc[0] + Y * cc[1] + Z * cc[2]
    A = ra * ra + rb * rb
    B = ra * sa + rb * sb
    C = sa * sa + sb * sb - R * R
    D = B * B - A * C
    ok = D >= 0
    t = (-B + np.sqrt(np.where(ok, D, 0.0))) / A
    th = np.arcsin(np.clip((t * rb + sb) / R, -1.0, 1.0)) / dth
    return th + 0.5 * n_u, (t * rc + sc) * inv_nv + 0.5 * n_v, t, ok


def cylcone_bp(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0):
    """Voxel-driven BP from a cylindrical detector: ``dev_cylcone_BP``
    (cuda/3d/cone_cyl_bp.cu:58-158) and its output scale (:246).  The ray
    from the source through the voxel meets the cylinder at parameter t
    (larger root of A t^2 + 2 B t + C = 0); the detector coordinates there
    are (asin(b / R) / angular step + n_u / 2, axial / |v| + n_v / 2)
    (texture coordinates); weight t^2.  A voxel whose ray misses the
    cylinder contributes nothing."""
    proj = np.asarray(proj, dtype=np.float64)
    n_v, nA, n_u = proj.shape
    vx, vy, vz = (f32(x) for x in vol_scale)
    fz, fy, fx = _voxel_grid(shape)
    fx, fy, fz = fx * float(vx), fy * float(vy), fz * float(vz)
    vol = np.zeros(shape)
    for a in range(nA):
        fU, fV, t, ok = cylcone_detector_coords(vecs[a], n_u, n_v, fx, fy, fz)
        vol += np.where(ok, tex2d(proj[:, a, :], fU, fV) * t * t, 0.0)
    return vol * float(f32(out_scale) * (vx * vy * vz))
