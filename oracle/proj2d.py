"""Oracle 2D projectors (float64), restating the reference CUDA sampling.

Test infrastructure only (see ``oracle/__init__.py``).

Layouts: volume ``[row][col]`` (row axis is world -y), sinogram
``[angle][det]``.  Vectors are the *normalised* ones (voxel units) produced by
``oracle.geometry.normalise_2d``; ``out_scale`` is the reference's
``SProjectorParams2D::fOutputScale`` (pixel length times any caller scale).

Every function returns the contribution that the reference *adds* to its
output buffer.
"""
import math

import numpy as np

f32 = np.float32


def tex1d(line, c):
    """Ideal linear texture fetch with border 0: centre of texel i is i+0.5."""
    line = np.asarray(line, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    t = c - 0.5
    i0 = np.floor(t)
    w = t - i0
    i0 = i0.astype(np.int64)
    n = line.shape[-1]
    v0 = np.where((i0 >= 0) & (i0 < n), line[np.clip(i0, 0, n - 1)], 0.0)
    i1 = i0 + 1
    v1 = np.where((i1 >= 0) & (i1 < n), line[np.clip(i1, 0, n - 1)], 0.0)
    return (1.0 - w) * v0 + w * v1


def tex2d(img, x, y):
    """Ideal bilinear fetch of ``img[y][x]`` (x = column coordinate)."""
    img = np.asarray(img, dtype=np.float64)
    h, w_ = img.shape
    tx = np.asarray(x, dtype=np.float64) - 0.5
    ty = np.asarray(y, dtype=np.float64) - 0.5
    x0 = np.floor(tx)
    y0 = np.floor(ty)
    wx = tx - x0
    wy = ty - y0
    x0 = x0.astype(np.int64)
    y0 = y0.astype(np.int64)

    def at(yy, xx):
        ok = (xx >= 0) & (xx < w_) & (yy >= 0) & (yy < h)
        return np.where(ok, img[np.clip(yy, 0, h - 1), np.clip(xx, 0, w_ - 1)], 0.0)

    return ((1 - wy) * ((1 - wx) * at(y0, x0) + wx * at(y0, x0 + 1))
            + wy * ((1 - wx) * at(y0 + 1, x0) + wx * at(y0 + 1, x0 + 1)))


def _scatter(n, idx, vals):
    out = np.bincount(idx.ravel(), weights=vals.ravel(), minlength=n + 1)
    return out[:n]


def tex1d_T(n, c, vals):
    """Transpose of ``tex1d``: a line of length n receiving every sample's
    value times its two linear weights (taps outside the line dropped)."""
    t = np.asarray(c, dtype=np.float64) - 0.5
    i0 = np.floor(t)
    w = t - i0
    i0 = i0.astype(np.int64)
    vals = np.broadcast_to(np.asarray(vals, dtype=np.float64), t.shape)
    out = np.zeros(n)
    for i, wt in ((i0, 1.0 - w), (i0 + 1, w)):
        ok = (i >= 0) & (i < n)
        out += _scatter(n, np.where(ok, i, n), np.where(ok, wt * vals, 0.0))
    return out


def tex2d_T(shape, x, y, vals):
    """Transpose of ``tex2d``: an image ``[h][w]`` receiving every sample's
    value times its four bilinear weights."""
    h, w_ = shape
    tx = np.asarray(x, dtype=np.float64) - 0.5
    ty = np.asarray(y, dtype=np.float64) - 0.5
    x0 = np.floor(tx)
    y0 = np.floor(ty)
    wx = tx - x0
    wy = ty - y0
    x0 = x0.astype(np.int64)
    y0 = y0.astype(np.int64)
    vals = np.broadcast_to(np.asarray(vals, dtype=np.float64), tx.shape)
    out = np.zeros(h * w_)
    for yy, xx, wt in ((y0, x0, (1 - wy) * (1 - wx)), (y0, x0 + 1, (1 - wy) * wx),
                       (y0 + 1, x0, wy * (1 - wx)), (y0 + 1, x0 + 1, wy * wx)):
        ok = (xx >= 0) & (xx < w_) & (yy >= 0) & (yy < h)
        out += _scatter(h * w_, np.where(ok, yy * w_ + xx, h * w_), np.where(ok, wt * vals, 0.0))
    return out.reshape(h, w_)


def par_parameters(vec, n_det):
    """``getParParameters`` (src/GeometryUtil2D.cpp:111-134), in float64.

    Returns (angle, det_size, offset) of the equivalent circular geometry.
    """
    rx, ry, dsx, dsy, ux, uy = (float(x) for x in vec)
    t = (ux * rx + uy * ry) / (rx * rx + ry * ry)
    ux -= t * rx
    uy -= t * ry
    angle = math.atan2(uy, ux)
    norm2 = uy * uy + ux * ux
    det_size = math.sqrt(norm2)
    offset = -0.5 * n_det - (dsy * uy + dsx * ux) / norm2
    return angle, det_size, offset


def par_fp(vol, vecs, n_det, out_scale=1.0, rays=1):
    """2D parallel Joseph FP.

    ``FP_simple_internal`` + ``FPhorizontal_simple`` / ``FPvertical_simple``
    (cuda/2d/par_fp.cu:53-125, 130-200, 233-301).  Orientation per angle:
    vertical iff |rayX| <= |rayY| (par_fp.cu:272).  The oracle uses exact
    cos/sin of the recovered angle (the reference uses fast intrinsics).
    rays > 1: sub-rays at detector offsets (k + 0.5) / rays - 0.5, averaged
    (par_fp.cu:94-111).
    """
    if rays > 1:
        return sum(par_fp(vol, _shift_det(vecs, o), n_det, out_scale) for o in _sub_det(rays)) / rays
    vol = np.asarray(vol, dtype=np.float64)
    H, W = vol.shape
    nA = vecs.shape[0]
    sino = np.zeros((nA, n_det))
    det = np.arange(n_det, dtype=np.float64)
    for a in range(nA):
        theta, dsize, off = par_parameters(vecs[a], n_det)
        c, s = math.cos(theta), math.sin(theta)
        vertical = abs(f32(vecs[a][0])) <= abs(f32(vecs[a][1]))
        acc = np.zeros(n_det)
        if vertical:
            det_step = dsize / c
            slice_step = s / c
            dist = out_scale / abs(c)
            p0 = (det - 0.5 * n_det + 0.5 - off) * det_step + 0.5 * W
            for r in range(H):
                fp = p0 + (r - 0.5 * H + 0.5) * slice_step
                acc += tex1d(vol[r], fp)
        else:
            det_step = -dsize / s
            slice_step = c / s
            dist = out_scale / abs(s)
            p0 = (det - 0.5 * n_det + 0.5 - off) * det_step + 0.5 * H
            for col in range(W):
                fp = p0 + (col - 0.5 * W + 0.5) * slice_step
                acc += tex1d(vol[:, col], fp)
        sino[a] = acc * dist
    return sino


def _sub_det(rays):
    """Sub-ray detector offsets relative to the pixel centre."""
    return (np.arange(rays) + 0.5) / rays - 0.5


def _shift_det(vecs, o):
    """Vectors with the detector moved by o pixels along U (the parallel
    orientation test uses only the ray direction, so it is unchanged)."""
    v = np.array(vecs, dtype=np.float64)
    v[:, 2:4] += o * v[:, 4:6]
    return v


def par_fp_T(sino, vecs, shape, out_scale=1.0):
    """Exact transpose of ``par_fp`` (the matched adjoint: the same samples
    and weights, scattered)."""
    sino = np.asarray(sino, dtype=np.float64)
    H, W = shape
    n_det = sino.shape[1]
    vol = np.zeros((H, W))
    det = np.arange(n_det, dtype=np.float64)
    for a in range(vecs.shape[0]):
        theta, dsize, off = par_parameters(vecs[a], n_det)
        c, s = math.cos(theta), math.sin(theta)
        vertical = abs(f32(vecs[a][0])) <= abs(f32(vecs[a][1]))
        if vertical:
            vals = sino[a] * (out_scale / abs(c))
            p0 = (det - 0.5 * n_det + 0.5 - off) * (dsize / c) + 0.5 * W
            for r in range(H):
                vol[r] += tex1d_T(W, p0 + (r - 0.5 * H + 0.5) * (s / c), vals)
        else:
            vals = sino[a] * (out_scale / abs(s))
            p0 = (det - 0.5 * n_det + 0.5 - off) * (-dsize / s) + 0.5 * H
            for col in range(W):
                vol[:, col] += tex1d_T(H, p0 + (col - 0.5 * W + 0.5) * (c / s), vals)
    return vol


def par_bp(sino, vecs, shape, out_scale=1.0, ss=1):
    """2D parallel voxel-driven BP.

    ``devBP`` + ``transferConstants`` (cuda/2d/par_bp.cu:52-87, 163-191).
    T = x c - y_row s + off, weight |r|/|d| per angle.  ss > 1: ss x ss
    sub-pixels at offsets (k + 0.5) / ss, averaged (``devBP_SS``, :90-137).
    """
    sino = np.asarray(sino, dtype=np.float64)
    H, W = shape
    vol = np.zeros((H, W))
    for ox in (np.arange(ss) + 0.5) / ss:
        for oy in (np.arange(ss) + 0.5) / ss:
            X = np.arange(W) - 0.5 * W + ox
            Y = np.arange(H) - 0.5 * H + oy
            fx, fy = np.meshgrid(X, Y)
            for a in range(vecs.shape[0]):
                rx, ry, dsx, dsy, ux, uy = (float(x) for x in vecs[a])
                d = ux * ry - uy * rx
                cs = ry / d
                sn = -rx / d
                off = (dsy * rx - dsx * ry) / d
                scale = math.sqrt(rx * rx + ry * ry) / abs(d)
                T = fx * cs - fy * sn + off
                vol += tex1d(sino[a], T) * scale
    return vol * out_scale / (ss * ss)


def _fma32(a, b, c):
    """float32 fused multiply-add, emulated in float64 (a*b exact for the
    integer * float operands used here)."""
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def fan_ray_is_horizontal(vec, det):
    """Per-ray orientation of ``FanFPhorizontal`` (cuda/2d/fan_fp.cu:79-83).

    The reference adds 0.5 in *world* units (a quirk kept verbatim) and
    compares in float32; ``DetS + det*U`` is formed as one fused
    multiply-add, as nvcc contracts it.
    """
    sx, sy, dsx, dsy, ux, uy = (f32(x) for x in vec)
    fdx = abs(f32(_fma32(det, ux, dsx) + f32(0.5)) - sx)
    fdy = abs(f32(_fma32(det, uy, dsy) + f32(0.5)) - sy)
    return not (fdy > fdx)


def fan_fp(vol, vecs, n_det, out_scale=1.0, rays=1):
    """2D fan-beam Joseph FP.

    ``FanFPhorizontal`` / ``FanFPvertical`` (cuda/2d/fan_fp.cu:55-193).
    Horizontal rays sample columns (y = alpha x + beta), vertical rays sample
    rows (x = alpha y + beta).  rays > 1: sub-rays through det + (k + 0.5) /
    rays, averaged; the orientation is decided once per detector
    (fan_fp.cu:79-96).
    """
    if rays > 1:
        return sum(_fan_fp_one(vol, vecs, n_det, out_scale, o) for o in (np.arange(rays) + 0.5) / rays) / rays
    return _fan_fp_one(vol, vecs, n_det, out_scale, 0.5)


def _fan_fp_one(vol, vecs, n_det, out_scale, off):
    vol = np.asarray(vol, dtype=np.float64)
    H, W = vol.shape
    nA = vecs.shape[0]
    sino = np.zeros((nA, n_det))
    cols = np.arange(W, dtype=np.float64)
    rows = np.arange(H, dtype=np.float64)
    tdet = np.arange(n_det)
    for a in range(nA):
        sx, sy, dsx, dsy, ux, uy = (float(x) for x in vecs[a])
        hor = np.array([fan_ray_is_horizontal(vecs[a], t) for t in tdet])
        fdet = tdet + off
        dx = dsx + fdet * ux
        dy = dsy + fdet * uy
        if hor.any():
            h = np.nonzero(hor)[0]
            alpha = (sy - dy[h]) / (sx - dx[h])
            beta = sy - alpha * sx
            dist = np.sqrt(alpha * alpha + 1.0) * out_scale
            fy = -alpha[:, None] * (cols[None, :] - 0.5 * W + 0.5) - beta[:, None] + 0.5 * H
            v = tex2d(vol, np.broadcast_to(cols[None, :] + 0.5, fy.shape), fy).sum(axis=1)
            sino[a, h] = v * dist
        if (~hor).any():
            g = np.nonzero(~hor)[0]
            alpha = (sx - dx[g]) / (sy - dy[g])
            beta = sx - alpha * sy
            dist = np.sqrt(alpha * alpha + 1.0) * out_scale
            fxx = -alpha[:, None] * (rows[None, :] - 0.5 * H + 0.5) + beta[:, None] + 0.5 * W
            v = tex2d(vol, fxx, np.broadcast_to(rows[None, :] + 0.5, fxx.shape)).sum(axis=1)
            sino[a, g] = v * dist
    return sino


def fan_fp_T(sino, vecs, shape, out_scale=1.0):
    """Exact transpose of ``fan_fp`` (matched adjoint)."""
    sino = np.asarray(sino, dtype=np.float64)
    H, W = shape
    n_det = sino.shape[1]
    vol = np.zeros((H, W))
    cols = np.arange(W, dtype=np.float64)
    rows = np.arange(H, dtype=np.float64)
    tdet = np.arange(n_det)
    for a in range(vecs.shape[0]):
        sx, sy, dsx, dsy, ux, uy = (float(x) for x in vecs[a])
        hor = np.array([fan_ray_is_horizontal(vecs[a], t) for t in tdet])
        fdet = tdet + 0.5
        dx = dsx + fdet * ux
        dy = dsy + fdet * uy
        if hor.any():
            h = np.nonzero(hor)[0]
            alpha = (sy - dy[h]) / (sx - dx[h])
            beta = sy - alpha * sx
            vals = sino[a, h] * np.sqrt(alpha * alpha + 1.0) * out_scale
            fy = -alpha[:, None] * (cols[None, :] - 0.5 * W + 0.5) - beta[:, None] + 0.5 * H
            vol += tex2d_T((H, W), np.broadcast_to(cols[None, :] + 0.5, fy.shape), fy,
                           np.broadcast_to(vals[:, None], fy.shape))
        if (~hor).any():
            g = np.nonzero(~hor)[0]
            alpha = (sx - dx[g]) / (sy - dy[g])
            beta = sx - alpha * sy
            vals = sino[a, g] * np.sqrt(alpha * alpha + 1.0) * out_scale
            fxx = -alpha[:, None] * (rows[None, :] - 0.5 * H + 0.5) + beta[:, None] + 0.5 * W
            vol += tex2d_T((H, W), fxx, np.broadcast_to(rows[None, :] + 0.5, fxx.shape),
                           np.broadcast_to(vals[:, None], fxx.shape))
    return vol


def fan_bp_params(vec, fbp=False):
    """``DevFanParams`` (cuda/2d/fan_bp.cu:225-279), float64."""
    sx, sy, dsx, dsy, ux, uy = (float(x) for x in vec)

    def det2(ax, ay, bx, by):
        return ax * by - ay * bx

    smdx, smdy = sx - dsx, sy - dsy
    if not fbp:
        scale = math.hypot(ux, uy) / det2(ux, uy, smdx, smdy)
    else:
        scale = 1.0 / det2(ux, uy, sx, sy)
    num_c = scale * det2(sx, sy, dsx, dsy)
    num_x = scale * smdy
    num_y = -scale * smdx
    den_c = scale * det2(ux, uy, sx, sy)
    den_x = scale * uy
    den_y = -scale * ux
    return num_c, num_x, num_y, den_c, den_x, den_y


def fan_bp(sino, vecs, shape, out_scale=1.0, fbp_weighted=False, ss=1):
    """2D fan-beam voxel-driven BP: ``devFanBP<FBPWEIGHT>``
    (cuda/2d/fan_bp.cu:57-103).  Weight fr (BP) or fr^2 with DenC = 1 (FBP).
    ss > 1 (BP only): ss x ss sub-pixels, averaged (``devFanBP_SS``, :106-161).
    """
    sino = np.asarray(sino, dtype=np.float64)
    H, W = shape
    if fbp_weighted:
        ss = 1
    vol = np.zeros((H, W))
    for ox in (np.arange(ss) + 0.5) / ss:
        for oy in (np.arange(ss) + 0.5) / ss:
            X = np.arange(W) - 0.5 * W + ox
            Y = -(np.arange(H) - 0.5 * H + oy)
            fx, fy = np.meshgrid(X, Y)
            for a in range(vecs.shape[0]):
                nc, nx, ny, dc, dx, dy = fan_bp_params(vecs[a], fbp_weighted)
                num = nc + nx * fx + ny * fy
                den = (1.0 if fbp_weighted else dc) + dx * fx + dy * fy
                fr = 1.0 / den
                T = num * fr
                w = fr * fr if fbp_weighted else fr
                vol += tex1d(sino[a], T) * w
    return vol * out_scale / (ss * ss)


def sart_bp(row, vec, shape, kind, out_scale=1.0):
    """SART's BP of one sinogram row (``devBP_SART`` / ``devFanBP_SART``,
    cuda/2d/par_bp.cu:139-159, cuda/2d/fan_bp.cu:167-199): the detector
    position of every pixel as in the full BP, without the per-angle or fr
    weight."""
    row = np.asarray(row, dtype=np.float64)
    H, W = shape
    X = np.arange(W) - 0.5 * W + 0.5
    if kind == "parallel":
        Y = np.arange(H) - 0.5 * H + 0.5
        fx, fy = np.meshgrid(X, Y)
        rx, ry, dsx, dsy, ux, uy = (float(x) for x in vec)
        d = ux * ry - uy * rx
        T = fx * (ry / d) - fy * (-rx / d) + (dsy * rx - dsx * ry) / d
    else:
        Y = -(np.arange(H) - 0.5 * H + 0.5)
        fx, fy = np.meshgrid(X, Y)
        nc, nx, ny, dc, dx, dy = fan_bp_params(vec, False)
        T = (nc + nx * fx + ny * fy) / (dc + dx * fx + dy * fy)
    return tex1d(row, T) * out_scale
