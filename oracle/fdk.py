"""Oracle FDK / FBP pieces (float64).

Test infrastructure only (see ``oracle/__init__.py``).
"""
import math

import numpy as np

from . import proj2d, proj3d

f32 = np.float32


def next_pow2(n):
    """``calcNextPowerOfTwo`` (src/Filters.cpp:662-669)."""
    x = 1
    while x < n:
        x *= 2
    return x


def ramp_kernel(n):
    """Spatial Ram-Lak kernel of ``generateRampFilter`` (src/Filters.cpp:41-68):
    h[0] = 1/4, h[i odd] = -1/(pi * min(i, n-i))^2, h[i even] = 0."""
    h = np.zeros(n)
    h[0] = 0.25
    i = np.arange(1, n, 2)
    j = np.where(2 * i > n, n - i, i)
    h[i] = -1.0 / (math.pi * j) ** 2
    return h


def ramp_filter(n_det, d=1.0):
    """Half-spectrum filter of ``genFilter`` for FILTER_RAMLAK
    (src/Filters.cpp:133-160, 473-485): 2 Re(DFT(h))[k], k = 0..N/2, zeroed
    where 2 pi k / N > pi d.  N = nextpow2(2 nDet)."""
    n = next_pow2(2 * n_det)
    H = 2.0 * np.real(np.fft.fft(ramp_kernel(n)))[: n // 2 + 1]
    w = 2.0 * math.pi * np.arange(n // 2 + 1) / n
    H[w > math.pi * d] = 0.0
    return H


def filter_rows(rows, H):
    """Zero-pad each row to N, multiply the spectrum by the real filter and
    keep the first len(row) samples (cuda/3d/fdk.cu:351-398,
    cuda/2d/fft.cu:52-131): R2C, x filter, C2R, / N."""
    rows = np.asarray(rows, dtype=np.float64)
    n = 2 * (len(H) - 1)
    m = rows.shape[-1]
    pad = np.zeros(rows.shape[:-1] + (n,))
    pad[..., :m] = rows
    spec = np.fft.rfft(pad, axis=-1) * H
    return np.fft.irfft(spec, n=n, axis=-1)[..., :m]


def fdk_geometry(vecs, n_u, n_v):
    """Scalar cone parameters extracted from angles[0] by ``FDK``
    (cuda/3d/fdk.cu:427-445), all in normalised units."""
    p = vecs[0]
    dcx = p[3] + 0.5 * n_u * p[6]
    dcy = p[4] + 0.5 * n_u * p[7]
    dcz = p[5] + 0.5 * n_v * p[11]
    sod = math.sqrt(p[0] ** 2 + p[1] ** 2)
    odd = math.sqrt(dcx ** 2 + dcy ** 2)
    du = math.sqrt(p[6] ** 2 + p[7] ** 2)
    dv = abs(p[11])
    zshift = dcz - p[2]
    return sod, odd, du, dv, zshift


def fdk_preweight(proj, vecs):
    """``devFDK_preweight`` (cuda/3d/fdk.cu:70-113): p *= W / sqrt(D^2+u^2+v^2)
    with W = D * D/(du SOD) * (pi/2) / nA, D = SOD + ODD."""
    proj = np.array(proj, dtype=np.float64)
    n_v, n_a, n_u = proj.shape
    sod, odd, du, dv, zshift = fdk_geometry(vecs, n_u, n_v)
    D = sod + odd
    w2 = D / (du * sod)
    w = D * w2 * (math.pi / 2.0) / n_a
    U = (np.arange(n_u) - 0.5 * n_u + 0.5) * du
    V = (np.arange(n_v) - 0.5 * n_v + 0.5) * dv + zshift
    weight = w / np.sqrt(D * D + U[None, :] ** 2 + V[:, None] ** 2)
    return proj * weight[:, None, :]


def fdk(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, d=1.0):
    """Full FDK (cuda/3d/fdk.cu:415-478): pre-weight, Ram-Lak row filter,
    FDK-weighted cone BP."""
    n_v, n_a, n_u = proj.shape
    p = fdk_preweight(proj, vecs)
    p = filter_rows(p, ramp_filter(n_u, d))
    return proj3d.cone_bp(p, vecs, shape, vol_scale, out_scale, kernel="fdk")


def fbp_fan(sino, vecs, shape, out_scale, pixel_area):
    """Fan-beam FBP (cuda/2d/fbp.cu:106-148, src/CudaFilteredBackProjection
    Algorithm.cpp:174-239): the fan sinogram is pre-weighted as a one-row
    cone, filtered, and back-projected with the FBP (fr^2) weighting;
    output scale / pixel area."""
    n_a, n_det = sino.shape
    sod, odd, du = fan_parameters(vecs[-1], n_det)
    D = sod + odd
    w = D * (D / (du * sod)) * (math.pi / 2.0) / n_a
    U = (np.arange(n_det) - 0.5 * n_det + 0.5) * du
    p = np.asarray(sino, dtype=np.float64) * (w / np.sqrt(D * D + U * U))[None, :]
    p = filter_rows(p, ramp_filter(n_det))
    return proj2d.fan_bp(p, vecs, shape, out_scale / pixel_area, fbp_weighted=True)


def fan_parameters(vec, n_det):
    """``getFanParameters`` (src/GeometryUtil2D.cpp:138-177): origin-source
    distance, origin-detector distance and detector pixel size."""
    sx, sy, dsx, dsy, ux, uy = (float(x) for x in vec)
    sod = math.sqrt(sx * sx + sy * sy)
    du = math.sqrt(ux * ux + uy * uy)
    t = (-dsx * ux - dsy * uy) / (ux * ux + uy * uy)
    odd = math.sqrt((dsx + t * ux) ** 2 + (dsy + t * uy) ** 2)
    return sod, odd, du


def fbp_par(sino, vecs, shape, out_scale, pixel_area):
    """Parallel-beam FBP (cuda/2d/fbp.cu:150-189): filter, then BP with
    output scale (pi/2)/nA / pixel area."""
    n_a, n_det = sino.shape
    p = filter_rows(sino, ramp_filter(n_det))
    return proj2d.par_bp(p, vecs, shape, out_scale * (math.pi / 2) / n_a / pixel_area)

