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
    """Zero-pad each row to N, multiply the spectrum by the filter and keep
    the first len(row) samples (cuda/3d/fdk.cu:351-398,
    cuda/2d/fft.cu:52-131): R2C, x filter, C2R, / N.

    H is a real half spectrum (N/2 + 1 bins, every row), or an array
    [n_spec, N/2 + 1] (possibly complex) of per-row half spectra: row r of
    the flattened rows uses H[r % n_spec]."""
    rows = np.asarray(rows, dtype=np.float64)
    H = np.asarray(H)
    n = 2 * (H.shape[-1] - 1)
    m = rows.shape[-1]
    pad = np.zeros(rows.shape[:-1] + (n,))
    pad[..., :m] = rows
    spec = np.fft.rfft(pad, axis=-1)
    if H.ndim == 1:
        spec = spec * H
    else:
        flat = spec.reshape(-1, spec.shape[-1])
        flat = flat * H[np.arange(flat.shape[0]) % H.shape[0]]
        spec = flat.reshape(spec.shape)
    return np.fft.irfft(spec, n=n, axis=-1)[..., :m]


# windows of genFilter (src/Filters.cpp:151-472) for bins k >= 1 of
# nb = N/2 + 1 bins, w = 2 pi k / N
FILTER_TYPES = ("ram-lak", "shepp-logan", "cosine", "hamming", "hann", "tukey", "lanczos",
                "triangular", "gaussian", "barlett-hann", "blackman", "nuttall", "blackman-harris",
                "blackman-nuttall", "flat-top", "kaiser", "parzen")
_COSINE_SUMS = {"nuttall": (0.355768, 0.487396, 0.144232, 0.012604, 0.0),
                "blackman-harris": (0.35875, 0.48829, 0.14128, 0.01168, 0.0),
                "blackman-nuttall": (0.3635819, 0.4891775, 0.1365995, 0.0106411, 0.0),
                "flat-top": (1.0, 1.93, 1.29, 0.388, 0.032)}


def filter_window(name, nb, n, d=1.0, param=-1.0):
    """Window values for bins k = 1 .. nb-1 (float64)."""
    k = np.arange(1, nb, dtype=np.float64)
    w = 2.0 * math.pi * k / n
    nm1 = nb - 1.0
    if name == "ram-lak":
        return np.ones_like(k)
    if name == "shepp-logan":
        x = w / 2.0 / d
        return np.sin(x) / x
    if name == "cosine":
        return np.cos(w / 2.0 / d)
    if name == "hamming":
        return 0.54 + 0.46 * np.cos(w / d)
    if name == "hann":
        return (1.0 + np.cos(w / d)) / 2.0
    if name == "tukey":
        alpha = 0.5 if param < 0 else param
        half = nb / 2.0
        et, den = alpha * half, (1.0 - alpha) * half
        inside = (half - et <= k) & (k <= half + et)
        return np.where(inside, 1.0, 0.5 * (1.0 + np.cos(math.pi * (k - et) / den)))
    if name == "lanczos":
        x = math.pi * (2.0 * k / nm1 - 1.0)
        return np.where(np.abs(x) > 0.001, np.sin(x) / np.where(x == 0, 1.0, x), 1.0)
    if name == "triangular":
        return 2.0 / nm1 * (nm1 / 2.0 - np.abs(k - nm1 / 2.0))
    if name == "gaussian":
        sigma = 0.4 if param < 0 else param
        q = nm1 / 2.0
        return np.exp(-0.5 * ((k - q) / (sigma * q)) ** 2)
    if name == "barlett-hann":
        return 0.62 - 0.48 * np.abs(k / nm1 - 0.5) - 0.38 * np.cos(2.0 * math.pi * k / nm1)
    if name == "blackman":
        alpha = 0.16 if param < 0 else param
        return ((1.0 - alpha) / 2.0 - 0.5 * np.cos(math.pi * k / nm1)
                + alpha / 2.0 * np.cos(2.0 * math.pi * k / nm1))
    if name in _COSINE_SUMS:
        a = _COSINE_SUMS[name]
        b = math.pi * k / nm1
        return a[0] - a[1] * np.cos(2 * b) + a[2] * np.cos(4 * b) - a[3] * np.cos(6 * b) + a[4] * np.cos(8 * b)
    if name == "kaiser":
        from scipy.special import j0   # the reference evaluates J0 (not I0)
        alpha = 3.0 if param < 0 else param
        x = 2.0 * k / nm1 - 1.0
        return j0(math.pi * alpha * np.sqrt(1.0 - x * x)) / j0(math.pi * alpha)
    if name == "parzen":
        q = k / nm1
        return np.where(q <= 0.5, 1.0 - 6.0 * q * q * (1.0 - q), 2.0 * (1.0 - q) ** 3)
    raise ValueError("unknown filter %r" % name)


def windowed_filter(name, n_det, d=1.0, param=-1.0):
    """``genFilter`` half spectrum: Ram-Lak x window (bins >= 1), zeroed
    where 2 pi k / N > pi d."""
    n = next_pow2(2 * n_det)
    nb = n // 2 + 1
    H = 2.0 * np.real(np.fft.fft(ramp_kernel(n)))[:nb]
    H[1:] *= filter_window(name, nb, n, d, param)
    w = 2.0 * math.pi * np.arange(nb) / n
    H[w > math.pi * d] = 0.0
    return H


def custom_filter(data, n_det, real_space):
    """Per-row half spectra of user filters (cuda/2d/fft.cu:447-595):
    Fourier-domain rows are used as given; a real-space kernel of width W
    is centred at W/2, wrapped into N samples and transformed."""
    data = np.atleast_2d(np.asarray(data, dtype=np.float64))
    n = next_pow2(2 * n_det)
    if not real_space:
        return data
    width = data.shape[1]
    used = min(width, n)
    start = (width - used) // 2
    shift = width // 2
    out = np.zeros((data.shape[0], n))
    j = np.arange(start, start + used)
    out[:, (j + n - shift) % n] = data[:, j]
    return np.fft.rfft(out, axis=-1)


def parker_weight(beta, gamma, fan):
    """Parker weight of ray (beta, gamma) for half fan angle `fan`
    (devFDK_ParkerWeight, cuda/3d/fdk.cu:136-156), vectorised."""
    beta, gamma = np.broadcast_arrays(np.asarray(beta, np.float64), np.asarray(gamma, np.float64))
    w = np.zeros(beta.shape)
    r1 = (beta > 0) & (beta <= 2.0 * (fan + gamma))
    r2 = ~r1 & (beta > 0) & (beta <= math.pi + 2 * gamma)
    r3 = ~r1 & ~r2 & (beta > 0) & (beta <= math.pi + 2 * fan)
    w[r1] = np.sin((math.pi / 4.0) * beta[r1] / (fan + gamma[r1])) ** 2
    w[r2] = 1.0
    w[r3] = np.sin((math.pi / 4.0) * (math.pi + 2.0 * fan - beta[r3]) / (fan - gamma[r3])) ** 2
    return w


def parker_weights(angles, n_u, sod, odd, du):
    """[angle][u] short-scan weights of FDK_PreWeight (cuda/3d/fdk.cu:201-251):
    angles made relative to the start of the scanned range (wrapped into
    [0, 2 pi)), scaled by range / pi."""
    angles = np.asarray(angles, dtype=np.float64)
    na = len(angles)
    if na < 2:
        return np.ones((na, n_u))
    da = angles[1] - angles[0]
    while da < -math.pi:
        da += 2 * math.pi
    while da >= math.pi:
        da -= 2 * math.pi
    base = angles[0] if da >= 0 else angles[-1]
    rel = np.mod(angles - base, 2 * math.pi)
    rng = abs(rel[-1] - rel[0]) / (na - 1) * na
    D = sod + odd
    fan = abs(math.atan(du * (n_u * 0.5) / D))
    gamma = np.arctan((np.arange(n_u) - 0.5 * n_u + 0.5) * du / D)
    return parker_weight(rel[:, None], gamma[None, :], fan) * (rng / math.pi)


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


def fdk(proj, vecs, shape, vol_scale=(1.0, 1.0, 1.0), out_scale=1.0, d=1.0, H=None,
        short_scan=False):
    """Full FDK (cuda/3d/fdk.cu:415-478): pre-weight (and Parker weights for
    a short scan, with the angles -atan2(srcX, srcY) + pi of fdk.cu:447-450),
    row filter (Ram-Lak unless H is given; per-angle spectra index the
    angle), FDK-weighted cone BP."""
    n_v, n_a, n_u = proj.shape
    p = fdk_preweight(proj, vecs)
    if short_scan:
        sod, odd, du, dv, zshift = fdk_geometry(vecs, n_u, n_v)
        ang = -np.arctan2(vecs[:, 0], vecs[:, 1]) + math.pi
        p = p * parker_weights(ang, n_u, sod, odd, du)[None, :, :]
    p = filter_rows(p, ramp_filter(n_u, d) if H is None else H)
    return proj3d.cone_bp(p, vecs, shape, vol_scale, out_scale, kernel="fdk")


def fbp_fan(sino, vecs, shape, out_scale, pixel_area, H=None, filt=None, short_scan=False):
    """Fan-beam FBP (cuda/2d/fbp.cu:106-148, src/CudaFilteredBackProjection
    Algorithm.cpp:174-239): the fan sinogram is pre-weighted as a one-row
    cone, filtered, and back-projected with the FBP (fr^2) weighting;
    output scale / pixel area.  H: filter half spectra (default Ram-Lak);
    filt = 'none' skips the filter."""
    n_a, n_det = sino.shape
    sod, odd, du = fan_parameters(vecs[-1], n_det)
    D = sod + odd
    w = D * (D / (du * sod)) * (math.pi / 2.0) / n_a
    U = (np.arange(n_det) - 0.5 * n_det + 0.5) * du
    p = np.asarray(sino, dtype=np.float64) * (w / np.sqrt(D * D + U * U))[None, :]
    if short_scan:
        # each angle's getFanParameters angle atan2(uy, ux) (cuda/2d/fbp.cu:117-128)
        ang = np.arctan2(vecs[:, 5], vecs[:, 4])
        p = p * parker_weights(ang, n_det, sod, odd, du)
    if H is not None or filt is None:
        p = filter_rows(p, ramp_filter(n_det) if H is None else H)
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


def fbp_par(sino, vecs, shape, out_scale, pixel_area, H=None, filt=None):
    """Parallel-beam FBP (cuda/2d/fbp.cu:150-189): filter, then BP with
    output scale (pi/2)/nA / pixel area.  H / filt as for ``fbp_fan``."""
    n_a, n_det = sino.shape
    p = np.asarray(sino, dtype=np.float64)
    if H is not None or filt is None:
        p = filter_rows(p, ramp_filter(n_det) if H is None else H)
    return proj2d.par_bp(p, vecs, shape, out_scale * (math.pi / 2) / n_a / pixel_area)

