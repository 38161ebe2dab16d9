"""Seeded ellipse / ellipsoid phantoms used by the benchmark and tests.

Centres uniform in the inscribed disk / ball at <= 0.8 radius, semi-axes
uniform in [0.05, 0.4] * N, rotation uniform in [0, pi), density uniform in
[0, 1], summed and clipped to [0, 1].
"""
import numpy as np


def ellipses_2d(n_rows, n_cols=None, count=10, seed=0):
    n_cols = n_rows if n_cols is None else n_cols
    rng = np.random.default_rng(seed)
    n = min(n_rows, n_cols)
    y, x = np.mgrid[0:n_rows, 0:n_cols].astype(np.float64)
    x = x - (n_cols - 1) / 2.0
    y = y - (n_rows - 1) / 2.0
    img = np.zeros((n_rows, n_cols))
    for _ in range(count):
        r = 0.8 * (n / 2.0) * np.sqrt(rng.random())
        t = 2 * np.pi * rng.random()
        cx, cy = r * np.cos(t), r * np.sin(t)
        a, b = rng.uniform(0.05, 0.4, 2) * n
        phi = np.pi * rng.random()
        rho = rng.random()
        c, s = np.cos(phi), np.sin(phi)
        u = (x - cx) * c + (y - cy) * s
        v = -(x - cx) * s + (y - cy) * c
        img[(u / a) ** 2 + (v / b) ** 2 <= 1.0] += rho
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def ellipses_2d_device(n, count=10, seed=0, device="cuda"):
    """Square n x n version of ellipses_2d rasterised on a torch device."""
    import torch

    rng = np.random.default_rng(seed)
    ar = torch.arange(n, device=device, dtype=torch.float32) - (n - 1) / 2.0
    y, x = ar.view(n, 1), ar.view(1, n)
    img = torch.zeros((n, n), dtype=torch.float32, device=device)
    for _ in range(count):
        r = 0.8 * (n / 2.0) * np.sqrt(rng.random())
        t = 2 * np.pi * rng.random()
        cx, cy = r * np.cos(t), r * np.sin(t)
        a, b = rng.uniform(0.05, 0.4, 2) * n
        phi = np.pi * rng.random()
        rho = rng.random()
        c, s = float(np.cos(phi)), float(np.sin(phi))
        u = (x - cx) * c + (y - cy) * s
        v = -(x - cx) * s + (y - cy) * c
        img += ((u / a) ** 2 + (v / b) ** 2 <= 1.0).to(torch.float32) * float(rho)
    return img.clamp_(0.0, 1.0)


def _ellipsoid_params(n, count, seed):
    rng = np.random.default_rng(seed)
    params = []
    for _ in range(count):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = d * 0.8 * (n / 2.0) * rng.random() ** (1.0 / 3.0)
        ax = rng.uniform(0.05, 0.4, 3) * n
        phi = np.pi * rng.random()
        rho = rng.random()
        params.append((c, ax, phi, rho))
    return params


def ellipsoids_3d_device(n, count=20, seed=0, device="cuda"):
    """Same phantom as ellipsoids_3d, rasterised on a torch device."""
    import torch

    ar = torch.arange(n, device=device, dtype=torch.float32) - (n - 1) / 2.0
    zz, yy, xx = ar.view(n, 1, 1), ar.view(1, n, 1), ar.view(1, 1, n)
    vol = torch.zeros((n, n, n), dtype=torch.float32, device=device)
    for c, ax, phi, rho in _ellipsoid_params(n, count, seed):
        cs, sn = float(np.cos(phi)), float(np.sin(phi))
        u = (xx - float(c[0])) * cs + (yy - float(c[1])) * sn
        v = -(xx - float(c[0])) * sn + (yy - float(c[1])) * cs
        w = zz - float(c[2])
        inside = (u / float(ax[0])) ** 2 + (v / float(ax[1])) ** 2 + (w / float(ax[2])) ** 2 <= 1.0
        vol += inside.to(torch.float32) * float(rho)
    return vol.clamp_(0.0, 1.0)


def ellipsoids_3d(n, count=20, seed=0, chunk=64):
    """n x n x n ellipsoid phantom, [z][y][x]; built slab-wise to bound memory."""
    params = _ellipsoid_params(n, count, seed)
    vol = np.zeros((n, n, n), dtype=np.float32)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    xx = xx - (n - 1) / 2.0
    yy = yy - (n - 1) / 2.0
    for z0 in range(0, n, chunk):
        z1 = min(n, z0 + chunk)
        zz = (np.arange(z0, z1, dtype=np.float64) - (n - 1) / 2.0)[:, None, None]
        slab = np.zeros((z1 - z0, n, n))
        for c, ax, phi, rho in params:
            cs, sn = np.cos(phi), np.sin(phi)
            u = (xx - c[0]) * cs + (yy - c[1]) * sn
            v = -(xx - c[0]) * sn + (yy - c[1]) * cs
            w = zz - c[2]
            slab[((u / ax[0]) ** 2 + (v / ax[1]) ** 2)[None] + (w / ax[2]) ** 2 <= 1.0] += rho
        vol[z0:z1] = np.clip(slab, 0.0, 1.0)
    return vol


# ------------------------------------------------------------ Shepp-Logan --
# The Shepp-Logan ellipses (Shepp & Logan 1974) in the reference's units:
# centre (x, y[, z]) and semi-axes as fractions of the half extent, rotation
# in degrees, density; the modified densities (Toft) replace the originals
# (src/SheppLogan.cpp:187-246).  One table serves 2D (z ignored).
#                x      y       z    ax      ay     az    rot  value  modified
_SHEPP_LOGAN = ((0.0,   0.0,    0.0, 0.69,   0.92,  0.81,   0,  2.00,  1.0),
                (0.0,  -0.0184, 0.0, 0.6624, 0.874, 0.78,   0, -0.98, -0.8),
                (0.22,  0.0,    0.0, 0.11,   0.31,  0.22, -18, -0.02, -0.2),
                (-0.22, 0.0,    0.0, 0.16,   0.41,  0.28,  18, -0.02, -0.2),
                (0.0,   0.35,   0.0, 0.21,   0.25,  0.41,   0,  0.01,  0.1),
                (0.0,   0.1,    0.0, 0.046,  0.046, 0.05,   0,  0.01,  0.1),
                (0.0,  -0.1,    0.0, 0.046,  0.046, 0.05,   0,  0.01,  0.1),
                (-0.08, -0.605, 0.0, 0.046,  0.023, 0.05,   0,  0.01,  0.1),
                (0.0,  -0.605,  0.0, 0.023,  0.023, 0.02,   0,  0.01,  0.1),
                (0.06, -0.605,  0.0, 0.023,  0.046, 0.02,   0,  0.01,  0.1))


def _row_spans(ex, ey, ax, ay, rot, width, height):
    """Per image row, the columns [lo, hi) inside one ellipse, and the
    implicit-function coefficients (A, B(y), C(y)) of its boundary
    A x^2 + B x + C = 0.  Pixel centres sit at integer (x, y); rows whose
    span would start left of column 0 are left out, as the reference's
    unsigned column arithmetic leaves them out (src/SheppLogan.cpp:56-116)."""
    a = (0.5 + 0.5 * ex) * width - 0.5
    b = (0.5 - 0.5 * ey) * height - 0.5
    c2 = (0.5 * ax * width) ** 2
    d2 = (0.5 * ay * height) ** 2
    th = rot * np.pi / 180.0
    co, si = np.cos(th), np.sin(th)
    cc, ss, cs = co * co, si * si, co * si
    # ((x-a)cos - (y-b)sin)^2 / c^2 + ((x-a)sin + (y-b)cos)^2 / d^2 - 1,
    # expanded in x and y
    qxx = cc / c2 + ss / d2
    qx = -2 * a * cc / c2 + 2 * b * cs / c2 - 2 * b * cs / d2 - 2 * a * ss / d2
    qxy = -2 * cs / c2 + 2 * cs / d2
    q0 = (a * a * cc / c2 + b * b * cc / d2 - 2 * a * b * cs / c2 + 2 * a * b * cs / d2
          + b * b * ss / c2 + a * a * ss / d2 - 1)
    qy = -2 * b * cc / d2 + 2 * a * cs / c2 - 2 * a * cs / d2 - 2 * b * ss / c2
    qyy = cc / d2 + ss / c2
    y = np.arange(height, dtype=np.float64)
    A = qxx
    B = y * qxy + qx
    C = y * y * qyy + y * qy + q0
    disc = B * B - 4 * A * C
    ok = disc >= 0
    root = np.sqrt(np.where(ok, disc, 0.0))
    lo = np.ceil(0.5 * (-B - root) / A)
    hi = np.floor(0.5 * (-B + root) / A) + 1
    ok &= (lo >= 0) & (lo < width) & (hi > 0)
    lo = np.where(ok, lo, 0).astype(np.int64)
    hi = np.where(ok, np.minimum(hi, width), 0).astype(np.int64)
    return ok, lo, hi, A, B, C


def shepp_logan_2d(rows, cols, modified=True):
    """[rows][cols] float32 Shepp-Logan phantom on pixel centres, negative
    sums clipped to 0 (generateSheppLogan, src/SheppLogan.cpp:187-222)."""
    img = np.zeros((rows, cols), dtype=np.float32)
    x = np.arange(cols)[None, :]
    for ex, ey, _, ax, ay, _, rot, val, mval in _SHEPP_LOGAN:
        ok, lo, hi, _, _, _ = _row_spans(ex, ey, ax, ay, rot, cols, rows)
        inside = ok[:, None] & (x >= lo[:, None]) & (x < hi[:, None])
        img[inside] += np.float32(mval if modified else val)
    img[img < 0] = 0.0
    return img


def shepp_logan_3d(slices, rows, cols, modified=True):
    """[slices][rows][cols] float32 Shepp-Logan phantom: ellipsoids whose
    x-y section is the 2D ellipse; voxel (x, y) of a section spans the
    slices within ((z - e) / f)^2 <= -q(x, y), q the section's implicit
    function (generateSheppLogan3D, src/SheppLogan.cpp:118-184, 224-246)."""
    vol = np.zeros((slices, rows, cols), dtype=np.float32)
    x = np.arange(cols, dtype=np.float64)[None, :]
    zz = np.arange(slices)[:, None, None]
    for ex, ey, ez, ax, ay, az, rot, val, mval in _SHEPP_LOGAN:
        ok, lo, hi, A, B, C = _row_spans(ex, ey, ax, ay, rot, cols, rows)
        e = (0.5 - 0.5 * ez) * slices - 0.5
        f = 0.5 * az * slices
        inside = ok[:, None] & (x >= lo[:, None]) & (x < hi[:, None])
        q = (A * x) * x + B[:, None] * x + C[:, None]
        with np.errstate(invalid="ignore"):
            half = f * np.sqrt(-q)
        zlo = np.ceil(e - half)
        zhi = np.floor(e + half) + 1
        inside &= (zlo >= 0) & (zlo < slices) & (zhi > 0)   # NaN (q > 0) fails these
        zlo = np.where(inside, zlo, 0).astype(np.int64)
        zhi = np.where(inside, np.minimum(zhi, slices), 0).astype(np.int64)
        mask = inside[None] & (zz >= zlo[None]) & (zz < zhi[None])
        vol[mask] += np.float32(mval if modified else val)
    vol[vol < 0] = 0.0
    return vol
