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
