"""Known answers for the 2D linear-interpolation (Joseph) projectors, after
the reference's kernel test (tests/python/functional/test_line2d.py:399-513):
a pixel-aligned rectangle of ones in a random volume window, random
parallel / fan (vector) geometries, and per ray the exact line integral of
the linearly interpolated image along the ray's sampling axis:
sum over the rectangle's slices of the trapezoid weight of the crossing
point times the path length per slice."""
import math

import numpy as np


def random_case(kind, rng):
    """(vol_geom args, proj_geom args, data, rectangle) of one random case."""
    shape = rng.integers(8, 64, size=2)              # (nx, ny)
    rect_min = [int(rng.integers(0, a)) for a in shape]
    rect_max = [int(rng.integers(rect_min[i] + 1, shape[i] + 1)) for i in range(2)]
    pix = 0.5 + rng.random()                           # square pixels
    origin = 10 * rng.random(size=2)
    nx, ny = int(shape[0]), int(shape[1])
    vol = (ny, nx, origin[0] - 0.5 * nx * pix, origin[0] + 0.5 * nx * pix,
           origin[1] - 0.5 * ny * pix, origin[1] + 0.5 * ny * pix)
    data = np.zeros((ny, nx), dtype=np.float32)
    data[rect_min[1]:rect_max[1], rect_min[0]:rect_max[0]] = 1
    ndet = int(rng.integers(8, 64))
    if kind == "parallel":
        w = 0.8 + 0.4 * rng.random()
        proj = ("parallel", w, ndet, np.linspace(0, 2 * np.pi, int(rng.integers(8, 64)), endpoint=False))
    elif kind == "fanflat":
        w = 0.6 + 0.8 * rng.random()
        proj = ("fanflat", w, ndet, np.linspace(0, 2 * np.pi, int(rng.integers(8, 64)), endpoint=False),
                256 * (0.5 + rng.random()), 256 * rng.random())
    else:
        w = 0.6 + 0.8 * rng.random()
        vecs = np.zeros((16, 6))
        for i in range(16):
            a1 = 2 * np.pi * rng.random()
            a2 = a1 + 0.5 * rng.random()                 # oblique detector
            detc = 10 * rng.random(size=2)
            detu = [math.cos(a1) * w, math.sin(a1) * w]
            if kind == "parallel_vec":
                ln = 0.6 + 0.8 * rng.random()
                first = [math.sin(a2) * ln, -math.cos(a2) * ln]
            else:
                d = 256 * (0.5 + rng.random())
                first = [math.sin(a2) * d, -math.cos(a2) * d]
            vecs[i] = [first[0], first[1], detc[0], detc[1], detu[0], detu[1]]
        proj = (kind, ndet, vecs)
    box = (origin[0] + (-0.5 * nx + rect_min[0]) * pix, origin[0] + (-0.5 * nx + rect_max[0]) * pix,
           origin[1] + (0.5 * ny - rect_max[1]) * pix, origin[1] + (0.5 * ny - rect_min[1]) * pix)
    return vol, proj, data, (box, rect_min, rect_max, origin, pix, nx, ny)


def _rays(proj):
    """Per ray, two points (a, b) on its centre line, in sinogram order."""
    kind = proj[0]
    if kind in ("parallel", "fanflat"):
        w, ndet, angles = proj[1], proj[2], proj[3]
        rows = []
        for t in angles:
            u = np.array([math.cos(t), math.sin(t)]) * w
            if kind == "parallel":
                first, detc = np.array([math.sin(t), -math.cos(t)]), np.zeros(2)
            else:
                first = np.array([math.sin(t), -math.cos(t)]) * proj[4]
                detc = np.array([-math.sin(t), math.cos(t)]) * proj[5]
            rows.append((first, detc, u))
    else:
        ndet = proj[1]
        rows = [(v[0:2], v[2:4], v[4:6]) for v in proj[2]]
    par = kind.startswith("parallel")
    for first, detc, u in rows:
        base = detc + (0.5 - 0.5 * ndet) * u
        for i in range(ndet):
            det = base + i * u
            yield (det - first if par else first), det


def _trapezoid(t, lo, hi, width):
    """Linear interpolation at t of a unit box on pixel centres in [lo, hi]."""
    if t < lo - 0.5 * width or t >= hi + 0.5 * width:
        return 0.0
    if t < lo + 0.5 * width:
        return (t - (lo - 0.5 * width)) / width
    if t < hi - 0.5 * width:
        return 1.0
    return (hi + 0.5 * width - t) / width


def linear_kernel_sinogram(proj, rect, shape, vertical=None):
    """Known-answer sinogram; `vertical` (per ray, optional) overrides the
    sampling orientation |dy| >= |dx| of the exact centre line."""
    (xmin, xmax, ymin, ymax), rect_min, rect_max, origin, pix, nx, ny = rect
    out = np.zeros(shape)
    flat = out.reshape(-1)
    for k, (a, b) in enumerate(_rays(proj)):
        xd, yd = b[0] - a[0], b[1] - a[1]
        acc = 0.0
        horizontal = abs(xd) > abs(yd) if vertical is None else not vertical[k]
        if horizontal:              # steps along x: slices are columns
            length = math.sqrt(1.0 + (yd / xd) ** 2) * pix
            for j in range(rect_min[0], rect_max[0]):
                x = origin[0] + (-0.5 * nx + j + 0.5) * pix
                y = a[1] + (x - a[0]) * yd / xd
                acc += _trapezoid(y, ymin, ymax, pix) * length
        else:                        # steps along y: slices are rows
            length = math.sqrt(1.0 + (xd / yd) ** 2) * pix
            for j in range(rect_min[1], rect_max[1]):
                y = origin[1] + (0.5 * ny - j - 0.5) * pix
                x = a[0] + (y - a[1]) * xd / yd
                acc += _trapezoid(x, xmin, xmax, pix) * length
        flat[k] = acc
    return out


def fan_kernel_vertical(normalised_rows, ndet):
    """Per-ray sampling orientation of the fan FP (cuda/2d/fan_fp.cu:79-83):
    float32 |dy| > |dx| of DetS + det * U + 0.5 - src in normalised units.
    Within about half a voxel of 45 degrees it differs from the exact centre
    line's orientation, and the two linear-kernel integrals differ there by
    more than the tolerance, so the fan known answers follow the kernel."""
    out = []
    f32 = np.float32
    for r in np.asarray(normalised_rows).astype(f32):
        sx, sy, dsx, dsy, ux, uy = r
        for d in range(ndet):
            tx = f32(f32(d) * ux + dsx) + f32(0.5) - sx
            ty = f32(f32(d) * uy + dsy) + f32(0.5) - sy
            out.append(bool(abs(ty) > abs(tx)))
    return out


def oracle_vectors(proj, vol_geom):
    """(oracle kind, normalised vectors, output scale) of a random case."""
    from oracle import geometry as OG
    kind = proj[0]
    if kind == "parallel":
        ov, ok = OG.gen_par_projections(proj[3], proj[2], proj[1]), "par"
    elif kind == "fanflat":
        ov, ok = OG.gen_fan_projections(proj[3], proj[2], proj[4], proj[5], proj[1]), "fan"
    elif kind == "parallel_vec":
        ov, ok = OG.parvec_from_astra(proj[2], proj[1]), "par"
    else:
        ov, ok = OG.fanvec_from_astra(proj[2], proj[1]), "fan"
    nv, osc = OG.normalise_2d(ov, ok, vol_geom)
    return ok, nv, osc
